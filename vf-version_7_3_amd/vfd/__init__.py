"""Python host binding of libvfd.so (the MI355X cytokine-diffusion engine).

Thin ctypes layer over the C-ABI in include/vfd.h and include/vf_diffusion.h.
Device memory and streams come from torch (plumbing only): every compute call
goes to the HIP kernels in libvfd.so, and importing this module fails loudly
when the library has not been built.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# VFD_LIBRARY selects another build of the library (same C-ABI)
LIB_PATH = os.environ.get("VFD_LIBRARY") or os.path.join(_PKG, "libvfd.so")

VFD_OK = 0
VFD_ERR_INVALID, VFD_ERR_UNSUPPORTED, VFD_ERR_DEVICE = -1, -2, -3
VFD_ERR_NOMEM, VFD_ERR_BUDGET, VFD_ERR_NOTFOUND, VFD_ERR_STATE = -4, -5, -6, -7
F32, F64 = 4, 8
PAD_CONSTANT_VF, PAD_ZERO = 3, 100
KSPEC_REAL, KSPEC_COMPLEX_STD, KSPEC_SEPARABLE = 0, 1, 2
PASS_XFWD, PASS_YFWD, PASS_ZCONV, PASS_YINV, PASS_XINV = 0, 1, 2, 3, 4
PASS_NAMES = ["xfwd", "yfwd", "zconv", "yinv", "xinv"]
# pass orders (vfd_plan_info.order); the x-last order runs the same five pass
# indices as z R2C, y forward, x fused, y inverse, z C2R
ORDER_XFIRST, ORDER_XLAST = 0, 1
XLAST_PASS_NAMES = ["zr2c", "yfwd", "xconv", "yinv", "zc2r"]
# the single launches of the z pass (per-kernel timing); together they are PASS_ZCONV
PASS_Z_DIRECT8, PASS_Z_DIRECT12, PASS_Z_DIRECT16, PASS_Z_FFT = 5, 6, 7, 8
Z_PART_NAMES = ["z_direct8", "z_direct12", "z_direct16", "z_fft"]


class VfdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vfd error {code}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(f"libvfd.so not built at {LIB_PATH}: run __graft_entry__.build() "
                      "(make -C vf-version_7_3_amd)")

lib = ctypes.CDLL(LIB_PATH)

_i, _d, _p, _sz = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t


class PlanDesc(ctypes.Structure):
    _fields_ = [("nz", _i), ("ny", _i), ("nx", _i), ("pz", _i), ("py", _i), ("px", _i),
                ("precision", _i), ("pad_mode", _i), ("kspec_format", _i)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("nz", _i), ("ny", _i), ("nx", _i), ("pz", _i), ("py", _i), ("px", _i),
                ("L", _i), ("hx", _i), ("ld", ctypes.c_long), ("rs_y", _i), ("rs_z", _i),
                ("work_bytes", _sz), ("kspec_bytes", _sz), ("bytes_per_field_tick", _d), ("order", _i),
                ("y_inv_fold", _i)]


class Field(ctypes.Structure):
    _fields_ = [("inp", _p), ("out", _p), ("work", _p), ("kspec", _p), ("base", _d),
                ("epi", _i), ("reserved", _i)]


lib.vfd_last_error.restype = ctypes.c_char_p
lib.vfd_version.restype = _i
lib.vfd_length_supported.argtypes = [_i]
lib.vfd_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(_p)]
lib.vfd_plan_destroy.argtypes = [_p]
lib.vfd_plan_destroy.restype = None
lib.vfd_plan_get_info.argtypes = [_p, ctypes.POINTER(PlanInfo)]
lib.vfd_diffuse.argtypes = [_p, ctypes.POINTER(Field), _i, _p]
lib.vfd_diffuse_pass.argtypes = [_p, ctypes.POINTER(Field), _i, _i, _p]
lib.vfd_plan_pass_bytes.argtypes = [_p, _i]
lib.vfd_plan_pass_bytes.restype = _d
lib.vfd_plan_autotune.argtypes = [_p, _i, _p, ctypes.POINTER(ctypes.c_float)]
lib.vfd_plan_get_schedule.argtypes = [_p, ctypes.POINTER(_i)]
lib.vfd_plan_set_schedule.argtypes = [_p, ctypes.POINTER(_i)]
lib.vfd_gauss_spectrum.argtypes = [_p, _d, _d, _p, _p]
lib.vfd_field_mass.argtypes = [_p, ctypes.POINTER(_p), _i, ctypes.POINTER(_d), _p]
lib.vfd_r2c_3d.argtypes = [_p, _p, _p, _p]
lib.vfd_diffuse_update.argtypes = [_p, ctypes.POINTER(Field), ctypes.POINTER(_p), ctypes.POINTER(ctypes.c_float),
                                   _p, _p, ctypes.POINTER(_p), ctypes.POINTER(_d), _p]
lib.vfd_diffuse_update_noalloc.argtypes = [_p, ctypes.POINTER(Field), ctypes.POINTER(_p),
                                           ctypes.POINTER(ctypes.c_float), _p, _p, ctypes.POINTER(_p), _p, _p]
lib.vfd_diffuse_sample.argtypes =[_p, ctypes.POINTER(Field), _i, ctypes.POINTER(_i), _i, ctypes.POINTER(_p), _p]
lib.vfd_sample_chem.argtypes = [_p, _i, _i, _i, _i, _i, _p, _p]
lib.vfd_chem_update.argtypes = [ctypes.POINTER(_p), ctypes.POINTER(_p), ctypes.POINTER(_p),
                                ctypes.POINTER(ctypes.c_float), _p, _p, ctypes.POINTER(_p), ctypes.c_long,
                                ctypes.POINTER(_d), _p]
lib.vfd_c2r_3d.argtypes = [_p, _p, _p, _p]
lib.vfd_stream_copy.argtypes = [_p, _p, _sz, _p]
lib.vfd_stream_mix.argtypes = [_p, _p, _p, _p, _sz, _p]


def stream_copy(dst, src, stream=None):
    """Copy src to dst (device tensors of equal byte size, a multiple of 16)
    with the library's 16-byte non-temporal streaming kernel."""
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() != nbytes:
        raise ValueError("stream_copy: tensors differ in size")
    check(lib.vfd_stream_copy(_ptr(dst), _ptr(src), nbytes, _stream(stream)))


def stream_mix(dst0, src0, dst1=None, src1=None, stream=None):
    """The streaming kernel for other read:write mixes: dst1 given, src0 is
    written to dst0 and dst1 (1 read : 2 writes); src1 given, dst0 = src0 +
    src1 as 32-bit integer lanes (2 reads : 1 write).  Equal byte sizes."""
    nbytes = src0.numel() * src0.element_size()
    for t in (dst0, dst1, src1):
        if t is not None and t.numel() * t.element_size() != nbytes:
            raise ValueError("stream_mix: tensors differ in size")
    check(lib.vfd_stream_mix(_ptr(dst0), None if dst1 is None else _ptr(dst1), _ptr(src0),
                             None if src1 is None else _ptr(src1), nbytes, _stream(stream)))


def check(rc):
    if rc != VFD_OK:
        raise VfdError(rc, lib.vfd_last_error().decode())
    return rc


def last_error():
    return lib.vfd_last_error().decode()


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class Plan:
    """One grid shape: data box (nz, ny, nx), FFT box (pz, py, px)."""

    def __init__(self, shape, box=None, precision=F32, pad_mode=PAD_ZERO,
                 kspec_format=KSPEC_REAL):
        nz, ny, nx = shape
        pz, py, px = box if box is not None else (0, 0, 0)
        self.desc = PlanDesc(nz, ny, nx, pz, py, px, precision, pad_mode, kspec_format)
        h = ctypes.c_void_p()
        check(lib.vfd_plan_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self.h = h
        self.info = PlanInfo()
        check(lib.vfd_plan_get_info(self.h, ctypes.byref(self.info)))
        self.precision = precision
        self.kspec_format = kspec_format

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:
            lib.vfd_plan_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        return (self.info.nz, self.info.ny, self.info.nx)

    @property
    def box(self):
        return (self.info.pz, self.info.py, self.info.px)

    # ---- device buffers (torch plumbing) ----
    def _dtype(self):
        import torch
        return torch.float32 if self.precision == F32 else torch.float64

    def new_work(self, device="cuda"):
        import torch
        return torch.empty(self.info.work_bytes, dtype=torch.uint8, device=device)

    def new_kspec(self, device="cuda"):
        import torch
        return torch.empty(self.info.kspec_bytes, dtype=torch.uint8, device=device)

    def gauss_spectrum(self, sigma2, decay, out=None, stream=None):
        out = self.new_kspec() if out is None else out
        check(lib.vfd_gauss_spectrum(self.h, float(sigma2), float(decay), _ptr(out), _stream(stream)))
        return out

    @staticmethod
    def field_array(fields):
        """fields: list of dicts with keys in, out, work, kspec (torch tensors),
        optional base, epi -> ctypes array of vfd_field.  The array keeps the
        tensors referenced, so their memory lives as long as the array does."""
        arr = (Field * len(fields))()
        for i, f in enumerate(fields):
            arr[i] = Field(_ptr(f["in"]), _ptr(f["out"]), _ptr(f["work"]), _ptr(f["kspec"]),
                           float(f.get("base", 0.0)), int(f.get("epi", 0)), 0)
        arr._tensors = [f[k] for f in fields for k in ("in", "out", "work", "kspec")]
        return arr

    def diffuse(self, fields, stream=None):
        arr = fields if isinstance(fields, ctypes.Array) else self.field_array(fields)
        check(lib.vfd_diffuse(self.h, arr, len(arr), _stream(stream)))

    def diffuse_pass(self, fields, pass_index, stream=None):
        arr = fields if isinstance(fields, ctypes.Array) else self.field_array(fields)
        check(lib.vfd_diffuse_pass(self.h, arr, len(arr), int(pass_index), _stream(stream)))

    def diffuse_update(self, fields, deltas, decay, frag_col, frag_ela, grads, stream=None, totals=True):
        """One tick of the 8 chem fields with the chem update fused: fields[i]['out']
        receives deltas[i] + t * decay[i], deltas are cleared, grads (NEU, MAC, FIB)
        and the 8 totals are formed.  Returns the totals unless totals=False."""
        arr = fields if isinstance(fields, ctypes.Array) else self.field_array(fields)
        if len(arr) != 8 or len(deltas) != 8 or len(decay) != 8 or len(grads) != 3:
            raise ValueError("diffuse_update: 8 fields, 8 deltas, 8 decays and 3 gradient fields")
        dv = (ctypes.c_void_p * 8)(*[x.data_ptr() for x in deltas])
        gv = (ctypes.c_void_p * 3)(*[x.data_ptr() for x in grads])
        dec = (ctypes.c_float * 8)(*[float(x) for x in decay])
        out = (ctypes.c_double * 8)()
        check(lib.vfd_diffuse_update(self.h, arr, dv, dec, _ptr(frag_col), _ptr(frag_ela), gv,
                                     out if totals else None, _stream(stream)))
        return list(out) if totals else None

    def diffuse_update_noalloc(self, fields, deltas, decay, frag_col, frag_ela, grads, d_totals=None, stream=None):
        """diffuse_update without device allocation: the 8 totals go to the
        FP64 device tensor d_totals (8 elements, optional); asynchronous."""
        arr = fields if isinstance(fields, ctypes.Array) else self.field_array(fields)
        if len(arr) != 8 or len(deltas) != 8 or len(decay) != 8 or len(grads) != 3:
            raise ValueError("diffuse_update_noalloc: 8 fields, 8 deltas, 8 decays and 3 gradient fields")
        if d_totals is not None and (d_totals.numel() < 8 or d_totals.element_size() != 8 or not d_totals.is_cuda):
            raise ValueError("diffuse_update_noalloc: d_totals must be a device tensor of 8 FP64 values")
        dv = (ctypes.c_void_p * 8)(*[x.data_ptr() for x in deltas])
        gv = (ctypes.c_void_p * 3)(*[x.data_ptr() for x in grads])
        dec = (ctypes.c_float * 8)(*[float(x) for x in decay])
        check(lib.vfd_diffuse_update_noalloc(self.h, arr, dv, dec, _ptr(frag_col), _ptr(frag_ela), gv,
                                             _ptr(d_totals), _stream(stream)))

    def diffuse_sample(self, fields, chem_index, stride, vis, stream=None):
        """diffuse() with the visualisation downsample of each field's output
        (sample_chem with chem_index[i]) written into vis[i]."""
        arr = fields if isinstance(fields, ctypes.Array) else self.field_array(fields)
        n = len(arr)
        if len(chem_index) != n or len(vis) != n:
            raise ValueError("diffuse_sample: one chem index and one output array per field")
        ci = (ctypes.c_int * max(1, n))(*[int(c) for c in chem_index])
        vv = (ctypes.c_void_p * max(1, n))(*[x.data_ptr() for x in vis])
        check(lib.vfd_diffuse_sample(self.h, arr, n, ci, int(stride), vv, _stream(stream)))

    def pass_bytes(self, pass_index):
        return lib.vfd_plan_pass_bytes(self.h, int(pass_index))

    def autotune(self, nfields=8, stream=None):
        """Pick the fastest compiled schedule per pass; returns (variants, ms)."""
        ms = (ctypes.c_float * 5)()
        check(lib.vfd_plan_autotune(self.h, int(nfields), _stream(stream), ms))
        return self.schedule, list(ms)

    @property
    def schedule(self):
        v = (ctypes.c_int * 5)()
        check(lib.vfd_plan_get_schedule(self.h, v))
        return list(v)

    @schedule.setter
    def schedule(self, variants):
        v = (ctypes.c_int * 5)(*variants)
        check(lib.vfd_plan_set_schedule(self.h, v))

    def mass(self, tensors, stream=None):
        n = len(tensors)
        ptrs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in tensors])
        out = (ctypes.c_double * n)()
        check(lib.vfd_field_mass(self.h, ptrs, n, out, _stream(stream)))
        return list(out)

    def r2c(self, real, cplx, stream=None):
        check(lib.vfd_r2c_3d(self.h, _ptr(real), _ptr(cplx), _stream(stream)))

    def c2r(self, cplx, real, stream=None):
        check(lib.vfd_c2r_3d(self.h, _ptr(cplx), _ptr(real), _stream(stream)))


def length_supported(q):
    return bool(lib.vfd_length_supported(int(q)))


def chem_update(t, d, p, decay, frag_col, frag_ela, grads, stream=None, totals=True):
    """WHChemical::update (gpuUpdate) on the device for 8 chems of n voxels:
    p = d + t * decay, d = 0, gradients (NEU, MAC, FIB) from p and the ECM
    fragment fields.  t, d, p: 8 FP32 device tensors each (p may be t);
    grads: 3 FP32 device tensors.  Returns the 8 compensated totals of p (and
    synchronizes the stream) unless totals=False."""
    n = t[0].numel()
    for a in list(t) + list(d) + list(p) + [frag_col, frag_ela] + list(grads):
        if a.numel() != n or not a.is_contiguous():
            raise ValueError("chem_update: every array must be contiguous with the same voxel count")
    P8 = ctypes.c_void_p * 8
    tv, dv, pv = P8(*[x.data_ptr() for x in t]), P8(*[x.data_ptr() for x in d]), P8(*[x.data_ptr() for x in p])
    gv = (ctypes.c_void_p * 3)(*[x.data_ptr() for x in grads])
    dec = (ctypes.c_float * 8)(*[float(x) for x in decay])
    out = (ctypes.c_double * 8)()
    check(lib.vfd_chem_update(tv, dv, pv, dec, _ptr(frag_col), _ptr(frag_ela), gv, n,
                              out if totals else None, _stream(stream)))
    return list(out) if totals else None


def sample_chem(field, chem_index, stride=3, out=None, stream=None):
    """sampleChem: every stride-th voxel of a [D][H][W] FP32 device field times
    the chem's display scale, clamped at 1 -> [D//stride][H//stride][W//stride]."""
    import torch
    D, H, W = field.shape
    if out is None:
        out = torch.empty((D // stride, H // stride, W // stride), dtype=torch.float32, device=field.device)
    check(lib.vfd_sample_chem(_ptr(field), D, H, W, int(chem_index), int(stride), _ptr(out), _stream(stream)))
    return out
