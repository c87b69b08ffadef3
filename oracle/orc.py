"""ctypes front end of the CPU oracle (oracle/vfd_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU baseline),
never as the measured or shipped path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

PAD_CLAMP, PAD_RIGHTWALL, PAD_MIRROR, PAD_CONSTANT_VF, PAD_ZERO = 0, 1, 2, 3, 100

_d = ctypes.c_double
_i = ctypes.c_int
_l = ctypes.c_long
_p = ctypes.c_void_p


def build():
    """Compile the oracle with gcc (oracle/Makefile)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_snap_transform_size.argtypes = [_i]
        L.orc_snap_transform_size.restype = _i
        L.orc_fft_c2c.argtypes = [_i, _p, _i]
        L.orc_rfft3.argtypes = [_i, _i, _i, _p, _p]
        L.orc_irfft3.argtypes = [_i, _i, _i, _p, _p]
        L.orc_pad_constant_vf3d.argtypes = [_p, _p, _i, _i, _i, _i, _i, _i, _i, _d]
        L.orc_pad_zero3d.argtypes = [_p, _p, _i, _i, _i, _i, _i, _i]
        L.orc_pad_clamp3d.argtypes = [_p, _p] + [_i] * 9
        L.orc_pad_kernel3d.argtypes = [_p, _p] + [_i] * 9
        L.orc_unpad3d.argtypes = [_p, _p] + [_i] * 6
        L.orc_extract_center3d.argtypes = [_p, _p] + [_i] * 6
        L.orc_kernel_spectrum3d.argtypes = [_p] + [_i] * 9 + [_p]
        L.orc_fft_diffuse3d.argtypes = [_p, _i, _i, _i, _p, _i, _i, _i, _i, _i, _d, _p]
        L.orc_direct_conv3d.argtypes = [_p, _i, _i, _i, _p] + [_i] * 10 + [_i, _d, _p]
        L.orc_gauss_window1d.argtypes = [_d, _i, _p]
        L.orc_gauss_radius.argtypes = [_d]
        L.orc_gauss_radius.restype = _i
        L.orc_direct_separable3d.argtypes = [_p, _i, _i, _i, _d, _d, _p]
        L.orc_direct_separable3d_f64.argtypes = [_p, _i, _i, _i, _d, _d, _p]
        L.orc_gauss_window3d.argtypes = [_d, _d, _p, _i]
        L.orc_compute_kernel3d.argtypes = [_i, _d, _d, _d, _i] + [_i] * 9 + [_i, _p]
        L.orc_stencil_heat_window.argtypes = [_d, _d, _d, _i, _i, _i, _i, _i, _p]
        L.orc_neumaier_sum_f32.argtypes = [_p, _l]
        L.orc_neumaier_sum_f32.restype = _d
        L.orc_neumaier_sum_f64.argtypes = [_p, _l]
        L.orc_neumaier_sum_f64.restype = _d
        L.orc_synth_field.argtypes = [_p, _i, _i, _i, ctypes.c_uint64, _d, _d, _d, _d, _i, _d, _d]
        L.orc_synth_params.argtypes = [ctypes.c_uint64, _i, _d, _d, _d, _d,
                                       ctypes.POINTER(_d), ctypes.POINTER(_d)]
        L.cpu_fft_diffuse3d_f32.argtypes = [_p, _p, _i, _i, _i, _p, _i]
        L.orc_gauss_spectrum_real.argtypes = [_d, _d, _i, _i, _i, _d, _p]
        L.orc_num_threads.restype = _i
        L.orc_sample_chem.argtypes = [_p, _i, _i, _i, _i, _i, _p]
        L.orc_chem_update.argtypes = [ctypes.POINTER(_p), ctypes.POINTER(_p), ctypes.POINTER(_p),
                                      ctypes.POINTER(ctypes.c_float), _p, _p, _p, _p, _p, ctypes.c_long,
                                      ctypes.POINTER(_d)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def snap_transform_size(n):
    return lib().orc_snap_transform_size(int(n))


def fft(x, inverse=False):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    rc = lib().orc_fft_c2c(a.size, _ptr(a), 1 if inverse else 0)
    assert rc == 0
    return a


def rfft3(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    nz, ny, nx = a.shape
    out = np.empty((nz, ny, nx // 2 + 1), dtype=np.complex128)
    assert lib().orc_rfft3(nz, ny, nx, _ptr(a), _ptr(out)) == 0
    return out


def irfft3(h, nx):
    h = np.ascontiguousarray(h, dtype=np.complex128)
    nz, ny, _ = h.shape
    out = np.empty((nz, ny, nx), dtype=np.float64)
    assert lib().orc_irfft3(nz, ny, nx, _ptr(h), _ptr(out)) == 0
    return out


def pad_constant_vf(u, box, epi, base):
    u = np.ascontiguousarray(u, dtype=np.float32)
    D, H, W = u.shape
    FD, FH, FW = box
    out = np.empty(box, dtype=np.float64)
    lib().orc_pad_constant_vf3d(_ptr(out), _ptr(u), FD, FH, FW, D, H, W, epi, base)
    return out


def pad_zero(u, box):
    u = np.ascontiguousarray(u, dtype=np.float32)
    D, H, W = u.shape
    out = np.empty(box, dtype=np.float64)
    lib().orc_pad_zero3d(_ptr(out), _ptr(u), *box, D, H, W)
    return out


def pad_clamp(d, box, KZ, KY, KX):
    d = np.ascontiguousarray(d, dtype=np.float64)
    D, H, W = d.shape
    out = np.empty(box, dtype=np.float64)
    lib().orc_pad_clamp3d(_ptr(out), _ptr(d), *box, D, H, W, KZ, KY, KX)
    return out


def pad_kernel(k, box, KZ, KY, KX):
    k = np.ascontiguousarray(k, dtype=np.float64)
    out = np.zeros(box, dtype=np.float64)
    lib().orc_pad_kernel3d(_ptr(out), _ptr(k), *box, *k.shape, KZ, KY, KX)
    return out


def unpad(p, shape):
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty(shape, dtype=np.float64)
    lib().orc_unpad3d(_ptr(out), _ptr(p), *shape, *p.shape)
    return out


def extract_center(d, wshape):
    d = np.ascontiguousarray(d, dtype=np.float64)
    out = np.empty(wshape, dtype=np.float64)
    lib().orc_extract_center3d(_ptr(out), _ptr(d), *d.shape, *wshape)
    return out


def kernel_spectrum(window, centre, box):
    """computeKernelSpectrum3D: window centre to origin, R2C on the box."""
    w = np.ascontiguousarray(window, dtype=np.float64)
    FD, FH, FW = box
    out = np.empty((FD, FH, FW // 2 + 1), dtype=np.complex128)
    assert lib().orc_kernel_spectrum3d(_ptr(w), *w.shape, *centre, FD, FH, FW, _ptr(out)) == 0
    return out


def fft_diffuse3d(u, spec_half, box, pad_mode=PAD_ZERO, epi=0, base=0.0):
    """fftDiffuse3D in FP64: pad -> R2C -> x K^/(FD FH FW) -> C2R -> unpad."""
    u = np.ascontiguousarray(u, dtype=np.float32)
    s = np.ascontiguousarray(spec_half, dtype=np.complex128)
    D, H, W = u.shape
    out = np.empty((D, H, W), dtype=np.float64)
    rc = lib().orc_fft_diffuse3d(_ptr(u), D, H, W, _ptr(s), *box, pad_mode, epi, base, _ptr(out))
    assert rc == 0, rc
    return out


def direct_conv3d(u, window, centre, box, pad_mode=PAD_ZERO, epi=0, base=0.0):
    u = np.ascontiguousarray(u, dtype=np.float32)
    w = np.ascontiguousarray(window, dtype=np.float64)
    out = np.empty(u.shape, dtype=np.float64)
    rc = lib().orc_direct_conv3d(_ptr(u), *u.shape, _ptr(w), *w.shape, *centre, *box,
                                 pad_mode, epi, base, _ptr(out))
    assert rc == 0, rc
    return out


def gauss_radius(sigma2):
    return lib().orc_gauss_radius(float(sigma2))


def gauss_window1d(sigma2, R=None):
    if R is None:
        R = gauss_radius(sigma2)
    g = np.empty(2 * R + 1, dtype=np.float64)
    lib().orc_gauss_window1d(float(sigma2), R, _ptr(g))
    return g


def gauss_window3d(sigma2, decay, R=None):
    if R is None:
        R = gauss_radius(sigma2)
    K = 2 * R + 1
    w = np.empty((K, K, K), dtype=np.float64)
    lib().orc_gauss_window3d(float(sigma2), float(decay), _ptr(w), R)
    return w


def direct_separable(u, sigma2, decay):
    """Separable direct Gaussian convolution in FP64.  An FP32 input is the
    reference's field type; an FP64 input (a previous FP64 tick) is used as is."""
    if np.asarray(u).dtype == np.float64:
        u = np.ascontiguousarray(u)
        fn = lib().orc_direct_separable3d_f64
    else:
        u = np.ascontiguousarray(u, dtype=np.float32)
        fn = lib().orc_direct_separable3d
    out = np.empty(u.shape, dtype=np.float64)
    assert fn(_ptr(u), *u.shape, float(sigma2), float(decay), _ptr(out)) == 0
    return out


def compute_kernel3d(kernel_radius, lam, gamma, dt, niter, dom, box, window, quirk=1):
    w = np.empty(window, dtype=np.float64)
    rc = lib().orc_compute_kernel3d(int(kernel_radius), float(lam), float(gamma), float(dt),
                                    int(niter), *dom, *box, *window, int(quirk), _ptr(w))
    assert rc == 0
    return w


def stencil_heat_window(lam, gamma, dt, M, F, window):
    w = np.empty(window, dtype=np.float64)
    assert lib().orc_stencil_heat_window(float(lam), float(gamma), float(dt), int(M), int(F),
                                         *window, _ptr(w)) == 0
    return w


def neumaier_sum(v):
    v = np.ascontiguousarray(v)
    if v.dtype == np.float32:
        return lib().orc_neumaier_sum_f32(_ptr(v), v.size)
    v = v.astype(np.float64, copy=False)
    return lib().orc_neumaier_sum_f64(_ptr(v), v.size)


def synth_field(shape, key, mu=-2.0, sigma=1.0, lo=0.0, hi=0.0, nsrc=50, amp=(1.0, 100.0)):
    """Counter-based synthetic field: lognormal (or log-uniform when lo>0)
    background plus nsrc point sources with U[amp] amplitude."""
    D, H, W = shape
    f = np.empty(shape, dtype=np.float32)
    lib().orc_synth_field(_ptr(f), D, H, W, ctypes.c_uint64(key), mu, sigma, lo, hi, nsrc,
                          amp[0], amp[1])
    return f


def synth_params(key, field, drange=(0.5, 5.0), lrange=(0.001, 0.05)):
    Dv, lv = _d(), _d()
    lib().orc_synth_params(ctypes.c_uint64(key), field, drange[0], drange[1], lrange[0],
                           lrange[1], ctypes.byref(Dv), ctypes.byref(lv))
    return Dv.value, lv.value


def gauss_spectrum_real(sigma2, decay, box, scale):
    Pz, Py, Px = box
    out = np.empty((Pz, Py, Px // 2 + 1), dtype=np.float32)
    assert lib().orc_gauss_spectrum_real(float(sigma2), float(decay), Pz, Py, Px, float(scale),
                                         _ptr(out)) == 0
    return out


def cpu_diffuse_f32(u, kreal, nthreads=0):
    """FP32 CPU port of the pruned pipeline (the timed CPU baseline)."""
    u = np.ascontiguousarray(u, dtype=np.float32)
    k = np.ascontiguousarray(kreal, dtype=np.float32)
    out = np.empty_like(u)
    assert lib().cpu_fft_diffuse3d_f32(_ptr(u), _ptr(out), *u.shape, _ptr(k), int(nthreads)) == 0
    return out


def num_threads():
    return lib().orc_num_threads()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def chem_update(t, d, decay, frag_col, frag_ela):
    """WHChemical::update (gpuUpdate) restated: returns (p, d_after, (neu, mac, fib), totals)
    for 8 FP32 arrays t, d of n voxels."""
    n = t[0].size
    t = [np.ascontiguousarray(x, dtype=np.float32).ravel() for x in t]
    d = [np.ascontiguousarray(x, dtype=np.float32).ravel().copy() for x in d]
    p = [np.empty(n, dtype=np.float32) for _ in range(8)]
    fc = np.ascontiguousarray(frag_col, dtype=np.float32).ravel()
    fe = np.ascontiguousarray(frag_ela, dtype=np.float32).ravel()
    neu, mac, fib = (np.empty(n, dtype=np.float32) for _ in range(3))
    P8 = ctypes.c_void_p * 8
    tot = (_d * 8)()
    dec = (ctypes.c_float * 8)(*[float(x) for x in decay])
    assert lib().orc_chem_update(P8(*[a.ctypes.data for a in t]), P8(*[a.ctypes.data for a in d]),
                                 P8(*[a.ctypes.data for a in p]), dec, _ptr(fc), _ptr(fe), _ptr(neu),
                                 _ptr(mac), _ptr(fib), n, tot) == 0
    return p, d, (neu, mac, fib), list(tot)


def sample_chem(u, chem, stride=3):
    u = np.ascontiguousarray(u, dtype=np.float32)
    D, H, W = u.shape
    out = np.empty((D // stride, H // stride, W // stride), dtype=np.float32)
    assert lib().orc_sample_chem(_ptr(u), D, H, W, int(chem), int(stride), _ptr(out)) == 0
    return out
