"""C-ABI checks that need no GPU: libvfd.so loads, exports every function the
public headers declare, and rejects bad plans before touching the device.
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "vf-version_7_3_amd", "libvfd.so")

_DECL = re.compile(r"^\s*(?:extern\s+)?(?:const\s+)?[A-Za-z_][\w\s\*]*?[\s\*]([A-Za-z_]\w*)\s*\(", re.M)
_SKIP = {"iDivUp", "iAlignUp", "if", "for", "while", "return", "sizeof"}


C_HEADERS = ("vfd.h", "vf_diffusion.h")          # the extern "C" interface of libvfd.so
CXX_HEADERS = ("DiffusionHelper.h", "Diffuser.h")  # C++ surface compiled into the ABM's build
CXX_CHECK = os.path.join(ROOT, "tests", "cxx", "libdiffuser_check.so")


def declared_functions(headers=C_HEADERS):
    names = {}
    for fn in headers:
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in _DECL.finditer(text):
            name = m.group(1)
            if name in _SKIP:
                continue
            names[name] = fn
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libvfd.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def test_headers_declare_expected_entry_points():
    names = declared_functions()
    for must in ("fftDiffuse3D", "computeKernel3D", "computeKernelSpectrum3D", "computeKernel3DBatch",
                 "snapTransformSize", "H2D", "D2H", "padDataConstantVF3D", "vfd_plan_create", "vfd_diffuse",
                 "vfd_gauss_spectrum", "vfd_plan_autotune"):
        assert must in names, must


def test_library_exports_every_declared_function(lib):
    exported = exported_symbols()
    missing = [(n, h) for n, h in declared_functions().items() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    assert "gpu_id" in exported      # extern int gpu_id[N_CHEM] (convolutionFFT_common.h:384)


def test_cxx_surface_builds_and_exports_reference_names():
    """diffusion_helper::* and Diffuser (DiffusionHelper.h:30-43, Diffuser.h:24-70)
    compile against the ABM-side header set and export every declared name."""
    if not os.path.exists(CXX_CHECK):
        pytest.fail("tests/cxx/libdiffuser_check.so is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-DC", "--defined-only", CXX_CHECK], capture_output=True, text=True,
                         check=True).stdout
    for name in ("findGPUs", "initKernelDimensions", "allocKernelBuffers", "deallocConvCtxBuffers",
                 "initChemDimensions", "allocChemBuffers", "prepAndComputeKernelSpectrum", "prepKernel",
                 "printContext"):
        assert f"diffusion_helper::{name}(" in out, name
    for sig in ("Diffuser::Diffuser()", "Diffuser::Diffuser(int, CCTX*, CCTX*, WHChemical*)",
                "Diffuser::diffuseChemGPU(int)", "Diffuser::~Diffuser()"):
        assert sig in out, sig
    names = declared_functions(CXX_HEADERS)
    assert {"initKernelDimensions", "allocChemBuffers", "diffuseChemGPU"} <= set(names)


def test_snap_transform_size_matches_reference_rule(lib, orc):
    lib.snapTransformSize.argtypes = [ctypes.c_int]
    for n in list(range(1, 70)) + [100, 597, 1000, 1024, 1025, 1390, 2049, 5000]:
        assert lib.snapTransformSize(n) == orc.snap_transform_size(n), n


def test_length_support_table(lib):
    lib.vfd_length_supported.argtypes = [ctypes.c_int]
    for q in (2, 3, 4, 5, 6, 8, 9, 12, 16, 20, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 616, 768, 1024,
              1152, 1280, 1536, 2048, 2304):
        assert lib.vfd_length_supported(q) == 1, q
    for q in (1, 7, 11, 13, 17, 100, 1000, 4096, 3072):
        assert lib.vfd_length_supported(q) == 0, q


class PlanDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nz", "ny", "nx", "pz", "py", "px", "precision", "pad_mode",
                                           "kspec_format")]


@pytest.mark.parametrize("desc,code,needle", [
    ((0, 4, 4, 0, 0, 0, 4, 100, 0), -1, "positive"),          # empty box
    ((4, 4, 4, 2, 8, 8, 4, 100, 0), -1, "smaller"),           # FFT box smaller than data
    ((4, 4, 5, 8, 8, 9, 4, 100, 0), -2, "even"),               # odd px
    ((4, 4, 4, 8, 8, 8, 2, 100, 0), -1, "precision"),
    ((4, 4, 4, 8, 8, 8, 4, 7, 0), -1, "pad mode"),
    ((4, 4, 4, 8, 8, 8, 4, 100, 9), -1, "spectrum format"),
    ((4, 4, 4, 8, 8, 34, 4, 100, 0), -2, "not compiled"),     # px/2 = 17
    ((4, 4, 4, 8, 26, 8, 4, 100, 0), -2, "not compiled"),     # py = 26 = 2*13
    ((4, 4, 4, 8, 8, 8, 4, 3, 2), -1, "separable"),           # separable spectrum needs the zero pad
])
def test_plan_validation_before_device(lib, desc, code, needle):
    """Invalid dimensions are rejected with a code and a message, before any
    device call (this container has no GPU)."""
    lib.vfd_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(ctypes.c_void_p)]
    lib.vfd_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    rc = lib.vfd_plan_create(ctypes.byref(PlanDesc(*desc)), ctypes.byref(h))
    assert rc == code
    assert needle in lib.vfd_last_error().decode()
    assert not h.value


def test_null_arguments_are_errors(lib):
    lib.vfd_plan_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert lib.vfd_plan_create(None, None) == -1
    lib.vfd_diffuse.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    assert lib.vfd_diffuse(None, None, 0, None) == -1


def test_legacy_bad_arguments_do_not_exit(lib):
    """fftDiffuse3D with null buffers returns false instead of exiting."""
    lib.fftDiffuse3D.restype = ctypes.c_bool
    from vfd import legacy  # noqa: F401  (argtypes)
    cc = legacy.CCtx()
    assert not legacy.lib.fftDiffuse3D(None, None, 0, 0, cc, 1, 0.0, 0, 0)


def test_resident_tick_argument_checks_before_device(lib):
    """vfdTickChems / vfdTickChemsBegin / vfdTickChemsEnd validate on the host
    (no device call): null arguments, levels shared between chems, and an End
    without a pending Begin are errors with a message, not exits."""
    from vfd import legacy
    import vfd
    cc = legacy.CCtx()
    cc.DD, cc.DH, cc.DW = 4, 4, 4
    cc.FFTD, cc.FFTH, cc.FFTW = 8, 8, 8
    f8 = (ctypes.c_float * 8)(*([0.5] * 8))
    buf = ctypes.create_string_buffer(4 * 64)
    p = ctypes.cast(buf, ctypes.c_void_p).value
    dv = (ctypes.c_void_p * 8)(*([p] * 8))
    gv = (ctypes.c_void_p * 3)(*([p] * 3))
    tot = (ctypes.c_double * 8)()
    assert not legacy.lib.vfdTickChems(cc, 4, None, dv, f8, p, p, gv, tot)
    assert "null argument" in vfd.last_error()
    for ic in range(8):
        cc.d_data[ic] = p                      # one buffer for every chem
        cc.h_ibuffs[ic] = p
        cc.d_kernelspectrum_h[ic] = p
    assert not legacy.lib.vfdTickChems(cc, 4, f8, dv, f8, p, p, gv, tot)
    assert "share d_data" in vfd.last_error()
    assert not legacy.lib.vfdTickChemsBegin(cc, 4, f8)
    assert "share d_data" in vfd.last_error()
    bufs = [ctypes.create_string_buffer(4 * 64) for _ in range(8)]
    for ic in range(8):
        cc.d_data[ic] = ctypes.cast(bufs[ic], ctypes.c_void_p).value
    assert not legacy.lib.vfdTickChemsEnd(cc, dv, f8, p, p, gv, tot)
    assert "no vfdTickChemsBegin is pending" in vfd.last_error()


def test_release_workspaces_with_nothing_held(lib):
    """vfdReleaseWorkspaces with no workspace made is a no-op returning 0 (no device call)."""
    lib.vfdReleaseWorkspaces.restype = ctypes.c_int
    assert lib.vfdReleaseWorkspaces() == 0


def test_convolution_clamp_cpu_definition(lib):
    """convolutionClampToBorderCPU (a host routine of the library) against its
    definition evaluated independently with numpy: clamped (edge-replicated)
    data shifted per kernel tap, each product rounded to FP32, summed in FP64."""
    import numpy as np
    rng = np.random.default_rng(5)
    H, W, kH, kW, KY, KX = 7, 9, 3, 5, 1, 3          # off-centre window, clamps on every side
    data = rng.standard_normal((H, W)).astype(np.float32)
    k = rng.standard_normal((kH, kW)).astype(np.float32)
    res = np.zeros_like(data)
    P = ctypes.c_void_p
    lib.convolutionClampToBorderCPU(P(res.ctypes.data), P(data.ctypes.data), P(k.ctypes.data), H, W, kH, kW, KY, KX)
    padded = np.pad(data, ((kH, kH), (kW, kW)), mode="edge")
    acc = np.zeros((H, W), np.float64)
    for r in range(kH):
        for c in range(kW):
            # u(clamp(y + KY - r), clamp(x + KX - c)) for every (y, x)
            sh = padded[kH + KY - r: kH + KY - r + H, kW + KX - c: kW + KX - c + W]
            acc += (sh * k[r, c]).astype(np.float32).astype(np.float64)
    assert np.allclose(res, acc.astype(np.float32), rtol=1e-6, atol=1e-6)
    # identity kernel: the data itself, exactly
    kid = np.zeros((kH, kW), np.float32)
    kid[KY, KX] = 1.0
    lib.convolutionClampToBorderCPU(P(res.ctypes.data), P(data.ctypes.data), P(kid.ctypes.data), H, W, kH, kW, KY, KX)
    assert np.array_equal(res, data)
