"""The C++ drop-in surface end to end: diffusion_helper init (kernel and chem
dimensions, buffers, device kernel build) and Diffuser::diffuseChemGPU ticks,
driven as the ABM drives them (woundHealingWorld.cpp:1009-1049, 2403-2417),
through tests/cxx/libdiffuser_check.so.

Each chem's result after every tick is checked against the FP64 oracle's
fftDiffuse3D operator (constant-VF pad with that chem's baseline, epithelium
at the right wall) applied with the spectrum the layer built: rel-L2 <= 1e-5
per tick (FP32 path).
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cxx", "libdiffuser_check.so")
REL_F32 = 1e-5


@pytest.fixture(scope="module")
def check_lib(vfd):
    if not os.path.exists(CHECK):
        pytest.fail("tests/cxx/libdiffuser_check.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(CHECK)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.diffuser_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, fp, fp, fp,
                                   ctypes.c_int, ctypes.c_int, fp, fp, fp, fp, ctypes.POINTER(ctypes.c_int)]
    return lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


@pytest.mark.parametrize("shape,dt,ticks", [((16, 20, 24), 10.0, 2), ((12, 18, 30), 12.0, 1)])
def test_diffuser_ticks_match_oracle(check_lib, orc, shape, dt, ticks):
    nz, ny, nx = shape
    print(f"case diffuser shape={shape} dt={dt} ticks={ticks}", flush=True)
    rng = np.random.default_rng(5)
    lam = rng.uniform(0.01, 0.03, 8).astype(np.float32)       # stencil lambda = D dt / dx^2
    gam = rng.uniform(1e-5, 1e-4, 8).astype(np.float32)       # decay rate per second
    base = rng.uniform(0.0, 0.5, 8).astype(np.float32)
    wall = nx - 3
    u0 = np.stack([orc.synth_field(shape, 300 + ic, nsrc=4) for ic in range(8)]).astype(np.float32)
    fields = np.ascontiguousarray(u0.copy())
    info = (ctypes.c_int * 13)()
    # the chem FFT box is snapTransformSize(n + window - 1) per axis; allocate
    # for the largest box these cases can reach
    cap = 8 * 128 * 128 * 65
    spectra = np.zeros(2 * cap, dtype=np.float32)
    rc = check_lib.diffuser_check(nx, ny, nz, dt, _fp(lam), _fp(gam), _fp(base), wall, ticks, _fp(fields),
                                  _fp(spectra), None, None, info)
    assert rc == 0
    kfft, kwin, krad, niter, FD, FH, FW, KD, KH, KW, KZ, KY, KX = list(info)
    print(f"kernel box {kfft} window {kwin} radius {krad} niter {niter}; chem box {FD}x{FH}x{FW} "
          f"window {KD}x{KH}x{KW} centre ({KZ},{KY},{KX})", flush=True)
    assert krad == int(300 / dt) and niter == 6
    assert FD == orc.snap_transform_size(nz + KD - 1) and FW == orc.snap_transform_size(nx + KW - 1)
    sn = FD * FH * (FW // 2 + 1)
    assert 8 * sn <= cap
    spec = spectra[: 2 * 8 * sn].view(np.complex64).reshape(8, FD, FH, FW // 2 + 1).astype(np.complex128)
    out = fields.reshape(8, nz, ny, nx)
    for ic in range(8):
        ref = u0[ic].astype(np.float64)
        for _ in range(ticks):
            ref = orc.fft_diffuse3d(ref, spec[ic], (FD, FH, FW), pad_mode=orc.PAD_CONSTANT_VF, epi=wall,
                                    base=float(base[ic]))
        err = np.abs(out[ic] - ref)
        rel = float(np.linalg.norm(err) / np.linalg.norm(ref))
        print(f"chem {ic}: max-abs {err.max():.3e} rel-L2 {rel:.3e}", flush=True)
        assert rel <= REL_F32 * ticks
    # DC gain of each spectrum = the kernel window's mass: decay (1 - gamma dt)
    # per substep times the window's coverage of the heat kernel, in (0.5, 1]
    dc = spec[:, 0, 0, 0].real
    assert np.all(dc > 0.5) and np.all(dc <= 1.0 + 1e-6)
