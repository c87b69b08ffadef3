"""The ABM's production configuration (HUMAN_VF) through the C++ drop-in
surface, and the legacy host<->device copies.

* H2D / D2H (Diffusion.cu:72-122, the M40 build: the spectrum stays on the
  device): exact copies of np floats between c_ctx's host and device buffers.
* HUMAN_VF init and one tick, driven as the ABM drives them
  (woundHealingWorld.cpp:1009-1049, 2403-2417; DiffusionHelper.cpp:29-163,
  381-461): kernel box 616^3, window 147^3, kernelRadius 150, niter 6, dt 2 s,
  chem box 1152x1536x256 for the 110x1390x1006 grid, the production
  diffusivities and half-lives.
    - every chem's kernel window (computeKernel3DBatch, built on the device in
      FP64) against the stencil heat kernel in closed form,
          K(x) = P^-3 sum_k S(k)^M cos(theta . x),
          S(k) = 1 - 6 lambda - gamma dt + 2 lambda (cos th_z + cos th_y + cos th_x),
      M = niter (kernelRadius - 1) = 894 stencil steps, P = 616: the window's
      sum, a set of voxels, and its mirror and axis-exchange symmetry;
    - the tick's output on sub-boxes (a corner of the baseChem region, the
      epithelium boundary, around a source, the far faces) against an FP64
      convolution of the constant-VF padded field with the verified window.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK_HUMAN = os.path.join(ROOT, "tests", "cxx", "libdiffuser_check_human.so")

# production chem order (enums.h:124-133): TNF, TGF, FGF, MMP8, IL1beta, IL6, IL8, IL10;
# diffusivities (um^2/s, config_VocalFold.txt:20-29) and half-lives (13 min; IL10 30 min,
# woundHealingWorld.cpp:90, 965); patch width 1.65 mm / 109 (World.cpp:68); dt = 2 s
D_UM2 = [15.0, 13.0, 13.0, 13.0, 15.0, 13.5, 15.0, 15.0]
HALF_LIFE_S = [780.0] * 7 + [1800.0]
DX_UM = 1650.0 / 109.0
DT = 2.0
NX, NY, NZ = 110, 1390, 1006
RIGHT_WALL = 106                     # bmcx - 1
P_K, WIN, KR, NITER = 616, 147, 150, 6


def test_h2d_d2h_round_trip(vfd, orc):
    """H2D copies np floats of h_ibuffs[ic] into d_data[ic]; D2H copies
    d_data[ic] into h_obuffs[ic]; nothing past np is touched; a bad chem
    index is refused with a message."""
    import torch
    from vfd import legacy as leg
    n = 1 << 20
    rng = np.random.default_rng(4)
    host_in = [rng.standard_normal(n).astype(np.float32) for _ in range(2)]
    host_out = [np.full(n, -7.0, np.float32) for _ in range(2)]
    dev = [torch.full((n,), 3.0, device="cuda") for _ in range(2)]
    cc = leg.CCtx()
    for ic in range(2):
        cc.h_ibuffs[ic] = host_in[ic].ctypes.data
        cc.h_obuffs[ic] = host_out[ic].ctypes.data
        cc.d_data[ic] = dev[ic].data_ptr()
        cc.dev_id[ic] = 0
    npart = n - 5
    for ic in range(2):
        leg.lib.H2D(ic, ctypes.byref(cc), npart)
    torch.cuda.synchronize()
    for ic in range(2):
        got = dev[ic].cpu().numpy()
        assert np.array_equal(got[:npart], host_in[ic][:npart]) and np.all(got[npart:] == 3.0), ic
        dev[ic].mul_(2.0)
    torch.cuda.synchronize()
    for ic in range(2):
        leg.lib.D2H(ic, ctypes.byref(cc), npart)
        assert np.array_equal(host_out[ic][:npart], 2.0 * host_in[ic][:npart]), ic
        assert np.all(host_out[ic][npart:] == -7.0), ic
    leg.lib.H2D(8, ctypes.byref(cc), npart)
    assert "H2D" in vfd.last_error()


def _params():
    lam = np.array([d * DT / DX_UM ** 2 for d in D_UM2], np.float32)      # stencil lambda = D dt / dx^2
    gam = np.array([np.log(2.0) / h for h in HALF_LIFE_S], np.float32)    # decay rate per second
    return lam, gam


class HeatKernel:
    """The free-space heat kernel of the 7-point stencil after M steps,
    evaluated from its spectrum on the P-box; only |k| <= 120 per axis
    contributes (S^M < e^-170 beyond)."""

    def __init__(self, lam, gam, dt, M, P=P_K, kmax=120):
        lam, gdt = float(np.float32(lam)), float(np.float32(gam)) * float(np.float32(dt))
        self.P = P
        self.th = 2.0 * np.pi * np.arange(-kmax, kmax + 1) / P
        c = np.cos(self.th)
        s = (1.0 - 6.0 * lam - gdt) + 2.0 * lam * (c[:, None, None] + c[None, :, None] + c[None, None, :])
        assert s.min() > 0.0
        self.SM = np.exp(M * np.log(s))
        self.total = (1.0 - gdt) ** M       # S(0)^M: the kernel's whole mass

    def at(self, dz, dy, dx):
        return np.einsum("abc,a,b,c->", self.SM, np.cos(self.th * dz), np.cos(self.th * dy),
                         np.cos(self.th * dx)) / self.P ** 3

    def window_sum(self, half):
        d = np.cos(self.th[:, None] * np.arange(-half, half + 1)[None, :]).sum(axis=1)
        return np.einsum("abc,a,b,c->", self.SM, d, d, d) / self.P ** 3


@pytest.fixture(scope="module")
def human_run(vfd, orc):
    """The HUMAN_VF init and one tick through libdiffuser_check_human.so."""
    if not os.path.exists(CHECK_HUMAN):
        pytest.fail("tests/cxx/libdiffuser_check_human.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(CHECK_HUMAN)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.diffuser_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, fp, fp, fp,
                                   ctypes.c_int, ctypes.c_int, fp, fp, fp, fp, ctypes.POINTER(ctypes.c_int)]
    lam, gam = _params()
    base = np.array([2e-4, 1e-4, 5e-5, 3e-4, 1e-4, 2e-4, 4e-4, 1e-4], np.float32)
    print(f"case human_vf init+tick grid {NX}x{NY}x{NZ}: synthesizing 8 fields", flush=True)
    n = NX * NY * NZ
    fields = np.empty(8 * n, np.float32)
    for ic in range(8):
        fields[ic * n:(ic + 1) * n] = orc.synth_field((NZ, NY, NX), 700 + ic, nsrc=50).reshape(-1)
    u0 = fields[7 * n:8 * n].reshape(NZ, NY, NX).copy()
    windows = np.zeros(8 * WIN ** 3, np.float32)
    rows = np.zeros(8 * 2 * (256 // 2 + 1), np.float32)
    info = (ctypes.c_int * 13)()
    f = lambda a: a.ctypes.data_as(fp)
    print("case human_vf: init (kernel build) and one diffuseChemGPU tick", flush=True)
    rc = lib.diffuser_check(NX, NY, NZ, DT, f(lam), f(gam), f(base), RIGHT_WALL, 1, f(fields), None, f(windows),
                            f(rows), info)
    assert rc == 0
    out = fields[7 * n:8 * n].reshape(NZ, NY, NX)
    rows = rows.view(np.complex64).reshape(8, 256 // 2 + 1).astype(np.complex128)
    return list(info), windows.reshape(8, WIN, WIN, WIN), rows, lam, gam, base, u0, out


def test_human_vf_dimensions(human_run):
    info = human_run[0]
    kfft, kwin, krad, niter, FD, FH, FW, KD, KH, KW, KZ, KY, KX = info
    print(f"kernel box {kfft} window {kwin} radius {krad} niter {niter}; chem box {FD}x{FH}x{FW} "
          f"window {KD}x{KH}x{KW} centre ({KZ},{KY},{KX})", flush=True)
    assert (kfft, kwin, krad, niter) == (P_K, WIN, KR, NITER)
    assert (FD, FH, FW) == (1152, 1536, 256)
    assert (KD, KH, KW) == (WIN, WIN, WIN) and (KZ, KY, KX) == (73, 73, 73)


def test_human_vf_kernel_window(human_run):
    """The kernel window against the closed-form heat kernel: sum, voxels,
    symmetries.  The kernel context has one window buffer per GPU, aliased by
    every chem (DiffusionHelper.cpp:166-240): each chem's spectrum is built
    from it right after that chem's window, so after the batch every slot
    holds the last chem's (IL10) window, which is checked here; every chem's
    kernel is checked through its spectrum below."""
    _, windows, _, lam, gam, _, _, _ = human_run
    ic = 7
    for j in range(7):
        assert np.array_equal(windows[j], windows[ic])
    w = windows[ic].astype(np.float64)
    M = NITER * (KR - 1)
    hk = HeatKernel(lam[ic], gam[ic], DT, M)
    h = WIN // 2
    wmax = float(w.max())
    # symmetry: mirror on every axis, exchange of axes (an isotropic stencil)
    for a in (w[::-1], w[:, ::-1], w[:, :, ::-1], w.transpose(1, 0, 2), w.transpose(2, 1, 0)):
        assert np.abs(a - w).max() <= 1e-6 * wmax
    # the centre is the maximum
    assert w[h, h, h] == w.max()
    ws, ref_sum = float(w.sum()), hk.window_sum(h)
    cover = ref_sum / hk.total
    print(f"chem {ic}: lambda {lam[ic]:.5f} gamma {gam[ic]:.4e}: window sum {ws:.9f}, closed form {ref_sum:.9f} "
          f"= (1 - gamma dt)^M {hk.total:.6f} x coverage {cover:.9f}", flush=True)
    assert abs(ws - ref_sum) <= 1e-6 * ref_sum
    assert 0.99999 < cover <= 1.0
    for d in [(0, 0, 0), (1, 0, 0), (0, 2, 1), (5, 3, 2), (12, 0, 7), (20, 20, 0), (31, 5, 17), (45, 0, 0),
              (60, 10, 3), (73, 73, 73)]:
        ref = hk.at(*d)
        got = w[h + d[0], h + d[1], h + d[2]]
        tol = 2e-6 * abs(ref) + 1e-7 * wmax
        assert abs(got - ref) <= tol, (ic, d, got, ref)


@pytest.mark.parametrize("ic", range(8))
def test_human_vf_spectrum_row(human_run, ic):
    """Every chem's spectrum (computeKernelSpectrum3D of its window on the
    1152x1536x256 chem box) at kz = ky = 0 against the closed form
      K^(0, 0, kx) = sum over the window of K(x) cos(2 pi kx x_x / 256)
                   = P^-3 sum_q S(q)^M D(th_z) D(th_y) E_kx(th_x),
    D(th) = sum_{|j| <= 73} cos(th j), E_kx(th) = sum_{|j| <= 73} cos(th j) cos(2 pi kx j / 256):
    the decay and diffusivity of each chem, the window's truncation, and the
    centring of the window at the origin of the box."""
    _, _, rows, lam, gam, _, _, _ = human_run
    hk = HeatKernel(lam[ic], gam[ic], DT, NITER * (KR - 1))
    h = WIN // 2
    j = np.arange(-h, h + 1)
    d = np.cos(hk.th[:, None] * j[None, :]).sum(axis=1)
    row = rows[ic]
    dc = float(row[0].real)
    for kx in (0, 1, 2, 5, 9, 17, 40, 128):
        e = (np.cos(hk.th[:, None] * j[None, :]) * np.cos(2.0 * np.pi * kx * j / 256.0)[None, :]).sum(axis=1)
        ref = np.einsum("abc,a,b,c->", hk.SM, d, d, e) / hk.P ** 3
        got = row[kx]
        assert abs(got.real - ref) <= 1e-6 * dc and abs(got.imag) <= 1e-6 * dc, (ic, kx, got, ref)
        if kx in (0, 9):
            print(f"chem {ic} K^(0,0,{kx}) = {got.real:.9f} (closed form {ref:.9f}), imag {got.imag:.2e}",
                  flush=True)


def test_human_vf_tick_sub_boxes(human_run):
    """One production tick (constant-VF pad with the chem's baseline, the
    epithelium at x = 106) against the FP64 convolution of the padded field
    with its window, on sub-boxes; chem 7, whose window is verified above."""
    from scipy.signal import fftconvolve
    _, windows, _, _, _, base, u0, out = human_run
    R = WIN // 2
    shape = (NZ, NY, NX)
    for ic in (7,):
        w = windows[ic].astype(np.float64)
        u = u0
        b = float(base[ic])

        def expected(lo, n):
            rng = [np.arange(l - R, l + m + R) for l, m in zip(lo, n)]
            inside = [(r > 0) & (r < lim) for r, lim in zip(rng, (NZ - 1, NY - 1, RIGHT_WALL))]
            cl = [np.clip(r, 0, lim - 1) for r, lim in zip(rng, shape)]
            blk = u[np.ix_(cl[0], cl[1], cl[2])].astype(np.float64)
            mask = inside[0][:, None, None] & inside[1][None, :, None] & inside[2][None, None, :]
            return fftconvolve(np.where(mask, blk, b), w, mode="valid")

        for lo, n in [((0, 0, 96), (6, 6, 14)), ((498, 690, 40), (6, 6, 6)), ((1000, 1384, 104), (6, 6, 6)),
                      ((300, 1000, 0), (6, 6, 6))]:
            ref = expected(lo, n)
            got = out[lo[0]:lo[0] + n[0], lo[1]:lo[1] + n[1], lo[2]:lo[2] + n[2]].astype(np.float64)
            err = np.abs(got - ref)
            rel = float(np.linalg.norm(err) / np.linalg.norm(ref))
            print(f"chem {ic} sub-box at {lo}: max-abs {err.max():.3e} rel-L2 {rel:.3e}", flush=True)
            assert rel <= 1e-5
