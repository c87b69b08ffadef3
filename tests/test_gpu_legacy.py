"""The reference-compatible interface (include/vf_diffusion.h) on the GPU,
called through the C-ABI with the reference's argument meaning, checked
against the FP64 oracle (restatement of the same reference functions).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def leg(vfd):
    from vfd import legacy
    return legacy


def _t(a, dev="cuda"):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _rel(got, ref):
    got = np.asarray(got, np.float64)
    rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"  max-abs {np.abs(got - ref).max():.3e} rel-L2 {rel:.3e}", flush=True)
    return rel


@pytest.mark.parametrize("shape,box,K,epi,base", [
    ((10, 12, 9), (24, 24, 24), 7, 7, 0.5),        # D,H,W data; FD,FH,FW box
    ((21, 30, 11), (48, 64, 32), 9, 9, 0.125),
    ((30, 40, 12), (72, 80, 36), 9, 11, 0.0),     # radix-3/5 box lengths
])
def test_fftDiffuse3D_constant_vf(vfd, orc, leg, shape, box, K, epi, base):
    import torch
    print(f"case fftDiffuse3D shape={shape} box={box}", flush=True)
    rng = np.random.default_rng(11)
    u = orc.synth_field(shape, 5, nsrc=4)
    w = rng.random((K, K, K))
    w = (w + w[::-1, ::-1, ::-1]) / (2 * w.sum())
    c = K // 2
    spec = orc.kernel_spectrum(w, (c, c, c), box)
    ref = orc.fft_diffuse3d(u, spec, box, orc.PAD_CONSTANT_VF, epi, base)
    cc = leg.CCtx()
    cc.DD, cc.DH, cc.DW = shape
    cc.FFTD, cc.FFTH, cc.FFTW = box
    cc.KD = cc.KH = cc.KW = K
    cc.KZ = cc.KY = cc.KX = c
    d = _t(u)
    ds = _t(spec.astype(np.complex64))
    pf, pi = ctypes.c_int(), ctypes.c_int()
    assert leg.lib.vfdPlan3d(ctypes.byref(pf), *box, 0x2a) == 0
    assert leg.lib.vfdPlan3d(ctypes.byref(pi), *box, 0x2c) == 0
    ok = leg.lib.fftDiffuse3D(ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(ds.data_ptr()), pf, pi, cc,
                              epi, base, 0, 0)
    assert ok
    torch.cuda.synchronize()
    assert _rel(d.cpu().numpy(), ref) <= 1e-5
    # a second call reuses the cached workspace and stays identical
    d2 = _t(u)
    assert leg.lib.fftDiffuse3D(ctypes.c_void_p(d2.data_ptr()), ctypes.c_void_p(ds.data_ptr()), pf, pi, cc,
                                epi, base, 1, 1)
    torch.cuda.synchronize()
    assert torch.equal(d, d2)
    assert leg.lib.vfdDestroy(pf) == 0 and leg.lib.vfdDestroy(pi) == 0


def test_fftDiffuse3D_rejects_bad_plan(vfd, leg):
    import torch
    cc = leg.CCtx()
    cc.DD = cc.DH = cc.DW = 4
    cc.FFTD = cc.FFTH = cc.FFTW = 8
    d = torch.zeros((4, 4, 4), device="cuda")
    s = torch.zeros((8, 8, 5), dtype=torch.complex64, device="cuda")
    p = ctypes.c_int()
    assert leg.lib.vfdPlan3d(ctypes.byref(p), 8, 8, 16, 0x2a) == 0   # wrong box
    assert not leg.lib.fftDiffuse3D(ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(s.data_ptr()), p, 0, cc,
                                    3, 0.0, 0, 0)
    assert "box" in vfd.last_error()
    leg.lib.vfdDestroy(p)


@pytest.mark.parametrize("quirk", [1, 0])
def test_computeKernel3D_and_spectrum(vfd, orc, leg, quirk):
    """computeKernel3D (stencil heat-kernel window) and computeKernelSpectrum3D
    vs the FP64 restatements, on a small kernel context."""
    import torch
    kr, niter, lam, gamma, dt = 5, 2, 0.11, 0.002, 2.0
    dom = (kr - 1) * 2 + 1
    kbox = orc.snap_transform_size(dom + 3 - 1)
    win = 7
    print(f"case computeKernel3D quirk={quirk} dom={dom} box={kbox}", flush=True)
    ref_w = orc.compute_kernel3d(kr, lam, gamma, dt, niter, (dom,) * 3, (kbox,) * 3, (win,) * 3, quirk=quirk)
    kc = leg.CCtx()
    kc.dt, kc.kernelRadius, kc.niter = dt, kr, niter
    kc.KD = kc.KH = kc.KW = 3
    kc.KZ = kc.KY = kc.KX = 1
    kc.DD = kc.DH = kc.DW = dom
    kc.FFTD = kc.FFTH = kc.FFTW = kbox
    kc.windowD = kc.windowH = kc.windowW = win
    dwin = torch.zeros(win ** 3, dtype=torch.float32, device="cuda")
    hwin = np.zeros(win ** 3, np.float32)
    kc.d_data[0] = dwin.data_ptr()
    kc.h_data[0] = hwin.ctypes.data
    leg.lib.vfdSetKernelPowerQuirk(quirk)
    try:
        assert leg.lib.computeKernel3D(kr, lam, gamma, dt, kc, 0, 0, 0)
    finally:
        leg.lib.vfdSetKernelPowerQuirk(1)
    got = hwin.reshape(win, win, win)
    assert np.array_equal(got, dwin.cpu().numpy().reshape(win, win, win))
    assert _rel(got, ref_w) <= 1e-6
    # spectrum on a chem box
    box = (24, 20, 16)
    cc = leg.CCtx()
    cc.KD = cc.KH = cc.KW = win
    cc.KZ = cc.KY = cc.KX = win // 2
    cc.FFTD, cc.FFTH, cc.FFTW = box
    dspec = torch.zeros((box[0], box[1], box[2] // 2 + 1), dtype=torch.complex64, device="cuda")
    hspec = np.zeros(dspec.shape, np.complex64)
    cc.d_kernelspectrum_h[0] = dspec.data_ptr()
    cc.h_kernelspectrum[0] = hspec.ctypes.data
    assert leg.lib.computeKernelSpectrum3D(kc, cc, 0)
    ref_s = orc.kernel_spectrum(got.astype(np.float64), (win // 2,) * 3, box)
    assert np.abs(hspec - ref_s).max() <= 1e-6 * np.abs(ref_s).max()


def test_pad_helpers_bit_exact(vfd, orc, leg):
    """3D pad / unpad / extract helpers are pure index maps: bit-exact."""
    import torch
    rng = np.random.default_rng(2)
    D, H, W = 5, 6, 7
    box = (9, 10, 12)
    u = rng.random((D, H, W)).astype(np.float32)
    du = _t(u)
    out = torch.empty(box, device="cuda")
    leg.lib.padDataConstantVF3D(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(du.data_ptr()), *box, D, H, W,
                                3, 3, 3, 1, 1, 1, 5, 0.75)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), orc.pad_constant_vf(u, box, 5, 0.75).astype(np.float32))
    leg.lib.padDataClampToBorder3D(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(du.data_ptr()), *box, D, H, W,
                                   3, 3, 3, 1, 2, 1)
    torch.cuda.synchronize()
    ref = orc.pad_clamp(u.astype(np.float64), box, 1, 2, 1).astype(np.float32)
    assert np.array_equal(out.cpu().numpy(), ref)
    k = rng.random((3, 4, 5)).astype(np.float32)
    kd = _t(k)
    out.zero_()
    leg.lib.padKernel3D(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(kd.data_ptr()), *box, 3, 4, 5, 1, 2, 2)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), orc.pad_kernel(k.astype(np.float64), box, 1, 2, 2).astype(np.float32))
    big = _t(rng.random(box).astype(np.float32))
    crop = torch.empty((D, H, W), device="cuda")
    leg.lib.unpadResult3D(ctypes.c_void_p(crop.data_ptr()), ctypes.c_void_p(big.data_ptr()), D, H, W, *box)
    torch.cuda.synchronize()
    assert torch.equal(crop, big[:D, :H, :W])
    win = torch.empty((3, 4, 5), device="cuda")
    leg.lib.extractCenter3D(ctypes.c_void_p(win.data_ptr()), ctypes.c_void_p(big.data_ptr()), *box, 3, 4, 5)
    torch.cuda.synchronize()
    ref = orc.extract_center(big.cpu().numpy().astype(np.float64), (3, 4, 5)).astype(np.float32)
    assert np.array_equal(win.cpu().numpy(), ref)


def test_modulate_and_power(vfd, leg):
    import torch
    rng = np.random.default_rng(4)
    fD, fH, fW = 3, 4, 6
    n = (fD, fH, fW // 2 + 1)
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    da, db = _t(a), _t(b)
    leg.lib.modulateAndNormalize3D(ctypes.c_void_p(da.data_ptr()), ctypes.c_void_p(db.data_ptr()), fD, fH, fW, 1)
    torch.cuda.synchronize()
    ref = (a.astype(np.complex128) * b) / (fD * fH * fW)
    assert np.abs(da.cpu().numpy() - ref).max() < 1e-6 * np.abs(ref).max()
    # complexPower keeps the reference's atan(b/a) phase: exact for Re > 0
    z = (np.abs(rng.standard_normal(n)) + 0.5 + 0.3j * rng.standard_normal(n)).astype(np.complex64)
    dz = _t(z)
    leg.lib.complexPower(ctypes.c_void_p(dz.data_ptr()), fD, fH, fW, 1, 3)
    torch.cuda.synchronize()
    ref = z.astype(np.complex128) ** 3
    assert np.abs(dz.cpu().numpy() - ref).max() < 1e-5 * np.abs(ref).max()
    # Re < 0: |z|^n cis(n atan(b/a)) = -(z^n) for odd n
    zn = np.full(n, -2.0 + 0.0j, np.complex64)
    dzn = _t(zn)
    leg.lib.complexPower(ctypes.c_void_p(dzn.data_ptr()), fD, fH, fW, 1, 3)
    torch.cuda.synchronize()
    assert np.allclose(dzn.cpu().numpy(), 8.0)


def test_2d_pads_known_answers(vfd, leg):
    """2D pad rules (convolutionFFT.cuh:179-363) on a 2x3 field in a 5x6 box."""
    import torch
    src = np.array([[1, 2, 3], [4, 5, 6]], np.float32)
    ds = _t(src)
    out = torch.empty((5, 6), device="cuda")
    P = ctypes.c_void_p
    leg.lib.padDataMirror(P(out.data_ptr()), P(ds.data_ptr()), 5, 6, 2, 3, 3, 3, 1, 1)
    torch.cuda.synchronize()
    # rows: y<2 data, y=2 mirror of row 1, y>=3 -> fH-y-1 (1, 0); cols likewise
    expect = np.array([[1, 2, 3, 3, 2, 1], [4, 5, 6, 6, 2, 1], [4, 5, 6, 6, 2, 1],
                       [4, 5, 6, 6, 5, 4], [1, 2, 3, 3, 2, 1]], np.float32)
    got = out.cpu().numpy()
    # build the expectation from the index rule itself, element by element
    exp2 = np.empty((5, 6), np.float32)
    for y in range(5):
        dy = y if y < 2 else (2 - (y - 2) - 1 if y < 3 else 5 - y - 1)
        for x in range(6):
            dx = x if x < 3 else (3 - (x - 3) - 1 if x < 4 else 6 - x - 1)
            exp2[y, x] = src[dy, dx]
    assert np.array_equal(got, exp2)
    assert got[0, 3] == 3 and got[2, 0] == 4 and got[4, 5] == src[0, 0]
    del expect
    leg.lib.padDataClampToBorder(P(out.data_ptr()), P(ds.data_ptr()), 5, 6, 2, 3, 3, 3, 1, 1)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert list(got[0]) == [1, 2, 3, 3, 1, 1] and list(got[2]) == [4, 5, 6, 6, 4, 4] and list(got[4]) == [1, 2, 3, 3, 1, 1]
    leg.lib.padDataConstantVF(P(out.data_ptr()), P(ds.data_ptr()), 5, 6, 2, 3, 3, 3, 1, 1, 3, 9.0)
    torch.cuda.synchronize()
    assert np.all(out.cpu().numpy() == 9.0)     # H=2: no row with 0<y<1
    leg.lib.padDataRightWall(P(out.data_ptr()), P(ds.data_ptr()), 5, 6, 2, 3, 3, 3, 1, 1)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert list(got[0][:3]) == [1, 2, 3] and got[0][3] == 3 and got[0][4] == np.float32(0.0001)
    assert np.all(got[2][:3] == np.float32(0.0001))   # constant band below the data


@pytest.mark.parametrize("fH,fW,H,W,KY,KX", [(9, 11, 3, 4, 2, 2), (8, 13, 3, 4, 1, 6), (7, 7, 2, 2, 4, 1)])
def test_right_wall_pad_case_table(vfd, leg, fH, fW, H, W, KY, KX):
    """padDataRightWall over the whole box against its case table: data; to
    the right a band mirrored in x (rows: data, then mirrored in y, then the
    far-side mirror fH-1-y); 0.0001 everywhere else and where a mirror index
    falls outside the data (bands wider than the data in the second and third
    cases, far-side rows beyond the data in the first)."""
    import torch
    print(f"case right_wall box={fH}x{fW} data={H}x{W} centre=({KY},{KX})", flush=True)
    src = (np.arange(H * W, dtype=np.float32) + 1).reshape(H, W)
    ds = _t(src)
    out = torch.empty((fH, fW), device="cuda")
    P = ctypes.c_void_p
    leg.lib.padDataRightWall(P(out.data_ptr()), P(ds.data_ptr()), fH, fW, H, W, 3, 3, KY, KX)
    torch.cuda.synchronize()
    c = np.float32(0.0001)
    expect = np.full((fH, fW), c, np.float32)
    ys, xs = np.arange(fH), np.arange(fW)
    expect[:H, :W] = src
    band_x = xs[(xs >= W) & (xs < W + KX)]
    for y in ys:
        sy = y if y < H else (2 * H - 1 - y if y < H + KY else fH - 1 - y)
        for x in band_x:
            sx = 2 * W - 1 - x
            if 0 <= sy < H and sx >= 0:
                expect[y, x] = src[sy, sx]
    assert np.array_equal(out.cpu().numpy(), expect)


def test_convolution_clamp_cpu(vfd, leg):
    """convolutionClampToBorderCPU (host): identity kernel and clamp edges."""
    data = np.arange(12, dtype=np.float32).reshape(3, 4)
    k = np.zeros((3, 3), np.float32)
    k[1, 1] = 1.0
    res = np.zeros_like(data)
    leg.lib.convolutionClampToBorderCPU(res.ctypes.data, data.ctypes.data, k.ctypes.data, 3, 4, 3, 3, 1, 1)
    assert np.array_equal(res, data)
    k2 = np.zeros((3, 3), np.float32)
    k2[1, 0] = 1.0   # picks data[y][x+1] with clamp
    leg.lib.convolutionClampToBorderCPU(res.ctypes.data, data.ctypes.data, k2.ctypes.data, 3, 4, 3, 3, 1, 1)
    assert list(res[0]) == [1, 2, 3, 3]


def _gauss_dft(g, P, nk):
    """DFT over a length-P box of the centred symmetric window g (length 2R+1):
    G[k] = g0 + 2 sum_j g_j cos(2 pi j k / P), k < nk, FP64."""
    R = (len(g) - 1) // 2
    k = np.arange(nk, dtype=np.float64)[:, None]
    j = np.arange(1, R + 1, dtype=np.float64)[None, :]
    return g[R] + 2.0 * (np.cos(2.0 * np.pi * j * k / P) * g[R + 1:][None, :]).sum(axis=1)


def test_fftDiffuse3D_production_shape(vfd, orc, leg):
    """The ABM's production configuration through the reference C interface:
    a 110x1390x1006 (x, y, z) chem field on the 256x1536x1152 FFT box,
    constant-VF pad (epithelium at x = 106, baseChem on the faces and beyond),
    complex R2C-layout spectrum of a sigma = 15.3-patch Gaussian (radius 92,
    the production kernel's width).  Sub-boxes at a corner of the baseChem
    region, around a source, and at the far faces are checked against an FP64
    direct convolution of the padded field with the oracle's window."""
    import torch
    D, H, W = 1006, 1390, 110
    box = (1152, 1536, 256)
    epi, base, decay = 106, 0.37, 0.98
    sigma2 = 15.3 ** 2
    print(f"case fftDiffuse3D production D,H,W={D},{H},{W} box={box}", flush=True)
    g = orc.gauss_window1d(sigma2)
    R = (len(g) - 1) // 2
    assert R <= min(box[0] - D, box[1] - H, box[2] - W)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(17)
    u = torch.exp(torch.randn((D, H, W), generator=gen, device="cuda") - 2.0)
    src = [(500, 700, 50), (120, 300, 20), (900, 1200, 100)]
    for z, y, x in src:
        u[z, y, x] += 5000.0
    ud = u.clone()
    gz = torch.from_numpy(_gauss_dft(g, box[0], box[0])).cuda()
    gy = torch.from_numpy(_gauss_dft(g, box[1], box[1])).cuda()
    gx = torch.from_numpy(_gauss_dft(g, box[2], box[2] // 2 + 1)).cuda()
    ks = ((decay * gz)[:, None, None] * gy[None, :, None] * gx[None, None, :]).to(torch.complex64)
    cc = leg.CCtx()
    cc.DD, cc.DH, cc.DW = D, H, W
    cc.FFTD, cc.FFTH, cc.FFTW = box
    cc.KD = cc.KH = cc.KW = 2 * R + 1
    cc.KZ = cc.KY = cc.KX = R
    pf, pi = ctypes.c_int(), ctypes.c_int()
    assert leg.lib.vfdPlan3d(ctypes.byref(pf), *box, 0x2a) == 0
    assert leg.lib.vfdPlan3d(ctypes.byref(pi), *box, 0x2c) == 0
    assert leg.lib.fftDiffuse3D(ctypes.c_void_p(ud.data_ptr()), ctypes.c_void_p(ks.data_ptr()), pf, pi, cc,
                                epi, ctypes.c_float(base), 0, 0)
    torch.cuda.synchronize()
    assert leg.lib.vfdDestroy(pf) == 0 and leg.lib.vfdDestroy(pi) == 0
    del ks

    def expected(lo, n):
        # padded field on [lo - R, lo + n + R) per axis: the field inside
        # 0<z<D-1, 0<y<H-1, 0<x<epi, baseChem everywhere else
        rng = [np.arange(l - R, l + m + R) for l, m in zip(lo, n)]
        inside = [(r > 0) & (r < lim) for r, lim in zip(rng, (D - 1, H - 1, epi))]
        cl = [np.clip(r, 0, lim - 1) for r, lim in zip(rng, (D, H, W))]
        blk = u[cl[0][0]:cl[0][-1] + 1, cl[1][0]:cl[1][-1] + 1, cl[2][0]:cl[2][-1] + 1].double().cpu().numpy()
        blk = blk[np.ix_(cl[0] - cl[0][0], cl[1] - cl[1][0], cl[2] - cl[2][0])]
        mask = inside[0][:, None, None] & inside[1][None, :, None] & inside[2][None, None, :]
        a = np.where(mask, blk, base)
        for ax in (2, 1, 0):                           # x, y, z: valid part of the correlation
            m = a.shape[ax] - 2 * R
            a = sum(g[j] * np.take(a, np.arange(j, j + m), axis=ax) for j in range(2 * R + 1))
        return decay * a

    for lo, n in [((0, 0, 96), (12, 12, 14)), ((494, 694, 44), (12, 12, 12)),
                  ((994, 1378, 0), (12, 12, 12)), ((114, 294, 14), (12, 12, 12))]:
        ref = expected(lo, n)
        got = ud[lo[0]:lo[0] + n[0], lo[1]:lo[1] + n[1], lo[2]:lo[2] + n[2]].double().cpu().numpy()
        print(f"  sub-box at {lo}:", end="", flush=True)
        assert _rel(got, ref) <= 1e-5


@pytest.mark.parametrize("split", [False, True])
def test_vfdTickChems_matches_diffuse_then_update(vfd, orc, leg, split):
    """The resident tick (vfdTickChems: deltas in, diffusion with the update
    fused, levels and gradients out; or its split form vfdTickChemsBegin /
    vfdTickChemsEnd around the agents' step) against the reference flow:
    vfdDiffuseChems (pChem -> tChem) then the host update of the oracle.
    Levels, gradients and cleared deltas bit-exact, totals to 1e-12; two ticks,
    so the second starts from the levels kept on the device."""
    import torch
    shape, box, epi = (10, 12, 14), (24, 24, 32), 12
    print(f"case vfdTickChems split={split} shape={shape} box={box}", flush=True)
    rng = np.random.default_rng(21)
    n = int(np.prod(shape))
    K = 5
    specs = []
    for ic in range(8):
        w = rng.random((K, K, K))
        w = (w + w[::-1, ::-1, ::-1]) / (2 * w.sum())
        specs.append(_t(orc.kernel_spectrum(w, (2, 2, 2), box).astype(np.complex64)))
    base = (ctypes.c_float * 8)(*[0.01 * (ic + 1) for ic in range(8)])
    decay = [0.319, 0.00003052, 0.67, 0.2, 0.2, 0.7937, 0.9284, 0.2]   # woundHealingWorld.cpp:2636-2643
    dec = (ctypes.c_float * 8)(*decay)
    pin = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).pin_memory()

    def cctx(d_data, h_in, h_out):
        cc = leg.CCtx()
        cc.DD, cc.DH, cc.DW = shape
        cc.FFTD, cc.FFTH, cc.FFTW = box
        cc.KD = cc.KH = cc.KW = K
        cc.KZ = cc.KY = cc.KX = 2
        for ic in range(8):
            cc.d_data[ic] = d_data[ic].data_ptr()
            cc.h_ibuffs[ic] = h_in[ic].data_ptr()
            cc.h_obuffs[ic] = h_out[ic].data_ptr() if h_out else 0
            cc.d_kernelspectrum_h[ic] = specs[ic].data_ptr()
            cc.dev_id[ic] = 0
        return cc

    levels = [rng.lognormal(-2, 1, shape).astype(np.float32) for _ in range(8)]
    dev_levels = [_t(x) for x in levels]
    h_levels = [pin(np.zeros(shape)) for _ in range(8)]
    cc_tick = cctx(dev_levels, h_levels, None)
    for tick in range(2):
        deltas = [(rng.standard_normal(shape) * 0.01).astype(np.float32) for _ in range(8)]
        fc, fe = rng.random(shape).astype(np.float32), rng.random(shape).astype(np.float32)
        # reference flow
        h_p = [pin(x) for x in levels]
        h_t = [pin(np.zeros(shape)) for _ in range(8)]
        scratch = [torch.empty(shape, dtype=torch.float32, device="cuda") for _ in range(8)]
        assert leg.lib.vfdDiffuseChems(cctx(scratch, h_p, h_t), 8, epi, base), vfd.last_error()
        p_ref, d_ref, g_ref, tot_ref = orc.chem_update([x.numpy().reshape(-1) for x in h_t],
                                                       [x.reshape(-1) for x in deltas], decay,
                                                       fc.reshape(-1), fe.reshape(-1))
        # resident tick
        h_d = [pin(x) for x in deltas]
        h_fc, h_fe = pin(fc), pin(fe)
        h_g = [pin(np.zeros(shape)) for _ in range(3)]
        tot = (ctypes.c_double * 8)()
        dv = (ctypes.c_void_p * 8)(*[x.data_ptr() for x in h_d])
        gv = (ctypes.c_void_p * 3)(*[x.data_ptr() for x in h_g])
        if split:
            assert leg.lib.vfdTickChemsBegin(cc_tick, epi, base), vfd.last_error()
            # no other diffusion while the split tick is pending
            assert not leg.lib.vfdDiffuseChems(cctx(scratch, h_p, h_t), 8, epi, base)
            assert "pending" in vfd.last_error()
            # End on another host thread, as after the agents' join
            import threading
            res = []
            th = threading.Thread(target=lambda: res.append(
                (leg.lib.vfdTickChemsEnd(cc_tick, dv, dec, h_fc.data_ptr(), h_fe.data_ptr(), gv, tot),
                 vfd.last_error())))
            th.start()
            th.join()
            assert res[0][0], res[0][1]
        else:
            assert leg.lib.vfdTickChems(cc_tick, epi, base, dv, dec, h_fc.data_ptr(), h_fe.data_ptr(), gv, tot), \
                vfd.last_error()
        for ic in range(8):
            got = h_levels[ic].numpy().reshape(-1)
            assert np.array_equal(got, p_ref[ic]), (tick, ic)
            assert np.array_equal(dev_levels[ic].cpu().numpy().reshape(-1), p_ref[ic]), (tick, ic)
            assert not h_d[ic].numpy().any(), (tick, ic)
            assert abs(tot[ic] - tot_ref[ic]) <= 1e-12 * abs(tot_ref[ic]), (tick, ic)
        for g in range(3):
            assert np.array_equal(h_g[g].numpy().reshape(-1), g_ref[g]), (tick, g)
        levels = [x.reshape(shape) for x in p_ref]


def test_vfdTickChems_rejects_shared_levels(vfd, leg):
    import torch
    shape, box = (4, 4, 4), (8, 8, 8)
    cc = leg.CCtx()
    cc.DD, cc.DH, cc.DW = shape
    cc.FFTD, cc.FFTH, cc.FFTW = box
    d = torch.zeros(shape, dtype=torch.float32, device="cuda")
    h = torch.zeros(shape, dtype=torch.float32).pin_memory()
    k = torch.zeros((8, 8, 5), dtype=torch.complex64, device="cuda")
    for ic in range(8):
        cc.d_data[ic] = d.data_ptr()           # one buffer shared by every chem: refused
        cc.h_ibuffs[ic] = h.data_ptr()
        cc.d_kernelspectrum_h[ic] = k.data_ptr()
    f8 = (ctypes.c_float * 8)(*([0.5] * 8))
    dv = (ctypes.c_void_p * 8)(*([h.data_ptr()] * 8))
    gv = (ctypes.c_void_p * 3)(*([h.data_ptr()] * 3))
    assert not leg.lib.vfdTickChems(cc, 4, f8, dv, f8, h.data_ptr(), h.data_ptr(), gv, (ctypes.c_double * 8)())
    assert "share d_data" in vfd.last_error()


def test_release_workspaces_then_diffuse_again(vfd, orc, leg):
    """vfdReleaseWorkspaces frees the per-box workspace; the next fftDiffuse3D
    makes it again and gives the same result."""
    import torch
    shape, box, K, epi, base = (10, 12, 9), (24, 24, 24), 7, 7, 0.5
    u = orc.synth_field(shape, 5, nsrc=4)
    rng = np.random.default_rng(11)
    w = rng.random((K, K, K))
    w = (w + w[::-1, ::-1, ::-1]) / (2 * w.sum())
    spec = _t(orc.kernel_spectrum(w, (3, 3, 3), box).astype(np.complex64))
    cc = leg.CCtx()
    cc.DD, cc.DH, cc.DW = shape
    cc.FFTD, cc.FFTH, cc.FFTW = box
    outs = []
    for _ in range(2):
        d = _t(u)
        assert leg.lib.fftDiffuse3D(ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(spec.data_ptr()), 0, 0, cc, epi,
                                    base, 0, 0), vfd.last_error()
        torch.cuda.synchronize()
        outs.append(d)
        leg.lib.vfdReleaseWorkspaces.restype = ctypes.c_int
        assert leg.lib.vfdReleaseWorkspaces() == 0, vfd.last_error()
    assert torch.equal(outs[0], outs[1])


# ---------------------------------------------------------------- 2D operator
def _cc2d(leg, H, W, fH, fW, KH, KW, KY, KX, win=(1, 1)):
    cc = leg.CCtx()
    cc.DD, cc.DH, cc.DW = 1, H, W
    cc.FFTD, cc.FFTH, cc.FFTW = 1, fH, fW
    cc.KD, cc.KH, cc.KW = 1, KH, KW
    cc.KZ, cc.KY, cc.KX = 0, KY, KX
    cc.windowD, cc.windowH, cc.windowW = 1, win[0], win[1]
    return cc


@pytest.mark.parametrize("H,W,fH,fW,KY,KX,R,win", [
    (16, 20, 32, 32, 1, 1, 6, (9, 9)),
    (15, 18, 24, 36, 1, 1, 4, (7, 5)),
    (12, 12, 16, 16, 0, 2, 3, (5, 5)),      # stencil centre off its middle: a phase in S^
    (10, 14, 16, 32, 1, 1, 0, (3, 3)),      # no convolution round: the stencil itself
])
def test_computeKernel2D_brute_force(vfd, orc, leg, H, W, fH, fW, KY, KX, R, win):
    """computeKernel (Diffusion.cu:303-675) against a brute-force FP64 restatement:
    the stencil at the domain centre, clamp-to-border pad once (oracle), R
    circular convolutions with the stencil by shifted sums on the FFT box,
    crop, centred window (oracle)."""
    import torch
    print(f"case computeKernel2D H,W={H},{W} box={fH},{fW} centre={KY},{KX} R={R}", flush=True)
    lam, gamma, dt = np.float32(0.11), np.float32(0.002), np.float32(1.5)
    cc = _cc2d(leg, H, W, fH, fW, 3, 3, KY, KX, win)
    d_win = torch.full(win, -1.0, dtype=torch.float32, device="cuda")
    assert leg.lib.computeKernel(ctypes.c_void_p(d_win.data_ptr()), R, lam, gamma, dt, cc), vfd.last_error()
    c0 = 1.0 - 4.0 * float(lam) - float(gamma) * float(dt)
    hk = np.zeros((3, 3))
    hk[0, 1] = hk[1, 0] = hk[1, 2] = hk[2, 1] = float(lam)
    hk[1, 1] = c0
    d0 = np.zeros((H, W))
    d0[H // 2 - 1:H // 2 + 2, W // 2 - 1:W // 2 + 2] = hk
    p = orc.pad_clamp(d0[None], (1, fH, fW), 0, KY, KX)[0]
    for _ in range(R):
        p = sum(hk[i, j] * np.roll(p, (i - KY, j - KX), axis=(0, 1)) for i in range(3) for j in range(3))
    ref = orc.extract_center(p[:H, :W][None], (1,) + win)[0]
    assert _rel(d_win.cpu().numpy(), ref) <= 1e-6


def test_computeKernelSpectrum2D_packed(vfd, orc, leg):
    """computeKernelSpectrum (Diffusion.cu:677-771): the C2C transform of the
    padded kernel read as FFTH x FFTW/2 complex, against numpy's fft2."""
    import torch
    fH, fW, KH, KW, KY, KX = 24, 32, 5, 7, 2, 3
    k = np.random.default_rng(5).random((KH, KW)).astype(np.float32)
    cc = _cc2d(leg, 8, 8, fH, fW, KH, KW, KY, KX)
    out = torch.zeros((fH, fW // 2), dtype=torch.complex64, device="cuda")
    kd = _t(k)
    assert leg.lib.computeKernelSpectrum(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(kd.data_ptr()), cc, cc), \
        vfd.last_error()
    p = orc.pad_kernel(k[None], (1, fH, fW), 0, KY, KX)[0]
    ref = np.fft.fft2(p[:, 0::2] + 1j * p[:, 1::2])
    got = out.cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"  packed spectrum rel-L2 {err:.3e}", flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("H,W,fH,fW,epi,base", [(20, 26, 32, 48, 22, 0.3), (17, 9, 24, 16, 100, 1.25), (9, 12, 16, 16, 0, 0.5)])
def test_fftDiffuse2D(vfd, orc, leg, H, W, fH, fW, epi, base):
    """fftDiffuse2D (Diffusion.cu:773-907) with a packed spectrum of a
    non-symmetric 5x5 kernel: the constant-VF pad (the bit-exact 2D pad helper),
    numpy's circular convolution on the box, crop; and the same through the
    spectrum this library's computeKernelSpectrum builds."""
    import torch
    print(f"case fftDiffuse2D H,W={H},{W} box={fH},{fW} epi={epi}", flush=True)
    rng = np.random.default_rng(H * W)
    u = rng.random((H, W)).astype(np.float32)
    k = rng.random((5, 5)).astype(np.float32)
    kp = orc.pad_kernel(k[None], (1, fH, fW), 0, 2, 2)[0]
    packed = np.fft.fft2(kp[:, 0::2] + 1j * kp[:, 1::2]).astype(np.complex64)
    cc = _cc2d(leg, H, W, fH, fW, 5, 5, 2, 2)
    ud = _t(u)
    pad = torch.empty((fH, fW), dtype=torch.float32, device="cuda")
    P = ctypes.c_void_p
    leg.lib.padDataConstantVF(P(pad.data_ptr()), P(ud.data_ptr()), fH, fW, H, W, 5, 5, 2, 2, epi, base)
    torch.cuda.synchronize()
    p = pad.cpu().numpy().astype(np.float64)
    ref = np.real(np.fft.ifft2(np.fft.fft2(p) * np.fft.fft2(kp)))[:H, :W]
    for src in ("numpy", "library"):
        d = ud.clone()
        if src == "numpy":
            ks = _t(packed)
        else:
            ks = torch.zeros((fH, fW // 2), dtype=torch.complex64, device="cuda")
            assert leg.lib.computeKernelSpectrum(P(ks.data_ptr()), P(_t(k).data_ptr()), cc, cc), vfd.last_error()
        assert leg.lib.fftDiffuse2D(P(d.data_ptr()), P(ks.data_ptr()), cc, epi, ctypes.c_float(base)), vfd.last_error()
        print(f"  spectrum from {src}:", end="", flush=True)
        assert _rel(d.cpu().numpy(), ref) <= 1e-5
