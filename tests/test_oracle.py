"""The CPU oracle, pinned against things it does not share code with:
the O(n^2) DFT definition, numpy golden vectors (tests/golden), closed forms,
brute-force direct convolution and hand-computed known answers.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_vectors.npz")


def dft_matrix(n, inverse=False):
    k = np.arange(n)
    s = 1.0 if inverse else -1.0
    return np.exp(s * 2j * np.pi * np.outer(k, k) / n)


# --------------------------------------------------------------- 1D FFT
@pytest.mark.parametrize("n", list(range(1, 65)) + [96, 97, 121, 128, 143])
def test_fft_matches_dft_definition(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = dft_matrix(n) @ x
    assert np.abs(orc.fft(x) - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    refi = dft_matrix(n, True) @ x
    assert np.abs(orc.fft(x, inverse=True) - refi).max() <= 1e-12 * max(1.0, np.abs(refi).max())


def test_fft_golden_vectors(orc):
    g = np.load(GOLDEN)
    for key in g.files:
        if key.startswith("fft_in_"):
            n = key[len("fft_in_"):]
            out = orc.fft(g[key])
            assert np.abs(out - g["fft_out_" + n]).max() < 1e-11, n


def test_fft_identities(orc):
    n = 48
    imp = np.zeros(n, complex)
    imp[0] = 1.0
    assert np.allclose(orc.fft(imp), np.ones(n), atol=1e-14)          # impulse -> constant
    k0 = 5
    cosv = np.cos(2 * np.pi * k0 * np.arange(n) / n)
    spec = orc.fft(cosv)
    expect = np.zeros(n, complex)
    expect[k0] = expect[n - k0] = n / 2                                 # cosine -> two bins
    assert np.abs(spec - expect).max() < 1e-12
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = orc.fft(x)
    assert abs(np.sum(np.abs(X) ** 2) - n * np.sum(np.abs(x) ** 2)) < 1e-9  # Parseval
    assert np.abs(orc.fft(X, inverse=True) / n - x).max() < 1e-13        # round trip
    sh = np.roll(x, 3)                                                   # shift theorem
    assert np.abs(orc.fft(sh) - X * np.exp(-2j * np.pi * 3 * np.arange(n) / n)).max() < 1e-11


# --------------------------------------------------------------- 3D real
def test_rfft3_golden_and_roundtrip(orc):
    g = np.load(GOLDEN)
    for key in g.files:
        if key.startswith("rfftn_in_"):
            tag = key[len("rfftn_in_"):]
            a = g[key]
            h = orc.rfft3(a)
            assert np.abs(h - g["rfftn_out_" + tag]).max() < 1e-11, tag
            back = orc.irfft3(h, a.shape[2]) / a.size
            assert np.abs(back - a).max() < 1e-12, tag


# ------------------------------------------------------------ pad rules
def test_pad_constant_vf_known_answer(orc):
    """Interior 0<z<D-1, 0<y<H-1, 0<x<epi keeps the field; everything else in
    the FFT box is baseChem (convolutionFFT.cuh:617-637)."""
    D, H, W = 4, 5, 6
    u = (np.arange(D * H * W, dtype=np.float32) + 1.0).reshape(D, H, W)
    epi = 4
    box = (6, 7, 9)
    p = orc.pad_constant_vf(u, box, epi, -7.0)
    keep = p != -7.0
    assert keep.sum() == (D - 2) * (H - 2) * (epi - 1)      # 2*3*3 interior voxels
    zz, yy, xx = np.nonzero(keep)
    assert zz.min() == 1 and zz.max() == D - 2
    assert yy.min() == 1 and yy.max() == H - 2
    assert xx.min() == 1 and xx.max() == epi - 1
    assert p[2, 3, 2] == u[2, 3, 2] and p[1, 1, 3] == u[1, 1, 3]
    # epi = W keeps x = 1..W-1; an epi beyond the row (where the reference reads
    # the next row) is bounded at W: voxels past the row are baseChem
    p2 = orc.pad_constant_vf(u, box, W, 0.0)
    assert (p2 != 0).sum() == (D - 2) * (H - 2) * (W - 1)
    p3 = orc.pad_constant_vf(u, box, W + 2, 0.0)
    assert np.array_equal(p3, p2)


def test_pad_clamp_known_answer(orc):
    """[0,N) data, [N,N+K) last slice, [N+K,P) first slice (convolutionFFT.cuh:537-582)."""
    d = np.arange(3, dtype=np.float64).reshape(1, 1, 3) + 10.0  # x: 10, 11, 12
    p = orc.pad_clamp(d, (1, 1, 8), 0, 0, 2)
    assert list(p[0, 0]) == [10, 11, 12, 12, 12, 10, 10, 10]


def test_pad_kernel_wrap_known_answer(orc):
    """Kernel centre (KZ,KY,KX) goes to the origin; negative offsets wrap."""
    k = np.zeros((1, 1, 3))
    k[0, 0] = [1.0, 2.0, 3.0]
    p = orc.pad_kernel(k, (1, 1, 6), 0, 0, 1)
    assert list(p[0, 0]) == [2.0, 3.0, 0.0, 0.0, 0.0, 1.0]


def test_unpad_and_extract_center(orc):
    a = np.arange(4 * 5 * 6, dtype=np.float64).reshape(4, 5, 6)
    assert np.array_equal(orc.unpad(a, (2, 3, 4)), a[:2, :3, :4])
    c = orc.extract_center(a, (2, 3, 2))
    assert np.array_equal(c, a[1:3, 1:4, 2:4])


def test_snap_transform_size(orc):
    known = {1: 16, 13: 16, 16: 16, 17: 32, 33: 64, 100: 128, 1000: 1024, 1024: 1024, 1025: 2048,
             2049: 4096, 597: 1024}
    for n, want in known.items():
        assert orc.snap_transform_size(n) == want, n


# -------------------------------------------------------------- operator
@pytest.mark.parametrize("pad", ["zero", "cvf"])
@pytest.mark.parametrize("shape,box,K,centre", [
    ((4, 5, 6), (8, 10, 12), (3, 3, 3), (1, 1, 1)),
    ((5, 4, 7), (9, 8, 10), (3, 2, 4), (1, 0, 2)),   # non-symmetric, off-centre window
    ((6, 6, 5), (12, 9, 15), (5, 3, 3), (2, 1, 1)),
])
def test_fft_diffuse_matches_direct_convolution(orc, pad, shape, box, K, centre):
    rng = np.random.default_rng(7)
    u = rng.random(shape).astype(np.float32) + 0.1
    w = rng.random(K)
    mode = orc.PAD_ZERO if pad == "zero" else orc.PAD_CONSTANT_VF
    epi, base = (shape[2] - 1, 0.37) if pad == "cvf" else (0, 0.0)
    spec = orc.kernel_spectrum(w, centre, box)
    v = orc.fft_diffuse3d(u, spec, box, mode, epi, base)
    ref = orc.direct_conv3d(u, w, centre, box, mode, epi, base)
    assert np.abs(v - ref).max() < 1e-12 * np.abs(ref).max()


def test_operator_identities(orc):
    shape, box = (6, 7, 8), (12, 14, 16)
    rng = np.random.default_rng(3)
    w = rng.random((3, 3, 3))
    # delta field at p -> kernel centred at p
    u = np.zeros(shape, np.float32)
    u[3, 3, 4] = 1.0
    v = orc.fft_diffuse3d(u, orc.kernel_spectrum(w, (1, 1, 1), box), box)
    assert np.abs(v[2:5, 2:5, 3:6] - w).max() < 1e-14
    assert abs(v.sum() - w.sum()) < 1e-13
    # delta kernel -> identity
    d = np.zeros((1, 1, 1))
    d[0, 0, 0] = 1.0
    u2 = rng.random(shape).astype(np.float32)
    v2 = orc.fft_diffuse3d(u2, orc.kernel_spectrum(d, (0, 0, 0), box), box)
    assert np.abs(v2 - u2).max() < 1e-14


def test_gauss_direct_separable_vs_fft(orc):
    shape = (12, 16, 20)
    s2, dec = 2.0 * 1.7, np.exp(-0.03)
    R = orc.gauss_radius(s2)
    assert R == int(np.ceil(6 * np.sqrt(s2)))
    u = orc.synth_field(shape, 9, nsrc=6)
    box = tuple(2 * s for s in shape)
    spec = orc.kernel_spectrum(orc.gauss_window3d(s2, dec), (R, R, R), box)
    v = orc.fft_diffuse3d(u, spec, box)
    d = orc.direct_separable(u, s2, dec)
    assert orc.rel_l2(v, d) < 1e-13
    # the real separable spectrum equals the R2C of the 3D window (imag ~ 0)
    ks = orc.gauss_spectrum_real(s2, dec, box, 1.0)
    assert np.abs(spec.imag).max() < 1e-13
    assert np.abs(ks - spec.real).max() < 1e-6 * np.abs(spec.real).max()


def test_gauss_mass_and_symmetry(orc):
    """A blob far from the box edges keeps decay * mass; the operator
    commutes with mirroring x (even kernel)."""
    shape = (40, 40, 40)
    u = np.zeros(shape, np.float32)
    u[18:22, 19:21, 20:23] = 3.0
    s2, dec = 2.0 * 0.8, np.exp(-0.02)
    v = orc.direct_separable(u, s2, dec)
    assert abs(v.sum() - dec * u.sum()) < 1e-10 * u.sum()
    vm = orc.direct_separable(np.ascontiguousarray(u[:, :, ::-1]), s2, dec)
    assert np.abs(vm[:, :, ::-1] - v).max() < 1e-14


def test_gauss_window_normalized(orc):
    for s2 in (1.0, 2.3, 10.0):
        g = orc.gauss_window1d(s2)
        assert abs(g.sum() - 1.0) < 1e-15
        R = (len(g) - 1) // 2
        assert np.allclose(g, g[::-1])
        assert g[R] == g.max()
        assert abs(g[R + 1] / g[R] - np.exp(-1.0 / (2 * s2))) < 1e-14


# ------------------------------------------------------- kernel build
def test_compute_kernel3d_matches_closed_form_heat_kernel(orc):
    """computeKernel3D (Diffusion.cu:1078-1531) on a domain wide enough that
    the clamp-to-border pad sees only ~0, vs inverse DFT of S(theta)^M."""
    kr, niter = 5, 2
    lam, gamma, dt = 0.1, 0.002, 2.0
    dom = ((kr - 1) * 2 + 1 + 20,) * 3          # 29^3
    box = (orc.snap_transform_size(dom[0] + 3 - 1),) * 3   # 32^3
    win = (9, 9, 9)
    w = orc.compute_kernel3d(kr, lam, gamma, dt, niter, dom, box, win, quirk=0)
    M = 1 + (kr - 2) + (niter - 1) * (kr - 1)   # stencil in the data + powers
    ref = orc.stencil_heat_window(lam, gamma, dt, M, 64, win)
    assert np.abs(w - ref).max() < 1e-12 * np.abs(ref).max()


def test_compute_kernel3d_mass_check(orc):
    """TEST_KERNEL (Diffusion.cu:1439-1454): with no decay the window sums to
    1 +- 1e-3 once it spans the kernel."""
    kr, niter = 6, 3
    dom = ((kr - 1) * 2 + 1,) * 3
    box = (orc.snap_transform_size(dom[0] + 2),) * 3
    w = orc.compute_kernel3d(kr, 0.05, 0.0, 2.0, niter, dom, box, (11, 11, 11))
    assert abs(w.sum() - 1.0) < 1e-3


def test_cpow_quirk_changes_sign_when_spectrum_negative(orc):
    """cpow's atan(b/a) (convolutionFFT.cuh:658-659) raises |S| instead of S
    where S < 0 (lambda > 1/12); the principal power flips sign for odd n."""
    kr, niter = 5, 1                 # first-iteration power kr-2 = 3 (odd)
    dom = (9, 9, 9)
    box = (16, 16, 16)
    wq = orc.compute_kernel3d(kr, 0.15, 0.0, 1.0, niter, dom, box, (9, 9, 9), quirk=1)
    wp = orc.compute_kernel3d(kr, 0.15, 0.0, 1.0, niter, dom, box, (9, 9, 9), quirk=0)
    assert np.abs(wq - wp).max() > 1e-6
    assert abs(wp.sum() - 1.0) < 1e-12   # principal power conserves mass (S(0) = 1)


# ---------------------------------------------------------- CPU port
@pytest.mark.parametrize("shape", [(8, 16, 32), (6, 12, 24), (9, 18, 12)])
def test_cpu_port_matches_fp64(orc, shape):
    s2, dec = 2.0 * 0.6, np.exp(-0.01)
    u = orc.synth_field(shape, 21, nsrc=4)
    box = tuple(2 * s for s in shape)
    ks = orc.gauss_spectrum_real(s2, dec, box, 1.0 / np.prod(box))
    c = orc.cpu_diffuse_f32(u, ks, 2)
    d = orc.direct_separable(u, s2, dec)
    assert orc.rel_l2(c, d) < 1e-6


def test_synth_deterministic(orc):
    a = orc.synth_field((4, 5, 6), 42)
    b = orc.synth_field((4, 5, 6), 42)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, orc.synth_field((4, 5, 6), 43))
    D, lam = orc.synth_params(42, 3)
    assert 0.5 <= D <= 5.0 and 0.001 <= lam <= 0.05


def test_separable_f64_input_matches_f32_entry(orc):
    """The FP64-input separable entry point (used to chain FP64 ticks) equals
    the FP32-input one on FP32-representable data, and keeps an FP64 field's
    low bits (an FP32 round trip of the input changes the result)."""
    u = orc.synth_field((6, 7, 9), 17, nsrc=3)
    a = orc.direct_separable(u, 1.3, 0.97)
    b = orc.direct_separable(u.astype(np.float64), 1.3, 0.97)
    assert np.array_equal(a, b)
    r1 = orc.direct_separable(a, 1.3, 0.97)
    r1f = orc.direct_separable(a.astype(np.float32), 1.3, 0.97)
    assert not np.array_equal(r1, r1f)
    assert orc.rel_l2(r1f, r1) < 1e-7


def test_chem_update_known_answers(orc):
    """WHChemical::update (gpuUpdate): hand-computed levels, gradients, totals."""
    n = 5
    t = [np.full(n, k + 1.0, np.float32) for k in range(8)]
    d = [np.full(n, 0.5, np.float32) for _ in range(8)]
    p, d_after, (neu, mac, fib), tot = orc.chem_update(t, d, [0.5] * 8, np.ones(n), np.full(n, 2.0))
    for k in range(8):
        assert np.all(p[k] == 0.5 + 0.5 * (k + 1))
        assert np.all(d_after[k] == 0.0)
        assert tot[k] == n * (0.5 + 0.5 * (k + 1))
    # IL1beta 3.0 + TNF 1.0 + TGF 1.5 + FGF 2.0 + frag_col 1.0 = 8.5
    assert np.all(fib == 1.5) and np.all(neu == 8.5 + 3.5 + 4.0) and np.all(mac == 8.5 + 2.0)


def test_chem_update_rounds_product_before_add(orc):
    """level = d + t*decay with the product rounded to FP32 first (no fused
    multiply-add): (1+2^-12)^2 rounds to 1+2^-11, so d = -(1+2^-11) gives 0,
    where a fused evaluation would give 2^-24."""
    e = np.float32(1 + 2.0 ** -12)
    t = [np.array([e], np.float32) for _ in range(8)]
    d = [np.array([-(1 + 2.0 ** -11)], np.float32) for _ in range(8)]
    p, _, _, _ = orc.chem_update(t, d, [float(e)] * 8, np.zeros(1), np.zeros(1))
    assert all(x[0] == 0.0 for x in p)


def test_sample_chem_known_answers(orc):
    """sampleChem (volumeRender_kernel.cu:215-300): stride-3 picks, per-chem
    display scale, clamp above at 1 (negative values pass through)."""
    u = np.zeros((7, 8, 10), np.float32)
    u[3, 6, 9] = 1e-5          # picked (3, 6, 9 are multiples of 3)
    u[3, 3, 0] = -2e-5         # picked, negative: scaled, not clamped
    u[1, 1, 1] = 5.0           # not a multiple of 3: never sampled
    u[0, 0, 3] = 1.0           # picked, clamped to 1
    s = orc.sample_chem(u, 0, 3)
    assert s.shape == (2, 2, 3)
    assert s[0, 0, 1] == 1.0
    assert s[1, 1, 0] == np.float32(-2e-5) * np.float32(50000.0)
    # (3, 6, 9) -> (1, 2, 3) lies outside D/3 x H/3 x W/3 = 2 x 2 x 3 (the reference keeps dx < W/3)
    assert np.count_nonzero(s) == 2
    s2 = orc.sample_chem(u, 2, 3)
    assert s2[0, 0, 1] == 1.0 and s2[1, 1, 0] == np.float32(-2e-5) * np.float32(1e6)
