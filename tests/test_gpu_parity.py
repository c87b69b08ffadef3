"""HIP path vs the FP64 oracle through the C-ABI (include/vfd.h).

Bar: rel-L2 <= 1e-5 (FP32 path) / 1e-11 (FP64 path) against the FP64 oracle,
mass drift |sum_gpu - sum_oracle| / sum_oracle <= 1e-6 (FP32) with Neumaier
sums; max-abs is printed for every case.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_F32 = 1e-5
REL_F64 = 1e-11
MASS_F32 = 1e-6


def _report(tag, got, ref):
    err = np.abs(got.astype(np.float64) - ref)
    rel = float(np.linalg.norm(err) / np.linalg.norm(ref))
    print(f"{tag}: max-abs {err.max():.3e} rel-L2 {rel:.3e}", flush=True)
    return rel


def _native_case(vfd, orc, shape, nfields, key, precision, nsrc=5):
    import torch
    plan = vfd.Plan(shape, precision=precision)
    dt = torch.float32 if precision == vfd.F32 else torch.float64
    us, outs, works, ks, refs, params = [], [], [], [], [], []
    # kernel radius ceil(6 sigma) must stay <= n-1 on every axis (linear convolution)
    dmax = min(5.0, 0.98 * ((min(shape) - 1) / 6.0) ** 2 / 2.0)
    for f in range(nfields):
        u = orc.synth_field(shape, key * 100 + f, nsrc=nsrc)
        D, lam = orc.synth_params(key, f, drange=(0.5, dmax))
        s2, dec = 2.0 * D, float(np.exp(-lam))
        params.append((s2, dec))
        refs.append(orc.direct_separable(u, s2, dec))
        us.append(torch.from_numpy(u.astype(np.float32 if precision == vfd.F32 else np.float64)).cuda())
        outs.append(torch.empty_like(us[-1]))
        works.append(plan.new_work())
        ks.append(plan.gauss_spectrum(s2, dec))
    plan.diffuse([dict(**{"in": us[i]}, out=outs[i], work=works[i], kspec=ks[i]) for i in range(nfields)])
    torch.cuda.synchronize()
    mass = plan.mass(outs)
    return plan, [o.cpu().numpy() for o in outs], refs, mass


@pytest.mark.parametrize("shape", [(32, 64, 64), (16, 20, 24), (12, 24, 48), (9, 18, 36), (8, 8, 12)])
def test_native_gauss_f32(vfd, orc, shape):
    print(f"case native_gauss_f32 shape={shape}", flush=True)
    nf = 8 if shape == (32, 64, 64) else 3
    plan, outs, refs, mass = _native_case(vfd, orc, shape, nf, 11, vfd.F32)
    for i, (o, r) in enumerate(zip(outs, refs)):
        rel = _report(f"field {i}", o, r)
        assert rel <= REL_F32
        mref = orc.neumaier_sum(r)
        assert abs(mass[i] - mref) / mref <= MASS_F32


@pytest.mark.parametrize("shape", [(16, 20, 24), (12, 24, 48)])
def test_native_gauss_f64(vfd, orc, shape):
    print(f"case native_gauss_f64 shape={shape}", flush=True)
    plan, outs, refs, mass = _native_case(vfd, orc, shape, 2, 5, vfd.F64)
    for i, (o, r) in enumerate(zip(outs, refs)):
        assert _report(f"field {i}", o, r) <= REL_F64


@pytest.mark.parametrize("shape,box,epi,base", [
    ((7, 9, 10), (16, 12, 16), 8, 0.25),
    ((6, 5, 12), (8, 16, 24), 12, 1.5),
    ((11, 13, 14), (24, 32, 36), 9, 0.0),
])
def test_constant_vf_complex_spectrum(vfd, orc, shape, box, epi, base):
    """Reference semantics: CONSTANT_VF pad, complex R2C-layout spectrum,
    arbitrary (non-2N) FFT box, scale 1/(FD FH FW)."""
    import torch
    print(f"case constant_vf shape={shape} box={box}", flush=True)
    rng = np.random.default_rng(3)
    u = orc.synth_field(shape, 77, nsrc=3)
    K = 5
    win = rng.random((K, K, K))
    win = 0.5 * (win + win[::-1, ::-1, ::-1])
    spec = orc.kernel_spectrum(win, (2, 2, 2), box)
    ref = orc.fft_diffuse3d(u, spec, box, orc.PAD_CONSTANT_VF, epi, base)
    plan = vfd.Plan(shape, box, vfd.F32, vfd.PAD_CONSTANT_VF, vfd.KSPEC_COMPLEX_STD)
    ud = torch.from_numpy(u).cuda()
    out = torch.empty_like(ud)
    ks = torch.from_numpy(spec.astype(np.complex64)).cuda()
    plan.diffuse([{"in": ud, "out": out, "work": plan.new_work(), "kspec": ks, "base": base, "epi": epi}])
    torch.cuda.synchronize()
    assert _report("cvf", out.cpu().numpy(), ref) <= REL_F32


@pytest.mark.parametrize("box", [(8, 12, 16), (6, 10, 18), (16, 32, 64)])
def test_r2c_c2r_roundtrip(vfd, box):
    import torch
    print(f"case r2c box={box}", flush=True)
    rng = np.random.default_rng(1)
    a = rng.standard_normal(box)
    plan = vfd.Plan(box, box, vfd.F64)
    ad = torch.from_numpy(a).cuda()
    hx = box[2] // 2 + 1
    c = torch.empty((box[0], box[1], hx), dtype=torch.complex128, device="cuda")
    plan.r2c(ad, c)
    torch.cuda.synchronize()
    ref = np.fft.rfftn(a)
    got = c.cpu().numpy()
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
    back = torch.empty_like(ad)
    plan.c2r(c, back)
    torch.cuda.synchronize()
    assert np.abs(back.cpu().numpy() / np.prod(box) - a).max() < 1e-12


@pytest.mark.parametrize("shape", [(16, 20, 24), (12, 24, 48), (32, 64, 64)])
@pytest.mark.parametrize("precision", [4, 8])
def test_separable_spectrum_bit_identical_to_real(vfd, orc, shape, precision):
    """VFD_KSPEC_SEPARABLE (FP64 tables, K^ formed in the z pass) gives the
    same bits as the streamed 3D real spectrum, and matches the oracle."""
    import torch
    print(f"case separable_vs_real shape={shape} precision={precision}", flush=True)
    dt = torch.float32 if precision == 4 else torch.float64
    pr = vfd.Plan(shape, precision=precision, kspec_format=vfd.KSPEC_REAL)
    ps = vfd.Plan(shape, precision=precision, kspec_format=vfd.KSPEC_SEPARABLE)
    assert ps.info.kspec_bytes < 64 * 1024
    ps.schedule = pr.schedule
    dmax = min(5.0, 0.98 * ((min(shape) - 1) / 6.0) ** 2 / 2.0)
    ent_r, ent_s, refs = [], [], []
    for f in range(3):
        u = orc.synth_field(shape, 900 + f, nsrc=5)
        D, lam = orc.synth_params(13, f, drange=(0.5, dmax))
        s2, dec = 2.0 * D, float(np.exp(-lam))
        refs.append(orc.direct_separable(u, s2, dec))
        ud = torch.from_numpy(u.astype(np.float64)).to(dt).cuda()
        ent_r.append({"in": ud, "out": torch.empty_like(ud), "work": pr.new_work(), "kspec": pr.gauss_spectrum(s2, dec)})
        ent_s.append({"in": ud, "out": torch.empty_like(ud), "work": ps.new_work(), "kspec": ps.gauss_spectrum(s2, dec)})
    pr.diffuse(ent_r)
    ps.diffuse(ent_s)
    torch.cuda.synchronize()
    for f in range(3):
        a, b = ent_r[f]["out"].cpu().numpy(), ent_s[f]["out"].cpu().numpy()
        assert np.array_equal(a, b), f"field {f}"
        assert _report(f"field {f}", b, refs[f]) <= (REL_F32 if precision == 4 else REL_F64)
