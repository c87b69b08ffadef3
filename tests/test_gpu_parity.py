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
    ((6, 7, 8), (12, 16, 16), 11, 0.75),     # epiBoundary beyond the row: bounded at W
    ((6, 7, 8), (12, 16, 16), 0, 0.5),       # no interior: the whole box is baseChem
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
def test_separable_spectrum_matches_real(vfd, orc, shape, precision):
    """VFD_KSPEC_SEPARABLE against the streamed 3D real spectrum.  FP64: the z
    pass forms K^ from the FP64 tables and the ticks are bit-identical.  FP32:
    the z pass applies the z window as a direct convolution (the same operator:
    gz is the DFT of that window), so the two agree to FP32 rounding, and both
    match the oracle."""
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
        if precision == 8:
            assert np.array_equal(a, b), f"field {f}"
        else:
            rel = float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(a.astype(np.float64)))
            print(f"  field {f}: separable (direct z) vs real (FFT z) rel-L2 {rel:.3e}", flush=True)
            assert rel <= 1e-6, f"field {f}"
        assert _report(f"field {f}", b, refs[f]) <= (REL_F32 if precision == 4 else REL_F64)


def test_f64_arrays_beyond_4gib(vfd, orc):
    """The FP64 path addresses its arrays with 64-bit offsets: at 1536x768x192
    the FP64 work array is 7.4 GB.  One tick there agrees with the FP32 path
    (whose arrays stay below 4 GiB) to FP32 accuracy."""
    import torch
    shape = (192, 768, 1536)
    print(f"case f64_beyond_4gib shape={shape}", flush=True)
    p64 = vfd.Plan(shape, precision=vfd.F64, kspec_format=vfd.KSPEC_SEPARABLE)
    p32 = vfd.Plan(shape, precision=vfd.F32, kspec_format=vfd.KSPEC_SEPARABLE)
    assert p64.info.work_bytes > (1 << 32)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    u = torch.rand(shape, generator=g, device="cuda")
    u.view(-1)[::9973] += 100.0
    u64 = u.double()
    o32, o64 = torch.empty_like(u), torch.empty_like(u64)
    p32.diffuse([{"in": u, "out": o32, "work": p32.new_work(), "kspec": p32.gauss_spectrum(3.1, 0.97)}])
    p64.diffuse([{"in": u64, "out": o64, "work": p64.new_work(), "kspec": p64.gauss_spectrum(3.1, 0.97)}])
    torch.cuda.synchronize()
    rel = float(torch.linalg.vector_norm(o32.double() - o64) / torch.linalg.vector_norm(o64))
    m32, m64 = p32.mass([o32])[0], p64.mass([o64])[0]
    print(f"rel-L2 FP32 vs FP64 {rel:.3e}, mass drift {abs(m32 - m64) / m64:.3e}", flush=True)
    assert rel <= REL_F32
    assert abs(m32 - m64) / m64 <= MASS_F32
    # the last plane of the last rows (the top of the 7.4 GB array) is filled
    assert torch.isfinite(o64[-1]).all() and float(o64[-1].abs().sum()) > 0


def test_precision_many_ticks(vfd, orc):
    """100 ticks on a wide-dynamic-range field (log-uniform 1e-6..1e3 plus
    sharp sources): FP64 GPU against the chained FP64 oracle, FP32 GPU against
    FP64 GPU; max-abs, rel-L2 and mass drift reported per checkpoint."""
    import torch
    shape = (24, 40, 48)
    print(f"case precision_many_ticks shape={shape}", flush=True)
    rng = np.random.default_rng(11)
    u = (10.0 ** rng.uniform(-6, 3, shape)).astype(np.float32)
    u.reshape(-1)[rng.integers(0, u.size, 20)] += np.float32(800.0)
    s2, dec = 2.0 * 1.7, float(np.exp(-0.01))
    p64 = vfd.Plan(shape, precision=vfd.F64, kspec_format=vfd.KSPEC_SEPARABLE)
    p32 = vfd.Plan(shape, precision=vfd.F32, kspec_format=vfd.KSPEC_SEPARABLE)
    f64 = torch.from_numpy(u.astype(np.float64)).cuda()
    f32 = torch.from_numpy(u).cuda()
    # field_array keeps raw pointers: the work and spectrum tensors must stay referenced
    w64, k64 = p64.new_work(), p64.gauss_spectrum(s2, dec)
    w32, k32 = p32.new_work(), p32.gauss_spectrum(s2, dec)
    a64 = p64.field_array([{"in": f64, "out": f64, "work": w64, "kspec": k64}])
    a32 = p32.field_array([{"in": f32, "out": f32, "work": w32, "kspec": k32}])
    ref = u.astype(np.float64)
    for tick in range(1, 101):
        p64.diffuse(a64)
        p32.diffuse(a32)
        ref = orc.direct_separable(ref, s2, dec)
        if tick in (1, 10, 100):
            torch.cuda.synchronize()
            g64, g32 = f64.cpu().numpy(), f32.cpu().numpy().astype(np.float64)
            r64 = orc.rel_l2(g64, ref)
            r32 = orc.rel_l2(g32, g64)
            m64, m32 = p64.mass([f64])[0], p32.mass([f32])[0]
            mref = orc.neumaier_sum(ref)
            print(f"tick {tick}: FP64 vs oracle rel-L2 {r64:.3e} max-abs {np.abs(g64 - ref).max():.3e} "
                  f"mass drift {abs(m64 - mref) / mref:.3e}; FP32 vs FP64 rel-L2 {r32:.3e} "
                  f"max-abs {np.abs(g32 - g64).max():.3e} mass drift {abs(m32 - m64) / m64:.3e}", flush=True)
            assert r64 <= 1e-12 * tick
            assert abs(m64 - mref) / mref <= 1e-12
            # per-tick bounds: FP32 twiddle rounding gives each frequency a
            # systematic gain error, so drift accumulates linearly in ticks
            assert r32 <= REL_F32 * tick
            assert abs(m32 - m64) / m64 <= MASS_F32 * tick


def test_cpu_reference_workload_8_fields(vfd, orc):
    """The CPU-runnable reference workload: 8 FP32 fields on 64x64x32,
    lognormal(-2, 1) background plus 50 point sources of U[1, 100], D ~ U[0.5, 5],
    lambda ~ U[0.001, 0.05], one tick, against the FP64 oracle (rel-L2 <= 1e-5,
    mass drift <= 1e-6 per field), through the native separable path."""
    import torch
    shape = (32, 64, 64)
    print(f"case cpu_reference_workload shape={shape}", flush=True)
    plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
    ents, refs = [], []
    for f in range(8):
        u = orc.synth_field(shape, 2026 + f, nsrc=50)
        D, lam = orc.synth_params(2026, f)
        s2, dec = 2.0 * D, float(np.exp(-lam))
        refs.append(orc.direct_separable(u, s2, dec))
        ud = torch.from_numpy(u).cuda()
        ents.append({"in": ud, "out": torch.empty_like(ud), "work": plan.new_work(),
                     "kspec": plan.gauss_spectrum(s2, dec)})
    arr = plan.field_array(ents)
    plan.diffuse(arr)
    torch.cuda.synchronize()
    mass = plan.mass([e["out"] for e in ents])
    for f in range(8):
        got = ents[f]["out"].cpu().numpy()
        assert _report(f"field {f}", got, refs[f]) <= REL_F32
        mref = orc.neumaier_sum(refs[f])
        assert abs(mass[f] - mref) / mref <= MASS_F32


@pytest.mark.parametrize("shape", [(1, 1, 2), (1, 4, 6), (3, 1, 8), (2, 3, 2), (5, 7, 3), (1, 2, 40), (64, 1, 1),
                                   (1, 1, 1)])
def test_degenerate_grids(vfd, orc, shape):
    """Grids with 1- or 2-voxel axes and odd extents: the default FFT box is
    the smallest compiled length >= 2n, and a kernel radius up to P - n per axis
    still gives the linear convolution."""
    import torch
    print(f"case degenerate shape={shape}", flush=True)
    for precision in (4, 8):
        plan = vfd.Plan(shape, precision=precision, kspec_format=vfd.KSPEC_SEPARABLE)
        box = plan.box
        assert all(p >= 2 * n for p, n in zip(box, shape))
        r_max = min(p - n for p, n in zip(box, shape))
        s2 = 0.95 * (r_max / 6.0) ** 2
        assert orc.gauss_radius(s2) <= r_max
        u = orc.synth_field(shape, 77, nsrc=1)
        ref = orc.direct_separable(u, s2, 0.97)
        dt = torch.float32 if precision == 4 else torch.float64
        ud = torch.from_numpy(u).to(dt).cuda()
        out = torch.empty_like(ud)
        plan.diffuse([{"in": ud, "out": out, "work": plan.new_work(), "kspec": plan.gauss_spectrum(s2, 0.97)}])
        torch.cuda.synchronize()
        assert _report(f"{shape} prec {precision}", out.cpu().numpy(), ref) <= (REL_F32 if precision == 4 else REL_F64)


def test_kernel_radius_beyond_padding_is_refused(vfd):
    """A Gaussian whose radius exceeds the zero padding would wrap around:
    the spectrum builder refuses it instead of returning a circular result."""
    plan = vfd.Plan((8, 8, 12), box=(12, 16, 24), kspec_format=vfd.KSPEC_SEPARABLE)
    with pytest.raises(vfd.VfdError) as ei:
        plan.gauss_spectrum(4.0, 0.99)      # radius ceil(6*2) = 12 > 12 - 8 on z
    assert ei.value.code == vfd.VFD_ERR_INVALID and "padding" in str(ei.value)


def test_more_than_32_fields_and_empty_batch(vfd, orc):
    """Batches above the 32 fields of one launch run in chunks; an empty batch
    is a no-op."""
    import torch
    shape = (4, 6, 8)
    print(f"case many_fields shape={shape}", flush=True)
    plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
    plan.diffuse([])
    ents, refs = [], []
    for f in range(40):
        u = orc.synth_field(shape, 500 + f, nsrc=2)
        s2 = 0.1 + 0.005 * f          # radius ceil(6 sigma) <= 4 = P - n on z
        refs.append(orc.direct_separable(u, s2, 0.99))
        ud = torch.from_numpy(u).cuda()
        ents.append({"in": ud, "out": torch.empty_like(ud), "work": plan.new_work(),
                     "kspec": plan.gauss_spectrum(s2, 0.99)})
    plan.diffuse(ents)
    torch.cuda.synchronize()
    for f in range(40):
        assert orc.rel_l2(ents[f]["out"].cpu().numpy(), refs[f]) <= REL_F32, f


@pytest.mark.parametrize("shape,box,epi,base,s2", [
    ((11, 13, 14), (24, 32, 36), 9, 0.75, 1.0),
    ((7, 9, 10), (16, 24, 24), 8, 2.5, 2.2),
])
def test_constant_vf_real_spectrum(vfd, orc, shape, box, epi, base, s2):
    """CONSTANT_VF pad with the native real spectrum, whose 1/(FD FH FW) and
    decay are folded into the stored values: the restore term base * sum(k)
    must come out as base * decay (the oracle's complex spectrum is stored
    unnormalized, like the reference's)."""
    import torch
    print(f"case constant_vf_real shape={shape} box={box} base={base}", flush=True)
    dec = 0.93
    u = orc.synth_field(shape, 78, nsrc=3)
    win = orc.gauss_window3d(s2, dec)
    R = (win.shape[0] - 1) // 2
    spec = orc.kernel_spectrum(win, (R, R, R), box)
    ref = orc.fft_diffuse3d(u, spec, box, orc.PAD_CONSTANT_VF, epi, base)
    plan = vfd.Plan(shape, box, vfd.F32, vfd.PAD_CONSTANT_VF, vfd.KSPEC_REAL)
    ud = torch.from_numpy(u).cuda()
    out = torch.empty_like(ud)
    ks = plan.gauss_spectrum(s2, dec)
    plan.diffuse([{"in": ud, "out": out, "work": plan.new_work(), "kspec": ks, "base": base, "epi": epi}])
    torch.cuda.synchronize()
    assert _report("cvf real", out.cpu().numpy(), ref) <= REL_F32


def test_separable_table_at_reused_address(vfd, orc):
    """The z-pass form follows the radius each separable table stores, not
    where the table lives: a table copied into memory that held a table of
    another radius (as the caching allocator hands back), and a fresh copy at
    a new address, both match the oracle."""
    import torch
    shape = (32, 64, 64)
    print(f"case separable_table_reused shape={shape}", flush=True)
    plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
    u = orc.synth_field(shape, 4321, nsrc=5)
    ud = torch.from_numpy(u).cuda()
    work = plan.new_work()
    small, large = (1.0, 0.98), (9.0, 0.95)      # radius 6 (direct z pass), 18 (FFT z pass)
    assert orc.gauss_radius(small[0]) <= 8 and orc.gauss_radius(large[0]) > 16
    slot = plan.gauss_spectrum(*small)            # this memory first holds the radius-6 table
    cases = []
    for name, (s2, dec) in (("large into small's memory", large), ("small into large's memory", small)):
        src = plan.gauss_spectrum(s2, dec)
        slot.copy_(src)
        del src
        cases.append((name, slot, s2, dec))
        out = torch.empty_like(ud)
        plan.diffuse([{"in": ud, "out": out, "work": work, "kspec": slot}])
        torch.cuda.synchronize()
        assert _report(name, out.cpu().numpy(), orc.direct_separable(u, s2, dec)) <= REL_F32
    fresh = slot.clone()                          # a copy at an address no table was written to
    out = torch.empty_like(ud)
    plan.diffuse([{"in": ud, "out": out, "work": work, "kspec": fresh}])
    torch.cuda.synchronize()
    assert _report("fresh copy", out.cpu().numpy(), orc.direct_separable(u, *small)) <= REL_F32
    # both forms in one launch: a batch mixing the two radii
    ks = [plan.gauss_spectrum(*small), plan.gauss_spectrum(*large)]
    outs = [torch.empty_like(ud), torch.empty_like(ud)]
    works = [work, plan.new_work()]
    plan.diffuse([{"in": ud, "out": outs[i], "work": works[i], "kspec": ks[i]} for i in range(2)])
    torch.cuda.synchronize()
    for i, (s2, dec) in enumerate((small, large)):
        assert _report(f"mixed batch {i}", outs[i].cpu().numpy(), orc.direct_separable(u, s2, dec)) <= REL_F32


@pytest.mark.parametrize("nbytes", [16, 4096 + 48, (1 << 20) + 16 * 37])
def test_stream_reference_kernels(vfd, nbytes):
    """The streaming reference kernels move exactly the bytes asked for: the
    copy, the one-read-two-writes form and the two-reads-one-write form
    (32-bit integer lanes added), on sizes that end inside a work-group; the
    bytes after the end stay untouched."""
    import torch
    print(f"case stream_reference_kernels nbytes={nbytes}", flush=True)
    n = nbytes // 4
    g = torch.Generator(device="cpu").manual_seed(nbytes)
    a = torch.randint(0, 1 << 30, (n + 64,), generator=g, dtype=torch.int32).cuda()
    b = torch.randint(0, 1 << 30, (n + 64,), generator=g, dtype=torch.int32).cuda()
    c = torch.full((n + 64,), -7, dtype=torch.int32, device="cuda")
    d = torch.full((n + 64,), -9, dtype=torch.int32, device="cuda")
    vfd.stream_copy(c[:n], a[:n])
    torch.cuda.synchronize()
    assert torch.equal(c[:n], a[:n]) and bool((c[n:] == -7).all())
    vfd.stream_mix(c[:n], a[:n], dst1=d[:n])
    torch.cuda.synchronize()
    assert torch.equal(c[:n], a[:n]) and torch.equal(d[:n], a[:n]) and bool((d[n:] == -9).all())
    vfd.stream_mix(c[:n], a[:n], src1=b[:n])
    torch.cuda.synchronize()
    assert torch.equal(c[:n], a[:n] + b[:n]) and bool((c[n:] == -7).all())
