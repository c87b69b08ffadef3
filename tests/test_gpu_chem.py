"""Post-diffusion chem update (WHChemical::update, gpuUpdate;
WHChemical.cpp:395-450) on the device against the oracle: levels, cleared
deltas and gradients bit-exact (FP32 with the product rounded before the add),
compensated totals to 1e-12 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 1000, 1 << 20, 3_000_001])
def test_chem_update_matches_oracle(vfd, orc, n):
    import torch
    print(f"case chem_update n={n}", flush=True)
    rng = np.random.default_rng(n)
    t = [rng.lognormal(-2, 1, n).astype(np.float32) for _ in range(8)]
    d = [(rng.standard_normal(n) * 0.01).astype(np.float32) for _ in range(8)]
    decay = [0.319, 0.00003052, 0.67, 0.2, 0.2, 0.7937, 0.9284, 0.2]   # woundHealingWorld.cpp:2636-2643
    fc = rng.random(n).astype(np.float32)
    fe = rng.random(n).astype(np.float32)
    p_ref, d_ref, g_ref, tot_ref = orc.chem_update(t, d, decay, fc, fe)
    dev = lambda a: torch.from_numpy(a).cuda()
    td, dd = [dev(x) for x in t], [dev(x) for x in d]
    pd = [torch.empty_like(x) for x in td]
    gd = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(3)]
    tot = vfd.chem_update(td, dd, pd, decay, dev(fc), dev(fe), gd)
    for k in range(8):
        assert np.array_equal(pd[k].cpu().numpy(), p_ref[k]), k
        assert np.array_equal(dd[k].cpu().numpy(), d_ref[k]), k
        assert abs(tot[k] - tot_ref[k]) <= 1e-12 * abs(tot_ref[k]) + 1e-300, k
    for g in range(3):
        assert np.array_equal(gd[g].cpu().numpy(), g_ref[g]), g


def test_chem_update_in_place(vfd, orc):
    """p may alias t (the diffused fields become the new levels in place)."""
    import torch
    n = 4097
    rng = np.random.default_rng(9)
    t = [rng.random(n).astype(np.float32) for _ in range(8)]
    d = [rng.random(n).astype(np.float32) for _ in range(8)]
    fc, fe = rng.random(n).astype(np.float32), rng.random(n).astype(np.float32)
    p_ref, _, g_ref, _ = orc.chem_update(t, d, [0.5] * 8, fc, fe)
    td = [torch.from_numpy(x).cuda() for x in t]
    dd = [torch.from_numpy(x).cuda() for x in d]
    gd = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(3)]
    vfd.chem_update(td, dd, td, [0.5] * 8, torch.from_numpy(fc).cuda(), torch.from_numpy(fe).cuda(), gd,
                    totals=False)
    torch.cuda.synchronize()
    for k in range(8):
        assert np.array_equal(td[k].cpu().numpy(), p_ref[k])


@pytest.mark.parametrize("shape,stride", [((32, 48, 64), 3), ((20, 17, 33), 2), ((64, 64, 64), 8)])
def test_sample_chem_matches_oracle(vfd, orc, shape, stride):
    import torch
    print(f"case sample_chem shape={shape} stride={stride}", flush=True)
    rng = np.random.default_rng(stride)
    u = (rng.lognormal(-12, 3, shape) * np.where(rng.random(shape) < 0.1, -1, 1)).astype(np.float32)
    ud = torch.from_numpy(u).cuda()
    for chem in range(8):
        got = vfd.sample_chem(ud, chem, stride).cpu().numpy()
        assert np.array_equal(got, orc.sample_chem(u, chem, stride)), chem


@pytest.mark.parametrize("shape", [(32, 48, 64), (20, 17, 33), (128, 512, 1024)])
def test_fused_diffuse_update_bit_identical(vfd, shape):
    """vfd_diffuse_update (update fused into the x-inverse epilogue) against
    vfd_diffuse followed by vfd_chem_update on the same inputs: levels,
    cleared deltas, gradients and totals all bit-identical."""
    import torch
    print(f"case fused_diffuse_update shape={shape}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(sum(shape))
    plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
    decay = [0.319, 0.00003052, 0.67, 0.2, 0.2, 0.7937, 0.9284, 0.2]   # woundHealingWorld.cpp:2636-2643
    u = [torch.rand(shape, generator=g, device="cuda") for _ in range(8)]
    d = [torch.randn(shape, generator=g, device="cuda") * 0.01 for _ in range(8)]
    fc, fe = torch.rand(shape, generator=g, device="cuda"), torch.rand(shape, generator=g, device="cuda")
    ks = [plan.gauss_spectrum(0.3 + 0.1 * i, 0.99) for i in range(8)]
    # reference order: diffuse into t, then the update p = d + t * decay
    t = [torch.empty_like(x) for x in u]
    plan.diffuse([{"in": u[i], "out": t[i], "work": plan.new_work(), "kspec": ks[i]} for i in range(8)])
    d_ref = [x.clone() for x in d]
    p_ref = [torch.empty_like(x) for x in u]
    g_ref = [torch.empty_like(fc) for _ in range(3)]
    tot_ref = vfd.chem_update(t, d_ref, p_ref, decay, fc, fe, g_ref)
    # fused
    d_fu = [x.clone() for x in d]
    p_fu = [torch.empty_like(x) for x in u]
    g_fu = [torch.empty_like(fc) for _ in range(3)]
    tot_fu = plan.diffuse_update([{"in": u[i], "out": p_fu[i], "work": plan.new_work(), "kspec": ks[i]}
                                  for i in range(8)], d_fu, decay, fc, fe, g_fu)
    torch.cuda.synchronize()
    for k in range(8):
        assert torch.equal(p_fu[k], p_ref[k]), k
        assert torch.equal(d_fu[k], d_ref[k]) and not d_fu[k].any(), k
        assert tot_fu[k] == tot_ref[k], k
    for j in range(3):
        assert torch.equal(g_fu[j], g_ref[j]), j


def test_fused_diffuse_update_in_place(vfd):
    """The levels may overwrite the input fields (out aliases in), as the ABM's tick does."""
    import torch
    shape = (24, 40, 56)
    g = torch.Generator(device="cuda").manual_seed(5)
    plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
    decay = [0.5] * 8
    u = [torch.rand(shape, generator=g, device="cuda") for _ in range(8)]
    d = [torch.rand(shape, generator=g, device="cuda") for _ in range(8)]
    fc, fe = torch.rand(shape, generator=g, device="cuda"), torch.rand(shape, generator=g, device="cuda")
    ks = [plan.gauss_spectrum(0.5, 0.97) for _ in range(8)]
    t = [torch.empty_like(x) for x in u]
    plan.diffuse([{"in": u[i], "out": t[i], "work": plan.new_work(), "kspec": ks[i]} for i in range(8)])
    p_ref, g_ref = [torch.empty_like(x) for x in u], [torch.empty_like(fc) for _ in range(3)]
    vfd.chem_update(t, [x.clone() for x in d], p_ref, decay, fc, fe, g_ref, totals=False)
    g_fu = [torch.empty_like(fc) for _ in range(3)]
    plan.diffuse_update([{"in": u[i], "out": u[i], "work": plan.new_work(), "kspec": ks[i]} for i in range(8)],
                        d, decay, fc, fe, g_fu, totals=False)
    torch.cuda.synchronize()
    for k in range(8):
        assert torch.equal(u[k], p_ref[k]), k
    for j in range(3):
        assert torch.equal(g_fu[j], g_ref[j]), j


def test_fused_diffuse_update_refuses_bad_arguments(vfd):
    import torch
    plan64 = vfd.Plan((8, 8, 8), precision=vfd.F64)
    z = torch.zeros((8, 8, 8), dtype=torch.float32, device="cuda")
    ents = [{"in": z, "out": z, "work": plan64.new_work(), "kspec": plan64.new_kspec()} for _ in range(8)]
    with pytest.raises(vfd.VfdError):
        plan64.diffuse_update(ents, [z] * 8, [0.5] * 8, z, z, [z] * 3)


@pytest.mark.parametrize("shape,stride", [((32, 48, 64), 3), ((20, 17, 33), 2), ((7, 9, 10), 4), ((128, 512, 1024), 3)])
def test_fused_sample_bit_identical(vfd, shape, stride):
    """vfd_diffuse_sample (the sampleChem downsample fused into the x-inverse
    epilogue) against vfd_diffuse followed by vfd_sample_chem: the diffused
    fields and the downsampled arrays bit-identical, for even and odd nx."""
    import torch
    print(f"case fused_sample shape={shape} stride={stride}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(sum(shape) + stride)
    plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
    nf = 3
    u = [torch.rand(shape, generator=g, device="cuda") * 1e-4 for _ in range(nf)]
    ks = [plan.gauss_spectrum(0.4 + 0.2 * i, 0.99) for i in range(nf)]
    chems = [0, 3, 7]
    ref_out = [torch.empty_like(x) for x in u]
    plan.diffuse([{"in": u[i], "out": ref_out[i], "work": plan.new_work(), "kspec": ks[i]} for i in range(nf)])
    ref_vis = [vfd.sample_chem(ref_out[i], chems[i], stride) for i in range(nf)]
    out = [torch.empty_like(x) for x in u]
    vis = [torch.full_like(r, -1.0) for r in ref_vis]
    plan.diffuse_sample([{"in": u[i], "out": out[i], "work": plan.new_work(), "kspec": ks[i]} for i in range(nf)],
                        chems, stride, vis)
    torch.cuda.synchronize()
    for i in range(nf):
        assert torch.equal(out[i], ref_out[i]), i
        assert torch.equal(vis[i], ref_vis[i]), i
    assert any(bool((r == 1.0).any()) for r in ref_vis) or shape[0] < 16  # the clamp is exercised


def test_fused_sample_refuses_bad_arguments(vfd):
    import torch
    plan = vfd.Plan((8, 8, 8))
    z = torch.zeros((8, 8, 8), dtype=torch.float32, device="cuda")
    ent = [{"in": z, "out": z, "work": plan.new_work(), "kspec": plan.new_kspec()}]
    with pytest.raises(vfd.VfdError):
        plan.diffuse_sample(ent, [8], 2, [z])
    with pytest.raises(vfd.VfdError):
        plan.diffuse_sample(ent, [0], 0, [z])
