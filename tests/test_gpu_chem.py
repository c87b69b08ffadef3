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
