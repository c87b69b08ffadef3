"""vfd_diffuse_update_noalloc (the allocation-free entry point) against
vfd_diffuse_update through the C-ABI: two calls in a row, the second on a
smaller grid, each bit-identical to the allocating entry point (levels,
cleared deltas, gradients, and the 8 compensated totals)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(shape, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    u = [torch.rand(shape, generator=g, device="cuda") for _ in range(8)]
    d = [torch.rand(shape, generator=g, device="cuda") * 0.01 for _ in range(8)]
    fc = torch.rand(shape, generator=g, device="cuda")
    fe = torch.rand(shape, generator=g, device="cuda")
    return u, d, fc, fe


def _tick(vfd, plan, u, d, fc, fe, decay, ks, noalloc):
    import torch
    lv = [x.clone() for x in u]
    dl = [x.clone() for x in d]
    grads = [torch.empty_like(fc) for _ in range(3)]
    ents = [{"in": lv[i], "out": lv[i], "work": plan.new_work(), "kspec": ks[i]} for i in range(8)]
    arr = plan.field_array(ents)
    if noalloc:
        tot = torch.full((8,), -1.0, dtype=torch.float64, device="cuda")
        plan.diffuse_update_noalloc(arr, dl, decay, fc, fe, grads, d_totals=tot)
        torch.cuda.synchronize()
        totals = tot.cpu().numpy()
    else:
        totals = np.array(plan.diffuse_update(arr, dl, decay, fc, fe, grads))
        torch.cuda.synchronize()
    return lv, dl, grads, totals


def test_noalloc_matches_diffuse_update(vfd):
    # the first grid has 393,216 voxels (1024 partial sums per chem, the
    # array's capacity), the second 53,760 (210): the second call reads only
    # what it wrote over the first call's contents
    decay = [0.319, 0.00003052, 0.67, 0.2, 0.2, 0.7937, 0.9284, 0.2]
    for call, shape in enumerate([(64, 64, 96), (24, 40, 56)]):
        print(f"case noalloc call {call} shape={shape}", flush=True)
        plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
        ks = [plan.gauss_spectrum(0.6 + 0.4 * i, 0.97) for i in range(8)]
        u, d, fc, fe = _inputs(shape, 100 + call)
        ref = _tick(vfd, plan, u, d, fc, fe, decay, ks, noalloc=False)
        got = _tick(vfd, plan, u, d, fc, fe, decay, ks, noalloc=True)
        for k in range(8):
            assert np.array_equal(got[0][k].cpu().numpy(), ref[0][k].cpu().numpy()), ("levels", call, k)
            assert float(got[1][k].abs().max()) == 0.0, ("deltas cleared", call, k)
        for j in range(3):
            assert np.array_equal(got[2][j].cpu().numpy(), ref[2][j].cpu().numpy()), ("gradient", call, j)
        assert np.array_equal(got[3], ref[3]), ("totals", call, got[3], ref[3])
        print(f"call {call}: totals {got[3][:3]} ... bit-identical", flush=True)


def test_noalloc_leaves_device_memory_unchanged(vfd):
    """Repeated calls take no device memory: the free-memory figure of the
    device is the same before and after (torch holds its own blocks)."""
    import torch
    shape = (32, 48, 64)
    plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
    ks = [plan.gauss_spectrum(1.0, 0.97) for _ in range(8)]
    u, d, fc, fe = _inputs(shape, 7)
    grads = [torch.empty_like(fc) for _ in range(3)]
    works = [plan.new_work() for _ in range(8)]
    arr = plan.field_array([{"in": u[i], "out": u[i], "work": works[i], "kspec": ks[i]} for i in range(8)])
    tot = torch.empty(8, dtype=torch.float64, device="cuda")
    plan.diffuse_update_noalloc(arr, d, [0.5] * 8, fc, fe, grads, d_totals=tot)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(5):
        plan.diffuse_update_noalloc(arr, d, [0.5] * 8, fc, fe, grads, d_totals=tot)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    assert torch.isfinite(tot).all()
