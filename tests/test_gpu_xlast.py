"""The x-last pass order (z R2C of column pairs, y, x fused with the Hermitian
extension of the R2C-layout spectrum, y inverse, z C2R) through the C-ABI.

Same operator as the x-first order, so both are checked against the FP64
oracle (orc.fft_diffuse3d: pad -> R2C -> x K^/(FD FH FW) -> C2R -> unpad,
restating src/Diffusion/Diffusion.cu:1622-1934) with rel-L2 <= 1e-5, on boxes
with odd and even nx, constant-VF and zero pads, and a kernel that is not
symmetric (its spectrum is complex, so the mirrored half must be conjugated).
The fused chem update and visualisation downsample of the z C2R pass are
bit-identical to the unfused forms, as in the x-first order.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_F32 = 1e-5


def _rel(tag, got, ref):
    err = np.abs(got.astype(np.float64) - ref)
    rel = float(np.linalg.norm(err) / np.linalg.norm(ref))
    print(f"{tag}: max-abs {err.max():.3e} rel-L2 {rel:.3e}", flush=True)
    return rel


def _plan(vfd, monkeypatch, order, shape, box, pad):
    monkeypatch.setenv("VFD_ORDER", str(order))
    plan = vfd.Plan(shape, box, vfd.F32, pad, vfd.KSPEC_COMPLEX_STD)
    monkeypatch.delenv("VFD_ORDER")
    return plan


def _spectrum(orc, box, key, symmetric=False):
    rng = np.random.default_rng(key)
    win = rng.random((5, 5, 5))
    if symmetric:
        win = 0.5 * (win + win[::-1, ::-1, ::-1])
    win /= win.sum()
    return orc.kernel_spectrum(win, (2, 2, 2), box)


@pytest.mark.parametrize("shape,box,pad,epi,base", [
    ((20, 24, 11), (32, 32, 64), 3, 8, 0.3),        # odd nx: the last pair has one column
    ((50, 40, 30), (64, 48, 64), 100, 0, 0.0),      # zero pad
    ((100, 60, 21), (128, 72, 64), 3, 30, 1.5),     # epiBoundary beyond the row: bounded at nx
    ((25, 30, 40), (32, 40, 192), 3, 0, 0.5),       # no interior: the whole box is baseChem
    ((30, 17, 100), (32, 18, 256), 3, 97, 0.2),     # radix-3 y box
    ((20, 600, 6), (32, 1024, 32), 3, 15, 0.4),     # y rows past P/2: folded split, Q = 512
    ((12, 800, 5), (32, 1536, 32), 100, 0, 0.0),    # folded split, Q = 768, zero pad
    ((20, 12, 700), (32, 16, 1152), 3, 15, 0.3),    # z pairs of 1152-long columns: split, rows past 576 folded
    ((10, 8, 900), (32, 16, 1536), 100, 0, 0.0),    # z pairs of 1536-long columns, zero pad
])
def test_xlast_matches_oracle_and_xfirst(vfd, orc, monkeypatch, shape, box, pad, epi, base):
    import torch
    print(f"case xlast shape={shape} box={box} pad={pad}", flush=True)
    u = orc.synth_field(shape, 91, nsrc=4)
    spec = _spectrum(orc, box, sum(shape))
    ref = orc.fft_diffuse3d(u, spec, box, pad, epi, base)
    ks = torch.from_numpy(spec.astype(np.complex64)).cuda()
    ud = torch.from_numpy(u).cuda()
    outs = {}
    for order in (vfd.ORDER_XLAST, vfd.ORDER_XFIRST):
        plan = _plan(vfd, monkeypatch, order, shape, box, pad)
        assert plan.info.order == order
        out = torch.full_like(ud, float("nan"))
        plan.diffuse([{"in": ud, "out": out, "work": plan.new_work(), "kspec": ks, "base": base, "epi": epi}])
        torch.cuda.synchronize()
        outs[order] = out.cpu().numpy()
        assert np.isfinite(outs[order]).all()
        assert _rel(f"order {order} vs oracle", outs[order], ref) <= REL_F32
    assert _rel("x-last vs x-first", outs[vfd.ORDER_XLAST], outs[vfd.ORDER_XFIRST].astype(np.float64)) <= REL_F32


def test_xlast_selection(vfd, monkeypatch):
    """Taken by itself where it saves bytes (the production box), never for
    FP64, non-complex spectra or a box padded alike on every axis; VFD_ORDER=0
    forces x-first."""
    monkeypatch.delenv("VFD_ORDER", raising=False)
    prod = vfd.Plan((1006, 1390, 110), (1152, 1536, 256), vfd.F32, vfd.PAD_CONSTANT_VF, vfd.KSPEC_COMPLEX_STD)
    assert prod.info.order == vfd.ORDER_XLAST
    # 8N + 32 (B1 + B2) + 8 hz py px bytes; B1 = 110*1390*577, B2 = 110*1536*577
    n = 110 * 1390 * 1006
    b1, b2 = 110 * 1390 * 577, 110 * 1536 * 577
    assert prod.info.bytes_per_field_tick == pytest.approx(8 * n + 32 * (b1 + b2) + 8 * 577 * 1536 * 256)
    assert sum(vfd.lib.vfd_plan_pass_bytes(prod.h, p) for p in range(5)) == pytest.approx(
        prod.info.bytes_per_field_tick)
    assert vfd.Plan((64, 64, 64), None, vfd.F32, vfd.PAD_ZERO, vfd.KSPEC_COMPLEX_STD).info.order == vfd.ORDER_XFIRST
    assert vfd.Plan((1006, 1390, 110), (1152, 1536, 256), vfd.F64, vfd.PAD_CONSTANT_VF,
                    vfd.KSPEC_COMPLEX_STD).info.order == vfd.ORDER_XFIRST
    assert vfd.Plan((1006, 1390, 110), (1152, 1536, 256), vfd.F32, vfd.PAD_ZERO,
                    vfd.KSPEC_SEPARABLE).info.order == vfd.ORDER_XFIRST
    monkeypatch.setenv("VFD_ORDER", "0")
    assert vfd.Plan((1006, 1390, 110), (1152, 1536, 256), vfd.F32, vfd.PAD_CONSTANT_VF,
                    vfd.KSPEC_COMPLEX_STD).info.order == vfd.ORDER_XFIRST


@pytest.mark.parametrize("shape,box", [((20, 17, 33), (32, 18, 64)), ((32, 48, 64), (64, 64, 128))])
def test_xlast_fused_update_and_sample_bit_identical(vfd, orc, monkeypatch, shape, box):
    """The chem update (WHChemical::update) and the sampleChem downsample fused
    into the z C2R pass against the unfused forms on the same plan."""
    import torch
    print(f"case xlast fused shape={shape} box={box}", flush=True)
    plan = _plan(vfd, monkeypatch, vfd.ORDER_XLAST, shape, box, vfd.PAD_CONSTANT_VF)
    assert plan.info.order == vfd.ORDER_XLAST
    g = torch.Generator(device="cuda").manual_seed(sum(shape))
    decay = [0.319, 0.00003052, 0.67, 0.2, 0.2, 0.7937, 0.9284, 0.2]   # woundHealingWorld.cpp:2636-2643
    u = [torch.rand(shape, generator=g, device="cuda") for _ in range(8)]
    d = [torch.randn(shape, generator=g, device="cuda") * 0.01 for _ in range(8)]
    fc, fe = torch.rand(shape, generator=g, device="cuda"), torch.rand(shape, generator=g, device="cuda")
    ks = [torch.from_numpy(_spectrum(orc, box, 7 + i).astype(np.complex64)).cuda() for i in range(8)]
    ent = lambda i, o: {"in": u[i], "out": o, "work": plan.new_work(), "kspec": ks[i], "base": 0.01 * i,
                        "epi": shape[2] - 3}
    t = [torch.empty_like(x) for x in u]
    plan.diffuse([ent(i, t[i]) for i in range(8)])
    d_ref = [x.clone() for x in d]
    p_ref = [torch.empty_like(x) for x in u]
    g_ref = [torch.empty_like(fc) for _ in range(3)]
    tot_ref = vfd.chem_update(t, d_ref, p_ref, decay, fc, fe, g_ref)
    d_fu = [x.clone() for x in d]
    p_fu = [torch.empty_like(x) for x in u]
    g_fu = [torch.empty_like(fc) for _ in range(3)]
    tot_fu = plan.diffuse_update([ent(i, p_fu[i]) for i in range(8)], d_fu, decay, fc, fe, g_fu)
    stride, chems = 2, [0, 3, 7]
    vis = [torch.full_like(vfd.sample_chem(t[i], chems[k], stride), -1.0) for k, i in enumerate((0, 1, 2))]
    out = [torch.empty_like(u[i]) for i in range(3)]
    plan.diffuse_sample([ent(i, out[i]) for i in range(3)], chems, stride, vis)
    torch.cuda.synchronize()
    for k in range(8):
        assert torch.equal(p_fu[k], p_ref[k]), k
        assert torch.equal(d_fu[k], d_ref[k]) and not d_fu[k].any(), k
        assert tot_fu[k] == tot_ref[k], k
    for j in range(3):
        assert torch.equal(g_fu[j], g_ref[j]), j
    for k in range(3):
        assert torch.equal(out[k], t[k]), k
        assert torch.equal(vis[k], vfd.sample_chem(t[k], chems[k], stride)), k


def test_y_fold_selection(vfd):
    """The FP32 inverse y pass of columns longer than half their box takes
    the folded even/odd split where it is compiled (P/2 = 512 or 768), the
    ABM's production box among them; FP64 plans and columns within half their
    box keep the plain split."""
    print("case y_fold_selection", flush=True)
    box = (32, 1024, 32)
    assert vfd.Plan((20, 600, 6), box, vfd.F32, vfd.PAD_ZERO, vfd.KSPEC_COMPLEX_STD).info.y_inv_fold == 1
    assert vfd.Plan((20, 600, 6), box, vfd.F64, vfd.PAD_ZERO, vfd.KSPEC_COMPLEX_STD).info.y_inv_fold == 0
    assert vfd.Plan((20, 500, 6), box, vfd.F32, vfd.PAD_ZERO, vfd.KSPEC_COMPLEX_STD).info.y_inv_fold == 0
    prod = vfd.Plan((1006, 1390, 110), (1152, 1536, 256), vfd.F32, vfd.PAD_CONSTANT_VF, vfd.KSPEC_COMPLEX_STD)
    assert prod.info.y_inv_fold == 1
