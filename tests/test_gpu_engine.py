"""Engine: resident runs, batched ticks, hipGraph replay, plain per-field form,
and the continuous-batching server, through the C-ABI (include/vfd.h).

Batching must not change results: every field goes through the same kernels
with the same per-shape schedule whether it is ticked alone, in a batch of
mixed shapes, inside a replayed graph, one field at a time, or through the
server, so those comparisons are bit-exact.  Accuracy is checked against the
FP64 oracle over several ticks (rel-L2 <= 1e-5 per tick for FP32).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_F32 = 1e-5
REL_F64 = 1e-11
GiB = 1 << 30


@pytest.fixture(scope="module")
def eng_mod(vfd):
    import vfd.engine as e
    return e


def _params(orc, shape, nf, key):
    dmax = min(5.0, 0.98 * ((min(shape) - 1) / 6.0) ** 2 / 2.0)
    s2, dc = [], []
    for f in range(nf):
        D, lam = orc.synth_params(key, f, drange=(0.5, dmax))
        s2.append(2.0 * D)
        dc.append(float(np.exp(-lam)))
    return s2, dc


def _make_run(eng, orc, shape, nf, key, precision=4):
    import torch
    s2, dc = _params(orc, shape, nf, key)
    rid = eng.create_run(shape, s2, dc, precision=precision)
    us = []
    for f in range(nf):
        u = orc.synth_field(shape, key * 1000 + f, nsrc=5)
        us.append(u)
        t = torch.from_numpy(u.astype(np.float32 if precision == 4 else np.float64)).cuda()
        eng.write_field(rid, f, t)
    torch.cuda.synchronize()
    return rid, us, s2, dc


def _fields(eng, rid, nf):
    import torch
    out = [eng.read_field(rid, f) for f in range(nf)]
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def test_engine_ticks_match_oracle(vfd, orc, eng_mod):
    import torch
    print("case engine_ticks_match_oracle shape=(16,20,24) nf=3 ticks=3", flush=True)
    shape, nf = (16, 20, 24), 3
    with eng_mod.Engine(4 * GiB) as eng:
        rid, us, s2, dc = _make_run(eng, orc, shape, nf, 7)
        refs = list(us)
        for t in range(3):
            eng.step([rid])
            torch.cuda.synchronize()
            refs = [orc.direct_separable(r, s2[f], dc[f]) for f, r in enumerate(refs)]
            got = _fields(eng, rid, nf)
            for f in range(nf):
                rel = orc.rel_l2(got[f], refs[f])
                print(f"tick {t} field {f}: rel-L2 {rel:.3e}", flush=True)
                assert rel <= REL_F32 * (t + 1)


def test_engine_f64_run(vfd, orc, eng_mod):
    import torch
    print("case engine_f64_run shape=(12,24,48)", flush=True)
    shape, nf = (12, 24, 48), 2
    with eng_mod.Engine(4 * GiB) as eng:
        rid, us, s2, dc = _make_run(eng, orc, shape, nf, 9, precision=8)
        eng.step([rid], ticks=2)
        torch.cuda.synchronize()
        for f in range(nf):
            ref = orc.direct_separable(orc.direct_separable(us[f], s2[f], dc[f]), s2[f], dc[f])
            got = eng.read_field(rid, f).cpu().numpy()
            rel = orc.rel_l2(got, ref)
            print(f"field {f}: rel-L2 {rel:.3e}", flush=True)
            assert rel <= REL_F64


def test_batched_graph_plain_serve_bit_exact(vfd, orc, eng_mod):
    """The same run, ticked alone, in a mixed-shape batch, by graph replay,
    one field at a time, and through the server, gives identical bits."""
    import torch
    print("case batched_graph_plain_serve_bit_exact", flush=True)
    A, B = (16, 20, 24), (12, 24, 48)
    with eng_mod.Engine(8 * GiB) as eng:
        alone, _, _, _ = _make_run(eng, orc, A, 3, 21)
        batched, _, _, _ = _make_run(eng, orc, A, 3, 21)
        graphed, _, _, _ = _make_run(eng, orc, A, 3, 21)
        plain, _, _, _ = _make_run(eng, orc, A, 3, 21)
        served, _, _, _ = _make_run(eng, orc, A, 3, 21)
        other_a, _, _, _ = _make_run(eng, orc, A, 2, 22)
        other_b, _, _, _ = _make_run(eng, orc, B, 4, 23)
        other_b2, _, _, _ = _make_run(eng, orc, B, 1, 24)
        st = eng.stats()
        assert st["runs"] == 8 and st["shapes"] == 2
        for tick in range(3):
            eng.step([alone])
            eng.step([other_b, batched, other_a, other_b2])
            eng.step([graphed, other_b], graph=True)   # captured once, replayed on later ticks
            eng.step_plain([plain])
            torch.cuda.synchronize()
            ref = _fields(eng, alone, 3)
            for name, rid in (("batched", batched), ("graph", graphed), ("plain", plain)):
                got = _fields(eng, rid, 3)
                for f in range(3):
                    assert np.array_equal(got[f], ref[f]), f"{name} tick {tick} field {f}"
        eng.serve_start(slots=2)
        try:
            for tick in range(3):
                tickets = [eng.submit(r) for r in (other_b, served, other_a)]
                lat = [eng.wait(t) for t in tickets]
                assert all(x > 0.0 for x in lat)
        finally:
            eng.serve_stop()
        torch.cuda.synchronize()
        ref = _fields(eng, alone, 3)
        got = _fields(eng, served, 3)
        for f in range(3):
            assert np.array_equal(got[f], ref[f]), f"served field {f}"
        st = eng.stats()
        assert st["requests_done"] == 9 and st["requests_pending"] == 0


def test_graph_replays_match_eager(vfd, orc, eng_mod):
    """Several replays of one cached graph, with eager work and
    synchronizations in between, each checked against the eager result."""
    import torch
    print("case graph_replays_match_eager", flush=True)
    A = (9, 18, 36)
    with eng_mod.Engine(4 * GiB) as eng:
        e_run, _, _, _ = _make_run(eng, orc, A, 2, 31)
        g_run, _, _, _ = _make_run(eng, orc, A, 2, 31)
        for rep in range(4):
            eng.step([g_run], ticks=2, graph=True)
            torch.cuda.synchronize()
            eng.step([e_run], ticks=2)
            torch.cuda.synchronize()
            a, b = _fields(eng, e_run, 2), _fields(eng, g_run, 2)
            for f in range(2):
                assert np.array_equal(a[f], b[f]), f"replay {rep} field {f}"


def test_serving_order_and_many_requests(vfd, orc, eng_mod):
    """Back-to-back requests of one run apply in submission order: k requests
    equal k eager ticks."""
    import torch
    print("case serving_order_and_many_requests", flush=True)
    A = (8, 8, 12)
    with eng_mod.Engine(2 * GiB) as eng:
        ref_run, _, _, _ = _make_run(eng, orc, A, 2, 41)
        srv_run, _, _, _ = _make_run(eng, orc, A, 2, 41)
        eng.step([ref_run], ticks=5)
        eng.serve_start(slots=3, max_batch_fields=64)
        try:
            tickets = [eng.submit(srv_run) for _ in range(5)]
            lat = [eng.wait(t) for t in tickets]
        finally:
            eng.serve_stop()
        torch.cuda.synchronize()
        assert all(x > 0.0 for x in lat)
        a, b = _fields(eng, ref_run, 2), _fields(eng, srv_run, 2)
        for f in range(2):
            assert np.array_equal(a[f], b[f])


def test_budget_refusal(vfd, orc, eng_mod):
    print("case budget_refusal", flush=True)
    shape = (32, 64, 64)
    need = eng_mod.run_bytes(shape, 4)
    assert need > 0
    with eng_mod.Engine(need // 2) as eng:
        with pytest.raises(vfd.VfdError) as ei:
            eng.create_run(shape, [2.0] * 4, [0.99] * 4)
        assert ei.value.code == vfd.VFD_ERR_BUDGET
        st = eng.stats()
        assert st["used_bytes"] == 0 and st["runs"] == 0


def test_run_lifecycle_errors(vfd, orc, eng_mod):
    print("case run_lifecycle_errors", flush=True)
    with eng_mod.Engine(1 * GiB) as eng:
        rid, _, _, _ = _make_run(eng, orc, (8, 8, 12), 1, 51)
        with pytest.raises(vfd.VfdError) as ei:
            eng.step([rid + 100])
        assert ei.value.code == vfd.VFD_ERR_NOTFOUND
        with pytest.raises(vfd.VfdError):
            eng.submit(rid)                     # not serving
        used = eng.stats()["used_bytes"]
        assert used > 0
        eng.destroy_run(rid)
        assert eng.stats()["used_bytes"] == 0


def test_admission_waits_for_memory(vfd, orc, eng_mod):
    """vfd_run_create_wait: a run that does not fit waits until another run is
    destroyed, then is admitted; it gives up with VFD_ERR_BUDGET at its
    timeout; a run larger than the whole budget is refused at once."""
    import threading
    import time
    print("case admission_waits_for_memory", flush=True)
    shape, nf = (32, 64, 64), 4
    run = eng_mod.run_bytes(shape, nf)
    work = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE).info.work_bytes
    s2, dc = [2.0] * nf, [0.99] * nf
    with eng_mod.Engine(2 * run + work + run // 2) as eng:   # room for two runs and the shape's work buffer
        a = eng.create_run(shape, s2, dc)
        eng.create_run(shape, s2, dc)
        with pytest.raises(vfd.VfdError) as ei:
            eng.create_run(shape, s2, dc)                    # no wait: refused
        assert ei.value.code == vfd.VFD_ERR_BUDGET
        t0 = time.perf_counter()
        with pytest.raises(vfd.VfdError) as ei:
            eng.create_run_wait(shape, s2, dc, 0.3)          # waits, then gives up
        assert ei.value.code == vfd.VFD_ERR_BUDGET and time.perf_counter() - t0 >= 0.25
        with pytest.raises(vfd.VfdError) as ei:
            eng.create_run_wait((64, 128, 128), [2.0] * 32, [0.99] * 32, 30.0)   # never fits
        assert ei.value.code == vfd.VFD_ERR_BUDGET
        got = []
        th = threading.Thread(target=lambda: got.append(eng.create_run_wait(shape, s2, dc, 60.0)))
        th.start()
        time.sleep(0.5)
        assert th.is_alive() and not got                     # still waiting for memory
        eng.destroy_run(a)
        th.join(30.0)
        assert not th.is_alive() and len(got) == 1
        st = eng.stats()
        assert st["runs"] == 2 and st["used_bytes"] <= st["budget_bytes"]
        assert st["used_bytes"] <= st["peak_used_bytes"] <= st["budget_bytes"]
        assert st["runs_admitted_after_wait"] == 1           # the timed-out waits admitted nothing


def test_admission_frees_idle_work_buffers(vfd, orc, eng_mod):
    """A batch grows its shape's pool of work buffers (one per field here);
    once the batch is done those buffers are idle, and a run that fits only
    without them is admitted after they are freed (each shape keeps one)."""
    import torch
    print("case admission_frees_idle_work_buffers", flush=True)
    shape, nf = (32, 64, 64), 4
    run = eng_mod.run_bytes(shape, nf)
    work = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE).info.work_bytes
    s2, dc = [2.0] * nf, [0.99] * nf
    with eng_mod.Engine(2 * run + 2 * nf * work + run // 2) as eng:
        a = eng.create_run(shape, s2, dc)
        b = eng.create_run(shape, s2, dc)
        eng.step([a, b])
        torch.cuda.synchronize()
        grown = eng.stats()["used_bytes"] - 2 * run
        assert grown % work == 0 and grown // work == 2 * nf   # one work buffer per field of the batch
        c = eng.create_run(shape, s2, dc)                      # fits only without the idle buffers
        st = eng.stats()
        assert st["used_bytes"] == 3 * run + work and st["runs"] == 3
        assert st["runs_admitted_after_wait"] == 0
        eng.step([a, b, c])                                     # still ticks, with the buffers it can get
        torch.cuda.synchronize()
        assert eng.stats()["used_bytes"] <= st["budget_bytes"]
