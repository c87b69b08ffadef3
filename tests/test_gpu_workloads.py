"""Parity at the workloads' own sizes, through the C-ABI (include/vfd.h).

* The 8-field 1024x512x128 tick on its 2048x1024x256 box: lognormal(-2, 1)
  background plus 50 point sources, D ~ U[0.5, 5], lambda ~ U[0.001, 0.05].
  The data key is chosen so that the fields' z windows (R = ceil(6 sqrt(2D)),
  here 8 to 19) take every z-pass form: the direct pass of radius bound 8, 12
  and 16 and the FFT form beyond 16.
    - tick 1 of every field against the FP64 oracle (full field), FP32 and the
      full-FP64 GPU path;
    - ticks 1..10 of the FP32 path against the FP64 GPU path: the one-tick
      error from the same FP32 state each tick, and the trajectory;
    - the separable spectrum (direct z pass) against the streamed real
      spectrum (FFT z pass), field by field.
* The batched-ensemble grids (64x64x32 ... 512x256x64) through the engine,
  and a 1024x512x128 run served among smaller runs: tick 1 against the
  oracle, and bit-identical alone and in the mix.

Bars: rel-L2 <= 1e-5 (FP32) / 1e-11 (FP64) and mass drift <= 1e-6 per tick
against the FP64 reference; max-abs is printed for every case.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_F32 = 1e-5
REL_F64 = 1e-11
MASS_F32 = 1e-6
GiB = 1 << 30

HEADLINE = (128, 512, 1024)   # nz, ny, nx
NFIELDS = 8
KEY = 1                       # z radii 11, 8, 15, 15, 18, 9, 19, 16
ENSEMBLE_GRIDS = [(32, 64, 64), (64, 128, 128), (64, 128, 256), (64, 256, 256), (64, 256, 512)]


def _zform(R):
    return 8 if R <= 8 else 12 if R <= 12 else 16 if R <= 16 else 0


def _params(orc, key, nf):
    out = []
    for f in range(nf):
        D, lam = orc.synth_params(key, f)
        out.append((2.0 * D, float(np.exp(-lam))))
    return out


def _errors(orc, got, ref):
    got = got.astype(np.float64)
    err = np.abs(got - ref)
    rel = float(np.linalg.norm(err) / np.linalg.norm(ref))
    return rel, float(err.max())


@pytest.fixture(scope="module")
def headline(orc):
    """The 8 fields, their (sigma2, decay) and the FP64 oracle's tick 1."""
    print(f"case headline fixture shape={HEADLINE} key={KEY}: oracle tick 1", flush=True)
    params = _params(orc, KEY, NFIELDS)
    radii = [orc.gauss_radius(s2) for s2, _ in params]
    assert {_zform(R) for R in radii} == {0, 8, 12, 16}, radii
    us, refs = [], []
    for f, (s2, dec) in enumerate(params):
        u = orc.synth_field(HEADLINE, KEY * 1000 + f, nsrc=50)
        us.append(u)
        refs.append(orc.direct_separable(u, s2, dec))
    return us, params, radii, refs


def _run_plan(vfd, plan, fields_in, params, ticks=1, outs=None, works=None, ks=None):
    """`ticks` ticks of every field in place (out = in after the first)."""
    import torch
    dt = torch.float32 if plan.precision == vfd.F32 else torch.float64
    if outs is None:
        outs = [torch.from_numpy(u).to(device="cuda", dtype=dt) for u in fields_in]
    if works is None:
        works = [plan.new_work() for _ in outs]
    if ks is None:
        ks = [plan.gauss_spectrum(s2, dec) for s2, dec in params]
    arr = plan.field_array([{"in": o, "out": o, "work": w, "kspec": k} for o, w, k in zip(outs, works, ks)])
    for _ in range(ticks):
        plan.diffuse(arr)
    torch.cuda.synchronize()
    return outs, works, ks


def test_headline_tick1_vs_oracle(vfd, orc, headline):
    us, params, radii, refs = headline
    plan = vfd.Plan(HEADLINE, kspec_format=vfd.KSPEC_SEPARABLE)
    sched, _ = plan.autotune(NFIELDS)
    print(f"case headline_tick1 schedule={sched}", flush=True)
    outs, _, _ = _run_plan(vfd, plan, us, params)
    mass = plan.mass(outs)
    for f in range(NFIELDS):
        rel, mx = _errors(orc, outs[f].cpu().numpy(), refs[f])
        mref = orc.neumaier_sum(refs[f])
        drift = abs(mass[f] - mref) / mref
        print(f"FP32 field {f} R={radii[f]} form={_zform(radii[f]) or 'fft'}: rel-L2 {rel:.3e} "
              f"max-abs {mx:.3e} mass drift {drift:.3e}", flush=True)
        assert rel <= REL_F32 and drift <= MASS_F32
    del outs
    plan64 = vfd.Plan(HEADLINE, precision=vfd.F64, kspec_format=vfd.KSPEC_SEPARABLE)
    for f in range(NFIELDS):   # one field at a time: 2.2 GB of FP64 work each
        o64, _, _ = _run_plan(vfd, plan64, [us[f]], [params[f]])
        rel, mx = _errors(orc, o64[0].cpu().numpy(), refs[f])
        print(f"FP64 field {f}: rel-L2 {rel:.3e} max-abs {mx:.3e}", flush=True)
        assert rel <= REL_F64


def test_headline_ten_ticks_vs_f64(vfd, orc, headline):
    """FP32 against the FP64 GPU path (pinned to the oracle at tick 1 above).
    Each tick: the one-tick error, FP32 tick vs FP64 tick of the same FP32
    state, must meet the per-tick bars; the trajectory against the chained
    FP64 ticks is reported and bounded by the per-tick bars accumulated."""
    import torch
    us, params, radii, _ = headline
    print("case headline_ten_ticks", flush=True)
    plan = vfd.Plan(HEADLINE, kspec_format=vfd.KSPEC_SEPARABLE)
    plan64 = vfd.Plan(HEADLINE, precision=vfd.F64, kspec_format=vfd.KSPEC_SEPARABLE)
    f32, w32, k32 = _run_plan(vfd, plan, us, params, ticks=0)
    f64 = [torch.from_numpy(u).to(device="cuda", dtype=torch.float64) for u in us]
    k64 = [plan64.gauss_spectrum(s2, dec) for s2, dec in params]
    w64 = plan64.new_work()
    step = torch.empty_like(f64[0])
    arr32 = plan.field_array([{"in": o, "out": o, "work": w, "kspec": k} for o, w, k in zip(f32, w32, k32)])
    for tick in range(1, 11):
        # the FP64 tick of the current FP32 state (one-tick reference)
        one = []
        for f in range(NFIELDS):
            step.copy_(f32[f].double())
            plan64.diffuse([{"in": step, "out": step, "work": w64, "kspec": k64[f]}])
            one.append(plan64.mass([step])[0])
            one[-1] = (one[-1], step.cpu().numpy())
            plan64.diffuse([{"in": f64[f], "out": f64[f], "work": w64, "kspec": k64[f]}])
        plan.diffuse(arr32)
        torch.cuda.synchronize()
        m32 = plan.mass(f32)
        m64 = plan64.mass(f64)
        worst = [0.0, 0.0, 0.0, 0.0]
        for f in range(NFIELDS):
            got = f32[f].cpu().numpy()
            rel1, _ = _errors(orc, got, one[f][1])
            d1 = abs(m32[f] - one[f][0]) / one[f][0]
            relt, mxt = _errors(orc, got, f64[f].cpu().numpy())
            dt = abs(m32[f] - m64[f]) / m64[f]
            worst = [max(worst[0], rel1), max(worst[1], d1), max(worst[2], relt), max(worst[3], dt)]
            assert rel1 <= REL_F32 and d1 <= MASS_F32, (tick, f, rel1, d1)
            assert relt <= REL_F32 and dt <= tick * MASS_F32, (tick, f, relt, dt)
            if tick in (1, 10):
                print(f"tick {tick} field {f}: one-tick rel-L2 {rel1:.3e} mass {d1:.3e}; "
                      f"trajectory rel-L2 {relt:.3e} max-abs {mxt:.3e} mass {dt:.3e}", flush=True)
        print(f"tick {tick} worst: one-tick rel-L2 {worst[0]:.3e} mass {worst[1]:.3e}; "
              f"trajectory rel-L2 {worst[2]:.3e} mass {worst[3]:.3e}", flush=True)


def test_headline_direct_z_matches_fft_z(vfd, orc, headline):
    """Per field: the separable table's z pass (direct for R <= 16, FFT beyond)
    against the streamed real spectrum, whose z pass is always the FFT form."""
    us, params, radii, _ = headline
    print("case headline_direct_vs_fft_z", flush=True)
    sep = vfd.Plan(HEADLINE, kspec_format=vfd.KSPEC_SEPARABLE)
    real = vfd.Plan(HEADLINE, kspec_format=vfd.KSPEC_REAL)
    for f in range(NFIELDS):
        a, _, _ = _run_plan(vfd, sep, [us[f]], [params[f]])
        b, _, _ = _run_plan(vfd, real, [us[f]], [params[f]])
        rel, mx = _errors(orc, a[0].cpu().numpy(), b[0].cpu().numpy().astype(np.float64))
        print(f"field {f} R={radii[f]}: separable vs real rel-L2 {rel:.3e} max-abs {mx:.3e}", flush=True)
        assert rel <= 1e-6


def _engine_run(eng, shape, us, params):
    import torch
    rid = eng.create_run(shape, [p[0] for p in params], [p[1] for p in params])
    for f, u in enumerate(us):
        eng.write_field(rid, f, torch.from_numpy(u).cuda())
    torch.cuda.synchronize()
    return rid


def _read(eng, rid, nf):
    import torch
    out = [eng.read_field(rid, f) for f in range(nf)]
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def _check_vs_oracle(orc, tag, got, us, params):
    for f in range(len(got)):
        ref = orc.direct_separable(us[f], *params[f])
        rel, mx = _errors(orc, got[f], ref)
        mref = orc.neumaier_sum(ref)
        drift = abs(orc.neumaier_sum(got[f]) - mref) / mref
        print(f"{tag} field {f}: rel-L2 {rel:.3e} max-abs {mx:.3e} mass drift {drift:.3e}", flush=True)
        assert rel <= REL_F32 and drift <= MASS_F32


def test_ensemble_grids_through_engine(vfd, orc):
    """One 8-field run of each ensemble grid, batched in one graph-replayed
    tick, against the oracle and against the same run ticked alone."""
    import vfd.engine as em
    print("case ensemble_grids_through_engine", flush=True)
    data = []
    with em.Engine(64 * GiB) as eng:
        mix, alone = [], []
        for gi, shape in enumerate(ENSEMBLE_GRIDS):
            params = _params(orc, 100 + gi, NFIELDS)
            us = [orc.synth_field(shape, (100 + gi) * 1000 + f, nsrc=50) for f in range(NFIELDS)]
            data.append((shape, us, params))
            mix.append(_engine_run(eng, shape, us, params))
            alone.append(_engine_run(eng, shape, us, params))
        eng.step(mix, graph=True)
        for rid in alone:
            eng.step([rid])
        for gi, (shape, us, params) in enumerate(data):
            got = _read(eng, mix[gi], NFIELDS)
            ref = _read(eng, alone[gi], NFIELDS)
            for f in range(NFIELDS):
                assert np.array_equal(got[f], ref[f]), (shape, f)
            _check_vs_oracle(orc, f"grid {shape}", got, us, params)


def test_served_headline_run_in_mix(vfd, orc, headline):
    """A 1024x512x128 run served among runs of the ensemble grids, with
    several batches in flight: tick 1 against the oracle, and the same bits as
    the same run served alone."""
    import vfd.engine as em
    us, params, _, refs = headline
    print("case served_headline_run_in_mix", flush=True)
    with em.Engine(200 * 10**9) as eng:
        big_mix = _engine_run(eng, HEADLINE, us, params)
        big_alone = _engine_run(eng, HEADLINE, us, params)
        small = []
        for gi, shape in enumerate(ENSEMBLE_GRIDS):
            params_s = _params(orc, 200 + gi, NFIELDS)
            us_s = [orc.synth_field(shape, (200 + gi) * 1000 + f, nsrc=50) for f in range(NFIELDS)]
            small.append(_engine_run(eng, shape, us_s, params_s))
        eng.serve_start(slots=4)
        try:
            tickets = [eng.submit(r) for r in small[:3]] + [eng.submit(big_mix)] + \
                      [eng.submit(r) for r in small[3:]] + [eng.submit(r) for r in small]
            lat = [eng.wait(t) for t in tickets]
            t_alone = eng.submit(big_alone)
            lat.append(eng.wait(t_alone))
        finally:
            eng.serve_stop()
        assert all(x > 0.0 for x in lat)
        print(f"served latencies ms: {[round(x, 2) for x in lat]}", flush=True)
        got = _read(eng, big_mix, NFIELDS)
        ref = _read(eng, big_alone, NFIELDS)
        for f in range(NFIELDS):
            assert np.array_equal(got[f], ref[f]), f
            rel, mx = _errors(orc, got[f], refs[f])
            mref = orc.neumaier_sum(refs[f])
            drift = abs(orc.neumaier_sum(got[f]) - mref) / mref
            print(f"served field {f}: rel-L2 {rel:.3e} max-abs {mx:.3e} mass drift {drift:.3e}", flush=True)
            assert rel <= REL_F32 and drift <= MASS_F32
