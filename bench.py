#!/usr/bin/env python3
"""bench.py -- voxel-updates/s of the 8-field cytokine-diffusion step.

Workload (default): 8 FP32 fields on a 1024x512x128 patch grid (x fastest),
zero-padded linear convolution on the 2048x1024x256 FFT box, per-field
normalized Gaussian Green's kernel (sigma^2 = 2 D dt, D ~ U[0.5,5], dt = 1)
times exp(-lambda dt) (lambda ~ U[0.001,0.05]); lognormal(-2,1) background
plus 50 point sources of amplitude U[1,100].  Inputs are synthetic and
generated on the device; kernel spectra are precomputed and resident.

A step = one tick of all fields (5 HBM passes per field, all fields batched
into each launch; the z pass is four launches: the direct form for window
radius bounds 8, 12, 16 and the FFT form, each acting on its own fields).  W
untimed warmup steps, then K timed steps bracketed by torch.cuda.synchronize().
Per-kernel durations come from HIP events recorded between the launches on the
launch stream over the timed region; roofline.kernel is the single kernel with
the largest average duration, roofline.step_frac the whole step.  copy_gbs is
the library's 16-byte non-temporal streaming copy measured in the same run;
fft_z_step repeats the step with the spectrum streamed as the real 3D array
(the FFT form of the z pass for every field).  The CPU baseline is the
oracle's FP32 OpenMP port timed on a bounded sample (five fields of the same
workload, one tick each) on the host cores.

Other workloads (one JSON line each, same fields where they apply):
  --workload ensemble  64 resident runs of uneven grids drawn from
                       {64x64x32, 128x128x64, 256x128x64, 256x256x64, 512x256x64},
                       8 fields each: batched ticks replayed as a hipGraph vs the
                       plain per-field form (one field at a time, per-pass
                       launches, no graph).
  --workload copy      1 GiB device-to-device copy: measured HBM bandwidth.
  --workload dropin    the reference-compatible drop-in tick: 8 fields in pinned
                       host memory, constant-VF pad, R2C-layout complex spectra,
                       through vfdDiffuseChems (copies overlapped with the
                       diffusion) and through the serial per-chem
                       H2D -> fftDiffuse3D -> D2H loop of the reference.
                       --shape 1006,1390,110 --box 1152,1536,256 is the ABM's
                       production configuration (HUMAN_VF).
  --workload update    the post-diffusion chem update (WHChemical::update) of the
                       8 fields on the device: levels, cleared deltas, 3 gradients.
  --workload precision 8 fields on 1536x768x192, log-uniform background in
                       [1e-6, 1e3] plus 500 sharp sources: the FP32 path against
                       the full-FP64 GPU path over 100 ticks (max-abs, rel-L2,
                       mass drift per field) and both against the FP64 oracle
                       after the first tick.
  --workload serve     Poisson request stream (mean 200/s) over resident runs
                       drawn from that set plus 1024x512x128, admitted against a
                       200 GB HBM budget; continuous batching; throughput and
                       p50/p99 per-request latency.
  --workload admit     serving with run churn: client threads open runs, tick
                       them, close them; a budget (--budget-gb) smaller than the
                       clients' runs makes opening runs wait for memory
                       (vfd_run_create_wait); admission waits beside latency.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vf-version_7_3_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shape", default="128,512,1024", help="nz,ny,nx of the data box")
    ap.add_argument("--box", default=None, help="dropin: FFT box FD,FH,FW (default twice the data box)")
    ap.add_argument("--flows", default="pipelined,serial,resident,tick,split",
                    help="dropin: flows to time (the first one named is the reported value); tick = "
                         "vfdTickChems, the diffusion and the chem update with resident levels; split = "
                         "vfdTickChemsEnd alone, the diffusion having run during the agents' step")
    ap.add_argument("--fields", type=int, default=8)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fft-z", action="store_true",
                    help="step: skip the second measurement with the real (FFT z pass) spectrum")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default (or VFD_SCHEDULE) schedule instead of autotuning")
    ap.add_argument("--kspec", default="separable", choices=["separable", "real"],
                    help="kernel-spectrum format: FP64 1D tables formed in the z pass, or the 3D real array")
    ap.add_argument("--workload", default="step",
                    choices=["step", "ensemble", "serve", "admit", "copy", "precision", "update", "dropin"])
    ap.add_argument("--ticks", type=int, default=100, help="precision: ticks of FP32 and FP64 side by side")
    ap.add_argument("--runs", type=int, default=64, help="ensemble/serve: resident runs")
    ap.add_argument("--grids", default=None,
                    help="ensemble: indices into the grid set to draw runs from (default: all five)")
    ap.add_argument("--forms", default="graph,eager,plain,reference",
                    help="ensemble: tick forms to time (graph, eager, plain, reference); one form alone is "
                         "how each is profiled")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: mean requests/s")
    ap.add_argument("--duration", type=float, default=10.0, help="serve: seconds of arrivals")
    ap.add_argument("--budget-gb", type=float, default=None,
                    help="serve/admit: HBM budget in GB (default 200 for serve, 8 for admit)")
    ap.add_argument("--slots", type=int, default=4, help="serve: batches in flight")
    ap.add_argument("--clients", type=int, default=16, help="admit: client threads, each opening runs in turn")
    ap.add_argument("--ticks-per-run", type=int, default=10, help="admit: ticks each run is served before it closes")
    return ap.parse_args()


ENSEMBLE_GRIDS = [(32, 64, 64), (64, 128, 128), (64, 128, 256), (64, 256, 256), (64, 256, 512)]  # nz,ny,nx
SERVE_GRIDS = ENSEMBLE_GRIDS + [(128, 512, 1024)]


def _fill_run(eng, rid, shape, nf, seed):
    import torch
    fields, _ = synth_fields(shape, nf, seed, "cuda:0")
    for f, t in enumerate(fields):
        eng.write_field(rid, f, t)
    torch.cuda.synchronize()


def _run_params(rng, nf):
    import numpy as np
    D = rng.uniform(0.5, 5.0, nf)
    lam = rng.uniform(0.001, 0.05, nf)
    return list(2.0 * D), list(np.exp(-lam))


def _header(args, metric, value, unit, ms_per_step, config, dtype="f32"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (device-generated lognormal background + 50 point sources; "
                    "random D, lambda per field)", "config": config}


def bench_copy(args):
    """Device-to-device copy of 1 GiB with the library's streaming kernel
    (16-byte non-temporal loads and stores, one per lane): the measured HBM
    bandwidth that sits beside the 8 TB/s spec peak, and the known-size kernel
    the HBM counters (FETCH_SIZE, WRITE_SIZE) are calibrated on.  torch's
    copy_ of the same bytes is reported beside it."""
    import torch
    import vfd
    n = (1 << 30) // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(args.steps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / args.steps

    ms = timed(lambda: vfd.stream_copy(dst, src, stream))
    ms_torch = timed(lambda: dst.copy_(src))
    gbs = 2.0 * n * 4 / (ms * 1e-3) / 1e9
    out = _header(args, "HBM copy bandwidth (1 GiB device-to-device, read + write bytes)", gbs, "GB/s", ms,
                  {"workload": "vfd_stream_copy of 1 GiB (16-byte non-temporal loads/stores, one per lane)"})
    out["frac_of_spec"] = gbs / HBM_PEAK_GBS
    out["torch_copy_gbs"] = 2.0 * n * 4 / (ms_torch * 1e-3) / 1e9
    out["mix_gbs"] = mix_bandwidth(vfd, torch, stream)
    print(json.dumps(out), flush=True)


def reference_flow_tick(picks, nf, seed, ticks=2):
    """One tick of every field of every run the way the reference's
    diffuseChemGPU drives the GPU: host->device, each pass followed by a
    device synchronization, device->host, one field at a time."""
    import numpy as np
    import torch
    import vfd
    rng = np.random.default_rng(seed + 1)
    plans, work, dev, spec = {}, {}, {}, {}
    host = []
    for g in sorted(set(int(x) for x in picks)):
        shape = ENSEMBLE_GRIDS[g]
        plans[g] = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
        work[g] = plans[g].new_work()
        dev[g] = torch.empty(shape, dtype=torch.float32, device="cuda:0")   # one device field per shape
        s2, dc = _run_params(rng, 1)
        spec[g] = plans[g].gauss_spectrum(s2[0], dc[0])
    for g in picks:
        host.append([torch.rand(ENSEMBLE_GRIDS[g], dtype=torch.float32).pin_memory() for _ in range(nf)])
    torch.cuda.synchronize()

    def tick():
        for i, g in enumerate(picks):
            g = int(g)
            ent = plans[g].field_array([{"in": dev[g], "out": dev[g], "work": work[g], "kspec": spec[g]}])
            for f in range(nf):
                dev[g].copy_(host[i][f])
                torch.cuda.synchronize()
                for p in range(5):
                    plans[g].diffuse_pass(ent, p)
                    torch.cuda.synchronize()
                host[i][f].copy_(dev[g])
                torch.cuda.synchronize()

    tick()
    a = time.perf_counter()
    for _ in range(ticks):
        tick()
    return (time.perf_counter() - a) / ticks


def precision_fields(shape, nfields, seed, device):
    """Log-uniform background in [1e-6, 1e3] plus 500 sharp sources per field."""
    import numpy as np
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    n = shape[0] * shape[1] * shape[2]
    out = []
    for _ in range(nfields):
        f = torch.rand(shape, generator=g, device=device, dtype=torch.float64)
        f = torch.pow(10.0, f * 9.0 - 6.0)
        idx = torch.randint(0, n, (500,), generator=g, device=device)
        amp = torch.rand(500, generator=g, device=device, dtype=torch.float64) * 500.0 + 500.0
        f.view(-1).index_add_(0, idx, amp)
        out.append(f.float())   # the FP32 field both paths start from
    return out


def bench_precision(args):
    """FP32 vs full-FP64 GPU path over many ticks; the FP64 oracle after tick 1."""
    import numpy as np
    import torch
    import vfd
    shape = (192, 768, 1536) if args.shape == "128,512,1024" else tuple(int(s) for s in args.shape.split(","))
    nf = args.fields
    rng = np.random.default_rng(args.seed)
    s2, dc = _run_params(rng, nf)
    p32 = vfd.Plan(shape, precision=vfd.F32, kspec_format=vfd.KSPEC_SEPARABLE)
    p64 = vfd.Plan(shape, precision=vfd.F64, kspec_format=vfd.KSPEC_SEPARABLE)
    u0 = precision_fields(shape, nf, args.seed, "cuda:0")
    f32 = [u.clone() for u in u0]
    f64 = [u.double() for u in u0]
    k32 = [p32.gauss_spectrum(a, b) for a, b in zip(s2, dc)]
    k64 = [p64.gauss_spectrum(a, b) for a, b in zip(s2, dc)]
    w32, w64 = p32.new_work(), p64.new_work()   # fields run one at a time: one work buffer each
    arr32 = [p32.field_array([{"in": f32[i], "out": f32[i], "work": w32, "kspec": k32[i]}]) for i in range(nf)]
    arr64 = [p64.field_array([{"in": f64[i], "out": f64[i], "work": w64, "kspec": k64[i]}]) for i in range(nf)]
    report = {}

    def compare(tick):
        m32, m64 = p32.mass(f32), p64.mass(f64)
        rows = []
        for i in range(nf):
            a, b = f32[i].double(), f64[i]
            err = (a - b).abs()
            rows.append({"field": i, "max_abs": float(err.max()), "rel_l2": float(torch.linalg.vector_norm(a - b) /
                                                                                torch.linalg.vector_norm(b)),
                         "mass_drift_fp32_vs_fp64": abs(m32[i] - m64[i]) / abs(m64[i])})
        report[str(tick)] = rows

    t0 = time.perf_counter()
    oracle = {}
    for tick in range(1, args.ticks + 1):
        for i in range(nf):
            p32.diffuse(arr32[i])
            p64.diffuse(arr64[i])
        if tick == 1:
            torch.cuda.synchronize()
            # FP64 oracle (separable direct convolution) on field 0 after one tick
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import orc
            ref = orc.direct_separable(u0[0].cpu().numpy(), s2[0], dc[0])
            oracle = {"field": 0, "rel_l2_fp64_gpu": orc.rel_l2(f64[0].cpu().numpy(), ref),
                      "rel_l2_fp32_gpu": orc.rel_l2(f32[0].cpu().numpy(), ref)}
        if tick in (1, 10, args.ticks):
            torch.cuda.synchronize()
            compare(tick)
        print(f"precision: tick {tick} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    last = report[str(args.ticks)]
    worst = max(r["rel_l2"] for r in last)
    out = _header(args, f"precision stress: worst per-field rel-L2 of FP32 vs the FP64 GPU path after {args.ticks} ticks",
                  worst, "rel-L2", wall * 1e3 / args.ticks,
                  {"workload": f"{nf} fields, {shape[2]}x{shape[1]}x{shape[0]} grid, log-uniform background in "
                               "[1e-6, 1e3] plus 500 sharp sources of amplitude U[500, 1000], zero-padded linear "
                               "convolution, Gaussian kernels as the tissue-scale workload",
                   "ticks": args.ticks}, dtype="f32 vs f64")
    out["higher_is_better"] = False
    out["per_tick"] = report
    out["oracle_tick1"] = oracle
    out["max_rel_l2"] = worst
    out["max_mass_drift"] = max(r["mass_drift_fp32_vs_fp64"] for r in last)
    # FP32 twiddle rounding gives every frequency a small systematic gain, so
    # the drift against FP64 grows linearly with ticks: report it per tick too
    out["max_mass_drift_per_tick"] = out["max_mass_drift"] / args.ticks
    out["max_rel_l2_per_tick"] = worst / args.ticks
    print(json.dumps(out), flush=True)


def bench_update(args):
    """vfd_chem_update over the 8 fields of the tissue-scale grid."""
    import torch
    import vfd
    shape = tuple(int(s) for s in args.shape.split(","))
    n = shape[0] * shape[1] * shape[2]
    g = torch.Generator(device="cuda:0")
    g.manual_seed(args.seed)
    t = [torch.rand(n, generator=g, device="cuda:0") for _ in range(8)]
    d = [torch.rand(n, generator=g, device="cuda:0") * 0.01 for _ in range(8)]
    fc, fe = torch.rand(n, generator=g, device="cuda:0"), torch.rand(n, generator=g, device="cuda:0")
    grads = [torch.empty(n, device="cuda:0") for _ in range(3)]
    decay = [0.319, 0.00003052, 0.67, 0.2, 0.2, 0.7937, 0.9284, 0.2]
    for _ in range(args.warmup):
        vfd.chem_update(t, d, t, decay, fc, fe, grads, totals=False)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        vfd.chem_update(t, d, t, decay, fc, fe, grads, totals=False)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    t0 = time.perf_counter()
    vfd.chem_update(t, d, t, decay, fc, fe, grads, totals=True)
    with_totals_ms = (time.perf_counter() - t0) * 1e3
    bytes_per_call = 148.0 * n       # reads 8 t + 8 d + 2 fragments, writes 8 p + 8 d + 3 gradients (FP32)
    gbs = bytes_per_call / (ms * 1e-3) / 1e9
    out = _header(args, "voxel-updates/s of the post-diffusion chem update (8 chems, levels + gradients + totals)",
                  n / (ms * 1e-3), "voxel-updates/s", ms,
                  {"workload": f"WHChemical::update (gpuUpdate) of 8 FP32 fields on {shape[2]}x{shape[1]}x{shape[0]}, "
                               "in place (p aliases t)"})
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", f"hbm_traffic_chem_update_{shape[2]}x{shape[1]}x{shape[0]}.json")) as f:
            pb = json.load(f)["per_launch_bytes"]["chem_update"]
            traffic = float(pb["read"] + pb["write"])
    except (OSError, ValueError, KeyError):
        pass
    out["roofline"] = {"bound": "hbm", "kernel": "chem_update_kernel", "achieved": gbs, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
                       "algorithmic_bytes_per_call": bytes_per_call}
    out["with_totals_host_ms"] = with_totals_ms
    # the whole tick: diffusion then the update, against the update fused into
    # the x-inverse pass (vfd_diffuse_update), both without the totals download
    plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_SEPARABLE)
    u = [x.view(shape) for x in t]
    tt = [torch.empty_like(x) for x in u]
    ks = [plan.gauss_spectrum(0.3 + 0.1 * i, 0.99) for i in range(8)]
    works = [plan.new_work() for _ in range(8)]
    split = plan.field_array([{"in": u[i], "out": tt[i], "work": works[i], "kspec": ks[i]} for i in range(8)])
    fused = plan.field_array([{"in": u[i], "out": u[i], "work": works[i], "kspec": ks[i]} for i in range(8)])
    plan.autotune(8)
    dv = [x.view(shape) for x in d]

    def tick_split():
        plan.diffuse(split)
        vfd.chem_update(tt, d, t, decay, fc, fe, grads, totals=False)

    def tick_fused():
        plan.diffuse_update(fused, dv, decay, fc, fe, grads, totals=False)

    for name, fn in (("tick_split_ms", tick_split), ("tick_fused_ms", tick_fused)):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        a.record()
        for _ in range(args.steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        out[name] = a.elapsed_time(b) / args.steps
    print(json.dumps(out), flush=True)


def bench_dropin(args):
    """The reference's per-tick flow through the reference-compatible C interface."""
    import ctypes
    import numpy as np
    import torch
    import vfd
    from vfd import legacy
    nz, ny, nx = (int(s) for s in args.shape.split(","))
    nf = 8
    FD, FH, FW = (int(s) for s in args.box.split(",")) if args.box else (2 * nz, 2 * ny, 2 * nx)
    cc = legacy.CCtx()
    cc.DD, cc.DH, cc.DW = nz, ny, nx
    cc.FFTD, cc.FFTH, cc.FFTW = FD, FH, FW
    rng = np.random.default_rng(args.seed)
    host = [torch.rand((nz, ny, nx), dtype=torch.float32).pin_memory() for _ in range(nf)]
    dev = [torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda:0") for _ in range(nf)]
    spec = []
    for ic in range(nf):
        # a Gaussian spectrum in the reference's complex R2C layout (values do not affect timing)
        s2 = 2.0 * rng.uniform(0.5, 5.0)
        k = [torch.exp(-2.0 * (np.pi ** 2) * s2 * (torch.fft.fftfreq(n, device="cuda:0") ** 2)) for n in (FD, FH)]
        kx = torch.exp(-2.0 * (np.pi ** 2) * s2 * (torch.fft.rfftfreq(FW, device="cuda:0") ** 2))
        sp = (k[0][:, None, None] * k[1][None, :, None] * kx[None, None, :]).to(torch.complex64).contiguous()
        spec.append(sp)
        cc.h_ibuffs[ic] = host[ic].data_ptr()
        cc.h_obuffs[ic] = host[ic].data_ptr()
        cc.d_data[ic] = dev[ic].data_ptr()
        cc.d_kernelspectrum_h[ic] = sp.data_ptr()
        cc.dev_id[ic] = 0
    base = (ctypes.c_float * nf)(*[0.01 * (ic + 1) for ic in range(nf)])
    epi = nx - 4
    torch.cuda.synchronize()

    def pipelined():
        if not legacy.lib.vfdDiffuseChems(cc, nf, epi, base):
            raise vfd.VfdError(-1, vfd.last_error())

    def serial():
        for ic in range(nf):
            dev[ic].copy_(host[ic])
            torch.cuda.synchronize()
            if not legacy.lib.fftDiffuse3D(dev[ic].data_ptr(), spec[ic].data_ptr(), 0, 0, cc, epi,
                                           ctypes.c_float(base[ic]), ic, 0):
                raise vfd.VfdError(-1, vfd.last_error())
            host[ic].copy_(dev[ic])
            torch.cuda.synchronize()

    # the resident tick (vfdTickChems): levels stay in d_data, deltas in, levels
    # and gradients out, the chem update fused into the diffusion
    tick_state = {}

    def tick():
        if not tick_state:
            tick_state["d"] = [torch.zeros((nz, ny, nx), dtype=torch.float32).pin_memory() for _ in range(nf)]
            tick_state["frag"] = [torch.rand((nz, ny, nx), dtype=torch.float32).pin_memory() for _ in range(2)]
            tick_state["g"] = [torch.empty((nz, ny, nx), dtype=torch.float32).pin_memory() for _ in range(3)]
            tick_state["dv"] = (ctypes.c_void_p * nf)(*[x.data_ptr() for x in tick_state["d"]])
            tick_state["gv"] = (ctypes.c_void_p * 3)(*[x.data_ptr() for x in tick_state["g"]])
            tick_state["decay"] = (ctypes.c_float * nf)(0.319, 0.00003052, 0.67, 0.2, 0.2, 0.7937, 0.9284, 0.2)
            tick_state["tot"] = (ctypes.c_double * nf)()
        st = tick_state
        if not legacy.lib.vfdTickChems(cc, epi, base, st["dv"], st["decay"], st["frag"][0].data_ptr(),
                                       st["frag"][1].data_ptr(), st["gv"], st["tot"]):
            raise vfd.VfdError(-1, vfd.last_error())

    def split_end():
        # vfdTickChemsBegin, the device left to finish the diffusion (the agents'
        # step on the host), then vfdTickChemsEnd alone is timed
        tick()  # allocates the state on first use
        st = tick_state
        if not legacy.lib.vfdTickChemsBegin(cc, epi, base):
            raise vfd.VfdError(-1, vfd.last_error())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if not legacy.lib.vfdTickChemsEnd(cc, st["dv"], st["decay"], st["frag"][0].data_ptr(),
                                          st["frag"][1].data_ptr(), st["gv"], st["tot"]):
            raise vfd.VfdError(-1, vfd.last_error())
        split_times.append(time.perf_counter() - t0)

    split_times = []

    def resident():
        for ic in range(nf):
            if not legacy.lib.fftDiffuse3D(dev[ic].data_ptr(), spec[ic].data_ptr(), 0, 0, cc, epi,
                                           ctypes.c_float(base[ic]), ic, 0):
                raise vfd.VfdError(-1, vfd.last_error())

    def timed(fn, reps):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    K = max(1, args.steps // 10)
    flows = {"pipelined": ("vfdDiffuseChems", pipelined), "serial": ("serial_h2d_fftDiffuse3D_d2h", serial),
             "resident": ("device_resident_fftDiffuse3D", resident),
             "tick": ("vfdTickChems_diffusion_and_update", tick),
             "split": ("vfdTickChemsEnd_after_agents", split_end)}
    names = [f.strip() for f in args.flows.split(",") if f.strip()]
    if not names or any(n not in flows for n in names):
        raise SystemExit(f"--flows: choose from {','.join(flows)}")
    t = {n: timed(flows[n][1], K) for n in names}
    if "split" in t:  # only the End call of each timed split tick
        t["split"] = sum(split_times[-K:]) / K
    vox = float(nx * ny * nz) * nf
    first = t[names[0]]
    metric = ("voxel-updates/s of the drop-in tick through the reference C interface"
              + (" (PCIe-inclusive)" if names[0] != "resident" else " (fields resident on the device)"))
    out = _header(args, metric, vox / first, "voxel-updates/s", first * 1e3,
                  {"workload": f"8 FP32 fields {nx}x{ny}x{nz} in pinned host memory, constant-VF pad (epi {epi}), "
                               f"complex R2C-layout spectra on the {FW}x{FH}x{FD} box, one tick"})
    for n in names:
        out[flows[n][0]] = {"ms_per_tick": t[n] * 1e3, "voxel_updates_s": vox / t[n]}
    out["pcie_bytes_per_tick"] = 2.0 * vox * 4
    # the plan fftDiffuse3D builds for this box: its pass order and algorithmic
    # HBM bytes (every pass streams its arrays once)
    plan = vfd.Plan((nz, ny, nx), (FD, FH, FW), vfd.F32, vfd.PAD_CONSTANT_VF, vfd.KSPEC_COMPLEX_STD)
    hbm = nf * plan.info.bytes_per_field_tick
    out["pass_order"] = "x-last" if plan.info.order == vfd.ORDER_XLAST else "x-first"
    out["hbm_bytes_per_tick"] = hbm
    if "resident" in t:
        out["device_resident_fftDiffuse3D"]["hbm_gbs"] = hbm / t["resident"] / 1e9
        out["device_resident_fftDiffuse3D"]["hbm_frac"] = hbm / t["resident"] / 1e9 / HBM_PEAK_GBS
    if "pipelined" in t and "serial" in t:
        out["speedup_vs_serial"] = t["serial"] / t["pipelined"]
    print(json.dumps(out), flush=True)


def bench_ensemble(args):
    """Batched (graph) vs plain per-field ticks over uneven resident runs."""
    import numpy as np
    import torch
    import vfd.engine as ve
    rng = np.random.default_rng(args.seed)
    nf = args.fields
    allowed = [int(g) for g in args.grids.split(",")] if args.grids else list(range(len(ENSEMBLE_GRIDS)))
    picks = np.array(allowed)[rng.integers(0, len(allowed), args.runs)]
    eng = ve.Engine(int(250e9))
    runs, vox = [], 0.0
    t0 = time.perf_counter()
    for i, g in enumerate(picks):
        shape = ENSEMBLE_GRIDS[g]
        s2, dc = _run_params(rng, nf)
        rid = eng.create_run(shape, s2, dc)
        _fill_run(eng, rid, shape, nf, args.seed + 7919 * i)
        runs.append(rid)
        vox += float(np.prod(shape)) * nf
    setup_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream()

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / steps

    K = args.steps
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    if not forms or any(f not in ("graph", "eager", "plain", "reference") for f in forms):
        raise SystemExit("--forms: choose from graph,eager,plain,reference")
    t = {}
    if "graph" in forms:
        t["graph"] = timed(lambda: eng.step(runs, graph=True, stream=stream), K)
    if "eager" in forms:
        t["eager"] = timed(lambda: eng.step(runs, graph=False, stream=stream), K)
    if "plain" in forms:
        t["plain"] = timed(lambda: eng.step_plain(runs, stream=stream), max(1, K // 2))
    if "reference" in forms:
        t["reference"] = reference_flow_tick(picks, nf, args.seed)
    counts = {f"{s[2]}x{s[1]}x{s[0]}": int((picks == i).sum()) for i, s in enumerate(ENSEMBLE_GRIDS)}
    bytes_tick = 0.0
    for i, s in enumerate(ENSEMBLE_GRIDS):
        import vfd
        p = vfd.Plan(s, kspec_format=vfd.KSPEC_SEPARABLE)
        bytes_tick += p.info.bytes_per_field_tick * nf * counts[f"{s[2]}x{s[1]}x{s[0]}"]
    head = forms[0]
    out = _header(args, f"voxel-updates/s, ensemble tick (64 uneven runs x 8 fields), {head} form", vox / t[head],
                  "voxel-updates/s", t[head] * 1e3,
                  {"workload": f"{args.runs} resident runs x {nf} FP32 fields, grids drawn from "
                               "{64x64x32,128x128x64,256x128x64,256x256x64,512x256x64}, one tick per step",
                   "runs_per_grid": counts, "voxels_per_tick": vox, "forms": forms})
    names = {"graph": "batched_graph", "eager": "batched_eager", "plain": "plain_per_field",
             "reference": "reference_flow"}
    for f, tt in t.items():
        out[names[f]] = {"ms_per_tick": tt * 1e3, "voxel_updates_s": vox / tt, "algorithmic_gbs": bytes_tick / tt / 1e9}
    if "reference" in t:
        out["reference_flow"]["what"] = ("the reference's per-field drop-in flow on these kernels: pinned host field "
                                         "-> device, 5 pass launches each followed by a device synchronization, "
                                         "device -> host, one field at a time (Diffuser.cpp:300-421, "
                                         "Diffusion.cu:1802-1870)")
    if "graph" in t and "plain" in t:
        out["speedup_vs_plain"] = t["plain"] / t["graph"]
    if "graph" in t and "reference" in t:
        out["speedup_vs_reference_flow"] = t["reference"] / t["graph"]
    # every pass streams its arrays once: the bytes of a tick and the time they
    # take at the measured streaming bandwidth (16-byte copy kernel)
    out["algorithmic_bytes_per_tick"] = bytes_tick
    out["algorithmic_gbs"] = bytes_tick / t[head] / 1e9
    copy = copy_bandwidth(vfd, torch, stream)
    out["copy_gbs"] = copy
    out["hbm_floor_ms"] = bytes_tick / (copy * 1e9) * 1e3
    if "plain" in t:
        out["max_speedup_vs_plain_at_floor"] = t["plain"] / (bytes_tick / (copy * 1e9))
    out["setup_s"] = setup_s
    out["engine"] = eng.stats()
    eng.close()
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Poisson request stream against resident runs; continuous batching."""
    import threading
    import numpy as np
    import vfd
    import vfd.engine as ve
    rng = np.random.default_rng(args.seed)
    nf = args.fields
    args.budget_gb = args.budget_gb or 200.0
    budget = int(args.budget_gb * 1e9)
    eng = ve.Engine(budget)
    runs, vox_of, refused = [], {}, 0
    for i in range(args.runs):
        shape = SERVE_GRIDS[int(rng.integers(0, len(SERVE_GRIDS)))]
        s2, dc = _run_params(rng, nf)
        try:
            rid = eng.create_run(shape, s2, dc)
        except vfd.VfdError as e:
            if e.code != vfd.VFD_ERR_BUDGET:
                raise
            refused += 1
            continue
        _fill_run(eng, rid, shape, nf, args.seed + 104729 * i)
        runs.append(rid)
        vox_of[rid] = float(np.prod(shape)) * nf
    eng.serve_start(slots=args.slots, max_batch_fields=256)
    # warmup: one request per run (captures first-use costs)
    for t in [eng.submit(r) for r in runs]:
        eng.wait(t)
    n_req = max(1, int(args.rate * args.duration))
    gaps = rng.exponential(1.0 / args.rate, n_req)
    which = rng.integers(0, len(runs), n_req)
    tickets = [None] * n_req
    lat = [0.0] * n_req
    ready = threading.Semaphore(0)

    def collect():
        for k in range(n_req):
            ready.acquire()
            lat[k] = eng.wait(tickets[k])

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    due = t0
    for k in range(n_req):
        due += gaps[k]
        d = due - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets[k] = eng.submit(runs[which[k]])
        ready.release()
    th.join()
    wall = time.perf_counter() - t0
    st = eng.stats()
    eng.serve_stop()
    vox = sum(vox_of[runs[w]] for w in which)
    out = _header(args, "voxel-updates/s served (Poisson 200/s, continuous batching); p50/p99 latency",
                  vox / wall, "voxel-updates/s", wall * 1e3 / n_req,
                  {"workload": f"Poisson stream mean {args.rate}/s for {args.duration}s over resident runs of "
                               f"{nf} FP32 fields, grids from the ensemble set plus 1024x512x128, "
                               f"HBM budget {args.budget_gb} GB",
                   "runs_admitted": len(runs), "runs_refused": refused, "slots": args.slots})
    out["requests"] = n_req
    out["requests_per_s"] = n_req / wall
    out["latency_ms"] = {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                         "mean": float(np.mean(lat)), "max": float(np.max(lat))}
    out["engine"] = st
    eng.close()
    print(json.dumps(out), flush=True)


def bench_admit(args):
    """Serving with run churn: client threads each open a run (a grid drawn from
    the serving set), tick it --ticks-per-run times through the server, close
    it and open the next; the HBM budget holds only part of the clients' runs,
    so an opening run waits in vfd_run_create_wait until closed runs free
    memory.  Reports served throughput, per-tick p50/p99 latency and the
    admission waits."""
    import threading
    import numpy as np
    import vfd
    import vfd.engine as ve
    nf = args.fields
    args.budget_gb = args.budget_gb or 8.0
    budget = int(args.budget_gb * 1e9)
    eng = ve.Engine(budget)
    eng.serve_start(slots=args.slots, max_batch_fields=256)
    # every grid once, so that first-use work (plans, autotuning) is not timed
    for i, shape in enumerate(SERVE_GRIDS):
        rid = eng.create_run_wait(shape, [2.0] * nf, [0.99] * nf, 600.0)
        eng.wait(eng.submit(rid))
        eng.destroy_run(rid)
    lock = threading.Lock()
    lat, waits, vox = [], [], [0.0]
    errors = []
    per_client = max(1, args.runs // args.clients)

    def client(c):
        rng = np.random.default_rng(args.seed + 7919 * c)
        try:
            for r in range(per_client):
                shape = SERVE_GRIDS[int(rng.integers(0, len(SERVE_GRIDS)))]
                s2, dc = _run_params(rng, nf)
                t0 = time.perf_counter()
                rid = eng.create_run_wait(shape, s2, dc, 600.0)
                w = time.perf_counter() - t0
                _fill_run(eng, rid, shape, nf, args.seed + 104729 * (c * per_client + r))
                mine = []
                for _ in range(args.ticks_per_run):
                    mine.append(eng.wait(eng.submit(rid)))
                eng.destroy_run(rid)
                with lock:
                    lat.extend(mine)
                    waits.append(w * 1e3)
                    vox[0] += float(np.prod(shape)) * nf * args.ticks_per_run
        except Exception as e:  # reported after the join
            errors.append(repr(e))

    ths = [threading.Thread(target=client, args=(c,)) for c in range(args.clients)]
    t0 = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    wall = time.perf_counter() - t0
    st = eng.stats()
    eng.serve_stop()
    eng.close()
    if errors:
        raise SystemExit(f"admit: client failed: {errors[0]}")
    lat, waits = np.array(lat), np.array(waits)
    out = _header(args, "voxel-updates/s served with run churn (admission waits for HBM); p50/p99 tick latency",
                  vox[0] / wall, "voxel-updates/s", wall * 1e3 / max(1, len(lat)),
                  {"workload": f"{args.clients} clients x {per_client} runs each, {args.ticks_per_run} ticks per run, "
                               f"{nf} FP32 fields, grids from the ensemble set plus 1024x512x128, "
                               f"HBM budget {args.budget_gb} GB", "slots": args.slots})
    out["runs"] = int(len(waits))
    out["ticks"] = int(len(lat))
    out["latency_ms"] = {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                         "max": float(lat.max())}
    out["admission_wait_ms"] = {"p50": float(np.percentile(waits, 50)), "p99": float(np.percentile(waits, 99)),
                                "max": float(waits.max())}
    out["engine"] = st
    print(json.dumps(out), flush=True)


def measured_traffic(shape, nfields, kspec, pass_name):
    """HBM bytes per launch of one pass, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs of this workload (scaled by the copy-kernel calibration);
    None when no measurement of this workload is committed."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{shape[2]}x{shape[1]}x{shape[0]}_{nfields}f.json")
    try:
        with open(path) as f:
            m = json.load(f)
    except (OSError, ValueError):
        return None
    if m.get("kspec") != kspec or m.get("fields") != nfields or list(m.get("shape", [])) != list(shape):
        return None
    b = m.get("per_launch_bytes", {}).get(pass_name)
    return None if b is None else float(b["read"] + b["write"])


def copy_bandwidth(vfd, torch, stream, nbytes=1 << 30, reps=7):
    """GB/s of the library's 16-byte non-temporal streaming copy (read + write
    bytes over time), median of reps, measured on `stream` with HIP events."""
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    vfd.stream_copy(b, a, stream)
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        vfd.stream_copy(b, a, stream)
        e1.record(stream)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    ts.sort()
    del a, b
    return 2.0 * nbytes / (ts[len(ts) // 2] * 1e-3) / 1e9


def mix_bandwidth(vfd, torch, stream, nbytes=1 << 30, reps=7):
    """GB/s (all bytes moved over time, median of reps) of the library's
    streaming kernel for three read:write mixes: r1w1 (the copy), r1w2 (one
    buffer read, two written: the mix of the x and y forward passes) and r2w1
    (their inverses)."""
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    b, c = torch.empty_like(a), torch.empty_like(a)
    a.fill_(1)
    b.fill_(2)
    forms = {"r1w1": (lambda: vfd.stream_mix(c, a, stream=stream), 2),
             "r1w2": (lambda: vfd.stream_mix(b, a, dst1=c, stream=stream), 3),
             "r2w1": (lambda: vfd.stream_mix(c, a, src1=b, stream=stream), 3)}
    out = {}
    for name, (fn, nbuf) in forms.items():
        fn()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        ts.sort()
        out[name] = nbuf * nbytes / (ts[len(ts) // 2] * 1e-3) / 1e9
    del a, b, c
    return out


# read:write mix of each launch of the step (x-first order)
LAUNCH_MIX = {"xfwd": "r1w2", "yfwd": "r1w2", "z_direct8": "r1w1", "z_direct12": "r1w1", "z_direct16": "r1w1",
              "z_fft": "r1w1", "yinv": "r2w1", "xinv": "r2w1"}


def fft_z_step(vfd, torch, np, shape, fields, params, model_bytes, stream, warmup, steps):
    """The same step with the spectrum streamed as the 3D real array: every
    field takes the FFT form of the z pass (the form a non-separable kernel,
    such as the ABM's stencil heat kernel, needs).  Fields are copies."""
    plan = vfd.Plan(shape, kspec_format=vfd.KSPEC_REAL)
    fs = [f.clone() for f in fields]
    works = [plan.new_work() for _ in fs]
    ks = [plan.gauss_spectrum(s2, dec, stream=stream) for (s2, dec) in params]
    arr = plan.field_array([{"in": f, "out": f, "work": w, "kspec": k} for f, w, k in zip(fs, works, ks)])
    schedule = plan.autotune(len(fs), stream)[0]
    for _ in range(warmup):
        plan.diffuse(arr, stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        plan.diffuse(arr, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nvox = float(np.prod(shape)) * len(fs)
    out = {"kernel_spectrum": "real", "steps": steps, "ms_per_step": ms, "voxel_updates_s": nvox / (ms * 1e-3),
           "bytes_per_step": model_bytes, "frac": model_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "schedule": dict(zip(vfd.PASS_NAMES, schedule))}
    del arr, fs, works, ks
    torch.cuda.empty_cache()
    return out


def synth_fields(shape, nfields, seed, device):
    import numpy as np
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    rng = np.random.default_rng(seed)
    n = shape[0] * shape[1] * shape[2]
    fields, params = [], []
    for _ in range(nfields):
        f = torch.randn(shape, generator=g, device=device, dtype=torch.float32)
        f.mul_(1.0).add_(-2.0).exp_()
        idx = torch.randint(0, n, (50,), generator=g, device=device)
        amp = torch.rand(50, generator=g, device=device) * 99.0 + 1.0
        f.view(-1).index_add_(0, idx, amp)
        fields.append(f)
        D = rng.uniform(0.5, 5.0)
        lam = rng.uniform(0.001, 0.05)
        params.append((2.0 * D * 1.0, float(np.exp(-lam * 1.0))))
    return fields, params


def cpu_baseline(shape, seed, nfields=5):
    """Oracle FP32 OpenMP port on the first fields of the same workload, one
    tick each (about 11 s of host work at 1024x512x128 on 16 threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import orc
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    box = tuple(2 * s for s in shape)
    dt = 0.0
    for f in range(nfields):
        u = orc.synth_field(shape, seed + f, nsrc=50)
        D, lam = orc.synth_params(seed, f)
        ks = orc.gauss_spectrum_real(2.0 * D, float(np.exp(-lam)), box, 1.0 / float(np.prod(box)))
        t0 = time.perf_counter()
        orc.cpu_diffuse_f32(u, ks, nthreads)
        dt += time.perf_counter() - t0
    nvox = float(np.prod(shape)) * nfields
    # the CPU-runnable reference case: 8 fields on 64x64x32, one tick, median of 5
    small = (32, 64, 64)
    sbox = tuple(2 * s for s in small)
    us, ks = [], []
    for f in range(8):
        us.append(orc.synth_field(small, seed + 100 + f, nsrc=50))
        D, lam = orc.synth_params(seed + 100, f)
        ks.append(orc.gauss_spectrum_real(2.0 * D, float(np.exp(-lam)), sbox, 1.0 / float(np.prod(sbox))))
    reps = []
    for _ in range(5):
        t0 = time.perf_counter()
        for u, k in zip(us, ks):
            orc.cpu_diffuse_f32(u, k, nthreads)
        reps.append(time.perf_counter() - t0)
    ref_s = float(np.median(reps))
    return {"value": nvox / dt, "unit": "voxel-updates/s", "cores": nthreads, "kind": "port",
            "sample": f"{nfields} fields x 1 tick of the {shape[2]}x{shape[1]}x{shape[0]} workload "
                      f"({dt:.2f} s of host compute; oracle FP32 OpenMP port of the same pruned 5-pass "
                      f"algorithm, {nthreads} threads)",
            "reference_case": {"workload": "8 fields x 1 tick on 64x64x32", "median_s": ref_s,
                               "voxel_updates_s": float(np.prod(small)) * 8 / ref_s, "reps": 5}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs on one GPU "
              "(fields shard across GPUs with no exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import vfd

    torch.cuda.set_device(0)
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "admit":
        return bench_admit(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "update":
        return bench_update(args)
    if args.workload == "dropin":
        return bench_dropin(args)
    shape = tuple(int(s) for s in args.shape.split(","))
    nz, ny, nx = shape
    dev = "cuda:0"
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    kfmt = vfd.KSPEC_SEPARABLE if args.kspec == "separable" else vfd.KSPEC_REAL
    plan = vfd.Plan(shape, kspec_format=kfmt)
    # the pruned 5-pass byte model with a streamed real K^ array (8N + 32 A1 + 32 A2 + 4 A3)
    model_bytes = vfd.Plan(shape).info.bytes_per_field_tick * args.fields
    fields, params = synth_fields(shape, args.fields, args.seed, dev)
    works = [plan.new_work(dev) for _ in fields]
    kspecs = [plan.gauss_spectrum(s2, dec, stream=stream) for (s2, dec) in params]
    arr = plan.field_array([{"in": f, "out": f, "work": w, "kspec": k}
                            for f, w, k in zip(fields, works, kspecs)])
    torch.cuda.synchronize()
    # schedule autotuner (per pass: radix 16/8 and work-group size), untimed
    schedule = plan.schedule if args.no_autotune else plan.autotune(len(fields), stream)[0]

    for _ in range(args.warmup):
        plan.diffuse(arr, stream)
    torch.cuda.synchronize()

    # every launch of a step on its own, in the order vfd_diffuse issues them
    # (the z pass is its four launches), with a HIP event between launches
    launches = [("xfwd", vfd.PASS_XFWD), ("yfwd", vfd.PASS_YFWD)] + \
               [(n, vfd.PASS_Z_DIRECT8 + i) for i, n in enumerate(vfd.Z_PART_NAMES)] + \
               [("yinv", vfd.PASS_YINV), ("xinv", vfd.PASS_XINV)]
    nl = len(launches)
    K = args.steps
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(nl + 1)] for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        evs[k][0].record(stream)
        for i, (_, p) in enumerate(launches):
            plan.diffuse_pass(arr, p, stream)
            evs[k][i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    kern_ms = [0.0] * nl
    step_ms = []
    for k in range(K):
        for i in range(nl):
            kern_ms[i] += evs[k][i].elapsed_time(evs[k][i + 1])
        step_ms.append(evs[k][0].elapsed_time(evs[k][nl]))
    kern_ms = [t / K for t in kern_ms]
    nf = len(fields)
    ms_per_step = wall * 1e3 / K
    vox = float(nz * ny * nx) * nf
    value = vox * K / wall
    # fields per launch: the z launches act on the fields of their own form
    # (window radius R = ceil(6 sigma) <= 8, <= 12, <= 16 direct; beyond, FFT)
    radii = [int(np.ceil(6.0 * np.sqrt(s2))) for s2, _ in params]
    if kfmt == vfd.KSPEC_SEPARABLE:
        forms = [0 if r <= 8 else 1 if r <= 12 else 2 if r <= 16 else 3 for r in radii]
    else:
        forms = [3] * nf
    kern_fields = [nf, nf] + [forms.count(i) for i in range(4)] + [nf, nf]
    kern_bytes = [plan.pass_bytes(p) * kern_fields[i] for i, (_, p) in enumerate(launches)]
    kern_gbs = [b / (t * 1e-3) / 1e9 if t > 0 else 0.0 for b, t in zip(kern_bytes, kern_ms)]
    dom = int(np.argmax(kern_ms))
    step_bytes = plan.info.bytes_per_field_tick * nf
    mean_step_ms = float(np.mean(step_ms))
    copy_gbs = copy_bandwidth(vfd, torch, stream)
    mix_gbs = mix_bandwidth(vfd, torch, stream)
    fft_z = None if args.no_fft_z else fft_z_step(vfd, torch, np, shape, fields, params, model_bytes, stream,
                                                  args.warmup, min(K, 20))

    out = {
        "metric": "voxel-updates/s for 8-field step at 1024x512x128 (% HBM roofline); p50/p99 latency",
        "value": value,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device-generated lognormal background + 50 point sources; "
                "random D, lambda per field)",
        "config": {"workload": f"{nf} FP32 cytokine fields, {nx}x{ny}x{nz} patch grid, zero-padded "
                               f"linear convolution on {2*nx}x{2*ny}x{2*nz} FFT box, Gaussian "
                               "Green's kernel x decay, 1 tick per step",
                   "fields": nf, "grid": [nx, ny, nz], "fft_box": [2 * nx, 2 * ny, 2 * nz],
                   "kernel_spectrum": args.kspec,
                   "schedule": dict(zip(vfd.PASS_NAMES, schedule))},
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99))},
        "roofline": {
            "bound": "hbm",
            # the single kernel with the largest average duration
            "kernel": launches[dom][0],
            "achieved": kern_gbs[dom],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": kern_gbs[dom] / HBM_PEAK_GBS,
            "traffic": measured_traffic(shape, nf, args.kspec, launches[dom][0]),
            # the whole step: the bytes its launches move / its duration
            "step_frac": step_bytes / (mean_step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "kernels_ms": {n: t for (n, _), t in zip(launches, kern_ms)},
            "kernels_gbs": {n: g for (n, _), g in zip(launches, kern_gbs)},
            "kernel_fields": {n: c for (n, _), c in zip(launches, kern_fields)},
            "algorithmic_bytes_per_step": step_bytes,
            # measured streaming reference: 16-byte non-temporal copy of 1 GiB
            "copy_gbs": copy_gbs,
            "step_frac_of_copy": step_bytes / (mean_step_ms * 1e-3) / 1e9 / copy_gbs,
            # the same streaming kernel for each launch's read:write mix, and
            # each launch's GB/s as a fraction of its mix's
            "mix_gbs": mix_gbs,
            "kernels_frac_of_mix": {n: (g / mix_gbs[LAUNCH_MIX[n]] if g > 0 else None)
                                    for (n, _), g in zip(launches, kern_gbs)},
            # the five-pass byte model with a streamed real K^ (8N + 32 A1 + 32 A2 + 4 A3):
            # counts 4 A3 per field that the separable spectrum never reads
            "model_bytes_per_step": model_bytes,
            "model_frac": model_bytes / (mean_step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
    }
    if fft_z is not None:
        out["fft_z_step"] = fft_z
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(shape, args.seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
