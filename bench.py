"""Throughput of the parallel cascaded-biquad IIR path on one MI355X.

Metric (the reference's headline, README.md:34-39): GSamples/s and % of the
HBM roof at 2^25 float32 samples, orders 2-16.  One step = one filter call
(one grid-wide kernel launch) over 2^25 samples already resident in HBM.
`value` is the order-2 step of the configured algorithm (PHH 64x32 by
default: PH with halo carries); every order is
measured and reported under "orders".  Inputs rotate over enough buffers to
exceed twice the 256 MB Infinity Cache, so every launch streams from HBM
(the reference rotates against its L2, test_performance.cu:140-157).

    python bench.py [--steps K] [--warmup W] [--kernel PHH] [--geom 64x32]
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
BYTES_PER_SAMPLE = 8           # fp32 in + fp32 out
N_SAMPLES = 1 << 25
MALL_BYTES = 256 << 20


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # untimed launches before each timed run: the GPU's clocks take about
    # 20 ms of back-to-back launches to settle (order 2: 689-691 GS/s after 20
    # warmup launches, 711 after 400, same box)
    ap.add_argument("--warmup", type=int, default=400)
    # untimed launches before every configuration other than the headline one
    # (the other orders, the candidates, the reference grid, the ticket
    # schedule and the MFMA forms), so that each is timed at settled clocks
    # however short --warmup is: the MFMA-heavy block-state form at order 16
    # read 405 GS/s after 5 launches and 444 after 200 (same box); at 200 the
    # block-state form at orders 8 / 16 still read 547 / 445 against 566 / 470
    # after 400 (same box, same build)
    ap.add_argument("--settle", type=int, default=400)
    # PHH: PH whose chunk carries come from the 256 input samples before each
    # chunk where the filter forgets its state within them (orders 2 and 4 at
    # Wn = 0.1), so no chunk waits on another and the launch assumes no
    # co-residency under either schedule; at orders 8 and 16 the filter's
    # memory is longer and the plan runs PH unchanged ("halo": 0)
    ap.add_argument("--kernel", default="PHH")
    ap.add_argument("--geom", default="64x32")
    ap.add_argument("--schedule", default="persistent")
    ap.add_argument("--orders", default="2,4,8,16")
    ap.add_argument("--wn", type=float, default=0.1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--geom-steps", type=int, default=None,
                    help="timed launches per configuration of the reference-geometry grid "
                         "(default: --steps; 0 skips the grid)")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the FP64 gate of every reported configuration (scipy sosfilt float64 golden)")
    ap.add_argument("--check", action="store_true", help="ragged/sweep workloads: gate sampled channels")
    ap.add_argument("--latency-steps", type=int, default=20,
                    help="single launches, each after a synchronisation, timed for the one-call latency")
    ap.add_argument("--no-ticket", action="store_true", help="skip the ticket-schedule measurement")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the all-cores CPU baseline (0: all)")
    ap.add_argument("--bs-geom", default="64x64",
                    help="block-state form geometry GxL for the MFMA vs plain fp32 comparison (orders 8 and 16)")
    ap.add_argument("--bs-forms", default="BS:64x64,BS:16x128,BS_BF16:32x64,BS_BF16:16x128",
                    help="block-state forms (kernel:GxL) timed beside the plain fp32 kernel at orders 8 and 16: "
                         "BS = FP16 hi/lo split, BS_BF16 = BF16 hi/mid/lo split")
    ap.add_argument("--no-mfma", action="store_true", help="skip the MFMA vs plain fp32 comparison")
    ap.add_argument("--candidates", default="PH:64x32,PHJ:64x32,BS:64x64",
                    help="further configurations (kernel:BxN) timed and gated at every order beside --kernel; "
                         "PHJ = PH with paired sections, PHH = PH with the chunk carries from the 256 samples "
                         "before each chunk (where the filter forgets within them, else PH), "
                         "BS = block-state form (MFMA)")
    ap.add_argument("--workload", default="single", choices=["single", "ragged", "sweep"],
                    help="single: one 2^25-sample channel per order; ragged: the 4096-channel batch; "
                         "sweep: order 8 at Wn 0.02/0.1/0.4 over lengths 2^16..2^25")
    ap.add_argument("--batch-n", type=int, default=32, help="ragged batch: samples per lane (16, 32 or 64)")
    ap.add_argument("--batch-bs-min-sections", type=int, default=0,
                    help="ragged batch: channels with at least this many sections run the block-state form "
                         "(0: the library's choice, 8; 9: none)")
    ap.add_argument("--batch-bs-tiles", type=int, default=32,
                    help="ragged batch: block-state tiles of 64 samples per chunk (16, 32 or 64)")
    ap.add_argument("--copy-steps", type=int, default=400,
                    help="launches of the library's 16-byte streaming copy of one 2^25-sample buffer, timed as "
                         "the measured HBM ceiling before the filters (also the known-size kernel that scales "
                         "rocprof FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def time_order(iirx, torch, order, args, bufs, out, stream, kernel=None, geom=None, warmup=None):
    kernel = kernel or args.kernel
    warmup = args.warmup if warmup is None else warmup
    B, N = (int(v) for v in (geom or args.geom).lower().split("x"))
    sos, g = iirx.butterworth_sos(order, args.wn)
    if kernel.upper() == "PLR":  # --geom BLOCKxX, 0 = automatic
        plan = iirx.Plan("PLR", sos, N_SAMPLES, block_size=B, plr_x=N, schedule=args.schedule)
    else:
        plan = iirx.Plan(kernel, sos, N_SAMPLES, block_size=B, n_blocks=N, schedule=args.schedule)
    for i in range(warmup):
        plan.execute(bufs[i % len(bufs)], out, stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(args.steps):
        plan.execute(bufs[i % len(bufs)], out, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    plan.check()
    info = plan.info()
    res = {"sos": sos, "gain": g, "ms": ms, "wall_ms": 1e3 * wall / args.steps, "info": info,
           "last_input": (warmup + args.steps - 1) % len(bufs), "plan": plan}
    return res


# the reference's BLOCK_SIZE x N_BLOCKS grid (32x32, 32x64, 32x128, 64x64); a
# chunk is BLOCK_SIZE x N_BLOCKS samples on one wave64, so 32x128 is the 64x64
# chunk and kernel and is reported once, as 64x64
REFERENCE_GEOMETRIES = ("32x32", "32x64", "64x64")


def gated(gs, out, ref, args):
    """A measured configuration with its FP64 gate verdict; a configuration
    that fails the gate reports 0 GS/s, as the reference's tables do, with
    the measured figure beside it."""
    if args.no_check or ref is None:
        return {"gsamples_per_s": gs}
    ok, rel = gate(out.cpu().numpy(), ref)
    return {"gsamples_per_s": gs if ok else 0.0, "measured_gsamples_per_s": gs, "gate_pass": bool(ok),
            "gate_rel": rel}


def reference_geometries(iirx, torch, order, args, bufs, out, stream, ref):
    """The reference's own grid for one order: PH, STCR (loop and hand-unrolled)
    and DTCR at BLOCK_SIZE x N_BLOCKS = 32x32, 32x64, 64x64 (= 32x128), and PLR
    at its automatic geometry; GSamples/s over args.geom_steps timed launches
    with each configuration's gate verdict, or the status name when the
    configuration is rejected (CR conditioning)."""
    sos, g = iirx.butterworth_sos(order, args.wn)
    res = {}
    cases = [(k, u, geom) for k, u in (("PH", False), ("STCR", False), ("STCR", True), ("DTCR", False))
             for geom in REFERENCE_GEOMETRIES] + [("PLR", False, "auto")]
    for kernel, unrolled, geom in cases:
        name = f"{kernel}{'_unrolled' if unrolled else ''} {geom}"
        try:
            if kernel == "PLR":
                plan = iirx.Plan("PLR", sos, N_SAMPLES, schedule=args.schedule)
            else:
                B, N = (int(v) for v in geom.split("x"))
                plan = iirx.Plan(kernel, sos, N_SAMPLES, block_size=B, n_blocks=N, unrolled=unrolled,
                                 schedule=args.schedule)
        except iirx.IirxError as e:
            res[name] = str(e)
            continue
        for i in range(max(args.settle, args.warmup)):
            plan.execute(bufs[i % len(bufs)], out, stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in range(args.geom_steps):
            plan.execute(bufs[i % len(bufs)], out, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        plan.check()
        plan.close()
        res[name] = gated(round(N_SAMPLES / (e0.elapsed_time(e1) / args.geom_steps * 1e-3) / 1e9, 1), out, ref, args)
    return res


def time_plan(torch, plan, bufs, out, stream, steps, warmup=5):
    """ms per launch over `steps` back-to-back launches (HIP events on the
    launch stream), after `warmup` launches."""
    for i in range(warmup):
        plan.execute(bufs[i % len(bufs)], out, stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(steps):
        plan.execute(bufs[i % len(bufs)], out, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    plan.check()
    return e0.elapsed_time(e1) / steps


def latency(torch, plan, bufs, out, stream, steps):
    """One call's latency: a single launch after a synchronisation, HIP events
    around it, median over `steps` calls (the back-to-back figure hides the
    launch overhead and the ramp of a lone launch)."""
    ts = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for i in range(steps):
        torch.cuda.synchronize()
        e0.record(stream)
        plan.execute(bufs[i % len(bufs)], out, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    plan.check()
    ts.sort()
    return ts[len(ts) // 2]


def copy_ceiling(iirx, torch, args, dev, stream):
    """Measured HBM ceiling: the library's 16-byte streaming copy kernel over
    one 2^25-float buffer (read 4 B + write 4 B per sample, as the filters),
    inputs rotating over 4 buffers; nontemporal stores (mode 0) and stores
    through the L2 (mode 1, as the cascades store); torch's copy_ beside it.
    Run before the filters, its back-to-back launches also bring the clocks
    up."""
    a = torch.empty(N_SAMPLES, dtype=torch.float32, device=dev).normal_()
    bufs = [a.clone() for _ in range(4)]
    dst = torch.empty_like(a)
    res = {}
    for name, fn in (("stream_copy_nt_store", lambda i: iirx.copy_ceiling(bufs[i % 4], dst, N_SAMPLES, 0, stream)),
                     ("stream_copy_l2_store", lambda i: iirx.copy_ceiling(bufs[i % 4], dst, N_SAMPLES, 1, stream)),
                     ("torch_copy", lambda i: dst.copy_(bufs[i % 4]))):
        for i in range(5):
            fn(i)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in range(args.copy_steps):
            fn(i)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / max(args.copy_steps, 1)
        res[name] = {"ms": round(ms, 5), "gbs": round(BYTES_PER_SAMPLE * N_SAMPLES / (ms * 1e-3) / 1e9, 1)}
    assert torch.equal(dst, bufs[(args.copy_steps - 1) % 4])
    best = max(res["stream_copy_nt_store"]["gbs"], res["stream_copy_l2_store"]["gbs"])
    res["gbs"] = best
    res["frac_of_peak"] = round(best / HBM_PEAK_GBS, 4)
    return res


def measured_traffic(args):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    kernel/geometry (profiles/traffic.json), or None."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(f"{args.kernel}_{args.geom}_order2", None)
    except (OSError, ValueError):
        return None


def kernel_is_phh(name):
    return name.upper() in ("PHH", "PH_HALO")


def fp64_reference(sos, x):
    """FP64 cascade of the float32 sections over the float32 input
    (scipy.signal.sosfilt in float64, the golden of ref_generate.py:89)."""
    import numpy as np
    from scipy import signal
    return signal.sosfilt(np.asarray(sos, np.float64), np.asarray(x, np.float64))


def gate(y, ref, rel_tol=1e-4):
    """The reference's accuracy gate (test_accuracy.cu:178-190): max|y - ref| <= rel_tol * max|ref|."""
    import numpy as np
    y = np.asarray(y, np.float64)
    if not np.all(np.isfinite(y)):
        return False, None  # non-finite output (JSON has no inf / nan)
    err = float(np.max(np.abs(y - ref)))
    mref = float(np.max(np.abs(ref)))
    return err <= rel_tol * mref, err / mref


def _timed(fn, budget):
    """Passes of fn() for at least `budget` seconds (one pass at least): (passes, seconds)."""
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget:
            return reps, el


def cpu_baseline(iirx, args):
    """The oracle's float32 PH restatement (a port of the reference algorithm at
    its published 32x64 geometry), single thread, over the 2^25-sample
    order-2 workload, repeated for about --cpu-seconds; and the same
    restatement on all host cores (OpenMP over chunks, bit-identical output)."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import numpy as np
    import oracle_lib as O
    sos, g = iirx.butterworth_sos(2, args.wn)
    x = (g * iirx.fill_normal(1, N_SAMPLES).astype(np.float64)).astype(np.float32)
    reps, el = _timed(lambda: O.ph(sos, x, 32, 64), args.cpu_seconds)
    one = {"value": round(reps * N_SAMPLES / el / 1e9, 4), "unit": "GSamples/s", "cores": 1, "kind": "port",
           "sample": f"oracle PH float32 restatement (32x64, order 2, Wn={args.wn}) over 2^25 N(0,1) samples, "
                     f"{reps} passes in {el:.1f} s, single thread"}
    threads = args.cpu_threads or (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())
    reps, el = _timed(lambda: O.ph_mt(sos, x, 32, 64, threads), args.cpu_seconds / 2)
    allc = {"value": round(reps * N_SAMPLES / el / 1e9, 4), "unit": "GSamples/s", "cores": threads, "kind": "port",
            "sample": f"oracle PH float32 restatement (32x64, order 2, Wn={args.wn}) over 2^25 N(0,1) samples, "
                      f"{reps} passes in {el:.1f} s, OpenMP over chunks ({threads} threads, output identical "
                      f"to the single-thread restatement)"}
    return one, allc


def cpu_baselines_more(iirx, args):
    """The other CPU references the workload names, each over the same
    2^25-sample input at orders 2 and 16 (one pass at least): scipy.signal.
    sosfilt in float64 (the golden; single thread), the oracle's PH
    restatement on all host cores, and the oracle's float32 restatements of
    the reference's STCR, DTCR and PLR emulators (tools/stcr_emulator.py,
    dtcr_emulator.py, plr_emulator.py) at their published geometries, single
    thread."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import numpy as np
    from scipy import signal
    import oracle_lib as O
    n = N_SAMPLES
    threads = args.cpu_threads or (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())
    out = []
    for order in (2, 16):
        sos, g = iirx.butterworth_sos(order, args.wn)
        x = (g * iirx.fill_normal(1, n).astype(np.float64)).astype(np.float32)
        bt, at = (np.asarray(v, np.float64) for v in iirx.direct_form_taps(sos))
        cases = [("scipy.signal.sosfilt float64", "library", 1,
                  lambda: signal.sosfilt(sos.astype(np.float64), x.astype(np.float64))),
                 ("oracle STCR float32 restatement 32x64", "port", 1, lambda: O.stcr(sos, x, 32, 64)),
                 ("oracle DTCR float32 restatement 32x64", "port", 1, lambda: O.dtcr(sos, x, 32, 64))]
        if order == 2:
            cases.append(("oracle PLR float32 restatement 256x16", "port", 1, lambda: O.plr(bt, at, x, 256, 16)))
        else:
            cases.append(("oracle PH float32 restatement 32x64", "port", threads, lambda: O.ph_mt(sos, x, 32, 64, threads)))
        for name, kind, cores, fn in cases:
            reps, el = _timed(fn, 0.5)
            how = "single thread" if cores == 1 else f"OpenMP over chunks, {cores} threads"
            out.append({"value": round(reps * n / el / 1e9, 4), "unit": "GSamples/s", "cores": cores, "kind": kind,
                        "sample": f"{name}, order {order}, Wn={args.wn}, 2^25 N(0,1) samples, {reps} passes in "
                                  f"{el:.1f} s, {how}"})
    return out


def ragged_traffic(args):
    """HBM bytes per replay of the default ragged configuration, from the
    committed PMC summary (profiles/traffic.json), or None for other settings."""
    if args.batch_n != 32 or args.batch_bs_min_sections not in (0, 8) or args.batch_bs_tiles != 32:
        return None
    try:
        with open(os.path.join(HERE, "profiles", "traffic.json")) as f:
            return json.load(f).get("ragged4096_default")
    except (OSError, ValueError):
        return None


def bench_ragged(args):
    """The ragged batch: 4096 channels, lengths log-uniform in [2^10, 2^20],
    order uniform over {2, 4, 8, 16}, Butterworth Wn uniform in [0.02, 0.45]
    (seed 4), gain * N(0,1) inputs generated on the device, CSR-packed; one
    step = one replay of the hipGraph-captured persistent launch."""
    import numpy as np
    import torch
    import cuda_recursive_filtering_amd as iirx
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from workloads import ragged_channels
    dev = torch.device("cuda:0")
    offsets, sections, gains = ragged_channels(4096, seed=4)
    n = int(offsets[-1])
    torch.manual_seed(4)
    x = torch.randn(n, dtype=torch.float32, device=dev)
    lens = torch.from_numpy(np.diff(offsets)).to(dev)
    x.mul_(torch.repeat_interleave(torch.from_numpy(gains.astype(np.float32)).to(dev), lens))
    out = torch.empty_like(x)
    s = torch.cuda.Stream()
    b = iirx.Batch(offsets, sections, n_blocks=args.batch_n, schedule=args.schedule,
                   bs_min_sections=args.batch_bs_min_sections, bs_tiles=args.batch_bs_tiles)
    b.capture(x, out, s)
    for _ in range(args.warmup):
        b.replay(s)
    s.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(args.steps):
        b.replay(s)
    e1.record(s)
    s.synchronize()
    b.check()
    ms = e0.elapsed_time(e1) / args.steps
    # the same launch without the graph, for comparison
    e0.record(s)
    for _ in range(args.steps):
        b.execute(x, out, s)
    e1.record(s)
    s.synchronize()
    b.check()
    ms_eager = e0.elapsed_time(e1) / args.steps
    info = b.info()
    gbs = BYTES_PER_SAMPLE * n / (ms * 1e-3) / 1e9
    line = {
        "metric": "GSamples/s, ragged batch (4096 channels, orders 2-16, one captured launch)",
        "value": round(n / (ms * 1e-3) / 1e9, 3),
        "unit": "GSamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "ragged batch: 4096 channels, lengths log-uniform [2^10, 2^20], order uniform "
                               "{2,4,8,16}, Butterworth Wn uniform [0.02, 0.45] (seed 4), gain * N(0,1)",
                   "n_samples": n, "n_channels": 4096,
                   "kernel": ("PH per chunk" if not info["n_blockstate"] else
                              f"PH, block-state for >= {args.batch_bs_min_sections or 8} sections"),
                   "blockstate_channels": info["n_blockstate"], "samples_per_lane": args.batch_n,
                   "blockstate_tiles": args.batch_bs_tiles, "launches_per_step": info["n_launches"],
                   "chunks": info["num_chunks"], "grid": info["grid"], "schedule": args.schedule},
        "eager_ms_per_step": round(ms_eager, 4),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": ragged_traffic(args),
                     "algorithmic_bytes": BYTES_PER_SAMPLE * n},
    }
    if args.check:
        worst = 0.0
        for c in range(0, 4096, 128):
            a_, b_ = int(offsets[c]), int(offsets[c + 1])
            ok, rel = gate(out[a_:b_].cpu().numpy(), fp64_reference(sections[c], x[a_:b_].cpu().numpy()))
            worst = max(worst, rel)
            assert ok, (c, rel)
        line["gate_worst_rel"] = worst
    b.close()
    print(json.dumps(line), flush=True)


def bench_sweep(args):
    """Length / conditioning sweep: order-8 Butterworth at Wn in {0.02, 0.1, 0.4},
    lengths 2^16..2^25, N(0,1) + 0.5 sin(2 pi 0.05 n) (seed 2); GSamples/s of
    --kernel/--geom per (Wn, length), inputs resident in HBM (one buffer: the
    short lengths are cache-resident, as they are in use)."""
    import numpy as np
    import torch
    import cuda_recursive_filtering_amd as iirx
    B, N = (int(v) for v in args.geom.lower().split("x"))
    dev = torch.device("cuda:0")
    rows = []
    for wn in (0.02, 0.1, 0.4):
        sos, g = iirx.butterworth_sos(8, wn)
        for log2n in range(16, 26):
            n = 1 << log2n
            x = torch.from_numpy((g * (iirx.fill_normal(2, n).astype(np.float64)
                                       + 0.5 * np.sin(2 * np.pi * 0.05 * np.arange(n)))).astype(np.float32)).to(dev)
            out = torch.empty_like(x)
            with iirx.Plan(args.kernel, sos, n, block_size=B, n_blocks=N, schedule=args.schedule) as p:
                for _ in range(args.warmup):
                    p.execute(x, out)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(args.steps):
                    p.execute(x, out)
                e1.record()
                torch.cuda.synchronize()
                p.check()
                ms = e0.elapsed_time(e1) / args.steps
                info = p.info()
            rows.append({"wn": wn, "log2n": log2n, "ms": round(ms, 5), "gsamples_per_s": round(n / (ms * 1e-3) / 1e9, 2),
                         "chunks": info["num_chunks"], "grid": info["grid"]})
            print(f"# wn={wn} n=2^{log2n}: {rows[-1]}", file=sys.stderr, flush=True)
    big = [r for r in rows if r["log2n"] == 25 and r["wn"] == 0.1][0]
    print(json.dumps({"metric": "GSamples/s, order-8 length/conditioning sweep", "value": big["gsamples_per_s"],
                      "unit": "GSamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": big["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "order-8 Butterworth, Wn 0.02/0.1/0.4, 2^16..2^25 samples of N(0,1) + "
                                             "0.5 sin(2 pi 0.05 n) (seed 2); value = Wn 0.1 at 2^25",
                                 "kernel": args.kernel, "geometry": args.geom}, "sweep": rows}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "ragged":
        bench_ragged(args)
        return
    if args.workload == "sweep":
        bench_sweep(args)
        return
    import numpy as np
    import torch
    import cuda_recursive_filtering_amd as iirx
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    # the measured ceiling first: its launches also bring the clocks up
    ceiling = copy_ceiling(iirx, torch, args, dev, stream) if args.copy_steps > 0 else None
    if ceiling:
        print(f"# copy ceiling: {ceiling}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    nbuf = max(4, math.ceil(2 * MALL_BYTES / (4 * N_SAMPLES)) + 1)
    base = iirx.fill_normal(1, N_SAMPLES)
    orders = [int(o) for o in args.orders.split(",")]
    if args.geom_steps is None:
        args.geom_steps = args.steps
    per, mfma = {}, {}
    for order in orders:
        sos, g = iirx.butterworth_sos(order, args.wn)
        host = (g * base.astype(np.float64)).astype(np.float32)
        src = torch.from_numpy(host).to(dev)
        bufs = [src] + [src.clone() for _ in range(nbuf - 1)]
        out = torch.empty_like(src)
        # the headline (first order, the kernel) after exactly --warmup launches
        r = time_order(iirx, torch, order, args, bufs, out, stream,
                       warmup=args.warmup if order == orders[0] else max(args.settle, args.warmup))
        gs = N_SAMPLES / (r["ms"] * 1e-3) / 1e9
        gbs = BYTES_PER_SAMPLE * N_SAMPLES / (r["ms"] * 1e-3) / 1e9
        entry = {"gsamples_per_s": round(gs, 3), "ms": round(r["ms"], 5), "achieved_gbs": round(gbs, 1),
                 "roof_frac": round(gbs / HBM_PEAK_GBS, 4),
                 # textbook cascade: 4 FMA = 8 flop per sample and section (reported, not the roof)
                 "algorithmic_gflops": round(gs * 4 * order, 1), "vgprs": r["info"]["vgprs"],
                 "scratch": r["info"]["scratch_bytes"], "chunks": r["info"]["num_chunks"],
                 "grid": r["info"]["grid"]}
        if kernel_is_phh(args.kernel):  # 1: carries from the halo, 0: the filter's memory is longer and PH ran
            entry["halo"] = r["info"]["halo"]
        if ceiling:
            entry["ceiling_frac"] = round(gbs / ceiling["gbs"], 4)
        # the FP64 golden of this order's input (every buffer holds the same samples)
        ref = None if args.no_check else fp64_reference(sos, host)
        if ref is not None:
            ok, rel = gate(out.cpu().numpy(), ref)
            entry["gate_rel"] = rel
            entry["gate_pass"] = bool(ok)
        # one call on its own, after a synchronisation
        if args.latency_steps > 0:
            entry["latency_us"] = round(1e3 * latency(torch, r["plan"], bufs, out, stream, args.latency_steps), 2)
        # the forward-progress-safe schedule (atomic tickets; PHH with halo
        # carries needs none and runs the same no-wait launch under it)
        if not args.no_ticket and args.kernel.upper() != "PLR":
            B, N = (int(v) for v in args.geom.lower().split("x"))
            with iirx.Plan(args.kernel, sos, N_SAMPLES, block_size=B, n_blocks=N, schedule="ticket") as pt:
                ms_t = time_plan(torch, pt, bufs, out, stream, args.geom_steps or args.steps,
                                 warmup=max(args.settle, args.warmup))
            entry["ticket"] = gated(round(N_SAMPLES / (ms_t * 1e-3) / 1e9, 3), out, ref, args)
            entry["ticket"]["ms"] = round(ms_t, 5)
            if entry.get("halo") == 1:
                entry["ticket"]["no_wait"] = True
            # the block-state form under the ticket schedule (one ticket per
            # workgroup of chunks), the safe schedule's best at orders 8 and 16
            if order >= 8:
                with iirx.Plan("BS", sos, N_SAMPLES, block_size=64, n_blocks=64, schedule="ticket") as pb:
                    ms_b = time_plan(torch, pb, bufs, out, stream, args.geom_steps or args.steps,
                                     warmup=max(args.settle, args.warmup))
                entry["ticket_blockstate"] = gated(round(N_SAMPLES / (ms_b * 1e-3) / 1e9, 3), out, ref, args)
                entry["ticket_blockstate"]["ms"] = round(ms_b, 5)
                entry["ticket_blockstate"]["config"] = "BS 64x64"
        if not args.no_mfma and order in (8, 16):
            # MFMA block-state form beside the plain fp32 kernel, same input, error vs FP64
            refm = ref if ref is not None else fp64_reference(sos, host)
            rms = float(np.sqrt(np.mean(refm ** 2)))
            last = bufs[0]
            def errs(plan):
                plan.execute(last, out, stream)
                torch.cuda.synchronize()
                d = out.cpu().numpy().astype(np.float64) - refm
                return float(np.max(np.abs(d))) / rms, float(np.sqrt(np.mean(d ** 2))) / rms
            e_fp32, r_fp32 = errs(r["plan"])
            forms = []
            for spec in [f"BS:{args.bs_geom}"] + [f for f in args.bs_forms.split(",") if f and f != f"BS:{args.bs_geom}"]:
                kname, geom = spec.split(":")
                rb = time_order(iirx, torch, order, args, bufs, out, stream, kernel=kname, geom=geom,
                                warmup=max(args.settle, args.warmup))
                e_bs, r_bs = errs(rb["plan"])
                ok_bs, rel_bs = gate(out.cpu().numpy(), refm)
                rb["plan"].close()
                forms.append({"kernel": kname, "split": "bf16 hi/mid/lo" if kname == "BS_BF16" else "fp16 hi/lo",
                              "geometry": geom, "gsamples_per_s": round(N_SAMPLES / (rb["ms"] * 1e-3) / 1e9, 3),
                              "max_err_over_rms": e_bs, "rms_err_over_rms": r_bs, "gate_pass": bool(ok_bs),
                              "gate_rel": rel_bs})
            f0 = forms[0]
            mfma[str(order)] = {
                "bs_gsamples_per_s": f0["gsamples_per_s"], "bs_geometry": f0["geometry"], "bs_split": f0["split"],
                "fp32_gsamples_per_s": entry["gsamples_per_s"],
                "fp32_kernel": f"{args.kernel} {args.geom}" + (" (no halo: PH)" if entry.get("halo") == 0 else ""),
                "bs_max_err_over_rms": f0["max_err_over_rms"], "fp32_max_err_over_rms": e_fp32,
                "bs_rms_err_over_rms": f0["rms_err_over_rms"], "fp32_rms_err_over_rms": r_fp32,
                "bs_gate_pass": f0["gate_pass"], "bs_gate_rel": f0["gate_rel"], "forms": forms}
            print(f"# mfma order {order}: {mfma[str(order)]}", file=sys.stderr, flush=True)
        cands = {}
        for spec in [c for c in args.candidates.split(",") if c]:
            kname, geom = spec.split(":")
            try:
                rc = time_order(iirx, torch, order, args, bufs, out, stream, kernel=kname, geom=geom,
                                warmup=max(args.settle, args.warmup))
            except iirx.IirxError as e:
                cands[f"{kname} {geom}"] = str(e)
                continue
            cands[f"{kname} {geom}"] = gated(round(N_SAMPLES / (rc["ms"] * 1e-3) / 1e9, 3), out, ref, args)
            if kernel_is_phh(kname):  # 1: carries from the halo, 0: the filter's memory is longer and PH ran
                cands[f"{kname} {geom}"]["halo"] = rc["plan"].info()["halo"]
            rc["plan"].close()
        if cands:
            entry["candidates"] = cands
        if args.geom_steps > 0:
            entry["reference_geometries"] = reference_geometries(iirx, torch, order, args, bufs, out, stream, ref)
        per[str(order)] = entry
        r["plan"].close()
        del bufs, src, out
        torch.cuda.empty_cache()
        print(f"# order {order}: {entry}", file=sys.stderr, flush=True)

    best = {}
    for o, e in per.items():
        seen = [(f"{args.kernel} {args.geom}", e["gsamples_per_s"], e.get("gate_pass"))]
        for group in ("candidates", "reference_geometries"):
            for name, v in e.get(group, {}).items():
                if isinstance(v, dict):
                    seen.append((name, v.get("measured_gsamples_per_s", v["gsamples_per_s"]), v.get("gate_pass")))
        for f in (mfma.get(o) or {}).get("forms", []):
            seen.append((f"{f['kernel']} {f['geometry']}", f["gsamples_per_s"], f.get("gate_pass")))
        ok = [t for t in seen if t[2] is not False]
        if ok:
            name, gs, gp = max(ok, key=lambda t: t[1])
            best[o] = {"config": name, "gsamples_per_s": gs, "roof_frac": round(gs * BYTES_PER_SAMPLE / HBM_PEAK_GBS, 4),
                       "gate_pass": gp}
    o2 = per[str(orders[0])]
    line = {
        "metric": "GSamples/s & % HBM roof at 2^25, order 2-16; MFMA vs plain fp32 @ equal err",
        "value": o2["gsamples_per_s"],
        "unit": "GSamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_launches": max(args.settle, args.warmup),
        "ms_per_step": o2["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"order-{orders[0]} Butterworth lowpass Wn={args.wn}, 2^25 float32 N(0,1) "
                               f"samples (seed 1, gain pre-applied), one channel, zero initial state",
                   "kernel": args.kernel, "geometry": args.geom, "schedule": args.schedule,
                   "n_samples": N_SAMPLES,
                   "input_buffers": nbuf},
        "gate_pass": o2.get("gate_pass"),
        "gate_rel": o2.get("gate_rel"),
        "orders": per,
        "roofline": {"bound": "hbm", "achieved": o2["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": o2["roof_frac"], "traffic": measured_traffic(args),
                     "algorithmic_bytes": BYTES_PER_SAMPLE * N_SAMPLES},
        "copy_ceiling": ceiling,
        "mfma_vs_fp32": mfma or None,
        # the fastest configuration that passes the FP64 gate at each order, among
        # every one measured above (the kernel, the candidates, the reference grid,
        # the MFMA forms)
        "best_per_order": best,
    }
    line["host_cores"] = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    if not args.no_cpu:
        line["cpu_baseline"], line["cpu_baseline_all_cores"] = cpu_baseline(iirx, args)
        line["cpu_baselines_more"] = cpu_baselines_more(iirx, args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
