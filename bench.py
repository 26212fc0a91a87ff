"""Benchmark: Mrealisations/s of the null-collision path tracer on one MI355X.

Workload (default): the precision configuration -- htrdr-atmosphere
shortwave image, 1024x1024 pixels, 256 realisations per pixel, monochromatic
550 nm, cloud grid 256x256x64 over 25.6x25.6x3.2 km (base 1 km) with the
seeded lognormal LWC law masked to 50% cover, octree with 2^3 leaf blocks,
50 gas layers, Lambertian ground (albedo 0.3), sun at 30 deg zenith,
pinhole camera 60 deg fov tilted 60 deg down from (12800, -2000, 5000).

A step is one slice of `--spp-per-step` realisations per pixel over the whole
image; the default 16 steps x 16 spp = 256 spp is the full image. The scene,
its octree and the per-pixel accumulators are resident in HBM before the
timed region starts. The timed region is bracketed by device
synchronisations and every step is timed with HIP events on the stream its
kernels are launched on.

Printed: one JSON line (rank 0). `value` is whole-job throughput in
Mrealisations/s; `time_to_0p5pct_s` is the time the solver needs, at the
measured rate, to bring 99% of the pixels with E > 0 to a 0.5% relative
standard error (N_req(p) = (sigma_p / (0.005 E_p))^2 from the measured
per-pixel variance). `roofline` prices the dominant kernel's algorithmic
bytes against the 8 TB/s HBM peak: per realisation one leaf record per
octree-leaf lookup (8 B for the FP32 kernel: level word + FP16 bounds; 12 B
for the FP64 kernel: level-map word + FP32 bounds), 8 B per tentative
collision (4 B LWC + 4 B gas coefficient), 32 B per BVH node and 36 B per
triangle of a mesh ground, and 64 B per pixel per launch;
`cpu_baseline` is the CPU oracle timed on the host cores over a bounded
pixel sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrealisations/s & time-to-0.5%-std-error, FP32 batched vs FP64 plain"
SOLVER_NAMES = {0: "FP64 plain", 1: "FP32 wavefront"}
HBM_PEAK_GBS = 8000.0
# (solver, config) -> rocprofv3 summary of the kernel on that workload and the
# map width it was taken at, for its 16-realisation launches
PROFILED = {("wavefront", 1): ("traffic_wavefront_sw_image_1024.json", 1024),
            ("wavefront", 2): ("traffic_wavefront_lw_flux_512.json", 512),
            ("wavefront", 4): ("traffic_wavefront_mesh_flux_512.json", 512)}
# the uneven 64-scene batch, one step at full size: one batch-kernel launch
PROFILED_BATCH = "traffic_wavefront_batch_64.json"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=16)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--config", type=int, default=1)
    p.add_argument("--solver", choices=("plain", "wavefront"), default="wavefront")
    p.add_argument("--fp64-steps", type=int, default=1,
                   help="steps of the FP64 plain solver timed after the main run (0: skip)")
    p.add_argument("--spp-per-step", type=int, default=16)
    p.add_argument("--width", type=int, default=0, help="override the image width/height")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--autotune-ms", type=float, default=12000.0,
                   help="device-time budget of the launch-parameter autotuner run before the "
                        "warm-up (FP32 solver, 4x for --config 3; 0: default parameters)")
    p.add_argument("--launch-params", default="",
                   help="EV,CHUNK,GRID[,SEG]: fixed FP32 launch parameters instead of the autotuner "
                        "(0 = default for a field; profiles use the values the tuner picked)")
    p.add_argument("--batch-scale", type=float, default=1.0,
                   help="--config 3: scale of the realisations per pixel of every map")
    p.add_argument("--batch-sequential", action="store_true",
                   help="--config 3: also time the maps as back-to-back single-map launches")
    p.add_argument("--batch-steps", type=int, default=1, choices=(1, 2, 4),
                   help="--config 3: steps per run, each all/half/a quarter of every map's "
                        "realisations per pixel (one step is fastest: no barrier between steps)")
    p.add_argument("--batch-no-graph", action="store_true",
                   help="--config 3: launch the steps directly instead of replaying a HIP graph")
    p.add_argument("--no-graph", action="store_true",
                   help="FP32 solver: launch each timed step directly instead of replaying the "
                        "HIP graph of a one-map plan")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="only run 5 plain device copies of MIB MiB (profiler calibration)")
    return p.parse_args()


def calibrate_copy(mib):
    """Plain copies of a known size: rocprofv3's FETCH_SIZE / WRITE_SIZE of
    these dispatches give the scale of the traffic counters."""
    from htrdr_kmin_amd import _ffi as F
    from htrdr_kmin_amd import hip as H
    n = mib << 20
    a, b = H.DeviceBuffer(n), H.DeviceBuffer(n)
    a.zero()
    s = H.Stream()
    for _ in range(5):
        F.check("htgx_debug_copy", F.load().htgx_debug_copy(b.ptr, a.ptr, n, s.h))
    s.synchronize()
    print(json.dumps({"copy_bytes": n, "dispatches": 5}), flush=True)


def cpu_baseline(synth, desc, target_s):
    """Oracle (FP64 C/OpenMP restatement) timed on a strided pixel sample."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    from htrdr_kmin_amd import atmosphere as A
    threads = host_threads()
    sky = O.Sky(synth.scene_ptr)
    # large terrains: the oracle's own BVH (the same casts as its brute force)
    O.lib().oracle_sky_use_bvh(sky.h, 1)
    npix = desc.width * desc.height
    d = A.map_preset(1)
    for f, _ in d._fields_:
        setattr(d, f, getattr(desc, f))
    d.spp = 2
    d.spp_offset = 0
    # calibration on 1024 pixels (enough work to keep every thread busy),
    # then a sample sized for about target_s
    nc = min(npix, 1024)
    calib = np.arange(0, npix, max(1, npix // nc), dtype=np.uint64)[:nc]
    t0 = time.perf_counter()
    sky.draw_pixels(d, calib, threads)
    rate = len(calib) * d.spp / max(time.perf_counter() - t0, 1e-6)
    n = int(min(npix, max(64, rate * target_s / d.spp)))
    pix = np.linspace(0, npix - 1, n).astype(np.uint64)
    t0 = time.perf_counter()
    sky.draw_pixels(d, pix, threads)
    dt = time.perf_counter() - t0
    sky.close()
    return {"value": n * d.spp / dt / 1e6, "unit": "Mrealisations/s", "cores": threads,
            "kind": "port", "label": "restatement",
            "threads_from": THREADS_FROM,
            "sample": f"{n} pixels evenly strided over the {desc.width}x{desc.height} image x "
                      f"{d.spp} realisations, FP64 C/OpenMP oracle (the reference's algorithm "
                      f"restated; the reference itself cannot be built), {threads} threads, "
                      f"{dt:.1f} s"}


THREADS_FROM = ""


def host_threads():
    """Threads of the CPU baseline: every core this process may run on
    (its affinity mask), or OMP_NUM_THREADS when the host sets it (the CPU
    share of one GPU on a shared GPU node)"""
    global THREADS_FROM
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        THREADS_FROM = "OMP_NUM_THREADS"
        return int(env)
    THREADS_FROM = "sched_getaffinity"
    return len(os.sched_getaffinity(0))


def workload_name(synth, desc, S, steps):
    from htrdr_kmin_amd import _ffi as F
    sc = synth.scene
    c = sc.clouds
    kind = "image" if desc.sensor == F.SENSOR_CAMERA else "flux map"
    band = ("SW %g nm" % desc.wlen_range[0] if desc.wlen_range[0] == desc.wlen_range[1]
            else "%s %g-%g nm, %d bands x %d k-points" % (
                "LW" if desc.spectral_type == F.SPECTRAL_LW else "SW", desc.wlen_range[0],
                desc.wlen_range[1], sc.spectral.nbands, sc.spectral.nquads[0]))
    ground = {F.GROUND_PLANE: "plane ground", F.GROUND_NONE: "no ground",
              F.GROUND_MESH: "%d-triangle fBm ground" % sc.ground.ntris}[sc.ground.type]
    return ("htrdr-atmosphere %s %dx%d, %d realisations/px (%d steps x %d), %s, cloud grid "
            "%dx%dx%d over %.1fx%.1fx%.1f km, %d^3 leaf blocks, %s"
            % (kind, desc.width, desc.height, S * steps, steps, S, band, c.def_[0], c.def_[1],
               c.def_[2], (c.upper[0] - c.lower[0]) / 1e3, (c.upper[1] - c.lower[1]) / 1e3,
               (c.upper[2] - c.lower[2]) / 1e3, sc.leaf_block, ground))


def bench_batch(args):
    """--config 3: the uneven batched workload (64 seeded scenes of uneven
    grids, image sizes, realisations per pixel and k-points). A batch plan
    renders all of them in --batch-steps steps; each step is one persistent
    launch over the 64 maps (with its queue resets, clock ticks and merges),
    recorded once as a HIP graph and replayed. Optionally beside the same maps
    rendered by back-to-back single-map launches."""
    import numpy as np
    from htrdr_kmin_amd import _ffi as F
    from htrdr_kmin_amd import atmosphere as A
    from htrdr_kmin_amd import hip as H
    items = A.uneven_batch(seed=args.seed, n=64, spp_scale=args.batch_scale)
    t_setup = time.perf_counter()
    ctxs = [A.Context(s.scene_ptr) for s, _ in items]
    setup = {"ctx_create_s": round(time.perf_counter() - t_setup, 4), "scenes": len(ctxs)}
    bufs = [H.DeviceBuffer(d.width * d.height * A.ACCUM_DTYPE.itemsize) for _, d in items]
    stream = H.Stream()
    total = sum(d.width * d.height * d.spp for _, d in items)
    nsteps = args.batch_steps

    def copy_desc(d, **kw):
        w = A.map_preset(0)
        for f, _ in w._fields_:
            setattr(w, f, getattr(d, f))
        for k, v in kw.items():
            setattr(w, k, v)
        return w

    # counting warm-up per map (4 realisations per pixel): byte rates
    bytes_total = 0.0
    for c, (_, d), b in zip(ctxs, items, bufs):
        b.zero()
        c.set_counting(True)
        c.render_accum(copy_desc(d, spp=4), b.address, stream.h.value)
        stream.synchronize()
        k = c.counters()
        c.set_counting(False)
        rate = (8.0 * k["nleaf"] + 8.0 * k["ncoll"] + 32.0 * k["nbvh"] + 36.0 * k["ntri"]) / max(
            k["nrealisations"], 1)
        bytes_total += rate * d.width * d.height * d.spp + 64.0 * d.width * d.height * nsteps
    print(f"[bench] batch: {len(items)} maps, {total / 1e9:.3f} G realisations, {nsteps} steps",
          file=sys.stderr, flush=True)
    assert all(d.spp % (4 * nsteps) == 0 for _, d in items)
    step_descs = [copy_desc(d, spp=d.spp // nsteps) for _, d in items]
    step_items = [(c, d, b.address) for c, d, b in zip(ctxs, step_descs, bufs)]
    tuned = None
    if args.launch_params:
        vals = [int(x) for x in args.launch_params.split(",")]
        for c in ctxs:
            c.set_wavefront_params(*vals)
        tuned = {"params": ctxs[0].wavefront_params(), "trials": 0, "fixed": True}
    elif args.autotune_ms > 0:
        # trials: the whole batch at 1/8 of each map's realisations per pixel
        budget = 4.0 * args.autotune_ms
        params, rate, ntrials = A.batch_autotune(step_items, budget, spp_div=8)
        tuned = {"params": params, "trials": ntrials, "budget_ms": budget,
                 "trial_rate_mreal_s": round(rate / 1e6, 3)}
        print(f"[bench] autotune: {params} after {ntrials} trials ({rate / 1e6:.2f} Mreal/s)",
              file=sys.stderr, flush=True)
    plan = A.BatchPlan(step_items, use_graph=not args.batch_no_graph)
    # warm-up: one step of a plan of the smallest map
    small = min(range(len(items)), key=lambda i: items[i][1].width * items[i][1].height)
    warm = A.BatchPlan([(ctxs[small], copy_desc(items[small][1], spp=4), bufs[small].address)],
                       use_graph=not args.batch_no_graph)
    warm.step(stream.h.value)
    stream.synchronize()
    warm.close()
    for b in bufs:
        b.zero()
    e0, e1 = H.Event(), H.Event()
    H.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(nsteps):
        plan.step(stream.h.value)
    e1.record(stream)
    e1.synchronize()
    batch_s = e0.elapsed_ms(e1) / 1e3
    wall = time.perf_counter() - t0
    scene_ms = plan.scene_ms()
    plan.close()
    n_done = sum(int(b.download(np.zeros(d.width * d.height, A.ACCUM_DTYPE))["n"].sum())
                 for (_, d), b in zip(items, bufs))
    assert n_done == total, (n_done, total)
    print(f"[bench] batch: {nsteps} graph steps {batch_s:.2f} s", file=sys.stderr, flush=True)
    seq_s = None
    if args.batch_sequential:
        for b in bufs:
            b.zero()
        e0.record(stream)
        for c, (_, d), b in zip(ctxs, items, bufs):
            c.render_accum(d, b.address, stream.h.value)
        e1.record(stream)
        e1.synchronize()
        seq_s = e0.elapsed_ms(e1) / 1e3
        print(f"[bench] batch: back-to-back launches {seq_s:.2f} s", file=sys.stderr, flush=True)
    achieved = bytes_total / batch_s / 1e9
    out = {
        "metric": METRIC,
        "value": round(total / batch_s / 1e6, 4),
        "unit": "Mrealisations/s",
        "n_gpus": 1, "steps": nsteps, "warmup": 1,
        "ms_per_step": round(batch_s * 1e3 / nsteps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, seed %d)" % args.seed,
        "config": {
            "workload": "uneven batch: 64 SW scenes, cloud grids 32^2x32 to 256^2x96, images 64^2 to "
                        "512^2, 16-1024 realisations/px (x%g), 1-16 k-points, %d step(s) of one "
                        "persistent launch each%s" % (args.batch_scale, nsteps,
                                                     "" if args.batch_no_graph else
                                                     ", replayed from a HIP graph"),
            "config_index": 3, "realisations": total, "wall_s": round(wall, 3),
            "autotune": tuned,
        },
        "scene_completion_ms": [round(float(x), 1) for x in scene_ms],
        "back_to_back_launches_s": None if seq_s is None else round(seq_s, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None},
        "truncated_queries": {"collision_budget": sum(c.counters()["ntrunc_coll"] for c in ctxs),
                              "iterator_bound": sum(c.counters()["ntrunc_iter"] for c in ctxs)},
        "setup": setup,
        "cpu_baseline": None,
    }
    tfile = os.path.join(ROOT, "profiles", PROFILED_BATCH)
    if os.path.exists(tfile) and args.batch_scale == 1.0 and nsteps == 1:
        with open(tfile) as f:
            prof = json.load(f)
        r = out["roofline"]
        r["traffic"] = prof.get("bytes_per_launch")
        for k in ("limiter", "valu_busy", "valu_lane_utilisation", "l2_hit_rate", "source"):
            r[{"limiter": "measured_limiter", "source": "counters_source"}.get(k, k)] = prof.get(k)
    if not args.no_cpu:
        big = max(range(len(items)), key=lambda i: items[i][1].width * items[i][1].height)
        out["cpu_baseline"] = cpu_baseline(items[big][0], items[big][1], args.cpu_seconds)
    for c in ctxs:
        c.close()
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.config == 3:
        if args.gpus > 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
            sys.exit("bench.py runs on one GPU only; use --gpus 1")
        bench_batch(args)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        sys.exit("bench.py runs on one GPU only: the library has no multi-GPU path "
                 "(pixels would shard across GPUs by Morton tiles); use --gpus 1")

    import numpy as np
    from htrdr_kmin_amd import _ffi as F
    from htrdr_kmin_amd import atmosphere as A
    from htrdr_kmin_amd import hip as H

    synth = A.Synth(args.config, seed=args.seed)
    desc = A.map_preset(args.config)
    if args.width:
        desc.width = desc.height = args.width
    desc.seed = args.seed
    npix = desc.width * desc.height
    S = args.spp_per_step
    # scene set-up, outside the timed region: the octree build (on the device),
    # a mesh ground's BVH build (host) and the uploads to HBM
    t_setup = time.perf_counter()
    ctx = A.Context(synth.scene_ptr)
    setup = {"ctx_create_s": round(time.perf_counter() - t_setup, 4)}
    nl, _, nf = ctx.svo_info()
    setup.update({"octree_levels": nl, "octree_leaves": nf})
    acc = H.DeviceBuffer(npix * A.ACCUM_DTYPE.itemsize)
    stream = H.Stream()

    def run(solver, nsteps, acc, count):
        """Warm-up (optionally counting) then nsteps timed steps into acc."""
        d = A.map_preset(args.config)
        for f, _ in d._fields_:
            setattr(d, f, getattr(desc, f))
        d.solver = solver

        def step(k):
            d.spp = S
            d.spp_offset = k * S
            ctx.render_accum(d, acc.address, stream.h.value)

        acc.zero()
        counters = None
        t_w = time.perf_counter()
        ctx.set_counting(count)
        step(0)
        stream.synchronize()
        if count:
            counters = ctx.counters()
        ctx.set_counting(False)
        print(f"[bench] {SOLVER_NAMES[solver]}: warm-up step 0 {time.perf_counter() - t_w:.2f} s",
              file=sys.stderr, flush=True)
        for w in range(1, args.warmup):
            step(w)
            stream.synchronize()
        acc.zero()
        plan = None
        if solver == F.SOLVER_WAVEFRONT_FP32 and not args.no_graph:
            # the timed steps replay the HIP graph of a one-map plan: zeroed
            # chunk queues, the persistent kernel, the chunk merge and the
            # move to the next slice of realisations
            dp = A.map_preset(args.config)
            for f, _ in dp._fields_:
                setattr(dp, f, getattr(d, f))
            dp.spp, dp.spp_offset = S, 0
            plan = A.BatchPlan([(ctx, dp, acc.address)], use_graph=True)
        ev = [(H.Event(), H.Event()) for _ in range(nsteps)]
        H.synchronize()
        t0 = time.perf_counter()
        for k in range(nsteps):
            ev[k][0].record(stream)
            if plan is not None:
                plan.step(stream.h.value)
            else:
                step(k)
            ev[k][1].record(stream)
            ev[k][1].synchronize()
            print(f"[bench] {SOLVER_NAMES[solver]}: step {k}: {ev[k][0].elapsed_ms(ev[k][1]):.1f} ms",
                  file=sys.stderr, flush=True)
        stream.synchronize()
        H.synchronize()
        wall = time.perf_counter() - t0
        total_s = sum(a.elapsed_ms(b) for a, b in ev) / 1e3
        if plan is not None:
            plan.close()
        h = acc.download(np.zeros(npix, A.ACCUM_DTYPE))
        return total_s, wall, h, counters

    def t05(h, total_s, nreal):
        """time for 99% of the pixels with E > 0 to reach a 0.5% relative SE"""
        E, SE = A.accum_estimate(h)
        n = h["n"].astype(np.float64)
        pos = E > 0
        rel_var = (SE[pos] ** 2 * n[pos]) / E[pos] ** 2  # per-realisation (sigma/E)^2
        nreq = rel_var / 0.005 ** 2
        spp_req = float(np.quantile(nreq, 0.99)) if nreq.size else float("nan")
        return spp_req * npix * total_s / nreal, spp_req, float(np.mean(rel_var)), E, SE

    main_solver = F.SOLVER_WAVEFRONT_FP32 if args.solver == "wavefront" else F.SOLVER_PLAIN_FP64
    tuned = None
    if main_solver == F.SOLVER_WAVEFRONT_FP32 and args.launch_params:
        ctx.set_wavefront_params(*(int(x) for x in args.launch_params.split(",")))
        tuned = {"params": ctx.wavefront_params(), "trials": 0, "fixed": True}
    elif main_solver == F.SOLVER_WAVEFRONT_FP32 and args.autotune_ms > 0:
        dt = A.map_preset(args.config)
        for f, _ in dt._fields_:
            setattr(dt, f, getattr(desc, f))
        dt.solver = main_solver
        dt.spp = S  # the launches of the run
        params, rate, ntrials = ctx.autotune(dt, args.autotune_ms)
        tuned = {"params": params, "trials": ntrials, "budget_ms": args.autotune_ms,
                 "trial_rate_mreal_s": round(rate / 1e6, 3)}
        print(f"[bench] autotune: {params} after {ntrials} trials ({rate / 1e6:.2f} Mreal/s)",
              file=sys.stderr, flush=True)
    total_s, wall, h, counters = run(main_solver, args.steps, acc, True)
    trunc = ctx.counters()
    real = npix * S * args.steps
    value = real / total_s / 1e6
    t05_main, spp_req, relvar_main, E_main, SE_main = t05(h, total_s, real)

    # roofline of the dominant kernel (one launch = one step): algorithmic
    # bytes per realisation from the kernel's own event counters
    creal = max(counters["nrealisations"], 1)
    leaf_bytes = 8.0 if args.solver == "wavefront" else 12.0
    bytes_real = (leaf_bytes * counters["nleaf"] + 8.0 * counters["ncoll"] + 32.0 * counters["nbvh"]
                  + 36.0 * counters["ntri"]) / creal + 64.0 / S
    avg_launch_s = total_s / args.steps
    achieved = bytes_real * npix * S / avg_launch_s / 1e9
    # rocprofv3 counters of the same launches, summarised under profiles/
    # (named for the kernel and workload they cover)
    traffic, counters_prof = None, {}
    tname, twidth = PROFILED.get((args.solver, args.config), (None, 0))
    tfile = os.path.join(ROOT, "profiles", tname) if tname else None
    if tfile and os.path.exists(tfile) and S == 16 and desc.width == twidth:
        with open(tfile) as f:
            counters_prof = json.load(f)
        traffic = counters_prof.get("bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Mrealisations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_s * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.solver == "wavefront" else "f64",
        "data": "synthetic (seeded scene generator, seed %d)" % args.seed,
        "config": {
            "workload": workload_name(synth, desc, S, args.steps),
            "solver": "FP32 batched wavefront" if args.solver == "wavefront" else "FP64 plain one-path-per-thread",
            "config_index": args.config,
            "spp_per_step": S,
            "step_launch": ("HIP graph replay of a one-map plan (queue reset, persistent kernel, "
                            "chunk merge, offset advance)"
                            if args.solver == "wavefront" and not args.no_graph else "direct launches"),
            "wall_s": round(wall, 3),
            "autotune": tuned,
        },
        "time_to_0p5pct_s": round(t05_main, 3),
        "spp_required_0p5pct_q99": round(spp_req, 1),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": traffic,
            "bytes_per_realisation": round(bytes_real, 2),
            # what the counters of the same launches show limits the kernel
            # (the roofline above prices its algorithmic bytes against HBM)
            "measured_limiter": counters_prof.get("limiter"),
            "valu_busy": counters_prof.get("valu_busy"),
            "valu_lane_utilisation": counters_prof.get("valu_lane_utilisation"),
            "l2_hit_rate": counters_prof.get("l2_hit_rate"),
            "counters_source": counters_prof.get("source"),
            "events_per_realisation": {
                "leaf_lookups": round(counters["nleaf"] / creal, 2),
                "segments": round(counters["nseg"] / creal, 2),
                "collisions": round(counters["ncoll"] / creal, 2),
                "lwc_fetches": round(counters["nlwc"] / creal, 2),
                "scatter_surface_events": round(counters["nevents"] / creal, 3),
            },
        },
        "setup": setup,
        # queries of the FP32 solver cut short by its tentative-collision
        # budget or its iterator's no-progress bound (each would bias the
        # estimate; counted on every launch of the run)
        "truncated_queries": {"collision_budget": trunc["ntrunc_coll"],
                              "iterator_bound": trunc["ntrunc_iter"]},
        "cpu_baseline": None,
    }
    if args.fp64_steps > 0 and main_solver == F.SOLVER_WAVEFRONT_FP32:
        acc64 = H.DeviceBuffer(npix * A.ACCUM_DTYPE.itemsize)
        s64, _, h64, _ = run(F.SOLVER_PLAIN_FP64, args.fp64_steps, acc64, False)
        acc64.free()
        real64 = npix * S * args.fp64_steps
        E64, SE64 = A.accum_estimate(h64)
        # per-pixel agreement at equal sample counts: an FP32 pass (not timed)
        # over as many realisations as the FP64 leg, beyond the timed ones
        d32 = A.map_preset(args.config)
        for f, _ in d32._fields_:
            setattr(d32, f, getattr(desc, f))
        d32.solver = F.SOLVER_WAVEFRONT_FP32
        d32.spp = S * args.fp64_steps
        d32.spp_offset = S * (args.steps + args.warmup)
        acc.zero()
        ctx.render_accum(d32, acc.address, stream.h.value)
        stream.synchronize()
        E32, SE32 = A.accum_estimate(acc.download(np.zeros(npix, A.ACCUM_DTYPE)))
        sig = np.sqrt(SE32 ** 2 + SE64 ** 2)
        m = sig > 0
        z = (E32[m] - E64[m]) / sig[m]
        # Per-realisation variance of each solver pooled over the pixels and
        # normalised by the same precise means (the FP32 run's, many more
        # realisations): sum_p s2_p / sum_p E_p^2 with s2_p the unbiased
        # per-pixel sample variance. Dividing each pixel's s2 by its own
        # noisy 16-realisation mean would bias the FP64 side.
        t_real_main, t_real_64 = total_s / real, s64 / real64
        pos = E_main > 0
        def pooled(hh):
            n = hh["n"].astype(np.float64)
            Eh, SEh = A.accum_estimate(hh)
            s2 = SEh ** 2 * n * n / np.maximum(n - 1, 1)
            return float(s2[pos].sum() / (E_main[pos] ** 2).sum())
        relvar_main, relvar_64 = pooled(h), pooled(h64)
        speedup = (t_real_64 * relvar_64) / (t_real_main * relvar_main)
        out["fp64_plain"] = {
            "value": round(real64 / s64 / 1e6, 4), "unit": "Mrealisations/s",
            "steps": args.fp64_steps, "spp": S * args.fp64_steps,
            # the FP32 run's q99 pixel requirement, scaled by the time per
            # realisation and the pooled variance ratio
            "time_to_0p5pct_s": round(t05_main * speedup, 3),
            "pooled_rel_variance_per_realisation": round(relvar_64, 4),
        }
        out["fp32_pooled_rel_variance_per_realisation"] = round(relvar_main, 4)
        out["time_to_0p5pct_speedup_fp32_vs_fp64"] = round(speedup, 3)
        out["fp32_vs_fp64_per_pixel"] = {
            "spp_each": S * args.fp64_steps,
            "within_3_sigma": round(float(np.mean(np.abs(z) <= 3)), 5),
            "z_mean": round(float(z.mean()), 4), "z_std": round(float(z.std()), 4),
            "image_mean_rel_diff": round(float((E_main.mean() - E64.mean()) / E64.mean()), 6),
        }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(synth, desc, args.cpu_seconds)
    stream.destroy()
    acc.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
