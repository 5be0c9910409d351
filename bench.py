#!/usr/bin/env python3
"""Throughput of the flowVC advection hot path on one MI355X.

Workload (one "step"): the 3-D unstructured pulsatile pipe - cylinder R=1,
L=10 voxelised on 60x60x330 hexes, Kuhn 6-tet split (5.6M tets, 0.98M nodes),
50 velocity slices per period of u_z = 2(1-r^2)(1+0.5 sin 2 pi t) plus swirl
u_theta = 0.1 r (1-r^2) sin(pi z/10) and N(0, 0.01^2) noise - with
  * 4M tracers uniform in the inlet third, Z-order sorted, RK4 dt=1e-3 over one
    period (50 frames x 20 steps; particles that leave keep their slots, and
    the live ones are listed once more than an eighth of the lanes idle), and
  * a 128x128x640 FTLE grid over the same period (flow-map gradient, largest
    Cauchy-Green eigenvalue), outputs copied back to host memory.
Inputs are resident in HBM when the timed region starts (mesh, all 50 slices,
located initial conditions); each step re-launches the tracers and the FTLE
slide from their device-resident initial states.

value = particle-steps/s of the whole step (tracer + FTLE advection).
Printed: one JSON line (bench contract) with roofline and cpu_baseline.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "flowvc-1_amd"))

BYTES_PER_STEP_F32 = 704  # 4 containment tests x (16 B conn + 64 B element frame) + 4 stages x 4 nodes x 2 slices x 12 B
HBM_PEAK = 8.0e12
# VALU issue ceiling of MI355X: 256 CUs x 4 SIMDs, one SIMD-cycle each per
# cycle at the 2.4 GHz peak engine clock (a wave64 VALU instruction occupies
# its SIMD for 4 cycles, FP64 multiplies and FMAs for 8)
VALU_PEAK_SIMD_CYCLES = 256 * 4 * 2.4e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """Every processor this process may run on (the reference runs its loops
    on omp_get_num_procs() threads, ompVC/parameters.c:395)."""
    return len(os.sched_getaffinity(0))


def host_info():
    """nproc and the lscpu model, sockets and threads of the host."""
    info = {"nproc": host_cores()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k] = v
    except Exception:
        pass
    return info


def profile_figure(fname, key, kernel):
    """A figure of `kernel` recorded in a committed counter summary under
    profiles/ (lines `<key> <kernel> <value>`); None when absent."""
    try:
        with open(os.path.join(HERE, "profiles", fname)) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == key and w[1] == kernel:
                    return float(w[2])
    except OSError:
        pass
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the FP32 pipe workload, from the
    committed FETCH_SIZE / WRITE_SIZE counter passes (profiles/pmc_traffic_pipe.md,
    scaled by the copy-kernel calibration there); None when not measured."""
    v = profile_figure("pmc_traffic_pipe.md", "traffic_bytes_per_launch", kernel)
    return None if v is None else int(v)


def roofline_valu(kernel, steps_per_launch, avg_ms, bpp, default_workload):
    """Roofline of the dominant RK4 kernel against the nearest ceiling.
    Counter passes on the FP32 pipe workload (profiles/pmc_valu_pipe.md,
    pmc_l2_pipe.md, pmc_traffic_pipe.md, pmc_wait_pipe.md) put VALU issue
    highest (64-68% busy), L2-to-L1 reads at 28-37% of the L2's bandwidth
    and HBM at 2-4%; the waves wait on their own gathers more than half of
    their cycles (the limiter, reported beside the roofline).  The bound
    reported is VALU issue: achieved = VALU busy SIMD-cycles per
    particle-step (SQ_ACTIVE_INST_VALU x 4 over the particle-steps of the
    profiled launches) x particle-steps of this launch / its measured
    duration, against 1024 SIMDs x 2.4 GHz.  The algorithmic gather bytes
    (704 B per particle-step in FP32) are reported beside it with the
    measured HBM traffic (cache-served: their HBM fraction exceeds 1)."""
    cyc = profile_figure("pmc_valu_pipe.md", "valu_cycles_per_particle_step", kernel) if default_workload else None
    secs = avg_ms * 1e-3
    gather = steps_per_launch * bpp / secs if secs > 0 else 0.0
    out = {"kernel": kernel, "avg_launch_ms": avg_ms, "particle_steps_per_launch": steps_per_launch,
           "gather_bytes_per_particle_step": bpp, "gather_GBps_algorithmic": gather / 1e9,
           "gather_frac_of_hbm_peak": gather / HBM_PEAK,
           "traffic": measured_traffic(kernel) if default_workload else None,
           "traffic_source": "profiles/pmc_traffic_pipe.md (FETCH_SIZE x2.000 + WRITE_SIZE x1.000 per launch)"}
    if cyc is None or secs <= 0:
        out.update(bound="valu", achieved=None, peak=VALU_PEAK_SIMD_CYCLES / 1e9, unit="G SIMD-cycles/s", frac=None)
        return out
    achieved = cyc * steps_per_launch / secs
    out.update(bound="valu", achieved=achieved / 1e9, peak=VALU_PEAK_SIMD_CYCLES / 1e9, unit="G SIMD-cycles/s",
               frac=achieved / VALU_PEAK_SIMD_CYCLES, valu_cycles_per_particle_step=cyc,
               valu_source="profiles/pmc_valu_pipe.md (SQ_ACTIVE_INST_VALU x 4 per particle-step)")
    wait = profile_figure("pmc_wait_pipe.md", "wave_wait_share", kernel)
    if wait is not None:
        out.update(limiter="gather latency: the kernel's waves wait on their loads %.0f%% of their cycles "
                           "(SQ_WAIT_ANY / SQ_WAVE_CYCLES, profiles/pmc_wait_pipe.md)" % (100 * wait))
    return out


def build_inputs(args, host, np, with_seeds=True):
    t0 = time.time()
    tube = host.pipe(args.nx, args.nx, args.nz, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(tube)
    slices = [host.tube_velocity(tube, xyz, s * 0.02, 0.01, 1234, s) for s in range(50)]
    seeds = host.tube_seeds(tube, args.tracers, 0.0, 10.0 / 3.0, 42) if with_seeds else None
    if getattr(args, "mesh_order", "zorder") == "zorder":
        from fvcpy import fvc
        tz = time.time()
        no, _, xyz, conn, adj = fvc.mesh_zorder(xyz, conn, adj)
        slices = [fvc.gather_rows(no, sl) for sl in slices]
        log("mesh renumbered in Z-order on the device (%.2f s with the 50 slices)" % (time.time() - tz))
    log("inputs: %d nodes, %d tets, %d tracers (%.1f s)" % (xyz.shape[0], conn.shape[0],
                                                           0 if seeds is None else seeds.shape[0], time.time() - t0))
    return tube, xyz, conn, adj, slices, seeds


def make_context(fvc, precision, xyz, conn, adj, slices):
    ctx = fvc.Context(3, [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0], precision, 50, coords=xyz, conn=conn, adj=adj)
    for s, sl in enumerate(slices):
        ctx.load_slice(s, s * 0.02, sl)
    return ctx


def element_order(np, recs):
    """Tracer records in the order of the element holding each (stable, so
    ties keep their Z-order; points outside last): the mesh is numbered along
    its own Z-curve, so neighbouring lanes gather from neighbouring element
    frames."""
    key = np.where(recs["ElementIndex"] < 0, 1 << 30, recs["ElementIndex"])
    return recs[np.argsort(key, kind="stable")]


def frames():
    """(tmin, tmax, lo, hi) of the 50 frames of one period (main.c:188-195; the
    last frame's upper slice wraps to file 0, velocity.c:77)."""
    for f in range(50):
        tmin = 0.0 + f * 0.02
        yield tmin, tmin + 0.02, f, (f + 1) % 50


def run_step(P, P_ic, F, ctx, n_tr, res):
    """One pass: tracers over the period, then the FTLE slide; returns counters."""
    out = {}
    t0 = time.perf_counter()
    P.copy_from(P_ic)
    P.launch_velocity(0.0, 0, 1, 0.0, 0.02)
    steps_tr = 0
    for tmin, tmax, lo, hi in frames():
        steps_tr += P.advect(tmin, tmax, lo, hi, tmin, tmax, 1e-3)
    ctx_sync(ctx)
    t1 = time.perf_counter()
    steps_ft = 0
    F.reset()
    for tmin, tmax, lo, hi in frames():
        F.frame(1, lo, hi, tmin, tmax, 1e-3)
        steps_ft += ctx.last_steps()
    wt, wot = F.output(1.0)
    ctx_sync(ctx)
    t2 = time.perf_counter()
    out.update(t_tr=t1 - t0, t_ft=t2 - t1, steps_tr=steps_tr, steps_ft=steps_ft, wt=wt)
    return out


def tracer_periods(P, P_ic, ctx, nper):
    """The tracers over `nper` periods of the periodic data (outside the timed
    steps): seconds, particle-steps and the live-list length (-1: the whole
    range launches) of each period, as particles leave through the outlet."""
    P.copy_from(P_ic)
    P.launch_velocity(0.0, 0, 1, 0.0, 0.02)
    out = []
    for k in range(nper):
        ctx_sync(ctx)
        t0 = time.perf_counter()
        steps = 0
        for tmin, tmax, lo, hi in frames():
            steps += P.advect(tmin + k, tmax + k, lo, hi, tmin + k, tmax + k, 1e-3)
        ctx_sync(ctx)
        secs = time.perf_counter() - t0
        left = int((P.get()["LeftDomain"] != 0).sum())
        out.append({"period": k, "seconds": secs, "particle_steps": int(steps), "left_domain": left,
                    "live_list": P.active()})
        log("tracer period %d: %.3f s, %d particle-steps, %d left, list %d" %
            (k, out[-1]["seconds"], steps, left, out[-1]["live_list"]))
    return out


def ctx_sync(ctx):
    from fvcpy import fvc
    fvc.lib().fvc_sync(ctx.h)


def cpu_baseline(args, np, xyz, conn, adj, slices, recs_sample, threads=None):
    """The oracle (CPU FP64 restatement, ompVC semantics, OpenMP dynamic chunks
    as tracers.c:222-223) on a bounded sample of the tracers, on every host
    core, or on `threads` of them."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle
    cores = threads or host_cores()
    om = oracle.Mesh(3, [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0], coords=xyz, conn=conn, adj=adj)
    pts = np.ascontiguousarray(recs_sample.copy())
    chunk = max(1, pts.shape[0] // cores)
    t_int = 0.0
    steps = 0
    for tmin, tmax, lo, hi in frames():
        om.set_slot(0, slices[lo])
        om.set_slot(1, slices[hi])
        om.set_interval(tmin, tmax)
        t0 = time.perf_counter()
        rc = om.advect(pts, tmin, tmax, 1e-3, nthreads=cores, chunk=chunk)
        t_int += time.perf_counter() - t0
        if rc:
            raise RuntimeError("oracle advect failed (%d)" % rc)
        steps += om.last_steps.value
    return dict(value=steps / t_int, seconds=t_int, cores=cores, steps=steps, pts=pts)


def make_releases(np, fvc, n_rel, lo_seeds, hi_seeds, seed):
    """64 releases: seed counts log-uniform in [lo, hi], start times uniform in
    the period, horizons uniform in [0.1, 1.0] periods, RK4 or AB3."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo_seeds), np.log(hi_seeds), n_rel)).astype(np.int64)
    starts = rng.uniform(0.0, 1.0, n_rel)
    horizons = rng.uniform(0.1, 1.0, n_rel)
    types = rng.choice([fvc.INT_RK4, fvc.INT_AB3], n_rel)
    return counts, starts, horizons, types


def release_table(counts, starts, horizons, types, scale=1):
    rel, b = [], 0
    for c, ts, h, it in zip(counts, starts, horizons, types):
        c = max(1, int(c) // scale)
        rel.append((b, c, float(ts), float(ts + h), int(it)))
        b += c
    return rel


def bench_batch(args, np, fvc, host):
    """Uneven releases fused into one persistent kernel (fvc_batch_*): the
    autotuner picks block size and steps per launch on a 1/8-size copy of the
    batch, then K timed runs of the full batch (reset from the device-resident
    initial state + run, launches replayed from a captured hipGraph)."""
    tube, xyz, conn, adj, slices, _ = build_inputs(args, host, np, with_seeds=False)
    counts, starts, horizons, types = make_releases(np, fvc, args.releases, 1e4, 2e6, 7)
    ctx = make_context(fvc, args.precision, xyz, conn, adj, slices)
    lo, hi = np.array([-1.0, -1.0, 0.0]), np.array([1.0, 1.0, 10.0])
    t0 = time.time()
    recs = []
    for r, c in enumerate(counts):
        sd = host.tube_seeds(tube, int(c), 0.0, 10.0 / 3.0, 1000 + r)
        sd = sd[ctx.zsort(sd, lo, hi)]
        r_recs = ctx.seed_list(sd, 1)
        recs.append(element_order(np, r_recs) if args.batch_order == "element" else r_recs)
    recs = np.concatenate(recs)
    rel = release_table(counts, starts, horizons, types)
    n = recs.shape[0]
    log("batch: %d releases, %d seeds (%d RK4 / %d AB3 releases), setup %.1f s" %
        (len(rel), n, int((types == fvc.INT_RK4).sum()), int((types == fvc.INT_AB3).sum()), time.time() - t0))
    if args.batch_config:
        # fixed block size and steps per launch, no tuning sweep
        bs, bu = (int(v) for v in args.batch_config.split(","))
        best, best_ms = {"block": bs, "steps_per_launch": bu, "use_graph": args.batch_graph, "waves_per_cu": 0}, None
        log("batch configuration given: %s (no autotune)" % best)
    else:
        # autotune on a 1/8 copy (same releases, an eighth of the seeds each)
        rel_t = release_table(counts, starts, horizons, types, scale=8)
        idx_t = np.concatenate([np.arange(b0, b0 + rt[1]) for (b0, _, _, _, _), rt in zip(rel, rel_t)])
        Pt = fvc.Particles(ctx, idx_t.size)
        Pt.set(recs[idx_t])
        Bt = fvc.Batch(ctx, Pt, rel_t, 0.0, 0.02, 50, 1e-3)
        t0 = time.time()
        blocks = [64, 128, 256, 512, 1024] if args.precision == "f32" else [256]
        budgets = [1, 4, 16, 64, 256, 0]
        best, best_ms = Bt.tune(blocks, budgets, args.batch_graph)
        Bt.close()
        del Pt
        log("autotune %.1f s over blocks %s x budgets %s: best %s (%.1f ms on the 1/8 batch)" %
            (time.time() - t0, blocks, budgets, best, best_ms))
    P = fvc.Particles(ctx, n)
    P.set(recs)
    B = fvc.Batch(ctx, P, rel, 0.0, 0.02, 50, 1e-3)
    for w in range(args.warmup):
        B.reset()
        st, nl = B.run(**best)
        log("warmup %d: %d particle-steps, %d launches" % (w, st, nl))
    fvc.lib().fvc_sync(ctx.h)
    tot_steps = 0
    launches = 0
    wall0 = time.perf_counter()
    for k in range(args.steps):
        B.reset()
        st, nl = B.run(**best)
        tot_steps += st
        launches = nl
    fvc.lib().fvc_sync(ctx.h)
    wall = time.perf_counter() - wall0
    # device time of one more run (events around all of its kernels)
    B.reset()
    ctx.kernel_time(reset=True)
    ctx.set_timing(True)
    B.run(**best)
    k_ms, _, k_steps = ctx.kernel_time(reset=True)
    ctx.set_timing(False)
    got = P.get()
    K = args.steps
    value = tot_steps / wall
    bpp = BYTES_PER_STEP_F32 if args.precision == "f32" else (512 if args.precision in ("bf16", "f16") else 1328)
    result = {
        "metric": "particle-steps/s + FTLE seeds/s at stated FP64 error; p50/p99 request latency",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"f32": "f32 (fp64 position/time state)", "bf16": "bf16 storage, f32 math",
                  "f16": "f16 storage, f32 math", "f64": "f64"}[args.precision],
        "data": "synthetic (deterministic generators, seeded)",
        "config": {
            "workload": "batching on the pipe mesh (%d tets): %d releases in one persistent kernel, seed counts "
                        "log-uniform [10k, 2M] (%d seeds), start uniform in the period, horizon uniform [0.1, 1] "
                        "periods, RK4/AB3 mix, dt=1e-3, autotuned block/steps-per-launch, %s" %
                        (conn.shape[0], len(rel), n, "hipGraph replay" if args.batch_graph else "direct launches"),
            "releases": len(rel), "seeds": n, "precision": args.precision,
            "particle_order": "each release's seeds by the element holding each" if args.batch_order == "element"
            else "each release's seeds in Z-order",
        },
        "batch": {"tuned": best, "tune_ms_eighth_batch": best_ms, "launch_groups": launches,
                  "particle_steps_per_run": tot_steps // K},
        "roofline": dict(roofline_valu("k_batch", k_steps, k_ms, bpp, args.precision == "f32"),
                         kernel="k_batch (whole run: batch kernels + compaction, device time)"),
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle
        cores = host_cores()
        om = oracle.Mesh(3, [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0], coords=xyz, conn=conn, adj=adj)
        stride = max(1, args.batch_cpu_stride)

        def frame_of(t):
            f = int(np.floor(t / 0.02))
            while f * 0.02 > t:
                f -= 1
            while (f + 1) * 0.02 <= t:
                f += 1
            return f

        # frame-major sweep (slots set once per frame), every release advanced
        # over its part of each frame exactly as the batch does
        st = []
        for (b0, c, ts, te, it) in rel:
            idx = np.arange(b0, b0 + c, stride)
            f0 = frame_of(ts)
            f1 = f0
            while te - (f1 + 1) * 0.02 >= 1e-10:
                f1 += 1
            st.append(dict(idx=idx, pts=np.ascontiguousarray(recs[idx].copy()), f0=f0, f1=f1, ts=ts, te=te, it=it,
                           hist=np.zeros(idx.size, oracle.HIST_DTYPE)))
        t_int, steps = 0.0, 0
        for f in range(min(r["f0"] for r in st), max(r["f1"] for r in st) + 1):
            tmin = f * 0.02
            tmax = tmin + 0.02
            om.set_slot(0, slices[f % 50])
            om.set_slot(1, slices[(f + 1) % 50])
            om.set_interval(tmin, tmax)
            for r in st:
                if not (r["f0"] <= f <= r["f1"]):
                    continue
                pts = r["pts"]
                if f == r["f0"]:
                    live = np.nonzero(pts["LeftDomain"] == 0)[0]
                    sub = pts[live].copy()
                    v, _ = om.velocity(sub, r["ts"])
                    sub["V"] = v
                    pts[live] = sub
                a, bb = max(r["ts"], tmin), min(r["te"], tmax)
                t1 = time.perf_counter()
                if r["it"] == fvc.INT_AB3:
                    rc, ns = oracle.ab_advect(om, pts, r["hist"], a, bb, 1e-3, 3, nthreads=cores)
                else:
                    rc = om.advect(pts, a, bb, 1e-3, nthreads=cores, chunk=max(1, pts.shape[0] // (4 * cores)))
                    ns = om.last_steps.value
                t_int += time.perf_counter() - t1
                steps += ns
                if rc:
                    raise RuntimeError("oracle failed (%d)" % rc)
        errs = []
        for r in st:
            g, pts = got[r["idx"]], r["pts"]
            ok = (g["LeftDomain"] == 0) & (pts["LeftDomain"] == 0) & (pts["LeftDomainTime"] > -1e-10)
            if ok.any():
                errs.append(np.abs(g["X"][ok] - pts["X"][ok]).max(axis=1))
        err = np.concatenate(errs) if errs else np.zeros(1)
        result["cpu_baseline"] = {
            "value": steps / t_int, "unit": "particle-steps/s", "cores": cores, "kind": "port", "host": host_info(),
            "sample": "every %d-th seed of each release, same releases/horizons/integrators, integrate-only "
                      "time %.1f s, oracle FP64 restatement (OpenMP)" % (stride, t_int)}
        result["pos_err_vs_fp64"] = {"max": float(err.max()), "p99": float(np.percentile(err, 99)),
                                     "tolerance": 1e-4, "domain_size": 10.0, "points": int(err.size)}
        result["speedup_vs_cpu"] = value / (steps / t_int)
    print(json.dumps(result), flush=True)


def serve_requests(np, host, tube, rate, window, seed):
    """Poisson arrivals at `rate`/s over `window` s; each an FTLE or tracer job
    with seed count log-uniform in [1k, 4M] and horizon uniform in [0.05, 1]
    period, starting anywhere in the period."""
    rng = np.random.default_rng(seed)
    lo = np.array([tube.lo[0], tube.lo[1], tube.lo[2]])
    hi = np.array([tube.hi[0], tube.hi[1], tube.hi[2]])
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > window:
            break
        kind = "ftle" if rng.uniform() < 0.5 else "tracer"
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(4e6))))
        ts = float(rng.uniform(0.0, 1.0))
        te = ts + float(rng.uniform(0.05, 1.0))
        if kind == "tracer":
            spec = host.tube_seeds(tube, n, lo[2], hi[2], int(rng.integers(1 << 30)))
        else:
            ext = hi - lo
            h = (ext.prod() / n) ** (1.0 / 3.0)
            res = [max(2, int(round(e / h))) for e in ext]
            spec = (lo.tolist(), hi.tolist(), res)
        out.append(dict(t=t, kind=kind, spec=spec, ts=ts, te=te))
    return out


def bench_serve(args, np, fvc, host, window=None, emit=True):
    """Served mode on the Y-bifurcation: requests arrive by a Poisson process
    and are admitted into the server's HBM pool as memory frees up; p50/p99
    latency (arrival to result) and aggregate particle-steps/s, beside the
    batch-at-a-time path (each request alone, first come first served, the
    same arithmetic: results are bit-identical).  `window` overrides the
    arrival window; with emit=False the result is returned, not printed."""
    from concurrent.futures import ThreadPoolExecutor
    window = args.serve_window if window is None else window
    t0 = time.time()
    tube = host.ybif(184, 65, 256)
    xyz, conn, adj = host.tube_mesh(tube)
    with ThreadPoolExecutor(16) as ex:
        slices = list(ex.map(lambda s: host.tube_velocity(tube, xyz, s * 0.01, 0.01, 4321, s), range(100)))
    bounds = [float(xyz[:, 0].min()), float(xyz[:, 0].max()), float(xyz[:, 1].min()), float(xyz[:, 1].max()),
              float(xyz[:, 2].min()), float(xyz[:, 2].max())]
    ctx = fvc.Context(3, bounds, args.serve_precision, 100, coords=xyz, conn=conn, adj=adj)
    for s, sl in enumerate(slices):
        ctx.load_slice(s, s * 0.01, sl)
    del slices
    reqs = serve_requests(np, host, tube, args.serve_rate, window, 11)
    log("serve setup %.1f s: %d nodes, %d tets, 100 %s slices, %d requests over %.0f s" %
        (time.time() - t0, xyz.shape[0], conn.shape[0], args.serve_precision, len(reqs), window))
    S = fvc.Server(ctx, 0.0, 0.01, 100, 1e-3, hbm_fraction=0.6, steps_per_launch=args.serve_budget,
                   use_graph=not args.serve_no_graph)
    cap, pool_bytes = S.capacity()

    def submit(r):
        if r["kind"] == "tracer":
            return S.submit_tracer(r["spec"], r["ts"], r["te"], fvc.INT_RK4)
        mn, mx, res = r["spec"]
        return S.submit_ftle(mn, mx, res, r["ts"], r["te"])

    # warm-up: one small request of each kind (code-object loading)
    for r in (dict(kind="tracer", spec=reqs[0]["spec"][:1000] if reqs[0]["kind"] == "tracer" else
                   host.tube_seeds(tube, 1000, -3.0, 3.0, 5), ts=0.1, te=0.15),
              dict(kind="ftle", spec=([-1, -1, -3], [1, 1, -1], [8, 8, 8]), ts=0.1, te=0.15)):
        rid = submit(r)
        S.run_until_idle()
        S.take(rid)
    ids = [None] * len(reqs)
    i = 0
    peak_running = 0
    ticks, t_poll = 0, 0.0
    clock0 = S.clock()
    while True:
        now = S.clock() - clock0
        while i < len(reqs) and reqs[i]["t"] <= now:
            ids[i] = submit(reqs[i])
            i += 1
        tp = time.perf_counter()
        r, q = S.poll()
        t_poll += time.perf_counter() - tp
        ticks += 1
        peak_running = max(peak_running, r)
        if r == 0 and q == 0:
            if i == len(reqs):
                break
            time.sleep(max(0.0, min(0.01, reqs[i]["t"] - (S.clock() - clock0))))
    sstats = S.stats()
    lat, steps, t_end = [], 0, 0.0
    served = {}
    for k, rid in enumerate(ids):
        inf = S.info(rid)
        if inf["state"] != fvc.RS_DONE:
            raise RuntimeError("request %d ended in state %d" % (k, inf["state"]))
        lat.append(inf["t_done"] - clock0 - reqs[k]["t"])
        steps += inf["particle_steps"]
        t_end = max(t_end, inf["t_done"] - clock0)
        res = S.take(rid)
        if k < args.serve_check:
            served[k] = res
    S.close()
    lat = np.array(lat)
    makespan = t_end - reqs[0]["t"]
    log("served %d requests: makespan %.1f s, p50 %.1f ms, p99 %.1f ms, peak running %d" %
        (len(reqs), makespan, 1e3 * np.median(lat), 1e3 * np.percentile(lat, 99), peak_running))
    # batch-at-a-time: every request alone through the plain path, FIFO
    service = []
    same = 0
    for k, r in enumerate(reqs if args.serve_baseline else []):
        t1 = time.perf_counter()
        if r["kind"] == "tracer":
            recs = ctx.seed_list(r["spec"], 1)
            P = fvc.Particles(ctx, max(1, recs.shape[0]))
            P.set(recs)
            B = fvc.Batch(ctx, P, [(0, recs.shape[0], r["ts"], r["te"], fvc.INT_RK4)], 0.0, 0.01, 100, 1e-3)
            B.run(256, 0, 0)
            out = P.get(recs.shape[0])
            B.close()
            P.close()
        else:
            mn, mx, res = r["spec"]
            F = fvc.FtleSlide(ctx, mn, mx, res, 1, r["ts"], r["te"])
            f = int(np.floor(r["ts"] / 0.01))
            while f * 0.01 > r["ts"]:
                f -= 1
            while (f + 1) * 0.01 <= r["ts"]:
                f += 1
            while F.frame(1, f % 100, (f + 1) % 100, f * 0.01, f * 0.01 + 0.01, 1e-3) != 2:
                f += 1
            out = F.output(r["te"] - r["ts"])
            F.close()
        service.append(time.perf_counter() - t1)
        if k in served:
            a = served[k]
            if r["kind"] == "tracer":
                same += int(np.array_equal(a["X"], out["X"]) and np.array_equal(a["ElementIndex"],
                                                                                out["ElementIndex"]) and
                            np.array_equal(a["LeftDomainTime"], out["LeftDomainTime"]))
            else:
                same += int(np.array_equal(a[0], out[0]) and np.array_equal(a[1], out[1]))
    svc_by_kind = {kd: float(sum(sv for r, sv in zip(reqs, service) if r["kind"] == kd)) for kd in ("tracer", "ftle")}
    fin, blat = 0.0, []
    for r, sv in zip(reqs, service):
        fin = max(fin, r["t"]) + sv
        blat.append(fin - r["t"])
    blat = np.array(blat if blat else [float("nan")])
    value = steps / makespan
    result = {
        "metric": "particle-steps/s + FTLE seeds/s at stated FP64 error; p50/p99 request latency",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": len(reqs), "warmup": 2,
        "ms_per_step": 1e3 * makespan / max(1, len(reqs)), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "%s storage, f32 math (fp64 position/time state)" % args.serve_precision,
        "data": "synthetic (deterministic generators, seeded)",
        "config": {
            "workload": "serving on a Y-bifurcation (parent R=1, daughters 0.7 at +/-30 deg; %d tets, %d nodes), "
                        "100 %s slices resident, Poisson arrivals %.0f/s for %.0f s (%d requests), FTLE or tracer, "
                        "seeds log-uniform [1k, 4M], horizon uniform [0.05, 1] period, RK4 dt=1e-3" %
                        (conn.shape[0], xyz.shape[0], args.serve_precision, args.serve_rate, window,
                         len(reqs)),
            "requests": len(reqs), "rate_per_s": args.serve_rate, "steps_per_launch": args.serve_budget,
        },
        "p50_latency_ms": 1e3 * float(np.median(lat)),
        "p99_latency_ms": 1e3 * float(np.percentile(lat, 99)),
        "batch_at_a_time": {"p50_latency_ms": 1e3 * float(np.median(blat)),
                            "p99_latency_ms": 1e3 * float(np.percentile(blat, 99)),
                            "busy_s": float(sum(service)), "busy_s_by_kind": svc_by_kind,
                            "throughput_particle_steps_per_s": steps / max(fin - reqs[0]["t"], 1e-9)}
        if args.serve_baseline else None,
        "served": {"makespan_s": makespan, "peak_running": peak_running, "ticks": ticks, "poll_s": t_poll,
                   "host_phases_s": sstats,
                   "latency_p50_ms_by_kind": {kd: 1e3 * float(np.median([l for l, r in zip(lat, reqs)
                                                                         if r["kind"] == kd] or [0]))
                                              for kd in ("tracer", "ftle")},
                   "pool_particles": cap,
                   "pool_gb": pool_bytes / 1e9, "results_identical_to_alone": "%d/%d" % (same, len(served))},
        "roofline": None,
    }
    ctx.close()
    if emit:
        print(json.dumps(result), flush=True)
    return result


def bench_traffic(args, np, fvc, host):
    """Short run for hardware-counter passes: a plain device copy of 1 GiB
    (k_copy_probe: reads and writes 1 GiB each, the calibration of the byte
    counters) followed by `--frames` frames of the pipe tracer pass (k_advect)
    and of the FTLE slide (k_ftle_advect) on the full mesh, tracer set and
    grid."""
    tube, xyz, conn, adj, slices, seeds = build_inputs(args, host, np)
    ctx = make_context(fvc, args.precision, xyz, conn, adj, slices)
    lo, hi = np.array([-1.0, -1.0, 0.0]), np.array([1.0, 1.0, 10.0])
    seeds = seeds[ctx.zsort(seeds, lo, hi)]
    recs = element_order(np, ctx.seed_list(seeds, 1))
    P = fvc.Particles(ctx, recs.shape[0])
    P.set(recs)
    ctx.copy_probe(1 << 30)
    P.launch_velocity(0.0, 0, 1, 0.0, 0.02)
    steps = 0
    for k, (tmin, tmax, lo_, hi_) in enumerate(frames()):
        if k >= args.frames:
            break
        steps += P.advect(tmin, tmax, lo_, hi_, tmin, tmax, 1e-3)
    res = [int(v) for v in args.ftle.split(",")]
    F = fvc.FtleSlide(ctx, [-1.0, -1.0, 0.0], [1.0, 1.0, 10.0], res, 1, 0.0, 1.0)
    fsteps = 0
    for k, (tmin, tmax, lo_, hi_) in enumerate(frames()):
        if k >= args.frames:
            break
        F.frame(1, lo_, hi_, tmin, tmax, 1e-3)
        fsteps += ctx.last_steps()
    fvc.lib().fvc_sync(ctx.h)
    print(json.dumps({"workload": "traffic", "copy_bytes": 1 << 30, "frames": args.frames,
                      "tracers": int(recs.shape[0]), "particle_steps": int(steps),
                      "particle_steps_per_launch": steps / max(1, args.frames), "ftle_points": F.n,
                      "ftle_particle_steps_per_launch": fsteps / max(1, args.frames)}), flush=True)


def dg_slide(fvc, np, ctx, res, T=10.0):
    """One forward FTLE slide of the double gyre over [0, T]: 2T frames of 0.5
    between consecutive slices, RK4 dt=0.01.  Returns the slide, FTLE (with
    integration time T), particle-steps and seconds."""
    F = fvc.FtleSlide(ctx, [0.0, 0.0, 0.0], [2.0, 1.0, 0.0], res, 1, 0.0, T)
    fvc.lib().fvc_sync(ctx.h)
    t0 = time.perf_counter()
    steps = 0
    for f in range(int(round(2 * T))):
        F.frame(1, f, f + 1, 0.5 * f, 0.5 * (f + 1), 0.01)
        steps += ctx.last_steps()
    wt, _ = F.output(T)
    fvc.lib().fvc_sync(ctx.h)
    return F, wt, steps, time.perf_counter() - t0


def bench_dg(args, np, fvc, host):
    """The CPU-runnable configuration and its precision variants: double gyre
    (A=0.1, eps=0.25, omega=2 pi/10) on 256x128 nodes, 21 slices over [0, 10];
    FTLE on a 512x256 grid (and a 1024x1024 grid, the 1M-seed comparison set)
    with RK4 dt=0.01 over T=10.  FP64 first (bit-identical to the oracle), then
    FP32, BF16 and FP16 slices; errors against FP64: FTLE relative error,
    final flow-map positions, velocity-grid cell and exit agreement."""
    nx, ny = 256, 128
    t0 = time.time()
    slices = [host.double_gyre_slice(nx, ny, 0.5 * s) for s in range(21)]
    log("double gyre slices: %.1f s" % (time.time() - t0))
    out = {"workload": "dg", "config": {"workload": "double gyre 256x128 nodes, 21 slices, RK4 dt=0.01 over T=10",
                                        "ftle_grids": [[512, 256], [1024, 1024]]}, "results": []}
    dx, dy = 2.0 / (nx - 1), 1.0 / (ny - 1)
    for res in ([512, 256, 1], [1024, 1024, 1]):
        ref = None
        for prec in ("f64", "f32", "bf16", "f16"):
            ctx = fvc.Context(2, [0.0, 2.0, 0.0, 1.0, 0.0, 0.0], prec, 21, cart_min=[0.0, 0.0, 0.0],
                              cart_max=[2.0, 1.0, 0.0], cart_res=[nx, ny, 1])
            for s, sl in enumerate(slices):
                ctx.load_slice(s, 0.5 * s, sl)
            dg_slide(fvc, np, ctx, res)  # warm-up
            F, wt, steps, secs = dg_slide(fvc, np, ctx, res)
            rec = F.records()
            r = {"grid": res[:2], "precision": prec, "seeds": F.n, "seconds": secs, "particle_steps": int(steps),
                 "particle_steps_per_s": steps / secs, "ftle_seeds_per_s": F.n / secs}
            if prec == "f64":
                ref = (wt, rec)
            else:
                w64, r64 = ref
                big = np.abs(w64) > 1e-3
                rel = np.abs(wt[big] - w64[big]) / np.abs(w64[big])
                both = (rec["LeftDomain"] == 0) & (r64["LeftDomain"] == 0)
                dpos = np.linalg.norm(rec["X"][both] - r64["X"][both], axis=1)
                cell = lambda X: np.stack([np.floor(X[:, 0] / dx), np.floor(X[:, 1] / dy)], 1)
                r.update(ftle_rel_p50=float(np.percentile(rel, 50)), ftle_rel_p99=float(np.percentile(rel, 99)),
                         ftle_rel_p999=float(np.percentile(rel, 99.9)), ftle_rel_max=float(rel.max()),
                         ftle_abs_max=float(np.abs(wt - w64).max()), pos_err_max=float(dpos.max()),
                         pos_err_p99=float(np.percentile(dpos, 99)),
                         cell_agreement=float((cell(rec["X"][both]) == cell(r64["X"][both])).all(1).mean()),
                         exit_agreement=float((rec["LeftDomain"] == r64["LeftDomain"]).mean()))
            out["results"].append(r)
            log(json.dumps(r))
            F.close()
            ctx.close()
    if not args.no_cpu:  # the oracle on every 16th point of the 512x256 grid, all 20 frames
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle
        om = oracle.Mesh(2, [0.0, 2.0, 0.0, 1.0, 0.0, 0.0], cart_min=[0.0, 0.0, 0.0], cart_max=[2.0, 1.0, 0.0],
                         cart_res=[nx, ny, 1])
        g = np.stack(np.meshgrid(np.linspace(0.0, 2.0, 512), np.linspace(0.0, 1.0, 256), indexing="ij"), -1)
        pts = np.zeros(512 * 256 // 16, oracle.POINT_DTYPE)
        pts["X"][:, :2] = g.reshape(-1, 2)[::16]
        cores = host_cores()
        t_int, steps = 0.0, 0
        for f in range(20):
            om.set_slot(0, slices[f])
            om.set_slot(1, slices[f + 1])
            om.set_interval(0.5 * f, 0.5 * (f + 1))
            t1 = time.perf_counter()
            rc = om.advect(pts, 0.5 * f, 0.5 * (f + 1), 0.01, nthreads=cores, chunk=max(1, pts.shape[0] // cores))
            t_int += time.perf_counter() - t1
            if rc:
                raise RuntimeError("oracle advect failed (%d)" % rc)
            steps += om.last_steps.value
        out["cpu_baseline"] = {"value": steps / t_int, "unit": "particle-steps/s", "cores": cores, "kind": "port",
                               "host": host_info(),
                               "sample": "%d of the 512x256 FTLE grid points (every 16th), T=10, FP64 oracle"
                                         % pts.shape[0]}
        # ompVC as configured for this deck: its FTLE loop parallelises over
        # output slides (UPPERLOOP, parameters.c:400-401, ftle.c:68), and one
        # slide runs on one thread
        pts1 = np.zeros(512 * 256 // 64, oracle.POINT_DTYPE)
        pts1["X"][:, :2] = g.reshape(-1, 2)[::64]
        t_one, steps_one = 0.0, 0
        for f in range(20):
            om.set_slot(0, slices[f])
            om.set_slot(1, slices[f + 1])
            om.set_interval(0.5 * f, 0.5 * (f + 1))
            t1 = time.perf_counter()
            rc = om.advect(pts1, 0.5 * f, 0.5 * (f + 1), 0.01, nthreads=1, chunk=pts1.shape[0])
            t_one += time.perf_counter() - t1
            if rc:
                raise RuntimeError("oracle advect failed (%d)" % rc)
            steps_one += om.last_steps.value
        out["cpu_baseline_as_configured"] = {
            "value": steps_one / t_one, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "%d of the 512x256 FTLE grid points (every 64th), T=10, FP64 oracle on one thread, as "
                      "ompVC runs one FTLE slide" % pts1.shape[0]}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="f32", choices=["f32", "bf16", "f16", "f64"])
    ap.add_argument("--tracers", type=int, default=4_000_000)
    ap.add_argument("--ftle", default="128,128,640")
    ap.add_argument("--nx", type=int, default=60)
    ap.add_argument("--nz", type=int, default=330)
    ap.add_argument("--cpu-sample", type=int, default=1 << 18)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mesh-order", default="zorder", choices=["native", "zorder"],
                    help="element and node numbering of the pipe mesh: renumbered in Z-order by fvc_mesh_zorder "
                         "before upload (cudaVC's ZOrderSort step; untimed setup), or as generated (voxel scan order)")
    ap.add_argument("--no-f64-check", action="store_true")
    ap.add_argument("--tracer-periods", type=int, default=0,
                    help="pipe workload: after the timed steps, run the tracers over this many periods and report "
                         "each period (particles leaving, live-particle lists)")
    ap.add_argument("--workload", default="pipe", choices=["pipe", "batch", "serve", "traffic", "dg"],
                    help="pipe: tracers + FTLE grid over one period; batch: uneven releases in one persistent "
                         "kernel; serve: Poisson request stream on the Y-bifurcation; dg: the 2-D double-gyre "
                         "FTLE and its BF16/FP16 variants")
    ap.add_argument("--frames", type=int, default=4, help="traffic workload: tracer frames after the copy")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-window", type=float, default=15.0)
    ap.add_argument("--serve-leg", type=float, default=5.0,
                    help="pipe workload: after it, the served Y-bifurcation workload over an arrival window of this "
                         "many seconds for p50/p99 request latency (0: skip)")
    ap.add_argument("--serve-budget", type=int, default=32)
    ap.add_argument("--serve-precision", default="bf16", choices=["f32", "bf16", "f16"])
    ap.add_argument("--serve-check", type=int, default=40)
    ap.add_argument("--serve-no-graph", action="store_true", help="server launches each tick directly")
    ap.add_argument("--serve-baseline", type=int, default=1, choices=[0, 1],
                    help="also run every request alone (batch-at-a-time comparison)")
    ap.add_argument("--releases", type=int, default=64)
    ap.add_argument("--batch-cpu-stride", type=int, default=10)
    ap.add_argument("--batch-order", default="element", choices=["element", "zorder"],
                    help="batch workload: each release's seeds in Z-order, then ordered by the element holding each "
                         "(as the pipe tracers), or in Z-order only")
    ap.add_argument("--batch-config", default="",
                    help="batch workload: 'block,steps_per_launch' to run without the autotuner sweep")
    ap.add_argument("--batch-graph", type=int, default=1, choices=[0, 1],
                    help="replay the launch sequence from a captured hipGraph (1) or launch it directly (0)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the advection path runs on a single device (particles shard by "
              "Z-order range across devices, but no multi-GPU path is implemented); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from fvcpy import fvc, host

    if args.workload == "batch":
        bench_batch(args, np, fvc, host)
        return
    if args.workload == "serve":
        bench_serve(args, np, fvc, host)
        return
    if args.workload == "traffic":
        bench_traffic(args, np, fvc, host)
        return
    if args.workload == "dg":
        bench_dg(args, np, fvc, host)
        return
    res = [int(v) for v in args.ftle.split(",")]
    tube, xyz, conn, adj, slices, seeds = build_inputs(args, host, np)
    t0 = time.time()
    ctx = make_context(fvc, args.precision, xyz, conn, adj, slices)
    lo, hi = np.array([-1.0, -1.0, 0.0]), np.array([1.0, 1.0, 10.0])
    perm = ctx.zsort(seeds, lo, hi)
    seeds = seeds[perm]
    recs = element_order(np, ctx.seed_list(seeds, 1))
    n_tr = recs.shape[0]
    P_ic = fvc.Particles(ctx, n_tr)
    P_ic.set(recs)
    P = fvc.Particles(ctx, n_tr)
    F = fvc.FtleSlide(ctx, [-1.0, -1.0, 0.0], [1.0, 1.0, 10.0], res, 1, 0.0, 1.0)
    log("device setup %.1f s: %d tracers inside, FTLE %d points (%d inside)" %
        (time.time() - t0, int((recs["LeftDomain"] == 0).sum()), F.n, F.inside))

    for w in range(args.warmup):
        r = run_step(P, P_ic, F, ctx, n_tr, res)
        log("warmup %d: tracers %.3f s, FTLE %.3f s" % (w, r["t_tr"], r["t_ft"]))
    ctx.set_timing(True)
    tot_t = tot_steps = tr_t = ft_t = tr_steps = ft_steps = 0.0
    k_tr = [0.0, 0, 0]
    k_ft = [0.0, 0, 0]
    last = None
    fvc.lib().fvc_sync(ctx.h)
    wall0 = time.perf_counter()
    for s in range(args.steps):
        ctx.kernel_time(reset=True)
        r = run_step(P, P_ic, F, ctx, n_tr, res)
        last = r
        tr_t += r["t_tr"]
        ft_t += r["t_ft"]
        tr_steps += r["steps_tr"]
        ft_steps += r["steps_ft"]
        log("step %d: tracers %.3f s (%d steps), FTLE %.3f s (%d steps)" % (s, r["t_tr"], r["steps_tr"], r["t_ft"],
                                                                          r["steps_ft"]))
    fvc.lib().fvc_sync(ctx.h)
    wall = time.perf_counter() - wall0
    ctx.set_timing(False)
    # kernel-level timing of one more tracer pass and one FTLE pass (events on the library stream)
    ctx.kernel_time(reset=True)
    ctx.set_timing(True)
    P.copy_from(P_ic)
    P.launch_velocity(0.0, 0, 1, 0.0, 0.02)
    for tmin, tmax, lo_, hi_ in frames():
        P.advect(tmin, tmax, lo_, hi_, tmin, tmax, 1e-3)
    k_tr = list(ctx.kernel_time(reset=True))
    F.reset()
    for tmin, tmax, lo_, hi_ in frames():
        F.frame(1, lo_, hi_, tmin, tmax, 1e-3)
    F.output(1.0)
    k_ft = list(ctx.kernel_time(reset=True))
    ctx.set_timing(False)

    K = args.steps
    ms_per_step = 1000.0 * wall / K
    total_steps = (tr_steps + ft_steps) / K
    value = total_steps / (wall / K)
    ftle_seeds = F.inside
    # dominant kernel for the roofline
    dom = ("k_advect", k_tr) if k_tr[0] >= k_ft[0] else ("k_ftle_advect", k_ft)
    bpp = BYTES_PER_STEP_F32 if args.precision in ("f32",) else (
        512 if args.precision in ("bf16", "f16") else 1328)
    avg_ms = dom[1][0] / max(dom[1][1], 1)
    steps_per_launch = dom[1][2] / max(dom[1][1], 1)
    result = {
        "metric": "particle-steps/s + FTLE seeds/s at stated FP64 error; p50/p99 request latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32 (fp64 position/time state)", "bf16": "bf16 storage, f32 math",
                  "f16": "f16 storage, f32 math", "f64": "f64"}[args.precision],
        "data": "synthetic (deterministic generators, seeded)",
        "config": {
            "workload": "pipe R=1 L=10, 60x60x330 voxels Kuhn-split (%d tets, %d nodes%s), 50 slices/period, "
                        "%d tracers + %dx%dx%d FTLE grid, RK4 dt=1e-3 over one period" %
                        (conn.shape[0], xyz.shape[0],
                         ", renumbered in Z-order" if args.mesh_order == "zorder" else "",
                         n_tr, res[0], res[1], res[2]),
            "tracers": n_tr, "ftle_grid": res, "frames": 50, "dt": 1e-3, "precision": args.precision,
            "mesh_order": args.mesh_order, "particle_order": "tracers by the element holding each; FTLE points in the grid's Z-order",
        },
        "tracer_particle_steps_per_s": tr_steps / tr_t if tr_t > 0 else None,
        "ftle_seeds_per_s": ftle_seeds / (ft_t / K) if ft_t > 0 else None,
        "ftle_particle_steps_per_s": ft_steps / ft_t if ft_t > 0 else None,
        "p50_latency_ms": None,
        "p99_latency_ms": None,
        "roofline": roofline_valu(dom[0], steps_per_launch, avg_ms, bpp, args.precision == "f32" and
                                  args.tracers == 4_000_000 and res == [128, 128, 640] and args.nx == 60),
    }

    # accuracy against FP64 and the CPU baseline on a bounded sample
    stride = max(1, n_tr // max(1, args.cpu_sample))
    sample_idx = np.arange(0, n_tr, stride)
    gpu_final, gpu_left = P.get_xyz()
    if not args.no_cpu:
        cb = cpu_baseline(args, np, xyz, conn, adj, slices, recs[sample_idx])
        pts = cb["pts"]
        live = (pts["LeftDomain"] == 0) & (pts["LeftDomainTime"] > -1e-10) & (gpu_left[sample_idx] == 0)
        err = np.abs(gpu_final[sample_idx][live] - pts["X"][live]).max(axis=1) if live.any() else np.zeros(1)
        result["cpu_baseline"] = {
            "value": cb["value"], "unit": "particle-steps/s", "cores": cb["cores"], "kind": "port", "host": host_info(),
            "sample": "%d of the %d tracers (every %d-th in element order), one period, integrate-only time %.1f s, "
                      "oracle FP64 restatement of ompVC (OpenMP dynamic chunks)" % (sample_idx.size, n_tr, stride,
                                                                                     cb["seconds"]),
        }
        result["tracer_pos_err_vs_fp64"] = {"max": float(err.max()), "p99": float(np.percentile(err, 99)),
                                            "tolerance": 1e-5 * 10.0, "domain_size": 10.0,
                                            "exit_agreement": float((gpu_left[sample_idx] ==
                                                                     pts["LeftDomain"]).mean())}
        result["speedup_vs_cpu"] = (tr_steps / tr_t) / cb["value"] if tr_t > 0 else None
        # ompVC as configured for the FTLE slide: its FTLE loop parallelises
        # over output slides (UPPERLOOP, parameters.c:400-401), one slide here
        # runs on one thread; a sixteenth of the sample on one core
        c1 = cpu_baseline(args, np, xyz, conn, adj, slices, recs[sample_idx[::16]], threads=1)
        result["cpu_baseline_as_configured"] = {
            "value": c1["value"], "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "%d tracers (every 16th of the all-core sample), one period, integrate-only time %.1f s, one "
                      "thread, as ompVC runs the FTLE slide" % (c1["pts"].shape[0], c1["seconds"])}
        result["ftle_speedup_vs_cpu_as_configured"] = (ft_steps / ft_t) / c1["value"] if ft_t > 0 else None
    if not args.no_f64_check and args.precision != "f64":
        t0 = time.time()
        red = P.get()
        del P, P_ic
        c64 = make_context(fvc, "f64", xyz, conn, adj, slices)
        # every tracer against the FP64 path (bit-identical to the oracle): position error and
        # agreement of the located element
        P64 = fvc.Particles(c64, n_tr)
        P64.set(recs)
        P64.launch_velocity(0.0, 0, 1, 0.0, 0.02)
        for tmin, tmax, lo_, hi_ in frames():
            P64.advect(tmin, tmax, lo_, hi_, tmin, tmax, 1e-3)
        ref = P64.get()
        del P64
        live = (ref["LeftDomain"] == 0) & (red["LeftDomain"] == 0) & (ref["LeftDomainTime"] > -1e-10)
        perr = np.abs(red["X"][live] - ref["X"][live]).max(axis=1)
        result["tracer_vs_fp64_all"] = {
            "points": int(n_tr), "pos_err_max": float(perr.max()), "pos_err_p99": float(np.percentile(perr, 99)),
            "element_agreement": float((red["ElementIndex"] == ref["ElementIndex"]).mean()),
            "exit_agreement": float((red["LeftDomain"] == ref["LeftDomain"]).mean()),
            "tolerance": 1e-5 * 10.0}
        F64 = fvc.FtleSlide(c64, [-1.0, -1.0, 0.0], [1.0, 1.0, 10.0], res, 1, 0.0, 1.0)
        for tmin, tmax, lo_, hi_ in frames():
            F64.frame(1, lo_, hi_, tmin, tmax, 1e-3)
        w64, _ = F64.output(1.0)
        w32 = last["wt"]
        ok = w64 > 0
        rel_all = np.zeros_like(w64)
        rel_all[ok] = np.abs(w32[ok] - w64[ok]) / np.maximum(np.abs(w64[ok]), 1e-3 * np.median(np.abs(w64[ok])))
        rel = rel_all[ok]
        # points whose FTLE stencil (the point and its six grid neighbours) left
        # the domain at different times in the two runs: there the early FTLE
        # integrates over a different time, whatever the arithmetic precision
        r32, r64 = F.records(), F64.records()
        flip = (r32["LeftDomain"] != r64["LeftDomain"]) | (r32["LeftDomainTime"] != r64["LeftDomainTime"])
        flip = flip.reshape(res[0], res[1], res[2])
        near = flip.copy()
        for ax in range(3):
            lo_s = [slice(None)] * 3
            hi_s = [slice(None)] * 3
            lo_s[ax], hi_s[ax] = slice(0, -1), slice(1, None)
            near[tuple(lo_s)] |= flip[tuple(hi_s)]
            near[tuple(hi_s)] |= flip[tuple(lo_s)]
        near_kji = near.transpose(2, 1, 0).ravel()  # FTLE outputs are in k, j, i order
        same = ok & ~near_kji
        rs = rel_all[same]
        result["ftle_rel_err_vs_fp64"] = {"p50": float(np.median(rel)), "p99": float(np.percentile(rel, 99)),
                                          "p999": float(np.percentile(rel, 99.9)), "max": float(rel.max()),
                                          "above_tolerance": int((rel > 1e-3).sum()), "tolerance": 1e-3,
                                          "floor": "1e-3 x median |FTLE|", "points": int(ok.sum()),
                                          "exit_time_differs_in_stencil": int((ok & near_kji).sum()),
                                          "same_exit_max": float(rs.max()) if rs.size else None,
                                          "same_exit_above_tolerance": int((rs > 1e-3).sum())}
        worst = np.argsort(rel_all)[-3:]  # the largest relative errors: FP64 and FP32 values beside the floor
        result["ftle_rel_err_vs_fp64"]["largest"] = [
            {"ftle_fp64": float(w64[q]), "ftle_fp32": float(w32[q]), "rel": float(rel_all[q])} for q in worst]
        result["ftle_rel_err_vs_fp64"]["floor_value"] = float(1e-3 * np.median(np.abs(w64[ok])))
        result["ftle_rel_err_vs_fp64"]["median_ftle"] = float(np.median(np.abs(w64[ok])))
        log("FP64 FTLE reference %.1f s" % (time.time() - t0))
    if args.tracer_periods > 0:
        result["tracer_periods"] = tracer_periods(P, P_ic, ctx, args.tracer_periods)
    if args.serve_leg > 0:
        # request latency: the served workload (Y-bifurcation, BF16 slices,
        # Poisson arrivals) over a shorter arrival window, after the pipe's
        # device memory is released
        F.close()
        ctx.close()
        sr = bench_serve(args, np, fvc, host, window=args.serve_leg, emit=False)
        result["p50_latency_ms"] = sr["p50_latency_ms"]
        result["p99_latency_ms"] = sr["p99_latency_ms"]
        result["serving"] = {"workload": sr["config"]["workload"], "requests": sr["steps"],
                             "particle_steps_per_s": sr["value"],
                             "p50_latency_ms": sr["p50_latency_ms"], "p99_latency_ms": sr["p99_latency_ms"],
                             "batch_at_a_time": sr["batch_at_a_time"],
                             "results_identical_to_alone": sr["served"]["results_identical_to_alone"],
                             "makespan_s": sr["served"]["makespan_s"], "peak_running": sr["served"]["peak_running"]}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
