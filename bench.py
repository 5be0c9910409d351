#!/usr/bin/env python3
"""Throughput of the flowVC advection hot path on one MI355X.

Workload (one "step"): the 3-D unstructured pulsatile pipe - cylinder R=1,
L=10 voxelised on 60x60x330 hexes, Kuhn 6-tet split (5.6M tets, 0.98M nodes),
50 velocity slices per period of u_z = 2(1-r^2)(1+0.5 sin 2 pi t) plus swirl
u_theta = 0.1 r (1-r^2) sin(pi z/10) and N(0, 0.01^2) noise - with
  * 4M tracers uniform in the inlet third, Z-order sorted, RK4 dt=1e-3 over one
    period (50 frames x 20 steps), exits compacted by the location walk, and
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

BYTES_PER_STEP_F32 = 704  # 5 containment tests x (16 B conn + 4 x 12 B coords) + 4 probes x 4 nodes x 2 slices x 12 B
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_inputs(args, host, np, with_seeds=True):
    t0 = time.time()
    tube = host.pipe(args.nx, args.nx, args.nz, 1.0, 10.0)
    xyz, conn, adj = host.tube_mesh(tube)
    slices = [host.tube_velocity(tube, xyz, s * 0.02, 0.01, 1234, s) for s in range(50)]
    seeds = host.tube_seeds(tube, args.tracers, 0.0, 10.0 / 3.0, 42) if with_seeds else None
    log("inputs: %d nodes, %d tets, %d tracers (%.1f s)" % (xyz.shape[0], conn.shape[0],
                                                           0 if seeds is None else seeds.shape[0], time.time() - t0))
    return tube, xyz, conn, adj, slices, seeds


def make_context(fvc, precision, xyz, conn, adj, slices):
    ctx = fvc.Context(3, [-1.0, 1.0, -1.0, 1.0, 0.0, 10.0], precision, 50, coords=xyz, conn=conn, adj=adj)
    for s, sl in enumerate(slices):
        ctx.load_slice(s, s * 0.02, sl)
    return ctx


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


def ctx_sync(ctx):
    from fvcpy import fvc
    fvc.lib().fvc_sync(ctx.h)


def cpu_baseline(args, np, xyz, conn, adj, slices, recs_sample):
    """The oracle (CPU FP64 restatement, ompVC semantics, OpenMP dynamic chunks
    as tracers.c:222-223) on a bounded sample of the tracers."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle
    cores = min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1, 16)
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
        recs.append(ctx.seed_list(sd, 1))
    recs = np.concatenate(recs)
    rel = release_table(counts, starts, horizons, types)
    n = recs.shape[0]
    log("batch: %d releases, %d seeds (%d RK4 / %d AB3 releases), setup %.1f s" %
        (len(rel), n, int((types == fvc.INT_RK4).sum()), int((types == fvc.INT_AB3).sum()), time.time() - t0))
    # autotune on a 1/8 copy (same releases, an eighth of the seeds each)
    rel_t = release_table(counts, starts, horizons, types, scale=8)
    idx_t = np.concatenate([np.arange(b0, b0 + rt[1]) for (b0, _, _, _, _), rt in zip(rel, rel_t)])
    Pt = fvc.Particles(ctx, idx_t.size)
    Pt.set(recs[idx_t])
    Bt = fvc.Batch(ctx, Pt, rel_t, 0.0, 0.02, 50, 1e-3)
    t0 = time.time()
    blocks = [64, 128, 256, 512, 1024] if args.precision == "f32" else [256]
    budgets = [1, 4, 16, 64, 256, 0]
    best, best_ms = Bt.tune(blocks, budgets, 1)
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
    achieved = k_steps * bpp / (k_ms * 1e-3) if k_ms > 0 else 0.0
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
                        "periods, RK4/AB3 mix, dt=1e-3, autotuned block/steps-per-launch, hipGraph replay" %
                        (conn.shape[0], len(rel), n),
            "releases": len(rel), "seeds": n, "precision": args.precision,
        },
        "batch": {"tuned": best, "tune_ms_eighth_batch": best_ms, "launch_groups": launches,
                  "particle_steps_per_run": tot_steps // K},
        "roofline": {
            "kernel": "k_batch (whole run: batch kernels + compaction, device time)",
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": None, "bytes_per_particle_step": bpp,
            "avg_launch_ms": k_ms,
        },
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle
        cores = min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1, 16)
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
            while (f1 + 1) * 0.02 < te:
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
            "value": steps / t_int, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": "every %d-th seed of each release, same releases/horizons/integrators, integrate-only "
                      "time %.1f s, oracle FP64 restatement (OpenMP)" % (stride, t_int)}
        result["pos_err_vs_fp64"] = {"max": float(err.max()), "p99": float(np.percentile(err, 99)),
                                     "tolerance": 1e-4, "domain_size": 10.0, "points": int(err.size)}
        result["speedup_vs_cpu"] = value / (steps / t_int)
    print(json.dumps(result), flush=True)


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
    ap.add_argument("--no-f64-check", action="store_true")
    ap.add_argument("--workload", default="pipe", choices=["pipe", "batch"],
                    help="pipe: tracers + FTLE grid over one period; batch: uneven releases in one persistent kernel")
    ap.add_argument("--releases", type=int, default=64)
    ap.add_argument("--batch-cpu-stride", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and this build has no multi-GPU path "
              "(sharding plan in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from fvcpy import fvc, host

    if args.workload == "batch":
        bench_batch(args, np, fvc, host)
        return
    res = [int(v) for v in args.ftle.split(",")]
    tube, xyz, conn, adj, slices, seeds = build_inputs(args, host, np)
    t0 = time.time()
    ctx = make_context(fvc, args.precision, xyz, conn, adj, slices)
    lo, hi = np.array([-1.0, -1.0, 0.0]), np.array([1.0, 1.0, 10.0])
    perm = ctx.zsort(seeds, lo, hi)
    seeds = seeds[perm]
    recs = ctx.seed_list(seeds, 1)
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
    dom = ("k_advect (tracers)", k_tr) if k_tr[0] >= k_ft[0] else ("k_ftle_advect (FTLE grid)", k_ft)
    bpp = BYTES_PER_STEP_F32 if args.precision in ("f32",) else (
        512 if args.precision in ("bf16", "f16") else 1328)
    avg_ms = dom[1][0] / max(dom[1][1], 1)
    steps_per_launch = dom[1][2] / max(dom[1][1], 1)
    achieved = steps_per_launch * bpp / (avg_ms * 1e-3) if avg_ms > 0 else 0.0
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
            "workload": "pipe R=1 L=10, 60x60x330 voxels Kuhn-split (%d tets, %d nodes), 50 slices/period, "
                        "%d tracers + %dx%dx%d FTLE grid, RK4 dt=1e-3 over one period" %
                        (conn.shape[0], xyz.shape[0], n_tr, res[0], res[1], res[2]),
            "tracers": n_tr, "ftle_grid": res, "frames": 50, "dt": 1e-3, "precision": args.precision,
        },
        "tracer_particle_steps_per_s": tr_steps / tr_t if tr_t > 0 else None,
        "ftle_seeds_per_s": ftle_seeds / (ft_t / K) if ft_t > 0 else None,
        "ftle_particle_steps_per_s": ft_steps / ft_t if ft_t > 0 else None,
        "p50_latency_ms": None,
        "p99_latency_ms": None,
        "roofline": {
            "kernel": dom[0],
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": None,
            "bytes_per_particle_step": bpp,
            "avg_launch_ms": avg_ms,
            "particle_steps_per_launch": steps_per_launch,
        },
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
            "value": cb["value"], "unit": "particle-steps/s", "cores": cb["cores"], "kind": "port",
            "sample": "%d of the %d tracers (every %d-th after Z-order), one period, integrate-only time %.1f s, "
                      "oracle FP64 restatement of ompVC (OpenMP dynamic chunks)" % (sample_idx.size, n_tr, stride,
                                                                                     cb["seconds"]),
        }
        result["tracer_pos_err_vs_fp64"] = {"max": float(err.max()), "p99": float(np.percentile(err, 99)),
                                            "tolerance": 1e-5 * 10.0, "domain_size": 10.0,
                                            "exit_agreement": float((gpu_left[sample_idx] ==
                                                                     pts["LeftDomain"]).mean())}
        result["speedup_vs_cpu"] = (tr_steps / tr_t) / cb["value"] if tr_t > 0 else None
    if not args.no_f64_check and args.precision != "f64":
        t0 = time.time()
        del P, P_ic
        c64 = make_context(fvc, "f64", xyz, conn, adj, slices)
        F64 = fvc.FtleSlide(c64, [-1.0, -1.0, 0.0], [1.0, 1.0, 10.0], res, 1, 0.0, 1.0)
        for tmin, tmax, lo_, hi_ in frames():
            F64.frame(1, lo_, hi_, tmin, tmax, 1e-3)
        w64, _ = F64.output(1.0)
        w32 = last["wt"]
        ok = w64 > 0
        rel = np.abs(w32[ok] - w64[ok]) / np.maximum(np.abs(w64[ok]), 1e-3 * np.median(np.abs(w64[ok])))
        result["ftle_rel_err_vs_fp64"] = {"p50": float(np.median(rel)), "p99": float(np.percentile(rel, 99)),
                                          "max": float(rel.max()), "tolerance": 1e-3,
                                          "floor": "1e-3 x median |FTLE|", "points": int(ok.sum())}
        log("FP64 FTLE reference %.1f s" % (time.time() - t0))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
