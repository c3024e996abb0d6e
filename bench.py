#!/usr/bin/env python3
"""Benchmark of the laminography reconstruction hot path on one MI355X.

Workload (split-Bregman configuration): volume [32,128,128] x 3 -> padded
(179, 43, 179); detector 128 -> 127 -> padded 305 x 305; two gamma datasets
(0, 45 deg) of 180 projections; tilt 35 deg.  A step is one forward plus one
adjoint projection of every dataset (all 3 magnetisation components), with
the inputs resident in HBM.  Also reported: CG iterations/s of the Toeplitz
normal operator + ramp preconditioner at the same shape, the plain-form
reconstruction latency of the whole split-Bregman job (30 outer x 3 inner),
the roofline of the dominant kernel and the CPU oracle timed on this host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "fwd+adj NUFFT projections/s, CG iters/s vs plain form; p50/p99 recon latency"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cg-iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-recon", action="store_true", help="skip the full reconstruction latency")
    ap.add_argument("--cpu-projections", type=int, default=24)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (16 uneven jobs) measurement")
    ap.add_argument("--no-served", action="store_true", help="skip the served (Poisson arrivals) measurement")
    ap.add_argument("--batched-jobs", type=int, default=16)
    ap.add_argument("--served-jobs", type=int, default=12)
    ap.add_argument("--slots", type=int, default=4, help="concurrent jobs of the batched/served server")
    ap.add_argument("--profile", action="store_true",
                    help="only the timed fwd+adj steps plus a 1 GiB device-to-device copy (PMC calibration), "
                         "for rocprofv3 runs")
    return ap.parse_args()


def hbm_traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, when present."""
    path = os.path.join(HERE, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        data = json.load(open(path))
        return data.get(kernel, {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _simulate_jobs(la, synth, specs):
    """Measured projections of every job (GPU forward projector, src/forward.cpp recipe)."""
    def project(mp, scan, dp):
        pg = la.PolarGrid.per_projection(scan.alpha, scan.gamma, dp, dp)
        return la.forward(mp, pg)

    jobs = []
    for sp in specs:
        ds, _ = synth.simulate(sp.work, project)
        jobs.append((sp.work.method, ds, sp.work.recon_dims, synth.recon_params(sp.work), sp.arrival_s))
    return jobs


def _rel(a, b):
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def run_batched(la, synth, njobs, slots):
    """Reconstructions/hour of a mix of uneven jobs: the plain form (one job at a
    time, lam_mbir) against the batched form (all jobs submitted at once to a
    server: concurrent slots, hipGraph-replayed solver iterations, autotuned
    spreader tiles).  Results of the two forms are compared job by job."""
    from lamino_amd.serve import Server

    specs = synth.batched_specs(njobs)
    log(f"batched: simulating {len(specs)} jobs")
    jobs = _simulate_jobs(la, synth, specs)
    log("batched: plain form, one job at a time")
    t0 = time.perf_counter()
    plain = [la.api._mbir(m, ds, rd, prm) for m, ds, rd, prm, _ in jobs]
    t_plain = time.perf_counter() - t0
    log(f"batched: plain {t_plain:.2f} s; batched form with {slots} slots")
    with Server(slots=slots, mode="graph") as srv:
        t0 = time.perf_counter()
        ids = [srv.submit(m, ds, rd, prm) for m, ds, rd, prm, _ in jobs]
        res = [srv.wait(i) for i in ids]
        t_batch = time.perf_counter() - t0
        st = srv.stats()
    diffs = [_rel(x, p) for (x, _), p in zip(res, plain)]
    for (m, ds, rd, prm, _), (x, info) in zip(jobs, res):
        log(f"  job MBIR{m} vol {rd} x {len(ds)}x{ds[0][0].shape}: slot {info['slot']} "
            f"start {info['start_s']:.2f} end {info['end_s']:.2f} setup {info['stats']['setup_s']:.3f} "
            f"solve {info['stats']['solve_s']:.3f}")
    out = {"jobs": len(jobs), "slots": slots, "plain_s": round(t_plain, 3), "batched_s": round(t_batch, 3),
           "recon_per_hour_plain": round(3600 * len(jobs) / t_plain, 1),
           "recon_per_hour_batched": round(3600 * len(jobs) / t_batch, 1),
           "speedup": round(t_plain / t_batch, 3), "max_rel_diff_vs_plain": max(diffs),
           "peak_admitted_bytes": st["peak_in_use"],
           "mix": "16 jobs: volumes {[16,64,64],[32,128,128],[51,256,256]}x3, 1-3 gammas of 90-360 angles, "
                  "tilt 25-40 deg, MBIR1 (20 CG) or MBIR2 (30x3), seeded"}
    log("batched:", out)
    return out


def run_served(la, synth, njobs, slots):
    """Continuous batching under Poisson arrivals: jobs are submitted at their
    arrival times while others run; latency = arrival -> result on host."""
    from lamino_amd.serve import Server

    specs = synth.served_specs(njobs)
    log(f"served: simulating {len(specs)} jobs")
    jobs = _simulate_jobs(la, synth, specs)
    log(f"served: {len(jobs)} Poisson arrivals (mean 4 jobs/s) on {slots} slots")
    with Server(slots=slots, mode="graph") as srv:
        t0 = time.perf_counter()
        ids = []
        for m, ds, rd, prm, arrival in jobs:
            delay = t0 + arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(m, ds, rd, prm))
        res = [srv.wait(i) for i in ids]
        t_all = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([info["latency_s"] for _, info in res])
    nonfinite = [f"MBIR{m} vol {rd}" for (m, ds, rd, prm, _), (x, _) in zip(jobs, res) if not np.isfinite(x).all()]
    waits = np.array([info["start_s"] - info["submit_s"] for _, info in res])
    out = {"jobs": len(jobs), "slots": slots, "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 3), "mean_queue_wait_s": round(float(waits.mean()), 3),
           "throughput_jobs_per_s": round(len(jobs) / t_all, 4), "makespan_s": round(t_all, 3),
           "nonfinite_jobs": nonfinite, "peak_admitted_bytes": st["peak_in_use"],
           "mix": "Poisson arrivals 4 jobs/s, volumes up to [51,511,511]x3, 2-4 gammas x 360 angles, "
                  "split-Bregman or qGGMRF 50/50, seeded"}
    log("served:", out)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported on this build "
              "(the path shards across GPUs by design, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import lamino_amd as la
    from lamino_amd import _lib, synth
    from lamino_amd.api import _dims, _ptrs

    work = synth.config(2)
    vol, det = work.solver_shapes()           # (179, 43, 179), 305
    G = len(work.scans)
    P = len(work.scans[0].alpha)
    log(f"workload: {work.name}; padded volume {vol}, detector {det}x{det}, {G} x {P} projections")

    ctx = la.default_context()
    lib = ctx.lib
    eps = la.api.DEFAULT_EPS

    # inputs: the padded vortex field and one projection stack per dataset, all resident on the device
    m = np.zeros((3,) + vol, np.float32)
    rd = work.recon_dims
    o = [(v - n) // 2 for v, n in zip(vol, rd)]
    m[:, o[0]:o[0] + rd[0], o[1]:o[1] + rd[1], o[2]:o[2] + rd[2]] = synth.vortex_field(rd, work.seed)
    md = ctx.to_device(m)
    grids = [la.PolarGrid.per_projection(s.alpha, s.gamma, det, det) for s in work.scans]
    projs = [ctx.alloc((P, det, det), np.float32) for _ in range(G)]
    outs = [ctx.alloc((3,) + vol, np.float32) for _ in range(G)]
    dims = _dims(vol)

    def step():
        for g in range(G):
            a, b, c = md.comps()
            _lib.check(lib.lam_forward(ctx.h, grids[g].h, dims, eps, a, b, c, projs[g].ptr))
            x, y, z = outs[g].comps()
            _lib.check(lib.lam_adjoint(ctx.h, grids[g].h, dims, eps, projs[g].ptr, x, y, z))

    log("warmup (includes per-grid point binning)")
    for _ in range(max(args.warmup, 1)):
        step()
    ctx.sync()
    # The timed region records HIP events around the spreader only (the dominant
    # kernel, whose launch time the roofline needs): every timed scope adds two
    # event records to the stream, and timing all ~28 launches of a step costs
    # about 5 % of the step.
    _lib.check(lib.lam_timing_enable(ctx.h, 1))
    _lib.check(lib.lam_timing_only(ctx.h, b"spread"))
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    t1 = time.perf_counter()
    spread_tot, spread_cnt = ctypes.c_double(), ctypes.c_int64()
    _lib.check(lib.lam_timing_query(ctx.h, b"spread", ctypes.byref(spread_tot), ctypes.byref(spread_cnt)))
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = G * P / ((t1 - t0) / args.steps)
    log(f"fwd+adj: {ms_per_step:.3f} ms/step -> {value:.1f} projections/s")
    if args.profile:
        # a few CG iterations of the Toeplitz normal operator + ramp preconditioner
        T = la.ToeplitzOperator(grids, vol)
        x0 = ctx.alloc((3,) + vol, np.float32)
        x0.zero()
        xd = ctx.alloc((3,) + vol, np.float32)
        op = _lib.OP_FN(ctypes.cast(lib.lam_toep_op, ctypes.c_void_p).value)
        it = ctypes.c_int64()
        _lib.check(lib.lam_cgsolver(ctx.h, dims, op, T.h, _ptrs(outs[0].comps()), _ptrs(x0.comps()), 5, 0.0,
                                    _ptrs(xd.comps()), ctypes.byref(it)))
        ctx.sync()
        log(f"ran {it.value} CG iterations")
        # plain copy kernel of known size: the FETCH_SIZE / WRITE_SIZE calibration
        nbytes = 1 << 30
        src, dst = ctx.alloc((nbytes // 4,), np.float32), ctx.alloc((nbytes // 4,), np.float32)
        src.zero()
        for _ in range(2):
            _lib.check(lib.lam_copy_device(ctx.h, ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes))
        ctx.sync()
        log(f"copied {nbytes} bytes device-to-device twice")
        return

    # per-kernel breakdown: a separate diagnostic pass with every scope timed
    diag_steps = min(args.steps, 5)
    _lib.check(lib.lam_timing_only(ctx.h, None))
    _lib.check(lib.lam_timing_enable(ctx.h, 1))
    for _ in range(diag_steps):
        step()
    ctx.sync()
    kernels = {}
    for name in ("interp", "spread", "fft_axis", "deapod_pad", "crop_deapod", "bin_points",
                 "spread_rows2d", "y_expand", "y_contract", "interp_rows2d"):
        tot, cnt = ctypes.c_double(), ctypes.c_int64()
        _lib.check(lib.lam_timing_query(ctx.h, name.encode(), ctypes.byref(tot), ctypes.byref(cnt)))
        if cnt.value:
            kernels[name] = {"total_ms": tot.value, "launches": cnt.value, "avg_ms": tot.value / cnt.value}
    _lib.check(lib.lam_timing_enable(ctx.h, 0))
    log("kernel time per step (ms, diagnostic pass):",
        {k: round(v["total_ms"] / diag_steps, 3) for k, v in kernels.items()})

    # roofline of the dominant NUFFT kernel: algorithmic bytes per launch M*(12+24) + Nfine*24
    # (coordinates + 3 complex components per point, fine grid touched once)
    nf = [2 * n for n in vol]
    from lamino_amd.synth import pad_n  # noqa: F401  (shape rules live there)

    def smooth_even(n):
        while True:
            r = n
            for p_ in (2, 3, 5):
                while r % p_ == 0:
                    r //= p_
            if r == 1 and n % 2 == 0:
                return n
            n += 1

    nf = [smooth_even(max(2 * n, 12)) for n in vol]
    Nfine = nf[0] * nf[1] * nf[2]
    M = P * det * det
    bytes_alg = M * 36 + Nfine * 24
    dom = max(("interp", "spread"), key=lambda k: kernels.get(k, {"total_ms": 0})["total_ms"])
    # the dominant kernel's launch time from the timed region (the diagnostic pass's
    # figure only if the interpolator were to dominate)
    avg_ms = spread_tot.value / spread_cnt.value if dom == "spread" and spread_cnt.value else kernels[dom]["avg_ms"]
    achieved = bytes_alg / (avg_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": hbm_traffic_from_profiles(dom), "kernel": dom,
                "bytes_per_launch": bytes_alg, "avg_launch_ms": round(avg_ms, 4)}
    # The roofs that bind the NUFFT kernels.  These datasets share their tilt, so the
    # projectors run row-factorised: per half-plane point the 2-D spreader
    # read-modify-writes W^2 cells x 3 complex components (2 x 24 B each) in its LDS
    # box and the 2-D interpolator reads them once (LDS roof; peak = 256 CUs x 128 B/clk
    # x 2.4 GHz, an estimate from the CU's LDS width, not a published figure); the y
    # passes stream the row slabs and the fine grid (HBM roof): y_contract reads the
    # fine grid once and writes one (z, x) slab per row, y_expand reads the slabs
    # (plus halo strips, not counted) and read-modify-writes the fine grid.
    w = 6
    half_points = P * ((det * det + 1) // 2)           # per launch: one dataset
    rows = ((det * det + 1) // 2 - 1) // det + 1        # detector rows of the half set
    slab = nf[0] * nf[2] * 24
    lds_peak_tbs = 256 * 128 * 2.4e9 / 1e12
    binding = {}
    for k, per_pt in (("spread_rows2d", 2 * 24 * w ** 2), ("interp_rows2d", 24 * w ** 2)):
        if k in kernels:
            tbs = half_points * per_pt / (kernels[k]["avg_ms"] / 1e3) / 1e12
            binding[k] = {"roof": "lds", "achieved_tbs": round(tbs, 2), "peak_tbs": round(lds_peak_tbs, 1),
                          "frac": round(tbs / lds_peak_tbs, 3), "bytes_per_launch": half_points * per_pt}
    for k, nbytes in (("y_contract", Nfine * 24 + rows * slab), ("y_expand", rows * slab + 2 * Nfine * 24)):
        if k in kernels:
            gbs = nbytes / (kernels[k]["avg_ms"] / 1e3) / 1e9
            binding[k] = {"roof": "hbm", "achieved_gbs": round(gbs, 1), "peak_gbs": HBM_PEAK_GBS,
                          "frac": round(gbs / HBM_PEAK_GBS, 3), "bytes_per_launch": nbytes}
    roofline["binding"] = binding

    # CG iterations/s: Toeplitz normal operator + ramp preconditioner at the same shape
    log("building Toeplitz normal operator")
    tb0 = time.perf_counter()
    T = la.ToeplitzOperator(grids, vol)
    ctx.sync()
    toep_build_s = time.perf_counter() - tb0
    b = outs[0]
    xd = ctx.alloc((3,) + vol, np.float32)
    x0 = ctx.alloc((3,) + vol, np.float32)
    x0.zero()
    op = _lib.OP_FN(ctypes.cast(lib.lam_toep_op, ctypes.c_void_p).value)
    it = ctypes.c_int64()
    _lib.check(lib.lam_cgsolver(ctx.h, dims, op, T.h, _ptrs(b.comps()), _ptrs(x0.comps()), 2, 0.0,
                                _ptrs(xd.comps()), ctypes.byref(it)))
    ctx.sync()
    tc0 = time.perf_counter()
    _lib.check(lib.lam_cgsolver(ctx.h, dims, op, T.h, _ptrs(b.comps()), _ptrs(x0.comps()), args.cg_iters, 0.0,
                                _ptrs(xd.comps()), ctypes.byref(it)))
    ctx.sync()
    cg_s = time.perf_counter() - tc0
    cg_rate = it.value / cg_s
    log(f"CG: {it.value} iterations in {cg_s:.3f} s -> {cg_rate:.1f} iters/s (Toeplitz build {toep_build_s:.3f} s)")

    # plain-form reconstruction latency of the whole split-Bregman job
    recon = None
    if not args.no_recon:
        log("simulating measured projections and running the split-Bregman job")

        def project(mp, scan, dp):
            pg = la.PolarGrid.per_projection(scan.alpha, scan.gamma, dp, dp)
            return la.forward(mp, pg)

        datasets, _ = synth.simulate(work, project)
        prm = la.ReconParams(max_iters=work.params["max_iters"], inner_iters=work.params["inner_iters"],
                             lam=work.params["lam"], mu=work.params["mu"], tol=0.0, xtol=0.0)
        tr0 = time.perf_counter()
        rec, st = la.MBIR2(datasets, rd, prm, stats=True)
        tr = time.perf_counter() - tr0
        recon = {"recon_s": round(tr, 4), "setup_s": round(st["setup_s"], 4), "solve_s": round(st["solve_s"], 4),
                 "normal_ops": st["normal_ops"], "cg_iters": st["cg_iters"], "outer_iters": st["outer_iters"],
                 "cg_iters_per_s": round(st["cg_iters"] / st["solve_s"], 2),
                 "finite": bool(np.isfinite(rec).all())}
        # the same job in graph form (device scalars, replayed hipGraph iterations); the first
        # submission creates the worker's plans and captures the graphs, the second is timed
        from lamino_amd.serve import Server

        with Server(slots=1, mode="graph") as srv:
            srv.wait(srv.submit(2, datasets, rd, prm))
            tg0 = time.perf_counter()
            recg, info = srv.wait(srv.submit(2, datasets, rd, prm))
            tg = time.perf_counter() - tg0
        sg = info["stats"]
        recon["graph"] = {"recon_s": round(tg, 4), "setup_s": round(sg["setup_s"], 4),
                          "solve_s": round(sg["solve_s"], 4), "cg_iters": sg["cg_iters"],
                          "cg_iters_per_s": round(sg["cg_iters"] / sg["solve_s"], 2),
                          "rel_diff_vs_plain": _rel(recg, rec)}
        log("recon:", recon)

    batched = None if args.no_batched else run_batched(la, synth, args.batched_jobs, args.slots)
    served = None if args.no_served else run_served(la, synth, args.served_jobs, args.slots)

    # CPU baseline: the oracle's FP32 restatement of the reference path (w=7, sigma=2 ES NUFFT,
    # per-component type-1/type-2, single-threaded detector FFTs), all host cores
    cpu = None
    if not args.no_cpu:
        import oracle

        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
        oracle.set_threads(cores)
        sc = work.scans[0]
        Ps = min(args.cpu_projections, P)
        log(f"CPU oracle baseline on {cores} threads: forward+adjoint of {Ps} projections")
        tcp0 = time.perf_counter()
        pc = oracle.forward(m, sc.alpha[:Ps], sc.gamma[:Ps], det, det, eps=1.2e-6, prec=32)
        oracle.adjoint(pc, vol, sc.alpha[:Ps], sc.gamma[:Ps], eps=1.2e-6, prec=32)
        tcpu = time.perf_counter() - tcp0
        gp = projs[0].download()[:Ps]
        diff = float(np.linalg.norm(gp - pc) / np.linalg.norm(pc))
        cpu = {"value": round(Ps / tcpu, 3), "unit": "projections/s", "cores": cores, "kind": "port",
               "sample": f"forward+adjoint of {Ps} projections of the gamma=0 dataset at the padded shapes "
                         f"(FP32, ES width 7), {tcpu:.2f} s",
               "gpu_vs_cpu_forward_rel_l2": diff}
        log("cpu:", cpu)

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "projections/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "split-Bregman config: vol [32,128,128]x3 -> padded (179,43,179), det 128 -> 305x305, "
                               "gammas (0,45) x 180 projections, tilt 35 deg; step = forward + adjoint of all datasets",
                   "padded_volume": list(vol), "detector": det, "datasets": G, "projections": P,
                   "nufft_eps": eps},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cg_iters_per_s": round(cg_rate, 2),
        "toeplitz_build_s": round(toep_build_s, 4),
        "recon": recon,
        "batched": batched,
        "served": served,
        "kernels_ms_per_step": {k: round(v["total_ms"] / diag_steps, 4) for k, v in kernels.items()},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
