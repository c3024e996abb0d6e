#!/usr/bin/env python3
"""Benchmark of the laminography reconstruction hot path on one MI355X.

Headline workload (precision-study configuration, the largest single-GPU one):
volume [51,256,256] x 3 -> padded (359, 71, 359); detector 256 -> 255 ->
padded 615 x 615; three gamma datasets (0, 45, 90 deg) of 360 projections;
tilt 30 deg.  A step is one forward plus one adjoint projection of every
dataset (all 3 magnetisation components), inputs resident in HBM.  The point
binning of each grid is built once (warmup) and reused by every step; the
reference re-sorts the points in every call.

Also reported: the plain-CG and split-Bregman configurations' throughput
(extra keys), CG iterations/s of the Toeplitz normal operator + ramp
preconditioner at each shape, the plain-form reconstruction latency of each
configuration's job, the batched and served forms, the roofline of the
dominant kernel and the CPU oracle timed on this host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# The batched and served forms run 4 worker streams beside the main stream.  With
# HIP's default of 4 hardware queues per process two workers can land on one queue,
# where their work serialises (measured, 16-job mix: passes at 0.65-1.07 x the plain
# form among passes at 1.35 x); with 8 queues, 12 of 12 passes ran at 1.36-1.37 x.
# The runtime reads this once, at its initialisation, which happens after this line;
# an environment that sets fewer (HIP's default is 4) is raised to 8.
try:
    _queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
except ValueError:
    _queues = 4
if _queues < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

METRIC = "fwd+adj NUFFT projections/s, CG iters/s vs plain form; p50/p99 recon latency"
HBM_PEAK_GBS = 8000.0
KERNEL_SCOPES = ("interp", "spread", "fft_axis", "deapod_pad", "crop_deapod", "bin_points", "spread_rows2d",
                 "y_expand", "y_contract", "interp_rows2d", "toep_mul", "toep_xfused")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=3, choices=(1, 2, 3),
                    help="headline configuration (1 plain CG, 2 split-Bregman, 3 precision study)")
    ap.add_argument("--extra-configs", default="1,2", help="configurations also measured (extra keys)")
    ap.add_argument("--cg-iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-recon", action="store_true", help="skip the reconstruction latencies")
    ap.add_argument("--cpu-projections", type=int, default=0, help="CPU sample size (0: per configuration)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (16 uneven jobs) measurement")
    ap.add_argument("--no-served", action="store_true", help="skip the served (Poisson arrivals) measurement")
    ap.add_argument("--batched-jobs", type=int, default=16)
    ap.add_argument("--served-jobs", type=int, default=12)
    ap.add_argument("--slots", type=int, default=4, help="concurrent jobs of the batched/served server")
    ap.add_argument("--batched-plain-only", action="store_true",
                    help="run only the batched mix's jobs one at a time (the plain form), for a kernel trace "
                         "of the GPU's busy time in that form")
    ap.add_argument("--profile", action="store_true",
                    help="only the headline's timed fwd+adj steps, a few CG iterations and a 1 GiB "
                         "device-to-device copy (PMC calibration), for rocprofv3 runs")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hbm_traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` at `workload` from the committed rocprofv3 PMC summary."""
    path = os.path.join(HERE, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        data = json.load(open(path))
        ent = data.get(workload, {}).get(kernel, {})
        return ent.get("hbm_bytes_per_launch")
    except (OSError, ValueError, AttributeError):
        return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def smooth_even(n):
    while True:
        r = n
        for p_ in (2, 3, 5):
            while r % p_ == 0:
                r //= p_
        if r == 1 and n % 2 == 0:
            return n
        n += 1


def toep_len(n):
    """The Toeplitz L-grid length of an axis of n voxels: smallest 2,3,5-smooth L >= 2n - 1."""
    L = max(1, 2 * n - 1)
    while True:
        r = L
        for p_ in (2, 3, 5):
            while r % p_ == 0:
                r //= p_
        if r == 1:
            return L
        L += 1


def _rel(a, b):
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


class Projectors:
    """One configuration's forward + adjoint step with everything resident on the device."""

    def __init__(self, la, work):
        from lamino_amd import synth

        self.la, self.work = la, work
        self.vol, self.det = work.solver_shapes()
        self.G = len(work.scans)
        self.P = len(work.scans[0].alpha)
        self.ctx = la.default_context()
        m = np.zeros((3,) + self.vol, np.float32)
        rd = work.recon_dims
        o = [(v - n) // 2 for v, n in zip(self.vol, rd)]
        m[:, o[0]:o[0] + rd[0], o[1]:o[1] + rd[1], o[2]:o[2] + rd[2]] = synth.vortex_field(rd, work.seed)
        self.m = m
        self.md = self.ctx.to_device(m)
        self.grids = [la.PolarGrid.per_projection(s.alpha, s.gamma, self.det, self.det) for s in work.scans]
        # the row tiling of the projectors, chosen by timing candidates on the first dataset's
        # grid (the others share its shapes and take the same tiling)
        t0 = time.perf_counter()
        tile, tile_ms = la.autotune_bins(self.grids[0], self.vol)
        self.tuning = {"row_tiling": [tile[0], tile[2]], "spread_interp_ms": round(tile_ms, 3),
                       "tuning_s": round(time.perf_counter() - t0, 3)}
        self.projs = [self.ctx.alloc((self.P, self.det, self.det), np.float32) for _ in range(self.G)]
        self.grid_arr = (ctypes.c_void_p * self.G)(*[g.h for g in self.grids])
        self.proj_arr = (ctypes.c_void_p * self.G)(*[p.ptr for p in self.projs])
        self.outs = [self.ctx.alloc((3,) + self.vol, np.float32) for _ in range(self.G)]

    def step(self):
        """The forward of every dataset from the volume (one fine-grid transform for all of
        them, lam_forward_multi), then the adjoint of every dataset's projections."""
        from lamino_amd import _lib
        from lamino_amd.api import _dims

        lib, ctx, eps = self.ctx.lib, self.ctx, self.la.api.DEFAULT_EPS
        dims = _dims(self.vol)
        a, b, c = self.md.comps()
        _lib.check(lib.lam_forward_multi(ctx.h, self.grid_arr, self.G, dims, eps, a, b, c, self.proj_arr))
        for g in range(self.G):
            x, y, z = self.outs[g].comps()
            _lib.check(lib.lam_adjoint(ctx.h, self.grids[g].h, dims, eps, self.projs[g].ptr, x, y, z))

    def timed(self, steps, warmup):
        """(ms per step, spread launch total ms, spread launches): events around the spreader only."""
        from lamino_amd import _lib

        lib, ctx = self.ctx.lib, self.ctx
        # the first step builds every grid's point bins (key, bucket scatter, in-bucket order)
        _lib.check(lib.lam_timing_enable(ctx.h, 1))
        _lib.check(lib.lam_timing_only(ctx.h, b"bin_points"))
        self.step()
        tot, cnt = ctypes.c_double(), ctypes.c_int64()
        _lib.check(lib.lam_timing_query(ctx.h, b"bin_points", ctypes.byref(tot), ctypes.byref(cnt)))
        self.bin_ms = {"total_ms": round(tot.value, 3), "builds": cnt.value}
        for _ in range(max(warmup - 1, 0)):
            self.step()
        ctx.sync()
        # Every timed scope adds two event records to the stream; timing all ~28 launches
        # of a step costs about 5 % of it, so the timed region records the spreader only.
        _lib.check(lib.lam_timing_enable(ctx.h, 1))
        _lib.check(lib.lam_timing_only(ctx.h, b"spread"))
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
        ctx.sync()
        t1 = time.perf_counter()
        tot, cnt = ctypes.c_double(), ctypes.c_int64()
        _lib.check(lib.lam_timing_query(ctx.h, b"spread", ctypes.byref(tot), ctypes.byref(cnt)))
        _lib.check(lib.lam_timing_only(ctx.h, None))
        _lib.check(lib.lam_timing_enable(ctx.h, 0))
        return (t1 - t0) * 1e3 / steps, tot.value, cnt.value

    def diagnostic(self, steps):
        """Per-kernel time per step from a separate pass with every scope timed."""
        from lamino_amd import _lib

        lib, ctx = self.ctx.lib, self.ctx
        _lib.check(lib.lam_timing_only(ctx.h, None))
        _lib.check(lib.lam_timing_enable(ctx.h, 1))   # enabling clears earlier records
        for _ in range(steps):
            self.step()
        ctx.sync()
        kernels = {}
        for name in KERNEL_SCOPES:
            tot, cnt = ctypes.c_double(), ctypes.c_int64()
            _lib.check(lib.lam_timing_query(ctx.h, name.encode(), ctypes.byref(tot), ctypes.byref(cnt)))
            if cnt.value:
                kernels[name] = {"total_ms": tot.value, "launches": cnt.value, "avg_ms": tot.value / cnt.value}
        _lib.check(lib.lam_timing_enable(ctx.h, 0))
        return kernels

    def cg(self, iters):
        """CG iterations/s of the Toeplitz normal operator (all datasets) + ramp preconditioner."""
        from lamino_amd import _lib
        from lamino_amd.api import _dims, _ptrs

        la, ctx, lib = self.la, self.ctx, self.ctx.lib
        tb0 = time.perf_counter()
        T = la.ToeplitzOperator(self.grids, self.vol)
        ctx.sync()
        build_s = time.perf_counter() - tb0
        dims = _dims(self.vol)
        b = self.outs[0]
        xd = ctx.alloc((3,) + self.vol, np.float32)
        x0 = ctx.alloc((3,) + self.vol, np.float32)
        x0.zero()
        op = _lib.OP_FN(ctypes.cast(lib.lam_toep_op, ctypes.c_void_p).value)
        it = ctypes.c_int64()
        _lib.check(lib.lam_cgsolver(ctx.h, dims, op, T.h, _ptrs(b.comps()), _ptrs(x0.comps()), 2, 0.0,
                                    _ptrs(xd.comps()), ctypes.byref(it)))
        ctx.sync()
        tc0 = time.perf_counter()
        _lib.check(lib.lam_cgsolver(ctx.h, dims, op, T.h, _ptrs(b.comps()), _ptrs(x0.comps()), iters, 0.0,
                                    _ptrs(xd.comps()), ctypes.byref(it)))
        ctx.sync()
        cg_s = time.perf_counter() - tc0
        ms = 1e3 * cg_s / it.value
        # algorithmic bytes of one PCG iteration (3 components): the Toeplitz normal
        # operator on the half spectrum (volume in, 3 transforms forward and back,
        # 6 kernel spectra, the multiply's read and write, volume out), the ramp
        # preconditioner (volume in and out) and the fused vector step (x, p, r, Ap
        # read; x, r written)
        nv = int(np.prod(self.vol))
        ng = int(np.prod([toep_len(n) for n in self.vol]))
        op = 12 * nv + 3 * 2 * 16 * (ng // 2) + 6 * 4 * (ng // 2) + 3 * 8 * (ng // 2) * 2 + 12 * nv
        it_bytes = op + 2 * 12 * nv + (4 * 12 + 2 * 12) * nv
        gbs = it_bytes / (ms / 1e3) / 1e9
        return {"cg_iters_per_s": round(it.value / cg_s, 2), "cg_ms_per_iter": round(ms, 4),
                "toeplitz_build_s": round(build_s, 4),
                "cg_roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_iteration": it_bytes,
                                "l_grid": [toep_len(n) for n in self.vol]}}

    def release(self):
        self.grids = self.projs = self.outs = self.md = self.grid_arr = self.proj_arr = None


def roofline(pr, kernels, spread_ms, spread_cnt, workload_key):
    """Roofline of the dominant NUFFT kernel (the spreader): algorithmic bytes per launch
    M*(12 + 24) + Nfine*24 (coordinates + 3 complex components per point, fine grid once)
    over its average launch time measured with HIP events on the library's stream."""
    vol, det, P = pr.vol, pr.det, pr.P
    nf = [smooth_even(max(2 * n, 12)) for n in vol]
    Nfine = nf[0] * nf[1] * nf[2]
    M = P * det * det
    bytes_alg = M * 36 + Nfine * 24
    avg_ms = spread_ms / spread_cnt if spread_cnt else kernels.get("spread", {}).get("avg_ms", float("nan"))
    achieved = bytes_alg / (avg_ms / 1e3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": hbm_traffic_from_profiles("spread", workload_key),
           "kernel": "spread", "bytes_per_launch": bytes_alg, "avg_launch_ms": round(avg_ms, 4)}
    # The roofs that bind the row-factorised kernels: per half-plane point the 2-D spreader
    # read-modify-writes W^2 cells x 3 complex components (2 x 24 B each) in its LDS box and
    # the 2-D interpolator reads them once (LDS roof; peak = 256 CUs x 128 B/clk x 2.4 GHz,
    # an estimate from the CU's LDS width, not a published figure); the y passes stream the
    # row slabs and the fine grid (HBM roof).
    w = 6
    half_points = P * ((det * det + 1) // 2)
    rows = ((det * det + 1) // 2 - 1) // det + 1
    slab = nf[0] * nf[2] * 24
    lds_peak_tbs = 256 * 128 * 2.4e9 / 1e12
    binding = {}
    for k, per_pt in (("spread_rows2d", 2 * 24 * w ** 2), ("interp_rows2d", 24 * w ** 2)):
        if k in kernels:
            n_launch = kernels[k]["launches"] / max(1, kernels["spread" if k.startswith("spread") else "interp"]["launches"])
            tbs = half_points * per_pt / (kernels[k]["avg_ms"] * n_launch / 1e3) / 1e12
            binding[k] = {"roof": "lds", "achieved_tbs": round(tbs, 2), "peak_tbs": round(lds_peak_tbs, 1),
                          "frac": round(tbs / lds_peak_tbs, 3)}
    for k, nbytes in (("y_contract", Nfine * 24 + rows * slab), ("y_expand", rows * slab + Nfine * 24)):
        if k in kernels:
            n_launch = kernels[k]["launches"] / max(1, kernels["spread" if k == "y_expand" else "interp"]["launches"])
            gbs = nbytes / (kernels[k]["avg_ms"] * n_launch / 1e3) / 1e9
            binding[k] = {"roof": "hbm", "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 3)}
    out["binding"] = binding
    return out


def run_recon(la, synth, work, graph: bool):
    """Plain-form latency of the configuration's whole job (lam_mbir), and the graph form's
    (second submission to a one-slot server, after one that creates plans and graphs)."""
    def project(mp, scan, dp):
        return la.forward(mp, la.PolarGrid.per_projection(scan.alpha, scan.gamma, dp, dp))

    datasets, _ = synth.simulate(work, project)
    prm = synth.recon_params(work)
    tr0 = time.perf_counter()
    rec, st = la.api._mbir(work.method, datasets, work.recon_dims, prm, stats=True)
    tr = time.perf_counter() - tr0
    out = {"method": f"MBIR{work.method}", "recon_s": round(tr, 4), "setup_s": round(st["setup_s"], 4),
           "solve_s": round(st["solve_s"], 4), "normal_ops": st["normal_ops"], "cg_iters": st["cg_iters"],
           "outer_iters": st["outer_iters"], "finite": bool(np.isfinite(rec).all())}
    if work.method == 3:
        out["s_per_iteration"] = round(st["solve_s"] / max(1, st["outer_iters"]), 5)
        out["lipschitz"] = st["lipschitz"]
    elif st["cg_iters"]:
        out["cg_iters_per_s"] = round(st["cg_iters"] / st["solve_s"], 2)
    if graph:
        from lamino_amd.serve import Server

        with Server(slots=1, mode="graph") as srv:
            srv.wait(srv.submit(work.method, datasets, work.recon_dims, prm))
            tg0 = time.perf_counter()
            recg, info = srv.wait(srv.submit(work.method, datasets, work.recon_dims, prm))
            tg = time.perf_counter() - tg0
        sg = info["stats"]
        out["graph"] = {"recon_s": round(tg, 4), "setup_s": round(sg["setup_s"], 4), "solve_s": round(sg["solve_s"], 4),
                        "rel_diff_vs_plain": _rel(recg, rec)}
        if sg["cg_iters"]:
            out["graph"]["cg_iters_per_s"] = round(sg["cg_iters"] / sg["solve_s"], 2)
        if work.method == 3:
            out["graph"]["s_per_iteration"] = round(sg["solve_s"] / max(1, sg["outer_iters"]), 5)
            out["graph"]["normal_ops"] = sg["normal_ops"]
    return out


def _simulate_jobs(la, synth, specs):
    """Measured projections of every job (GPU forward projector, src/forward.cpp recipe)."""
    def project(mp, scan, dp):
        return la.forward(mp, la.PolarGrid.per_projection(scan.alpha, scan.gamma, dp, dp))

    jobs = []
    for sp in specs:
        ds, _ = synth.simulate(sp.work, project)
        jobs.append((sp.work.method, ds, sp.work.recon_dims, synth.recon_params(sp.work), sp.arrival_s))
    return jobs


def run_batched(la, synth, njobs, slots):
    """Reconstructions/hour of a mix of uneven jobs: the plain form (one job at a time,
    lam_mbir) against the batched form (all jobs submitted at once to a server: concurrent
    slots, hipGraph-replayed solver iterations, autotuned spreader tiles)."""
    from lamino_amd.serve import Server

    specs = synth.batched_specs(njobs)
    log(f"batched: simulating {len(specs)} jobs")
    jobs = _simulate_jobs(la, synth, specs)
    # Untimed warm-up: one-iteration copies of the jobs through a graph-form server tune
    # every job shape's projector tiling (a process-wide table), so that both timed forms
    # run the same tilings and their results differ only by the setup's summation order.
    import dataclasses
    with Server(slots=slots, mode="graph") as srv:
        for i in [srv.submit(m, ds, rd, dataclasses.replace(prm, max_iters=1)) for m, ds, rd, prm, _ in jobs]:
            srv.wait(i)
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
    out = {"jobs": len(jobs), "slots": slots, "plain_s": round(t_plain, 3), "batched_s": round(t_batch, 3),
           "recon_per_hour_plain": round(3600 * len(jobs) / t_plain, 1),
           "recon_per_hour_batched": round(3600 * len(jobs) / t_batch, 1),
           "speedup": round(t_plain / t_batch, 3), "max_rel_diff_vs_plain": max(diffs),
           "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
           "peak_admitted_bytes": st["peak_in_use"],
           "mix": "16 jobs: volumes {[16,64,64],[32,128,128],[51,256,256]}x3, 1-3 gammas of 90-360 angles, "
                  "tilt 25-40 deg, MBIR1 (20 CG) or MBIR2 (30x3), seeded"}
    log("batched:", out)
    return out


def run_batched_plain(la, synth, njobs):
    """The batched mix's jobs one at a time (the plain form) and nothing else.  A 3x5x7
    fft2 just before the pass marks its start in a kernel trace; the GPU's busy time is the
    union of the kernel intervals after it."""
    specs = synth.batched_specs(njobs)
    jobs = _simulate_jobs(la, synth, specs)
    la.default_context().sync()
    la.fft2(np.ones((3, 5, 7), np.complex64), -1)
    t0 = time.perf_counter()
    for m, ds, rd, prm, _ in jobs:
        la.api._mbir(m, ds, rd, prm)
    log(f"batched mix, plain form: {time.perf_counter() - t0:.3f} s")


def run_served(la, synth, njobs, slots):
    """Continuous batching under Poisson arrivals: jobs are submitted at their arrival
    times while others run; latency = arrival -> result on host."""
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
    finite = np.array([bool(np.isfinite(x).all()) for x, _ in res])
    nonfinite = [f"MBIR{m} vol {rd}" for (m, ds, rd, prm, _), ok in zip(jobs, finite) if not ok]
    waits = np.array([info["start_s"] - info["submit_s"] for _, info in res])
    arrival_span = max(j[4] for j in jobs)
    thr = len(jobs) / t_all
    rate = len(jobs) / arrival_span if arrival_span > 0 else float("inf")
    out = {"jobs": len(jobs), "slots": slots, "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 3), "mean_queue_wait_s": round(float(waits.mean()), 3),
           "throughput_jobs_per_s": round(thr, 4), "arrival_rate_jobs_per_s": round(rate, 4),
           "saturated": bool(thr < rate), "makespan_s": round(t_all, 3), "nonfinite_jobs": nonfinite,
           "peak_admitted_bytes": st["peak_in_use"],
           "mix": "Poisson arrivals 4 jobs/s, volumes up to [51,511,511]x3, 2-4 gammas x 360 angles, "
                  "split-Bregman (mu 10) or qGGMRF (mu = the normal operator's Lipschitz estimate) 50/50, seeded"}
    if finite.any():
        out["p50_latency_finite_s"] = round(float(np.percentile(lat[finite], 50)), 3)
        out["p99_latency_finite_s"] = round(float(np.percentile(lat[finite], 99)), 3)
    log("served:", out)
    return out


def cpu_baseline(pr, ps):
    """The oracle's FP32 restatement of the reference path (ES width 7, sigma = 2 NUFFT per
    component with per-call point setup, single-threaded detector FFTs), all host threads,
    on a bounded sample: forward + adjoint of ps projections of the first dataset."""
    import oracle

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    oracle.set_threads(cores)
    sc = pr.work.scans[0]
    Ps = min(ps, pr.P)
    log(f"CPU oracle baseline on {cores} threads: forward+adjoint of {Ps} projections")
    t0 = time.perf_counter()
    pc = oracle.forward(pr.m, sc.alpha[:Ps], sc.gamma[:Ps], pr.det, pr.det, eps=1.2e-6, prec=32)
    oracle.adjoint(pc, pr.vol, sc.alpha[:Ps], sc.gamma[:Ps], eps=1.2e-6, prec=32)
    tcpu = time.perf_counter() - t0
    gp = pr.projs[0].download()[:Ps]
    return {"value": round(Ps / tcpu, 3), "unit": "projections/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"forward+adjoint of {Ps} projections of the gamma=0 dataset at the padded shapes "
                      f"{list(pr.vol)} / {pr.det}x{pr.det} (FP32, ES width 7), {tcpu:.2f} s",
            "gpu_vs_cpu_forward_rel_l2": _rel(gp, pc)}


WORKLOAD_KEY = {1: "plain_cg_89x21x89_det89", 2: "split_bregman_179x43x179_det305",
                3: "precision_study_359x71x359_det615"}


def describe(work, pr):
    return (f"{work.name}; padded volume {list(pr.vol)}, detector {pr.det}x{pr.det}, {pr.G} x {pr.P} projections; "
            "step = forward of all datasets (one shared fine-grid transform) + adjoint of each (3 components); "
            "row tiling autotuned and point binning "
            "built once per grid before the timed steps")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported on this build "
              "(the path shards across GPUs by design, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import lamino_amd as la
    from lamino_amd import _lib, synth

    if args.batched_plain_only:
        run_batched_plain(la, synth, args.batched_jobs)
        return

    work = synth.config(args.config)
    pr = Projectors(la, work)
    log(f"headline workload: {describe(work, pr)}")
    ms_per_step, spread_ms, spread_cnt = pr.timed(args.steps, args.warmup)
    value = pr.G * pr.P / (ms_per_step / 1e3)
    log(f"fwd+adj: {ms_per_step:.3f} ms/step -> {value:.1f} projections/s")

    if args.profile:
        cg = pr.cg(5)
        log(f"CG: {cg}")
        ctx, lib = pr.ctx, pr.ctx.lib
        nbytes = 1 << 30
        src, dst = ctx.alloc((nbytes // 4,), np.float32), ctx.alloc((nbytes // 4,), np.float32)
        src.zero()
        for _ in range(2):
            _lib.check(lib.lam_copy_device(ctx.h, ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes))
        ctx.sync()
        log(f"copied {nbytes} bytes device-to-device twice")
        return

    diag_steps = min(args.steps, 3)
    kernels = pr.diagnostic(diag_steps)
    log("kernel time per step (ms, diagnostic pass):",
        {k: round(v["total_ms"] / diag_steps, 3) for k, v in kernels.items()})
    roof = roofline(pr, kernels, spread_ms, spread_cnt, WORKLOAD_KEY[args.config])
    log("roofline:", roof)
    cg = pr.cg(args.cg_iters)
    log("CG:", cg)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pr, args.cpu_projections or {1: 24, 2: 12, 3: 4}[args.config])
        log("cpu:", cpu)
    head = {"workload": describe(work, pr), "padded_volume": list(pr.vol), "detector": pr.det, "datasets": pr.G,
            "projections": pr.P, "nufft_eps": la.api.DEFAULT_EPS, "point_binning": "cached per grid (built by the autotuner and in warmup)",
            "bin_build_ms": pr.bin_ms, "autotune": pr.tuning}
    kernels_ms = {k: round(v["total_ms"] / diag_steps, 4) for k, v in kernels.items()}
    pr.release()

    extra = {}
    for ci in [int(c) for c in args.extra_configs.split(",") if c.strip()]:
        if ci == args.config:
            continue
        w = synth.config(ci)
        p2 = Projectors(la, w)
        ms2, _, _ = p2.timed(max(5, args.steps), args.warmup)
        extra[f"config{ci}"] = {"workload": describe(w, p2), "projections_per_s": round(p2.G * p2.P / (ms2 / 1e3), 2),
                                "ms_per_step": round(ms2, 4), "autotune": p2.tuning, **p2.cg(args.cg_iters)}
        log(f"config {ci}:", extra[f"config{ci}"])
        p2.release()

    recon = None
    if not args.no_recon:
        recon = {}
        for ci in sorted({args.config, *[int(c) for c in args.extra_configs.split(",") if c.strip()]}):
            w = synth.config(ci)
            log(f"reconstruction job of configuration {ci} ({w.name})")
            recon[f"config{ci}"] = run_recon(la, synth, w, graph=True)
            log(recon[f"config{ci}"])

    batched = None if args.no_batched else run_batched(la, synth, args.batched_jobs, args.slots)
    served = None if args.no_served else run_served(la, synth, args.served_jobs, args.slots)

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "projections/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": head,
        "roofline": roof, "cpu_baseline": cpu,
        "cg_iters_per_s": cg["cg_iters_per_s"], "cg_ms_per_iter": cg["cg_ms_per_iter"],
        "cg_roofline": cg["cg_roofline"], "toeplitz_build_s": cg["toeplitz_build_s"],
        "extra_configs": extra, "recon": recon, "batched": batched, "served": served,
        "kernels_ms_per_step": kernels_ms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
