"""Batched / served reconstruction jobs (lam_server_*) on the GPU.

- graph and async modes give the plain lam_mbir result (same kernels and
  reduction order; only the setup backprojection's atomic summation order
  may differ), and every result is within 1e-3 of the FP64 oracle;
- a job gives the same result alone and in a mix of concurrent jobs
  (relative difference <= 1e-5);
- admission keeps the sum of admitted footprints within the HBM budget and
  each job's measured device peak within its estimate;
- a failing job reports its own error and does not stop the server.
"""
import numpy as np
import pytest

from conftest import rel, say

pytestmark = pytest.mark.gpu

REC_TOL = 1e-3
MIX_TOL = 1e-5


def _job(seed, rd=(9, 5, 11), nds=1, P=8, det=9):
    rng = np.random.default_rng(seed)
    ds = []
    for g in range(nds):
        proj = rng.standard_normal((P, det, det)).astype(np.float32) + 3.0
        alpha = np.full(P, np.radians(30.0 + 4 * g))
        gamma = np.radians(45.0 * g) + 2 * np.pi * np.arange(P) / P
        ds.append((proj, alpha, gamma))
    return ds, rd


def _params(method):
    import lamino_amd as la

    if method == 1:
        return la.ReconParams(max_iters=6, tol=0.0)
    if method == 2:
        return la.ReconParams(max_iters=3, inner_iters=2, lam=0.1, mu=10.0, tol=0.0, xtol=0.0)
    return la.ReconParams(max_iters=3, inner_iters=3, sigma=1000.0, p=1.2, mu=10.0, tol=0.0, xtol=0.0)


@pytest.mark.parametrize("mode", ["graph", "async"])
@pytest.mark.parametrize("method", [1, 2, 3])
def test_server_modes_match_plain_and_oracle(ctx, orc, mode, method):
    import lamino_amd as la
    from lamino_amd.serve import Server

    ds, rd = _job(10 + method, nds=2 if method > 1 else 1)
    prm = _params(method)
    plain = la.api._mbir(method, ds, rd, prm)
    with Server(slots=2, mode=mode) as s:
        jid = s.submit(method, ds, rd, prm)
        x, info = s.wait(jid)
    say(f"method {method} {mode}: vs plain {rel(x, plain):.2e}, info {info['stats']}")
    assert info["status"] == 0
    assert rel(x, plain) < MIX_TOL
    kw = dict(max_iters=prm.max_iters, inner_iters=prm.inner_iters, lam=prm.lam, mu=prm.mu, sigma=prm.sigma,
              p=prm.p, tol=prm.tol, xtol=prm.xtol)
    xo = orc.mbir(method, ds, rd, **kw)
    assert rel(x, xo) < REC_TOL


@pytest.mark.parametrize("mode", ["graph", "async"])
@pytest.mark.parametrize("case", ["rejections", "xtol"])
def test_mbir3_device_line_search_follows_plain(ctx, orc, mode, case):
    """MBIR3's graph form decides the Armijo test on the device.  The reference's test
    (un-squared norms, src/nagopt.cpp:164-171) rejects trials once the step is about
    2.4/L (mu = L / 24 as the Lipschitz constant; the iterate then grows, so 4
    iterations): the step shrinks and the momentum waits.  With xtol = 0.1 and a step
    of 0.4/L the solve stops at its x-change test before max_iters.  The trials, outer
    iterations and iterate follow the plain (host-driven) form's and the oracle's."""
    import lamino_amd as la
    from lamino_amd.serve import Server

    ds, rd = _job(31, nds=2)
    _, s0 = la.api._mbir(3, ds, rd, la.ReconParams(max_iters=1, mu=0.0, sigma=1000.0, p=1.2), stats=True)
    L = s0["lipschitz"]
    if case == "rejections":
        prm = la.ReconParams(max_iters=4, inner_iters=5, sigma=1000.0, p=1.2, mu=L / 24, tol=0.0, xtol=0.0)
    else:
        prm = la.ReconParams(max_iters=10, inner_iters=3, sigma=1000.0, p=1.2, mu=L / 4, tol=0.0, xtol=0.1)
    plain, pst = la.api._mbir(3, ds, rd, prm, stats=True)
    with Server(slots=1, mode=mode) as s:
        x, info = s.wait(s.submit(3, ds, rd, prm))
    st = info["stats"]
    say(f"MBIR3 {mode} {case} L {L:.4g}: vs plain {rel(x, plain):.2e}; plain {pst['normal_ops']} ops "
        f"{pst['outer_iters']} outer, {mode} {st['normal_ops']} ops {st['outer_iters']} outer")
    assert info["status"] == 0 and np.isfinite(x).all() and np.isfinite(plain).all()
    if case == "rejections":
        assert pst["normal_ops"] > pst["outer_iters"]          # some trials were rejected
    else:
        assert pst["outer_iters"] < prm.max_iters              # stopped by the x change
    assert st["normal_ops"] == pst["normal_ops"] and st["outer_iters"] == pst["outer_iters"]
    assert rel(x, plain) < MIX_TOL
    kw = dict(max_iters=prm.max_iters, inner_iters=prm.inner_iters, lam=prm.lam, mu=prm.mu, sigma=prm.sigma,
              p=prm.p, tol=prm.tol, xtol=prm.xtol)
    assert rel(x, orc.mbir(3, ds, rd, **kw)) < REC_TOL


def test_result_same_alone_and_in_mix(ctx):
    from lamino_amd.serve import Server

    jobs = [(2, *_job(20, rd=(9, 5, 11), nds=2)), (1, *_job(21, rd=(13, 7, 13))), (3, *_job(22, nds=2)),
            (2, *_job(23, rd=(11, 5, 9), nds=1)), (1, *_job(24))]
    target = 0
    m, ds, rd = jobs[target]
    with Server(slots=1, mode="graph") as s:
        alone, _ = s.wait(s.submit(m, ds, rd, _params(m)))
    with Server(slots=4, mode="graph") as s:
        ids = [s.submit(mm, d, r, _params(mm)) for mm, d, r in jobs]
        res = [s.wait(i) for i in ids]
        st = s.stats()
    starts = [r[1]["start_s"] for r in res]
    ends = [r[1]["end_s"] for r in res]
    overlap = any(starts[i] < ends[j] and starts[j] < ends[i] for i in range(len(res)) for j in range(i))
    say(f"mix: stats {st}, overlapping jobs: {overlap}")
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert rel(res[target][0], alone) < MIX_TOL
    for (mm, d, r), (x, info) in zip(jobs, res):
        assert np.isfinite(x).all()
        assert info["footprint_peak"] <= info["footprint_est"], info


def test_budget_admission_serialises_jobs_that_do_not_fit(ctx):
    import lamino_amd as la
    from lamino_amd.serve import Server

    jobs = [_job(30 + k, rd=(11, 5, 11)) for k in range(3)]
    prm = _params(1)
    est = max(la.api.mbir_footprint(1, ds, rd, prm) for ds, rd in jobs)
    budget = int(est * 1.5)   # room for one job at a time
    with Server(budget_bytes=budget, slots=3, mode="graph") as s:
        ids = [s.submit(1, ds, rd, prm) for ds, rd in jobs]
        infos = [s.wait(i)[1] for i in ids]
        st = s.stats()
    say("budget", budget, "peak admitted", st["peak_in_use"])
    assert st["peak_in_use"] <= budget
    spans = sorted((i["start_s"], i["end_s"]) for i in infos)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1, spans


def test_over_budget_job_is_rejected(ctx):
    from lamino_amd import _lib
    from lamino_amd.serve import Server

    ds, rd = _job(40)
    with Server(budget_bytes=1 << 20, slots=1, mode="graph") as s:
        with pytest.raises(_lib.LaminoError) as e:
            s.submit(1, ds, rd, _params(1))
    assert e.value.code == 9   # LAM_ERR_BUDGET


def test_failing_job_reports_its_error_and_server_continues(ctx):
    from lamino_amd import _lib
    from lamino_amd.serve import Server

    ds, rd = _job(50)
    bad = [(np.zeros_like(ds[0][0]), ds[0][1], ds[0][2])]   # max(proj) = 0: division by zero
    with Server(slots=2, mode="graph") as s:
        j_bad = s.submit(1, bad, rd, _params(1))
        j_ok = s.submit(1, ds, rd, _params(1))
        with pytest.raises(_lib.LaminoError) as e:
            s.wait(j_bad)
        x, info = s.wait(j_ok)
        st = s.stats()
    assert e.value.code == 7   # LAM_ERR_DIVZERO
    assert info["status"] == 0 and np.isfinite(x).all()
    assert st["failed"] == 1 and st["completed"] == 1


def test_autotuned_tiles_keep_projector_parity(ctx, orc):
    """The tile shape the autotuner picks changes only FP32 summation order."""
    import lamino_amd as la

    dims, R, C, P = (21, 9, 23), 31, 33, 40
    rng = np.random.default_rng(61)
    al = np.full(P, np.radians(35.0))
    ga = 2 * np.pi * np.arange(P) / P
    pg = la.PolarGrid.per_projection(al, ga, R, C)
    proj = rng.standard_normal((P, R, C)).astype(np.float32)
    m = rng.standard_normal((3,) + dims).astype(np.float32)
    a0 = la.adjoint(proj, pg, dims)
    y0 = la.forward(m, pg)
    T, ms = la.autotune_bins(pg, dims)
    say("autotuned tiles", T, f"{ms:.3f} ms")
    assert all(t >= 2 for t in T) and ms > 0
    a1 = la.adjoint(proj, pg, dims)
    y1 = la.forward(m, pg)
    assert rel(a1, a0) < 1e-5 and rel(y1, y0) < 1e-5
    assert rel(a1, orc.adjoint(proj, dims, al, ga)) < 1e-4


def test_row_path_autotuner_chooses_a_tiling_and_keeps_parity(ctx, orc):
    """On the row path (projections sharing the tilt) the autotuner times several 2-D
    tilings of the (z, x) fine plane, each with its own point bins, and keeps the
    fastest.  Any tiling changes only FP32 summation order."""
    import lamino_amd as la

    dims, R, C, P = (89, 21, 89), 89, 89, 24
    rng = np.random.default_rng(67)
    al = np.full(P, np.radians(30.0))
    ga = 2 * np.pi * np.arange(P) / P
    pg = la.PolarGrid.per_projection(al, ga, R, C)
    proj = rng.standard_normal((P, R, C)).astype(np.float32)
    m = rng.standard_normal((3,) + dims).astype(np.float32)
    a0 = la.adjoint(proj, pg, dims)
    y0 = la.forward(m, pg)
    T, ms = la.autotune_bins(pg, dims)
    say("row tiling (Tz, W, Tx)", T, f"{ms:.3f} ms")
    nf = 180   # fine length of 89 voxels (2 * 89 -> the next even 2,3,5-smooth length)
    assert T[1] == 6 and ms > 0
    assert nf % T[0] == 0 and nf % T[2] == 0 and T[2] % 5 == 0 and min(T[0], T[2]) >= 6
    a1 = la.adjoint(proj, pg, dims)
    y1 = la.forward(m, pg)
    assert rel(a1, a0) < 1e-5 and rel(y1, y0) < 1e-5
    assert rel(a1, orc.adjoint(proj, dims, al, ga)) < 1e-4
    assert rel(y1, orc.forward(m, al, ga, R, C)) < 1e-4


@pytest.mark.parametrize("cls", ["[16,64,64] MBIR1 1x90", "[32,128,128] MBIR2 2x180",
                                 "[51,256,256] MBIR3 3x360", "[51,511,511] MBIR2 2x360"])
def test_full_size_job_per_volume_class_graph_equals_plain(ctx, cls):
    """One job of each volume class of the batched and served mixes at its real shape (the
    largest: volume [51,511,511] -> padded (721,71,721), detector 511 -> 1233) through the
    graph-form server: equal to the plain form, finite, and the measured device peak within
    the admission estimate.  Iterations are cut to a few; the shapes are what is tested."""
    import lamino_amd as la
    from lamino_amd import synth
    from lamino_amd.serve import Server

    vol, meth, gp = cls.split()
    v = [int(t) for t in vol.strip("[]").split(",")]
    G, P = (int(t) for t in gp.split("x"))
    method = int(meth[-1])
    params = {1: dict(max_iters=5, tol=0.0),
              2: dict(lam=0.1, mu=10.0, max_iters=2, inner_iters=3, tol=0.0, xtol=0.0),
              3: dict(sigma=1000.0, p=1.2, mu=0.0, max_iters=2, inner_iters=2, tol=0.0, xtol=0.0)}[method]
    work = synth.Workload(cls, method, synth.recon_dims_internal(v), synth.oddify(v[1]),
                          [synth.lamino_scan(P, 30.0, 45.0 * g) for g in range(G)], synth.SEED_BASE + 77, noise=0.01,
                          params=params)

    def project(mp, scan, dp):
        return la.forward(mp, la.PolarGrid.per_projection(scan.alpha, scan.gamma, dp, dp))

    ds, _ = synth.simulate(work, project)
    prm = synth.recon_params(work)
    say(f"{cls}: solver shapes {work.solver_shapes()}")
    plain = la.api._mbir(method, ds, work.recon_dims, prm)
    with Server(slots=1, mode="graph") as s:
        x, info = s.wait(s.submit(method, ds, work.recon_dims, prm))
    say(f"{cls}: graph vs plain {rel(x, plain):.2e}, peak {info['footprint_peak'] / 2**30:.2f} GiB, "
        f"est {info['footprint_est'] / 2**30:.2f} GiB, stats {info['stats']}")
    assert np.isfinite(plain).all() and np.isfinite(x).all()
    assert rel(x, plain) < MIX_TOL
    assert info["footprint_peak"] <= info["footprint_est"]
