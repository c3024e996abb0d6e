"""Parity at the benchmark configurations' real shapes, against the FP64 oracle.

  * plain CG configuration in full: volume [16,64,64] -> padded (89,21,89),
    detector 64 -> 63 -> 89, 90 projections at a 30 deg tilt: forward, adjoint
    (one row group and several), Toeplitz normal operator, and the 20-iteration
    MBIR1 job on the projections the oracle simulates;
  * split-Bregman configuration at full volume (179,43,179) and detector 305,
    on a projection subset of both gamma datasets: forward, adjoint, Toeplitz;
  * precision-study configuration at full volume (359,71,359) and detector
    615 on 3 x 4 projections (one per-projection grid): grid coordinates,
    forward, adjoint, ramp preconditioner, qGGMRF prior, and the normal
    operator through size-independent properties (symmetry, equality with the
    composed forward/adjoint the oracle checks);
  * FFT lengths of the largest served shapes and lines beyond 3072 points,
    preconditioner slices (71, 359) and (71, 721).

Tolerances: forward/adjoint 1e-4, Toeplitz 2e-5, FFT and preconditioner 5e-6,
reconstructions 1e-3 (relative L2).
"""
import os

import numpy as np
import pytest

from conftest import rel, say

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-4
TOEP_TOL = 2e-5
REC_TOL = 1e-3


def _padded_field(work, vol):
    from lamino_amd import synth

    m = np.zeros((3,) + tuple(vol), np.float32)
    rd = work.recon_dims
    o = [(v - n) // 2 for v, n in zip(vol, rd)]
    m[:, o[0]:o[0] + rd[0], o[1]:o[1] + rd[1], o[2]:o[2] + rd[2]] = synth.vortex_field(rd, work.seed)
    return m


def _subset(scan, k):
    idx = (np.arange(k) * len(scan.alpha)) // k
    return scan.alpha[idx], scan.gamma[idx]


# ---------------------------------------------------------------- plain CG configuration, in full
@pytest.fixture(scope="module")
def cfg1():
    from lamino_amd import synth

    work = synth.config(1)
    vol, det = work.solver_shapes()
    assert (vol, det) == ((89, 21, 89), 89)
    return work, vol, det


@pytest.mark.parametrize("slab_bytes", [None, 4 << 20])
def test_cfg1_forward_adjoint_full(ctx, orc, cfg1, slab_bytes):
    """All 90 projections; slab_bytes = 4 MiB splits the row path into several row groups."""
    import lamino_amd as la

    work, vol, det = cfg1
    sc = work.scans[0]
    m = _padded_field(work, vol)
    old = os.environ.get("LAMINO_NUFFT_SLAB_BYTES")
    try:
        if slab_bytes:
            os.environ["LAMINO_NUFFT_SLAB_BYTES"] = str(slab_bytes)
        pg = la.PolarGrid.per_projection(sc.alpha, sc.gamma, det, det)
        y = la.forward(m, pg)
        yo = orc.forward(m, sc.alpha, sc.gamma, det, det, eps=1e-9)
        e1 = rel(y, yo)
        a = la.adjoint(yo.astype(np.float32), pg, vol)
        ao = orc.adjoint(yo, vol, sc.alpha, sc.gamma, eps=1e-9)
        e2 = rel(a, ao)
    finally:
        if old is None:
            os.environ.pop("LAMINO_NUFFT_SLAB_BYTES", None)
        else:
            os.environ["LAMINO_NUFFT_SLAB_BYTES"] = old
    say(f"cfg1 full (slab bytes {slab_bytes}): forward {e1:.2e}, adjoint {e2:.2e}")
    assert e1 < FWD_TOL and e2 < FWD_TOL


def test_cfg1_toeplitz_full(ctx, orc, cfg1):
    import lamino_amd as la

    work, vol, det = cfg1
    sc = work.scans[0]
    m = _padded_field(work, vol)
    pg = la.PolarGrid.per_projection(sc.alpha, sc.gamma, det, det)
    t = la.ToeplitzOperator([pg], vol)(m)
    so = orc.sysmat(m, [(sc.alpha, sc.gamma, det, det)], eps=1e-9)
    say(f"cfg1 Toeplitz vs oracle sysmat: {rel(t, so):.2e}")
    assert rel(t, so) < TOEP_TOL


def test_cfg1_mbir1_job_full(ctx, orc, cfg1):
    """The whole plain job: projections simulated by the oracle (src/forward.cpp recipe),
    20 CG iterations (tol 0) on the GPU against the oracle's MBIR1."""
    import lamino_amd as la
    from lamino_amd import synth

    work, vol, det = cfg1

    def project(mp, scan, dp):
        return orc.forward(mp, scan.alpha, scan.gamma, dp, dp, eps=1e-9)

    ds, _ = synth.simulate(work, project)
    prm = la.ReconParams(max_iters=20, tol=0.0)
    x = la.MBIR1(ds, work.recon_dims, prm)
    xo = orc.mbir(1, ds, work.recon_dims, max_iters=20, tol=0.0, eps=1e-6)
    say(f"cfg1 MBIR1 20 it: rel {rel(x, xo):.2e}")
    assert np.isfinite(x).all()
    assert rel(x, xo) < REC_TOL


# ---------------------------------------------------------------- split-Bregman configuration
def test_cfg2_full_shapes_on_a_projection_subset(ctx, orc):
    """Volume (179,43,179), detector 305, 6 projections of each gamma dataset (0, 45 deg)."""
    import lamino_amd as la
    from lamino_amd import synth

    work = synth.config(2)
    vol, det = work.solver_shapes()
    assert (vol, det) == ((179, 43, 179), 305)
    m = _padded_field(work, vol)
    sets, grids = [], []
    for sc in work.scans:
        al, ga = _subset(sc, 6)
        sets.append((al, ga, det, det))
        pg = la.PolarGrid.per_projection(al, ga, det, det)
        grids.append(pg)
        y = la.forward(m, pg)
        yo = orc.forward(m, al, ga, det, det, eps=1e-9)
        a = la.adjoint(yo.astype(np.float32), pg, vol)
        ao = orc.adjoint(yo, vol, al, ga, eps=1e-9)
        say(f"cfg2 gamma {np.degrees(ga[0]):.0f}: forward {rel(y, yo):.2e}, adjoint {rel(a, ao):.2e}")
        assert rel(y, yo) < FWD_TOL and rel(a, ao) < FWD_TOL
    t = la.ToeplitzOperator(grids, vol)(m)
    so = orc.sysmat(m, sets, eps=1e-8)
    say(f"cfg2 Toeplitz (2 datasets) vs oracle sysmat: {rel(t, so):.2e}")
    assert rel(t, so) < TOEP_TOL


# ---------------------------------------------------------------- precision-study configuration
@pytest.fixture(scope="module")
def cfg3():
    from lamino_amd import synth

    work = synth.config(3)
    vol, det = work.solver_shapes()
    assert (vol, det) == ((359, 71, 359), 615)
    al = np.concatenate([_subset(sc, 4)[0] for sc in work.scans])
    ga = np.concatenate([_subset(sc, 4)[1] for sc in work.scans])
    return work, vol, det, al, ga


def test_cfg3_grid_coordinates(ctx, orc, cfg3):
    import lamino_amd as la

    work, vol, det, al, ga = cfg3
    pg = la.PolarGrid.per_projection(al, ga, det, det)
    xo, yo, zo = orc.grid_coords(al, ga, det, det, prec=64)
    e = max(rel(pg.x, xo), rel(pg.y, yo), rel(pg.z, zo))
    say(f"cfg3 grid coordinates: {e:.2e}")
    assert e < 1e-6


def test_cfg3_forward_adjoint_full_shapes(ctx, orc, cfg3):
    import lamino_amd as la

    work, vol, det, al, ga = cfg3
    m = _padded_field(work, vol)
    pg = la.PolarGrid.per_projection(al, ga, det, det)
    y = la.forward(m, pg)
    yo = orc.forward(m, al, ga, det, det, eps=1e-7)
    say(f"cfg3 forward (3 x 4 projections): {rel(y, yo):.2e}")
    assert rel(y, yo) < FWD_TOL
    a = la.adjoint(yo.astype(np.float32), pg, vol)
    ao = orc.adjoint(yo, vol, al, ga, eps=1e-7)
    say(f"cfg3 adjoint: {rel(a, ao):.2e}")
    assert rel(a, ao) < FWD_TOL


def test_cfg3_normal_operator_properties(ctx, cfg3):
    """The Toeplitz operator at (359,71,359) equals adjoint(forward(.)) of the projectors
    checked above (the identity the oracle pins at small size) and is symmetric."""
    import lamino_amd as la

    work, vol, det, al, ga = cfg3
    grids = []
    for g in range(3):
        s = slice(4 * g, 4 * g + 4)
        grids.append(la.PolarGrid.per_projection(al[s], ga[s], det, det))
    T = la.ToeplitzOperator(grids, vol)
    m = _padded_field(work, vol)
    rng = np.random.default_rng(31)
    v = rng.standard_normal((3,) + vol).astype(np.float32)
    tm = T(m)
    comp = sum(la.adjoint(la.forward(m, pg), pg, vol).astype(np.float64) for pg in grids)
    say(f"cfg3 Toeplitz vs adjoint(forward): {rel(tm, comp):.2e}")
    assert rel(tm, comp) < FWD_TOL
    tv = T(v)
    lhs = float(np.sum(tm.astype(np.float64) * v))
    rhs = float(np.sum(m.astype(np.float64) * tv))
    say(f"cfg3 Toeplitz symmetry: {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)


@pytest.mark.parametrize("dims", [(4, 71, 359), (3, 71, 721)])
def test_precond_served_slices(ctx, orc, dims):
    import lamino_amd as la

    rng = np.random.default_rng(32)
    x = rng.standard_normal(dims).astype(np.float32)
    y = la.RampPreconditioner(dims).apply(x)
    e = rel(y, orc.precond(x))
    say(f"precond {dims}: {e:.2e}")
    assert e < 5e-6


def test_cfg3_precond_and_prior_full_volume(ctx, orc, cfg3):
    import lamino_amd as la

    work, vol, det, al, ga = cfg3
    m = _padded_field(work, vol)
    y = la.RampPreconditioner(vol).apply(m[0])
    e = rel(y, orc.precond(m[0]))
    say(f"cfg3 precond full volume: {e:.2e}")
    assert e < 5e-6
    g = np.zeros(vol, np.float32)
    got = la.qggmrf(g.copy(), m[1], 1000.0, 1.2)
    want = orc.qggmrf(g, m[1], 1000.0, 1.2, prec=32)
    err = np.max(np.abs(got - want)) / max(1e-30, np.max(np.abs(want)))
    say(f"cfg3 qGGMRF full volume: max err / max {err:.2e}")
    assert err <= 2e-6


@pytest.mark.parametrize("shape", [(1, 2, 1233), (1, 2, 1458), (1, 1, 1567), (1, 1, 3459), (1, 1233, 3)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft2_served_lengths(ctx, shape, sign):
    """1233 = 3^2 * 137 (Bluestein, m = 2500), 1458 = 2 * 3^6, 1567 (prime, Bluestein
    m = 3200 > 3072: the second instantiation), 3459 = 3 * 1153 (Rader with Q - 1 = 1152
    smooth, beyond 3072 points), and 1233 along the strided axis."""
    import lamino_amd as la

    rng = np.random.default_rng(33)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    y = la.fft2(x, sign)
    ref = np.fft.fft2(x.astype(np.complex128)) if sign < 0 else np.fft.ifft2(x.astype(np.complex128)) * shape[1] * shape[2]
    e = rel(y, ref)
    say(f"fft2 {shape} sign {sign}: {e:.2e}")
    assert e < 5e-6
