"""Per-stage precision study against the FP64 oracle (the precision-study
configuration's structure: three gamma datasets (0, 45, 90 deg) at a 30 deg
tilt, qGGMRF with sigma=1000, p=1.2, q=2, step 0.1/mu with mu=10), at a
reduced size so that the CPU oracle finishes in seconds.

Every stage runs the device path in FP32 (FP64 scalars) on the same seeded
inputs as the oracle (FP64, NUFFT eps 1e-12); the relative L2 error of each
stage is printed as a table and held to the tolerance beside it:

  stage               tolerance   note
  grid coordinates    1e-6        FP32 storage of FP64 coordinates
  interp (type 2)     1e-4        ES width 6 (eps 1e-5)
  spread (type 1)     1e-4
  detector FFT        5e-6        vs numpy FP64
  forward             1e-4
  adjoint             1e-4
  normal operator     2e-5        exact Toeplitz form
  preconditioner      5e-6
  qGGMRF prior        1e-5
  iterate (MBIR3)     1e-3        5 Nesterov iterations
"""
import numpy as np
import pytest

from conftest import rel, say

pytestmark = pytest.mark.gpu

DIMS = (17, 7, 17)
DET = 21
P = 24
TILT = 30.0
GAMMAS = (0.0, 45.0, 90.0)


def _scans():
    from lamino_amd import synth

    return [synth.lamino_scan(P, TILT, g) for g in GAMMAS]


def test_per_stage_error_vs_fp64_oracle(ctx, orc):
    import lamino_amd as la

    rng = np.random.default_rng(2026)
    scans = _scans()
    sc = scans[1]
    pg = la.PolarGrid.per_projection(sc.alpha, sc.gamma, DET, DET)
    rows = []

    def stage(name, got, want, tol):
        e = rel(got, want)
        rows.append((name, e, tol))
        return e

    gx, gy, gz = orc.grid_coords(sc.alpha, sc.gamma, DET, DET)
    stage("grid coordinates", np.stack([pg.x, pg.y, pg.z]), np.stack([gx, gy, gz]), 1e-6)

    f = (rng.standard_normal(DIMS) + 1j * rng.standard_normal(DIMS)).astype(np.complex64)
    stage("interp (type 2)", la.nufft3d2(f, pg), orc.nufft(2, DIMS, sc.alpha, sc.gamma, DET, DET, f, eps=1e-12), 1e-4)
    c = (rng.standard_normal((P, DET, DET)) + 1j * rng.standard_normal((P, DET, DET))).astype(np.complex64)
    stage("spread (type 1)", la.nufft3d1(c, DIMS, pg), orc.nufft(1, DIMS, sc.alpha, sc.gamma, DET, DET, c, eps=1e-12),
          1e-4)

    x2 = (rng.standard_normal((P, DET, DET)) + 1j * rng.standard_normal((P, DET, DET))).astype(np.complex64)
    stage("detector FFT", la.fft2(x2, -1), np.fft.fft2(x2.astype(np.complex128), axes=(1, 2)), 5e-6)

    m = rng.standard_normal((3,) + DIMS).astype(np.float32)
    y = la.forward(m, pg)
    yo = orc.forward(m, sc.alpha, sc.gamma, DET, DET, eps=1e-12)
    stage("forward", y, yo, 1e-4)
    stage("adjoint", la.adjoint(yo.astype(np.float32), pg, DIMS), orc.adjoint(yo, DIMS, sc.alpha, sc.gamma, eps=1e-12),
          1e-4)

    grids = [la.PolarGrid.per_projection(s.alpha, s.gamma, DET, DET) for s in scans]
    T = la.ToeplitzOperator(grids, DIMS)
    stage("normal operator", T(m), orc.sysmat(m, [(s.alpha, s.gamma, DET, DET) for s in scans], eps=1e-12), 2e-5)

    u = rng.standard_normal(DIMS).astype(np.float32)
    stage("preconditioner", la.RampPreconditioner(DIMS).apply(u), orc.precond(u), 5e-6)
    g0 = rng.standard_normal(DIMS).astype(np.float32)
    stage("qGGMRF prior", la.qggmrf(g0.copy(), u, 1000.0, 1.2), orc.qggmrf(g0.astype(np.float64), u, 1000.0, 1.2),
          1e-5)

    # the iterate: 8 projections per gamma keep the reference's step 0.1/mu stable (more
    # projections raise the normal operator's norm past mu and the reference diverges)
    from lamino_amd import synth

    it_scans = [synth.lamino_scan(8, TILT, g) for g in GAMMAS]
    datasets = [(rng.standard_normal((8, 15, 15)).astype(np.float32) + 3.0, s.alpha, s.gamma) for s in it_scans]
    rd = (11, 5, 11)
    prm = la.ReconParams(sigma=1000.0, p=1.2, mu=10.0, max_iters=5, inner_iters=3, tol=0.0, xtol=0.0)
    xg = la.MBIR3(datasets, rd, prm)
    xo = orc.mbir(3, datasets, rd, max_iters=5, inner_iters=3, sigma=1000.0, p=1.2, mu=10.0, tol=0.0, xtol=0.0,
                  eps=1e-9)
    assert np.isfinite(xo).all(), "oracle iterate diverged at this size"
    stage("iterate (MBIR3, 5 it.)", xg, xo, 1e-3)

    say("\n| stage | rel. L2 error vs FP64 oracle | tolerance |\n|---|---|---|")
    for name, e, tol in rows:
        say(f"| {name} | {e:.2e} | {tol:.0e} |")
    bad = [(n, e, t) for n, e, t in rows if not e < t]
    assert not bad, bad
