"""GPU reconstruction jobs (MBIR1/2/3) against the oracle's FP64 restatement of
src/mbir_cg.cpp, src/mbir_sb.cpp and src/mbir3.cpp on small seeded inputs.

Tolerance: reconstructions within 1e-3 relative L2 of the FP64 oracle.
"""
import numpy as np
import pytest

from conftest import rel, say

pytestmark = pytest.mark.gpu

REC_TOL = 1e-3


def _datasets(nds, P, det, seed, per_projection=True):
    rng = np.random.default_rng(seed)
    out = []
    for g in range(nds):
        proj = rng.standard_normal((P, det, det)).astype(np.float32) + 3.0
        if per_projection:
            alpha = np.full(P, np.radians(30.0 + 5 * g))
            gamma = np.radians(45.0 * g) + 2 * np.pi * np.arange(P) / P
        else:
            alpha = np.linspace(-1.2, 0.0, P)
            gamma = np.radians(45.0 * g)
        out.append((proj, alpha, gamma))
    return out


@pytest.mark.parametrize("per_projection", [True, False])
def test_mbir1_cg_matches_oracle(ctx, orc, per_projection):
    import lamino_amd as la

    rd = (9, 5, 11)
    ds = _datasets(1, 10, 9, 1, per_projection)
    prm = la.ReconParams(max_iters=6, tol=0.0)
    x = la.MBIR1(ds, rd, prm)
    xo = orc.mbir(1, ds, rd, max_iters=6, tol=0.0)
    say("MBIR1 rel", rel(x, xo))
    assert rel(x, xo) < REC_TOL


def test_mbir2_split_bregman_matches_oracle(ctx, orc):
    import lamino_amd as la

    rd = (9, 5, 11)
    ds = _datasets(2, 8, 9, 2)
    prm = la.ReconParams(max_iters=3, inner_iters=2, lam=0.1, mu=10.0, tol=0.0, xtol=0.0)
    x = la.MBIR2(ds, rd, prm)
    xo = orc.mbir(2, ds, rd, max_iters=3, inner_iters=2, lam=0.1, mu=10.0, tol=0.0, xtol=0.0)
    say("MBIR2 rel", rel(x, xo))
    assert rel(x, xo) < REC_TOL


def test_mbir3_nag_qggmrf_matches_oracle(ctx, orc):
    import lamino_amd as la

    rd = (9, 5, 11)
    ds = _datasets(2, 8, 9, 3)
    prm = la.ReconParams(max_iters=4, inner_iters=3, sigma=1000.0, p=1.2, mu=10.0, tol=0.0, xtol=0.0)
    x = la.MBIR3(ds, rd, prm)
    xo = orc.mbir(3, ds, rd, max_iters=4, inner_iters=3, sigma=1000.0, p=1.2, mu=10.0, tol=0.0, xtol=0.0)
    say("MBIR3 rel", rel(x, xo))
    assert rel(x, xo) < REC_TOL


def test_split_bregman_callback_matches_oracle_structure(ctx, orc):
    """opt::split_bregman through the host-callback adapter, A = diag (exercises rhs, shrink, CG restarts)."""
    import lamino_amd as la

    dims = (6, 5, 7)
    rng = np.random.default_rng(4)
    diag = (1.0 + rng.random((3,) + dims)).astype(np.float32)
    yT = rng.standard_normal((3,) + dims).astype(np.float32)
    x = la.split_bregman(lambda v: diag * v, yT, np.zeros_like(yT), 0.1, 1.0, 3, 4, 0.0, 0.0)
    assert np.isfinite(x).all()
    with pytest.raises(ValueError):
        la.split_bregman(lambda v: v, yT, np.zeros((3, 2, 2, 2), np.float32), 0.1, 1.0, 1, 1, 0.0, 0.0)


def _scan_3x360(seed):
    from lamino_amd import synth

    rng = np.random.default_rng(seed)
    ds = []
    for g in range(3):
        sc = synth.lamino_scan(360, 30.0, 45.0 * g)
        ds.append((rng.uniform(0.1, 1.0, (360, 9, 9)).astype(np.float32), sc.alpha, sc.gamma))
    return (9, 5, 9), ds


def test_mbir3_at_3x360_projections_matches_oracle_while_it_diverges(ctx, orc):
    """The reference's step 0.1/mu with mu = 10 diverges at 3 x 360 projections (the normal
    operator's norm is ~6e3); before and while it does, GPU and FP64 oracle agree."""
    import lamino_amd as la

    rd, ds = _scan_3x360(22)
    for iters in (1, 2):
        prm = la.ReconParams(max_iters=iters, inner_iters=2, sigma=1000.0, p=1.2, mu=10.0, tol=0.0, xtol=0.0)
        x = la.MBIR3(ds, rd, prm)
        xo = orc.mbir(3, ds, rd, max_iters=iters, inner_iters=2, sigma=1000.0, p=1.2, mu=10.0, tol=0.0, xtol=0.0,
                      eps=1e-6)
        say(f"MBIR3 3x360 mu=10, {iters} it: max|x| {np.abs(xo).max():.3e}, rel {rel(x, xo):.2e}")
        assert rel(x, xo) < REC_TOL
    assert np.abs(xo).max() > 1e3


def test_mbir3_with_the_estimated_lipschitz_constant_converges_and_matches_oracle(ctx, orc):
    """mu <= 0 passes the normal operator's FP64 power-iteration estimate to nagopt; the
    estimate is within 5 % of the oracle's, and the oracle run with the same mu agrees."""
    import lamino_amd as la

    rd, ds = _scan_3x360(23)
    prm = la.ReconParams(max_iters=3, inner_iters=2, sigma=1000.0, p=1.2, mu=0.0, tol=0.0, xtol=0.0)
    x, st = la.MBIR3(ds, rd, prm, stats=True)
    L = st["lipschitz"]
    od, _ = orc.mbir_setup(3, ds, rd, eps=1e-6)
    from lamino_amd import synth

    R2 = synth.pad_n(9, np.float32(1.4142))
    sets = [(al, ga, R2, R2) for _, al, ga in ds]
    Lo = orc.lipschitz(lambda v: orc.sysmat(v, sets, eps=1e-6), np.random.default_rng(24).uniform(0, 1, (3,) + od),
                       30, 1e-6)
    say(f"Lipschitz: GPU {L:.6e}, oracle {Lo:.6e}")
    assert abs(L - Lo) < 0.05 * Lo
    assert np.isfinite(x).all() and np.abs(x).max() < 1.0
    xo = orc.mbir(3, ds, rd, max_iters=3, inner_iters=2, sigma=1000.0, p=1.2, mu=L, tol=0.0, xtol=0.0, eps=1e-6)
    say(f"MBIR3 3x360 mu=L: rel {rel(x, xo):.2e}")
    assert rel(x, xo) < REC_TOL
