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
