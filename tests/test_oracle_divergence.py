"""MBIR3 at the precision-study projection count, on the CPU oracle (CPU only).

src/mbir3.cpp:123-125 passes ReconParams::mu (default 10, include/config.h)
to nagopt as the Lipschitz constant, so the step is 0.1/mu = 0.01 whatever the
normal operator's norm (src/nagopt.cpp:138).  With 3 x 360 projections the
normal operator's largest eigenvalue is in the thousands, far beyond 2/0.01,
and the reference algorithm diverges; with mu set to the operator's Lipschitz
constant it does not.  A small volume keeps the FP64 oracle within seconds.
"""
import math

import numpy as np

from lamino_amd import synth


def _job():
    rng = np.random.default_rng(20)
    ds = []
    for g in range(3):
        sc = synth.lamino_scan(360, 30.0, 45.0 * g)
        ds.append((rng.uniform(0.1, 1.0, (360, 9, 9)), sc.alpha, sc.gamma))
    return (9, 5, 9), ds


def test_reference_step_diverges_at_3x360_projections_and_the_lipschitz_step_does_not(orc):
    rd, ds = _job()
    eps = 1e-4
    od, _ = orc.mbir_setup(3, ds, rd, eps=eps)
    R2 = synth.pad_n(9, np.float32(1.4142))
    sets = [(al, ga, R2, R2) for _, al, ga in ds]
    A = lambda v: orc.sysmat(v, sets, eps=eps)     # noqa: E731
    # power iteration from below: |A x| <= lambda_max for unit x, so 4 steps bound it from below
    L = orc.lipschitz(A, np.random.default_rng(21).uniform(0, 1, (3,) + od), 4, 1e-3)
    assert 0.1 / 10.0 * L > 2.0 * 10      # the reference's step is > 20x beyond the stable 2/L
    kw = dict(max_iters=2, inner_iters=2, sigma=1000.0, p=1.2, tol=0.0, xtol=0.0, eps=eps)
    x_ref = orc.mbir(3, ds, rd, mu=10.0, **kw)
    x_lip = orc.mbir(3, ds, rd, mu=L, **kw)
    assert np.isfinite(x_lip).all() and np.abs(x_lip).max() < 1.0
    assert np.abs(x_ref).max() > 1e4 * np.abs(x_lip).max()
    assert math.isfinite(L)
