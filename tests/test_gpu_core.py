"""GPU parity of the core operators against the CPU oracle (small shapes).

Tolerances: forward/adjoint/NUFFT results within 1e-4 relative L2 of the FP64
oracle (NUFFT tolerance 1e-9); FFT, stencils and preconditioner at FP32
round-off.
"""
import math

import numpy as np
import pytest

from conftest import rel, say

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-4


@pytest.mark.parametrize("shape", [(2, 7, 9), (3, 89, 89), (2, 43, 179), (1, 305, 305), (2, 90, 360), (1, 21, 45),
                                   (1, 8, 10), (2, 1, 16), (1, 615, 17), (2, 61, 122), (1, 3, 1021), (1, 2, 3600),
                                   (1, 41, 82), (2, 123, 1230), (1, 915, 4), (1, 3, 915), (1, 183, 61), (1, 194, 6)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft2_matches_numpy(ctx, shape, sign):
    """Covers every plan kind of the pass: 2,3,5-smooth lengths in register
    radices up to 12 (9, 10, 16, 45, 90, 360, 3600 = 12*12*5*5); the Good-Thomas
    split with Rader's algorithm over the prime factor (7, 17, 21 = 3*7, 194 = 2*97) or with
    the direct DFT stage for the factors 41 and 61 (41, 82, 123, 1230 = 30*41,
    615 = 15*41 and 915 = 15*61 with their radix-15 stage, 61, 122, 183, 305);
    Bluestein (43, 89, 179, 1021), both pass orientations."""
    import lamino_amd as la

    say("fft2", shape, sign)
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    y = la.fft2(x, sign)
    ref = np.fft.fft2(x.astype(np.complex128)) if sign < 0 else np.fft.ifft2(x.astype(np.complex128)) * shape[1] * shape[2]
    assert rel(y, ref) < 5e-6


@pytest.mark.parametrize("shape", [(3, 7, 9), (2, 8, 10)])
def test_fft2_shifted_matches_reference_composition(ctx, orc, shape):
    import lamino_amd as la

    rng = np.random.default_rng(2)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    for sign in (-1, 1):
        want = orc.shift2(x, inverse=False)
        want = np.fft.fft2(want) if sign < 0 else np.fft.ifft2(want) * shape[1] * shape[2]
        want = orc.shift2(want, inverse=True)
        assert rel(la.fft2(x, sign, shifted=True), want) < 5e-6


@pytest.mark.parametrize("dims", [(5, 7, 9), (6, 8, 10), (4, 21, 89), (3, 43, 179), (1, 7, 9)])
def test_precond_matches_oracle(ctx, orc, dims):
    import lamino_amd as la

    say("precond", dims)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(dims).astype(np.float32)
    y = la.RampPreconditioner(dims).apply(x)
    assert rel(y, orc.precond(x)) < 5e-6


@pytest.mark.parametrize("dims", [(6, 5, 7), (9, 12, 11)])
def test_stencils_match_oracle(ctx, orc, dims):
    import lamino_amd as la

    rng = np.random.default_rng(4)
    u = rng.standard_normal(dims).astype(np.float32)
    v = rng.standard_normal((3,) + dims).astype(np.float32)
    np.testing.assert_array_equal(la.grad_u(u), orc.grad_u(u, prec=32).astype(np.float32))
    np.testing.assert_array_equal(la.divergence(v), orc.divergence(v, prec=32).astype(np.float32))
    lap_ref = orc.divergence(orc.grad_u(u, prec=32).astype(np.float32), prec=32).astype(np.float32)
    np.testing.assert_array_equal(la.laplacian(u), lap_ref)


@pytest.mark.parametrize("dims", [(6, 5, 7), (9, 12, 11), (35, 33, 70)])
def test_qggmrf_matches_oracle(ctx, orc, dims):
    """The device kernel evaluates each neighbour pair once and hands the far end the
    negated value (the δ = 0 constant unnegated); (35, 33, 70) spans several of its
    14 x 30 (j, k) tiles and 16-plane chunks, with a flat block across their edges."""
    import lamino_amd as la

    rng = np.random.default_rng(5)
    x = rng.standard_normal(dims).astype(np.float32)
    x[2, 2, 2] = x[2, 2, 3]  # delta = 0 branch
    if min(dims) > 20:
        x[10:20, 10:20, 25:40] = 0.5
    g = rng.standard_normal(dims).astype(np.float32)
    want = orc.qggmrf(g, x, 1000.0, 1.2, prec=32)
    got = la.qggmrf(g.copy(), x, 1000.0, 1.2)
    assert np.max(np.abs(got - want)) <= 2e-6 * max(1.0, np.max(np.abs(want)))


def _geom(P, seed=7):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, np.pi, P), rng.uniform(0, 2 * np.pi, P)


@pytest.mark.parametrize("dims,R,C,P", [((5, 3, 7), 7, 9, 4), ((6, 4, 8), 8, 9, 3)])
def test_nufft_tiny_vs_brute_force(ctx, orc, dims, R, C, P):
    import lamino_amd as la

    say("nufft tiny", dims, R, C, P)
    al, ga = _geom(P)
    pg = la.PolarGrid.per_projection(al, ga, R, C)
    rng = np.random.default_rng(8)
    f = (rng.standard_normal(dims) + 1j * rng.standard_normal(dims))
    c = la.nufft3d2(f, pg)
    assert rel(c, orc.nufft(2, dims, al, ga, R, C, f, brute=True)) < FWD_TOL
    cin = rng.standard_normal((P, R, C)) + 1j * rng.standard_normal((P, R, C))
    fo = la.nufft3d1(cin, dims, pg)
    assert rel(fo, orc.nufft(1, dims, al, ga, R, C, cin, brute=True)) < FWD_TOL


@pytest.mark.parametrize("literal", [True, False])
def test_forward_adjoint_vs_oracle(ctx, orc, literal):
    import lamino_amd as la

    dims, R, C, P = (9, 5, 11), 13, 15, 12
    rng = np.random.default_rng(9)
    if literal:
        th = np.linspace(-1.2, 0.0, P)
        pg = la.PolarGrid(th, R, C, 0.6)
        al, ga = th, np.full(P, 0.6)
    else:
        al, ga = np.full(P, math.radians(35)), np.linspace(0, 2 * np.pi, P, endpoint=False) + 0.3
        pg = la.PolarGrid.per_projection(al, ga, R, C)
    m = rng.standard_normal((3,) + dims).astype(np.float32)
    y = la.forward(m, pg)
    yo = orc.forward(m, al, ga, R, C)
    say("forward rel", rel(y, yo))
    assert rel(y, yo) < FWD_TOL
    p = rng.standard_normal((P, R, C)).astype(np.float32)
    a = la.adjoint(p, pg, dims)
    ao = orc.adjoint(p, dims, al, ga)
    say("adjoint rel", rel(a, ao))
    assert rel(a, ao) < FWD_TOL


@pytest.mark.parametrize("R,C,P", [(13, 15, 7), (13, 15, 2), (15, 11, 1), (12, 14, 6), (13, 14, 5)])
def test_forward_adjoint_detector_pairs_vs_oracle(ctx, orc, R, C, P):
    """Odd detector planes go through the detector transform two projections at a
    time (an odd count leaves the last one alone); even sizes take the one-plane
    path.  Both must match the oracle's plane-by-plane projector."""
    import lamino_amd as la

    dims = (9, 5, 11)
    rng = np.random.default_rng(90 + P)
    al, ga = np.full(P, math.radians(30)), np.linspace(0, 2 * np.pi, P, endpoint=False) + 0.2
    pg = la.PolarGrid.per_projection(al, ga, R, C)
    m = rng.standard_normal((3,) + dims).astype(np.float32)
    y = la.forward(m, pg)
    yo = orc.forward(m, al, ga, R, C)
    say(f"R={R} C={C} P={P}: forward rel", rel(y, yo))
    assert rel(y, yo) < FWD_TOL
    # every plane on its own, so that a pair's planes cannot trade content
    for k in range(P):
        assert rel(y[k], yo[k]) < 10 * FWD_TOL, k
    p = rng.standard_normal((P, R, C)).astype(np.float32)
    a = la.adjoint(p, pg, dims)
    ao = orc.adjoint(p, dims, al, ga)
    say(f"R={R} C={C} P={P}: adjoint rel", rel(a, ao))
    assert rel(a, ao) < FWD_TOL


@pytest.mark.parametrize("tilt,slab", [(30.0, None), (30.0, "1"), (-25.0, None), (90.0, None), (0.0, "1")])
def test_row_factorised_projectors(ctx, orc, monkeypatch, tilt, slab):
    """Projections that share the tilt run the projectors row-factorised (2-D spread /
    gather per detector row plus a y pass).  Checked against the oracle and against
    the direct 3-D path (LAMINO_NUFFT_ROWS=0); a 1-byte slab budget forces one row
    per group.  Tilts cover both orders of the rows' first y cells, rows spanning
    the whole y period (90 deg) and all rows on one y cell (0 deg)."""
    import lamino_amd as la

    dims, R, C, P = (9, 5, 11), 13, 15, 7
    rng = np.random.default_rng(123)
    al, ga = np.full(P, math.radians(tilt)), np.linspace(0, 2 * np.pi, P, endpoint=False) + 0.4
    if slab:
        monkeypatch.setenv("LAMINO_NUFFT_SLAB_BYTES", slab)
    pg = la.PolarGrid.per_projection(al, ga, R, C)
    m = rng.standard_normal((3,) + dims).astype(np.float32)
    p = rng.standard_normal((P, R, C)).astype(np.float32)
    y, a = la.forward(m, pg), la.adjoint(p, pg, dims)
    yo, ao = orc.forward(m, al, ga, R, C), orc.adjoint(p, dims, al, ga)
    say(f"tilt {tilt} slab {slab}: forward rel", rel(y, yo), "adjoint rel", rel(a, ao))
    assert rel(y, yo) < FWD_TOL
    assert rel(a, ao) < FWD_TOL
    monkeypatch.setenv("LAMINO_NUFFT_ROWS", "0")
    y3, a3 = la.forward(m, pg), la.adjoint(p, pg, dims)
    say("  vs direct path:", rel(y, y3), rel(a, a3))
    assert rel(y, y3) < 3e-5
    assert rel(a, a3) < 3e-5


@pytest.mark.parametrize("dims", [(9, 5, 11), (8, 6, 10)])
def test_sysmat_and_toeplitz_vs_oracle(ctx, orc, dims):
    """The quadratic form sums one of each mirrored pair of spectral points twice and the
    self-mirrored points once; the two shapes give spectral grids (18, 9, 21) and (15, 12,
    20), odd and even lengths on each axis."""
    import lamino_amd as la

    R, C, P = 13, 15, 10
    al1, ga1 = np.full(P, 0.5), np.linspace(0, 2 * np.pi, P, endpoint=False)
    al2, ga2 = np.linspace(-1, 1, P), np.full(P, 0.8)
    g1 = la.PolarGrid.per_projection(al1, ga1, R, C)
    g2 = la.PolarGrid.per_projection(al2, ga2, R + 2, C + 4)
    rng = np.random.default_rng(10)
    x = rng.standard_normal((3,) + dims).astype(np.float32)
    s1 = la.sysmat(x, g1)
    so1 = orc.sysmat(x, [(al1, ga1, R, C)])
    say("sysmat rel", rel(s1, so1))
    assert rel(s1, so1) < FWD_TOL
    T = la.ToeplitzOperator([g1, g2], dims, keep_per_grid=True)
    so = orc.sysmat(x, [(al1, ga1, R, C), (al2, ga2, R + 2, C + 4)])
    t = T(x)
    say("toeplitz rel", rel(t, so))
    assert rel(t, so) < 2e-5
    q = T.quad(x)
    so2 = orc.sysmat(x, [(al2, ga2, R + 2, C + 4)])
    want = np.array([[np.vdot(x[i].astype(np.float64), so1[i]) for i in range(3)],
                     [np.vdot(x[i].astype(np.float64), so2[i]) for i in range(3)]])
    assert np.max(np.abs(q - want)) / np.max(np.abs(want)) < 2e-5


def test_cg_diag_matches_oracle(ctx, orc):
    """tests/test_cg.cpp:16-54 pattern: diagonal operator through the host-callback adapter."""
    import lamino_amd as la

    dims = (4, 5, 6)
    rng = np.random.default_rng(11)
    diag = (1.0 + np.arange(3 * np.prod(dims)).reshape((3,) + dims) / 10.0).astype(np.float32)
    b = rng.standard_normal((3,) + dims).astype(np.float32)
    x0 = np.zeros_like(b)
    x, it = la.cgsolver(lambda v: diag * v, b, x0, 25, 0.0)
    xo, ito = orc.cg_diag(diag, b, x0, 25, 0.0)
    assert it == ito
    assert rel(x, xo) < 1e-4


@pytest.mark.parametrize("rows", ["1", "0"])
def test_point_bins_are_identical_across_builds(ctx, orc, rows):
    """The bins come from a bucket scatter with atomic cursors followed by an in-bucket rank
    by (cell, point index), so two builds of the same geometry give the same order: the
    row path's forward and adjoint (no atomics anywhere) and the 3-D path's interpolation
    are bit-identical across freshly built grids."""
    import os

    import lamino_amd as la

    dims, R, C, P = (21, 9, 23), 31, 33, 40
    al = np.full(P, np.radians(35.0))
    ga = 2 * np.pi * np.arange(P) / P
    rng = np.random.default_rng(71)
    m = rng.standard_normal((3,) + dims).astype(np.float32)
    proj = rng.standard_normal((P, R, C)).astype(np.float32)
    old = os.environ.get("LAMINO_NUFFT_ROWS")
    os.environ["LAMINO_NUFFT_ROWS"] = rows
    try:
        out = []
        for _ in range(2):
            pg = la.PolarGrid.per_projection(al, ga, R, C)
            out.append((la.forward(m, pg), la.adjoint(proj, pg, dims)))
    finally:
        if old is None:
            os.environ.pop("LAMINO_NUFFT_ROWS", None)
        else:
            os.environ["LAMINO_NUFFT_ROWS"] = old
    np.testing.assert_array_equal(out[0][0], out[1][0])
    if rows == "1":
        np.testing.assert_array_equal(out[0][1], out[1][1])
    assert rel(out[0][0], orc.forward(m, al, ga, R, C)) < FWD_TOL
    assert rel(out[0][1], orc.adjoint(proj, dims, al, ga)) < FWD_TOL


@pytest.mark.parametrize("dims", [(21, 9, 23), (20, 8, 22)])
def test_toeplitz_fused_spectral_pass_matches_separate_passes(ctx, orc, dims):
    """The normal operator's fused x-transform + multiply pass (the default) against the
    separate transform and multiply passes (LAMINO_TOEP_FUSED=0 at creation) and the
    oracle's sysmat; odd and even L-grid lengths (self-mirrored spectrum rows)."""
    import os

    import lamino_amd as la

    R, C, P = 15, 17, 20
    al = np.full(P, np.radians(30.0))
    ga = 2 * np.pi * np.arange(P) / P
    pg = la.PolarGrid.per_projection(al, ga, R, C)
    x = np.random.default_rng(72).standard_normal((3,) + dims).astype(np.float32)
    fused = la.ToeplitzOperator([pg], dims)(x)
    old = os.environ.get("LAMINO_TOEP_FUSED")
    os.environ["LAMINO_TOEP_FUSED"] = "0"
    try:
        sep = la.ToeplitzOperator([pg], dims)(x)
    finally:
        if old is None:
            os.environ.pop("LAMINO_TOEP_FUSED", None)
        else:
            os.environ["LAMINO_TOEP_FUSED"] = old
    say(f"Toeplitz {dims}: fused vs separate {rel(fused, sep):.2e}")
    assert rel(fused, sep) < 1e-6
    assert rel(fused, orc.sysmat(x, [(al, ga, R, C)])) < 2e-5


@pytest.mark.parametrize("tilts", [(30.0, 30.0, 30.0), (30.0, 38.0)])
def test_multi_dataset_projectors_equal_the_per_dataset_calls(ctx, orc, tilts):
    """lam_forward_multi / lam_adjoint_multi run one fine-grid transform for all
    gamma datasets of a volume.  Same tilt (one shared y window): the forward is
    bit-identical to per-dataset calls and the adjoint equals their sum to FP32
    rounding.  Different tilts (different windows): the transform covers every
    plane and both stay equal.  The summed adjoint also matches the oracle."""
    import lamino_amd as la

    dims, R, C, P = (21, 9, 23), 31, 33, 12
    rng = np.random.default_rng(71)
    m = rng.standard_normal((3,) + dims).astype(np.float32)
    grids, scans = [], []
    for k, t in enumerate(tilts):
        al = np.full(P, np.radians(t))
        ga = np.radians(45.0 * k) + 2 * np.pi * np.arange(P) / P
        grids.append(la.PolarGrid.per_projection(al, ga, R, C))
        scans.append((al, ga))
    multi = la.forward_multi(m, grids)
    single = [la.forward(m, g) for g in grids]
    for a, b in zip(multi, single):
        assert np.array_equal(a, b)
    projs = [rng.standard_normal((P, R, C)).astype(np.float32) for _ in grids]
    summed = la.adjoint_multi(projs, grids, dims)
    each = sum(la.adjoint(p, g, dims).astype(np.float64) for p, g in zip(projs, grids))
    say(f"tilts {tilts}: adjoint_multi vs sum of adjoints {rel(summed, each):.2e}")
    assert rel(summed, each) < 1e-6
    want = sum(orc.adjoint(p, dims, al, ga) for p, (al, ga) in zip(projs, scans))
    assert rel(summed, want) < FWD_TOL
