"""Pins the CPU oracle (oracle/oracle.cpp) -- CPU only.

Every oracle function is checked against something independent of it:
  * tests/golden/reference_fixtures.npz -- outputs of the reference's own
    sources (polar_grid.h, projection.h, fftutils.h, padding.h,
    finitediff.cpp, qggmrf.cpp) compiled by tests/golden/make_golden.py;
  * the live reference build (oracle/_ref) where present;
  * numpy.fft (library routine) for FFTs, shifts and the ramp filter;
  * brute-force NUDFT, closed forms (voxel-delta projection, Toeplitz kernel),
    adjointness and the identity adjoint(forward) = sysmat.
"""
import math
import os

import numpy as np
import pytest

from conftest import rel

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "reference_fixtures.npz"))


# ---------------------------------------------------------------- FFT and shifts
@pytest.mark.parametrize("n", [1, 2, 3, 7, 12, 21, 45, 89, 90, 143, 305, 360, 615, 1233])
def test_fft_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    assert rel(orc.fft1d(x, -1), np.fft.fft(x)) < 1e-13
    assert rel(orc.fft1d(x, +1), np.fft.ifft(x) * n) < 1e-13


@pytest.mark.parametrize("name", ["odd", "even"])
def test_shifts_match_reference_fixture(orc, name):
    a = FIX[f"shift_{name}_in"]
    np.testing.assert_array_equal(orc.shift2(a, False), FIX[f"shift_{name}_fwd"])
    np.testing.assert_array_equal(orc.shift2(a, True), FIX[f"shift_{name}_inv"])
    # fftshift2 is numpy's ifftshift, ifftshift2 numpy's fftshift
    np.testing.assert_array_equal(FIX[f"shift_{name}_fwd"], np.fft.ifftshift(a, axes=(1, 2)))
    np.testing.assert_array_equal(FIX[f"shift_{name}_inv"], np.fft.fftshift(a, axes=(1, 2)))


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("gi", [0, 1, 2])
def test_grid_matches_reference_fixture(orc, prec, gi):
    theta = FIX[f"grid{prec}_{gi}_theta"]
    gamma = float(FIX[f"grid{prec}_{gi}_gamma"][0])
    R, C = (int(v) for v in FIX[f"grid{prec}_{gi}_rc"])
    x, y, z = orc.grid_coords(theta, np.full(theta.size, gamma), R, C, prec=prec)
    np.testing.assert_array_equal(np.stack([x, y, z]), FIX[f"grid{prec}_{gi}_xyz"])


def test_grid_plane_basis_invariants(orc):
    rng = np.random.default_rng(3)
    al, ga = rng.uniform(-3, 3, 9), rng.uniform(-3, 3, 9)
    x, y, z = orc.grid_coords(al, ga, 5, 7)
    for p in range(9):
        u = np.array([math.cos(ga[p]), 0, -math.sin(ga[p])])
        v = np.array([math.cos(al[p]) * math.sin(ga[p]), -math.sin(al[p]), math.cos(al[p]) * math.cos(ga[p])])
        d = np.array([math.sin(ga[p]) * math.sin(al[p]), math.cos(al[p]), math.cos(ga[p]) * math.sin(al[p])])
        assert abs(u @ v) < 1e-14 and abs(u @ u - 1) < 1e-14 and abs(v @ v - 1) < 1e-14
        np.testing.assert_allclose(np.cross(u, v), -d, atol=1e-14)
        pts = np.stack([x[p].ravel(), y[p].ravel(), z[p].ravel()], 1)
        assert np.max(np.abs(pts @ d)) < 1e-13   # every point lies in the plane normal to d
        # the point (R-1-j, C-1-k) is the negative of (j, k)
        np.testing.assert_allclose(pts[::-1], -pts, atol=1e-13)


def test_beam_dir_matches_reference_fixture():
    ga, al = FIX["beam_gamma_alpha"]
    want = FIX["beam_dir64"]
    got = np.stack([[math.sin(g) * math.sin(a), math.cos(a), math.cos(g) * math.sin(a)] for g, a in zip(ga, al)])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)


# ---------------------------------------------------------------- NUFFT
@pytest.mark.parametrize("eps,tol", [(1e-9, 5e-9), (1e-6, 5e-6), (1.2e-6, 5e-6)])
def test_nufft_vs_brute_force(orc, eps, tol):
    dims, P, R, C = (5, 3, 7), 4, 7, 9
    rng = np.random.default_rng(11)
    al, ga = rng.uniform(0, np.pi, P), rng.uniform(0, 2 * np.pi, P)
    f = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
    c = rng.standard_normal((P, R, C)) + 1j * rng.standard_normal((P, R, C))
    assert rel(orc.nufft(2, dims, al, ga, R, C, f, eps=eps), orc.nufft(2, dims, al, ga, R, C, f, brute=True)) < tol
    assert rel(orc.nufft(1, dims, al, ga, R, C, c, eps=eps), orc.nufft(1, dims, al, ga, R, C, c, brute=True)) < tol


def test_nufft_single_mode_is_plane_wave(orc):
    """f = delta at mode k0 -> c_p = exp(-i k0 . x_p) exactly (brute force and fast)."""
    dims, P, R, C = (5, 3, 7), 3, 7, 9
    al, ga = np.array([0.3, 1.0, 2.0]), np.array([0.1, 0.7, 2.5])
    f = np.zeros(dims, complex)
    idx = (3, 0, 5)
    f[idx] = 1.0
    k = np.array([idx[2] - dims[2] // 2, idx[1] - dims[1] // 2, idx[0] - dims[0] // 2])
    x, y, z = orc.grid_coords(al, ga, R, C)
    want = np.exp(-1j * (k[0] * x + k[1] * y + k[2] * z))
    assert rel(orc.nufft(2, dims, al, ga, R, C, f, brute=True), want) < 1e-14
    assert rel(orc.nufft(2, dims, al, ga, R, C, f, eps=1e-9), want) < 5e-9


# ---------------------------------------------------------------- projectors and normal operator
def _dirichlet(v, n):
    den = np.sin(np.pi * v / n)
    small = np.abs(den) < 1e-12
    return np.where(small, n * 1.0, np.sin(np.pi * v) / np.where(small, 1.0, den))


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_forward_voxel_delta_closed_form(orc, comp):
    dims, P, R, C = (5, 3, 7), 5, 7, 9
    rng = np.random.default_rng(1)
    al, ga = rng.uniform(0, np.pi, P), rng.uniform(0, np.pi, P)
    D = (1, -1, 2)
    m = np.zeros((3,) + dims)
    m[comp, dims[0] // 2 + D[0], dims[1] // 2 + D[1], dims[2] // 2 + D[2]] = 1
    p = orc.forward(m, al, ga, R, C)
    Dx, Dy, Dz = D[2], D[1], D[0]
    want = np.zeros((P, R, C))
    for j in range(P):
        u = np.array([np.cos(ga[j]), 0, -np.sin(ga[j])])
        v = np.array([np.cos(al[j]) * np.sin(ga[j]), -np.sin(al[j]), np.cos(al[j]) * np.cos(ga[j])])
        d = [np.sin(ga[j]) * np.sin(al[j]), np.cos(al[j]), np.cos(ga[j]) * np.sin(al[j])]
        Du, Dv = Dx * u[0] + Dz * u[2], Dx * v[0] + Dy * v[1] + Dz * v[2]
        rr, cc = np.arange(R)[:, None], np.arange(C)[None, :]
        want[j] = d[comp] * _dirichlet(rr - (R - 1) / 2 - Dv, R) / R * _dirichlet(cc - (C - 1) / 2 - Du, C) / C
    assert np.max(np.abs(p - want)) < 1e-9


def test_forward_adjoint_are_adjoint_and_compose_to_sysmat(orc):
    dims, P, R, C = (5, 3, 7), 5, 7, 9
    rng = np.random.default_rng(2)
    al, ga = rng.uniform(0, np.pi, P), rng.uniform(0, np.pi, P)
    m = rng.standard_normal((3,) + dims)
    y = rng.standard_normal((P, R, C))
    lhs = np.vdot(orc.forward(m, al, ga, R, C), y)
    rhs = np.vdot(m, orc.adjoint(y, dims, al, ga))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    s = orc.sysmat(m, [(al, ga, R, C)])
    assert rel(orc.adjoint(orc.forward(m, al, ga, R, C), dims, al, ga), s) < 1e-12


def test_toeplitz_closed_form_equals_sysmat(orc):
    dims, P, R, C = (3, 3, 5), 4, 7, 9
    rng = np.random.default_rng(4)
    al, ga = rng.uniform(0, np.pi, P), rng.uniform(0, np.pi, P)
    m = rng.standard_normal((3,) + dims)
    n1, n2, n3 = dims
    Ds = np.array([(dx, dy, dz) for dz in range(-n1 + 1, n1) for dy in range(-n2 + 1, n2) for dx in range(-n3 + 1, n3)])
    K = dict(zip(map(tuple, Ds), orc.toeplitz_kernel(al, ga, R, C, Ds)))
    pair = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    out = np.zeros_like(m)
    for i1 in range(n1):
        for i2 in range(n2):
            for i3 in range(n3):
                for j1 in range(n1):
                    for j2 in range(n2):
                        for j3 in range(n3):
                            k6 = K[(i3 - j3, i2 - j2, i1 - j1)]
                            for a in range(3):
                                for b in range(3):
                                    out[a, i1, i2, i3] += k6[pair[(min(a, b), max(a, b))]] * m[b, j1, j2, j3]
    assert rel(out, orc.sysmat(m, [(al, ga, R, C)])) < 1e-9


# ---------------------------------------------------------------- preconditioner, stencils, prior
@pytest.mark.parametrize("shape", [(4, 7, 9), (3, 8, 10), (2, 21, 89)])
def test_precond_matches_numpy_rfft2(orc, shape):
    x = np.random.default_rng(5).standard_normal(shape)
    n2, n3 = shape[1:]
    xf = np.arange(n3 // 2 + 1) / n3
    yf = np.array([j / n2 if j <= n2 // 2 else (j - n2) / n2 for j in range(n2)])
    F = np.sqrt(xf[None, :] ** 2 + yf[:, None] ** 2)
    want = np.fft.irfft2(np.fft.rfft2(x) * F, s=(n2, n3))
    assert np.max(np.abs(orc.precond(x) - want)) < 1e-12


@pytest.mark.parametrize("prec", [32, 64])
def test_stencils_and_prior_match_reference_fixture(orc, prec):
    u, v, g0 = FIX["fd_u"], FIX["fd_v"], FIX["fd_g0"]
    np.testing.assert_array_equal(orc.grad_u(u, prec), FIX[f"grad_u_{prec}"])
    np.testing.assert_array_equal(orc.divergence(v, prec), FIX[f"divergence_{prec}"])
    np.testing.assert_array_equal(orc.divergence(orc.grad_u(u, prec), prec), FIX[f"laplacian_{prec}"])
    np.testing.assert_array_equal(orc.qggmrf(g0, u, 1000.0, 1.2, prec), FIX[f"qggmrf_{prec}"])
    np.testing.assert_array_equal(orc.qggmrf(g0, u, 2.0, 1.1, prec), FIX[f"qggmrf_s2_{prec}"])


def test_divergence_is_not_the_adjoint_of_grad(orc):
    """The reference's interior-only stencils are not adjoint (the laplacian is non-symmetric)."""
    rng = np.random.default_rng(6)
    u, d = rng.standard_normal((6, 5, 7)), rng.standard_normal((3, 6, 5, 7))
    a = np.vdot(orc.grad_u(u), d)
    b = -np.vdot(u, orc.divergence(d))
    assert abs(a - b) > 1e-3 * abs(a)


def test_planted_sign_error_is_caught(orc):
    """A dropped minus in the divergence (planted here, on the CPU) fails the fixture check."""
    v = FIX["fd_v"]
    wrong = -orc.divergence(v, 64)
    assert not np.array_equal(wrong, FIX["divergence_64"])


# ---------------------------------------------------------------- solvers
def test_cg_on_diagonal_operator_reaches_the_preconditioned_optimum(orc):
    """tests/test_cg.cpp:16-54 (A = diag): the ramp preconditioner zeroes the DC of every
    (n2, n3) slice (include/precond.h:45-61), so PCG from x0 = 0 stays in the zero-DC
    subspace V and converges to argmin_{x in V} ||x - A^-1 b||_A = Q (Q'AQ)^-1 Q'b."""
    from scipy.linalg import null_space

    dims = (3, 4, 5)
    rng = np.random.default_rng(7)
    diag = 1.0 + np.arange(3 * np.prod(dims)).reshape((3,) + dims) / np.prod(dims)
    b = rng.standard_normal((3,) + dims)
    N = b.size
    B = np.zeros((3 * dims[0], N))
    for c in range(3):
        for s in range(dims[0]):
            m = np.zeros((3,) + dims)
            m[c, s] = 1.0
            B[c * dims[0] + s] = m.ravel()
    Q = null_space(B)
    A = np.diag(diag.ravel())
    x_opt = Q @ np.linalg.solve(Q.T @ A @ Q, Q.T @ b.ravel())
    x, it = orc.cg_diag(diag, b, np.zeros_like(b), 500, 1e-12)
    assert it < 500                     # stopped by |pAp| < 1e-10 (src/conjgrad.cpp:73-76)
    assert np.max(np.abs(B @ x.ravel())) < 1e-12          # stays in the zero-DC subspace
    xs_ = (b / diag).ravel()
    obj = lambda v: float((v - xs_) @ (diag.ravel() * (v - xs_)))   # noqa: E731
    assert abs(obj(x.ravel()) - obj(x_opt)) < 1e-9 * obj(x_opt)
    assert rel(x.ravel(), x_opt) < 1e-5   # the absolute |pAp| break stops just short of it
    # A-norm error is monotone over the iterations
    xs = b / diag
    errs = []
    for k in (1, 2, 5, 10, 20):
        xk, _ = orc.cg_diag(diag, b, np.zeros_like(b), k, 0.0)
        e = (xk - xs).ravel()
        errs.append(float(e @ (diag.ravel() * e)))
    assert all(e2 <= e1 * (1 + 1e-12) for e1, e2 in zip(errs, errs[1:]))
