"""CPU oracle for the laminography reconstruction path -- TEST INFRASTRUCTURE.

Thin ctypes wrapper over ``oracle/liboracle.so`` (a C++ restatement of the
reference, see ``oracle/oracle.cpp`` for the file:line map) and, where it has
been built, ``oracle/_ref/libref.so`` (the reference's own geometry, padding,
shift, finite-difference and qGGMRF sources, compiled where they lie).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker / baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_longlong)


def build(ref_dir: str = "/root/reference") -> None:
    """Compile liboracle.so (and _ref/libref.so when the reference is present)."""
    targets = ["all"]
    if os.path.isdir(os.path.join(ref_dir, "src")):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", _HERE, f"REF={ref_dir}", *targets], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        _LIB.orc_last_error.restype = ctypes.c_char_p
    return _LIB


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libref.so"))


def ref():
    global _REF
    if _REF is None:
        _REF = ctypes.CDLL(os.path.join(_HERE, "_ref", "libref.so"))
    return _REF


def _d(a):
    return a.ctypes.data_as(_dp)


def _l(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(_lp)


def _chk(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def _cplx_in(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


def set_threads(n: int) -> int:
    return lib().orc_set_threads(ctypes.c_int(n))


# ---------------------------------------------------------------- FFT / shifts
def fft1d(x: np.ndarray, sign: int) -> np.ndarray:
    x = np.array(x, dtype=np.complex128, copy=True)
    n = x.shape[-1]
    batch = x.size // n
    buf = x.view(np.float64)
    _chk(lib().orc_fft1d(_d(buf), ctypes.c_longlong(n), ctypes.c_longlong(batch), ctypes.c_int(sign)))
    return x


def shift2(x: np.ndarray, inverse: bool) -> np.ndarray:
    x = np.array(x, dtype=np.complex128, copy=True)
    dims, dp = _l(x.shape)
    _chk(lib().orc_shift2(_d(x.view(np.float64)), dp, ctypes.c_int(int(inverse))))
    return x


# ---------------------------------------------------------------- geometry
def _angles(alpha, gamma, P):
    alpha = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, dtype=np.float64), (P,)))
    gamma = np.ascontiguousarray(np.broadcast_to(np.asarray(gamma, dtype=np.float64), (P,)))
    return alpha, gamma


def grid_coords(alpha, gamma, R, C, prec=64):
    """Per-projection (alpha_p, gamma_p) polar grid; returns x, y, z of shape (P,R,C)."""
    P = len(np.atleast_1d(alpha))
    a, g = _angles(alpha, gamma, P)
    x = np.empty((P, R, C)); y = np.empty_like(x); z = np.empty_like(x)
    _chk(lib().orc_grid_coords(prec, _d(a), _d(g), ctypes.c_longlong(P), ctypes.c_longlong(R),
                               ctypes.c_longlong(C), _d(x), _d(y), _d(z)))
    return x, y, z


def nufft(type_: int, dims, alpha, gamma, R, C, data, eps=1e-9, brute=False, prec=64):
    P = len(np.atleast_1d(alpha))
    a, g = _angles(alpha, gamma, P)
    dd, dp = _l(dims)
    src = _cplx_in(data)
    if type_ == 2:
        out = np.empty((P, R, C), dtype=np.complex128)
    else:
        out = np.empty(tuple(dims), dtype=np.complex128)
    _chk(lib().orc_nufft(prec, type_, int(brute), ctypes.c_double(eps), dp, _d(a), _d(g), ctypes.c_longlong(P),
                         ctypes.c_longlong(R), ctypes.c_longlong(C), _d(src), _d(out.view(np.float64))))
    return out


def nufft_info(dims, eps):
    dd, dp = _l(dims)
    nf = np.zeros(3, dtype=np.int64)
    w = ctypes.c_int(0)
    _chk(lib().orc_nufft_info(ctypes.c_double(eps), dp, nf.ctypes.data_as(_lp), ctypes.byref(w)))
    return tuple(int(v) for v in nf), w.value


# ---------------------------------------------------------------- projectors
def forward(m, alpha, gamma, R, C, eps=1e-9, prec=64, fft_threads=False):
    m = np.ascontiguousarray(m, dtype=np.float64)
    dims = m.shape[1:]
    P = len(np.atleast_1d(alpha))
    a, g = _angles(alpha, gamma, P)
    dd, dp = _l(dims)
    out = np.empty((P, R, C))
    _chk(lib().orc_forward(prec, ctypes.c_double(eps), int(fft_threads), dp, _d(m), _d(a), _d(g),
                           ctypes.c_longlong(P), ctypes.c_longlong(R), ctypes.c_longlong(C), _d(out)))
    return out


def adjoint(proj, dims, alpha, gamma, eps=1e-9, prec=64, fft_threads=False):
    proj = np.ascontiguousarray(proj, dtype=np.float64)
    P, R, C = proj.shape
    a, g = _angles(alpha, gamma, P)
    dd, dp = _l(dims)
    out = np.empty((3,) + tuple(dims))
    _chk(lib().orc_adjoint(prec, ctypes.c_double(eps), int(fft_threads), dp, _d(proj), _d(a), _d(g),
                           ctypes.c_longlong(P), ctypes.c_longlong(R), ctypes.c_longlong(C), _d(out)))
    return out


def _pack_sets(sets):
    """sets: list of (alpha, gamma, R, C)."""
    Ps = np.array([len(np.atleast_1d(s[0])) for s in sets], dtype=np.int64)
    Rs = np.array([s[2] for s in sets], dtype=np.int64)
    Cs = np.array([s[3] for s in sets], dtype=np.int64)
    al = np.concatenate([_angles(s[0], s[1], int(P))[0] for s, P in zip(sets, Ps)])
    ga = np.concatenate([_angles(s[0], s[1], int(P))[1] for s, P in zip(sets, Ps)])
    return Ps, Rs, Cs, al, ga


def sysmat(x, sets, eps=1e-9, prec=64):
    """Sum over datasets of the reference sysmat; sets = [(alpha, gamma, R, C), ...]."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    dims = x.shape[1:]
    Ps, Rs, Cs, al, ga = _pack_sets(sets)
    dd, dp = _l(dims)
    out = np.empty_like(x)
    _chk(lib().orc_sysmat(prec, ctypes.c_double(eps), dp, _d(x), len(sets), Ps.ctypes.data_as(_lp),
                          Rs.ctypes.data_as(_lp), Cs.ctypes.data_as(_lp), _d(al), _d(ga), _d(out)))
    return out


def toeplitz_kernel(alpha, gamma, R, C, deltas):
    P = len(np.atleast_1d(alpha))
    a, g = _angles(alpha, gamma, P)
    dl = np.ascontiguousarray(deltas, dtype=np.int64).reshape(-1, 3)
    out = np.empty((dl.shape[0], 6))
    _chk(lib().orc_toeplitz_kernel(_d(a), _d(g), ctypes.c_longlong(P), ctypes.c_longlong(R), ctypes.c_longlong(C),
                                   dl.ctypes.data_as(_lp), ctypes.c_longlong(dl.shape[0]), _d(out)))
    return out


# ---------------------------------------------------------------- elementwise
def precond(x, prec=64):
    x = np.ascontiguousarray(x, dtype=np.float64)
    dd, dp = _l(x.shape)
    out = np.empty_like(x)
    _chk(lib().orc_precond(prec, dp, _d(x), _d(out)))
    return out


def grad_u(u, prec=64):
    u = np.ascontiguousarray(u, dtype=np.float64)
    dd, dp = _l(u.shape)
    out = np.empty((3,) + u.shape)
    _chk(lib().orc_grad_u(prec, dp, _d(u), _d(out)))
    return out


def divergence(v, prec=64):
    v = np.ascontiguousarray(v, dtype=np.float64)
    dd, dp = _l(v.shape[1:])
    out = np.empty(v.shape[1:])
    _chk(lib().orc_divergence(prec, dp, _d(v), _d(out)))
    return out


def qggmrf(g, x, sigma, p, prec=64):
    g = np.array(g, dtype=np.float64, copy=True)
    x = np.ascontiguousarray(x, dtype=np.float64)
    dd, dp = _l(x.shape)
    _chk(lib().orc_qggmrf(prec, dp, _d(g), _d(x), ctypes.c_double(sigma), ctypes.c_double(p)))
    return g


def cg_diag(diag, b, x0, max_iter, tol, prec=64):
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    dd, dp = _l(diag.shape[1:])
    out = np.empty_like(b)
    it = ctypes.c_int(0)
    _chk(lib().orc_cg_diag(prec, dp, _d(diag), _d(b), _d(x0), ctypes.c_longlong(max_iter), ctypes.c_double(tol),
                           _d(out), ctypes.byref(it)))
    return out, it.value


def mbir(method, datasets, recon_dims, *, max_iters, inner_iters=3, lam=0.1, mu=10.0, sigma=1000.0, p=1.2,
         tol=0.0, xtol=0.0, eps=1e-9, prec=64):
    """datasets: list of (proj (P,R,C), alpha (P,), gamma (P,) or scalar)."""
    Ps = np.array([d[0].shape[0] for d in datasets], dtype=np.int64)
    Rs = np.array([d[0].shape[1] for d in datasets], dtype=np.int64)
    Cs = np.array([d[0].shape[2] for d in datasets], dtype=np.int64)
    proj = np.concatenate([np.ascontiguousarray(d[0], dtype=np.float64).ravel() for d in datasets])
    al = np.concatenate([_angles(d[1], d[2], int(P))[0] for d, P in zip(datasets, Ps)])
    ga = np.concatenate([_angles(d[1], d[2], int(P))[1] for d, P in zip(datasets, Ps)])
    dd, dp = _l(recon_dims)
    out = np.empty((3,) + tuple(recon_dims))
    _chk(lib().orc_mbir(prec, ctypes.c_double(eps), method, dp, len(datasets), Ps.ctypes.data_as(_lp),
                        Rs.ctypes.data_as(_lp), Cs.ctypes.data_as(_lp), _d(proj), _d(al), _d(ga),
                        ctypes.c_longlong(max_iters), ctypes.c_longlong(inner_iters), ctypes.c_double(lam),
                        ctypes.c_double(mu), ctypes.c_double(sigma), ctypes.c_double(p), ctypes.c_double(tol),
                        ctypes.c_double(xtol), _d(out)))
    return out


# ---------------------------------------------------------------- solvers on Python operators
_OP_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _dp, _dp)
_LOSS_CB = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, _dp)


def _wrap_op(fn, shape):
    """ctypes callback y = fn(x) on [3, *shape] double arrays."""
    n = 3 * int(np.prod(shape))

    def cb(_user, xp, yp):
        x = np.ctypeslib.as_array(xp, shape=(n,)).reshape((3,) + tuple(shape)).copy()
        y = np.asarray(fn(x), dtype=np.float64).reshape(-1)
        np.ctypeslib.as_array(yp, shape=(n,))[:] = y

    return _OP_CB(cb)


def _wrap_loss(fn, shape):
    n = 3 * int(np.prod(shape))

    def cb(_user, xp):
        x = np.ctypeslib.as_array(xp, shape=(n,)).reshape((3,) + tuple(shape)).copy()
        return float(fn(x))

    return _LOSS_CB(cb)


def cgsolver(A, b, x0, max_iter, tol, prec=64):
    """src/conjgrad.cpp:36-117 with the operator A(x) given in Python ([3, n1, n2, n3])."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    dd, dp = _l(b.shape[1:])
    cb = _wrap_op(A, b.shape[1:])
    out = np.empty_like(b)
    it = ctypes.c_int(0)
    _chk(lib().orc_cgsolver_cb(prec, dp, cb, None, _d(b), _d(x0), ctypes.c_longlong(max_iter),
                               ctypes.c_double(tol), _d(out), ctypes.byref(it)))
    return out, it.value


def split_bregman(A, yT, x0, lam, mu, outer_max, inner_max, tol=0.0, xtol=0.0, prec=64):
    """src/bregman.cpp:38-134 with the operator A(x) given in Python."""
    yT = np.ascontiguousarray(yT, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    dd, dp = _l(yT.shape[1:])
    cb = _wrap_op(A, yT.shape[1:])
    out = np.empty_like(yT)
    _chk(lib().orc_split_bregman_cb(prec, dp, cb, None, _d(yT), _d(x0), ctypes.c_double(lam), ctypes.c_double(mu),
                                    ctypes.c_longlong(outer_max), ctypes.c_longlong(inner_max), ctypes.c_double(tol),
                                    ctypes.c_double(xtol), _d(out)))
    return out


def nagopt(grad, loss, x, max_iters, lipschitz, tol=0.0, xtol=0.0, max_inner=3, prec=64):
    """src/nagopt.cpp:115-209 with grad(x) and loss(x) given in Python; returns the result
    (the reference also mutates x; the input array here is left alone)."""
    x = np.array(x, dtype=np.float64, copy=True)
    dd, dp = _l(x.shape[1:])
    gcb = _wrap_op(grad, x.shape[1:])
    lcb = _wrap_loss(loss, x.shape[1:])
    _chk(lib().orc_nagopt_cb(prec, dp, gcb, lcb, None, _d(x), ctypes.c_longlong(max_iters),
                             ctypes.c_double(lipschitz), ctypes.c_double(tol), ctypes.c_double(xtol),
                             ctypes.c_longlong(max_inner)))
    return x


def lipschitz(A, x0, max_iter=20, tol=1e-6, prec=64):
    """src/lipschitz.cpp:30-53 (power iteration) over the 3-component vector, from start x0."""
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    dd, dp = _l(x0.shape[1:])
    cb = _wrap_op(A, x0.shape[1:])
    L = ctypes.c_double(0)
    _chk(lib().orc_lipschitz_cb(prec, dp, cb, None, _d(x0), ctypes.c_longlong(max_iter), ctypes.c_double(tol),
                                ctypes.byref(L)))
    return L.value


def _pack_datasets(datasets):
    Ps = np.array([d[0].shape[0] for d in datasets], dtype=np.int64)
    Rs = np.array([d[0].shape[1] for d in datasets], dtype=np.int64)
    Cs = np.array([d[0].shape[2] for d in datasets], dtype=np.int64)
    proj = np.concatenate([np.ascontiguousarray(d[0], dtype=np.float64).ravel() for d in datasets])
    al = np.concatenate([_angles(d[1], d[2], int(P))[0] for d, P in zip(datasets, Ps)])
    ga = np.concatenate([_angles(d[1], d[2], int(P))[1] for d, P in zip(datasets, Ps)])
    return Ps, Rs, Cs, proj, al, ga


def mbir_setup(method, datasets, recon_dims, eps=1e-9, prec=64):
    """The drivers' shared set-up: (padded dims, yT [3, *padded])."""
    Ps, Rs, Cs, proj, al, ga = _pack_datasets(datasets)
    dd, dp = _l(recon_dims)
    od = np.zeros(3, np.int64)
    _chk(lib().orc_mbir_setup(prec, ctypes.c_double(eps), method, dp, len(datasets), Ps.ctypes.data_as(_lp),
                              Rs.ctypes.data_as(_lp), Cs.ctypes.data_as(_lp), _d(proj), _d(al), _d(ga),
                              od.ctypes.data_as(_lp), None))
    yT = np.empty((3,) + tuple(int(v) for v in od))
    _chk(lib().orc_mbir_setup(prec, ctypes.c_double(eps), method, dp, len(datasets), Ps.ctypes.data_as(_lp),
                              Rs.ctypes.data_as(_lp), Cs.ctypes.data_as(_lp), _d(proj), _d(al), _d(ga),
                              od.ctypes.data_as(_lp), _d(yT)))
    return tuple(int(v) for v in od), yT


def mbir3_closures(datasets, recon_dims, x, sigma=1000.0, p=1.2, eps=1e-9, prec=64):
    """MBIR3's gradient and loss closures (src/mbir3.cpp:86-117) at the padded iterate x."""
    Ps, Rs, Cs, proj, al, ga = _pack_datasets(datasets)
    dd, dp = _l(recon_dims)
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.empty_like(x)
    loss = ctypes.c_double(0)
    _chk(lib().orc_mbir3_closures(prec, ctypes.c_double(eps), dp, len(datasets), Ps.ctypes.data_as(_lp),
                                  Rs.ctypes.data_as(_lp), Cs.ctypes.data_as(_lp), _d(proj), _d(al), _d(ga),
                                  ctypes.c_double(sigma), ctypes.c_double(p), _d(x), _d(g), ctypes.byref(loss)))
    return g, loss.value


# ---------------------------------------------------------------- reference (_ref)
class Ref:
    """The reference's own compiled sources (oracle/_ref/libref.so)."""

    @staticmethod
    def grid_coords(theta, gamma, R, C, prec=64):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        P = theta.size
        x = np.empty((P, R, C)); y = np.empty_like(x); z = np.empty_like(x)
        ref().ref_grid_coords(prec, _d(theta), ctypes.c_longlong(P), ctypes.c_longlong(R), ctypes.c_longlong(C),
                              ctypes.c_double(gamma), _d(x), _d(y), _d(z))
        return x, y, z

    @staticmethod
    def beam_dir(gamma, alpha, prec=64):
        out = np.empty(3)
        ref().ref_beam_dir(prec, ctypes.c_double(gamma), ctypes.c_double(alpha), _d(out))
        return out

    @staticmethod
    def shift2(x, inverse):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        dd, dp = _l(x.shape)
        out = np.empty_like(x)
        ref().ref_shift2(dp, _d(x.view(np.float64)), int(inverse), _d(out.view(np.float64)))
        return out

    @staticmethod
    def pad2d(x, factor, prec=64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        dd, dp = _l(x.shape)
        od = np.zeros(3, dtype=np.int64)
        ref().ref_pad2d(prec, dp, _d(x), ctypes.c_double(factor), None, od.ctypes.data_as(_lp))
        out = np.empty(tuple(od))
        ref().ref_pad2d(prec, dp, _d(x), ctypes.c_double(factor), _d(out), od.ctypes.data_as(_lp))
        return out

    @staticmethod
    def crop3d(x, new_dims):
        x = np.ascontiguousarray(x, dtype=np.float64)
        dd, dp = _l(x.shape)
        nd, ndp = _l(new_dims)
        out = np.empty(tuple(new_dims))
        ref().ref_crop3d(dp, _d(x), ndp, _d(out))
        return out

    @staticmethod
    def grad_u(u, prec=64):
        u = np.ascontiguousarray(u, dtype=np.float64)
        dd, dp = _l(u.shape)
        out = np.empty((3,) + u.shape)
        ref().ref_grad_u(prec, dp, _d(u), _d(out))
        return out

    @staticmethod
    def divergence(v, prec=64):
        v = np.ascontiguousarray(v, dtype=np.float64)
        dd, dp = _l(v.shape[1:])
        out = np.empty(v.shape[1:])
        ref().ref_divergence(prec, dp, _d(v), _d(out))
        return out

    @staticmethod
    def laplacian(u, prec=64):
        u = np.ascontiguousarray(u, dtype=np.float64)
        dd, dp = _l(u.shape)
        out = np.empty_like(u)
        ref().ref_laplacian(prec, dp, _d(u), _d(out))
        return out

    @staticmethod
    def qggmrf(g, x, sigma, p, prec=64):
        g = np.array(g, dtype=np.float64, copy=True)
        x = np.ascontiguousarray(x, dtype=np.float64)
        dd, dp = _l(x.shape)
        ref().ref_qggmrf(prec, dp, _d(g), _d(x), ctypes.c_double(sigma), ctypes.c_double(p))
        return g
