"""Python mirror of the reference's operator interface, over the gfx950 C ABI.

Names, argument meaning and errors follow the reference (lbl-camera/lamino):
  PolarGrid                  include/polar_grid.h:33-110
  beam_dir_vector            include/projection.h:36-41
  forward / adjoint          include/projection.h:52-66, src/projection.cpp:41-126
  nufft3d1 / nufft3d2        include/nufft.h:37-65
  sysmat/gradient/residual   include/tomocam.h:50-82, src/gradient.cpp:33-127
  RampPreconditioner         include/precond.h:33-87
  grad_u/divergence/laplacian include/bregman.h:15-36
  qggmrf, cgsolver, split_bregman, nagopt, lipschitz  include/optimize.h:74-153
  MBIR1/MBIR2/MBIR3          include/tomocam.h:101-154
Host arrays are numpy; every operation runs on the GPU through
liblamino_amd.so (device buffers are managed here).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Callable, Sequence

import numpy as np

from . import _lib
from ._lib import check

# ---------------------------------------------------------------- context / memory


class Context:
    """A device + stream + plan cache (lam_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.lam_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            check(self.lib.lam_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.lib.lam_ctx_sync(self.h))

    def alloc(self, shape, dtype=np.float32) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def to_device(self, a: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d


class DeviceArray:
    """Device memory owned through lam_malloc / lam_free."""

    def __init__(self, ctx: Context, shape, dtype=np.float32):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, int) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(ctx.lib.lam_malloc(ctx.h, max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.ctx.lib.lam_free(self.ctx.h, ctypes.c_void_p(self.ptr))
        except Exception:
            pass
        self.ptr = None

    def at(self, offset_elems: int = 0) -> int:
        return self.ptr + offset_elems * self.dtype.itemsize

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(self.ctx.lib.lam_upload(self.ctx.h, ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                                      self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(self.ctx.lib.lam_download(self.ctx.h, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                                        self.nbytes))
        return out

    def zero(self):
        check(self.ctx.lib.lam_memset(self.ctx.h, ctypes.c_void_p(self.ptr), 0, self.nbytes))

    def comps(self):
        """Device pointers of the 3 components of a (3, n1, n2, n3) array."""
        n = int(np.prod(self.shape[1:]))
        return [self.at(i * n) for i in range(self.shape[0])]


_DEFAULT: Context | None = None


def default_context() -> Context:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


def _dims(d) -> ctypes.Array:
    d = tuple(int(v) for v in d)
    if len(d) != 3:
        raise ValueError("dims must have 3 entries")
    return (ctypes.c_int64 * 3)(*d)


def _ptrs(ps):
    return (ctypes.c_void_p * 3)(*[ctypes.c_void_p(p) for p in ps])


DEFAULT_EPS = 1e-5

# ---------------------------------------------------------------- geometry


def beam_dir_vector(gamma: float, alpha: float):
    """include/projection.h:36-41."""
    return np.array([math.sin(gamma) * math.sin(alpha), math.cos(alpha), math.cos(gamma) * math.sin(alpha)])


class PolarGrid:
    """include/polar_grid.h:33-110; coordinates are generated on the device.

    ``PolarGrid(angles, nrows, ncols, gamma)`` is the reference constructor;
    ``PolarGrid.per_projection(alpha, gamma, nrows, ncols)`` takes one
    (alpha_p, gamma_p) pair per projection.
    """

    def __init__(self, angles, nrows: int, ncols: int, gamma: float, ctx: Context | None = None, _pp=None):
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        if _pp is None:
            th = np.ascontiguousarray(angles, dtype=np.float64)
            check(self.ctx.lib.lam_grid_polar(self.ctx.h, th.ctypes.data_as(ctypes.c_void_p), th.size, int(nrows),
                                              int(ncols), float(gamma), ctypes.byref(h)))
            self.theta = th
            self.gamma = np.full(th.size, float(gamma))
            self.literal_gamma = float(gamma)
        else:
            al, ga = _pp
            al = np.ascontiguousarray(al, dtype=np.float64)
            ga = np.ascontiguousarray(np.broadcast_to(ga, al.shape), dtype=np.float64)
            check(self.ctx.lib.lam_grid_polar_pp(self.ctx.h, al.ctypes.data_as(ctypes.c_void_p),
                                                 ga.ctypes.data_as(ctypes.c_void_p), al.size, int(nrows), int(ncols),
                                                 ctypes.byref(h)))
            self.theta = al
            self.gamma = ga
            self.literal_gamma = None
        self.h = h
        self.nrows, self.ncols = int(nrows), int(ncols)
        self._xyz = None

    @classmethod
    def per_projection(cls, alpha, gamma, nrows, ncols, ctx: Context | None = None):
        return cls(None, nrows, ncols, 0.0, ctx=ctx, _pp=(alpha, gamma))

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.lam_grid_destroy(self.h)
        except Exception:
            pass

    def dims(self):
        return (self.theta.size, self.nrows, self.ncols)

    def size(self):
        return self.theta.size * self.nrows * self.ncols

    def nprojs(self):
        return self.theta.size

    def angles(self):
        return self.theta

    def angle(self, i):
        return self.theta[i]

    def _coords(self):
        if self._xyz is None:
            shape = self.dims()
            x = np.empty(shape, np.float32); y = np.empty_like(x); z = np.empty_like(x)
            check(self.ctx.lib.lam_grid_coords_host(self.h, x.ctypes.data_as(ctypes.c_void_p),
                                                    y.ctypes.data_as(ctypes.c_void_p),
                                                    z.ctypes.data_as(ctypes.c_void_p)))
            self._xyz = (x, y, z)
        return self._xyz

    @property
    def x(self):
        return self._coords()[0]

    @property
    def y(self):
        return self._coords()[1]

    @property
    def z(self):
        return self._coords()[2]


# ---------------------------------------------------------------- NUFFT and projectors


def nufft3d2(fz: np.ndarray, pg: PolarGrid, eps: float = DEFAULT_EPS) -> np.ndarray:
    """Type 2: c_p = sum_k f_k exp(-i k.x_p); fz complex (n1, n2, n3) -> (P, R, C)."""
    ctx = pg.ctx
    f = ctx.to_device(np.asarray(fz, dtype=np.complex64))
    c = ctx.alloc(pg.dims(), np.complex64)
    check(ctx.lib.lam_nufft3d2(ctx.h, pg.h, _dims(fz.shape), eps, ctypes.c_void_p(c.ptr), ctypes.c_void_p(f.ptr)))
    return c.download()


def nufft3d1(cz: np.ndarray, dims, pg: PolarGrid, eps: float = DEFAULT_EPS) -> np.ndarray:
    """Type 1: f_k = sum_p c_p exp(+i k.x_p); cz complex (P, R, C) -> (n1, n2, n3)."""
    ctx = pg.ctx
    c = ctx.to_device(np.asarray(cz, dtype=np.complex64))
    f = ctx.alloc(tuple(dims), np.complex64)
    check(ctx.lib.lam_nufft3d1(ctx.h, pg.h, _dims(dims), eps, ctypes.c_void_p(c.ptr), ctypes.c_void_p(f.ptr)))
    return f.download()


def forward(magnetization, pg: PolarGrid, gamma: float | None = None, eps: float = DEFAULT_EPS) -> np.ndarray:
    """src/projection.cpp:41-77; magnetization (3, n1, n2, n3) -> projections (P, R, C)."""
    ctx = pg.ctx
    m = ctx.to_device(np.asarray(magnetization, dtype=np.float32))
    out = ctx.alloc(pg.dims(), np.float32)
    a, b, c = m.comps()
    check(ctx.lib.lam_forward(ctx.h, pg.h, _dims(m.shape[1:]), eps, a, b, c, out.ptr))
    return out.download()


def autotune_bins(pg: PolarGrid, recon_dims, eps: float = DEFAULT_EPS):
    """Times spreading tile shapes for this grid and volume and keeps the fastest
    for later calls with the same shapes; returns ((Tz, Ty, Tx), spread+interp ms)."""
    T = (ctypes.c_int * 3)()
    ms = ctypes.c_double()
    check(pg.ctx.lib.lam_autotune_bins(pg.ctx.h, pg.h, _dims(recon_dims), float(eps), T, ctypes.byref(ms)))
    return tuple(T), ms.value


def adjoint(proj, pg: PolarGrid, recon_dims, gamma: float | None = None, eps: float = DEFAULT_EPS) -> np.ndarray:
    """src/projection.cpp:89-126; projections (P, R, C) -> (3, n1, n2, n3)."""
    ctx = pg.ctx
    p = ctx.to_device(np.asarray(proj, dtype=np.float32))
    out = ctx.alloc((3,) + tuple(recon_dims), np.float32)
    a, b, c = out.comps()
    check(ctx.lib.lam_adjoint(ctx.h, pg.h, _dims(recon_dims), eps, p.ptr, a, b, c))
    return out.download()


def _grid_array(grids):
    arr = (ctypes.c_void_p * len(grids))(*[g.h for g in grids])
    return arr


def forward_multi(magnetization, grids, eps: float = DEFAULT_EPS) -> list:
    """The datasets loop of src/mbir3.cpp:66-83 for the forward: one projection stack per
    grid from one volume, with one fine-grid transform for all of them."""
    ctx = grids[0].ctx
    m = ctx.to_device(np.asarray(magnetization, dtype=np.float32))
    outs = [ctx.alloc(g.dims(), np.float32) for g in grids]
    ptrs = (ctypes.c_void_p * len(grids))(*[o.ptr for o in outs])
    a, b, c = m.comps()
    check(ctx.lib.lam_forward_multi(ctx.h, _grid_array(grids), len(grids), _dims(m.shape[1:]), eps, a, b, c, ptrs))
    return [o.download() for o in outs]


def adjoint_multi(projs, grids, recon_dims, eps: float = DEFAULT_EPS) -> np.ndarray:
    """yT = sum_g adjoint(proj_g, grid_g) (src/mbir3.cpp:66-83) with one fine-grid
    transform for all datasets."""
    ctx = grids[0].ctx
    ps = [ctx.to_device(np.asarray(p, dtype=np.float32)) for p in projs]
    ptrs = (ctypes.c_void_p * len(ps))(*[p.ptr for p in ps])
    out = ctx.alloc((3,) + tuple(recon_dims), np.float32)
    a, b, c = out.comps()
    check(ctx.lib.lam_adjoint_multi(ctx.h, _grid_array(grids), len(grids), _dims(recon_dims), eps, ptrs, a, b, c, 0))
    return out.download()


def sysmat(x, pg: PolarGrid, gamma: float | None = None, eps: float = DEFAULT_EPS) -> np.ndarray:
    """src/gradient.cpp:33-89 (NUFFT form)."""
    ctx = pg.ctx
    xd = ctx.to_device(np.asarray(x, dtype=np.float32))
    yd = ctx.alloc(xd.shape, np.float32)
    check(ctx.lib.lam_sysmat(ctx.h, pg.h, _dims(xd.shape[1:]), eps, _ptrs(xd.comps()), _ptrs(yd.comps())))
    return yd.download()


def gradient(f, yT, pg: PolarGrid, gamma: float | None = None, eps: float = DEFAULT_EPS) -> np.ndarray:
    """src/gradient.cpp:96-105: sysmat(f) - yT."""
    return sysmat(f, pg, gamma, eps) - np.asarray(yT, dtype=np.float32)


def residual(f, yT, pg: PolarGrid, yTy: float, gamma: float | None = None, eps: float = DEFAULT_EPS) -> float:
    """src/gradient.cpp:115-127: sum_i sqrt(f_i.A f_i - 2 f_i.yT_i + yTy)."""
    Af = sysmat(f, pg, gamma, eps).astype(np.float64)
    f = np.asarray(f, dtype=np.float64)
    yT = np.asarray(yT, dtype=np.float64)
    err = 0.0
    for i in range(3):
        err += math.sqrt(float(np.vdot(f[i], Af[i]) - 2.0 * np.vdot(f[i], yT[i]) + yTy)) \
            if (np.vdot(f[i], Af[i]) - 2.0 * np.vdot(f[i], yT[i]) + yTy) >= 0 else float("nan")
    return err


class ToeplitzOperator:
    """Exact Toeplitz form of sum_g sysmat_g for a list of grids (device resident)."""

    def __init__(self, grids: Sequence[PolarGrid], dims, keep_per_grid: bool = False):
        self.ctx = grids[0].ctx
        self.dims = tuple(int(v) for v in dims)
        self.grids = list(grids)
        arr = (ctypes.c_void_p * len(grids))(*[g.h for g in grids])
        h = ctypes.c_void_p()
        check(self.ctx.lib.lam_toep_create(self.ctx.h, arr, len(grids), _dims(dims), int(keep_per_grid),
                                           ctypes.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.lam_toep_destroy(self.h)
        except Exception:
            pass

    def apply_device(self, x: DeviceArray, y: DeviceArray, mu: float = 0.0):
        check(self.ctx.lib.lam_toep_apply(self.h, _ptrs(x.comps()), _ptrs(y.comps()), float(mu)))

    def __call__(self, x, mu: float = 0.0) -> np.ndarray:
        xd = self.ctx.to_device(np.asarray(x, dtype=np.float32))
        yd = self.ctx.alloc(xd.shape, np.float32)
        self.apply_device(xd, yd, mu)
        return yd.download()

    def quad(self, x) -> np.ndarray:
        xd = self.ctx.to_device(np.asarray(x, dtype=np.float32))
        q = np.zeros(3 * len(self.grids))
        check(self.ctx.lib.lam_toep_quad(self.h, _ptrs(xd.comps()), q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return q.reshape(len(self.grids), 3)


# ---------------------------------------------------------------- spectral / stencils


def fft2(x: np.ndarray, sign: int = -1, shifted: bool = False, ctx: Context | None = None) -> np.ndarray:
    """include/fft.h:34-114 (batched unnormalised 2-D over the last two axes)."""
    ctx = ctx or default_context()
    x = np.asarray(x, dtype=np.complex64)
    d = ctx.to_device(x)
    o = ctx.alloc(x.shape, np.complex64)
    check(ctx.lib.lam_fft2(ctx.h, _dims(x.shape), sign, int(shifted), ctypes.c_void_p(d.ptr), ctypes.c_void_p(o.ptr)))
    return o.download()


class RampPreconditioner:
    """include/precond.h:33-87."""

    def __init__(self, dims, ctx: Context | None = None):
        self.dims = tuple(int(v) for v in dims)
        self.ctx = ctx or default_context()

    def apply(self, x: np.ndarray) -> np.ndarray:
        d = self.ctx.to_device(np.asarray(x, dtype=np.float32))
        o = self.ctx.alloc(d.shape, np.float32)
        check(self.ctx.lib.lam_precond(self.ctx.h, _dims(d.shape), d.ptr, o.ptr))
        return o.download()


def grad_u(u: np.ndarray, ctx: Context | None = None) -> np.ndarray:
    ctx = ctx or default_context()
    d = ctx.to_device(np.asarray(u, dtype=np.float32))
    o = ctx.alloc((3,) + d.shape, np.float32)
    a, b, c = o.comps()
    check(ctx.lib.lam_grad_u(ctx.h, _dims(d.shape), d.ptr, a, b, c))
    return o.download()


def divergence(v: np.ndarray, ctx: Context | None = None) -> np.ndarray:
    ctx = ctx or default_context()
    d = ctx.to_device(np.asarray(v, dtype=np.float32))
    o = ctx.alloc(d.shape[1:], np.float32)
    a, b, c = d.comps()
    check(ctx.lib.lam_divergence(ctx.h, _dims(d.shape[1:]), a, b, c, o.ptr))
    return o.download()


def laplacian(u: np.ndarray, ctx: Context | None = None) -> np.ndarray:
    ctx = ctx or default_context()
    d = ctx.to_device(np.asarray(u, dtype=np.float32))
    o = ctx.alloc(d.shape, np.float32)
    check(ctx.lib.lam_laplacian(ctx.h, _dims(d.shape), d.ptr, o.ptr))
    return o.download()


def qggmrf(g: np.ndarray, x: np.ndarray, sigma: float, p: float, ctx: Context | None = None) -> np.ndarray:
    """src/qggmrf.cpp:59-95: returns g + prior gradient (the reference updates g in place)."""
    ctx = ctx or default_context()
    gd = ctx.to_device(np.asarray(g, dtype=np.float32))
    xd = ctx.to_device(np.asarray(x, dtype=np.float32))
    check(ctx.lib.lam_qggmrf(ctx.h, _dims(xd.shape), gd.ptr, xd.ptr, float(sigma), float(p)))
    out = gd.download()
    if isinstance(g, np.ndarray) and g.dtype == np.float32 and g.flags.writeable:
        g[...] = out
    return out


def dot(a: np.ndarray, b: np.ndarray, ctx: Context | None = None) -> float:
    ctx = ctx or default_context()
    ad = ctx.to_device(np.asarray(a, dtype=np.float32).ravel())
    bd = ctx.to_device(np.asarray(b, dtype=np.float32).ravel())
    out = ctypes.c_double()
    check(ctx.lib.lam_dot(ctx.h, ad.shape[0], ad.ptr, bd.ptr, ctypes.byref(out)))
    return out.value


def demag(m, ctx: Context | None = None) -> np.ndarray:
    """opt::demag (include/demag.h:13-79): H = F^-1[-k (k . F m) / |k|^2] of the three
    components m (3, n1, n2, n3), DC zeroed, k in numpy's fftfreq order on every axis."""
    ctx = ctx or default_context()
    m = np.ascontiguousarray(m, dtype=np.float32)
    md = ctx.to_device(m)
    hd = ctx.alloc(m.shape, np.float32)
    check(ctx.lib.lam_demag(ctx.h, _dims(m.shape[1:]), *md.comps(), *hd.comps()))
    return hd.download()


# ---------------------------------------------------------------- solvers


def _vec_op(ctx: Context, fn: Callable, dims):
    """Wrap a numpy VecArray -> VecArray callable as a device operator callback."""
    n = int(np.prod(dims))
    shape = (3,) + tuple(dims)

    def cb(user, xp, yp):
        try:
            xs = np.empty(shape, np.float32)
            for i in range(3):
                check(ctx.lib.lam_download(ctx.h, xs[i].ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(xp[i]),
                                           4 * n))
            ys = np.ascontiguousarray(np.asarray(fn(xs), dtype=np.float32).reshape(shape))
            for i in range(3):
                check(ctx.lib.lam_upload(ctx.h, ctypes.c_void_p(yp[i]), ys[i].ctypes.data_as(ctypes.c_void_p),
                                         4 * n))
            return 0
        except Exception:  # pragma: no cover - reported through the status code
            return 1

    return _lib.OP_FN(cb)


def cgsolver(A: Callable, b, x0, max_iters: int, tol: float, lam: float = 0.0, ctx: Context | None = None):
    """src/conjgrad.cpp:36-117 (lambda unused there as well).  Returns (x, iterations)."""
    ctx = ctx or default_context()
    b = np.asarray(b, dtype=np.float32)
    dims = b.shape[1:]
    bd, x0d = ctx.to_device(b), ctx.to_device(np.asarray(x0, dtype=np.float32))
    xd = ctx.alloc(b.shape, np.float32)
    cb = _vec_op(ctx, A, dims)
    it = ctypes.c_int64()
    check(ctx.lib.lam_cgsolver(ctx.h, _dims(dims), cb, None, _ptrs(bd.comps()), _ptrs(x0d.comps()), int(max_iters),
                               float(tol), _ptrs(xd.comps()), ctypes.byref(it)))
    return xd.download(), it.value


def split_bregman(A: Callable, yT, x0, lam: float, mu: float, outer_max: int, inner_max: int, tol: float,
                  xtol: float, ctx: Context | None = None):
    """src/bregman.cpp:38-134."""
    ctx = ctx or default_context()
    yT = np.asarray(yT, dtype=np.float32)
    x0 = np.asarray(x0, dtype=np.float32)
    if x0.shape != yT.shape:
        raise ValueError("yT is not the raw data, but backprojected data, so it should have the same dimensions as x0")
    dims = yT.shape[1:]
    yd, x0d = ctx.to_device(yT), ctx.to_device(x0)
    xd = ctx.alloc(yT.shape, np.float32)
    cb = _vec_op(ctx, A, dims)
    check(ctx.lib.lam_split_bregman(ctx.h, _dims(dims), cb, None, _ptrs(yd.comps()), _ptrs(x0d.comps()), float(lam),
                                    float(mu), int(outer_max), int(inner_max), float(tol), float(xtol),
                                    _ptrs(xd.comps())))
    return xd.download()


def nagopt(grad: Callable, loss: Callable, x: np.ndarray, max_iters: int, lipschitz: float, tol: float, xtol: float,
           max_inner_iters: int = 20, ctx: Context | None = None):
    """src/nagopt.cpp:115-209; x (float32, (3, ...)) is updated in place and returned."""
    ctx = ctx or default_context()
    dims = x.shape[1:]
    n = int(np.prod(dims))
    xd = ctx.to_device(np.asarray(x, dtype=np.float32))
    gcb = _vec_op(ctx, grad, dims)
    shape = (3,) + tuple(dims)

    def lcb(user, xp, out):
        try:
            xs = np.empty(shape, np.float32)
            for i in range(3):
                check(ctx.lib.lam_download(ctx.h, xs[i].ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(xp[i]),
                                           4 * n))
            out[0] = float(loss(xs))
            return 0
        except Exception:  # pragma: no cover
            return 1

    lfn = _lib.LOSS_FN(lcb)
    check(ctx.lib.lam_nagopt(ctx.h, _dims(dims), gcb, lfn, None, _ptrs(xd.comps()), int(max_iters), float(lipschitz),
                             float(tol), float(xtol), int(max_inner_iters)))
    res = xd.download()
    if x.dtype == np.float32 and x.flags.writeable:
        x[...] = res
    return res


def lipschitz(A: Callable, dims, max_iters: int = 20, tol: float = 1e-5, seed: int = 0,
              ctx: Context | None = None) -> float:
    """src/lipschitz.cpp:30-53 (power iteration), FP64 on the device."""
    ctx = ctx or default_context()
    cb = _vec_op(ctx, A, dims)
    L = ctypes.c_double()
    check(ctx.lib.lam_lipschitz(ctx.h, _dims(dims), cb, None, int(max_iters), float(tol), int(seed), ctypes.byref(L)))
    return L.value


# ---------------------------------------------------------------- MBIR jobs


@dataclasses.dataclass
class ReconParams:
    """include/config.h:146-158 defaults."""

    max_iters: int = 50
    inner_iters: int = 3
    sigma: float = 1000.0
    p: float = 1.2
    lam: float = 0.1
    mu: float = 10.0
    tol: float = 1e-5
    xtol: float = 1e-5
    pad_factor: float = float(np.float32(1.4142))
    eps: float = DEFAULT_EPS


def _pack_job(method: int, datasets, params: ReconParams):
    """lam_dataset array and lam_recon_params of a job, plus the arrays they point into."""
    keep = []
    arr = (_lib.lam_dataset * len(datasets))()
    for i, ds in enumerate(datasets):
        proj, angles, gamma = ds
        proj = np.ascontiguousarray(proj, dtype=np.float32)
        al = np.ascontiguousarray(angles, dtype=np.float64)
        keep += [proj, al]
        arr[i].nproj, arr[i].nrows, arr[i].ncols = proj.shape
        arr[i].proj = proj.ctypes.data_as(ctypes.c_void_p)
        arr[i].alpha = al.ctypes.data_as(ctypes.c_void_p)
        if np.ndim(gamma) == 0:
            arr[i].gamma = None
            arr[i].gamma_const = float(gamma)
        else:
            ga = np.ascontiguousarray(np.broadcast_to(gamma, al.shape), dtype=np.float64)
            keep.append(ga)
            arr[i].gamma = ga.ctypes.data_as(ctypes.c_void_p)
    pr = _lib.lam_recon_params(method, params.max_iters, params.inner_iters, params.lam, params.mu, params.sigma,
                               params.p, params.tol, params.xtol, params.pad_factor, params.eps)
    return arr, pr, keep


def mbir_footprint(method: int, datasets, recon_dims, params: ReconParams) -> int:
    """Upper bound of the device bytes one job of these shapes allocates."""
    arr, pr, _keep = _pack_job(method, datasets, params)
    out = ctypes.c_size_t()
    check(_lib.load().lam_mbir_footprint(arr, len(datasets), _dims(recon_dims), ctypes.byref(pr), ctypes.byref(out)))
    return out.value


def _mbir(method: int, datasets, recon_dims, params: ReconParams, ctx: Context | None = None, stats=False):
    ctx = ctx or default_context()
    arr, pr, _keep = _pack_job(method, datasets, params)
    out = np.empty((3,) + tuple(recon_dims), np.float32)
    st = _lib.lam_recon_stats()
    check(ctx.lib.lam_mbir(ctx.h, arr, len(datasets), _dims(recon_dims), ctypes.byref(pr),
                           out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)))
    if stats:
        return out, {f: getattr(st, f) for f, _ in st._fields_}
    return out


def MBIR1(datasets, recon_dims, params: ReconParams, **kw):
    """src/mbir_cg.cpp:47-128: unconstrained preconditioned CG."""
    return _mbir(1, datasets, recon_dims, params, **kw)


def MBIR2(datasets, recon_dims, params: ReconParams, **kw):
    """src/mbir_sb.cpp:46-114: split-Bregman isotropic vector TV."""
    return _mbir(2, datasets, recon_dims, params, **kw)


def MBIR3(datasets, recon_dims, params: ReconParams, **kw):
    """src/mbir3.cpp:46-135: Nesterov + qGGMRF."""
    return _mbir(3, datasets, recon_dims, params, **kw)
