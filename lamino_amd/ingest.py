"""Request ingest and result output: the reference's `recon` inputs and outputs.

  load_config        include/config.h:40-55, 87-141, 161-245, 260-300 (TOML:
                     [[input]] datasets, [recon_params], [output])
  read_angles        include/config.h:61-83 (degrees when max > 2 pi)
  read_tiff          include/tiff.h:31-85 (the detector odd-ification quirk)
  mask_infs_nans     include/mask.h:22-32 (on the device)
  transpose120       include/array_ops.h:117-137 as used by src/recon.cpp:76-78
                     (on the device)
  write_tiff3 / write_tiff_vectors / write_vti
                     include/tiff.h:87-188, include/write_vti.h:36-111
  recon              src/recon.cpp:30-112: config -> MBIR1/2/3 -> transpose ->
                     files

The readers and writers are the library's own native code (lam_tiff_*,
lam_vti_*, lam_read_angles); TOML is parsed with tomli.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os
import sys

import numpy as np

from . import _lib
from ._lib import check
from .api import ReconParams, _dims, default_context

F32_PI = math.pi


class ConfigError(RuntimeError):
    """The reference's std::runtime_error for a bad configuration."""


# ---------------------------------------------------------------- files
def read_tiff(path: str) -> np.ndarray:
    """tiff::read: (pages, rows, cols) float32 with the reference's odd-ification."""
    lib = _lib.load()
    d = (ctypes.c_int64 * 3)()
    check(lib.lam_tiff_info(os.fsencode(path), d))
    out = np.empty(tuple(d), np.float32)
    check(lib.lam_tiff_read(os.fsencode(path), out.ctypes.data_as(ctypes.c_void_p), d))
    return out


def write_tiff(path: str, data: np.ndarray) -> None:
    """tiff::write: one float32 page per slice."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    check(_lib.load().lam_tiff_write(os.fsencode(path), a.ctypes.data_as(ctypes.c_void_p), _dims(a.shape)))


def _three(m):
    cs = [np.ascontiguousarray(m[i], dtype=np.float32) for i in range(3)]
    if not (cs[0].shape == cs[1].shape == cs[2].shape):
        raise ValueError("the three components must have the same shape")
    return cs, [c.ctypes.data_as(ctypes.c_void_p) for c in cs]


def write_tiff3(name: str, m) -> list:
    """tiff::write3: <name without .tif...>{0,1,2}.tiff; returns the paths."""
    cs, ps = _three(m)
    check(_lib.load().lam_tiff_write3(os.fsencode(name), *ps, _dims(cs[0].shape)))
    base = name[:name.rfind(".tif")] if ".tif" in name else name
    return [f"{base}{i}.tiff" for i in range(3)]


def write_tiff_vectors(path: str, m) -> None:
    """tiff::write_vectors: 3 float samples per pixel, LZW."""
    cs, ps = _three(m)
    check(_lib.load().lam_tiff_write_vectors(os.fsencode(path), *ps, _dims(cs[0].shape)))


def write_vti(path: str, m, field_name: str = "magnetization") -> str:
    """vti::write_vectors; returns the written path (extension replaced by .vti)."""
    cs, ps = _three(m)
    check(_lib.load().lam_vti_write_vectors(os.fsencode(path), *ps, _dims(cs[0].shape), field_name.encode()))
    root, _ = os.path.splitext(path)
    return root + ".vti"


def read_angles(path: str) -> np.ndarray:
    """read_angles_file<float>: radians (converted from degrees when max > 2 pi)."""
    lib = _lib.load()
    n = ctypes.c_int64()
    check(lib.lam_read_angles(os.fsencode(path), None, 0, ctypes.byref(n)))
    out = np.empty(n.value, np.float64)
    check(lib.lam_read_angles(os.fsencode(path), out.ctypes.data_as(_lib.c_dp), n.value, ctypes.byref(n)))
    return out


# ---------------------------------------------------------------- device steps
def mask_infs_nans(a: np.ndarray, ctx=None) -> np.ndarray:
    """mask_infs_nans on the device: non-finite values -> 0."""
    ctx = ctx or default_context()
    a = np.ascontiguousarray(a, dtype=np.float32)
    d = ctx.to_device(a)
    check(ctx.lib.lam_mask_infs_nans(ctx.h, int(a.size), d.ptr))
    return d.download()


def transpose120(vol: np.ndarray, ctx=None) -> np.ndarray:
    """array::transpose(a, {1, 2, 0}) on the device: [n1][n2][n3] -> [n2][n3][n1]."""
    ctx = ctx or default_context()
    v = np.ascontiguousarray(vol, dtype=np.float32)
    d = ctx.to_device(v)
    o = ctx.alloc((v.shape[1], v.shape[2], v.shape[0]), np.float32)
    check(ctx.lib.lam_transpose120(ctx.h, _dims(v.shape), d.ptr, o.ptr))
    return o.download()


# ---------------------------------------------------------------- configuration
@dataclasses.dataclass
class OutputParams:
    """include/config.h:260-300."""

    filepath: str = "./recon.tiff"
    formats: list = dataclasses.field(default_factory=lambda: ["tiff", "vti"])

    def has_format(self, fmt: str) -> bool:
        return fmt in self.formats


@dataclasses.dataclass
class Request:
    """One reconstruction request: what src/recon.cpp builds from its TOML file."""

    method: int                  # 1 unconstrained CG, 2 split-Bregman, 3 qGGMRF
    datasets: list               # [(proj (P, R, C) float32, angles (P,) rad, gamma rad)]
    recon_dims: tuple            # [recon_params] recon_dims, odd-ified (d0, d1, d2)
    params: ReconParams
    output: OutputParams

    @property
    def internal_dims(self) -> tuple:
        """src/recon.cpp:60: (n1, n2, n3) = (d2, d0, d1)."""
        d = self.recon_dims
        return (d[2], d[0], d[1])


def _toml(path: str) -> dict:
    import tomli

    if not os.path.exists(path):
        raise ConfigError(f"TOML file does not exist: {path}")
    try:
        with open(path, "rb") as f:
            return tomli.load(f)
    except tomli.TOMLDecodeError as e:
        raise ConfigError(f"Failed to parse TOML file '{path}': {e}") from e


def _num(v) -> bool:
    return isinstance(v, (int, float)) and not isinstance(v, bool)


def _f32(v) -> float:
    return float(np.float32(v))


def parse_input_datasets(config: dict, ctx=None) -> list:
    """include/config.h:87-141: every [[input]] as (projections, angles, gamma in rad)."""
    inputs = config.get("input")
    if not isinstance(inputs, list):
        raise ConfigError("Missing [[input]] array in TOML file")
    out = []
    for entry in inputs:
        if not isinstance(entry, dict):
            raise ConfigError("Invalid [[input]] entry")
        if not all(k in entry for k in ("filename", "angles", "gamma")):
            raise ConfigError("[[input]] entry must have 'filename', 'angles', and 'gamma' fields")
        fn = entry["filename"]
        if not isinstance(fn, str):
            raise ConfigError("[[input]] 'filename' must be a string")
        if not os.path.exists(fn):
            raise ConfigError(f"Projection file does not exist: {fn}")
        af = entry["angles"]
        if not isinstance(af, str):
            raise ConfigError("[[input]] 'angles' must be a string")
        if not os.path.exists(af):
            raise ConfigError(f"Angles file does not exist: {af}")
        g = entry["gamma"]
        if not _num(g):
            raise ConfigError("[[input]] 'gamma' field must be a number")
        proj = mask_infs_nans(read_tiff(fn), ctx)
        angles = read_angles(af)
        gamma = _f32(np.float32(g) * np.float64(math.pi) / np.float32(180.0))
        out.append((proj, angles, gamma))
    return out


def parse_recon_params(config: dict):
    """include/config.h:161-245: (method, recon dims (odd), ReconParams).  Note the
    reference reads `max_iters` (not the `max_outer_iters` its template writes) and
    `inner_iters` only under the split-Bregman table."""
    recon = config.get("recon_params")
    if not isinstance(recon, dict):
        raise ConfigError("Missing [recon_params] section in config file")
    p = ReconParams()
    mi = recon.get("max_iters", 50)
    p.max_iters = int(mi) if isinstance(mi, int) and not isinstance(mi, bool) and mi >= 0 else 50
    dims = recon.get("recon_dims")
    if not (isinstance(dims, list) and len(dims) == 3):
        raise ConfigError("[recon_params] 'recon_dims' must be an array of three integers")
    rd = []
    for i, v in enumerate(dims):
        t = int(v) if isinstance(v, int) and not isinstance(v, bool) and v > 0 else 0
        if t == 0:
            raise ConfigError(f"[recon_params] 'recon_dims[{i}]' must be a positive integer")
        rd.append(t - 1 if t % 2 == 0 else t)
    if np.float32(rd[0]) / np.float32(rd[2]) > np.float32(0.15):
        print(f"Warning: recon_dims[0] / recon_dims[2] = {rd[0] / rd[2]:.2f} > 0.15. In laminography, the thickness "
              "(recon_dims[0]) is expected to be much smaller than the in-plane dimensions (recon_dims[1], "
              "recon_dims[2]). This may lead to increased memory usage.", file=sys.stderr)
    p.tol = _f32(recon["tol"]) if _num(recon.get("tol")) else _f32(1e-5)
    p.xtol = _f32(recon["xtol"]) if _num(recon.get("xtol")) else _f32(1e-5)
    p.sigma, p.p, p.lam, p.mu = _f32(1000.0), _f32(1.2), _f32(0.1), _f32(10.0)
    method = 1
    reg = recon.get("regularizer")
    if isinstance(reg, dict):
        m = reg.get("method", "split_bregman")
        m = m if isinstance(m, str) else "split_bregman"
        if m == "qGGMRF":
            method = 3
            t = reg.get("qGGMRF")
            if not isinstance(t, dict):
                raise ConfigError("Missing [recon_params.regularizer.qGGMRF] section in config file")
            p.sigma = _f32(t["sigma"]) if _num(t.get("sigma")) else _f32(1000.0)
            p.p = _f32(t["p"]) if _num(t.get("p")) else _f32(1.2)
        elif m == "split_bregman":
            method = 2
            t = reg.get("split_bregman")
            if not isinstance(t, dict):
                raise ConfigError("Missing [recon_params.regularizer.split_bregman] section in config file")
            p.lam = _f32(t["lambda"]) if _num(t.get("lambda")) else _f32(0.1)
            p.mu = _f32(t["mu"]) if _num(t.get("mu")) else _f32(10.0)
            ii = t.get("inner_iters", 3)
            p.inner_iters = int(ii) if isinstance(ii, int) and not isinstance(ii, bool) and ii >= 0 else 3
        else:
            raise ConfigError("[recon_params] 'regularizer' must be either 'qGGMRF' or 'split_bregman'")
    return method, tuple(rd), p


def parse_output(config: dict) -> OutputParams:
    """include/config.h:260-300."""
    out = config.get("output")
    if out is None:
        raise ConfigError("Missing [output] section in config file")
    fp = out.get("filename", "./recon.tiff") if isinstance(out, dict) else "./recon.tiff"
    fp = fp if isinstance(fp, str) else "./recon.tiff"
    fmts = out.get("formats") if isinstance(out, dict) else None
    if isinstance(fmts, list):
        got = []
        for f in fmts:
            if not isinstance(f, str):
                raise ConfigError("[output] 'formats' array must contain strings")
            if f not in ("tiff", "vti"):
                raise ConfigError(f"[output] invalid format '{f}'. Must be 'tiff' or 'vti'")
            if f not in got:
                got.append(f)
    else:
        got = ["tiff", "vti"]
    return OutputParams(fp, got)


def load_config(path: str, ctx=None) -> Request:
    """A TOML configuration file as a reconstruction request (src/recon.cpp:41-45)."""
    cfg = _toml(path)
    datasets = parse_input_datasets(cfg, ctx)
    method, rd, params = parse_recon_params(cfg)
    return Request(method, datasets, rd, params, parse_output(cfg))


def write_outputs(req: Request, result) -> list:
    """src/recon.cpp:100-110: the transposed result as tiff::write3 and/or vti."""
    d = os.path.dirname(req.output.filepath)
    if d and not os.path.exists(d):
        os.makedirs(d, exist_ok=True)
    written = []
    if req.output.has_format("tiff"):
        written += write_tiff3(req.output.filepath, result)
    if req.output.has_format("vti"):
        written.append(write_vti(req.output.filepath, result))
    return written


def recon(path: str, ctx=None):
    """src/recon.cpp: load the configuration, reconstruct (MBIR1/2/3 by the regularizer),
    transpose {1, 2, 0} on the device, write the files.  Returns (result (3, d0, d1, d2),
    written paths)."""
    from .api import _mbir

    ctx = ctx or default_context()
    req = load_config(path, ctx)
    x = _mbir(req.method, req.datasets, req.internal_dims, req.params, ctx=ctx)
    out = np.stack([transpose120(x[i], ctx) for i in range(3)])
    return out, write_outputs(req, out)


# ---------------------------------------------------------------- forward simulation
def forward_simulate(m, angles, gamma: float, out_dims, eps: float = 1e-5, ctx=None) -> np.ndarray:
    """src/forward.cpp:123-183 on the device: components m (3, d0, d1, d2) as read from
    TIFF -> transpose {2, 0, 1} -> pad3d by sqrt(2) - 1 -> forward projection on
    PolarGrid(angles, padded(rows), padded(cols), gamma) -> crop2d to out_dims."""
    ctx = ctx or default_context()
    cs = [np.ascontiguousarray(m[i], dtype=np.float32) for i in range(3)]
    a = np.ascontiguousarray(angles, dtype=np.float64)
    rows, cols = int(out_dims[0]), int(out_dims[1])
    out = np.empty((a.size, rows, cols), np.float32)
    check(ctx.lib.lam_forward_simulate(ctx.h, *[c.ctypes.data_as(ctypes.c_void_p) for c in cs], _dims(cs[0].shape),
                                       a.ctypes.data_as(_lib.c_dp), a.size, float(gamma), rows, cols, float(eps),
                                       out.ctypes.data_as(ctypes.c_void_p)))
    return out


def forward_from_config(path: str, ctx=None):
    """src/forward.cpp main: [input] data_path + component_1..3 (TIFF), [angles] filename,
    [orientation] gamma (degrees), [output] basedir / filename and dims = [rows, cols].
    Writes the projections as a TIFF stack; returns (projections, output path)."""
    cfg = _toml(path)
    try:
        inp, ang, ori, out = cfg["input"], cfg["angles"], cfg["orientation"], cfg["output"]
        base = inp["data_path"]
        comps = [inp["component_1"], inp["component_2"], inp["component_3"]]
        angles_file = ang["filename"]
        # float gamma * M_PI / 180.0f (src/forward.cpp:44), rounded back to float
        gamma = float(np.float32(float(np.float32(ori["gamma"])) * math.pi / float(np.float32(180.0))))
        out_file, out_dir = out["filename"], out["basedir"]
        dims = out["dims"]
    except (KeyError, TypeError) as e:
        raise ConfigError(f"missing key in forward-simulation config: {e}") from e
    if not (isinstance(dims, list) and len(dims) == 2 and all(isinstance(v, int) for v in dims)):
        raise ConfigError("Output dimensions must be an array of two integers")
    os.makedirs(out_dir, exist_ok=True)
    for c in comps:
        if not os.path.exists(os.path.join(base, c)):
            raise ConfigError(f"One or more components do not exist in data path: {base}")
    angles = read_angles(angles_file)
    m = [read_tiff(os.path.join(base, c)) for c in comps]
    proj = forward_simulate(m, angles, gamma, dims, ctx=ctx)
    dst = os.path.join(out_dir, out_file)
    write_tiff(dst, proj)
    return proj, dst


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("Usage: python -m lamino_amd.ingest <input.toml>", file=sys.stderr)
        return 1
    out, written = recon(argv[0])
    print("Reconstruction saved to", ", ".join(written))
    return 0


if __name__ == "__main__":
    sys.exit(main())
