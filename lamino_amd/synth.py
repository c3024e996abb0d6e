"""Synthetic workloads: seeded vortex magnetisation fields, scan geometries and
the reference's shape rules, for benchmarks and tests.

Shape rules followed:
  recon dims odd-ified (n even -> n-1)                     include/config.h:184-186
  detector odd-ified on read                               include/tiff.h:54-59
  (th, h, w) -> (n1, n2, n3) = (w, th, h)                   src/recon.cpp:60
  volume padding n + 2*floor(n*(PAD-1)/2), PAD = 1.4142f    src/mbir_sb.cpp:51-61
  detector padding by (PAD-1) in MBIR1, by PAD in MBIR2/3   src/mbir_cg.cpp:83, src/mbir_sb.cpp:78
  data generation: pad3d by sqrt(2)-1, project at the padded
  detector, crop2d to the odd detector                      src/forward.cpp:138-175
Randomness comes from SplitMix64 keyed by (seed, stream) so the same inputs
are produced on any host.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

PAD_FACTOR = float(np.float32(1.4142))
FWD_PADDING = 1.41421356237  # src/forward.cpp:16-17

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int, stream: int = 0):
        self.state = (seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03 + 0x632BE59BD9B4E019) & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def u64_array(self, n: int) -> np.ndarray:
        """The next n outputs, vectorised (SplitMix64 is counter based)."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & _MASK
        return z

    def normal_array(self, n: int) -> np.ndarray:
        """Box-Muller normals from a block of uniforms."""
        m = (n + 1) // 2
        u = ((self.u64_array(2 * m) >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / (1 << 53))
        r = np.sqrt(-2.0 * np.log(u[:m]))
        t = 2 * np.pi * u[m:]
        return np.concatenate([r * np.cos(t), r * np.sin(t)])[:n]


def oddify(n: int) -> int:
    return n - 1 if n % 2 == 0 else n


def pad_n(n: int, f: float) -> int:
    """n + 2*floor(n*f/2) with the product in float32 (include/padding.h:98)."""
    return int(n + 2 * (int(np.float32(n) * np.float32(f)) // 2))


def fwd_padded(n: int) -> int:
    """src/forward.cpp:149-152."""
    return int(n + 2 * int((n * (FWD_PADDING - 1)) / 2))


def recon_dims_internal(toml_dims) -> tuple:
    """[d0, d1, d2] of a config -> odd-ified (n1, n2, n3) = (d2, d0, d1)."""
    d = [oddify(int(v)) for v in toml_dims]
    return (d[2], d[0], d[1])


def solver_dims(method: int, recon_dims, detector: int):
    vpad = np.float32(PAD_FACTOR - 1.0) if method == 1 else np.float32(np.float32(PAD_FACTOR) - np.float32(1.0))
    dpad = np.float32(PAD_FACTOR - 1.0) if method == 1 else np.float32(PAD_FACTOR)
    vol = tuple(pad_n(n, vpad) for n in recon_dims)
    det = pad_n(detector, dpad)
    return vol, det


def vortex_field(dims, seed: int, nvortex: int = 8) -> np.ndarray:
    """Sum of Gaussian-core vortices on an (n1, n2, n3) grid, |m| <= 1, FP64 -> FP32.

    Components are (m0, m1, m2) along the (n3, n2, n1) axes -- the pairing of
    beam_dir_vector with the NUFFT axes (include/projection.h:36-41,
    include/nufft.h:56-57); n2 is the film normal.
    """
    n1, n2, n3 = dims
    rng = SplitMix64(seed, 1)
    z, y, x = np.meshgrid(np.arange(n1, dtype=np.float64), np.arange(n2, dtype=np.float64),
                          np.arange(n3, dtype=np.float64), indexing="ij")
    m = np.zeros((3, n1, n2, n3))
    inplane = max(n1, n3)
    for _ in range(nvortex):
        c = np.array([rng.uniform(2, max(2.0, n3 - 3)), rng.uniform(min(2.0, (n2 - 1) / 2), max(2.0, n2 - 3)),
                      rng.uniform(2, max(2.0, n1 - 3))])
        rho = rng.uniform(2.0, max(3.0, inplane / 8.0))
        if rng.uniform() < 0.5:
            a = np.array([0.0, 1.0, 0.0])
        else:
            cz = rng.uniform(-1.0, 1.0)
            ph = rng.uniform(0.0, 2 * np.pi)
            s = math.sqrt(max(0.0, 1 - cz * cz))
            a = np.array([s * math.cos(ph), s * math.sin(ph), cz])
        chi = 1.0 if rng.uniform() < 0.5 else -1.0
        pol = 1.0 if rng.uniform() < 0.5 else -1.0
        r = np.stack([x - c[0], y - c[1], z - c[2]])
        along = a[0] * r[0] + a[1] * r[1] + a[2] * r[2]
        perp = r - along[None] * a[:, None, None, None]
        s2 = np.sum(perp * perp, axis=0)
        s = np.sqrt(s2)
        env = np.exp(-np.sum(r * r, axis=0) / (2 * (3 * rho) ** 2))
        core = np.exp(-s2 / rho ** 2)
        cross = np.stack([a[1] * r[2] - a[2] * r[1], a[2] * r[0] - a[0] * r[2], a[0] * r[1] - a[1] * r[0]])
        with np.errstate(invalid="ignore", divide="ignore"):
            swirl = np.where(s > 1e-12, (1 - core) / np.where(s > 1e-12, s, 1.0), 0.0)
        m += env[None] * (chi * cross * swirl[None] + pol * a[:, None, None, None] * core[None])
    norm = np.sqrt(np.sum(m * m, axis=0)).max()
    m /= max(1.0, norm)
    return m.astype(np.float32)


@dataclasses.dataclass
class Scan:
    """One gamma dataset of a laminography scan: per-projection (alpha, gamma)."""

    alpha: np.ndarray
    gamma: np.ndarray


def lamino_scan(nproj: int, tilt_deg: float, gamma_deg: float) -> Scan:
    """Rotation psi_j = 2 pi j / P; beam at the tilt from the film normal: alpha = tilt, gamma = offset + psi."""
    psi = 2 * np.pi * np.arange(nproj) / nproj
    return Scan(np.full(nproj, math.radians(tilt_deg)), math.radians(gamma_deg) + psi)


@dataclasses.dataclass
class Workload:
    name: str
    method: int
    recon_dims: tuple           # odd-ified internal (n1, n2, n3)
    detector: int               # odd detector size
    scans: list                 # list[Scan]
    seed: int
    noise: float = 0.0
    params: dict = dataclasses.field(default_factory=dict)

    def solver_shapes(self):
        return solver_dims(self.method, self.recon_dims, self.detector)


SEED_BASE = 20261202


def config(idx: int) -> Workload:
    """The configurations of the benchmark plan (1-based index)."""
    if idx == 1:
        return Workload("plain MBIR1 CG: vol [16,64,64]x3, det 64, 1 gamma x 90 proj, tilt 30", 1,
                        recon_dims_internal([16, 64, 64]), oddify(64), [lamino_scan(90, 30, 0)], SEED_BASE + 1,
                        params=dict(max_iters=20, tol=0.0))
    if idx == 2:
        return Workload("split-Bregman MBIR2: vol [32,128,128]x3, det 128, gammas (0,45) x 180 proj, tilt 35", 2,
                        recon_dims_internal([32, 128, 128]), oddify(128),
                        [lamino_scan(180, 35, 0), lamino_scan(180, 35, 45)], SEED_BASE + 2, noise=0.01,
                        params=dict(lam=0.1, mu=10.0, max_iters=30, inner_iters=3, tol=0.0, xtol=0.0))
    if idx == 3:
        return Workload("qGGMRF MBIR3: vol [51,256,256]x3, det 256, gammas (0,45,90) x 360 proj, tilt 30", 3,
                        recon_dims_internal([51, 256, 256]), oddify(256),
                        [lamino_scan(360, 30, g) for g in (0, 45, 90)], SEED_BASE + 3,
                        params=dict(sigma=1000.0, p=1.2, mu=0.0, max_iters=50, inner_iters=3, tol=0.0, xtol=0.0))
    raise ValueError(idx)


def simulate(work: Workload, project) -> list:
    """Projection stacks of the workload's vortex field, following src/forward.cpp:138-175.

    ``project(m_padded, scan, det_padded) -> (P, det_padded, det_padded)`` is the
    forward projector to use (the GPU one in benchmarks, the oracle in tests).
    Returns (datasets, m): the list of (proj (P, det, det) float32, alpha, gamma)
    and the unpadded field.
    """
    m = vortex_field(work.recon_dims, work.seed)
    pad = [fwd_padded(n) for n in work.recon_dims]
    mp = np.zeros((3,) + tuple(pad), np.float32)
    o = [(p - n) // 2 for p, n in zip(pad, work.recon_dims)]
    mp[:, o[0]:o[0] + work.recon_dims[0], o[1]:o[1] + work.recon_dims[1], o[2]:o[2] + work.recon_dims[2]] = m
    dp = fwd_padded(work.detector)
    out = []
    for gi, sc in enumerate(work.scans):
        p = np.asarray(project(mp, sc, dp), dtype=np.float32)
        c = (dp - work.detector) // 2
        p = np.ascontiguousarray(p[:, c:c + work.detector, c:c + work.detector])
        if work.noise > 0:
            rng = SplitMix64(work.seed, 100 + gi)
            p = p + (work.noise * float(np.abs(p).max()) * rng.normal_array(p.size).reshape(p.shape)).astype(np.float32)
        out.append((p.astype(np.float32), sc.alpha, sc.gamma))
    return out, m


# ---------------------------------------------------------------- batched / served job mixes
VOLUME_CLASSES = ([16, 64, 64], [32, 128, 128], [51, 256, 256])
SERVED_CLASSES = (([16, 64, 64], 0.35), ([32, 128, 128], 0.35), ([51, 256, 256], 0.2), ([51, 511, 511], 0.1))
SB_PARAMS = dict(lam=0.1, mu=10.0, max_iters=30, inner_iters=3, tol=0.0, xtol=0.0)
CG_PARAMS = dict(max_iters=20, tol=0.0)
# mu = 0: nagopt gets the normal operator's Lipschitz estimate as its Lipschitz constant.  The
# reference passes mu (default 10) there (src/mbir3.cpp:123-125), a step of 0.01 that diverges
# once 2-4 datasets of 360 projections put the operator's norm in the thousands.
QG_PARAMS = dict(sigma=1000.0, p=1.2, mu=0.0, max_iters=50, inner_iters=3, tol=0.0, xtol=0.0)


@dataclasses.dataclass
class JobSpec:
    work: Workload
    arrival_s: float = 0.0


def batched_specs(njobs: int = 16) -> list:
    """Batched configuration: jobs of uneven size, volumes drawn from the three
    plan volumes, 1-3 gamma datasets of 90-360 angles, tilt 25-40 deg, CG
    (20 iterations) or split-Bregman (30 x 3); job seed = base + 1000 + job id."""
    rng = SplitMix64(SEED_BASE + 4, 0)
    out = []
    for j in range(njobs):
        v = VOLUME_CLASSES[min(2, int(rng.uniform() * 3))]
        G = 1 + min(2, int(rng.uniform() * 3))
        P = 90 * (1 + min(3, int(rng.uniform() * 4)))
        tilt = rng.uniform(25.0, 40.0)
        method = 1 if rng.uniform() < 0.5 else 2
        scans = [lamino_scan(P, tilt, 45.0 * g) for g in range(G)]
        params = dict(CG_PARAMS) if method == 1 else dict(SB_PARAMS)
        name = f"batched job {j}: vol {v}, {G} x {P} proj, tilt {tilt:.1f}, MBIR{method}"
        out.append(JobSpec(Workload(name, method, recon_dims_internal(v), oddify(v[1]), scans,
                                    SEED_BASE + 1000 + j, noise=0.01, params=params)))
    return out


def served_specs(njobs: int = 12, rate: float = 4.0) -> list:
    """Served configuration: Poisson arrivals (mean `rate` jobs/s), volumes up to
    [51,511,511], 2-4 gamma datasets of 360 angles, detector = volume width,
    split-Bregman or qGGMRF 50/50; job seed = base + 2000 + job id."""
    rng = SplitMix64(SEED_BASE + 5, 0)
    out = []
    t = 0.0
    for j in range(njobs):
        t += -math.log(max(1e-12, 1.0 - rng.uniform())) / rate
        u = rng.uniform()
        acc = 0.0
        v = SERVED_CLASSES[-1][0]
        for vol, p in SERVED_CLASSES:
            acc += p
            if u < acc:
                v = vol
                break
        G = 2 + min(2, int(rng.uniform() * 3))
        tilt = rng.uniform(25.0, 40.0)
        method = 2 if rng.uniform() < 0.5 else 3
        scans = [lamino_scan(360, tilt, 45.0 * g) for g in range(G)]
        params = dict(SB_PARAMS) if method == 2 else dict(QG_PARAMS)
        name = f"served job {j}: vol {v}, {G} x 360 proj, tilt {tilt:.1f}, MBIR{method}"
        out.append(JobSpec(Workload(name, method, recon_dims_internal(v), oddify(v[1]), scans,
                                    SEED_BASE + 2000 + j, noise=0.01, params=params), arrival_s=t))
    return out


def recon_params(work: Workload):
    """ReconParams of a workload's solver settings."""
    from .api import ReconParams

    return ReconParams(**work.params)
