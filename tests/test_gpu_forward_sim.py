"""The forward-simulation generator (src/forward.cpp) and the demagnetising field
operator (include/demag.h) on the GPU.

forward_simulate is checked against its composition from pieces pinned
elsewhere: numpy's transpose, the pad3d size rule with its float factor, the
oracle's forward projection on the reference's PolarGrid(angles, R, C, gamma),
and the symmetric crop (1e-4).  demag is checked against the closed form for
a plane wave, H = -k (k . a) / |k|^2 cos(k . r), and a numpy restatement of
include/demag.h on a random field (1e-5)."""
import math
import os

import numpy as np
import pytest

from conftest import rel, say

pytestmark = pytest.mark.gpu


def _pad_n(n):
    f = np.float32(1.41421356237 - 1)
    return int(n + 2 * (int(np.float32(n) * f) // 2))


def _det_pad(n):
    return int(n + 2 * int((n * (1.41421356237 - 1)) / 2))


def test_forward_simulate_matches_its_composition(ctx, orc):
    from lamino_amd import ingest

    rng = np.random.default_rng(101)
    d = (5, 9, 11)
    m = rng.standard_normal((3,) + d).astype(np.float32)
    angles = np.radians(np.arange(12) * 30.0 - 165.0)
    gamma, rows, cols = 0.5, 9, 11
    got = ingest.forward_simulate(m, angles, gamma, (rows, cols), eps=1e-9, ctx=ctx)
    t = np.transpose(m, (0, 3, 1, 2))                      # {2, 0, 1} per component
    pd = tuple(_pad_n(n) for n in t.shape[1:])
    mp = np.zeros((3,) + pd)
    o = [(p - n) // 2 for p, n in zip(pd, t.shape[1:])]
    mp[:, o[0]:o[0] + t.shape[1], o[1]:o[1] + t.shape[2], o[2]:o[2] + t.shape[3]] = t
    R2, C2 = _det_pad(rows), _det_pad(cols)
    full = orc.forward(mp, angles, np.full(angles.size, gamma), R2, C2, eps=1e-9)
    want = full[:, (R2 - rows) // 2:(R2 - rows) // 2 + rows, (C2 - cols) // 2:(C2 - cols) // 2 + cols]
    say(f"forward simulation vs composition: {rel(got, want):.2e}")
    assert got.shape == (12, rows, cols)
    assert rel(got, want) < 1e-4


def test_forward_simulation_from_a_config(ctx, tmp_path):
    from lamino_amd import ingest

    rng = np.random.default_rng(102)
    m = rng.standard_normal((3, 5, 7, 9)).astype(np.float32)
    for i in range(3):
        ingest.write_tiff(str(tmp_path / f"m{i}.tif"), m[i])
    deg = np.arange(10) * 36.0
    (tmp_path / "ang.txt").write_text(" ".join(str(v) for v in deg))
    cfg = tmp_path / "f.toml"
    cfg.write_text(f'[input]\ndata_path = "{tmp_path}"\ncomponent_1 = "m0.tif"\ncomponent_2 = "m1.tif"\n'
                   f'component_3 = "m2.tif"\n[angles]\nfilename = "{tmp_path / "ang.txt"}"\n'
                   f'[orientation]\ngamma = 45\n[output]\nfilename = "proj.tiff"\nbasedir = "{tmp_path / "out"}"\n'
                   'dims = [7, 9]\n')
    proj, dst = ingest.forward_from_config(str(cfg), ctx)
    gamma = float(np.float32(float(np.float32(45)) * math.pi / 180.0))
    angles = ingest.read_angles(str(tmp_path / "ang.txt"))
    np.testing.assert_array_equal(proj, ingest.forward_simulate(m, angles, gamma, (7, 9), ctx=ctx))
    np.testing.assert_array_equal(ingest.read_tiff(dst), proj)


def _demag_numpy(m):
    n = m.shape[1:]
    k = np.meshgrid(*[2 * np.pi * np.fft.fftfreq(s) for s in n], indexing="ij")
    mh = [np.fft.fftn(m[i]) for i in range(3)]
    k2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    k2[0, 0, 0] = 1.0
    km = k[0] * mh[0] + k[1] * mh[1] + k[2] * mh[2]
    out = []
    for i in range(3):
        h = -km * k[i] / k2
        h[0, 0, 0] = 0
        out.append(np.real(np.fft.ifftn(h)))
    return np.stack(out)


@pytest.mark.parametrize("shape", [(8, 6, 10), (9, 7, 11)])
def test_demag_plane_wave_and_random_field(ctx, shape):
    import lamino_amd as la

    n1, n2, n3 = shape
    f = (2, 1, 3)
    k0 = np.array([2 * np.pi * f[0] / n1, 2 * np.pi * f[1] / n2, 2 * np.pi * f[2] / n3])
    a = np.array([0.3, -0.7, 0.5])
    i0, i1, i2 = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    phase = np.cos(k0[0] * i0 + k0[1] * i1 + k0[2] * i2)
    m = np.stack([a[c] * phase for c in range(3)]).astype(np.float32)
    want = np.stack([-k0[c] * (k0 @ a) / (k0 @ k0) * phase for c in range(3)])
    got = la.demag(m, ctx)
    say(f"demag plane wave {shape}: {rel(got, want):.2e}")
    assert rel(got, want) < 1e-5
    r = np.random.default_rng(103).standard_normal((3,) + shape).astype(np.float32)
    e = rel(la.demag(r, ctx), _demag_numpy(r.astype(np.float64)))
    say(f"demag random field {shape} vs numpy: {e:.2e}")
    assert e < 1e-5
