"""Writes tests/golden/reference_fixtures.npz from the reference's own compiled
sources (oracle/_ref/libref.so, built from include/polar_grid.h,
include/projection.h, include/fftutils.h, include/padding.h,
src/finitediff.cpp and src/qggmrf.cpp where they lie).  Inputs are seeded and
small; the file holds inputs and the reference's outputs only.

Run in a container that has the reference checked out:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def main():
    oracle.build()
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (reference sources absent)")
    R = oracle.Ref
    rng = np.random.default_rng(20261202)
    out = {}
    # PolarGrid coordinates (reference constructor), FP32 and FP64
    theta = np.concatenate([np.linspace(-70, 0, 5), [90.0, 180.0, 33.3]]) * np.pi / 180
    for prec in (32, 64):
        for gi, (gamma, Rr, Cc) in enumerate([(np.pi / 4, 7, 9), (0.0, 8, 6), (1.1, 5, 5)]):
            x, y, z = R.grid_coords(theta, gamma, Rr, Cc, prec=prec)
            out[f"grid{prec}_{gi}_theta"] = theta
            out[f"grid{prec}_{gi}_gamma"] = np.array([gamma])
            out[f"grid{prec}_{gi}_rc"] = np.array([Rr, Cc])
            out[f"grid{prec}_{gi}_xyz"] = np.stack([x, y, z])
    # beam_dir_vector
    ga = rng.uniform(-np.pi, np.pi, 6)
    al = rng.uniform(-np.pi, np.pi, 6)
    out["beam_gamma_alpha"] = np.stack([ga, al])
    out["beam_dir64"] = np.stack([R.beam_dir(g, a, 64) for g, a in zip(ga, al)])
    out["beam_dir32"] = np.stack([R.beam_dir(g, a, 32) for g, a in zip(ga, al)])
    # fftshift2 / ifftshift2 on odd and even planes
    for name, shape in (("odd", (2, 7, 9)), ("even", (2, 8, 6))):
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        out[f"shift_{name}_in"] = a
        out[f"shift_{name}_fwd"] = R.shift2(a, False)
        out[f"shift_{name}_inv"] = R.shift2(a, True)
    # pad2d (factors of MBIR1 and MBIR2/3) and crop3d
    p = rng.standard_normal((2, 63, 63))
    out["pad_in"] = p
    for prec in (32, 64):
        out[f"pad2d_{prec}_04142"] = R.pad2d(p, float(np.float32(1.4142) - 1.0), prec)
        out[f"pad2d_{prec}_14142"] = R.pad2d(p, float(np.float32(1.4142)), prec)
    v = rng.standard_normal((11, 9, 13))
    out["crop_in"] = v
    out["crop_out"] = R.crop3d(v, (7, 5, 9))
    # finite differences and qGGMRF
    u = rng.standard_normal((7, 6, 9))
    u[3, 3, 3] = u[3, 3, 4]           # a delta == 0 pair for qGGMRF
    vv = rng.standard_normal((3, 7, 6, 9))
    g0 = rng.standard_normal((7, 6, 9))
    out["fd_u"] = u
    out["fd_v"] = vv
    out["fd_g0"] = g0
    for prec in (32, 64):
        out[f"grad_u_{prec}"] = R.grad_u(u, prec)
        out[f"divergence_{prec}"] = R.divergence(vv, prec)
        out[f"laplacian_{prec}"] = R.laplacian(u, prec)
        out[f"qggmrf_{prec}"] = R.qggmrf(g0, u, 1000.0, 1.2, prec)
        out[f"qggmrf_s2_{prec}"] = R.qggmrf(g0, u, 2.0, 1.1, prec)
    np.savez_compressed(os.path.join(HERE, "reference_fixtures.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
