"""Host-side logic (CPU only): shape rules, synthetic inputs, the C ABI surface."""
import os
import re
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIX = np.load(os.path.join(HERE, "golden", "reference_fixtures.npz"))


def test_config_shapes_follow_reference_rules():
    """recon dims odd-ified and reordered, padded by n + 2 floor(n f / 2) (include/config.h:184-186,
    src/recon.cpp:60, src/mbir_sb.cpp:51-61, src/mbir_cg.cpp:53-83)."""
    from lamino_amd import synth

    assert synth.config(1).solver_shapes() == ((89, 21, 89), 89)
    assert synth.config(2).solver_shapes() == ((179, 43, 179), 305)
    assert synth.config(3).solver_shapes() == ((359, 71, 359), 615)
    # served maximum [51, 511, 511], detector 512, MBIR2/3
    assert synth.solver_dims(2, synth.recon_dims_internal([51, 511, 511]), synth.oddify(512)) == ((721, 71, 721), 1233)


@pytest.mark.parametrize("prec", [32, 64])
def test_pad_rule_matches_reference_pad2d(prec):
    from lamino_amd import synth

    p = FIX["pad_in"]
    if prec == 32:
        p = p.astype(np.float32).astype(np.float64)
    for key, f in (("04142", np.float32(1.4142) - 1.0), ("14142", np.float32(1.4142))):
        want = FIX[f"pad2d_{prec}_{key}"]
        n = synth.pad_n(p.shape[1], float(f))
        assert want.shape[1:] == (n, n)
        d = (n - p.shape[1]) // 2
        np.testing.assert_array_equal(want[:, d:d + p.shape[1], d:d + p.shape[2]], p)
        assert np.count_nonzero(want) == np.count_nonzero(p)


def test_crop3d_offsets_match_reference():
    v, out = FIX["crop_in"], FIX["crop_out"]
    o = [(a - b) // 2 for a, b in zip(v.shape, out.shape)]
    np.testing.assert_array_equal(v[o[0]:o[0] + 7, o[1]:o[1] + 5, o[2]:o[2] + 9], out)


def test_splitmix_and_vortex_field_are_deterministic():
    from lamino_amd import synth

    a = synth.vortex_field((15, 7, 17), 5)
    b = synth.vortex_field((15, 7, 17), 5)
    np.testing.assert_array_equal(a, b)
    assert np.sqrt((a.astype(np.float64) ** 2).sum(0)).max() <= 1.0 + 1e-6
    r = synth.SplitMix64(1, 2)
    first = [r.next_u64() for _ in range(4)]
    np.testing.assert_array_equal(synth.SplitMix64(1, 2).u64_array(4), np.array(first, dtype=np.uint64))


def _declared():
    txt = open(os.path.join(ROOT, "include", "lamino.h")).read()
    return sorted(set(re.findall(r"\b(lam_[a-z0-9_]+)\s*\(", txt)) - {"lam_op_fn", "lam_loss_fn", "lam_status"})


def test_library_exports_every_declared_symbol():
    from lamino_amd import _lib

    lib = _lib.load()
    decl = _declared()
    missing = [s for s in decl if not hasattr(lib, s)]
    assert not missing, missing
    assert set(decl) == set(_lib.exported_symbols()), set(decl) ^ set(_lib.exported_symbols())


def test_library_carries_gfx950_code_objects():
    from lamino_amd import _lib

    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_context_creation_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from lamino_amd import _lib
    from lamino_amd.api import Context

    with pytest.raises(_lib.LaminoError):
        Context(0)


TOMOCAM_HEADERS = ["array.h", "array_ops.h", "bregman.h", "config.h", "device.h", "dtypes.h", "logger.h", "nufft.h", "slice.h",
                   "optimize.h", "polar_grid.h", "precond.h", "projection.h", "tomocam.h"]


@pytest.mark.parametrize("header", TOMOCAM_HEADERS)
def test_dropin_header_compiles_on_its_own(header):
    """Each include/tomocam/ header is self-contained C++20 (g++, no HIP headers needed)."""
    src = f'#include "tomocam/{header}"\nint main() {{ return 0; }}\n'
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"), "-x", "c++", "-"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_dropin_test_binary_links_and_fails_loudly_without_gpu():
    """tests/cxx/test_dropin resolves the library and the oracle; without a GPU the
    first device call raises instead of computing anything on the host."""
    import torch

    exe = os.path.join(ROOT, "tests", "cxx", "test_dropin")
    subprocess.run(["make", "-s", "-C", os.path.dirname(exe)], check=True)
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "liblamino_amd.so" in ldd and "liboracle.so" in ldd and "not found" not in ldd, ldd
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([exe, "grid_coords"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "exception" in r.stdout and "FAIL" in r.stdout, r.stdout


def _ds(P, det, seed=0):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal((P, det, det)).astype(np.float32), np.full(P, 0.5), 0.0)]


def test_mbir_footprint_estimate_is_host_only_and_monotone():
    """lam_mbir_footprint needs no device and grows with every shape it depends on."""
    import lamino_amd as la

    prm = la.ReconParams()
    base = la.api.mbir_footprint(2, _ds(8, 9), (9, 5, 11), prm)
    assert base > 64 << 20   # includes the fixed slack for plans and reductions
    assert la.api.mbir_footprint(2, _ds(16, 9), (9, 5, 11), prm) > base          # more projections
    assert la.api.mbir_footprint(2, _ds(8, 15), (9, 5, 11), prm) > base          # larger detector
    assert la.api.mbir_footprint(2, _ds(8, 9), (17, 5, 11), prm) > base          # larger volume
    assert la.api.mbir_footprint(2, _ds(8, 9) * 2, (9, 5, 11), prm) > base       # more datasets
    # MBIR3 keeps one Toeplitz spectrum set per dataset: a second dataset costs it more than MBIR2
    d2 = la.api.mbir_footprint(2, _ds(8, 9) * 2, (9, 5, 11), prm) - base
    d3 = la.api.mbir_footprint(3, _ds(8, 9) * 2, (9, 5, 11), prm) - la.api.mbir_footprint(3, _ds(8, 9), (9, 5, 11), prm)
    assert d3 > d2 > 0


def test_server_creation_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from lamino_amd import _lib
    from lamino_amd.serve import Server

    with pytest.raises(_lib.LaminoError):
        Server(slots=2)


def test_fastdiv_matches_integer_division(tmp_path):
    """lamino_amd/csrc/fastdiv.h (the FFT passes' index splits) against / and %."""
    exe = str(tmp_path / "test_fastdiv")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "lamino_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cxx", "test_fastdiv.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
