"""Request ingest and output on the GPU: mask_infs_nans and the {1, 2, 0}
transpose on the device against numpy, and the reference's `recon` chain
(TOML -> TIFF stacks + angle files -> MBIR -> transpose -> TIFF/VTI files)
against the same reconstruction called directly and against the FP64 oracle
(reconstructions within 1e-3)."""
import math
import os

import numpy as np
import pytest

from conftest import rel, say

pytestmark = pytest.mark.gpu


def test_mask_infs_nans_on_device(ctx):
    from lamino_amd import ingest

    rng = np.random.default_rng(90)
    a = rng.standard_normal((5, 7, 9)).astype(np.float32)
    a[0, 0, 0], a[1, 2, 3], a[4, 6, 8] = np.nan, np.inf, -np.inf
    want = np.where(np.isfinite(a), a, 0).astype(np.float32)
    np.testing.assert_array_equal(ingest.mask_infs_nans(a, ctx), want)


@pytest.mark.parametrize("shape", [(1, 1, 1), (33, 7, 45), (89, 21, 89), (359, 71, 359)])
def test_transpose120_on_device(ctx, shape):
    from lamino_amd import ingest

    v = np.random.default_rng(91).standard_normal(shape).astype(np.float32)
    np.testing.assert_array_equal(ingest.transpose120(v, ctx), np.transpose(v, (1, 2, 0)))


def test_recon_from_a_toml_request(ctx, orc, tmp_path):
    import lamino_amd as la
    from lamino_amd import ingest, synth

    toml_dims = [8, 24, 24]                 # odd-ified to (7, 23, 23); internal (23, 7, 23)
    rd_int = synth.recon_dims_internal(toml_dims)
    det = 23
    rng = np.random.default_rng(92)
    files = []
    truth = []
    for g, gdeg in enumerate((0.0, 45.0)):
        P = 24
        deg = np.arange(P) * (360.0 / P)
        alpha = np.full(P, math.radians(30.0))
        gamma_rad = float(np.float32(np.float32(gdeg) * np.float64(math.pi) / np.float32(180.0)))
        psi = (deg.astype(np.float32).astype(np.float64) * math.pi / 180.0).astype(np.float32).astype(np.float64)
        # the reference's literal grid: angles = psi, constant gamma
        proj = rng.uniform(0.5, 1.5, (P, det + 1, det + 1)).astype(np.float32)   # even: the reader trims to 23
        proj[0, 3, 4] = np.nan
        tif = tmp_path / f"g{g}.tif"
        ingest.write_tiff(str(tif), proj)
        ang = tmp_path / f"g{g}.txt"
        ang.write_text("\n".join(f"{v}" for v in deg))
        files.append((tif, ang, gdeg))
        trimmed = np.nan_to_num(proj[:, :det, :det], nan=0.0)
        truth.append((trimmed, psi, gamma_rad))
    out = tmp_path / "out" / "recon.tiff"
    cfg = tmp_path / "c.toml"
    body = "".join(f'[[input]]\nfilename = "{t}"\nangles = "{a}"\ngamma = {g}\n\n' for t, a, g in files)
    body += f'[output]\nfilename = "{out}"\nformats = ["tiff", "vti"]\n\n'
    body += ("[recon_params]\nmax_iters = 3\ntol = 0\nxtol = 0\nrecon_dims = [8, 24, 24]\n"
             "[recon_params.regularizer]\nmethod = \"split_bregman\"\n"
             "[recon_params.regularizer.split_bregman]\nlambda = 0.1\nmu = 10.0\ninner_iters = 2\n")
    cfg.write_text(body)
    req = ingest.load_config(str(cfg), ctx)
    assert req.method == 2 and req.internal_dims == rd_int
    for (p, a, g), (tp, ta, tg) in zip(req.datasets, truth):
        np.testing.assert_array_equal(p, tp)
        np.testing.assert_allclose(a, ta, rtol=0, atol=0)
        assert g == tg
    result, written = ingest.recon(str(cfg), ctx)
    assert result.shape == (3, 7, 23, 23)
    direct = la.MBIR2(truth, rd_int, la.ReconParams(max_iters=3, inner_iters=2, lam=float(np.float32(0.1)), mu=10.0,
                                                     tol=0.0, xtol=0.0))
    # the literal grid (angles as alpha, constant gamma) takes the 3-D spreader, whose box
    # flush adds with atomics: two runs agree to FP32 summation order
    assert rel(result, np.stack([np.transpose(direct[i], (1, 2, 0)) for i in range(3)])) < 1e-5
    xo = orc.mbir(2, [(p, a, g) for p, a, g in truth], rd_int, max_iters=3, inner_iters=2,
                  lam=float(np.float32(0.1)), mu=10.0, tol=0.0, xtol=0.0)
    e = rel(direct, xo)
    say(f"recon from TOML vs oracle: {e:.2e}")
    assert e < 1e-3
    assert sorted(os.path.basename(w) for w in written) == ["recon.vti", "recon0.tiff", "recon1.tiff", "recon2.tiff"]
    for i in range(3):
        np.testing.assert_array_equal(ingest.read_tiff(str(tmp_path / "out" / f"recon{i}.tiff")), result[i])


def test_served_requests_write_their_files_while_the_next_jobs_run(ctx, tmp_path):
    """lam_server_submit_output: three requests on two slots; each job's result leaves the
    device transposed {1, 2, 0}, its files are written by the server's writer thread, and
    what is written equals the plain form's reconstruction in the same order."""
    import lamino_amd as la
    from lamino_amd import ingest
    from lamino_amd.serve import Server

    rng = np.random.default_rng(93)
    jobs = []
    for k, (rd, P, det) in enumerate([((11, 5, 13), 16, 13), ((9, 5, 9), 12, 11), ((13, 7, 11), 20, 15)]):
        ds = []
        for g in range(2):
            al = np.full(P, math.radians(30.0))
            ga = math.radians(45.0 * g) + 2 * np.pi * np.arange(P) / P
            ds.append((rng.uniform(0.5, 1.5, (P, det, det)).astype(np.float32), al, ga))
        jobs.append((rd, ds))
    prm = la.ReconParams(max_iters=3, inner_iters=2, lam=0.1, mu=10.0, tol=0.0, xtol=0.0)
    with Server(slots=2, mode="graph") as s:
        ids = [s.submit_output(2, ds, rd, prm, str(tmp_path / f"job{k}" / "rec.tiff"), formats=("tiff", "vti"))
               for k, (rd, ds) in enumerate(jobs)]
        res = [s.wait(i) for i in ids]
        with pytest.raises(Exception):
            s.wait(ids[0])          # a job is waited on once
    for k, ((rd, ds), (x, info)) in enumerate(zip(jobs, res)):
        assert info["status"] == 0
        plain = la.api._mbir(2, ds, rd, prm)
        want = np.stack([np.transpose(plain[i], (1, 2, 0)) for i in range(3)])
        assert x.shape == want.shape
        assert rel(x, want) < 1e-5
        for i in range(3):
            np.testing.assert_array_equal(ingest.read_tiff(str(tmp_path / f"job{k}" / f"rec{i}.tiff")), x[i])
        assert os.path.getsize(tmp_path / f"job{k}" / "rec.vti") > 3 * 8 * x[0].size
