"""Batched and served reconstruction jobs (lam_server_* of include/lamino.h).

A Server runs MBIR jobs on `slots` device workers at once, admitting queued
jobs in arrival order as a worker frees and the HBM budget allows.  Modes:
"plain" (the lam_mbir solver loop), "async" (solver scalars on the device),
"graph" (async, iterations replayed from captured hipGraphs).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check
from .api import ReconParams, _dims, _pack_job

_MODES = {"plain": _lib.LAM_SERVE_PLAIN, "async": _lib.LAM_SERVE_ASYNC, "graph": _lib.LAM_SERVE_GRAPH}


def _info_dict(info: _lib.lam_job_info) -> dict:
    d = {f: getattr(info, f) for f, _ in info._fields_ if f != "stats"}
    d["stats"] = {f: getattr(info.stats, f) for f, _ in info.stats._fields_}
    d["latency_s"] = info.end_s - info.submit_s
    return d


class Server:
    def __init__(self, device: int = 0, budget_bytes: int = 0, slots: int = 4, mode: str = "graph"):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.lam_server_create(int(device), int(budget_bytes), int(slots), _MODES[mode], ctypes.byref(h)))
        self.h = h
        self._jobs = {}

    def submit(self, method: int, datasets, recon_dims, params: ReconParams) -> int:
        """Queues one MBIR job (method 1, 2 or 3); returns its id."""
        arr, pr, _keep = _pack_job(method, datasets, params)
        out = np.empty((3,) + tuple(int(v) for v in recon_dims), np.float32)
        jid = ctypes.c_int64()
        check(self.lib.lam_server_submit(self.h, arr, len(datasets), _dims(recon_dims), ctypes.byref(pr),
                                         out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(jid)))
        self._jobs[jid.value] = out
        return jid.value

    def submit_output(self, method: int, datasets, recon_dims, params: ReconParams, path: str,
                      formats=("tiff", "vti"), transpose120: bool = True, keep_result: bool = True) -> int:
        """Queues a job whose result is written to files (TIFF write3 / VTI / vector TIFF)
        by the server's writer thread while the next jobs run; with transpose120 the result
        is array::transpose(., {1, 2, 0}) of the volume, src/recon.cpp's output order."""
        flags = {"tiff": _lib.LAM_OUT_TIFF, "vti": _lib.LAM_OUT_VTI, "tiff_vectors": _lib.LAM_OUT_TIFF_VECTORS}
        fm = 0
        for f in formats:
            fm |= flags[f]
        arr, pr, _keep = _pack_job(method, datasets, params)
        rd = tuple(int(v) for v in recon_dims)
        shape = (3, rd[1], rd[2], rd[0]) if transpose120 else (3,) + rd
        out = np.empty(shape, np.float32) if keep_result else None
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        spec = _lib.lam_output(os.fsencode(path), fm, 1 if transpose120 else 0)
        jid = ctypes.c_int64()
        check(self.lib.lam_server_submit_output(self.h, arr, len(datasets), _dims(rd), ctypes.byref(pr),
                                                ctypes.byref(spec),
                                                out.ctypes.data_as(ctypes.c_void_p) if out is not None else None,
                                                ctypes.byref(jid)))
        self._jobs[jid.value] = out
        return jid.value

    def submit_request(self, req, keep_result: bool = True) -> int:
        """An ingested request (lamino_amd.ingest.Request) with its output files."""
        return self.submit_output(req.method, req.datasets, req.internal_dims, req.params, req.output.filepath,
                                  tuple(req.output.formats), True, keep_result)

    def wait(self, job_id: int):
        """Blocks until the job is done; returns (result, info dict).  A job is waited on once."""
        info = _lib.lam_job_info()
        try:
            check(self.lib.lam_server_wait(self.h, int(job_id), ctypes.byref(info)))
        finally:
            out = self._jobs.pop(job_id, None)
        return out, _info_dict(info)

    def stats(self) -> dict:
        st = _lib.lam_server_stats()
        check(self.lib.lam_server_query(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self):
        if self.h:
            check(self.lib.lam_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
