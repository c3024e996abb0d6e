"""The C++ drop-in layer (include/tomocam/*.h) on the GPU.

tests/cxx/test_dropin calls the reference's own C++ API -- tomocam::forward,
adjoint, gradient, residual, nufft::nufft3d1/3d2, opt::cgsolver,
RampPreconditioner, grad_u/divergence/laplacian, qggmrf, MBIR1/MBIR2 -- the
way the reference's tests/test_grad.cpp, test_loss.cpp and test_cg.cpp call
it, and compares each result with the oracle (FP64) in-process.  Tolerances
are in the program: 1e-4 relative for projections and NUFFTs, 1e-3 for
reconstructions, 5e-6 for the preconditioner, 1e-6 for the stencils.  It also
checks that misuse raises the exception type and message the reference raises.
"""
import os
import subprocess

import pytest

from conftest import ROOT, say

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cxx", "test_dropin")
CASES = ["grid_coords", "forward_adjoint", "gradient", "residual", "nufft", "stencils", "cgsolver", "mbir",
         "errors", "double"]


@pytest.fixture(scope="module")
def dropin_run():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run __graft_entry__.build() first")
    say("running", BIN)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    say(r.stdout)
    if r.stderr:
        say(r.stderr)
    sections, cur = {}, None
    for line in r.stdout.splitlines():
        if line.startswith("== "):
            cur = line[3:].strip()
            sections[cur] = []
        elif line.startswith("## "):
            cur = None
        elif cur is not None:
            sections[cur].append(line)
    return r.returncode, sections


@pytest.mark.parametrize("case", CASES)
def test_dropin_case(dropin_run, case):
    _, sections = dropin_run
    lines = sections.get(case)
    assert lines, f"case {case} did not run"
    assert all(l.rstrip().endswith("PASS") for l in lines if l.strip()), "\n".join(lines)


def test_dropin_exit_status(dropin_run):
    rc, _ = dropin_run
    assert rc == 0
