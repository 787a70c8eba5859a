"""The C++17 header API of the additions (bem::batch_solve, SwimmerSolver::gradient),
exercised by tests/cpp/test_extensions.cpp (built by squirmer-bem-3dtube_amd/Makefile),
run on the GPU: batch_solve against SwimmerSolver::solve_free on the same meshes, the
refined batch against FP64, and the adjoint's mode derivatives against the linearity of
the solution in (B1, B2)."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_header_extensions(tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "test_extensions")
    assert os.path.exists(exe), "tests/cpp/test_extensions is not built: run __graft_entry__.build()"
    print("running tests/cpp/test_extensions", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "all checks passed" in r.stdout and "FAIL" not in r.stdout
