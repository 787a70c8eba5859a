"""The reference's own test programs (tests/test_duct_flow.cpp, test_squirmer.cpp,
test_trajectory.cpp), compiled unchanged against this repository's C++ headers and
libsqbem (oracle/Makefile target `dropin`, built where the reference sources exist), run
on the GPU: a source-level drop-in check. Each program prints PASS/FAIL per physics check
and exits nonzero on any failure.

test_squirmer --full writes the files the reference keeps as its outputs/reference/squirmer
results; they are compared with those at the golden-file tolerance as well.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

REFBIN = os.path.join(ROOT, "oracle", "_ref")


def _run(name, args, cwd, timeout=600):
    exe = os.path.join(REFBIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{name} not built (the reference's test sources were absent at build time)")
    print(f"running {name} {' '.join(args)}", flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=timeout, cwd=cwd)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "FAIL" not in r.stdout
    return r.stdout


def test_reference_test_duct_flow(tmp_path):
    out = _run("test_duct_flow", [], tmp_path)
    assert "checks passed" in out


def _read(path):
    with open(path) as f:
        lines = [ln.strip() for ln in f if ln.strip() and not ln.startswith("#")]
    return lines[0].split(","), np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])


def test_reference_test_squirmer_full(tmp_path):
    out = _run("test_squirmer", ["--full"], tmp_path)
    assert "checks passed" in out
    res = tmp_path / "outputs" / "tests" / "squirmer"
    for name in ["axis.csv", "offaxis.csv", "reciprocal.csv", "field.csv", "blake.csv"]:
        head, got = _read(res / name)
        h2, ref = _read(os.path.join(GOLDEN, "reference_outputs", name))
        assert head == h2 and got.shape == ref.shape, name
        fin = np.isfinite(ref)
        assert np.array_equal(fin, np.isfinite(got)), name
        # 10 printed digits; columns that vanish by symmetry hold the reference's round-off,
        # so a floor of 1e-8 of the file's largest value applies (tighter per-column checks
        # of the same files are in test_gpu_golden.py)
        floor = 1e-8 * np.nanmax(np.abs(ref[:, 1:]))
        for c in range(ref.shape[1]):
            m = fin[:, c]
            err = np.abs(got[m, c] - ref[m, c])
            assert np.all(err <= 2e-9 * np.abs(ref[m, c]) + floor), (name, head[c], err.max())


def test_reference_test_trajectory(tmp_path):
    out = _run("test_trajectory", [], tmp_path, timeout=900)
    assert "checks passed" in out
