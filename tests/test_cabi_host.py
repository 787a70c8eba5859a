"""Host-side checks of libsqbem (no GPU needed): the library loads and exports every entry
point declared in include/*.h, the host mesh generators produce the oracle's meshes bit for
bit, and the configuration/CLI front end behaves like the reference's."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "squirmer-bem-3dtube_amd")


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(sqbem_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol(sqhost):
    lib = ctypes.CDLL(os.path.join(PKG, "libsqbem.so"))
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.parametrize("args", [(1.0, 4.0, 0.3, 24, 6), (1 / 0.3, 4 / 0.3, 1.0, 30, 4), (1.0, 4.0, 0.4, 64, 16)])
def test_host_tube_mesh_bit_identical_to_oracle(sqhost, O, args):
    g, o = sqhost.swimmer_tube_mesh(*args), O.swimmer_tube_mesh(*args)
    assert np.array_equal(g.tris, o.tris) and np.array_equal(g.tags, o.tags)
    for k in ["nodes", "centroid", "normal", "area", "h"]:
        assert np.array_equal(getattr(g, k), getattr(o, k)), k


@pytest.mark.parametrize("lev,c", [(2, (0, 0, 0)), (3, (0.7, 0, 0)), (4, (0.15, 0, 0))])
def test_host_icosphere_bit_identical_to_oracle(sqhost, O, lev, c):
    g, o = sqhost.icosphere(0.4, lev, c), O.icosphere(0.4, lev, c)
    assert np.array_equal(g.tris, o.tris)
    for k in ["nodes", "centroid", "normal", "area", "h"]:
        assert np.array_equal(getattr(g, k), getattr(o, k)), k


def test_host_duct_mesh_and_merge(sqhost, O):
    g, o = sqhost.duct_mesh(1.0, 4.0, 24, 24, 6), O.duct_mesh(1.0, 4.0, 24, 24, 6)
    assert np.array_equal(g.tris, o.tris) and np.array_equal(g.centroid, o.centroid)
    gm = sqhost.merge(sqhost.swimmer_tube_mesh(1.0, 4.0, 0.3, 12, 3), sqhost.icosphere(0.3, 1))
    om = O.merge(O.swimmer_tube_mesh(1.0, 4.0, 0.3, 12, 3), O.icosphere(0.3, 1))
    assert np.array_equal(gm.tris, om.tris) and np.array_equal(gm.normal, om.normal)
    t = sqhost.icosphere(0.3, 1)
    t.translate((0.1, -0.2, 0.3))
    u = O.icosphere(0.3, 1)
    u.translate((0.1, -0.2, 0.3))
    assert np.array_equal(t.centroid, u.centroid)


def test_mesh_generator_argument_errors(sqhost):
    lib = sqhost.lib()
    assert not lib.sqbem_mesh_swimmer_tube(-1.0, 4.0, 0.3, 24, 4)
    assert "invalid" in sqhost.last_error()
    assert not lib.sqbem_mesh_icosphere(1.0, -1, 0.0, 0.0, 0.0)


def test_cli_usage_and_errors(tmp_path):
    exe = os.path.join(PKG, "squirmer_bem")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "usage:" in r.stderr
    r = subprocess.run([exe, str(tmp_path / "missing.in")], capture_output=True, text=True)
    assert r.returncode == 1 and "error: cannot open input file:" in r.stderr
    bad = tmp_path / "bad.in"
    bad.write_text("mode = nonsense  # comment\n")
    r = subprocess.run([exe, str(bad), f"output_dir={tmp_path}/o"], capture_output=True, text=True)
    assert r.returncode == 1
    assert "unknown mode 'nonsense' (duct|kinematics|field|trajectory)" in r.stderr
