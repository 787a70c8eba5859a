"""Pins the CPU oracle (oracle/bem_oracle.c + oracle/oracle.py) before it is trusted.

  * against the reference's own headers built from source (oracle/_ref, when present in
    this container) and the vectors they wrote into tests/golden/ref_vectors.npz;
  * against closed forms: Dunavant-7 points/weights, Gauss-Legendre nodes, exactness of the
    degree-5 rule, cap area pi R^2, polyhedron volume 4 pi a^3 / 3, closed-surface sum A n = 0,
    the Gauss identity for the double layer, free-space Stokes drag and swimming speed;
  * against the reference's golden outputs (outputs/ CSVs copied to tests/golden).
Each check is chosen so that a plausible slip (a dropped term, a sign, a transposed index,
a wrong rule threshold) fails it.
"""
import csv
import ctypes
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

REFLIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")
VEC = np.load(os.path.join(GOLDEN, "ref_vectors.npz"))


def read_csv(name):
    rows = []
    with open(os.path.join(GOLDEN, "reference_outputs", name)) as f:
        lines = [ln for ln in f if not ln.startswith("#")]
    rd = csv.DictReader(lines)
    for r in rd:
        rows.append({k: float(v) for k, v in r.items()})
    return rows


MESHES = {
    "tube_cfg1": lambda O: O.swimmer_tube_mesh(1.0, 4.0, 0.3, 24, 6),
    "tube_axis03": lambda O: O.swimmer_tube_mesh(1.0 / 0.3, 4.0 / 0.3, 1.0, 24, 4),
    "duct": lambda O: O.duct_mesh(1.0, 4.0, 24, 24, 6),
    "sphere_l2": lambda O: O.icosphere(0.3, 2),
    "sphere_l3_off": lambda O: O.icosphere(1.0, 3, (0.7, 0.0, 0.0)),
}


# ------------------------------------------------------------------ reference-build vectors
@pytest.mark.parametrize("name", list(MESHES))
def test_mesh_matches_reference_build(O, name):
    M = MESHES[name](O)
    assert np.array_equal(M.tris, VEC[f"{name}/tris"]), "connectivity must be bit-exact"
    assert np.array_equal(M.tags, VEC[f"{name}/tags"])
    # coordinates: the reference build contracts some products into FMAs; ulp-level only
    assert np.abs(M.nodes - VEC[f"{name}/nodes"]).max() <= 4e-16 * max(1.0, np.abs(M.nodes).max())
    for k in ["centroid", "normal", "area", "h"]:
        ref = VEC[f"{name}/{k}"]
        assert np.abs(getattr(M, k) - ref).max() <= 4e-15 * max(1.0, np.abs(ref).max()), k


@pytest.mark.parametrize("name", list(MESHES))
def test_rule_levels_match_reference_build(O, name):
    M = MESHES[name](O)
    lev, dist = O.levels(M, M.centroid, self_src=np.arange(M.P()))
    assert np.array_equal(lev, VEC[f"{name}/levels"]), "rule choice differs from the reference"
    # distance from the rule thresholds: decisions are robust to rounding-level differences
    off = ~np.eye(M.P(), dtype=bool)
    for thr in (0.75, 1.5, 3.0):
        assert np.abs(dist[off] / thr - 1).min() > 1e-9


@pytest.mark.parametrize("name", list(MESHES))
def test_panel_integrals_match_reference_build(O, name):
    M = MESHES[name](O)
    pairs = VEC[f"{name}/pairs"]
    for q, (i, p) in enumerate(pairs):
        SL, DL = O.panel_integrals(M, p, M.centroid[i], i == p)
        sref, dref = VEC[f"{name}/SL"][q].reshape(3, 3), VEC[f"{name}/DL"][q].reshape(3, 3)
        assert np.abs(SL - sref).max() <= 1e-13 * np.abs(sref).max()
        scale = max(np.abs(dref).max(), np.abs(sref).max() / max(np.linalg.norm(M.centroid[p] - M.centroid[i]), 1e-3))
        assert np.abs(DL - dref).max() <= 1e-13 * scale
    kinds = set(int(v) for v in VEC[f"{name}/levels"][pairs[:, 0], pairs[:, 1]])
    assert 4 in kinds and 0 in kinds


def test_rules_and_gauss_legendre_match_reference_build(O):
    for lev in range(4):
        L, W = O.rule(lev)
        assert np.allclose(L, VEC[f"rule{lev}/L"], rtol=0, atol=2e-16)
        assert np.allclose(W, VEC[f"rule{lev}/W"], rtol=0, atol=1e-18)
    x, w = O.gauss_legendre01(10)
    assert np.allclose(x, VEC["gl10/x"], atol=1e-16) and np.allclose(w, VEC["gl10/w"], atol=1e-16)


def test_field_matches_reference_build(O):
    M = O.icosphere(1.0, 2)
    v = O.velocity_at(M, VEC["field/u"], VEC["field/f"], 1.0, VEC["field/pts"])
    assert np.abs(v - VEC["field/out"]).max() <= 1e-13 * np.abs(VEC["field/out"]).max()


@pytest.mark.skipif(not os.path.exists(REFLIB), reason="reference build absent (built only where the sources are)")
def test_live_reference_build_random_pairs(O):
    """Random target points (not centroids) against the live reference build."""
    L = ctypes.CDLL(REFLIB)
    vp, i_, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.ref_swimmer_tube_mesh.restype = vp
    L.ref_swimmer_tube_mesh.argtypes = [d, d, d, i_, i_]
    L.ref_panel_integrals.argtypes = [vp, i_, vp, i_, vp, vp]
    h = L.ref_swimmer_tube_mesh(1.0, 4.0, 0.3, 16, 4)
    M = O.swimmer_tube_mesh(1.0, 4.0, 0.3, 16, 4)
    rng = np.random.default_rng(1)
    for _ in range(200):
        p = int(rng.integers(M.P()))
        x0 = np.ascontiguousarray(M.centroid[p] + rng.normal(scale=0.2, size=3))
        SLr, DLr = np.zeros(9), np.zeros(9)
        L.ref_panel_integrals(h, p, x0.ctypes.data_as(vp), 0, SLr.ctypes.data_as(vp), DLr.ctypes.data_as(vp))
        SL, DL = O.panel_integrals(M, p, x0, False)
        assert np.abs(SL.ravel() - SLr).max() <= 1e-13 * np.abs(SLr).max()
        assert np.abs(DL.ravel() - DLr).max() <= 1e-12 * max(np.abs(DLr).max(), np.abs(SLr).max())


# ------------------------------------------------------------------ closed forms
def test_dunavant7_closed_form_and_exactness(O):
    L, W = O.rule(0)
    s15 = math.sqrt(15.0)
    alpha, beta = (6 + s15) / 21, (6 - s15) / 21
    # the reference's constants are 15-digit truncations of the closed forms:
    # points (1 - 2t, t, t) and rotations for t = alpha, beta; centroid weight 9/40
    assert abs(L[1, 0] - (1 - 2 * alpha)) < 1e-14 and abs(L[1, 1] - alpha) < 1e-14
    assert abs(L[4, 0] - (1 - 2 * beta)) < 1e-14 and abs(L[4, 1] - beta) < 1e-14
    assert abs(W[0] - 9 / 40) < 1e-15
    assert abs(W[1] - (155 + s15) / 1200) < 1e-14 and abs(W[4] - (155 - s15) / 1200) < 1e-14
    for lev in range(4):
        Ll, Wl = O.rule(lev)
        assert abs(Wl.sum() - 1) < 1e-14 and np.allclose(Ll.sum(axis=1), 1, atol=1e-15)
        # degree-5 exactness on the reference triangle (x = L1, y = L2): int x^a y^b = a! b! / (a+b+2)!
        for a in range(6):
            for b in range(6 - a):
                exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2) * 2
                got = np.sum(Wl * Ll[:, 1] ** a * Ll[:, 2] ** b)
                assert abs(got - exact) < 1e-13, (lev, a, b)


def test_gauss_legendre_vs_numpy(O):
    x, w = O.gauss_legendre01(10)
    xr, wr = np.polynomial.legendre.leggauss(10)
    order = np.argsort(-xr)  # the reference maps x = (1 - z)/2 with z descending
    assert np.allclose(x, 0.5 * (1 - xr[order]), atol=1e-15)
    assert np.allclose(w, 0.5 * wr[order], atol=1e-15)  # Jacobian 1/2 of [-1,1] -> [0,1]


@pytest.mark.parametrize("nth,nr", [(12, 3), (24, 6), (64, 16)])
def test_tube_mesh_counts_and_invariants(O, nth, nr):
    R, L = 1.0, 4.0
    M = O.swimmer_tube_mesh(R, L, 0.3, nth, nr)
    z = O.graded_z(L, 0.6 * 2 * math.pi * R / nth, 1.4 * 2 * math.pi * R / nth, 0.3 + 0.5 * R)
    nz = len(z) - 1
    assert M.P() == 2 * nth * nz + 2 * nth * (2 * nr - 1)
    assert z[0] == -L / 2 and z[-1] == L / 2 and np.all(np.diff(z) > 0)
    for tag in (1, 2):  # each cap's area is exactly pi R^2 (area-corrected polygon radius)
        assert abs(M.area[M.tags == tag].sum() - math.pi * R * R) < 1e-12
    wall = M.tags == 0
    g = 2 * math.pi / nth
    Rp = R * math.sqrt(g / math.sin(g))
    assert abs(M.area[wall].sum() - nth * 2 * Rp * math.sin(g / 2) * L) < 1e-11
    assert np.abs((M.area[:, None] * M.normal).sum(axis=0)).max() < 1e-12  # closed surface
    c = M.centroid
    assert np.all((M.normal[wall, 0] * c[wall, 0] + M.normal[wall, 1] * c[wall, 1]) > 0)  # outward
    assert np.all(M.normal[M.tags == 1, 2] < 0) and np.all(M.normal[M.tags == 2, 2] > 0)


@pytest.mark.parametrize("lev", [0, 1, 2, 3])
def test_icosphere_invariants(O, lev):
    a, c = 0.7, np.array([0.3, -0.2, 0.5])
    M = O.icosphere(a, lev, c)
    assert M.P() == 20 * 4 ** lev and len(M.nodes) == 10 * 4 ** lev + 2
    # polyhedron volume = 4 pi a^3 / 3 (divergence theorem with the inward normals)
    vol = -np.sum(M.area * np.sum((M.centroid - c) * M.normal, axis=1)) / 3
    assert abs(vol - 4 * math.pi * a ** 3 / 3) < 1e-12
    assert np.all(np.sum((M.centroid - c) * M.normal, axis=1) < 0)  # normals into the sphere
    assert np.abs((M.area[:, None] * M.normal).sum(axis=0)).max() < 1e-13


def test_stokeslet_brute_force_and_symmetry(O):
    """Far-field 7-point integral vs a 4^6-subdivided midpoint brute force; SL symmetric."""
    M = O.icosphere(1.0, 0)
    p = 3
    V = M.V[p]
    x0 = M.centroid[p] + 5.0 * M.normal[p] + np.array([0.3, -0.1, 0.2])
    SL, DL = O.panel_integrals(M, p, x0)
    n = 256  # brute force: midpoints of n^2 sub-triangles
    SLb, DLb = np.zeros((3, 3)), np.zeros((3, 3))
    e1, e2 = V[1] - V[0], V[2] - V[0]
    A = M.area[p]
    for i in range(n):
        for j in range(n - i):
            for up in (0, 1):
                if up and j == n - i - 1:
                    continue
                u, v = ((i + 1 / 3, j + 1 / 3) if not up else (i + 2 / 3, j + 2 / 3))
                x = V[0] + (u / n) * e1 + (v / n) * e2
                r = x - x0
                rn = np.linalg.norm(r)
                w = A / n ** 2
                SLb += w * (np.eye(3) / rn + np.outer(r, r) / rn ** 3)
                DLb += -6 * w * np.dot(r, M.normal[p]) * np.outer(r, r) / rn ** 5
    assert np.abs(SL - SLb).max() < 1e-6 * np.abs(SLb).max()
    assert np.abs(DL - DLb).max() < 1e-6 * np.abs(DLb).max()
    assert np.allclose(SL, SL.T, rtol=1e-14, atol=0) and np.allclose(DL, DL.T, rtol=1e-14, atol=0)


def test_self_panel_duffy_vs_brute_force(O):
    """Duffy self integral over a flat triangle vs polar integration about the target.

    In polar coordinates about x0 in the panel plane, int (I/r + r r^T/r^3) dS =
    int dphi rho(phi) (I + e e^T), e the in-plane unit direction and rho the distance to
    the far edge; its trace 4 int rho dphi has a closed form per edge, h ln tan((al + phi)/2).
    The reference's rule (3 sub-triangles, 10x10 Gauss-Legendre) is accurate to ~1e-5 in the
    trace and ~4e-5 per entry on this panel; a dropped term or Jacobian is an O(1) error."""
    M = O.icosphere(1.0, 1)
    p = 5
    x0 = M.centroid[p]
    SL, DL = O.panel_integrals(M, p, x0, True)
    assert np.all(DL == 0)
    V = M.V[p]
    tot, full = 0.0, np.zeros((3, 3))
    xg, wg = np.polynomial.legendre.leggauss(200)
    for e in range(3):
        a, b = V[e] - x0, V[(e + 1) % 3] - x0
        ang = math.acos(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b))
        al = math.acos(np.dot(-a, b - a) / np.linalg.norm(a) / np.linalg.norm(b - a))
        h = np.linalg.norm(a) * math.sin(al)  # distance from x0 to the edge line
        tot += h * (math.log(math.tan((al + ang) / 2)) - math.log(math.tan(al / 2)))
        ua = a / np.linalg.norm(a)
        ub = b - np.dot(b, ua) * ua
        ub /= np.linalg.norm(ub)
        phis = 0.5 * ang * (xg + 1)
        rho = h / np.sin(al + phis)
        for phi, r, w in zip(phis, rho, 0.5 * ang * wg):
            ev = math.cos(phi) * ua + math.sin(phi) * ub
            full += w * r * (np.eye(3) + np.outer(ev, ev))
    assert abs(np.trace(full) - 4 * tot) < 1e-12 * 4 * tot  # the polar brute force itself
    assert abs(np.trace(SL) - 4 * tot) < 2e-5 * 4 * tot
    assert np.abs(SL - full).max() < 1e-4 * np.abs(full).max()
    assert np.allclose(SL, SL.T, rtol=0, atol=1e-15)


def test_gauss_identity_double_layer(O):
    """Stresslet identity: over a closed surface with normals into the body, the summed DL
    kernel -6 (r.n) r r^T / r^5 seen from a point on the surface is +4 pi I (half of the
    8 pi I interior value); every target of a level-3 sphere, off-diagonals vanish."""
    M = O.icosphere(1.0, 3)
    P = M.P()
    _, _, dls = O.block(M, M.centroid, self_src=np.arange(P), dlsum=True)
    d = dls.reshape(P, 3, 3) / (4 * math.pi)
    assert np.abs(d - np.eye(3)).max() < 1e-6
    # a point inside the body sees 8 pi I, one outside 0
    _, _, din = O.block(M, np.array([[0.1, -0.2, 0.3]]), dlsum=True)
    _, _, dout = O.block(M, np.array([[1.5, 0.4, -0.2]]), dlsum=True)
    assert np.abs(din.reshape(3, 3) / (8 * math.pi) - np.eye(3)).max() < 1e-6
    assert np.abs(dout).max() < 1e-6 * 8 * math.pi


def test_free_space_drag_and_speed(O):
    """Stokes drag 6 pi mu a U and U = 2 B1 / 3 (test_squirmer.cpp:53-55 tolerances)."""
    s = O.SwimmerSolver(None, O.icosphere(1.0, 2), [0, 0, 0])
    K = -s.force(s.solve_prescribed([0, 0, 1], [0, 0, 0]))[2] / (6 * math.pi)
    e = np.array([0, 0, 1.0])
    m = s.solve_free([s.slip(e, 1, 0), s.slip(e, 0, 1)])
    assert abs(K - 1) < 5e-3
    assert abs(m[0].U[2] / (2 / 3) - 1) < 6e-3
    assert np.linalg.norm(m[1].U) + np.linalg.norm(m[1].Om) < 1e-10


# ------------------------------------------------------------------ golden outputs
def test_axis_csv_row(O):
    """outputs/reference/squirmer/axis.csv at a/R = 0.3 (sphere level 3, n_theta = 24)."""
    g = [r for r in read_csv("axis.csv") if abs(r["a_over_R"] - 0.3) < 1e-12][0]
    a, lam = 1.0, 0.3
    tube = O.Tube(a / lam, 4 * a / lam, a, 24)
    s = O.SwimmerSolver(tube, O.icosphere(a, 3), [0, 0, 0])
    K = -s.force(s.solve_prescribed([0, 0, 1], [0, 0, 0]))[2] / (6 * math.pi)
    e = np.array([0, 0, 1.0])
    m = s.solve_free([s.slip(e, 1, 0), s.slip(e, 0, 1)])
    assert abs(K - g["drag_K"]) <= 2e-9 * abs(g["drag_K"])
    assert abs(m[0].U[2] - g["U_B1"]) <= 2e-9 * abs(g["U_B1"])
    assert abs(m[1].U[2] - g["U_B2"]) <= 1e-9 * abs(g["U_B1"])


def test_duct_flow_golden(O):
    """outputs/duct_flow: Q for every dP, and every panel's velocity and traction."""
    d = O.DuctFlow(1.0, 4.0, 24, 24, 6)
    for r in read_csv("flow_rate.csv"):
        _, _, Q = d.solve(r["dP"], 0.0)
        assert abs(Q - r["Q"]) <= 2e-9 * abs(r["Q"])
    u, f, _ = d.solve(1.0, 0.0)
    rows = read_csv("panels.csv")
    M = d.mesh
    cen = np.array([[r["cx"], r["cy"], r["cz"]] for r in rows])
    assert np.abs(M.centroid - cen).max() < 1e-9
    fr = np.array([[r["fx"], r["fy"], r["fz"]] for r in rows])
    ur = np.array([[r["ux"], r["uy"], r["uz"]] for r in rows])
    assert np.abs(f - fr).max() <= 2e-9 * np.abs(fr).max()
    assert np.abs(u - ur).max() <= 2e-9 * np.abs(ur).max()


# the benchmark configurations' meshes (R = 1): (tube args, sphere level, a, centre)
BENCH_MESHES = {
    "cfg1": ((1.0, 4.0, 0.3, 24, 6), 2, 0.3, (0.0, 0.0, 0.0)),
    "cfg2": ((1.0, 6.0, 0.5, 32, 8), 3, 0.5, (0.15, 0.0, 0.0)),
    "cfg4": ((1.0, 8.0, 0.3, 24, 6), 3, 0.3, (0.14, 0.0, 0.0)),
    "46k": ((1.0, 4.0, 0.4, 64, 16), 4, 0.4, (0.15, 0.0, 0.0)),
    "cfg5": ((1.0, 8.0, 0.4, 64, 16), 4, 0.4, (0.15, 0.0, 0.0)),
}


def min_threshold_margin(O, M, targets, self_src=None, chunk=512):
    """Smallest |d / threshold - 1| over all (target, panel) pairs, d = |c_p - x0| / h_p,
    thresholds 0.75, 1.5, 3 (src/utils/kernels.hpp:74-75); self pairs excluded."""
    best = np.inf
    for i0 in range(0, len(targets), chunk):
        x0 = targets[i0:i0 + chunk]
        ss = None if self_src is None else self_src[i0:i0 + chunk]
        _, dist = O.levels(M, x0, self_src=ss)
        if self_src is not None:
            dist[np.arange(len(x0)), ss] = np.inf
        for thr in (0.75, 1.5, 3.0):
            best = min(best, float(np.abs(dist / thr - 1).min()))
    return best


@pytest.mark.slow
@pytest.mark.parametrize("name", list(BENCH_MESHES))
def test_rule_threshold_margins_on_benchmark_meshes(O, name):
    """The rule choice is discontinuous, and the reference's build contracts the distance
    |c_p - x0| into FMAs while the GPU evaluates it without contraction. On every block the
    benchmark configurations assemble (tube-tube, tube rows x sphere panels, sphere rows x
    tube panels, sphere-sphere), every pair's distance ratio sits farther from each threshold
    than rounding can move it (1e-12 relative; a rounding difference is ~1e-16), so both
    builds pick the same rule for every pair."""
    targs, lev, a, c = BENCH_MESHES[name]
    tm = O.swimmer_tube_mesh(*targs)
    sm = O.icosphere(a, lev, c)
    margins = {
        "tube-tube": min_threshold_margin(O, tm, tm.centroid, self_src=np.arange(tm.P(), dtype=np.int32)),
        "tube rows x sphere panels": min_threshold_margin(O, sm, tm.centroid),
        "sphere rows x tube panels": min_threshold_margin(O, tm, sm.centroid),
        "sphere-sphere": min_threshold_margin(O, sm, sm.centroid, self_src=np.arange(sm.P(), dtype=np.int32)),
    }
    print(name, margins)
    for block, m in margins.items():
        assert m > 1e-12, (block, m)
