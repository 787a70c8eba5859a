"""The GPU path against the reference's own output files (tests/golden/reference_outputs,
copied from the reference checkout's outputs/, printed there with 10 significant digits).

Each test reruns the computation that wrote the file, with the reference's inputs, through
the public interfaces of this library: the Python binding (mirroring swimmer.hpp,
duct_flow.hpp, field.hpp, trajectory.hpp) and the squirmer_bem command-line driver
(mirroring main.cpp with the reference's input files, tests/golden/inputs).

Tolerance: the files hold 10 significant digits (rounding 5e-10 relative) and two runs of
the reference agree to ~1e-10 of a column's scale, so values are compared at 2e-9 of the
value, and quantities that vanish by symmetry at 2e-9 of their vector's scale.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

E_Z = np.array([0.0, 0.0, 1.0])
RTOL = 2e-9
EXE = os.path.join(ROOT, "squirmer-bem-3dtube_amd", "squirmer_bem")


def read_csv(path):
    """Rows of a reference CSV (comment lines start with '#') as a header list + float array."""
    with open(path) as f:
        lines = [ln.strip() for ln in f if ln.strip() and not ln.startswith("#")]
    head = lines[0].split(",")
    return head, np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])


def golden(name):
    return read_csv(os.path.join(GOLDEN, "reference_outputs", name))


def col(head, data, name):
    return data[:, head.index(name)]


def close(got, ref, scale=None, rtol=RTOL):
    got, ref = np.asarray(got, dtype=float), np.asarray(ref, dtype=float)
    tol = rtol * (np.abs(ref) if scale is None else scale)
    return np.all(np.abs(got - ref) <= tol + 1e-300), np.max(np.abs(got - ref) - tol)


# ------------------------------------------------------------------ squirmer (test_squirmer.cpp)
def test_axis_csv_all_rows(sq):
    """axis.csv: towed drag and squirmer speeds on the axis, a/R = 0.1 .. 0.6, sphere level 3,
    n_theta = 24, L = 4R (tests/test_squirmer.cpp:83-102)."""
    head, g = golden("axis.csv")
    a = 1.0
    for row in g:
        lam = row[head.index("a_over_R")]
        tube = sq.Tube(a / lam, 4 * a / lam, a, 24)
        s = sq.SwimmerSolver(tube, sq.icosphere(a, 3), (0.0, 0.0, 0.0))
        K = -s.force(s.solve_prescribed([0, 0, 1], [0, 0, 0]))[2] / (6 * math.pi)
        m = s.solve_free([s.slip(E_Z, 1, 0), s.slip(E_Z, 0, 1)])
        UB1 = row[head.index("U_B1")]
        ok, d = close(K, row[head.index("drag_K")])
        assert ok, ("drag_K", lam, d)
        ok, d = close(m[0].U[2], UB1)
        assert ok, ("U_B1", lam, d)
        # U_B2 vanishes by symmetry: the file holds round-off (~1e-9 of U_B1)
        assert abs(m[1].U[2] - row[head.index("U_B2")]) <= 2e-8 * abs(UB1), lam
        assert np.linalg.norm(m[1].U) + np.linalg.norm(m[1].Om) < 1e-6  # test_squirmer.cpp:102


def test_offaxis_csv_all_rows(sq):
    """offaxis.csv: a/R = 0.3, n_theta = 36, level 3, beta = 0 .. 0.85
    (tests/test_squirmer.cpp:148-172)."""
    head, g = golden("offaxis.csv")
    a, R = 1.0, 1.0 / 0.3
    tube = sq.Tube(R, 4 * R, a, 36)
    sU = np.abs(col(head, g, "UB1z")).max()  # vector scales: swimming speed, B1 rotation rate
    sO = np.abs(col(head, g, "OB1y")).max()
    for row in g:
        beta = row[0]
        b = beta * (R - a)
        s = sq.SwimmerSolver(tube, sq.icosphere(a, 3, (b, 0.0, 0.0)), (b, 0.0, 0.0))
        m = s.solve_free([s.slip(E_Z, 1, 0), s.slip(E_Z, 0, 1)])
        for mode in range(2):
            U, Om = m[mode].U, m[mode].Om
            for name, v, sc in [(f"UB{mode + 1}x", U[0], sU), (f"UB{mode + 1}z", U[2], sU),
                                (f"OB{mode + 1}y", Om[1], sO)]:
                ref = row[head.index(name)]
                assert abs(v - ref) <= RTOL * (abs(ref) + sc), (beta, name, v, ref)


def test_reciprocal_csv(sq):
    """reciprocal.csv: U direct and from the Lorentz reciprocal theorem with the towed-sphere
    tractions (tests/test_squirmer.cpp:104-118)."""
    head, g = golden("reciprocal.csv")
    a, R = 1.0, 1.0 / 0.3
    tube = sq.Tube(R, 4 * R, a, 24)
    s = sq.SwimmerSolver(tube, sq.icosphere(a, 3), (0.0, 0.0, 0.0))
    sl = s.slip(E_Z, 1, 0)
    m = s.solve_free([sl])
    tow = s.solve_prescribed([0, 0, 1], [0, 0, 0])
    Fz = s.force(tow)[2]
    Urec = float(np.sum(np.sum(sl * tow.f_sphere, axis=1) * s.S.area)) / Fz
    assert close(m[0].U[2], g[0, 0])[0]
    assert close(Urec, g[0, 1])[0]


def test_field_csv(sq):
    """field.csv: lab-frame velocity on a 25 x 57 grid in y = 0 around the on-axis squirmer,
    a/R = 0.3, both modes (tests/test_squirmer.cpp:120-146)."""
    head, g = golden("field.csv")
    a, R = 1.0, 1.0 / 0.3
    tube = sq.Tube(R, 4 * R, a, 24)
    s = sq.SwimmerSolver(tube, sq.icosphere(a, 3), (0.0, 0.0, 0.0))
    sl1, sl2 = s.slip(E_Z, 1, 0), s.slip(E_Z, 0, 1)
    m1 = s.solve_free([sl1])[0]
    m2 = s.solve_free([sl2])[0]
    nx, nz = 25, 57
    pts, inside = [], []
    for i in range(nx):
        for j in range(nz):
            p = np.array([-0.97 * R + 1.94 * R * i / (nx - 1), 0.0, -2.2 * R + 4.4 * R * j / (nz - 1)])
            inside.append(np.linalg.norm(p) < 1.03 * a)
            pts.append(p)
    pts, inside = np.array(pts), np.array(inside)
    assert np.allclose(pts[:, 0], col(head, g, "x"), rtol=1e-9) and np.allclose(pts[:, 2], col(head, g, "z"), rtol=1e-9)
    v1 = s.velocity(m1, sl1, pts[~inside])
    v2 = s.velocity(m2, sl2, pts[~inside])
    gi = g[~inside]
    assert np.all(np.isnan(g[inside][:, 2:]))
    for v, cx, cz in [(v1, "ux_B1", "uz_B1"), (v2, "ux_B2", "uz_B2")]:
        ref = np.stack([gi[:, head.index(cx)], gi[:, head.index(cz)]], axis=1)
        got = v[:, [0, 2]]
        scale = np.abs(ref).max()
        # pointwise: 2e-9 of the value plus 2e-9 of the field's scale (far points are ~1e-6 U)
        assert np.all(np.abs(got - ref) <= RTOL * np.abs(ref) + RTOL * scale), np.abs(got - ref).max() / scale


def test_blake_csv_unbounded(sq):
    """blake.csv: the unbounded pusher (B2/B1 = -3) on shells r = 1.1, 1.5, 3 a
    (tests/test_squirmer.cpp:47-81); the file's Blake columns are the closed form."""
    head, g = golden("blake.csv")
    s = sq.SwimmerSolver(None, sq.icosphere(1.0, 3), (0.0, 0.0, 0.0))
    sl1, sl2 = s.slip(E_Z, 1, 0), s.slip(E_Z, 0, 1)
    m = s.solve_free([sl1, sl2])
    r, th = col(head, g, "r"), col(head, g, "theta")
    pts = np.stack([r * np.sin(th), 0 * r, r * np.cos(th)], axis=1)
    ub = s.velocity(m[0], sl1, pts) - 3.0 * s.velocity(m[1], sl2, pts)
    rh = pts / r[:, None]
    eth = np.sum(rh * E_Z, axis=1)[:, None] * rh - E_Z
    eth /= np.linalg.norm(eth, axis=1)[:, None]
    ur, ut = np.sum(ub * rh, axis=1), np.sum(ub * eth, axis=1)
    scale = np.abs(col(head, g, "ur_bem")).max()
    assert np.all(np.abs(ur - col(head, g, "ur_bem")) <= RTOL * (np.abs(col(head, g, "ur_bem")) + scale))
    assert np.all(np.abs(ut - col(head, g, "ut_bem")) <= RTOL * (np.abs(col(head, g, "ut_bem")) + scale))
    # and the closed form (Blake 1971) within the reference's own checks at level 3
    err = np.hypot(ur - col(head, g, "ur_blake"), ut - col(head, g, "ut_blake")) / (2.0 / 3.0)
    assert err[r > 1.4].max() < 1e-2 and err[r > 2.9].max() < 5e-3


# ------------------------------------------------------------------ duct flow (test_duct_flow.cpp / main.cpp)
def test_duct_python_binding_golden(sq):
    """flow_rate.csv, panels.csv, profile.csv (inputs/duct_flow.in) via DuctFlow."""
    d = sq.DuctFlow(1.0, 4.0, 24, 24, 6)
    head, g = golden("flow_rate.csv")
    for row in g:
        s = d.solve(row[0], 0.0)
        assert close(s.Q, row[1])[0], row[0]
    s = d.solve(1.0, 0.0)
    head, g = golden("panels.csv")
    M = d.mesh
    assert np.allclose(M.centroid, g[:, 0:3], rtol=0, atol=1e-9)
    assert np.array_equal(M.tags, g[:, head.index("tag")].astype(int))
    for k, name in enumerate(["fx", "fy", "fz"]):
        assert close(s.f[:, k], col(head, g, name), scale=np.abs(g[:, 11:14]).max())[0], name
    for k, name in enumerate(["ux", "uy", "uz"]):
        assert close(s.u[:, k], col(head, g, name), scale=np.abs(g[:, 8:11]).max())[0], name


def _run_cli(args, tmp_path):
    r = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_cli_duct_reference_input(tmp_path):
    out = tmp_path / "duct"
    _run_cli([os.path.join(GOLDEN, "inputs", "duct_flow.in"), f"output_dir={out}"], tmp_path)
    for name in ["flow_rate.csv", "panels.csv", "profile.csv"]:
        head, g = golden(name)
        h2, got = read_csv(out / name)
        assert h2 == head and got.shape == g.shape, name
        for c in range(g.shape[1]):
            scale = np.abs(g[:, c]).max()
            ok, d = close(got[:, c], g[:, c], scale=np.abs(g[:, c]) + scale)
            assert ok, (name, head[c], d)


def test_cli_kinematics_reference_input(tmp_path):
    """inputs/squirmer_kinematics.in is the off-axis sweep of offaxis.csv (n_theta = 36,
    level 3, a/R = 0.3); kinematics.csv holds the same U, Omega for a = 1."""
    out = tmp_path / "kin"
    _run_cli([os.path.join(GOLDEN, "inputs", "squirmer_kinematics.in"), f"output_dir={out}"], tmp_path)
    hk, k = read_csv(out / "kinematics.csv")
    head, g = golden("offaxis.csv")
    assert np.allclose(col(hk, k, "beta"), g[:, 0])
    for mode in (1, 2):
        scale = np.abs(col(hk, k, "UB1z")).max()
        for name, kname in [(f"UB{mode}x", f"UB{mode}x"), (f"UB{mode}z", f"UB{mode}z"), (f"OB{mode}y", f"OB{mode}y")]:
            ok, d = close(col(hk, k, kname), col(head, g, name), scale=np.abs(col(head, g, name)) + scale)
            assert ok, (name, d)
    # towed drag only on the axis; the sphere-on-axis B2 mode does not move
    K = col(hk, k, "drag_K")
    assert np.isfinite(K[0]) and np.all(np.isnan(K[1:]))


@pytest.mark.parametrize("towed", [1, 0])
def test_cli_kinematics_axis_sweep(tmp_path, towed):
    """The a/R sweep of axis.csv (level 3, n_theta = 24, L = 4R, on the axis) through
    squirmer_bem's kinematics mode. towed_drag = 1: every row needs its towed problem, so
    each configuration is solved by its own SwimmerSolver; towed_drag = 0: all six are one
    batched solve (sqbem_batch_solve, both modes as instances), drag_K NaN."""
    out = tmp_path / "kin"
    _run_cli([os.path.join(GOLDEN, "inputs", "squirmer_kinematics.in"), f"output_dir={out}",
              "a_over_R=0.1 0.2 0.3 0.4 0.5 0.6", "beta=0", "n_theta=24", f"towed_drag={towed}"], tmp_path)
    hk, k = read_csv(out / "kinematics.csv")
    head, g = golden("axis.csv")
    assert np.allclose(col(hk, k, "a_over_R"), col(head, g, "a_over_R"))
    ok, d = close(col(hk, k, "UB1z"), col(head, g, "U_B1"))
    assert ok, d
    UB1 = np.abs(col(head, g, "U_B1"))
    assert np.all(np.abs(col(hk, k, "UB2z") - col(head, g, "U_B2")) <= 2e-8 * UB1)
    if towed:
        ok, d = close(col(hk, k, "drag_K"), col(head, g, "drag_K"))
        assert ok, d
    else:
        assert np.all(np.isnan(col(hk, k, "drag_K")))


def test_cli_field_matches_field_csv(tmp_path):
    """inputs/squirmer_field.in with the grid of field.csv (z_extent = 2.2 R)."""
    out = tmp_path / "field"
    R = 1.0 / 0.3
    _run_cli([os.path.join(GOLDEN, "inputs", "squirmer_field.in"), f"output_dir={out}",
              f"z_extent={2.2 * R!r}"], tmp_path)
    head, g = golden("field.csv")
    for mode in (1, 2):
        hf, f = read_csv(out / f"field_B{mode}.csv")
        assert np.allclose(col(hf, f, "x"), col(head, g, "x"), rtol=1e-9)
        assert np.allclose(col(hf, f, "z"), col(head, g, "z"), rtol=1e-9, atol=1e-12)
        inside = np.isnan(col(head, g, f"ux_B{mode}"))
        assert np.array_equal(np.isnan(col(hf, f, "ux")), inside)
        for c in ("x", "z"):
            ref = col(head, g, f"u{c}_B{mode}")[~inside]
            got = col(hf, f, f"u{c}")[~inside]
            scale = np.abs(ref).max()
            assert np.all(np.abs(got - ref) <= RTOL * (np.abs(ref) + scale)), (mode, c)


# ------------------------------------------------------------------ trajectories (trajectory.hpp)
TRAJ = [  # golden file, input file, status the run ends with
    ("neutral.csv", "trajectory_neutral.in", "finished"),
    ("pusher.csv", "trajectory_pusher.in", "wall contact"),
    ("puller_a3.csv", "trajectory_puller_a3.in", "finished"),
    ("puller_a5.csv", "trajectory_puller_a5.in", "finished"),
    ("neutral_yaw15.csv", "trajectory_neutral_3d.in", "finished"),
    ("puller_a3_yaw15.csv", "trajectory_puller_a3_3d.in", "finished"),
]


def _input_keys(inp):
    cfg = {}
    for ln in open(os.path.join(GOLDEN, "inputs", inp)):
        ln = ln.split("#")[0].strip()
        if "=" in ln:
            k, v = (x.strip() for x in ln.split("=", 1))
            cfg[k] = v
    return cfg


def _check_traj(head, got, g):
    """Every row of a whole reference trajectory. The file holds 10 digits (5e-10 of each
    value); the reference's own two code revisions agree to 1.7e-10 of a column's largest
    value over 300 steps (neutral.csv vs neutral_rev2.csv); rounding differences of the
    evaluations (1e-13) are carried along the path by the integration, allowed to grow by
    2e-12 of the column scale per step."""
    assert got.shape == g.shape, (got.shape, g.shape)
    assert np.allclose(got[:, 0], g[:, 0], rtol=0, atol=1e-9)
    steps = np.arange(len(g))[:, None]
    worst = {}
    groups = {"x": ["x", "y", "z"], "e": ["ex", "ey", "ez"], "U": ["Ux", "Uy", "Uz"],
              "Om": ["Omx", "Omy", "Omz"], "beta": ["beta"]}
    for gname, cols in groups.items():
        idx = [head.index(c) for c in cols]
        scale = np.abs(g[:, idx]).max()
        err = np.abs(got[:, idx] - g[:, idx])
        tol = RTOL * np.abs(g[:, idx]) + (1e-9 + 2e-12 * steps) * scale
        worst[gname] = float((err / scale).max())
        assert np.all(err <= tol), (gname, int(np.argmax((err - tol).max(axis=1))), (err / scale).max())
    print("largest difference / column scale:", worst, flush=True)


@pytest.mark.parametrize("gname,inp,status", TRAJ[:3])
def test_trajectory_cli_reference_inputs_whole_run(tmp_path, gname, inp, status):
    """The reference's input files through squirmer_bem to their end (t_max or wall contact)."""
    head, g = golden(gname)
    out = tmp_path / "traj"
    stdout = _run_cli([os.path.join(GOLDEN, "inputs", inp), f"output_dir={out}"], tmp_path)
    assert f"{status}: t = " in stdout, stdout[-500:]
    h2, got = read_csv(out / f"{_input_keys(inp)['name']}.csv")
    assert h2 == head
    _check_traj(head, got, g)
    assert not (out / f"{_input_keys(inp)['name']}.ckpt").exists()


@pytest.mark.parametrize("gname,inp,status", TRAJ[3:])
def test_trajectory_python_binding_whole_run(sq, tmp_path, gname, inp, status):
    head, g = golden(gname)
    cfg = _input_keys(inp)
    P = sq.TrajectoryParams(alpha=float(cfg["alpha"]), a_over_R=float(cfg["a_over_R"]), beta0=float(cfg["beta0"]),
                            pitch0_deg=float(cfg["pitch0_deg"]), yaw0_deg=float(cfg["yaw0_deg"]),
                            dt=float(cfg["dt"]), t_max=float(cfg["t_max"]), L_over_R=float(cfg["L_over_R"]),
                            n_theta=int(cfg["n_theta"]), sphere_level=int(cfg["sphere_level"]),
                            beta_stop=float(cfg["beta_stop"]))
    assert sq.run_trajectory(P, str(tmp_path / "t")) == status
    h2, got = read_csv(tmp_path / "t.csv")
    assert h2 == head
    _check_traj(head, got, g)

