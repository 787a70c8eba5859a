"""The oracle's solver restatement (oracle/oracle.py) against 17-digit outputs of the
reference's own solver headers (tests/golden/ref_solver_*.npz, written by
tests/golden/make_ref_solver.py from oracle/_ref/libref_solver.so, the reference's
src/swimmer.hpp, src/duct_flow.hpp and src/trajectory.hpp built from source against an
LP64 OpenBLAS). CPU only.

Tolerances: the two computations run the same algorithm in FP64 with different operation
order in the BLAS calls (OpenBLAS inside the reference, SciPy's OpenBLAS and numpy in the
oracle) and different FMA contraction in the quadrature (the reference build contracts, the
oracle's C core does not), so they agree to rounding amplified by the conditioning of the
bordered system: 1e-12 norm-wise for U, Omega and tractions (measured 2e-15 to 3e-14),
1e-11 over a 12-step trajectory.

Also pins: Gs^-1 Gc (the Schur operator's rigid-body action) through the six towed
problems of the tiny case; the unbounded branch (the -I completion term); the duct's mixed
boundary conditions; run_trajectory's RK4 start-up, AB4 steps, renormalisation and
checkpoint-resume (the reference's paused-and-resumed CSV equals its uninterrupted one).
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

import oracle as O


def fixture(name):
    return np.load(os.path.join(GOLDEN, f"ref_solver_{name}.npz"))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def oracle_solve(d, towed=False, field=False):
    a, lev, c, e = float(d["a"]), int(d["level"]), d["center"], d["e"]
    if int(d["unbounded"]):
        tube = None
    elif "z" in d:
        tube = O.Tube(float(d["R"]), float(d["L"]), a, int(d["n_theta"]), int(d["n_r"]),
                      mesh=O.tube_mesh(float(d["R"]), d["z"], int(d["n_theta"]), int(d["n_r"])))
    else:
        tube = O.Tube(float(d["R"]), float(d["L"]), a, int(d["n_theta"]), int(d["n_r"]))
    s = O.SwimmerSolver(tube, O.icosphere(a, lev, c), c)
    assert s.Ns == int(d["Ns"]) and (tube is None or tube.Nt == int(d["Nt"]))
    sol = s.solve_free([s.slip(e, m[0], m[1]) for m in d["modes"]])
    return s, sol


def check_modes(d, sol, tol=1e-12):
    # every mode against the largest (U, Omega) of the set: modes that vanish by symmetry
    # carry only discretisation noise, whose relative error means nothing
    scale = max(np.linalg.norm(np.r_[d["U"][k], d["Om"][k]]) for k in range(len(d["U"])))
    for k, g in enumerate(sol):
        err = np.linalg.norm(np.r_[g.U - d["U"][k], g.Om - d["Om"][k]]) / scale
        assert err <= tol, (k, err)
        assert rel(g.f_sphere, d["f_sphere"][k]) <= tol
        if "f_tube" in d:
            assert rel(g.f_tube, d["f_tube"][k]) <= tol


@pytest.mark.parametrize("name", ["tiny", "tiny_unbounded"])
def test_oracle_tiny_cases_and_schur_action(name):
    d = fixture(name)
    s, sol = oracle_solve(d)
    check_modes(d, sol)
    # six towed problems: f_s = -Gs^-1 Gc w for the unit rigid motions w, and F = -sum A f_s
    for w, F, fs in zip(d["towed_w"], d["towed_F"], d["towed_f_sphere"]):
        t = s.solve_prescribed(w[:3], w[3:])
        assert np.linalg.norm(s.force(t) - F) <= 1e-12 * np.linalg.norm(d["towed_F"], axis=1).max()
        assert rel(t.f_sphere, fs) <= 1e-12
    for k, g in enumerate(sol):
        v = s.velocity(g, s.slip(d["e"], *d["modes"][k]), d["field_pts"])
        assert np.abs(v - d["field_u"][k]).max() <= 1e-12 * np.abs(d["field_u"]).max()


def test_oracle_cfg1_kinematics_towed_and_field():
    d = fixture("cfg1")
    s, sol = oracle_solve(d)
    check_modes(d, sol)
    t = s.solve_prescribed(np.array([0.0, 0.0, 1.0]), np.zeros(3))
    assert abs(s.force(t)[2] - d["towed_F"][0][2]) <= 1e-12 * abs(d["towed_F"][0][2])
    v = s.velocity(sol[0], s.slip(d["e"], 1, 0), d["field_pts"])
    assert np.abs(v - d["field_u"][0]).max() <= 1e-12 * np.abs(d["field_u"][0]).max()


@pytest.mark.slow
@pytest.mark.parametrize("name", ["cfg2", "cfg4", "batch_smallest"])
def test_oracle_mid_size_configs(name):
    d = fixture(name)
    _, sol = oracle_solve(d)
    check_modes(d, sol)


def test_oracle_duct():
    d = fixture("duct")
    duct = O.DuctFlow(float(d["R"]), float(d["L"]), int(d["n_theta"]), int(d["n_z"]), int(d["n_r"]), float(d["mu"]))
    u, f, Q = duct.solve(float(d["p_in"]), float(d["p_out"]))
    assert rel(u, d["u"]) <= 1e-12 and rel(f, d["f"]) <= 1e-12
    assert abs(Q - float(d["Q"])) <= 1e-12 * abs(float(d["Q"]))


def parse_csv(text):
    rows = [ln for ln in str(text).splitlines() if ln and not ln.startswith("#") and not ln.startswith("t,")]
    return np.array([[float(v) for v in ln.split(",")] for ln in rows])


def test_oracle_run_trajectory_against_reference_steps():
    d = fixture("traj")
    P = O.TrajectoryParams(alpha=float(d["alpha"]), a_over_R=float(d["a_over_R"]), beta0=float(d["beta0"]),
                           pitch0_deg=float(d["pitch0_deg"]), yaw0_deg=float(d["yaw0_deg"]), dt=float(d["dt"]),
                           t_max=float(d["dt"]) * int(d["n_steps"]), L_over_R=float(d["L_over_R"]),
                           n_theta=int(d["n_theta"]), sphere_level=int(d["sphere_level"]),
                           beta_stop=float(d["beta_stop"]))
    # the reference resumed from its own checkpoints prints the same rows as one run
    assert str(d["csv_full"]) == str(d["csv_resumed"])
    ref_rows = parse_csv(d["csv_full"])
    n = int(d["n_steps"])
    rows, status, y = O.run_trajectory(P, max_steps=n - 1)
    # states after steps 1 .. n-1 at 17 digits (the reference's checkpoints)
    states = rows[1:, 1:7]
    scale = np.abs(d["states"]).max(axis=0)
    err = np.abs(states - d["states"][:n - 2]).max(axis=0) / scale
    assert err.max() <= 1e-11, err
    assert np.abs(y - d["states"][n - 2]).max() <= 1e-11 * np.abs(d["states"]).max()
    # every printed row (t, state, U, Omega, beta at 10 digits) of the uninterrupted run
    rows_full, status, _ = O.run_trajectory(P)
    assert status == "finished" and len(rows_full) == len(ref_rows) == n
    colmax = np.abs(ref_rows).max(axis=0)
    assert (np.abs(rows_full - ref_rows) <= 6e-10 * np.maximum(np.abs(ref_rows), 1e-3 * colmax) + 1e-14).all()
    # the newest derivative in each checkpoint is (U, Omega x e) at the state before the step
    for k in range(n - 1):
        U, Om = rows_full[k, 7:10], rows_full[k, 10:13]
        e = rows_full[k, 4:7] / np.linalg.norm(rows_full[k, 4:7])
        fd = np.r_[U, np.cross(Om, e)]
        assert np.abs(fd - d["deriv_newest"][k]).max() <= 1e-9 * np.abs(d["deriv_newest"]).max()


def test_reference_trajectory_revisions_agree():
    """The reference's two code revisions' runs of the neutral path (outputs/reference/
    trajectories/neutral.csv and outputs/trajectories/neutral.csv) agree to 1.7e-10 of each
    column's largest value: the premise of the GPU trajectory tolerance."""
    def read(name):
        lines = [ln for ln in open(os.path.join(GOLDEN, "reference_outputs", name)) if ln.strip() and ln[0] != "#"]
        return np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])

    a, b = read("neutral.csv"), read("neutral_rev2.csv")
    assert a.shape == b.shape == (300, 14)
    assert (np.abs(a - b) / np.abs(a).max(axis=0).clip(1e-300)).max() < 2e-10
