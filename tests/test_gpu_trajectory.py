"""run_trajectory's restart, pause and wall-contact paths on the GPU (src/trajectory.hpp:
69-104, 111-117, 143, 149), and its steps against the reference's own integrator at 17
digits.

* Pause and resume: a budget of 1e-9 s lets every call take exactly one step (the budget is
  only checked once an evaluation has been timed) and then pause, keeping <prefix>.ckpt;
  the next call resumes from it and appends to the CSV. Resuming from the 17-digit
  checkpoint continues the same arithmetic, so the CSV of the paused-and-resumed run equals
  the uninterrupted run's byte for byte, and the checkpoint is gone once the run finishes.
* Wall contact: with beta_stop between two rows' offsets the run ends before the first step
  from a state beyond it, with exactly the rows before that state, and no checkpoint.
* The reference's run_trajectory, advanced one step per call the same way
  (tests/golden/ref_solver_traj.npz, written from the reference built from source), gives
  the state after every step at 17 digits: the GPU's states agree to 1e-10.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def rows_of(path):
    with open(path) as f:
        lines = [ln for ln in f if ln.strip() and not ln.startswith("#")]
    assert lines[0].strip() == "t,x,y,z,ex,ey,ez,Ux,Uy,Uz,Omx,Omy,Omz,beta"
    return np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])


def ref_params(sq, d, **over):
    kw = dict(alpha=float(d["alpha"]), a_over_R=float(d["a_over_R"]), beta0=float(d["beta0"]),
              pitch0_deg=float(d["pitch0_deg"]), yaw0_deg=float(d["yaw0_deg"]), dt=float(d["dt"]),
              t_max=float(d["dt"]) * int(d["n_steps"]), L_over_R=float(d["L_over_R"]), n_theta=int(d["n_theta"]),
              sphere_level=int(d["sphere_level"]), beta_stop=float(d["beta_stop"]))
    kw.update(over)
    return sq.TrajectoryParams(**kw)


def test_trajectory_pause_and_resume_equals_one_run(sq, tmp_path):
    d = np.load(os.path.join(GOLDEN, "ref_solver_traj.npz"))
    n = int(d["n_steps"])
    full = str(tmp_path / "full")
    assert sq.run_trajectory(ref_params(sq, d), full) == "finished"
    assert not os.path.exists(full + ".ckpt")
    part = str(tmp_path / "part")
    for k in range(n):
        st = sq.run_trajectory(ref_params(sq, d, budget_s=1e-9), part)
        print(f"call {k}: {st}", flush=True)
        if k < n - 1:
            assert st == "paused"
            assert os.path.exists(part + ".ckpt")
            head = open(part + ".ckpt").readline().split()
            assert int(head[1]) == k + 1 and int(head[2]) == min(k + 1, 3)
        else:
            assert st == "finished"
    assert not os.path.exists(part + ".ckpt")
    a, b = open(full + ".csv").read(), open(part + ".csv").read()
    assert a == b, "the resumed run differs from the uninterrupted one"
    assert a.count("\n") == n + 2  # the parameter comment, the header, n rows


def test_trajectory_states_match_reference_at_17_digits(sq, tmp_path):
    d = np.load(os.path.join(GOLDEN, "ref_solver_traj.npz"))
    n = int(d["n_steps"])
    pre = str(tmp_path / "t")
    states = []
    for k in range(n - 1):
        assert sq.run_trajectory(ref_params(sq, d, budget_s=1e-9), pre) == "paused"
        tok = open(pre + ".ckpt").read().split()
        states.append([float(v) for v in tok[3:9]])
    states = np.array(states)
    scale = np.abs(d["states"]).max(axis=0)
    err = (np.abs(states - d["states"]) / scale).max(axis=0)
    print("state error / scale after each step:", err, flush=True)
    assert err.max() <= 1e-10
    # the rows (10 digits) of the reference's uninterrupted run
    rows = [ln for ln in str(d["csv_full"]).splitlines() if ln and ln[0] != "#" and not ln.startswith("t,")]
    ref = np.array([[float(v) for v in ln.split(",")] for ln in rows])
    got = rows_of(pre + ".csv")
    assert got.shape == (n - 1, 14)
    colmax = np.abs(ref).max(axis=0)
    assert (np.abs(got - ref[:n - 1]) <= 2e-9 * np.abs(ref[:n - 1]) + 1e-9 * colmax).all()
    assert open(pre + ".csv").readline().strip() == str(d["csv_full"]).splitlines()[0]


def test_trajectory_wall_contact_stop(sq, tmp_path):
    d = np.load(os.path.join(GOLDEN, "ref_solver_traj.npz"))
    # toward the wall: the offset grows every step
    base = dict(beta0=0.6, pitch0_deg=-30.0, yaw0_deg=0.0, t_max=8 * float(d["dt"]))
    free = str(tmp_path / "free")
    assert sq.run_trajectory(ref_params(sq, d, beta_stop=10.0, **base), free) == "finished"
    beta = rows_of(free + ".csv")[:, 13]
    assert np.all(np.diff(beta) > 0), beta
    j = 5
    stop = 0.5 * (beta[j - 1] + beta[j])  # between the offsets of rows j - 1 and j
    wall = str(tmp_path / "wall")
    assert sq.run_trajectory(ref_params(sq, d, beta_stop=stop, **base), wall) == "wall contact"
    assert not os.path.exists(wall + ".ckpt")
    got = rows_of(wall + ".csv")
    assert got.shape[0] == j
    assert np.array_equal(got, rows_of(free + ".csv")[:j])


@pytest.mark.parametrize("prec", [0, 1])
def test_moved_solver_equals_solver_on_translated_mesh(sq, prec):
    """The trajectory's solver: sphere-sphere blocks copied from a template assembled at the
    origin, the sphere translated by the shift. Against a solver assembled from scratch on
    the translated mesh (the reference's SwimmerSolver(&tube, S.translate(c), c)): the blocks
    differ only by the rounding of the translated coordinates, so U, Omega and tractions
    agree to 1e-12 (FP64; refined: both refined to 1e-12, agree to 1e-10)."""
    a = 1.0
    R = a / 0.3
    tube = sq.Tube(R, 4 * R, a, 16, 4, precision=prec)
    blocks = sq.SphereBlocks(sq.icosphere(a, 2), np.zeros(3))
    e = np.array([0.3, -0.2, 0.9])
    e /= np.linalg.norm(e)
    tol = 1e-12 if prec == 0 else 1e-10
    for shift in ([-1.1, 0.4, 0.0], [0.3, 0.2, 0.0]):
        c = np.array(shift)
        fresh = sq.SwimmerSolver(tube, sq.icosphere(a, 2, c), c)
        moved = sq.SwimmerSolver.moved(tube, blocks, c)
        assert np.array_equal(moved.S.centroid, fresh.S.centroid) and np.array_equal(moved.xc, fresh.xc)
        A = fresh.solve_free([fresh.slip(e, 1, 0), fresh.slip(e, 0, 1)])
        B = moved.solve_free([moved.slip(e, 1, 0), moved.slip(e, 0, 1)])
        scale = max(np.linalg.norm(np.r_[x.U, x.Om]) for x in A)
        for x, y in zip(A, B):
            assert np.linalg.norm(np.r_[x.U - y.U, x.Om - y.Om]) <= tol * scale
            assert np.linalg.norm(x.f_sphere - y.f_sphere) <= tol * np.linalg.norm(x.f_sphere)
            assert np.linalg.norm(x.f_tube - y.f_tube) <= tol * np.linalg.norm(x.f_tube)
