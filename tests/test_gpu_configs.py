"""Every benchmark configuration on the GPU, against the reference's own solver.

Reference values are 17-digit outputs of the reference's src/swimmer.hpp built from source
(tests/golden/ref_solver_*.npz, see tests/golden/make_ref_solver.py). Bars:
  * FP64 GPU path vs the reference: 1e-10 norm-wise for U, Omega, sphere and tube tractions
    (modes that vanish by symmetry are measured against the largest mode of the set);
  * refined modes (FP32, FP16-MFMA and BF16-MFMA factors + FP64 refinement) vs the FP64 GPU
    path: 1e-8, with the refinement's final relative residual <= 1e-12;
  * adjoint gradients vs central differences (h = 1e-4): 1e-5 of each row's largest entry.

Configurations (R = 1):
  1  L/R=4, n_theta=24, n_r=6, level 2, a=0.3, on axis, e=+z (5,430 DOF), FP64
  2  L/R=6, n_theta=32, n_r=8, level 3, a=0.5, b=0.15, 30-degree tilt, B2/B1=-3 (13,254 DOF)
  3  256 configurations from seed 0 (4,710-19,014 DOF), batched, FP32 + FP64 refinement
  4  L/R=8, n_theta=24, n_r=6, level 3, a=0.3, b=0.14, 10-degree tilt (9,750 DOF): gradient
     of (U_z, Omega_x); trajectory adjoint over 60 steps of dt=0.01, checkpoint every 50
  5  L/R=8, n_theta=64, n_r=16, level 4, a=0.4, b=0.15, B2/B1=+3 (57,222 DOF), FP32 + IR vs
     FP64, velocity on the masked 41 x 41 x 201 grid
  46k  L/R=4, n_theta=64, n_r=16, level 4, a=0.4, b=0.15, B2/B1=+3 (46,470 DOF), the
     benchmark's mesh: FP64 vs the reference, BF16 + IR vs FP64
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FP64, FP32_IR, FP16_IR, BF16_IR = 0, 1, 2, 3


def fixture(name):
    path = os.path.join(GOLDEN, f"ref_solver_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{os.path.basename(path)} not generated (tests/golden/make_ref_solver.py --large)")
    return np.load(path)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def tube_of(sq, d, prec=FP64):
    if int(d["unbounded"]):
        return None
    R, L, a, nth, nr = float(d["R"]), float(d["L"]), float(d["a"]), int(d["n_theta"]), int(d["n_r"])
    mesh = sq.tube_mesh(R, d["z"], nth, nr) if "z" in d else None
    return sq.Tube(R, L, a, nth, nr, precision=prec, mesh=mesh)


def solve(sq, d, prec=FP64, tube=None):
    """GPU solve of a fixture's configuration: (solver, solutions of the fixture's modes)."""
    print(f"solve {d['Nt'] if 'Nt' in d else ''} + {d['Ns']} panels, precision {prec}", flush=True)
    if tube is None:
        tube = tube_of(sq, d, prec)
    a, c, e = float(d["a"]), d["center"], d["e"]
    s = sq.SwimmerSolver(tube, sq.icosphere(a, int(d["level"]), c), c, precision=prec)
    sol = s.solve_free([s.slip(e, m[0], m[1]) for m in d["modes"]])
    if "combine" in d:
        w = d["combine"]
        sol = [sq.SwimmerSolution(sum(wi * q.U for wi, q in zip(w, sol)), sum(wi * q.Om for wi, q in zip(w, sol)),
                                  sum(wi * q.f_sphere for wi, q in zip(w, sol)),
                                  sum(wi * q.f_tube for wi, q in zip(w, sol)))]
    return s, sol


def errors(sol, U, Om, fs, ft=None):
    """Norm-wise errors of a set of solutions against reference arrays (per mode), the
    (U, Omega) error measured against the largest (U, Omega) of the set."""
    scale = max(np.linalg.norm(np.r_[U[k], Om[k]]) for k in range(len(U)))
    out = []
    for k, g in enumerate(sol):
        e = {"UOm": float(np.linalg.norm(np.r_[g.U - U[k], g.Om - Om[k]]) / scale), "fs": rel(g.f_sphere, fs[k])}
        if ft is not None:
            e["ft"] = rel(g.f_tube, ft[k])
        out.append(e)
    return out


def assert_errors(errs, tol):
    print(errs, flush=True)
    for e in errs:
        for k, v in e.items():
            assert v <= tol, (k, v, tol)


def vs_fixture(sol, d):
    return errors(sol, d["U"], d["Om"], d["f_sphere"], d["f_tube"] if "f_tube" in d else None)


def vs_solution(sol, ref):
    return errors(sol, [r.U for r in ref], [r.Om for r in ref], [r.f_sphere for r in ref],
                  [r.f_tube for r in ref] if len(ref[0].f_tube) else None)


# ------------------------------------------------------------------ configuration 1 + tiny
def test_cfg1_fp64_vs_reference_towed_and_field(sq):
    d = fixture("cfg1")
    s, sol = solve(sq, d)
    assert_errors(vs_fixture(sol, d), 1e-10)
    t = s.solve_prescribed(np.array([0.0, 0.0, 1.0]), np.zeros(3))
    Fz = s.force(t)[2]
    assert abs(Fz - d["towed_F"][0][2]) <= 1e-10 * abs(d["towed_F"][0][2]), (Fz, d["towed_F"][0])
    v = s.velocity(sol[0], s.slip(d["e"], 1, 0), d["field_pts"])
    assert np.abs(v - d["field_u"][0]).max() <= 1e-10 * np.abs(d["field_u"][0]).max()


@pytest.mark.parametrize("name", ["tiny", "tiny_unbounded"])
def test_tiny_schur_action_through_towed_problems(sq, name):
    """Gs^-1 Gc (the rigid-body columns of the Schur complement) pinned through the six
    towed problems w = unit rigid motions, confined and unbounded, off axis and tilted."""
    d = fixture(name)
    s, sol = solve(sq, d)
    assert_errors(vs_fixture(sol, d), 1e-10)
    Fscale = np.linalg.norm(d["towed_F"], axis=1).max()
    for w, F, fs in zip(d["towed_w"], d["towed_F"], d["towed_f_sphere"]):
        t = s.solve_prescribed(w[:3], w[3:])
        assert np.linalg.norm(s.force(t) - F) <= 1e-10 * Fscale
        assert rel(t.f_sphere, fs) <= 1e-10
    for k, g in enumerate(sol):
        v = s.velocity(g, s.slip(d["e"], *d["modes"][k]), d["field_pts"])
        assert np.abs(v - d["field_u"][k]).max() <= 1e-10 * np.abs(d["field_u"]).max()


# ------------------------------------------------------------------ configuration 2
def test_cfg2_fp64_vs_reference(sq):
    d = fixture("cfg2")
    assert 3 * (int(d["Nt"]) + int(d["Ns"])) + 6 == 13254
    _, sol = solve(sq, d)
    assert_errors(vs_fixture(sol, d), 1e-10)
    # the configuration's swimmer, B2/B1 = -3
    U = sol[0].U - 3 * sol[1].U
    Om = sol[0].Om - 3 * sol[1].Om
    Ur, Omr = d["U"][0] - 3 * d["U"][1], d["Om"][0] - 3 * d["Om"][1]
    assert np.linalg.norm(np.r_[U - Ur, Om - Omr]) <= 1e-10 * np.linalg.norm(np.r_[Ur, Omr])


@pytest.mark.parametrize("prec", [FP32_IR, FP16_IR, BF16_IR])
def test_cfg2_refined_modes_vs_fp64(sq, prec):
    d = fixture("cfg2")
    _, ref = solve(sq, d)
    s, sol = solve(sq, d, prec)
    print(f"precision {prec}: {s.ir_iters} refinement steps, residual {s.ir_residual:.2e}", flush=True)
    assert 0 < s.ir_residual <= 1e-12
    assert_errors(vs_solution(sol, ref), 1e-8)
    assert_errors(vs_fixture(sol, d), 1e-8)


# ------------------------------------------------------------------ configuration 3
def test_cfg3_batch_256_seed0(sq):
    params = sq.draw_batch_params(256, seed=0)
    inst = sq.batch_instances(params)
    dof = np.array([3 * (q.tube.P() + q.sphere.P()) + 6 for q in inst])
    assert dof.min() == 4710 and dof.max() == 19014, (dof.min(), dof.max())
    print("batch of 256, FP32 + FP64 refinement", flush=True)
    r = sq.batch_solve(inst, precision=FP32_IR)
    print(r.stats, flush=True)
    assert np.all(r.resid > 0) and np.all(r.resid <= 1e-12), r.resid.max()
    # the smallest instance against the reference
    d = fixture("batch_smallest")
    i = int(d["index"])
    assert i == int(np.argmin(dof))
    got = np.r_[r.U[i], r.Om[i]]
    ref = np.r_[d["U"][0], d["Om"][0]]
    assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref), np.linalg.norm(got - ref) / np.linalg.norm(ref)
    # four instances against single FP64 solves
    rng = np.random.default_rng(5)
    pick = [int(np.argmin(dof)), int(np.argmax(dof))] + [int(k) for k in rng.choice(256, 2, replace=False)]
    for k in pick:
        q, pr = inst[k], params[k]
        print(f"instance {k}: {dof[k]} DOF", flush=True)
        t = sq.Tube(1.0, pr["L"], pr["a"], pr["n_theta"], pr["n_r"], mesh=q.tube)
        s = sq.SwimmerSolver(t, q.sphere, q.center)
        one = s.solve_free([s.slip(q.e, 1.0, q.B2)])[0]
        ref = np.r_[one.U, one.Om]
        err = np.linalg.norm(np.r_[r.U[k], r.Om[k]] - ref) / np.linalg.norm(ref)
        assert err <= 1e-8, (k, err)


# ------------------------------------------------------------------ configuration 4
def test_cfg4_fp64_vs_reference(sq):
    d = fixture("cfg4")
    assert 3 * (int(d["Nt"]) + int(d["Ns"])) + 6 == 9750
    _, sol = solve(sq, d)
    assert_errors(vs_fixture(sol, d), 1e-10)


def test_cfg4_gradient_vs_central_differences(sq):
    """d(U_z, Omega_x)/d(x_c, y_c, z_c, pitch, yaw, B2/B1) at 9,750 DOF by the adjoint,
    against central differences of the FP64 solve.

    With the configuration's step h = 1e-4 some tube-sphere pairs change quadrature rule
    between x_c - h and x_c + h (their distance ratios lie within h / h_panel of a rule
    threshold, src/utils/kernels.hpp:74-75), so those differences include the rules' small
    jump; they must still agree to 1e-5. The frozen-level check then takes the largest step
    at which no pair changes rule, where the differences see only the smooth branch the
    adjoint differentiates, and checks the centre columns there too."""
    import test_gpu_adjoint as A

    tube = sq.Tube(1.0, 8.0, 0.3, 24, 6)
    assert 3 * (tube.Nt + 1280) + 6 == 9750
    p = np.array([0.2 * 0.7, 0.0, 0.0, math.radians(10.0), 0.0, -1.0])
    model = A.Model(sq, tube, 0.3, 3, h=1e-4)
    A.check_against_fd(model, p)
    for h in (3e-5, 1e-5, 3e-6, 1e-6, 3e-7):
        model.h = h
        if model.no_level_change(p[:3]):
            break
    else:
        pytest.fail("no step down to 3e-7 keeps every pair's quadrature rule")
    print(f"frozen-level step h = {h:g}", flush=True)
    _, adj = model.value_and_jacobian(p)
    fd = np.zeros((2, 3))
    for k in range(3):
        dp = np.zeros(6)
        dp[k] = h
        fd[:, k] = (model.value(p + dp) - model.value(p - dp)) / (2 * h)
    print("adjoint (centre)\n", adj[:, :3], "\nfinite differences\n", fd, flush=True)
    for i in range(2):
        assert np.max(np.abs(adj[i, :3] - fd[i])) <= 1e-5 * np.max(np.abs(adj[i])), (i, adj[i, :3], fd[i])


def test_cfg4_trajectory_adjoint_across_a_checkpoint(sq):
    """d(final beta, final tilt)/d(beta0, tilt0) over 60 steps of dt = 0.01 with a checkpoint
    every 50 steps (so the backward sweep recomputes two segments), against central
    differences of the same FP64 integration."""
    def P(beta0, pitch0):
        return sq.TrajectoryParams(alpha=-1.0, a_over_R=0.3, beta0=beta0, pitch0_deg=pitch0, yaw0_deg=0.0, dt=0.01,
                                   L_over_R=8.0, n_theta=24, sphere_level=3, n_r=6)

    n, ck, h = 60, 50, 1e-4
    print("trajectory adjoint, 60 steps", flush=True)
    r = sq.trajectory_gradient(P(0.2, 10.0), n, ck)
    assert r["evaluations"] == 2 * (n + 9)

    def final(b0, p0):
        q = sq.trajectory_gradient(P(b0, p0), n, ck, adjoint=False)
        return np.array([q["beta_final"], q["tilt_final_deg"]])

    fd = np.stack([(final(0.2 + h, 10.0) - final(0.2 - h, 10.0)) / (2 * h),
                   (final(0.2, 10.0 + h) - final(0.2, 10.0 - h)) / (2 * h)], axis=1)
    adj = np.array([[r["dbeta_dbeta0"], r["dbeta_dpitch0"]], [r["dtilt_dbeta0"], r["dtilt_dpitch0"]]])
    print("adjoint\n", adj, "\nfinite differences\n", fd, flush=True)
    for i in range(2):
        assert np.max(np.abs(adj[i] - fd[i])) <= 1e-5 * np.max(np.abs(fd[i])), (i, adj[i], fd[i])


# ------------------------------------------------------------------ configuration 5 and 46k
def masked_grid(R, L, xc, a):
    g = np.linspace(-R, R, 41)
    z = np.linspace(-L / 2, L / 2, 201)
    X, Y, Z = np.meshgrid(g, g, z, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    keep = (np.hypot(pts[:, 0], pts[:, 1]) < 0.97 * R) & (np.linalg.norm(pts - xc, axis=1) >= 1.03 * a)
    return np.ascontiguousarray(pts[keep])


def large_case(sq, d, prec_mixed):
    """FP64 solve (vs the reference fixture when present) and a refined solve vs FP64."""
    s64, sol64 = solve(sq, d)
    assert_errors(vs_fixture(sol64, d), 1e-10)
    smx, solmx = solve(sq, d, prec_mixed)
    print(f"precision {prec_mixed}: {smx.ir_iters} refinement steps, residual {smx.ir_residual:.2e}", flush=True)
    assert 0 < smx.ir_residual <= 1e-12
    assert_errors(vs_solution(solmx, sol64), 1e-8)
    return s64, sol64, smx, solmx


def test_cfg5_refined_vs_fp64_and_masked_field(sq, O):
    d = fixture("cfg5")
    assert 3 * (int(d["Nt"]) + int(d["Ns"])) + 6 == 57222
    s64, sol64, smx, solmx = large_case(sq, d, FP32_IR)
    R, L, a, xc = 1.0, 8.0, 0.4, d["center"]
    pts = masked_grid(R, L, xc, a)
    assert len(pts) == 235264
    slip = s64.slip(d["e"], 1.0, 3.0)
    print("field on the masked 41 x 41 x 201 grid", flush=True)
    u64 = s64.velocity(sol64[0], slip, pts)
    umx = smx.velocity(solmx[0], slip, pts)
    assert rel(umx, u64) <= 1e-8
    # a few hundred grid points against the oracle's velocity_at on the same solution
    tm = s64.tube.mesh
    allm = O.merge(O.swimmer_tube_mesh(R, L, a, 64, 16), O.icosphere(a, 4, xc))
    assert allm.P() == tm.P() + s64.Ns
    uu = np.zeros((allm.P(), 3))
    ff = np.zeros((allm.P(), 3))
    ff[:tm.P()] = sol64[0].f_tube
    ff[tm.P():] = sol64[0].f_sphere
    uu[tm.P():] = sol64[0].U + np.cross(sol64[0].Om, s64.S.centroid - xc) + slip
    sel = np.random.default_rng(3).choice(len(pts), 300, replace=False)
    uo = O.velocity_at(allm, uu, ff, 1.0, pts[sel])
    assert np.abs(u64[sel] - uo).max() <= 1e-10 * np.abs(uo).max()


def test_46k_fp64_vs_reference_and_bf16_vs_fp64(sq):
    d = fixture("46k")
    assert 3 * (int(d["Nt"]) + int(d["Ns"])) + 6 == 46470
    large_case(sq, d, BF16_IR)
