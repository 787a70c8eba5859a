"""Adjoint sensitivities against central finite differences of the FP64 solver.

The library's adjoint (sqbem_swimmer_gradient: transposed solves through the tube's and the
bordered matrix's LU factors, plus the contraction of the translation derivative of the
tube-sphere coupling kernels) is wrapped in a torch.autograd.Function whose forward pass is
the free-swimming solve and whose backward pass is the adjoint. Its gradients of (U_z,
Omega_x) with respect to the sphere centre, two orientation angles and B2/B1 are compared
with central differences (step 1e-4) of the FP64 solve itself, which reproduces the
reference's solutions (test_gpu_golden.py, test_gpu_parity.py).

The discrete solution is only piecewise smooth in the sphere centre: a panel pair whose
distance ratio crosses a quadrature-rule threshold (d = 0.75, 1.5, 3; src/utils/kernels.hpp:
74-75) switches rule. The adjoint differentiates the branch at the base point, so each test
first checks that no tube-sphere pair changes rule level between x_c - h and x_c + h. With
near pairs present, some pair lies within ~1e-5 of a threshold for almost every centre, so
that case uses h = 2e-6 at a centre whose closest pair is 2e-5 away; the all-far case uses
h = 1e-4.

Tolerance: 1e-5 relative to the largest entry of each gradient row (central differences:
O(h^2) truncation, and rounding of ~1e-13 / h).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

H = 1e-4
RTOL = 1e-5
G_UZ_OMX = np.array([[0, 0, 1, 0, 0, 0], [0, 0, 0, 1, 0, 0]], dtype=np.float64)


def orient(pitch, yaw):
    """e = (cos(yaw) sin(pitch), sin(yaw), cos(yaw) cos(pitch)) (src/trajectory.hpp:86-87)."""
    return np.array([math.cos(yaw) * math.sin(pitch), math.sin(yaw), math.cos(yaw) * math.cos(pitch)])


def orient_jac(pitch, yaw):
    """d e / d (pitch, yaw) as a 3 x 2 matrix."""
    return np.array([[math.cos(yaw) * math.cos(pitch), -math.sin(yaw) * math.sin(pitch)],
                     [0.0, math.cos(yaw)],
                     [-math.cos(yaw) * math.sin(pitch), -math.sin(yaw) * math.cos(pitch)]])


def rule_levels(targets, src):
    """Quadrature level of every (target, source panel) pair, as the assembly picks it."""
    d = src.centroid[None, :, :] - targets[:, None, :]
    dist = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]) / src.h[None, :]
    return np.where(dist < 0.75, 3, np.where(dist < 1.5, 2, np.where(dist < 3.0, 1, 0)))


class Model:
    """A confined squirmer: parameters (x_c, y_c, z_c, pitch, yaw, alpha) -> (U_z, Omega_x)."""

    def __init__(self, sq, tube, a, level, h=H):
        self.sq, self.tube, self.a, self.level, self.h = sq, tube, a, level, h

    def solver(self, xc):
        S = self.sq.icosphere(self.a, self.level, center=xc)
        return self.sq.SwimmerSolver(self.tube, S, xc)

    def value(self, p):
        sw = self.solver(p[:3])
        sol = sw.solve_free([sw.slip(orient(p[3], p[4]), 1.0, p[5])])[0]
        return np.array([sol.U[2], sol.Om[0]])

    def value_and_jacobian(self, p):
        sw = self.solver(p[:3])
        r = sw.gradient(orient(p[3], p[4]), 1.0, p[5], G_UZ_OMX)
        jac = np.zeros((2, 6))
        jac[:, :3] = r["dxc"]
        jac[:, 3:5] = r["de"] @ orient_jac(p[3], p[4])
        jac[:, 5] = r["dB"][:, 1]  # B2 = alpha, B1 = 1
        return np.array([r["U"][2], r["Om"][0]]), jac

    def no_level_change(self, xc):
        if self.tube is None:
            return True
        tm = self.tube.mesh
        for k in range(3):
            ref_t = ref_s = None
            for s in (-1.0, 1.0):
                c = np.array(xc, dtype=np.float64)
                c[k] += s * self.h
                S = self.sq.icosphere(self.a, self.level, center=c)
                lt, ls = rule_levels(tm.centroid, S), rule_levels(S.centroid, tm)
                if ref_t is None:
                    ref_t, ref_s = lt, ls
                elif not (np.array_equal(lt, ref_t) and np.array_equal(ls, ref_s)):
                    return False
        return True


class SwimmerFunctional(torch.autograd.Function):
    """(U_z, Omega_x) of the confined squirmer; backward through the library's adjoint."""

    @staticmethod
    def forward(ctx, params, model):
        J, jac = model.value_and_jacobian(params.detach().cpu().numpy().astype(np.float64))
        ctx.save_for_backward(torch.from_numpy(jac))
        return torch.from_numpy(J)

    @staticmethod
    def backward(ctx, gout):
        (jac,) = ctx.saved_tensors
        return gout @ jac, None


def fd_jacobian(model, p):
    jac = np.zeros((2, 6))
    for k in range(6):
        dp = np.zeros(6)
        dp[k] = model.h
        jac[:, k] = (model.value(p + dp) - model.value(p - dp)) / (2 * model.h)
    return jac


def check_against_fd(model, p):
    params = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    J = SwimmerFunctional.apply(params, model)
    rows = [torch.autograd.grad(J[i], params, retain_graph=True)[0].numpy() for i in range(2)]
    adj = np.stack(rows)
    fd = fd_jacobian(model, np.asarray(p, dtype=np.float64))
    print("adjoint\n", adj, "\nfinite differences\n", fd, flush=True)
    np.testing.assert_allclose(J.detach().numpy(), model.value(np.asarray(p)), rtol=1e-12, atol=0)
    for i in range(2):
        scale = np.max(np.abs(fd[i]))
        err = np.max(np.abs(adj[i] - fd[i]))
        assert err <= RTOL * scale, (i, err / scale, adj[i], fd[i])


def test_swimmer_gradient_confined_far_pairs(sq):
    # a/R = 0.15 in a 24-panel tube: every tube-sphere pair uses the 7-point rule
    tube = sq.Tube(1.0, 4.0, 0.15, 24, 4)
    model = Model(sq, tube, 0.15, 1)
    p = [0.09, -0.05, 0.07, 0.35, 0.2, -1.5]
    assert model.no_level_change(p[:3])
    check_against_fd(model, p)


def test_swimmer_gradient_confined_near_pairs(sq):
    # a/R = 0.3 in a 16-panel tube, level-2 sphere: 13k sphere-tube pairs use the 28-point
    # rule, so the derivative of the near-field quadrature is exercised
    tube = sq.Tube(1.0, 4.0, 0.3, 16, 4)
    model = Model(sq, tube, 0.3, 2, h=2e-6)
    p = [0.215, 0.092, -0.119, -0.4, 0.3, 2.5]
    assert model.no_level_change(p[:3])
    check_against_fd(model, p)


def test_swimmer_gradient_unbounded(sq):
    # free space: translation invariant (dJ/dx_c = 0); orientation and mode derivatives
    model = Model(sq, None, 0.3, 1)
    p = [0.0, 0.0, 0.0, 0.5, -0.2, 1.0]
    value, jac = model.value_and_jacobian(np.asarray(p))
    assert np.all(jac[:, :3] == 0.0)
    fd = fd_jacobian(model, np.asarray(p))
    for i in range(2):
        scale = np.max(np.abs(fd[i]))
        assert np.max(np.abs(jac[i] - fd[i])) <= RTOL * scale + 1e-10, (jac[i], fd[i])


def test_swimmer_gradient_rejects_mixed_precision(sq):
    tube = sq.Tube(1.0, 4.0, 0.3, 12, 3, precision=sq.FP32_IR)
    S = sq.icosphere(0.3, 1)
    sw = sq.SwimmerSolver(tube, S, np.zeros(3))
    with pytest.raises(RuntimeError, match="FP64 swimmer"):
        sw.gradient(np.array([0, 0, 1.0]), 1.0, 0.0, G_UZ_OMX)


def traj_params(sq, beta0, pitch0):
    # a/R = 0.15: the tube-sphere pairs stay in the far field along the path
    return sq.TrajectoryParams(alpha=-1.0, a_over_R=0.15, beta0=beta0, pitch0_deg=pitch0, yaw0_deg=0.0, dt=0.05,
                               L_over_R=4.0, n_theta=24, sphere_level=1, n_r=4)


def test_trajectory_forward_matches_run_trajectory(sq, tmp_path):
    # the gradient's forward pass is run_trajectory's integration (its CSV row at step n is
    # the state before step n, so t_max = (n + 1) dt prints the state after n steps)
    n = 6
    P = traj_params(sq, 0.2, 10.0)
    P.t_max = (n + 1) * P.dt + 1e-12
    sq.run_trajectory(P, str(tmp_path / "tr"))
    with open(tmp_path / "tr.csv") as f:
        lines = [ln for ln in f if ln.strip() and not ln.startswith("#")]
    last = np.array([float(v) for v in lines[-1].split(",")])
    beta = math.hypot(last[1], last[2]) / (1 / 0.15 - 1)
    tilt = math.degrees(math.atan2(math.hypot(last[4], last[5]), last[6]))
    r = sq.trajectory_gradient(traj_params(sq, 0.2, 10.0), n, 4, adjoint=False)
    assert abs(r["beta_final"] - beta) <= 1e-9 * abs(beta)
    assert abs(r["tilt_final_deg"] - tilt) <= 1e-9 * abs(tilt)


def test_trajectory_gradient_fd(sq):
    # 10 steps (3 RK4 start-up steps, 7 AB4 steps), checkpoints every 4 steps
    n, ck = 10, 4
    r = sq.trajectory_gradient(traj_params(sq, 0.2, 10.0), n, ck)
    print(r, flush=True)
    assert r["evaluations"] == 2 * (n + 9)  # forward + recomputation (RK4 steps evaluate 4 times)

    def final(beta0, pitch0):
        q = sq.trajectory_gradient(traj_params(sq, beta0, pitch0), n, ck, adjoint=False)
        return np.array([q["beta_final"], q["tilt_final_deg"]])

    d_beta0 = (final(0.2 + H, 10.0) - final(0.2 - H, 10.0)) / (2 * H)
    d_pitch0 = (final(0.2, 10.0 + H) - final(0.2, 10.0 - H)) / (2 * H)
    adj = np.array([[r["dbeta_dbeta0"], r["dbeta_dpitch0"]], [r["dtilt_dbeta0"], r["dtilt_dpitch0"]]])
    fd = np.stack([d_beta0, d_pitch0], axis=1)
    print("adjoint\n", adj, "\nfinite differences\n", fd, flush=True)
    for i in range(2):
        scale = np.max(np.abs(fd[i]))
        assert np.max(np.abs(adj[i] - fd[i])) <= RTOL * scale, (i, adj[i], fd[i])


def reference_value_fn(kind, R, L, a, nth, nr, level):
    """(x_c, pitch, yaw, alpha) -> (U_z, Omega_x) by the reference's own FP64 solver (built
    from source, oracle/_ref/libref_solver.so) or by the oracle's restatement of it."""
    if kind == "reference":
        import ref_solver as RS

        if not RS.available():
            pytest.skip("reference solver build absent (make -C oracle refsolver)")
        tube = RS.Tube(R, L, a, nth, nr) if nth else None

        def value(p):
            s = RS.SwimmerSolver(tube, a, level, p[:3])
            sol = s.solve_free(orient(p[3], p[4]), [(1.0, p[5])])[0]
            return np.array([sol.U[2], sol.Om[0]])
    else:
        import oracle as O

        tube = O.Tube(R, L, a, nth, nr) if nth else None

        def value(p):
            s = O.SwimmerSolver(tube, O.icosphere(a, level, p[:3]), p[:3])
            sol = s.solve_free([s.slip(orient(p[3], p[4]), 1.0, p[5])])[0]
            return np.array([sol.U[2], sol.Om[0]])
    return value


@pytest.mark.parametrize("kind", ["reference", "oracle"])
@pytest.mark.parametrize("confined", [True, False])
def test_swimmer_gradient_vs_fd_of_reference_solver(sq, kind, confined):
    """The GPU adjoint against central differences (h = 1e-4) of the reference's FP64 solver
    (the reference's own headers built from source) and of the oracle: confined with far
    tube-sphere pairs only (no rule can change within the step), and unbounded."""
    R, L, a, nth, nr, level = 1.0, 4.0, 0.15, (24 if confined else 0), 4, 1
    tube = sq.Tube(R, L, a, nth, nr) if confined else None
    model = Model(sq, tube, a, level)
    p = np.array([0.09, -0.05, 0.07, 0.35, 0.2, -1.5] if confined else [0.0, 0.0, 0.0, 0.5, -0.2, 1.0])
    assert model.no_level_change(p[:3])
    _, jac = model.value_and_jacobian(p)
    value = reference_value_fn(kind, R, L, a, nth, nr, level)
    fd = np.zeros((2, 6))
    for k in range(6):
        dp = np.zeros(6)
        dp[k] = H
        fd[:, k] = (value(p + dp) - value(p - dp)) / (2 * H)
    print(kind, "adjoint\n", jac, "\nfinite differences\n", fd, flush=True)
    for i in range(2):
        scale = np.max(np.abs(fd[i]))
        assert np.max(np.abs(jac[i] - fd[i])) <= RTOL * scale + (0 if confined else 1e-10), (i, jac[i], fd[i])
