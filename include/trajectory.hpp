// trajectory.hpp -- time integration of a squirmer's path in a tube:
//
//   dX/dt = U(X, e),   de/dt = Omega(X, e) x e
//
// Moving frame: the tube is centred on the swimmer's axial position, so its factors are
// built once and stay resident in HBM; each evaluation assembles only the sphere-dependent
// couplings and Schur complement on the device. Linearity in the squirming modes:
// U = U[B1=1] + alpha U[B2=1], both modes solved together. Three RK4 steps start
// Adams-Bashforth 4; the state is checkpointed every step for restarts.
//
// API, CSV schema and checkpoint format of src/trajectory.hpp:23-151.
#pragma once
#include <array>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "swimmer.hpp"
#include "utils/io.hpp"

namespace bem {

struct TrajectoryParams {
    double alpha = 0;         // B2/B1 (pusher < 0 < puller)
    double a_over_R = 0.3;    // confinement
    double beta0 = 0.5;       // initial offset from the axis, in units of (R - a)
    double pitch0_deg = 0;    // initial tilt toward +x (radial direction)
    double yaw0_deg = 0;      // initial tilt toward +y (azimuthal direction)
    double dt = 0.5;          // time step, units a/B1
    double t_max = 100;
    double L_over_R = 4;      // truncated tube length
    int n_theta = 30;         // tube panels around the circumference
    int sphere_level = 2;     // icosphere refinement (2 -> 320 panels)
    double beta_stop = 0.95;  // stop when the swimmer gets this close to the wall
    double budget_s = 0;      // wall-clock budget per call (0 = unlimited)
    int n_r = 4;              // rings per tube cap (the reference's Tube default)
    int precision = SQBEM_FP64;
};

namespace detail {

using PathState = std::array<double, 6>;  // centre (x, y, z) and orientation (ex, ey, ez)

// The right-hand side of the path equations, dX/dt = U, de/dt = Omega x e, at one state. In
// the moving frame the sphere sits at (x, y, 0): the tube (assembled and factored once) and
// the sphere template (its translation-invariant blocks, assembled once) stay on the device,
// and each evaluation builds only the tube-sphere couplings and the Schur complement of the
// translated sphere. Both squirming modes are solved together (U = U[B1=1] + alpha U[B2=1]).
class PathRhs {
public:
    struct Value {
        PathState rate{};
        V3 U, Om;
    };

    explicit PathRhs(const TrajectoryParams& P)
        : a_(1.0),
          R_(a_ / P.a_over_R),
          alpha_(P.alpha),
          tube_(R_, P.L_over_R * R_, a_, P.n_theta, P.n_r, 1.0, P.precision),
          sphere_(icosphere(a_, P.sphere_level), V3{0, 0, 0}, 1.0) {}

    // the solver at state y (kept by the caller when the adjoint needs it)
    std::unique_ptr<SwimmerSolver> solver(const PathState& y) const {
        return std::make_unique<SwimmerSolver>(&tube_, sphere_, V3{y[0], y[1], 0.0});
    }
    Value operator()(const PathState& y, std::unique_ptr<SwimmerSolver>* keep = nullptr) {
        auto s = solver(y);
        const V3 e = unit(V3{y[3], y[4], y[5]});
        const auto modes = s->solve_free({s->slip(e, 1, 0), s->slip(e, 0, 1)});
        Value v;
        v.U = modes[0].U + alpha_ * modes[1].U;
        v.Om = modes[0].Om + alpha_ * modes[1].Om;
        const V3 turn = cross(v.Om, e);
        v.rate = {v.U.x, v.U.y, v.U.z, turn.x, turn.y, turn.z};
        ++calls_;
        if (keep) *keep = std::move(s);
        return v;
    }
    // radial offset in units of the free gap R - a
    double beta(const PathState& y) const { return std::hypot(y[0], y[1]) / (R_ - a_); }
    double gap() const { return R_ - a_; }
    int calls() const { return calls_; }

private:
    double a_, R_, alpha_;
    Tube tube_;
    SphereBlocks sphere_;
    int calls_ = 0;
};

// state at t = 0: offset beta0 (R - a) toward -x, e tilted by pitch toward +x and yaw toward +y
inline PathState initial_state(const TrajectoryParams& P, double gap) {
    const double b = P.beta0 * gap, pitch = P.pitch0_deg * PI / 180, yaw = P.yaw0_deg * PI / 180;
    return {-b, 0.0, 0.0, std::cos(yaw) * std::sin(pitch), std::sin(yaw), std::cos(yaw) * std::cos(pitch)};
}

// y + h d
inline PathState shifted(const PathState& y, double h, const PathState& d) {
    PathState out;
    for (int i = 0; i < 6; ++i) out[i] = y[i] + h * d[i];
    return out;
}

// one step from y with f = rate(y): classical RK4 while fewer than three earlier rates are
// known (the start-up), then Adams-Bashforth 4 over f and the three earlier rates (history,
// oldest first), coefficients (55, -59, 37, -9) / 24. The orientation is not renormalised here.
template <typename Rate>
PathState integrate_step(const PathState& y, const PathState& f, const std::vector<PathState>& history, double dt,
                         Rate&& rate) {
    PathState z = y;
    if (history.size() < 3) {
        const PathState k2 = rate(shifted(y, dt / 2, f));
        const PathState k3 = rate(shifted(y, dt / 2, k2));
        const PathState k4 = rate(shifted(y, dt, k3));
        for (int i = 0; i < 6; ++i) z[i] += dt / 6 * (f[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
    } else {
        const PathState &f1 = history[2], &f2 = history[1], &f3 = history[0];
        for (int i = 0; i < 6; ++i) z[i] += dt / 24 * (55 * f[i] - 59 * f1[i] + 37 * f2[i] - 9 * f3[i]);
    }
    return z;
}

// keep |e| = 1
inline void renormalise(PathState& y) {
    const double n = std::sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    for (int i = 3; i < 6; ++i) y[i] /= n;
}

// Restart file of run_trajectory (src/trajectory.hpp:69-104): "t step n" on the first line,
// the state on the second, then the n (<= 3) most recent rates, oldest first, each value
// with 17 significant digits.
struct PathCheckpoint {
    double t = 0;
    int step = 0;
    PathState y{};
    std::vector<PathState> history;

    bool read(const std::string& path) {
        std::ifstream in(path);
        if (!in) return false;
        size_t n = 0;
        in >> t >> step >> n;
        for (double& v : y) in >> v;
        history.assign(n, PathState{});
        for (PathState& h : history)
            for (double& v : h) in >> v;
        return true;
    }
    void write(const std::string& path) const {
        std::FILE* f = std::fopen(path.c_str(), "w");
        if (!f) return;
        std::fprintf(f, "%.17g %d %zu\n", t, step, history.size());
        for (double v : y) std::fprintf(f, "%.17g ", v);
        std::fprintf(f, "\n");
        for (const PathState& h : history) {
            for (double v : h) std::fprintf(f, "%.17g ", v);
            std::fprintf(f, "\n");
        }
        std::fclose(f);
    }
};

}  // namespace detail

// Runs (or resumes) a trajectory: writes <out_prefix>.csv, keeps <out_prefix>.ckpt while
// paused. Returns "finished", "wall contact" or "paused". Behaviour of src/trajectory.hpp:
// 42-151: the CSV row of a step holds the state at its start and the velocity there; the
// run stops before a step from a state with beta > beta_stop, or (budget_s > 0) before a
// step whose projected end, the elapsed time plus 1.3 times the mean time per evaluation
// for each evaluation of that step (4 during the RK4 start-up, else 1), passes the budget.
inline std::string run_trajectory(const TrajectoryParams& P, const std::string& out_prefix, bool verbose = true) {
    using detail::PathState;
    Timer clock;
    detail::PathRhs rhs(P);
    const std::string ckpt_path = out_prefix + ".ckpt";
    detail::PathCheckpoint run;
    const bool resumed = run.read(ckpt_path);
    if (resumed) {
        if (verbose) std::printf("restart at t = %.3f (step %d)\n", run.t, run.step);
    } else {
        run.y = detail::initial_state(P, rhs.gap());
    }
    char params[256];
    std::snprintf(params, sizeof params,
                  "alpha=%g a_over_R=%g beta0=%g pitch0_deg=%g yaw0_deg=%g dt=%g n_theta=%d sphere_level=%d", P.alpha,
                  P.a_over_R, P.beta0, P.pitch0_deg, P.yaw0_deg, P.dt, P.n_theta, P.sphere_level);
    Csv csv(out_prefix + ".csv", "t,x,y,z,ex,ey,ez,Ux,Uy,Uz,Omx,Omy,Omz,beta", {params}, resumed);

    auto rate_of = [&](const PathState& y) { return rhs(y).rate; };
    std::string status = "finished";
    for (; run.t < P.t_max - 1e-9; ++run.step, run.t += P.dt) {
        const double beta = rhs.beta(run.y);
        if (beta > P.beta_stop) {
            status = "wall contact";
            break;
        }
        if (P.budget_s > 0 && rhs.calls() > 0) {
            const double spent = clock.seconds(), per_call = spent / rhs.calls();
            const int calls_next = run.history.size() < 3 ? 4 : 1;
            if (spent + 1.3 * per_call * calls_next > P.budget_s) {
                status = "paused";
                break;
            }
        }
        const auto here = rhs(run.y);
        const PathState& y = run.y;
        csv.row({run.t, y[0], y[1], y[2], y[3], y[4], y[5], here.U.x, here.U.y, here.U.z, here.Om.x, here.Om.y,
                 here.Om.z, beta});
        run.y = detail::integrate_step(run.y, here.rate, run.history, P.dt, rate_of);
        detail::renormalise(run.y);
        run.history.push_back(here.rate);
        if (run.history.size() > 3) run.history.erase(run.history.begin());
        detail::PathCheckpoint next = run;  // the checkpoint after the step
        next.t += P.dt;
        ++next.step;
        next.write(ckpt_path);
    }
    if (verbose) {
        const int n = rhs.calls();
        std::printf("%s: t = %.2f, beta = %.3f, z = %.2f  (%d solves, %.1f s, %.2f s/solve)\n", status.c_str(), run.t,
                    rhs.beta(run.y), run.y[2], n, clock.seconds(), n ? clock.seconds() / n : 0.0);
    }
    if (status != "paused") std::remove(ckpt_path.c_str());
    return status;
}

// ---------------------------------------------------------------- trajectory sensitivities
// An addition with no counterpart in the reference. n_steps steps of the scheme above from
// (beta0, pitch0_deg, yaw0_deg), then the derivatives of beta_final = |(x, y)| / (R - a)
// and tilt_final_deg (angle of e from the tube axis) with respect to beta0 and pitch0_deg by
// the discrete adjoint of the integrator: the forward pass keeps the state every ckpt_every
// steps; backward, each segment is recomputed from its checkpoint with every evaluation's
// solver kept in device memory, and the adjoint state is swept back through AB4 / RK4 /
// renormalisation, one adjoint solve (SwimmerSolver::gradient) per evaluation.
struct TrajectoryGradient {
    double beta_final = 0, tilt_final_deg = 0;
    double dbeta_dbeta0 = 0, dbeta_dpitch0 = 0, dtilt_dbeta0 = 0, dtilt_dpitch0 = 0;
    int evaluations = 0;
};

inline TrajectoryGradient trajectory_gradient(const TrajectoryParams& P, int n_steps, int ckpt_every = 50,
                                              bool adjoint = true, bool verbose = false) {
    Timer clock;
    using State = std::array<double, 6>;
    using Adj = std::array<State, 2>;  // adjoint states of (beta_final, tilt_final_deg)
    if (P.precision != SQBEM_FP64) throw std::runtime_error("trajectory_gradient: FP64 only");
    if (n_steps < 1 || ckpt_every < 1) throw std::runtime_error("trajectory_gradient: bad step counts");
    const double a = 1, R = a / P.a_over_R;
    detail::PathRhs rhs(P);

    struct Eval {
        State s{}, d{};
        V3 U, Om;
        std::unique_ptr<SwimmerSolver> sw;  // kept for the adjoint sweep
    };
    // the evaluation of run_trajectory (same right-hand side)
    auto evaluate = [&](const State& y, bool keep) {
        Eval r;
        r.s = y;
        const auto v = rhs(y, keep ? &r.sw : nullptr);
        r.U = v.U;
        r.Om = v.Om;
        r.d = v.rate;
        return r;
    };
    // abar . df/ds at an evaluation, f(s) = (U, Omega x e), e = s_e / |s_e|, centre (s0, s1, 0):
    // abar_U . U + abar_e . (Omega x e) = abar_U . U + (e x abar_e) . Omega, plus the explicit
    // e-dependence (abar_e x Omega) . de
    auto vjp = [&](Eval& ev, const Adj& ab) {
        const V3 se{ev.s[3], ev.s[4], ev.s[5]};
        const double ne = norm(se);
        const V3 e = unit(se);
        std::vector<std::array<double, 6>> g(2);
        for (int j = 0; j < 2; ++j) {
            const V3 ae{ab[j][3], ab[j][4], ab[j][5]}, w = cross(e, ae);
            g[j] = {ab[j][0], ab[j][1], ab[j][2], w.x, w.y, w.z};
        }
        const auto gr = ev.sw->gradient(e, 1.0, P.alpha, g);
        ev.sw.reset();
        Adj out{};
        for (int j = 0; j < 2; ++j) {
            const V3 ae{ab[j][3], ab[j][4], ab[j][5]};
            const V3 ge = gr.d[j].de + cross(ae, ev.Om);
            const V3 pe = (1.0 / ne) * (ge - dot(e, ge) * e);
            out[j] = {gr.d[j].dxc.x, gr.d[j].dxc.y, 0.0, pe.x, pe.y, pe.z};
        }
        return out;
    };
    struct StepRec {
        Eval f, k2, k3, k4;
        State z{};  // state before the renormalisation of e
        bool rk = false;
    };
    // one step of run_trajectory's loop (same arithmetic); rec keeps the evaluations
    auto step = [&](State& y, std::vector<State>& hist, StepRec* rec) {
        const bool keep = rec != nullptr;
        Eval f = evaluate(y, keep);
        State z{};
        const bool rk = hist.size() < 3;
        if (rk) {
            auto axpy = [](const State& s, double h, const State& d) {
                State o;
                for (int i = 0; i < 6; ++i) o[i] = s[i] + h * d[i];
                return o;
            };
            Eval k2 = evaluate(axpy(y, P.dt / 2, f.d), keep);
            Eval k3 = evaluate(axpy(y, P.dt / 2, k2.d), keep);
            Eval k4 = evaluate(axpy(y, P.dt, k3.d), keep);
            for (int i = 0; i < 6; ++i) z[i] = y[i] + P.dt / 6 * (f.d[i] + 2 * k2.d[i] + 2 * k3.d[i] + k4.d[i]);
            if (rec) {
                rec->k2 = std::move(k2);
                rec->k3 = std::move(k3);
                rec->k4 = std::move(k4);
            }
        } else {
            const State &f1 = hist[2], &f2 = hist[1], &f3 = hist[0];
            for (int i = 0; i < 6; ++i) z[i] = y[i] + P.dt / 24 * (55 * f.d[i] - 59 * f1[i] + 37 * f2[i] - 9 * f3[i]);
        }
        y = z;
        const double n = std::sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
        for (int i = 3; i < 6; ++i) y[i] /= n;
        hist.push_back(f.d);
        if (hist.size() > 3) hist.erase(hist.begin());
        if (rec) {
            rec->z = z;
            rec->rk = rk;
            rec->f = std::move(f);
        }
    };
    auto beta_of = [&](const State& y) { return std::hypot(y[0], y[1]) / (R - a); };

    // ---- forward, keeping checkpoints
    const double b = P.beta0 * (R - a), xi = P.pitch0_deg * PI / 180, chi = P.yaw0_deg * PI / 180;
    State y = {-b, 0, 0, std::cos(chi) * std::sin(xi), std::sin(chi), std::cos(chi) * std::cos(xi)};
    struct Ckpt {
        State y;
        std::vector<State> hist;
    };
    std::vector<Ckpt> ck;
    std::vector<State> hist;
    for (int n = 0; n < n_steps; ++n) {
        if (n % ckpt_every == 0) {
            ck.push_back({y, hist});
            if (verbose && n > 0) std::fprintf(stderr, "forward: step %d of %d, %.1f s\n", n, n_steps, clock.seconds());
        }
        if (beta_of(y) > P.beta_stop) throw std::runtime_error("trajectory_gradient: wall contact before the last step");
        step(y, hist, nullptr);
    }
    TrajectoryGradient G;
    const double hyp = std::hypot(y[0], y[1]), rho = std::hypot(y[3], y[4]), ez = y[5];
    G.beta_final = hyp / (R - a);
    G.tilt_final_deg = std::atan2(rho, ez) * 180 / PI;
    if (!adjoint) {
        G.evaluations = rhs.calls();
        return G;
    }
    if (hyp == 0 || rho == 0) throw std::runtime_error("trajectory_gradient: final state on the axis");
    // ---- adjoint: d/dy_N of beta_final and tilt_final_deg
    Adj yb{};
    yb[0] = {y[0] / (hyp * (R - a)), y[1] / (hyp * (R - a)), 0, 0, 0, 0};
    const double dd = 180 / PI / (rho * rho + ez * ez);
    yb[1] = {0, 0, 0, dd * ez * y[3] / rho, dd * ez * y[4] / rho, -dd * rho};
    std::vector<Adj> Fb(n_steps, Adj{});  // adjoint of f(y_n), the derivative that enters steps n..n+3
    auto add = [](Adj& dst, double c, const Adj& src) {
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 6; ++i) dst[j][i] += c * src[j][i];
    };
    for (int c = (int)ck.size() - 1; c >= 0; --c) {
        const int n0 = c * ckpt_every, n1 = std::min(n_steps, n0 + ckpt_every);
        State ys = ck[c].y;
        std::vector<State> hs = ck[c].hist;
        std::vector<StepRec> rec(n1 - n0);
        for (int n = n0; n < n1; ++n) step(ys, hs, &rec[n - n0]);
        for (int n = n1 - 1; n >= n0; --n) {
            StepRec& r = rec[n - n0];
            // through the renormalisation y_e = z_e / |z_e|
            const V3 ze{r.z[3], r.z[4], r.z[5]};
            const double nz = norm(ze);
            const V3 u = unit(ze);
            Adj zb = yb;
            for (int j = 0; j < 2; ++j) {
                const V3 ye{yb[j][3], yb[j][4], yb[j][5]};
                const V3 q = (1.0 / nz) * (ye - dot(u, ye) * u);
                zb[j][3] = q.x;
                zb[j][4] = q.y;
                zb[j][5] = q.z;
            }
            Adj yn = zb;  // z = y_n + ...
            if (!r.rk) {
                const double h = P.dt / 24;
                add(Fb[n], 55 * h, zb);
                add(Fb[n - 1], -59 * h, zb);
                add(Fb[n - 2], 37 * h, zb);
                add(Fb[n - 3], -9 * h, zb);
                add(yn, 1.0, vjp(r.f, Fb[n]));
            } else {
                Adj k1b = Fb[n], k2b{}, k3b{}, k4b{};
                add(k1b, P.dt / 6, zb);
                add(k2b, P.dt / 3, zb);
                add(k3b, P.dt / 3, zb);
                add(k4b, P.dt / 6, zb);
                const Adj s4 = vjp(r.k4, k4b);  // k4 = f(y + dt k3)
                add(yn, 1.0, s4);
                add(k3b, P.dt, s4);
                const Adj s3 = vjp(r.k3, k3b);  // k3 = f(y + dt/2 k2)
                add(yn, 1.0, s3);
                add(k2b, P.dt / 2, s3);
                const Adj s2 = vjp(r.k2, k2b);  // k2 = f(y + dt/2 k1)
                add(yn, 1.0, s2);
                add(k1b, P.dt / 2, s2);
                add(yn, 1.0, vjp(r.f, k1b));
            }
            yb = yn;
        }
        if (verbose)
            std::fprintf(stderr, "adjoint: segment %d of %zu done, %.1f s\n", (int)ck.size() - c, ck.size(),
                         clock.seconds());
    }
    // y_0 = (-beta0 (R - a), 0, 0, cos(chi) sin(xi), sin(chi), cos(chi) cos(xi))
    const double dxi = PI / 180;
    G.dbeta_dbeta0 = -(R - a) * yb[0][0];
    G.dtilt_dbeta0 = -(R - a) * yb[1][0];
    G.dbeta_dpitch0 = dxi * std::cos(chi) * (std::cos(xi) * yb[0][3] - std::sin(xi) * yb[0][5]);
    G.dtilt_dpitch0 = dxi * std::cos(chi) * (std::cos(xi) * yb[1][3] - std::sin(xi) * yb[1][5]);
    G.evaluations = rhs.calls();
    return G;
}

}  // namespace bem
