// swimmer.hpp -- a spherical squirmer in a closed rigid tube (or in unbounded fluid),
// solved on the GPU by libsqbem.
//
// Boundary conditions: tube wall and caps u = 0; sphere u = U + Omega x (x - xc) + u_s(x);
// unknowns: tractions everywhere, U, Omega; closure: force-free and torque-free.
// Linear system (t = tube, s = sphere), T the fixed tube-tube block:
//   [ T  B  Ct ] [f_t ]   [r_t]
//   [ D  E  Cs ] [f_s ] = [r_s]
//   [ 0  F  0  ] [UOm ]   [ 0 ]
// T is assembled, deflated and LU-factorised once in HBM (Tube); the swimmer is eliminated
// through the Schur complement G = [E | Cs] - D T^-1 [B | Ct] of size 3 Ns + 6, formed
// with MFMA GEMMs against the resident tube factors.
//
// API of src/swimmer.hpp:33-289 (Tube, SwimmerSolution, SwimmerSolver). Extension: an
// optional `precision` (SQBEM_FP32_IR, SQBEM_FP16_IR, SQBEM_BF16_IR) factors in low
// precision and recovers FP64 accuracy by iterative refinement against the FP64 operator.
#pragma once
#include <array>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <vector>

#include "sqbem.h"
#include "utils/field.hpp"
#include "utils/io.hpp"
#include "utils/kernels.hpp"
#include "utils/linalg.hpp"
#include "utils/mesh.hpp"

namespace bem {

// ---------------------------------------------------------------- tube
struct Tube {
    Mesh mesh;
    int Nt = 0;
    double R = 1, L = 4, mu = 1;
    LUFactor lu;  // view onto the factors held in HBM
    int precision = SQBEM_FP64;
    double t_assembly = 0, t_factor = 0;  // seconds, filled by times()

    Tube() = default;
    Tube(double R_, double L_, double a, int n_theta, int n_r = 4, double mu_ = 1, int precision_ = SQBEM_FP64)
        : R(R_), L(L_), mu(mu_), precision(precision_) {
        std::shared_ptr<sqbem_mesh> m(sqbem_mesh_swimmer_tube(R, L, a, n_theta, n_r), sqbem_mesh_free);
        if (!m) throw std::runtime_error(sqbem_last_error());
        mesh = detail::from_sqbem(m.get());
        Nt = mesh.P();
        sqbem_tube* t = sqbem_tube_create(m.get(), mu, precision);
        if (!t) throw std::runtime_error(sqbem_last_error());
        handle_.reset(t, sqbem_tube_free);
        // the device may still be factoring: the factors are waited for where first used
        lu.attach(sqbem_tube_lu(t));
    }
    // assembly and factorisation seconds (waits for the factorisation to finish)
    void times() { sqbem_check(sqbem_tube_info(handle_.get(), nullptr, &t_assembly, &t_factor)); }

    // Wielandt deflation of the null mode f = c m of panels [p0, p0 + np) of M, applied to
    // the host matrix A at row/column offset off: A += u v^T with u = m / (8 pi mu sqrt(At))
    // and v = m A_p (the device solver adds the same term inside its assembly)
    static void deflate(const Mesh& M, int p0, int np, Mat& A, int off, double mu) {
        double At = 0;
        for (int p = p0; p < p0 + np; ++p) At += M.area[p];
        const double scale = 1.0 / (EIGHT_PI * mu * std::sqrt(At));
        std::vector<double> u(3 * (size_t)np), v(3 * (size_t)np);
        for (int p = 0; p < np; ++p)
            for (int c = 0; c < 3; ++c) {
                u[3 * p + c] = scale * M.normal[p0 + p][c];
                v[3 * p + c] = M.normal[p0 + p][c] * M.area[p0 + p];
            }
        for (int j = 0; j < 3 * np; ++j) {
            double* col = &A(off, off + j);
            for (int i = 0; i < 3 * np; ++i) col[i] += u[i] * v[j];
        }
    }

    const sqbem_tube* handle() const { return handle_.get(); }

private:
    std::shared_ptr<sqbem_tube> handle_;
};

// ---------------------------------------------------------------- sphere template
// An addition: the translation-invariant sphere-sphere blocks of a sphere mesh centred at
// `center`, assembled once on the device (sqbem_sphere_create). A SwimmerSolver built from
// it with a shift solves the same sphere translated by that shift without reassembling
// them (a trajectory moves one sphere through a fixed tube).
class SphereBlocks {
public:
    SphereBlocks(const Mesh& sphere, V3 center, double mu = 1.0) : S_(sphere), c_(center), mu_(mu) {
        auto m = detail::to_sqbem(S_);
        const double c[3] = {c_.x, c_.y, c_.z};
        sqbem_sphere* h = sqbem_sphere_create(m.get(), c, mu_);
        if (!h) throw std::runtime_error(sqbem_last_error());
        h_.reset(h, sqbem_sphere_free);
    }
    const Mesh& mesh() const { return S_; }
    V3 center() const { return c_; }
    double mu() const { return mu_; }
    const sqbem_sphere* handle() const { return h_.get(); }
    std::shared_ptr<sqbem_sphere> share() const { return h_; }

private:
    Mesh S_;
    V3 c_;
    double mu_;
    std::shared_ptr<sqbem_sphere> h_;
};

// ---------------------------------------------------------------- swimmer
struct SwimmerSolution {
    V3 U, Om;
    std::vector<V3> f_sphere, f_tube;  // tractions (f = sigma.m, m out of the fluid)
};

class SwimmerSolver {
public:
    const Tube* tube;  // nullptr -> unbounded fluid
    Mesh S;            // sphere mesh
    V3 xc;
    double mu;
    int Ns, Nt;
    double t_assembly = 0, t_schur = 0;
    int ir_iters = 0;          // refinement steps of the last solve (refined precisions)
    double ir_residual = 0.0;  // its final relative residual

    SwimmerSolver(const Tube* t, const Mesh& sphere, V3 center, double mu_ = 1.0, int precision = -1)
        : tube(t), S(sphere), xc(center), mu(mu_) {
        Ns = S.P();
        Nt = tube ? tube->Nt : 0;
        const int prec = precision >= 0 ? precision : (tube ? tube->precision : SQBEM_FP64);
        auto m = detail::to_sqbem(S);
        const double c[3] = {xc.x, xc.y, xc.z};
        sqbem_swimmer* s = sqbem_swimmer_create(tube ? tube->handle() : nullptr, m.get(), c, mu, prec);
        if (!s) throw std::runtime_error(sqbem_last_error());
        h_.reset(s, sqbem_swimmer_free);
        times();
    }

    // the sphere of `blocks` translated by `shift` (an addition; equal to the constructor
    // above on the translated mesh, to rounding)
    SwimmerSolver(const Tube* t, const SphereBlocks& blocks, V3 shift, int precision = -1)
        : tube(t), S(blocks.mesh()), xc(blocks.center() + shift), mu(blocks.mu()) {
        S.translate(shift);
        Ns = S.P();
        Nt = tube ? tube->Nt : 0;
        const int prec = precision >= 0 ? precision : (tube ? tube->precision : SQBEM_FP64);
        const double d[3] = {shift.x, shift.y, shift.z};
        sqbem_swimmer* s = sqbem_swimmer_create_moved(tube ? tube->handle() : nullptr, blocks.handle(), d, prec);
        if (!s) throw std::runtime_error(sqbem_last_error());
        h_.reset(s, sqbem_swimmer_free);
        keep_ = blocks.share();
        times();
    }

    // squirming slip on every sphere panel for orientation e and modes B1, B2
    std::vector<V3> slip(V3 e, double B1, double B2) const {
        std::vector<V3> u(Ns);
        for (int p = 0; p < Ns; ++p) u[p] = squirmer_slip(S.centroid[p], xc, e, B1, B2);
        return u;
    }

    // force- and torque-free swimming, one solution per slip field
    std::vector<SwimmerSolution> solve_free(const std::vector<std::vector<V3>>& slips) {
        const int k = (int)slips.size();
        std::vector<SwimmerSolution> out(k);
        for (int j0 = 0; j0 < k; j0 += 8) {  // the device solve takes up to 8 fields at once
            const int kk = std::min(8, k - j0);
            std::vector<double> u((size_t)kk * 3 * Ns), U(3 * kk), Om(3 * kk), fs((size_t)kk * 3 * Ns),
                ft((size_t)kk * 3 * std::max(Nt, 1));
            for (int j = 0; j < kk; ++j)
                for (int p = 0; p < Ns; ++p)
                    for (int c = 0; c < 3; ++c) u[(size_t)j * 3 * Ns + 3 * p + c] = slips[j0 + j][p][c];
            sqbem_check(sqbem_swimmer_solve_free(h_.get(), kk, u.data(), U.data(), Om.data(), fs.data(),
                                                 Nt ? ft.data() : nullptr));
            for (int j = 0; j < kk; ++j) out[j0 + j] = unpack(U.data() + 3 * j, Om.data() + 3 * j, fs, ft, j);
        }
        times();
        return out;
    }

    // rigid sphere with prescribed (U, Omega) and no slip ("towed" problem)
    SwimmerSolution solve_prescribed(V3 U, V3 Om) {
        std::vector<double> fs(3 * (size_t)Ns), ft(3 * (size_t)std::max(Nt, 1));
        const double u[3] = {U.x, U.y, U.z}, w[3] = {Om.x, Om.y, Om.z};
        sqbem_check(sqbem_swimmer_solve_prescribed(h_.get(), u, w, fs.data(), Nt ? ft.data() : nullptr));
        times();
        return unpack(u, w, fs, ft, 0);
    }

    // hydrodynamic force ON the sphere (its normals point into it): F = -sum_p A_p f_p
    V3 force(const SwimmerSolution& s) const {
        double F[3] = {0, 0, 0};
        for (int p = 0; p < Ns; ++p)
            for (int c = 0; c < 3; ++c) F[c] -= S.area[p] * s.f_sphere[p][c];
        return V3{F[0], F[1], F[2]};
    }

    // velocity at points in the fluid (lab frame), from the boundary data on the merged
    // surface: tube panels first (zero velocity, tube tractions), then the sphere's panels
    // (rigid motion plus slip, sphere tractions)
    std::vector<V3> velocity(const SwimmerSolution& s, const std::vector<V3>& slip_field,
                             const std::vector<V3>& pts) const {
        std::vector<V3> f(s.f_tube.begin(), s.f_tube.begin() + Nt);
        f.insert(f.end(), s.f_sphere.begin(), s.f_sphere.end());
        std::vector<V3> u(Nt, V3{});
        u.reserve(Nt + Ns);
        for (int p = 0; p < Ns; ++p) u.push_back(s.U + cross(s.Om, S.centroid[p] - xc) + slip_field[p]);
        return velocity_at(tube ? merge(tube->mesh, S) : S, u, f, mu, pts);
    }

    // Adjoint gradients (an addition; the reference has no derivatives). For each weight
    // vector g_j on (U, Omega): J_j = g_j . (U, Omega) of the free-swimming solution with
    // slip(e, B1, B2), and dJ_j/d(x_c, e, B1, B2). FP64 swimmers.
    struct Gradient {
        V3 dxc, de;
        double dB1 = 0, dB2 = 0;
    };
    struct GradientResult {
        V3 U, Om;
        std::vector<Gradient> d;
    };
    GradientResult gradient(V3 e, double B1, double B2, const std::vector<std::array<double, 6>>& g) {
        const int k = (int)g.size();
        std::vector<double> gw(6 * (size_t)k), dxc(3 * (size_t)k), de(3 * (size_t)k), dB(2 * (size_t)k);
        for (int j = 0; j < k; ++j)
            for (int c = 0; c < 6; ++c) gw[6 * (size_t)j + c] = g[j][c];
        const double ev[3] = {e.x, e.y, e.z};
        double U[3], Om[3];
        sqbem_check(sqbem_swimmer_gradient(h_.get(), ev, B1, B2, k, gw.data(), U, Om, dxc.data(), de.data(),
                                           dB.data(), nullptr));
        GradientResult r;
        r.U = V3{U[0], U[1], U[2]};
        r.Om = V3{Om[0], Om[1], Om[2]};
        r.d.resize(k);
        for (int j = 0; j < k; ++j) {
            r.d[j].dxc = V3{dxc[3 * j], dxc[3 * j + 1], dxc[3 * j + 2]};
            r.d[j].de = V3{de[3 * j], de[3 * j + 1], de[3 * j + 2]};
            r.d[j].dB1 = dB[2 * j];
            r.d[j].dB2 = dB[2 * j + 1];
        }
        return r;
    }

    const sqbem_swimmer* handle() const { return h_.get(); }

private:
    std::shared_ptr<sqbem_sphere> keep_;  // a moved solver's template (declared first: outlives h_)
    std::shared_ptr<sqbem_swimmer> h_;

    void times() { sqbem_check(sqbem_swimmer_times(h_.get(), &t_assembly, &t_schur, &ir_iters, &ir_residual)); }
    SwimmerSolution unpack(const double* U, const double* Om, const std::vector<double>& fs,
                           const std::vector<double>& ft, int j) const {
        SwimmerSolution s;
        s.U = V3{U[0], U[1], U[2]};
        s.Om = V3{Om[0], Om[1], Om[2]};
        s.f_sphere.resize(Ns);
        for (int p = 0; p < Ns; ++p) {
            const double* q = fs.data() + (size_t)j * 3 * Ns + 3 * p;
            s.f_sphere[p] = V3{q[0], q[1], q[2]};
        }
        if (Nt) {
            s.f_tube.resize(Nt);
            for (int p = 0; p < Nt; ++p) {
                const double* q = ft.data() + (size_t)j * 3 * Nt + 3 * p;
                s.f_tube[p] = V3{q[0], q[1], q[2]};
            }
        }
        return s;
    }
};

// ---------------------------------------------------------------- batched solves
// Free-swimming U, Omega of many independent configurations in one call (an addition: the
// reference loops over SwimmerSolver one configuration at a time, src/main.cpp:82-100).
// Configuration i: tube mesh, sphere mesh placed at its centre, squirming direction e and
// modes B1, B2 (slip squirmer_slip at the sphere centroids). Identical tube meshes share
// one factorisation; the device pipeline is described at sqbem_batch_solve.
struct BatchConfig {
    const Mesh* tube = nullptr;
    const Mesh* sphere = nullptr;
    V3 center, e{0, 0, 1};
    double B1 = 1, B2 = 0;
};
struct BatchResult {
    std::vector<V3> U, Om;
    std::vector<int> iters;       // refinement steps per configuration (0 in FP64 mode)
    std::vector<double> resid;    // final relative residual of the full system
    std::array<double, 12> stats; // see sqbem_batch_solve
};
inline BatchResult batch_solve(const std::vector<BatchConfig>& cfg, double mu = 1, int precision = SQBEM_FP64,
                               double mem_budget_gb = 0) {
    const int n = (int)cfg.size();
    // one handle per distinct mesh object (identical contents are also shared by the library)
    std::vector<std::shared_ptr<sqbem_mesh>> keep;
    std::vector<const Mesh*> seen;
    std::vector<sqbem_mesh*> handle_of;
    auto handle = [&](const Mesh* M) -> const sqbem_mesh* {
        if (!M) throw std::runtime_error("batch_solve: every configuration needs a tube and a sphere");
        for (size_t j = 0; j < seen.size(); ++j)
            if (seen[j] == M) return handle_of[j];
        keep.push_back(detail::to_sqbem(*M));
        seen.push_back(M);
        handle_of.push_back(keep.back().get());
        return keep.back().get();
    };
    std::vector<const sqbem_mesh*> tubes(n), spheres(n);
    std::vector<double> c(3 * (size_t)n), e(3 * (size_t)n), B(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        tubes[i] = handle(cfg[i].tube);
        spheres[i] = handle(cfg[i].sphere);
        c[3 * i] = cfg[i].center.x; c[3 * i + 1] = cfg[i].center.y; c[3 * i + 2] = cfg[i].center.z;
        e[3 * i] = cfg[i].e.x; e[3 * i + 1] = cfg[i].e.y; e[3 * i + 2] = cfg[i].e.z;
        B[2 * i] = cfg[i].B1;
        B[2 * i + 1] = cfg[i].B2;
    }
    std::vector<double> U(3 * (size_t)n), Om(3 * (size_t)n);
    BatchResult r;
    r.iters.resize(n);
    r.resid.resize(n);
    sqbem_check(sqbem_batch_solve(n, tubes.data(), spheres.data(), c.data(), e.data(), B.data(), mu, precision,
                                  mem_budget_gb, U.data(), Om.data(), r.iters.data(), r.resid.data(),
                                  r.stats.data()));
    r.U.resize(n);
    r.Om.resize(n);
    for (int i = 0; i < n; ++i) {
        r.U[i] = V3{U[3 * i], U[3 * i + 1], U[3 * i + 2]};
        r.Om[i] = V3{Om[3 * i], Om[3 * i + 1], Om[3 * i + 2]};
    }
    return r;
}

}  // namespace bem
