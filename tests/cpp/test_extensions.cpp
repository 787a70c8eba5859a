// test_extensions.cpp -- the C++17 header API of the additions beyond the reference
// (bem::batch_solve, SwimmerSolver::gradient), on the GPU through libsqbem. Prints one
// PASS/FAIL line per check and exits nonzero on any failure.
//
//  1. batch_solve of three configurations (two sharing one tube mesh object, one on an
//     identical copy) equals SwimmerSolver::solve_free on the same meshes, FP64, to 1e-12;
//  2. the same batch in FP16 MFMA + FP64 refinement agrees with FP64 to 1e-8;
//  3. the adjoint's dJ/dB2 equals J of the B2-only solution (J is linear in B1, B2), and
//     dJ/dB1 equals J of the B1-only solution, to 1e-10.
#include <cmath>
#include <cstdio>
#include <vector>

#include "swimmer.hpp"

static int failures = 0;
static void check(bool ok, const char* what, double v) {
    std::printf("%s  %-58s %.3e\n", ok ? "PASS" : "FAIL", what, v);
    if (!ok) ++failures;
}
static double rel(const bem::V3& a, const bem::V3& b, const bem::V3& c, const bem::V3& d) {
    const double num = bem::norm(a - c) * bem::norm(a - c) + bem::norm(b - d) * bem::norm(b - d);
    const double den = bem::norm(c) * bem::norm(c) + bem::norm(d) * bem::norm(d);
    return std::sqrt(num / den);
}

int main() {
    const double R = 1.0, L = 4.0;
    const int nth = 12, nr = 3;
    bem::Mesh tube = bem::swimmer_tube_mesh(R, L, 0.3, nth, nr);
    bem::Mesh tube_copy = bem::swimmer_tube_mesh(R, L, 0.3, nth, nr);
    struct Cfg {
        double a;
        int level;
        bem::V3 xc, e;
        double B2;
        const bem::Mesh* tube;
    };
    const std::vector<Cfg> cfgs = {
        {0.3, 1, {0.1, 0.0, 0.0}, {0.0, 0.0, 1.0}, -1.0, &tube},
        {0.35, 2, {0.0, 0.15, 0.1}, bem::unit(bem::V3{0.3, 0.2, 0.9}), 2.0, &tube},
        {0.25, 1, {0.0, 0.0, -0.2}, bem::unit(bem::V3{1.0, 0.0, 1.0}), 0.5, &tube_copy},
    };
    std::vector<bem::Mesh> spheres;
    for (const Cfg& c : cfgs) spheres.push_back(bem::icosphere(c.a, c.level, c.xc));
    std::vector<bem::BatchConfig> bc;
    for (size_t i = 0; i < cfgs.size(); ++i)
        bc.push_back({cfgs[i].tube, &spheres[i], cfgs[i].xc, cfgs[i].e, 1.0, cfgs[i].B2});

    const bem::BatchResult r64 = bem::batch_solve(bc, 1.0, SQBEM_FP64);
    check(r64.stats[4] == 1.0, "batch: identical tube meshes share one factorisation", r64.stats[4]);
    const bem::Tube T(R, L, 0.3, nth, nr);
    double worst = 0;
    for (size_t i = 0; i < cfgs.size(); ++i) {
        bem::SwimmerSolver s(&T, spheres[i], cfgs[i].xc);
        const auto sol = s.solve_free({s.slip(cfgs[i].e, 1.0, cfgs[i].B2)})[0];
        worst = std::max(worst, rel(r64.U[i], r64.Om[i], sol.U, sol.Om));
    }
    check(worst <= 1e-12, "batch_solve (FP64) = SwimmerSolver::solve_free", worst);

    const bem::BatchResult r16 = bem::batch_solve(bc, 1.0, SQBEM_FP16_IR);
    double w16 = 0, res = 0;
    for (size_t i = 0; i < cfgs.size(); ++i) {
        w16 = std::max(w16, rel(r16.U[i], r16.Om[i], r64.U[i], r64.Om[i]));
        res = std::max(res, r16.resid[i]);
    }
    check(w16 <= 1e-8, "batch_solve (FP16 + FP64 refinement) vs FP64", w16);
    check(res <= 1e-12 && res > 0.0, "batch refinement residual", res);

    // adjoint: J = g . (U, Omega) is linear in (B1, B2)
    bem::SwimmerSolver s(&T, spheres[1], cfgs[1].xc);
    const std::vector<std::array<double, 6>> g = {{0, 0, 1, 0, 0, 0}, {0, 0, 0, 1, 0, 0}};
    const auto gr = s.gradient(cfgs[1].e, 1.0, cfgs[1].B2, g);
    const auto s1 = s.solve_free({s.slip(cfgs[1].e, 1.0, 0.0)})[0];
    const auto s2 = s.solve_free({s.slip(cfgs[1].e, 0.0, 1.0)})[0];
    const double J1[2] = {s1.U.z, s1.Om.x}, J2[2] = {s2.U.z, s2.Om.x};
    double eb = 0;
    for (int j = 0; j < 2; ++j) {
        const double sc = std::max(std::fabs(J1[j]), std::fabs(J2[j]));
        eb = std::max(eb, std::fabs(gr.d[j].dB1 - J1[j]) / sc);
        eb = std::max(eb, std::fabs(gr.d[j].dB2 - J2[j]) / sc);
    }
    check(eb <= 1e-10, "adjoint dJ/dB1, dJ/dB2 = J of the single-mode solutions", eb);

    std::printf("%s\n", failures ? "some checks failed" : "all checks passed");
    return failures ? 1 : 0;
}
