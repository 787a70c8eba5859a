// test_extensions.cpp -- the C++17 header API of the additions beyond the reference
// (bem::batch_solve, SwimmerSolver::gradient), on the GPU through libsqbem. Prints one
// PASS/FAIL line per check and exits nonzero on any failure.
//
//  1. batch_solve of three configurations (two sharing one tube mesh object, one on an
//     identical copy) equals SwimmerSolver::solve_free on the same meshes, FP64, to 1e-12;
//  2. the same batch in FP16 MFMA + FP64 refinement agrees with FP64 to 1e-8;
//  3. the adjoint's dJ/dB2 equals J of the B2-only solution (J is linear in B1, B2), and
//     dJ/dB1 equals J of the B1-only solution, to 1e-10;
//  4. LUFactor::factor leaves the packed factors and 1-based pivots in LU and piv, as
//     dgetrf_ does (P A = L U to 1e-13 on a matrix that needs row interchanges); a Tube's
//     lu is a 0 x 0 Mat until download();
//  5. Tube::lu.solve of an FP32-factored tube returns the FP64 solution (refined against
//     the FP64 operator) to 1e-10 of the FP64 tube's.
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

    // LUFactor fills LU and piv (src/utils/linalg.hpp:62-71)
    {
        const int n = 7;
        bem::Mat A(n, n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) A(i, j) = std::sin(1.0 + i * 0.7 + j * j * 0.3) + (i == n - 1 - j ? 0.5 : 0.0);
        A(0, 0) = 1e-3;  // forces an interchange in the first column
        bem::LUFactor lu;
        lu.factor(A);
        bool shape = lu.LU.n == n && lu.LU.m == n && lu.LU.a.size() == (size_t)n * n && lu.piv.size() == (size_t)n;
        double err = shape ? 0.0 : 1.0;
        if (shape) {
            bem::Mat PA = A;  // apply the interchanges in order, as dlaswp does
            for (int i = 0; i < n; ++i) {
                const int p = lu.piv[i] - 1;
                if (p < 0 || p >= n) { err = 1.0; break; }
                for (int j = 0; j < n; ++j) std::swap(PA(i, j), PA(p, j));
            }
            double amax = 0;
            for (double v : A.a) amax = std::max(amax, std::fabs(v));
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    double s = 0;
                    for (int k = 0; k <= std::min(i, j); ++k) s += (k == i ? 1.0 : lu.LU(i, k)) * lu.LU(k, j);
                    err = std::max(err, std::fabs(s - PA(i, j)) / amax);
                }
        }
        check(shape && lu.piv[0] != 1, "LUFactor::factor fills LU (n x n) and piv (interchange)", shape ? lu.piv[0] : -1);
        check(err <= 1e-13, "LUFactor: P A = L U from the returned LU and piv", err);
        check(T.lu.LU.n == 0 && T.lu.LU.m == 0 && T.lu.LU.a.empty() && T.lu.piv.empty(),
              "Tube::lu is an empty 0 x 0 Mat until download()", (double)T.lu.LU.a.size());
    }
    // Tube::lu.solve of a refined-mode tube gives the FP64 solution
    {
        const bem::Tube T32(R, L, 0.3, nth, nr, 1.0, SQBEM_FP32_IR);
        const int n = 3 * T.Nt;
        bem::Mat b64(n, 2);
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < n; ++i) b64(i, j) = std::cos(0.37 * i + 1.3 * j);
        bem::Mat b32 = b64;
        T.lu.solve(b64);
        T32.lu.solve(b32);
        double num = 0, den = 0;
        for (size_t e = 0; e < b64.a.size(); ++e) {
            num += (b32.a[e] - b64.a[e]) * (b32.a[e] - b64.a[e]);
            den += b64.a[e] * b64.a[e];
        }
        const double d = std::sqrt(num / den);
        check(d <= 1e-10, "Tube::lu.solve (FP32 factors, refined) = FP64 tube", d);
    }

    std::printf("%s\n", failures ? "some checks failed" : "all checks passed");
    return failures ? 1 : 0;
}
