// quadrature.hpp -- triangle rules and Gauss-Legendre nodes used by the panel integrals.
//
// API of src/utils/quadrature.hpp:13-99: Rule, dunavant7, subdivided, rules,
// gauss_legendre01. The rule tables are uploaded once per device by libsqbem; the device
// assembly kernels read them from constant memory.
#pragma once
#include <array>
#include <cmath>
#include <vector>
#include "vec3.hpp"

namespace bem {

struct Rule {
    std::vector<std::array<double, 3>> L;  // barycentric coordinates of the points
    std::vector<double> W;                 // weights, summing to 1
};

// Degree-5 seven-point triangle rule (Dunavant 1985). The constants are the 15-digit
// values the reference uses (src/utils/quadrature.hpp:23-24), kept identical so that both
// codes integrate with the same points.
inline Rule dunavant7() {
    constexpr double a1 = 0.059715871789770, b1 = 0.470142064105115, w1 = 0.132394152788506;
    constexpr double a2 = 0.797426985353087, b2 = 0.101286507323456, w2 = 0.125939180544827;
    constexpr double c = 1. / 3;
    Rule r;
    r.L = {{c, c, c}, {a1, b1, b1}, {b1, a1, b1}, {b1, b1, a1}, {a2, b2, b2}, {b2, a2, b2}, {b2, b2, a2}};
    r.W = {0.225, w1, w1, w1, w2, w2, w2};
    return r;
}

// The seven-point rule applied on each of the 4^level congruent sub-triangles obtained by
// repeated midpoint splitting (corner, corner, corner, centre order).
inline Rule subdivided(int level) {
    using Bary = std::array<double, 3>;
    using Tri = std::array<Bary, 3>;
    auto half = [](const Bary& p, const Bary& q) {
        return Bary{(p[0] + q[0]) / 2, (p[1] + q[1]) / 2, (p[2] + q[2]) / 2};
    };
    std::vector<Tri> sub{Tri{Bary{1, 0, 0}, Bary{0, 1, 0}, Bary{0, 0, 1}}};
    for (int l = 0; l < level; ++l) {
        std::vector<Tri> next;
        next.reserve(4 * sub.size());
        for (const Tri& t : sub) {
            const Bary ab = half(t[0], t[1]), bc = half(t[1], t[2]), ca = half(t[2], t[0]);
            next.push_back({t[0], ab, ca});
            next.push_back({ab, t[1], bc});
            next.push_back({ca, bc, t[2]});
            next.push_back({ab, bc, ca});
        }
        sub.swap(next);
    }
    const Rule base = dunavant7();
    Rule r;
    for (const Tri& t : sub)
        for (size_t q = 0; q < base.W.size(); ++q) {
            Bary l{0, 0, 0};
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) l[c] += base.L[q][k] * t[k][c];
            r.L.push_back(l);
            r.W.push_back(base.W[q] / sub.size());
        }
    return r;
}

// rules()[0..3] = 7, 28, 112, 448 points
inline const std::vector<Rule>& rules() {
    static const std::vector<Rule> table = {dunavant7(), subdivided(1), subdivided(2), subdivided(3)};
    return table;
}

// n-point Gauss-Legendre rule mapped to [0, 1]: Newton iteration on P_n from the
// cosine estimate of each root; nodes x = (1 - z)/2 (descending z).
inline void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(PI * (i + 0.75) / (n + 0.5));
        double dp = 0;  // P_n'(z)
        for (int it = 0; it < 100; ++it) {
            double pn = 1, pm = 0;  // P_j, P_{j-1}
            for (int j = 1; j <= n; ++j) {
                const double pmm = pm;
                pm = pn;
                pn = ((2 * j - 1) * z * pm - (j - 1) * pmm) / j;
            }
            dp = n * (z * pn - pm) / (z * z - 1);
            const double step = pn / dp;
            z -= step;
            if (std::fabs(step) < 1e-15) break;
        }
        x[i] = 0.5 * (1 - z);
        w[i] = 1.0 / ((1 - z * z) * dp * dp);
    }
}

}  // namespace bem
