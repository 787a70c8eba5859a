// duct_flow.hpp -- pressure-driven Stokes flow in a finite rigid tube, solved on the GPU.
//
// Tube 0 < z < L, radius R; wall u = 0; inlet u_x = u_y = 0, f_z = +p_in; outlet
// u_x = u_y = 0, f_z = -p_out. Unknowns: wall tractions; on the caps u_z, f_x, f_y.
// Exact answer (Hagen-Poiseuille): u_z(r) = G (R^2 - r^2)/(4 mu), G = dP/L, Q = pi R^4 G/(8 mu).
// The mixed-BC matrix is assembled and LU-factorised once in HBM; every pressure drop is a
// new right-hand side.
//
// API of src/duct_flow.hpp:22-109.
#pragma once
#include <cmath>
#include <memory>
#include <stdexcept>
#include <vector>

#include "sqbem.h"
#include "utils/field.hpp"
#include "utils/kernels.hpp"
#include "utils/linalg.hpp"
#include "utils/mesh.hpp"

namespace bem {

class DuctFlow {
public:
    Mesh mesh;
    double R, L, mu;

    struct Solution {
        std::vector<V3> u, f;
        double Q = 0;  // volume flow rate through the outlet
    };

    DuctFlow(double R_, double L_, int n_theta, int n_z, int n_r, double mu_ = 1.0) : R(R_), L(L_), mu(mu_) {
        std::shared_ptr<sqbem_mesh> m(sqbem_mesh_duct(R, L, n_theta, n_z, n_r), sqbem_mesh_free);
        if (!m) throw std::runtime_error(sqbem_last_error());
        mesh = detail::from_sqbem(m.get());
        sqbem_duct* d = sqbem_duct_create(m.get(), mu);
        if (!d) throw std::runtime_error(sqbem_last_error());
        h_.reset(d, sqbem_duct_free);
    }

    // true if the velocity component (panel p, component b) is prescribed
    bool u_known(int p, int b) const { return !(mesh.tags[p] != WALL && b == 2); }

    Solution solve(double p_in, double p_out) const {
        const int P = mesh.P();
        std::vector<double> u(3 * (size_t)P), f(3 * (size_t)P);
        Solution s;
        sqbem_check(sqbem_duct_solve(h_.get(), p_in, p_out, u.data(), f.data(), &s.Q));
        s.u.resize(P);
        s.f.resize(P);
        for (int p = 0; p < P; ++p) {
            s.u[p] = V3{u[3 * p], u[3 * p + 1], u[3 * p + 2]};
            s.f[p] = V3{f[3 * p], f[3 * p + 1], f[3 * p + 2]};
        }
        return s;
    }

    std::vector<V3> velocity(const Solution& s, const std::vector<V3>& pts) const {
        return velocity_at(mesh, s.u, s.f, mu, pts);
    }

    double exact_uz(double r, double dP) const { return dP / L * (R * R - r * r) / (4 * mu); }
    double exact_Q(double dP) const { return PI * std::pow(R, 4) * dP / L / (8 * mu); }

private:
    std::shared_ptr<sqbem_duct> h_;
};

}  // namespace bem
