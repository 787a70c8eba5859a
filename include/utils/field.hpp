// field.hpp -- velocity at points inside the fluid from the boundary solution:
//
//   u_j(x0) = 1/(8 pi mu) sum_p SL_p,ij f_p,i  -  1/(8 pi) sum_p DL_p,ij u_p,i
//
// API of src/utils/field.hpp:18-34; evaluated by the libsqbem field kernel (the assembly
// kernel fused with the matrix-vector product, no matrix stored).
#pragma once
#include <stdexcept>
#include <vector>

#include "kernels.hpp"

namespace bem {

inline std::vector<V3> velocity_at(const Mesh& M, const std::vector<V3>& u, const std::vector<V3>& f, double mu,
                                   const std::vector<V3>& points) {
    std::vector<V3> out(points.size());
    if (points.empty()) return out;
    auto h = detail::to_sqbem(M);
    std::vector<double> uu(3 * u.size()), ff(3 * f.size()), pp(3 * points.size()), oo(3 * points.size());
    for (size_t p = 0; p < u.size(); ++p)
        for (int c = 0; c < 3; ++c) uu[3 * p + c] = u[p][c];
    for (size_t p = 0; p < f.size(); ++p)
        for (int c = 0; c < 3; ++c) ff[3 * p + c] = f[p][c];
    for (size_t n = 0; n < points.size(); ++n)
        for (int c = 0; c < 3; ++c) pp[3 * n + c] = points[n][c];
    if (sqbem_velocity_at(h.get(), uu.data(), ff.data(), mu, pp.data(), (int)points.size(), oo.data()) != SQBEM_OK)
        throw std::runtime_error(sqbem_last_error());
    for (size_t n = 0; n < points.size(); ++n) out[n] = V3{oo[3 * n], oo[3 * n + 1], oo[3 * n + 2]};
    return out;
}

}  // namespace bem
