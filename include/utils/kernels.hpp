// kernels.hpp -- free-space Stokes Green's functions integrated over flat panels.
//
//   Stokeslet   G_ij(r) = d_ij/|r| + r_i r_j/|r|^3,   stresslet T_ijk(r) = -6 r_i r_j r_k/|r|^5
//   SL_ij = int_p G_ij dS,   DL_ij = int_p T_ijk m_k dS       (row-major 3x3, symmetric)
//
// Rule by d = |c_p - x0|/h_p: d >= 3 -> 7 points, [1.5, 3) -> 28, [0.75, 1.5) -> 112,
// < 0.75 -> 448; the panel containing x0 -> Duffy 3 x 10 x 10 (SL only, DL = 0).
//
// API of src/utils/kernels.hpp:67-82, 107-111. panel_integrals evaluates one pair on the GPU
// (the batched form used by the solvers is sqbem_assemble); squirmer_slip is host algebra.
#pragma once
#include <memory>
#include <stdexcept>

#include "../sqbem.h"
#include "mesh.hpp"
#include "quadrature.hpp"
#include "vec3.hpp"

namespace bem {

namespace detail {
// libsqbem copy of a host mesh (geometry recomputed by the library from nodes and panels)
inline std::shared_ptr<sqbem_mesh> to_sqbem(const Mesh& M) {
    std::vector<double> nodes(3 * M.nodes.size());
    for (size_t i = 0; i < M.nodes.size(); ++i)
        for (int c = 0; c < 3; ++c) nodes[3 * i + c] = M.nodes[i][c];
    std::vector<int> tris(3 * (size_t)M.P());
    for (int p = 0; p < M.P(); ++p)
        for (int k = 0; k < 3; ++k) tris[3 * (size_t)p + k] = M.tris[p][k];
    sqbem_mesh* h = sqbem_mesh_from_arrays(nodes.data(), (int)M.nodes.size(), tris.data(),
                                           M.tags.empty() ? nullptr : M.tags.data(), M.P());
    if (!h) throw std::runtime_error(sqbem_last_error());
    return std::shared_ptr<sqbem_mesh>(h, sqbem_mesh_free);
}

// host Mesh from a libsqbem mesh (nodes, panels, tags, geometry)
inline Mesh from_sqbem(const sqbem_mesh* h) {
    int nn = 0, np = 0;
    if (sqbem_mesh_sizes(h, &nn, &np) != SQBEM_OK) throw std::runtime_error(sqbem_last_error());
    std::vector<double> nodes(3 * (size_t)nn);
    std::vector<int> tris(3 * (size_t)np), tags(np);
    if (sqbem_mesh_get(h, nodes.data(), tris.data(), tags.data(), nullptr, nullptr, nullptr, nullptr, nullptr) !=
        SQBEM_OK)
        throw std::runtime_error(sqbem_last_error());
    Mesh m;
    m.nodes.resize(nn);
    for (int i = 0; i < nn; ++i) m.nodes[i] = V3{nodes[3 * i], nodes[3 * i + 1], nodes[3 * i + 2]};
    m.tris.resize(np);
    for (int p = 0; p < np; ++p) m.tris[p] = {tris[3 * p], tris[3 * p + 1], tris[3 * p + 2]};
    m.tags = tags;
    m.geometry();
    return m;
}
}  // namespace detail

// SL and DL (row-major 3x3) of panel p of M at target x0; self: x0 lies in panel p
inline void panel_integrals(const Mesh& M, int p, const V3& x0, bool self, double* SL, double* DL) {
    auto h = detail::to_sqbem(M);
    const double x[3] = {x0.x, x0.y, x0.z};
    if (sqbem_panel_integrals(h.get(), p, x, self ? 1 : 0, SL, DL) != SQBEM_OK)
        throw std::runtime_error(sqbem_last_error());
}

// Tangential slip of the Blake / Ishikawa-Pedley squirmer:
//   u_s = (B1 + B2 cos th) sin th e_th,  cos th = rhat . e,  sin th e_th = cos th rhat - e
inline V3 squirmer_slip(V3 x, V3 center, V3 e, double B1, double B2) {
    const V3 rh = unit(x - center);
    const double ct = dot(rh, e);
    return (B1 + B2 * ct) * (ct * rh - e);
}

}  // namespace bem
