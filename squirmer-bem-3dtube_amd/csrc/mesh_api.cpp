// mesh_api.cpp -- C exports of the host mesh generators (include/utils/mesh.hpp).
// Compiled with floating-point contraction off so panel geometry, and with it every
// quadrature-rule decision, is reproducible on any host.
#include <cstring>
#include <string>

#include "../../include/sqbem.h"
#include "../../include/utils/mesh.hpp"
#include "mesh_handle.hpp"

namespace sq {
void set_last_error(const std::string& msg);
}

template <typename F>
static sqbem_mesh* make(F&& f) {
    try {
        return new sqbem_mesh{f()};
    } catch (const std::exception& e) {
        sq::set_last_error(e.what());
        return nullptr;
    }
}

extern "C" {

sqbem_mesh* sqbem_mesh_swimmer_tube(double R, double L, double a, int n_theta, int n_r) {
    if (!(R > 0) || !(L > 0) || n_theta < 3 || n_r < 1) {
        sq::set_last_error("sqbem_mesh_swimmer_tube: invalid arguments");
        return nullptr;
    }
    return make([&] { return bem::swimmer_tube_mesh(R, L, a, n_theta, n_r); });
}

sqbem_mesh* sqbem_mesh_tube(double R, const double* z, int nz_nodes, int n_theta, int n_r) {
    if (!(R > 0) || !z || nz_nodes < 2 || n_theta < 3 || n_r < 1) {
        sq::set_last_error("sqbem_mesh_tube: invalid arguments");
        return nullptr;
    }
    return make([&] { return bem::tube_mesh(R, std::vector<double>(z, z + nz_nodes), n_theta, n_r); });
}

sqbem_mesh* sqbem_mesh_duct(double R, double L, int n_theta, int n_z, int n_r) {
    if (!(R > 0) || !(L > 0) || n_theta < 3 || n_z < 1 || n_r < 1) {
        sq::set_last_error("sqbem_mesh_duct: invalid arguments");
        return nullptr;
    }
    return make([&] { return bem::tube_mesh(R, bem::uniform_z(0, L, n_z), n_theta, n_r); });
}

sqbem_mesh* sqbem_mesh_icosphere(double a, int level, double cx, double cy, double cz) {
    if (!(a > 0) || level < 0 || level > 7) {
        sq::set_last_error("sqbem_mesh_icosphere: invalid arguments");
        return nullptr;
    }
    return make([&] { return bem::icosphere(a, level, bem::V3{cx, cy, cz}); });
}

sqbem_mesh* sqbem_mesh_from_arrays(const double* nodes, int nn, const int* tris, const int* tags,
                                   int np) {
    if (!nodes || !tris || nn <= 0 || np <= 0) {
        sq::set_last_error("sqbem_mesh_from_arrays: invalid arguments");
        return nullptr;
    }
    for (int i = 0; i < 3 * np; ++i)
        if (tris[i] < 0 || tris[i] >= nn) {
            sq::set_last_error("sqbem_mesh_from_arrays: node index out of range");
            return nullptr;
        }
    return make([&] {
        bem::Mesh m;
        m.nodes.resize(nn);
        for (int i = 0; i < nn; ++i) m.nodes[i] = bem::V3{nodes[3 * i], nodes[3 * i + 1], nodes[3 * i + 2]};
        m.tris.resize(np);
        for (int p = 0; p < np; ++p) m.tris[p] = {tris[3 * p], tris[3 * p + 1], tris[3 * p + 2]};
        m.tags.assign(np, 0);
        if (tags) m.tags.assign(tags, tags + np);
        m.geometry();
        return m;
    });
}

sqbem_mesh* sqbem_mesh_merge(const sqbem_mesh* a, const sqbem_mesh* b) {
    if (!a || !b) {
        sq::set_last_error("sqbem_mesh_merge: null mesh");
        return nullptr;
    }
    return make([&] { return bem::merge(a->m, b->m); });
}

int sqbem_mesh_translate(sqbem_mesh* m, double dx, double dy, double dz) {
    if (!m) {
        sq::set_last_error("sqbem_mesh_translate: null mesh");
        return SQBEM_ERR_ARG;
    }
    m->m.translate(bem::V3{dx, dy, dz});
    return SQBEM_OK;
}

int sqbem_mesh_sizes(const sqbem_mesh* m, int* nn, int* np) {
    if (!m) {
        sq::set_last_error("sqbem_mesh_sizes: null mesh");
        return SQBEM_ERR_ARG;
    }
    if (nn) *nn = (int)m->m.nodes.size();
    if (np) *np = m->m.P();
    return SQBEM_OK;
}

int sqbem_mesh_get(const sqbem_mesh* mh, double* nodes, int* tris, int* tags, double* V,
                   double* normal, double* centroid, double* area, double* h) {
    if (!mh) {
        sq::set_last_error("sqbem_mesh_get: null mesh");
        return SQBEM_ERR_ARG;
    }
    const bem::Mesh& m = mh->m;
    if (nodes)
        for (size_t i = 0; i < m.nodes.size(); ++i)
            for (int c = 0; c < 3; ++c) nodes[3 * i + c] = m.nodes[i][c];
    for (int p = 0; p < m.P(); ++p) {
        if (tris)
            for (int k = 0; k < 3; ++k) tris[3 * p + k] = m.tris[p][k];
        if (tags) tags[p] = m.tags[p];
        if (V)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) V[9 * p + 3 * k + c] = m.V[p][k][c];
        for (int c = 0; c < 3; ++c) {
            if (normal) normal[3 * p + c] = m.normal[p][c];
            if (centroid) centroid[3 * p + c] = m.centroid[p][c];
        }
        if (area) area[p] = m.area[p];
        if (h) h[p] = m.h[p];
    }
    return SQBEM_OK;
}

void sqbem_mesh_free(sqbem_mesh* m) { delete m; }

}  // extern "C"
