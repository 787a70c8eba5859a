// ref_driver.cpp -- C entry points around the reference's own geometry, quadrature, kernel
// and field headers (src/utils/{vec3,mesh,quadrature,kernels,field}.hpp), compiled from the
// reference sources where they lie into oracle/_ref/libref.so.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle restatement; never linked
// into the product. The reference's solver headers need LAPACK/BLAS, an external library,
// and are not built (see the oracle Makefile).
#include <cstring>
#include <vector>
#include "utils/field.hpp"
#include "utils/kernels.hpp"
#include "utils/mesh.hpp"
#include "utils/quadrature.hpp"

using namespace bem;

extern "C" {

void* ref_swimmer_tube_mesh(double R, double L, double a, int nth, int nr) {
    return new Mesh(swimmer_tube_mesh(R, L, a, nth, nr));
}
void* ref_duct_mesh(double R, double L, int nth, int nz, int nr) {
    return new Mesh(tube_mesh(R, uniform_z(0, L, nz), nth, nr));
}
void* ref_icosphere(double a, int level, double cx, double cy, double cz) {
    return new Mesh(icosphere(a, level, V3{cx, cy, cz}));
}
void* ref_merge(void* a, void* b) { return new Mesh(merge(*(Mesh*)a, *(Mesh*)b)); }
void ref_translate(void* m, double dx, double dy, double dz) { ((Mesh*)m)->translate(V3{dx, dy, dz}); }
void ref_mesh_free(void* m) { delete (Mesh*)m; }
int ref_mesh_nn(void* m) { return (int)((Mesh*)m)->nodes.size(); }
int ref_mesh_np(void* m) { return ((Mesh*)m)->P(); }

void ref_mesh_get(void* mh, double* nodes, int* tris, int* tags, double* V, double* normal,
                  double* centroid, double* area, double* h) {
    const Mesh& m = *(Mesh*)mh;
    for (size_t i = 0; i < m.nodes.size(); ++i)
        for (int c = 0; c < 3; ++c) nodes[3 * i + c] = m.nodes[i][c];
    for (int p = 0; p < m.P(); ++p) {
        for (int k = 0; k < 3; ++k) tris[3 * p + k] = m.tris[p][k];
        tags[p] = m.tags[p];
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) V[9 * p + 3 * k + c] = m.V[p][k][c];
        for (int c = 0; c < 3; ++c) {
            normal[3 * p + c] = m.normal[p][c];
            centroid[3 * p + c] = m.centroid[p][c];
        }
        area[p] = m.area[p];
        h[p] = m.h[p];
    }
}

int ref_graded_z(double L, double dzmin, double dzmax, double core, double* out, int cap) {
    auto z = graded_z(L, dzmin, dzmax, core);
    if ((int)z.size() > cap) return -1;
    for (size_t i = 0; i < z.size(); ++i) out[i] = z[i];
    return (int)z.size();
}

int ref_rule(int lev, double* L, double* W) {
    const Rule& r = rules()[lev];
    for (size_t q = 0; q < r.W.size(); ++q) {
        for (int c = 0; c < 3; ++c) L[3 * q + c] = r.L[q][c];
        W[q] = r.W[q];
    }
    return (int)r.W.size();
}

void ref_gauss_legendre01(int n, double* x, double* w) {
    std::vector<double> xx, ww;
    gauss_legendre01(n, xx, ww);
    for (int i = 0; i < n; ++i) { x[i] = xx[i]; w[i] = ww[i]; }
}

void ref_panel_integrals(void* mh, int p, const double* x0, int self, double* SL, double* DL) {
    panel_integrals(*(Mesh*)mh, p, V3{x0[0], x0[1], x0[2]}, self != 0, SL, DL);
}

// Distance ratio and rule level exactly as panel_integrals computes them
// (src/utils/kernels.hpp:74-75); target-major output i*np + p.
void ref_levels(void* mh, int p0, int np, const double* x0, int nt, signed char* lev, double* dist) {
    const Mesh& M = *(Mesh*)mh;
    for (int i = 0; i < nt; ++i)
        for (int p = 0; p < np; ++p) {
            V3 x{x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]};
            double d = norm(M.centroid[p0 + p] - x) / M.h[p0 + p];
            lev[(size_t)i * np + p] = (signed char)(d < 0.75 ? 3 : (d < 1.5 ? 2 : (d < 3.0 ? 1 : 0)));
            if (dist) dist[(size_t)i * np + p] = d;
        }
}

void ref_block(void* mh, int p0, int np, const double* x0, int nt, const int* self_src,
               double* SL, double* DL) {
    const Mesh& M = *(Mesh*)mh;
    for (int i = 0; i < nt; ++i)
        for (int p = 0; p < np; ++p) {
            bool self = self_src && self_src[i] == p0 + p;
            size_t o = ((size_t)i * np + p) * 9;
            panel_integrals(M, p0 + p, V3{x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]}, self, SL + o, DL + o);
        }
}

void ref_velocity_at(void* mh, const double* u, const double* f, double mu, const double* pts,
                     int npts, double* out) {
    const Mesh& M = *(Mesh*)mh;
    std::vector<V3> uu(M.P()), ff(M.P()), pp(npts);
    for (int p = 0; p < M.P(); ++p) {
        uu[p] = {u[3 * p], u[3 * p + 1], u[3 * p + 2]};
        ff[p] = {f[3 * p], f[3 * p + 1], f[3 * p + 2]};
    }
    for (int n = 0; n < npts; ++n) pp[n] = {pts[3 * n], pts[3 * n + 1], pts[3 * n + 2]};
    auto v = velocity_at(M, uu, ff, mu, pp);
    for (int n = 0; n < npts; ++n)
        for (int c = 0; c < 3; ++c) out[3 * n + c] = v[n][c];
}

void ref_squirmer_slip(const double* x, const double* c, const double* e, double B1, double B2,
                       double* out) {
    V3 u = squirmer_slip(V3{x[0], x[1], x[2]}, V3{c[0], c[1], c[2]}, V3{e[0], e[1], e[2]}, B1, B2);
    for (int k = 0; k < 3; ++k) out[k] = u[k];
}

}  // extern "C"
