/*
 * bem_oracle.c -- CPU restatement of the squirmer-in-tube BEM geometry, quadrature and
 * panel integrals, used ONLY as a test oracle and as the CPU baseline.
 *
 * TEST INFRASTRUCTURE: nothing in the product library links, loads or calls this file.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * Every function names the reference file:line it restates (paths relative to the
 * reference repository root). Floating-point contraction is disabled (-ffp-contract=off)
 * so results do not depend on the host CPU or compiler; the reference build contracts
 * some products into FMAs, which changes values at the rounding level only. The rule
 * choice (a discontinuous function of a distance) is compared against the reference
 * build pair by pair in the tests.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OM_PI 3.14159265358979323846
#define OM_EIGHT_PI (8.0 * OM_PI)

enum { OM_WALL = 0, OM_INLET = 1, OM_OUTLET = 2, OM_SPHERE = 10 };

typedef struct {
    int nn, np;
    double *nodes;    /* nn*3 */
    int *tris;        /* np*3 */
    int *tags;        /* np */
    double *V;        /* np*9: vertex coordinates per panel */
    double *normal;   /* np*3 */
    double *centroid; /* np*3 */
    double *area;     /* np */
    double *h;        /* np: sqrt(2*area) */
} om_mesh;

/* ---- 3-vector helpers (src/utils/vec3.hpp:18-26) ---- */
static inline void v_sub(const double *a, const double *b, double *o) {
    o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}
static inline double v_dot(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void v_cross(const double *a, const double *b, double *o) {
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}
static inline double v_norm(const double *a) { return sqrt(v_dot(a, a)); }
static inline void v_unit(double *a) {
    double s = 1.0 / v_norm(a);
    a[0] = s * a[0]; a[1] = s * a[1]; a[2] = s * a[2];
}

/* ---- mesh container ---- */
static om_mesh *mesh_alloc(int nn, int np) {
    om_mesh *m = (om_mesh *)calloc(1, sizeof(om_mesh));
    m->nn = nn; m->np = np;
    m->nodes = (double *)calloc((size_t)nn * 3 + 1, sizeof(double));
    m->tris = (int *)calloc((size_t)np * 3 + 1, sizeof(int));
    m->tags = (int *)calloc((size_t)np + 1, sizeof(int));
    m->V = (double *)calloc((size_t)np * 9 + 1, sizeof(double));
    m->normal = (double *)calloc((size_t)np * 3 + 1, sizeof(double));
    m->centroid = (double *)calloc((size_t)np * 3 + 1, sizeof(double));
    m->area = (double *)calloc((size_t)np + 1, sizeof(double));
    m->h = (double *)calloc((size_t)np + 1, sizeof(double));
    return m;
}

void om_mesh_free(om_mesh *m) {
    if (!m) return;
    free(m->nodes); free(m->tris); free(m->tags); free(m->V); free(m->normal);
    free(m->centroid); free(m->area); free(m->h); free(m);
}

/* Mesh::geometry, src/utils/mesh.hpp:61-72 */
static void mesh_geometry(om_mesh *m) {
    for (int p = 0; p < m->np; ++p) {
        double *V = m->V + 9 * p;
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) V[3 * k + c] = m->nodes[3 * m->tris[3 * p + k] + c];
        double e1[3], e2[3], cr[3];
        v_sub(V + 3, V, e1);
        v_sub(V + 6, V, e2);
        v_cross(e1, e2, cr);
        double A = 0.5 * v_norm(cr);
        m->area[p] = A;
        double s = 1.0 / (2 * A);
        for (int c = 0; c < 3; ++c) m->normal[3 * p + c] = s * cr[c];
        for (int c = 0; c < 3; ++c) m->centroid[3 * p + c] = (1.0 / 3) * ((V[c] + V[3 + c]) + V[6 + c]);
        m->h[p] = sqrt(2 * A);
    }
}

/* Mesh::orient, src/utils/mesh.hpp:76-81. kind selects the outward field:
 * 0: (x, y, 0) tube wall; 1: (0, 0, sg) caps; 2: center - x (sphere). */
static void mesh_orient(om_mesh *m, int kind, double sg, const double *center) {
    mesh_geometry(m);
    for (int p = 0; p < m->np; ++p) {
        const double *c = m->centroid + 3 * p;
        double o[3];
        if (kind == 0) { o[0] = c[0]; o[1] = c[1]; o[2] = 0; }
        else if (kind == 1) { o[0] = 0; o[1] = 0; o[2] = sg; }
        else { o[0] = center[0] - c[0]; o[1] = center[1] - c[1]; o[2] = center[2] - c[2]; }
        if (v_dot(m->normal + 3 * p, o) < 0) {
            int t = m->tris[3 * p + 1];
            m->tris[3 * p + 1] = m->tris[3 * p + 2];
            m->tris[3 * p + 2] = t;
        }
    }
    mesh_geometry(m);
}

/* merge, src/utils/mesh.hpp:95-103 */
om_mesh *om_merge(const om_mesh *a, const om_mesh *b) {
    om_mesh *m = mesh_alloc(a->nn + b->nn, a->np + b->np);
    memcpy(m->nodes, a->nodes, sizeof(double) * 3 * a->nn);
    memcpy(m->nodes + 3 * a->nn, b->nodes, sizeof(double) * 3 * b->nn);
    memcpy(m->tris, a->tris, sizeof(int) * 3 * a->np);
    for (int i = 0; i < 3 * b->np; ++i) m->tris[3 * a->np + i] = b->tris[i] + a->nn;
    memcpy(m->tags, a->tags, sizeof(int) * a->np);
    memcpy(m->tags + a->np, b->tags, sizeof(int) * b->np);
    mesh_geometry(m);
    return m;
}

/* Mesh::translate, src/utils/mesh.hpp:82-85 */
void om_translate(om_mesh *m, double dx, double dy, double dz) {
    for (int i = 0; i < m->nn; ++i) {
        m->nodes[3 * i] = m->nodes[3 * i] + dx;
        m->nodes[3 * i + 1] = m->nodes[3 * i + 1] + dy;
        m->nodes[3 * i + 2] = m->nodes[3 * i + 2] + dz;
    }
    mesh_geometry(m);
}

om_mesh *om_copy(const om_mesh *a) {
    om_mesh *m = mesh_alloc(a->nn, a->np);
    memcpy(m->nodes, a->nodes, sizeof(double) * 3 * a->nn);
    memcpy(m->tris, a->tris, sizeof(int) * 3 * a->np);
    memcpy(m->tags, a->tags, sizeof(int) * a->np);
    mesh_geometry(m);
    return m;
}

/* uniform_z, src/utils/mesh.hpp:112-116 */
int om_uniform_z(double z0, double z1, int nz, double *out) {
    for (int k = 0; k <= nz; ++k) out[k] = z0 + (z1 - z0) * k / nz;
    return nz + 1;
}

/* graded_z, src/utils/mesh.hpp:122-136. Returns the node count (<= cap), or -1. */
int om_graded_z(double L, double dzmin, double dzmax, double core, double growth, double *out,
                int cap) {
    int nh = 1, hcap = cap;
    double *half = (double *)malloc(sizeof(double) * (size_t)(hcap + 2));
    half[0] = 0.0;
    double dz = dzmin;
    while (half[nh - 1] < L / 2 - 1e-12) {
        if (half[nh - 1] >= core) { double g = dz * growth; dz = g < dzmax ? g : dzmax; }
        double nx = half[nh - 1] + dz;
        if (nh >= hcap) { free(half); return -1; }
        half[nh++] = nx < L / 2 ? nx : L / 2;
    }
    if (nh > 2 && L / 2 - half[nh - 2] < 0.3 * dz) {
        half[nh - 2] = half[nh - 1];
        --nh;
    }
    int n = 2 * nh - 1;
    if (n > cap) { free(half); return -1; }
    int k = 0;
    for (int i = nh - 1; i >= 0; --i) out[k++] = -half[i];
    for (int i = 1; i < nh; ++i) out[k++] = half[i];
    free(half);
    return n;
}

/* tube_mesh, src/utils/mesh.hpp:149-190 (wall, inlet cap, outlet cap, merged in that order) */
static om_mesh *tube_cap(double z0, int tag, double sg, int nth, int nr, double Rp, double al) {
    om_mesh *m = mesh_alloc(1 + nr * nth, nth + 2 * nth * (nr - 1));
    m->nodes[0] = 0; m->nodes[1] = 0; m->nodes[2] = z0;
    int q = 1;
    for (int k = 1; k <= nr; ++k)
        for (int i = 0; i < nth; ++i, ++q) {
            double rk = Rp * k / nr;
            m->nodes[3 * q] = rk * cos(i * al);
            m->nodes[3 * q + 1] = rk * sin(i * al);
            m->nodes[3 * q + 2] = z0;
        }
#define RING(k, i) (1 + (k) * nth + ((i) % nth))
    int t = 0;
    for (int i = 0; i < nth; ++i, ++t) {
        m->tris[3 * t] = 0; m->tris[3 * t + 1] = RING(0, i); m->tris[3 * t + 2] = RING(0, i + 1);
    }
    for (int k = 0; k < nr - 1; ++k)
        for (int i = 0; i < nth; ++i) {
            int a = RING(k, i), b = RING(k, i + 1), c = RING(k + 1, i + 1), d = RING(k + 1, i);
            m->tris[3 * t] = a; m->tris[3 * t + 1] = d; m->tris[3 * t + 2] = c; ++t;
            m->tris[3 * t] = a; m->tris[3 * t + 1] = c; m->tris[3 * t + 2] = b; ++t;
        }
#undef RING
    for (int p = 0; p < m->np; ++p) m->tags[p] = tag;
    mesh_orient(m, 1, sg, NULL);
    return m;
}

om_mesh *om_tube_mesh(double R, const double *z, int nzn, int nth, int nr) {
    double al = 2 * OM_PI / nth;
    double Rp = R * sqrt(al / sin(al));
    int nz = nzn - 1;
    om_mesh *wall = mesh_alloc(nth * (nz + 1), 2 * nth * nz);
    int q = 0;
    for (int i = 0; i < nth; ++i)
        for (int k = 0; k <= nz; ++k, ++q) {
            wall->nodes[3 * q] = Rp * cos(i * al);
            wall->nodes[3 * q + 1] = Rp * sin(i * al);
            wall->nodes[3 * q + 2] = z[k];
        }
#define WID(i, k) (((i) % nth) * (nz + 1) + (k))
    int t = 0;
    for (int i = 0; i < nth; ++i)
        for (int k = 0; k < nz; ++k) {
            int a = WID(i, k), b = WID(i + 1, k), c = WID(i + 1, k + 1), d = WID(i, k + 1);
            wall->tris[3 * t] = a; wall->tris[3 * t + 1] = b; wall->tris[3 * t + 2] = c; ++t;
            wall->tris[3 * t] = a; wall->tris[3 * t + 1] = c; wall->tris[3 * t + 2] = d; ++t;
        }
#undef WID
    for (int p = 0; p < wall->np; ++p) wall->tags[p] = OM_WALL;
    mesh_orient(wall, 0, 0, NULL);
    om_mesh *in = tube_cap(z[0], OM_INLET, -1.0, nth, nr, Rp, al);
    om_mesh *out = tube_cap(z[nz], OM_OUTLET, 1.0, nth, nr, Rp, al);
    om_mesh *wi = om_merge(wall, in);
    om_mesh *all = om_merge(wi, out);
    om_mesh_free(wall); om_mesh_free(in); om_mesh_free(out); om_mesh_free(wi);
    return all;
}

/* swimmer_tube_mesh, src/utils/mesh.hpp:194-197 */
om_mesh *om_swimmer_tube_mesh(double R, double L, double a, int nth, int nr) {
    double h = 2 * OM_PI * R / nth;
    int cap = 100000;
    double *z = (double *)malloc(sizeof(double) * cap);
    int n = om_graded_z(L, 0.6 * h, 1.4 * h, a + 0.5 * R, 1.25, z, cap);
    om_mesh *m = n > 0 ? om_tube_mesh(R, z, n, nth, nr) : NULL;
    free(z);
    return m;
}

/* duct mesh: tube_mesh(R, uniform_z(0, L, n_z), ...), src/duct_flow.hpp:32-36 */
om_mesh *om_duct_mesh(double R, double L, int nth, int nz, int nr) {
    double *z = (double *)malloc(sizeof(double) * (size_t)(nz + 1));
    om_uniform_z(0, L, nz, z);
    om_mesh *m = om_tube_mesh(R, z, nz + 1, nth, nr);
    free(z);
    return m;
}

/* icosphere, src/utils/mesh.hpp:208-245. Edge midpoints are numbered in order of first
 * use, exactly as the reference's ordered midpoint cache does. */
om_mesh *om_icosphere(double a, int level, double cx, double cy, double cz, int tag) {
    double t = (1 + sqrt(5.0)) / 2;
    int nv_final = 10 * (1 << (2 * level)) + 2;
    int nf_final = 20 * (1 << (2 * level));
    double *v = (double *)malloc(sizeof(double) * 3 * (size_t)nv_final);
    const double v0[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                              {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (int i = 0; i < 12; ++i) {
        for (int c = 0; c < 3; ++c) v[3 * i + c] = v0[i][c];
        v_unit(v + 3 * i);
    }
    int nv = 12;
    const int f0[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                           {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                           {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    int *F = (int *)malloc(sizeof(int) * 3 * (size_t)nf_final);
    int *NF = (int *)malloc(sizeof(int) * 3 * (size_t)nf_final);
    int nf = 20;
    for (int i = 0; i < 20; ++i)
        for (int c = 0; c < 3; ++c) F[3 * i + c] = f0[i][c];
    /* adjacency cache: for vertex lo, up to 8 (hi, mid) entries */
    int *ej = (int *)malloc(sizeof(int) * 8 * (size_t)nv_final);
    int *em = (int *)malloc(sizeof(int) * 8 * (size_t)nv_final);
    int *ec = (int *)malloc(sizeof(int) * (size_t)nv_final);
    for (int l = 0; l < level; ++l) {
        memset(ec, 0, sizeof(int) * (size_t)nv_final);
        int nnf = 0;
        for (int f = 0; f < nf; ++f) {
            int mids[3];
            int pairs[3][2] = {{F[3 * f], F[3 * f + 1]}, {F[3 * f + 1], F[3 * f + 2]}, {F[3 * f + 2], F[3 * f]}};
            for (int e = 0; e < 3; ++e) {
                int i = pairs[e][0], j = pairs[e][1];
                int lo = i < j ? i : j, hi = i < j ? j : i, found = -1;
                for (int q = 0; q < ec[lo]; ++q)
                    if (ej[8 * lo + q] == hi) { found = em[8 * lo + q]; break; }
                if (found < 0) {
                    double *w = v + 3 * nv;
                    for (int c = 0; c < 3; ++c) w[c] = v[3 * i + c] + v[3 * j + c];
                    v_unit(w);
                    found = nv++;
                    ej[8 * lo + ec[lo]] = hi;
                    em[8 * lo + ec[lo]] = found;
                    ec[lo]++;
                }
                mids[e] = found;
            }
            int ab = mids[0], bc = mids[1], ca = mids[2];
            int *o = NF + 3 * nnf;
            o[0] = F[3 * f]; o[1] = ab; o[2] = ca;
            o[3] = F[3 * f + 1]; o[4] = bc; o[5] = ab;
            o[6] = F[3 * f + 2]; o[7] = ca; o[8] = bc;
            o[9] = ab; o[10] = bc; o[11] = ca;
            nnf += 4;
        }
        int *tmp = F; F = NF; NF = tmp;
        nf = nnf;
    }
    double vol = 0;
    for (int f = 0; f < nf; ++f) {
        double cr[3];
        v_cross(v + 3 * F[3 * f + 1], v + 3 * F[3 * f + 2], cr);
        vol += v_dot(v + 3 * F[3 * f], cr);
    }
    double s = cbrt(4 * OM_PI / 3 / (fabs(vol) / 6));
    om_mesh *m = mesh_alloc(nv, nf);
    double as = a * s, ctr[3] = {cx, cy, cz};
    for (int i = 0; i < nv; ++i)
        for (int c = 0; c < 3; ++c) m->nodes[3 * i + c] = ctr[c] + as * v[3 * i + c];
    memcpy(m->tris, F, sizeof(int) * 3 * (size_t)nf);
    for (int p = 0; p < nf; ++p) m->tags[p] = tag;
    mesh_orient(m, 2, 0, ctr);
    free(v); free(F); free(NF); free(ej); free(em); free(ec);
    return m;
}

/* ---- quadrature (src/utils/quadrature.hpp) ---- */
typedef struct { int n; double L[448][3]; double W[448]; } om_rule;

/* dunavant7, src/utils/quadrature.hpp:22-30: the degree-5 seven-point rule with the
 * reference's 15-digit constants. */
static void dunavant7(om_rule *r) {
    const double a1 = 0.059715871789770, b1 = 0.470142064105115, w1 = 0.132394152788506;
    const double a2 = 0.797426985353087, b2 = 0.101286507323456, w2 = 0.125939180544827;
    const double L[7][3] = {{1. / 3, 1. / 3, 1. / 3}, {a1, b1, b1}, {b1, a1, b1}, {b1, b1, a1},
                            {a2, b2, b2}, {b2, a2, b2}, {b2, b2, a2}};
    const double W[7] = {0.225, w1, w1, w1, w2, w2, w2};
    r->n = 7;
    for (int q = 0; q < 7; ++q) {
        for (int c = 0; c < 3; ++c) r->L[q][c] = L[q][c];
        r->W[q] = W[q];
    }
}

/* subdivided, src/utils/quadrature.hpp:37-64: the seven-point rule on the 4^level
 * sub-triangles, sub-triangles in the order of the nested splitting. */
static void subdivided(int level, om_rule *r) {
    int nt = 1;
    static double T[64 * 9], NT[64 * 9];
    const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(T, I, sizeof I);
    for (int l = 0; l < level; ++l) {
        int k = 0;
        for (int t = 0; t < nt; ++t) {
            const double *a = T + 9 * t, *b = a + 3, *c = a + 6;
            double ab[3], bc[3], ca[3];
            for (int q = 0; q < 3; ++q) {
                ab[q] = (a[q] + b[q]) / 2; bc[q] = (b[q] + c[q]) / 2; ca[q] = (c[q] + a[q]) / 2;
            }
            const double *tri[4][3] = {{a, ab, ca}, {ab, b, bc}, {ca, bc, c}, {ab, bc, ca}};
            for (int s = 0; s < 4; ++s, ++k)
                for (int v = 0; v < 3; ++v)
                    for (int q = 0; q < 3; ++q) NT[9 * k + 3 * v + q] = tri[s][v][q];
        }
        nt = k;
        memcpy(T, NT, sizeof(double) * 9 * (size_t)nt);
    }
    om_rule d;
    dunavant7(&d);
    int n = 0;
    for (int t = 0; t < nt; ++t)
        for (int q = 0; q < 7; ++q, ++n) {
            double l[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) l[c] += d.L[q][k] * T[9 * t + 3 * k + c];
            for (int c = 0; c < 3; ++c) r->L[n][c] = l[c];
            r->W[n] = d.W[q] / nt;
        }
    r->n = n;
}

static om_rule g_rules[4];
static double g_gl_x[10], g_gl_w[10];
static int g_init = 0;

/* gauss_legendre01, src/utils/quadrature.hpp:79-99: Newton on P_n from a cosine guess */
void om_gauss_legendre01(int n, double *x, double *w) {
    for (int i = 0; i < n; ++i) {
        double z = cos(OM_PI * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-15) break;
        }
        x[i] = 0.5 * (1 - z);
        w[i] = 1.0 / ((1 - z * z) * pp * pp);
    }
}

/* rules(), src/utils/quadrature.hpp:69-72 */
static void init_tables(void) {
    if (g_init) return;
    dunavant7(&g_rules[0]);
    for (int l = 1; l <= 3; ++l) subdivided(l, &g_rules[l]);
    om_gauss_legendre01(10, g_gl_x, g_gl_w);
    g_init = 1;
}

int om_rule_get(int lev, double *L, double *W) {
    init_tables();
    const om_rule *r = &g_rules[lev];
    for (int q = 0; q < r->n; ++q) {
        for (int c = 0; c < 3; ++c) L[3 * q + c] = r->L[q][c];
        W[q] = r->W[q];
    }
    return r->n;
}

/* ---- kernels (src/utils/kernels.hpp) ---- */

/* accum_kernels, src/utils/kernels.hpp:29-40 */
static inline void accum_kernels(const double *r, double w, const double *m, double *SL,
                                 double *DL) {
    double r2 = v_dot(r, r), rn = sqrt(r2), ir = 1 / rn, ir3 = ir / r2;
    double d = -6 * w * v_dot(r, m) * ir3 / r2, s = w * ir3;
    for (int i = 0; i < 3; ++i) {
        SL[4 * i] += w * ir;
        for (int j = 0; j < 3; ++j) {
            SL[3 * i + j] += s * r[i] * r[j];
            DL[3 * i + j] += d * r[i] * r[j];
        }
    }
}

/* self_single_layer, src/utils/kernels.hpp:50-65: three Duffy-mapped sub-triangles
 * (x0, v_e, v_e+1), 10x10 Gauss-Legendre each. */
static void self_single_layer(const double *V, const double *x0, double *SL) {
    for (int k = 0; k < 9; ++k) SL[k] = 0;
    double dummy[9];
    const double zero[3] = {0, 0, 0};
    for (int e = 0; e < 3; ++e) {
        double e1[3], e2[3], cr[3];
        v_sub(V + 3 * e, x0, e1);
        v_sub(V + 3 * ((e + 1) % 3), x0, e2);
        v_cross(e1, e2, cr);
        double J = v_norm(cr);
        for (int a = 0; a < 10; ++a)
            for (int b = 0; b < 10; ++b) {
                double u = g_gl_x[a], v = g_gl_x[b], r[3];
                for (int c = 0; c < 3; ++c) r[c] = u * ((1 - v) * e1[c] + v * e2[c]);
                accum_kernels(r, g_gl_w[a] * g_gl_w[b] * u * J, zero, SL, dummy);
            }
    }
}

/* rule level from the distance ratio, src/utils/kernels.hpp:74-75 */
static inline int rule_level(const om_mesh *M, int p, const double *x0, double *dout) {
    double dv[3];
    v_sub(M->centroid + 3 * p, x0, dv);
    double d = v_norm(dv) / M->h[p];
    if (dout) *dout = d;
    return d < 0.75 ? 3 : (d < 1.5 ? 2 : (d < 3.0 ? 1 : 0));
}

/* panel_integrals, src/utils/kernels.hpp:67-82 */
void om_panel_integrals(const om_mesh *M, int p, const double *x0, int self, double *SL,
                        double *DL) {
    init_tables();
    for (int k = 0; k < 9; ++k) SL[k] = DL[k] = 0;
    if (self) {
        self_single_layer(M->V + 9 * p, x0, SL);
        return;
    }
    int lev = rule_level(M, p, x0, NULL);
    const om_rule *R = &g_rules[lev];
    const double *V = M->V + 9 * p;
    for (int q = 0; q < R->n; ++q) {
        double x[3], r[3];
        for (int c = 0; c < 3; ++c)
            x[c] = (R->L[q][0] * V[c] + R->L[q][1] * V[3 + c]) + R->L[q][2] * V[6 + c];
        v_sub(x, x0, r);
        accum_kernels(r, R->W[q] * M->area[p], M->normal + 3 * p, SL, DL);
    }
}

/* Rule level and distance ratio for every (target, panel) pair; lev = 4 marks a self pair.
 * Output is target-major: index i*np + p. */
void om_levels(const om_mesh *M, int p0, int np, const double *x0, int nt, const int *self_src,
               signed char *lev, double *dist) {
    for (int i = 0; i < nt; ++i)
        for (int p = 0; p < np; ++p) {
            double d;
            int l = rule_level(M, p0 + p, x0 + 3 * i, &d);
            if (self_src && self_src[i] == p0 + p) l = 4;
            lev[(size_t)i * np + p] = (signed char)l;
            if (dist) dist[(size_t)i * np + p] = d;
        }
}

/* One dense coupling block, as the loops of src/swimmer.hpp:52-57, 188-221 and
 * src/duct_flow.hpp:90-102 fill it: for target i and panel p0+p,
 *   outSL(3i+a, 3p+b) = SL_ab / sl_div          (column-major, leading dim ldSL)
 *   outDL(3i+a, 3p+b) = -DL_ab / dl_div
 *   dlsum[9i+k]     += DL_k                     (raw DL, summed over the panels)
 * Any output may be NULL. self_src[i] is the panel that contains target i, or -1. */
void om_block(const om_mesh *M, int p0, int np, const double *x0, int nt, const int *self_src,
              double sl_div, double *outSL, long ldSL, double dl_div, double *outDL, long ldDL,
              double *dlsum) {
    init_tables();
    double SL[9], DL[9];
    for (int i = 0; i < nt; ++i)
        for (int p = 0; p < np; ++p) {
            int self = self_src && self_src[i] == p0 + p;
            om_panel_integrals(M, p0 + p, x0 + 3 * i, self, SL, DL);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    size_t r = (size_t)(3 * i + a), c = (size_t)(3 * p + b);
                    if (outSL) outSL[r + c * (size_t)ldSL] = SL[3 * a + b] / sl_div;
                    if (outDL) outDL[r + c * (size_t)ldDL] = -DL[3 * a + b] / dl_div;
                }
            if (dlsum)
                for (int k = 0; k < 9; ++k) dlsum[9 * i + k] += DL[k];
        }
}

/* velocity_at, src/utils/field.hpp:18-34 */
void om_velocity_at(const om_mesh *M, const double *u, const double *f, double mu,
                    const double *pts, int npts, double *out) {
    init_tables();
    double SL[9], DL[9];
    for (int n = 0; n < npts; ++n) {
        double s[3] = {0, 0, 0};
        for (int p = 0; p < M->np; ++p) {
            om_panel_integrals(M, p, pts + 3 * n, 0, SL, DL);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    s[j] += SL[3 * i + j] * f[3 * p + i] / (OM_EIGHT_PI * mu) -
                            DL[3 * i + j] * u[3 * p + i] / OM_EIGHT_PI;
        }
        for (int j = 0; j < 3; ++j) out[3 * n + j] = s[j];
    }
}

/* squirmer_slip, src/utils/kernels.hpp:107-111 */
void om_squirmer_slip(const double *x, const double *center, const double *e, double B1,
                      double B2, double *out) {
    double rh[3];
    v_sub(x, center, rh);
    v_unit(rh);
    double ct = v_dot(rh, e), s = B1 + B2 * ct;
    for (int c = 0; c < 3; ++c) out[c] = s * (ct * rh[c] - e[c]);
}

/* accessors for the Python side */
int om_mesh_nn(const om_mesh *m) { return m->nn; }
int om_mesh_np(const om_mesh *m) { return m->np; }
void om_mesh_get(const om_mesh *m, double *nodes, int *tris, int *tags, double *V, double *normal,
                 double *centroid, double *area, double *h) {
    if (nodes) memcpy(nodes, m->nodes, sizeof(double) * 3 * m->nn);
    if (tris) memcpy(tris, m->tris, sizeof(int) * 3 * m->np);
    if (tags) memcpy(tags, m->tags, sizeof(int) * m->np);
    if (V) memcpy(V, m->V, sizeof(double) * 9 * m->np);
    if (normal) memcpy(normal, m->normal, sizeof(double) * 3 * m->np);
    if (centroid) memcpy(centroid, m->centroid, sizeof(double) * 3 * m->np);
    if (area) memcpy(area, m->area, sizeof(double) * m->np);
    if (h) memcpy(h, m->h, sizeof(double) * m->np);
}
