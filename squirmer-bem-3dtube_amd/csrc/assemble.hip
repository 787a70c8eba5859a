// assemble.hip -- panel-integral assembly of the Stokeslet/stresslet blocks on gfx950.
//
// Replaces the host loops over panel_integrals (src/utils/kernels.hpp:67-82) in
// src/swimmer.hpp:52-57,188-221, src/duct_flow.hpp:90-102 and src/utils/field.hpp:18-34.
//
// Kernel structure (one thread per target, workgroup = 256 targets x 32 source panels):
//  * the 32 source panels' geometry and their 7 precomputed far-field quadrature points
//    are staged in LDS;
//  * far pairs (d = |c_p - x0|/h_p >= 3, not self) are integrated by the owning thread
//    with the 7-point rule and written immediately (3 contiguous rows per column, so a
//    wave writes 3 x 64 contiguous doubles per output column);
//  * near and self pairs (28/112/448-point rules, 300-point Duffy rule) are queued in a
//    per-thread bit mask and then integrated by the whole wave together, one pair at a
//    time, the quadrature points spread over the 64 lanes and summed with a butterfly
//    reduction. This keeps the lanes busy whatever the mix of rule sizes.
//  * the rule distance d is evaluated without contraction, so the rule choice equals the
//    host computation of the same expression bit for bit.
//  * row sums of the stresslet integral (the "completed double layer" diagonal) are kept
//    per (target, source tile) and reduced over tiles in a fixed order: deterministic.
#include <cmath>
#include <vector>

#include "../../include/utils/quadrature.hpp"
#include "assemble.hpp"
#include "common.hpp"
#include "mesh_handle.hpp"

namespace sq {

// rules()[1..3] concatenated (28 + 112 + 448 points) as (L0, L1, L2, W); the 7-point rule;
// 10-point Gauss-Legendre on [0, 1]
__constant__ double c_rule[588 * 4];
__constant__ double c_rule7[7 * 4];
__constant__ double c_glx[10];
__constant__ double c_glw[10];
__constant__ int c_rule_off[4] = {0, 0, 28, 140};
__constant__ int c_rule_n[4] = {7, 28, 112, 448};

void upload_tables() {
    const auto& R = bem::rules();
    std::vector<double> r(588 * 4), r7(7 * 4);
    int q = 0;
    for (int l = 1; l <= 3; ++l)
        for (size_t k = 0; k < R[l].W.size(); ++k, ++q) {
            for (int c = 0; c < 3; ++c) r[4 * q + c] = R[l].L[k][c];
            r[4 * q + 3] = R[l].W[k];
        }
    for (int k = 0; k < 7; ++k) {
        for (int c = 0; c < 3; ++c) r7[4 * k + c] = R[0].L[k][c];
        r7[4 * k + 3] = R[0].W[k];
    }
    std::vector<double> gx, gw;
    bem::gauss_legendre01(10, gx, gw);
    SQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_rule), r.data(), r.size() * sizeof(double)));
    SQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_rule7), r7.data(), r7.size() * sizeof(double)));
    SQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_glx), gx.data(), 10 * sizeof(double)));
    SQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_glw), gw.data(), 10 * sizeof(double)));
}

// ------------------------------------------------------------------ device mesh
DMesh::DMesh(const bem::Mesh& m) { upload(m); }

__global__ void far_points_kernel(int np, const double* __restrict__ V, const double* __restrict__ area,
                                  double* __restrict__ q7) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = V[(size_t)k * np + p];
    const double A = area[p];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const double l0 = c_rule7[4 * q], l1 = c_rule7[4 * q + 1], l2 = c_rule7[4 * q + 2];
#pragma unroll
        for (int c = 0; c < 3; ++c) q7[(size_t)(4 * q + c) * np + p] = l0 * v[c] + l1 * v[3 + c] + l2 * v[6 + c];
        q7[(size_t)(4 * q + 3) * np + p] = c_rule7[4 * q + 3] * A;
    }
}

void DMesh::upload(const bem::Mesh& m) {
    np = m.P();
    std::vector<double> hv((size_t)9 * np), hn((size_t)3 * np), hc((size_t)3 * np), ha(np), hh(np), cen_aos(3 * (size_t)np),
        nrm_aos(3 * (size_t)np);
    for (int p = 0; p < np; ++p) {
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) hv[(size_t)(3 * k + c) * np + p] = m.V[p][k][c];
        for (int c = 0; c < 3; ++c) {
            hn[(size_t)c * np + p] = m.normal[p][c];
            hc[(size_t)c * np + p] = m.centroid[p][c];
            cen_aos[3 * (size_t)p + c] = m.centroid[p][c];
            nrm_aos[3 * (size_t)p + c] = m.normal[p][c];
        }
        ha[p] = m.area[p];
        hh[p] = m.h[p];
    }
    V.alloc(hv.size()); V.upload(hv.data(), hv.size());
    nrm.alloc(hn.size()); nrm.upload(hn.data(), hn.size());
    cen.alloc(hc.size()); cen.upload(hc.data(), hc.size());
    area.alloc(np); area.upload(ha.data(), np);
    h.alloc(np); h.upload(hh.data(), np);
    cen_xyz.alloc(cen_aos.size()); cen_xyz.upload(cen_aos.data(), cen_aos.size());
    nrm_xyz.alloc(nrm_aos.size()); nrm_xyz.upload(nrm_aos.data(), nrm_aos.size());
    q7.alloc((size_t)28 * np);
    hipLaunchKernelGGL(far_points_kernel, dim3(ceil_div(np, 256)), dim3(256), 0, ctx().s(), np, V.p, area.p, q7.p);
    check_launch("far_points");
    total_area = 0;
    for (int p = 0; p < np; ++p) total_area += m.area[p];
}

// ------------------------------------------------------------------ pair kernels
constexpr int TS = 32;   // source panels per workgroup tile
constexpr int TT = 256;  // targets per workgroup

struct Acc {
    double w0;                                      // sum of w/r (Stokeslet diagonal)
    double s00, s01, s02, s11, s12, s22;            // sum of w r_i r_j / r^3
    double d00, d01, d02, d11, d12, d22;            // sum of -6 w (r.m) r_i r_j / r^5
    __device__ void zero() { w0 = s00 = s01 = s02 = s11 = s12 = s22 = d00 = d01 = d02 = d11 = d12 = d22 = 0.0; }
};

// one quadrature point: accum_kernels, src/utils/kernels.hpp:29-40 (symmetric form)
__device__ __forceinline__ void accum_point(Acc& a, double rx, double ry, double rz, double w, double mx, double my,
                                            double mz, bool with_dl) {
    const double r2 = rx * rx + ry * ry + rz * rz;
    // 1/|r| by the FP64 reciprocal square root (within an ulp) and its powers by products: a
    // square root and three divisions per point cost as much as the rest of the point
    const double ir = rsqrt(r2);
    const double ir2 = ir * ir;
    const double ir3 = ir * ir2;
    const double s = w * ir3;
    a.w0 += w * ir;
    const double sx = s * rx, sy = s * ry, sz = s * rz;
    a.s00 += sx * rx; a.s01 += sx * ry; a.s02 += sx * rz;
    a.s11 += sy * ry; a.s12 += sy * rz; a.s22 += sz * rz;
    if (with_dl) {
        const double dd = -6.0 * w * (rx * mx + ry * my + rz * mz) * ir3 * ir2;
        const double dx = dd * rx, dy = dd * ry, dz = dd * rz;
        a.d00 += dx * rx; a.d01 += dx * ry; a.d02 += dx * rz;
        a.d11 += dy * ry; a.d12 += dy * rz; a.d22 += dz * rz;
    }
}

// rule level from d = |c - x0| / h, evaluated with contraction off so it matches the
// host evaluation of the same expression exactly (src/utils/kernels.hpp:74-75)
__device__ __attribute__((noinline)) int rule_level(double cx, double cy, double cz, double x, double y, double z,
                                                     double h) {
#pragma clang fp contract(off)
    const double dx = cx - x, dy = cy - y, dz = cz - z;
    const double d = sqrt(dx * dx + dy * dy + dz * dz) / h;
    return d < 0.75 ? 3 : (d < 1.5 ? 2 : (d < 3.0 ? 1 : 0));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

struct SrcTile {
    double q7[TS][28];   // x, y, z, w*A for the 7 far-field points
    double V[TS][9];
    double cen[TS][3];
    double nrm[TS][3];
    double area[TS];
    double h[TS];
};

__device__ __forceinline__ void load_tile(SrcTile& t, const AsmArgs& a, int pt0, int ns) {
    const int np = a.src_np;
    for (int e = threadIdx.x; e < ns * 28; e += blockDim.x) {
        int s = e % ns, k = e / ns;
        t.q7[s][k] = a.q7[(size_t)k * np + pt0 + s];
    }
    for (int e = threadIdx.x; e < ns * 9; e += blockDim.x) {
        int s = e % ns, k = e / ns;
        t.V[s][k] = a.V[(size_t)k * np + pt0 + s];
    }
    for (int e = threadIdx.x; e < ns * 3; e += blockDim.x) {
        int s = e % ns, k = e / ns;
        t.cen[s][k] = a.cen[(size_t)k * np + pt0 + s];
        t.nrm[s][k] = a.nrm[(size_t)k * np + pt0 + s];
    }
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        t.area[s] = a.area[pt0 + s];
        t.h[s] = a.h[pt0 + s];
    }
}

// Store one pair's 3x3 blocks (MODE_BLOCKS) or fold it into the velocity (MODE_FIELD).
template <int MODE>
__device__ __forceinline__ void consume(const AsmArgs& a, const Acc& q, int i, int pcol, int pabs, double* dls,
                                        double* vel, const double* tn, const SrcTile& t, int s) {
    const double SLm[3][3] = {{q.w0 + q.s00, q.s01, q.s02}, {q.s01, q.w0 + q.s11, q.s12}, {q.s02, q.s12, q.w0 + q.s22}};
    const double DLm[3][3] = {{q.d00, q.d01, q.d02}, {q.d01, q.d11, q.d12}, {q.d02, q.d12, q.d22}};
    const double isl = 1.0 / a.sl_div, idl = 1.0 / a.dl_div;  // 1/(8 pi mu), 1/(8 pi)
    if (MODE == MODE_BLOCKS) {
        if (a.SL) {
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double* col = a.SL + (size_t)(3 * pcol + b) * a.ld_sl + 3 * (size_t)i;
                float* col32 = a.SL32 ? a.SL32 + (size_t)(3 * pcol + b) * a.ld_sl32 + 3 * (size_t)i : nullptr;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    double v = SLm[r][b] * isl;
                    if (a.defl_sc != 0.0) v += a.defl_sc * tn[r] * t.nrm[s][b] * t.area[s];
                    col[r] = v;
                    if (col32) col32[r] = (float)v;
                }
            }
        }
        if (a.DL) {
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double* col = a.DL + (size_t)(3 * pcol + b) * a.ld_dl + 3 * (size_t)i;
#pragma unroll
                for (int r = 0; r < 3; ++r) col[r] = -DLm[r][b] * idl;
            }
        }
        if (dls) {
            dls[0] += DLm[0][0]; dls[1] += DLm[0][1]; dls[2] += DLm[0][2];
            dls[3] += DLm[1][0]; dls[4] += DLm[1][1]; dls[5] += DLm[1][2];
            dls[6] += DLm[2][0]; dls[7] += DLm[2][1]; dls[8] += DLm[2][2];
        }
    } else {
        // u_j += SL_ij f_i / (8 pi mu) - DL_ij u_i / (8 pi), src/utils/field.hpp:27-30
        const double* f = a.fld_f + 3 * (size_t)pabs;
        const double* u = a.fld_u + 3 * (size_t)pabs;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int ii = 0; ii < 3; ++ii) acc += SLm[ii][j] * f[ii] * isl - DLm[ii][j] * u[ii] * idl;
            vel[j] += acc;
        }
    }
}

// One workgroup's share of an assembly: targets [bx TT, bx TT + TT) against source tile
// `tile` (TS panels from p0 + tile TS), staged in the LDS tile t.
template <int MODE>
__device__ __forceinline__ void assemble_tile(const AsmArgs& a, SrcTile& t, int bx, int tile) {
    const int pt0 = a.p0 + tile * TS;
    const int ns = min(TS, a.p0 + a.np - pt0);
    const int i = bx * TT + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < a.nt;
    load_tile(t, a, pt0, ns);
    __syncthreads();

    double x0 = 0, y0 = 0, z0 = 0;
    int self = -1;
    double tn[3] = {0, 0, 0};
    if (valid) {
        x0 = a.x0[3 * (size_t)i];
        y0 = a.x0[3 * (size_t)i + 1];
        z0 = a.x0[3 * (size_t)i + 2];
        if (a.self_src) self = a.self_src[i];
        if (a.defl_sc != 0.0) {
            tn[0] = a.defl_tn[3 * (size_t)i];
            tn[1] = a.defl_tn[3 * (size_t)i + 1];
            tn[2] = a.defl_tn[3 * (size_t)i + 2];
        }
    }
    const bool want_dl = a.DL || a.dls_part || MODE == MODE_FIELD;
    double dls[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double vel[3] = {0, 0, 0};
    double* dlsp = (MODE == MODE_BLOCKS && a.dls_part) ? dls : nullptr;
    unsigned near_mask = 0u;

    // far pairs: 7-point rule, owner thread
    if (valid) {
        for (int s = 0; s < ns; ++s) {
            const int p = pt0 + s;
            const int lev = (p == self) ? 4 : rule_level(t.cen[s][0], t.cen[s][1], t.cen[s][2], x0, y0, z0, t.h[s]);
            if (lev != 0) {
                near_mask |= 1u << s;
                continue;
            }
            Acc q;
            q.zero();
#pragma unroll
            for (int k = 0; k < 7; ++k)
                accum_point(q, t.q7[s][4 * k] - x0, t.q7[s][4 * k + 1] - y0, t.q7[s][4 * k + 2] - z0, t.q7[s][4 * k + 3],
                            t.nrm[s][0], t.nrm[s][1], t.nrm[s][2], want_dl);
            consume<MODE>(a, q, i, p - a.p0, p, dlsp, vel, tn, t, s);
        }
    }

    // near and self pairs: the wave integrates one queued pair at a time
    for (;;) {
        const unsigned long long pending = __ballot(near_mask != 0u);
        if (pending == 0ull) break;
        const int owner = __ffsll((long long)pending) - 1;
        const int s = __shfl(__ffs((int)near_mask) - 1, owner);
        if (lane == owner) near_mask &= ~(1u << s);
        const double ox = __shfl(x0, owner), oy = __shfl(y0, owner), oz = __shfl(z0, owner);
        const int p = pt0 + s;
        const bool is_self = __shfl(self, owner) == p;
        Acc q;
        q.zero();
        if (is_self) {
            // Duffy rule, src/utils/kernels.hpp:50-65: sub-triangles (x0, v_e, v_e+1)
            for (int k = lane; k < 300; k += 64) {
                const int e = k / 100, ia = (k / 10) % 10, ib = k % 10;
                const int e2i = (e + 1) % 3;
                const double e1x = t.V[s][3 * e] - ox, e1y = t.V[s][3 * e + 1] - oy, e1z = t.V[s][3 * e + 2] - oz;
                const double e2x = t.V[s][3 * e2i] - ox, e2y = t.V[s][3 * e2i + 1] - oy, e2z = t.V[s][3 * e2i + 2] - oz;
                const double cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
                const double J = sqrt(cx * cx + cy * cy + cz * cz);
                const double u = c_glx[ia], v = c_glx[ib];
                const double rx = u * ((1 - v) * e1x + v * e2x);
                const double ry = u * ((1 - v) * e1y + v * e2y);
                const double rz = u * ((1 - v) * e1z + v * e2z);
                accum_point(q, rx, ry, rz, c_glw[ia] * c_glw[ib] * u * J, 0.0, 0.0, 0.0, false);
            }
        } else {
            const int lev = rule_level(t.cen[s][0], t.cen[s][1], t.cen[s][2], ox, oy, oz, t.h[s]);
            const int off = c_rule_off[lev], n = c_rule_n[lev];
            const double A = t.area[s];
            for (int k = lane; k < n; k += 64) {
                const double* r = c_rule + 4 * (off + k);
                const double px = r[0] * t.V[s][0] + r[1] * t.V[s][3] + r[2] * t.V[s][6];
                const double py = r[0] * t.V[s][1] + r[1] * t.V[s][4] + r[2] * t.V[s][7];
                const double pz = r[0] * t.V[s][2] + r[1] * t.V[s][5] + r[2] * t.V[s][8];
                accum_point(q, px - ox, py - oy, pz - oz, r[3] * A, t.nrm[s][0], t.nrm[s][1], t.nrm[s][2], want_dl);
            }
        }
        q.w0 = wave_sum(q.w0);
        q.s00 = wave_sum(q.s00); q.s01 = wave_sum(q.s01); q.s02 = wave_sum(q.s02);
        q.s11 = wave_sum(q.s11); q.s12 = wave_sum(q.s12); q.s22 = wave_sum(q.s22);
        if (want_dl && !is_self) {
            q.d00 = wave_sum(q.d00); q.d01 = wave_sum(q.d01); q.d02 = wave_sum(q.d02);
            q.d11 = wave_sum(q.d11); q.d12 = wave_sum(q.d12); q.d22 = wave_sum(q.d22);
        }
        if (lane == owner) consume<MODE>(a, q, i, p - a.p0, p, dlsp, vel, tn, t, s);
    }

    if (!valid) return;
    if (MODE == MODE_BLOCKS && a.dls_part) {
#pragma unroll
        for (int k = 0; k < 9; ++k) a.dls_part[((size_t)tile * 9 + k) * a.nt + i] = dls[k];
    }
    if (MODE == MODE_FIELD) {
#pragma unroll
        for (int k = 0; k < 3; ++k) a.vel_part[((size_t)tile * 3 + k) * a.nt + i] = vel[k];
    }
}

template <int MODE>
__global__ void __launch_bounds__(TT) assemble_kernel(AsmArgs a) {
    __shared__ SrcTile t;
    assemble_tile<MODE>(a, t, blockIdx.x, blockIdx.y);
}

// Persistent assembly of many jobs (target set x source panel range, each with its own
// outputs) in one launch. Work item w in [0, total) is tile w - start[j] of the job j with
// start[j] <= w < start[j + 1], taken target-tile-fastest; workgroups draw items from an
// atomic queue (counter zeroed before the launch), so the costly near-field tiles of one
// job do not hold up the others. Every workgroup leaves once the queue is drained.
__global__ void __launch_bounds__(TT) assemble_batch_kernel(const AsmArgs* __restrict__ jobs,
                                                            const long* __restrict__ start, int njobs,
                                                            unsigned long long* queue) {
    __shared__ SrcTile t;
    __shared__ long s_w;
    const long total = start[njobs];
    for (;;) {
        __syncthreads();  // the previous item's readers of t are done
        if (threadIdx.x == 0) s_w = (long)atomicAdd(queue, 1ull);
        __syncthreads();
        const long w = s_w;
        if (w >= total) break;
        int lo = 0, hi = njobs - 1;  // last job with start <= w
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= w) lo = mid;
            else hi = mid - 1;
        }
        const AsmArgs& a = jobs[lo];
        const long local = w - start[lo];
        const int gx = (a.nt + TT - 1) / TT;
        assemble_tile<MODE_BLOCKS>(a, t, (int)(local % gx), (int)(local / gx));
    }
}

// sum the per-tile partials in tile order: out[w*i + k] += sum_t part[(t*w + k)*nt + i]
__global__ void reduce_tiles_kernel(int nt, int ntiles, int width, const double* __restrict__ part, double* out,
                                    int accumulate) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    for (int k = 0; k < width; ++k) {
        double s = 0.0;
        for (int t = 0; t < ntiles; ++t) s += part[((size_t)t * width + k) * nt + i];
        double* o = out + (size_t)width * i + k;
        *o = accumulate ? *o + s : s;
    }
}

// ------------------------------------------------------------------ centre derivative
// Contraction of the translation derivative of the coupling blocks with an adjoint vector:
//   g_k(i) = sigma sum_p sum_q w_q [ a_sl lam_i . (d_k K_SL(r_q)) F_p + a_dl lam_i . (d_k K_DL(r_q)) V ]
// with r = y_q - x0 (source point minus target), K_SL(r) = I/|r| + r r^T/|r|^3 and
// K_DL(r) = -6 (r.m) r r^T/|r|^5 (the integrands of accum_kernels, src/utils/kernels.hpp:29-40).
// sigma = +1 when the source panels move with the sphere centre, -1 when the targets do.
// V is the source panel's vector (src_v) or the target's own (tgt_v). Quadrature rules and
// levels are those of the assembly, so this is the exact derivative of the assembled blocks
// (away from a change of rule level). Self pairs do not occur (tube-sphere couplings only).
struct AdjArgs {
    int nt;
    const double* x0;
    const double* lam;
    const double* tgt_v;
    int p0, np, src_np;
    const double *V, *nrm, *cen, *area, *h, *q7;
    const double* src_f;
    const double* src_v;
    double a_sl, a_dl, sigma;
    double* part;
};

// one quadrature point of the derivative contraction (r = y - x0, w = weight x area)
__device__ __forceinline__ void dpoint(double* g, double rx, double ry, double rz, double w, const double* L,
                                       const double* F, const double* Vv, double mx, double my, double mz,
                                       double a_sl, double a_dl) {
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double ir = rsqrt(r2), ir2 = ir * ir;
    const double ir3 = ir * ir2, ir5 = ir3 * ir2, ir7 = ir5 * ir2;
    const double rL = rx * L[0] + ry * L[1] + rz * L[2];
    const double rF = rx * F[0] + ry * F[1] + rz * F[2];
    const double LF = L[0] * F[0] + L[1] * F[1] + L[2] * F[2];
    const double rV = rx * Vv[0] + ry * Vv[1] + rz * Vv[2];
    const double rm = rx * mx + ry * my + rz * mz;
    const double r[3] = {rx, ry, rz}, m[3] = {mx, my, mz};
    // lam . d_k K_SL F = -(L.F) r_k/|r|^3 + (L_k (r.F) + F_k (r.L))/|r|^3 - 3 (r.L)(r.F) r_k/|r|^5
    // lam . d_k K_DL V = -6 (m_k (r.L)(r.V) + (r.m)(L_k (r.V) + V_k (r.L)))/|r|^5
    //                    + 30 (r.m)(r.L)(r.V) r_k/|r|^7
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double tsl = (-LF * r[k] + L[k] * rF + F[k] * rL) * ir3 - 3.0 * rL * rF * r[k] * ir5;
        const double tdl = -6.0 * (m[k] * rL * rV + rm * (L[k] * rV + Vv[k] * rL)) * ir5 + 30.0 * rm * rL * rV * r[k] * ir7;
        g[k] += w * (a_sl * tsl + a_dl * tdl);
    }
}

__global__ void __launch_bounds__(TT) adjoint_xc_kernel(AdjArgs a) {
    __shared__ SrcTile t;
    __shared__ double tf[TS][3], tv[TS][3];
    const int tile = blockIdx.y;
    const int pt0 = a.p0 + tile * TS;
    const int ns = min(TS, a.p0 + a.np - pt0);
    const int i = blockIdx.x * TT + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < a.nt;
    AsmArgs la{};
    la.src_np = a.src_np;
    la.V = a.V; la.nrm = a.nrm; la.cen = a.cen; la.area = a.area; la.h = a.h; la.q7 = a.q7;
    load_tile(t, la, pt0, ns);
    for (int e = threadIdx.x; e < ns * 3; e += blockDim.x) {
        const int s = e / 3, k = e % 3;
        tf[s][k] = a.src_f ? a.src_f[3 * (size_t)(pt0 + s) + k] : 0.0;
        tv[s][k] = a.src_v ? a.src_v[3 * (size_t)(pt0 + s) + k] : 0.0;
    }
    __syncthreads();
    double x0 = 0, y0 = 0, z0 = 0, L[3] = {0, 0, 0}, Vt[3] = {0, 0, 0};
    if (valid) {
        x0 = a.x0[3 * (size_t)i];
        y0 = a.x0[3 * (size_t)i + 1];
        z0 = a.x0[3 * (size_t)i + 2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            L[k] = a.lam[3 * (size_t)i + k];
            Vt[k] = a.tgt_v ? a.tgt_v[3 * (size_t)i + k] : 0.0;
        }
    }
    double g[3] = {0, 0, 0};
    unsigned near_mask = 0u;
    if (valid) {
        for (int s = 0; s < ns; ++s) {
            const int lev = rule_level(t.cen[s][0], t.cen[s][1], t.cen[s][2], x0, y0, z0, t.h[s]);
            if (lev != 0) {
                near_mask |= 1u << s;
                continue;
            }
            const double* Vv = a.tgt_v ? Vt : tv[s];
#pragma unroll
            for (int k = 0; k < 7; ++k)
                dpoint(g, t.q7[s][4 * k] - x0, t.q7[s][4 * k + 1] - y0, t.q7[s][4 * k + 2] - z0, t.q7[s][4 * k + 3], L,
                       tf[s], Vv, t.nrm[s][0], t.nrm[s][1], t.nrm[s][2], a.a_sl, a.a_dl);
        }
    }
    // near pairs: the wave integrates one queued pair at a time
    for (;;) {
        const unsigned long long pending = __ballot(near_mask != 0u);
        if (pending == 0ull) break;
        const int owner = __ffsll((long long)pending) - 1;
        const int s = __shfl(__ffs((int)near_mask) - 1, owner);
        if (lane == owner) near_mask &= ~(1u << s);
        const double ox = __shfl(x0, owner), oy = __shfl(y0, owner), oz = __shfl(z0, owner);
        double oL[3], oV[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            oL[k] = __shfl(L[k], owner);
            oV[k] = a.tgt_v ? __shfl(Vt[k], owner) : tv[s][k];
        }
        const int lev = rule_level(t.cen[s][0], t.cen[s][1], t.cen[s][2], ox, oy, oz, t.h[s]);
        const int off = c_rule_off[lev], n = c_rule_n[lev];
        const double A = t.area[s];
        double q[3] = {0, 0, 0};
        for (int k = lane; k < n; k += 64) {
            const double* r = c_rule + 4 * (off + k);
            const double px = r[0] * t.V[s][0] + r[1] * t.V[s][3] + r[2] * t.V[s][6];
            const double py = r[0] * t.V[s][1] + r[1] * t.V[s][4] + r[2] * t.V[s][7];
            const double pz = r[0] * t.V[s][2] + r[1] * t.V[s][5] + r[2] * t.V[s][8];
            dpoint(q, px - ox, py - oy, pz - oz, r[3] * A, oL, tf[s], oV, t.nrm[s][0], t.nrm[s][1], t.nrm[s][2], a.a_sl,
                   a.a_dl);
        }
        q[0] = wave_sum(q[0]);
        q[1] = wave_sum(q[1]);
        q[2] = wave_sum(q[2]);
        if (lane == owner) {
            g[0] += q[0];
            g[1] += q[1];
            g[2] += q[2];
        }
    }
    if (!valid) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) a.part[((size_t)tile * 3 + k) * a.nt + i] = a.sigma * g[k];
}

// ------------------------------------------------------------------ host side
static DBuf<double> g_part;

static double* part_buffer(size_t n) {
    if (g_part.n < n) g_part.alloc(std::max<size_t>(n, 1 << 20));
    return g_part.p;
}

void assemble(const DMesh& src, int p0, int np, const double* x0, int nt, const int* self_src, double sl_div,
              double* SL, long ld_sl, double dl_div, double* DL, long ld_dl, double* dlsum, double defl_sc,
              const double* defl_tn, float* SL32, long ld_sl32) {
    if (nt <= 0 || np <= 0) return;
    SQ_REQUIRE(p0 >= 0 && p0 + np <= src.np, SQBEM_ERR_ARG, "assemble: source panel range out of bounds");
    SQ_REQUIRE(!SL || ld_sl >= 3L * nt, SQBEM_ERR_ARG, "assemble: SL leading dimension too small");
    SQ_REQUIRE(!DL || ld_dl >= 3L * nt, SQBEM_ERR_ARG, "assemble: DL leading dimension too small");
    SQ_REQUIRE(!SL32 || (SL && ld_sl32 >= 3L * nt), SQBEM_ERR_ARG, "assemble: SL32 without SL, or its leading dimension too small");
    AsmArgs a{};
    a.nt = nt;
    a.x0 = x0;
    a.self_src = self_src;
    a.p0 = p0;
    a.np = np;
    a.src_np = src.np;
    a.V = src.V.p; a.nrm = src.nrm.p; a.cen = src.cen.p; a.area = src.area.p; a.h = src.h.p; a.q7 = src.q7.p;
    a.sl_div = sl_div; a.dl_div = dl_div;
    a.SL = SL; a.ld_sl = ld_sl; a.DL = DL; a.ld_dl = ld_dl;
    a.SL32 = SL32; a.ld_sl32 = ld_sl32;
    a.defl_sc = defl_sc; a.defl_tn = defl_tn;
    const int ntiles = (int)ceil_div(np, TS);
    if (dlsum) a.dls_part = part_buffer((size_t)ntiles * 9 * nt);
    dim3 grid(ceil_div(nt, TT), ntiles);
    // algorithmic bytes: the stored 3x3 blocks; flops: 72 per far-field point (7 per pair)
    ProfScope prof(PROF_ASSEMBLE, 72.0 * 7.0 * nt * np,
                   8.0 * 9.0 * nt * np * ((SL ? 1 : 0) + (DL ? 1 : 0)));
    hipLaunchKernelGGL(assemble_kernel<MODE_BLOCKS>, grid, dim3(TT), 0, ctx().s(), a);
    check_launch("assemble");
    if (dlsum) {
        hipLaunchKernelGGL(reduce_tiles_kernel, dim3(ceil_div(nt, 256)), dim3(256), 0, ctx().s(), nt, ntiles, 9,
                           a.dls_part, dlsum, 1);
        check_launch("reduce_tiles");
    }
}

int asm_tiles(int np) { return (int)ceil_div(np, TS); }

// out[9i+k] = (0 + sum_t part1) + sum_t part2: the order of two accumulating reduce_tiles
__global__ void reduce_dls_batch_kernel(const DlsReduce* __restrict__ jobs) {
    const DlsReduce& r = jobs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.nt) return;
    for (int k = 0; k < 9; ++k) {
        double s = 0.0;
        for (int t = 0; t < r.ntiles1; ++t) s += r.part1[((size_t)t * 9 + k) * r.nt + i];
        double o = 0.0 + s;
        if (r.part2) {
            double s2 = 0.0;
            for (int t = 0; t < r.ntiles2; ++t) s2 += r.part2[((size_t)t * 9 + k) * r.nt + i];
            o = o + s2;
        }
        r.out[9 * (size_t)i + k] = o;
    }
}

void assemble_batch_prepare(const std::vector<AsmJob>& jobs, AsmBatch& plan) {
    std::vector<AsmArgs> args;
    std::vector<long> start(1, 0);
    plan.flops = plan.bytes = 0;
    for (const AsmJob& j : jobs) {
        if (j.nt <= 0 || j.np <= 0) continue;
        SQ_REQUIRE(j.p0 >= 0 && j.p0 + j.np <= j.src->np, SQBEM_ERR_ARG, "assemble: source panel range out of bounds");
        SQ_REQUIRE(!j.SL || j.ld_sl >= 3L * j.nt, SQBEM_ERR_ARG, "assemble: SL leading dimension too small");
        SQ_REQUIRE(!j.DL || j.ld_dl >= 3L * j.nt, SQBEM_ERR_ARG, "assemble: DL leading dimension too small");
        AsmArgs a{};
        a.nt = j.nt;
        a.x0 = j.x0;
        a.self_src = j.self_src;
        a.p0 = j.p0;
        a.np = j.np;
        a.src_np = j.src->np;
        a.V = j.src->V.p; a.nrm = j.src->nrm.p; a.cen = j.src->cen.p; a.area = j.src->area.p; a.h = j.src->h.p;
        a.q7 = j.src->q7.p;
        a.sl_div = j.sl_div; a.dl_div = j.dl_div;
        a.SL = j.SL; a.ld_sl = j.ld_sl; a.DL = j.DL; a.ld_dl = j.ld_dl;
        a.dls_part = j.dls_part;
        a.defl_sc = j.defl_sc; a.defl_tn = j.defl_tn;
        args.push_back(a);
        start.push_back(start.back() + (long)ceil_div(j.nt, TT) * asm_tiles(j.np));
        plan.flops += 72.0 * 7.0 * j.nt * j.np;
        plan.bytes += 8.0 * 9.0 * j.nt * j.np * ((j.SL ? 1 : 0) + (j.DL ? 1 : 0));
    }
    plan.njobs = (int)args.size();
    plan.total = start.back();
    if (args.empty()) return;
    plan.jobs.alloc(args.size());
    plan.start.alloc(start.size());
    if (!plan.queue.p) plan.queue.alloc(1);
    plan.jobs.upload(args.data(), args.size());
    plan.start.upload(start.data(), start.size());
    sync();  // args and start are host staging
}

void assemble_batch_launch(const AsmBatch& plan) {
    if (!plan.njobs) return;
    fill_zero(plan.queue.p, sizeof(unsigned long long));  // a kernel: the launch may be captured
    const unsigned grid = (unsigned)std::min<long>(plan.total, 8L * std::max(ctx().cu_count, 1));
    ProfScope prof(PROF_ASSEMBLE, plan.flops, plan.bytes);
    hipLaunchKernelGGL(assemble_batch_kernel, dim3(grid), dim3(TT), 0, ctx().s(), plan.jobs.p, plan.start.p,
                       plan.njobs, plan.queue.p);
    check_launch("assemble_batch");
}

void assemble_batch(const std::vector<AsmJob>& jobs) {
    AsmBatch plan;
    assemble_batch_prepare(jobs, plan);
    assemble_batch_launch(plan);
    sync();  // the plan's tables are released on return
}

void reduce_dls_prepare(const std::vector<DlsReduce>& jobs, DlsBatch& plan) {
    plan.n = (int)jobs.size();
    plan.ntmax = 1;
    if (jobs.empty()) return;
    for (const DlsReduce& r : jobs) plan.ntmax = std::max(plan.ntmax, r.nt);
    plan.jobs.alloc(jobs.size());
    plan.jobs.upload(jobs.data(), jobs.size());
    sync();  // jobs is host staging
}

void reduce_dls_launch(const DlsBatch& plan) {
    if (!plan.n) return;
    hipLaunchKernelGGL(reduce_dls_batch_kernel, dim3(ceil_div(plan.ntmax, 256), (unsigned)plan.n), dim3(256), 0,
                       ctx().s(), plan.jobs.p);
    check_launch("reduce_dls_batch");
}

void reduce_dls_batch(const std::vector<DlsReduce>& jobs) {
    DlsBatch plan;
    reduce_dls_prepare(jobs, plan);
    reduce_dls_launch(plan);
    sync();
}

void field(const DMesh& src, const double* pts, int npts, const double* u, const double* f, double mu, double* out) {
    if (npts <= 0) return;
    AsmArgs a{};
    a.nt = npts;
    a.x0 = pts;
    a.p0 = 0;
    a.np = src.np;
    a.src_np = src.np;
    a.V = src.V.p; a.nrm = src.nrm.p; a.cen = src.cen.p; a.area = src.area.p; a.h = src.h.p; a.q7 = src.q7.p;
    a.sl_div = bem::EIGHT_PI * mu;
    a.dl_div = bem::EIGHT_PI;
    a.fld_u = u;
    a.fld_f = f;
    const int ntiles = (int)ceil_div(src.np, TS);
    a.vel_part = part_buffer((size_t)ntiles * 3 * npts);
    dim3 grid(ceil_div(npts, TT), ntiles);
    hipLaunchKernelGGL(assemble_kernel<MODE_FIELD>, grid, dim3(TT), 0, ctx().s(), a);
    check_launch("field");
    hipLaunchKernelGGL(reduce_tiles_kernel, dim3(ceil_div(npts, 256)), dim3(256), 0, ctx().s(), npts, ntiles, 3,
                       a.vel_part, out, 0);
    check_launch("reduce_tiles");
}

void adjoint_xc(const DMesh& src, const double* x0, int nt, const double* lam, const double* tgt_v,
                const double* src_f, const double* src_v, double a_sl, double a_dl, double sigma, double* out) {
    if (nt <= 0 || src.np <= 0) return;
    AdjArgs a{};
    a.nt = nt;
    a.x0 = x0;
    a.lam = lam;
    a.tgt_v = tgt_v;
    a.p0 = 0;
    a.np = src.np;
    a.src_np = src.np;
    a.V = src.V.p; a.nrm = src.nrm.p; a.cen = src.cen.p; a.area = src.area.p; a.h = src.h.p; a.q7 = src.q7.p;
    a.src_f = src_f;
    a.src_v = src_v;
    a.a_sl = a_sl;
    a.a_dl = a_dl;
    a.sigma = sigma;
    const int ntiles = (int)ceil_div(src.np, TS);
    a.part = part_buffer((size_t)ntiles * 3 * nt);
    dim3 grid(ceil_div(nt, TT), ntiles);
    hipLaunchKernelGGL(adjoint_xc_kernel, grid, dim3(TT), 0, ctx().s(), a);
    check_launch("adjoint_xc");
    hipLaunchKernelGGL(reduce_tiles_kernel, dim3(ceil_div(nt, 256)), dim3(256), 0, ctx().s(), nt, ntiles, 3, a.part,
                       out, 1);
    check_launch("reduce_tiles");
}

}  // namespace sq

// ------------------------------------------------------------------ C ABI
using namespace sq;

struct sqbem_dmesh {
    DMesh d;
};

extern "C" {

sqbem_dmesh* sqbem_dmesh_upload(const sqbem_mesh* m) {
    return guarded_ptr<sqbem_dmesh>([&] {
        SQ_REQUIRE(m != nullptr, SQBEM_ERR_ARG, "sqbem_dmesh_upload: null mesh");
        auto* d = new sqbem_dmesh;
        try {
            d->d.upload(m->m);
            sync();
        } catch (...) {
            delete d;
            throw;
        }
        return d;
    });
}

void sqbem_dmesh_free(sqbem_dmesh* d) { delete d; }

int sqbem_assemble(const sqbem_dmesh* src, int p0, int np, const double* x0, int nt, const int* self_src,
                   double sl_div, double* SL, long ld_sl, double dl_div, double* DL, long ld_dl, double* dlsum) {
    return guarded([&] {
        SQ_REQUIRE(src && x0, SQBEM_ERR_ARG, "sqbem_assemble: null argument");
        assemble(src->d, p0, np, x0, nt, self_src, sl_div, SL, ld_sl, dl_div, DL, ld_dl, dlsum, 0.0, nullptr);
    });
}

int sqbem_panel_integrals(const sqbem_mesh* m, int p, const double* x0, int self, double* SLo, double* DLo) {
    return guarded([&] {
        SQ_REQUIRE(m && x0 && SLo && DLo, SQBEM_ERR_ARG, "sqbem_panel_integrals: null argument");
        SQ_REQUIRE(p >= 0 && p < m->m.P(), SQBEM_ERR_ARG, "sqbem_panel_integrals: panel out of range");
        DMesh d(m->m);
        DBuf<double> dx(3), sl(9), dl(9);
        DBuf<int> ds(1);
        dx.upload(x0, 3);
        int si = self ? p : -1;
        ds.upload(&si, 1);
        assemble(d, p, 1, dx.p, 1, ds.p, 1.0, sl.p, 3, -1.0, dl.p, 3, nullptr, 0.0, nullptr);
        double a[9], b[9];
        sl.download(a, 9);
        dl.download(b, 9);
        // column-major (r + 3 c) -> row-major [3 r + c], as panel_integrals returns
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                SLo[3 * r + c] = a[r + 3 * c];
                DLo[3 * r + c] = b[r + 3 * c];
            }
    });
}

}  // extern "C"
