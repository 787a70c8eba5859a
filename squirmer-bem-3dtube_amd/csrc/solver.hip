// solver.hip -- device-resident solvers behind the C ABI: Tube (T factor), SwimmerSolver
// (coupling assembly, Schur complement, bordered force/torque-free solve, iterative
// refinement for the mixed-precision modes), DuctFlow and the field evaluation.
//
// Follows src/swimmer.hpp:33-289, src/duct_flow.hpp:22-109 and src/utils/field.hpp:18-34.
// Everything stays in HBM between calls; only slip fields, rigid-body velocities and
// tractions cross to the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/utils/kernels.hpp"
#include "assemble.hpp"
#include "common.hpp"
#include "dense.hpp"
#include "lu_handle.hpp"
#include "mesh_handle.hpp"
#include "solver.hpp"

namespace sq {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}


// ------------------------------------------------------------------ small kernels
__global__ void iota_kernel(int n, int* a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// AuBs(3i+a, 3i+b) += dlsum_i[3a+b] / (8 pi) - (unbounded ? delta_ab : 0), src/swimmer.hpp:218-220
__global__ void completion_diag_kernel(int Ns, double* A, long lda, const double* __restrict__ dls, double eight_pi,
                                       double minus_identity) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Ns) return;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            A[(3 * (size_t)i + a) + (3 * (size_t)i + b) * lda] += dls[9 * (size_t)i + 3 * a + b] / eight_pi -
                                                                  (a == b ? minus_identity : 0.0);
}

// C(row, 0..5) = sum_col A(row, col) K(col, 0..5) for the rigid-body map K of
// src/swimmer.hpp:223-231 (u = U + Omega x r); partial sums over panel chunks.
constexpr int KP_CHUNKS = 16;
__global__ void kproj_kernel(int nrows, int Ns, const double* __restrict__ A, long lda, const double* __restrict__ cen,
                             double xcx, double xcy, double xcz, double* __restrict__ part) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    int chunk = blockIdx.y;
    if (row >= nrows) return;
    int per = (Ns + KP_CHUNKS - 1) / KP_CHUNKS;
    int pa = chunk * per, pb = min(Ns, pa + per);
    double c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
    for (int p = pa; p < pb; ++p) {
        const double a0 = A[row + (size_t)(3 * p) * lda];
        const double a1 = A[row + (size_t)(3 * p + 1) * lda];
        const double a2 = A[row + (size_t)(3 * p + 2) * lda];
        const double rx = cen[3 * (size_t)p] - xcx, ry = cen[3 * (size_t)p + 1] - xcy, rz = cen[3 * (size_t)p + 2] - xcz;
        c0 += a0;
        c1 += a1;
        c2 += a2;
        c3 += -a1 * rz + a2 * ry;
        c4 += a0 * rz - a2 * rx;
        c5 += -a0 * ry + a1 * rx;
    }
    double* o = part + (size_t)chunk * 6 * nrows + row;
    o[0] = c0; o[(size_t)nrows] = c1; o[2 * (size_t)nrows] = c2;
    o[3 * (size_t)nrows] = c3; o[4 * (size_t)nrows] = c4; o[5 * (size_t)nrows] = c5;
}
__global__ void kproj_reduce(int nrows, const double* __restrict__ part, double* C, long ldc) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    for (int c = 0; c < 6; ++c) {
        double s = 0;
        for (int k = 0; k < KP_CHUNKS; ++k) s += part[((size_t)k * 6 + c) * nrows + row];
        C[row + (size_t)c * ldc] = s;
    }
}

// max |A(i,j)| over an m x n block (non-negative doubles order like their bit patterns)
template <typename T>
__global__ void absmax_kernel_t(int m, int n, const T* __restrict__ A, long lda, unsigned long long* out) {
    // workgroups stride over columns, threads over the rows of a column (no divisions)
    double mx = 0.0;
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        const T* col = A + (size_t)j * lda;
        for (int i = threadIdx.x; i < m; i += blockDim.x) mx = fmax(mx, fabs((double)col[i]));
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

// bordered matrix of solve_free, src/swimmer.hpp:107-125: copy [Gs | Gc] (ns x (ns + 6))
// into the first ns rows, zero the last 6 rows, then the scaled force/torque rows. The copy
// can also find max |Gs| (mx, as absmax_kernel_t), which sets the rows' scale: one pass over
// G instead of two. Workgroups stride over columns, threads over rows, four loads in flight.
template <typename T>
__global__ void border_copy_kernel(int ns, const T* __restrict__ G, T* __restrict__ M, unsigned long long* mx) {
    const int n = ns + 6;
    double m = 0.0;
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        const T* g = G + (size_t)j * ns;
        T* out = M + (size_t)j * n;
        for (int i0 = threadIdx.x; i0 < n; i0 += 4 * 256) {
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = i0 + 256 * u;
                v[u] = i < ns ? g[i] : T(0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = i0 + 256 * u;
                if (i < n) out[i] = v[u];
                if (j < ns) m = fmax(m, fabs((double)v[u]));
            }
        }
    }
    if (mx) {
        for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
        if ((threadIdx.x & 63) == 0) atomicMax(mx, (unsigned long long)__double_as_longlong(m));
    }
}
template <typename T>
__global__ void border_rows_kernel(int Ns, T* M, const double* __restrict__ cen, const double* __restrict__ area,
                                   double sc, const double* scp, double xcx, double xcy, double xcz) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Ns) return;
    const size_t ns = 3 * (size_t)Ns, n = ns + 6;
    const double rx = cen[3 * (size_t)p] - xcx, ry = cen[3 * (size_t)p + 1] - xcy, rz = cen[3 * (size_t)p + 2] - xcz;
    const double A = area[p] * (scp ? *scp : sc);  // the scale from device memory when given
    auto at = [&](size_t r, size_t c) -> T& { return M[r + c * n]; };
    for (int a = 0; a < 3; ++a) at(ns + a, 3 * p + a) = (T)A;
    at(ns + 3, 3 * p + 1) = (T)(-A * rz); at(ns + 3, 3 * p + 2) = (T)(A * ry);
    at(ns + 4, 3 * p + 0) = (T)(A * rz);  at(ns + 4, 3 * p + 2) = (T)(-A * rx);
    at(ns + 5, 3 * p + 0) = (T)(-A * ry); at(ns + 5, 3 * p + 1) = (T)(A * rx);
}

// res_c(a, k) = -sc sum_p F rows . f_s (constraint residual of the bordered system)
__global__ void constraint_apply_kernel(int Ns, int nrhs, const double* __restrict__ fs, long ldf,
                                        const double* __restrict__ cen, const double* __restrict__ area, double sc,
                                        double xcx, double xcy, double xcz, double sign, double* out, long ldo) {
    // one workgroup per right-hand side, 256 threads
    const int k = blockIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int p = threadIdx.x; p < Ns; p += blockDim.x) {
        const double rx = cen[3 * (size_t)p] - xcx, ry = cen[3 * (size_t)p + 1] - xcy, rz = cen[3 * (size_t)p + 2] - xcz;
        const double A = area[p] * sc;
        const double f0 = fs[3 * (size_t)p + k * ldf], f1 = fs[3 * (size_t)p + 1 + k * ldf], f2 = fs[3 * (size_t)p + 2 + k * ldf];
        acc[0] += A * f0; acc[1] += A * f1; acc[2] += A * f2;
        acc[3] += -A * rz * f1 + A * ry * f2;
        acc[4] += A * rz * f0 - A * rx * f2;
        acc[5] += -A * ry * f0 + A * rx * f1;
    }
    __shared__ double red[6][256];
    for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x + ldo * k] = sign * red[threadIdx.x][0];
}

// exact widening of a matrix entry to FP64 (BF16 through FP32)
__device__ __forceinline__ double widen(double v) { return v; }
__device__ __forceinline__ double widen(float v) { return (double)v; }
__device__ __forceinline__ double widen(__bf16 v) { return (double)(float)v; }

// y = alpha A x + beta y for a few right-hand sides (nrhs <= 8), split over column chunks
constexpr int GV_CHUNKS = 32;
template <typename T, typename TX = T>
__global__ void gemv_part_kernel(int m, int n, int nrhs, const T* __restrict__ A, long lda, const TX* __restrict__ X,
                                 long ldx, double* __restrict__ part) {
    // rows row0 and row0 + 256 of a 512-row block (a column's 4 KB run per workgroup)
    const int row0 = blockIdx.x * 512 + threadIdx.x;
    int chunk = blockIdx.y;
    if (row0 >= m) return;
    const bool two = row0 + 256 < m;
    const int row1 = two ? row0 + 256 : row0;
    int per = (n + GV_CHUNKS - 1) / GV_CHUNKS;
    int ja = chunk * per, jb = min(n, ja + per);
    double acc[2][8] = {{0, 0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0, 0}};
    int j = ja;
    // eight columns a round, their loads issued together (HBM needs many bytes in flight);
    // every row still sums its columns in order
    for (; j + 8 <= jb; j += 8) {
        double a0[8], a1[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a0[u] = widen(A[row0 + (size_t)(j + u) * lda]);
            a1[u] = widen(A[row1 + (size_t)(j + u) * lda]);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (r < nrhs) {
                    const double x = (double)X[j + u + (size_t)r * ldx];
                    acc[0][r] += a0[u] * x;
                    acc[1][r] += a1[u] * x;
                }
    }
    for (; j < jb; ++j) {
        const double a0 = widen(A[row0 + (size_t)j * lda]), a1 = widen(A[row1 + (size_t)j * lda]);
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if (r < nrhs) {
                const double x = (double)X[j + (size_t)r * ldx];
                acc[0][r] += a0 * x;
                acc[1][r] += a1 * x;
            }
    }
    for (int r = 0; r < nrhs; ++r) {
        part[((size_t)chunk * nrhs + r) * m + row0] = acc[0][r];
        if (two) part[((size_t)chunk * nrhs + r) * m + row1] = acc[1][r];
    }
}
// right-hand side blockIdx.y: the chunks' partial sums of a row, all loads issued before the
// (in-order) sum
template <typename T>
__global__ void gemv_reduce_kernel(int m, int nrhs, const double* __restrict__ part, double alpha, double beta, T* Y,
                                   long ldy) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (row >= m || r >= nrhs) return;
    double v[GV_CHUNKS];
#pragma unroll
    for (int k = 0; k < GV_CHUNKS; ++k) v[k] = part[((size_t)k * nrhs + r) * m + row];
    double s = 0;
#pragma unroll
    for (int k = 0; k < GV_CHUNKS; ++k) s += v[k];
    T* y = Y + row + (size_t)r * ldy;
    *y = (T)(beta == 0.0 ? alpha * s : alpha * s + beta * (double)*y);
}

// partial sums of the column chunks, one buffer per stream (products may run on both at once)
static DBuf<double> g_gv_main, g_gv_side;
template <typename T>
void gemv(int m, int n, int nrhs, double alpha, const T* A, long lda, const T* X, long ldx, double beta, T* Y,
                 long ldy) {
    if (m <= 0 || nrhs <= 0) return;
    SQ_REQUIRE(nrhs <= 8, SQBEM_ERR_ARG, "gemv: at most 8 right-hand sides");
    size_t need = (size_t)GV_CHUNKS * nrhs * m;
    DBuf<double>& g_gv = ctx().redirect ? g_gv_side : g_gv_main;
    if (g_gv.n < need) g_gv.alloc(std::max<size_t>(need, 1 << 20));
    if (n <= 0) {
        scale_matrix<T>(m, nrhs, (T)beta, Y, (int)ldy);
        return;
    }
    ProfScope prof(PROF_GEMV, 2.0 * m * n * nrhs, sizeof(T) * ((double)m * n + (double)n * nrhs + 2.0 * m * nrhs));
    hipLaunchKernelGGL(gemv_part_kernel<T>, dim3(ceil_div(m, 512), GV_CHUNKS), dim3(256), 0, ctx().s(), m, n, nrhs, A,
                       lda, X, ldx, g_gv.p);
    hipLaunchKernelGGL(gemv_reduce_kernel<T>, dim3(ceil_div(m, 256), nrhs), dim3(256), 0, ctx().s(), m, nrhs, g_gv.p, alpha,
                       beta, Y, ldy);
    check_launch("gemv");
}

// The same product with a BF16 A (column-major) and FP32 x, y: the refinement's corrections
// multiply by BF16 copies of D and Y (half the bytes of their FP32 forms; the corrections'
// accuracy is set by the 16-bit factorisations, not by these products).
void gemv_cols16(int m, int n, int nrhs, double alpha, const __bf16* A, long lda, const float* X, long ldx, double beta,
                 float* Y, long ldy) {
    if (m <= 0 || nrhs <= 0) return;
    SQ_REQUIRE(nrhs <= 8 && n > 0, SQBEM_ERR_ARG, "gemv_cols16: 1 to 8 right-hand sides, n > 0");
    size_t need = (size_t)GV_CHUNKS * nrhs * m;
    DBuf<double>& g_gv = ctx().redirect ? g_gv_side : g_gv_main;
    if (g_gv.n < need) g_gv.alloc(std::max<size_t>(need, 1 << 20));
    ProfScope prof(PROF_GEMV, 2.0 * m * n * nrhs, 2.0 * m * n + 4.0 * ((double)n * nrhs + 2.0 * m * nrhs));
    hipLaunchKernelGGL((gemv_part_kernel<__bf16, float>), dim3(ceil_div(m, 512), GV_CHUNKS), dim3(256), 0, ctx().s(),
                       m, n, nrhs, A, lda, X, ldx, g_gv.p);
    hipLaunchKernelGGL(gemv_reduce_kernel<float>, dim3(ceil_div(m, 256), nrhs), dim3(256), 0, ctx().s(), m, nrhs, g_gv.p,
                       alpha, beta, Y, ldy);
    check_launch("gemv_cols16");
}

// y(i, r) = alpha sum_j A(i, j) x(j, r) + beta y(i, r) for a row-major BF16 A (row i at
// A + i lda, lda a multiple of 8, A 16-byte aligned): one wave per row, 16-byte loads of 8
// entries a lane, FP64 sums (nrhs <= 8)
__global__ void __launch_bounds__(256) gemv_rows16_kernel(int m, int n, int nrhs, const __bf16* __restrict__ A,
                                                          long lda, const float* __restrict__ X, long ldx,
                                                          double alpha, double beta, float* Y, long ldy) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= m) return;  // whole waves
    const __bf16* a = A + (size_t)row * lda;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int n8 = n & ~7;
    for (int j = lane * 8; j < n8; j += 512) {
        const uint4 w = *reinterpret_cast<const uint4*>(a + j);
        const unsigned u[4] = {w.x, w.y, w.z, w.w};
        float av[8];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            av[2 * e] = __uint_as_float(u[e] << 16);
            av[2 * e + 1] = __uint_as_float(u[e] & 0xffff0000u);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if (r < nrhs) {
                const float* x = X + j + (size_t)r * ldx;
#pragma unroll
                for (int e = 0; e < 8; ++e) acc[r] += (double)av[e] * (double)x[e];
            }
    }
    for (int j = n8 + lane; j < n; j += 64) {
        const double av = (double)(float)a[j];
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if (r < nrhs) acc[r] += av * (double)X[j + (size_t)r * ldx];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if (r >= nrhs) break;
        double v = acc[r];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) {
            float* y = Y + row + (size_t)r * ldy;
            *y = (float)(beta == 0.0 ? alpha * v : alpha * v + beta * (double)*y);
        }
    }
}
void gemv_rows16(int m, int n, int nrhs, double alpha, const __bf16* A, long lda, const float* X, long ldx,
                 double beta, float* Y, long ldy) {
    if (m <= 0 || nrhs <= 0) return;
    SQ_REQUIRE(nrhs <= 8 && n > 0 && lda % 8 == 0 && ((uintptr_t)A & 15) == 0, SQBEM_ERR_ARG,
               "gemv_rows16: 1 to 8 right-hand sides, n > 0, 16-byte aligned rows");
    ProfScope prof(PROF_GEMV, 2.0 * m * n * nrhs, 2.0 * m * n + 4.0 * ((double)n * nrhs + 2.0 * m * nrhs));
    hipLaunchKernelGGL(gemv_rows16_kernel, dim3(ceil_div(m, 4)), dim3(256), 0, ctx().s(), m, n, nrhs, A, lda, X, ldx,
                       alpha, beta, Y, ldy);
    check_launch("gemv_rows16");
}

// Z = X + alpha Y elementwise over n values
__global__ void axpy_kernel(size_t n, double alpha, const double* __restrict__ y, double* x) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        x[e] += alpha * y[e];
}
void axpy(size_t n, double alpha, const double* y, double* x) {
    if (!n) return;
    unsigned blocks = (unsigned)std::min<size_t>(ceil_div(n, 256), 4096);
    hipLaunchKernelGGL(axpy_kernel, dim3(blocks), dim3(256), 0, ctx().s(), n, alpha, y, x);
    check_launch("axpy");
}

static void device_iota(DBuf<int>& b, int n) {
    b.alloc(n);
    hipLaunchKernelGGL(iota_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx().s(), n, b.p);
    check_launch("iota");
}

// Rigid-body projection C = A K (columns of C at ldc).
static void kproj(int nrows, int Ns, const double* A, long lda, const double* cen_xyz, const double* xc, double* C,
                  long ldc) {
    if (nrows <= 0) return;
    DBuf<double> part((size_t)KP_CHUNKS * 6 * nrows);
    hipLaunchKernelGGL(kproj_kernel, dim3(ceil_div(nrows, 256), KP_CHUNKS), dim3(256), 0, ctx().s(), nrows, Ns, A, lda,
                       cen_xyz, xc[0], xc[1], xc[2], part.p);
    hipLaunchKernelGGL(kproj_reduce, dim3(ceil_div(nrows, 256)), dim3(256), 0, ctx().s(), nrows, part.p, C, ldc);
    check_launch("kproj");
    sync();  // part is released on return
}

template void gemv<double>(int, int, int, double, const double*, long, const double*, long, double, double*, long);
template void gemv<float>(int, int, int, double, const float*, long, const float*, long, double, float*, long);

void completion_diag(int Ns, double* A, long lda, const double* dls, double minus_identity) {
    if (Ns <= 0) return;
    hipLaunchKernelGGL(completion_diag_kernel, dim3(ceil_div(Ns, 256)), dim3(256), 0, ctx().s(), Ns, A, lda, dls,
                       bem::EIGHT_PI, minus_identity);
    check_launch("completion_diag");
}

size_t kproj_part_elems(int nrows) { return (size_t)KP_CHUNKS * 6 * std::max(nrows, 1); }

void kproj_async(int nrows, int Ns, const double* A, long lda, const double* cen_xyz, const double* xc, double* C,
                 long ldc, double* part) {
    if (nrows <= 0) return;
    hipLaunchKernelGGL(kproj_kernel, dim3(ceil_div(nrows, 256), KP_CHUNKS), dim3(256), 0, ctx().s(), nrows, Ns, A, lda,
                       cen_xyz, xc[0], xc[1], xc[2], part);
    hipLaunchKernelGGL(kproj_reduce, dim3(ceil_div(nrows, 256)), dim3(256), 0, ctx().s(), nrows, part, C, ldc);
    check_launch("kproj");
}

template <typename T>
void absmax_async(int m, int n, const T* A, long lda, unsigned long long* out) {
    if (m <= 0 || n <= 0) return;
    unsigned blocks = (unsigned)std::min(n, 4096);
    hipLaunchKernelGGL(absmax_kernel_t<T>, dim3(blocks), dim3(256), 0, ctx().s(), m, n, A, lda, out);
    check_launch("absmax");
}
template void absmax_async<double>(int, int, const double*, long, unsigned long long*);
template void absmax_async<float>(int, int, const float*, long, unsigned long long*);

template <typename T>
static void border_copy(int ns, const T* G, T* M, unsigned long long* mx) {
    hipLaunchKernelGGL(border_copy_kernel<T>, dim3((unsigned)std::min(ns + 6, 4096)), dim3(256), 0, ctx().s(), ns, G, M,
                       mx);
    check_launch("border_copy");
}
template <typename T>
static void border_rows(int Ns, T* M, const double* cen_xyz, const double* area, double sc, const double* xc,
                        const double* sc_dev) {
    hipLaunchKernelGGL(border_rows_kernel<T>, dim3(ceil_div(Ns, 256)), dim3(256), 0, ctx().s(), Ns, M, cen_xyz, area, sc,
                       sc_dev, xc[0], xc[1], xc[2]);
    check_launch("border_rows");
}
template <typename T>
void border_build(int Ns, const T* G, T* M, const double* cen_xyz, const double* area, double sc, const double* xc,
                  const double* sc_dev) {
    border_copy<T>(3 * Ns, G, M, nullptr);
    border_rows<T>(Ns, M, cen_xyz, area, sc, xc, sc_dev);
}
// [Gs | Gc] into the bordered matrix M (n = ns + 6) and max |Gs|, on the host, from the same pass
template <typename T>
static double border_copy_absmax(int ns, const T* G, T* M) {
    DBuf<unsigned long long> r(1);
    r.zero();
    border_copy<T>(ns, G, M, r.p);
    unsigned long long bits = 0;
    r.download(&bits, 1);
    double v;
    std::memcpy(&v, &bits, sizeof v);
    return v;
}
template void border_build<double>(int, const double*, double*, const double*, const double*, double, const double*,
                                   const double*);
template void border_build<float>(int, const float*, float*, const double*, const double*, double, const double*,
                                  const double*);

void constraint_apply(int Ns, int nrhs, const double* fs, long ldf, const double* cen_xyz, const double* area,
                      double sc, const double* xc, double* out, long ldo) {
    if (nrhs <= 0) return;
    hipLaunchKernelGGL(constraint_apply_kernel, dim3(nrhs), dim3(256), 0, ctx().s(), Ns, nrhs, fs, ldf, cen_xyz, area,
                       sc, xc[0], xc[1], xc[2], -1.0, out, ldo);
    check_launch("constraint_apply");
}

}  // namespace sq

using namespace sq;

// ------------------------------------------------------------------ tube
void sq::tube_build(sqbem_tube* t) {
    double t0 = now_s();
    t->Nt = t->mesh.P();
    t->nt = 3 * t->Nt;
    const int Nt = t->Nt, nt = t->nt;
    t->dm.upload(t->mesh);
    device_iota(t->self, Nt);
    t->T64.alloc((size_t)nt * nt);
    // the refined modes factor an FP32 copy, written by the assembly beside the FP64 matrix
    if (mixed(t->prec)) t->T32.alloc((size_t)nt * nt);
    // Wielandt deflation of the f = c m null mode, fused into the assembly
    // (Tube::deflate, src/swimmer.hpp:64-74): += sc n_i[a] n_p[b] A_p, sc = 1/(8 pi mu sqrt(At))
    double At = 0;
    for (int p = 0; p < Nt; ++p) At += t->mesh.area[p];
    const double sc = 1.0 / (bem::EIGHT_PI * t->mu * std::sqrt(At));
    // Large refined tubes: the panels whose columns lie beyond the no-pivot factorisation's
    // first split are assembled on the side stream, beside the factorisation of the left half
    // (which reads and writes only its own columns); the factorisation waits for them before
    // it first touches one. The assembly allocates nothing, so the side stream holds no block
    // the library stream could take.
    const bool split = mixed(t->prec) && nt >= 8192 && ctx().side && !ctx().redirect && !capturing();
    // panel boundaries at or past the factorisation's first split and its left part's split
    const int w1 = split ? getrf_top_split(nt) : nt;
    const int P2 = split ? std::min(Nt, (w1 + 2) / 3) : Nt;
    const int P1 = split ? std::min(P2, (getrf_top_split(w1) + 2) / 3) : Nt;
    assemble(t->dm, 0, P1, t->dm.cen_xyz.p, Nt, t->self.p, bem::EIGHT_PI * t->mu, t->T64.p, nt, 1.0, nullptr, 0,
             nullptr, sc, t->dm.nrm_xyz.p, t->T32.p, nt);
    int nparts = 0;
    if (split && P1 < Nt) {
        if (!t->fork_ev) SQ_HIP(hipEventCreateWithFlags(&t->fork_ev, hipEventDisableTiming));
        for (hipEvent_t& e : t->asm_ev)
            if (!e) SQ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        SQ_HIP(hipEventRecord(t->fork_ev, ctx().s()));  // after the uploads
        SQ_HIP(hipStreamWaitEvent(ctx().side, t->fork_ev, 0));
        const int bounds[3] = {P1, P2, Nt};
        ctx().redirect = ctx().side;
        try {
            for (int q = 0; q < 2; ++q) {
                if (bounds[q + 1] <= bounds[q]) continue;
                assemble(t->dm, bounds[q], bounds[q + 1] - bounds[q], t->dm.cen_xyz.p, Nt, t->self.p,
                         bem::EIGHT_PI * t->mu, t->T64.p + (size_t)3 * bounds[q] * nt, nt, 1.0, nullptr, 0, nullptr,
                         sc, t->dm.nrm_xyz.p, t->T32.p + (size_t)3 * bounds[q] * nt, nt);
                SQ_HIP(hipEventRecord(t->asm_ev[nparts], ctx().side));
                t->asm_from[nparts++] = 3 * bounds[q];
            }
        } catch (...) {
            ctx().redirect = nullptr;
            side_drain();
            throw;
        }
        ctx().redirect = nullptr;
    }
    sync();
    double t1 = now_s();
    t->ipiv.alloc(nt);
    // the factorisation is issued here and finished by tube_ready (its first use)
    if (mixed(t->prec)) {
        t->aux32.alloc(getrf_aux_elems(nt));
        t->job32.reset(new GetrfJob<float>);
        GetrfJob<float>& j = *t->job32;
        j.kind = t->prec;
        j.n = j.lda = nt;
        j.A = t->T32.p;
        j.ipiv = t->ipiv.p;
        j.aux = t->aux32.p;
        for (int q = 0; q < nparts; ++q) {  // columns written by the side stream's assembly
            j.cols_event[q] = t->asm_ev[q];
            j.cols_from[q] = t->asm_from[q];
        }
        j.ncols = nparts;
        // the chained solves' tiled copy is made when a swimmer's Schur complement is done
        // (beside its bordered factorisation) or when a solve first needs it
        j.defer_tiles = true;
        // the pivoting fallback rebuilds the FP32 matrix from the FP64 one (no copy is kept)
        j.restore = [t, nt] { convert_matrix<double, float>(nt, nt, t->T64.p, nt, t->T32.p, nt); };
        if (t->prec == SQBEM_BF16_IR && nt >= 512) {
            // BF16 copies of the factors for the 16-bit products of the factorisation and of
            // the Schur complement's solves
            t->sh_ld = (nt + 63) / 64 * 64;
            t->SL.alloc((size_t)nt * t->sh_ld);
            t->SU.alloc((size_t)nt * t->sh_ld);
            t->SUr.alloc((size_t)nt * t->sh_ld);
            j.sh_lo = t->SL.p;
            j.sh_up = t->SU.p;
            j.sh_upr = t->SUr.p;
            j.sh_ld = t->sh_ld;
        }
        getrf_begin<float>(j);
    } else {
        t->aux64.alloc(getrf_aux_elems(nt));
        t->job64.reset(new GetrfJob<double>);
        GetrfJob<double>& j = *t->job64;
        j.kind = SQBEM_FP64;
        j.n = j.lda = nt;
        j.A = t->T64.p;
        j.ipiv = t->ipiv.p;
        j.aux = t->aux64.p;
        j.defer_tiles = true;
        getrf_begin<double>(j);
    }
    t->t_factor0 = t1;
    t->view.owner = t;
    t->view.n = nt;
    t->view.kind = t->prec;
    t->view.LU = mixed(t->prec) ? nullptr : t->T64.p;
    t->view.LU32 = mixed(t->prec) ? t->T32.p : nullptr;
    t->view.piv = t->ipiv.p;
    t->view.aux = mixed(t->prec) ? nullptr : t->aux64.p;
    t->view.aux32 = mixed(t->prec) ? t->aux32.p : nullptr;
    t->view.A64 = mixed(t->prec) ? t->T64.p : nullptr;
    t->t_assembly = t1 - t0;
    t->t_factor = 0;
}

void sq::tube_ready(const sqbem_tube* tc) {
    if (!tc || (!tc->job32 && !tc->job64)) return;
    sqbem_tube* t = const_cast<sqbem_tube*>(tc);
    int info = 0;
    try {
        info = t->job32 ? getrf_end<float>(*t->job32) : getrf_end<double>(*t->job64);
        t->np_path = t->job32 ? t->job32->path : t->job64->path;
    } catch (...) {
        t->job32.reset();
        t->job64.reset();
        throw;
    }
    t->job32.reset();
    t->job64.reset();
    t->t_factor = now_s() - t->t_factor0;
    SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgetrf failed");
}

// ------------------------------------------------------------------ sphere template
// The sphere-sphere blocks of SwimmerSolver::build (src/swimmer.hpp:209-222) depend only on
// the sphere's shape, not on where it sits: E (deflated SL/(8 pi mu)), the DL part of AuBs,
// and the sphere panels' share of each row's stresslet sum. A trajectory moves one sphere
// through a fixed tube, so these are assembled once and copied into every solver; only the
// tube-sphere couplings, the completion diagonal and the Schur complement are rebuilt.
struct sqbem_sphere {
    bem::Mesh S;
    double xc[3] = {0, 0, 0};
    double mu = 1;
    int Ns = 0, ns = 0;
    DBuf<double> E, AuBs, dlsum;  // ns x ns, ns x ns, 9 Ns
};

static void sphere_template_build(sqbem_sphere* t) {
    t->Ns = t->S.P();
    t->ns = 3 * t->Ns;
    const int Ns = t->Ns, ns = t->ns;
    const double e8 = bem::EIGHT_PI;
    DMesh dm;
    dm.upload(t->S);
    DBuf<int> self;
    device_iota(self, Ns);
    double Asum = 0;
    for (int p = 0; p < Ns; ++p) Asum += t->S.area[p];
    t->E.alloc((size_t)ns * ns);
    t->AuBs.alloc((size_t)ns * ns);
    t->dlsum.alloc((size_t)9 * Ns);
    t->dlsum.zero();
    assemble(dm, 0, Ns, dm.cen_xyz.p, Ns, self.p, e8 * t->mu, t->E.p, ns, e8, t->AuBs.p, ns, t->dlsum.p,
             1.0 / (e8 * t->mu * std::sqrt(Asum)), dm.nrm_xyz.p);
    sync();
}

// ------------------------------------------------------------------ swimmer
struct sqbem_swimmer {
    const sqbem_tube* tube = nullptr;
    const sqbem_sphere* tmpl = nullptr;  // sphere-sphere blocks to copy instead of assembling
    bem::Mesh S;
    DMesh dm;
    double xc[3] = {0, 0, 0};
    DBuf<double> xc_d;
    double mu = 1;
    int prec = SQBEM_FP64;
    int Ns = 0, Nt = 0, ns = 0, nt = 0;
    // FP64 operator pieces (src/swimmer.hpp:290-292 naming)
    DBuf<double> AuBt, AuBs, D, Y, G;  // Y = T^-1 [B | Ct] (FP64 mode); G = [Gs | Gc]
    DBuf<double> B, Ct, E, Cs;         // mixed: kept in FP64 for the refinement residual
    DBuf<float> Y32, G32, D32, M32;    // mixed: low-precision Schur pieces and bordered factors
    DBuf<__bf16> D16;                  // BF16 mode: D row-major in BF16 (the Schur product's A)
    int D16ld = 0;
    DBuf<__bf16> Y16;                  // BF16 mode: Y column-major in BF16, as the solves finalise it
    int Y16ld = 0;
    DBuf<int> Mpiv;
    DBuf<float> Maux;                  // getrf's solve aid for the M32 factors
    DBuf<float> Gs32;                  // mixed, towed problem: factors of Gs alone
    DBuf<int> Gspiv;
    DBuf<float> Gsaux;
    double sc32 = 0;
    double t_assembly = 0, t_schur = 0;
    int ir_iters = 0;
    double ir_residual = 0;
    ~sqbem_swimmer() { side_drain(); }  // the bordered factors' tiled copy may still be written
};

static void swimmer_build(sqbem_swimmer* s) {
    double t0 = now_s();
    const sqbem_tube* tb = s->tube;
    s->Ns = s->S.P();
    s->Nt = tb ? tb->Nt : 0;
    s->ns = 3 * s->Ns;
    s->nt = 3 * s->Nt;
    const int Ns = s->Ns, Nt = s->Nt, ns = s->ns, nt = s->nt;
    const double mu = s->mu, e8 = bem::EIGHT_PI;
    const bool mx = mixed(s->prec);
    // The coupling blocks need the meshes only, not the tube's factors: they are assembled
    // on the side stream while the tube's factorisation (issued when the tube was built) may
    // still be running on the library stream, into the CUs its narrow steps leave idle.
    DBuf<int> self;
    DBuf<double> dlsum;
    {
    SideScope side;
    s->dm.upload(s->S);
    s->xc_d.alloc(3);
    s->xc_d.upload(s->xc, 3);
    device_iota(self, Ns);
    dlsum.alloc((size_t)9 * Ns);
    dlsum.zero();
    // the refined modes' low-precision copies the Schur complement starts from, written by
    // the assembly beside the FP64 blocks: G32 = [E | Cs], Y32 = [B | Ct] (solved in place),
    // D32 = D
    if (mx) {
        s->G32.alloc((size_t)ns * (ns + 6));
        if (tb) {
            s->Y32.alloc((size_t)nt * (ns + 6));
            s->D32.alloc((size_t)ns * nt);
        }
    }
    // tube rows <- sphere sources: B = SL/(8 pi mu), AuBt = -DL/(8 pi)   (src/swimmer.hpp:188-198)
    if (tb) {
        s->AuBt.alloc((size_t)nt * ns);
        double* Bdst;
        if (mx) {
            s->B.alloc((size_t)nt * ns);
            Bdst = s->B.p;
        } else {
            s->Y.alloc((size_t)nt * (ns + 6));
            Bdst = s->Y.p;
        }
        assemble(s->dm, 0, Ns, tb->dm.cen_xyz.p, Nt, nullptr, e8 * mu, Bdst, nt, e8, s->AuBt.p, nt, nullptr, 0.0,
                 nullptr, mx ? s->Y32.p : nullptr, nt);
        // sphere rows <- tube sources: D = SL/(8 pi mu), DL summed (src/swimmer.hpp:201-207)
        s->D.alloc((size_t)ns * nt);
        assemble(tb->dm, 0, Nt, s->dm.cen_xyz.p, Ns, nullptr, e8 * mu, s->D.p, ns, 1.0, nullptr, 0, dlsum.p, 0.0,
                 nullptr, mx ? s->D32.p : nullptr, ns);
    }
    // sphere rows <- sphere sources: E (deflated, src/swimmer.hpp:222), AuBs, DL sums
    s->AuBs.alloc((size_t)ns * ns);
    double* Edst;
    if (mx) {
        s->E.alloc((size_t)ns * ns);
        Edst = s->E.p;
    } else {
        s->G.alloc((size_t)ns * (ns + 6));
        Edst = s->G.p;
    }
    if (s->tmpl) {  // translation-invariant: copied from the sphere's template
        convert_matrix<double, double>(ns, ns, s->tmpl->E.p, ns, Edst, ns);
        convert_matrix<double, double>(ns, ns, s->tmpl->AuBs.p, ns, s->AuBs.p, ns);
        axpy((size_t)9 * Ns, 1.0, s->tmpl->dlsum.p, dlsum.p);
    } else {
        double Asum = 0;
        for (int p = 0; p < Ns; ++p) Asum += s->S.area[p];
        const double defl = 1.0 / (e8 * mu * std::sqrt(Asum));
        assemble(s->dm, 0, Ns, s->dm.cen_xyz.p, Ns, self.p, e8 * mu, Edst, ns, e8, s->AuBs.p, ns, dlsum.p, defl,
                 s->dm.nrm_xyz.p, mx ? s->G32.p : nullptr, ns);
    }
    // completed double layer (src/swimmer.hpp:218-220)
    hipLaunchKernelGGL(completion_diag_kernel, dim3(ceil_div(Ns, 256)), dim3(256), 0, ctx().s(), Ns, s->AuBs.p,
                       (long)ns, dlsum.p, e8, tb ? 0.0 : 1.0);
    check_launch("completion_diag");
    // rigid-body columns Ct = AuBt K, Cs = AuBs K (src/swimmer.hpp:223-234)
    if (mx) {
        if (tb) {
            s->Ct.alloc((size_t)nt * 6);
            kproj(nt, Ns, s->AuBt.p, nt, s->dm.cen_xyz.p, s->xc, s->Ct.p, nt);
        }
        s->Cs.alloc((size_t)ns * 6);
        kproj(ns, Ns, s->AuBs.p, ns, s->dm.cen_xyz.p, s->xc, s->Cs.p, ns);
    } else {
        if (tb) kproj(nt, Ns, s->AuBt.p, nt, s->dm.cen_xyz.p, s->xc, s->Y.p + (size_t)nt * ns, nt);
        kproj(ns, Ns, s->AuBs.p, ns, s->dm.cen_xyz.p, s->xc, s->G.p + (size_t)ns * ns, ns);
    }
    if (mx) {
        // the rigid-body columns' low-precision copies (and E's, when it came from the
        // sphere's template), still beside the tube's factorisation
        if (s->tmpl) convert_matrix<double, float>(ns, ns, s->E.p, ns, s->G32.p, ns);
        convert_matrix<double, float>(ns, 6, s->Cs.p, ns, s->G32.p + (size_t)ns * ns, ns);
        if (tb) {
            convert_matrix<double, float>(nt, 6, s->Ct.p, nt, s->Y32.p + (size_t)nt * ns, nt);
            if (s->prec == SQBEM_BF16_IR) {
                s->D16ld = shadow_rows_ld(nt);
                s->D16.alloc(shadow_rows_elems(ns, nt));
                shadow_rows(ns, nt, s->D32.p, ns, s->D16.p);
            }
        }
    }
    sync();
    }  // the library stream waits for the side stream from here
    double t1 = now_s();
    s->t_assembly = t1 - t0;
    tube_ready(tb);  // the factors the Schur complement uses
    t1 = now_s();
    // Schur complement (src/swimmer.hpp:235-249): Y = T^-1 [B | Ct]; [Gs | Gc] -= D Y
    if (!mx) {
        if (tb) {
            getrs<double>(SQBEM_FP64, 0, nt, ns + 6, tb->T64.p, nt, tb->ipiv.p, s->Y.p, nt, tb->aux64.p);
            dgemm(ns, ns + 6, nt, -1.0, s->D.p, ns, s->Y.p, nt, 1.0, s->G.p, ns);
            chain_tiles_issue(tb->aux64.p);  // for the solves, after the Schur complement's products
        }
    } else {
        const int kind = s->prec;
        if (tb) {
            // BF16: the 16-bit products of the solves and of D Y read BF16 copies of their
            // operands (the tube's factors, the solved right-hand sides as the solves'
            // leaves finalise them, D) instead of converting them
            // (Y's copy is kept: the refinement's corrections multiply by it)
            const bool sh = kind == SQBEM_BF16_IR && tb->np_path == 1 && tb->SL.p;
            const int yld = (nt + 63) / 64 * 64;
            s->Y16.release();
            s->Y16ld = 0;
            if (sh) {
                s->Y16.alloc((size_t)yld * (ns + 6));
                s->Y16ld = yld;
            }
            {
                ShadowScope f(sh ? tb->T32.p : nullptr, nt, nt, nt, tb->SL.p, tb->SU.p, tb->sh_ld, true, tb->SUr.p);
                ShadowScope y(sh ? s->Y32.p : nullptr, nt, nt, ns + 6, nullptr, s->Y16.p, yld, false);
                ShadowScope d(sh ? s->D32.p : nullptr, ns, ns, nt, s->D16.p, nullptr, s->D16ld, false);
                getrs<float>(kind, 0, nt, ns + 6, tb->T32.p, nt, tb->ipiv.p, s->Y32.p, nt, tb->aux32.p);
                if (kind == SQBEM_FP32_IR)
                    sgemm(ns, ns + 6, nt, -1.0f, s->D32.p, ns, s->Y32.p, nt, 1.0f, s->G32.p, ns);
                else
                    gemm16(kind, ns, ns + 6, nt, -1.0f, s->D32.p, ns, s->Y32.p, nt, 1.0f, s->G32.p, ns);
            }
        }
        // bordered low-precision factors, reused by every refinement step
        const int n = ns + 6;
        s->M32.alloc((size_t)n * n);
        const double amax = border_copy_absmax<float>(ns, s->G32.p, s->M32.p);
        double Amax = 0;
        for (int p = 0; p < Ns; ++p) Amax = std::max(Amax, 5.0 * s->S.area[p]);
        s->sc32 = amax / Amax;
        border_rows<float>(Ns, s->M32.p, s->dm.cen_xyz.p, s->dm.area.p, s->sc32, s->xc, nullptr);
        // the tube factors' tiled copy for the refinement's solves: on the side stream beside the
        // bordered factorisation (latency-bound), after the Schur complement's products
        if (tb) chain_tiles_issue(tb->aux32.p);
        // rebuilds M from G (the pivoting fallback)
        const std::function<void()> border = [&] {
            border_build<float>(Ns, s->G32.p, s->M32.p, s->dm.cen_xyz.p, s->dm.area.p, s->sc32, s->xc);
        };
        s->Mpiv.alloc(n);
        s->Maux.alloc(getrf_aux_elems(n));
        DBuf<__bf16> Msh;  // BF16 copies of the bordered factors for its 16-bit products
        const int mld = (n + 63) / 64 * 64;
        if (kind == SQBEM_BF16_IR && n >= 512) Msh.alloc((size_t)2 * n * mld);
        int info = getrf<float>(kind, n, s->M32.p, n, s->Mpiv.p, s->Maux.p, &border, Msh.p,
                                Msh.p ? Msh.p + (size_t)n * mld : nullptr, mld);
        SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgesv failed");
    }
    sync();
    s->t_schur = now_s() - t1;
}

// Reduced right-hand sides for nslip slip fields u (device, ns x k):
//   rs = -AuBs u - D T^-1 (-AuBt u),   rt = T^-1 (-AuBt u)      (src/swimmer.hpp:256-269)
static void reduced_rhs(sqbem_swimmer* s, int k, const double* u, double* rs, long ldrs, double* rt) {
    const int ns = s->ns, nt = s->nt;
    gemv<double>(ns, ns, k, -1.0, s->AuBs.p, ns, u, ns, 0.0, rs, ldrs);
    if (s->tube) {
        gemv<double>(nt, ns, k, -1.0, s->AuBt.p, nt, u, ns, 0.0, rt, nt);
        getrs<double>(SQBEM_FP64, 0, nt, k, s->tube->T64.p, nt, s->tube->ipiv.p, rt, nt, s->tube->aux64.p);
        gemv<double>(ns, nt, k, -1.0, s->D.p, ns, rt, nt, 1.0, rs, ldrs);
    }
}

// LU of the bordered matrix of solve_free (src/swimmer.hpp:107-125): [Gs | Gc] over the
// force/torque rows scaled by max|Gs| / max(5 A), FP64
static void bordered_factor_fp64(sqbem_swimmer* s, DBuf<double>& M, DBuf<int>& piv) {
    const int Ns = s->Ns, ns = s->ns, n = ns + 6;
    M.alloc((size_t)n * n);
    const double amax = border_copy_absmax<double>(ns, s->G.p, M.p);
    double Amax = 0;
    for (int p = 0; p < Ns; ++p) Amax = std::max(Amax, 5.0 * s->S.area[p]);
    const double sc = amax / Amax;
    border_rows<double>(Ns, M.p, s->dm.cen_xyz.p, s->dm.area.p, sc, s->xc, nullptr);
    piv.alloc(n);
    int info = getrf<double>(SQBEM_FP64, n, M.p, n, piv.p);
    SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgesv failed");
}

// FP64 path of solve_free / solve_prescribed (src/swimmer.hpp:107-153)
static void solve_fp64(sqbem_swimmer* s, int k, const double* u_host, bool towed, const double* w_host, double* X_host,
                       double* ft_host) {
    const int ns = s->ns, nt = s->nt, n = ns + 6;
    DBuf<double> X((size_t)n * k), rt((size_t)std::max(nt, 1) * k);
    if (!towed) {
        DBuf<double> u((size_t)ns * k);
        u.upload(u_host, (size_t)ns * k);
        X.zero();
        reduced_rhs(s, k, u.p, X.p, n, rt.p);
        DBuf<double> M;
        DBuf<int> piv;
        bordered_factor_fp64(s, M, piv);
        getrs<double>(SQBEM_FP64, 0, n, k, M.p, n, piv.p, X.p, n);
    } else {
        // towed sphere: Gs f = -Gc w, solved on a copy of Gs (src/swimmer.hpp:142-153)
        DBuf<double> w(6), Gs((size_t)ns * ns);
        w.upload(w_host, 6);
        convert_matrix<double, double>(ns, ns, s->G.p, ns, Gs.p, ns);
        gemv<double>(ns, 6, 1, -1.0, s->G.p + (size_t)ns * ns, ns, w.p, 6, 0.0, X.p, n);
        SQ_HIP(hipMemcpyAsync(X.p + ns, w.p, 6 * sizeof(double), hipMemcpyDeviceToDevice, ctx().s()));
        DBuf<int> piv(ns);
        int info = getrf<double>(SQBEM_FP64, ns, Gs.p, ns, piv.p);
        SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgesv failed");
        getrs<double>(SQBEM_FP64, 0, ns, 1, Gs.p, ns, piv.p, X.p, n);
        if (nt) rt.zero();
    }
    // f_t = T^-1 r_t - Y x (unpack, src/swimmer.hpp:274-288)
    if (s->tube) gemv<double>(nt, n, k, -1.0, s->Y.p, nt, X.p, n, 1.0, rt.p, nt);
    X.download(X_host, (size_t)n * k);
    if (s->tube && ft_host) rt.download(ft_host, (size_t)nt * k);
}

// ---- mixed precision: iterative refinement on the full bordered system
//   [ T  B  Ct ] [f_t]   [r_t]
//   [ D  E  Cs ] [f_s] = [r_s]         r_t = -AuBt u, r_s = -AuBs u
//   [ 0 scF 0  ] [ w ]   [ 0 ]
// Residuals in FP64 from the FP64 operator; corrections from the FP32 factors of T and
// of the bordered Schur matrix (or of Gs alone for the towed problem, whose w is given and
// whose unknowns are f_t, f_s only). The Schur elimination is exact algebra, so the
// refined solution is the reference's solution to the refinement tolerance.
//
// x, b, r: device, N = nt + ns + 6 rows per right-hand side, blocks [t; s; c].
// Rows of the tube (r_t): T x_t + B x_s + Ct w, the larger part of the FP64 operator.
static void residual_tube_rows(sqbem_swimmer* s, int k, const double* x, const double* b, double* r, bool towed) {
    const int ns = s->ns, nt = s->nt, N = nt + ns + 6;
    if (!s->tube) return;
    convert_matrix<double, double>(nt, k, b, N, r, N);
    gemv<double>(nt, nt, k, -1.0, s->tube->T64.p, nt, x, N, 1.0, r, N);
    gemv<double>(nt, ns, k, -1.0, s->B.p, nt, x + nt, N, 1.0, r, N);
    if (!towed) gemv<double>(nt, 6, k, -1.0, s->Ct.p, nt, x + nt + ns, N, 1.0, r, N);
}
// Rows of the sphere (r_s: D x_t + E x_s + Cs w) and of the constraints (r_c); c: 6 k
// values of scratch that outlive the launches.
static void residual_sphere_rows(sqbem_swimmer* s, int k, const double* x, const double* b, double* r, bool towed,
                                 double* c) {
    const int ns = s->ns, nt = s->nt, N = nt + ns + 6;
    const double* xt = x;
    const double* xs = x + nt;
    const double* xw = xs + ns;
    double* rs = r + nt;
    double* rc = rs + ns;
    convert_matrix<double, double>(ns + 6, k, b + nt, N, rs, N);
    if (s->tube) gemv<double>(ns, nt, k, -1.0, s->D.p, ns, xt, N, 1.0, rs, N);
    gemv<double>(ns, ns, k, -1.0, s->E.p, ns, xs, N, 1.0, rs, N);
    if (!towed) {
        gemv<double>(ns, 6, k, -1.0, s->Cs.p, ns, xw, N, 1.0, rs, N);
        hipLaunchKernelGGL(constraint_apply_kernel, dim3(k), dim3(256), 0, ctx().s(), s->Ns, k, xs, (long)N,
                           s->dm.cen_xyz.p, s->dm.area.p, s->sc32, s->xc[0], s->xc[1], s->xc[2], 1.0, c, 6L);
        check_launch("constraint_apply");
        for (int j = 0; j < k; ++j) axpy(6, -1.0, c + 6 * j, rc + (size_t)j * N);
    }
}

// The sphere and constraint rows of a refinement residual run on the side stream: the next
// correction reads them only after its tube solves (chains of latency-bound hand-overs,
// which leave most of the HBM bandwidth to this pass) on the library stream. Nothing is
// allocated or released on the side stream here (the partial sums have a buffer of its
// own, the scratch comes from the caller), and the library stream waits for the pass
// before anything reads those rows and before the caller's buffers are released.
struct ResidualFork {
    hipEvent_t fork = nullptr, join = nullptr;
    bool pending = false;
    ResidualFork() {
        SQ_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        SQ_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    }
    ~ResidualFork() {
        wait();
        (void)hipEventDestroy(fork);
        (void)hipEventDestroy(join);
    }
    ResidualFork(const ResidualFork&) = delete;
    ResidualFork& operator=(const ResidualFork&) = delete;
    void wait() {
        if (!pending) return;
        pending = false;
        (void)hipStreamWaitEvent(ctx().s(), join, 0);
    }
};

// r = b - A x in FP64 (every right-hand side a pass: the FP64 operator is read once per
// refinement step). With rf, the sphere and constraint rows go to the side stream and rf
// holds their completion until rf->wait().
static void full_residual(sqbem_swimmer* s, int k, const double* x, const double* b, double* r, bool towed,
                          double* c, ResidualFork* rf) {
    if (rf && ctx().side && !ctx().redirect && !capturing()) {
        SQ_HIP(hipEventRecord(rf->fork, ctx().s()));
        SQ_HIP(hipStreamWaitEvent(ctx().side, rf->fork, 0));
        ctx().redirect = ctx().side;
        try {
            residual_sphere_rows(s, k, x, b, r, towed, c);
        } catch (...) {
            ctx().redirect = nullptr;
            (void)hipEventRecord(rf->join, ctx().side);
            rf->pending = true;
            throw;
        }
        ctx().redirect = nullptr;
        SQ_HIP(hipEventRecord(rf->join, ctx().side));
        rf->pending = true;
    } else {
        residual_sphere_rows(s, k, x, b, r, towed, c);
    }
    residual_tube_rows(s, k, x, b, r, towed);
}

// d = approx A^-1 r from the low-precision factors
// rf: a residual whose sphere rows may still be in the making on the side stream
static void approx_solve(sqbem_swimmer* s, int k, const double* r, double* d, bool towed, ResidualFork& rf) {
    const int ns = s->ns, nt = s->nt, N = nt + ns + 6, n = ns + 6;
    const int kind = s->prec;
    const int nv = towed ? ns : n;  // unknowns of the reduced system
    DBuf<float> z32((size_t)std::max(nt, 1) * k), v32((size_t)n * k);
    // z = T^-1 r_t ; v = [r_s - D z ; r_c]
    if (nt) {
        convert_matrix<double, float>(nt, k, r, N, z32.p, nt);
        getrs<float>(kind, 0, nt, k, s->tube->T32.p, nt, s->tube->ipiv.p, z32.p, nt, s->tube->aux32.p);
    }
    rf.wait();
    convert_matrix<double, float>(nv, k, r + nt, N, v32.p, n);
    if (nt) {
        if (s->D16.p)
            gemv_rows16(ns, nt, k, -1.0, s->D16.p, s->D16ld, z32.p, nt, 1.0, v32.p, n);
        else
            gemv<float>(ns, nt, k, -1.0, s->D32.p, ns, z32.p, nt, 1.0, v32.p, n);
    }
    if (towed)
        getrs<float>(kind, 0, ns, k, s->Gs32.p, ns, s->Gspiv.p, v32.p, n, s->Gsaux.p);
    else
        getrs<float>(kind, 0, n, k, s->M32.p, n, s->Mpiv.p, v32.p, n, s->Maux.p);
    // f_t = z - Y [f_s; w]   (towed: w is not an unknown, only the f_s columns of Y)
    if (nt) {
        if (s->Y16.p)
            gemv_cols16(nt, nv, k, -1.0, s->Y16.p, s->Y16ld, v32.p, n, 1.0, z32.p, nt);
        else
            gemv<float>(nt, nv, k, -1.0, s->Y32.p, nt, v32.p, n, 1.0, z32.p, nt);
    }
    fill_zero(d, sizeof(double) * (size_t)N * k);
    if (nt) convert_matrix<float, double>(nt, k, z32.p, nt, d, N);
    convert_matrix<float, double>(nv, k, v32.p, n, d + nt, N);
}

// sum of squares of column blockIdx.x of a (N rows, leading dimension N): fixed-order
// partial sums, so the same vector always gives the same norm
__global__ void __launch_bounds__(256) col_norm2_kernel(int N, const double* __restrict__ a, double* out) {
    const double* c = a + (size_t)blockIdx.x * N;
    double s = 0;
    for (int i = threadIdx.x; i < N; i += 256) s += c[i] * c[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
}

// the 2-norms of the k columns of a, on the device; only k values cross to the host
// the squared 2-norms of the k columns of a into sq (device, k values), without waiting
static void col_norms_launch(int N, int k, const double* a, double* sq) {
    hipLaunchKernelGGL(col_norm2_kernel, dim3(k), dim3(256), 0, ctx().s(), N, a, sq);
    check_launch("col_norm2");
}
// the first cnt of them on the host, as norms
static void col_norms_read(const DBuf<double>& sq, int cnt, std::vector<double>& out) {
    out.assign(cnt, 0.0);
    sq.download(out.data(), cnt);
    sync();
    for (double& v : out) v = std::sqrt(v);
}

static void solve_mixed(sqbem_swimmer* s, int k, const double* u_host, bool towed, const double* w_host,
                        double* X_host, double* ft_host) {
    const int ns = s->ns, nt = s->nt, N = nt + ns + 6, n = ns + 6;
    DBuf<double> b((size_t)N * k), x((size_t)N * k), r((size_t)N * k), d((size_t)N * k), c((size_t)6 * k);
    b.zero();
    x.zero();
    if (!towed) {
        DBuf<double> u((size_t)ns * k);
        u.upload(u_host, (size_t)ns * k);
        // every slip field in one pass over AuBt and AuBs
        for (int j0 = 0; j0 < k; j0 += 8) {
            const int kk = std::min(8, k - j0);
            const double* uj = u.p + (size_t)j0 * ns;
            double* bj = b.p + (size_t)j0 * N;
            if (nt) gemv<double>(nt, ns, kk, -1.0, s->AuBt.p, nt, uj, ns, 0.0, bj, N);
            gemv<double>(ns, ns, kk, -1.0, s->AuBs.p, ns, uj, ns, 0.0, bj + nt, N);
        }
    } else {
        SQ_REQUIRE(k == 1, SQBEM_ERR_ARG, "towed solve takes one right-hand side");
        // Gs factor for the towed corrections (built on first use)
        if (!s->Gs32.p) {
            s->Gs32.alloc((size_t)ns * ns);
            const std::function<void()> copy = [&] {
                convert_matrix<float, float>(ns, ns, s->G32.p, ns, s->Gs32.p, ns);
            };
            copy();
            s->Gspiv.alloc(ns);
            s->Gsaux.alloc(getrf_aux_elems(ns));
            int info = getrf<float>(s->prec, ns, s->Gs32.p, ns, s->Gspiv.p, s->Gsaux.p, &copy);
            SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgesv failed");
        }
        DBuf<double> w(6);
        w.upload(w_host, 6);
        if (nt) gemv<double>(nt, 6, 1, -1.0, s->Ct.p, nt, w.p, 6, 0.0, b.p, N);
        gemv<double>(ns, 6, 1, -1.0, s->Cs.p, ns, w.p, 6, 0.0, b.p + nt, N);
    }
    // norms of b (first k) and of the residual (next k): b's are read with the first check's
    std::vector<double> nrm;
    DBuf<double> nrm_d((size_t)2 * k);
    col_norms_launch(N, k, b.p, nrm_d.p);
    convert_matrix<double, double>(N, k, b.p, N, r.p, N);
    const double tol = SQBEM_IR_TOL;
    // The residual is first read back after min_check steps, so the steps before run without
    // a host round trip: a BF16 approximate solve has never brought the relative residual to
    // 1e-12 in fewer than three steps (four on every benchmark workload), FP16 and FP32 ones
    // in fewer than two. A step more than needed leaves the solution at its FP64 accuracy.
    const int min_check = s->prec == SQBEM_BF16_IR ? 3 : 2;
    int it = 0;
    double worst = 1;
    ResidualFork rf;  // after the buffers: waited for before they are released
    for (; it < 40; ++it) {
        approx_solve(s, k, r.p, d.p, towed, rf);
        axpy((size_t)N * k, 1.0, d.p, x.p);
        full_residual(s, k, x.p, b.p, r.p, towed, c.p, &rf);
        if (it + 1 < min_check) continue;
        rf.wait();
        col_norms_launch(N, k, r.p, nrm_d.p + k);
        col_norms_read(nrm_d, 2 * k, nrm);
        worst = 0;
        for (int j = 0; j < k; ++j) worst = std::max(worst, nrm[j] > 0 ? nrm[k + j] / nrm[j] : nrm[k + j]);
        if (worst <= tol) {
            ++it;
            break;
        }
    }
    s->ir_iters = it;
    s->ir_residual = worst;
    SQ_REQUIRE(worst <= tol, SQBEM_ERR_CONVERGE,
               "iterative refinement did not reach relative residual 1e-12 (" + std::to_string(worst) + " after " +
                   std::to_string(it) + " steps)");
    std::vector<double> hx((size_t)N * k);
    SQ_HIP(hipMemcpyAsync(hx.data(), x.p, hx.size() * sizeof(double), hipMemcpyDeviceToHost, ctx().s()));
    sync();
    for (int j = 0; j < k; ++j) {
        std::memcpy(X_host + (size_t)j * n, hx.data() + (size_t)j * N + nt, sizeof(double) * n);
        if (towed)
            for (int c = 0; c < 6; ++c) X_host[(size_t)j * n + ns + c] = w_host[c];
        if (ft_host && nt) std::memcpy(ft_host + (size_t)j * nt, hx.data() + (size_t)j * N, sizeof(double) * nt);
    }
}

// ------------------------------------------------------------------ adjoint
// y(j, r) = alpha sum_i A(i, j) x(i, r) + beta y(j, r): one wave per column of A
__global__ void gemv_t_kernel(int m, int n, int k, double alpha, const double* __restrict__ A, long lda,
                              const double* __restrict__ X, long ldx, double beta, double* Y, long ldy) {
    const int col = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    if (col >= n) return;  // whole waves
    const double* a = A + (size_t)col * lda;
    for (int r = 0; r < k; ++r) {
        const double* x = X + (size_t)r * ldx;
        double acc = 0.0;
        for (int i = lane; i < m; i += 64) acc += a[i] * x[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) {
            double* y = Y + col + (size_t)r * ldy;
            *y = beta == 0.0 ? alpha * acc : alpha * acc + beta * *y;
        }
    }
}
static void gemv_t(int m, int n, int k, double alpha, const double* A, long lda, const double* X, long ldx,
                   double beta, double* Y, long ldy) {
    if (n <= 0 || k <= 0) return;
    ProfScope prof(PROF_GEMV, 2.0 * m * n * k, 8.0 * ((double)m * n + (double)m * k + 2.0 * n * k));
    hipLaunchKernelGGL(gemv_t_kernel, dim3(ceil_div((size_t)n * 64, 256)), dim3(256), 0, ctx().s(), m, n, k, alpha, A,
                       lda, X, ldx, beta, Y, ldy);
    check_launch("gemv_t");
}

// Gradients of nfun functionals J = g_j . (U, Omega) of the free-swimming solution for one
// slip field u (FP64 swimmer). Full system (unknowns f_t, f_s, w = (U, Omega); rows tube,
// sphere, constraints):
//   [ T  B   Ct ] [f_t]   [-AuBt u]
//   [ D  E   Cs ] [f_s] = [-AuBs u]        Ct = AuBt K, Cs = AuBs K
//   [ 0  scF 0  ] [ w ]   [   0   ]
// Adjoint A^T lam = dJ/dz by the same Schur elimination, transposed:
//   M^T [lam_s; lam_c] = [0; g]   (M the bordered matrix of solve_free, its LU reused)
//   lam_t = -T^-T D^T lam_s       (the tube LU reused)
// Then dJ/du = -(AuBt^T lam_t + AuBs^T lam_s) (u enters the right-hand side only), and for
// the sphere centre x_c, which moves the sphere panels but not the tube,
//   dJ/dx_c = -lam^T (dA/dx_c z - db/dx_c)
//           = -[ lam_t . (dB f_s + dAuBt v) + lam_s . (dD f_t + d(AuBs diagonal) v) ],
// v = K w + u the sphere's surface velocity: the sphere-sphere blocks, K and the
// constraint rows are invariant under a translation of the sphere.
static void adjoint_fp64(sqbem_swimmer* s, const double* u_host, int nfun, const double* g_host, double* X_host,
                         double* ft_host, double* dxc_host, double* du_host) {
    const int Ns = s->Ns, Nt = s->Nt, ns = s->ns, nt = s->nt, n = ns + 6;
    const double e8 = bem::EIGHT_PI, mu = s->mu;
    DBuf<double> u(ns), X(n), rt((size_t)std::max(nt, 1));
    u.upload(u_host, ns);
    X.zero();
    reduced_rhs(s, 1, u.p, X.p, n, rt.p);
    DBuf<double> M;
    DBuf<int> piv;
    bordered_factor_fp64(s, M, piv);
    getrs<double>(SQBEM_FP64, 0, n, 1, M.p, n, piv.p, X.p, n);
    if (s->tube) gemv<double>(nt, n, 1, -1.0, s->Y.p, nt, X.p, n, 1.0, rt.p, nt);  // f_t
    // adjoint of the reduced system
    std::vector<double> hl((size_t)n * nfun, 0.0);
    for (int j = 0; j < nfun; ++j)
        for (int c = 0; c < 6; ++c) hl[(size_t)j * n + ns + c] = g_host[6 * j + c];
    DBuf<double> lam((size_t)n * nfun), lt((size_t)std::max(nt, 1) * nfun), psi((size_t)ns * nfun);
    lam.upload(hl.data(), hl.size());
    getrs<double>(SQBEM_FP64, 1, n, nfun, M.p, n, piv.p, lam.p, n);
    if (s->tube) {
        gemv_t(ns, nt, nfun, -1.0, s->D.p, ns, lam.p, n, 0.0, lt.p, nt);
        getrs<double>(SQBEM_FP64, 1, nt, nfun, s->tube->T64.p, nt, s->tube->ipiv.p, lt.p, nt);
    }
    // dJ/du = -(AuBs^T lam_s + AuBt^T lam_t)
    gemv_t(ns, ns, nfun, -1.0, s->AuBs.p, ns, lam.p, n, 0.0, psi.p, ns);
    if (s->tube) gemv_t(nt, ns, nfun, -1.0, s->AuBt.p, nt, lt.p, nt, 1.0, psi.p, ns);
    X.download(X_host, n);
    if (ft_host && nt) rt.download(ft_host, nt);
    psi.download(du_host, (size_t)ns * nfun);
    for (int j = 0; j < 3 * nfun; ++j) dxc_host[j] = 0.0;
    if (!s->tube) {
        sync();
        return;  // no confinement: translation invariant
    }
    // sphere surface velocity v = U + Omega x (c - x_c) + u
    std::vector<double> hv((size_t)ns);
    const double* w = X_host + ns;
    for (int p = 0; p < Ns; ++p) {
        const bem::V3 r = s->S.centroid[p] - bem::V3{s->xc[0], s->xc[1], s->xc[2]};
        const double om[3] = {w[3], w[4], w[5]};
        const double cr[3] = {om[1] * r.z - om[2] * r.y, om[2] * r.x - om[0] * r.z, om[0] * r.y - om[1] * r.x};
        for (int c = 0; c < 3; ++c) hv[3 * (size_t)p + c] = w[c] + cr[c] + u_host[3 * (size_t)p + c];
    }
    DBuf<double> v(ns), gt((size_t)3 * Nt), gs((size_t)3 * Ns);
    v.upload(hv.data(), ns);
    std::vector<double> ht((size_t)3 * Nt), hs((size_t)3 * Ns);
    for (int j = 0; j < nfun; ++j) {
        gt.zero();
        gs.zero();
        // tube targets, sphere sources moving with x_c: B = SL/(8 pi mu) on f_s, AuBt = -DL/(8 pi) on v
        adjoint_xc(s->dm, s->tube->dm.cen_xyz.p, Nt, lt.p + (size_t)j * nt, nullptr, X.p, v.p, 1.0 / (e8 * mu),
                   -1.0 / e8, +1.0, gt.p);
        // sphere targets moving with x_c, tube sources: D = SL/(8 pi mu) on f_t, and the tube
        // part of the completed double layer, +DL/(8 pi) on the target's own v
        adjoint_xc(s->tube->dm, s->dm.cen_xyz.p, Ns, lam.p + (size_t)j * n, v.p, rt.p, nullptr, 1.0 / (e8 * mu),
                   1.0 / e8, -1.0, gs.p);
        gt.download(ht.data(), ht.size());
        gs.download(hs.data(), hs.size());
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
            for (int i = 0; i < Nt; ++i) acc += ht[3 * (size_t)i + k];
            for (int i = 0; i < Ns; ++i) acc += hs[3 * (size_t)i + k];
            dxc_host[3 * j + k] = -acc;
        }
    }
}

// ------------------------------------------------------------------ duct
struct sqbem_duct {
    bem::Mesh mesh;
    DMesh dm;
    double mu = 1;
    int P = 0, n = 0;
    DBuf<double> LU, Fk;
    DBuf<int> ipiv;
    std::vector<int> cap;
};

// column selection of src/duct_flow.hpp:97-99: column 3p+b from A_f when u is known
__global__ void duct_select_kernel(int n, const int* __restrict__ tags, const double* __restrict__ Af,
                                   const double* __restrict__ Au, double* M) {
    size_t total = (size_t)n * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        size_t j = e / n;
        int p = (int)(j / 3), b = (int)(j % 3);
        bool known = !(tags[p] != bem::WALL && b == 2);
        M[e] = known ? Af[e] : Au[e];
    }
}
__global__ void duct_diag_kernel(int P, const int* __restrict__ tags, const double* __restrict__ dls, double* M,
                                 double eight_pi) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const size_t n = 3 * (size_t)P;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            bool known = !(tags[i] != bem::WALL && b == 2);
            if (!known) M[(3 * (size_t)i + a) + (3 * (size_t)i + b) * n] += dls[9 * (size_t)i + 3 * a + b] / eight_pi;
        }
}
__global__ void gather_cols_kernel(int n, int ncols, const int* __restrict__ cols, const double* __restrict__ A,
                                   double* out) {
    size_t total = (size_t)n * ncols;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        size_t i = e % n, c = e / n;
        out[e] = A[i + (size_t)cols[c] * n];
    }
}

static void duct_build(sqbem_duct* d) {
    d->P = d->mesh.P();
    d->n = 3 * d->P;
    const int P = d->P, n = d->n;
    d->dm.upload(d->mesh);
    DBuf<int> self;
    device_iota(self, P);
    DBuf<double> Af((size_t)n * n), Au((size_t)n * n), dls((size_t)9 * P);
    dls.zero();
    assemble(d->dm, 0, P, d->dm.cen_xyz.p, P, self.p, bem::EIGHT_PI * d->mu, Af.p, n, bem::EIGHT_PI, Au.p, n, dls.p, 0.0,
             nullptr);
    DBuf<int> tags(P);
    tags.upload(d->mesh.tags.data(), P);
    d->LU.alloc((size_t)n * n);
    unsigned blocks = std::min<unsigned>(ceil_div((size_t)n * n, 256), 65535);
    hipLaunchKernelGGL(duct_select_kernel, dim3(blocks), dim3(256), 0, ctx().s(), n, tags.p, Af.p, Au.p, d->LU.p);
    hipLaunchKernelGGL(duct_diag_kernel, dim3(ceil_div(P, 256)), dim3(256), 0, ctx().s(), P, tags.p, dls.p, d->LU.p,
                       bem::EIGHT_PI);
    check_launch("duct");
    d->cap.clear();
    for (int p = 0; p < P; ++p)
        if (d->mesh.tags[p] != bem::WALL) d->cap.push_back(p);
    const int nc = (int)d->cap.size();
    std::vector<int> cols(nc);
    for (int k = 0; k < nc; ++k) cols[k] = 3 * d->cap[k] + 2;
    DBuf<int> cols_d(std::max(nc, 1));
    if (nc) cols_d.upload(cols.data(), nc);
    d->Fk.alloc((size_t)n * std::max(nc, 1));
    if (nc) {
        hipLaunchKernelGGL(gather_cols_kernel, dim3(std::min<unsigned>(ceil_div((size_t)n * nc, 256), 65535)), dim3(256),
                           0, ctx().s(), n, nc, cols_d.p, Af.p, d->Fk.p);
        check_launch("gather_cols");
    }
    d->ipiv.alloc(n);
    int info = getrf<double>(SQBEM_FP64, n, d->LU.p, n, d->ipiv.p);
    SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgetrf failed");
    sync();
}

// ------------------------------------------------------------------ C ABI
extern "C" {

sqbem_tube* sqbem_tube_create(const sqbem_mesh* mesh, double mu, int precision) {
    return guarded_ptr<sqbem_tube>([&] {
        SQ_REQUIRE(mesh && mesh->m.P() > 0, SQBEM_ERR_ARG, "sqbem_tube_create: empty mesh");
        SQ_REQUIRE(precision >= SQBEM_FP64 && precision <= SQBEM_BF16_IR, SQBEM_ERR_ARG, "unknown precision");
        auto* t = new sqbem_tube;
        try {
            t->mesh = mesh->m;
            t->mu = mu;
            t->prec = precision;
            tube_build(t);
        } catch (...) {
            delete t;
            throw;
        }
        return t;
    });
}

int sqbem_tube_info(const sqbem_tube* t, int* Nt, double* ta, double* tf) {
    return guarded([&] {
        SQ_REQUIRE(t, SQBEM_ERR_ARG, "null tube");
        if (tf) tube_ready(t);  // the factorisation time is known once it has finished
        if (Nt) *Nt = t->Nt;
        if (ta) *ta = t->t_assembly;
        if (tf) *tf = t->t_factor;
    });
}

const sqbem_lu* sqbem_tube_lu(const sqbem_tube* t) { return t ? &t->view : nullptr; }

void sqbem_tube_free(sqbem_tube* t) { delete t; }

sqbem_swimmer* sqbem_swimmer_create(const sqbem_tube* tube, const sqbem_mesh* sphere, const double* center, double mu,
                                    int precision) {
    return guarded_ptr<sqbem_swimmer>([&] {
        SQ_REQUIRE(sphere && sphere->m.P() > 0 && center, SQBEM_ERR_ARG, "sqbem_swimmer_create: bad arguments");
        SQ_REQUIRE(precision >= SQBEM_FP64 && precision <= SQBEM_BF16_IR, SQBEM_ERR_ARG, "unknown precision");
        SQ_REQUIRE(!tube || mixed(tube->prec) == mixed(precision), SQBEM_ERR_ARG,
                   "swimmer precision must match the tube's (FP64 vs refined)");
        SQ_REQUIRE(!tube || tube->prec == precision || !mixed(precision), SQBEM_ERR_ARG,
                   "swimmer precision must match the tube's");
        auto* s = new sqbem_swimmer;
        try {
            s->tube = tube;
            s->S = sphere->m;
            for (int c = 0; c < 3; ++c) s->xc[c] = center[c];
            s->mu = mu;
            s->prec = precision;
            swimmer_build(s);
        } catch (...) {
            delete s;
            throw;
        }
        return s;
    });
}

sqbem_sphere* sqbem_sphere_create(const sqbem_mesh* sphere, const double* center, double mu) {
    return guarded_ptr<sqbem_sphere>([&] {
        SQ_REQUIRE(sphere && sphere->m.P() > 0 && center, SQBEM_ERR_ARG, "sqbem_sphere_create: bad arguments");
        auto* t = new sqbem_sphere;
        try {
            t->S = sphere->m;
            for (int c = 0; c < 3; ++c) t->xc[c] = center[c];
            t->mu = mu;
            sphere_template_build(t);
        } catch (...) {
            delete t;
            throw;
        }
        return t;
    });
}

void sqbem_sphere_free(sqbem_sphere* t) { delete t; }

sqbem_swimmer* sqbem_swimmer_create_moved(const sqbem_tube* tube, const sqbem_sphere* sph, const double* shift,
                                          int precision) {
    return guarded_ptr<sqbem_swimmer>([&] {
        SQ_REQUIRE(sph && shift, SQBEM_ERR_ARG, "sqbem_swimmer_create_moved: bad arguments");
        SQ_REQUIRE(precision >= SQBEM_FP64 && precision <= SQBEM_BF16_IR, SQBEM_ERR_ARG, "unknown precision");
        SQ_REQUIRE(!tube || tube->prec == precision || !mixed(precision), SQBEM_ERR_ARG,
                   "swimmer precision must match the tube's");
        SQ_REQUIRE(!tube || mixed(tube->prec) == mixed(precision), SQBEM_ERR_ARG,
                   "swimmer precision must match the tube's (FP64 vs refined)");
        SQ_REQUIRE(!tube || tube->mu == sph->mu, SQBEM_ERR_ARG, "sphere template and tube differ in viscosity");
        auto* s = new sqbem_swimmer;
        try {
            s->tube = tube;
            s->tmpl = sph;
            s->S = sph->S;
            s->S.translate(bem::V3{shift[0], shift[1], shift[2]});
            for (int c = 0; c < 3; ++c) s->xc[c] = sph->xc[c] + shift[c];
            s->mu = sph->mu;
            s->prec = precision;
            swimmer_build(s);
        } catch (...) {
            delete s;
            throw;
        }
        return s;
    });
}

static void unpack_out(sqbem_swimmer* s, int k, const std::vector<double>& X, const std::vector<double>& ft, double* U,
                       double* Om, double* fs, double* ftube) {
    const int ns = s->ns, n = ns + 6, nt = s->nt;
    for (int j = 0; j < k; ++j) {
        const double* x = X.data() + (size_t)j * n;
        if (U)
            for (int c = 0; c < 3; ++c) U[3 * j + c] = x[ns + c];
        if (Om)
            for (int c = 0; c < 3; ++c) Om[3 * j + c] = x[ns + 3 + c];
        if (fs) std::memcpy(fs + (size_t)j * ns, x, sizeof(double) * ns);
        if (ftube && nt) std::memcpy(ftube + (size_t)j * nt, ft.data() + (size_t)j * nt, sizeof(double) * nt);
    }
}

int sqbem_swimmer_solve_free(sqbem_swimmer* s, int nslip, const double* slips, double* U, double* Om, double* fs,
                             double* ft) {
    return guarded([&] {
        SQ_REQUIRE(s && slips && nslip > 0 && nslip <= 8, SQBEM_ERR_ARG, "solve_free: 1..8 slip fields required");
        const int n = s->ns + 6;
        std::vector<double> X((size_t)n * nslip), FT((size_t)std::max(s->nt, 1) * nslip);
        if (mixed(s->prec))
            solve_mixed(s, nslip, slips, false, nullptr, X.data(), FT.data());
        else
            solve_fp64(s, nslip, slips, false, nullptr, X.data(), FT.data());
        unpack_out(s, nslip, X, FT, U, Om, fs, ft);
    });
}

int sqbem_swimmer_solve_prescribed(sqbem_swimmer* s, const double* U, const double* Om, double* fs, double* ft) {
    return guarded([&] {
        SQ_REQUIRE(s && U && Om, SQBEM_ERR_ARG, "solve_prescribed: null argument");
        const int n = s->ns + 6;
        double w[6] = {U[0], U[1], U[2], Om[0], Om[1], Om[2]};
        std::vector<double> X((size_t)n), FT((size_t)std::max(s->nt, 1));
        if (mixed(s->prec))
            solve_mixed(s, 1, nullptr, true, w, X.data(), FT.data());
        else
            solve_fp64(s, 1, nullptr, true, w, X.data(), FT.data());
        unpack_out(s, 1, X, FT, nullptr, nullptr, fs, ft);
    });
}

int sqbem_swimmer_gradient(sqbem_swimmer* s, const double* e, double B1, double B2, int nfun, const double* g,
                           double* U, double* Om, double* dxc, double* de, double* dB, double* du) {
    return guarded([&] {
        SQ_REQUIRE(s && e && g && nfun > 0 && nfun <= 8, SQBEM_ERR_ARG, "gradient: 1..8 functionals required");
        SQ_REQUIRE(!mixed(s->prec), SQBEM_ERR_ARG, "gradient: the adjoint runs on an FP64 swimmer");
        const int Ns = s->Ns, ns = s->ns, n = ns + 6;
        const bem::V3 ev{e[0], e[1], e[2]}, xc{s->xc[0], s->xc[1], s->xc[2]};
        std::vector<double> u(ns), X(n), dxc_h((size_t)3 * nfun), du_h((size_t)ns * nfun);
        for (int p = 0; p < Ns; ++p) {
            const bem::V3 v = bem::squirmer_slip(s->S.centroid[p], xc, ev, B1, B2);
            u[3 * (size_t)p] = v.x;
            u[3 * (size_t)p + 1] = v.y;
            u[3 * (size_t)p + 2] = v.z;
        }
        adjoint_fp64(s, u.data(), nfun, g, X.data(), nullptr, dxc_h.data(), du_h.data());
        if (U)
            for (int c = 0; c < 3; ++c) U[c] = X[ns + c];
        if (Om)
            for (int c = 0; c < 3; ++c) Om[c] = X[ns + 3 + c];
        if (dxc) std::memcpy(dxc, dxc_h.data(), sizeof(double) * 3 * nfun);
        if (du) std::memcpy(du, du_h.data(), sizeof(double) * ns * nfun);
        // chain rule through u_p = (B1 + B2 c)(c rh - e), c = rh . e (src/utils/kernels.hpp:107-111)
        for (int j = 0; j < nfun; ++j) {
            double ge[3] = {0, 0, 0}, g1 = 0, g2 = 0;
            const double* gu = du_h.data() + (size_t)j * ns;
            for (int p = 0; p < Ns; ++p) {
                const bem::V3 rh = bem::unit(s->S.centroid[p] - xc);
                const double c = bem::dot(rh, ev);
                const double r[3] = {rh.x, rh.y, rh.z}, ee[3] = {ev.x, ev.y, ev.z};
                const double* gp = gu + 3 * (size_t)p;
                double gt = 0, gr = 0;  // g . (c rh - e), g . rh
                for (int a = 0; a < 3; ++a) {
                    gt += gp[a] * (c * r[a] - ee[a]);
                    gr += gp[a] * r[a];
                }
                g1 += gt;
                g2 += c * gt;
                for (int k = 0; k < 3; ++k) ge[k] += B2 * r[k] * gt + (B1 + B2 * c) * (r[k] * gr - gp[k]);
            }
            if (de)
                for (int k = 0; k < 3; ++k) de[3 * j + k] = ge[k];
            if (dB) {
                dB[2 * j] = g1;
                dB[2 * j + 1] = g2;
            }
        }
    });
}

int sqbem_swimmer_times(const sqbem_swimmer* s, double* ta, double* ts, int* it, double* res) {
    return guarded([&] {
        SQ_REQUIRE(s, SQBEM_ERR_ARG, "null swimmer");
        if (ta) *ta = s->t_assembly;
        if (ts) *ts = s->t_schur;
        if (it) *it = s->ir_iters;
        if (res) *res = s->ir_residual;
    });
}

void sqbem_swimmer_free(sqbem_swimmer* s) { delete s; }

int sqbem_velocity_at(const sqbem_mesh* m, const double* u, const double* f, double mu, const double* pts, int npts,
                      double* out) {
    return guarded([&] {
        SQ_REQUIRE(m && u && f && (npts == 0 || (pts && out)), SQBEM_ERR_ARG, "velocity_at: null argument");
        if (npts == 0) return;
        DMesh d(m->m);
        const int P = m->m.P();
        DBuf<double> du((size_t)3 * P), df((size_t)3 * P), dp((size_t)3 * npts), dout((size_t)3 * npts);
        du.upload(u, (size_t)3 * P);
        df.upload(f, (size_t)3 * P);
        dp.upload(pts, (size_t)3 * npts);
        field(d, dp.p, npts, du.p, df.p, mu, dout.p);
        dout.download(out, (size_t)3 * npts);
    });
}

sqbem_duct* sqbem_duct_create(const sqbem_mesh* mesh, double mu) {
    return guarded_ptr<sqbem_duct>([&] {
        SQ_REQUIRE(mesh && mesh->m.P() > 0, SQBEM_ERR_ARG, "sqbem_duct_create: empty mesh");
        auto* d = new sqbem_duct;
        try {
            d->mesh = mesh->m;
            d->mu = mu;
            duct_build(d);
        } catch (...) {
            delete d;
            throw;
        }
        return d;
    });
}

int sqbem_duct_solve(const sqbem_duct* d, double p_in, double p_out, double* u, double* f, double* Q) {
    return guarded([&] {
        SQ_REQUIRE(d, SQBEM_ERR_ARG, "null duct");
        const int P = d->P, n = d->n, nc = (int)d->cap.size();
        std::vector<double> fk(nc);
        for (int k = 0; k < nc; ++k) fk[k] = d->mesh.tags[d->cap[k]] == bem::INLET ? p_in : -p_out;
        DBuf<double> fkd(std::max(nc, 1)), rhs(n);
        if (nc) fkd.upload(fk.data(), nc);
        gemv<double>(n, nc, 1, -1.0, d->Fk.p, n, fkd.p, std::max(nc, 1), 0.0, rhs.p, n);
        getrs<double>(SQBEM_FP64, 0, n, 1, d->LU.p, n, d->ipiv.p, rhs.p, n);
        std::vector<double> x(n);
        rhs.download(x.data(), n);
        std::vector<double> uu((size_t)3 * P, 0.0), ff((size_t)3 * P, 0.0);
        for (int p = 0; p < P; ++p)
            for (int b = 0; b < 3; ++b) {
                bool known = !(d->mesh.tags[p] != bem::WALL && b == 2);
                (known ? ff : uu)[3 * (size_t)p + b] = x[3 * (size_t)p + b];
            }
        double q = 0;
        for (int k = 0; k < nc; ++k) {
            int p = d->cap[k];
            ff[3 * (size_t)p + 2] = fk[k];
            if (d->mesh.tags[p] == bem::OUTLET) q += uu[3 * (size_t)p + 2] * d->mesh.area[p];
        }
        if (u) std::memcpy(u, uu.data(), sizeof(double) * 3 * P);
        if (f) std::memcpy(f, ff.data(), sizeof(double) * 3 * P);
        if (Q) *Q = q;
    });
}

void sqbem_duct_free(sqbem_duct* d) { delete d; }

}  // extern "C"
