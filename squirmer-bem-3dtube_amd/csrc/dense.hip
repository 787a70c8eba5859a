// dense.hip -- dense linear algebra for gfx950: MFMA GEMMs, partial-pivoting LU, triangular
// solves. Replaces the OpenBLAS calls of src/utils/linalg.hpp:17-77 (dgemm_, dgetrf_,
// dgetrs_, dgesv_).
//
// Design:
//  * GEMM (C = alpha A B + beta C, column-major): 128x128 workgroup tiles, four waves of
//    64x64, double-buffered LDS, MFMA 16x16x4 for FP64 and FP32, MFMA 16x16x32 for
//    FP16/BF16 with FP32 accumulation and per-row/per-column power-of-two scaling.
//    Workgroup ids are remapped so each XCD (own L2) works on a contiguous, grouped
//    range of output tiles.
//  * LU: recursive (Toledo) split of the columns, so almost all flops are large GEMMs.
//    Leaves of <= 32 columns are factored by one cooperative kernel: the panel rows are
//    spread over up to one workgroup per CU, held in LDS, and each column's pivot is
//    chosen with one grid-wide barrier (double-buffered candidate rows). Pivot search,
//    swap and scaling are done in FP64 whatever the storage type.
//  * TRSM: recursive split with 64-row leaf kernels (one thread per right-hand side,
//    the triangle in LDS), GEMM updates in between.
#include <algorithm>
#include <array>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <limits>
#include <climits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include <hipblaslt/hipblaslt-ext.hpp>
#include <hipblaslt/hipblaslt.h>
#include <rocblas/rocblas.h>

#include "common.hpp"
#include "dense.hpp"

namespace sq {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ tile scheduling
// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range
// of logical tiles, then walk tiles in groups of GROUP_M tile-rows so neighbours in time
// share A row-panels and B column-panels in that XCD's L2.
__device__ __forceinline__ void tile_of(int bid, int nblocks, int tiles_m, int& tm, int& tn) {
    constexpr int XCD = 8, GROUP_M = 8;
    int xcd = bid % XCD, idx = bid / XCD;
    int per = nblocks / XCD, rem = nblocks % XCD;
    int logical = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    int tiles_n = nblocks / tiles_m;
    int in_group = GROUP_M * tiles_n;
    int g = logical / in_group;
    int first = g * GROUP_M;
    int gsz = min(tiles_m - first, GROUP_M);
    int r = logical % in_group;
    tm = first + r % gsz;
    tn = r / gsz;
}

// Instruction order of one k step of a pipelined GEMM main loop: each of its NM MFMAs is
// followed by one of the step's memory operations, the next step's fragment reads (NR) first,
// then the step's share of the next stage's LDS writes (NW) and of a later tile's global
// loads (NG); what is left over follows the last MFMA. Only the order changes, never the
// values.
template <int NM, int NR, int NW, int NG>
__device__ __forceinline__ void mfma_interleave() {
    if constexpr (NM > 0) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
        if constexpr (NR > 0) {
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
            mfma_interleave<NM - 1, NR - 1, NW, NG>();
        } else if constexpr (NW > 0) {
            __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // one LDS write
            mfma_interleave<NM - 1, 0, NW - 1, NG>();
        } else if constexpr (NG > 0) {
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // one global load
            mfma_interleave<NM - 1, 0, 0, NG - 1>();
        } else {
            mfma_interleave<NM - 1, 0, 0, 0>();
        }
    } else {
        if constexpr (NR > 0) __builtin_amdgcn_sched_group_barrier(0x100, NR, 0);
        if constexpr (NW > 0) __builtin_amdgcn_sched_group_barrier(0x200, NW, 0);
        if constexpr (NG > 0) __builtin_amdgcn_sched_group_barrier(0x020, NG, 0);
    }
}

// f(integral_constant<int, 0>), ..., f(integral_constant<int, N - 1>) in order
template <int N, int I = 0, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

// ------------------------------------------------------------------ FP64 / FP32 GEMM
template <typename T>
struct MfmaOp;
// Accumulator row of value r in lane group lk = lane/16 (column is lane%16), measured on
// gfx950: FP64 16x16x4 interleaves the groups (row lk + 4r), the FP32/FP16/BF16 forms
// keep them in blocks (row 4lk + r).
template <>
struct MfmaOp<double> {
    using acc = d4v;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct MfmaOp<float> {
    using acc = f4v;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

constexpr int GBM = 128, GBN = 128;
// k tile of the FP32/FP64 GEMM: 32 (FP32) or 16 (FP64) values, so that a short product
// (k = 128: the leaves of the factorisation and of the triangular solves) waits on few
// global-load round trips; both keep the two LDS stages within 70 KiB. Row padding: FP32
// rows of 34 words put the 16 rows x 2 k of a half-wave's operand read in 32 distinct
// banks (33-word rows pair them up: measured 0.4 conflicts per LDS cycle)
template <typename T>
struct GemmK {
    static constexpr int KB = sizeof(T) == 4 ? 32 : 16, PAD = sizeof(T) == 4 ? KB + 2 : KB + 1;
};

// MFMA 16x16x4 operand layout: lane l supplies A[l%16][l/16] and B[l/16][l%16]; the four
// accumulator values of lane l are D[Op::row(l/16, r)][l%16].
// Workgroup tile TM x TN (TM = 128: TN = 128, or 64 for products whose 128 x 128 tiles would
// leave CUs idle, four waves of 64 x TN/2; TM = 64: four waves of 64 x TN/4, for products
// whose 128-row tiles would give the CUs under two workgroups each).
// The LDS block of gemm_nn_tile, one per (element type, size): tiles of equal size (128 x 64
// and 64 x 128) used by one kernel share it.
template <typename T, int ELEMS>
__device__ __forceinline__ T* gemm_lds() {
    __shared__ __attribute__((aligned(16))) T block[ELEMS];
    return block;
}

template <typename T, int TN = GBN, int TM = GBM>
__device__ __forceinline__ void gemm_nn_tile(int m, int n, int k, T alpha, const T* __restrict__ A, int lda,
                                             const T* __restrict__ B, int ldb, T beta, T* __restrict__ C, int ldc,
                                             int tm, int tn, int* unit_flag = nullptr, __bf16* sh_col = nullptr,
                                             int sh_cld = 0, __bf16* sh_row = nullptr, int sh_rld = 0) {
    using Op = MfmaOp<T>;
    using Acc = typename Op::acc;
    static_assert(TM == 128 || TM == 64, "row tile of 64 or 128");
    constexpr int KB = GemmK<T>::KB, PAD = GemmK<T>::PAD, AS = 256 / TM, NQ = TM * KB / 256, JS = 256 / KB,
                  NQB = TN * KB / 256;
    constexpr int JT = TM == 128 ? TN / 32 : TN / 64;  // 16-column MFMA tiles per wave
    // one LDS block: the two operand stages during the k loop, the C tile in the epilogue
    constexpr int SMEM = 2 * TM * PAD + 2 * TN * PAD;
    T* smem = gemm_lds<T, SMEM>();
    T(*As)[TM][PAD] = reinterpret_cast<T(*)[TM][PAD]>(smem);
    T(*Bs)[TN][PAD] = reinterpret_cast<T(*)[TN][PAD]>(smem + 2 * TM * PAD);

    const int m0 = tm * TM, n0 = tn * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = TM == 128 ? (wave & 1) * 64 : 0, wn = TM == 128 ? (wave >> 1) * (TN / 2) : wave * (TN / 4);

    // global -> register staging: A column kk has TM contiguous rows; B column j has KB
    // contiguous k values. Loads read clamped (always valid) addresses and are masked when
    // written into LDS, so no instruction waits for a load before the LDS write needs it.
    const int a_i = tid % TM, a_k0 = tid / TM;     // rows a_i, k = a_k0 + AS q
    const int b_k = tid % KB, b_j0 = tid / KB;      // k = b_k, cols b_j0 + JS q
    const int gi = m0 + a_i, gic = min(gi, m - 1);
    T ra[NQ], rb[NQB];
    auto load_a = [&](int k0, int q) { ra[q] = A[gic + (size_t)min(k0 + a_k0 + AS * q, k - 1) * lda]; };
    auto load_b = [&](int k0, int q) {
        rb[q] = B[min(k0 + b_k, k - 1) + (size_t)min(n0 + b_j0 + JS * q, n - 1) * ldb];
    };
    auto store = [&](int buf, int k0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) As[buf][a_i][a_k0 + AS * q] = (gi < m && k0 + a_k0 + AS * q < k) ? ra[q] : T(0);
#pragma unroll
        for (int q = 0; q < NQB; ++q)
            Bs[buf][b_j0 + JS * q][b_k] = (k0 + b_k < k && n0 + b_j0 + JS * q < n) ? rb[q] : T(0);
    };

    Acc acc[4][JT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < JT; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    const int nk = (k + KB - 1) / KB;
#pragma unroll
    for (int q = 0; q < NQ; ++q) load_a(0, q);
#pragma unroll
    for (int q = 0; q < NQB; ++q) load_b(0, q);
    store(0, 0);
    __syncthreads();
    const int lr = lane & 15, lk = lane >> 4;
    // k tile kt in S steps of 4: each step reads the next step's fragments (two fragment
    // sets) and issues 1/S of tile kt + 1's global loads; tile kt + 1 goes into the other
    // LDS stage after the last step
    constexpr int S = KB / 4, QA = NQ / S, QB = NQB / S;
    static_assert(QA * S == NQ && QB * S == NQB, "loads do not divide over the k steps");
    auto step = [&](auto load_tag, int kt) {
        constexpr bool LD = decltype(load_tag)::value;
        const int buf = kt & 1, k1 = (kt + 1) * KB;
        T af[2][4], bf[2][JT];
        auto frags = [&](int s, T* a_, T* b_) {
#pragma unroll
            for (int i = 0; i < 4; ++i) a_[i] = As[buf][wm + i * 16 + lr][4 * s + lk];
#pragma unroll
            for (int j = 0; j < JT; ++j) b_[j] = Bs[buf][wn + j * 16 + lr][4 * s + lk];
        };
        frags(0, af[0], bf[0]);
        auto sub = [&](auto s_tag) {
            constexpr int s = decltype(s_tag)::value;
            if constexpr (s + 1 < S) frags(s + 1, af[(s + 1) & 1], bf[(s + 1) & 1]);
            if constexpr (LD) {
#pragma unroll
                for (int q = 0; q < QA; ++q) load_a(k1, s * QA + q);
#pragma unroll
                for (int q = 0; q < QB; ++q) load_b(k1, s * QB + q);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < JT; ++j) acc[i][j] = Op::mma(af[s & 1][i], bf[s & 1][j], acc[i][j]);
            mfma_interleave<4 * JT, s + 1 < S ? 4 + JT : 0, 0, LD ? QA + QB : 0>();
        };
        static_for<S>(sub);
        if constexpr (LD) store(buf ^ 1, k1);
        __syncthreads();
    };
    int kt = 0;
    for (; kt + 1 < nk; ++kt) step(std::true_type{}, kt);
    step(std::false_type{}, kt);

    // epilogue through LDS (the operand stages are free once the k loop has ended): the
    // accumulators are written into a column-major tile of CW columns at a time, then every
    // thread updates C down columns, consecutive rows on consecutive lanes, with all of its
    // C values read before any result is stored
    constexpr int LDC = TM + 4, CW = SMEM / LDC >= TN ? TN : TN / 2;
    static_assert(CW * LDC <= SMEM, "epilogue tile exceeds the LDS block");
    T* Ct = smem;
    const int er = tid % TM, ec = tid / TM;  // row, first column (then every AS-th)
    int bad = 0;
#pragma unroll
    for (int c0 = 0; c0 < TN; c0 += CW) {
        // the LDS may still be read by the k loop's last step (c0 = 0) or by the previous
        // chunk's readers
        __syncthreads();
#pragma unroll
        for (int j = 0; j < JT; ++j) {
            const int cl = wn + j * 16 + lr - c0;
            if (cl < 0 || cl >= CW) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) Ct[cl * LDC + wm + i * 16 + Op::row(lk, r)] = acc[i][j][r];
        }
        __syncthreads();
        const int row = m0 + er;
        T cv[CW / AS];
        if (beta != T(0)) {
#pragma unroll
            for (int q = 0; q < CW / AS; ++q)
                cv[q] = C[min(row, m - 1) + (size_t)min(n0 + c0 + ec + AS * q, n - 1) * ldc];
        }
#pragma unroll
        for (int q = 0; q < CW / AS; ++q) {
            const int col = n0 + c0 + ec + AS * q;
            if (row < m && col < n) {
                const T v = alpha * Ct[(ec + AS * q) * LDC + er];
                const T out = beta == T(0) ? v : v + beta * cv[q];
                C[row + (size_t)col * ldc] = out;
                if (sh_col) sh_col[row + (size_t)col * sh_cld] = (__bf16)out;
                if (unit_flag && !(fabs((double)out) <= 1.0)) bad = 1;
            }
        }
        if (sh_row) {
            // the row-major BF16 copy (beta = 0: the stored values are alpha times the tile),
            // consecutive columns on consecutive lanes
            for (int e = tid; e < TM * CW; e += 256) {
                const int r = e / CW, c = e % CW, row = m0 + r, col = n0 + c0 + c;
                if (row < m && col < n) sh_row[(size_t)row * sh_rld + col] = (__bf16)(alpha * Ct[c * LDC + r]);
            }
        }
    }
    // unit_flag |= (some stored |value| > 1 or not finite): the no-pivot LU's multiplier check
    if (unit_flag && __syncthreads_or(bad) && tid == 0) atomicOr(unit_flag, 1);
}

template <typename T, int TN = GBN, int TM = GBM>
__global__ void __launch_bounds__(256) gemm_nn_kernel(int m, int n, int k, T alpha, const T* __restrict__ A,
                                                      int lda, const T* __restrict__ B, int ldb, T beta,
                                                      T* __restrict__ C, int ldc, int tiles_m, int* unit_flag,
                                                      __bf16* sh_col, int sh_cld, __bf16* sh_row, int sh_rld) {
    int tm, tn;
    tile_of(blockIdx.x, gridDim.x, tiles_m, tm, tn);
    gemm_nn_tile<T, TN, TM>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, unit_flag, sh_col, sh_cld, sh_row,
                            sh_rld);
}

// Many independent GEMMs C_j = alpha_j A_j B_j + beta_j C_j in one launch: workgroup
// (x, y) computes tile x (row tiles fastest) of job y and returns when job y has fewer.
template <typename T>
__global__ void __launch_bounds__(256) gemm_nn_batch_kernel(const GemmJob<T>* __restrict__ jobs) {
    const GemmJob<T>& g = jobs[blockIdx.y];
    const int tiles_m = (g.m + GBM - 1) / GBM, tiles_n = (g.n + GBN - 1) / GBN;
    if ((int)blockIdx.x >= tiles_m * tiles_n) return;
    gemm_nn_tile<T>(g.m, g.n, g.k, g.alpha, g.A, g.lda, g.B, g.ldb, g.beta, g.C, g.ldc, blockIdx.x % tiles_m,
                    blockIdx.x / tiles_m);
}

// ------------------------------------------------------------------ 16-bit MFMA GEMM
// FP32 operands in global memory, converted to FP16/BF16 when staged into LDS, MFMA
// 16x16x32 with FP32 accumulation. Optional per-row scale of A and per-column scale of B
// (powers of two, so the unscaling in the epilogue is exact) keep FP16 values in range.
constexpr int HBK = 64, HPAD = HBK + 8;

template <bool BF16>
struct Half;
typedef float f16x __attribute__((ext_vector_type(16)));
template <>
struct Half<false> {
    using t = _Float16;
    using v8 = h8v;
    static __device__ __forceinline__ f4v mma(v8 a, v8 b, f4v c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ f16x mma32(v8 a, v8 b, f16x c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Half<true> {
    using t = __bf16;
    using v8 = b8v;
    static __device__ __forceinline__ f4v mma(v8 a, v8 b, f4v c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ f16x mma32(v8 a, v8 b, f16x c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

// MFMA 16x16x32 operand layout: lane l supplies A[l%16][8*(l/16)+0..7] and
// B[8*(l/16)+0..7][l%16]; accumulator value r of lane l is D[4*(l/16)+r][l%16].
template <bool BF16>
__global__ void __launch_bounds__(256) gemm16_nn_kernel(int m, int n, int k, float alpha,
                                                        const float* __restrict__ A, int lda,
                                                        const float* __restrict__ B, int ldb, float beta,
                                                        float* __restrict__ C, int ldc, int tiles_m,
                                                        const float* __restrict__ rs,
                                                        const float* __restrict__ cs) {
    using H = Half<BF16>;
    using ht = typename H::t;
    using v8 = typename H::v8;
    __shared__ __attribute__((aligned(16))) ht As[2][GBM][HPAD];
    __shared__ __attribute__((aligned(16))) ht Bs[2][GBN][HPAD];

    int tm, tn;
    tile_of(blockIdx.x, gridDim.x, tiles_m, tm, tn);
    const int m0 = tm * GBM, n0 = tn * GBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;

    // A tile 128 x 64: thread holds row a_i (consecutive threads, consecutive rows of a
    // column) at 8 consecutive k of chunks a_c0 + 2q, stored as one 16-byte LDS write each
    const int a_i = tid & 127, a_c0 = tid >> 7;
    // B tile 64 x 128: 16 threads per column, 4 consecutive k each, one 8-byte write each
    const int b_k = (tid & 15) * 4, b_j0 = tid >> 4;
    float ra[4][8];
    float rb[8][4];
    const float sa = (rs && m0 + a_i < m) ? rs[m0 + a_i] : 1.0f;

    auto load = [&](int k0) {
        const int gi = m0 + a_i;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + 8 * (a_c0 + 2 * q) + e;
                ra[q][e] = (gi < m && gk < k) ? A[gi + (size_t)gk * lda] : 0.0f;
            }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int gj = n0 + b_j0 + 16 * q;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int gk = k0 + b_k + e;
                rb[q][e] = (gk < k && gj < n) ? B[gk + (size_t)gj * ldb] : 0.0f;
            }
        }
    };
    typedef ht h4v __attribute__((ext_vector_type(4)));
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (ht)(ra[q][e] * sa);
            *(v8*)&As[buf][a_i][8 * (a_c0 + 2 * q)] = v;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int j = b_j0 + 16 * q;
            const float sb = (cs && n0 + j < n) ? cs[n0 + j] : 1.0f;
            h4v v;
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = (ht)(rb[q][e] * sb);
            *(h4v*)&Bs[buf][j][b_k] = v;
        }
    };

    f4v acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4v{0, 0, 0, 0};

    const int nk = (k + HBK - 1) / HBK;
    load(0);
    store(0);
    __syncthreads();
    const int lr = lane & 15, lk = lane >> 4;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load((kt + 1) * HBK);
#pragma unroll
        for (int s = 0; s < HBK / 32; ++s) {
            v8 af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = *(const v8*)&As[buf][wm + i * 16 + lr][32 * s + 8 * lk];
#pragma unroll
            for (int j = 0; j < 4; ++j) bf[j] = *(const v8*)&Bs[buf][wn + j * 16 + lr][32 * s + 8 * lk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = H::mma(af[i], bf[j], acc[i][j]);
        }
        if (kt + 1 < nk) store(buf ^ 1);
        __syncthreads();
    }

    // epilogue through LDS, 64 columns (the tiles of waves 0-1, then of waves 2-3) at a
    // time in the A operand stages: column-major tile, then C updated down columns,
    // consecutive rows on consecutive lanes, the chunk's C values read before any store
    constexpr int LDC = GBM + 1, CW = 64, NQ = CW / 2;
    static_assert(CW * LDC * sizeof(float) <= sizeof(As), "epilogue tile exceeds the A stages");
    float* Ct = reinterpret_cast<float*>(&As[0][0][0]);
    const int er = tid & (GBM - 1), ec = tid >> 7;
    const int row = m0 + er;
    const float ur = rs ? 1.0f / rs[min(row, m - 1)] : 1.0f;
#pragma unroll
    for (int c0 = 0; c0 < GBN; c0 += CW) {
        __syncthreads();
        if (wn == c0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) Ct[(j * 16 + lr) * LDC + wm + i * 16 + 4 * lk + r] = acc[i][j][r];
        }
        __syncthreads();
        float cv[NQ];
        if (beta != 0.0f) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) cv[q] = C[min(row, m - 1) + (size_t)min(n0 + c0 + ec + 2 * q, n - 1) * ldc];
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int cl = ec + 2 * q, col = n0 + c0 + cl;
            if (row < m && col < n) {
                const float uc = cs ? 1.0f / cs[col] : 1.0f;
                const float v = alpha * (Ct[cl * LDC + er] * ur * uc);
                C[row + (size_t)col * ldc] = beta == 0.0f ? v : v + beta * cv[q];
            }
        }
    }
}

// Power-of-two scale per row of A (m x k) or per column of B (k x n) so the largest
// magnitude maps to about 2^14 (FP16 max is 65504). Row maxima: 64 x 64 tiles, one atomic
// max per row and tile on the float bit pattern (non-negative floats order like integers).
__global__ void row_absmax_kernel(int m, int k, const float* __restrict__ A, int lda, unsigned* __restrict__ mx) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c0 = blockIdx.y * 64 + (threadIdx.x >> 6) * 16;
    if (i >= m) return;
    float v = 0.0f;
    for (int c = c0; c < min(k, c0 + 16); ++c) v = fmaxf(v, fabsf(A[i + (size_t)c * lda]));
    atomicMax(mx + i, __float_as_uint(v));
}
__global__ void bits_to_pow2_scale(int m, const unsigned* __restrict__ mx, float* __restrict__ s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const float v = __uint_as_float(mx[i]);
    s[i] = v > 0.0f ? exp2f(14.0f - (float)ilogbf(v)) : 1.0f;
}
__global__ void pow2_scale_cols(int k, int n, const float* __restrict__ B, int ldb, float* __restrict__ s) {
    // one wave per column
    int j = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    int lane = threadIdx.x & 63;
    if (j >= n) return;
    float mx = 0.0f;
    for (int r = lane; r < k; r += 64) mx = fmaxf(mx, fabsf(B[r + (size_t)j * ldb]));
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if (lane == 0) s[j] = mx > 0.0f ? exp2f(14.0f - (float)ilogbf(mx)) : 1.0f;
}

// ------------------------------------------------------------------ 16-bit GEMM on converted operands
// For large products the FP32 operands are converted once into padded 16-bit copies:
// A as Ah[i * kp + k] (row i contiguous in k) and B as Bh[j * kp + k] (column j contiguous
// in k), rows and columns padded to multiples of 256 and k to a multiple of 64 with zeros,
// FP16 values scaled per row / column by powers of two. The GEMM streams 16-byte chunks
// into LDS with no conversion and no bounds checks in its main loop.
//   workgroup tile 256 x 256, four waves of 128 x 128, MFMA 32x32x16 (4 x 4 tiles a wave),
//   k tile 64, two LDS stages (one barrier per k tile: the next tile's chunks are written
//   into the other stage while this one is read), 144 KiB of LDS, one workgroup per CU.
// MFMA 32x32x16 layout (measured on gfx950): lane l supplies A[l%32][8(l/32)+0..7] and
// B[8(l/32)+0..7][l%32]; accumulator value r of lane l is D[8(r/4) + 4(l/32) + r%4][l%32].
constexpr int QBK = 64, QPAD = QBK + 8;

// Workgroup tile TM x TN (256 x 256 for large products, 128 x 128 for mid-sized ones), four
// waves of TM/2 x TN/2, each a grid of 32 x 32 MFMA tiles. PF register sets stage the global
// loads: with PF = 1 a k tile is loaded one tile before it is stored into LDS, with PF = 2
// two tiles before, so a load has two k tiles of MFMA work to arrive in (short-k products
// wait on the load latency otherwise).
template <bool BF16, int TM, int TN, int PF>
__global__ void __launch_bounds__(256) gemm16q_kernel(int m, int n, int kp, float alpha,
                                                      const typename Half<BF16>::t* __restrict__ Ah,
                                                      const typename Half<BF16>::t* __restrict__ Bh, float beta,
                                                      float* __restrict__ C, int ldc, int tiles_m,
                                                      const float* __restrict__ rs, const float* __restrict__ cs) {
    using H = Half<BF16>;
    using ht = typename H::t;
    using v8 = typename H::v8;
    constexpr int IM = TM / 64, JN = TN / 64;        // 32 x 32 tiles per wave
    constexpr int CA = TM * (QBK / 8) / 256, CB = TN * (QBK / 8) / 256;  // 16-byte chunks a thread
    // two stages of each operand tile; separate static arrays tell the compiler that A and B
    // tiles never alias (up to 144 KiB of the CU's 160 KiB of LDS)
    __shared__ __attribute__((aligned(16))) ht As_[2][TM][QPAD];
    __shared__ __attribute__((aligned(16))) ht Bs_[2][TN][QPAD];

    int tm, tn;
    tile_of(blockIdx.x, gridDim.x, tiles_m, tm, tn);
    const int m0 = tm * TM, n0 = tn * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * (TM / 2), wn = (wave >> 1) * (TN / 2);
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    // chunk c = tid + 256 q of a panel: row c >> 3, 8-element k chunk c & 7
    u4 ra[CA], rb[CB];                              // register set 0
    u4 ra1[PF == 2 ? CA : 1], rb1[PF == 2 ? CB : 1];  // register set 1 (PF = 2)
    auto load_into = [&](int k0, u4* A_, u4* B_) {
#pragma unroll
        for (int q = 0; q < CA; ++q) {
            const int c = tid + 256 * q, row = c >> 3, kc = c & 7;
            A_[q] = *(const u4*)(Ah + (size_t)(m0 + row) * kp + k0 + 8 * kc);
        }
#pragma unroll
        for (int q = 0; q < CB; ++q) {
            const int c = tid + 256 * q, row = c >> 3, kc = c & 7;
            B_[q] = *(const u4*)(Bh + (size_t)(n0 + row) * kp + k0 + 8 * kc);
        }
    };
    auto store_from = [&](int stage, const u4* A_, const u4* B_) {
#pragma unroll
        for (int q = 0; q < CA; ++q) {
            const int c = tid + 256 * q, row = c >> 3, kc = c & 7;
            *(u4*)&As_[stage][row][8 * kc] = A_[q];
        }
#pragma unroll
        for (int q = 0; q < CB; ++q) {
            const int c = tid + 256 * q, row = c >> 3, kc = c & 7;
            *(u4*)&Bs_[stage][row][8 * kc] = B_[q];
        }
    };
    f16x acc[IM][JN];
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const int nk = kp / QBK;
    const int lr = lane & 31, lk = lane >> 5;
    // k tile kt: its four 16-deep steps each read the next step's fragments (two fragment
    // sets), write a quarter of tile kt + 1 into the other LDS stage (last read before the
    // previous barrier) and load a quarter of tile kt + 1 + PF into the register set just
    // written out (tile t waits in set t % PF); mfma_interleave places one of these operations
    // after each MFMA.
    constexpr int QA = CA / 4, QB = CB / 4;
    auto step = [&](auto store_tag, auto load_tag, auto set_tag, int kt) {
        constexpr bool ST = decltype(store_tag)::value, LD = decltype(load_tag)::value;
        constexpr int SET = decltype(set_tag)::value;
        u4* ra_ = SET == 0 ? ra : ra1;
        u4* rb_ = SET == 0 ? rb : rb1;
        const int st = kt & 1;
        v8 af[2][IM], bf[2][JN];
        auto frags = [&](int s2, v8* a_, v8* b_) {
#pragma unroll
            for (int i = 0; i < IM; ++i) a_[i] = *(const v8*)&As_[st][wm + i * 32 + lr][16 * s2 + 8 * lk];
#pragma unroll
            for (int j = 0; j < JN; ++j) b_[j] = *(const v8*)&Bs_[st][wn + j * 32 + lr][16 * s2 + 8 * lk];
        };
        frags(0, af[0], bf[0]);
        auto sub = [&](auto s2_tag) {
            constexpr int s2 = decltype(s2_tag)::value;
            if constexpr (s2 < 3) frags(s2 + 1, af[(s2 + 1) & 1], bf[(s2 + 1) & 1]);
            if constexpr (ST) {
#pragma unroll
                for (int qq = 0; qq < QA; ++qq) {
                    const int q = s2 * QA + qq, c = tid + 256 * q, row = c >> 3, kc = c & 7;
                    *(u4*)&As_[st ^ 1][row][8 * kc] = ra_[q];
                }
#pragma unroll
                for (int qq = 0; qq < QB; ++qq) {
                    const int q = s2 * QB + qq, c = tid + 256 * q, row = c >> 3, kc = c & 7;
                    *(u4*)&Bs_[st ^ 1][row][8 * kc] = rb_[q];
                }
            }
            if constexpr (LD) {
                const int k0 = (kt + 1 + PF) * QBK;
#pragma unroll
                for (int qq = 0; qq < QA; ++qq) {
                    const int q = s2 * QA + qq, c = tid + 256 * q, row = c >> 3, kc = c & 7;
                    ra_[q] = *(const u4*)(Ah + (size_t)(m0 + row) * kp + k0 + 8 * kc);
                }
#pragma unroll
                for (int qq = 0; qq < QB; ++qq) {
                    const int q = s2 * QB + qq, c = tid + 256 * q, row = c >> 3, kc = c & 7;
                    rb_[q] = *(const u4*)(Bh + (size_t)(n0 + row) * kp + k0 + 8 * kc);
                }
            }
#pragma unroll
            for (int i = 0; i < IM; ++i)
#pragma unroll
                for (int j = 0; j < JN; ++j) acc[i][j] = H::mma32(af[s2 & 1][i], bf[s2 & 1][j], acc[i][j]);
            mfma_interleave<IM * JN, s2 < 3 ? IM + JN : 0, ST ? QA + QB : 0, LD ? QA + QB : 0>();
        };
        sub(std::integral_constant<int, 0>{});
        sub(std::integral_constant<int, 1>{});
        sub(std::integral_constant<int, 2>{});
        sub(std::integral_constant<int, 3>{});
        __syncthreads();
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, PF == 2 ? 1 : 0>;
    load_into(0, ra, rb);
    store_from(0, ra, rb);
    if (nk > 1) load_into(QBK, PF == 2 ? ra1 : ra, PF == 2 ? rb1 : rb);
    if (PF == 2 && nk > 2) load_into(2 * QBK, ra, rb);
    __syncthreads();
    int kt = 0;
    if constexpr (PF == 1) {
        for (; kt + 2 < nk; ++kt) step(T_{}, T_{}, S0{}, kt);
        if (kt + 1 < nk) {
            step(T_{}, F_{}, S0{}, kt);
            ++kt;
        }
        step(F_{}, F_{}, S0{}, kt);
    } else {
        for (; kt + 4 < nk; kt += 2) {  // tiles kt + 1 (set 1) and kt + 2 (set 0) stored
            step(T_{}, T_{}, S1{}, kt);
            step(T_{}, T_{}, S0{}, kt + 1);
        }
        if (kt + 3 < nk) {  // kt even: the last load, tile kt + 3, into set 0 after storing set 1
            step(T_{}, T_{}, S1{}, kt);
            ++kt;
        }
        for (; kt + 1 < nk; ++kt) {  // stores only
            if ((kt + 1) & 1) step(T_{}, F_{}, S1{}, kt);
            else step(T_{}, F_{}, S0{}, kt);
        }
        step(F_{}, F_{}, S0{}, kt);
    }
    // epilogue through LDS, 64 columns at a time in the A operand stages (free once the k
    // loop has ended): the chunk's accumulators are written as a column-major tile, then
    // every thread updates C down columns, consecutive rows on consecutive lanes, with the
    // chunk's C values read before any result is stored (unscaling in the same order)
    constexpr int LDC = TM + 1, CW = 64, CS = 256 / TM, NQ = CW / CS, QG = NQ < 32 ? NQ : 32;
    static_assert(CW * LDC * sizeof(float) <= sizeof(As_), "epilogue tile exceeds the A stages");
    float* Ct = reinterpret_cast<float*>(&As_[0][0][0]);
    const int er = tid % TM, ec = tid / TM;  // row; first column, then every CS-th
    const int row = m0 + er;
    const float ur = rs ? 1.0f / rs[min(row, m - 1)] : 1.0f;
#pragma unroll
    for (int c0 = 0; c0 < TN; c0 += CW) {
        __syncthreads();  // the previous chunk's readers are done
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const int cl = wn + j * 32 - c0;  // tile j lies wholly inside or outside the chunk
            if (cl < 0 || cl >= CW) continue;
#pragma unroll
            for (int i = 0; i < IM; ++i)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    Ct[(cl + lr) * LDC + wm + i * 32 + 8 * (r >> 2) + 4 * lk + (r & 3)] = acc[i][j][r];
        }
        __syncthreads();
#pragma unroll
        for (int q0 = 0; q0 < NQ; q0 += QG) {  // QG columns' C values in flight at a time
            float cv[QG];
            if (beta != 0.0f) {
#pragma unroll
                for (int q = 0; q < QG; ++q)
                    cv[q] = C[min(row, m - 1) + (size_t)min(n0 + c0 + ec + CS * (q0 + q), n - 1) * ldc];
            }
#pragma unroll
            for (int q = 0; q < QG; ++q) {
                const int cl = ec + CS * (q0 + q), col = n0 + c0 + cl;
                if (row < m && col < n) {
                    const float uc = cs ? 1.0f / cs[col] : 1.0f;
                    const float v = alpha * (Ct[cl * LDC + er] * ur * uc);
                    C[row + (size_t)col * ldc] = beta == 0.0f ? v : v + beta * cv[q];
                }
            }
        }
    }
}

// ------------------------------------------------------------------ 16-bit GEMM, operands DMA'd into LDS
// The same product as gemm16q_kernel on the same converted operands, with the k tiles
// copied from global memory into LDS by the CU's load-to-LDS path (global_load_lds_dwordx4:
// one wave instruction fills one contiguous 1 KiB LDS block, lane l at byte 16 l). No VGPRs
// stage the operands, so NS stages can be in flight: the loads of tile t + NS - 1 are
// issued as tile t starts. A stage holds TM + TN rows of KT 16-bit values, KT / 8 chunks of
// 16 bytes a row; chunk kc of row r sits in slot kc ^ swz(r), which puts the lanes of a
// fragment read (consecutive rows, one chunk each) on different 16-byte bank groups.
// Each stage is its own LDS array, so the compiler's own bookkeeping of the DMA writes can
// tell the stages apart and never waits on the loads of a later stage before reading one.
typedef __attribute__((address_space(3))) void lds_void_t;

// s_waitcnt immediate that waits until at most n vector-memory operations are outstanding
// (gfx9 encoding: vmcnt in bits 3:0 and 15:14; expcnt 7 and lgkmcnt 15 mean no wait)
constexpr unsigned vmcnt_imm(unsigned n) { return (n & 15u) | ((n >> 4) << 14) | (7u << 4) | (15u << 8); }

template <bool BF16, int TM, int TN, int KT, int NS>
__global__ void __launch_bounds__(256) gemm16d_kernel(int m, int n, int kp, float alpha,
                                                      const typename Half<BF16>::t* __restrict__ Ah,
                                                      const typename Half<BF16>::t* __restrict__ Bh, float beta,
                                                      float* __restrict__ C, int ldc, int tiles_m,
                                                      const float* __restrict__ rs, const float* __restrict__ cs) {
    using H = Half<BF16>;
    using ht = typename H::t;
    using v8 = typename H::v8;
    constexpr int IM = TM / 64, JN = TN / 64;  // 32 x 32 MFMA tiles a wave (2 x 2 waves)
    constexpr int CPR = KT / 8;                 // 16-byte chunks per row of a stage
    constexpr int RPB = 64 / CPR;               // rows per 1 KiB DMA block
    constexpr int BA = TM / RPB, BB = TN / RPB; // DMA blocks of A and of B per stage
    constexpr int LPW = (BA + BB) / 4;          // DMA instructions per wave (per thread) per stage
    constexpr int SB = (TM + TN) * KT;          // 16-bit values per stage
    static_assert((BA + BB) % 4 == 0 && KT % 16 == 0 && NS >= 3 && NS <= 4, "gemm16d shape");
    static_assert(NS * SB * 2 <= 160 * 1024, "gemm16d stages exceed the LDS");
    constexpr int LDC = TM + 1, CW = 32;  // epilogue tile: CW columns of TM (+1 pad) FP32 values
    constexpr int S0N = SB > CW * LDC * 2 ? SB : CW * LDC * 2;
    __shared__ __attribute__((aligned(1024))) ht S0[S0N];
    __shared__ __attribute__((aligned(1024))) ht S1[SB];
    __shared__ __attribute__((aligned(1024))) ht S2[SB];
    __shared__ __attribute__((aligned(1024))) ht S3[NS > 3 ? SB : 8];
    auto stage = [&](auto s_) -> ht* {
        constexpr int s = decltype(s_)::value;
        if constexpr (s == 0) return S0;
        else if constexpr (s == 1) return S1;
        else if constexpr (s == 2) return S2;
        else return S3;
    };
    // conflict-free for both ways the LDS may serve a 16-byte read: 8 lanes on 32 banks of
    // 4 bytes, or 16 lanes on 64 banks (checked over every 8- and 16-row group)
    auto swz = [](int row) { return CPR == 4 ? ((row >> 1) ^ (row >> 3)) & 3 : (row ^ (row >> 3)) & 7; };

    int tm, tn;
    tile_of(blockIdx.x, gridDim.x, tiles_m, tm, tn);
    const int m0 = tm * TM, n0 = tn * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * (TM / 2), wn = (wave >> 1) * (TN / 2);

    // this thread's source chunk of each of its LPW DMA instructions (block w LPW + q of the
    // stage: A blocks first, then B), and the wave-uniform destination of each in a stage
    const ht* src[LPW];
    int dst[LPW];
#pragma unroll
    for (int q = 0; q < LPW; ++q) {
        const int blk = wave * LPW + q;
        const bool isA = blk < BA;
        const int b = isA ? blk : blk - BA;
        const int row = b * RPB + lane / CPR, kc = (lane % CPR) ^ swz(row);
        src[q] = (isA ? Ah + (size_t)(m0 + row) * kp : Bh + (size_t)(n0 + row) * kp) + 8 * kc;
        dst[q] = (isA ? 0 : TM * KT) + b * (RPB * KT);
    }
    auto issue = [&](auto s_, int k0) {
        ht* S = stage(s_);
#pragma unroll
        for (int q = 0; q < LPW; ++q)
            __builtin_amdgcn_global_load_lds((const void*)(src[q] + k0), (lds_void_t*)(S + dst[q]), 16, 0, 0);
    };
    auto barrier = [&]() {
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    };

    f16x acc[IM][JN];
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const int lr = lane & 31, lk = lane >> 5;
    auto frags = [&](const ht* S, int ks, v8* a_, v8* b_) {
        const int kc = 2 * ks + lk;
#pragma unroll
        for (int i = 0; i < IM; ++i) {
            const int row = wm + i * 32 + lr;
            a_[i] = *(const v8*)(S + row * KT + 8 * (kc ^ swz(row)));
        }
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const int row = wn + j * 32 + lr;
            b_[j] = *(const v8*)(S + TM * KT + row * KT + 8 * (kc ^ swz(row)));
        }
    };
    constexpr int KS = KT / 16;  // MFMA k steps a tile (even: the next tile's first step reuses set 0)
    static_assert(KS % 2 == 0, "gemm16d: an even number of k steps a tile");
    const int nk = kp / KT;
    v8 af[2][IM], bf[2][JN];
    // Tile t in stage s = t % NS, its first fragments already in set 0. Each k step reads the
    // next step's fragments while its MFMAs run; before the last step the wave waits for its
    // own loads of tile t + 1 and meets the others at the barrier, after which tile t + 1 is
    // in LDS for every wave and every wave is done reading stage (t - 1) % NS. The last step
    // then issues tile t + NS - 1 into that stage (past the last tile: the last tile again,
    // into a stage nobody reads, so every tile waits the same count) and reads tile t + 1's
    // first fragments.
    auto tile = [&](auto s_, int t) {
        constexpr int s = decltype(s_)::value;
        const ht* S = stage(s_);
        static_for<KS>([&](auto ks_) {
            constexpr int ks = decltype(ks_)::value;
            if constexpr (ks + 1 < KS) {
                frags(S, ks + 1, af[(ks + 1) & 1], bf[(ks + 1) & 1]);
            } else {
                __builtin_amdgcn_s_waitcnt(vmcnt_imm((NS - 3) * LPW));
                barrier();
                issue(std::integral_constant<int, (s + NS - 1) % NS>{}, min(t + NS - 1, nk - 1) * KT);
                frags(stage(std::integral_constant<int, (s + 1) % NS>{}), 0, af[0], bf[0]);
            }
#pragma unroll
            for (int i = 0; i < IM; ++i)
#pragma unroll
                for (int j = 0; j < JN; ++j) acc[i][j] = H::mma32(af[ks & 1][i], bf[ks & 1][j], acc[i][j]);
            mfma_interleave<IM * JN, IM + JN, 0, ks + 1 < KS ? 0 : LPW>();
        });
    };

    static_for<NS - 1>([&](auto p_) { issue(p_, min((int)decltype(p_)::value, nk - 1) * KT); });
    __builtin_amdgcn_s_waitcnt(vmcnt_imm((NS - 2) * LPW));  // tile 0
    barrier();
    frags(S0, 0, af[0], bf[0]);
    int t0 = 0;
    for (; t0 + NS <= nk; t0 += NS) static_for<NS>([&](auto s_) { tile(s_, t0 + decltype(s_)::value); });
    static_for<NS - 1>([&](auto s_) {  // the last nk % NS tiles
        if (t0 + decltype(s_)::value < nk) tile(s_, t0 + decltype(s_)::value);
    });

    // epilogue through LDS, 32 columns (one MFMA tile column) at a time in stage 0, which is
    // sized for it (every DMA has landed and every wave is past its last fragment read): the
    // chunk's accumulators as a column-major tile, then C updated down columns, consecutive
    // rows on consecutive lanes, the chunk's C values read before any result is stored
    // (unscaling in the same order as gemm16q)
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
    constexpr int CS = 256 / TM, NQ = CW / CS, QG = NQ < 16 ? NQ : 16;
    float* Ct = reinterpret_cast<float*>(S0);
    const int er = tid % TM, ec = tid / TM;
    const int row = m0 + er;
    const float ur = rs ? 1.0f / rs[min(row, m - 1)] : 1.0f;
#pragma unroll
    for (int c0 = 0; c0 < TN; c0 += CW) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            if (wn + j * 32 != c0) continue;  // wave-uniform
#pragma unroll
            for (int i = 0; i < IM; ++i)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    Ct[lr * LDC + wm + i * 32 + 8 * (r >> 2) + 4 * lk + (r & 3)] = acc[i][j][r];
        }
        __syncthreads();
#pragma unroll
        for (int q0 = 0; q0 < NQ; q0 += QG) {
            float cv[QG];
            if (beta != 0.0f) {
#pragma unroll
                for (int q = 0; q < QG; ++q)
                    cv[q] = C[min(row, m - 1) + (size_t)min(n0 + c0 + ec + CS * (q0 + q), n - 1) * ldc];
            }
#pragma unroll
            for (int q = 0; q < QG; ++q) {
                const int cl = ec + CS * (q0 + q), col = n0 + c0 + cl;
                if (row < m && col < n) {
                    const float uc = cs ? 1.0f / cs[col] : 1.0f;
                    const float v = alpha * (Ct[cl * LDC + er] * ur * uc);
                    C[row + (size_t)col * ldc] = beta == 0.0f ? v : v + beta * cv[q];
                }
            }
        }
    }
}

// Operand conversion of gemm16q_kernel, both operands in one launch (zero beyond m, n, k):
//   workgroups [0, nrb):  Ah[i * kp + k] = A(i, k) rs[i], one 64 x 64 tile each, transposed
//                         through LDS; a thread issues its 16 loads together and writes 16
//                         consecutive k of one row as two 16-byte chunks;
//   workgroups [nrb, ..): Bh[j * kp + k] = B(k, j) cs[j], 8 consecutive k of one column a
//                         thread, one 16-byte store (float4 loads where B is 16-byte aligned).
// kp is a multiple of 64 and (np kp / 8) < 2^32 (checked by the host).
template <bool BF16>
__global__ void __launch_bounds__(256) convert16_kernel(int m, int n, int k, const float* __restrict__ A, int lda,
                                                        const float* __restrict__ rs, const float* __restrict__ B,
                                                        int ldb, const float* __restrict__ cs,
                                                        typename Half<BF16>::t* __restrict__ Ah,
                                                        typename Half<BF16>::t* __restrict__ Bh, int kp, int np,
                                                        int nrb, int b_vec) {
    using H = Half<BF16>;
    using ht = typename H::t;
    using v8 = typename H::v8;
    __shared__ float tile[64][65];  // tile[k][i]
    const int t = threadIdx.x;
    if ((int)blockIdx.x < nrb) {
        const int kb = kp / 64;
        const int i0 = (blockIdx.x / kb) * 64, k0 = (blockIdx.x % kb) * 64;
        const int tx = t & 63, ty = t >> 6;
        const int i = i0 + tx;
        const float sc = rs ? rs[min(i, m - 1)] : 1.0f;
        const float* Ai = A + min(i, m - 1);
        float v[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) v[c] = Ai[(size_t)min(k0 + ty + 4 * c, k - 1) * lda];
#pragma unroll
        for (int c = 0; c < 16; ++c) tile[ty + 4 * c][tx] = (i < m && k0 + ty + 4 * c < k) ? v[c] * sc : 0.0f;
        __syncthreads();
        const int r = t >> 2, q = 16 * (t & 3);
        v8 o0, o1;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            o0[e] = (ht)tile[q + e][r];
            o1[e] = (ht)tile[q + 8 + e][r];
        }
        v8* dst = (v8*)(Ah + (size_t)(i0 + r) * kp + k0 + q);
        dst[0] = o0;
        dst[1] = o1;
        return;
    }
    const unsigned q = (unsigned)kp / 8u;
    const unsigned c = (blockIdx.x - (unsigned)nrb) * 256u + (unsigned)t;
    if (c >= q * (unsigned)np) return;
    const int j = (int)(c / q), k0 = 8 * (int)(c - (unsigned)j * q);
    const float sc = j < n ? (cs ? cs[j] : 1.0f) : 0.0f;
    const float* Bj = B + (size_t)min(j, n - 1) * ldb;
    float v[8];
    if (b_vec && j < n && k0 + 8 <= k) {
        const f4v x0 = *(const f4v*)(Bj + k0), x1 = *(const f4v*)(Bj + k0 + 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v[e] = x0[e] * sc;
            v[4 + e] = x1[e] * sc;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = Bj[min(k0 + e, k - 1)];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (j < n && k0 + e < k) ? v[e] * sc : 0.0f;
    }
    v8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = (ht)v[e];
    *(v8*)(Bh + (size_t)j * kp + k0) = o;
}

// ------------------------------------------------------------------ BF16 shadows
// Registered FP32 matrices with BF16 copies of their final values. A factor registration
// holds the part below the 128 x 128 diagonal blocks row-major (lo) and the part above them
// column-major (up): every 16-bit product of the factorisation and of its triangular solves
// takes L blocks as the A operand (k-contiguous rows) and U blocks or solved right-hand
// sides as the B operand (k-contiguous columns); only the U-side solves take U blocks as A
// (column-major). A dense registration holds a whole matrix, column-major or row-major.
namespace {
struct ShadowMap {
    const float* base;
    int ld, rows, cols;
    __bf16* lo;   // (i, j) at lo[i lds + j]
    __bf16* up;   // (i, j) at up[i + j lds]
    int lds;
    bool factor;
    int id;
    __bf16* upr;  // a factor's part above the diagonal blocks again, row-major (optional)
};
std::vector<ShadowMap> g_shadows;
int g_shadow_next = 1;
constexpr int SHB = 128;  // diagonal block of a registered factor

// the registration holding p (same leading dimension) and p's row and column in it
const ShadowMap* shadow_at(const float* p, int ld, int rows, int cols, int& r0, int& c0) {
    for (auto it = g_shadows.rbegin(); it != g_shadows.rend(); ++it) {
        const ShadowMap& m = *it;
        if (m.ld != ld || p < m.base) continue;
        const size_t off = (size_t)(p - m.base), c = off / (size_t)ld, r = off % (size_t)ld;
        if (c + (size_t)cols > (size_t)m.cols || r + (size_t)rows > (size_t)m.rows) continue;
        r0 = (int)r;
        c0 = (int)c;
        return &m;
    }
    return nullptr;
}

// an operand's BF16 copy: trans = stored as its transpose (a row-major copy)
struct Sh16 {
    const __bf16* p;
    int ld;
    bool trans;
};
// the copy of the rows x cols operand at (p, ld), if registered and aligned for 16-byte
// loads; a_operand: the product's A (k-contiguous rows preferred) or its B (k-contiguous
// columns preferred)
bool shadow_operand(const float* p, int ld, int rows, int cols, bool a_operand, Sh16& o) {
    if (g_shadows.empty() || rows <= 0 || cols <= 0) return false;
    int r0, c0;
    const ShadowMap* m = shadow_at(p, ld, rows, cols, r0, c0);
    if (!m) return false;
    const bool lower = r0 / SHB > (c0 + cols - 1) / SHB, upper = c0 / SHB > (r0 + rows - 1) / SHB;
    const Sh16 lo_rows = {m->lo ? m->lo + (size_t)r0 * m->lds + c0 : nullptr, m->lds, true};
    const Sh16 up_rows = {m->upr ? m->upr + (size_t)r0 * m->lds + c0 : nullptr, m->lds, true};
    const Sh16 up_cols = {m->up ? m->up + r0 + (size_t)c0 * m->lds : nullptr, m->lds, false};
    if (!m->factor)
        o = m->lo ? lo_rows : up_cols;
    else if (lower && m->lo)
        o = lo_rows;
    else if (upper && (m->up || m->upr))
        o = (a_operand && m->upr) || !m->up ? up_rows : up_cols;
    else
        return false;
    return (uintptr_t)o.p % 16 == 0 && o.ld % 8 == 0;
}
// where a kernel that finalises the rows x cols block at (p, ld) writes its copy:
// column-major (the up part of a factor, or a dense column-major registration) or row-major
// (the lo or upr part of a factor); null when the block has none
__bf16* shadow_out(const float* p, int ld, int rows, int cols, bool row_major, int* sld) {
    if (g_shadows.empty() || rows <= 0 || cols <= 0) return nullptr;
    int r0, c0;
    const ShadowMap* m = shadow_at(p, ld, rows, cols, r0, c0);
    if (!m) return nullptr;
    *sld = m->lds;
    const bool lower = r0 / SHB > (c0 + cols - 1) / SHB, upper = c0 / SHB > (r0 + rows - 1) / SHB;
    if (row_major) {
        if (!m->factor) return nullptr;
        if (lower) return m->lo ? m->lo + (size_t)r0 * m->lds + c0 : nullptr;
        return upper && m->upr ? m->upr + (size_t)r0 * m->lds + c0 : nullptr;
    }
    return m->up && (!m->factor || upper) ? m->up + r0 + (size_t)c0 * m->lds : nullptr;
}
}  // namespace

ShadowScope::ShadowScope(const float* base, int ld, int rows, int cols, __bf16* lo, __bf16* up, int lds, bool factor,
                         __bf16* upr)
    : id(0) {
    if (!base || (!lo && !up && !upr)) return;
    id = g_shadow_next++;
    g_shadows.push_back(ShadowMap{base, ld, rows, cols, lo, up, lds, factor, id, upr});
}
ShadowScope::~ShadowScope() {
    if (!id) return;
    for (auto it = g_shadows.begin(); it != g_shadows.end(); ++it)
        if (it->id == id) {
            g_shadows.erase(it);
            break;
        }
}

// ------------------------------------------------------------------ host GEMM wrappers
// ------------------------------------------------------------------ library FP64/FP32 GEMM
// Large plain FP64 and FP32 products (no operand overlapping C, no fused epilogue) go to
// rocBLAS, whose DGEMM runs at about 75 TFLOP/s on gfx950 against about 44 for the FP64 MFMA
// tile above. Like the 16-bit library GEMM it runs on the library stream only and never
// under graph capture; false sends the product to the hand-written kernels.
static int g_library_off = 0;  // LibraryGemmOff scopes alive
LibraryGemmOff::LibraryGemmOff() { ++g_library_off; }
LibraryGemmOff::~LibraryGemmOff() { --g_library_off; }
namespace rb {
static rocblas_handle g_handle = nullptr;
static bool g_broken = false;
static rocblas_handle handle() {
    if (!g_handle && !g_broken && rocblas_create_handle(&g_handle) != rocblas_status_success) {
        g_broken = true;
        g_handle = nullptr;
    }
    return g_handle;
}
}  // namespace rb
// whether the rows x cols blocks at X and Y (leading dimensions ldx, ldy) share an element:
// blocks of one matrix (same leading dimension) compare as rectangles (a block beside or
// above another in the same columns does not overlap it), others by address range
template <typename T>
static bool blocks_overlap(const T* X, int ldx, int mx, int nx, const T* Y, int ldy, int my, int ny) {
    if (mx <= 0 || nx <= 0 || my <= 0 || ny <= 0) return false;
    if (ldx == ldy && mx <= ldx && my <= ldy) {
        const long long d = (long long)(Y - X), ld = ldx;
        const long long dc = d >= 0 ? d / ld : -((-d + ld - 1) / ld), dr = d - dc * ld;  // 0 <= dr < ld
        if (dr + my <= ld) {
            const bool rows = dr < mx && 0 < dr + my, cols = dc < nx && 0 < dc + ny;
            return rows && cols;
        }
    }
    const char* x0 = (const char*)X;
    const char* x1 = (const char*)(X + (size_t)(nx - 1) * ldx + mx);
    const char* y0 = (const char*)Y;
    const char* y1 = (const char*)(Y + (size_t)(ny - 1) * ldy + my);
    return x0 < y1 && y0 < x1;
}

template <typename T>
static bool gemm_library(int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C,
                         int ldc) {
    if (ctx().redirect || capturing() || g_library_off) return false;
    rocblas_handle h = rb::handle();
    if (!h || rocblas_set_stream(h, ctx().s()) != rocblas_status_success) return false;
    rocblas_status st;
    if constexpr (sizeof(T) == 8)
        st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta, C,
                           ldc);
    else
        st = rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta, C,
                           ldc);
    SQ_REQUIRE(st == rocblas_status_success, SQBEM_ERR_HIP, "gemm: rocBLAS call failed");
    return true;
}

// sh_col / sh_row (optional): BF16 copies of the stored C, column-major at leading
// dimension sh_cld / row-major at sh_rld (beta = 0), addressed like C
template <typename T>
static void gemm_launch(int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta,
                        T* C, int ldc, int* unit_flag = nullptr, __bf16* sh_col = nullptr, int sh_cld = 0,
                        __bf16* sh_row = nullptr, int sh_rld = 0) {
    SQ_REQUIRE(!sh_row || beta == T(0), SQBEM_ERR_ARG, "gemm: a row-major copy needs beta = 0");
    if (m <= 0 || n <= 0) return;
    if (k <= 0 || alpha == T(0)) {
        // C = beta C
        if (beta == T(1)) return;
        scale_matrix(m, n, beta, C, ldc);
        return;
    }
    const int tiles_m = (int)ceil_div(m, GBM);
    ProfScope prof(sizeof(T) == 8 ? PROF_DGEMM : PROF_SGEMM, 2.0 * m * n * k,
                   (double)sizeof(T) * ((double)m * k + (double)k * n + (beta != T(0) ? 2.0 : 1.0) * m * n));
    // Smaller tiles when 128 x 128 ones would leave the CUs short of work. An in-place
    // product relies on one workgroup reading all of C's rows and columns that its own
    // results overwrite before it writes them: with A overlapping C (L21 = A21 U11^-1) a
    // workgroup must hold every column of its rows, with B overlapping C (X = inv B) every
    // row of its columns.
    const char* a0 = (const char*)A;
    const char* a1 = (const char*)(A + (size_t)(k - 1) * lda + m);
    const char* b0 = (const char*)B;
    const char* b1 = (const char*)(B + (size_t)(n - 1) * ldb + k);
    const char* c0 = (const char*)C;
    const char* c1 = (const char*)(C + (size_t)(n - 1) * ldc + m);
    const bool overlap = a0 < c1 && c0 < a1, overlap_b = b0 < c1 && c0 < b1;
    if (!unit_flag && !sh_col && !sh_row && m >= 256 && n >= 256 && (double)m * n * k >= 1.3e8 &&
        !blocks_overlap(A, lda, m, k, (const T*)C, ldc, m, n) && !blocks_overlap(B, ldb, k, n, (const T*)C, ldc, m, n) &&
        gemm_library<T>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc))
        return;
    const int cus = std::max(ctx().cu_count, 1);
    const int tiles128 = (int)(tiles_m * ceil_div(n, GBN));
    if (!overlap_b && n <= GBN && tiles128 < 2 * cus) {
        // 64 x 128 tiles: twice the workgroups, every one holding all columns of its rows
        const int tm64 = (int)ceil_div(m, 64);
        hipLaunchKernelGGL((gemm_nn_kernel<T, GBN, 64>), dim3(tm64), dim3(256), 0, ctx().s(), m, n, k, alpha, A, lda,
                           B, ldb, beta, C, ldc, tm64, unit_flag, sh_col, sh_cld, sh_row, sh_rld);
    } else if (!overlap && tiles128 < cus) {
        // 128 x 64 tiles (a one-row-tile product with many columns, e.g. the inverse-block
        // leaves of the triangular solves)
        const int tiles_n = (int)ceil_div(n, 64);
        hipLaunchKernelGGL((gemm_nn_kernel<T, 64>), dim3(tiles_m * tiles_n), dim3(256), 0, ctx().s(), m, n, k, alpha,
                           A, lda, B, ldb, beta, C, ldc, tiles_m, unit_flag, sh_col, sh_cld, sh_row, sh_rld);
    } else {
        const int tiles_n = (int)ceil_div(n, GBN);
        hipLaunchKernelGGL((gemm_nn_kernel<T, GBN>), dim3(tiles_m * tiles_n), dim3(256), 0, ctx().s(), m, n, k, alpha,
                           A, lda, B, ldb, beta, C, ldc, tiles_m, unit_flag, sh_col, sh_cld, sh_row, sh_rld);
    }
    check_launch("gemm");
}

void dgemm(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb, double beta,
           double* C, int ldc) {
    gemm_launch<double>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
void sgemm(int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb, float beta,
           float* C, int ldc) {
    gemm_launch<float>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

struct ScaleWork {
    DBuf<float> buf;
    float* get(size_t n) {
        if (buf.n < n) buf.alloc(std::max<size_t>(n, 1 << 16));
        return buf.p;
    }
};
// the 16-bit products' working buffers, one set per stream: work issued on the side stream
// runs beside the library stream's products
static ScaleWork g_scale_main, g_scale_side;
static DBuf<char> g_q16_main, g_q16_side;
static ScaleWork& scale_work() { return ctx().redirect ? g_scale_side : g_scale_main; }
static DBuf<char>& q16_work() { return ctx().redirect ? g_q16_side : g_q16_main; }

// ------------------------------------------------------------------ tile autotuner
// The 16-bit GEMM on converted operands comes in four workgroup tiles (128 or 256 rows by
// 128 or 256 columns). The fastest depends on the shape: how many workgroups the product
// gives the 256 CUs, how long its k loop is, how much of C is padding. The choice is
// learned per shape class (operand type and the binary exponents of m, n and k) from live
// calls: the first calls of a class each run another candidate between two events, and
// once every candidate's time is in, the class uses the fastest. Every candidate computes
// the same product, so exploring costs time, never correctness. sqbem_tune(0) restores
// the fixed rule.
namespace tune {
// 0: 128 x 128, 1: 128 x 256, 2: 256 x 128, 3: 256 x 256 (loads staged one k tile ahead);
// 4: 128 x 128, 5: 128 x 256, 6: 256 x 128 (two k tiles ahead); operand tiles DMA'd into
// LDS: 7: 256 x 256 and 8: 256 x 128 with four 32-deep stages, 9: 128 x 128 with four
// 64-deep stages; 10-13: hipBLASLt's BF16 GEMM on the same operands with the first four
// algorithms its heuristic offers for the shape (FP16 runs 3 in their place)
constexpr int NC = 14;
struct Cls {
    int tried = 0, best = -1;
    float ms[NC] = {};
    bool done[NC] = {};
    hipEvent_t e0[NC] = {}, e1[NC] = {};
};
static std::map<unsigned, Cls> g_table;
static int g_mode = 1;

static int lg(int v) {
    int l = 0;
    while ((1 << (l + 1)) <= v && l < 30) ++l;
    return l;
}

static void poll(Cls& c) {
    bool all = c.tried == NC;
    for (int i = 0; i < c.tried; ++i) {
        if (c.done[i]) continue;
        if (hipEventQuery(c.e1[i]) == hipSuccess) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, c.e0[i], c.e1[i]) == hipSuccess) c.ms[i] = ms;
            (void)hipEventDestroy(c.e0[i]);
            (void)hipEventDestroy(c.e1[i]);
            c.e0[i] = c.e1[i] = nullptr;
            c.done[i] = true;
        } else {
            all = false;
        }
    }
    if (all && c.best < 0) {
        int b = 0;
        for (int i = 1; i < NC; ++i)
            if (c.ms[i] < c.ms[b]) b = i;
        c.best = b;
    }
}

// the tile of a decided class, else the rule; looks up without adding a class or events
static int decided(int kind, int m, int n, int k, int rule) {
    if (!g_mode) return rule;
    auto it = g_table.find((unsigned)kind << 24 | (unsigned)lg(m) << 16 | (unsigned)lg(n) << 8 | (unsigned)lg(k));
    if (it == g_table.end() || it->second.best < 0) return rule;
    return it->second.best;
}

// Candidate for one call, with its timing events around the call when it explores.
struct Probe {
    int cand;
    Cls* cls = nullptr;
    int slot = -1;
    Probe(int kind, int m, int n, int k, int rule) : cand(rule) {
        if (!g_mode) return;
        if (capturing()) {  // recorded, not run: no timing events, and the class stays as it is
            cand = decided(kind, m, n, k, rule);
            return;
        }
        Cls& c = g_table[(unsigned)kind << 24 | (unsigned)lg(m) << 16 | (unsigned)lg(n) << 8 | (unsigned)lg(k)];
        poll(c);
        if (c.best >= 0) {
            cand = c.best;
        } else if (c.tried < NC) {
            slot = c.tried++;
            cand = slot;
            cls = &c;
            SQ_HIP(hipEventCreate(&c.e0[slot]));
            SQ_HIP(hipEventCreate(&c.e1[slot]));
            SQ_HIP(hipEventRecord(c.e0[slot], ctx().s()));
        }
    }
    ~Probe() {
        if (cls) (void)hipEventRecord(cls->e1[slot], ctx().s());
    }
};


static void reset(int mode) {
    SQ_HIP(hipStreamSynchronize(ctx().s()));
    for (auto& kv : g_table)
        for (int i = 0; i < NC; ++i) {
            if (kv.second.e0[i]) (void)hipEventDestroy(kv.second.e0[i]);
            if (kv.second.e1[i]) (void)hipEventDestroy(kv.second.e1[i]);
        }
    g_table.clear();
    g_mode = mode;
}

// decided classes as text, one per line: key, best tile, the four times (ms)
static void save(const char* path) {
    SQ_HIP(hipStreamSynchronize(ctx().s()));
    FILE* f = std::fopen(path, "w");
    SQ_REQUIRE(f, SQBEM_ERR_ARG, std::string("cannot open tune table for writing: ") + path);
    std::fprintf(f, "# sqbem gemm16q tile table: key(kind<<24|lg m<<16|lg n<<8|lg k) best ms[0..%d]\n", NC - 1);
    for (auto& kv : g_table) {
        poll(kv.second);
        if (kv.second.best < 0) continue;
        const Cls& c = kv.second;
        std::fprintf(f, "%u %d", kv.first, c.best);
        for (int i = 0; i < NC; ++i) std::fprintf(f, " %.6g", c.ms[i]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

// adds the classes of a saved table as decided (online learning continues for the others)
static int load(const char* path) {
    FILE* f = std::fopen(path, "r");
    SQ_REQUIRE(f, SQBEM_ERR_ARG, std::string("cannot open tune table: ") + path);
    char line[256];
    int n = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        unsigned key;
        int best, pos = 0, used = 0;
        float ms[NC] = {};
        if (std::sscanf(line, "%u %d%n", &key, &best, &pos) != 2 || best < 0 || best >= NC) continue;
        // older tables hold fewer candidates: the missing times stay 0 (unused, best is given)
        for (int i = 0; i < NC && std::sscanf(line + pos, " %g%n", &ms[i], &used) == 1; ++i) pos += used;
        Cls& c = g_table[key];
        for (int i = 0; i < NC; ++i) {
            if (c.e0[i]) (void)hipEventDestroy(c.e0[i]);
            if (c.e1[i]) (void)hipEventDestroy(c.e1[i]);
            c.e0[i] = c.e1[i] = nullptr;
            c.ms[i] = ms[i];
            c.done[i] = true;
        }
        c.tried = NC;
        c.best = best;
        ++n;
    }
    std::fclose(f);
    return n;
}

static void info(int* classes, int* decided, int* counts) {
    SQ_HIP(hipStreamSynchronize(ctx().s()));
    int d = 0;
    if (counts)
        for (int i = 0; i < NC; ++i) counts[i] = 0;
    for (auto& kv : g_table) {
        poll(kv.second);
        if (kv.second.best >= 0) {
            ++d;
            if (counts) ++counts[kv.second.best];
        }
    }
    if (classes) *classes = (int)g_table.size();
    if (decided) *decided = d;
}
}  // namespace tune

// Workspace of one 16-bit GEMM: the converted, zero-padded operands (large shapes) and,
// for FP16, the power-of-two row and column scales with the row absmax words. ws == null:
// the library's own growing buffers, the tile autotuner and kernel timing; otherwise the
// caller's region of ws_bytes (no allocation, no events: a decided tile or the fixed rule).
struct G16Work {
    unsigned char* ws = nullptr;
    size_t ws_bytes = 0;
};

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// bytes of a caller workspace for gemm16(kind, m, n, k): operands padded to 256 rows and
// columns (every tile candidate fits) and 64-deep k, then the FP16 scales
size_t gemm16_work_bytes(int kind, int m, int n, int k) {
    if (m <= 0 || n <= 0 || k <= 0) return 0;
    size_t b = 0;
    if (m >= 256 && n >= 256 && k >= 128) {
        const size_t mp = (size_t)(m + 255) / 256 * 256, np = (size_t)(n + 255) / 256 * 256,
                     kp = (size_t)(k + QBK - 1) / QBK * QBK;
        b = align256((mp + np) * kp * 2);
    }
    if (kind != SQBEM_BF16_IR) b += align256(sizeof(float) * ((size_t)2 * m + n));
    return b;
}

template <bool BF16, int TM, int TN, int PF>
static void gemm16q(int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb, float beta,
                    float* C, int ldc, float* rs, float* cs, const G16Work& w) {
    using ht = typename Half<BF16>::t;
    const int mp = (m + TM - 1) / TM * TM, np = (n + TN - 1) / TN * TN, kp = (k + QBK - 1) / QBK * QBK;
    const size_t need = ((size_t)mp + np) * kp * sizeof(ht);
    ht* Ah;
    if (w.ws) {
        Ah = (ht*)w.ws;  // gemm16_work_bytes reserved the 256-padded operands at its start
    } else {
        DBuf<char>& q16 = q16_work();
        if (q16.n < need) q16.alloc(need);
        Ah = (ht*)q16.p;
    }
    ht* Bh = Ah + (size_t)mp * kp;
    const size_t nrb = (size_t)(mp / 64) * (kp / 64), chunks = (size_t)np * kp / 8;
    if (chunks >= ((size_t)1 << 32) || nrb + ceil_div(chunks, 256) >= ((size_t)1 << 31))
        throw std::runtime_error("gemm16: operand too large for the 16-bit conversion");
    const int b_vec = ((uintptr_t)B % 16 == 0 && ldb % 4 == 0) ? 1 : 0;
    hipLaunchKernelGGL(convert16_kernel<BF16>, dim3((unsigned)(nrb + ceil_div(chunks, 256))), dim3(256), 0,
                       ctx().s(), m, n, k, A, lda, rs, B, ldb, cs, Ah, Bh, kp, np, (int)nrb, b_vec);
    check_launch("convert16");
    const int tiles_m = mp / TM, tiles_n = np / TN;
    const double fl = 2.0 * m * n * k, by = 2.0 * ((double)m * k + (double)k * n) + 4.0 * (beta != 0.0f ? 2.0 : 1.0) * m * n;
    constexpr int cand = PF == 2 ? (TM == 128 ? (TN == 128 ? 4 : 5) : 6)
                                 : (TM == 128 ? (TN == 128 ? 0 : 1) : (TN == 128 ? 2 : 3));
    ProfScope prof(w.ws ? -1 : PROF_GEMM16Q, fl, by);
    ProfScope prof_tile(w.ws ? -1 : PROF_G16Q_C0 + cand, fl, by);
    hipLaunchKernelGGL((gemm16q_kernel<BF16, TM, TN, PF>), dim3(tiles_m * tiles_n), dim3(256), 0, ctx().s(), m, n, kp,
                       alpha, Ah, Bh, beta, C, ldc, tiles_m, rs, cs);
    check_launch("gemm16q");
}

// the same product with gemm16d_kernel (operand tiles DMA'd into LDS, NS stages of KT)
template <bool BF16, int TM, int TN, int KT, int NS>
static void gemm16d(int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb, float beta,
                    float* C, int ldc, float* rs, float* cs, const G16Work& w, int cand) {
    using ht = typename Half<BF16>::t;
    const int mp = (m + TM - 1) / TM * TM, np = (n + TN - 1) / TN * TN, kp = (k + QBK - 1) / QBK * QBK;
    const size_t need = ((size_t)mp + np) * kp * sizeof(ht);
    ht* Ah;
    if (w.ws) {
        Ah = (ht*)w.ws;
    } else {
        DBuf<char>& q16 = q16_work();
        if (q16.n < need) q16.alloc(need);
        Ah = (ht*)q16.p;
    }
    ht* Bh = Ah + (size_t)mp * kp;
    const size_t nrb = (size_t)(mp / 64) * (kp / 64), chunks = (size_t)np * kp / 8;
    if (chunks >= ((size_t)1 << 32) || nrb + ceil_div(chunks, 256) >= ((size_t)1 << 31))
        throw std::runtime_error("gemm16: operand too large for the 16-bit conversion");
    const int b_vec = ((uintptr_t)B % 16 == 0 && ldb % 4 == 0) ? 1 : 0;
    hipLaunchKernelGGL(convert16_kernel<BF16>, dim3((unsigned)(nrb + ceil_div(chunks, 256))), dim3(256), 0,
                       ctx().s(), m, n, k, A, lda, rs, B, ldb, cs, Ah, Bh, kp, np, (int)nrb, b_vec);
    check_launch("convert16");
    const int tiles_m = mp / TM, tiles_n = np / TN;
    const double fl = 2.0 * m * n * k, by = 2.0 * ((double)m * k + (double)k * n) + 4.0 * (beta != 0.0f ? 2.0 : 1.0) * m * n;
    ProfScope prof(w.ws ? -1 : PROF_GEMM16Q, fl, by);
    ProfScope prof_tile(w.ws ? -1 : PROF_G16Q_C0 + cand, fl, by);
    hipLaunchKernelGGL((gemm16d_kernel<BF16, TM, TN, KT, NS>), dim3(tiles_m * tiles_n), dim3(256), 0, ctx().s(), m, n,
                       kp, alpha, Ah, Bh, beta, C, ldc, tiles_m, rs, cs);
    check_launch("gemm16d");
}

// ------------------------------------------------------------------ library candidate (BF16)
// hipBLASLt's BF16 GEMM with FP32 accumulation and FP32 C/D, on the same converted operands
// (a plain GEMM: Ah holds A^T as kp x mp column-major, Bh holds B as kp x np column-major,
// so D = alpha Ah^T Bh + beta C is its "TN" form). One plan (descriptors and the heuristic's
// first algorithm) per shape, kept for the process; the workspace is the library's own.
namespace lt {
constexpr int kAlgos = 4;  // the heuristic's first kAlgos algorithms are autotuner candidates
struct Plan {
    hipblasLtMatmulDesc_t desc = nullptr;
    hipblasLtMatrixLayout_t a = nullptr, b = nullptr, c = nullptr;
    hipblasLtMatmulAlgo_t algo[kAlgos]{};
    std::string kernel[kAlgos];  // the library's name for each algorithm's kernel (profiling)
    int nalgo = 0;
    bool ok = false;
};
static hipblasLtHandle_t g_handle = nullptr;
static bool g_broken = false;  // the library could not be set up: the candidate falls back
static std::map<std::array<long long, 9>, Plan> g_plans;
static DBuf<char> g_ws_main, g_ws_side;  // the library's workspace, one per stream
constexpr size_t kWorkspace = size_t(64) << 20;
static void* ws() {
    DBuf<char>& w = ctx().redirect ? g_ws_side : g_ws_main;
    if (!w.p) w.alloc(kWorkspace);
    return w.p;
}

static bool good(hipblasStatus_t st) { return st == HIPBLAS_STATUS_SUCCESS; }

// the plan of D(m x n, ldc) = alpha op(A) op(B) + beta C, BF16 A and B: ta = A stored as
// its transpose (k x m at lda), else m x k; tb = B stored as its transpose (n x k at ldb),
// else k x n
static const Plan* plan(int m, int n, int k, bool ta, int lda, bool tb, int ldb, int ldc, bool beta0) {
    if (g_broken) return nullptr;
    if (!g_handle) {
        if (!good(hipblasLtCreate(&g_handle))) {
            g_broken = true;
            return nullptr;
        }
    }
    const std::array<long long, 9> key{m, n, k, ta ? 1 : 0, lda, tb ? 1 : 0, ldb, ldc, beta0 ? 1 : 0};
    auto it = g_plans.find(key);
    if (it != g_plans.end()) return it->second.ok ? &it->second : nullptr;
    Plan& P = g_plans[key];
    const hipblasOperation_t tA = ta ? HIPBLAS_OP_T : HIPBLAS_OP_N, tB = tb ? HIPBLAS_OP_T : HIPBLAS_OP_N;
    bool ok = good(hipblasLtMatmulDescCreate(&P.desc, HIPBLAS_COMPUTE_32F, HIP_R_32F)) &&
              good(hipblasLtMatmulDescSetAttribute(P.desc, HIPBLASLT_MATMUL_DESC_TRANSA, &tA, sizeof tA)) &&
              good(hipblasLtMatmulDescSetAttribute(P.desc, HIPBLASLT_MATMUL_DESC_TRANSB, &tB, sizeof tB)) &&
              good(hipblasLtMatrixLayoutCreate(&P.a, HIP_R_16BF, ta ? k : m, ta ? m : k, lda)) &&
              good(hipblasLtMatrixLayoutCreate(&P.b, HIP_R_16BF, tb ? n : k, tb ? k : n, ldb)) &&
              good(hipblasLtMatrixLayoutCreate(&P.c, HIP_R_32F, m, n, ldc));
    if (ok) {
        hipblasLtMatmulPreference_t pref = nullptr;
        uint64_t ws = kWorkspace;
        hipblasLtMatmulHeuristicResult_t res[kAlgos];
        int found = 0;
        ok = good(hipblasLtMatmulPreferenceCreate(&pref)) &&
             good(hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &ws,
                                                        sizeof ws)) &&
             good(hipblasLtMatmulAlgoGetHeuristic(g_handle, P.desc, P.a, P.b, P.c, P.c, pref, kAlgos, res, &found));
        for (int i = 0; ok && i < found && i < kAlgos; ++i) {
            if (!good(res[i].state) || res[i].workspaceSize > kWorkspace) continue;
            P.algo[P.nalgo] = res[i].algo;
            try {
                P.kernel[P.nalgo] = hipblaslt_ext::getKernelNameFromAlgo(g_handle, res[i].algo);
            } catch (...) {
            }
            if (P.kernel[P.nalgo].empty()) P.kernel[P.nalgo] = "hipblaslt_gemm_" + std::to_string(i);
            ++P.nalgo;
        }
        ok = ok && P.nalgo > 0;
        if (pref) (void)hipblasLtMatmulPreferenceDestroy(pref);
    }
    P.ok = ok;
    return ok ? &P : nullptr;
}
}  // namespace lt

// candidate 10: the library GEMM on the converted operands (BF16 only; false when the
// library has no plan for the shape, and the caller then runs an own kernel)
static bool gemm16lt(int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb, float beta,
                     float* C, int ldc, int which) {
    const int mp = (m + 63) / 64 * 64, np = (n + 63) / 64 * 64, kp = (k + QBK - 1) / QBK * QBK;
    const lt::Plan* P = lt::plan(m, n, k, true, kp, false, kp, ldc, beta == 0.0f);
    if (!P) return false;
    const size_t need = ((size_t)mp + np) * kp * sizeof(__bf16);
    DBuf<char>& q16 = q16_work();
    if (q16.n < need) q16.alloc(need);
    __bf16* Ah = (__bf16*)q16.p;
    __bf16* Bh = Ah + (size_t)mp * kp;
    const size_t nrb = (size_t)(mp / 64) * (kp / 64), chunks = (size_t)np * kp / 8;
    if (chunks >= ((size_t)1 << 32) || nrb + ceil_div(chunks, 256) >= ((size_t)1 << 31))
        throw std::runtime_error("gemm16: operand too large for the 16-bit conversion");
    const int b_vec = ((uintptr_t)B % 16 == 0 && ldb % 4 == 0) ? 1 : 0;
    hipLaunchKernelGGL(convert16_kernel<true>, dim3((unsigned)(nrb + ceil_div(chunks, 256))), dim3(256), 0,
                       ctx().s(), m, n, k, A, lda, nullptr, B, ldb, nullptr, Ah, Bh, kp, np, (int)nrb, b_vec);
    check_launch("convert16");
    const double fl = 2.0 * m * n * k, by = 2.0 * ((double)m * k + (double)k * n) + 4.0 * (beta != 0.0f ? 2.0 : 1.0) * m * n;
    const int a = which < P->nalgo ? which : 0;
    ProfScope prof(PROF_GEMM16Q, fl, by);
    ProfScope prof_tile(PROF_G16Q_C0 + 10 + a, fl, by, P->kernel[a].c_str());
    const hipblasStatus_t st = hipblasLtMatmul(lt::g_handle, P->desc, &alpha, Ah, P->a, Bh, P->b, &beta, C, P->c, C,
                                               P->c, &P->algo[a], lt::ws(), lt::kWorkspace, ctx().s());
    SQ_REQUIRE(st == HIPBLAS_STATUS_SUCCESS, SQBEM_ERR_HIP, "gemm16: hipblasLtMatmul failed");
    return true;
}

// the library GEMM straight on registered BF16 copies of both operands (no conversion);
// false when the library has no plan for the shape
static bool gemm16lt_shadow(int m, int n, int k, float alpha, const Sh16& a, const Sh16& b, float beta, float* C,
                            int ldc, int which) {
    const lt::Plan* P = lt::plan(m, n, k, a.trans, a.ld, b.trans, b.ld, ldc, beta == 0.0f);
    if (!P) return false;
    const double fl = 2.0 * m * n * k, by = 2.0 * ((double)m * k + (double)k * n) + 4.0 * (beta != 0.0f ? 2.0 : 1.0) * m * n;
    const int al = which < P->nalgo ? which : 0;
    ProfScope prof(PROF_GEMM16Q, fl, by);
    ProfScope prof_tile(PROF_G16Q_C0 + 10 + al, fl, by, P->kernel[al].c_str());
    const hipblasStatus_t st = hipblasLtMatmul(lt::g_handle, P->desc, &alpha, a.p, P->a, b.p, P->b, &beta, C, P->c, C,
                                               P->c, &P->algo[al], lt::ws(), lt::kWorkspace, ctx().s());
    SQ_REQUIRE(st == HIPBLAS_STATUS_SUCCESS, SQBEM_ERR_HIP, "gemm16: hipblasLtMatmul failed");
    return true;
}

int shadow_rows_ld(int k) { return (k + 63) / 64 * 64; }
size_t shadow_rows_elems(int m, int k) { return (size_t)(m + 63) / 64 * 64 * (size_t)shadow_rows_ld(k); }
// the A operand conversion of the 16-bit products alone: rows of A, k-contiguous, BF16
void shadow_rows(int m, int k, const float* A, int lda, __bf16* out) {
    if (m <= 0 || k <= 0) return;
    const int kp = shadow_rows_ld(k), mp = (m + 63) / 64 * 64;
    const size_t nrb = (size_t)(mp / 64) * (kp / 64);
    SQ_REQUIRE(nrb < ((size_t)1 << 31), SQBEM_ERR_ARG, "shadow_rows: matrix too large");
    hipLaunchKernelGGL(convert16_kernel<true>, dim3((unsigned)nrb), dim3(256), 0, ctx().s(), m, 0, k, A, lda, nullptr,
                       nullptr, 0, nullptr, out, nullptr, kp, 0, (int)nrb, 0);
    check_launch("shadow_rows");
}

// FP16 operand scales (powers of two per row of A and per column of B) into sp[0, 2m + n)
static void fp16_scales(int m, int n, int k, const float* A, int lda, const float* B, int ldb, float* sp, float** rs,
                        float** cs) {
    *rs = sp;
    *cs = sp + m;
    unsigned* mx = (unsigned*)sp + m + n;
    fill_zero(mx, sizeof(unsigned) * m);
    hipLaunchKernelGGL(row_absmax_kernel, dim3(ceil_div(m, 64), ceil_div(k, 64)), dim3(256), 0, ctx().s(), m, k, A,
                       lda, mx);
    hipLaunchKernelGGL(bits_to_pow2_scale, dim3(ceil_div(m, 256)), dim3(256), 0, ctx().s(), m, mx, *rs);
    hipLaunchKernelGGL(pow2_scale_cols, dim3(ceil_div(n, 4)), dim3(256), 0, ctx().s(), k, n, B, ldb, *cs);
    check_launch("fp16 scales");
}

static void gemm16_impl(int kind, int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb,
                        float beta, float* C, int ldc, const G16Work& w) {
    if (m <= 0 || n <= 0) return;
    if (k <= 0 || alpha == 0.0f) {
        if (beta != 1.0f) scale_matrix(m, n, beta, C, ldc);
        return;
    }
    const bool bf = kind == SQBEM_BF16_IR;
    const bool big = m >= 256 && n >= 256 && k >= 128;
    // The library's GEMM kernels are stream-K kernels: one persistent workgroup per CU (or
    // two), some of which wait for others' partial tiles, so they assume the whole grid is
    // resident. Two of them running at once on different streams can each hold CUs that the
    // other's waiting workgroups need. Library kernels therefore run on the library stream
    // only; products issued on the side stream use this file's kernels, whose workgroups
    // never wait on one another (beside those, a stream-K kernel's missing workgroups start
    // as soon as CUs free up).
    const bool side = ctx().redirect != nullptr;
    if (bf && !w.ws && !side) {
        // both operands have BF16 copies: the library GEMM reads them as they are (the
        // algorithm the autotuner chose for the class when it chose the library)
        Sh16 sa, sb;
        if (shadow_operand(A, lda, m, k, true, sa) && shadow_operand(B, ldb, k, n, false, sb)) {
            const int cand = big ? tune::decided(1, m, n, k, -1) : -1;
            if (gemm16lt_shadow(m, n, k, alpha, sa, sb, beta, C, ldc, cand >= 10 ? cand - 10 : 0)) return;
        }
    }
    float *rs = nullptr, *cs = nullptr;
    if (!bf) {
        float* sp;
        if (w.ws) {
            const size_t mp = (size_t)(m + 255) / 256 * 256, np = (size_t)(n + 255) / 256 * 256,
                         kp = (size_t)(k + QBK - 1) / QBK * QBK;
            sp = (float*)(w.ws + (big ? align256((mp + np) * kp * 2) : 0));
        } else {
            sp = scale_work().get((size_t)2 * m + n);
        }
        fp16_scales(m, n, k, A, lda, B, ldb, sp, &rs, &cs);
    }
    if (big) {
        // fixed rule: 256 x 256 tiles once they fill the device, 128 x 128 below that;
        // the autotuner replaces it per shape class once it has timed every tile
        const int rule = ceil_div(m, 256) * ceil_div(n, 256) >= 192 ? 3 : 0;
        int cand = rule;
        std::unique_ptr<tune::Probe> probe;
        if (w.ws || side) {
            // no timing: the decided tile, and an own kernel where the library was chosen
            // (it may allocate; beside the library stream it must not run)
            cand = tune::decided(bf ? 1 : 0, m, n, k, rule);
            if (cand >= 10) cand = 3;
        } else {
            probe.reset(new tune::Probe(bf ? 1 : 0, m, n, k, rule));
            cand = probe->cand;
        }
        if (cand >= 10) {
            if (bf && !side && gemm16lt(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cand - 10)) return;
            cand = 3;  // FP16 (scaled operands), side stream, or no library plan: the 256 x 256 kernel
        }
        auto run = [&](auto tag_bf) {
            constexpr bool BFv = decltype(tag_bf)::value;
            switch (cand) {
                case 0: gemm16q<BFv, 128, 128, 1>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w); break;
                case 1: gemm16q<BFv, 128, 256, 1>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w); break;
                case 2: gemm16q<BFv, 256, 128, 1>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w); break;
                case 4: gemm16q<BFv, 128, 128, 2>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w); break;
                case 5: gemm16q<BFv, 128, 256, 2>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w); break;
                case 6: gemm16q<BFv, 256, 128, 2>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w); break;
                case 7: gemm16d<BFv, 256, 256, 32, 4>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w, 7); break;
                case 8: gemm16d<BFv, 256, 128, 32, 4>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w, 8); break;
                case 9: gemm16d<BFv, 128, 128, 64, 4>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w, 9); break;
                default: gemm16q<BFv, 256, 256, 1>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, rs, cs, w); break;
            }
        };
        if (bf) run(std::true_type{});
        else run(std::false_type{});
        return;
    }
    int tiles_m = (int)ceil_div(m, GBM), tiles_n = (int)ceil_div(n, GBN);
    ProfScope prof(w.ws ? -1 : PROF_GEMM16, 2.0 * m * n * k,
                   4.0 * ((double)m * k + (double)k * n + (beta != 0.0f ? 2.0 : 1.0) * m * n));
    if (bf)
        hipLaunchKernelGGL(gemm16_nn_kernel<true>, dim3(tiles_m * tiles_n), dim3(256), 0, ctx().s(), m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, rs, cs);
    else
        hipLaunchKernelGGL(gemm16_nn_kernel<false>, dim3(tiles_m * tiles_n), dim3(256), 0, ctx().s(), m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, rs, cs);
    check_launch("gemm16");
}

void gemm16(int kind, int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb,
            float beta, float* C, int ldc) {
    gemm16_impl(kind, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, G16Work{});
}

// ------------------------------------------------------------------ no-allocation 16-bit GEMM
// Scratch of sqbem_gemm16_noalloc: one fixed region per device, in the code object's
// device globals (each device holds its own copy until the process ends). Host code
// reaches it only through the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[SQBEM_GEMM16_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

void gemm16_scratch_setup(int device) {
    SQ_REQUIRE(device >= 0 && device < kMaxDevices, SQBEM_ERR_ARG, "gemm16 scratch: device id out of range");
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    SQ_REQUIRE(e == hipSuccess && p, SQBEM_ERR_HIP,
               std::string("gemm16 scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (unsigned char*)p;
}

void gemm16_noalloc(int kind, int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb,
                    float beta, float* C, int ldc) {
    const size_t need = gemm16_work_bytes(kind, m, n, k);
    SQ_REQUIRE(need <= (size_t)SQBEM_GEMM16_SCRATCH_BYTES, SQBEM_ERR_ARG,
               "gemm16_noalloc: the product needs " + std::to_string(need) + " bytes of scratch, more than the " +
                   std::to_string((size_t)SQBEM_GEMM16_SCRATCH_BYTES) + " the library holds");
    const int dev = ctx().device;
    SQ_REQUIRE(dev >= 0 && dev < kMaxDevices && g_scratch_ptr[dev], SQBEM_ERR_STATE,
               "gemm16_noalloc: the library is not initialised on this device (call sqbem_init first)");
    G16Work w;
    w.ws = g_scratch_ptr[dev];
    w.ws_bytes = SQBEM_GEMM16_SCRATCH_BYTES;
    gemm16_impl(kind, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, w);
}

// ------------------------------------------------------------------ elementwise helpers
template <typename T>
__global__ void scale_matrix_kernel(int m, int n, T beta, T* C, int ldc) {
    size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t total = (size_t)m * n;
    for (; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(idx % m), j = (int)(idx / m);
        T* c = C + i + (size_t)j * ldc;
        *c = beta == T(0) ? T(0) : beta * *c;
    }
}
template <typename T>
void scale_matrix(int m, int n, T beta, T* C, int ldc) {
    size_t total = (size_t)m * n;
    unsigned blocks = (unsigned)std::min<size_t>(ceil_div(total, 256), 65535);
    hipLaunchKernelGGL(scale_matrix_kernel<T>, dim3(blocks), dim3(256), 0, ctx().s(), m, n, beta, C, ldc);
    check_launch("scale_matrix");
}
template void scale_matrix<double>(int, int, double, double*, int);
template void scale_matrix<float>(int, int, float, float*, int);

template <typename Ti, typename To>
__global__ void __launch_bounds__(256) convert_kernel(int m, int n, const Ti* __restrict__ A, int lda,
                                                      To* __restrict__ B, int ldb) {
    // workgroup (x, y): rows 1024 x .. 1024 x + 1023 of columns y, y + gridDim.y, ...; a thread
    // moves rows r, r + 256, r + 512, r + 768 (consecutive rows on consecutive lanes, no
    // division per element)
    const int i0 = blockIdx.x * 1024 + threadIdx.x;
    for (int j = blockIdx.y; j < n; j += gridDim.y) {
        const Ti* a = A + (size_t)j * lda;
        To* b = B + (size_t)j * ldb;
        Ti v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = a[min(i0 + 256 * u, m - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + 256 * u < m) b[i0 + 256 * u] = (To)v[u];
    }
}
template <typename Ti, typename To>
void convert_matrix(int m, int n, const Ti* A, int lda, To* B, int ldb) {
    if (m <= 0 || n <= 0) return;
    const dim3 grid((unsigned)ceil_div(m, 1024), (unsigned)std::min(n, 65535));
    hipLaunchKernelGGL((convert_kernel<Ti, To>), grid, dim3(256), 0, ctx().s(), m, n, A, lda, B, ldb);
    check_launch("convert_matrix");
}
template void convert_matrix<double, float>(int, int, const double*, int, float*, int);
template void convert_matrix<float, double>(int, int, const float*, int, double*, int);
template void convert_matrix<double, double>(int, int, const double*, int, double*, int);
template void convert_matrix<float, float>(int, int, const float*, int, float*, int);

// ------------------------------------------------------------------ panel LU
constexpr int LEAF = 32;       // columns per panel kernel
constexpr int PANEL_WG = 256;  // threads per workgroup
constexpr int IB = LEAF;       // diagonal blocks inverted for the triangular solves
constexpr int IBP = IB + 1;

// cleared once the runtime refuses a cooperative launch (or by sqbem_coop(0)): the pivoted
// panels then take the barrier-free path
static bool g_coop_ok = true;

// Workspace of one factorisation: candidates of the cooperative pivot search, a saved
// copy of the current panel, one verification flag per panel, and the inverses of the
// IB x IB diagonal blocks of L (unit lower) used by the triangular solves inside getrf.
struct PanelWork {
    DBuf<double> cand;  // 2 x G x (LEAF + 2): (|pivot|, row, row values)
    DBuf<double> rowj;  // 2 x LEAF: the row being replaced
    DBuf<unsigned> bar; // 2 ping-pong barrier counters
    DBuf<int> status;   // [0] first zero pivot (1-based), [1] barrier timeout
    DBuf<int> flags;    // per panel: 1 when row interchanges were needed
    DBuf<char> backup;  // m x LEAF copy of the panel in its storage type
    DBuf<char> inv;     // Linv blocks then Uinv blocks, IB x IB each, storage type
    DBuf<double> serial; // m x LEAF working panel of the barrier-free fallback (FP64)
    DBuf<int> pivrow;   // the fallback's pivot row of the current column
    int gmax = 0;
    void ensure(int g, int n, size_t elem) {
        if (gmax < g || !cand.p) {
            gmax = std::max(g, 256);
            cand.alloc((size_t)2 * gmax * (LEAF + 2));
            rowj.alloc(2 * LEAF);
        }
        if (!bar.p) bar.alloc(2);
        if (!status.p) status.alloc(2);
        const size_t nb = (size_t)(n + IB - 1) / IB;
        if (flags.n < nb + 1) flags.alloc(nb + 1);
        if (backup.n < (size_t)n * LEAF * elem) backup.alloc((size_t)n * LEAF * elem);
        if (inv.n < 2 * nb * IB * IB * elem) inv.alloc(2 * nb * IB * IB * elem);
        if (!g_coop_ok && serial.n < (size_t)n * LEAF) serial.alloc((size_t)n * LEAF);
        if (!pivrow.p) pivrow.alloc(1);
    }
};
static PanelWork g_panel;
static int g_bar_parity = 0;

// Inverses of one IB x IB diagonal block B (LDS, row-major; strict lower part = L with a unit
// diagonal, upper part = U), padded to IB with the identity beyond w. Threads 0..IB-1 solve
// L x = e_j, threads IB..2IB-1 solve U x = e_j; columns go to Xl / Xu (LDS) and then to
// Linv / Uinv (global, column-major IB x IB).
template <typename T>
__device__ void invert_block(const double (*B)[IBP], int w, double (*Xl)[IBP], double (*Xu)[IBP], T* Linv, T* Uinv) {
    const int t = threadIdx.x;
    if (t < IB) {
        const int j = t;
        for (int i = 0; i < IB; ++i) {
            double x;
            if (i < j) x = 0.0;
            else if (i == j) x = 1.0;
            else {
                double s = 0.0;
                if (i < w)
                    for (int k = j; k < i; ++k) s += B[i][k] * Xl[k][j];
                x = -s;
            }
            Xl[i][j] = x;
        }
    } else if (t < 2 * IB) {
        const int j = t - IB;
        for (int i = IB - 1; i >= 0; --i) {
            double x = 0.0;
            if (i <= j) {
                double s = (i == j) ? 1.0 : 0.0;
                if (i < w)
                    for (int k = i + 1; k <= j && k < w; ++k) s -= B[i][k] * Xu[k][j];
                x = s / (i < w ? B[i][i] : 1.0);
            }
            Xu[i][j] = x;
        }
    }
    __syncthreads();
    for (int e = t; e < IB * IB; e += blockDim.x) {
        const int i = e % IB, j = e / IB;
        if (Linv) Linv[i + j * IB] = (T)Xl[i][j];
        if (Uinv) Uinv[i + j * IB] = (T)Xu[i][j];
    }
}

// Inverses of all diagonal blocks of an LU factor (one workgroup per block)
template <typename T>
__global__ void __launch_bounds__(64) invert_diag_kernel(int n, const T* __restrict__ LU, int lda, T* Linv, T* Uinv) {
    __shared__ double Bk[IB][IBP], Xl[IB][IBP], Xu[IB][IBP];
    const int b = blockIdx.x, r0 = b * IB, w = min(IB, n - r0);
    for (int e = threadIdx.x; e < IB * IB; e += blockDim.x) {
        const int i = e % IB, j = e / IB;
        Bk[i][j] = (i < w && j < w) ? (double)LU[(r0 + i) + (size_t)(r0 + j) * lda] : 0.0;
    }
    __syncthreads();
    invert_block<T>(Bk, w, Xl, Xu, Linv ? Linv + (size_t)b * IB * IB : nullptr,
                    Uinv ? Uinv + (size_t)b * IB * IB : nullptr);
}

// Speculative panel factorisation without row interchanges. Every workgroup factors the
// w x w top block redundantly (no pivoting), then its own rows below it as
// L_r = A_r U11^-1. Partial pivoting keeps the diagonal pivot exactly when every
// multiplier satisfies |l| <= 1 (LAPACK takes the first of equal magnitudes), which is
// checked here; otherwise *flag is raised and the cooperative pivoting kernel redoes the
// panel from the saved copy. Workgroup 0 also writes ipiv (identity) and L11^-1.
constexpr int SPEC_ROWS = 64;  // rows per workgroup of the speculative pass
template <typename T>
__global__ void __launch_bounds__(PANEL_WG) getf2_spec(T* A, int lda, int m, int w, int* ipiv, int base,
                                                       T* backup, int* flag, T* Linv) {
    __shared__ double TB[IB][IBP];
    __shared__ double Ui[IB][IBP];
    __shared__ double Xl[IB][IBP];
    __shared__ double Ch[IB][SPEC_ROWS];  // this workgroup's rows, column-major
    __shared__ int bad;
    const int g = blockIdx.x, tid = threadIdx.x;
    const int r0 = g * SPEC_ROWS, nr = max(0, min(m, r0 + SPEC_ROWS) - r0);
    if (tid == 0) bad = 0;
    for (int e = tid; e < IB * IB; e += PANEL_WG) {
        const int i = e % IB, j = e / IB;
        TB[i][j] = (i < w && j < w) ? (double)A[i + (size_t)j * lda] : (i == j ? 1.0 : 0.0);
    }
    for (int e = tid; e < nr * w; e += PANEL_WG) {
        const int lr = e % nr, c = e / nr;
        const T v = A[(r0 + lr) + (size_t)c * lda];
        backup[(r0 + lr) + (size_t)c * m] = v;
        Ch[c][lr] = (double)v;
    }
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        const double p = TB[j][j];
        if (tid == 0 && !(p != 0.0 && p == p)) bad = 1;
        const double rinv = 1.0 / p;
        for (int i = j + 1 + tid; i < w; i += PANEL_WG) {
            const double l = TB[i][j] * rinv;
            TB[i][j] = l;
            if (!(fabs(l) <= 1.0)) bad = 1;
        }
        __syncthreads();
        const int nn = w - j - 1;
        for (int e = tid; e < nn * nn; e += PANEL_WG) {
            const int i = j + 1 + e % nn, c = j + 1 + e / nn;
            TB[i][c] -= TB[i][j] * TB[j][c];
        }
        __syncthreads();
    }
    // U11^-1 (all workgroups) and L11^-1 (workgroup 0) by column solves
    if (tid < IB) {
        const int j = tid;
        for (int i = IB - 1; i >= 0; --i) {
            double x = 0.0;
            if (i <= j) {
                double s = (i == j) ? 1.0 : 0.0;
                for (int k = i + 1; k <= j; ++k) s -= TB[i][k] * Ui[k][j];
                x = s / TB[i][i];
            }
            Ui[i][j] = x;
        }
    } else if (g == 0 && tid >= 64 && tid < 64 + IB) {
        const int j = tid - 64;
        for (int i = 0; i < IB; ++i) {
            double x;
            if (i < j) x = 0.0;
            else if (i == j) x = 1.0;
            else {
                double s = 0.0;
                for (int k = j; k < i; ++k) s += TB[i][k] * Xl[k][j];
                x = -s;
            }
            Xl[i][j] = x;
        }
    }
    __syncthreads();
    // L(r, c) = sum_{k <= c} A(r, k) U11^-1(k, c) for the rows below the top block
    for (int e = tid; e < nr * w; e += PANEL_WG) {
        const int lr = e % nr, c = e / nr, r = r0 + lr;
        double v;
        if (r < w) {
            v = TB[r][c];
        } else {
            v = 0.0;
            for (int k = 0; k <= c; ++k) v += Ch[k][lr] * Ui[k][c];
            if (!(fabs(v) <= 1.0)) bad = 1;
        }
        A[r + (size_t)c * lda] = (T)v;
    }
    if (g == 0) {
        for (int j = tid; j < w; j += PANEL_WG) ipiv[j] = base + j;
        for (int e = tid; e < IB * IB; e += PANEL_WG) {
            const int i = e % IB, j = e / IB;
            Linv[i + j * IB] = (T)Xl[i][j];
        }
    }
    __syncthreads();
    if (tid == 0 && bad) atomicOr(flag, 1);
}

// Bounded grid barrier: counter reaches target when every workgroup has arrived. A
// sticky timeout flag ends all waits if a workgroup never arrives, so the kernel exits.
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned target, int* timeout) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (++spins > (1u << 21)) {
                __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
    }
    __syncthreads();
}

// Partial-pivoting panel factorisation (LAPACK dgetf2 order: pivot = first row of largest
// magnitude, swap, scale by the reciprocal pivot, rank-1 update), run only when the
// speculative pass flagged the panel: it restarts from the saved copy. The rows are
// spread over the workgroups and held in LDS; each column's pivot is chosen with one grid
// barrier (double-buffered candidate rows). ipiv[j] = base + pivot row (0-based).
// Workgroup 0 finishes with the inverse of the new L11.
template <typename T>
__global__ void __launch_bounds__(PANEL_WG) getf2_coop(T* A, int lda, int m, int w, int rpw, int* ipiv, int base,
                                                       double* cand, double* rowj, unsigned* bar, unsigned* bar_next,
                                                       int* status, const int* flag, const T* backup, T* Linv) {
    extern __shared__ double lds[];  // max(rpw x w, 2 IB IBP), then scratch
    __shared__ double TBs[IB][IBP];
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    if (g == 0 && tid == 0) *bar_next = 0u;  // reset the other counter for the next launch
    if (*flag == 0) return;                  // speculative factors verified: nothing to do
    const int r0 = g * rpw, nr = max(0, min(m, r0 + rpw) - r0);
    const int area = max(rpw * w, 2 * IB * IBP);
    double* prow = lds + area;              // pivot row values [LEAF]
    double* redv = prow + LEAF;             // reduction scratch [PANEL_WG]
    int* redi = (int*)(redv + PANEL_WG);    // [PANEL_WG]
    __shared__ int s_piv;
    __shared__ int s_best;
    __shared__ double s_pval;

    for (int e = tid; e < nr * w; e += PANEL_WG) {
        int lr = e % nr, c = e / nr;
        lds[c * rpw + lr] = (double)backup[(r0 + lr) + (size_t)c * m];
    }
    __syncthreads();

    const int csz = LEAF + 2;
    for (int j = 0; j < w; ++j) {
        double* cj = cand + (size_t)(j & 1) * G * csz;
        double* rj = rowj + (j & 1) * LEAF;
        // local candidate: largest |a| in column j among rows >= j (first on ties)
        double best = -1.0;
        int brow = 0x7fffffff;
        for (int lr = tid; lr < nr; lr += PANEL_WG) {
            int r = r0 + lr;
            if (r < j) continue;
            double v = fabs(lds[j * rpw + lr]);
            if (v > best || (v == best && r < brow)) { best = v; brow = r; }
        }
        redv[tid] = best;
        redi[tid] = brow;
        __syncthreads();
        for (int off = PANEL_WG / 2; off > 0; off >>= 1) {
            if (tid < off) {
                double v = redv[tid + off];
                int r = redi[tid + off];
                if (v > redv[tid] || (v == redv[tid] && r < redi[tid])) { redv[tid] = v; redi[tid] = r; }
            }
            __syncthreads();
        }
        const double lbest = redv[0];
        const int lrow = redi[0];
        __syncthreads();
        if (tid == 0) {
            cj[(size_t)g * csz] = lbest;
            cj[(size_t)g * csz + 1] = (double)lrow;
        }
        if (lbest >= 0.0)
            for (int c = tid; c < w; c += PANEL_WG) cj[(size_t)g * csz + 2 + c] = lds[c * rpw + (lrow - r0)];
        if (j >= r0 && j < r0 + nr)
            for (int c = tid; c < w; c += PANEL_WG) rj[c] = lds[c * rpw + (j - r0)];

        grid_barrier(bar, (unsigned)(j + 1) * G, status + 1);

        // global pivot: every workgroup reduces the G candidates identically
        best = -1.0;
        brow = 0x7fffffff;
        int bg = -1;
        for (int q = tid; q < G; q += PANEL_WG) {
            double v = cj[(size_t)q * csz];
            int r = (int)cj[(size_t)q * csz + 1];
            if (v > best || (v == best && v >= 0.0 && r < brow)) { best = v; brow = r; bg = q; }
        }
        redv[tid] = best;
        redi[tid] = bg;
        __syncthreads();
        for (int off = PANEL_WG / 2; off > 0; off >>= 1) {
            if (tid < off) {
                double v = redv[tid + off];
                int q2 = redi[tid + off], q1 = redi[tid];
                int ra = q1 >= 0 ? (int)cj[(size_t)q1 * csz + 1] : 0x7fffffff;
                int rb = q2 >= 0 ? (int)cj[(size_t)q2 * csz + 1] : 0x7fffffff;
                if (v > redv[tid] || (v == redv[tid] && rb < ra)) { redv[tid] = v; redi[tid] = q2; }
            }
            __syncthreads();
        }
        if (tid == 0) {
            // no finite candidate anywhere (column all NaN): keep row j, report a zero pivot
            s_best = redi[0];
            s_piv = redi[0] >= 0 ? (int)cj[(size_t)redi[0] * csz + 1] : j;
        }
        __syncthreads();
        const int pr = s_piv;
        const double* pvals = s_best >= 0 ? cj + (size_t)s_best * csz + 2 : rj;
        for (int c = tid; c < w; c += PANEL_WG) prow[c] = pvals[c];
        __syncthreads();
        if (tid == 0) {
            s_pval = prow[j];
            if (g == 0) {
                ipiv[j] = base + pr;
                if (!(prow[j] != 0.0 && prow[j] == prow[j]) && status[0] == 0) status[0] = base + j + 1;
            }
        }
        // swap rows j and pr inside the owning workgroups
        if (pr != j) {
            if (j >= r0 && j < r0 + nr)
                for (int c = tid; c < w; c += PANEL_WG) lds[c * rpw + (j - r0)] = prow[c];
            if (pr >= r0 && pr < r0 + nr)
                for (int c = tid; c < w; c += PANEL_WG) lds[c * rpw + (pr - r0)] = rj[c];
        }
        __syncthreads();
        const double pv = s_pval;
        const double rinv = pv != 0.0 ? 1.0 / pv : 0.0;
        for (int lr = tid; lr < nr; lr += PANEL_WG) {
            if (r0 + lr <= j) continue;
            double l = lds[j * rpw + lr] * rinv;
            lds[j * rpw + lr] = l;
            for (int c = j + 1; c < w; ++c) lds[c * rpw + lr] -= l * prow[c];
        }
        __syncthreads();
    }
    if (g == 0)
        for (int e = tid; e < IB * IB; e += PANEL_WG) {
            const int i = e % IB, j = e / IB;
            TBs[i][j] = (i < w && j < w) ? lds[j * rpw + i] : 0.0;
        }
    for (int e = tid; e < nr * w; e += PANEL_WG) {
        int lr = e % nr, c = e / nr;
        A[(r0 + lr) + (size_t)c * lda] = (T)lds[c * rpw + lr];
    }
    if (g == 0) {
        __syncthreads();
        double(*Xl)[IBP] = (double(*)[IBP])lds;
        double(*Xu)[IBP] = (double(*)[IBP])(lds + IB * IBP);
        invert_block<T>(TBs, w, Xl, Xu, Linv, nullptr);
    }
}

// Barrier-free form of getf2_coop, for when the runtime refuses a cooperative launch (whose
// residency guarantee the grid barrier needs): the same arithmetic (the panel in FP64, the
// first largest |a| of each column as pivot, rows swapped, multipliers l = a * (1 / pivot),
// rank-one updates) on a working copy in global memory, one column at a time in three
// launches (pivot search, swap, update), then the factors stored and the new L11 inverted.
// Every kernel returns at once when the speculative pass verified its panel (flag 0).
template <typename T>
__global__ void pser_load(const int* flag, const T* __restrict__ backup, size_t mw, double* P) {
    if (*flag == 0) return;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < mw; e += (size_t)gridDim.x * blockDim.x)
        P[e] = (double)backup[e];
}
__global__ void __launch_bounds__(1024) pser_pivot(const int* flag, const double* __restrict__ P, int m, int j,
                                                   int* pivrow) {
    if (*flag == 0) return;
    __shared__ double rv[1024];
    __shared__ int ri[1024];
    const int tid = threadIdx.x;
    double best = -1.0;
    int brow = 0x7fffffff;
    for (int r = j + tid; r < m; r += 1024) {
        const double v = fabs(P[r + (size_t)j * m]);
        if (v > best || (v == best && r < brow)) { best = v; brow = r; }
    }
    rv[tid] = best;
    ri[tid] = brow;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (tid < off) {
            const double v = rv[tid + off];
            const int r = ri[tid + off];
            if (v > rv[tid] || (v == rv[tid] && r < ri[tid])) { rv[tid] = v; ri[tid] = r; }
        }
        __syncthreads();
    }
    // no finite candidate (column all NaN): keep row j, a zero pivot is reported by the swap
    if (tid == 0) *pivrow = rv[0] >= 0.0 ? ri[0] : j;
}
__global__ void pser_swap(const int* flag, double* P, int m, int w, int j, const int* pivrow, int* ipiv, int base,
                          int* status) {
    if (*flag == 0) return;
    const int pr = *pivrow, c = threadIdx.x;
    if (c < w && pr != j) {
        const double t = P[j + (size_t)c * m];
        P[j + (size_t)c * m] = P[pr + (size_t)c * m];
        P[pr + (size_t)c * m] = t;
    }
    if (c == 0) {
        ipiv[j] = base + pr;
        const double pv = P[pr + (size_t)j * m];  // the value moved to row j
        if (!(pv != 0.0 && pv == pv) && status[0] == 0) status[0] = base + j + 1;
    }
}
__global__ void pser_update(const int* flag, double* P, int m, int w, int j) {
    if (*flag == 0) return;
    const int r = j + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const double pv = P[j + (size_t)j * m];
    const double rinv = pv != 0.0 ? 1.0 / pv : 0.0;
    const double l = P[r + (size_t)j * m] * rinv;
    P[r + (size_t)j * m] = l;
    for (int c = j + 1; c < w; ++c) P[r + (size_t)c * m] -= l * P[j + (size_t)c * m];
}
template <typename T>
__global__ void pser_store(const int* flag, const double* __restrict__ P, int m, int w, T* A, int lda) {
    if (*flag == 0) return;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)m * w;
         e += (size_t)gridDim.x * blockDim.x)
        A[(e % m) + (e / m) * (size_t)lda] = (T)P[e];
}
template <typename T>
__global__ void __launch_bounds__(PANEL_WG) pser_invert(const int* flag, const double* __restrict__ P, int m, int w,
                                                        T* Linv) {
    if (*flag == 0) return;
    __shared__ double TBs[IB][IBP];
    __shared__ double X[2 * IB * IBP];
    for (int e = threadIdx.x; e < IB * IB; e += PANEL_WG) {
        const int i = e % IB, j = e / IB;
        TBs[i][j] = (i < w && j < w) ? P[i + (size_t)j * m] : 0.0;
    }
    __syncthreads();
    invert_block<T>(TBs, w, (double(*)[IBP])X, (double(*)[IBP])(X + IB * IBP), Linv, nullptr);
}

template <typename T>
static void panel_leaf_serial(T* A, int lda, int m, int w, int* ipiv, int base, const T* backup, const int* flag,
                              T* Linv) {
    double* P = g_panel.serial.p;
    int* pr = g_panel.pivrow.p;
    int* status = g_panel.status.p;
    const size_t mw = (size_t)m * w;
    const unsigned gl = (unsigned)std::min<size_t>(std::max<size_t>(ceil_div(mw, 256), 1), 4096);
    hipLaunchKernelGGL(pser_load<T>, dim3(gl), dim3(256), 0, ctx().s(), flag, backup, mw, P);
    for (int j = 0; j < w; ++j) {
        hipLaunchKernelGGL(pser_pivot, dim3(1), dim3(1024), 0, ctx().s(), flag, P, m, j, pr);
        hipLaunchKernelGGL(pser_swap, dim3(1), dim3(64), 0, ctx().s(), flag, P, m, w, j, pr, ipiv, base, status);
        if (m - j - 1 > 0)
            hipLaunchKernelGGL(pser_update, dim3((unsigned)ceil_div(m - j - 1, 256)), dim3(256), 0, ctx().s(), flag, P,
                               m, w, j);
    }
    hipLaunchKernelGGL(pser_store<T>, dim3(gl), dim3(256), 0, ctx().s(), flag, P, m, w, A, lda);
    hipLaunchKernelGGL(pser_invert<T>, dim3(1), dim3(PANEL_WG), 0, ctx().s(), flag, P, m, w, Linv);
    check_launch("panel (barrier-free)");
}

template <typename T>
static void panel_leaf(T* A, int lda, int m, int w, int* ipiv, int base) {
    // one workgroup per >= 128 rows, at most one per CU (the fallback is cooperative)
    int G = std::max(1, std::min((int)ceil_div(m, 128), std::max(ctx().cu_count, 1)));
    int rpw = (int)ceil_div(m, G);
    G = (int)ceil_div(m, rpw);
    const size_t area = std::max<size_t>((size_t)rpw * w, 2 * IB * IBP);
    size_t shmem = (area + LEAF + PANEL_WG) * sizeof(double) + PANEL_WG * sizeof(int);
    // default dynamic LDS limit per workgroup (64 KiB) minus the static part: ~60k rows
    SQ_REQUIRE(shmem + IB * IBP * sizeof(double) <= 64 * 1024, SQBEM_ERR_ARG,
               "panel too tall for LDS (more than ~60k rows)");
    const int blk = base / IB;
    T* backup = (T*)g_panel.backup.p;
    int* flag = g_panel.flags.p + blk;
    T* Linv = (T*)g_panel.inv.p + (size_t)blk * IB * IB;
    {
        ProfScope prof(PROF_PANEL, (double)m * w * w, 3.0 * sizeof(T) * m * w);
        hipLaunchKernelGGL(getf2_spec<T>, dim3(ceil_div(m, SPEC_ROWS)), dim3(PANEL_WG), 0, ctx().s(), A, lda, m, w,
                           ipiv, base, backup, flag, Linv);
        check_launch("getf2_spec");
    }
    unsigned* bar = g_panel.bar.p + g_bar_parity;
    unsigned* bar_next = g_panel.bar.p + (g_bar_parity ^ 1);
    g_bar_parity ^= 1;
    double* cand = g_panel.cand.p;
    double* rowj = g_panel.rowj.p;
    int* status = g_panel.status.p;
    const int* cflag = flag;
    const T* cbackup = backup;
    void* args[] = {&A, &lda, &m, &w, &rpw, &ipiv, &base, &cand, &rowj, &bar, &bar_next, &status, &cflag, &cbackup,
                    &Linv};
    ProfScope prof(PROF_PANEL, 0.0, 0.0);
    // The cooperative launch guarantees that all G <= CU-count workgroups are resident, as
    // the grid barrier needs. The runtime refuses it when another process on the device holds
    // the cooperative (global wave sync) queue; then, and from then on, the panel takes the
    // barrier-free path, which needs no residency.
    if (g_coop_ok) {
        const hipError_t e = hipLaunchCooperativeKernel((const void*)getf2_coop<T>, dim3(G), dim3(PANEL_WG), args,
                                                        (unsigned)shmem, ctx().s());
        if (e == hipSuccess) return;
        (void)hipGetLastError();  // launch refused (nothing was enqueued): clear it
        g_coop_ok = false;
        if (g_panel.serial.n < (size_t)m * LEAF) g_panel.serial.alloc((size_t)m * LEAF);
    }
    panel_leaf_serial<T>(A, lda, m, w, ipiv, base, cbackup, cflag, Linv);
}

// ------------------------------------------------------------------ row interchanges
// For every column: rows j and p = ipiv[ibase+j] - vbase swapped, j = j0..j1-1 in order
// (dir > 0) or in reverse order (dir < 0, the inverse permutation).
template <typename T>
__global__ void laswp_kernel(T* A, int lda, int ncols, const int* __restrict__ ipiv, int ibase, int vbase, int j0,
                             int j1, int dir) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    T* col = A + (size_t)c * lda;
    if (dir > 0) {
        for (int j = j0; j < j1; ++j) {
            int p = ipiv[ibase + j] - vbase;
            if (p != j) { T t = col[j]; col[j] = col[p]; col[p] = t; }
        }
    } else {
        for (int j = j1 - 1; j >= j0; --j) {
            int p = ipiv[ibase + j] - vbase;
            if (p != j) { T t = col[j]; col[j] = col[p]; col[p] = t; }
        }
    }
}

// Long pivot ranges: the interchanges that move a row are first compacted into an ordered
// list, then the swap kernel walks only that list, so identity pivots (the no-pivot
// factorisation) cost one short pass over ipiv. Workgroup c takes pivots 1024 c .. 1024 c +
// 1023 of the range (coalesced), places its moves in order in its own slot region (ballot
// and per-wave counts) and records their number; the last workgroup to finish (a device-scope
// counter, which it resets for the next launch) lays the regions out one after another.
__global__ void __launch_bounds__(1024) piv_compact_kernel(const int* __restrict__ ipiv, int ibase, int vbase, int j0,
                                                           int j1, int* list, int* count, int* ccount, int* slots,
                                                           unsigned* done) {
    __shared__ int wsum[16];
    __shared__ int s_last;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int c = blockIdx.x, nch = gridDim.x;
    const int j = j0 + c * 1024 + tid;
    const int p = j < j1 ? ipiv[ibase + j] - vbase : j;
    const bool mv = j < j1 && p != j;
    const unsigned long long m = __ballot(mv);
    if (lane == 0) wsum[wv] = __popcll(m);
    __syncthreads();
    if (mv) {
        int k = __popcll(m & ((1ull << lane) - 1ull));
        for (int q = 0; q < wv; ++q) k += wsum[q];
        slots[c * 2048 + 2 * k] = j;
        slots[c * 2048 + 2 * k + 1] = p;
    }
    if (tid == 0) {
        int tot = 0;
        for (int q = 0; q < 16; ++q) tot += wsum[q];
        ccount[c] = tot;
    }
    __threadfence();  // this chunk's slots and count, device-wide, before it is counted
    __syncthreads();
    if (tid == 0) s_last = atomicAdd(done, 1u) == (unsigned)(nch - 1);
    __syncthreads();
    if (!s_last) return;
    __threadfence();  // every chunk's writes, seen by the last workgroup
    int base = 0;
    for (int cc = 0; cc < nch; ++cc) {
        const int nc = __hip_atomic_load(&ccount[cc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int e = tid; e < 2 * nc; e += 1024) list[2 * base + e] = slots[cc * 2048 + e];
        base += nc;
    }
    if (tid == 0) {
        *count = base;
        *done = 0u;
    }
}

template <typename T>
__global__ void laswp_list_kernel(T* A, int lda, int ncols, const int* __restrict__ list, const int* __restrict__ count,
                                  int dir) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int nsw = *count;
    if (c >= ncols || nsw == 0) return;
    T* col = A + (size_t)c * lda;
    for (int q = 0; q < nsw; ++q) {
        const int qq = dir > 0 ? q : nsw - 1 - q;
        const int j = list[2 * qq], p = list[2 * qq + 1];
        const T t = col[j];
        col[j] = col[p];
        col[p] = t;
    }
}

// [0] count, [1] the compaction's completion counter (zero between launches), then the
// (row, pivot row) pairs, the per-chunk counts and the per-chunk slot regions
static DBuf<int> g_swaps;

template <typename T>
static void laswp(T* A, int lda, int ncols, const int* ipiv, int ibase, int vbase, int j0, int j1, int dir = 1) {
    if (ncols <= 0 || j1 <= j0) return;
    ProfScope prof(PROF_LASWP, 0.0, 0.0);
    if (j1 - j0 > 256) {
        const size_t len = (size_t)(j1 - j0), nch = (len + 1023) / 1024;
        const size_t need = 2 + 2 * len + nch + nch * 2048;
        if (g_swaps.n < need) {
            g_swaps.alloc(need);
            g_swaps.zero();
        }
        int* list = g_swaps.p + 2;
        int* ccount = list + 2 * len;
        hipLaunchKernelGGL(piv_compact_kernel, dim3((unsigned)nch), dim3(1024), 0, ctx().s(), ipiv, ibase, vbase, j0,
                           j1, list, g_swaps.p, ccount, ccount + nch, (unsigned*)(g_swaps.p + 1));
        hipLaunchKernelGGL(laswp_list_kernel<T>, dim3(ceil_div(ncols, 256)), dim3(256), 0, ctx().s(), A, lda, ncols,
                           list, g_swaps.p, dir);
        check_launch("laswp");
        return;
    }
    hipLaunchKernelGGL(laswp_kernel<T>, dim3(ceil_div(ncols, 256)), dim3(256), 0, ctx().s(), A, lda, ncols, ipiv,
                       ibase, vbase, j0, j1, dir);
    check_launch("laswp");
}

// ------------------------------------------------------------------ triangular leaves
// B (w x nrhs, w <= 64) <- T^-1 B for T the w x w triangle at L. One wave per 64
// right-hand sides: the triangle and the 64 columns are staged in LDS (coalesced), each
// lane solves one column. The triangle is padded to 64 with an identity.
constexpr int TL = 64;
template <typename T, bool LOWER, bool UNIT, bool TRANS>
__global__ void __launch_bounds__(64) trsm_leaf(int w, int nrhs, const T* __restrict__ L, int ldl, T* B, int ldb) {
    __shared__ T Ls[TL][TL + 1];
    __shared__ T Xs[TL][64];
    const int tid = threadIdx.x, c0 = blockIdx.x * 64;
    const int nc = min(64, nrhs - c0);
    for (int e = tid; e < TL * TL; e += 64) {
        int i = e % TL, j = e / TL;
        T v = (i == j) ? T(1) : T(0);
        if (i < w && j < w) v = TRANS ? L[j + (size_t)i * ldl] : L[i + (size_t)j * ldl];
        Ls[i][j] = v;
    }
    for (int e = tid; e < TL * 64; e += 64) {
        int i = e % TL, c = e / TL;
        Xs[i][c] = (i < w && c < nc) ? B[i + (size_t)(c0 + c) * ldb] : T(0);
    }
    __syncthreads();
    if (LOWER) {
        for (int i = 0; i < TL; ++i) {
            T s = Xs[i][tid];
            for (int j = 0; j < i; ++j) s -= Ls[i][j] * Xs[j][tid];
            Xs[i][tid] = UNIT ? s : s / Ls[i][i];
        }
    } else {
        for (int i = TL - 1; i >= 0; --i) {
            T s = Xs[i][tid];
            for (int j = i + 1; j < TL; ++j) s -= Ls[i][j] * Xs[j][tid];
            Xs[i][tid] = UNIT ? s : s / Ls[i][i];
        }
    }
    __syncthreads();
    for (int e = tid; e < TL * 64; e += 64) {
        int i = e % TL, c = e / TL;
        if (i < w && c < nc) B[i + (size_t)(c0 + c) * ldb] = Xs[i][c];
    }
}

template <typename T>
static void gemm_kind(int kind, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C,
                      int ldc);
template <>
void gemm_kind<double>(int, int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
                       double beta, double* C, int ldc) {
    dgemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
template <>
void gemm_kind<float>(int kind, int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb,
                      float beta, float* C, int ldc) {
    if (kind == SQBEM_FP16_IR || kind == SQBEM_BF16_IR)
        gemm16(kind, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else
        sgemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

static int split_point(int w, int unit) {
    int h = (w / 2 + unit - 1) / unit * unit;
    return std::min(std::max(h, unit), w - 1);
}

// Recursive triangular solve with the n x n triangle at L (leading dim ldl):
//   LOWER && UNIT : B <- L^-1 B (unit lower, the L of an LU)
//   !LOWER        : B <- U^-1 B (upper, non-unit)
// TRANS uses the transposed triangle (for the A^T solves of the adjoint).
template <typename T, bool LOWER, bool UNIT, bool TRANS>
static void trsm_rec(int kind, int n, int nrhs, const T* L, int ldl, T* B, int ldb) {
    if (n <= 0 || nrhs <= 0) return;
    if (n <= TL) {
        ProfScope prof(PROF_TRSM, (double)n * n * nrhs, sizeof(T) * (2.0 * n * nrhs + (double)n * n));
        hipLaunchKernelGGL((trsm_leaf<T, LOWER, UNIT, TRANS>), dim3(ceil_div(nrhs, 64)), dim3(64), 0, ctx().s(), n,
                           nrhs, L, ldl, B, ldb);
        check_launch("trsm_leaf");
        return;
    }
    const int n1 = split_point(n, TL), n2 = n - n1;
    // effective triangle blocks: L11 = L[0:n1,0:n1], L22 = L[n1:,n1:], off-diagonal X
    const T* L11 = L;
    const T* L22 = L + n1 + (size_t)n1 * ldl;
    T* B1 = B;
    T* B2 = B + n1;
    if (LOWER) {
        // (effective lower) X1 = L11^-1 B1; B2 -= L21 X1; X2 = L22^-1 B2
        trsm_rec<T, LOWER, UNIT, TRANS>(kind, n1, nrhs, L11, ldl, B1, ldb);
        if (!TRANS)
            gemm_kind<T>(kind, n2, nrhs, n1, T(-1), L + n1, ldl, B1, ldb, T(1), B2, ldb);
        else
            gemm_tn<T>(kind, n2, nrhs, n1, T(-1), L + (size_t)n1 * ldl, ldl, B1, ldb, T(1), B2, ldb);
        trsm_rec<T, LOWER, UNIT, TRANS>(kind, n2, nrhs, L22, ldl, B2, ldb);
    } else {
        // (effective upper) X2 = U22^-1 B2; B1 -= U12 X2; X1 = U11^-1 B1
        trsm_rec<T, LOWER, UNIT, TRANS>(kind, n2, nrhs, L22, ldl, B2, ldb);
        if (!TRANS)
            gemm_kind<T>(kind, n1, nrhs, n2, T(-1), L + (size_t)n1 * ldl, ldl, B2, ldb, T(1), B1, ldb);
        else
            gemm_tn<T>(kind, n1, nrhs, n2, T(-1), L + n1, ldl, B2, ldb, T(1), B1, ldb);
        trsm_rec<T, LOWER, UNIT, TRANS>(kind, n1, nrhs, L11, ldl, B1, ldb);
    }
}

// ------------------------------------------------------------------ transposed-A GEMM
// C = alpha A^T B + beta C (A is k x m). Used only by the transposed triangular solves
// (adjoint); FP64/FP32 via a transposing copy into a workspace, then the NN kernel.
struct TransWork {
    DBuf<char> buf;
    void* get(size_t bytes) {
        if (buf.n < bytes) buf.alloc(std::max<size_t>(bytes, 1 << 20));
        return buf.p;
    }
};
static TransWork g_trans;

template <typename T>
__global__ void transpose_kernel(int rows, int cols, const T* __restrict__ A, int lda, T* __restrict__ B, int ldb) {
    __shared__ T tile[32][33];
    int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
    for (int r = ty; r < 32; r += 8) {
        int i = bx + tx, j = by + r;
        if (i < rows && j < cols) tile[r][tx] = A[i + (size_t)j * lda];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int j = by + tx, i = bx + r;
        if (i < rows && j < cols) B[j + (size_t)i * ldb] = tile[tx][r];
    }
}

template <typename T>
void gemm_tn(int kind, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C, int ldc) {
    if (m <= 0 || n <= 0) return;
    T* At = (T*)g_trans.get((size_t)m * k * sizeof(T));
    hipLaunchKernelGGL(transpose_kernel<T>, dim3(ceil_div(k, 32), ceil_div(m, 32)), dim3(256), 0, ctx().s(), k, m, A,
                       lda, At, m);
    check_launch("transpose");
    gemm_kind<T>(kind, m, n, k, alpha, At, m, B, ldb, beta, C, ldc);
}
template void gemm_tn<double>(int, int, int, int, double, const double*, int, const double*, int, double, double*,
                              int);
template void gemm_tn<float>(int, int, int, int, float, const float*, int, const float*, int, float, float*, int);

// ------------------------------------------------------------------ inverse-based triangular solves
// B <- T^-1 B for the n x n triangle T at L whose IB x IB diagonal blocks have their
// inverses at inv (consecutive blocks, column-major IB x IB). Recursive halving at
// multiples of IB; a leaf is one in-place GEMM with the block inverse (m <= IB < 128: one
// row tile, every workgroup reads its columns of B completely before writing them).
template <typename T>
static void gemm_exact(int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C, int ldc);
template <>
void gemm_exact<double>(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
                        double beta, double* C, int ldc) {
    dgemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
template <>
void gemm_exact<float>(int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb, float beta,
                       float* C, int ldc) {
    sgemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// The bottom two levels of trsm_inv in one launch, for 128 < n <= 256 with 128 x 128
// inverted diagonal blocks: workgroup j takes columns 64 j .. 64 j + 63 of B through that
// level's three products (lower: X1 = inv1 B1, B2 -= L21 X1, X2 = inv2 B2; upper: the
// mirror image), each one tile of the workgroup in the storage precision. Every product
// reads its operands completely before it writes its tile (in place where C is B), and the
// barrier between two products orders the workgroup's global writes before its next reads.
constexpr int TRSM2_B = 128, TRSM2_TN = 64;
template <typename T, bool LOWER>
__global__ void __launch_bounds__(256) trsm2_kernel(int n, int nrhs, const T* L, int ldl, const T* inv, T* B,
                                                    int ldb) {
    const int n1 = TRSM2_B, n2 = n - TRSM2_B, tn = blockIdx.x;
    const T* inv2 = inv + (size_t)TRSM2_B * TRSM2_B;
    if (LOWER) {
        gemm_nn_tile<T, TRSM2_TN>(n1, nrhs, n1, T(1), inv, TRSM2_B, B, ldb, T(0), B, ldb, 0, tn);
        __syncthreads();
        gemm_nn_tile<T, TRSM2_TN>(n2, nrhs, n1, T(-1), L + n1, ldl, B, ldb, T(1), B + n1, ldb, 0, tn);
        __syncthreads();
        gemm_nn_tile<T, TRSM2_TN>(n2, nrhs, n2, T(1), inv2, TRSM2_B, B + n1, ldb, T(0), B + n1, ldb, 0, tn);
    } else {
        gemm_nn_tile<T, TRSM2_TN>(n2, nrhs, n2, T(1), inv2, TRSM2_B, B + n1, ldb, T(0), B + n1, ldb, 0, tn);
        __syncthreads();
        gemm_nn_tile<T, TRSM2_TN>(n1, nrhs, n2, T(-1), L + (size_t)n1 * ldl, ldl, B + n1, ldb, T(1), B, ldb, 0, tn);
        __syncthreads();
        gemm_nn_tile<T, TRSM2_TN>(n1, nrhs, n1, T(1), inv, TRSM2_B, B, ldb, T(0), B, ldb, 0, tn);
    }
}

// FP32 form of trsm2_kernel with the workgroup's TN-column panel of B (n <= 256 rows) held
// in LDS through the three products: the panel is read once and written once, only the
// operands inv1, L21 (U12) and inv2 pass through two LDS stages. The operands do not depend
// on the panel, so each is loaded whole into registers before the product that needs it
// (the first two at the start, the third while the first product runs): a workgroup waits
// for global memory about once instead of once per 32-deep k tile. The products are the
// same MFMA sums in the same order as gemm_nn_tile's (k tiles of 32, steps of 4), so the
// results equal trsm2_kernel's.
template <bool LOWER, int TN, bool MID16 = false>
__global__ void __launch_bounds__(256, TN <= 32 ? 2 : 1) trsm2p_kernel(int n, int nrhs, const float* L, int ldl,
                                                                       const float* inv, float* B, int ldb,
                                                                       __bf16* sh, int shld, __bf16* shr, int shrld) {
    using Op = MfmaOp<float>;
    using Acc = typename Op::acc;
    constexpr int KB = GemmK<float>::KB, PAD = GemmK<float>::PAD, NQ = KB / 2, S = KB / 4;
    constexpr int NKT = TRSM2_B / KB;  // k tiles of a product (k <= 128)
    // waves of 64 rows x TN/2 columns (2 x 2), or, for 16-column panels, of 32 rows x 16
    // columns (4 x 1): MI 16-row and JT 16-column MFMA tiles each
    constexpr int MI = TN == 16 ? 2 : 4, JT = TN == 16 ? 1 : TN / 32;
    constexpr int NP = 2 * TRSM2_B + 34;  // panel column stride (34 mod 64 words)
    __shared__ __attribute__((aligned(16))) float Pn[TN * NP];        // Pn[c * NP + r] = B(r, c)
    __shared__ __attribute__((aligned(16))) float As[2][GBM][PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = TN == 16 ? wave * 32 : (wave & 1) * 64, wn = TN == 16 ? 0 : (wave >> 1) * (TN / 2);
    const int lr = lane & 15, lk = lane >> 4;
    const int n0 = blockIdx.x * TN, n1 = TRSM2_B, n2 = n - TRSM2_B;
    const float* inv2 = inv + (size_t)TRSM2_B * TRSM2_B;
    // the three operands in the order used: A (mr x kd, column-major at lda)
    const float* opA[3];
    int opld[3], opm[3], opk[3];
    if (LOWER) {
        opA[0] = inv;    opld[0] = TRSM2_B; opm[0] = n1; opk[0] = n1;  // X1 = inv1 B1
        opA[1] = L + n1; opld[1] = ldl;     opm[1] = n2; opk[1] = n1;  // B2 -= L21 X1
        opA[2] = inv2;   opld[2] = TRSM2_B; opm[2] = n2; opk[2] = n2;  // X2 = inv2 B2
    } else {
        opA[0] = inv2;                    opld[0] = TRSM2_B; opm[0] = n2; opk[0] = n2;  // X2 = inv2 B2
        opA[1] = L + (size_t)n1 * ldl;    opld[1] = ldl;     opm[1] = n1; opk[1] = n2;  // B1 -= U12 X2
        opA[2] = inv;                     opld[2] = TRSM2_B; opm[2] = n1; opk[2] = n1;  // X1 = inv1 B1
    }
    // thread t holds row t % 128 and k values t / 128 + 2 q of every k tile of an operand
    const int a_i = tid & 127, a_k0 = tid >> 7;
    float ra[NKT][NQ], rb[NKT][NQ];
    auto fetch = [&](float (&r)[NKT][NQ], int o) {
        const float* A = opA[o];
        const int lda = opld[o], ai = min(a_i, opm[o] - 1), kd = opk[o];
#pragma unroll
        for (int kt = 0; kt < NKT; ++kt)
#pragma unroll
            for (int q = 0; q < NQ; ++q) r[kt][q] = A[ai + (size_t)min(kt * KB + a_k0 + 2 * q, kd - 1) * lda];
    };
    fetch(ra, 0);
    fetch(rb, 1);
    // the panel: thread t moves row t of every column (zero beyond n and nrhs)
    {
        float v[TN];
#pragma unroll
        for (int c = 0; c < TN; ++c) v[c] = B[min(tid, n - 1) + (size_t)min(n0 + c, nrhs - 1) * ldb];
#pragma unroll
        for (int c = 0; c < TN; ++c) Pn[c * NP + tid] = (tid < n && n0 + c < nrhs) ? v[c] : 0.0f;
    }
    // acc = A(0:mr, 0:kd) P(ko:ko+kd, :) with A held in r
    Acc acc[MI][JT];
    auto product = [&](const float (&r)[NKT][NQ], int o, int ko) {
        const int mr = opm[o], kd = opk[o];
        const int nk = (kd + KB - 1) / KB;
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < JT; ++j) acc[i][j] = Acc{0, 0, 0, 0};
        auto store = [&](int buf, int kt) {
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                As[buf][a_i][a_k0 + 2 * q] = (a_i < mr && kt * KB + a_k0 + 2 * q < kd) ? r[kt][q] : 0.0f;
        };
        store(0, 0);
        __syncthreads();  // also orders the panel's earlier writes before its reads below
#pragma unroll
        for (int kt = 0; kt < NKT; ++kt) {
            if (kt < nk) {
                const int buf = kt & 1;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    float af[MI], bf[JT];
#pragma unroll
                    for (int i = 0; i < MI; ++i) af[i] = As[buf][wm + i * 16 + lr][4 * s + lk];
#pragma unroll
                    for (int j = 0; j < JT; ++j) bf[j] = Pn[(wn + j * 16 + lr) * NP + ko + kt * KB + 4 * s + lk];
#pragma unroll
                    for (int i = 0; i < MI; ++i)
#pragma unroll
                        for (int j = 0; j < JT; ++j) acc[i][j] = Op::mma(af[i], bf[j], acc[i][j]);
                }
                if (kt + 1 < nk) store(buf ^ 1, kt + 1);
                __syncthreads();
            }
        }
    };
    // MID16: the off-diagonal product (o = 1) with BF16 MFMA (16x16x32: lane (lr, lk) holds
    // A(row, 8 lk .. 8 lk + 7) and B(8 lk .. 8 lk + 7, col), rounded to BF16 as the 16-bit
    // products above these levels round their operands), one MFMA step per 32-deep k tile
    auto product16 = [&](const float (&r)[NKT][NQ], int o, int ko) {
        const int mr = opm[o], kd = opk[o];
        const int nk = (kd + KB - 1) / KB;
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < JT; ++j) acc[i][j] = Acc{0, 0, 0, 0};
        auto store = [&](int buf, int kt) {
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                As[buf][a_i][a_k0 + 2 * q] = (a_i < mr && kt * KB + a_k0 + 2 * q < kd) ? r[kt][q] : 0.0f;
        };
        store(0, 0);
        __syncthreads();
#pragma unroll
        for (int kt = 0; kt < NKT; ++kt) {
            if (kt < nk) {
                const int buf = kt & 1;
                b8v af[MI], bv[JT];
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int e = 0; e < 8; ++e) af[i][e] = (__bf16)As[buf][wm + i * 16 + lr][8 * lk + e];
#pragma unroll
                for (int j = 0; j < JT; ++j)
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        bv[j][e] = (__bf16)Pn[(wn + j * 16 + lr) * NP + ko + kt * KB + 8 * lk + e];
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < JT; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bv[j], acc[i][j], 0, 0, 0);
                if (kt + 1 < nk) store(buf ^ 1, kt + 1);
                __syncthreads();
            }
        }
    };
    // rows ro .. ro + mr - 1 of the panel: = acc (set) or -= acc (sub, as gemm's -acc + c)
    auto put = [&](int ro, int mr, bool sub) {
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < JT; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = wm + i * 16 + Op::row(lk, r);
                    if (row < mr) {
                        float* p = &Pn[(wn + j * 16 + lr) * NP + ro + row];
                        *p = sub ? -acc[i][j][r] + *p : acc[i][j][r];
                    }
                }
    };
    // panel rows each product reads and writes: lower 0 | 0 -> n1 | n1; upper n1 | n1 -> 0 | 0
    const int ko0 = LOWER ? 0 : n1, ko1 = LOWER ? 0 : n1, ko2 = LOWER ? n1 : 0;
    const int ro0 = LOWER ? 0 : n1, ro1 = LOWER ? n1 : 0, ro2 = LOWER ? n1 : 0;
    product(ra, 0, ko0);  // its reads of the panel ended at its last barrier
    fetch(ra, 2);         // the third operand, in flight during the second product
    put(ro0, opm[0], false);
    __syncthreads();
    if constexpr (MID16)
        product16(rb, 1, ko1);
    else
        product(rb, 1, ko1);
    put(ro1, opm[1], true);
    __syncthreads();
    product(ra, 2, ko2);
    put(ro2, opm[2], false);
    __syncthreads();
    if (tid < n) {
#pragma unroll
        for (int c = 0; c < TN; ++c)
            if (n0 + c < nrhs) {
                const float v = Pn[c * NP + tid];
                B[tid + (size_t)(n0 + c) * ldb] = v;
                if (sh) sh[tid + (size_t)(n0 + c) * shld] = (__bf16)v;  // BF16 copy (column-major)
            }
    }
    if (shr) {
        // the row-major BF16 copy: consecutive columns of a row on consecutive lanes
        for (int e = tid; e < n * TN; e += 256) {
            const int r = e / TN, c = e % TN;
            if (n0 + c < nrhs) shr[(size_t)r * shrld + n0 + c] = (__bf16)Pn[c * NP + r];
        }
    }
}

// Triangular solve through inverted BS x BS diagonal blocks (consecutive in inv, leading
// dimension BS): recursive split, each leaf one in-place GEMM X = inv_kk B_k in the exact
// (storage) precision, the off-diagonal updates in the precision of kind.
template <typename T, bool LOWER, int BS = IB>
static void trsm_inv(int kind, int n, int nrhs, const T* L, int ldl, const T* inv, T* B, int ldb) {
    if (n <= 0 || nrhs <= 0) return;
    // a leaf's results are final rows of X: their BF16 copy, where B's matrix has one
    int shld = 0, shrld = 0;
    __bf16 *sh = nullptr, *shr = nullptr;
    if constexpr (sizeof(T) == 4) {
        if (n <= 2 * BS) {
            sh = shadow_out((const float*)B, ldb, n, nrhs, false, &shld);
            shr = shadow_out((const float*)B, ldb, n, nrhs, true, &shrld);
        }
    }
    if (n <= BS) {
        ProfScope prof(PROF_TRSM, 2.0 * n * n * nrhs, sizeof(T) * (2.0 * n * nrhs + (double)n * n));
        if (sh || shr)
            gemm_launch<T>(n, nrhs, n, T(1), inv, BS, B, ldb, T(0), B, ldb, nullptr, sh, shld, shr, shrld);
        else
            gemm_exact<T>(n, nrhs, n, T(1), inv, BS, B, ldb, T(0), B, ldb);
        return;
    }
    if constexpr (BS == TRSM2_B) {
        if (n <= 2 * BS) {  // split_point gives n1 = BS here
            ProfScope prof(PROF_TRSM, 2.0 * n * n * nrhs, sizeof(T) * (2.0 * n * nrhs + (double)n * n));
            if constexpr (sizeof(T) == 4) {
                // narrower panels when 64-column ones would leave CUs idle (narrow right-hand
                // sides inside the LU): 32 columns, or 16 when even those fill under half the
                // CUs; more workgroups, less work each
                // BF16 mode: the off-diagonal product in BF16, like the updates above it
                const int cus = std::max(ctx().cu_count, 1);
                auto go = [&](auto mid_tag) {
                    constexpr bool M16 = decltype(mid_tag)::value;
                    if (ceil_div(nrhs, 32) < cus / 2)
                        hipLaunchKernelGGL((trsm2p_kernel<LOWER, 16, M16>), dim3((unsigned)ceil_div(nrhs, 16)),
                                           dim3(256), 0, ctx().s(), n, nrhs, L, ldl, inv, B, ldb, sh, shld, shr,
                                           shrld);
                    else if (ceil_div(nrhs, TRSM2_TN) < cus)
                        hipLaunchKernelGGL((trsm2p_kernel<LOWER, 32, M16>), dim3((unsigned)ceil_div(nrhs, 32)),
                                           dim3(256), 0, ctx().s(), n, nrhs, L, ldl, inv, B, ldb, sh, shld, shr,
                                           shrld);
                    else
                        hipLaunchKernelGGL((trsm2p_kernel<LOWER, TRSM2_TN, M16>),
                                           dim3((unsigned)ceil_div(nrhs, TRSM2_TN)), dim3(256), 0, ctx().s(), n, nrhs,
                                           L, ldl, inv, B, ldb, sh, shld, shr, shrld);
                };
                if (kind == SQBEM_BF16_IR)
                    go(std::true_type{});
                else
                    go(std::false_type{});
            }
            else
                hipLaunchKernelGGL((trsm2_kernel<T, LOWER>), dim3((unsigned)ceil_div(nrhs, TRSM2_TN)), dim3(256), 0,
                                   ctx().s(), n, nrhs, L, ldl, inv, B, ldb);
            check_launch("trsm2");
            return;
        }
    }
    const int n1 = split_point(n, BS), n2 = n - n1;
    const T* inv2 = inv + (size_t)(n1 / BS) * BS * BS;
    if (LOWER) {
        trsm_inv<T, LOWER, BS>(kind, n1, nrhs, L, ldl, inv, B, ldb);
        gemm_kind<T>(kind, n2, nrhs, n1, T(-1), L + n1, ldl, B, ldb, T(1), B + n1, ldb);
        trsm_inv<T, LOWER, BS>(kind, n2, nrhs, L + n1 + (size_t)n1 * ldl, ldl, inv2, B + n1, ldb);
    } else {
        trsm_inv<T, LOWER, BS>(kind, n2, nrhs, L + n1 + (size_t)n1 * ldl, ldl, inv2, B + n1, ldb);
        gemm_kind<T>(kind, n1, nrhs, n2, T(-1), L + (size_t)n1 * ldl, ldl, B + n1, ldb, T(1), B, ldb);
        trsm_inv<T, LOWER, BS>(kind, n1, nrhs, L, ldl, inv, B, ldb);
    }
}

// Columns g_cols_from and beyond of the matrix being factored are written by work on another
// stream that completes at g_cols_ev: the no-pivot recursion waits for it before it first
// touches such a column (GetrfJob::cols_event).
static hipEvent_t g_cols_ev[2] = {nullptr, nullptr};
static long g_cols_from[2] = {LONG_MAX, LONG_MAX};
static void cols_needed(long upto) {
    for (int i = 0; i < 2; ++i) {
        if (g_cols_from[i] >= upto) continue;
        const hipEvent_t ev = g_cols_ev[i];
        g_cols_ev[i] = nullptr;
        g_cols_from[i] = LONG_MAX;
        SQ_HIP(hipStreamWaitEvent(ctx().s(), ev, 0));
    }
}

// ------------------------------------------------------------------ no-pivot blocked LU
// Matrices of this code's operators never need row interchanges (the multipliers of
// partial pivoting stay below ~0.7), so getrf first factors without pivoting in
// 128-column blocks and verifies |l| <= 1 for every multiplier, which is exactly the
// condition under which partial pivoting keeps every diagonal pivot (LAPACK's first-
// largest tie rule included). If any multiplier fails (or a pivot is zero or not finite),
// the matrix is restored from its copy and factored by the pivoting path.
//
// lu128_kernel factors one <= 128 x 128 diagonal block without pivoting and forms the
// inverses of its unit-lower L and upper U, in one workgroup of four waves, in the block's
// own precision (FP64 blocks in FP64; FP32 blocks, whose trailing updates are FP32 or 16-bit
// products, in FP32 like LAPACK's sgetrf). The block lives in LDS (row-major, 132 KiB in
// FP64) for the whole kernel. The factor runs in four
// 32-column panels. A panel is factored without barriers: every wave holds the panel's top
// 32 rows (lanes 0-31, the same rows in every wave) and 32 of the rows below it (lanes
// 32-63), one row per lane in registers, so the pivot row is always in its own wave and is
// read out of lane j (v_readlane). Then U12 = L11^-1 A12 (one column per thread) and the
// trailing update A22 -= L21 U12 (register tiles). The inverses follow in LAPACK trtri
// order in 32 x 32 blocks: the eight diagonal blocks by substitution (FP32: waves 0-1 the
// four L blocks, waves 2-3 the four U blocks; FP64: wave b, lanes 0-31 L's block b, lanes
// 32-63 U's block b, U taken in reversed order so that both are lower triangular; the same
// operations in the same order either way), then each block column from two triangular
// products. Every loop whose
// bounds depend on a panel or block index is unrolled with compile-time indices, so the
// register arrays stay in VGPRs and the loads of a step are independent of its sums.
// The padded part of a smaller block is the identity.
constexpr int NB = 128;
// the in-place GEMMs (L21 = A21 U11^-1, X = inv B) rely on one tile spanning the block
static_assert(GBM >= NB && GBN >= NB, "GEMM tile narrower than the inverted blocks");
constexpr int LU_LD = NB + 1;  // LDS row stride (elements)
constexpr int LU_THREADS = 256;

__device__ __forceinline__ double readlane_c(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_c(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
// Reciprocal of a pivot: FP64 blocks divide (correctly rounded, as LAPACK's dgetf2); FP32
// blocks, whose factors only precondition the FP64 refinement, take the hardware reciprocal
// (within 1 ulp): one instruction instead of a ten-instruction division on every step of
// the panels' dependency chain.
__device__ __forceinline__ double pivot_recip(double v) { return 1.0 / v; }
__device__ __forceinline__ float pivot_recip(float v) { return __builtin_amdgcn_rcpf(v); }

// Panel P (columns 32P .. 32P+31, rows 32P ..) of the factor. Waves 0 .. NW-1 take part.
template <int P, typename C>
__device__ __forceinline__ void lu128_panel(C* M, int t, int& mybad) {
    constexpr int p0 = 32 * P, NREST = 96 - 32 * P, NW = NREST >= 32 ? NREST / 32 : 1;
    const int lane = t & 63, wv = t >> 6;
    if (wv >= NW) return;
    const bool top = lane < 32;
    const bool valid = top || NREST > 0;
    const int row = valid ? (top ? p0 + lane : p0 + 32 + 32 * wv + (lane - 32)) : p0;
    C* Mr = M + row * LU_LD + p0;
    C a[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) a[c] = Mr[c];
    // Branch-free (selects, no divergent branches): the whole unrolled panel is one basic
    // block, so the scheduler overlaps each step's pivot reciprocal with the previous step's
    // updates. The values are those of the guarded form.
    int bad = 0;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const C piv = readlane_c(a[j], j);  // lane j holds row p0 + j in every wave
        bad |= (int)!(piv != C(0) && fabs(piv) <= std::numeric_limits<C>::max());
        const C rp = pivot_recip(piv);
        const bool below = valid && (!top || lane > j);
        const C lm = below ? a[j] * rp : C(0);
        bad |= (int)(below & !(fabs(lm) <= C(1)));
        a[j] = below ? lm : a[j];
#pragma unroll
        for (int c = j + 1; c < 32; ++c) a[c] = fma(-lm, readlane_c(a[c], j), a[c]);  // lm = 0 above
    }
    mybad |= bad;
    if (valid && (wv == 0 || !top)) {
#pragma unroll
        for (int c = 0; c < 32; ++c) Mr[c] = a[c];
    }
}

// U12 = L11^-1 A12 for panel P: rows 32P .. 32P+31, one column right of the panel per thread.
// Column-oriented: step q reads column q of L11 at once (independent loads) and updates every
// row below it, so row i still takes its updates in the order q = 0 .. i-1.
template <int P, typename C>
__device__ __forceinline__ void lu128_u12(C* M, int t) {
    constexpr int p0 = 32 * P, NC = 96 - 32 * P;
    if (t >= NC) return;
    const int col = p0 + 32 + t;
    C x[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) x[i] = M[(p0 + i) * LU_LD + col];
#pragma unroll
    for (int q = 0; q < 31; ++q) {
#pragma unroll
        for (int i = 0; i < 32; ++i)
            if (i > q) x[i] = fma(-M[(p0 + i) * LU_LD + p0 + q], x[q], x[i]);
    }
#pragma unroll
    for (int i = 1; i < 32; ++i) M[(p0 + i) * LU_LD + col] = x[i];
}

// A22 -= L21 U12 after panel P: rows and columns 32P+32 .., inner dimension 32. Thread
// (ty, tx) of a 16 x 16 grid owns rows ty + 16 ii and columns tx + 16 jj.
template <int P, typename C>
__device__ __forceinline__ void lu128_update(C* M, int t) {
    constexpr int p0 = 32 * P, r0 = p0 + 32, R = 96 - 32 * P, MT = R / 16;
    const int ty = t >> 4, tx = t & 15;
    C acc[MT][MT];
#pragma unroll
    for (int ii = 0; ii < MT; ++ii)
#pragma unroll
        for (int jj = 0; jj < MT; ++jj) acc[ii][jj] = M[(r0 + ty + 16 * ii) * LU_LD + r0 + tx + 16 * jj];
#pragma unroll 4
    for (int k = 0; k < 32; ++k) {
        C a[MT], b[MT];
#pragma unroll
        for (int ii = 0; ii < MT; ++ii) a[ii] = M[(r0 + ty + 16 * ii) * LU_LD + p0 + k];
#pragma unroll
        for (int jj = 0; jj < MT; ++jj) b[jj] = M[(p0 + k) * LU_LD + r0 + tx + 16 * jj];
#pragma unroll
        for (int ii = 0; ii < MT; ++ii)
#pragma unroll
            for (int jj = 0; jj < MT; ++jj) acc[ii][jj] = fma(-a[ii], b[jj], acc[ii][jj]);
    }
#pragma unroll
    for (int ii = 0; ii < MT; ++ii)
#pragma unroll
        for (int jj = 0; jj < MT; ++jj) M[(r0 + ty + 16 * ii) * LU_LD + r0 + tx + 16 * jj] = acc[ii][jj];
}

// Block column J of the inverse of L (S = 2 - J) or of U (S = J - 1), diagonal blocks
// already inverted; R = 32 (S + 1) rows. Thread (ty, tx) of a 32 x (NT / 32) grid owns rows
// ty + 32 ii and columns tx + (NT / 32) jj. Two barriers, whatever NT.
//   L: Tmp = X(r0.., r0..) L(r0.., cJ), X unit lower;  X(r0.., cJ) = -Tmp X(J, J), r0 = cJ + 32
//   U: Tmp = X(0..R, 0..R) U(0..R, cJ), X upper;       X(0..R, cJ)  = -Tmp X(J, J), R = cJ
template <bool LOWER, int J, typename C, int NT = LU_THREADS>
__device__ __forceinline__ void lu128_inv_column(C* M, C* Tmp, int t) {
    constexpr int cj = 32 * J, R = LOWER ? 96 - cj : cj, RB = R / 32, r0 = LOWER ? cj + 32 : 0;
    constexpr int TX = NT / 32, JJ = 32 / TX;  // thread (ty, tx) owns columns tx + TX jj
    const int ty = t / TX, tx = t % TX;
    C acc[RB][JJ];
#pragma unroll
    for (int ii = 0; ii < RB; ++ii)
#pragma unroll
        for (int jj = 0; jj < JJ; ++jj) acc[ii][jj] = C(0);
#pragma unroll
    for (int kk = 0; kk < RB; ++kk) {
#pragma unroll 8
        for (int qq = 0; qq < 32; ++qq) {
            const int q = 32 * kk + qq;  // index within the R x R triangle
            C b[JJ];
#pragma unroll
            for (int jj = 0; jj < JJ; ++jj) b[jj] = M[(r0 + q) * LU_LD + cj + tx + TX * jj];
#pragma unroll
            for (int ii = 0; ii < RB; ++ii) {
                // X(i, q), i = ty + 32 ii: only chunks kk <= ii (L) or kk >= ii (U) are nonzero;
                // in the diagonal chunk the triangle is applied by a select
                if (LOWER ? kk > ii : kk < ii) continue;
                const C m = M[(r0 + ty + 32 * ii) * LU_LD + r0 + q];
                C a = m;
                if (kk == ii) a = LOWER ? (qq < ty ? m : (qq == ty ? C(1) : C(0))) : (qq >= ty ? m : C(0));
#pragma unroll
                for (int jj = 0; jj < JJ; ++jj) acc[ii][jj] = fma(a, b[jj], acc[ii][jj]);
            }
        }
    }
#pragma unroll
    for (int ii = 0; ii < RB; ++ii)
#pragma unroll
        for (int jj = 0; jj < JJ; ++jj) Tmp[(ty + 32 * ii) * 33 + tx + TX * jj] = acc[ii][jj];
    __syncthreads();
#pragma unroll
    for (int ii = 0; ii < RB; ++ii)
#pragma unroll
        for (int jj = 0; jj < JJ; ++jj) acc[ii][jj] = C(0);
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        C tq[RB];
#pragma unroll
        for (int ii = 0; ii < RB; ++ii) tq[ii] = Tmp[(ty + 32 * ii) * 33 + q];
#pragma unroll
        for (int jj = 0; jj < JJ; ++jj) {
            // X(J, J)(q, c), c = tx + TX jj: unit lower needs q >= c, upper q <= c
            if (LOWER ? q < TX * jj : q > TX * jj + TX - 1) continue;
            const int c = tx + TX * jj;
            const C m = M[(cj + q) * LU_LD + cj + c];
            const C x = LOWER ? (q > c ? m : (q == c ? C(1) : C(0))) : (q <= c ? m : C(0));
#pragma unroll
            for (int ii = 0; ii < RB; ++ii) acc[ii][jj] = fma(tq[ii], x, acc[ii][jj]);
        }
    }
#pragma unroll
    for (int ii = 0; ii < RB; ++ii)
#pragma unroll
        for (int jj = 0; jj < JJ; ++jj) M[(r0 + ty + 32 * ii) * LU_LD + cj + tx + TX * jj] = -acc[ii][jj];
    __syncthreads();
}

// FP32 blocks: the trailing updates and the inverses' block columns as MFMA products
// (mfma_f32_16x16x4f32: lane l supplies A(l%16, l/16) and B(l/16, l%16) of a 16 x 4 x 16
// step and holds D(4 (l/16) + r, l%16), r = 0..3), one 16 x 16 output tile per wave at a
// time, operands read from the block in LDS. The sums are the same as the VALU forms',
// four terms at a time.
__device__ __forceinline__ f4v lu_mma(float a, float b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// A22 -= L21 U12 after panel P (rows and columns 32P+32 .., inner dimension 32): wave w
// takes PW consecutive 16 x 16 tiles of A22 and runs them as independent accumulators, the
// operand reads of all its tiles issued before the products.
template <int P>
__device__ __forceinline__ void lu128_update_mfma(float* M, int t) {
    constexpr int p0 = 32 * P, r0 = p0 + 32, NT = (96 - 32 * P) / 16, PW = (NT * NT + 3) / 4;
    const int lane = t & 63, wv = t >> 6, lr = lane & 15, lk = lane >> 4;
    f4v acc[PW];
    float a[PW][8], b[PW][8];
#pragma unroll
    for (int u = 0; u < PW; ++u) {
        const int tile = min(wv * PW + u, NT * NT - 1);  // a surplus tile repeats the last, unstored
        const int i0 = r0 + 16 * (tile / NT), j0 = r0 + 16 * (tile % NT);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[u][r] = M[(i0 + 4 * lk + r) * LU_LD + j0 + lr];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            a[u][q] = -M[(i0 + lr) * LU_LD + p0 + 4 * q + lk];
            b[u][q] = M[(p0 + 4 * q + lk) * LU_LD + j0 + lr];
        }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int u = 0; u < PW; ++u) acc[u] = lu_mma(a[u][q], b[u][q], acc[u]);
#pragma unroll
    for (int u = 0; u < PW; ++u) {
        const int tile = wv * PW + u;
        if (tile < NT * NT) {
            const int i0 = r0 + 16 * (tile / NT), j0 = r0 + 16 * (tile % NT);
#pragma unroll
            for (int r = 0; r < 4; ++r) M[(i0 + 4 * lk + r) * LU_LD + j0 + lr] = acc[u][r];
        }
    }
}

// Block column J of the inverse of L (unit lower) or of U, as lu128_inv_column, by the two
// waves of its chain: wave cw owns the 16-column half cw of the block column and runs its
// R / 16 row tiles as independent accumulators; Tmp is the chain's R x 32 product (row
// stride 33).
//   L: Tmp = X(r0.., r0..) L(r0.., cJ);  X(r0.., cJ) = -Tmp X(J, J)   (r0 = cJ + 32, R = 96 - cJ)
//   U: Tmp = X(0..R, 0..R) U(0..R, cJ);  X(0..R, cJ) = -Tmp X(J, J)   (R = cJ)
// k steps where a tile's triangular operand is zero are skipped at compile time; inside
// the remaining ones the triangle (and the unit diagonal of L's inverse) is applied by
// selects.
template <bool LOWER, int J>
__device__ __forceinline__ void lu128_inv_column_mfma(float* M, float* Tmp, int cw, int lane) {
    constexpr int cj = 32 * J, R = LOWER ? 96 - cj : cj, r0 = LOWER ? cj + 32 : 0, NTI = R / 16, KS = R / 4;
    const int lr = lane & 15, lk = lane >> 4;
    const int c = 16 * cw + lr;  // this lane's column of the block column (B operand, output)
    f4v acc[NTI];
#pragma unroll
    for (int ti = 0; ti < NTI; ++ti) acc[ti] = f4v{0.f, 0.f, 0.f, 0.f};
    // Tmp = X(r0.., r0..) B, X(i, q) nonzero for q <= i (L) or q >= i (U): row tile ti needs
    // k steps s < 4 ti + 4 (L) or s >= 4 ti (U)
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int q = 4 * s + lk;
        const float bv = M[(r0 + q) * LU_LD + cj + c];
#pragma unroll
        for (int ti = 0; ti < NTI; ++ti) {
            if (LOWER ? s < 4 * ti + 4 : s >= 4 * ti) {
                const int i = 16 * ti + lr;
                const float m = M[(r0 + i) * LU_LD + r0 + q];
                const float av = LOWER ? (q < i ? m : (q == i ? 1.f : 0.f)) : (q >= i ? m : 0.f);
                acc[ti] = lu_mma(av, bv, acc[ti]);
            }
        }
    }
#pragma unroll
    for (int ti = 0; ti < NTI; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r) Tmp[(16 * ti + 4 * lk + r) * 33 + c] = acc[ti][r];
    __syncthreads();
    // X(r0.., cJ) = -Tmp X(J, J), X(J, J)(q, c) nonzero for q >= c (L) or q <= c (U): this
    // wave's column half needs k steps s >= 4 cw (L) or s < 4 cw + 4 (U)
#pragma unroll
    for (int ti = 0; ti < NTI; ++ti) acc[ti] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        if (LOWER ? s >= 4 * cw : s < 4 * cw + 4) {
            const int q = 4 * s + lk;
            const float m = M[(cj + q) * LU_LD + cj + c];
            const float bv = LOWER ? (q > c ? m : (q == c ? 1.f : 0.f)) : (q <= c ? m : 0.f);
#pragma unroll
            for (int ti = 0; ti < NTI; ++ti) acc[ti] = lu_mma(Tmp[(16 * ti + lr) * 33 + q], bv, acc[ti]);
        }
    }
#pragma unroll
    for (int ti = 0; ti < NTI; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r) M[(r0 + 16 * ti + 4 * lk + r) * LU_LD + cj + c] = -acc[ti][r];
    __syncthreads();
}

// factor = 1: factor the block at A (grid of one), check it, write L\U back, and the
// inverses. factor = 0: the blocks are already factored; workgroup b inverts diagonal
// block b (A advanced by b NB (lda + 1), inverses by b NB^2).
template <typename T>
__device__ __forceinline__ void lu128_block(T* A, int lda, int n, T* Linv, T* Uinv, int* flag, int factor, int b) {
    __shared__ T M[NB * LU_LD];  // the block, row-major: M[i * LU_LD + j] (132 KiB in FP64)
    __shared__ T Tmp[(sizeof(T) == 4 ? 2 : 1) * 96 * 33];  // products of the inversion (one per chain in FP32)
    __shared__ int bad;
    A += (size_t)b * NB * ((size_t)lda + 1);
    Linv += (size_t)b * NB * NB;
    Uinv += (size_t)b * NB * NB;
    const int w = min(NB, n - b * NB);
    const int t = threadIdx.x;
    if (t == 0) bad = 0;
    // Thread t moves row t % 128 of columns t / 128 + 2 m. Loads are issued 16 (FP32: 32) at a time
    // from clamped (always valid) addresses and masked afterwards: a guarded load would be
    // waited for before the next one is issued.
    const int ri = t & (NB - 1), rk = t >> 7;
    const T* Ai = A + min(ri, w - 1);
    constexpr int LG = sizeof(T) == 4 ? 32 : 16;  // loads in flight a round
#pragma unroll 1
    for (int m0 = 0; m0 < NB / 2; m0 += LG) {
        T v[LG];
#pragma unroll
        for (int m = 0; m < LG; ++m) v[m] = Ai[(size_t)min(rk + 2 * (m0 + m), w - 1) * lda];
#pragma unroll
        for (int m = 0; m < LG; ++m) {
            const int k = rk + 2 * (m0 + m);
            M[ri * LU_LD + k] = (ri < w && k < w) ? v[m] : (ri == k ? T(1) : T(0));
        }
    }
    __syncthreads();
    if (factor) {
        int mybad = 0;
        lu128_panel<0, T>(M, t, mybad);
        __syncthreads();
        lu128_u12<0, T>(M, t);
        __syncthreads();
        if constexpr (sizeof(T) == 4) lu128_update_mfma<0>(M, t);
        else lu128_update<0, T>(M, t);
        __syncthreads();
        lu128_panel<1, T>(M, t, mybad);
        __syncthreads();
        lu128_u12<1, T>(M, t);
        __syncthreads();
        if constexpr (sizeof(T) == 4) lu128_update_mfma<1>(M, t);
        else lu128_update<1, T>(M, t);
        __syncthreads();
        lu128_panel<2, T>(M, t, mybad);
        __syncthreads();
        lu128_u12<2, T>(M, t);
        __syncthreads();
        if constexpr (sizeof(T) == 4) lu128_update_mfma<2>(M, t);
        else lu128_update<2, T>(M, t);
        __syncthreads();
        lu128_panel<3, T>(M, t, mybad);
        if (mybad) bad = 1;
        __syncthreads();
#pragma unroll 1
        for (int m0 = 0; m0 < NB / 2; m0 += 16) {
            T v[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) v[m] = M[ri * LU_LD + rk + 2 * (m0 + m)];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int k = rk + 2 * (m0 + m);
                if (ri < w && k < w) A[ri + (size_t)k * lda] = v[m];
            }
        }
        __syncthreads();  // every wave has read L\U before the diagonal blocks are inverted in place
    }
    if constexpr (sizeof(T) == 4) {
        // the four 32 x 32 diagonal blocks in place: waves 0-1 invert L_bb, waves 2-3 U_bb;
        // lane c of half h of wave w forms column c of block b = 2 (w & 1) + h by substitution
        // (row i of inv(L) from rows 0 .. i-1, row i of inv(U) from rows 31 .. i+1 and a
        // division by U(i, i)). Every operand sits at a compile-time offset from the block.
        const int lane = t & 63, wv = t >> 6, c = lane & 31;
        T* Bb = M + 32 * (2 * (wv & 1) + (lane >> 5)) * (LU_LD + 1);
        T x[32];
        if (wv < 2) {
#pragma unroll
            for (int i = 0; i < 32; ++i) {
                T v = (i == c) ? T(1) : T(0);
#pragma unroll
                for (int q = 0; q < 32; ++q)
                    if (q < i) v = fma(-Bb[i * LU_LD + q], x[q], v);
                x[i] = v;
            }
#pragma unroll
            for (int i = 1; i < 32; ++i)
                if (i > c) Bb[i * LU_LD + c] = x[i];
        } else {
#pragma unroll
            for (int ii = 0; ii < 32; ++ii) {
                const int i = 31 - ii;
                T v = (i == c) ? T(1) : T(0);
#pragma unroll
                for (int qq = 0; qq < 32; ++qq) {
                    const int q = 31 - qq;
                    if (q > i) v = fma(-Bb[i * LU_LD + q], x[q], v);
                }
                x[i] = v * pivot_recip(Bb[i * LU_LD + i]);
            }
#pragma unroll
            for (int i = 0; i < 32; ++i)
                if (i <= c) Bb[i * LU_LD + c] = x[i];
        }
    } else {
        // FP64 (the split above spills there): lane c of the lower half of wave b forms
        // column c of inv(L_bb), lane c of the upper half column c of inv(Ur) with
        // Ur(i, q) = U_bb(31-i, 31-q); the same operations in the same order as above
        const int lane = t & 63, o = 32 * (t >> 6), c = lane & 31;
        const bool up = lane >= 32;
        const int base = up ? (o + 31) * LU_LD + o + 31 : o * LU_LD + o;
        const int s = up ? -1 : 1;  // element (i, q) of the lower-triangular operand: base + s (i LU_LD + q)
        T x[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            T v = (i == c) ? T(1) : T(0);
#pragma unroll
            for (int q = 0; q < i; ++q) v = fma(-M[base + s * (i * LU_LD + q)], x[q], v);
            x[i] = up ? v / M[base + s * (i * LU_LD + i)] : v;
        }
#pragma unroll
        for (int i = 0; i < 32; ++i)
            if (up ? i >= c : i > c) M[base + s * (i * LU_LD + c)] = x[i];
    }
    __syncthreads();
    if constexpr (sizeof(T) == 4) {
        // the L and U chains write disjoint parts of M (strictly lower, upper) and read only
        // their own: waves 0-1 run L's block columns, waves 2-3 U's, two barriers a step each
        const int cw = (t >> 6) & 1, lane = t & 63;
        if (t < 128) {
            lu128_inv_column_mfma<true, 2>(M, Tmp, cw, lane);
            lu128_inv_column_mfma<true, 1>(M, Tmp, cw, lane);
            lu128_inv_column_mfma<true, 0>(M, Tmp, cw, lane);
        } else {
            lu128_inv_column_mfma<false, 1>(M, Tmp + 96 * 33, cw, lane);
            lu128_inv_column_mfma<false, 2>(M, Tmp + 96 * 33, cw, lane);
            lu128_inv_column_mfma<false, 3>(M, Tmp + 96 * 33, cw, lane);
        }
    } else {
        lu128_inv_column<true, 2, T>(M, Tmp, t);
        lu128_inv_column<true, 1, T>(M, Tmp, t);
        lu128_inv_column<true, 0, T>(M, Tmp, t);
        lu128_inv_column<false, 1, T>(M, Tmp, t);
        lu128_inv_column<false, 2, T>(M, Tmp, t);
        lu128_inv_column<false, 3, T>(M, Tmp, t);
    }
#pragma unroll 1
    for (int m0 = 0; m0 < NB / 2; m0 += 16) {  // column-major outputs, consecutive rows a wave
        T v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = M[ri * LU_LD + rk + 2 * (m0 + m)];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int c = rk + 2 * (m0 + m);
            Linv[ri + (size_t)c * NB] = ri > c ? v[m] : (ri == c ? T(1) : T(0));
            Uinv[ri + (size_t)c * NB] = ri <= c ? v[m] : T(0);
        }
    }
    if (factor && t == 0 && bad) atomicOr(flag, 1);
}

template <typename T>
__global__ void __launch_bounds__(LU_THREADS) lu128_kernel(T* A, int lda, int n, T* Linv, T* Uinv, int* flag,
                                                           int factor) {
    lu128_block<T>(A, lda, n, Linv, Uinv, flag, factor, blockIdx.x);
}

// one <= 128 x 128 diagonal block per workgroup, each of its own matrix (factor and invert)
template <typename T>
struct Lu128Job {
    T* A;
    int lda, n;
    T* Linv;
    T* Uinv;
    int* flag;
};
template <typename T>
__global__ void __launch_bounds__(LU_THREADS) lu128_batch_kernel(const Lu128Job<T>* __restrict__ jobs) {
    const Lu128Job<T>& j = jobs[blockIdx.x];
    lu128_block<T>(j.A, j.lda, j.n, j.Linv, j.Uinv, j.flag, 1, 0);
}

template <typename T>
static void lu128_launch(int grid, T* A, int lda, int n, T* Linv, T* Uinv, int* flag, int factor) {
    hipLaunchKernelGGL(lu128_kernel<T>, dim3(grid), dim3(LU_THREADS), 0, ctx().s(), A, lda, n, Linv, Uinv, flag,
                       factor);
    check_launch("lu128");
}



// The two products that follow the first diagonal block of a pair of leaves (w = 128 + w2
// columns, w2 <= 128) in one launch, as they depend on that block alone: workgroups below
// tiles_l form L21 = A21 U11^-1 in 64-row tiles (in place, multipliers checked, the BF16
// copy row-major where the factor has one), the others U12 = L11^-1 A12 in 64-column tiles
// (in place, each tile holding all 128 rows of its columns; the BF16 copy column-major).
// Both tile shapes have the same LDS block.
template <typename T>
__global__ void __launch_bounds__(256) l21_u12_kernel(int m, int w2, T* A, int lda, const T* __restrict__ Linv,
                                                      const T* __restrict__ Uinv, int* flag, int tiles_l,
                                                      __bf16* sh_row, int sh_rld, __bf16* sh_col, int sh_cld,
                                                      __bf16* sh_urow, int sh_urld) {
    const int b = blockIdx.x;
    if (b < tiles_l)
        gemm_nn_tile<T, NB, 64>(m - NB, NB, NB, T(1), A + NB, lda, Uinv, NB, T(0), A + NB, lda, b, 0, flag, nullptr,
                                0, sh_row, sh_rld);
    else
        gemm_nn_tile<T, 64, NB>(NB, w2, NB, T(1), Linv, NB, A + (size_t)NB * lda, lda, T(0),
                                A + (size_t)NB * lda, lda, 0, b - tiles_l, nullptr, sh_col, sh_cld, sh_urow, sh_urld);
}

template <typename T>
static void getrf_np(int kind, T* A, int lda, int m, int w, int base, T* Linv, T* Uinv, int* flag) {
    if (w <= 2 * NB) cols_needed((long)base + w);  // a leaf or a pair of leaves: its own columns
    if (w > NB && w <= 2 * NB && m > NB) {
        // a pair of leaves: the first diagonal block, its L21 and U12 together, the update,
        // then the second leaf
        const int w2 = w - NB;
        const size_t off = (size_t)(base / NB) * NB * NB;
        {
            ProfScope prof(PROF_PANEL, (2.0 / 3.0) * NB * NB * NB, 2.0 * sizeof(T) * NB * NB);
            lu128_launch<T>(1, A, lda, NB, Linv + off, Uinv + off, flag, 1);
        }
        int srl = 0, scl = 0, sul = 0;
        __bf16 *sr = nullptr, *sc = nullptr, *su = nullptr;
        if constexpr (sizeof(T) == 4) {
            sr = shadow_out((const float*)(A + NB), lda, m - NB, NB, true, &srl);
            sc = shadow_out((const float*)(A + (size_t)NB * lda), lda, NB, w2, false, &scl);
            su = shadow_out((const float*)(A + (size_t)NB * lda), lda, NB, w2, true, &sul);
        }
        T* A12 = A + (size_t)NB * lda;
        const int tiles_l = (int)ceil_div(m - NB, 64), tiles_u = (int)ceil_div(w2, 64);
        {
            ProfScope prof(sizeof(T) == 8 ? PROF_DGEMM : PROF_SGEMM, 2.0 * (m - NB) * NB * NB + 2.0 * NB * NB * w2,
                           sizeof(T) * (2.0 * (m - NB) * NB + 2.0 * NB * w2 + 2.0 * NB * NB));
            hipLaunchKernelGGL(l21_u12_kernel<T>, dim3((unsigned)(tiles_l + tiles_u)), dim3(256), 0, ctx().s(), m, w2,
                               A, lda, Linv + off, Uinv + off, flag, tiles_l, sr, srl, sc, scl, su, sul);
            check_launch("l21_u12");
        }
        gemm_kind<T>(kind, m - NB, w2, NB, T(-1), A + NB, lda, A12, lda, T(1), A12 + NB, lda);
        getrf_np<T>(kind, A12 + NB, lda, m - NB, w2, base + NB, Linv, Uinv, flag);
        return;
    }
    if (w <= NB) {
        const size_t off = (size_t)(base / NB) * NB * NB;
        {
            ProfScope prof(PROF_PANEL, (2.0 / 3.0) * w * w * w, 2.0 * sizeof(T) * w * w);
            lu128_launch<T>(1, A, lda, w, Linv + off, Uinv + off, flag, 1);
        }
        if (m > w) {
            // L21 = A21 U11^-1, in place (one column tile, each row tile read before written),
            // with every multiplier checked against |l| <= 1 as it is stored (and its BF16
            // copy written row-major where the factor has one)
            int sld = 0;
            __bf16* sr = nullptr;
            if constexpr (sizeof(T) == 4) sr = shadow_out((const float*)(A + w), lda, m - w, w, true, &sld);
            gemm_launch<T>(m - w, w, w, T(1), A + w, lda, Uinv + off, NB, T(0), A + w, lda, flag, nullptr, 0, sr,
                           sld);
        }
        return;
    }
    const int w1 = split_point(w, NB), w2 = w - w1;
    getrf_np<T>(kind, A, lda, m, w1, base, Linv, Uinv, flag);
    T* A12 = A + (size_t)w1 * lda;
    cols_needed((long)base + w);  // the right part's columns from here on
    trsm_inv<T, true, NB>(kind, w1, w2, A, lda, Linv + (size_t)(base / NB) * NB * NB, A12, lda);
    gemm_kind<T>(kind, m - w1, w2, w1, T(-1), A + w1, lda, A12, lda, T(1), A12 + w1, lda);
    getrf_np<T>(kind, A12 + w1, lda, m - w1, w2, base + w1, Linv, Uinv, flag);
}

// ------------------------------------------------------------------ skinny triangular solves
// B (n x nrhs, nrhs <= 8) <- T^-1 B in one launch, reading the triangle once. Workgroup k,
// in the order workgroups start (an atomic ticket), owns the k-th 128-row block in solve
// order. It streams its block row of the triangle tile by tile (the loads do not depend
// on the solution, so they are issued before any wait), applies each solved block x_j as
// soon as the completion counter shows it, then forms its own x_k and publishes it.
// Blocks complete in ticket order, so one monotonic counter serves as every block's flag,
// and a workgroup only waits on tickets taken before its own: the chain cannot deadlock
// whatever the residency. Both counters grow by nb per launch and are never reset; a wait
// is bounded and a timeout is reported through the device error word. Sums run in FP64.
struct TrsvState {
    unsigned long long ticket, done;
};
static DBuf<TrsvState> g_trsv;
static unsigned long long g_trsv_base = 0;
// the tagged words of the single-system chain (trsv_tag_body), and the base past which the
// counters and words are started over, far below where a 32-bit tag would wrap
static DBuf<unsigned long long> g_trsv_words;
constexpr unsigned long long kTagLimit = 0xF0000000ull;

// after a chain reported an error its counters may be out of step: start them over (and
// clear the tagged words, whose tags restart with the counters)
void trsv_reset_counters() {
    if (g_trsv.p) g_trsv.zero();
    g_trsv_words.zero();
    g_trsv_base = 0;
    g_swaps.zero();  // the pivot compaction's completion counter
}

// B (n x nrhs, nrhs <= 8) <- T^-1 B with the inverted 128 x 128 diagonal blocks of T at
// hand (Dinv: nb blocks, NB x NB each, consecutive): the same ticket-ordered chain as
// trsv_chain with 128-row blocks, where a block's own solve is the product of its inverse
// (staged in LDS while the workgroup waits) with the accumulated right-hand side, spread
// over all 256 threads. Thread t owns row t % 128 and column half t / 128 of every tile.
template <typename T, bool LOWER, int NR>
__device__ __forceinline__ void trsv_inv_body(int n, int nrhs, const T* __restrict__ Lt, const T* __restrict__ Dinv,
                                              T* B, int ldb, TrsvState* st, unsigned long long base, int* err,
                                              unsigned char* tv_smem) {
    constexpr int GQ = sizeof(T) == 4 ? 2 : 1;
    constexpr int DLD = NB + 1;
    T* Dv = reinterpret_cast<T*>(tv_smem);                                    // [NB][DLD]
    // the tile buffer (main loop) and the reduction buffers (after it) share one region
    T(*Xs)[NR][NB] = reinterpret_cast<T(*)[NR][NB]>(tv_smem + sizeof(T) * NB * DLD);  // [GQ][NR][NB]
    double(*part)[NR][NB] = reinterpret_cast<double(*)[NR][NB]>(tv_smem + sizeof(T) * NB * DLD);  // [2][NR][NB]
    double(*rv)[NB] = reinterpret_cast<double(*)[NB]>(tv_smem + sizeof(T) * NB * DLD +
                                                      sizeof(double) * 2 * NR * NB);  // [NR][NB]
    __shared__ int s_k;
    __shared__ unsigned long long s_seen;
    const int tid = threadIdx.x, row = tid & (NB - 1), hf = tid >> 7;
    const int nb = (n + NB - 1) / NB;
    if (tid == 0) {
        const unsigned long long tk =
            __hip_atomic_fetch_add(&st->ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // a ticket below the launch's base means the host and device counters disagree: the
        // block cannot be named, so the call fails instead of touching rows outside B
        s_k = tk < base ? -1 : (tk - base >= (unsigned long long)nb ? nb : (int)(tk - base));
        if (s_k < 0) __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const int k = s_k;
    if (k < 0 || k >= nb) return;  // counter mismatch, or surplus workgroup of a batched launch
    const int blk = LOWER ? k : nb - 1 - k;
    const int r0 = blk * NB, rb = min(NB, n - r0);
    {
        const T* D = Dinv + (size_t)blk * NB * NB;
        for (int e = tid; e < NB * NB; e += 256) Dv[(e % NB) * DLD + e / NB] = D[e];
    }
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0.0;
    const bool rowok = row < rb;
    // this block's right-hand sides, read now: only this workgroup writes these rows of B,
    // so the read leaves the chain's critical path
    T bown[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) bown[r] = (hf == 0 && r < nrhs && rowok) ? B[(r0 + row) + (size_t)r * ldb] : T(0);
    // tiles of this block row, contiguous in the tiled copy (zero beyond n)
    const T* Lrow = Lt + (size_t)blk * nb * NB * NB + (size_t)(hf * 64) * NB + row;
    unsigned long long seen = base;
    for (int q0 = 0; q0 < k; q0 += GQ) {
        const int nq = min(GQ, k - q0);
        T l[GQ][64];
        // unconditional loads (a tile past the group's end re-reads its last one, unused), so
        // all of them are in flight together: a guarded load would be waited for one by one
#pragma unroll
        for (int g = 0; g < GQ; ++g) {
            const int qq = q0 + min(g, nq - 1);
            const int jb = LOWER ? qq : nb - 1 - qq;
            const T* tile = Lrow + (size_t)jb * NB * NB;
#pragma unroll
            for (int c = 0; c < 64; ++c) l[g][c] = tile[(size_t)c * NB];
        }
        int qd = q0;
        while (qd < q0 + nq) {
            if (seen < base + qd + 1) {
                if (tid == 0) {
                    unsigned long long d;
                    unsigned spins = 0;
                    while ((d = __hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) <
                           base + qd + 1) {
                        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                        if (++spins > (1u << 22)) {
                            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    s_seen = d >= base + qd + 1 ? d : base + (unsigned long long)k;
                }
                __syncthreads();
                seen = s_seen;
            }
            const int upto = min(q0 + nq, (int)(seen - base));
            for (int e = tid; e < (upto - qd) * NR * NB; e += 256) {
                const int g = e / (NR * NB), r = (e / NB) % NR, i = e % NB;
                const int jb = LOWER ? qd + g : nb - 1 - (qd + g);
                const int rr = jb * NB + i;
                Xs[qd - q0 + g][r][i] = (r < nrhs && rr < n) ? B[rr + (size_t)r * ldb] : T(0);
            }
            __syncthreads();
#pragma unroll
            for (int g = 0; g < GQ; ++g) {
                if (g >= qd - q0 && g < upto - q0) {
#pragma unroll
                    for (int c = 0; c < 64; ++c) {
                        const double lv = (double)l[g][c];
#pragma unroll
                        for (int r = 0; r < NR; ++r) acc[r] += lv * (double)Xs[g][r][hf * 64 + c];
                    }
                }
            }
            __syncthreads();
            qd = upto;
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) part[hf][r][row] = acc[r];
    __syncthreads();
    if (hf == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
            rv[r][row] = (r < nrhs && rowok) ? (double)bown[r] - (part[0][r][row] + part[1][r][row]) : 0.0;
    }
    __syncthreads();
    // x = Dinv r: thread (row, hf) sums over the columns of its half
    double xs[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) xs[r] = 0.0;
    for (int c = 0; c < 64; ++c) {
        const int cc = hf * 64 + c;
        const double dv = (double)Dv[row * DLD + cc];
#pragma unroll
        for (int r = 0; r < NR; ++r) xs[r] += dv * rv[r][cc];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) part[hf][r][row] = xs[r];
    __syncthreads();
    if (hf == 0 && rowok) {
        for (int r = 0; r < nrhs; ++r) B[(r0 + row) + (size_t)r * ldb] = (T)(part[0][r][row] + part[1][r][row]);
        __threadfence();
    }
    __syncthreads();
    if (tid == 0)
        __hip_atomic_fetch_max(&st->done, base + (unsigned long long)k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// The chained solve of one system (trsv_inv) hands solved values from block to block in
// tagged words instead of through a completion counter: every value travels as 32-bit words
// (one for FP32, two for FP64), each in the low half of a 64-bit word whose high half is the
// tag of the block that wrote it (the launch's base + its position in solve order + 1).
// Writer and readers use relaxed device-scope 64-bit atomics, so a reader that sees its tag
// has that launch's value: no fence, counter or cache invalidation sits on the chain. Tags
// grow from launch to launch (the host restarts them, and clears the words, long before
// they wrap), so a word left by an earlier launch never carries a tag a reader waits for.
// Workgroup k forms y_k = Dinv_k (b_k - sum_{q < k-1} T_kq x_q) before it waits for block
// k-1 (x_q of earlier blocks arrive while it streams its block row), then only
// x_k = y_k - M_k x_{k-1} remains, with M_k = Dinv_k T_{k,k-1} formed once per
// factorisation (chain_premul_kernel). Waits are bounded; a timeout, or a ticket below the
// launch's base, is reported through the device error word.
template <typename T, bool LOWER, int NR>
__device__ __forceinline__ void trsv_tag_body(int n, int nrhs, const T* __restrict__ Lt, const T* __restrict__ Dinv,
                                              const T* __restrict__ Mk, T* B, int ldb, unsigned long long* XT,
                                              TrsvState* st, unsigned long long base, int* err,
                                              unsigned char* tv_smem) {
    constexpr int GQ = sizeof(T) == 4 ? 2 : 1;
    constexpr int W = sizeof(T) / 4;  // 32-bit words a value
    constexpr int DLD = NB + 1;
    T* Dv = reinterpret_cast<T*>(tv_smem);                                              // [NB][DLD]
    T(*Xs)[NR][NB] = reinterpret_cast<T(*)[NR][NB]>(tv_smem + sizeof(T) * NB * DLD);  // [GQ][NR][NB]
    // the partial sums of the two column halves, and the right-hand side of the block solve;
    // part shares its bytes with Xs
    double(*part)[NR][NB] = reinterpret_cast<double(*)[NR][NB]>(tv_smem + sizeof(T) * NB * DLD);
    double(*rv)[NB] = reinterpret_cast<double(*)[NB]>(tv_smem + sizeof(T) * NB * DLD + sizeof(double) * 2 * NR * NB);
    __shared__ int s_k;
    const int tid = threadIdx.x, row = tid & (NB - 1), hf = tid >> 7;
    const int nb = (n + NB - 1) / NB;
    if (tid == 0) {
        const unsigned long long tk =
            __hip_atomic_fetch_add(&st->ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_k = tk < base ? -1 : (tk - base >= (unsigned long long)nb ? nb : (int)(tk - base));
        if (s_k < 0) __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const int k = s_k;
    if (k < 0 || k >= nb) return;
    const int blk = LOWER ? k : nb - 1 - k;
    const int r0 = blk * NB, rb = min(NB, n - r0);
    {
        const T* D = Dinv + (size_t)blk * NB * NB;
        for (int e = tid; e < NB * NB; e += 256) Dv[(e % NB) * DLD + e / NB] = D[e];
    }
    const bool rowok = row < rb;
    T bown[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) bown[r] = (hf == 0 && r < nrhs && rowok) ? B[(r0 + row) + (size_t)r * ldb] : T(0);
    // row `row`, columns of half hf, of M_k (column-major 128 x 128 like the tiles)
    T mk[64];
    if (k > 0) {
        const T* M = Mk + (size_t)blk * NB * NB + (size_t)(hf * 64) * NB + row;
#pragma unroll
        for (int c = 0; c < 64; ++c) mk[c] = M[(size_t)c * NB];
    }
    int failed = 0;
    // x of solve-order blocks q0 .. q0+nq-1 into Xs: each thread polls its own words
    auto fetch = [&](int q0, int nq) {
        for (int e = tid; e < nq * NR * NB; e += 256) {
            const int g = e / (NR * NB), r = (e / NB) % NR, i = e % NB;
            const int q = q0 + g, jb = LOWER ? q : nb - 1 - q;
            const int rr = jb * NB + i;
            T xv = T(0);
            if (r < nrhs && rr < n) {
                const unsigned tag = (unsigned)(base + (unsigned long long)q + 1);
                unsigned bits[W];
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    const unsigned long long* word = XT + ((size_t)r * n + rr) * W + w;
                    unsigned long long v = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    unsigned spins = 0;
                    while ((unsigned)(v >> 32) != tag && !failed) {
                        if (++spins > (1u << 22)) {
                            failed = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        v = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    bits[w] = (unsigned)v;
                }
                if constexpr (W == 1) xv = __uint_as_float(bits[0]);
                else xv = __hiloint2double((int)bits[1], (int)bits[0]);
            }
            Xs[g][r][i] = xv;
        }
    };
    // tiles of this block row, contiguous in the tiled copy; blocks before k-1 are summed
    // before block k-1 is waited for
    const T* Lrow = Lt + (size_t)blk * nb * NB * NB + (size_t)(hf * 64) * NB + row;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0.0;
    const int kp = k > 0 ? k - 1 : 0;
    for (int q0 = 0; q0 < kp; q0 += GQ) {
        const int nq = min(GQ, kp - q0);
        T l[GQ][64];
        // unconditional loads (a tile past the group's end re-reads its last one, unused)
#pragma unroll
        for (int g = 0; g < GQ; ++g) {
            const int qq = q0 + min(g, nq - 1);
            const int jb = LOWER ? qq : nb - 1 - qq;
            const T* tile = Lrow + (size_t)jb * NB * NB;
#pragma unroll
            for (int c = 0; c < 64; ++c) l[g][c] = tile[(size_t)c * NB];
        }
        fetch(q0, nq);
        __syncthreads();
#pragma unroll
        for (int g = 0; g < GQ; ++g) {
            if (g < nq) {
#pragma unroll
                for (int c = 0; c < 64; ++c) {
                    const double lv = (double)l[g][c];
#pragma unroll
                    for (int r = 0; r < NR; ++r) acc[r] += lv * (double)Xs[g][r][hf * 64 + c];
                }
            }
        }
        __syncthreads();
    }
    // y = Dinv (b - sum), thread (row, hf) summing over the columns of its half
#pragma unroll
    for (int r = 0; r < NR; ++r) part[hf][r][row] = acc[r];
    __syncthreads();
    if (hf == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
            rv[r][row] = (r < nrhs && rowok) ? (double)bown[r] - (part[0][r][row] + part[1][r][row]) : 0.0;
    }
    __syncthreads();
    double y[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) y[r] = 0.0;
    for (int c = 0; c < 64; ++c) {
        const int cc = hf * 64 + c;
        const double dv = (double)Dv[row * DLD + cc];
#pragma unroll
        for (int r = 0; r < NR; ++r) y[r] += dv * rv[r][cc];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) part[hf][r][row] = y[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NR; ++r) y[r] = part[0][r][row] + part[1][r][row];
    __syncthreads();  // part is read by every thread before Xs (its bytes) is written
    if (k > 0) {
        fetch(k - 1, 1);
        __syncthreads();
        double s2[NR];
        if constexpr (sizeof(T) == 4) {
            // FP32 factors: this product, the only work left on the chain's critical path,
            // accumulates in FP32 (twice the FP64 rate) over four independent partial sums;
            // its rounding stays far below that of the FP32 factors themselves
            float f2[NR][4];
#pragma unroll
            for (int r = 0; r < NR; ++r)
#pragma unroll
                for (int u = 0; u < 4; ++u) f2[r][u] = 0.0f;
#pragma unroll
            for (int c = 0; c < 64; ++c) {
#pragma unroll
                for (int r = 0; r < NR; ++r) f2[r][c & 3] = fmaf(mk[c], Xs[0][r][hf * 64 + c], f2[r][c & 3]);
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) s2[r] = (double)((f2[r][0] + f2[r][1]) + (f2[r][2] + f2[r][3]));
        } else {
#pragma unroll
            for (int r = 0; r < NR; ++r) s2[r] = 0.0;
#pragma unroll
            for (int c = 0; c < 64; ++c) {
                const double mv = (double)mk[c];
#pragma unroll
                for (int r = 0; r < NR; ++r) s2[r] += mv * (double)Xs[0][r][hf * 64 + c];
            }
        }
        __syncthreads();  // Xs read by every thread before part (its bytes) is written
#pragma unroll
        for (int r = 0; r < NR; ++r) part[hf][r][row] = s2[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; ++r) y[r] -= part[0][r][row] + part[1][r][row];
    }
    if (failed) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (hf == 0 && rowok) {
        const unsigned long long tag = (unsigned long long)(unsigned)(base + (unsigned long long)k + 1) << 32;
        for (int r = 0; r < nrhs; ++r) {
            const T x = (T)y[r];
            B[(r0 + row) + (size_t)r * ldb] = x;
            unsigned bits[W];
            if constexpr (W == 1) {
                bits[0] = __float_as_uint(x);
            } else {
                bits[0] = (unsigned)__double2loint(x);
                bits[1] = (unsigned)__double2hiint(x);
            }
#pragma unroll
            for (int w = 0; w < W; ++w)
                __hip_atomic_store(XT + ((size_t)r * n + r0 + row) * W + w, tag | bits[w], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <typename T, bool LOWER, int NR>
__global__ void __launch_bounds__(256) trsv_inv_chain(int n, int nrhs, const T* __restrict__ Lt,
                                                      const T* __restrict__ Dinv, const T* __restrict__ Mk, T* B,
                                                      int ldb, unsigned long long* XT, TrsvState* st,
                                                      unsigned long long base, int* err) {
    extern __shared__ __attribute__((aligned(16))) unsigned char tv_smem[];
    trsv_tag_body<T, LOWER, NR>(n, nrhs, Lt, Dinv, Mk, B, ldb, XT, st, base, err, tv_smem);
}

// Independent chained solves, one right-hand side each, in one launch: row y of the grid
// serves job y with its own (zeroed) ticket and completion counters; workgroups beyond a
// job's block count leave at once.
template <typename T>
struct TrsvJob {
    int n;
    const T* Lt;
    const T* Dinv;
    T* B;
};
template <typename T, bool LOWER>
__global__ void __launch_bounds__(256) trsv_inv_batch_kernel(const TrsvJob<T>* __restrict__ jobs, TrsvState* st,
                                                             int* err) {
    extern __shared__ __attribute__((aligned(16))) unsigned char tv_smem[];
    const TrsvJob<T>& j = jobs[blockIdx.y];
    trsv_inv_body<T, LOWER, 1>(j.n, 1, j.Lt, j.Dinv, j.B, j.n, st + blockIdx.y, 0ull, err, tv_smem);
}

template <typename T, bool LOWER, int NR>
static void trsv_inv_launch(int n, int nrhs, const T* Lt, const T* Dinv, const T* Mk, T* B, int ldb,
                            unsigned long long base, int* err) {
    constexpr int GQ = sizeof(T) == 4 ? 2 : 1;
    constexpr size_t xs = sizeof(T) * GQ * NR * NB, red = sizeof(double) * 3 * NR * NB;
    constexpr size_t lds = sizeof(T) * NB * (NB + 1) + (xs > red ? xs : red);
    static_assert(lds + 64 <= 160 * 1024, "trsv_inv_chain exceeds the LDS of a CU");
    static bool lds_set = false;
    if (!lds_set) {
        SQ_HIP(hipFuncSetAttribute((const void*)trsv_inv_chain<T, LOWER, NR>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        lds_set = true;
    }
    const int nb = (n + NB - 1) / NB;
    hipLaunchKernelGGL((trsv_inv_chain<T, LOWER, NR>), dim3(nb), dim3(256), lds, ctx().s(), n, nrhs, Lt, Dinv, Mk, B,
                       ldb, g_trsv_words.p, g_trsv.p, base, err);
}

// one launch of the chained solve through inverted diagonal blocks; Lt is the tiled copy
// of the factors (tile_copy), Dinv the inverted diagonal blocks of the triangle used and Mk
// the products M_k = Dinv_k T_{k,k-1} (in solve order) of chain_premul
template <typename T, bool LOWER>
static void trsv_inv(int n, int nrhs, const T* Lt, const T* Dinv, const T* Mk, T* B, int ldb) {
    SQ_REQUIRE(nrhs >= 1 && nrhs <= 8, SQBEM_ERR_ARG, "trsv: 1..8 right-hand sides");
    // the launch's tags come from a host counter, which a captured graph would freeze
    SQ_REQUIRE(!capturing(), SQBEM_ERR_STATE, "trsv: the chained solve cannot be captured in a graph");
    const int nb = (n + NB - 1) / NB;
    const size_t words = (size_t)n * 8 * 2;  // up to 8 right-hand sides of two words
    if (!g_trsv.p || g_trsv_words.n < words || g_trsv_base + (unsigned long long)nb > kTagLimit) {
        // fresh words carry tag 0, below every tag a launch waits for
        if (g_trsv_words.n < words) g_trsv_words.alloc(words);
        if (!g_trsv.p) g_trsv.alloc(1);
        trsv_reset_counters();
    }
    int* err = device_error_word();
    const unsigned long long base = g_trsv_base;
    mark_device_error_check();
    side_join();  // a tiled factor copy may still be in the making on the side stream
    ProfScope prof(PROF_TRSV, (double)n * n * nrhs, sizeof(T) * 0.5 * n * n);
    if (nrhs == 1) trsv_inv_launch<T, LOWER, 1>(n, nrhs, Lt, Dinv, Mk, B, ldb, base, err);
    else if (nrhs == 2) trsv_inv_launch<T, LOWER, 2>(n, nrhs, Lt, Dinv, Mk, B, ldb, base, err);
    else if (nrhs <= 4) trsv_inv_launch<T, LOWER, 4>(n, nrhs, Lt, Dinv, Mk, B, ldb, base, err);
    else trsv_inv_launch<T, LOWER, 8>(n, nrhs, Lt, Dinv, Mk, B, ldb, base, err);
    check_launch("trsv_inv_chain");
    // the device counter advances by nb only once the launch is queued: a throw above leaves
    // host base and device ticket in step
    g_trsv_base = base + (unsigned long long)nb;
}

// Copy of an n x n matrix as 128 x 128 tiles, tile (I, J) at (I nb + J) 128^2, column-major
// inside (zero beyond n): a block row of tiles is then one contiguous range, which the
// chained solves stream far faster than 512-byte pieces of columns 4n bytes apart.
template <typename T>
__global__ void __launch_bounds__(256) tile_copy_kernel(int n, const T* __restrict__ A, int lda, T* __restrict__ At) {
    const int nb = (n + NB - 1) / NB;
    const int r = threadIdx.x & 127, c1 = threadIdx.x >> 7;
    for (int t = blockIdx.x; t < nb * nb; t += gridDim.x) {
        const int I = t % nb, J = t / nb;
        T* dst = At + ((size_t)I * nb + J) * NB * NB;
        const int gi = I * NB + r;
        for (int c0 = c1; c0 < NB; c0 += 32) {  // 16 loads in flight, then their stores
            T v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int gj = J * NB + c0 + 2 * q;
                v[q] = (gi < n && gj < n) ? A[gi + (size_t)gj * lda] : T(0);
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) dst[(size_t)(c0 + 2 * q) * NB + r] = v[q];
        }
    }
}

// On the side stream the copy runs beside the library stream's work with one workgroup on
// at most half the CUs, so the kernels queued there keep getting CUs (a workgroup per tile
// would fill the device and hold them off for the whole copy).
template <typename T>
static void tile_copy(int n, const T* A, int lda, T* At, hipStream_t stream = nullptr) {
    const int nb = (n + NB - 1) / NB;
    const unsigned tiles = (unsigned)nb * nb;
    const unsigned grid = stream ? std::min(tiles, (unsigned)std::max(ctx().cu_count / 2, 1)) : tiles;
    hipLaunchKernelGGL(tile_copy_kernel<T>, dim3(grid), dim3(256), 0, stream ? stream : ctx().s(), n, A, lda, At);
    check_launch("tile_copy");
}

// The products the single-system chain applies after its one wait, one 128 x 128 x 128
// product per workgroup (grid nb x 2): M^L_k = Linv_k L_{k,k-1} (k >= 1) and
// M^U_k = Uinv_k U_{k,k+1} (k <= nb - 2), from the inverted diagonal blocks and the tiled
// copy, column-major 128 x 128 at ML + k 128^2 and MU + k 128^2. Thread (ty, tx) of a
// 16 x 16 grid owns rows ty + 16 a and columns tx + 16 b; the k dimension is staged through
// LDS in KC-deep slices; FP64 sums for FP64 blocks, FP32 for FP32 ones.
template <typename T>
__global__ void __launch_bounds__(256) chain_premul_kernel(int nb, const T* __restrict__ Linv,
                                                           const T* __restrict__ Uinv, const T* __restrict__ Lt,
                                                           T* __restrict__ ML, T* __restrict__ MU) {
    constexpr int KC = sizeof(T) == 4 ? 32 : 16;
    __shared__ T As[NB][KC + 1];
    __shared__ T Bs[KC][NB + 1];
    const int k = blockIdx.x;
    const bool up = blockIdx.y != 0;
    if (up ? k >= nb - 1 : k == 0) return;
    const T* A = (up ? Uinv : Linv) + (size_t)k * NB * NB;
    const T* Bt = Lt + ((size_t)k * nb + (up ? k + 1 : k - 1)) * NB * NB;
    T* C = (up ? MU : ML) + (size_t)k * NB * NB;
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    T acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = T(0);
    for (int k0 = 0; k0 < NB; k0 += KC) {
        for (int e = t; e < NB * KC; e += 256) {
            As[e % NB][e / NB] = A[(size_t)(k0 + e / NB) * NB + e % NB];  // A(i, k0 + kk)
            Bs[e % KC][e / KC] = Bt[(size_t)(e / KC) * NB + k0 + e % KC];  // B(k0 + kk, j)
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < KC; ++kk) {
            T av[8], bv[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) av[a] = As[ty + 16 * a][kk];
#pragma unroll
            for (int b = 0; b < 8; ++b) bv[b] = Bs[kk][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) C[(size_t)(tx + 16 * b) * NB + ty + 16 * a] = acc[a][b];
}

// what the single-system chained solves read besides the inverted diagonal blocks (aux =
// getrf's solve aid): the tiled copy of the factors, then the chain products
template <typename T>
static void chain_prepare(int n, const T* A, int lda, T* aux, hipStream_t stream = nullptr) {
    const int nb = (n + NB - 1) / NB;
    const size_t blk = (size_t)NB * NB;
    T* Lt = aux + 2 * nb * blk;
    tile_copy<T>(n, A, lda, Lt, stream);
    T* ML = Lt + (size_t)nb * nb * blk;
    hipLaunchKernelGGL(chain_premul_kernel<T>, dim3((unsigned)nb, 2), dim3(256), 0, stream ? stream : ctx().s(), nb,
                       aux, aux + nb * blk, Lt, ML, ML + nb * blk);
    check_launch("chain_premul");
}

namespace {
std::map<const void*, std::function<void()>> g_pending_tiles;
}
void chain_tiles_issue(const void* aux) {
    auto it = g_pending_tiles.find(aux);
    if (it == g_pending_tiles.end()) return;
    const std::function<void()> fn = std::move(it->second);
    g_pending_tiles.erase(it);
    fn();
}
void chain_tiles_forget(const void* aux) { g_pending_tiles.erase(aux); }

// ------------------------------------------------------------------ recursive LU
template <typename T>
static void getrf_rec(int kind, T* A, int lda, int m, int w, int* ipiv, int base) {
    if (w <= LEAF) {
        panel_leaf<T>(A, lda, m, w, ipiv + base, base);
        return;
    }
    const int w1 = split_point(w, LEAF), w2 = w - w1;
    getrf_rec<T>(kind, A, lda, m, w1, ipiv, base);
    T* A12 = A + (size_t)w1 * lda;
    laswp<T>(A12, lda, w2, ipiv, base, base, 0, w1);
    const T* Linv = (const T*)g_panel.inv.p + (size_t)(base / IB) * IB * IB;
    trsm_inv<T, true>(kind, w1, w2, A, lda, Linv, A12, lda);
    gemm_kind<T>(kind, m - w1, w2, w1, T(-1), A + w1, lda, A12, lda, T(1), A12 + w1, lda);
    getrf_rec<T>(kind, A12 + w1, lda, m - w1, w2, ipiv, base + w1);
    laswp<T>(A, lda, w1, ipiv, base, base, w1, std::min(w, m));
}

template <typename T>
__global__ void ipiv_to_one_based(int n, int* ipiv, int add) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ipiv[i] += add;
}

__global__ void iota_pivots_kernel(int n, int* ipiv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ipiv[i] = i + 1;
}

// orders at least this large try the no-pivot factorisation first
constexpr int NP_MIN = 512;
static thread_local int g_last_getrf_path = 0;  // 0 pivoting, 1 no-pivot verified, 2 no-pivot rejected

// aux (optional, getrf_aux_elems(n) elements): receives the solve aid for getrs: the
// inverted 128 x 128 diagonal blocks of L then of U, the tiled copy of the factors, then the
// chained solves' products M^L_k and M^U_k (chain_premul_kernel).
template <typename T>
void getrf_begin(GetrfJob<T>& j) {
    j.started = true;
    if (j.n <= 0) return;
    SQ_REQUIRE(j.lda >= j.n, SQBEM_ERR_ARG, "getrf: lda < n");
    if (j.ncols > 0) {
        for (int i = 0; i < 2; ++i) {
            g_cols_ev[i] = i < j.ncols ? j.cols_event[i] : nullptr;
            g_cols_from[i] = i < j.ncols ? j.cols_from[i] : LONG_MAX;
        }
        // the pivoted path, and the copy kept for it, read every column at once
        if (j.n < NP_MIN || !j.restore) cols_needed(LONG_MAX);
    }
    if (j.n < NP_MIN) return;
    const int n = j.n;
    const size_t nb = (size_t)(n + NB - 1) / NB;
    if (!j.restore) {
        j.backup.alloc((size_t)n * n);
        SQ_HIP(hipMemcpy2DAsync(j.backup.p, (size_t)n * sizeof(T), j.A, (size_t)j.lda * sizeof(T),
                                (size_t)n * sizeof(T), n, hipMemcpyDeviceToDevice, ctx().s()));
    }
    j.flag.alloc(1);
    j.flag.zero();
    T* Linv = j.aux;
    if (!Linv) {
        j.inv.alloc(2 * nb * NB * NB);
        Linv = j.inv.p;
    }
    // BF16 copies of the factors, written as the no-pivot recursion finalises them and read
    // by its 16-bit products (registered while its launches are issued)
    std::unique_ptr<ShadowScope> sh;
    if constexpr (sizeof(T) == 4) {
        if (j.kind == SQBEM_BF16_IR && j.sh_lo && j.sh_up)
            sh.reset(new ShadowScope((const float*)j.A, j.lda, n, n, j.sh_lo, j.sh_up, j.sh_ld ? j.sh_ld : n, true,
                                     j.sh_upr));
    }
    getrf_np<T>(j.kind, j.A, j.lda, n, n, 0, Linv, Linv + nb * NB * NB, j.flag.p);
    cols_needed(LONG_MAX);  // nothing later in the stream may run before those columns are whole
}
int getrf_top_split(int n) { return split_point(n, NB); }

template <typename T>
int getrf_end(GetrfJob<T>& j) {
    SQ_REQUIRE(j.started, SQBEM_ERR_STATE, "getrf: end without begin");
    j.started = false;
    const int n = j.n, lda = j.lda;
    T* A = j.A;
    T* aux = j.aux;
    int* ipiv = j.ipiv;
    j.path = 0;
    if (n <= 0) return 0;
    g_last_getrf_path = 0;
    const size_t nb = (size_t)(n + NB - 1) / NB;
    if (n >= NP_MIN) {
        int bad = 0;
        j.flag.download(&bad, 1);
        j.inv.release();
        if (!bad) {
            hipLaunchKernelGGL(iota_pivots_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx().s(), n, ipiv);
            check_launch("iota_pivots");
            if (aux) {
                // the tiled copy is read only by the chained solves (which wait for it):
                // made on the side stream, beside what follows on the library stream (now,
                // or when the caller issues it)
                if (j.defer_tiles) {
                    g_pending_tiles[aux] = [n, A, lda, aux] {
                        side_defer();
                        chain_prepare<T>(n, A, lda, aux, ctx().side);
                    };
                } else {
                    side_defer();
                    chain_prepare<T>(n, A, lda, aux, ctx().side);
                }
            }
            j.backup.release();  // stream-ordered reuse only
            g_last_getrf_path = 1;
            j.path = 1;
            return 0;
        }
        if (j.restore)
            j.restore();
        else
            SQ_HIP(hipMemcpy2DAsync(A, (size_t)lda * sizeof(T), j.backup.p, (size_t)n * sizeof(T),
                                    (size_t)n * sizeof(T), n, hipMemcpyDeviceToDevice, ctx().s()));
        g_last_getrf_path = 2;
        j.path = 2;
    }
    g_panel.ensure(256, n, sizeof(T));
    g_panel.bar.zero();
    g_panel.status.zero();
    g_panel.flags.zero();
    g_bar_parity = 0;
    getrf_rec<T>(j.kind, A, lda, n, n, ipiv, 0);
    hipLaunchKernelGGL(ipiv_to_one_based<T>, dim3(ceil_div(n, 256)), dim3(256), 0, ctx().s(), n, ipiv, 1);
    check_launch("ipiv");
    int st[2];
    g_panel.status.download(st, 2);
    SQ_REQUIRE(st[1] == 0, SQBEM_ERR_HIP, "getrf: panel barrier timed out");
    if (aux && st[0] == 0) {
        lu128_launch<T>((int)nb, A, lda, n, aux, aux + nb * NB * NB, nullptr, 0);
        chain_prepare<T>(n, A, lda, aux);
    }
    return st[0];
}

template <typename T>
int getrf(int kind, int n, T* A, int lda, int* ipiv, T* aux, const std::function<void()>* restore, __bf16* sh_lo,
          __bf16* sh_up, int sh_ld) {
    GetrfJob<T> j;
    j.kind = kind;
    j.n = n;
    j.lda = lda;
    j.A = A;
    j.ipiv = ipiv;
    j.aux = aux;
    j.sh_lo = sh_lo;
    j.sh_up = sh_up;
    j.sh_ld = sh_ld;
    if (restore) j.restore = *restore;
    getrf_begin<T>(j);
    return getrf_end<T>(j);
}
template int getrf<double>(int, int, double*, int, int*, double*, const std::function<void()>*, __bf16*, __bf16*,
                           int);
template int getrf<float>(int, int, float*, int, int*, float*, const std::function<void()>*, __bf16*, __bf16*, int);
template void getrf_begin<double>(GetrfJob<double>&);
template void getrf_begin<float>(GetrfJob<float>&);
template int getrf_end<double>(GetrfJob<double>&);
template int getrf_end<float>(GetrfJob<float>&);

// number of panels that needed row interchanges in the last getrf (0 = diagonal pivots)
int last_getrf_pivoted_panels(int n) {
    const int nb = (n + LEAF - 1) / LEAF;
    std::vector<int> f(nb);
    SQ_HIP(hipMemcpyAsync(f.data(), g_panel.flags.p, nb * sizeof(int), hipMemcpyDeviceToHost, ctx().s()));
    sync();
    int c = 0;
    for (int v : f) c += v != 0;
    return c;
}

static DBuf<char> g_getrs_inv;

// B <- A^-1 B (trans = 0) or A^-T B (trans = 1) from getrf factors; ipiv 1-based. aux, when
// given, is getrf's solve aid for these factors (inverted 128 x 128 diagonal blocks of L and
// U, then the tiled copy of the factors); otherwise what a solve needs is formed here.
template <typename T>
void getrs(int kind, int trans, int n, int nrhs, const T* LU, int lda, const int* ipiv, T* B, int ldb,
           const T* aux) {
    if (n <= 0 || nrhs <= 0) return;
    if (!trans) {
        laswp<T>(B, ldb, nrhs, ipiv, 0, 1, 0, n, +1);
        const int nb = (n + NB - 1) / NB;
        const size_t ninv = 2 * (size_t)nb * NB * NB;
        if (!aux) {
            const size_t need = (nrhs <= 8 ? getrf_aux_elems(n) : ninv) * sizeof(T);
            if (g_getrs_inv.n < need) g_getrs_inv.alloc(need);
            T* w = (T*)g_getrs_inv.p;
            lu128_launch<T>(nb, const_cast<T*>(LU), lda, n, w, w + (size_t)nb * NB * NB, nullptr, 0);
            if (nrhs <= 8) chain_prepare<T>(n, LU, lda, w);
            aux = w;
        }
        const T* Linv = aux;
        const T* Uinv = aux + (size_t)nb * NB * NB;
        if (nrhs <= 8) {
            chain_tiles_issue(aux);  // a tiled copy left pending by the factorisation
            const T* Lt = aux + ninv;
            const T* ML = Lt + (size_t)nb * nb * NB * NB;
            trsv_inv<T, true>(n, nrhs, Lt, Linv, ML, B, ldb);
            trsv_inv<T, false>(n, nrhs, Lt, Uinv, ML + (size_t)nb * NB * NB, B, ldb);
            return;
        }
        // recursive solves whose leaves are single in-place GEMMs with the inverses
        trsm_inv<T, true, NB>(kind, n, nrhs, LU, lda, Linv, B, ldb);
        trsm_inv<T, false, NB>(kind, n, nrhs, LU, lda, Uinv, B, ldb);
    } else {
        // A^T = U^T L^T P: solve U^T (lower, non-unit), L^T (upper, unit), then P^T
        trsm_rec<T, true, false, true>(kind, n, nrhs, LU, lda, B, ldb);
        trsm_rec<T, false, true, true>(kind, n, nrhs, LU, lda, B, ldb);
        laswp<T>(B, ldb, nrhs, ipiv, 0, 1, 0, n, -1);
    }
}
template void getrs<double>(int, int, int, int, const double*, int, const int*, double*, int, const double*);
template void getrs<float>(int, int, int, int, const float*, int, const int*, float*, int, const float*);

// ------------------------------------------------------------------ batched factorisations
// A batch is planned on the host first: the recursion over the largest matrix is walked
// once, every step's per-matrix work (clipped to that matrix's order; matrices already
// finished drop out) becomes one job table, and all tables go to the device in a single
// copy before the launches are issued in order. A step is then one launch for the whole
// batch, so a batch of k matrices costs the launches of one factorisation, and the small
// diagonal-block kernels of different matrices run side by side.
template <typename T>
__global__ void __launch_bounds__(256) check_unit_bound_batch_kernel(const GemmJob<T>* __restrict__ jobs) {
    const GemmJob<T>& g = jobs[blockIdx.y];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const size_t total = (size_t)g.m * g.n;
    int mine = 0;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx % g.m), j = (int)(idx / g.m);
        if (!(fabs((double)g.C[i + (size_t)j * g.ldc]) <= 1.0)) mine = 1;
    }
    if (mine) bad = 1;
    __syncthreads();
    if (threadIdx.x == 0 && bad) atomicOr((int*)g.A, 1);  // A carries the flag of the job
}

namespace {
enum PlanKind { PK_LU128, PK_GEMM, PK_CHECK };
struct PlanStep {
    int kind;
    size_t off;
    int count, gridx;
};
struct Plan {
    std::vector<unsigned char> bytes;
    std::vector<PlanStep> steps;
    template <typename J>
    size_t add(const std::vector<J>& v) {
        size_t off = (bytes.size() + 63) / 64 * 64;
        bytes.resize(off + v.size() * sizeof(J));
        std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(J));
        return off;
    }
};

template <typename T>
struct NpPlanner {
    Plan& P;
    const std::vector<BatchMat<T>>& M;
    struct TJ {
        const T* L;
        int ldl;
        const T* inv;
        T* B;
        int ldb, nrhs;
    };
    static size_t nbk(int n) { return (size_t)(n + NB - 1) / NB; }
    T* Linv(int i, int base) const { return M[i].aux + (size_t)(base / NB) * NB * NB; }
    T* Uinv(int i, int base) const { return M[i].aux + (nbk(M[i].n) + (size_t)(base / NB)) * NB * NB; }
    void gemm(const std::vector<GemmJob<T>>& g, int kind = PK_GEMM) {
        if (g.empty()) return;
        int tiles = 1;
        for (const auto& j : g) tiles = std::max(tiles, (int)(ceil_div(j.m, GBM) * ceil_div(j.n, GBN)));
        if (kind == PK_CHECK) tiles = 64;
        P.steps.push_back({kind, P.add(g), (int)g.size(), tiles});
    }
    void trsm(int n, const std::vector<TJ>& J) {
        if (J.empty()) return;
        std::vector<GemmJob<T>> g;
        if (n <= NB) {
            // X = inv B in place: one row tile, each workgroup reads its columns before writing
            for (const TJ& j : J) g.push_back({n, j.nrhs, n, T(1), T(0), j.inv, NB, j.B, j.ldb, j.B, j.ldb});
            gemm(g);
            return;
        }
        const int n1 = split_point(n, NB), n2 = n - n1;
        trsm(n1, J);
        for (const TJ& j : J) g.push_back({n2, j.nrhs, n1, T(-1), T(1), j.L + n1, j.ldl, j.B, j.ldb, j.B + n1, j.ldb});
        gemm(g);
        std::vector<TJ> J2;
        for (const TJ& j : J)
            J2.push_back({j.L + n1 + (size_t)n1 * j.ldl, j.ldl, j.inv + (size_t)(n1 / NB) * NB * NB, j.B + n1, j.ldb,
                          j.nrhs});
        trsm(n2, J2);
    }
    void rec(int base, int w) {
        const int nm = (int)M.size();
        if (w <= NB) {
            std::vector<Lu128Job<T>> lu;
            std::vector<GemmJob<T>> g, chk;
            for (int i = 0; i < nm; ++i) {
                const int wi = std::min(w, M[i].n - base);
                if (wi <= 0) continue;
                T* Ad = M[i].A + base + (size_t)base * M[i].lda;
                lu.push_back({Ad, M[i].lda, wi, Linv(i, base), Uinv(i, base), M[i].flag});
                const int mi = M[i].n - base - wi;
                if (mi > 0) {
                    // L21 = A21 U11^-1 in place (one column tile)
                    g.push_back({mi, wi, wi, T(1), T(0), Ad + wi, M[i].lda, Uinv(i, base), NB, Ad + wi, M[i].lda});
                    chk.push_back({mi, wi, 0, T(0), T(0), (const T*)M[i].flag, 0, nullptr, 0, Ad + wi, M[i].lda});
                }
            }
            if (!lu.empty()) P.steps.push_back({PK_LU128, P.add(lu), (int)lu.size(), (int)lu.size()});
            gemm(g);
            gemm(chk, PK_CHECK);
            return;
        }
        const int w1 = split_point(w, NB), w2 = w - w1;
        rec(base, w1);
        std::vector<TJ> J;
        std::vector<GemmJob<T>> g;
        for (int i = 0; i < nm; ++i) {
            const int w2i = std::min(w2, M[i].n - base - w1);
            if (w2i <= 0) continue;
            const int ld = M[i].lda;
            T* Ad = M[i].A + base + (size_t)base * ld;
            T* A12 = Ad + (size_t)w1 * ld;
            J.push_back({Ad, ld, Linv(i, base), A12, ld, w2i});
            g.push_back({M[i].n - base - w1, w2i, w1, T(-1), T(1), Ad + w1, ld, A12, ld, A12 + w1, ld});
        }
        trsm(w1, J);
        gemm(g);
        rec(base + w1, w2);
    }
};

}  // namespace

template <typename T>
void getrf_np_batch_launch(const BatchLuPlan<T>& bp);

namespace {
// a plan run once: tables uploaded, launches issued, then a wait (the tables are staging)
template <typename T>
void run_plan(const Plan& P) {
    if (P.steps.empty()) return;
    BatchLuPlan<T> bp;
    for (const PlanStep& st : P.steps) bp.steps.push_back({st.kind, st.off, st.count, st.gridx});
    bp.tab.alloc(P.bytes.size());
    bp.tab.upload(P.bytes.data(), P.bytes.size());
    getrf_np_batch_launch<T>(bp);  // no matrices: the steps only
    sync();
}
}  // namespace

template <typename T>
void getrf_np_batch_prepare(const std::vector<BatchMat<T>>& mats, BatchLuPlan<T>& bp) {
    int nmax = 0;
    for (const auto& m : mats) {
        SQ_REQUIRE(m.n >= 0 && m.lda >= std::max(1, m.n) && m.A && m.aux && m.flag, SQBEM_ERR_ARG,
                   "getrf_np_batch: bad matrix");
        nmax = std::max(nmax, m.n);
    }
    bp.mats = mats;
    bp.steps.clear();
    bp.flops = 0;
    if (!nmax) return;
    Plan P;
    NpPlanner<T> pl{P, mats};
    pl.rec(0, nmax);
    for (const PlanStep& st : P.steps) bp.steps.push_back({st.kind, st.off, st.count, st.gridx});
    for (const auto& m : mats) bp.flops += (2.0 / 3.0) * m.n * (double)m.n * m.n;
    if (!P.bytes.empty()) {
        bp.tab.alloc(P.bytes.size());
        bp.tab.upload(P.bytes.data(), P.bytes.size());
        sync();  // the host plan is staging
    }
}

template <typename T>
void getrf_np_batch_launch(const BatchLuPlan<T>& bp) {
    if (bp.steps.empty()) return;
    for (const auto& m : bp.mats) fill_zero(m.flag, sizeof(int));
    {
        ProfScope prof(bp.mats.empty() ? -1 : PROF_PANEL, bp.flops, 0.0);  // no matrices: a plain GEMM plan
        for (const BatchLuStep& st : bp.steps) {
            const void* tab = bp.tab.p + st.off;
            if (st.kind == PK_LU128) {
                hipLaunchKernelGGL(lu128_batch_kernel<T>, dim3(st.count), dim3(LU_THREADS), 0, ctx().s(),
                                   (const Lu128Job<T>*)tab);
            } else if (st.kind == PK_GEMM) {
                hipLaunchKernelGGL(gemm_nn_batch_kernel<T>, dim3(st.gridx, st.count), dim3(256), 0, ctx().s(),
                                   (const GemmJob<T>*)tab);
            } else {
                hipLaunchKernelGGL(check_unit_bound_batch_kernel<T>, dim3(st.gridx, st.count), dim3(256), 0, ctx().s(),
                                   (const GemmJob<T>*)tab);
            }
            check_launch("batched factorisation step");
        }
    }
    for (const auto& m : bp.mats)
        if (m.n > 0) tile_copy<T>(m.n, m.A, m.lda, m.aux + 2 * NpPlanner<T>::nbk(m.n) * NB * NB);
}

template <typename T>
void getrf_np_batch(const std::vector<BatchMat<T>>& mats) {
    BatchLuPlan<T> bp;
    getrf_np_batch_prepare<T>(mats, bp);
    getrf_np_batch_launch<T>(bp);
    sync();  // the plan's table is released on return
}
template void getrf_np_batch<float>(const std::vector<BatchMat<float>>&);
template void getrf_np_batch<double>(const std::vector<BatchMat<double>>&);
template void getrf_np_batch_prepare<float>(const std::vector<BatchMat<float>>&, BatchLuPlan<float>&);
template void getrf_np_batch_prepare<double>(const std::vector<BatchMat<double>>&, BatchLuPlan<double>&);
template void getrf_np_batch_launch<float>(const BatchLuPlan<float>&);
template void getrf_np_batch_launch<double>(const BatchLuPlan<double>&);

// rows of a right-hand side gathered through a permutation: tmp = v[perm], then v = tmp
template <typename T>
struct PermJob {
    int n;
    const int* perm;
    T* v;
    T* tmp;
};
template <typename T>
__global__ void perm_gather_kernel(const PermJob<T>* __restrict__ jobs) {
    const PermJob<T>& j = jobs[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < j.n; i += gridDim.x * blockDim.x) j.tmp[i] = j.v[j.perm[i]];
}
template <typename T>
__global__ void perm_store_kernel(const PermJob<T>* __restrict__ jobs) {
    const PermJob<T>& j = jobs[blockIdx.y];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < j.n; i += gridDim.x * blockDim.x) j.v[i] = j.tmp[i];
}

template <typename T, bool LOWER>
static constexpr size_t trsv_batch_lds() {
    constexpr int GQ = sizeof(T) == 4 ? 2 : 1;
    constexpr size_t xs = sizeof(T) * GQ * NB, red = sizeof(double) * 3 * NB;
    return sizeof(T) * NB * (NB + 1) + (xs > red ? xs : red);
}
// the kernel's LDS limit, set once, outside any graph capture (from the prepare step)
template <typename T, bool LOWER>
static void trsv_batch_setup() {
    static bool lds_set = false;
    if (!lds_set) {
        SQ_HIP(hipFuncSetAttribute((const void*)trsv_inv_batch_kernel<T, LOWER>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)trsv_batch_lds<T, LOWER>()));
        lds_set = true;
    }
}
template <typename T, bool LOWER>
static void trsv_batch_launch(int njobs, const void* dev_jobs, int nbmax, TrsvState* state) {
    constexpr size_t lds = trsv_batch_lds<T, LOWER>();
    trsv_batch_setup<T, LOWER>();
    fill_zero(state, sizeof(TrsvState) * njobs);  // a kernel, not a memset: safe in captured graphs
    hipLaunchKernelGGL((trsv_inv_batch_kernel<T, LOWER>), dim3(nbmax, (unsigned)njobs), dim3(256), lds, ctx().s(),
                       (const TrsvJob<T>*)dev_jobs, state, device_error_word());
    check_launch("trsv_inv_batch");
}

template <typename T>
void getrs_np_batch_prepare(const std::vector<BatchSolve<T>>& s, BatchSolvePlan<T>& p) {
    std::vector<TrsvJob<T>> L, U;
    std::vector<PermJob<T>> P;
    p.njobs = (int)s.size();
    p.nbmax = 0;
    p.flops = p.bytes = 0;
    size_t tmp_need = 0;
    for (const auto& q : s) {
        const size_t nb = (size_t)(q.n + NB - 1) / NB;
        p.nbmax = std::max(p.nbmax, (int)nb);
        L.push_back({q.n, q.aux + 2 * nb * NB * NB, q.aux, q.B});
        U.push_back({q.n, q.aux + 2 * nb * NB * NB, q.aux + nb * NB * NB, q.B});
        p.flops += 2.0 * q.n * (double)q.n;
        p.bytes += sizeof(T) * (double)q.n * q.n;
        if (q.perm) tmp_need += q.n;
    }
    if (s.empty() || !p.nbmax) return;
    if (tmp_need) p.tmp.alloc(tmp_need);
    size_t toff = 0;
    for (const auto& q : s)
        if (q.perm) {
            P.push_back({q.n, q.perm, q.B, p.tmp.p + toff});
            toff += q.n;
        }
    p.nperm = (int)P.size();
    p.permmax = 1;
    for (const auto& j : P) p.permmax = std::max(p.permmax, j.n);
    p.half = (L.size() * sizeof(TrsvJob<T>) + 63) / 64 * 64;
    const size_t pbytes = P.size() * sizeof(PermJob<T>);
    std::vector<unsigned char> h(2 * p.half + pbytes);
    std::memcpy(h.data(), L.data(), L.size() * sizeof(TrsvJob<T>));
    std::memcpy(h.data() + p.half, U.data(), U.size() * sizeof(TrsvJob<T>));
    if (pbytes) std::memcpy(h.data() + 2 * p.half, P.data(), pbytes);
    p.jobs.alloc(h.size());
    p.state.alloc(std::max<size_t>(s.size(), 1));
    SQ_HIP(hipMemcpyAsync(p.jobs.p, h.data(), h.size(), hipMemcpyHostToDevice, ctx().s()));
    (void)device_error_word();  // allocated here, outside any capture
    trsv_batch_setup<T, true>();
    trsv_batch_setup<T, false>();
    sync();                     // h is staging
}

template <typename T>
void getrs_np_batch_launch(const BatchSolvePlan<T>& p) {
    if (!p.njobs || !p.nbmax) return;
    if (p.nperm) {
        const void* pj = p.jobs.p + 2 * p.half;
        const unsigned gx = ceil_div(p.permmax, 256);
        hipLaunchKernelGGL(perm_gather_kernel<T>, dim3(gx, p.nperm), dim3(256), 0, ctx().s(), (const PermJob<T>*)pj);
        hipLaunchKernelGGL(perm_store_kernel<T>, dim3(gx, p.nperm), dim3(256), 0, ctx().s(), (const PermJob<T>*)pj);
        check_launch("perm");
    }
    mark_device_error_check();
    ProfScope prof(PROF_TRSV, p.flops, p.bytes);
    trsv_batch_launch<T, true>(p.njobs, p.jobs.p, p.nbmax, p.state.p);
    trsv_batch_launch<T, false>(p.njobs, p.jobs.p + p.half, p.nbmax, p.state.p);
}

template <typename T>
void getrs_np_batch(const std::vector<BatchSolve<T>>& s) {
    BatchSolvePlan<T> p;
    getrs_np_batch_prepare<T>(s, p);
    getrs_np_batch_launch<T>(p);
    sync();  // surfaces a timed-out wait; p's buffers go back to the cache afterwards
}
template void getrs_np_batch<float>(const std::vector<BatchSolve<float>>&);
template void getrs_np_batch<double>(const std::vector<BatchSolve<double>>&);
template void getrs_np_batch_prepare<float>(const std::vector<BatchSolve<float>>&, BatchSolvePlan<float>&);
template void getrs_np_batch_prepare<double>(const std::vector<BatchSolve<double>>&, BatchSolvePlan<double>&);
template void getrs_np_batch_launch<float>(const BatchSolvePlan<float>&);
template void getrs_np_batch_launch<double>(const BatchSolvePlan<double>&);

template <typename T>
void gemm_batch(const std::vector<GemmJob<T>>& jobs) {
    std::vector<GemmJob<T>> g;
    for (const auto& j : jobs)
        if (j.m > 0 && j.n > 0) {
            SQ_REQUIRE(j.k > 0, SQBEM_ERR_ARG, "gemm_batch: k must be positive");
            g.push_back(j);
        }
    if (g.empty()) return;
    Plan P;
    int tiles = 1;
    double flops = 0;
    for (const auto& j : g) {
        tiles = std::max(tiles, (int)(ceil_div(j.m, GBM) * ceil_div(j.n, GBN)));
        flops += 2.0 * j.m * (double)j.n * j.k;
    }
    P.steps.push_back({PK_GEMM, P.add(g), (int)g.size(), tiles});
    ProfScope prof(sizeof(T) == 8 ? PROF_DGEMM : PROF_SGEMM, flops, 0.0);
    run_plan<T>(P);
}
template void gemm_batch<float>(const std::vector<GemmJob<float>>&);
template void gemm_batch<double>(const std::vector<GemmJob<double>>&);

}  // namespace sq

// ------------------------------------------------------------------ C ABI
using namespace sq;

extern "C" {

int sqbem_dgemm(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb, double beta,
                double* C, int ldc) {
    return guarded([&] {
        SQ_REQUIRE(m >= 0 && n >= 0 && k >= 0 && lda >= std::max(1, m) && ldb >= std::max(1, k) &&
                       ldc >= std::max(1, m),
                   SQBEM_ERR_ARG, "gemm: dimension mismatch");
        dgemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    });
}

int sqbem_sgemm(int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb, float beta,
                float* C, int ldc) {
    return guarded([&] {
        SQ_REQUIRE(m >= 0 && n >= 0 && k >= 0 && lda >= std::max(1, m) && ldb >= std::max(1, k) &&
                       ldc >= std::max(1, m),
                   SQBEM_ERR_ARG, "gemm: dimension mismatch");
        sgemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    });
}

int sqbem_gemm16(int kind, int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb,
                 float beta, float* C, int ldc) {
    return guarded([&] {
        SQ_REQUIRE(kind == SQBEM_FP16_IR || kind == SQBEM_BF16_IR, SQBEM_ERR_ARG, "gemm16: kind must be FP16 or BF16");
        SQ_REQUIRE(m >= 0 && n >= 0 && k >= 0 && lda >= std::max(1, m) && ldb >= std::max(1, k) &&
                       ldc >= std::max(1, m),
                   SQBEM_ERR_ARG, "gemm: dimension mismatch");
        gemm16(kind, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    });
}

int sqbem_gemm16_noalloc(int kind, int m, int n, int k, float alpha, const float* A, int lda, const float* B,
                         int ldb, float beta, float* C, int ldc) {
    // no implicit initialisation here (it allocates): the caller has called sqbem_init
    try {
        SQ_REQUIRE(ctx().device >= 0, SQBEM_ERR_STATE, "gemm16_noalloc: call sqbem_init first");
        int cur = -1;
        SQ_HIP(hipGetDevice(&cur));
        SQ_REQUIRE(cur == ctx().device, SQBEM_ERR_STATE, "gemm16_noalloc: another device is current");
        SQ_REQUIRE(kind == SQBEM_FP16_IR || kind == SQBEM_BF16_IR, SQBEM_ERR_ARG, "gemm16: kind must be FP16 or BF16");
        SQ_REQUIRE(m >= 0 && n >= 0 && k >= 0 && lda >= std::max(1, m) && ldb >= std::max(1, k) &&
                       ldc >= std::max(1, m),
                   SQBEM_ERR_ARG, "gemm: dimension mismatch");
        gemm16_noalloc(kind, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
        return SQBEM_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return SQBEM_ERR_ARG;
    }
}

unsigned long sqbem_gemm16_noalloc_bytes(int kind, int m, int n, int k) { return (unsigned long)gemm16_work_bytes(kind, m, n, k); }

int sqbem_dgetrf(int n, double* A, int lda, int* ipiv, int* info) {
    return guarded([&] {
        SQ_REQUIRE(n >= 0 && lda >= std::max(1, n), SQBEM_ERR_ARG, "dgetrf: bad dimensions");
        int i = getrf<double>(SQBEM_FP64, n, A, lda, ipiv);
        if (info) *info = i;
        SQ_REQUIRE(i == 0, SQBEM_ERR_SINGULAR, "dgetrf failed");
    });
}

int sqbem_dgetrs(int n, int nrhs, const double* LU, int lda, const int* ipiv, double* B, int ldb) {
    return guarded([&] {
        SQ_REQUIRE(n >= 0 && nrhs >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), SQBEM_ERR_ARG,
                   "dgetrs: bad dimensions");
        getrs<double>(SQBEM_FP64, 0, n, nrhs, LU, lda, ipiv, B, ldb);
    });
}

int sqbem_dgetrs_trans(int n, int nrhs, const double* LU, int lda, const int* ipiv, double* B, int ldb) {
    return guarded([&] {
        SQ_REQUIRE(n >= 0 && nrhs >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), SQBEM_ERR_ARG,
                   "dgetrs: bad dimensions");
        getrs<double>(SQBEM_FP64, 1, n, nrhs, LU, lda, ipiv, B, ldb);
    });
}

int sqbem_dgesv(int n, int nrhs, double* A, int lda, int* ipiv, double* B, int ldb, int* info) {
    return guarded([&] {
        SQ_REQUIRE(n >= 0 && nrhs >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), SQBEM_ERR_ARG,
                   "dgesv: bad dimensions");
        int i = getrf<double>(SQBEM_FP64, n, A, lda, ipiv);
        if (info) *info = i;
        SQ_REQUIRE(i == 0, SQBEM_ERR_SINGULAR, "dgesv failed");
        getrs<double>(SQBEM_FP64, 0, n, nrhs, A, lda, ipiv, B, ldb);
    });
}

int sqbem_sgetrf(int kind, int n, float* A, int lda, int* ipiv, int* info) {
    return guarded([&] {
        SQ_REQUIRE(n >= 0 && lda >= std::max(1, n), SQBEM_ERR_ARG, "sgetrf: bad dimensions");
        SQ_REQUIRE(kind == SQBEM_FP32_IR || kind == SQBEM_FP16_IR || kind == SQBEM_BF16_IR, SQBEM_ERR_ARG,
                   "sgetrf: kind must be FP32_IR, FP16_IR or BF16_IR");
        int i = getrf<float>(kind, n, A, lda, ipiv);
        if (info) *info = i;
        SQ_REQUIRE(i == 0, SQBEM_ERR_SINGULAR, "dgetrf failed");
    });
}

int sqbem_tune(int mode) {
    return guarded([&] {
        SQ_REQUIRE(mode == 0 || mode == 1, SQBEM_ERR_ARG, "sqbem_tune: mode must be 0 (fixed rule) or 1 (online)");
        tune::reset(mode);
    });
}

int sqbem_getrf_path(void) { return g_last_getrf_path; }

int sqbem_coop(int enable) {
    return guarded([&] {
        SQ_REQUIRE(enable == 0 || enable == 1, SQBEM_ERR_ARG, "sqbem_coop: enable must be 0 or 1");
        g_coop_ok = enable != 0;
    });
}

int sqbem_tune_info(int* classes, int* decided, int* counts) {
    return guarded([&] { tune::info(classes, decided, counts); });
}

int sqbem_tune_save(const char* path) {
    return guarded([&] {
        SQ_REQUIRE(path, SQBEM_ERR_ARG, "sqbem_tune_save: null path");
        tune::save(path);
    });
}

int sqbem_tune_load(const char* path, int* classes) {
    return guarded([&] {
        SQ_REQUIRE(path, SQBEM_ERR_ARG, "sqbem_tune_load: null path");
        const int n = tune::load(path);
        if (classes) *classes = n;
    });
}

int sqbem_sgetrs(int n, int nrhs, const float* LU, int lda, const int* ipiv, float* B, int ldb) {
    return guarded([&] {
        SQ_REQUIRE(n >= 0 && nrhs >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), SQBEM_ERR_ARG,
                   "sgetrs: bad dimensions");
        getrs<float>(SQBEM_FP32_IR, 0, n, nrhs, LU, lda, ipiv, B, ldb);
    });
}

}  // extern "C"
