// dense.hpp -- internal interface of the dense linear algebra (dense.hip).
#pragma once
#include <cstddef>
#include <functional>
#include <vector>

#include "common.hpp"

namespace sq {

// C = alpha A B + beta C on the library stream (column-major, device pointers)
void dgemm(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb, double beta,
           double* C, int ldc);
void sgemm(int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb, float beta,
           float* C, int ldc);
void gemm16(int kind, int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb,
            float beta, float* C, int ldc);
// C = alpha A^T B + beta C, A stored k x m
template <typename T>
void gemm_tn(int kind, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C, int ldc);
template <typename T>
void scale_matrix(int m, int n, T beta, T* C, int ldc);
template <typename Ti, typename To>
void convert_matrix(int m, int n, const Ti* A, int lda, To* B, int ldb);

// LU with partial pivoting; ipiv (device, n) 1-based on return; returns LAPACK info.
// kind selects the GEMM of the FP32 variant (SQBEM_FP32_IR / FP16_IR / BF16_IR).
// aux (optional, getrf_aux_elems(n) elements): receives a solve aid for getrs (inverted
// 128 x 128 diagonal blocks of L and U, a tiled copy of the factors, and the products the
// chained solves apply after their one wait).
// restore (optional): rebuilds A as it was on entry (e.g. from the matrix it was converted
// from); given, the first attempt without row interchanges keeps no copy of A and calls it
// instead when it has to fall back to partial pivoting.
template <typename T>
int getrf(int kind, int n, T* A, int lda, int* ipiv, T* aux = nullptr,
          const std::function<void()>* restore = nullptr, __bf16* sh_lo = nullptr, __bf16* sh_up = nullptr,
          int sh_ld = 0);
// The same in two halves (getrf = begin, then end): begin issues the factorisation without
// row interchanges (orders of 512 and more) and returns without waiting; end waits for its
// verdict, finishes, or refactors with partial pivoting, and returns the LAPACK info. In
// between the stream may carry other work that leaves A, ipiv and aux alone. The job owns
// the verdict word and, without restore, the copy of A (and, without aux, the inverses).
// sh_lo, sh_up (optional, BF16 kind, n x sh_ld each, sh_ld >= n, 0 = n): receive BF16
// copies of the factors as the no-pivot factorisation finalises them (sh_lo: the part below
// the 128 x 128 diagonal blocks, row-major, (i, j) at sh_lo[i sh_ld + j]; sh_up: the part
// above them, column-major), and its 16-bit products read their operands from there instead
// of converting them. path: 1 when the factors came from the verified no-pivot
// factorisation (the copies are then whole), 2 when it fell back to pivoting.
template <typename T>
struct GetrfJob {
    int kind = 0, n = 0, lda = 0;
    T* A = nullptr;
    int* ipiv = nullptr;
    T* aux = nullptr;
    std::function<void()> restore;
    __bf16* sh_lo = nullptr;
    __bf16* sh_up = nullptr;
    __bf16* sh_upr = nullptr;  // optional: the part above the diagonal blocks again, row-major
    int sh_ld = 0;
    int path = 0;
    // the tiled copy of the factors for the chained solves is not made by getrf_end but left
    // pending (chain_tiles_issue), so that it can run beside later latency-bound work
    bool defer_tiles = false;
    // columns cols_from[i] .. n-1 of A are still being written by work that completes at
    // cols_event[i] (another stream; ncols ranges, cols_from increasing): the factorisation
    // waits for each before it first touches one of its columns
    hipEvent_t cols_event[2] = {nullptr, nullptr};
    int cols_from[2] = {0, 0};
    int ncols = 0;
    DBuf<int> flag;
    DBuf<T> backup, inv;
    bool started = false;
};
template <typename T>
void getrf_begin(GetrfJob<T>& job);
// the first split of the no-pivot factorisation of order n (its left part, columns
// 0 .. split-1, is factored before any later column is touched)
int getrf_top_split(int n);
template <typename T>
int getrf_end(GetrfJob<T>& job);
// A pending tiled copy (GetrfJob::defer_tiles) of the factors whose solve aid is aux: issue
// starts it on the side stream (no-op when none is pending; getrs issues it itself before a
// chained solve needs it); forget drops it, before the factors or aux are released.
void chain_tiles_issue(const void* aux);
void chain_tiles_forget(const void* aux);
inline size_t getrf_aux_elems(int n) {
    const size_t nb = (size_t)((n + 127) / 128);
    return (4 * nb + nb * nb) * 128 * 128;
}
// While one is alive, plain FP64/FP32 products use this library's kernels instead of
// rocBLAS: work that is also captured into graphs (where rocBLAS is never used) then runs
// the same kernels eagerly and in a graph.
struct LibraryGemmOff {
    LibraryGemmOff();
    ~LibraryGemmOff();
    LibraryGemmOff(const LibraryGemmOff&) = delete;
    LibraryGemmOff& operator=(const LibraryGemmOff&) = delete;
};

// ---- BF16 copies of final values ("shadows") for the 16-bit products (BF16 kind)
// While a scope is alive, a 16-bit product whose operands lie in a registered FP32 matrix
// reads them from its BF16 copy (no conversion pass), and the kernels that finalise values
// of the matrix (the L21 products and the triangular-solve leaves) write the copy as well.
// A factor registration covers the parts below (lo, row-major) and above (up,
// column-major) the 128 x 128 diagonal blocks; a dense one the whole matrix, column-major
// (up) or row-major (lo); a factor may also hold its upper part row-major (upr), which the
// U-side solves read as their products' A operand. The caller registers only copies whose
// values are, or will be made, final before any product reads them.
struct ShadowScope {
    ShadowScope(const float* base, int ld, int rows, int cols, __bf16* lo, __bf16* up, int lds, bool factor,
                __bf16* upr = nullptr);
    ~ShadowScope();
    ShadowScope(const ShadowScope&) = delete;
    ShadowScope& operator=(const ShadowScope&) = delete;
    int id;
};
// row-major BF16 copy of an m x k FP32 matrix (leading dimension shadow_rows_ld(k))
size_t shadow_rows_elems(int m, int k);
int shadow_rows_ld(int k);
void shadow_rows(int m, int k, const float* A, int lda, __bf16* out);

// B <- A^-1 B (trans = 0) or A^-T B (trans = 1)
// aux (optional): getrf's solve aid for these factors; formed here when absent
template <typename T>
void getrs(int kind, int trans, int n, int nrhs, const T* LU, int lda, const int* ipiv, T* B, int ldb,
           const T* aux = nullptr);

// ---- batched forms (one launch per step for all matrices of a batch)
// C_j = alpha_j A_j B_j + beta_j C_j (column-major, storage-precision MFMA)
template <typename T>
struct GemmJob {
    int m, n, k;
    T alpha, beta;
    const T* A;
    int lda;
    const T* B;
    int ldb;
    T* C;
    int ldc;
};
template <typename T>
void gemm_batch(const std::vector<GemmJob<T>>& jobs);
// LU without row interchanges of independent square matrices of any orders, in storage
// precision; aux (getrf_aux_elems(n)) receives the solve aid as getrf's; *flag is set
// when a multiplier exceeds 1 in magnitude or is not finite, i.e. when partial pivoting
// would have interchanged rows (the factors are then not the pivoted ones: refactor).
template <typename T>
struct BatchMat {
    T* A;
    int lda, n;
    T* aux;
    int* flag;
};
template <typename T>
void getrf_np_batch(const std::vector<BatchMat<T>>& mats);
// the same in two halves: prepare plans the launches and uploads their job tables (and
// synchronises); launch only issues kernels (the flags are cleared by a kernel), so it may
// be captured into a graph while the plan lives
struct BatchLuStep {
    int kind;
    size_t off;
    int count, gridx;
};
template <typename T>
struct BatchLuPlan {
    std::vector<BatchMat<T>> mats;
    std::vector<BatchLuStep> steps;
    DBuf<unsigned char> tab;
    double flops = 0;
};
template <typename T>
void getrf_np_batch_prepare(const std::vector<BatchMat<T>>& mats, BatchLuPlan<T>& plan);
template <typename T>
void getrf_np_batch_launch(const BatchLuPlan<T>& plan);
// B_j <- A_j^-1 B_j, one right-hand side each, from factors with getrf's solve aid (aux):
// rows first gathered through perm (device, n entries: B <- B[perm], the row interchanges
// of a pivoted factorisation as one permutation) unless perm is null
template <typename T>
struct BatchSolve {
    int n;
    const T* aux;
    T* B;
    const int* perm = nullptr;
};
template <typename T>
void getrs_np_batch(const std::vector<BatchSolve<T>>& s);
// the same split in two: prepare uploads the job tables (and synchronises); launch only
// issues kernels (no allocation, copy or synchronisation), so it may be captured in a
// graph and replayed while the plan lives
struct TrsvState;
template <typename T>
struct BatchSolvePlan {
    int njobs = 0, nbmax = 0, nperm = 0, permmax = 1;
    size_t half = 0;
    double flops = 0, bytes = 0;
    DBuf<unsigned char> jobs;
    DBuf<TrsvState> state;
    DBuf<T> tmp;
};
template <typename T>
void getrs_np_batch_prepare(const std::vector<BatchSolve<T>>& s, BatchSolvePlan<T>& p);
template <typename T>
void getrs_np_batch_launch(const BatchSolvePlan<T>& p);

}  // namespace sq
