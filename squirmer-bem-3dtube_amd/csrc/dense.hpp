// dense.hpp -- internal interface of the dense linear algebra (dense.hip).
#pragma once
#include <cstddef>

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
// 128 x 128 diagonal blocks of L and U, and a tiled copy of the factors).
template <typename T>
int getrf(int kind, int n, T* A, int lda, int* ipiv, T* aux = nullptr);
inline size_t getrf_aux_elems(int n) {
    const size_t nb = (size_t)((n + 127) / 128);
    return (2 * nb + nb * nb) * 128 * 128;
}
// B <- A^-1 B (trans = 0) or A^-T B (trans = 1)
// aux (optional): getrf's solve aid for these factors; formed here when absent
template <typename T>
void getrs(int kind, int trans, int n, int nrhs, const T* LU, int lda, const int* ipiv, T* B, int ldb,
           const T* aux = nullptr);

}  // namespace sq
