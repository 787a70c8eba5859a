// linalg.hpp -- column-major dense matrix and the dense solves, executed by libsqbem on the
// GPU (MFMA GEMM, partial-pivoting LU with factors kept in HBM).
//
// API of src/utils/linalg.hpp:31-78: Mat, gemm, solve_dense, LUFactor. Error texts are the
// reference's ("gemm: dimension mismatch", "dgesv failed", "dgetrf failed").
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sqbem.h"

namespace bem {

inline void sqbem_check(int rc) {
    if (rc != SQBEM_OK) throw std::runtime_error(sqbem_last_error());
}

// Column-major dense matrix: element (i, j) at a[i + j*n]
struct Mat {
    int n = 0, m = 0;
    std::vector<double> a;
    Mat() = default;
    Mat(int n_, int m_) : n(n_), m(m_), a((size_t)n_ * m_, 0.0) {}
    double& operator()(int i, int j) { return a[i + (size_t)j * n]; }
    double operator()(int i, int j) const { return a[i + (size_t)j * n]; }
};

// C = alpha A B + beta C
inline void gemm(double alpha, const Mat& A, const Mat& B, double beta, Mat& C) {
    if (A.m != B.n || C.n != A.n || C.m != B.m) throw std::runtime_error("gemm: dimension mismatch");
    if (A.n == 0 || B.m == 0 || A.m == 0) return;
    sqbem_check(sqbem_host_dgemm(A.n, B.m, A.m, alpha, A.a.data(), A.n, B.a.data(), B.n, beta, C.a.data(), C.n));
}

// B <- A^{-1} B (A is copied, so it is left unchanged)
inline void solve_dense(Mat A, Mat& B) {
    if (A.n == 0 || B.m == 0) return;
    sqbem_check(sqbem_host_dgesv(A.n, B.m, A.a.data(), A.n, B.a.data(), B.n));
}

// LU factorisation kept (in HBM) for repeated solves with different right-hand sides.
// factor() also copies the packed factors and 1-based pivots back into LU and piv, as the
// reference's dgetrf_ leaves them (src/utils/linalg.hpp:62-71). A view onto factors owned
// elsewhere (a Tube's) leaves LU an empty 0 x 0 Mat and piv empty until download().
struct LUFactor {
    Mat LU;
    std::vector<int> piv;

    void factor(Mat A) {
        int info = 0;
        sqbem_lu* h = sqbem_lu_create(A.n, A.a.data(), A.n, &info);
        if (!h) throw std::runtime_error(info ? "dgetrf failed" : sqbem_last_error());
        own_.reset(h, sqbem_lu_free);
        view_ = h;
        download();
    }
    void solve(Mat& B) const {  // B <- A^{-1} B
        if (B.m == 0) return;
        if (!view_) throw std::runtime_error("LUFactor::solve before factor");
        sqbem_check(sqbem_lu_solve(view_, B.m, B.a.data(), B.n));
    }
    void download() {
        if (!view_) return;
        const int n = sqbem_lu_size(view_);
        LU = Mat(n, n);
        piv.assign(n, 0);
        sqbem_check(sqbem_lu_get(view_, LU.a.data(), piv.data()));
    }
    // non-owning view onto factors owned elsewhere (a Tube)
    void attach(const sqbem_lu* v) {
        own_.reset();
        view_ = v;
        LU = Mat();
        piv.clear();
    }
    const sqbem_lu* handle() const { return view_; }

private:
    std::shared_ptr<sqbem_lu> own_;
    const sqbem_lu* view_ = nullptr;
};

}  // namespace bem
