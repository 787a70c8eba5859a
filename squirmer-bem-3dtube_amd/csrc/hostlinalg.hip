// hostlinalg.hip -- host-matrix entry points of the dense algebra (the C++ Mat API of
// src/utils/linalg.hpp:31-78): operands are staged through HBM, the work runs on the
// device kernels of dense.hip, and LU factors stay resident in HBM.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "dense.hpp"
#include "lu_handle.hpp"
#include "solver.hpp"

using namespace sq;

namespace {
// copy a host column-major block (m x n, leading dim ld) into a packed device buffer
void put(DBuf<double>& d, int m, int n, const double* h, int ld) {
    d.alloc((size_t)std::max(m, 1) * std::max(n, 1));
    if (m <= 0 || n <= 0) return;
    SQ_HIP(hipMemcpy2DAsync(d.p, sizeof(double) * m, h, sizeof(double) * ld, sizeof(double) * m, n,
                            hipMemcpyHostToDevice, ctx().s()));
}
void get(const DBuf<double>& d, int m, int n, double* h, int ld) {
    if (m <= 0 || n <= 0) return;
    SQ_HIP(hipMemcpy2DAsync(h, sizeof(double) * ld, d.p, sizeof(double) * m, sizeof(double) * m, n,
                            hipMemcpyDeviceToHost, ctx().s()));
    sync();
}
}  // namespace

namespace {
// 2-norms of the nrhs columns of a device n x nrhs matrix (leading dimension ld)
std::vector<double> col_norms(int n, int nrhs, const double* a, int ld) {
    std::vector<double> h((size_t)n * nrhs), out(nrhs, 0.0);
    SQ_HIP(hipMemcpy2DAsync(h.data(), sizeof(double) * n, a, sizeof(double) * ld, sizeof(double) * n, nrhs,
                            hipMemcpyDeviceToHost, ctx().s()));
    sync();
    for (int j = 0; j < nrhs; ++j) {
        double s2 = 0;
        for (int i = 0; i < n; ++i) s2 += h[(size_t)j * n + i] * h[(size_t)j * n + i];
        out[j] = std::sqrt(s2);
    }
    return out;
}
}  // namespace

void sqbem_lu::solve_device(int nrhs, double* B, int ldb) const {
    if (owner) tube_ready(owner);
    if (LU32) {
        DBuf<float> d32((size_t)n * nrhs);
        if (!A64) {  // no FP64 operator at hand: the FP32 factors' solution
            convert_matrix<double, float>(n, nrhs, B, ldb, d32.p, n);
            getrs<float>(kind, 0, n, nrhs, LU32, n, piv, d32.p, n, aux32);
            convert_matrix<float, double>(n, nrhs, d32.p, n, B, ldb);
            sync();
            return;
        }
        // iterative refinement: x += LU32^-1 r, r = b - A64 x (FP64), to SQBEM_IR_TOL
        DBuf<double> x((size_t)n * nrhs), r((size_t)n * nrhs), d((size_t)n * nrhs);
        x.zero();
        convert_matrix<double, double>(n, nrhs, B, ldb, r.p, n);
        const std::vector<double> bn = col_norms(n, nrhs, r.p, n);
        double worst = 1.0;
        int it = 0;
        for (; it < 40 && worst > SQBEM_IR_TOL; ++it) {
            convert_matrix<double, float>(n, nrhs, r.p, n, d32.p, n);
            getrs<float>(kind, 0, n, nrhs, LU32, n, piv, d32.p, n, aux32);
            convert_matrix<float, double>(n, nrhs, d32.p, n, d.p, n);
            axpy((size_t)n * nrhs, 1.0, d.p, x.p);
            convert_matrix<double, double>(n, nrhs, B, ldb, r.p, n);
            dgemm(n, nrhs, n, -1.0, A64, n, x.p, n, 1.0, r.p, n);
            const std::vector<double> rn = col_norms(n, nrhs, r.p, n);
            worst = 0.0;
            for (int j = 0; j < nrhs; ++j) worst = std::max(worst, bn[j] > 0 ? rn[j] / bn[j] : rn[j]);
        }
        SQ_REQUIRE(worst <= SQBEM_IR_TOL, SQBEM_ERR_CONVERGE,
                   "iterative refinement did not reach relative residual 1e-12 (" + std::to_string(worst) +
                       " after " + std::to_string(it) + " steps)");
        convert_matrix<double, double>(n, nrhs, x.p, n, B, ldb);
        sync();
    } else {
        getrs<double>(SQBEM_FP64, 0, n, nrhs, LU, n, piv, B, ldb, aux);
    }
}

extern "C" {

int sqbem_host_dgemm(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
                     double beta, double* C, int ldc) {
    return guarded([&] {
        SQ_REQUIRE(m >= 0 && n >= 0 && k >= 0 && lda >= std::max(1, m) && ldb >= std::max(1, k) &&
                       ldc >= std::max(1, m),
                   SQBEM_ERR_ARG, "gemm: dimension mismatch");
        if (m == 0 || n == 0 || k == 0) return;
        DBuf<double> a, b, c;
        put(a, m, k, A, lda);
        put(b, k, n, B, ldb);
        if (beta != 0.0)
            put(c, m, n, C, ldc);
        else
            c.alloc((size_t)m * n);
        dgemm(m, n, k, alpha, a.p, m, b.p, k, beta, c.p, m);
        get(c, m, n, C, ldc);
    });
}

int sqbem_host_dgesv(int n, int nrhs, const double* A, int lda, double* B, int ldb) {
    return guarded([&] {
        SQ_REQUIRE(n >= 0 && nrhs >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), SQBEM_ERR_ARG,
                   "dgesv: bad dimensions");
        if (n == 0 || nrhs == 0) return;
        DBuf<double> a, b;
        DBuf<int> piv(n);
        put(a, n, n, A, lda);
        put(b, n, nrhs, B, ldb);
        int info = getrf<double>(SQBEM_FP64, n, a.p, n, piv.p);
        SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgesv failed");
        getrs<double>(SQBEM_FP64, 0, n, nrhs, a.p, n, piv.p, b.p, n);
        get(b, n, nrhs, B, ldb);
    });
}

sqbem_lu* sqbem_lu_create(int n, const double* A, int lda, int* info) {
    return guarded_ptr<sqbem_lu>([&] {
        SQ_REQUIRE(n > 0 && A && lda >= n, SQBEM_ERR_ARG, "dgetrf: bad dimensions");
        auto* lu = new sqbem_lu;
        try {
            lu->n = n;
            put(lu->own, n, n, A, lda);
            lu->own_piv.alloc(n);
            lu->LU = lu->own.p;
            lu->piv = lu->own_piv.p;
            int i = getrf<double>(SQBEM_FP64, n, lu->own.p, n, lu->own_piv.p);
            if (info) *info = i;
            SQ_REQUIRE(i == 0, SQBEM_ERR_SINGULAR, "dgetrf failed");
        } catch (...) {
            delete lu;
            throw;
        }
        return lu;
    });
}

int sqbem_lu_solve(const sqbem_lu* lu, int nrhs, double* B, int ldb) {
    return guarded([&] {
        SQ_REQUIRE(lu && (nrhs == 0 || B) && ldb >= lu->n, SQBEM_ERR_ARG, "dgetrs: bad arguments");
        if (nrhs == 0) return;
        DBuf<double> b;
        put(b, lu->n, nrhs, B, ldb);
        lu->solve_device(nrhs, b.p, lu->n);
        get(b, lu->n, nrhs, B, ldb);
    });
}

int sqbem_lu_get(const sqbem_lu* lu, double* LUh, int* piv) {
    return guarded([&] {
        SQ_REQUIRE(lu, SQBEM_ERR_ARG, "null factor");
        if (lu->owner) tube_ready(lu->owner);
        const size_t nn = (size_t)lu->n * lu->n;
        if (LUh) {
            if (lu->LU32) {
                std::vector<float> tmp(nn);
                SQ_HIP(hipMemcpyAsync(tmp.data(), lu->LU32, nn * sizeof(float), hipMemcpyDeviceToHost, ctx().s()));
                sync();
                for (size_t e = 0; e < nn; ++e) LUh[e] = tmp[e];
            } else {
                SQ_HIP(hipMemcpyAsync(LUh, lu->LU, nn * sizeof(double), hipMemcpyDeviceToHost, ctx().s()));
            }
        }
        if (piv) SQ_HIP(hipMemcpyAsync(piv, lu->piv, lu->n * sizeof(int), hipMemcpyDeviceToHost, ctx().s()));
        sync();
    });
}

int sqbem_lu_size(const sqbem_lu* lu) { return lu ? lu->n : 0; }

void sqbem_lu_free(sqbem_lu* lu) { delete lu; }

}  // extern "C"
