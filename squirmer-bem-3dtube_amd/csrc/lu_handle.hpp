// lu_handle.hpp -- the opaque sqbem_lu of the C ABI: LU factors resident in HBM, either
// owned (sqbem_lu_create) or a view onto a tube's factors.
#pragma once
#include "common.hpp"

struct sqbem_tube;
namespace sq {
// finishes the tube's factorisation if it is still running (solver.hip)
void tube_ready(const sqbem_tube* t);
}  // namespace sq

struct sqbem_lu {
    const sqbem_tube* owner = nullptr;  // a view: the tube whose factors these are
    int n = 0;
    int kind = SQBEM_FP64;       // trailing-update kind of an FP32 factor
    const double* LU = nullptr;  // FP64 factors, or
    const float* LU32 = nullptr; // FP32 factors (refined modes)
    const int* piv = nullptr;    // 1-based pivots
    const double* aux = nullptr;   // getrf's solve aid for LU (or null)
    const float* aux32 = nullptr;  // same for LU32
    const double* A64 = nullptr;   // FP64 operator of FP32 factors (refinement residuals), or null
    sq::DBuf<double> own;
    sq::DBuf<int> own_piv;
    // B (device, n x nrhs) <- A^-1 B; FP32 factors with A64 set are refined to SQBEM_IR_TOL
    void solve_device(int nrhs, double* B, int ldb) const;
};
