// solver.hpp -- internals shared by the single-instance solvers (solver.hip) and the
// batched solver (batch.hip): the tube object and the small device helpers of the Schur
// complement and of the bordered force/torque-free system.
#pragma once
#include <memory>
#include "../../include/utils/mesh.hpp"
#include "assemble.hpp"
#include "common.hpp"
#include "dense.hpp"
#include "lu_handle.hpp"

// Tube::Tube + Tube::factor (src/swimmer.hpp:40-60): FP64 operator and its factors in HBM.
struct sqbem_tube {
    bem::Mesh mesh;
    sq::DMesh dm;
    double mu = 1;
    int prec = SQBEM_FP64;
    int Nt = 0, nt = 0;
    sq::DBuf<double> T64;  // FP64 mode: LU factors; mixed: the FP64 operator
    sq::DBuf<float> T32;   // mixed: FP32 LU factors
    sq::DBuf<int> ipiv;
    sq::DBuf<int> self;      // panel index of each collocation point (the assembly's self pairs)
    // large refined tubes: the columns beyond the factorisation's first split are assembled on
    // the side stream (fork_ev: its start, asm_ev: its end), beside the left half's factorisation
    hipEvent_t fork_ev = nullptr, asm_ev[2] = {nullptr, nullptr};
    int asm_from[2] = {0, 0};  // first column of each side-stream part
    sq::DBuf<double> aux64;  // getrf's solve aid for the factors (inverted blocks, tiled copy)
    sq::DBuf<float> aux32;
    // BF16 mode: BF16 copies of the factors below / above the diagonal blocks (row-major /
    // column-major, leading dimension sh_ld), whole when np_path = 1 (no-pivot factors)
    sq::DBuf<__bf16> SL, SU, SUr;  // SUr: the part above the diagonal blocks again, row-major
    int sh_ld = 0, np_path = 0;
    sqbem_lu view;  // non-owning handle onto the factors (Tube::lu)
    double t_assembly = 0, t_factor = 0;
    // The factorisation is issued by tube_build and finished by tube_ready, when the factors
    // are first needed: in between the device factors while the host goes on (a swimmer's
    // coupling blocks are assembled on the side stream meanwhile).
    mutable std::unique_ptr<sq::GetrfJob<float>> job32;
    mutable std::unique_ptr<sq::GetrfJob<double>> job64;
    mutable double t_factor0 = 0;
    ~sqbem_tube() {
        // a factorisation still running writes into the buffers released below
        if (job32 || job64) (void)hipStreamSynchronize(sq::ctx().s());
        sq::chain_tiles_forget(aux32.p);  // a tiled copy never issued
        sq::chain_tiles_forget(aux64.p);
        sq::side_drain();  // a tiled copy of the factors (or an assembly) may still run there
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        for (hipEvent_t e : asm_ev)
            if (e) (void)hipEventDestroy(e);
    }
};

namespace sq {

double now_s();
inline bool mixed(int prec) { return prec == SQBEM_FP32_IR || prec == SQBEM_FP16_IR || prec == SQBEM_BF16_IR; }

// assembles, deflates and factors t->mesh (t->mu, t->prec set by the caller)
void tube_build(sqbem_tube* t);
// waits for the factorisation tube_build issued, finishes it (the pivoting fallback if
// needed) and completes t->view and t->t_factor; throws if the matrix is singular. Cheap
// once done.
void tube_ready(const sqbem_tube* t);

// Y = alpha A X + beta Y for at most 8 right-hand sides (FP64 sums)
template <typename T>
void gemv(int m, int n, int nrhs, double alpha, const T* A, long lda, const T* X, long ldx, double beta, T* Y,
          long ldy);

// x += alpha y (n doubles, device)
void axpy(size_t n, double alpha, const double* y, double* x);

// AuBs(3i+a, 3i+b) += dls_i[3a+b] / (8 pi) - minus_identity delta_ab (completed double layer)
void completion_diag(int Ns, double* A, long lda, const double* dls, double minus_identity);

// C = A K for the rigid-body map K about xc (host), columns of C at ldc; part: workspace of
// kproj_part_elems(nrows) doubles. No synchronisation.
size_t kproj_part_elems(int nrows);
void kproj_async(int nrows, int Ns, const double* A, long lda, const double* cen_xyz, const double* xc, double* C,
                 long ldc, double* part);

// *out = max(*out, max |A(i,j)|) as the bit pattern of a non-negative double (no download)
template <typename T>
void absmax_async(int m, int n, const T* A, long lda, unsigned long long* out);

// bordered matrix M ((ns + 6)^2) of solve_free from G = [Gs | Gc] (ns x (ns + 6)) and the
// force/torque rows scaled by sc (src/swimmer.hpp:107-125)
template <typename T>
void border_build(int Ns, const T* G, T* M, const double* cen_xyz, const double* area, double sc, const double* xc,
                  const double* sc_dev = nullptr);  // sc_dev (device, one double): the scale, read there

// out(0..5, k) = -sc F f_s(:, k): the constraint rows applied to sphere tractions
void constraint_apply(int Ns, int nrhs, const double* fs, long ldf, const double* cen_xyz, const double* area,
                      double sc, const double* xc, double* out, long ldo);

}  // namespace sq
