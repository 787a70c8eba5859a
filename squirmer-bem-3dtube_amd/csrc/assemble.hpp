// assemble.hpp -- internal interface of the panel-integral kernels (assemble.hip).
#pragma once
#include "../../include/utils/mesh.hpp"
#include <vector>

#include "common.hpp"

namespace sq {

constexpr int MODE_BLOCKS = 0;  // write SL / DL blocks and DL row sums
constexpr int MODE_FIELD = 1;   // fold pairs into a velocity (representation formula)

// Device copy of a mesh's panel geometry in structure-of-arrays form, plus the seven
// far-field quadrature points of every panel (x, y, z, weight * area).
struct DMesh {
    int np = 0;
    double total_area = 0;
    DBuf<double> V;        // 9 x np: vertex k coordinate c at row 3k + c
    DBuf<double> nrm;      // 3 x np
    DBuf<double> cen;      // 3 x np
    DBuf<double> area;     // np
    DBuf<double> h;        // np
    DBuf<double> q7;       // 28 x np
    DBuf<double> cen_xyz;  // 3 np, xyz interleaved (targets of collocation)
    DBuf<double> nrm_xyz;  // 3 np, xyz interleaved
    DMesh() = default;
    explicit DMesh(const bem::Mesh& m);
    void upload(const bem::Mesh& m);
};

struct AsmArgs {
    int nt;
    const double* x0;
    const int* self_src;
    int p0, np, src_np;
    const double *V, *nrm, *cen, *area, *h, *q7;
    double sl_div, dl_div;
    double* SL;
    long ld_sl;
    float* SL32;  // optional FP32 copy of SL, written with it
    long ld_sl32;
    double* DL;
    long ld_dl;
    double* dls_part;
    double defl_sc;
    const double* defl_tn;
    const double* fld_u;
    const double* fld_f;
    double* vel_part;
};

// SL(3i+a, 3p+b) = SL_ab/sl_div (+ defl_sc tn_i[a] n_p[b] A_p), DL(...) = -DL_ab/dl_div,
// dlsum[9i+k] += DL_k; SL32 (optional) receives SL rounded to FP32. Device pointers; any
// output may be null.
void assemble(const DMesh& src, int p0, int np, const double* x0, int nt, const int* self_src, double sl_div,
              double* SL, long ld_sl, double dl_div, double* DL, long ld_dl, double* dlsum, double defl_sc,
              const double* defl_tn, float* SL32 = nullptr, long ld_sl32 = 0);

// One job of a batched assembly: the arguments of assemble(), except that the stresslet
// row sums stay per source tile in dls_part (asm_tiles(np) x 9 x nt doubles, or null) for
// reduce_dls_batch.
struct AsmJob {
    const DMesh* src;
    int p0, np;
    const double* x0;
    int nt;
    const int* self_src;
    double sl_div;
    double* SL;
    long ld_sl;
    double dl_div;
    double* DL;
    long ld_dl;
    double* dls_part;
    double defl_sc;
    const double* defl_tn;
};
int asm_tiles(int np);
// all jobs in one persistent launch (work queue over every job's workgroup tiles)
void assemble_batch(const std::vector<AsmJob>& jobs);
// the same in two halves: the job tables uploaded once (prepare, which synchronises), then
// launches only (launch: the work-queue counter is cleared by a kernel, so the launch can
// be captured into a graph while the plan lives)
struct AsmBatch {
    DBuf<AsmArgs> jobs;
    DBuf<long> start;
    DBuf<unsigned long long> queue;
    int njobs = 0;
    long total = 0;
    double flops = 0, bytes = 0;
};
void assemble_batch_prepare(const std::vector<AsmJob>& jobs, AsmBatch& plan);
void assemble_batch_launch(const AsmBatch& plan);
// out[9i + k] = (sum over tiles of part1) + (sum over tiles of part2), tiles in order (the
// row sums of two consecutive assemble() calls into one zeroed dlsum); part2 may be null
struct DlsReduce {
    int nt;
    int ntiles1;
    const double* part1;
    int ntiles2;
    const double* part2;
    double* out;
};
void reduce_dls_batch(const std::vector<DlsReduce>& jobs);
struct DlsBatch {
    DBuf<DlsReduce> jobs;
    int n = 0, ntmax = 1;
};
void reduce_dls_prepare(const std::vector<DlsReduce>& jobs, DlsBatch& plan);
void reduce_dls_launch(const DlsBatch& plan);

// velocity at npts points (device, xyz interleaved) from panel velocities u and
// tractions f (device, 3 per panel): out (device, 3 npts)
void field(const DMesh& src, const double* pts, int npts, const double* u, const double* f, double mu, double* out);

// out[3i + k] += sigma sum_p sum_q w_q [a_sl lam_i . d_k K_SL F_p + a_dl lam_i . d_k K_DL V]
// over the source panels of src for the targets x0 (3 per target): the derivative of the
// assembled coupling blocks under a translation of the sources (sigma = +1) or of the
// targets (sigma = -1), contracted with lam (3 per target) and with the source tractions
// src_f and velocities (src_v per source, or tgt_v per target). Device pointers.
void adjoint_xc(const DMesh& src, const double* x0, int nt, const double* lam, const double* tgt_v,
                const double* src_f, const double* src_v, double a_sl, double a_dl, double sigma, double* out);

}  // namespace sq
