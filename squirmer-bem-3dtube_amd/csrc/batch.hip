// batch.hip -- many independent squirmer-in-tube solves of uneven size in one call.
//
// The reference solves one configuration at a time: its kinematics sweeps build a
// SwimmerSolver per configuration and reuse the Tube across offsets (src/main.cpp:82-100,
// src/swimmer.hpp:83-139). Here a whole batch is one device pipeline:
//
//  * tube meshes with identical contents share one assembly and factorisation (a tube
//    whose mesh does not depend on the swimmer, e.g. uniform axial nodes, is factored once
//    for every configuration in it);
//  * instances on the same tube are processed in chunks sized to a memory budget; every
//    coupling block of a chunk (tube rows <- sphere, sphere rows <- tube, sphere <- sphere)
//    is integrated by one persistent work-queue launch (assemble_batch);
//  * the Schur right-hand sides [B_i | Ct_i] of all instances of a chunk sit side by side
//    in one nt x sum(ns_i + 6) matrix, so T^-1 [B | Ct] is one triangular solve with many
//    columns through the shared factors; D_i Y_i is one GEMM per instance;
//  * the bordered force/torque-free matrices are factored per instance;
//  * FP64 mode solves the reduced systems directly (src/swimmer.hpp:107-139, 256-288); the
//    refined modes iterate on every instance's full bordered system together, with the
//    shared-tube products T X of all instances as one GEMM, and one host round trip per
//    refinement step for all convergence checks.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "../../include/utils/kernels.hpp"
#include "assemble.hpp"
#include "common.hpp"
#include "dense.hpp"
#include "mesh_handle.hpp"
#include "solver.hpp"

namespace sq {

namespace {

struct Inst {
    int idx = 0;   // position in the caller's arrays
    int tube = 0;  // unique-tube index
    bem::Mesh S;
    double xc[3] = {0, 0, 0};
    double e[3] = {0, 0, 1};
    double B1 = 1, B2 = 0;
    int Ns = 0, ns = 0, m = 0;  // m = ns + 6 unknowns of the bordered system
};

// FNV-1a over the arrays that define a mesh (nodes, connectivity, tags)
uint64_t mesh_hash(const bem::Mesh& M) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    mix(M.nodes.data(), M.nodes.size() * sizeof(bem::V3));
    mix(M.tris.data(), M.tris.size() * sizeof(M.tris[0]));
    mix(M.tags.data(), M.tags.size() * sizeof(int));
    return h;
}

bool same_mesh(const bem::Mesh& a, const bem::Mesh& b) {
    return a.nodes.size() == b.nodes.size() && a.tris.size() == b.tris.size() &&
           std::memcmp(a.nodes.data(), b.nodes.data(), a.nodes.size() * sizeof(bem::V3)) == 0 &&
           std::memcmp(a.tris.data(), b.tris.data(), a.tris.size() * sizeof(a.tris[0])) == 0 && a.tags == b.tags;
}

// device bytes one instance needs inside a chunk on a tube of nt unknowns
double inst_bytes(int nt, int ns, int prec) {
    const double m = ns + 6.0, nb = (m + 127) / 128;
    const double aux = (2 * nb + nb * nb) * 128.0 * 128.0;
    double b = 8.0 * nt * (m + 2.0 * ns) + 8.0 * (ns * m + (double)ns * ns + m * m + aux) + 8.0 * 9 * ns / 3 * 2.0 +
               8.0 * 4 * (nt + m);
    if (mixed(prec)) b += 4.0 * nt * (m + ns) + 4.0 * (ns * m + m * m + aux) + 4.0 * 2 * (nt + m);
    // stresslet row-sum partials of the two jobs with sphere targets
    b += 8.0 * 9.0 * (ns / 3.0) * (asm_tiles(nt / 3) + asm_tiles(ns / 3));
    return b;
}

// ||r_i|| over the tube block Rt(:, i) and the instance's block of Rsc
__global__ void batch_norms_kernel(int nt, const double* __restrict__ Rt, const double* __restrict__ Rsc,
                                   const int* __restrict__ coff, const int* __restrict__ mlen, double* out) {
    const int i = blockIdx.x;
    double s = 0.0;
    for (int r = threadIdx.x; r < nt; r += blockDim.x) {
        const double v = Rt[(size_t)i * nt + r];
        s += v * v;
    }
    const double* q = Rsc + coff[i];
    for (int r = threadIdx.x; r < mlen[i]; r += blockDim.x) s += q[r] * q[r];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = sqrt(red[0]);
}

__global__ void batch_iota_kernel(int n, int* a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ void batch_axpy_kernel(size_t n, const double* __restrict__ y, double* x) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        x[e] += y[e];
}
void add_into(size_t n, const double* y, double* x) {
    if (!n) return;
    const unsigned blocks = (unsigned)std::min<size_t>(ceil_div(n, 256), 8192);
    hipLaunchKernelGGL(batch_axpy_kernel, dim3(blocks), dim3(256), 0, ctx().s(), n, y, x);
    check_launch("batch_axpy");
}

struct Stats {
    double t_tubes = 0, t_assembly = 0, t_schur = 0, t_solve = 0;
    int unique_tubes = 0, chunks = 0, max_iters = 0;
    double worst = 0;
};

// One chunk: instances I (all on tube tb), built, reduced and solved together.
void run_chunk(const sqbem_tube* tb, std::vector<Inst*>& I, int prec, double mu, double* U, double* Om, int* iters,
               double* resid, Stats& st) {
    const int k = (int)I.size();
    const int Nt = tb->Nt, nt = tb->nt;
    const double e8 = bem::EIGHT_PI;
    const bool mx = mixed(prec);
    std::vector<int> coff(k + 1, 0), soff(k + 1, 0), Noff(k + 1, 0);
    int maxNs = 1;
    for (int i = 0; i < k; ++i) {
        coff[i + 1] = coff[i] + I[i]->m;
        soff[i + 1] = soff[i] + I[i]->ns;
        Noff[i + 1] = Noff[i] + I[i]->Ns;
        maxNs = std::max(maxNs, I[i]->Ns);
    }
    const size_t C = coff[k], S = soff[k];
    // per-instance offsets into the concatenated buffers
    std::vector<size_t> oD(k), oG(k), oA(k), oP2(k), oP3(k);
    size_t nD = 0, nG = 0, nA = 0, nP = 0;
    for (int i = 0; i < k; ++i) {
        const size_t ns = I[i]->ns, m = I[i]->m, Ns = I[i]->Ns;
        oD[i] = nD; nD += ns * nt;
        oG[i] = nG; nG += ns * m;
        oA[i] = nA; nA += ns * ns;
        oP2[i] = nP; nP += (size_t)asm_tiles(Nt) * 9 * Ns;
        oP3[i] = nP; nP += (size_t)asm_tiles((int)Ns) * 9 * Ns;
    }

    // ---- coupling assembly (src/swimmer.hpp:188-234), one persistent launch
    double t0 = now_s();
    std::vector<DMesh> dm(k);
    for (int i = 0; i < k; ++i) dm[i].upload(I[i]->S);
    DBuf<int> self(maxNs);
    hipLaunchKernelGGL(batch_iota_kernel, dim3(ceil_div(maxNs, 256)), dim3(256), 0, ctx().s(), maxNs, self.p);
    check_launch("iota");
    DBuf<double> BC((size_t)nt * C), AuBt((size_t)nt * S), D(nD), GE(nG), AuBs(nA), part(nP), dls(9 * (size_t)Noff[k]);
    std::vector<AsmJob> jobs;
    std::vector<DlsReduce> red;
    for (int i = 0; i < k; ++i) {
        const Inst& q = *I[i];
        double Asum = 0;
        for (int p = 0; p < q.Ns; ++p) Asum += q.S.area[p];
        const double defl = 1.0 / (e8 * mu * std::sqrt(Asum));
        // tube rows <- sphere sources: B = SL/(8 pi mu), AuBt = -DL/(8 pi)
        jobs.push_back({&dm[i], 0, q.Ns, tb->dm.cen_xyz.p, Nt, nullptr, e8 * mu, BC.p + (size_t)coff[i] * nt, nt, e8,
                        AuBt.p + (size_t)soff[i] * nt, nt, nullptr, 0.0, nullptr});
        // sphere rows <- tube sources: D = SL/(8 pi mu), stresslet row sums
        jobs.push_back({&tb->dm, 0, Nt, dm[i].cen_xyz.p, q.Ns, nullptr, e8 * mu, D.p + oD[i], q.ns, 1.0, nullptr, 0,
                        part.p + oP2[i], 0.0, nullptr});
        // sphere rows <- sphere sources: E (deflated), AuBs, stresslet row sums
        jobs.push_back({&dm[i], 0, q.Ns, dm[i].cen_xyz.p, q.Ns, self.p, e8 * mu, GE.p + oG[i], q.ns, e8,
                        AuBs.p + oA[i], q.ns, part.p + oP3[i], defl, dm[i].nrm_xyz.p});
        red.push_back({q.Ns, asm_tiles(Nt), part.p + oP2[i], asm_tiles(q.Ns), part.p + oP3[i], dls.p + 9 * (size_t)Noff[i]});
    }
    assemble_batch(jobs);
    reduce_dls_batch(red);
    {
        DBuf<double> kp(kproj_part_elems(std::max(nt, 3 * maxNs)));
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            completion_diag(q.Ns, AuBs.p + oA[i], q.ns, dls.p + 9 * (size_t)Noff[i], 0.0);
            // Ct = AuBt K beside B_i, Cs = AuBs K beside E_i (src/swimmer.hpp:223-234)
            kproj_async(nt, q.Ns, AuBt.p + (size_t)soff[i] * nt, nt, dm[i].cen_xyz.p, q.xc,
                        BC.p + ((size_t)coff[i] + q.ns) * nt, nt, kp.p);
            kproj_async(q.ns, q.Ns, AuBs.p + oA[i], q.ns, dm[i].cen_xyz.p, q.xc, GE.p + oG[i] + (size_t)q.ns * q.ns, q.ns,
                        kp.p);
        }
        sync();
    }
    // right-hand sides of the full bordered system for the slip of each instance:
    // b_t = -AuBt u, b_s = -AuBs u, b_c = 0 (src/swimmer.hpp:256-269)
    DBuf<double> u(S), Bt((size_t)nt * k), Bsc(C);
    {
        std::vector<double> hu(S);
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            const bem::V3 c{q.xc[0], q.xc[1], q.xc[2]}, e{q.e[0], q.e[1], q.e[2]};
            for (int p = 0; p < q.Ns; ++p) {
                const bem::V3 v = bem::squirmer_slip(q.S.centroid[p], c, e, q.B1, q.B2);
                hu[soff[i] + 3 * (size_t)p] = v.x;
                hu[soff[i] + 3 * (size_t)p + 1] = v.y;
                hu[soff[i] + 3 * (size_t)p + 2] = v.z;
            }
        }
        u.upload(hu.data(), S);
        Bsc.zero();
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            gemv<double>(nt, q.ns, 1, -1.0, AuBt.p + (size_t)soff[i] * nt, nt, u.p + soff[i], q.ns, 0.0,
                         Bt.p + (size_t)i * nt, nt);
            gemv<double>(q.ns, q.ns, 1, -1.0, AuBs.p + oA[i], q.ns, u.p + soff[i], q.ns, 0.0, Bsc.p + coff[i], q.m);
        }
    }
    AuBt.release();
    AuBs.release();
    part.release();
    sync();
    double t1 = now_s();
    st.t_assembly += t1 - t0;

    // ---- Schur complement (src/swimmer.hpp:235-249) and bordered factors
    std::vector<double> Amax(k, 0.0);
    for (int i = 0; i < k; ++i)
        for (int p = 0; p < I[i]->Ns; ++p) Amax[i] = std::max(Amax[i], 5.0 * I[i]->S.area[p]);
    DBuf<unsigned long long> amax(k);
    amax.zero();
    std::vector<unsigned long long> abits(k);
    std::vector<double> sc(k);
    std::vector<DBuf<int>> Mpiv(k);
    // FP64 mode
    std::vector<DBuf<double>> M64(k);
    // refined modes
    DBuf<float> Y32, D32, G32;
    std::vector<DBuf<float>> M32(k), Maux(k);
    if (!mx) {
        getrs<double>(SQBEM_FP64, 0, nt, (int)C, tb->T64.p, nt, tb->ipiv.p, BC.p, nt, tb->aux64.p);  // Y in place
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            dgemm(q.ns, q.m, nt, -1.0, D.p + oD[i], q.ns, BC.p + (size_t)coff[i] * nt, nt, 1.0, GE.p + oG[i], q.ns);
            absmax_async<double>(q.ns, q.ns, GE.p + oG[i], q.ns, amax.p + i);
        }
        amax.download(abits.data(), k);
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            double a;
            std::memcpy(&a, &abits[i], sizeof a);
            sc[i] = a / Amax[i];
            M64[i].alloc((size_t)q.m * q.m);
            border_build<double>(q.Ns, GE.p + oG[i], M64[i].p, dm[i].cen_xyz.p, dm[i].area.p, sc[i], q.xc);
            Mpiv[i].alloc(q.m);
            const int info = getrf<double>(SQBEM_FP64, q.m, M64[i].p, q.m, Mpiv[i].p);
            SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgesv failed");
        }
    } else {
        Y32.alloc((size_t)nt * C);
        convert_matrix<double, float>(nt, (int)C, BC.p, nt, Y32.p, nt);
        getrs<float>(prec, 0, nt, (int)C, tb->T32.p, nt, tb->ipiv.p, Y32.p, nt, tb->aux32.p);
        D32.alloc(nD);
        G32.alloc(nG);
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            convert_matrix<double, float>(q.ns, nt, D.p + oD[i], q.ns, D32.p + oD[i], q.ns);
            convert_matrix<double, float>(q.ns, q.m, GE.p + oG[i], q.ns, G32.p + oG[i], q.ns);
            const float* Yi = Y32.p + (size_t)coff[i] * nt;
            if (prec == SQBEM_FP32_IR)
                sgemm(q.ns, q.m, nt, -1.0f, D32.p + oD[i], q.ns, Yi, nt, 1.0f, G32.p + oG[i], q.ns);
            else
                gemm16(prec, q.ns, q.m, nt, -1.0f, D32.p + oD[i], q.ns, Yi, nt, 1.0f, G32.p + oG[i], q.ns);
            absmax_async<float>(q.ns, q.ns, G32.p + oG[i], q.ns, amax.p + i);
        }
        amax.download(abits.data(), k);
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            double a;
            std::memcpy(&a, &abits[i], sizeof a);
            sc[i] = a / Amax[i];
            M32[i].alloc((size_t)q.m * q.m);
            border_build<float>(q.Ns, G32.p + oG[i], M32[i].p, dm[i].cen_xyz.p, dm[i].area.p, sc[i], q.xc);
            Mpiv[i].alloc(q.m);
            Maux[i].alloc(getrf_aux_elems(q.m));
            const int info = getrf<float>(prec, q.m, M32[i].p, q.m, Mpiv[i].p, Maux[i].p);
            SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgesv failed");
        }
        G32.release();
    }
    sync();
    double t2 = now_s();
    st.t_schur += t2 - t1;

    // ---- solves
    std::vector<double> X(C);
    if (!mx) {
        // z = T^-1 b_t for all instances at once; x_i = M_i^-1 [b_s - D_i z_i; 0]
        DBuf<double> Z((size_t)nt * k), Xd(C);
        convert_matrix<double, double>(nt, k, Bt.p, nt, Z.p, nt);
        getrs<double>(SQBEM_FP64, 0, nt, k, tb->T64.p, nt, tb->ipiv.p, Z.p, nt, tb->aux64.p);
        convert_matrix<double, double>((int)C, 1, Bsc.p, (int)C, Xd.p, (int)C);
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            gemv<double>(q.ns, nt, 1, -1.0, D.p + oD[i], q.ns, Z.p + (size_t)i * nt, nt, 1.0, Xd.p + coff[i], q.m);
            getrs<double>(SQBEM_FP64, 0, q.m, 1, M64[i].p, q.m, Mpiv[i].p, Xd.p + coff[i], q.m);
        }
        Xd.download(X.data(), C);
        for (int i = 0; i < k; ++i) {
            iters[I[i]->idx] = 0;
            resid[I[i]->idx] = 0.0;
        }
    } else {
        // iterative refinement of every instance's full bordered system
        //   [T B Ct; D E Cs; 0 scF 0] [f_t; f_s; w] = [b_t; b_s; 0]
        // residuals from the FP64 operator, corrections from the low-precision factors
        DBuf<double> Xt((size_t)nt * k), Xsc(C), Rt((size_t)nt * k), Rsc(C), Dt((size_t)nt * k), Dsc(C), nrm(k);
        DBuf<float> Z32((size_t)nt * k), V32(C);
        DBuf<int> dcoff(k), dmlen(k);
        {
            std::vector<int> ml(k);
            for (int i = 0; i < k; ++i) ml[i] = I[i]->m;
            dcoff.upload(coff.data(), k);
            dmlen.upload(ml.data(), k);
        }
        Xt.zero();
        Xsc.zero();
        convert_matrix<double, double>(nt, k, Bt.p, nt, Rt.p, nt);
        convert_matrix<double, double>((int)C, 1, Bsc.p, (int)C, Rsc.p, (int)C);
        std::vector<double> bn(k), rn(k), rel(k, 1.0);
        std::vector<int> itc(k, 0);
        hipLaunchKernelGGL(batch_norms_kernel, dim3(k), dim3(256), 0, ctx().s(), nt, Bt.p, Bsc.p, dcoff.p, dmlen.p,
                           nrm.p);
        check_launch("batch_norms");
        nrm.download(bn.data(), k);
        const double tol = 1e-12;
        int it = 0;
        double worst = 1.0;
        for (; it < 40; ++it) {
            // corrections: Z = T^-1 r_t; v_i = M_i^-1 [r_s - D_i z_i; r_c]; z_i -= Y_i v_i
            convert_matrix<double, float>(nt, k, Rt.p, nt, Z32.p, nt);
            getrs<float>(prec, 0, nt, k, tb->T32.p, nt, tb->ipiv.p, Z32.p, nt, tb->aux32.p);
            convert_matrix<double, float>((int)C, 1, Rsc.p, (int)C, V32.p, (int)C);
            for (int i = 0; i < k; ++i) {
                const Inst& q = *I[i];
                float* zi = Z32.p + (size_t)i * nt;
                float* vi = V32.p + coff[i];
                gemv<float>(q.ns, nt, 1, -1.0, D32.p + oD[i], q.ns, zi, nt, 1.0, vi, q.m);
                getrs<float>(prec, 0, q.m, 1, M32[i].p, q.m, Mpiv[i].p, vi, q.m, Maux[i].p);
                gemv<float>(nt, q.m, 1, -1.0, Y32.p + (size_t)coff[i] * nt, nt, vi, q.m, 1.0, zi, nt);
            }
            convert_matrix<float, double>(nt, k, Z32.p, nt, Dt.p, nt);
            convert_matrix<float, double>((int)C, 1, V32.p, (int)C, Dsc.p, (int)C);
            add_into((size_t)nt * k, Dt.p, Xt.p);
            add_into(C, Dsc.p, Xsc.p);
            // residual: R = B - A X, the shared T X of all instances as one product
            convert_matrix<double, double>(nt, k, Bt.p, nt, Rt.p, nt);
            dgemm(nt, k, nt, -1.0, tb->T64.p, nt, Xt.p, nt, 1.0, Rt.p, nt);
            convert_matrix<double, double>((int)C, 1, Bsc.p, (int)C, Rsc.p, (int)C);
            for (int i = 0; i < k; ++i) {
                const Inst& q = *I[i];
                const double* xs = Xsc.p + coff[i];
                double* rs = Rsc.p + coff[i];
                gemv<double>(nt, q.m, 1, -1.0, BC.p + (size_t)coff[i] * nt, nt, xs, q.m, 1.0, Rt.p + (size_t)i * nt, nt);
                gemv<double>(q.ns, nt, 1, -1.0, D.p + oD[i], q.ns, Xt.p + (size_t)i * nt, nt, 1.0, rs, q.m);
                gemv<double>(q.ns, q.m, 1, -1.0, GE.p + oG[i], q.ns, xs, q.m, 1.0, rs, q.m);
                constraint_apply(q.Ns, 1, xs, q.m, dm[i].cen_xyz.p, dm[i].area.p, sc[i], q.xc, rs + q.ns, q.m);
            }
            hipLaunchKernelGGL(batch_norms_kernel, dim3(k), dim3(256), 0, ctx().s(), nt, Rt.p, Rsc.p, dcoff.p, dmlen.p,
                               nrm.p);
            check_launch("batch_norms");
            nrm.download(rn.data(), k);
            worst = 0.0;
            for (int i = 0; i < k; ++i) {
                rel[i] = bn[i] > 0 ? rn[i] / bn[i] : rn[i];
                if (rel[i] > tol || itc[i] == 0) itc[i] = it + 1;
                worst = std::max(worst, rel[i]);
            }
            if (worst <= tol) {
                ++it;
                break;
            }
        }
        SQ_REQUIRE(worst <= 1e-9, SQBEM_ERR_CONVERGE,
                   "iterative refinement stalled at relative residual " + std::to_string(worst));
        Xsc.download(X.data(), C);
        for (int i = 0; i < k; ++i) {
            iters[I[i]->idx] = itc[i];
            resid[I[i]->idx] = rel[i];
        }
        st.max_iters = std::max(st.max_iters, it);
        st.worst = std::max(st.worst, worst);
    }
    for (int i = 0; i < k; ++i) {
        const Inst& q = *I[i];
        for (int c = 0; c < 3; ++c) {
            U[3 * q.idx + c] = X[coff[i] + q.ns + c];
            Om[3 * q.idx + c] = X[coff[i] + q.ns + 3 + c];
        }
    }
    st.t_solve += now_s() - t2;
}

}  // namespace
}  // namespace sq

using namespace sq;

extern "C" {

int sqbem_batch_solve(int n, const sqbem_mesh* const* tubes, const sqbem_mesh* const* spheres, const double* centers,
                      const double* e, const double* B, double mu, int precision, double mem_budget_gb, double* U,
                      double* Om, int* iters, double* resid, double* stats) {
    return guarded([&] {
        SQ_REQUIRE(n >= 0 && (n == 0 || (tubes && spheres && centers && e && B && U && Om)), SQBEM_ERR_ARG,
                   "sqbem_batch_solve: bad arguments");
        SQ_REQUIRE(precision >= SQBEM_FP64 && precision <= SQBEM_BF16_IR, SQBEM_ERR_ARG, "unknown precision");
        std::vector<int> it_local(std::max(n, 1));
        std::vector<double> res_local(std::max(n, 1));
        if (!iters) iters = it_local.data();
        if (!resid) resid = res_local.data();
        Stats st;
        // unique tubes by content
        std::vector<const bem::Mesh*> utube;
        std::vector<uint64_t> uhash;
        std::vector<Inst> inst(n);
        for (int i = 0; i < n; ++i) {
            SQ_REQUIRE(tubes[i] && tubes[i]->m.P() > 0, SQBEM_ERR_ARG, "sqbem_batch_solve: every instance needs a tube");
            SQ_REQUIRE(spheres[i] && spheres[i]->m.P() > 0, SQBEM_ERR_ARG, "sqbem_batch_solve: empty sphere mesh");
            const bem::Mesh& tm = tubes[i]->m;
            const uint64_t h = mesh_hash(tm);
            int u = -1;
            for (size_t j = 0; j < utube.size(); ++j)
                if (uhash[j] == h && (utube[j] == &tm || same_mesh(*utube[j], tm))) {
                    u = (int)j;
                    break;
                }
            if (u < 0) {
                u = (int)utube.size();
                utube.push_back(&tm);
                uhash.push_back(h);
            }
            Inst& q = inst[i];
            q.idx = i;
            q.tube = u;
            q.S = spheres[i]->m;
            for (int c = 0; c < 3; ++c) {
                q.xc[c] = centers[3 * i + c];
                q.e[c] = e[3 * i + c];
            }
            q.B1 = B[2 * i];
            q.B2 = B[2 * i + 1];
            q.Ns = q.S.P();
            q.ns = 3 * q.Ns;
            q.m = q.ns + 6;
        }
        st.unique_tubes = (int)utube.size();
        const double budget = (mem_budget_gb > 0 ? mem_budget_gb : 64.0) * 1e9;
        // per tube: build, run its chunks (largest instances first), release
        for (int u = 0; u < (int)utube.size(); ++u) {
            std::vector<Inst*> mine;
            for (Inst& q : inst)
                if (q.tube == u) mine.push_back(&q);
            std::stable_sort(mine.begin(), mine.end(), [](const Inst* a, const Inst* b) { return a->Ns > b->Ns; });
            double t0 = now_s();
            std::unique_ptr<sqbem_tube> tb(new sqbem_tube);
            tb->mesh = *utube[u];
            tb->mu = mu;
            tb->prec = precision;
            tube_build(tb.get());
            sync();
            st.t_tubes += now_s() - t0;
            size_t a = 0;
            while (a < mine.size()) {
                size_t b = a;
                double bytes = 0;
                while (b < mine.size()) {
                    const double need = inst_bytes(tb->nt, mine[b]->ns, precision);
                    if (b > a && bytes + need > budget) break;
                    bytes += need;
                    ++b;
                }
                std::vector<Inst*> chunk(mine.begin() + a, mine.begin() + b);
                run_chunk(tb.get(), chunk, precision, mu, U, Om, iters, resid, st);
                ++st.chunks;
                a = b;
            }
        }
        if (stats) {
            stats[0] = st.t_tubes;
            stats[1] = st.t_assembly;
            stats[2] = st.t_schur;
            stats[3] = st.t_solve;
            stats[4] = st.unique_tubes;
            stats[5] = st.chunks;
            stats[6] = st.max_iters;
            stats[7] = st.worst;
        }
    });
}

}  // extern "C"
