// batch.hip -- many independent squirmer-in-tube solves of uneven size in one call.
//
// The reference solves one configuration at a time: its kinematics sweeps build a
// SwimmerSolver per configuration and reuse the Tube across offsets (src/main.cpp:82-100,
// src/swimmer.hpp:83-139). Here a whole batch is one device pipeline:
//
//  * tube meshes with identical contents share one assembly and factorisation (a tube
//    whose mesh does not depend on the swimmer, e.g. uniform axial nodes, is factored once
//    for every configuration in it);
//  * instances on the same tube are processed in chunks sized to a memory budget, in one
//    workspace allocated once per call; every coupling block of a chunk (tube rows <-
//    sphere, sphere rows <- tube, sphere <- sphere) is integrated by one persistent
//    work-queue launch (assemble_batch);
//  * the Schur right-hand sides [B_i | Ct_i] of all instances of a chunk sit side by side
//    in one nt x sum(ns_i + 6) matrix. FP64 mode solves them through the shared LU factors
//    in one triangular solve; the refined modes form the tube's inverse once (FP32, from
//    its factors) and take T^-1 [B | Ct] as one large GEMM in the mode's MFMA precision;
//  * the bordered force/torque-free matrices of all instances are factored together by the
//    batched no-pivot LU (one launch per step for the whole chunk), each verified against
//    the partial-pivoting condition and refactored with pivoting if it fails;
//  * FP64 mode then solves the reduced systems directly (src/swimmer.hpp:107-139,
//    256-288); the refined modes iterate on every instance's full bordered system together:
//    the shared-tube products are single GEMMs, the per-instance products one batched
//    matrix-vector launch, the bordered solves one batched chained solve, and one host
//    round trip per refinement step serves all convergence checks.
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

// ------------------------------------------------------------------ workspace
// Bump allocation over one device buffer. A sizing pass (base == nullptr) only counts.
struct Arena {
    unsigned char* base = nullptr;
    size_t cap = 0, off = 0;
    template <typename T>
    T* take(size_t count) {
        off = (off + 255) / 256 * 256;
        T* p = base ? (T*)(base + off) : nullptr;
        off += count * sizeof(T);
        SQ_REQUIRE(!base || off <= cap, SQBEM_ERR_STATE, "batch workspace overflow");
        return p;
    }
};

// ------------------------------------------------------------------ batched small kernels
// y_j += alpha_j (A1_j x1_j + A2_j x2_j) for many independent jobs (one right-hand side
// each, FP64 sums): column chunks of every job in parallel, then a reduction per job.
struct GvJob {
    int m, n1, n2;
    const void* A1;
    long lda1;
    const void* x1;
    const void* A2;
    long lda2;
    const void* x2;
    void* y;
    double alpha;
    long poff;  // partial sums at part + poff, GV_CH x m
};
constexpr int GV_CH = 16;

template <typename T>
__global__ void gemv_batch_part(const GvJob* __restrict__ jobs, double* __restrict__ part) {
    const GvJob& g = jobs[blockIdx.z];
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= g.m) return;
    const int n = g.n1 + g.n2, per = (n + GV_CH - 1) / GV_CH;
    const int ja = blockIdx.y * per, jb = min(n, ja + per);
    const T* A1 = (const T*)g.A1;
    const T* x1 = (const T*)g.x1;
    const T* A2 = (const T*)g.A2;
    const T* x2 = (const T*)g.x2;
    double acc = 0.0;
    // columns j0 .. j1-1 of one matrix in order, eight loads in flight a round
    auto run = [&](const T* A, long lda, const T* x, int j0, int j1) {
        int j = j0;
        for (; j + 8 <= j1; j += 8) {
            double a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = (double)A[row + (size_t)(j + u) * lda];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += a[u] * (double)x[j + u];
        }
        for (; j < j1; ++j) acc += (double)A[row + (size_t)j * lda] * (double)x[j];
    };
    run(A1, g.lda1, x1, ja, min(jb, g.n1));
    run(A2, g.lda2, x2, max(ja, g.n1) - g.n1, jb - g.n1);
    part[g.poff + (size_t)blockIdx.y * g.m + row] = acc;
}

template <typename T>
__global__ void gemv_batch_reduce(const GvJob* __restrict__ jobs, const double* __restrict__ part) {
    const GvJob& g = jobs[blockIdx.y];
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= g.m) return;
    // every chunk's load issued before the (in-order) sum
    double v[GV_CH];
#pragma unroll
    for (int c = 0; c < GV_CH; ++c) v[c] = part[g.poff + (size_t)c * g.m + row];
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < GV_CH; ++c) s += v[c];
    T* y = (T*)g.y;
    y[row] = (T)((double)y[row] + g.alpha * s);
}

// r_c = -sc F f_s for many instances (the constraint rows of the bordered system)
struct CJob {
    int Ns;
    const double* fs;
    const double* cen;
    const double* area;
    const double* sc;  // device: the instance's bordered scale
    double xc[3];
    double* out;
};
__global__ void constraint_batch_kernel(const CJob* __restrict__ jobs) {
    const CJob& j = jobs[blockIdx.x];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double sc = *j.sc;
    for (int p = threadIdx.x; p < j.Ns; p += blockDim.x) {
        const double rx = j.cen[3 * (size_t)p] - j.xc[0], ry = j.cen[3 * (size_t)p + 1] - j.xc[1],
                     rz = j.cen[3 * (size_t)p + 2] - j.xc[2];
        const double A = j.area[p] * sc;
        const double f0 = j.fs[3 * (size_t)p], f1 = j.fs[3 * (size_t)p + 1], f2 = j.fs[3 * (size_t)p + 2];
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
    if (threadIdx.x < 6) j.out[threadIdx.x] = -red[threadIdx.x][0];
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

// sc_i = max |G_i| / Amax_i, the maximum as the bits of a non-negative double
__global__ void border_scale_kernel(int k, const unsigned long long* __restrict__ amax, const double* __restrict__ Amax,
                                    double* sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) sc[i] = __longlong_as_double((long long)amax[i]) / Amax[i];
}

__global__ void batch_iota_kernel(int n, int* a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ void identity_kernel(int n, float* A) {
    const size_t total = (size_t)n * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
        A[e] = (e % n) == (e / n) ? 1.0f : 0.0f;
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

// Job table of a batched small kernel, uploaded once (the launches that use it then copy
// nothing, so they can be replayed from a captured graph while the table lives).
template <typename J>
struct Table {
    DBuf<unsigned char> d;
    int n = 0;
    void put(const std::vector<J>& v) {
        n = (int)v.size();
        if (!n) return;
        d.alloc(v.size() * sizeof(J));
        SQ_HIP(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(J), hipMemcpyHostToDevice, ctx().s()));
        sync();  // v is staging
    }
    const J* get() const { return (const J*)d.p; }
};

// y_j += alpha_j (A1_j x1_j + A2_j x2_j) for a prepared set of jobs
struct GvPlan {
    Table<GvJob> tab;
    int mmax = 1;
    double flops = 0, bytes = 0;
};
template <typename T>
void gemv_prepare(GvPlan& p, std::vector<GvJob> jobs) {
    long off = 0;
    p.mmax = 1;
    p.flops = p.bytes = 0;
    for (GvJob& g : jobs) {
        g.poff = off;
        off += (long)GV_CH * g.m;
        p.mmax = std::max(p.mmax, g.m);
        p.bytes += sizeof(T) * (double)g.m * (g.n1 + g.n2);
        p.flops += 2.0 * g.m * (double)(g.n1 + g.n2);
    }
    p.tab.put(jobs);
}
template <typename T>
void gemv_launch(const GvPlan& p, double* part) {
    if (!p.tab.n) return;
    ProfScope prof(PROF_GEMV, p.flops, p.bytes);
    hipLaunchKernelGGL(gemv_batch_part<T>, dim3(ceil_div(p.mmax, 256), GV_CH, (unsigned)p.tab.n), dim3(256), 0,
                       ctx().s(), p.tab.get(), part);
    hipLaunchKernelGGL(gemv_batch_reduce<T>, dim3(ceil_div(p.mmax, 256), (unsigned)p.tab.n), dim3(256), 0, ctx().s(),
                       p.tab.get(), part);
    check_launch("gemv_batch");
}
template <typename T>
void gemv_batch(std::vector<GvJob> jobs, double* part) {
    GvPlan p;
    gemv_prepare<T>(p, std::move(jobs));
    gemv_launch<T>(p, part);
    sync();
}

size_t gemv_part_elems(int nt, const std::vector<int>& ns, const std::vector<int>& m) {
    size_t s = 0;
    for (size_t i = 0; i < ns.size(); ++i) s += (size_t)GV_CH * (nt + ns[i] + m[i]);
    return std::max<size_t>(s, 1);
}

struct Stats {
    double t_tubes = 0, t_assembly = 0, t_schur = 0, t_solve = 0;
    int unique_tubes = 0, chunks = 0, max_iters = 0, pivoted = 0, graph_replays = 0, build_graphs = 0;
    double worst = 0;
};

bool g_use_graph = true;  // capture each chunk's build stages and its refinement step as graphs

// A tube shared by instances: its factors, and for the refined modes its FP32 inverse.
struct SharedTube {
    std::unique_ptr<sqbem_tube> t;
    DBuf<float> inv;  // T^-1 (nt x nt), refined modes
};

void build_shared_tube(SharedTube& S, const bem::Mesh& mesh, double mu, int prec) {
    S.t.reset(new sqbem_tube);
    S.t->mesh = mesh;
    S.t->mu = mu;
    S.t->prec = prec;
    tube_build(S.t.get());
    tube_ready(S.t.get());
    // the tiled copy for the chained solves, beside what follows
    chain_tiles_issue(S.t->aux32.p);
    chain_tiles_issue(S.t->aux64.p);
    if (mixed(prec)) {
        const int nt = S.t->nt;
        S.inv.alloc((size_t)nt * nt);
        hipLaunchKernelGGL(identity_kernel, dim3(8192), dim3(256), 0, ctx().s(), nt, S.inv.p);
        check_launch("identity");
        // FP32 updates whatever the mode: the inverse carries the factors' accuracy
        getrs<float>(SQBEM_FP32_IR, 0, nt, nt, S.t->T32.p, nt, S.t->ipiv.p, S.inv.p, nt, S.t->aux32.p);
    }
    sync();
}

// Device memory of one chunk (per-instance offsets into concatenated buffers, then the
// buffers themselves, taken from the arena in a fixed order; with an empty arena this only
// measures the chunk).
struct Work {
    int k = 0, maxNs = 1;
    std::vector<int> coff, soff, Noff, nsv, mv;
    std::vector<size_t> oD, oG, oA, oP2, oP3, oM, oX;
    size_t C = 0, S = 0;
    double *BC, *AuBt, *D, *GE, *AuBs, *part, *dls, *kp, *u, *Bt, *Bsc, *gvp, *nrm;
    int *self, *flags, *dcoff, *dmlen, *piv, *perm;
    unsigned long long* amax;
    double *M64 = nullptr, *X64 = nullptr, *Z64 = nullptr, *Xd = nullptr;
    float *Y32 = nullptr, *B32 = nullptr, *D32 = nullptr, *G32 = nullptr, *M32 = nullptr, *X32 = nullptr;
    double *Xt = nullptr, *Xsc = nullptr, *Rt = nullptr, *Rsc = nullptr, *Dt = nullptr, *Dsc = nullptr;
    float *Z32 = nullptr, *Zo32 = nullptr, *V32 = nullptr;
};

void layout(Arena& ar, int Nt, const std::vector<Inst*>& I, bool mx, Work& w) {
    const int k = (int)I.size(), nt = 3 * Nt;
    w.k = k;
    w.coff.assign(k + 1, 0);
    w.soff.assign(k + 1, 0);
    w.Noff.assign(k + 1, 0);
    w.nsv.resize(k);
    w.mv.resize(k);
    w.maxNs = 1;
    for (int i = 0; i < k; ++i) {
        w.coff[i + 1] = w.coff[i] + I[i]->m;
        w.soff[i + 1] = w.soff[i] + I[i]->ns;
        w.Noff[i + 1] = w.Noff[i] + I[i]->Ns;
        w.maxNs = std::max(w.maxNs, I[i]->Ns);
        w.nsv[i] = I[i]->ns;
        w.mv[i] = I[i]->m;
    }
    w.C = w.coff[k];
    w.S = w.soff[k];
    for (auto* v : {&w.oD, &w.oG, &w.oA, &w.oP2, &w.oP3, &w.oM, &w.oX}) v->resize(k);
    size_t nD = 0, nG = 0, nA = 0, nP = 0, nM = 0, nX = 0;
    for (int i = 0; i < k; ++i) {
        const size_t ns = I[i]->ns, m = I[i]->m, Ns = I[i]->Ns;
        w.oD[i] = nD; nD += ns * nt;
        w.oG[i] = nG; nG += ns * m;
        w.oA[i] = nA; nA += ns * ns;
        w.oP2[i] = nP; nP += (size_t)asm_tiles(Nt) * 9 * Ns;
        w.oP3[i] = nP; nP += (size_t)asm_tiles((int)Ns) * 9 * Ns;
        w.oM[i] = nM; nM += m * m;
        w.oX[i] = nX; nX += getrf_aux_elems((int)m);
    }
    const size_t C = w.C, S = w.S;
    ar.off = 0;
    w.BC = ar.take<double>((size_t)nt * C);    // [B_i | Ct_i], FP64 (Y in FP64 mode)
    w.AuBt = ar.take<double>((size_t)nt * S);  // -DL/(8 pi), tube rows <- sphere
    w.D = ar.take<double>(nD);                 // SL/(8 pi mu), sphere rows <- tube
    w.GE = ar.take<double>(nG);                // [E_i | Cs_i] (the Schur G in FP64 mode)
    w.AuBs = ar.take<double>(nA);
    w.part = ar.take<double>(nP);
    w.dls = ar.take<double>(9 * (size_t)w.Noff[k]);
    w.kp = ar.take<double>(kproj_part_elems(std::max(nt, 3 * w.maxNs)));
    w.u = ar.take<double>(S);
    w.Bt = ar.take<double>((size_t)nt * k);
    w.Bsc = ar.take<double>(C);
    w.gvp = ar.take<double>(gemv_part_elems(nt, w.nsv, w.mv));
    w.nrm = ar.take<double>(k);
    w.self = ar.take<int>(w.maxNs);
    w.flags = ar.take<int>(k);
    w.dcoff = ar.take<int>(k);
    w.dmlen = ar.take<int>(k);
    w.piv = ar.take<int>(C);
    w.perm = ar.take<int>(C);
    w.amax = ar.take<unsigned long long>(k);
    if (!mx) {
        w.M64 = ar.take<double>(nM);
        w.X64 = ar.take<double>(nX);
        w.Z64 = ar.take<double>((size_t)nt * k);
        w.Xd = ar.take<double>(C);
    } else {
        w.Y32 = ar.take<float>((size_t)nt * C);
        w.B32 = ar.take<float>((size_t)nt * C);
        w.D32 = ar.take<float>(nD);
        w.G32 = ar.take<float>(nG);
        w.M32 = ar.take<float>(nM);
        w.X32 = ar.take<float>(nX);
        w.Xt = ar.take<double>((size_t)nt * k);
        w.Xsc = ar.take<double>(C);
        w.Rt = ar.take<double>((size_t)nt * k);
        w.Rsc = ar.take<double>(C);
        w.Dt = ar.take<double>((size_t)nt * k);
        w.Dsc = ar.take<double>(C);
        w.Z32 = ar.take<float>((size_t)nt * k);
        w.Zo32 = ar.take<float>((size_t)nt * k);
        w.V32 = ar.take<float>(C);
    }
}

size_t chunk_bytes(int Nt, const std::vector<Inst*>& I, bool mx) {
    Arena a;
    Work w;
    layout(a, Nt, I, mx, w);
    return a.off;
}

// One chunk: instances I (all on tube S), built, reduced and solved together.
void run_chunk(SharedTube& ST, std::vector<Inst*>& I, int prec, double mu, Arena& ar, double* U, double* Om, int* iters,
               double* resid, Stats& st) {
    const sqbem_tube* tb = ST.t.get();
    const int k = (int)I.size();
    const int Nt = tb->Nt, nt = tb->nt;
    const double e8 = bem::EIGHT_PI;
    const bool mx = mixed(prec);
    Work w;
    layout(ar, Nt, I, mx, w);
    const std::vector<int>&coff = w.coff, &soff = w.soff, &Noff = w.Noff, &mv = w.mv;
    const std::vector<size_t>&oD = w.oD, &oG = w.oG, &oA = w.oA, &oP2 = w.oP2, &oP3 = w.oP3, &oM = w.oM, &oX = w.oX;
    const size_t C = w.C, S = w.S;
    const int maxNs = w.maxNs;
    double *BC = w.BC, *AuBt = w.AuBt, *D = w.D, *GE = w.GE, *AuBs = w.AuBs, *part = w.part, *dls = w.dls, *kp = w.kp,
           *u = w.u, *Bt = w.Bt, *Bsc = w.Bsc, *gvp = w.gvp, *nrm = w.nrm;
    int *self = w.self, *flags = w.flags, *dcoff = w.dcoff, *dmlen = w.dmlen, *piv = w.piv, *perm = w.perm;
    unsigned long long* amax = w.amax;
    double *M64 = w.M64, *X64 = w.X64, *Z64 = w.Z64, *Xd = w.Xd;
    float *Y32 = w.Y32, *B32 = w.B32, *D32 = w.D32, *G32 = w.G32, *M32 = w.M32, *X32 = w.X32;
    double *Xt = w.Xt, *Xsc = w.Xsc, *Rt = w.Rt, *Rsc = w.Rsc, *Dt = w.Dt, *Dsc = w.Dsc;
    float *Z32 = w.Z32, *Zo32 = w.Zo32, *V32 = w.V32;

    // ---- preparation on the host: every job table and upload the chunk's build needs,
    // then one wait; the build itself (coupling assembly, right-hand sides, Schur
    // complement, bordered factors) is launches only, two captured graphs when it records
    // nothing else, with one read-back at its end (the no-pivot verdicts and the norms)
    double t0 = now_s();
    std::vector<DMesh> dm(k);
    for (int i = 0; i < k; ++i) dm[i].upload(I[i]->S);
    std::vector<AsmJob> jobs;
    std::vector<DlsReduce> red;
    for (int i = 0; i < k; ++i) {
        const Inst& q = *I[i];
        double Asum = 0;
        for (int p = 0; p < q.Ns; ++p) Asum += q.S.area[p];
        const double defl = 1.0 / (e8 * mu * std::sqrt(Asum));
        // tube rows <- sphere sources: B = SL/(8 pi mu), AuBt = -DL/(8 pi)   (src/swimmer.hpp:188-198)
        jobs.push_back({&dm[i], 0, q.Ns, tb->dm.cen_xyz.p, Nt, nullptr, e8 * mu, BC + (size_t)coff[i] * nt, nt, e8,
                        AuBt + (size_t)soff[i] * nt, nt, nullptr, 0.0, nullptr});
        // sphere rows <- tube sources: D = SL/(8 pi mu), stresslet row sums (src/swimmer.hpp:201-207)
        jobs.push_back({&tb->dm, 0, Nt, dm[i].cen_xyz.p, q.Ns, nullptr, e8 * mu, D + oD[i], q.ns, 1.0, nullptr, 0,
                        part + oP2[i], 0.0, nullptr});
        // sphere rows <- sphere sources: E (deflated), AuBs, stresslet row sums (src/swimmer.hpp:210-222)
        jobs.push_back({&dm[i], 0, q.Ns, dm[i].cen_xyz.p, q.Ns, self, e8 * mu, GE + oG[i], q.ns, e8, AuBs + oA[i], q.ns,
                        part + oP3[i], defl, dm[i].nrm_xyz.p});
        red.push_back({q.Ns, asm_tiles(Nt), part + oP2[i], asm_tiles(q.Ns), part + oP3[i], dls + 9 * (size_t)Noff[i]});
    }
    AsmBatch asm_plan;
    assemble_batch_prepare(jobs, asm_plan);
    DlsBatch dls_plan;
    reduce_dls_prepare(red, dls_plan);
    // right-hand sides of the full bordered system for each instance's slip:
    // b_t = -AuBt u, b_s = -AuBs u, b_c = 0 (src/swimmer.hpp:256-269)
    GvPlan rhs_plan;
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
        SQ_HIP(hipMemcpyAsync(u, hu.data(), S * sizeof(double), hipMemcpyHostToDevice, ctx().s()));
        std::vector<GvJob> g;
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            g.push_back({nt, q.ns, 0, AuBt + (size_t)soff[i] * nt, nt, u + soff[i], nullptr, 0, nullptr,
                         Bt + (size_t)i * nt, -1.0, 0});
            g.push_back({q.ns, q.ns, 0, AuBs + oA[i], q.ns, u + soff[i], nullptr, 0, nullptr, Bsc + coff[i], -1.0, 0});
        }
        gemv_prepare<double>(rhs_plan, g);  // synchronises: hu is uploaded
    }
    // the bordered scale sc_i = max |G_i| / (5 max area_i) (the max |G_i| found on the
    // device), Amax_i from the host mesh
    DBuf<double> Amax_d(k), sc_d(k);
    {
        std::vector<double> Amax(k, 0.0);
        for (int i = 0; i < k; ++i)
            for (int p = 0; p < I[i]->Ns; ++p) Amax[i] = std::max(Amax[i], 5.0 * I[i]->S.area[p]);
        Amax_d.upload(Amax.data(), k);
    }
    BatchLuPlan<float> lu32;
    BatchLuPlan<double> lu64;
    if (mx) {
        std::vector<BatchMat<float>> bm;
        for (int i = 0; i < k; ++i) bm.push_back({M32 + oM[i], I[i]->m, I[i]->m, X32 + oX[i], flags + i});
        getrf_np_batch_prepare<float>(bm, lu32);
    } else {
        std::vector<BatchMat<double>> bm;
        for (int i = 0; i < k; ++i) bm.push_back({M64 + oM[i], I[i]->m, I[i]->m, X64 + oX[i], flags + i});
        getrf_np_batch_prepare<double>(bm, lu64);
    }
    if (mx) {
        std::vector<int> hc(coff.begin(), coff.begin() + k);
        SQ_HIP(hipMemcpyAsync(dcoff, hc.data(), sizeof(int) * k, hipMemcpyHostToDevice, ctx().s()));
        SQ_HIP(hipMemcpyAsync(dmlen, mv.data(), sizeof(int) * k, hipMemcpyHostToDevice, ctx().s()));
    }
    sync();  // every host staging buffer above is consumed
    const double t_prep = now_s() - t0;

    auto border = [&](int i) {
        const Inst& q = *I[i];
        if (mx) border_build<float>(q.Ns, G32 + oG[i], M32 + oM[i], dm[i].cen_xyz.p, dm[i].area.p, 0.0, q.xc, sc_d.p + i);
        else border_build<double>(q.Ns, GE + oG[i], M64 + oM[i], dm[i].cen_xyz.p, dm[i].area.p, 0.0, q.xc, sc_d.p + i);
    };
    // stage 1: coupling assembly (src/swimmer.hpp:188-234) and the right-hand sides
    auto stage_assembly = [&]() {
        hipLaunchKernelGGL(batch_iota_kernel, dim3(ceil_div(maxNs, 256)), dim3(256), 0, ctx().s(), maxNs, self);
        check_launch("iota");
        assemble_batch_launch(asm_plan);
        reduce_dls_launch(dls_plan);
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            completion_diag(q.Ns, AuBs + oA[i], q.ns, dls + 9 * (size_t)Noff[i], 0.0);
            // Ct = AuBt K beside B_i, Cs = AuBs K beside E_i (src/swimmer.hpp:223-234)
            kproj_async(nt, q.Ns, AuBt + (size_t)soff[i] * nt, nt, dm[i].cen_xyz.p, q.xc,
                        BC + ((size_t)coff[i] + q.ns) * nt, nt, kp);
            kproj_async(q.ns, q.Ns, AuBs + oA[i], q.ns, dm[i].cen_xyz.p, q.xc, GE + oG[i] + (size_t)q.ns * q.ns, q.ns,
                        kp);
        }
        fill_zero(Bt, sizeof(double) * (size_t)nt * k);
        fill_zero(Bsc, sizeof(double) * C);
        gemv_launch<double>(rhs_plan, gvp);
    };
    // stage 2: Schur complement (src/swimmer.hpp:235-249), bordered matrices and their
    // factors without pivoting (each verified against the partial-pivoting condition)
    auto stage_schur = [&]() {
        fill_zero(amax, sizeof(unsigned long long) * k);
        if (!mx) {
            getrs<double>(SQBEM_FP64, 0, nt, (int)C, tb->T64.p, nt, tb->ipiv.p, BC, nt, tb->aux64.p);  // Y in place
            for (int i = 0; i < k; ++i) {
                const Inst& q = *I[i];
                dgemm(q.ns, q.m, nt, -1.0, D + oD[i], q.ns, BC + (size_t)coff[i] * nt, nt, 1.0, GE + oG[i], q.ns);
                absmax_async<double>(q.ns, q.ns, GE + oG[i], q.ns, amax + i);
            }
        } else {
            // Y = T^-1 [B | Ct] for every instance: one GEMM through the shared inverse
            convert_matrix<double, float>(nt, (int)C, BC, nt, B32, nt);
            if (prec == SQBEM_FP32_IR)
                sgemm(nt, (int)C, nt, 1.0f, ST.inv.p, nt, B32, nt, 0.0f, Y32, nt);
            else
                gemm16(prec, nt, (int)C, nt, 1.0f, ST.inv.p, nt, B32, nt, 0.0f, Y32, nt);
            for (int i = 0; i < k; ++i) {
                const Inst& q = *I[i];
                convert_matrix<double, float>(q.ns, nt, D + oD[i], q.ns, D32 + oD[i], q.ns);
                convert_matrix<double, float>(q.ns, q.m, GE + oG[i], q.ns, G32 + oG[i], q.ns);
                const float* Yi = Y32 + (size_t)coff[i] * nt;
                if (prec == SQBEM_FP32_IR)
                    sgemm(q.ns, q.m, nt, -1.0f, D32 + oD[i], q.ns, Yi, nt, 1.0f, G32 + oG[i], q.ns);
                else
                    gemm16(prec, q.ns, q.m, nt, -1.0f, D32 + oD[i], q.ns, Yi, nt, 1.0f, G32 + oG[i], q.ns);
                absmax_async<float>(q.ns, q.ns, G32 + oG[i], q.ns, amax + i);
            }
        }
        hipLaunchKernelGGL(border_scale_kernel, dim3(ceil_div(k, 256)), dim3(256), 0, ctx().s(), k, amax, Amax_d.p,
                           sc_d.p);
        check_launch("border_scale");
        for (int i = 0; i < k; ++i) border(i);
        if (mx) getrf_np_batch_launch<float>(lu32);
        else getrf_np_batch_launch<double>(lu64);
        if (mx) {
            hipLaunchKernelGGL(batch_norms_kernel, dim3(k), dim3(256), 0, ctx().s(), nt, Bt, Bsc, dcoff, dmlen, nrm);
            check_launch("batch_norms");
        }
    };
    const bool use_graph = g_use_graph && !prof_enabled();
    hipEvent_t ev[3];
    for (auto& x : ev) SQ_HIP(hipEventCreate(&x));
    std::vector<int> hflag(k, 0);
    std::vector<double> bn(k, 0.0);
    try {
        GraphExec g1, g2;
        SQ_HIP(hipEventRecord(ev[0], ctx().s()));
        if (use_graph) st.build_graphs += launch_as_graph(stage_assembly, &g1.x);
        else stage_assembly();
        SQ_HIP(hipEventRecord(ev[1], ctx().s()));
        if (use_graph) st.build_graphs += launch_as_graph(stage_schur, &g2.x);
        else stage_schur();
        SQ_HIP(hipEventRecord(ev[2], ctx().s()));
        SQ_HIP(hipMemcpyAsync(hflag.data(), flags, sizeof(int) * k, hipMemcpyDeviceToHost, ctx().s()));
        if (mx) SQ_HIP(hipMemcpyAsync(bn.data(), nrm, sizeof(double) * k, hipMemcpyDeviceToHost, ctx().s()));
        sync();
    } catch (...) {
        for (auto& x : ev) (void)hipEventDestroy(x);
        throw;
    }
    float ms_asm = 0, ms_schur = 0;
    (void)hipEventElapsedTime(&ms_asm, ev[0], ev[1]);
    (void)hipEventElapsedTime(&ms_schur, ev[1], ev[2]);
    for (auto& x : ev) (void)hipEventDestroy(x);
    st.t_assembly += t_prep + 1e-3 * ms_asm;
    double t1 = now_s();
    // any bordered matrix that partial pivoting would interchange is refactored alone with
    // pivoting (src/utils/linalg.hpp:42-77)
    std::vector<char> pivoted(k, 0);
    for (int i = 0; i < k; ++i) {
        if (!hflag[i]) continue;
        pivoted[i] = 1;
        ++st.pivoted;
        border(i);
        const int info = mx ? getrf<float>(prec, I[i]->m, M32 + oM[i], I[i]->m, piv + coff[i], X32 + oX[i])
                            : getrf<double>(SQBEM_FP64, I[i]->m, M64 + oM[i], I[i]->m, piv + coff[i], X64 + oX[i]);
        SQ_REQUIRE(info == 0, SQBEM_ERR_SINGULAR, "dgesv failed");
        // LAPACK's sequential interchanges as one gather: (P v)[r] = v[perm[r]]
        const int m = I[i]->m;
        std::vector<int> ip(m), pr(m);
        SQ_HIP(hipMemcpyAsync(ip.data(), piv + coff[i], sizeof(int) * m, hipMemcpyDeviceToHost, ctx().s()));
        sync();
        for (int r = 0; r < m; ++r) pr[r] = r;
        for (int r = 0; r < m; ++r) std::swap(pr[r], pr[ip[r] - 1]);
        SQ_HIP(hipMemcpyAsync(perm + coff[i], pr.data(), sizeof(int) * m, hipMemcpyHostToDevice, ctx().s()));
        sync();
    }
    sync();
    double t2 = now_s();
    st.t_schur += 1e-3 * ms_schur + (t2 - t1);

    // bordered solves of all instances, B_i <- M_i^-1 B_i: one batched chained solve (the
    // pivoted ones gather their rows through the permutation first)
    auto bordered_plan = [&](auto* Xbase, auto* V, auto& plan) {
        using T = std::remove_pointer_t<decltype(V)>;
        std::vector<BatchSolve<T>> bs;
        for (int i = 0; i < k; ++i)
            bs.push_back({I[i]->m, Xbase + oX[i], V + coff[i], pivoted[i] ? perm + coff[i] : nullptr});
        getrs_np_batch_prepare<T>(bs, plan);
    };

    // ---- solves
    std::vector<double> X(C);
    if (!mx) {
        // z = T^-1 b_t for all instances at once; x_i = M_i^-1 [b_s - D_i z_i; 0]
        convert_matrix<double, double>(nt, k, Bt, nt, Z64, nt);
        getrs<double>(SQBEM_FP64, 0, nt, k, tb->T64.p, nt, tb->ipiv.p, Z64, nt, tb->aux64.p);
        convert_matrix<double, double>((int)C, 1, Bsc, (int)C, Xd, (int)C);
        std::vector<GvJob> g;
        for (int i = 0; i < k; ++i)
            g.push_back({I[i]->ns, nt, 0, D + oD[i], I[i]->ns, Z64 + (size_t)i * nt, nullptr, 0, nullptr, Xd + coff[i],
                         -1.0, 0});
        gemv_batch<double>(g, gvp);
        BatchSolvePlan<double> bp;
        bordered_plan(X64, Xd, bp);
        getrs_np_batch_launch<double>(bp);
        SQ_HIP(hipMemcpyAsync(X.data(), Xd, C * sizeof(double), hipMemcpyDeviceToHost, ctx().s()));
        sync();
        for (int i = 0; i < k; ++i) {
            iters[I[i]->idx] = 0;
            resid[I[i]->idx] = 0.0;
        }
    } else {
        // iterative refinement of every instance's full bordered system
        //   [T B Ct; D E Cs; 0 scF 0] [f_t; f_s; w] = [b_t; b_s; 0]
        // residuals from the FP64 operator, corrections from the low-precision Schur pieces
        fill_zero(Xt, sizeof(double) * (size_t)nt * k);
        fill_zero(Xsc, sizeof(double) * C);
        convert_matrix<double, double>(nt, k, Bt, nt, Rt, nt);
        convert_matrix<double, double>((int)C, 1, Bsc, (int)C, Rsc, (int)C);
        std::vector<double> rn(k), rel(k, 1.0);  // bn: the right-hand-side norms, read back with the verdicts
        std::vector<int> itc(k, 0);
        // job lists of the per-instance products (fixed for the whole refinement)
        std::vector<GvJob> gv_s, gv_t, gv_r;
        std::vector<CJob> cj;
        for (int i = 0; i < k; ++i) {
            const Inst& q = *I[i];
            float* zi = Zo32 + (size_t)i * nt;
            float* vi = V32 + coff[i];
            gv_s.push_back({q.ns, nt, 0, D32 + oD[i], q.ns, zi, nullptr, 0, nullptr, vi, -1.0, 0});
            gv_t.push_back({nt, q.m, 0, Y32 + (size_t)coff[i] * nt, nt, vi, nullptr, 0, nullptr, zi, -1.0, 0});
            const double* xs = Xsc + coff[i];
            gv_r.push_back({nt, q.m, 0, BC + (size_t)coff[i] * nt, nt, xs, nullptr, 0, nullptr, Rt + (size_t)i * nt,
                            -1.0, 0});
            gv_r.push_back({q.ns, nt, q.m, D + oD[i], q.ns, Xt + (size_t)i * nt, GE + oG[i], q.ns, xs, Rsc + coff[i],
                            -1.0, 0});
            cj.push_back({q.Ns, xs, dm[i].cen_xyz.p, dm[i].area.p, sc_d.p + i, {q.xc[0], q.xc[1], q.xc[2]},
                          Rsc + coff[i] + q.ns});
        }
        GvPlan ps, pt, pr;
        gemv_prepare<float>(ps, gv_s);
        gemv_prepare<float>(pt, gv_t);
        gemv_prepare<double>(pr, gv_r);
        Table<CJob> ct;
        ct.put(cj);
        BatchSolvePlan<float> bp;
        bordered_plan(X32, V32, bp);
        // one refinement step: launches only, all tables prepared above
        auto step = [&]() {
            // corrections: z = T^-1 r_t; v_i = M_i^-1 [r_s - D_i z_i; r_c]; z_i -= Y_i v_i
            convert_matrix<double, float>(nt, k, Rt, nt, Z32, nt);
            sgemm(nt, k, nt, 1.0f, ST.inv.p, nt, Z32, nt, 0.0f, Zo32, nt);
            convert_matrix<double, float>((int)C, 1, Rsc, (int)C, V32, (int)C);
            gemv_launch<float>(ps, gvp);
            getrs_np_batch_launch<float>(bp);
            gemv_launch<float>(pt, gvp);
            convert_matrix<float, double>(nt, k, Zo32, nt, Dt, nt);
            convert_matrix<float, double>((int)C, 1, V32, (int)C, Dsc, (int)C);
            add_into((size_t)nt * k, Dt, Xt);
            add_into(C, Dsc, Xsc);
            // residual R = B - A X; the shared T X of all instances as one product
            convert_matrix<double, double>(nt, k, Bt, nt, Rt, nt);
            dgemm(nt, k, nt, -1.0, tb->T64.p, nt, Xt, nt, 1.0, Rt, nt);
            convert_matrix<double, double>((int)C, 1, Bsc, (int)C, Rsc, (int)C);
            gemv_launch<double>(pr, gvp);
            hipLaunchKernelGGL(constraint_batch_kernel, dim3(k), dim3(256), 0, ctx().s(), ct.get());
            check_launch("constraint_batch");
            hipLaunchKernelGGL(batch_norms_kernel, dim3(k), dim3(256), 0, ctx().s(), nt, Rt, Rsc, dcoff, dmlen, nrm);
            check_launch("batch_norms");
        };
        // the step is a fixed launch sequence: capture it once and replay it (not while the
        // per-launch timing is on, whose events would be captured too)
        hipGraphExec_t exec = nullptr;
        if (g_use_graph && !prof_enabled()) {
            hipGraph_t graph = nullptr;
            SQ_HIP(hipStreamBeginCapture(ctx().s(), hipStreamCaptureModeThreadLocal));
            try {
                step();
            } catch (...) {
                (void)hipStreamEndCapture(ctx().s(), &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            SQ_HIP(hipStreamEndCapture(ctx().s(), &graph));
            const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            SQ_REQUIRE(ie == hipSuccess, SQBEM_ERR_HIP, std::string("refinement graph: ") + hipGetErrorString(ie));
        }
        const double tol = SQBEM_IR_TOL;
        int it = 0;
        double worst = 1.0;
        try {
            for (; it < 40; ++it) {
                if (exec) {
                    SQ_HIP(hipGraphLaunch(exec, ctx().s()));
                    ++st.graph_replays;
                } else {
                    step();
                }
                SQ_HIP(hipMemcpyAsync(rn.data(), nrm, sizeof(double) * k, hipMemcpyDeviceToHost, ctx().s()));
                mark_device_error_check();  // the chained solves report a timed-out wait here
                sync();
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
        } catch (...) {
            if (exec) (void)hipGraphExecDestroy(exec);
            throw;
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        SQ_REQUIRE(worst <= tol, SQBEM_ERR_CONVERGE,
                   "iterative refinement did not reach relative residual 1e-12 (" + std::to_string(worst) +
                       " after " + std::to_string(it) + " steps)");
        SQ_HIP(hipMemcpyAsync(X.data(), Xsc, C * sizeof(double), hipMemcpyDeviceToHost, ctx().s()));
        sync();
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

int sqbem_batch_options(int use_graph) {
    return guarded([&] { g_use_graph = use_graph != 0; });
}

int sqbem_batch_solve(int n, const sqbem_mesh* const* tubes, const sqbem_mesh* const* spheres, const double* centers,
                      const double* e, const double* B, double mu, int precision, double mem_budget_gb, double* U,
                      double* Om, int* iters, double* resid, double* stats) {
    return guarded([&] {
        SQ_REQUIRE(n >= 0 && (n == 0 || (tubes && spheres && centers && e && B && U && Om)), SQBEM_ERR_ARG,
                   "sqbem_batch_solve: bad arguments");
        SQ_REQUIRE(precision >= SQBEM_FP64 && precision <= SQBEM_BF16_IR, SQBEM_ERR_ARG, "unknown precision");
        // the chunks' stages run eagerly or as captured graphs: the same kernels either way
        LibraryGemmOff own_kernels;
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
        const bool mx = mixed(precision);
        // chunks: per tube, largest instances first, greedily up to the budget (at least one)
        std::vector<std::vector<std::vector<Inst*>>> chunks(utube.size());
        size_t need = 0;
        for (int u = 0; u < (int)utube.size(); ++u) {
            std::vector<Inst*> mine;
            for (Inst& q : inst)
                if (q.tube == u) mine.push_back(&q);
            std::stable_sort(mine.begin(), mine.end(), [](const Inst* a, const Inst* b) { return a->Ns > b->Ns; });
            const int Nt = utube[u]->P();
            std::vector<Inst*> cur;
            for (Inst* q : mine) {
                cur.push_back(q);
                if (cur.size() > 1 && (double)chunk_bytes(Nt, cur, mx) > budget) {
                    cur.pop_back();
                    need = std::max(need, chunk_bytes(Nt, cur, mx));
                    chunks[u].push_back(cur);
                    cur.assign(1, q);
                }
            }
            if (!cur.empty()) {
                need = std::max(need, chunk_bytes(Nt, cur, mx));
                chunks[u].push_back(cur);
            }
        }
        // one workspace for every chunk of the call
        DBuf<unsigned char> ws;
        double t0 = now_s();
        ws.alloc(need);
        Arena ar;
        ar.base = ws.p;
        ar.cap = need;
        st.t_assembly += now_s() - t0;
        for (int u = 0; u < (int)utube.size(); ++u) {
            double t1 = now_s();
            SharedTube S;
            build_shared_tube(S, *utube[u], mu, precision);
            st.t_tubes += now_s() - t1;
            for (auto& ch : chunks[u]) {
                run_chunk(S, ch, precision, mu, ar, U, Om, iters, resid, st);
                ++st.chunks;
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
            stats[8] = st.pivoted;
            stats[9] = (double)need;
            stats[10] = st.graph_replays;
            stats[11] = st.build_graphs;
        }
    });
}

}  // extern "C"
