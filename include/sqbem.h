/*
 * sqbem.h -- C ABI of libsqbem, the MI355X (gfx950) squirmer-in-tube BEM library.
 *
 * Plain C types only: host or device pointers, sizes, status codes. Every function
 * returns 0 on success or a nonzero status, with a message from sqbem_last_error().
 * Messages reuse the reference's error texts ("dgetrf failed", "dgesv failed",
 * "gemm: dimension mismatch") so the C++ headers can rethrow them unchanged.
 *
 * Matrices are column-major with LAPACK leading dimensions. Device pointers are HIP
 * device memory of the current device; device work is ordered on the library stream
 * (sqbem_set_stream) and host-returning calls synchronise it.
 *
 * Reference citations are file:line in byjesh1998/squirmer-bem-3dtube.
 */
#ifndef SQBEM_H
#define SQBEM_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SQBEM_OK 0
#define SQBEM_ERR_ARG 1      /* bad argument or dimension mismatch */
#define SQBEM_ERR_HIP 2      /* HIP runtime error (no device, launch failure, ...) */
#define SQBEM_ERR_SINGULAR 3 /* zero pivot: LAPACK info > 0 */
#define SQBEM_ERR_NOMEM 4    /* device allocation failed */
#define SQBEM_ERR_STATE 5    /* object used in a state that does not allow the call */
#define SQBEM_ERR_CONVERGE 6 /* iterative refinement did not reach SQBEM_IR_TOL */

/* Relative residual |b - A x| / |b| (2-norm per right-hand side, FP64 operator) that every
 * refined solve reaches, or fails with SQBEM_ERR_CONVERGE after 40 steps. */
#define SQBEM_IR_TOL 1e-12

/* arithmetic of the dense factorisations (operator assembly is FP64 in every mode) */
#define SQBEM_FP64 0    /* FP64 everywhere: the drop-in path */
#define SQBEM_FP32_IR 1 /* FP32 factors (FP32 MFMA updates), FP64 iterative refinement */
#define SQBEM_FP16_IR 2 /* FP32 factors, FP16 MFMA trailing updates with per-block scaling */
#define SQBEM_BF16_IR 3 /* FP32 factors, BF16 MFMA trailing updates */

/* ------------------------------------------------------------------ runtime */
const char* sqbem_last_error(void);
int sqbem_init(int device);
int sqbem_set_stream(void* hip_stream); /* NULL: the library's own stream */
int sqbem_sync(void);
int sqbem_device_info(char* name, int name_len, int* cu_count, double* hbm_gb);
/* Device memory the library releases is kept for reuse by later requests of the same size
 * (repeated solves then allocate nothing); sqbem_trim returns all of it to the driver and
 * reports how much (GB). An allocation that fails empties the cache and retries first. */
int sqbem_trim(double* released_gb);
/* Kernel timing (no reference counterpart; measurement only). sqbem_profile(mask) resets
 * the totals and records HIP events around every launch, on the library stream, of the
 * classes in mask (bit 1 << class; -1 = all classes; 0 = off); sqbem_profile_get returns
 * launches, algorithmic flops and bytes, and summed kernel time (ms) of one class. */
#define SQBEM_PROF_DGEMM 0
#define SQBEM_PROF_SGEMM 1
#define SQBEM_PROF_GEMM16 2
#define SQBEM_PROF_PANEL 3
#define SQBEM_PROF_TRSM 4
#define SQBEM_PROF_ASSEMBLE 5
#define SQBEM_PROF_GEMV 6
#define SQBEM_PROF_LASWP 7
#define SQBEM_PROF_GEMM16Q 8 /* 16-bit GEMM on pre-converted, padded operands (large shapes) */
#define SQBEM_PROF_TRSV 9    /* triangular solves with at most 8 right-hand sides (one launch each) */
/* gemm16q launches by autotuner candidate c = 0..13 (each also counted in SQBEM_PROF_GEMM16Q):
 * workgroup tiles 128x128, 128x256, 256x128, 256x256 with loads staged one k tile ahead
 * (c = 0..3), 128x128, 128x256, 256x128 staged two k tiles ahead (c = 4..6), and operand
 * tiles copied into LDS by DMA: 256x256 (7) and 256x128 (8) with four 32-deep stages,
 * 128x128 with four 64-deep stages (9), and hipBLASLt's BF16 GEMM on the same operands
 * with the first four algorithms of its heuristic for the shape (10-13; BF16 only, FP16 runs
 * candidate 3 in their place) */
#define SQBEM_PROF_G16Q(c) (10 + (c))
int sqbem_profile(int mask);
int sqbem_profile_get(int kernel_class, long* launches, double* flops, double* bytes, double* ms);
/* the same totals per kernel name for launches whose kernel the library names (the library
 * GEMM candidate, whose kernels hipBLASLt chooses per shape): one line per kernel,
 * "name\tlaunches\tflops\tbytes\tms\n", into buf (len bytes) */
int sqbem_profile_named(char* buf, int len);

/* ------------------------------------------------------------------ meshes (host)
 * Generators with the reference's node/panel order, src/utils/mesh.hpp:45-247. */
typedef struct sqbem_mesh sqbem_mesh;
/* swimmer_tube_mesh, src/utils/mesh.hpp:194-197 */
sqbem_mesh* sqbem_mesh_swimmer_tube(double R, double L, double a, int n_theta, int n_r);
/* tube_mesh over explicit axial nodes, src/utils/mesh.hpp:149-190 */
sqbem_mesh* sqbem_mesh_tube(double R, const double* z, int nz_nodes, int n_theta, int n_r);
/* tube_mesh(R, uniform_z(0, L, n_z), ...) as DuctFlow builds it, src/duct_flow.hpp:32-36 */
sqbem_mesh* sqbem_mesh_duct(double R, double L, int n_theta, int n_z, int n_r);
/* icosphere, src/utils/mesh.hpp:208-245 */
sqbem_mesh* sqbem_mesh_icosphere(double a, int level, double cx, double cy, double cz);
/* Mesh{nodes, tris, tags} + geometry(), src/utils/mesh.hpp:50-72 */
sqbem_mesh* sqbem_mesh_from_arrays(const double* nodes, int n_nodes, const int* tris,
                                   const int* tags, int n_panels);
/* merge, src/utils/mesh.hpp:95-103 */
sqbem_mesh* sqbem_mesh_merge(const sqbem_mesh* a, const sqbem_mesh* b);
/* Mesh::translate, src/utils/mesh.hpp:82-85 */
int sqbem_mesh_translate(sqbem_mesh* m, double dx, double dy, double dz);
int sqbem_mesh_sizes(const sqbem_mesh* m, int* n_nodes, int* n_panels);
/* copies out nodes[3nn], tris[3np], tags[np], V[9np], normal[3np], centroid[3np],
 * area[np], h[np]; any pointer may be NULL */
int sqbem_mesh_get(const sqbem_mesh* m, double* nodes, int* tris, int* tags, double* V,
                   double* normal, double* centroid, double* area, double* h);
void sqbem_mesh_free(sqbem_mesh* m);

/* ------------------------------------------------------------------ dense algebra
 * Device pointers; replace the OpenBLAS calls of src/utils/linalg.hpp:17-26. */
/* C = alpha A B + beta C, dgemm_('N','N'), src/utils/linalg.hpp:42-48 */
int sqbem_dgemm(int m, int n, int k, double alpha, const double* A, int lda, const double* B,
                int ldb, double beta, double* C, int ldc);
/* FP32 counterpart (FP32 MFMA) used by the mixed-precision factorisations */
int sqbem_sgemm(int m, int n, int k, float alpha, const float* A, int lda, const float* B,
                int ldb, float beta, float* C, int ldc);
/* FP32 in/out with 16-bit MFMA products (kind SQBEM_FP16_IR or SQBEM_BF16_IR) and FP32
 * accumulation; FP16 operands are scaled per row of A and per column of B by powers of
 * two, undone exactly in the epilogue */
int sqbem_gemm16(int kind, int m, int n, int k, float alpha, const float* A, int lda,
                 const float* B, int ldb, float beta, float* C, int ldc);
/* No counterpart in the reference: sqbem_gemm16 that allocates nothing, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Its workspace (converted operands, FP16 scales) is a
 * fixed scratch region of SQBEM_GEMM16_SCRATCH_BYTES per device inside the library, set up
 * by sqbem_init (which must have been called; this call never initialises). Products whose
 * workspace (sqbem_gemm16_noalloc_bytes) exceeds it are rejected with SQBEM_ERR_ARG, e.g.
 * m = n = k = 4096 needs 64 MiB + scales. Every call on a device shares that device's
 * region: calls that overlap in time on one device (two streams, two host threads) race.
 * The tile is the autotuner's decided one for the shape class, else the fixed rule; the
 * call neither explores nor records timing events. Same results as sqbem_gemm16. */
#define SQBEM_GEMM16_SCRATCH_BYTES (96u << 20)
int sqbem_gemm16_noalloc(int kind, int m, int n, int k, float alpha, const float* A, int lda,
                         const float* B, int ldb, float beta, float* C, int ldc);
unsigned long sqbem_gemm16_noalloc_bytes(int kind, int m, int n, int k);
/* P A = L U with partial pivoting, dgetrf_, src/utils/linalg.hpp:65-71. ipiv is a device
 * array of n 1-based pivot rows; *info (host) is 0 or the first zero-pivot column. */
int sqbem_dgetrf(int n, double* A, int lda, int* ipiv, int* info);
/* B <- A^-1 B from the dgetrf factors, dgetrs_('N'), src/utils/linalg.hpp:72-77 */
int sqbem_dgetrs(int n, int nrhs, const double* LU, int lda, const int* ipiv, double* B, int ldb);
/* B <- A^-T B from the dgetrf factors (dgetrs_('T')); used by the adjoint solves */
int sqbem_dgetrs_trans(int n, int nrhs, const double* LU, int lda, const int* ipiv, double* B,
                       int ldb);
/* solve A X = B overwriting A with its factors, dgesv_, src/utils/linalg.hpp:53-58 */
int sqbem_dgesv(int n, int nrhs, double* A, int lda, int* ipiv, double* B, int ldb, int* info);
/* FP32 factorisation; kind selects the trailing-update MFMA (SQBEM_FP32_IR/FP16_IR/BF16_IR) */
int sqbem_sgetrf(int kind, int n, float* A, int lda, int* ipiv, int* info);
int sqbem_sgetrs(int n, int nrhs, const float* LU, int lda, const int* ipiv, float* B, int ldb);
/* How the calling thread's most recent LU factorisation went (a diagnostic, no reference
 * counterpart; dgetrf_ always pivots): 0 partial pivoting throughout (orders below 512),
 * 1 factored without row interchanges, every multiplier verified |l| <= 1 (exactly the
 * factors partial pivoting gives), 2 that attempt was rejected and the matrix refactored
 * with partial pivoting. A swimmer's construction factors the tube, then the bordered
 * Schur matrix: afterwards this tells about the bordered matrix. */
int sqbem_getrf_path(void);
/* Cooperative launches of the pivoted LU's panel kernel (a test switch, no reference
 * counterpart; dgetrf_ has no such choice): 1 (the default) uses them while the runtime
 * accepts them; 0 makes every pivoted panel take the barrier-free path that the library
 * otherwise uses only once the runtime has refused a cooperative launch. */
int sqbem_coop(int enable);

/* Tile autotuner of the 16-bit GEMMs (no reference counterpart): mode 1 (the default)
 * learns the fastest of fourteen candidates (the SQBEM_PROF_G16Q list: workgroup tile and
 * load staging depth) per shape class (operand type, binary exponents of m, n, k) by
 * timing each candidate once on live calls; mode 0 uses the fixed rule (256 x 256 tiles
 * when they fill the device, else 128 x 128). Either call clears what was learned.
 * sqbem_tune_info: number of shape classes seen, how many are decided, and counts[14] of
 * decided classes per candidate; any pointer may be NULL. */
int sqbem_tune(int mode);
int sqbem_tune_info(int* classes, int* decided, int* counts);
/* Decided classes to a text file, and back: a loaded class uses its saved tile from the
 * first call (the package ships tune_gfx950.txt, learned on an MI355X); *classes (may be
 * NULL) receives how many classes were read. */
int sqbem_tune_save(const char* path);
int sqbem_tune_load(const char* path, int* classes);

/* Host-array forms of the same calls, for the C++ linalg.hpp API (Mat lives in host
 * memory): the library stages the operands through HBM. */
/* gemm(alpha, A, B, beta, C), src/utils/linalg.hpp:42-48 */
int sqbem_host_dgemm(int m, int n, int k, double alpha, const double* A, int lda, const double* B,
                     int ldb, double beta, double* C, int ldc);
/* solve_dense(A, B): B <- A^-1 B, A unchanged, src/utils/linalg.hpp:53-58 */
int sqbem_host_dgesv(int n, int nrhs, const double* A, int lda, double* B, int ldb);
/* LUFactor (src/utils/linalg.hpp:62-78) with the factors kept in HBM */
typedef struct sqbem_lu sqbem_lu;
/* LUFactor::factor: copies A (host) to the device and factors it; *info as dgetrf_ */
sqbem_lu* sqbem_lu_create(int n, const double* A, int lda, int* info);
/* LUFactor::solve: B (host, n x nrhs) <- A^-1 B */
int sqbem_lu_solve(const sqbem_lu* lu, int nrhs, double* B, int ldb);
/* copies the packed L\U factors (n x n, column-major) and 1-based pivots to the host */
int sqbem_lu_get(const sqbem_lu* lu, double* LU, int* piv);
int sqbem_lu_size(const sqbem_lu* lu);
void sqbem_lu_free(sqbem_lu* lu);

/* ------------------------------------------------------------------ panel integrals
 * Device copy of a mesh's panel geometry, read by the assembly kernels. */
typedef struct sqbem_dmesh sqbem_dmesh;
sqbem_dmesh* sqbem_dmesh_upload(const sqbem_mesh* m);
void sqbem_dmesh_free(sqbem_dmesh* d);
/* The loops over panel_integrals (src/utils/kernels.hpp:67-82) of src/swimmer.hpp:52-57,
 * 188-221 and src/duct_flow.hpp:90-102 as one device pass. For target i (x0: device,
 * 3*nt doubles, xyz interleaved) and panel p0+p of src:
 *   SL(3i+a, 3p+b) = SL_ab / sl_div,  DL(3i+a, 3p+b) = -DL_ab / dl_div,
 *   dlsum[9i+k] += DL_k (raw stresslet integral, summed over the panels).
 * self_src (device, nt ints, or NULL) names the panel containing target i (Duffy rule);
 * SL, DL, dlsum are device pointers or NULL. */
int sqbem_assemble(const sqbem_dmesh* src, int p0, int np, const double* x0, int nt,
                   const int* self_src, double sl_div, double* SL, long ld_sl, double dl_div,
                   double* DL, long ld_dl, double* dlsum);
/* one pair on the device, host in/out: panel_integrals, src/utils/kernels.hpp:67-82 */
int sqbem_panel_integrals(const sqbem_mesh* m, int p, const double* x0, int self, double* SL,
                          double* DL);

/* ------------------------------------------------------------------ solvers */
typedef struct sqbem_tube sqbem_tube;
typedef struct sqbem_swimmer sqbem_swimmer;
typedef struct sqbem_duct sqbem_duct;

/* Tube::Tube + Tube::factor (src/swimmer.hpp:40-60): assembles T = SL/(8 pi mu) with the
 * rank-one deflation (src/swimmer.hpp:64-74) on the device and factors it there. Returns
 * once the factorisation is issued: the device goes on factoring while the caller carries
 * on (a swimmer created on the tube meanwhile assembles its coupling blocks beside it), and
 * every use of the factors (a swimmer on the tube, the factor view's solve and get,
 * t_factor below) first waits for it to finish. A singular tube matrix is reported there
 * ("dgetrf failed"), not here. */
sqbem_tube* sqbem_tube_create(const sqbem_mesh* mesh, double mu, int precision);
/* Nt, and the seconds spent assembling and factoring; asking for t_factor waits for the
 * factorisation (any pointer may be NULL). */
int sqbem_tube_info(const sqbem_tube* t, int* Nt, double* t_assembly, double* t_factor);
/* the tube's factor (Tube::lu, src/swimmer.hpp:38): a view valid while the tube lives. In
 * the refined modes sqbem_lu_solve corrects the FP32 factors' solution by iterative
 * refinement against the FP64 operator to SQBEM_IR_TOL, so it returns FP64 solutions in
 * every mode; sqbem_lu_get returns the FP32 factors widened to FP64. */
const sqbem_lu* sqbem_tube_lu(const sqbem_tube* t);
void sqbem_tube_free(sqbem_tube* t);

/* SwimmerSolver::SwimmerSolver + build (src/swimmer.hpp:92-97, 180-251): coupling
 * assembly, rigid-body map, Schur complement against the tube factors. tube may be NULL
 * (unbounded fluid). The sphere mesh is copied. */
sqbem_swimmer* sqbem_swimmer_create(const sqbem_tube* tube, const sqbem_mesh* sphere,
                                    const double* center, double mu, int precision);
/* Sphere template (an addition; the reference reassembles these per solver): the
 * translation-invariant sphere-sphere blocks of SwimmerSolver::build (src/swimmer.hpp:
 * 209-222: E, the double-layer part of AuBs and the sphere panels' stresslet row sums) of a
 * sphere mesh centred at center, assembled once on the device. sqbem_swimmer_create_moved
 * builds the solver of that sphere translated by shift[3] (centre center + shift, mesh
 * translated as Mesh::translate does, src/utils/mesh.hpp:82-85), copying those blocks and
 * assembling only the tube-sphere couplings; the result equals sqbem_swimmer_create on the
 * translated mesh to rounding (the sphere blocks are computed in the template's position).
 * The template must outlive the solvers made from it and share the tube's viscosity. */
typedef struct sqbem_sphere sqbem_sphere;
sqbem_sphere* sqbem_sphere_create(const sqbem_mesh* sphere, const double* center, double mu);
void sqbem_sphere_free(sqbem_sphere* t);
sqbem_swimmer* sqbem_swimmer_create_moved(const sqbem_tube* tube, const sqbem_sphere* sphere,
                                          const double* shift, int precision);
/* solve_free (src/swimmer.hpp:107-139), one solution per slip field. slips: host,
 * nslip*Ns*3; outputs host: U[3*nslip], Om[3*nslip], f_sphere[nslip*Ns*3],
 * f_tube[nslip*Nt*3] (outputs may be NULL) */
int sqbem_swimmer_solve_free(sqbem_swimmer* s, int nslip, const double* slips, double* U,
                             double* Om, double* f_sphere, double* f_tube);
/* solve_prescribed (src/swimmer.hpp:142-153): towed rigid sphere, no slip */
int sqbem_swimmer_solve_prescribed(sqbem_swimmer* s, const double* U, const double* Om,
                                   double* f_sphere, double* f_tube);
/* t_assembly / t_schur (src/swimmer.hpp:90, 181, 235, 250), seconds; iterations and final
 * relative residual of the last refined solve (0 and 0 in FP64 mode) */
int sqbem_swimmer_times(const sqbem_swimmer* s, double* t_assembly, double* t_schur,
                        int* ir_iters, double* ir_residual);
void sqbem_swimmer_free(sqbem_swimmer* s);
/* Adjoint gradients (no counterpart in the reference, which has no derivatives): for the
 * free-swimming solution with slip squirmer_slip(c_p, x_c, e, B1, B2) (src/utils/kernels.hpp:
 * 107-111) and nfun (1..8) functionals J_j = g[6j..6j+5] . (U, Omega), returns U, Om (host,
 * 3 each), dxc[3*nfun] = dJ/dx_c (sphere centre; the tube stays), de[3*nfun] = dJ/de,
 * dB[2*nfun] = dJ/d(B1, B2), du[nfun*Ns*3] = dJ/d(slip field). Transposed solves reuse the
 * tube's LU and the bordered matrix's LU; the centre derivative contracts the translation
 * derivative of the tube-sphere coupling kernels with the adjoint state. FP64 swimmers
 * only; any output may be NULL. */
int sqbem_swimmer_gradient(sqbem_swimmer* s, const double* e, double B1, double B2, int nfun,
                           const double* g, double* U, double* Om, double* dxc, double* de,
                           double* dB, double* du);

/* Batched uneven solves (no single counterpart in the reference, whose kinematics sweeps
 * build one SwimmerSolver per configuration and reuse the Tube, src/main.cpp:82-100):
 * n independent free-swimming problems, instance i on tube mesh tubes[i] with sphere mesh
 * spheres[i] (already placed at its centre) centred at centers[3i..3i+2], squirming along
 * e[3i..3i+2] with modes B[2i], B[2i+1] (slip squirmer_slip at the sphere centroids,
 * src/utils/kernels.hpp:107-111; solve_free, src/swimmer.hpp:107-139). Tube meshes with
 * identical contents share one assembly and factorisation. Instances on one tube are
 * processed in chunks of at most mem_budget_gb of device memory (<= 0: 64 GB): one
 * persistent assembly launch for all their coupling blocks, one triangular solve through
 * the shared tube factors for all their Schur right-hand sides, and (refined modes) a
 * refinement of all their full systems together. Outputs (host): U[3n], Om[3n]; iters[n]
 * and resid[n] (refinement steps and final relative residual; 0 in FP64 mode; may be
 * NULL); stats[12] (may be NULL): seconds in tube assembly + factorisation (+ inverse),
 * coupling assembly (host preparation and device time), Schur + bordered factors, solves;
 * unique tubes, chunks, most refinement steps, worst final residual, bordered matrices that
 * needed pivoting, bytes of the chunk workspace, refinement steps replayed from a captured
 * HIP graph, build stages run as captured HIP graphs. Every instance needs a tube
 * (tubes[i] != NULL). */
/* Batched-solve options (no reference counterpart): use_graph != 0 (the default) runs each
 * chunk's build (coupling assembly and right-hand sides; Schur complement and bordered
 * factors: two stages, no host round trip inside either) as two captured HIP graphs, and
 * captures its refinement step once and replays it every step; a stage whose capture
 * records anything but kernel launches runs its launches directly instead. 0 launches the
 * same kernels one by one. The results are bit-identical either way. */
int sqbem_batch_options(int use_graph);
int sqbem_batch_solve(int n, const sqbem_mesh* const* tubes, const sqbem_mesh* const* spheres,
                      const double* centers, const double* e, const double* B, double mu, int precision,
                      double mem_budget_gb, double* U, double* Om, int* iters, double* resid,
                      double* stats);

/* velocity_at (src/utils/field.hpp:18-34) on the device: u, f host [3*np]; pts host
 * [3*npts]; out host [3*npts] */
int sqbem_velocity_at(const sqbem_mesh* m, const double* u, const double* f, double mu,
                      const double* pts, int npts, double* out);

/* DuctFlow (src/duct_flow.hpp:32-36, 82-108): mixed-BC assembly and factorisation */
sqbem_duct* sqbem_duct_create(const sqbem_mesh* mesh, double mu);
/* DuctFlow::solve (src/duct_flow.hpp:41-63): u, f host [3*np], *Q host */
int sqbem_duct_solve(const sqbem_duct* d, double p_in, double p_out, double* u, double* f,
                     double* Q);
void sqbem_duct_free(sqbem_duct* d);

/* run_trajectory (src/trajectory.hpp:23-151): RK4 start-up + Adams-Bashforth 4 with the
 * tube factors resident in HBM; writes <out_prefix>.csv (and <out_prefix>.ckpt while
 * paused), status receives "finished", "wall contact" or "paused". n_r is the rings per
 * cap (the reference's Tube default is 4); precision as for sqbem_tube_create. */
int sqbem_trajectory_run(double alpha, double a_over_R, double beta0, double pitch0_deg,
                         double yaw0_deg, double dt, double t_max, double L_over_R, int n_theta,
                         int sphere_level, double beta_stop, double budget_s, int n_r, int precision,
                         const char* out_prefix, int verbose, char* status, int status_len);

/* Trajectory sensitivities (no counterpart in the reference): integrates n_steps steps of
 * run_trajectory's scheme (src/trajectory.hpp:106-142; RK4 start-up, Adams-Bashforth 4,
 * renormalised e) from beta0 and the initial tilt pitch0_deg (yaw0_deg = 0 as given), then
 * differentiates beta_final = |(x, y)| / (R - a) and tilt_final_deg (angle of e from the
 * tube axis, degrees) with respect to beta0 and pitch0_deg by a discrete adjoint of the
 * integrator: checkpoints every ckpt_every steps, each segment recomputed forward with its
 * solvers kept in device memory, then swept backward (one adjoint solve per evaluation).
 * out[0..1] = beta_final, tilt_final_deg; out[2..5] = d beta_f/d beta0, d beta_f/d pitch0,
 * d tilt_f/d beta0, d tilt_f/d pitch0; out[6] = forward evaluations including recomputation.
 * adjoint = 0 runs the forward integration only (out[2..5] = 0); verbose != 0 prints
 * progress to stderr. FP64 only. */
int sqbem_trajectory_gradient(double alpha, double a_over_R, double beta0, double pitch0_deg,
                              double yaw0_deg, double dt, int n_steps, int ckpt_every,
                              double L_over_R, int n_theta, int sphere_level, int n_r, int adjoint,
                              int verbose, double* out);

#ifdef __cplusplus
}
#endif
#endif /* SQBEM_H */
