#!/usr/bin/env python3
"""Benchmark: cold squirmer-in-tube solves per second at 46k DOF on one MI355X.

One step = one cold solve of the whole hot path on the device: tube assembly + deflation
+ LU of T, coupling assembly, Schur complement, force/torque-free bordered solve for the
two squirming modes, with FP64 iterative refinement in the mixed-precision modes. The
host-side mesh generation of each solve is inside its timed step (it takes a few ms); all
matrices live in HBM.

Workload (46,470 DOF): tube R=1, L/R=4, n_theta=64, n_r=16, axially graded; icosphere level
4 (5120 panels), a/R=0.4, centre offset beta=0.25 (b = beta (R - a)), B2/B1 = +3, e = +z.

Beside it (detail.gradient): the adjoint gradient of (U_z, Omega_x) with respect to the
sphere centre, orientation and B2/B1 on the 9,750-DOF adjoint workload, in ms per solve,
checked against central differences; and (detail.large_field) the 57,222-DOF solve in
FP32 + FP64 refinement and in FP64 with the velocity on a masked 41 x 41 x 201 grid.

Beside those (detail.small_configs): the 5,430- and 13,254-DOF configurations in FP64 and
in the refined mode, next to the reference's own CPU solve of the same configurations
(cpu_baseline: the reference built from source, timed on this host's cores).

Prints one JSON line (rank 0). --gpus N > 1 exits with a message: this build runs on
one GPU.
"""
import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"dgemm": 78.6e12, "sgemm": 157.3e12, "gemm16": 2.5e15, "gemm16q": 2.5e15}
# the kernel each timing class measures (as rocprofv3 names it; {t} = true for BF16, false for FP16)
KERNEL_NAMES = {"dgemm": "gemm_nn_kernel<double>", "sgemm": "gemm_nn_kernel<float>", "gemm16": "gemm16_nn_kernel<{t}>",
                "gemm16q": "gemm16q_kernel (all candidates)"}
for _c, (_tm, _tn, _pf) in enumerate([(128, 128, 1), (128, 256, 1), (256, 128, 1), (256, 256, 1), (128, 128, 2),
                                       (128, 256, 2), (256, 128, 2)]):
    _k = "gemm16q_" + ["128x128", "128x256", "256x128", "256x256", "128x128_pf2", "128x256_pf2", "256x128_pf2"][_c]
    PEAKS[_k] = 2.5e15
    KERNEL_NAMES[_k] = "gemm16q_kernel<{t}, %d, %d, %d>" % (_tm, _tn, _pf)
for _name, (_tm, _tn, _kt, _ns) in [("256x256_dma32x4", (256, 256, 32, 4)), ("256x128_dma32x4", (256, 128, 32, 4)),
                                    ("128x128_dma64x4", (128, 128, 64, 4))]:
    PEAKS["gemm16q_" + _name] = 2.5e15
    KERNEL_NAMES["gemm16q_" + _name] = "gemm16d_kernel<{t}, %d, %d, %d, %d>" % (_tm, _tn, _kt, _ns)
for _a in range(4):
    PEAKS["gemm16q_hipblaslt%d" % _a] = 2.5e15
    KERNEL_NAMES["gemm16q_hipblaslt%d" % _a] = "hipBLASLt BF16 GEMM, heuristic algorithm %d" % _a
TIMED = list(KERNEL_NAMES)
MODES = {"fp64": 0, "fp32": 1, "fp16": 2, "bf16": 3}


def workload(args):
    R = 1.0
    a = args.a_over_R * R
    b = args.beta * (R - a)
    return dict(R=R, L=args.L_over_R * R, a=a, nth=args.n_theta, nr=args.n_r, lev=args.level,
                center=(b, 0.0, 0.0), alpha=args.alpha)


def cold_solve(pkg, W, prec, e):
    t0 = time.perf_counter()
    tube = pkg.Tube(W["R"], W["L"], W["a"], W["nth"], W["nr"], precision=prec)
    t1 = time.perf_counter()
    s = pkg.SwimmerSolver(tube, pkg.icosphere(W["a"], W["lev"], W["center"]), W["center"], precision=prec)
    t2 = time.perf_counter()
    sol = s.solve_free([s.slip(e, 1, 0), s.slip(e, 0, 1)])
    t3 = time.perf_counter()
    phases = {"tube_assembly_s": tube.t_assembly, "tube_factor_s": tube.t_factor,
              "coupling_assembly_s": s.t_assembly, "schur_s": s.t_schur, "solve_free_s": t3 - t2,
              "total_s": t3 - t0, "ir_iters": s.ir_iters, "ir_residual": s.ir_residual,
              "Nt": tube.Nt, "Ns": s.Ns}
    U = sol[0].U + W["alpha"] * sol[1].U
    Om = sol[0].Om + W["alpha"] * sol[1].Om
    fs = sol[0].f_sphere + W["alpha"] * sol[1].f_sphere
    del s, tube  # no reference cycles: both are freed here
    return U, Om, fs, phases


def orient(pitch, yaw):
    """e = (cos(yaw) sin(pitch), sin(yaw), cos(yaw) cos(pitch)), as the trajectory's initial e."""
    return np.array([math.cos(yaw) * math.sin(pitch), math.sin(yaw), math.cos(yaw) * math.cos(pitch)])


def grad_bench(pkg, reps=3):
    """Adjoint gradient of (U_z, Omega_x) with respect to the sphere centre, two orientation
    angles and B2/B1 on the adjoint workload (tube R=1, L/R=8, n_theta=24, n_r=6, graded;
    icosphere level 3, a/R=0.3, offset beta=0.2, tilt 10 degrees, B2/B1=-1; 9,750 DOF), FP64.
    Timed: the gradient call on a built solver (forward solve + transposed solves +
    derivative kernels), beside the build (assembly + Schur) and the forward solve alone.
    Checked: central differences (h = 1e-4) of the same FP64 solver."""
    R, L, a, nth, nr, lev = 1.0, 8.0, 0.3, 24, 6, 3
    xc = np.array([0.2 * (R - a), 0.0, 0.0])
    p = np.array([xc[0], xc[1], xc[2], math.radians(10.0), 0.0, -1.0])
    G = np.array([[0, 0, 1, 0, 0, 0], [0, 0, 0, 1, 0, 0]], dtype=np.float64)
    tube = pkg.Tube(R, L, a, nth, nr)

    def solver(c):
        return pkg.SwimmerSolver(tube, pkg.icosphere(a, lev, c), c)

    def value(q):
        sw = solver(q[:3])
        sol = sw.solve_free([sw.slip(orient(q[3], q[4]), 1.0, q[5])])[0]
        return np.array([sol.U[2], sol.Om[0]])

    pkg.sync()
    t0 = time.perf_counter()
    sw = solver(xc)
    pkg.sync()
    t_build = time.perf_counter() - t0
    e = orient(p[3], p[4])
    t_solve = t_grad = 1e9
    for _ in range(reps):
        t0 = time.perf_counter()
        sw.solve_free([sw.slip(e, 1.0, p[5])])
        t_solve = min(t_solve, time.perf_counter() - t0)
        t0 = time.perf_counter()
        r = sw.gradient(e, 1.0, p[5], G)
        t_grad = min(t_grad, time.perf_counter() - t0)
    # Jacobian in the parameters (x, y, z, pitch, yaw, alpha)
    de_dang = np.array([[math.cos(p[4]) * math.cos(p[3]), -math.sin(p[4]) * math.sin(p[3])],
                        [0.0, math.cos(p[4])],
                        [-math.cos(p[4]) * math.sin(p[3]), -math.sin(p[4]) * math.cos(p[3])]])
    jac = np.concatenate([r["dxc"], r["de"] @ de_dang, r["dB"][:, 1:2]], axis=1)
    h = 1e-4
    fd = np.zeros((2, 6))
    t0 = time.perf_counter()
    for k in range(6):
        dp = np.zeros(6)
        dp[k] = h
        fd[:, k] = (value(p + dp) - value(p - dp)) / (2 * h)
    t_fd = time.perf_counter() - t0
    rel = [float(np.max(np.abs(jac[i] - fd[i])) / np.max(np.abs(fd[i]))) for i in range(2)]
    return {"workload": "adjoint: tube L/R=8 n_theta=24 n_r=6 + icosphere level 3, a/R=0.3, beta=0.2, "
                        "tilt 10 deg, B2/B1=-1, FP64",
            "dof": 3 * (tube.Nt + sw.Ns) + 6,
            "grad_ms_per_solve": 1000 * t_grad, "forward_solve_ms": 1000 * t_solve,
            "build_ms": 1000 * t_build, "fd_12_solves_ms": 1000 * t_fd,
            "params": ["x_c", "y_c", "z_c", "pitch", "yaw", "B2/B1"], "functionals": ["U_z", "Omega_x"],
            "jacobian": jac.tolist(), "fd_rel_err_rows": rel}


def traj_grad_bench(pkg, n_steps, ckpt_every=50, h=1e-4):
    """Trajectory sensitivities on the adjoint workload's geometry (tube L/R=8, n_theta=24,
    n_r=6; icosphere level 3, a/R=0.3; B2/B1=-1): n_steps steps of dt=0.01 from beta0=0.2
    and a 10-degree tilt, d(final beta, final tilt)/d(beta0, tilt0) by the checkpointed
    adjoint (checkpoint every ckpt_every steps), against central differences (step h)."""
    def P(beta0, pitch0):
        return pkg.TrajectoryParams(alpha=-1.0, a_over_R=0.3, beta0=beta0, pitch0_deg=pitch0, yaw0_deg=0.0, dt=0.01,
                                    L_over_R=8.0, n_theta=24, sphere_level=3, n_r=6)

    t0 = time.perf_counter()
    r = pkg.trajectory_gradient(P(0.2, 10.0), n_steps, ckpt_every, verbose=True)
    t_adj = time.perf_counter() - t0

    def final(b0, p0):
        q = pkg.trajectory_gradient(P(b0, p0), n_steps, ckpt_every, adjoint=False, verbose=True)
        return np.array([q["beta_final"], q["tilt_final_deg"]])

    t0 = time.perf_counter()
    fd = np.stack([(final(0.2 + h, 10.0) - final(0.2 - h, 10.0)) / (2 * h),
                   (final(0.2, 10.0 + h) - final(0.2, 10.0 - h)) / (2 * h)], axis=1)
    t_fd = time.perf_counter() - t0
    adj = np.array([[r["dbeta_dbeta0"], r["dbeta_dpitch0"]], [r["dtilt_dbeta0"], r["dtilt_dpitch0"]]])
    return {"steps": n_steps, "ckpt_every": ckpt_every, "adjoint_s": t_adj, "fd_4_runs_s": t_fd,
            "evaluations": r["evaluations"], "final": [r["beta_final"], r["tilt_final_deg"]],
            "jacobian": adj.tolist(), "fd": fd.tolist(),
            "fd_rel_err_rows": [float(np.max(np.abs(adj[i] - fd[i])) / np.max(np.abs(fd[i]))) for i in range(2)]}


def field_bench(pkg):
    """The large case with field evaluation: tube R=1, L/R=8, n_theta=64, n_r=16 (graded),
    icosphere level 4, a/R=0.4, beta=0.25, B2/B1=+3 (57,222 DOF). Cold solve in FP32 + FP64
    refinement and in FP64, then the velocity on a 41 x 41 x 201 grid over [-R, R]^2 x
    [-L/2, L/2], masked to r_xy < 0.97 R and |x - x_c| >= 1.03 a (the reference's field-mode
    mask, src/main.cpp:117-129): 235,264 points x 19,072 panels, FP64, no matrix stored.
    Errors of the refined solve (U, Omega, tractions, field) against the FP64 solve."""
    R, L, a, nth, nr, lev, alpha = 1.0, 8.0, 0.4, 64, 16, 4, 3.0
    xc = np.array([0.25 * (R - a), 0.0, 0.0])
    e = np.array([0.0, 0.0, 1.0])
    g = np.linspace(-R, R, 41)
    z = np.linspace(-L / 2, L / 2, 201)
    X, Y, Z = np.meshgrid(g, g, z, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    keep = (np.hypot(pts[:, 0], pts[:, 1]) < 0.97 * R) & (np.linalg.norm(pts - xc, axis=1) >= 1.03 * a)
    pts = np.ascontiguousarray(pts[keep])
    out = {"workload": "large: tube L/R=8 n_theta=64 n_r=16 + icosphere level 4, a/R=0.4, beta=0.25, B2/B1=+3; "
                       "field on 41x41x201 masked", "points": int(len(pts))}
    res = {}
    for name, prec in (("fp32+ir", 1), ("fp64", 0)):
        pkg.sync()
        t0 = time.perf_counter()
        tube = pkg.Tube(R, L, a, nth, nr, precision=prec)
        S = pkg.icosphere(a, lev, xc)
        sw = pkg.SwimmerSolver(tube, S, xc, precision=prec)
        sol = sw.solve_free([sw.slip(e, 1, 0), sw.slip(e, 0, 1)])
        t1 = time.perf_counter()
        comb = pkg.SwimmerSolution(sol[0].U + alpha * sol[1].U, sol[0].Om + alpha * sol[1].Om,
                                   sol[0].f_sphere + alpha * sol[1].f_sphere, sol[0].f_tube + alpha * sol[1].f_tube)
        slip = sw.slip(e, 1, alpha)
        t2 = time.perf_counter()
        u = sw.velocity(comb, slip, pts)
        t3 = time.perf_counter()
        res[name] = (comb, u)
        out[name] = {"solve_s": t1 - t0, "field_s": t3 - t2, "ir_iters": sw.ir_iters}
        out["dof"] = 3 * (tube.Nt + sw.Ns) + 6
        npanels = tube.Nt + sw.Ns
        del sw, tube, S
        gc.collect()
    c32, u32 = res["fp32+ir"]
    c64, u64 = res["fp64"]
    rel = lambda x, y: float(np.linalg.norm(np.asarray(x) - np.asarray(y)) / np.linalg.norm(np.asarray(y)))
    out["rel_err_vs_fp64"] = {"U": rel(c32.U, c64.U), "Omega": rel(c32.Om, c64.Om),
                              "f_sphere": rel(c32.f_sphere, c64.f_sphere), "f_tube": rel(c32.f_tube, c64.f_tube),
                              "field": rel(u32, u64)}
    pairs = float(len(pts)) * npanels
    out["field_pairs"] = pairs
    # far-field count: 7 points x 72 flops per pair (near pairs use more points)
    out["field_tflops_far_count"] = 7 * 72 * pairs / out["fp64"]["field_s"] / 1e12
    out["speedup_fp32ir_vs_fp64"] = out["fp64"]["solve_s"] / out["fp32+ir"]["solve_s"]
    return out


def hbm_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter summary of this
    workload (FETCH_SIZE and WRITE_SIZE per dispatch, calibrated on a plain copy of known
    size), or None when the summary has no entry for it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_46k_bf16.json")
    try:
        tab = json.load(open(path))
    except (OSError, ValueError):
        return None
    e = tab.get("kernels", {}).get(kernel)
    if not e:
        return None
    return {"bytes_per_launch": e["hbm_bytes_per_launch"], "source": "profiles/hbm_traffic_46k_bf16.json"}


def calib_copy(nbytes):
    """One plain device-to-device copy of nbytes, as hipMemcpy2D of 1 MiB rows (the HIP
    runtime's copy kernel; a flat hipMemcpy of this size goes to a DMA engine, which the
    counters do not see), through the HIP runtime libsqbem is linked against (the loaded
    soname resolves to it). Not through torch: importing torch first would bind libsqbem's
    libhipblaslt.so.1 to torch's bundled copy, whose GEMM kernels differ from the system
    library's."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    row = 1 << 20
    nbytes = nbytes // row * row

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"calibration copy: {what} failed with hip error {rc}")

    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    ok(hip.hipMalloc(ctypes.byref(src), nbytes), "hipMalloc")
    try:
        ok(hip.hipMalloc(ctypes.byref(dst), nbytes), "hipMalloc")
        try:
            ok(hip.hipMemset(src, 0x3f, nbytes), "hipMemset")
            ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
            ok(hip.hipMemcpy2D(dst, row, src, row, row, nbytes // row, 3), "hipMemcpy2D")
            ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        finally:
            hip.hipFree(dst)
    finally:
        hip.hipFree(src)
    return nbytes


def batch_bench(pkg, n=256, seed=0, modes=(("fp32+ir", 1), ("fp16+ir", 2)), check=4):
    """Batched uneven solves: n configurations drawn from seed (a/R ~ U[0.1, 0.7], b/R ~
    U[0, 0.8 (1 - a/R)], B2/B1 ~ U[-5, 5], random e, sphere level {2, 3}, tube n_theta
    {16, 24, 32} with uniform n_z = 2 n_theta, L/R = 4, n_r = n_theta/4), solved in one
    sqbem_batch_solve call per mode (meshes built on the host beforehand). Checked: the
    smallest, the largest and a few random instances against single-instance FP64 solves."""
    params = pkg.draw_batch_params(n, seed=seed)
    inst = pkg.batch_instances(params)
    dof = np.array([3 * (q.tube.P() + q.sphere.P()) + 6 for q in inst])
    out = {"workload": f"{n} configurations from seed {seed}: a/R U[0.1,0.7], b/R U[0,0.8(1-a/R)], B2/B1 U[-5,5], "
                       "random e, level {2,3}, n_theta {16,24,32}, n_z=2 n_theta uniform, L/R=4, n_r=n_theta/4",
           "instances": n, "dof_min": int(dof.min()), "dof_max": int(dof.max()), "dof_total": int(dof.sum())}
    res = {}
    for name, prec in modes:
        pkg.sync()
        t0 = time.perf_counter()
        r = pkg.batch_solve(inst, precision=prec)
        t = time.perf_counter() - t0
        res[name] = r
        out[name] = {"s": t, "instances_per_s": n / t, "dof_per_s": float(dof.sum()) / t, **r.stats,
                     "iters_max": int(r.iters.max()), "resid_max": float(r.resid.max())}
    rng = np.random.default_rng(1)
    pick = [int(np.argmin(dof)), int(np.argmax(dof))] + [int(i) for i in rng.choice(n, size=max(check - 2, 0),
                                                                                    replace=False)]
    errs = {name: [] for name in res}
    for i in pick:
        q, pr = inst[i], params[i]
        t = pkg.Tube(1.0, pr["L"], pr["a"], pr["n_theta"], pr["n_r"], mesh=q.tube)
        s = pkg.SwimmerSolver(t, q.sphere, q.center)
        sol = s.solve_free([s.slip(q.e, 1.0, q.B2)])[0]
        ref = np.r_[sol.U, sol.Om]
        for name, r in res.items():
            errs[name].append(float(np.linalg.norm(np.r_[r.U[i], r.Om[i]] - ref) / np.linalg.norm(ref)))
        del s, t
        gc.collect()
    out["checked_instances"] = pick
    out["rel_err_vs_single_fp64"] = {name: max(v) for name, v in errs.items()}
    out["U_Omega_checksum"] = {name: float(np.sum(np.abs(r.U)) + np.sum(np.abs(r.Om))) for name, r in res.items()}
    return out


# the reference CPU path's own configurations beside the headline mesh (R = 1): the plain
# FP64 kinematics case and the mixed-precision single solve's configuration
SMALL_CFGS = {
    "cfg1": dict(R=1.0, L=4.0, a=0.3, nth=24, nr=6, lev=2, center=(0.0, 0.0, 0.0), pitch=0.0, alpha=0.0),
    "cfg2": dict(R=1.0, L=6.0, a=0.5, nth=32, nr=8, lev=3, center=(0.15, 0.0, 0.0), pitch=30.0, alpha=-3.0),
}
CPU_PROFILE = os.path.join(ROOT, "profiles", "cpu_reference_46k.json")


def host_cores():
    """Host threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    lim = os.environ.get("OMP_NUM_THREADS", "")
    if lim.isdigit() and int(lim) > 0:
        n = min(n, int(lim))
    return max(n, 1)


def reference_cold_solves(cfgs, progress=False):
    """Cold FP64 solves by the reference's own code (its solver headers built from source,
    oracle/_ref/libref_solver.so; OpenBLAS threads = host cores, assembly single-threaded
    as the reference's is), phase by phase."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_solver as RS

    cores = host_cores()
    RS.blas_threads(cores)
    runs = {}
    for name, c in cfgs.items():
        say = (lambda msg, n=name: print(f"[cpu reference {n}] {msg}", file=sys.stderr, flush=True)) if progress \
            else None
        runs[name] = RS.timed_cold_solve(c["R"], c["L"], c["a"], c["nth"], c["nr"], c["lev"], c["center"],
                                         orient(math.radians(c["pitch"]), 0.0), progress=say)
    return cores, RS.blas_config(), runs


def cpu_full(W):
    """The reference's full cold FP64 solve of the benchmark mesh and of the two small
    configurations, timed phase by phase on this host (minutes): the record behind the
    default run's CPU baseline (profiles/cpu_reference_46k.json)."""
    cfgs = dict(SMALL_CFGS)
    cfgs["46k"] = dict(R=W["R"], L=W["L"], a=W["a"], nth=W["nth"], nr=W["nr"], lev=W["lev"], center=W["center"],
                       pitch=0.0, alpha=W["alpha"])
    cores, blas, runs = reference_cold_solves(cfgs, progress=True)
    return {"what": "reference FP64 cold solves (src/swimmer.hpp, built from source), phase by phase",
            "cores": cores, "blas": blas, "runs": runs}


def cpu_baseline(W, quick=False):
    """The reference's CPU path timed on this host: one full cold FP64 solve of the workload
    (46,470 DOF; about 70 s on a 16-thread MI355X host, 12 s of it the two single-threaded
    assembly phases) and of the two small configurations, by the reference's own code, phase
    by phase. quick: only the small configurations are timed now, and the 46k figure is the
    recorded full solve (profiles/cpu_reference_46k.json, from bench.py --cpu-full) rescaled
    by how fast this host runs the 13,254-DOF solve compared with that recording."""
    cfgs = dict(SMALL_CFGS)
    if not quick:
        cfgs["46k"] = dict(R=W["R"], L=W["L"], a=W["a"], nth=W["nth"], nr=W["nr"], lev=W["lev"],
                           center=W["center"], pitch=0.0, alpha=W["alpha"])
    cores, blas, runs = reference_cold_solves(cfgs, progress=True)
    out = {"unit": "solves/s", "cores": cores, "kind": "reference", "blas": blas,
           "measured_now": {k: {kk: v for kk, v in r.items() if kk not in ("U", "Om")} for k, r in runs.items()}}
    if not quick:
        full = runs["46k"]
        out.update(value=1.0 / full["total_s"], seconds_per_solve=full["total_s"],
                   sample=(f"one full cold FP64 solve of the 46,470-DOF workload by the reference's own code "
                           f"(src/swimmer.hpp built from source, OpenBLAS on {cores} threads, assembly on one), "
                           f"timed now: {full['total_s']:.1f} s"))
        return out
    try:
        prof = json.load(open(CPU_PROFILE))
        full, rec2 = prof["runs"]["46k"], prof["runs"]["cfg2"]
    except (OSError, ValueError, KeyError):
        out.update(value=None, sample="no recorded 46k reference solve (run bench.py --cpu-full)")
        return out
    host = runs["cfg2"]["total_s"] / rec2["total_s"]
    secs = full["total_s"] * host
    out.update(value=1.0 / secs, seconds_per_solve=secs, host_speed_vs_record=host,
               recorded_46k_phases={k: v for k, v in full.items() if k.endswith("_s")},
               sample=(f"reference FP64 solves timed now: cfg1 {runs['cfg1']['total_s']:.2f} s, cfg2 "
                       f"{runs['cfg2']['total_s']:.2f} s; the 46,470-DOF solve recorded at {full['total_s']:.1f} s "
                       f"(bench.py --cpu-full), x {host:.3f} for this host"))
    return out


def small_cfg_gpu(pkg, prec, reps=3):
    """The two small configurations on the GPU, cold solves (best of reps), in FP64 and in
    the headline's refined mode, beside the reference's CPU times of the same solves."""
    out = {}
    for name, c in SMALL_CFGS.items():
        W = dict(R=c["R"], L=c["L"], a=c["a"], nth=c["nth"], nr=c["nr"], lev=c["lev"], center=c["center"],
                 alpha=c["alpha"])
        e = orient(math.radians(c["pitch"]), 0.0)
        res = {}
        for mode, p in (("fp64", 0), ("refined", prec)):
            best = None
            for _ in range(reps):
                U, Om, fs, ph = cold_solve(pkg, W, p, e)
                if best is None or ph["total_s"] < best[3]["total_s"]:
                    best = (U, Om, fs, ph)
            res[mode] = best
        ref = np.r_[res["fp64"][0], res["fp64"][1]]
        got = np.r_[res["refined"][0], res["refined"][1]]
        out[name] = {"dof": 3 * (res["fp64"][3]["Nt"] + res["fp64"][3]["Ns"]) + 6,
                     "fp64_s": res["fp64"][3]["total_s"], "refined_s": res["refined"][3]["total_s"],
                     "refined_ir_iters": res["refined"][3]["ir_iters"],
                     # (U, Omega) together: the on-axis case's Omega vanishes by symmetry
                     "rel_err_U_Omega": float(np.linalg.norm(got - ref) / np.linalg.norm(ref))}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=4,
                    help="untimed solves (the tile autotuner times each candidate once per shape class in these)")
    ap.add_argument("--mode", choices=list(MODES), default="bf16")
    ap.add_argument("--n_theta", type=int, default=64)
    ap.add_argument("--L_over_R", type=float, default=4.0)
    ap.add_argument("--n_r", type=int, default=16)
    ap.add_argument("--level", type=int, default=4)
    ap.add_argument("--a_over_R", type=float, default=0.4)
    ap.add_argument("--beta", type=float, default=0.25)
    ap.add_argument("--alpha", type=float, default=3.0)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 reference solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-grad", action="store_true", help="skip the adjoint-gradient timing")
    ap.add_argument("--no-field", action="store_true", help="skip the 57k-DOF solve + field grid")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-configuration batched solve")
    ap.add_argument("--no-small", action="store_true", help="skip the 5,430- and 13,254-DOF configurations")
    ap.add_argument("--batch-only", default="", metavar="MODES",
                    help="run only the 256-configuration batch in these modes (e.g. fp16,fp32) and print its line")
    ap.add_argument("--tune-table", default="shipped",
                    help="GEMM tile table to start from: 'shipped' (the package's tune_gfx950.txt), a path, "
                         "or 'none' (learn every class online in the warm-up)")
    ap.add_argument("--tune-save", default="", metavar="PATH", help="write the tile table after the run")
    ap.add_argument("--calib-copy-gib", type=float, default=0.0, metavar="GIB",
                    help="also run one plain device-to-device copy of GIB GiB (reference kernel for HBM counters)")
    ap.add_argument("--cpu-quick", action="store_true",
                    help="CPU baseline from the recorded full reference solve, rescaled by a 13,254-DOF solve "
                         "timed now (seconds instead of about a minute)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="only time the reference's full CPU solve of the workload and the two small "
                         "configurations (minutes; the record behind the CPU baseline) and print it")
    ap.add_argument("--traj-grad", type=int, default=0, metavar="STEPS",
                    help="also run the trajectory adjoint over STEPS steps (2000 for the full case; minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not implemented)"}))
        sys.exit(2)

    if args.cpu_full:
        print(json.dumps(cpu_full(workload(args))), flush=True)
        return

    import __graft_entry__ as ge

    pkg = ge.load_package()
    pkg.init(0)
    tune_loaded = 0
    if args.tune_table != "none":
        path = pkg.TUNE_TABLE if args.tune_table == "shipped" else args.tune_table
        if os.path.exists(path):
            tune_loaded = pkg.tune_load(path)
    if args.batch_only:
        names = {"fp64": ("fp64", 0), "fp32": ("fp32+ir", 1), "fp16": ("fp16+ir", 2), "bf16": ("bf16+ir", 3)}
        modes = tuple(names[m] for m in args.batch_only.split(","))
        b = batch_bench(pkg, modes=modes)
        first = b[modes[0][0]]
        print(json.dumps({"metric": "configurations/s, 256 batched uneven squirmer-in-tube solves", "value":
                          first["instances_per_s"], "unit": "configurations/s", "n_gpus": 1, "higher_is_better": True,
                          "dtype": modes[0][0], "data": "synthetic (seed 0)", "config": {"workload": b["workload"]},
                          "detail": b}), flush=True)
        return
    W = workload(args)
    e = np.array([0.0, 0.0, 1.0])
    prec = MODES[args.mode]

    for _ in range(args.warmup):
        cold_solve(pkg, W, prec, e)
    pkg.sync()
    # timed steps: no instrumentation inside (value, ms_per_step)
    t0 = time.perf_counter()
    runs = []
    for _ in range(args.steps):
        runs.append(cold_solve(pkg, W, prec, e))
    pkg.sync()
    elapsed = time.perf_counter() - t0
    # one more step of the same solve with HIP events around every GEMM launch on the library
    # stream: the kernel statistics and the roofline come from it (the events cost a few
    # microseconds of host time per launch, which the timed steps above do not carry)
    pkg.profile(True, classes=TIMED)
    t1 = time.perf_counter()
    cold_solve(pkg, W, prec, e)
    pkg.sync()
    profiled_ms = 1000.0 * (time.perf_counter() - t1)
    prof = {k: pkg.profile_get(k) for k in pkg.PROF_CLASSES}
    named = pkg.profile_named()
    pkg.profile(False)
    U, Om, fs, phases = runs[-1]
    ms_per_step = 1000.0 * elapsed / args.steps
    value = args.steps / elapsed
    dof = 3 * (phases["Nt"] + phases["Ns"]) + 6

    detail = {"dof": dof, "phases_last_step": phases, "precision": args.mode, "profiled_step_ms": profiled_ms,
              "tile_autotuner": dict(pkg.tune_info(), loaded_from_table=tune_loaded)}
    if not args.no_fp64 and prec != 0:
        U64, Om64, fs64, ph64 = cold_solve(pkg, W, 0, e)
        detail["fp64_solve_s"] = ph64["total_s"]
        detail["fp64_phases"] = ph64
        detail["speedup_vs_fp64"] = ph64["total_s"] / (ms_per_step / 1000.0)
        detail["rel_err_U"] = float(np.linalg.norm(U - U64) / np.linalg.norm(U64))
        detail["rel_err_Omega"] = float(np.linalg.norm(Om - Om64) / np.linalg.norm(Om64))
        detail["rel_err_f_sphere"] = float(np.linalg.norm(fs - fs64) / np.linalg.norm(fs64))
    detail["kernels"] = {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                             "tflops": (v["flops"] / (v["ms"] * 1e-3) / 1e12) if v["ms"] > 0 else 0.0}
                         for k, v in prof.items() if v["launches"]}
    detail["library_gemm_kernels"] = {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                                          "tflops": (v["flops"] / (v["ms"] * 1e-3) / 1e12) if v["ms"] > 0 else 0.0}
                                      for k, v in named.items()}

    # dominant kernel: the single kernel (one template instance, or one of the library GEMM's
    # kernels, as rocprofv3 lists them) with the most time in the profiled step
    tname = "true" if prec == 3 else "false"
    cands = {KERNEL_NAMES[k].format(t=tname): (prof[k], PEAKS[k]) for k in TIMED
             if k != "gemm16q" and not k.startswith("gemm16q_hipblaslt")}
    cands.update({name: (v, 2.5e15) for name, v in named.items()})
    kernel = max(cands, key=lambda k: cands[k][0]["ms"])
    kp, peak = cands[kernel]
    achieved = kp["flops"] / (kp["ms"] * 1e-3) if kp["ms"] > 0 else 0.0
    nl = max(kp["launches"], 1)
    roofline = {"bound": "mfma", "kernel": kernel, "achieved": achieved / 1e12,
                "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "launches": kp["launches"], "avg_launch_ms": kp["ms"] / nl, "flops_per_launch": kp["flops"] / nl,
                "algorithmic_bytes_per_launch": kp["bytes"] / nl}
    traffic = hbm_traffic(roofline["kernel"])
    if traffic:
        roofline["traffic"] = traffic["bytes_per_launch"]
        roofline["traffic_source"] = traffic["source"]

    if args.calib_copy_gib > 0:
        # a plain copy of known size: what FETCH_SIZE / WRITE_SIZE read for it calibrates
        # the same counters on the solver's kernels
        detail["calib_copy_bytes"] = calib_copy(int(args.calib_copy_gib * 2 ** 30))

    if not args.no_grad:
        try:
            detail["gradient"] = grad_bench(pkg)
        except Exception as ex:
            detail["gradient"] = {"error": str(ex)}

    if args.traj_grad > 0:
        detail["trajectory_gradient"] = traj_grad_bench(pkg, args.traj_grad)

    if not args.no_batch:
        try:
            detail["batch"] = batch_bench(pkg)
        except Exception as ex:
            detail["batch"] = {"error": str(ex)}

    if not args.no_field:
        try:
            detail["large_field"] = field_bench(pkg)
        except Exception as ex:
            detail["large_field"] = {"error": str(ex)}

    if not args.no_small:
        try:
            detail["small_configs"] = small_cfg_gpu(pkg, prec)
        except Exception as ex:
            detail["small_configs"] = {"error": str(ex)}

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(W, quick=args.cpu_quick)
            for name, g in detail.get("small_configs", {}).items():
                if isinstance(g, dict) and name in cpu.get("measured_now", {}):
                    g["cpu_reference_s"] = cpu["measured_now"][name]["total_s"]
        except Exception as ex:  # the baseline must not hide the GPU result
            cpu = {"error": str(ex)}

    line = {
        "metric": "solves/s at 46k DOF, mixed+IR vs FP64 GPU at <=1e-8 rel U,Omega err; grad ms/solve",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f64", 1: "f32+f64ir", 2: "f16mfma+f32+f64ir", 3: "bf16mfma+f32+f64ir"}[prec],
        "data": "synthetic (deterministic meshes)",
        "config": {"workload": (f"cold squirmer-in-tube solve, graded tube n_theta={args.n_theta} "
                                f"L/R={args.L_over_R:g} n_r={args.n_r} + icosphere level {args.level}, "
                                f"a/R={args.a_over_R:g}, beta={args.beta:g}, B2/B1={args.alpha:+g}"),
                   "dof": dof, "precision": args.mode},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": detail,
    }
    if args.tune_save:
        pkg.tune_save(args.tune_save)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
