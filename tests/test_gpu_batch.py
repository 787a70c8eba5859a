"""Batched uneven solves (sqbem_batch_solve) against the CPU oracle.

Five configurations on three distinct tubes: three on one uniform tube (two of them given
as separately generated but identical meshes, so the content-based sharing is exercised),
one on a second uniform tube and one on the reference's graded tube; spheres of levels 1
and 2, off-axis centres, tilted squirming directions, B2/B1 of both signs.

Bar: FP64 U and Omega of every instance within 1e-10 of the oracle's solve_free of the
same configuration (norm-wise over (U, Omega)); the refined modes within 1e-8 of FP64 with
a final relative residual <= 1e-12; chunking (one instance per chunk) changes nothing
beyond rounding (1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R, L = 1.0, 4.0


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


SPECS = [
    # (tube kind, n_theta, n_z or None, n_r, a, level, centre, e, B2/B1)
    ("uniform", 12, 16, 3, 0.30, 1, (0.10, 0.0, 0.0), (0.0, 0.0, 1.0), -1.0),
    ("uniform", 12, 16, 3, 0.40, 2, (0.20, 0.05, 0.10), unit((0.3, -0.2, 0.9)), 3.0),
    ("uniform", 12, 16, 3, 0.25, 1, (0.0, 0.20, 0.0), unit((0.5, 0.0, 0.5)), 0.0),
    ("uniform", 10, 20, 2, 0.30, 1, (0.0, 0.0, 0.30), unit((0.0, 1.0, 0.2)), 2.0),
    ("graded", 12, None, 3, 0.35, 1, (0.15, 0.0, 0.0), unit((0.2, 0.1, -1.0)), -2.0),
]


def uniform_z(L, nz):
    return np.array([-L / 2 + L * k / nz for k in range(nz + 1)])


def tube_meshes(M):
    """Tube mesh per spec with module M (binding or oracle); the first two uniform specs
    share one mesh object, the third gets its own identical copy."""
    out, first = [], None
    for j, (kind, nth, nz, nr, a, lev, c, e, al) in enumerate(SPECS):
        if kind == "graded":
            out.append(M.swimmer_tube_mesh(R, L, a, nth, nr))
        elif j == 1:
            out.append(first)
        else:
            m = M.tube_mesh(R, uniform_z(L, nz), nth, nr)
            first = m if j == 0 else first
            out.append(m)
    return out


def gpu_instances(sq):
    return [sq.BatchInstance(t, sq.icosphere(a, lev, c), np.array(c), np.asarray(e), 1.0, al)
            for t, (kind, nth, nz, nr, a, lev, c, e, al) in zip(tube_meshes(sq), SPECS)]


@pytest.fixture(scope="module")
def oracle_results(O):
    res = []
    for tm, (kind, nth, nz, nr, a, lev, c, e, al) in zip(tube_meshes(O), SPECS):
        t = O.Tube(R, L, a, nth, nr, mesh=tm)
        s = O.SwimmerSolver(t, O.icosphere(a, lev, c), c)
        sol = s.solve_free([s.slip(np.asarray(e), 1.0, al)])[0]
        res.append(np.r_[sol.U, sol.Om])
    return res


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def fp64_batch(sq):
    return sq.batch_solve(gpu_instances(sq), precision=sq.FP64)


def test_batch_fp64_matches_oracle(sq, fp64_batch, oracle_results):
    r = fp64_batch
    assert r.stats["unique_tubes"] == 3
    for i, ref in enumerate(oracle_results):
        err = rel(np.r_[r.U[i], r.Om[i]], ref)
        print(f"instance {i}: rel err {err:.2e}")
        assert err <= 1e-10, (i, err)
    assert np.all(r.iters == 0)


def test_batch_matches_single_instance_solver(sq, fp64_batch):
    """Instance 1 through the single-instance SwimmerSolver on the same meshes."""
    kind, nth, nz, nr, a, lev, c, e, al = SPECS[1]
    tm = sq.tube_mesh(R, uniform_z(L, nz), nth, nr)
    t = sq.Tube(R, L, a, nth, nr, mesh=tm)
    s = sq.SwimmerSolver(t, sq.icosphere(a, lev, c), c)
    sol = s.solve_free([s.slip(np.asarray(e), 1.0, al)])[0]
    assert rel(np.r_[fp64_batch.U[1], fp64_batch.Om[1]], np.r_[sol.U, sol.Om]) <= 1e-12


@pytest.mark.parametrize("prec", [1, 3])
def test_batch_refined_modes(sq, fp64_batch, prec):
    r = sq.batch_solve(gpu_instances(sq), precision=prec)
    assert r.stats["unique_tubes"] == 3
    assert np.all(r.iters >= 1) and np.all(r.resid <= 1e-12), (r.iters, r.resid)
    for i in range(len(SPECS)):
        err = rel(np.r_[r.U[i], r.Om[i]], np.r_[fp64_batch.U[i], fp64_batch.Om[i]])
        assert err <= 1e-8, (prec, i, err)


@pytest.mark.parametrize("prec", [0, 3])
def test_batch_chunking_invariant(sq, prec):
    whole = sq.batch_solve(gpu_instances(sq), precision=prec)
    split = sq.batch_solve(gpu_instances(sq), precision=prec, mem_budget_gb=1e-6)
    assert split.stats["chunks"] == len(SPECS) and whole.stats["chunks"] == 3
    tol = 1e-12 if prec == 0 else 1e-10
    for i in range(len(SPECS)):
        assert rel(np.r_[split.U[i], split.Om[i]], np.r_[whole.U[i], whole.Om[i]]) <= tol


def test_batch_rejects_missing_tube(sq):
    inst = gpu_instances(sq)[:1]
    inst[0].tube = None
    with pytest.raises(Exception):
        sq.batch_solve(inst)


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_batch_graph_replay_equals_eager(sq, prec):
    """Each chunk's build (coupling assembly and right-hand sides; Schur complement, bordered
    scales and no-pivot factors) runs as two captured HIP graphs, and its refinement step is
    replayed from a captured graph: bit for bit what the same kernels launched one by one
    give (every replay feeds the next step, so a wrong replay shows in the final iterate;
    eager work and synchronisation run between the replays, whose residual norms are read
    back after each). Instances whose bordered matrix needs pivoting are refactored after
    the graphs, from the verdicts they left on the device."""
    sq.batch_options(use_graph=False)
    try:
        eager = sq.batch_solve(gpu_instances(sq), precision=prec)
    finally:
        sq.batch_options(use_graph=True)
    graph = sq.batch_solve(gpu_instances(sq), precision=prec)
    again = sq.batch_solve(gpu_instances(sq), precision=prec)
    chunks = graph.stats["chunks"]
    print("chunks", chunks, "build graphs", graph.stats["build_graphs"], "pivoted", graph.stats["pivoted"], flush=True)
    assert eager.stats["graph_replays"] == 0 and eager.stats["build_graphs"] == 0
    if prec:
        assert graph.stats["graph_replays"] >= 3 * 1  # one capture per chunk, replayed every step
    # the assembly stage is kernel launches only; so is the Schur stage unless a library GEMM
    # records other nodes (then it runs its launches directly)
    assert chunks <= graph.stats["build_graphs"] <= 2 * chunks
    if prec in (0, 1, 2):  # this library's own GEMM kernels only
        assert graph.stats["build_graphs"] == 2 * chunks
    assert graph.stats["pivoted"] == eager.stats["pivoted"]
    for r in (graph, again):
        assert np.array_equal(r.iters, eager.iters)
        assert np.array_equal(r.U, eager.U) and np.array_equal(r.Om, eager.Om)
        assert np.array_equal(r.resid, eager.resid)
