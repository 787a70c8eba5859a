"""Bordered Schur matrices that partial pivoting would interchange (src/swimmer.hpp:107-125,
dgesv_). A large sphere near the wall, a/R = 0.7 at offset b/R = 0.22 in a 16-panel tube
with a level-2 sphere: LAPACK's factorisation of the FP64 bordered matrix swaps one row
(scipy.linalg.lu_factor of the oracle's bordered matrix; the largest multiplier is 0.97).
The factorisation without interchanges is then rejected and the matrix refactored with
partial pivoting: in FP64 from a copy of the matrix, in the refined modes rebuilt from the
FP32 Schur complement (no copy is kept). Both must give the FP64 answer: FP64 against
the oracle at 1e-10, refined against FP64 at 1e-8 with the residual at 1e-12; the towed
problem of the refined modes factors Gs alone the same way."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

A = 1.0
R = A / 0.7
OFFSET = 0.22 * R


def setup(sq, prec):
    tube = sq.Tube(R, 4 * R, A, 16, 4, precision=prec)
    c = np.array([OFFSET, 0.0, 0.0])
    s = sq.SwimmerSolver(tube, sq.icosphere(A, 2, c), c)
    return s, sq.getrf_path()


@pytest.fixture(scope="module")
def fp64(sq):
    s, _ = setup(sq, 0)
    e = np.array([0.0, 0.6, 0.8])
    sol = s.solve_free([s.slip(e, 1, 0), s.slip(e, 0, 1)])
    path = sq.getrf_path()  # FP64 solve_free factors the bordered matrix (the refined modes at construction)
    towed = s.solve_prescribed([0, 0, 1], [0, 0, 0])
    return s, path, e, sol, towed


def test_fp64_bordered_matrix_is_refactored_with_pivoting(fp64):
    s, path, e, sol, _ = fp64
    assert path == 2, path
    tube = O.Tube(R, 4 * R, A, 16, 4)
    c = np.array([OFFSET, 0.0, 0.0])
    so = O.SwimmerSolver(tube, O.icosphere(A, 2, c), c)
    ref = so.solve_free([so.slip(e, 1, 0), so.slip(e, 0, 1)])
    scale = max(np.linalg.norm(np.r_[r.U, r.Om]) for r in ref)
    for g, r in zip(sol, ref):
        assert np.linalg.norm(np.r_[g.U - r.U, g.Om - r.Om]) <= 1e-10 * scale
        assert np.linalg.norm(g.f_sphere - r.f_sphere) <= 1e-10 * np.linalg.norm(r.f_sphere)


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_refined_bordered_matrix_rebuilt_for_pivoting(sq, fp64, prec):
    _, _, e, ref, towed64 = fp64
    s, path = setup(sq, prec)
    print(f"precision {prec}: factorisation path {path}", flush=True)
    assert path == 2, path
    sol = s.solve_free([s.slip(e, 1, 0), s.slip(e, 0, 1)])
    assert s.ir_residual <= 1e-12
    scale = max(np.linalg.norm(np.r_[r.U, r.Om]) for r in ref)
    for g, r in zip(sol, ref):
        assert np.linalg.norm(np.r_[g.U - r.U, g.Om - r.Om]) <= 1e-8 * scale
        assert np.linalg.norm(g.f_sphere - r.f_sphere) <= 1e-8 * np.linalg.norm(r.f_sphere)
        assert np.linalg.norm(g.f_tube - r.f_tube) <= 1e-8 * np.linalg.norm(r.f_tube)
    t = s.solve_prescribed([0, 0, 1], [0, 0, 0])
    assert s.ir_residual <= 1e-12
    F, F64 = s.force(t), fp64[0].force(towed64)
    assert np.linalg.norm(F - F64) <= 1e-8 * np.linalg.norm(F64)
