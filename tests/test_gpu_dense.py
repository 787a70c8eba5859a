"""Dense kernels of libsqbem on the GPU against float64 references (numpy/scipy LAPACK).

GEMM: FP64 MFMA to rounding (k * eps), FP32 MFMA to FP32 accuracy, FP16/BF16 MFMA with
FP32 accumulation to the input rounding of the 16-bit formats. LU: partial-pivoting
factors and solves against scipy's LAPACK dgetrf/dgetrs (pivot rows compared exactly on
random matrices, backward error bounded by c n eps).
"""
import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a, dtype=None):
    a = np.asarray(a)
    dt = {np.float64: torch.float64, np.float32: torch.float32}[(dtype or a.dtype).type]
    t = torch.empty_strided(a.shape, (1, a.shape[0]) if a.ndim == 2 else (1,), dtype=dt, device="cuda")
    t.copy_(torch.from_numpy(np.asarray(a, dtype=dtype or a.dtype)))
    return t


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


# the last shape is large enough for the library (rocBLAS) path of plain FP64/FP32 products
SHAPES = [(1, 1, 1), (37, 53, 19), (128, 128, 16), (129, 130, 17), (300, 257, 129), (1000, 2, 700), (64, 960, 3),
          (1024, 1100, 700)]


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_dgemm(sq, m, n, k):
    rng = np.random.default_rng(m * 1000 + n + k)
    A, B, C = rng.standard_normal((m, k)), rng.standard_normal((k, n)), rng.standard_normal((m, n))
    alpha, beta = -1.25, 0.5
    dA, dB, dC = dev(A), dev(B), dev(C)
    torch.cuda.synchronize()
    sq.dgemm(alpha, dA, dB, beta, dC)
    sq.sync()
    ref = alpha * (A @ B) + beta * C
    scale = np.abs(A) @ np.abs(B)
    err = np.abs(host(dC) - ref) / (np.abs(alpha) * scale + np.abs(beta * C) + 1e-300)
    assert err.max() < 4 * k * 2.2e-16 + 1e-15, err.max()


def test_dgemm_library_path_submatrix(sq):
    """A large plain product (rocBLAS) on sub-blocks of one matrix: leading dimensions larger
    than the rows, beta = 0 with C starting as NaN, and C beside (not overlapping) A and B."""
    rng = np.random.default_rng(5)
    big = rng.standard_normal((1700, 1800))
    dbig = dev(big)
    A = dbig[3:803, 0:600]
    B = dbig[900:1500, 1000:1700]
    C = torch.full((800, 700), float("nan"), dtype=torch.float64, device="cuda").t().contiguous().t()
    torch.cuda.synchronize()
    sq.dgemm(1.0, A, B, 0.0, C)
    sq.sync()
    ref = big[3:803, 0:600] @ big[900:1500, 1000:1700]
    assert np.allclose(host(C), ref, rtol=1e-12, atol=1e-11)


def test_dgemm_submatrix_and_beta_zero(sq):
    rng = np.random.default_rng(1)
    big = rng.standard_normal((200, 200))
    dbig = dev(big)
    # C = A B on sub-blocks of one matrix with lda = 200; C starts as NaN (beta = 0 must not read it)
    A = dbig[10:60, 0:30]
    B = dbig[0:30, 100:140]
    C = torch.full((50, 40), float("nan"), dtype=torch.float64, device="cuda").t().contiguous().t()
    torch.cuda.synchronize()
    sq.dgemm(1.0, A, B, 0.0, C)
    sq.sync()
    ref = big[10:60, 0:30] @ big[0:30, 100:140]
    assert np.allclose(host(C), ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_sgemm(sq, m, n, k):
    rng = np.random.default_rng(7 + m + n + k)
    A, B, C = (rng.standard_normal(s).astype(np.float32) for s in [(m, k), (k, n), (m, n)])
    dA, dB, dC = dev(A), dev(B), dev(C)
    torch.cuda.synchronize()
    sq.sgemm(2.0, dA, dB, -1.0, dC)
    sq.sync()
    ref = 2.0 * (A.astype(np.float64) @ B.astype(np.float64)) - C.astype(np.float64)
    scale = 2.0 * (np.abs(A).astype(np.float64) @ np.abs(B)) + np.abs(C)
    err = np.abs(host(dC) - ref) / scale
    assert err.max() < 4 * k * 6e-8 + 1e-6, err.max()


@pytest.mark.parametrize("kind,tol", [(2, 2.5e-3), (3, 2e-2)])
@pytest.mark.parametrize("m,n,k", [(128, 128, 64), (300, 257, 129), (1000, 7, 700)])
def test_gemm16(sq, kind, tol, m, n, k):
    """16-bit MFMA products, FP32 accumulation: error within the 16-bit input rounding."""
    rng = np.random.default_rng(kind * 31 + m)
    # rows/columns of very different magnitude exercise the FP16 power-of-two scaling
    A = (rng.standard_normal((m, k)) * np.logspace(-6, 3, m)[:, None]).astype(np.float32)
    B = (rng.standard_normal((k, n)) * np.logspace(2, -5, n)[None, :]).astype(np.float32)
    C = np.zeros((m, n), dtype=np.float32)
    dA, dB, dC = dev(A), dev(B), dev(C)
    torch.cuda.synchronize()
    sq.gemm16(kind, 1.0, dA, dB, 0.0, dC)
    sq.sync()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    scale = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)
    err = np.abs(host(dC) - ref) / scale
    assert err.max() < tol, err.max()


@pytest.mark.parametrize("m,n,k", [(700, 555, 300), (300, 290, 128), (1100, 1030, 4160)])
@pytest.mark.parametrize("kind,tol", [(2, 2.5e-3), (3, 2e-2)])
def test_gemm16_autotuner_every_tile(sq, kind, tol, m, n, k):
    """The tile autotuner runs each of its fourteen candidates once on a fresh shape class
    (odd sizes, beta != 0) before settling: every call's product must be right, and the
    class must end decided. The k values give the load-to-LDS candidates (32- and 64-deep
    tiles, four stages) a partial last group of tiles (k = 300, 4160) and fewer tiles than
    stages (k = 128)."""
    rng = np.random.default_rng(7 + kind)
    A = (rng.standard_normal((m, k)) * np.logspace(-3, 2, m)[:, None]).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    C0 = rng.standard_normal((m, n)).astype(np.float32)
    ref = 0.5 * (A.astype(np.float64) @ B.astype(np.float64)) - 2.0 * C0
    scale = 0.5 * (np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)) + 2.0 * np.abs(C0)
    sq.tune(1)
    try:
        for call in range(16):
            dA, dB, dC = dev(A), dev(B), dev(C0)
            torch.cuda.synchronize()
            sq.gemm16(kind, 0.5, dA, dB, -2.0, dC)
            sq.sync()
            err = np.abs(host(dC) - ref) / scale
            assert err.max() < tol, (call, err.max())
        info = sq.tune_info()
        assert info["classes"] >= 1 and info["decided"] >= 1, info
    finally:
        sq.tune(1)


def test_gemm16_tune_table_save_and_load(sq, tmp_path):
    """A decided class survives save -> reset -> load and is used from the first call."""
    m, n, k = 640, 600, 320
    rng = np.random.default_rng(11)
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    sq.tune(1)
    try:
        for _ in range(15):  # fourteen candidates explored, then the decided one
            dA, dB, dC = dev(A), dev(B), dev(np.zeros((m, n), dtype=np.float32))
            torch.cuda.synchronize()
            sq.gemm16(3, 1.0, dA, dB, 0.0, dC)
            sq.sync()
        before = sq.tune_info()
        path = str(tmp_path / "tune.txt")
        sq.tune_save(path)
        sq.tune(1)
        assert sq.tune_info()["classes"] == 0
        assert sq.tune_load(path) == before["decided"] >= 1
        after = sq.tune_info()
        assert after["decided"] == before["decided"] and after["tiles"] == before["tiles"]
        dA, dB, dC = dev(A), dev(B), dev(np.zeros((m, n), dtype=np.float32))
        torch.cuda.synchronize()
        sq.gemm16(3, 1.0, dA, dB, 0.0, dC)
        sq.sync()
        assert np.abs(host(dC) - ref).max() < 2e-2 * np.abs(ref).max()
        assert sq.tune_info()["classes"] == after["classes"]
    finally:
        sq.tune(1)


def test_gemm16_fixed_rule_mode(sq):
    m, n, k = 512, 512, 256
    rng = np.random.default_rng(3)
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    sq.tune(0)
    try:
        dA, dB, dC = dev(A), dev(B), dev(np.zeros((m, n), dtype=np.float32))
        torch.cuda.synchronize()
        sq.gemm16(2, 1.0, dA, dB, 0.0, dC)
        sq.sync()
        ref = A.astype(np.float64) @ B.astype(np.float64)
        assert np.abs(host(dC) - ref).max() < 2.5e-3 * np.abs(ref).max()
        assert sq.tune_info()["classes"] == 0
    finally:
        sq.tune(1)


@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 64, 65, 100, 257, 1000])
def test_dgetrf_matches_lapack(sq, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    dA = dev(A)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    info = sq.dgetrf(dA, ipiv)
    assert info == 0
    lu, piv = sla.lu_factor(A)
    got_piv = host(ipiv) - 1
    assert np.array_equal(got_piv, piv), "pivot rows differ from LAPACK"
    assert np.allclose(host(dA), lu, rtol=1e-10, atol=1e-11 * np.abs(lu).max())


@pytest.mark.parametrize("n,nrhs", [(1, 1), (40, 3), (257, 966), (1500, 2)])
@pytest.mark.parametrize("trans", [False, True])
def test_dgetrs_backward_error(sq, n, nrhs, trans):
    rng = np.random.default_rng(n + nrhs)
    A = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    X = rng.standard_normal((n, nrhs))
    Bm = (A.T if trans else A) @ X
    dA, dB = dev(A), dev(Bm)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert sq.dgetrf(dA, ipiv) == 0
    sq.dgetrs(dA, ipiv, dB, trans=trans)
    sq.sync()
    Xg = host(dB)
    Aop = A.T if trans else A
    r = np.linalg.norm(Aop @ Xg - Bm, axis=0) / (np.linalg.norm(Aop, 2) * np.linalg.norm(Xg, axis=0))
    assert r.max() < 30 * n * 2.2e-16, r.max()


@pytest.mark.parametrize("n", [63, 64, 65, 1000, 4099])
@pytest.mark.parametrize("nrhs", [1, 2, 3, 5, 8])
def test_getrs_few_rhs_chained_solve(sq, n, nrhs):
    """The one-launch triangular solves (<= 8 right-hand sides): FP64 and FP32 factors,
    a leading dimension larger than n, two solves in a row (the kernel's counters carry
    over between launches), against the backward-error bound."""
    rng = np.random.default_rng(7 * n + nrhs)
    A = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    X = rng.standard_normal((n, nrhs))
    Bm = A @ X
    for dt, tol in [(np.float64, 30 * n * 2.2e-16), (np.float32, 30 * n * 1.2e-7)]:
        dA = dev(A.astype(dt))
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        assert (sq.dgetrf(dA, ipiv) if dt == np.float64 else sq.sgetrf(1, dA, ipiv)) == 0
        for rep in range(2):
            full = dev(np.vstack([Bm, np.full((5, nrhs), 7.0)]).astype(dt))  # ld = n + 5
            dB = full[:n]
            (sq.dgetrs if dt == np.float64 else sq.sgetrs)(dA, ipiv, dB)
            sq.sync()
            Xg = host(dB).astype(np.float64)
            assert np.all(host(full)[n:] == 7.0), "rows beyond n were written"
            r = np.linalg.norm(A @ Xg - Bm, axis=0) / (np.linalg.norm(A, 2) * np.linalg.norm(Xg, axis=0))
            assert r.max() < tol, (dt, rep, r.max())


@pytest.mark.parametrize("n", [512, 700, 1500])
@pytest.mark.parametrize("swap", [None, (600, 650)])
def test_getrf_no_pivot_path_and_fallback(sq, n, swap):
    """Column-dominant matrices need no row interchanges: the 128-block no-pivot path must
    give LAPACK's factors and identity pivots. With two rows exchanged, partial pivoting
    must move a row there: the no-pivot attempt is rejected and the pivoting path used."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 2.0 * n ** 0.5 * np.eye(n)
    A += np.diag(np.abs(A).sum(axis=0))  # strictly column dominant: GEPP keeps the diagonal
    if swap is not None and swap[1] < n:
        A[[swap[0], swap[1]]] = A[[swap[1], swap[0]]]
    lu, piv = sla.lu_factor(A)
    for dt in (np.float64, np.float32):
        dA = dev(A.astype(dt))
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        info = sq.dgetrf(dA, ipiv) if dt == np.float64 else sq.sgetrf(1, dA, ipiv)
        assert info == 0
        assert np.array_equal(host(ipiv) - 1, piv), "pivot rows differ from LAPACK"
        tol = 1e-11 if dt == np.float64 else 2e-5
        assert np.allclose(host(dA).astype(np.float64), lu, rtol=tol, atol=tol * np.abs(lu).max())
        # and a solve with many right-hand sides through the 128-block inverses
        X = rng.standard_normal((n, 20))
        dB = dev((A @ X).astype(dt))
        (sq.dgetrs if dt == np.float64 else sq.sgetrs)(dA, ipiv, dB)
        sq.sync()
        Xg = host(dB).astype(np.float64)
        r = np.linalg.norm(A @ Xg - A @ X, axis=0) / (np.linalg.norm(A, 2) * np.linalg.norm(Xg, axis=0))
        assert r.max() < 30 * n * (2.2e-16 if dt == np.float64 else 1.2e-7), (dt, r.max())


def test_dgetrf_singular_reports_info(sq):
    n = 70
    rng = np.random.default_rng(3)
    A = rng.standard_normal((n, n))
    A[:, 40] = 0.0  # exact zero column -> zero pivot at column 41 (1-based)
    dA = dev(A)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    info = sq.dgetrf(dA, ipiv)
    assert info == 41
    assert "dgetrf failed" in sq.last_error()


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_sgetrf_solves(sq, kind):
    n = 700
    rng = np.random.default_rng(kind)
    A = rng.standard_normal((n, n)) + 3 * n ** 0.5 * np.eye(n)
    X = rng.standard_normal((n, 2))
    Bm = (A @ X).astype(np.float32)
    dA, dB = dev(A.astype(np.float32)), dev(Bm)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert sq.sgetrf(kind, dA, ipiv) == 0
    sq.sgetrs(dA, ipiv, dB)
    sq.sync()
    err = np.linalg.norm(host(dB) - X) / np.linalg.norm(X)
    tol = {1: 1e-5, 2: 2e-2, 3: 5e-2}[kind]
    assert err < tol, err


def test_host_linalg_api(sq):
    """gemm / solve_dense / LUFactor on host matrices (src/utils/linalg.hpp:42-78)."""
    rng = np.random.default_rng(11)
    A = np.asfortranarray(rng.standard_normal((60, 60)) + 8 * np.eye(60))
    B = np.asfortranarray(rng.standard_normal((60, 4)))
    C = np.asfortranarray(rng.standard_normal((60, 4)))
    C0 = C.copy()
    sq.gemm(2.0, A, B, -1.0, C)
    assert np.allclose(C, 2 * A @ B - C0, rtol=1e-13, atol=1e-12)
    X = B.copy(order="F")
    sq.solve_dense(A, X)
    assert np.allclose(A @ X, B, rtol=1e-12, atol=1e-12)
    lu = sq.LUFactor()
    lu.factor(A)
    Y = B.copy(order="F")
    lu.solve(Y)
    assert np.allclose(Y, X, rtol=1e-13, atol=1e-13)
    with pytest.raises(RuntimeError, match="gemm: dimension mismatch"):
        sq.gemm(1.0, A, B, 0.0, np.zeros((5, 5), order="F"))


@pytest.mark.parametrize("n,nrhs", [(256, 16400), (200, 16384), (256, 6000), (200, 5000), (256, 1000), (160, 40)])
def test_sgetrs_fused_bottom_levels_all_panel_widths(sq, n, nrhs):
    """FP32 factors of order 128 < n <= 256 solve many right-hand sides in one fused launch
    of the triangular solves' bottom two levels: 64-column panels once ceil(nrhs / 64) fills
    every CU (nrhs >= 16,321 on 256 CUs), 32-column panels while ceil(nrhs / 32) fills at
    least half of them (nrhs >= 4,065), 16-column panels below that. Every width against the
    FP32 backward-error bound."""
    rng = np.random.default_rng(n + nrhs)
    A = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    X = rng.standard_normal((n, nrhs))
    Bm = A @ X
    dA = dev(A.astype(np.float32))
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert sq.sgetrf(1, dA, ipiv) == 0
    dB = dev(Bm.astype(np.float32))
    sq.sgetrs(dA, ipiv, dB)
    sq.sync()
    Xg = host(dB).astype(np.float64)
    r = np.linalg.norm(A @ Xg - Bm, axis=0) / (np.linalg.norm(A, 2) * np.linalg.norm(Xg, axis=0))
    assert r.max() < 30 * n * 1.2e-7, r.max()


@pytest.mark.parametrize("n", [5, 33, 65, 257, 1000, 1500])
def test_pivoted_panels_without_cooperative_launch(sq, n):
    """The pivoted LU's panels, made to take the barrier-free path (the one the library
    uses once the runtime refuses a cooperative launch): LAPACK's pivot rows and factors,
    FP64 and FP32, for random matrices (pivots throughout) and, at n >= 512, a column-
    dominant matrix with two rows exchanged (the no-pivot attempt rejected first)."""
    rng = np.random.default_rng(100 + n)
    mats = [rng.standard_normal((n, n))]
    if n >= 512:
        B = rng.standard_normal((n, n)) + 2.0 * n ** 0.5 * np.eye(n)
        B += np.diag(np.abs(B).sum(axis=0))
        B[[300, 310]] = B[[310, 300]]
        mats.append(B)
    sq.coop(0)
    try:
        for A in mats:
            lu, piv = sla.lu_factor(A)
            for dt in (np.float64, np.float32):
                dA = dev(A.astype(dt))
                ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
                torch.cuda.synchronize()
                info = sq.dgetrf(dA, ipiv) if dt == np.float64 else sq.sgetrf(1, dA, ipiv)
                assert info == 0
                if dt == np.float64:
                    # FP32 factors of a random matrix may legitimately choose other pivots on
                    # near-ties; FP64 ones must be LAPACK's
                    assert np.array_equal(host(ipiv) - 1, piv), "pivot rows differ from LAPACK"
                    assert np.allclose(host(dA), lu, rtol=1e-10, atol=1e-11 * np.abs(lu).max())
                X = rng.standard_normal((n, 3))
                dB = dev((A @ X).astype(dt))
                (sq.dgetrs if dt == np.float64 else sq.sgetrs)(dA, ipiv, dB)
                sq.sync()
                Xg = host(dB).astype(np.float64)
                r = np.linalg.norm(A @ Xg - A @ X, axis=0) / (np.linalg.norm(A, 2) * np.linalg.norm(Xg, axis=0))
                assert r.max() < 30 * n * (2.2e-16 if dt == np.float64 else 1.2e-7), (dt, r.max())
    finally:
        sq.coop(1)
