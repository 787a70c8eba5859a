"""FP64 parity of the GPU path against the CPU oracle (same meshes, same inputs).

Bar: panel integrals to 1e-13 relative (same rules, same points; only the summation order
and FMA contraction differ); U, Omega and panel tractions to 1e-10 relative, norm-wise
(components that vanish by symmetry are compared against the norm of the vector).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E_Z = np.array([0.0, 0.0, 1.0])


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_panel_integrals_all_rules(sq, O):
    """Far (7), 28, 112, 448-point and Duffy self pairs, on tube and sphere panels."""
    tube_o = O.swimmer_tube_mesh(1.0, 4.0, 0.3, 16, 4)
    tube_g = sq.swimmer_tube_mesh(1.0, 4.0, 0.3, 16, 4)
    sph_o = O.icosphere(0.3, 2, (0.2, 0.0, 0.1))
    sph_g = sq.icosphere(0.3, 2, (0.2, 0.0, 0.1))
    seen = set()
    rng = np.random.default_rng(0)
    for Mo, Mg in [(tube_o, tube_g), (sph_o, sph_g)]:
        lev, _ = O.levels(Mo, Mo.centroid)
        for L in range(4):
            ii, pp = np.nonzero(lev == L)
            pick = rng.choice(len(ii), size=min(6, len(ii)), replace=False) if len(ii) else []
            for q in pick:
                i, p = ii[q], pp[q]
                self_ = bool(i == p)
                so, do = O.panel_integrals(Mo, p, Mo.centroid[i], self_)
                sg, dg = sq.panel_integrals(Mg, p, Mg.centroid[i], self_)
                assert np.abs(sg - so).max() <= 1e-13 * np.abs(so).max()
                if not self_:
                    # DL ~ (r.m) A / r^3: for coplanar pairs it is pure cancellation, so
                    # compare against its natural size |SL| / distance as well
                    dist = np.linalg.norm(Mo.centroid[p] - Mo.centroid[i])
                    scale = max(np.abs(do).max(), np.abs(so).max() / dist)
                    assert np.abs(dg - do).max() <= 1e-13 * scale
                seen.add("self" if self_ else L)
    assert seen >= {0, 1, 2, 3, "self"}, seen


def test_tube_block_assembly(sq, O):
    """T = SL/(8 pi mu) + deflation for a whole tube, element by element."""
    torch = pytest.importorskip("torch")
    Mo = O.swimmer_tube_mesh(1.0, 4.0, 0.3, 12, 3)
    Mg = sq.swimmer_tube_mesh(1.0, 4.0, 0.3, 12, 3)
    P = Mo.P()
    T_o, _, _ = O.block(Mo, Mo.centroid, self_src=np.arange(P), sl_div=O.EIGHT_PI)
    dm = sq.DeviceMesh(Mg)
    x0 = torch.from_numpy(np.ascontiguousarray(Mg.centroid).reshape(-1)).cuda()
    self_src = torch.arange(P, dtype=torch.int32, device="cuda")
    T_g = torch.zeros((3 * P, 3 * P), dtype=torch.float64, device="cuda").t()
    DL_g = torch.zeros((3 * P, 3 * P), dtype=torch.float64, device="cuda").t()
    dls = torch.zeros(9 * P, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    sq.assemble(dm, x0, self_src, sl_div=O.EIGHT_PI, SL=T_g, dl_div=O.EIGHT_PI, DL=DL_g, dlsum=dls)
    sq.sync()
    _, DL_o, dls_o = O.block(Mo, Mo.centroid, self_src=np.arange(P), dl_div=O.EIGHT_PI, dlsum=True)
    assert np.abs(T_g.cpu().numpy() - T_o).max() <= 1e-13 * np.abs(T_o).max()
    assert np.abs(DL_g.cpu().numpy() - DL_o).max() <= 1e-13 * np.abs(DL_o).max()
    assert np.abs(dls.cpu().numpy().reshape(P, 9) - dls_o).max() <= 1e-12 * np.abs(dls_o).max()


def _oracle_solve(O, R, L, a, nth, nr, lev, center, slips_e=E_Z):
    t = O.Tube(R, L, a, nth, nr) if R else None
    s = O.SwimmerSolver(t, O.icosphere(a, lev, center), center)
    return s, s.solve_free([s.slip(slips_e, 1, 0), s.slip(slips_e, 0, 1)])


def _gpu_solve(sq, R, L, a, nth, nr, lev, center, prec=0, slips_e=E_Z):
    t = sq.Tube(R, L, a, nth, nr, precision=prec) if R else None
    s = sq.SwimmerSolver(t, sq.icosphere(a, lev, center), center, precision=prec)
    return s, s.solve_free([s.slip(slips_e, 1, 0), s.slip(slips_e, 0, 1)]), t


@pytest.mark.parametrize("center", [(0.0, 0.0, 0.0), (0.21, 0.0, 0.0)])
def test_swimmer_fp64_parity_cfg1(sq, O, center):
    """Config 1 (R=1, L/R=4, n_theta=24, n_r=6, level 2, a/R=0.3), on axis and off axis."""
    args = (1.0, 4.0, 0.3, 24, 6, 2, center)
    so, ro = _oracle_solve(O, *args)
    sg, rg, _ = _gpu_solve(sq, *args)
    # norm-wise over both modes: the B2 mode on the axis is zero by symmetry (~1e-9 noise)
    scale_u = max(np.linalg.norm(np.r_[r.U, r.Om]) for r in ro)
    scale_fs = max(np.linalg.norm(r.f_sphere) for r in ro)
    scale_ft = max(np.linalg.norm(r.f_tube) for r in ro)
    for g, r in zip(rg, ro):
        assert np.linalg.norm(np.r_[g.U - r.U, g.Om - r.Om]) < 1e-10 * scale_u
        assert np.linalg.norm(g.f_sphere - r.f_sphere) < 1e-10 * scale_fs
        assert np.linalg.norm(g.f_tube - r.f_tube) < 1e-10 * scale_ft


def test_swimmer_unbounded_and_towed(sq, O):
    """Free space (tube = nullptr) and the towed rigid sphere (solve_prescribed)."""
    so = O.SwimmerSolver(None, O.icosphere(1.0, 2), [0, 0, 0])
    sg = sq.SwimmerSolver(None, sq.icosphere(1.0, 2), [0, 0, 0])
    ro = so.solve_free([so.slip(E_Z, 1, 0), so.slip(E_Z, 0, 1)])
    rg = sg.solve_free([sg.slip(E_Z, 1, 0), sg.slip(E_Z, 0, 1)])
    assert rel(rg[0].U, ro[0].U) < 1e-10
    assert np.linalg.norm(rg[1].U) + np.linalg.norm(rg[1].Om) < 1e-10
    Ko = -so.force(so.solve_prescribed([0, 0, 1], [0, 0, 0]))[2] / (6 * np.pi)
    Kg = -sg.force(sg.solve_prescribed([0, 0, 1], [0, 0, 0]))[2] / (6 * np.pi)
    assert abs(Kg - Ko) < 1e-10 * abs(Ko)
    # towed inside the tube
    to = O.Tube(1.0, 4.0, 0.3, 16, 4)
    tg = sq.Tube(1.0, 4.0, 0.3, 16, 4)
    so = O.SwimmerSolver(to, O.icosphere(0.3, 1), [0, 0, 0])
    sg = sq.SwimmerSolver(tg, sq.icosphere(0.3, 1), [0, 0, 0])
    fo = so.solve_prescribed([0, 0, 1], [0, 0, 0])
    fg = sg.solve_prescribed([0, 0, 1], [0, 0, 0])
    assert rel(fg.f_sphere, fo.f_sphere) < 1e-10
    assert rel(fg.f_tube, fo.f_tube) < 1e-10


@pytest.mark.parametrize("prec,tol", [(1, 1e-8), (2, 1e-8), (3, 1e-8)])
def test_mixed_precision_refinement(sq, O, prec, tol):
    """FP32 / FP16-MFMA / BF16-MFMA factors with FP64 refinement reach the FP64 answer."""
    args = (1.0, 4.0, 0.3, 24, 6, 2, (0.21, 0.0, 0.0))
    _, r64, _ = _gpu_solve(sq, *args, prec=0)
    s, rmx, _ = _gpu_solve(sq, *args, prec=prec)
    for g, r in zip(rmx, r64):
        assert rel(np.r_[g.U, g.Om], np.r_[r.U, r.Om]) < tol
        assert rel(g.f_sphere, r.f_sphere) < 1e-7
    assert s.ir_residual <= 1e-12


def test_velocity_field_parity(sq, O):
    Mo = O.icosphere(1.0, 2)
    Mg = sq.icosphere(1.0, 2)
    rng = np.random.default_rng(5)
    u = rng.standard_normal((Mo.P(), 3))
    f = rng.standard_normal((Mo.P(), 3))
    pts = np.array([[0, 0, 1.1], [1.5, 0.2, 0.0], [0.0, -3.0, 0.5], [0.7, 0.7, 0.7]])
    vo = O.velocity_at(Mo, u, f, 1.0, pts)
    vg = sq.velocity_at(Mg, u, f, 1.0, pts)
    assert rel(vg, vo) < 1e-12


def test_duct_flow_parity(sq, O):
    do = O.DuctFlow(1.0, 4.0, 12, 12, 3)
    dg = sq.DuctFlow(1.0, 4.0, 12, 12, 3)
    uo, fo, Qo = do.solve(1.0, 0.0)
    sg = dg.solve(1.0, 0.0)
    assert abs(sg.Q - Qo) < 1e-10 * abs(Qo)
    assert rel(sg.f, fo) < 1e-10
    assert rel(sg.u, uo) < 1e-10


def test_device_memory_cache_reuse_and_trim(sq):
    """A second identical solve reuses the cached device blocks and gives the same answer
    bit for bit; sqbem_trim hands the cache back to the driver."""
    def solve():
        t = sq.Tube(1.0, 4.0, 0.3, 12, 3)
        s = sq.SwimmerSolver(t, sq.icosphere(0.3, 1, (0.1, 0.0, 0.0)), (0.1, 0.0, 0.0))
        r = s.solve_free([s.slip(E_Z, 1.0, 0.5)])[0]
        del s, t
        return np.r_[r.U, r.Om]

    a = solve()
    b = solve()
    assert np.array_equal(a, b)
    assert sq.trim() > 0.0
    assert sq.trim() == 0.0
    assert np.array_equal(solve(), a)
