"""CPU oracle for the squirmer-in-tube BEM hot path.

TEST INFRASTRUCTURE: only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import this module. The product library never calls it.

Geometry, quadrature and panel integrals come from ``bem_oracle.c`` (a C restatement
of src/utils/{mesh,quadrature,kernels,field}.hpp). The solver steps restate
src/swimmer.hpp, src/duct_flow.hpp and src/trajectory.hpp with numpy/scipy doing the
BLAS/LAPACK calls the reference delegates to OpenBLAS (dgetrf/dgetrs/dgesv/dgemm,
src/utils/linalg.hpp:42-77): scipy's LAPACK is an LP64 OpenBLAS, the same family of
library the reference links.
"""
from __future__ import annotations

import ctypes
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

_HERE = os.path.dirname(os.path.abspath(__file__))
PI = 3.14159265358979323846
EIGHT_PI = 8.0 * PI
WALL, INLET, OUTLET, SPHERE = 0, 1, 2, 10

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        L = ctypes.CDLL(path)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        P = ctypes.POINTER
        L.om_swimmer_tube_mesh.restype = vp
        L.om_swimmer_tube_mesh.argtypes = [d, d, d, i, i]
        L.om_duct_mesh.restype = vp
        L.om_duct_mesh.argtypes = [d, d, i, i, i]
        L.om_tube_mesh.restype = vp
        L.om_tube_mesh.argtypes = [d, vp, i, i, i]
        L.om_icosphere.restype = vp
        L.om_icosphere.argtypes = [d, i, d, d, d, i]
        L.om_merge.restype = vp
        L.om_merge.argtypes = [vp, vp]
        L.om_copy.restype = vp
        L.om_copy.argtypes = [vp]
        L.om_translate.argtypes = [vp, d, d, d]
        L.om_mesh_free.argtypes = [vp]
        L.om_mesh_nn.argtypes = [vp]
        L.om_mesh_np.argtypes = [vp]
        L.om_mesh_get.argtypes = [vp] + [vp] * 8
        L.om_graded_z.argtypes = [d, d, d, d, d, vp, i]
        L.om_rule_get.argtypes = [i, vp, vp]
        L.om_gauss_legendre01.argtypes = [i, vp, vp]
        L.om_panel_integrals.argtypes = [vp, i, vp, i, vp, vp]
        L.om_levels.argtypes = [vp, i, i, vp, i, vp, vp, vp]
        L.om_block.argtypes = [vp, i, i, vp, i, vp, d, vp, ctypes.c_long, d, vp, ctypes.c_long, vp]
        L.om_velocity_at.argtypes = [vp, vp, vp, d, vp, i, vp]
        L.om_squirmer_slip.argtypes = [vp, vp, vp, d, d, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Mesh:
    """Flat-triangle mesh (src/utils/mesh.hpp:50-91) backed by the C oracle."""

    def __init__(self, handle):
        self.h_ = handle
        self._load()

    def _load(self):
        L = lib()
        nn, npn = L.om_mesh_nn(self.h_), L.om_mesh_np(self.h_)
        self.nodes = np.zeros((nn, 3))
        self.tris = np.zeros((npn, 3), dtype=np.int32)
        self.tags = np.zeros(npn, dtype=np.int32)
        self.V = np.zeros((npn, 3, 3))
        self.normal = np.zeros((npn, 3))
        self.centroid = np.zeros((npn, 3))
        self.area = np.zeros(npn)
        self.h = np.zeros(npn)
        L.om_mesh_get(self.h_, _p(self.nodes), _p(self.tris), _p(self.tags), _p(self.V),
                      _p(self.normal), _p(self.centroid), _p(self.area), _p(self.h))

    def __del__(self):
        try:
            if self.h_:
                lib().om_mesh_free(self.h_)
        except Exception:
            pass

    def P(self):
        return len(self.tris)

    def total_area(self):
        s = 0.0
        for A in self.area:
            s += A
        return s

    def copy(self):
        return Mesh(lib().om_copy(self.h_))

    def translate(self, d):
        lib().om_translate(self.h_, float(d[0]), float(d[1]), float(d[2]))
        self._load()


def swimmer_tube_mesh(R, L, a, nth, nr=4):
    return Mesh(lib().om_swimmer_tube_mesh(R, L, a, nth, nr))


def duct_mesh(R, L, nth, nz, nr):
    return Mesh(lib().om_duct_mesh(R, L, nth, nz, nr))


def tube_mesh(R, z, nth, nr):
    z = np.ascontiguousarray(z, dtype=np.float64)
    return Mesh(lib().om_tube_mesh(R, _p(z), len(z), nth, nr))


def icosphere(a, level, center=(0.0, 0.0, 0.0), tag=SPHERE):
    return Mesh(lib().om_icosphere(a, level, float(center[0]), float(center[1]), float(center[2]), tag))


def merge(a, b):
    return Mesh(lib().om_merge(a.h_, b.h_))


def graded_z(L, dzmin, dzmax, core, growth=1.25):
    out = np.zeros(100000)
    n = lib().om_graded_z(L, dzmin, dzmax, core, growth, _p(out), len(out))
    return out[:n].copy()


def rule(lev):
    Lb = np.zeros((448, 3))
    W = np.zeros(448)
    n = lib().om_rule_get(lev, _p(Lb), _p(W))
    return Lb[:n].copy(), W[:n].copy()


def gauss_legendre01(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().om_gauss_legendre01(n, _p(x), _p(w))
    return x, w


def panel_integrals(M, p, x0, self_=False):
    SL, DL = np.zeros(9), np.zeros(9)
    x = np.ascontiguousarray(x0, dtype=np.float64)
    lib().om_panel_integrals(M.h_, p, _p(x), int(self_), _p(SL), _p(DL))
    return SL.reshape(3, 3), DL.reshape(3, 3)


def levels(M, x0, p0=0, np_=None, self_src=None):
    np_ = M.P() - p0 if np_ is None else np_
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1, 3)
    nt = len(x0)
    lev = np.zeros((nt, np_), dtype=np.int8)
    dist = np.zeros((nt, np_))
    ss = None if self_src is None else np.ascontiguousarray(self_src, dtype=np.int32)
    lib().om_levels(M.h_, p0, np_, _p(x0), nt, _p(ss), _p(lev), _p(dist))
    return lev, dist


def block(M, x0, self_src=None, p0=0, np_=None, sl_div=None, dl_div=None, dlsum=False,
          dlsum_init=None):
    """Column-major coupling blocks SL/sl_div and -DL/dl_div (om_block). The raw DL row
    sums accumulate onto dlsum_init when given (one running sum, as the reference)."""
    np_ = M.P() - p0 if np_ is None else np_
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1, 3)
    nt = len(x0)
    SLm = np.zeros((3 * nt, 3 * np_), order="F") if sl_div is not None else None
    DLm = np.zeros((3 * nt, 3 * np_), order="F") if dl_div is not None else None
    if dlsum_init is not None:
        ds = np.ascontiguousarray(dlsum_init, dtype=np.float64).reshape(nt, 9).copy()
    else:
        ds = np.zeros((nt, 9)) if dlsum else None
    ss = None if self_src is None else np.ascontiguousarray(self_src, dtype=np.int32)
    lib().om_block(M.h_, p0, np_, _p(x0), nt, _p(ss), sl_div or 1.0, _p(SLm), 3 * nt,
                   dl_div or 1.0, _p(DLm), 3 * nt, _p(ds))
    return SLm, DLm, ds


def squirmer_slip(x, center, e, B1, B2):
    out = np.zeros(3)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    cc = np.ascontiguousarray(center, dtype=np.float64)
    ee = np.ascontiguousarray(e, dtype=np.float64)
    lib().om_squirmer_slip(_p(xx), _p(cc), _p(ee), B1, B2, _p(out))
    return out


def velocity_at(M, u, f, mu, pts):
    """src/utils/field.hpp:18-34."""
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1, 3)
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(pts)
    lib().om_velocity_at(M.h_, _p(u), _p(f), mu, _p(pts), len(pts), _p(out))
    return out


# ------------------------------------------------------------------ linear algebra
class LUFactor:
    """LUFactor, src/utils/linalg.hpp:62-78 (dgetrf / dgetrs('N'))."""

    def factor(self, A):
        self.lu_piv = sla.lu_factor(A, overwrite_a=True, check_finite=False)
        if np.any(np.diag(self.lu_piv[0]) == 0):
            raise RuntimeError("dgetrf failed")

    def solve(self, B):
        if B.shape[1] == 0:
            return B
        return sla.lu_solve(self.lu_piv, B, overwrite_b=True, check_finite=False)


def solve_dense(A, B):
    """solve_dense, src/utils/linalg.hpp:53-58 (dgesv on a copy of A)."""
    return sla.solve(A, B, check_finite=False)


def deflate(area, normal, p0, np_, A, off, mu):
    """Tube::deflate, src/swimmer.hpp:64-74: A(off+3i+a, off+3p+b) += sc n_i[a] n_p[b] A_p."""
    At = 0.0
    for p in range(p0, p0 + np_):
        At += area[p]
    sc = 1.0 / (EIGHT_PI * mu * math.sqrt(At))
    nrm = normal[p0:p0 + np_]
    u = (sc * nrm).reshape(-1)                              # sc * n_i[a]
    w = np.repeat(area[p0:p0 + np_], 3)                     # A_p per column
    A[off:off + 3 * np_, off:off + 3 * np_] += np.outer(u, nrm.reshape(-1)) * w[None, :]


class Tube:
    """Tube, src/swimmer.hpp:33-75: SL tube-tube block, deflated, LU-factorised once."""

    def __init__(self, R, L, a, n_theta, n_r=4, mu=1.0, mesh=None):
        self.R, self.L, self.mu = R, L, mu
        self.mesh = mesh if mesh is not None else swimmer_tube_mesh(R, L, a, n_theta, n_r)
        self.Nt = self.mesh.P()
        self.t_assembly = 0.0
        self.t_factor = 0.0
        self._factor()

    def _factor(self):
        t0 = time.perf_counter()
        M = self.mesh
        T, _, _ = block(M, M.centroid, self_src=np.arange(self.Nt), sl_div=EIGHT_PI * self.mu)
        deflate(M.area, M.normal, 0, self.Nt, T, 0, self.mu)
        self.T = T.copy(order="F") if keep_T else None
        t1 = time.perf_counter()
        self.lu = LUFactor()
        self.lu.factor(T)
        self.t_assembly = t1 - t0
        self.t_factor = time.perf_counter() - t1


keep_T = False


@dataclass
class SwimmerSolution:
    U: np.ndarray
    Om: np.ndarray
    f_sphere: np.ndarray
    f_tube: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))


class SwimmerSolver:
    """SwimmerSolver, src/swimmer.hpp:83-289."""

    def __init__(self, tube, sphere, center, mu=1.0):
        self.tube = tube
        self.S = sphere
        self.xc = np.asarray(center, dtype=np.float64)
        self.mu = mu
        self.Ns = sphere.P()
        self.Nt = tube.Nt if tube is not None else 0
        self.t_assembly = 0.0
        self.t_schur = 0.0
        self.phase = {}
        self._build()

    def slip(self, e, B1, B2):
        """src/swimmer.hpp:100-104."""
        return np.array([squirmer_slip(self.S.centroid[p], self.xc, e, B1, B2) for p in range(self.Ns)])

    # src/swimmer.hpp:180-251
    def _build(self):
        t0 = time.perf_counter()
        Ns, Nt, mu = self.Ns, self.Nt, self.mu
        ns, nt = 3 * Ns, 3 * Nt
        S = self.S
        dls = np.zeros((Ns, 9))
        if self.tube is not None:
            tm = self.tube.mesh
            B, AuBt, _ = block(S, tm.centroid, sl_div=EIGHT_PI * mu, dl_div=EIGHT_PI)
            D, _, dls = block(tm, S.centroid, sl_div=EIGHT_PI * mu, dlsum=True)
        else:
            B = np.zeros((0, ns), order="F")
            AuBt = np.zeros((0, ns), order="F")
            D = np.zeros((ns, 0), order="F")
        # one running DL sum per sphere row: tube panels first, then sphere panels
        E, AuBs, dlsum = block(S, S.centroid, self_src=np.arange(Ns), sl_div=EIGHT_PI * mu,
                               dl_div=EIGHT_PI, dlsum=True, dlsum_init=dls)
        t_blocks = time.perf_counter()
        for i in range(Ns):
            blk = dlsum[i].reshape(3, 3) / EIGHT_PI - (0.0 if self.tube is not None else np.eye(3))
            AuBs[3 * i:3 * i + 3, 3 * i:3 * i + 3] += blk
        deflate(S.area, S.normal, 0, Ns, E, 0, mu)
        K = np.zeros((ns, 6), order="F")
        r = S.centroid - self.xc
        for p in range(Ns):
            K[3 * p:3 * p + 3, 0:3] = np.eye(3)
            K[3 * p + 0, 4] = r[p, 2]; K[3 * p + 0, 5] = -r[p, 1]
            K[3 * p + 1, 3] = -r[p, 2]; K[3 * p + 1, 5] = r[p, 0]
            K[3 * p + 2, 3] = r[p, 1]; K[3 * p + 2, 4] = -r[p, 0]
        Ct = AuBt @ K
        Cs = AuBs @ K
        t1 = time.perf_counter()
        self.t_assembly = t1 - t0
        self.phase["coupling_assembly"] = t_blocks - t0
        self.Gs = E.copy(order="F")
        self.Gc = Cs.copy(order="F")
        if self.tube is not None:
            Y = np.asfortranarray(np.hstack([B, Ct]))
            Y = self.tube.lu.solve(Y)
            t2 = time.perf_counter()
            DY = D @ Y
            self.Gs -= DY[:, :ns]
            self.Gc -= DY[:, ns:]
            self.Y = Y
            self.phase["schur_getrs"] = t2 - t1
            self.phase["schur_gemm"] = time.perf_counter() - t2
        else:
            self.Y = None
        self.AuBt, self.AuBs, self.D, self.K = AuBt, AuBs, D, K
        self.E, self.Cs, self.Ct, self.B = E, Cs, Ct, B
        self.t_schur = time.perf_counter() - t1

    def bordered_matrix(self):
        """Constraint-bordered Schur matrix of solve_free, src/swimmer.hpp:107-125."""
        Ns = self.Ns
        ns, n = 3 * Ns, 3 * Ns + 6
        M = np.zeros((n, n), order="F")
        M[:ns, :ns] = self.Gs
        M[:ns, ns:] = self.Gc
        amax = np.max(np.abs(self.Gs)) if self.Gs.size else 0.0
        Amax = 0.0
        for A in self.S.area:
            Amax = max(Amax, 5.0 * A)
        sc = amax / Amax
        r = self.S.centroid - self.xc
        for p in range(Ns):
            A = self.S.area[p] * sc
            for a in range(3):
                M[ns + a, 3 * p + a] = A
            M[ns + 3, 3 * p + 1] = -A * r[p, 2]; M[ns + 3, 3 * p + 2] = A * r[p, 1]
            M[ns + 4, 3 * p + 0] = A * r[p, 2]; M[ns + 4, 3 * p + 2] = -A * r[p, 0]
            M[ns + 5, 3 * p + 0] = -A * r[p, 1]; M[ns + 5, 3 * p + 1] = A * r[p, 0]
        return M, sc

    def reduced_rhs(self, us):
        """src/swimmer.hpp:256-269."""
        u = np.asarray(us, dtype=np.float64).reshape(-1, 1)
        rs = -(self.AuBs @ u)
        if self.tube is not None:
            rt = -(self.AuBt @ u)
            rt = self.tube.lu.solve(np.asfortranarray(rt))
            rs = rs - self.D @ rt
        else:
            rt = np.zeros((0, 1))
        return rs[:, 0], rt

    def unpack(self, x, rt):
        """src/swimmer.hpp:274-288."""
        ns = 3 * self.Ns
        U = x[ns:ns + 3].copy()
        Om = x[ns + 3:ns + 6].copy()
        fs = x[:ns].reshape(-1, 3).copy()
        if self.tube is not None:
            ft = rt[:, 0] - self.Y @ x
            ftube = ft.reshape(-1, 3)
        else:
            ftube = np.zeros((0, 3))
        return SwimmerSolution(U, Om, fs, ftube)

    def solve_free(self, slips):
        """src/swimmer.hpp:107-139."""
        t0 = time.perf_counter()
        ns = 3 * self.Ns
        M, _ = self.bordered_matrix()
        nr = len(slips)
        Rh = np.zeros((ns + 6, nr), order="F")
        rts = []
        for k, s in enumerate(slips):
            rs, rt = self.reduced_rhs(s)
            Rh[:ns, k] = rs
            rts.append(rt)
        X = solve_dense(M, Rh)
        out = [self.unpack(X[:, k], rts[k]) for k in range(nr)]
        self.phase["solve_free"] = time.perf_counter() - t0
        return out

    def solve_prescribed(self, U, Om):
        """Towed rigid sphere, src/swimmer.hpp:142-153."""
        ns = 3 * self.Ns
        w = np.array(list(U) + list(Om), dtype=np.float64)
        rhs = -(self.Gc @ w)
        fs = solve_dense(self.Gs, rhs.reshape(-1, 1))[:, 0]
        x = np.concatenate([fs, w])
        rt = np.zeros((3 * self.Nt, 1))
        return self.unpack(x, rt)

    def force(self, s):
        """src/swimmer.hpp:156-160: F = -sum A f on the sphere."""
        F = np.zeros(3)
        for p in range(self.Ns):
            F = F - self.S.area[p] * s.f_sphere[p]
        return F

    def velocity(self, s, slip_field, pts):
        """src/swimmer.hpp:163-173."""
        allm = merge(self.tube.mesh, self.S) if self.tube is not None else self.S
        P = allm.P()
        u = np.zeros((P, 3))
        f = np.zeros((P, 3))
        Nt = self.Nt
        if Nt:
            f[:Nt] = s.f_tube
        r = self.S.centroid - self.xc
        u[Nt:] = s.U + np.cross(s.Om, r) + np.asarray(slip_field)
        f[Nt:] = s.f_sphere
        return velocity_at(allm, u, f, self.mu, pts)


class DuctFlow:
    """DuctFlow, src/duct_flow.hpp:22-109 (mixed traction/velocity unknowns)."""

    def __init__(self, R, L, n_theta, n_z, n_r, mu=1.0):
        self.R, self.L, self.mu = R, L, mu
        self.mesh = duct_mesh(R, L, n_theta, n_z, n_r)
        self._assemble()

    def u_known(self, p, b):
        return not (self.mesh.tags[p] != WALL and b == 2)

    def _known_mask(self):
        P = self.mesh.P()
        k = np.ones((P, 3), dtype=bool)
        k[self.mesh.tags != WALL, 2] = False
        return k.reshape(-1)

    def _assemble(self):
        """src/duct_flow.hpp:82-108."""
        M = self.mesh
        P = M.P()
        Af, Au, dls = block(M, M.centroid, self_src=np.arange(P), sl_div=EIGHT_PI * self.mu,
                            dl_div=EIGHT_PI, dlsum=True)
        known = self._known_mask()
        A = np.where(known[None, :], Af, Au)
        A = np.asfortranarray(A)
        for i in range(P):
            for a in range(3):
                for b in range(3):
                    if not self.u_known(i, b):
                        A[3 * i + a, 3 * i + b] += dls[i, 3 * a + b] / EIGHT_PI
        self.cap = np.nonzero(M.tags != WALL)[0]
        self.Fk = np.asfortranarray(Af[:, 3 * self.cap + 2])
        self.lu = LUFactor()
        self.lu.factor(A)

    def solve(self, p_in, p_out):
        """src/duct_flow.hpp:41-63."""
        M = self.mesh
        P = M.P()
        fk = np.array([p_in if M.tags[p] == INLET else -p_out for p in self.cap])
        rhs = -(self.Fk @ fk.reshape(-1, 1))
        rhs = self.lu.solve(np.asfortranarray(rhs))[:, 0]
        known = self._known_mask()
        u = np.where(known, 0.0, rhs).reshape(P, 3)
        f = np.where(known, rhs, 0.0).reshape(P, 3)
        Q = 0.0
        for k, p in enumerate(self.cap):
            f[p, 2] = fk[k]
            if M.tags[p] == OUTLET:
                Q += u[p, 2] * M.area[p]
        return u, f, Q

    def velocity(self, u, f, pts):
        return velocity_at(self.mesh, u, f, self.mu, pts)

    def exact_uz(self, r, dP):
        return dP / self.L * (self.R * self.R - r * r) / (4 * self.mu)

    def exact_Q(self, dP):
        return PI * self.R ** 4 * dP / self.L / (8 * self.mu)


@dataclass
class TrajectoryParams:
    """src/trajectory.hpp:23-36."""
    alpha: float = 0.0
    a_over_R: float = 0.3
    beta0: float = 0.5
    pitch0_deg: float = 0.0
    yaw0_deg: float = 0.0
    dt: float = 0.5
    t_max: float = 100.0
    L_over_R: float = 4.0
    n_theta: int = 30
    sphere_level: int = 2
    beta_stop: float = 0.95
    budget_s: float = 0.0


def run_trajectory(P, max_steps=None, n_r=4):
    """run_trajectory, src/trajectory.hpp:42-151 (no files; returns rows and status)."""
    a = 1.0
    R = a / P.a_over_R
    tube = Tube(R, P.L_over_R * R, a, P.n_theta, n_r)
    sph0 = icosphere(a, P.sphere_level)

    def evaluate(y):
        c = np.array([y[0], y[1], 0.0])
        e = np.array(y[3:6])
        e = (1.0 / math.sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2])) * e
        S = sph0.copy()
        S.translate(c)
        s = SwimmerSolver(tube, S, c)
        k = s.solve_free([s.slip(e, 1, 0), s.slip(e, 0, 1)])
        U = k[0].U + P.alpha * k[1].U
        Om = k[0].Om + P.alpha * k[1].Om
        de = np.cross(Om, e)
        return np.concatenate([U, de]), U, Om

    def beta_of(y):
        return math.hypot(y[0], y[1]) / (R - a)

    b = P.beta0 * (R - a)
    xi, chi = P.pitch0_deg * PI / 180, P.yaw0_deg * PI / 180
    y = np.array([-b, 0, 0, math.cos(chi) * math.sin(xi), math.sin(chi), math.cos(chi) * math.cos(xi)])
    t, step = 0.0, 0
    hist = []
    rows = []
    status = "finished"
    while t < P.t_max - 1e-9:
        if max_steps is not None and step >= max_steps:
            break
        if beta_of(y) > P.beta_stop:
            status = "wall contact"
            break
        f, U, Om = evaluate(y)
        rows.append([t, *y, *U, *Om, beta_of(y)])
        if len(hist) < 3:
            k2 = evaluate(y + P.dt / 2 * f)[0]
            k3 = evaluate(y + P.dt / 2 * k2)[0]
            k4 = evaluate(y + P.dt * k3)[0]
            y = y + P.dt / 6 * (f + 2 * k2 + 2 * k3 + k4)
        else:
            f1, f2, f3 = hist[2], hist[1], hist[0]
            y = y + P.dt / 24 * (55 * f - 59 * f1 + 37 * f2 - 9 * f3)
        n = math.sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5])
        y[3:6] /= n
        hist.append(f)
        if len(hist) > 3:
            hist.pop(0)
        t += P.dt
        step += 1
    return np.array(rows), status, y
