"""ctypes access to oracle/_ref/libref_solver.so: the reference's own FP64 solver headers
(src/swimmer.hpp, src/duct_flow.hpp, src/trajectory.hpp) built from the reference sources
by `make -C oracle refsolver` and linked against the image's LP64 OpenBLAS.

TEST INFRASTRUCTURE ONLY: the fixtures writer (tests/golden/make_ref_solver.py), the
oracle-pinning tests and bench.py's CPU-baseline leg use it; the product never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libref_solver.so")
_lib = None

_d, _i, _vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
_SIGS = {
    "ref_last_error": (ctypes.c_char_p, []),
    "ref_blas_threads": (None, [_i]),
    "ref_blas_get_threads": (_i, []),
    "ref_blas_config": (ctypes.c_char_p, []),
    "ref_tube": (_vp, [_d, _d, _d, _i, _i]),
    "ref_tube_split": (_vp, [_d, _d, _d, _i, _i, _vp, _i, _vp, _vp]),
    "ref_tube_np": (_i, [_vp]),
    "ref_tube_free": (None, [_vp]),
    "ref_swimmer": (_vp, [_vp, _d, _i, _vp]),
    "ref_swimmer_np": (_i, [_vp]),
    "ref_swimmer_times": (None, [_vp, _vp, _vp]),
    "ref_swimmer_free": (None, [_vp]),
    "ref_swimmer_solve_free": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ref_swimmer_towed": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ref_swimmer_velocity": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp, _i, _vp]),
    "ref_trajectory": (_i, [_d, _d, _d, _d, _d, _d, _d, _d, _i, _i, _d, _d, ctypes.c_char_p, _i, ctypes.c_char_p,
                            _i]),
    "ref_duct": (_i, [_d, _d, _i, _i, _i, _d, _d, _d, _vp, _vp, _vp, _vp]),
}


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"reference solver build missing: run `make -C {_HERE} refsolver` where the "
                               "reference sources are present")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().ref_last_error().decode())


def blas_threads(n=None):
    if n is not None:
        lib().ref_blas_threads(int(n))
    return lib().ref_blas_get_threads()


def blas_config():
    return lib().ref_blas_config().decode()


class Tube:
    """The reference's Tube. split=False: its own constructor (graded tube); split=True:
    the same object built in two timed halves (assembly, dgetrf), optionally on explicit
    axial nodes z (tube_mesh(R, z, n_theta, n_r))."""

    def __init__(self, R, L, a, n_theta, n_r=4, split=False, z=None):
        self.t_assembly = self.t_factor = 0.0
        if split or z is not None:
            ta, tf = ctypes.c_double(), ctypes.c_double()
            zz = None if z is None else np.ascontiguousarray(z, dtype=np.float64)
            h = lib().ref_tube_split(R, L, a, n_theta, n_r, _p(zz), 0 if zz is None else len(zz), ctypes.byref(ta),
                                     ctypes.byref(tf))
            self.t_assembly, self.t_factor = ta.value, tf.value
        else:
            h = lib().ref_tube(R, L, a, n_theta, n_r)
        if not h:
            raise RuntimeError(lib().ref_last_error().decode())
        self.h = h
        self.Nt = lib().ref_tube_np(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_tube_free(self.h)
            self.h = None


class Solution:
    def __init__(self, U, Om, fs, ft):
        self.U, self.Om, self.f_sphere, self.f_tube = U, Om, fs, ft


class SwimmerSolver:
    """The reference's SwimmerSolver(tube or None, icosphere(a, level, c), c)."""

    def __init__(self, tube, a, level, center):
        c = np.ascontiguousarray(center, dtype=np.float64)
        self.tube = tube
        self.h = lib().ref_swimmer(tube.h if tube is not None else None, a, level, _p(c))
        if not self.h:
            raise RuntimeError(lib().ref_last_error().decode())
        self.Ns = lib().ref_swimmer_np(self.h)
        self.Nt = tube.Nt if tube is not None else 0
        ta, ts = ctypes.c_double(), ctypes.c_double()
        lib().ref_swimmer_times(self.h, ctypes.byref(ta), ctypes.byref(ts))
        self.t_assembly, self.t_schur = ta.value, ts.value
        self.t_solve = 0.0

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_swimmer_free(self.h)
            self.h = None

    def solve_free(self, e, modes):
        """solve_free for slips slip(e, B1, B2), (B1, B2) in modes."""
        k = len(modes)
        e = np.ascontiguousarray(e, dtype=np.float64)
        B = np.ascontiguousarray(np.asarray(modes, dtype=np.float64).reshape(k, 2))
        U, Om = np.zeros((k, 3)), np.zeros((k, 3))
        fs = np.zeros((k, self.Ns, 3))
        ft = np.zeros((k, max(self.Nt, 1), 3))
        t = ctypes.c_double()
        _check(lib().ref_swimmer_solve_free(self.h, _p(e), k, _p(B), _p(U), _p(Om), _p(fs),
                                            _p(ft) if self.Nt else None, ctypes.byref(t)))
        self.t_solve = t.value
        return [Solution(U[j], Om[j], fs[j], ft[j][:self.Nt]) for j in range(k)]

    def towed(self, U, Om):
        """solve_prescribed(U, Om) and force(): (F, f_sphere, f_tube)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        Om = np.ascontiguousarray(Om, dtype=np.float64)
        F = np.zeros(3)
        fs = np.zeros((self.Ns, 3))
        ft = np.zeros((max(self.Nt, 1), 3))
        _check(lib().ref_swimmer_towed(self.h, _p(U), _p(Om), _p(F), _p(fs), _p(ft) if self.Nt else None))
        return F, fs, ft[:self.Nt]

    def velocity(self, sol, e, B1, B2, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(pts)
        arr = lambda x: np.ascontiguousarray(x, dtype=np.float64)
        ft = arr(sol.f_tube) if self.Nt else None
        _check(lib().ref_swimmer_velocity(self.h, _p(arr(sol.U)), _p(arr(sol.Om)), _p(arr(sol.f_sphere)), _p(ft),
                                          _p(arr(e)), B1, B2, _p(pts), len(pts), _p(out)))
        return out


def trajectory(out_prefix, alpha=0.0, a_over_R=0.3, beta0=0.5, pitch0_deg=0.0, yaw0_deg=0.0, dt=0.5, t_max=100.0,
               L_over_R=4.0, n_theta=30, sphere_level=2, beta_stop=0.95, budget_s=0.0, verbose=False):
    """run_trajectory (src/trajectory.hpp:42-151): writes out_prefix.csv (and .ckpt while
    paused); returns the status string."""
    st = ctypes.create_string_buffer(64)
    _check(lib().ref_trajectory(alpha, a_over_R, beta0, pitch0_deg, yaw0_deg, dt, t_max, L_over_R, n_theta,
                                sphere_level, beta_stop, budget_s, out_prefix.encode(), int(verbose), st, 64))
    return st.value.decode()


def duct(R, L, n_theta, n_z, n_r, mu, p_in, p_out):
    """DuctFlow(...).solve(p_in, p_out): (u [P, 3], f [P, 3], Q)."""
    P = 2 * n_theta * (n_z + 2 * n_r - 1)
    u, f = np.zeros((P, 3)), np.zeros((P, 3))
    Q, npan = ctypes.c_double(), ctypes.c_int()
    _check(lib().ref_duct(R, L, n_theta, n_z, n_r, mu, p_in, p_out, _p(u), _p(f), ctypes.byref(Q), ctypes.byref(npan)))
    assert npan.value == P
    return u, f, Q.value


def timed_cold_solve(R, L, a, n_theta, n_r, level, center, e, modes=((1.0, 0.0), (0.0, 1.0)), progress=None):
    """One cold FP64 solve by the reference's own code, phase by phase as its t_assembly /
    t_schur split them (src/swimmer.hpp:181, 235, 250) plus the tube's two halves: seconds of
    tube assembly, tube dgetrf, coupling assembly, Schur complement (dgetrs with 3 Ns + 6
    right-hand sides and D Y), bordered solve (solve_free: dgesv of the (3 Ns + 6) system
    and the unpacking), and the U, Omega of each mode."""
    import time

    say = progress or (lambda msg: None)
    t0 = time.perf_counter()
    tube = Tube(R, L, a, n_theta, n_r, split=True)
    say(f"tube {tube.Nt} panels: assembly {tube.t_assembly:.1f} s, dgetrf {tube.t_factor:.1f} s")
    s = SwimmerSolver(tube, a, level, center)
    say(f"swimmer {s.Ns} panels: coupling assembly {s.t_assembly:.1f} s, Schur {s.t_schur:.1f} s")
    sol = s.solve_free(e, list(modes))
    total = time.perf_counter() - t0
    say(f"bordered solve {s.t_solve:.1f} s, total {total:.1f} s")
    out = {"Nt": tube.Nt, "Ns": s.Ns, "dof": 3 * (tube.Nt + s.Ns) + 6,
           "tube_assembly_s": tube.t_assembly, "tube_factor_s": tube.t_factor,
           "coupling_assembly_s": s.t_assembly, "schur_s": s.t_schur, "bordered_solve_s": s.t_solve,
           "total_s": total, "U": [q.U.tolist() for q in sol], "Om": [q.Om.tolist() for q in sol]}
    del s, tube
    return out
