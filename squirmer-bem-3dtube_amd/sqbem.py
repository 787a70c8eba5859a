"""Python binding of libsqbem (the MI355X squirmer-in-tube BEM library) over its C ABI.

Mirrors the reference's solver interface (src/swimmer.hpp, src/duct_flow.hpp,
src/utils/field.hpp, src/utils/mesh.hpp): same class and method names, argument meaning
and error texts. Every numerical step runs in libsqbem on the GPU; importing this module
fails loudly when the library is missing or the device is unavailable, there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsqbem.so")

FP64, FP32_IR, FP16_IR, BF16_IR = 0, 1, 2, 3
PRECISION_NAMES = {FP64: "fp64", FP32_IR: "fp32+ir", FP16_IR: "fp16+ir", BF16_IR: "bf16+ir"}
PI = 3.14159265358979323846
EIGHT_PI = 8.0 * PI
WALL, INLET, OUTLET, SPHERE = 0, 1, 2, 10

_vp, _i, _d, _l = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
_f = ctypes.c_float
_SIGS = {
    "sqbem_last_error": (ctypes.c_char_p, []),
    "sqbem_init": (_i, [_i]),
    "sqbem_set_stream": (_i, [_vp]),
    "sqbem_sync": (_i, []),
    "sqbem_device_info": (_i, [ctypes.c_char_p, _i, _vp, _vp]),
    "sqbem_profile": (_i, [_i]),
    "sqbem_profile_get": (_i, [_i, _vp, _vp, _vp, _vp]),
    "sqbem_tune": (_i, [_i]),
    "sqbem_trim": (_i, [_vp]),
    "sqbem_tune_info": (_i, [_vp, _vp, _vp]),
    "sqbem_tune_save": (_i, [ctypes.c_char_p]),
    "sqbem_tune_load": (_i, [ctypes.c_char_p, _vp]),
    "sqbem_mesh_swimmer_tube": (_vp, [_d, _d, _d, _i, _i]),
    "sqbem_mesh_tube": (_vp, [_d, _vp, _i, _i, _i]),
    "sqbem_mesh_duct": (_vp, [_d, _d, _i, _i, _i]),
    "sqbem_mesh_icosphere": (_vp, [_d, _i, _d, _d, _d]),
    "sqbem_mesh_from_arrays": (_vp, [_vp, _i, _vp, _vp, _i]),
    "sqbem_mesh_merge": (_vp, [_vp, _vp]),
    "sqbem_mesh_translate": (_i, [_vp, _d, _d, _d]),
    "sqbem_mesh_sizes": (_i, [_vp, _vp, _vp]),
    "sqbem_mesh_get": (_i, [_vp] * 9),
    "sqbem_mesh_free": (None, [_vp]),
    "sqbem_dgemm": (_i, [_i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i]),
    "sqbem_sgemm": (_i, [_i, _i, _i, _f, _vp, _i, _vp, _i, _f, _vp, _i]),
    "sqbem_gemm16": (_i, [_i, _i, _i, _i, _f, _vp, _i, _vp, _i, _f, _vp, _i]),
    "sqbem_gemm16_noalloc": (_i, [_i, _i, _i, _i, _f, _vp, _i, _vp, _i, _f, _vp, _i]),
    "sqbem_profile_named": (_i, [ctypes.c_char_p, _i]),
    "sqbem_gemm16_noalloc_bytes": (ctypes.c_ulong, [_i, _i, _i, _i]),
    "sqbem_dgetrf": (_i, [_i, _vp, _i, _vp, _vp]),
    "sqbem_dgetrs": (_i, [_i, _i, _vp, _i, _vp, _vp, _i]),
    "sqbem_dgetrs_trans": (_i, [_i, _i, _vp, _i, _vp, _vp, _i]),
    "sqbem_dgesv": (_i, [_i, _i, _vp, _i, _vp, _vp, _i, _vp]),
    "sqbem_sgetrf": (_i, [_i, _i, _vp, _i, _vp, _vp]),
    "sqbem_sgetrs": (_i, [_i, _i, _vp, _i, _vp, _vp, _i]),
    "sqbem_host_dgemm": (_i, [_i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i]),
    "sqbem_host_dgesv": (_i, [_i, _i, _vp, _i, _vp, _i]),
    "sqbem_lu_create": (_vp, [_i, _vp, _i, _vp]),
    "sqbem_lu_solve": (_i, [_vp, _i, _vp, _i]),
    "sqbem_lu_get": (_i, [_vp, _vp, _vp]),
    "sqbem_lu_size": (_i, [_vp]),
    "sqbem_lu_free": (None, [_vp]),
    "sqbem_dmesh_upload": (_vp, [_vp]),
    "sqbem_dmesh_free": (None, [_vp]),
    "sqbem_assemble": (_i, [_vp, _i, _i, _vp, _i, _vp, _d, _vp, _l, _d, _vp, _l, _vp]),
    "sqbem_panel_integrals": (_i, [_vp, _i, _vp, _i, _vp, _vp]),
    "sqbem_tube_create": (_vp, [_vp, _d, _i]),
    "sqbem_tube_info": (_i, [_vp, _vp, _vp, _vp]),
    "sqbem_tube_lu": (_vp, [_vp]),
    "sqbem_tube_free": (None, [_vp]),
    "sqbem_swimmer_create": (_vp, [_vp, _vp, _vp, _d, _i]),
    "sqbem_sphere_create": (_vp, [_vp, _vp, _d]),
    "sqbem_sphere_free": (None, [_vp]),
    "sqbem_swimmer_create_moved": (_vp, [_vp, _vp, _vp, _i]),
    "sqbem_swimmer_solve_free": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "sqbem_swimmer_solve_prescribed": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "sqbem_swimmer_times": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "sqbem_swimmer_free": (None, [_vp]),
    "sqbem_swimmer_gradient": (_i, [_vp, _vp, _d, _d, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sqbem_trajectory_gradient": (_i, [_d, _d, _d, _d, _d, _d, _i, _i, _d, _i, _i, _i, _i, _i, _vp]),
    "sqbem_batch_solve": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _d, _i, _d, _vp, _vp, _vp, _vp, _vp]),
    "sqbem_batch_options": (_i, [_i]),
    "sqbem_getrf_path": (_i, []),
    "sqbem_coop": (_i, [_i]),
    "sqbem_velocity_at": (_i, [_vp, _vp, _vp, _d, _vp, _i, _vp]),
    "sqbem_duct_create": (_vp, [_vp, _d]),
    "sqbem_duct_solve": (_i, [_vp, _d, _d, _vp, _vp, _vp]),
    "sqbem_duct_free": (None, [_vp]),
    "sqbem_trajectory_run": (_i, [_d, _d, _d, _d, _d, _d, _d, _d, _i, _i, _d, _d, _i, _i,
                                  ctypes.c_char_p, _i, ctypes.c_char_p, _i]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libsqbem.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libsqbem.so not found at {path}: run __graft_entry__.build() first")
    L = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue  # symbols absent from an older build are reported by the export test
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def lib():
    return load()


def last_error() -> str:
    return lib().sqbem_last_error().decode()


def check(rc: int):
    if rc != 0:
        raise RuntimeError(last_error())


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(a.data_ptr())  # torch tensor on the device


def _obj(h):
    if not h:
        raise RuntimeError(last_error())
    return h


def init(device: int = 0):
    check(lib().sqbem_init(device))


def sync():
    check(lib().sqbem_sync())


def device_info():
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int()
    gb = ctypes.c_double()
    check(lib().sqbem_device_info(name, 256, ctypes.byref(cu), ctypes.byref(gb)))
    return name.value.decode(), cu.value, gb.value


TUNE_TILES = ["128x128", "128x256", "256x128", "256x256", "128x128/pf2", "128x256/pf2", "256x128/pf2",
              "256x256/dma32x4", "256x128/dma32x4", "128x128/dma64x4", "hipblaslt0", "hipblaslt1",
              "hipblaslt2", "hipblaslt3"]
PROF_CLASSES = {"dgemm": 0, "sgemm": 1, "gemm16": 2, "panel": 3, "trsm": 4, "assemble": 5, "gemv": 6, "laswp": 7,
                "gemm16q": 8, "trsv": 9}
PROF_CLASSES.update({"gemm16q_" + t.replace("/", "_"): 10 + c for c, t in enumerate(TUNE_TILES)})


def profile(enable=True, classes=None):
    """Reset and enable (or disable) per-kernel HIP-event timing inside libsqbem, for all
    kernel classes or for the named ones."""
    if not enable:
        mask = 0
    elif classes is None:
        mask = -1
    else:
        mask = 0
        for c in classes:
            mask |= 1 << PROF_CLASSES[c]
    check(lib().sqbem_profile(mask))


def profile_named():
    """Kernel-time totals per named kernel (the library GEMM's kernels): {name: dict}."""
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().sqbem_profile_named(buf, len(buf)))
    out = {}
    for ln in buf.value.decode().splitlines():
        name, l, f, b, ms = ln.split("\t")
        out[name] = {"launches": int(l), "flops": float(f), "bytes": float(b), "ms": float(ms)}
    return out


def profile_get(name):
    n, fl, by, ms = ctypes.c_long(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    check(lib().sqbem_profile_get(PROF_CLASSES[name], ctypes.byref(n), ctypes.byref(fl), ctypes.byref(by),
                                  ctypes.byref(ms)))
    return {"launches": n.value, "flops": fl.value, "bytes": by.value, "ms": ms.value}


# ------------------------------------------------------------------ meshes
def trim() -> float:
    """Return the device memory the library keeps for reuse to the driver; GB released."""
    gb = ctypes.c_double()
    check(lib().sqbem_trim(ctypes.byref(gb)))
    return gb.value


def tune(mode: int = 1):
    """Tile autotuner of the 16-bit GEMMs: 1 learns per shape class from live calls, 0 uses
    the fixed rule; either clears what was learned."""
    check(lib().sqbem_tune(int(mode)))


TUNE_TABLE = os.path.join(HERE, "tune_gfx950.txt")  # learned on an MI355X by bench.py --tune-save


def tune_save(path: str):
    check(lib().sqbem_tune_save(path.encode()))


def tune_load(path: str = TUNE_TABLE) -> int:
    """Adds the decided classes of a saved tile table; returns how many were read."""
    n = ctypes.c_int()
    check(lib().sqbem_tune_load(path.encode(), ctypes.byref(n)))
    return n.value


def tune_info() -> dict:
    cl, dec = ctypes.c_int(), ctypes.c_int()
    counts = np.zeros(len(TUNE_TILES), dtype=np.int32)
    check(lib().sqbem_tune_info(ctypes.byref(cl), ctypes.byref(dec), _ptr(counts)))
    return {"classes": cl.value, "decided": dec.value, "tiles": dict(zip(TUNE_TILES, counts.tolist()))}


class Mesh:
    """bem::Mesh (src/utils/mesh.hpp:50-91) owned by libsqbem; geometry read back as numpy."""

    _GEOMETRY = ("nodes", "tris", "tags", "V", "normal", "centroid", "area", "h")

    def __init__(self, handle):
        self.h_ = _obj(handle)

    def __getattr__(self, name):
        # the geometry arrays are read back from the library on first access (a tube or
        # swimmer built from the mesh needs only its handle)
        if name in Mesh._GEOMETRY and "h_" in self.__dict__:
            self._load()
            return self.__dict__[name]
        raise AttributeError(name)

    def _load(self):
        nn, npn = ctypes.c_int(), ctypes.c_int()
        check(lib().sqbem_mesh_sizes(self.h_, ctypes.byref(nn), ctypes.byref(npn)))
        nn, npn = nn.value, npn.value
        self.nodes = np.zeros((nn, 3))
        self.tris = np.zeros((npn, 3), dtype=np.int32)
        self.tags = np.zeros(npn, dtype=np.int32)
        self.V = np.zeros((npn, 3, 3))
        self.normal = np.zeros((npn, 3))
        self.centroid = np.zeros((npn, 3))
        self.area = np.zeros(npn)
        self.h = np.zeros(npn)
        check(lib().sqbem_mesh_get(self.h_, _ptr(self.nodes), _ptr(self.tris), _ptr(self.tags), _ptr(self.V),
                                   _ptr(self.normal), _ptr(self.centroid), _ptr(self.area), _ptr(self.h)))

    def __del__(self):
        try:
            if getattr(self, "h_", None) and _lib is not None:
                _lib.sqbem_mesh_free(self.h_)
        except Exception:
            pass

    def P(self):
        return len(self.tris)

    def total_area(self):
        s = 0.0
        for A in self.area:
            s += A
        return s

    def translate(self, d):
        check(lib().sqbem_mesh_translate(self.h_, float(d[0]), float(d[1]), float(d[2])))
        self._load()

    def copy(self):
        return Mesh(lib().sqbem_mesh_from_arrays(_ptr(np.ascontiguousarray(self.nodes)), len(self.nodes),
                                                 _ptr(np.ascontiguousarray(self.tris)),
                                                 _ptr(np.ascontiguousarray(self.tags)), self.P()))


def swimmer_tube_mesh(R, L, a, n_theta, n_r=4):
    return Mesh(lib().sqbem_mesh_swimmer_tube(R, L, a, n_theta, n_r))


def tube_mesh(R, z, n_theta, n_r):
    z = np.ascontiguousarray(z, dtype=np.float64)
    return Mesh(lib().sqbem_mesh_tube(R, _ptr(z), len(z), n_theta, n_r))


def duct_mesh(R, L, n_theta, n_z, n_r):
    return Mesh(lib().sqbem_mesh_duct(R, L, n_theta, n_z, n_r))


def icosphere(a, level, center=(0.0, 0.0, 0.0)):
    return Mesh(lib().sqbem_mesh_icosphere(a, level, float(center[0]), float(center[1]), float(center[2])))


def merge(a, b):
    return Mesh(lib().sqbem_mesh_merge(a.h_, b.h_))


# ------------------------------------------------------------------ panel integrals
def panel_integrals(M, p, x0, self_=False):
    """panel_integrals (src/utils/kernels.hpp:67-82) for one pair, on the device."""
    SL, DL = np.zeros(9), np.zeros(9)
    x = np.ascontiguousarray(x0, dtype=np.float64)
    check(lib().sqbem_panel_integrals(M.h_, p, _ptr(x), int(self_), _ptr(SL), _ptr(DL)))
    return SL.reshape(3, 3), DL.reshape(3, 3)


class DeviceMesh:
    def __init__(self, M):
        self.h_ = _obj(lib().sqbem_dmesh_upload(M.h_))
        self.np = M.P()

    def __del__(self):
        try:
            if getattr(self, "h_", None) and _lib is not None:
                _lib.sqbem_dmesh_free(self.h_)
        except Exception:
            pass


def assemble(dmesh, x0, self_src=None, p0=0, np_=None, sl_div=1.0, SL=None, dl_div=1.0, DL=None, dlsum=None):
    """Device block assembly (sqbem_assemble) on torch device tensors."""
    np_ = dmesh.np - p0 if np_ is None else np_
    nt = x0.numel() // 3
    ld_sl = SL.stride(1) if SL is not None else 0
    ld_dl = DL.stride(1) if DL is not None else 0
    check(lib().sqbem_assemble(dmesh.h_, p0, np_, _ptr(x0), nt, _ptr(self_src), sl_div, _ptr(SL), ld_sl, dl_div,
                               _ptr(DL), ld_dl, _ptr(dlsum)))


# ------------------------------------------------------------------ dense algebra (device)
def _ld(t):
    # column-major view: torch tensor of shape (m, n) with stride (1, ld)
    assert t.stride(0) == 1, "expected a column-major (Fortran-ordered) device tensor"
    return t.stride(1) if t.dim() == 2 and t.shape[1] > 1 else max(1, t.shape[0])


def dgemm(alpha, A, B, beta, C):
    m, k = A.shape
    n = B.shape[1]
    check(lib().sqbem_dgemm(m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta, _ptr(C), _ld(C)))


def sgemm(alpha, A, B, beta, C):
    m, k = A.shape
    n = B.shape[1]
    check(lib().sqbem_sgemm(m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta, _ptr(C), _ld(C)))


def gemm16(kind, alpha, A, B, beta, C):
    m, k = A.shape
    n = B.shape[1]
    check(lib().sqbem_gemm16(kind, m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta, _ptr(C), _ld(C)))


def gemm16_noalloc(kind, alpha, A, B, beta, C):
    """sqbem_gemm16 with the library's fixed per-device scratch (allocates nothing)."""
    m, k = A.shape
    n = B.shape[1]
    check(lib().sqbem_gemm16_noalloc(kind, m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta, _ptr(C), _ld(C)))


def dgetrf(A, ipiv):
    info = ctypes.c_int(0)
    rc = lib().sqbem_dgetrf(A.shape[0], _ptr(A), _ld(A), _ptr(ipiv), ctypes.byref(info))
    if rc != 0 and info.value == 0:
        check(rc)
    return info.value


def dgetrs(LU, ipiv, B, trans=False):
    fn = lib().sqbem_dgetrs_trans if trans else lib().sqbem_dgetrs
    check(fn(LU.shape[0], B.shape[1], _ptr(LU), _ld(LU), _ptr(ipiv), _ptr(B), _ld(B)))


def sgetrf(kind, A, ipiv):
    info = ctypes.c_int(0)
    rc = lib().sqbem_sgetrf(kind, A.shape[0], _ptr(A), _ld(A), _ptr(ipiv), ctypes.byref(info))
    if rc != 0 and info.value == 0:
        check(rc)
    return info.value


def sgetrs(LU, ipiv, B):
    check(lib().sqbem_sgetrs(LU.shape[0], B.shape[1], _ptr(LU), _ld(LU), _ptr(ipiv), _ptr(B), _ld(B)))


# host-matrix API (src/utils/linalg.hpp:42-77)
def gemm(alpha, A, B, beta, C):
    """C = alpha A B + beta C on numpy Fortran arrays (in place on C)."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    assert C.flags.f_contiguous and C.dtype == np.float64
    if A.shape[1] != B.shape[0] or C.shape[0] != A.shape[0] or C.shape[1] != B.shape[1]:
        raise RuntimeError("gemm: dimension mismatch")
    m, k = A.shape
    n = B.shape[1]
    check(lib().sqbem_host_dgemm(m, n, k, alpha, _ptr(A), max(1, m), _ptr(B), max(1, k), beta, _ptr(C), max(1, m)))


def solve_dense(A, B):
    """B <- A^-1 B (A unchanged), in place on the Fortran array B."""
    A = np.asfortranarray(A, dtype=np.float64)
    assert B.flags.f_contiguous and B.dtype == np.float64
    n = A.shape[0]
    check(lib().sqbem_host_dgesv(n, B.shape[1], _ptr(A), max(1, n), _ptr(B), max(1, n)))


class LUFactor:
    """LUFactor (src/utils/linalg.hpp:62-78) with the factors resident in HBM."""

    def __init__(self, handle=None, owner=None):
        self.h_ = handle
        self._owner = owner  # keeps a tube alive for a borrowed view
        self._owned = handle is not None and owner is None

    def factor(self, A):
        A = np.asfortranarray(A, dtype=np.float64)
        info = ctypes.c_int(0)
        self.h_ = _obj(lib().sqbem_lu_create(A.shape[0], _ptr(A), max(1, A.shape[0]), ctypes.byref(info)))
        self._owned = True

    def solve(self, B):
        """B <- A^-1 B in place (B: numpy Fortran array, n x nrhs)."""
        assert B.flags.f_contiguous and B.dtype == np.float64
        if B.shape[1] == 0:
            return B
        check(lib().sqbem_lu_solve(self.h_, B.shape[1], _ptr(B), B.shape[0]))
        return B

    def get(self):
        n = lib().sqbem_lu_size(self.h_)
        LU = np.zeros((n, n), order="F")
        piv = np.zeros(n, dtype=np.int32)
        check(lib().sqbem_lu_get(self.h_, _ptr(LU), _ptr(piv)))
        return LU, piv

    def __del__(self):
        try:
            if self._owned and self.h_ and _lib is not None:
                _lib.sqbem_lu_free(self.h_)
        except Exception:
            pass


# ------------------------------------------------------------------ solvers
class Tube:
    """Tube (src/swimmer.hpp:33-75): tube-tube block assembled, deflated and factorised once
    on the device."""

    def __init__(self, R, L, a, n_theta, n_r=4, mu=1.0, precision=FP64, mesh=None):
        self.R, self.L, self.mu, self.precision = R, L, mu, precision
        self.mesh = mesh if mesh is not None else swimmer_tube_mesh(R, L, a, n_theta, n_r)
        # returns while the device may still be factoring: the factors are waited for where
        # first used (a swimmer on this tube, lu, t_factor)
        self.h_ = _obj(lib().sqbem_tube_create(self.mesh.h_, mu, precision))
        Nt = ctypes.c_int()
        check(lib().sqbem_tube_info(self.h_, ctypes.byref(Nt), None, None))
        self.Nt = Nt.value

    @property
    def lu(self):
        """Tube::lu: a view onto the factors that keeps this tube alive (made on each access,
        so the tube holds no reference to it and is freed as soon as the caller drops it)."""
        return LUFactor(lib().sqbem_tube_lu(self.h_), owner=self)

    def _times(self):
        ta, tf = ctypes.c_double(), ctypes.c_double()
        check(lib().sqbem_tube_info(self.h_, None, ctypes.byref(ta), ctypes.byref(tf)))
        return ta.value, tf.value

    @property
    def t_assembly(self):
        """Seconds assembling the tube-tube block."""
        return self._times()[0]

    @property
    def t_factor(self):
        """Seconds from issuing the factorisation to its end (waits for it)."""
        return self._times()[1]

    def __del__(self):
        try:
            if getattr(self, "h_", None) and _lib is not None:
                _lib.sqbem_tube_free(self.h_)
        except Exception:
            pass


@dataclass
class SwimmerSolution:
    """src/swimmer.hpp:78-81."""
    U: np.ndarray
    Om: np.ndarray
    f_sphere: np.ndarray
    f_tube: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))


def squirmer_slip(x, center, e, B1, B2):
    """squirmer_slip (src/utils/kernels.hpp:107-111): (B1 + B2 cos th)(cos th rhat - e)."""
    x = np.asarray(x, dtype=np.float64)
    r = x - np.asarray(center, dtype=np.float64)
    rh = (1.0 / np.sqrt(np.sum(r * r, axis=-1, keepdims=True))) * r
    e = np.asarray(e, dtype=np.float64)
    ct = np.sum(rh * e, axis=-1, keepdims=True)
    return (B1 + B2 * ct) * (ct * rh - e)


class SphereBlocks:
    """Translation-invariant sphere-sphere blocks of a sphere mesh centred at `center`,
    assembled once on the device (an addition; see SwimmerSolver.moved)."""

    def __init__(self, sphere, center, mu=1.0):
        self.S = sphere
        self.xc = np.asarray(center, dtype=np.float64).copy()
        self.mu = mu
        c = np.ascontiguousarray(self.xc)
        self.h_ = _obj(lib().sqbem_sphere_create(sphere.h_, _ptr(c), mu))

    def __del__(self):
        try:
            if getattr(self, "h_", None) and _lib is not None:
                _lib.sqbem_sphere_free(self.h_)
        except Exception:
            pass


class SwimmerSolver:
    """SwimmerSolver (src/swimmer.hpp:83-289) on the device: couplings, Schur complement
    against the tube factors, force- and torque-free bordered solve."""

    def __init__(self, tube, sphere, center, mu=1.0, precision=None, _moved=None):
        self.tube = tube
        self.xc = np.asarray(center, dtype=np.float64).copy()
        self.mu = mu
        self.precision = precision if precision is not None else (tube.precision if tube is not None else FP64)
        self.Nt = tube.Nt if tube is not None else 0
        if _moved is None:
            self.S = sphere
            c = np.ascontiguousarray(self.xc)
            self.h_ = _obj(lib().sqbem_swimmer_create(tube.h_ if tube is not None else None, sphere.h_, _ptr(c), mu,
                                                      self.precision))
        else:
            blocks, shift = _moved
            self._blocks = blocks  # the template outlives the solver built from it
            self.S = sphere
            d = np.ascontiguousarray(shift, dtype=np.float64)
            self.h_ = _obj(lib().sqbem_swimmer_create_moved(tube.h_ if tube is not None else None, blocks.h_, _ptr(d),
                                                            self.precision))
        self.Ns = self.S.P()
        self._times()

    @classmethod
    def moved(cls, tube, blocks, shift, precision=None):
        """The sphere of `blocks` translated by `shift`, its sphere-sphere blocks copied from
        the template (an addition; equals the plain constructor on the translated mesh to
        rounding)."""
        shift = np.asarray(shift, dtype=np.float64)
        S = blocks.S.copy()
        S.translate(shift)
        return cls(tube, S, blocks.xc + shift, blocks.mu, precision, _moved=(blocks, shift))

    def _times(self):
        ta, ts, it, res = ctypes.c_double(), ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        check(lib().sqbem_swimmer_times(self.h_, ctypes.byref(ta), ctypes.byref(ts), ctypes.byref(it),
                                        ctypes.byref(res)))
        self.t_assembly, self.t_schur, self.ir_iters, self.ir_residual = ta.value, ts.value, it.value, res.value

    def __del__(self):
        try:
            if getattr(self, "h_", None) and _lib is not None:
                _lib.sqbem_swimmer_free(self.h_)
        except Exception:
            pass

    def slip(self, e, B1, B2):
        return squirmer_slip(self.S.centroid, self.xc, e, B1, B2)

    def solve_free(self, slips):
        k = len(slips)
        u = np.ascontiguousarray(np.stack([np.asarray(s, dtype=np.float64).reshape(-1) for s in slips]))
        U = np.zeros((k, 3))
        Om = np.zeros((k, 3))
        fs = np.zeros((k, self.Ns, 3))
        ft = np.zeros((k, max(self.Nt, 0), 3))
        check(lib().sqbem_swimmer_solve_free(self.h_, k, _ptr(u), _ptr(U), _ptr(Om), _ptr(fs),
                                             _ptr(ft) if self.Nt else None))
        self._times()
        return [SwimmerSolution(U[j].copy(), Om[j].copy(), fs[j].copy(), ft[j].copy()) for j in range(k)]

    def solve_prescribed(self, U, Om):
        Uv = np.ascontiguousarray(U, dtype=np.float64)
        Ov = np.ascontiguousarray(Om, dtype=np.float64)
        fs = np.zeros((self.Ns, 3))
        ft = np.zeros((max(self.Nt, 0), 3))
        check(lib().sqbem_swimmer_solve_prescribed(self.h_, _ptr(Uv), _ptr(Ov), _ptr(fs), _ptr(ft) if self.Nt else None))
        self._times()
        return SwimmerSolution(Uv.copy(), Ov.copy(), fs, ft)

    def gradient(self, e, B1, B2, g):
        """Adjoint gradients (an addition; the reference has no derivatives) of the
        functionals J_j = g[j] . (U, Omega) of the free-swimming solution with slip
        (e, B1, B2): returns U, Om and per functional dJ/dx_c, dJ/de, dJ/d(B1, B2) and
        dJ/d(slip field). FP64 swimmers."""
        gw = np.ascontiguousarray(np.atleast_2d(np.asarray(g, dtype=np.float64)))
        k = gw.shape[0]
        ev = np.ascontiguousarray(e, dtype=np.float64)
        U, Om = np.zeros(3), np.zeros(3)
        dxc, de, dB = np.zeros((k, 3)), np.zeros((k, 3)), np.zeros((k, 2))
        du = np.zeros((k, self.Ns, 3))
        check(lib().sqbem_swimmer_gradient(self.h_, _ptr(ev), float(B1), float(B2), k, _ptr(gw), _ptr(U), _ptr(Om),
                                           _ptr(dxc), _ptr(de), _ptr(dB), _ptr(du)))
        return {"U": U, "Om": Om, "dxc": dxc, "de": de, "dB": dB, "du": du}

    def force(self, s):
        """Force on the sphere, F = -sum A f (src/swimmer.hpp:156-160)."""
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


@dataclass
class BatchInstance:
    """One configuration of a batched solve: tube and sphere meshes (the sphere placed at
    its centre), centre, squirming direction e and modes B1, B2."""
    tube: Mesh
    sphere: Mesh
    center: np.ndarray
    e: np.ndarray
    B1: float = 1.0
    B2: float = 0.0


@dataclass
class BatchResult:
    U: np.ndarray       # (n, 3)
    Om: np.ndarray      # (n, 3)
    iters: np.ndarray   # (n,) refinement steps (0 in FP64 mode)
    resid: np.ndarray   # (n,) final relative residual of the full system
    stats: dict


def batch_solve(instances, mu=1.0, precision=FP64, mem_budget_gb=0.0) -> BatchResult:
    """Free-swimming U, Omega of many independent configurations in one call (an addition:
    the reference loops over SwimmerSolver, src/main.cpp:82-100). Each result equals
    SwimmerSolver(Tube(mesh=inst.tube), inst.sphere, inst.center).solve_free([slip(e, B1, B2)])
    to the solver's accuracy; identical tube meshes share one factorisation."""
    n = len(instances)
    tubes = (ctypes.c_void_p * max(n, 1))(*[q.tube.h_ if q.tube is not None else None for q in instances])
    spheres = (ctypes.c_void_p * max(n, 1))(*[q.sphere.h_ for q in instances])
    c = np.ascontiguousarray(np.array([np.asarray(q.center, dtype=np.float64) for q in instances]).reshape(-1, 3))
    e = np.ascontiguousarray(np.array([np.asarray(q.e, dtype=np.float64) for q in instances]).reshape(-1, 3))
    B = np.ascontiguousarray(np.array([[q.B1, q.B2] for q in instances], dtype=np.float64).reshape(-1, 2))
    U = np.zeros((n, 3))
    Om = np.zeros((n, 3))
    it = np.zeros(n, dtype=np.int32)
    res = np.zeros(n)
    st = np.zeros(12)
    check(lib().sqbem_batch_solve(n, ctypes.cast(tubes, ctypes.c_void_p), ctypes.cast(spheres, ctypes.c_void_p),
                                  _ptr(c), _ptr(e), _ptr(B), mu, precision, float(mem_budget_gb), _ptr(U), _ptr(Om),
                                  _ptr(it), _ptr(res), _ptr(st)))
    stats = {"tube_s": st[0], "assembly_s": st[1], "schur_s": st[2], "solve_s": st[3], "unique_tubes": int(st[4]),
             "chunks": int(st[5]), "max_iters": int(st[6]), "worst_residual": st[7], "pivoted": int(st[8]),
             "workspace_gb": st[9] / 1e9, "graph_replays": int(st[10]), "build_graphs": int(st[11])}
    return BatchResult(U, Om, it, res, stats)


def getrf_path():
    """How the most recent LU factorisation went: 0 partial pivoting throughout (orders below
    512), 1 without row interchanges (verified), 2 that attempt rejected, then pivoted."""
    return int(lib().sqbem_getrf_path())


def coop(enable):
    """1: the pivoted LU's panels use cooperative launches while the runtime accepts them
    (the default); 0: they take the barrier-free path (a test switch)."""
    check(lib().sqbem_coop(int(enable)))


def batch_options(use_graph=True):
    """use_graph: run each chunk's two build stages as captured HIP graphs, and capture its
    refinement step as a graph and replay it."""
    check(lib().sqbem_batch_options(1 if use_graph else 0))


def draw_batch_params(n=256, seed=0):
    """Random uneven configurations: a/R ~ U[0.1, 0.7], radial offset b/R ~ U[0, 0.8 (1 - a/R)],
    B2/B1 ~ U[-5, 5] (B1 = 1), e uniform on the unit sphere, sphere level from {2, 3}, tube
    n_theta from {16, 24, 32} with n_z = 2 n_theta, L/R = 4, n_r = n_theta / 4; numpy's
    default_rng(seed), drawn in that order."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.1, 0.7, n)
    b = rng.uniform(0.0, 1.0, n) * 0.8 * (1.0 - a)
    alpha = rng.uniform(-5.0, 5.0, n)
    e = rng.standard_normal((n, 3))
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    level = rng.choice([2, 3], n)
    nth = rng.choice([16, 24, 32], n)
    return [dict(a=float(a[i]), b=float(b[i]), alpha=float(alpha[i]), e=e[i].copy(), level=int(level[i]),
                 n_theta=int(nth[i]), n_z=2 * int(nth[i]), L=4.0, n_r=int(nth[i]) // 4) for i in range(n)]


def batch_instances(params, R=1.0, graded=False):
    """BatchInstance list for draw_batch_params output. graded=False: uniform axial nodes
    (n_z = 2 n_theta over [-L/2, L/2]), so the tube depends on n_theta only and the batch
    holds three distinct tubes; graded=True: the reference's swimmer_tube_mesh, graded
    around each sphere (one tube per configuration, n_z then follows from the grading)."""
    tubes = {}
    out = []
    for q in params:
        if graded:
            tm = swimmer_tube_mesh(R, q["L"] * R, q["a"] * R, q["n_theta"], q["n_r"])
        else:
            key = (q["n_theta"], q["n_z"], q["L"], q["n_r"])
            if key not in tubes:
                z = np.array([-q["L"] * R / 2 + q["L"] * R * k / q["n_z"] for k in range(q["n_z"] + 1)])
                tubes[key] = tube_mesh(R, z, q["n_theta"], q["n_r"])
            tm = tubes[key]
        c = np.array([q["b"] * R, 0.0, 0.0])
        out.append(BatchInstance(tm, icosphere(q["a"] * R, q["level"], c), c, np.asarray(q["e"]), 1.0, q["alpha"]))
    return out


def velocity_at(M, u, f, mu, pts):
    """velocity_at (src/utils/field.hpp:18-34) on the device."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(pts)
    check(lib().sqbem_velocity_at(M.h_, _ptr(u), _ptr(f), mu, _ptr(pts), len(pts), _ptr(out)))
    return out


class DuctFlow:
    """DuctFlow (src/duct_flow.hpp:22-109) on the device."""

    @dataclass
    class Solution:
        u: np.ndarray
        f: np.ndarray
        Q: float = 0.0

    def __init__(self, R, L, n_theta, n_z, n_r, mu=1.0):
        self.R, self.L, self.mu = R, L, mu
        self.mesh = duct_mesh(R, L, n_theta, n_z, n_r)
        self.h_ = _obj(lib().sqbem_duct_create(self.mesh.h_, mu))

    def __del__(self):
        try:
            if getattr(self, "h_", None) and _lib is not None:
                _lib.sqbem_duct_free(self.h_)
        except Exception:
            pass

    def u_known(self, p, b):
        return not (self.mesh.tags[p] != WALL and b == 2)

    def solve(self, p_in, p_out):
        P = self.mesh.P()
        u = np.zeros((P, 3))
        f = np.zeros((P, 3))
        Q = ctypes.c_double()
        check(lib().sqbem_duct_solve(self.h_, p_in, p_out, _ptr(u), _ptr(f), ctypes.byref(Q)))
        return DuctFlow.Solution(u, f, Q.value)

    def velocity(self, s, pts):
        return velocity_at(self.mesh, s.u, s.f, self.mu, pts)

    def exact_uz(self, r, dP):
        return dP / self.L * (self.R * self.R - r * r) / (4 * self.mu)

    def exact_Q(self, dP):
        return PI * self.R ** 4 * dP / self.L / (8 * self.mu)


@dataclass
class TrajectoryParams:
    """TrajectoryParams (src/trajectory.hpp:23-38); n_r and precision are extensions."""
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
    n_r: int = 4
    precision: int = FP64


def run_trajectory(P: TrajectoryParams, out_prefix: str, verbose: bool = False) -> str:
    """run_trajectory (src/trajectory.hpp:40-151) on the device: writes <out_prefix>.csv
    (and <out_prefix>.ckpt while paused); returns "finished", "wall contact" or "paused"."""
    status = ctypes.create_string_buffer(64)
    check(lib().sqbem_trajectory_run(P.alpha, P.a_over_R, P.beta0, P.pitch0_deg, P.yaw0_deg, P.dt,
                                     P.t_max, P.L_over_R, P.n_theta, P.sphere_level, P.beta_stop,
                                     P.budget_s, P.n_r, P.precision, out_prefix.encode(),
                                     1 if verbose else 0, status, len(status)))
    return status.value.decode()


def trajectory_gradient(P: TrajectoryParams, n_steps: int, ckpt_every: int = 50, adjoint: bool = True,
                        verbose: bool = False) -> dict:
    """Trajectory sensitivities (an addition): n_steps steps of run_trajectory's scheme, then
    d(beta_final, tilt_final_deg)/d(beta0, pitch0_deg) by the checkpointed discrete adjoint
    of the integrator (FP64). adjoint=False integrates only."""
    out = np.zeros(7)
    check(lib().sqbem_trajectory_gradient(P.alpha, P.a_over_R, P.beta0, P.pitch0_deg, P.yaw0_deg, P.dt, int(n_steps),
                                          int(ckpt_every), P.L_over_R, P.n_theta, P.sphere_level, P.n_r,
                                          1 if adjoint else 0, 1 if verbose else 0, _ptr(out)))
    return {"beta_final": out[0], "tilt_final_deg": out[1],
            "dbeta_dbeta0": out[2], "dbeta_dpitch0": out[3], "dtilt_dbeta0": out[4], "dtilt_dpitch0": out[5],
            "evaluations": int(out[6])}
