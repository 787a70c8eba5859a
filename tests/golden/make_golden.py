"""Writes the golden fixtures under tests/golden/ (run in the build container, where the
reference checkout and its reference build exist):

  * reference output data (CSV files the reference repository holds, produced by its own
    runs): copied verbatim, including all six trajectories and the second revision's run of
    the neutral trajectory (outputs/trajectories/neutral.csv, as neutral_rev2.csv), which
    shows how closely two runs of the reference agree;
  * vectors from the reference's own geometry/quadrature/kernel headers, compiled by
    oracle/Makefile into oracle/_ref/libref.so: meshes, rule levels and panel integrals for
    selected pairs of every rule, stored as .npz (no pickles).

    python tests/golden/make_golden.py
"""
import ctypes
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("SQBEM_REFERENCE", "/root/reference")
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")

CSV_FULL = [
    "outputs/reference/squirmer/axis.csv",
    "outputs/reference/squirmer/offaxis.csv",
    "outputs/reference/squirmer/reciprocal.csv",
    "outputs/reference/squirmer/blake.csv",
    "outputs/reference/squirmer/field.csv",
    "outputs/duct_flow/flow_rate.csv",
    "outputs/duct_flow/panels.csv",
    "outputs/duct_flow/profile.csv",
]
CSV_FULL += [
    "outputs/reference/trajectories/neutral.csv",
    "outputs/reference/trajectories/pusher.csv",
    "outputs/reference/trajectories/puller_a3.csv",
    "outputs/reference/trajectories/puller_a5.csv",
    "outputs/reference/traj3d/neutral_yaw15.csv",
    "outputs/reference/traj3d/puller_a3_yaw15.csv",
]
CSV_RENAMED = {"outputs/trajectories/neutral.csv": "neutral_rev2.csv"}
INPUTS = "inputs"  # the reference's key=value input files, copied verbatim


def copy_csvs():
    out = os.path.join(HERE, "reference_outputs")
    os.makedirs(out, exist_ok=True)
    for rel in CSV_FULL:
        shutil.copyfile(os.path.join(REF, rel), os.path.join(out, os.path.basename(rel)))
    for rel, name in CSV_RENAMED.items():
        shutil.copyfile(os.path.join(REF, rel), os.path.join(out, name))


def copy_inputs():
    out = os.path.join(HERE, "inputs")
    os.makedirs(out, exist_ok=True)
    for fn in sorted(os.listdir(os.path.join(REF, INPUTS))):
        if fn.endswith(".in"):
            shutil.copyfile(os.path.join(REF, INPUTS, fn), os.path.join(out, fn))


def ref_lib():
    L = ctypes.CDLL(REFLIB)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.ref_swimmer_tube_mesh.restype = vp
    L.ref_swimmer_tube_mesh.argtypes = [d, d, d, i, i]
    L.ref_duct_mesh.restype = vp
    L.ref_duct_mesh.argtypes = [d, d, i, i, i]
    L.ref_icosphere.restype = vp
    L.ref_icosphere.argtypes = [d, i, d, d, d]
    L.ref_mesh_nn.argtypes = [vp]
    L.ref_mesh_np.argtypes = [vp]
    L.ref_mesh_get.argtypes = [vp] * 9
    L.ref_levels.argtypes = [vp, i, i, vp, i, vp, vp]
    L.ref_panel_integrals.argtypes = [vp, i, vp, i, vp, vp]
    L.ref_rule.argtypes = [i, vp, vp]
    L.ref_gauss_legendre01.argtypes = [i, vp, vp]
    L.ref_graded_z.argtypes = [d, d, d, d, vp, i]
    L.ref_velocity_at.argtypes = [vp, vp, vp, d, vp, i, vp]
    return L


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def mesh_arrays(L, h):
    nn, npn = L.ref_mesh_nn(h), L.ref_mesh_np(h)
    out = dict(nodes=np.zeros((nn, 3)), tris=np.zeros((npn, 3), np.int32), tags=np.zeros(npn, np.int32),
               V=np.zeros((npn, 3, 3)), normal=np.zeros((npn, 3)), centroid=np.zeros((npn, 3)),
               area=np.zeros(npn), h=np.zeros(npn))
    L.ref_mesh_get(h, *[P(out[k]) for k in ["nodes", "tris", "tags", "V", "normal", "centroid", "area", "h"]])
    return out


MESHES = {
    # name: (kind, args)
    "tube_cfg1": ("tube", (1.0, 4.0, 0.3, 24, 6)),
    "tube_axis03": ("tube", (1.0 / 0.3, 4.0 / 0.3, 1.0, 24, 4)),
    "duct": ("duct", (1.0, 4.0, 24, 24, 6)),
    "sphere_l2": ("sphere", (0.3, 2, 0.0, 0.0, 0.0)),
    "sphere_l3_off": ("sphere", (1.0, 3, 0.7, 0.0, 0.0)),
}


def ref_vectors():
    L = ref_lib()
    arrays = {}
    rng = np.random.default_rng(2024)
    for name, (kind, args) in MESHES.items():
        h = {"tube": L.ref_swimmer_tube_mesh, "duct": L.ref_duct_mesh, "sphere": L.ref_icosphere}[kind](*args)
        m = mesh_arrays(L, h)
        for k, v in m.items():
            arrays[f"{name}/{k}"] = v
        # all-pairs rule levels at the panel centroids (self pairs marked 4)
        x0 = np.ascontiguousarray(m["centroid"])
        npn = len(x0)
        lev = np.zeros((npn, npn), np.int8)
        dist = np.zeros((npn, npn))
        L.ref_levels(h, 0, npn, P(x0), npn, P(lev), P(dist))
        np.fill_diagonal(lev, 4)
        arrays[f"{name}/levels"] = lev
        # panel integrals: up to 8 pairs per rule level, plus 8 self pairs
        pairs = []
        for lv in range(5):
            ii, pp = np.nonzero(lev == lv)
            if len(ii):
                sel = rng.choice(len(ii), size=min(8, len(ii)), replace=False)
                pairs += [(int(ii[s]), int(pp[s])) for s in sel]
        SL = np.zeros((len(pairs), 9))
        DL = np.zeros((len(pairs), 9))
        for q, (i, p) in enumerate(pairs):
            xi = np.ascontiguousarray(x0[i])
            L.ref_panel_integrals(h, p, P(xi), int(i == p), P(SL[q]), P(DL[q]))
        arrays[f"{name}/pairs"] = np.array(pairs, dtype=np.int32)
        arrays[f"{name}/SL"] = SL
        arrays[f"{name}/DL"] = DL
    for lv in range(4):
        Lb = np.zeros((448, 3))
        W = np.zeros(448)
        n = L.ref_rule(lv, P(Lb), P(W))
        arrays[f"rule{lv}/L"] = Lb[:n]
        arrays[f"rule{lv}/W"] = W[:n]
    gx, gw = np.zeros(10), np.zeros(10)
    L.ref_gauss_legendre01(10, P(gx), P(gw))
    arrays["gl10/x"] = gx
    arrays["gl10/w"] = gw
    # field of a random density on the level-2 sphere at a few exterior points
    hs = L.ref_icosphere(1.0, 2, 0.0, 0.0, 0.0)
    npn = L.ref_mesh_np(hs)
    u = rng.standard_normal((npn, 3))
    f = rng.standard_normal((npn, 3))
    pts = np.array([[0.0, 0.0, 1.1], [1.5, 0.2, 0.0], [0.0, -3.0, 0.5], [0.7, 0.7, 0.7]])
    out = np.zeros_like(pts)
    L.ref_velocity_at(hs, P(u), P(f), 1.0, P(pts), len(pts), P(out))
    arrays["field/u"], arrays["field/f"], arrays["field/pts"], arrays["field/out"] = u, f, pts, out
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **arrays)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit(f"reference checkout not found at {REF}")
    if not os.path.exists(REFLIB):
        sys.exit("build the reference build first: make -C oracle")
    copy_csvs()
    copy_inputs()
    ref_vectors()
    print("golden fixtures written to", HERE)
