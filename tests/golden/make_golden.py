"""Writes tests/golden/ref_vectors.npz from the reference's own code.

Runs oracle/_ref/libstt_ref.so (the reference's ProduceRay, CastingRay,
T_VolumeAccess, Matrix4x4, Quaternion and CImg sources compiled in this
container by oracle/Makefile) on small seeded inputs and stores inputs and
outputs.  The fixture lets the oracle and the HIP path be checked against the
reference wherever the reference tree is absent (the GPU box).

    make -C oracle ref && python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _oracle as O  # noqa: E402
from spacetimetomography_amd._lib import Geometry  # noqa: E402

OUT = os.path.join(HERE, "ref_vectors.npz")

# vol, det, spacing, ds, sid, sdd, oxyz, angle
CASES = [
    ((12, 12, 12), (20, 20), (1.0, 1.0, 1.0), 1.0, 40.0, 70.0, (0.0, 0.0, 0.0), 0.0),
    ((14, 10, 16), (24, 18), (1.0, 1.0, 1.0), 0.8, 35.0, 60.0, (0.7, -1.1, 0.4), 33.0),
    ((11, 13, 9), (22, 26), (0.8, 1.25, 1.1), 0.7, 30.0, 55.0, (-0.5, 0.3, 1.2), 151.5),
    ((16, 16, 16), (40, 40), (1.0, 1.0, 1.0), 1.0, 50.0, 80.0, (0.0, 0.0, 0.0), 90.0),     # rays miss at the rim
    ((10, 15, 12), (28, 20), (1.2, 0.9, 1.0), 0.9, 28.0, 52.0, (1.0, 0.0, -0.8), 237.0),
    ((13, 13, 13), (30, 30), (1.0, 1.0, 1.0), 0.6, 22.0, 45.0, (0.0, 2.0, 0.0), 301.25),  # wide cone
    # rays entering below voxel coordinate 0.5 on a negative-direction axis: the
    # block restart of CastingRay adds samples before the entry point
    ((13, 10, 12), (29, 29), (1.0, 1.0, 1.0), 0.55, 16.07, 28.56, (-0.18, 2.67, -3.84), 200.48),
    ((13, 8, 12), (21, 21), (1.0, 1.0, 1.0), 0.55, 16.3, 31.8, (-3.6, 3.8, -2.77), 202.62),
    ((11, 11, 11), (30, 30), (1.0, 1.0, 1.0), 0.634, 22.59, 43.78, (-2.62, -0.68, 1.74), 93.41),
]


def main():
    if O.ref is None:
        raise SystemExit("oracle/_ref/libstt_ref.so not built (make -C oracle ref)")
    rng = np.random.default_rng(20260101)
    data = {}
    f32 = O.f32
    for i, (vol, det, sp, ds, sid, sdd, oxyz, ang) in enumerate(CASES):
        g = Geometry.make(vol, det, sp, ds, sid, sdd, oxyz)
        q = np.zeros(4, np.float32)
        O.ref.r_quat_y(ang, f32(q))
        mats = np.zeros(89, np.float32)
        O.ref.r_view_matrices(C.byref(g), f32(q), f32(mats))
        V = int(np.prod(vol))
        D = det[0] * det[1]
        x = rng.random(V, dtype=np.float32)
        ax = np.zeros(D, np.float32)
        rl = rng.random(D, dtype=np.float32)  # stale lengths: kept on missed rays
        rl_in = rl.copy()
        w = np.zeros(D, np.float32)
        tn = np.zeros(D, np.float32)
        tf = np.zeros(D, np.float32)
        O.ref.r_fp(C.byref(g), f32(q), f32(x), f32(ax), f32(rl), f32(w), f32(tn), f32(tf))
        corr = rng.standard_normal(D).astype(np.float32)
        con = np.zeros(V, np.float32)
        ws = np.zeros(V, np.float32)
        O.ref.r_bp_ray_scatter(C.byref(g), f32(q), f32(corr), f32(con), f32(ws))
        geo = np.array([*vol, *det, ds, sid, sdd, *oxyz, ang], np.float64)
        data[f"c{i}_geo"] = geo
        data[f"c{i}_spacing"] = np.array(sp, np.float32)
        data[f"c{i}_q"] = q
        data[f"c{i}_mats"] = mats
        data[f"c{i}_x"] = x
        data[f"c{i}_raylen_in"] = rl_in
        data[f"c{i}_ax"] = ax
        data[f"c{i}_raylen"] = rl
        data[f"c{i}_weight"] = w
        data[f"c{i}_tnear"] = tn
        data[f"c{i}_tfar"] = tf
        data[f"c{i}_corr"] = corr
        data[f"c{i}_contrib"] = con
        data[f"c{i}_wsum"] = ws
    data["ncases"] = np.array(len(CASES))
    # CImg _cubic_atXYZ at random points, including outside the grid (Neumann clamp)
    img = rng.random(7 * 9 * 5, dtype=np.float32)
    pts = rng.uniform(-1.5, 10.0, (300, 3)).astype(np.float32)
    pts[:20] = np.round(pts[:20])  # grid nodes
    vals = np.array([O.ref.r_cubic_atxyz(f32(img), 7, 9, 5, *map(float, p)) for p in pts], np.float32)
    data["cubic_img"] = img
    data["cubic_pts"] = pts
    data["cubic_vals"] = vals
    # quaternions: one per angle, and the 10-per-round composition of RotateVolume
    angles = np.concatenate([[0.0, 36.0, 90.0, 137.5078, 180.0, 359.5], rng.uniform(0, 360, 10)]).astype(np.float32)
    quats = np.zeros((len(angles), 4), np.float32)
    for k, a in enumerate(angles):
        O.ref.r_quat_y(float(a), f32(quats[k]))
    rounds = np.zeros((10, 4), np.float32)
    O.ref.r_quat_y(17.0, f32(rounds[0]))
    for k in range(1, 10):
        O.ref.r_quat_rotate_y(f32(np.ascontiguousarray(rounds[k - 1])), 36.0, f32(rounds[k]))
    data["quat_angles"] = angles
    data["quats"] = quats
    data["round_start"] = np.array(17.0, np.float32)
    data["round_quats"] = rounds
    # fp32 cofactor inverse of random well-conditioned and near-singular matrices
    mats_in = rng.standard_normal((8, 16)).astype(np.float32)
    mats_in[7] = np.array([1, 2, 3, 4, 2, 4, 6, 8, 0, 1, 0, 1, 3, 1, 2, 0], np.float32)  # singular
    inv = np.zeros_like(mats_in)
    dets = np.zeros(8, np.float32)
    for k in range(8):
        dets[k] = O.ref.r_matrix_inverse(f32(np.ascontiguousarray(mats_in[k])), f32(inv[k]))
    data["inv_in"] = mats_in
    data["inv_out"] = inv
    data["inv_det"] = dets
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
