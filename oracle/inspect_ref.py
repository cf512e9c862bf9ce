"""CPU restatement of the reference's inspectors -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of deme_inspect (the C ABI under the facade's DEMInspector). Each
quantity follows the reference's query code in the rounding of its device arithmetic (owner position in double, the
component offset rotated in float, the world corner added in double, then float), vectorised over numpy arrays:

  sphere_world    DEMSphereQueryKernels.cu:14-37     sphere centre X, Y, Z in float
  owner_world     DEMOwnerQueryKernels.cu:63-77      clump centre X, Y, Z in float
  clump_max_z     AuxClasses.cpp:17-19, 62-67       max over spheres of Z + r
  clump_min_z     AuxClasses.cpp:21-23, 69-74       min of Z - r
  clump_max_absv  AuxClasses.cpp:25-49, 94-99       max of |R(q) (omegaBar x (R(q) relPos)) + v|
  clump_volume    AuxClasses.cpp:51-55, 87-93       sum over clumps of the template volume (exact sum, then float)

A region is a Python predicate over the float32 arrays X, Y, Z. Empty selections give the reduction's identity.
Pins (tests/test_inspect_ref.py): hand-built states with closed-form answers -- a sphere's top and bottom, a 3-4-5
speed, a rotating clump whose answer under the reference's frame mixing differs from the physical speed, and a
volume sum inside and outside a region.
"""
from __future__ import annotations

import math

import numpy as np

f32, f64 = np.float32, np.float64
FLT_MAX = float(np.finfo(np.float32).max)


def owner_pos_f64(params, st):
    vox = np.asarray(st["voxelID"], np.uint64)
    nx, ny = int(params.nvXp2), int(params.nvYp2)
    ix = (vox & np.uint64((1 << nx) - 1)).astype(f64)
    iy = ((vox >> np.uint64(nx)) & np.uint64((1 << ny) - 1)).astype(f64)
    iz = (vox >> np.uint64(nx + ny)).astype(f64)
    vs, l = f64(params.voxelSize), f64(params.l)
    return (ix * vs + np.asarray(st["locX"], f64) * l, iy * vs + np.asarray(st["locY"], f64) * l,
            iz * vs + np.asarray(st["locZ"], f64) * l)


def rotate_f32(v, q):
    """applyOriQToVector3<float, float> on arrays: v (3, n) and q (4, n) as float32."""
    x, y, z = (np.asarray(c, f32) for c in v)
    w, a, b, c = (np.asarray(t, f32) for t in q)
    one, two = f32(1), f32(2)
    X = (two * (w * w + a * a) - one) * x + (two * (a * b - w * c)) * y + (two * (a * c + w * b)) * z
    Y = (two * (a * b + w * c)) * x + (two * (w * w + b * b) - one) * y + (two * (b * c - w * a)) * z
    Z = (two * (a * c - w * b)) * x + (two * (b * c + w * a)) * y + (two * (w * w + c * c) - one) * z
    return X, Y, Z


def _quat(st, idx):
    return tuple(np.asarray(st[k], f32)[idx] for k in ("oriQw", "oriQx", "oriQy", "oriQz"))


def sphere_world(params, st, sph_owner, sph_relpos):
    """(X, Y, Z) of every sphere centre (float32) and its rotated offset."""
    o = np.asarray(sph_owner, np.int64)
    ox, oy, oz = (c[o] for c in owner_pos_f64(params, st))
    rel = np.asarray(sph_relpos, f32)
    r = rotate_f32((rel[:, 0], rel[:, 1], rel[:, 2]), _quat(st, o))
    L = [f64(f32(v)) for v in params.LBF]
    X = ((ox + r[0].astype(f64)) + L[0]).astype(f32)
    Y = ((oy + r[1].astype(f64)) + L[1]).astype(f32)
    Z = ((oz + r[2].astype(f64)) + L[2]).astype(f32)
    return (X, Y, Z), r


def owner_world(params, st, n):
    L = [f64(f32(v)) for v in params.LBF]
    p = owner_pos_f64(params, st)
    return tuple((p[k][:n] + L[k]).astype(f32) for k in range(3))


def _reduce(vals, flavor):
    if vals.size == 0:
        return f32(0.0) if flavor == "sum" else f32(-FLT_MAX if flavor == "max" else FLT_MAX)
    if flavor == "max":
        return f32(vals.max())
    if flavor == "min":
        return f32(vals.min())
    return f32(math.fsum(vals.astype(f64).tolist()))


def inspect(quantity, params, st, sph_owner, sph_radius, sph_relpos, n_clumps=None, volume=None, region=None):
    if quantity == "clump_volume":
        X, Y, Z = owner_world(params, st, n_clumps)
        keep = np.ones(n_clumps, bool) if region is None else np.asarray(region(X, Y, Z), bool)
        return _reduce(np.asarray(volume, f32)[keep], "sum")
    (X, Y, Z), r = sphere_world(params, st, sph_owner, sph_relpos)
    keep = np.ones(X.size, bool) if region is None else np.asarray(region(X, Y, Z), bool)
    rad = np.asarray(sph_radius, f32)
    if quantity == "clump_max_z":
        return _reduce((Z + rad)[keep], "max")
    if quantity == "clump_min_z":
        return _reduce((Z - rad)[keep], "min")
    if quantity == "clump_max_absv":
        o = np.asarray(sph_owner, np.int64)
        wx, wy, wz = (np.asarray(st[k], f32)[o] for k in ("omgX", "omgY", "omgZ"))
        rx, ry, rz = r
        p = rotate_f32((wy * rz - wz * ry, wz * rx - wx * rz, wx * ry - wy * rx), _quat(st, o))
        vx, vy, vz = (np.asarray(st[k], f32)[o] for k in ("vX", "vY", "vZ"))
        px, py, pz = p[0] + vx, p[1] + vy, p[2] + vz
        return _reduce(np.sqrt(px * px + py * py + pz * pz)[keep], "max")
    raise ValueError(f"{quantity} is not a known query type.")
