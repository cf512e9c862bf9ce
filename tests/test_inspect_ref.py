"""Pins of the restated inspectors (oracle/inspect_ref.py) on hand-built states with closed-form answers."""
from types import SimpleNamespace

import numpy as np
import pytest

import inspect_ref

f32 = np.float32
FLT_MAX = float(np.finfo(np.float32).max)
PARAMS = SimpleNamespace(nvXp2=21, nvYp2=21, voxelSize=2.0 ** -10, l=2.0 ** -26, LBF=(-0.5, -0.5, -0.5))


def vox(ix, iy, iz):
    return ix + (iy << 21) + (iz << 42)


def state(voxels, quats, vels, omegas):
    n = len(voxels)
    q, v, w = np.asarray(quats, f32), np.asarray(vels, f32), np.asarray(omegas, f32)
    return {"voxelID": np.asarray(voxels, np.uint64), "locX": np.zeros(n, np.uint16), "locY": np.zeros(n, np.uint16),
            "locZ": np.zeros(n, np.uint16), "oriQw": q[:, 0], "oriQx": q[:, 1], "oriQy": q[:, 2], "oriQz": q[:, 3],
            "vX": v[:, 0], "vY": v[:, 1], "vZ": v[:, 2], "omgX": w[:, 0], "omgY": w[:, 1], "omgZ": w[:, 2]}


@pytest.fixture
def two_owners():
    # owner 0: a ball of radius 1/16 at (0, 0, 1/8) moving at (3, 4, 0); owner 1: one sphere of radius 1/32 offset
    # (1/4, 0, 0) from a centre at (0, 0, -1/4), turned a quarter turn about z, moving at (1, 0, 0) and spinning at
    # 2 rad/s about its body z axis
    s = np.float32(np.sqrt(0.5))
    st = state([vox(512, 512, 640), vox(512, 512, 256)], [[1, 0, 0, 0], [s, 0, 0, s]], [[3, 4, 0], [1, 0, 0]],
               [[0, 0, 0], [0, 0, 2]])
    geo = dict(sph_owner=[0, 1], sph_radius=[0.0625, 0.03125], sph_relpos=[[0, 0, 0], [0.25, 0, 0]])
    return st, geo


def test_heights_are_the_sphere_tops_and_bottoms(two_owners):
    st, geo = two_owners
    assert inspect_ref.inspect("clump_max_z", PARAMS, st, **geo) == f32(0.1875)
    assert inspect_ref.inspect("clump_min_z", PARAMS, st, **geo) == f32(-0.28125)
    below = lambda X, Y, Z: Z < 0  # noqa: E731
    assert inspect_ref.inspect("clump_max_z", PARAMS, st, **geo, region=below) == f32(-0.21875)
    nowhere = lambda X, Y, Z: Z > 10  # noqa: E731
    assert inspect_ref.inspect("clump_max_z", PARAMS, st, **geo, region=nowhere) == f32(-FLT_MAX)
    assert inspect_ref.inspect("clump_min_z", PARAMS, st, **geo, region=nowhere) == f32(FLT_MAX)


def test_max_absv_follows_the_reference_estimate(two_owners):
    st, geo = two_owners
    assert inspect_ref.inspect("clump_max_absv", PARAMS, st, **geo, region=lambda X, Y, Z: Z > 0) == f32(5.0)
    # the spinning sphere: the body-frame spin crossed with the world-frame offset (-1/2, 0, 0), turned a quarter turn
    # to (0, -1/2, 0), plus (1, 0, 0): sqrt(5/4). The physical speed (and either frame used consistently) is 1/2.
    v = inspect_ref.inspect("clump_max_absv", PARAMS, st, **geo, region=lambda X, Y, Z: Z < 0)
    assert v == pytest.approx(np.sqrt(1.25), rel=1e-6)


def test_volume_sums_inside_the_region():
    st = state([vox(256, 512, 512), vox(768, 512, 512), vox(896, 512, 512)], [[1, 0, 0, 0]] * 3, [[0, 0, 0]] * 3,
               [[0, 0, 0]] * 3)
    geo = dict(sph_owner=[0, 1, 2], sph_radius=[0.01] * 3, sph_relpos=[[0, 0, 0]] * 3, n_clumps=3,
               volume=[1.0, 2.0, 4.0])
    assert inspect_ref.inspect("clump_volume", PARAMS, st, **geo) == f32(7.0)
    assert inspect_ref.inspect("clump_volume", PARAMS, st, **geo, region=lambda X, Y, Z: X > 0) == f32(6.0)
    assert inspect_ref.inspect("clump_volume", PARAMS, st, **geo, region=lambda X, Y, Z: X > 1) == f32(0.0)


def test_unknown_quantity_is_rejected(two_owners):
    st, geo = two_owners
    with pytest.raises(ValueError):
        inspect_ref.inspect("clump_max_x", PARAMS, st, **geo)
