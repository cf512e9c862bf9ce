"""Regression tests of fixed defects, through the C ABI:

- plane-owner sums with the largest number of external-object owners the ABI takes (8, each with its own planes)
  over tens of thousands of contacts, against the oracle's per-owner sums (the partial sums are sized per tile);
- graph replay of steps whose owner update reads the time (an expression prescription in t and a time-dependent
  ChangeFamilyWhen rule): the replayed steps must see the current time, so graphs on equal graphs off bit for bit;
- the device-wide prefix scan (deme_prefix_scan_u32) at sizes that reach each of its code paths (one tile, folded
  half-size tiles, full tiles with a scan of the tile sums, and a second level of tile sums), against numpy.
"""
import ctypes as C

import numpy as np
import pytest

from deme_amd import scenes
from deme_amd.abi import ANAL_DTYPE, DEME_OK, DeviceSystem, decode_positions, load

pytestmark = pytest.mark.gpu


def eight_plane_owners(n=16000):
    # a packed bed whose six box planes, a second floor plane and a tilted plane each belong to their own owner
    sc = scenes.sphere_settling(n=n, spacing=1.98, velocity_sigma=0.05, omega_sigma=3.0, seed=1011)
    nc = sc.n_clumps
    walls = sc.anal.copy()
    extra = np.zeros(2, ANAL_DTYPE)
    for k, (p, nrm) in enumerate((((0.0, 0.0, float(walls["relPos"][0][2]) + 0.001), (0.0, 0.0, 1.0)),
                                  ((0.0, -0.148, -0.25), (0.0, 0.6, 0.8)))):
        extra[k]["type"] = 0
        extra[k]["material"] = 0
        extra[k]["relPos"] = np.array(p, np.float32)
        extra[k]["rot"] = np.array(nrm, np.float32)
        extra[k]["mass"] = np.float32(1e15)
    # the side walls and the top moved in to just inside the outermost spheres, so every wall carries contacts
    X = decode_positions(sc.params, sc.state["voxelID"][:nc], sc.state["locX"][:nc], sc.state["locY"][:nc],
                         sc.state["locZ"][:nc])
    r = float(sc.sph_radius.max())
    lo, hi = X.min(0), X.max(0)
    for k, axis, edge in ((1, 0, lo[0] - 0.8 * r), (2, 0, hi[0] + 0.8 * r), (3, 1, lo[1] - 0.8 * r),
                          (4, 1, hi[1] + 0.8 * r), (5, 2, hi[2] + 0.8 * r)):
        p = walls["relPos"][k].copy()
        p[axis] = np.float32(edge)
        walls["relPos"][k] = p
    anal = np.concatenate([walls, extra])
    anal["owner"] = nc + np.arange(len(anal), dtype=np.uint32)
    n_ext = len(anal)
    assert n_ext == 8
    st = {k: np.concatenate([v[:nc], np.repeat(v[nc:nc + 1], n_ext, axis=0)]) for k, v in sc.state.items()}
    sc.anal = anal
    sc.n_owners = nc + n_ext
    sc.owner_mass = np.concatenate([sc.owner_mass[:nc], np.full(n_ext, 1e15, np.float32)])
    sc.owner_moi = np.ascontiguousarray(np.concatenate([sc.owner_moi[:nc], np.full((n_ext, 3), 1e15, np.float32)]))
    sc.state = st
    return sc


def test_eight_plane_owners_sum_like_the_oracle():
    import pyoracle

    sc = eight_plane_owners()
    dev = DeviceSystem(sc)
    dev.step(30, 1)
    ora = pyoracle.OracleSystem(sc)
    ora.set_state(dev.get_state())
    dev.cd_run()
    ora.cd_run()
    da, db, dt, _ = dev.contacts()
    oa, ob, ot, _ = ora.contacts()
    assert np.array_equal(da, oa) and np.array_equal(db, ob) and np.array_equal(dt, ot)
    assert len(da) > 10000
    plane_of = db[dt == 3]  # DEME_SPHERE_PLANE_CONTACT
    assert len(set(sc.anal["owner"][plane_of].tolist())) >= 6  # most plane owners carry contacts
    ora.set_wildcards(dev.wildcards())
    dev.force()
    ora.force()
    dev.reduce_owner()
    ora.reduce_owner()
    sd, so = dev.get_state(), ora.get_state()
    nc = sc.n_clumps
    for k in ("aX", "aY", "aZ", "alphaX", "alphaY", "alphaZ"):
        d, o = sd[k][nc:].astype(np.float64), so[k][nc:].astype(np.float64)
        for w in range(8):
            # each owner against its own scale (plane-owner accelerations are F / 1e15)
            assert abs(d[w] - o[w]) <= 1e-5 * max(abs(o[w]), 1e-35), (k, w)
        assert np.any(np.abs(o) > 0)


KEYS = ("voxelID", "locX", "locY", "locZ", "oriQw", "oriQx", "oriQy", "oriQz", "vX", "vY", "vZ", "omgX", "omgY", "omgZ",
        "family")


@pytest.mark.parametrize("freq", [3, 7])
def test_graph_replay_sees_the_current_time(freq):
    sc = scenes.sphere_settling(n=3000, spacing=2.02, velocity_sigma=0.2, beta=2e-4, seed=1012)
    sc.state["family"][:100] = 7
    sc.state["family"][100:sc.n_clumps] = 1
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    a.set_step_graphs(True)
    b.set_step_graphs(False)
    for d in (a, b):
        d.set_motion_rules(
            vel_strategy="case 7: { vX = 0.f; vY = 0.f; vZ = -2000.f * t; LinPrescribed = 1; RotPrescribed = 0; "
                         "break; }",
            family_change_rules="if (family_code == 1 && t > 1.05e-4f) newFamily = 2;")
    for chunk in (5, 9, 12):
        a.step(chunk, freq)
        b.step(chunk, freq)
        sa, sb = a.get_state(), b.get_state()
        for k in KEYS:
            assert np.array_equal(sa[k], sb[k]), (chunk, k)
    s = a.get_state()
    h = float(np.float32(sc.params.h))
    t_last = np.float32(sum([h] * 25))  # the time the last step saw
    assert np.all(s["vZ"][:100] == np.float32(-2000.0) * t_last)
    assert np.all(s["family"][100:sc.n_clumps] == 2)


def _scan_case(lib, ctx, n, rng):
    from devmem import DeviceArray

    x = rng.integers(0, 1 << 16, n, dtype=np.uint32) if n else np.zeros(0, np.uint32)
    if n > 3:
        x[rng.integers(0, n, 3)] = 0xFFFFFFFF  # large values: totals beyond 32 bits
    din, dout = DeviceArray.of(x), DeviceArray(4 * (n + 1))
    try:
        total = C.c_uint64(0)
        st = lib.deme_prefix_scan_u32(ctx, din.ptr if n else None, dout.ptr, n, C.byref(total))
        assert st == DEME_OK, lib.deme_last_error(ctx)
        got = dout.get(np.uint32, n + 1)
    finally:
        din.close()
        dout.close()
    ref = np.zeros(n + 1, np.uint64)
    ref[1:] = np.cumsum(x, dtype=np.uint64)
    assert total.value == int(ref[-1])
    assert np.array_equal(got, (ref & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def test_prefix_scan_every_path_against_numpy():
    lib = load()
    h = C.c_void_p()
    assert lib.deme_ctx_create(0, None, C.byref(h)) == DEME_OK
    rng = np.random.default_rng(7)
    try:
        for n in (0, 1, 4095, 4096, 4097, 2048 * 2048, 2048 * 2048 + 1, 4096 * 2048 + 1, 4096 * 4096 + 1):
            print(f"prefix scan n={n}", flush=True)
            _scan_case(lib, h, n, rng)
    finally:
        lib.deme_ctx_destroy(h)
