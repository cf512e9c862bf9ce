"""Mid-simulation changes through the C ABI: appended clumps (UpdateClumps) and clump size changes (ChangeClumpSizes).

Appending keeps the contact list and its histories: spheres appended far from the bed leave every old owner, contact
and history bit-identical to a run without them, and fall exactly as the same spheres do in a system of their own;
a sphere appended onto the bed is detected at the next step while the histories of all old contacts match the run
without it. Resized clumps give the oracle's contact list and inspector values on the resized geometry."""
import dataclasses

import numpy as np
import pytest

import inspect_ref
import pyoracle
from deme_amd import scenes
from deme_amd.abi import (DEME_ERR_INVALID_ARGUMENT, DEME_ERR_NOT_READY, DemeError, DeviceSystem, decode_positions,
                          encode_positions)

pytestmark = pytest.mark.gpu
FIELDS = ("voxelID", "locX", "locY", "locZ", "oriQw", "oriQx", "oriQy", "oriQz", "vX", "vY", "vZ", "omgX", "omgY",
          "omgZ", "family")


def bed():
    return scenes.sphere_settling(n=2000, velocity_sigma=0.05)


def new_spheres(sc, xyz):
    xyz = np.asarray(xyz, np.float32)
    n = len(xyz)
    vox, lx, ly, lz = encode_positions(sc.params, xyz)
    return dict(sph_owner=np.arange(n), sph_radius=np.full(n, sc.sph_radius[0]), sph_relpos=np.zeros((n, 3)),
                sph_mat=np.zeros(n), mass=np.full(n, sc.owner_mass[0]), moi=np.tile(sc.owner_moi[0], (n, 1)),
                state=dict(voxelID=vox, locX=lx, locY=ly, locZ=lz))


def test_far_appended_spheres_leave_the_bed_untouched():
    sc = bed()
    far = [[-0.1 + 0.04 * k, 0.0, 0.1] for k in range(5)]
    print("append: 5 far spheres after 20 steps, then 30 steps", flush=True)
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    a.step(20)
    a.append_clumps(**new_spheres(sc, far))
    a.step(30)
    b.step(50)
    sa, sb = a.get_state(), b.get_state()
    n = sc.n_clumps
    for k in FIELDS:
        assert np.array_equal(sa[k][:n], sb[k][:n]), k
        assert np.array_equal(sa[k][n + 5:], sb[k][n:]), k  # the wall moved up by five owner numbers
    ca, cb = a.contacts(), b.contacts()
    for x, y in zip(ca[:3], cb[:3]):
        assert np.array_equal(x, y)
    assert np.array_equal(a.wildcards(), b.wildcards())
    # the appended spheres fall as the same five spheres do alone in the same world
    alone = dataclasses.replace(sc, n_spheres=0, n_owners=1, n_clumps=0, sph_owner=sc.sph_owner[:0],
                                sph_radius=sc.sph_radius[:0], sph_relpos=sc.sph_relpos[:0], sph_mat=sc.sph_mat[:0],
                                owner_mass=sc.owner_mass[n:], owner_moi=sc.owner_moi[n:],
                                anal=sc.anal.copy(), state={k: v[n:] for k, v in sc.state.items()})
    alone.anal["owner"] = 0
    c = DeviceSystem(alone)
    c.append_clumps(**new_spheres(sc, far))
    c.step(30)
    sc_ = c.get_state()
    for k in FIELDS:
        assert np.array_equal(sa[k][n:n + 5], sc_[k][:5]), k
    z0 = decode_positions(sc.params, *new_spheres(sc, far)["state"].values())[:, 2]
    z1 = decode_positions(sc.params, sa["voxelID"][n:n + 5], sa["locX"][n:n + 5], sa["locY"][n:n + 5],
                          sa["locZ"][n:n + 5])[:, 2]
    t = 30 * float(sc.params.h)
    assert np.allclose(z0 - z1, 0.5 * 9.81 * t * t, rtol=0.05)  # free fall, whatever the step's discretisation


def test_sphere_appended_onto_the_bed_is_detected_and_old_histories_survive():
    sc = bed()
    print("append: one sphere onto the top of the bed after 20 steps, then 1 step", flush=True)
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    a.step(20)
    b.step(20)
    st = a.get_state()
    n = sc.n_clumps
    pos = decode_positions(sc.params, st["voxelID"][:n], st["locX"][:n], st["locY"][:n], st["locZ"][:n])
    top = int(np.argmax(pos[:, 2]))
    r = float(sc.sph_radius[0])
    a.append_clumps(**new_spheres(sc, [pos[top] + np.array([0, 0, 2 * r * 0.995])]))
    with pytest.raises(DemeError) as e:
        a.force()  # the list does not know the new sphere yet
    assert e.value.code == DEME_ERR_NOT_READY
    a.step(1)
    b.step(1)
    ia, ib, ta, _ = a.contacts()
    assert np.any((ia == top) & (ib == n) & (ta == 1))
    old = (ia != n) & ~((ib == n) & (ta == 1))
    jb, kb, tb, _ = b.contacts()
    assert np.array_equal(ia[old], jb) and np.array_equal(ib[old], kb) and np.array_equal(ta[old], tb)
    assert np.array_equal(a.wildcards()[:, old], b.wildcards())  # (4, C): the history components


def test_resized_clumps_match_the_oracle():
    sc = scenes.clump_settling(n_clumps=3000, spacing=(0.0070, 0.0046, 0.0046), velocity_sigma=0.05,
                               omega_sigma=5.0, random_orientation=True)
    print("change_clump_sizes: 60 of 3000 clumps after 10 steps", flush=True)
    dev = DeviceSystem(sc)
    dev.step(10)
    rng = np.random.default_rng(4)
    ids = rng.choice(sc.n_clumps, 60, replace=False)
    f = rng.uniform(0.8, 1.25, 60).astype(np.float32)
    before = dev.inspect("clump_max_z")
    dev.change_clump_sizes(ids, f)
    st = dev.get_state()
    s2 = dev.scene
    want = inspect_ref.inspect("clump_max_z", s2.params, st, s2.sph_owner, s2.sph_radius, s2.sph_relpos)
    got = dev.inspect("clump_max_z")
    assert got.view(np.uint32) == np.float32(want).view(np.uint32) and got != before
    ora = pyoracle.OracleSystem(s2)
    ora.set_state(st)
    dev.cd_run()
    ora.cd_run()
    for x, y in zip(dev.contacts()[:3], ora.contacts()[:3]):
        assert np.array_equal(x, y)
    with pytest.raises(DemeError) as e:
        dev.change_clump_sizes([sc.n_clumps], [1.1])  # the wall is not a clump
    assert e.value.code == DEME_ERR_INVALID_ARGUMENT
    with pytest.raises(DemeError) as e:
        dev.change_clump_sizes([0], [0.0])
    assert e.value.code == DEME_ERR_INVALID_ARGUMENT
