"""Parity of the HIP hot path (through the C ABI) with the CPU oracle on identical states.

Bars: integer outputs (bin counts, bin tables, contact lists, history maps, voxel ids) bit-exact; per-contact forces
within 1e-5 of the system force scale (max |F|) of an FP64 evaluation and within 1e-6 of the FP32 oracle; owner
accelerations within 1e-5 relative of the oracle; N-step positions within 1e-7 m.
"""
import numpy as np
import pytest

from deme_amd import scenes
from deme_amd.abi import DEME_ERR_BIN_OVERFLOW, DemeError, DeviceSystem

pytestmark = pytest.mark.gpu


def packed_spheres():
    return scenes.sphere_settling(n=3000, spacing=1.98, velocity_sigma=0.05, omega_sigma=5.0, Crr=0.02)


def packed_clumps():
    return scenes.clump_settling(n_clumps=3000, spacing=(0.0070, 0.0046, 0.0046), velocity_sigma=0.05,
                                 omega_sigma=5.0, random_orientation=True)


def loose_box():
    return scenes.sphere_settling(n=2000, velocity_sigma=0.3)


def expanded_clumps():
    return scenes.clump_settling(n_clumps=2000, spacing=(0.0072, 0.0049, 0.0049), beta_frac=0.2,
                                 velocity_sigma=0.05, omega_sigma=3.0, random_orientation=True)


def crowded_cells():
    # three large spheres widen the detection cells to ~90 small spheres each: blocks whose neighbour segments
    # exceed the staging budget sweep global memory instead of LDS
    return scenes.sphere_settling(n=3000, r=0.002, box=(0.1, 0.1, 0.2), spacing=2.0, velocity_sigma=0.05,
                                  n_big=3, r_big=0.01)


def polydisperse():
    # lognormal radii, four materials, a top plate in a prescribed family moving down
    return scenes.polydisperse_bed(n=3000, velocity_sigma=0.05)


def expanded_polydisperse():
    # the lognormal bed with a drift margin of 0.2 median radii: the sweep's float pre-reject under an expanded radius
    return scenes.polydisperse_bed(n=3000, velocity_sigma=0.05, beta=2e-4)


def masked_families():
    # DisableContactBetweenFamilies(1, 2) on a packed bed of spheres split at random between families 1 and 2
    sc = scenes.sphere_settling(n=3000, spacing=1.98, velocity_sigma=0.05, omega_sigma=5.0, seed=1010)
    rng = np.random.default_rng(5)
    sc.state["family"][:sc.n_clumps] = rng.integers(1, 3, sc.n_clumps)
    sc.mask[(1 + 2) * 2 // 2 + 1] = 1  # locateMaskPair(1, 2)
    return sc


SCENES = [packed_spheres, packed_clumps, loose_box, expanded_clumps, crowded_cells, polydisperse,
          expanded_polydisperse, masked_families]


def systems(make, pre_steps=0):
    import pyoracle

    sc = make()
    dev = DeviceSystem(sc)
    if pre_steps:
        dev.step(pre_steps, 1)
    ora = pyoracle.OracleSystem(sc)
    ora.set_state(dev.get_state())
    return sc, dev, ora


@pytest.mark.parametrize("make", SCENES)
def test_bin_counts_and_table_bit_exact(make):
    sc, dev, ora = systems(make, pre_steps=20)
    dev.cd_bin_count()
    ora.cd_bins()
    nb_d, na_d = dev.bin_counts()
    nb_o, na_o = ora.bin_counts()
    assert np.array_equal(nb_d, nb_o)
    assert np.array_equal(na_d, na_o)
    dev.cd_bin_fill()
    start, sph = dev.bins()
    ob, osph = ora.bin_pairs()
    p = sc.params
    bins = np.repeat(np.arange(p.nbX * p.nbY * p.nbZ, dtype=np.uint32), np.diff(start.astype(np.int64)))
    order = np.lexsort((sph, bins))
    assert np.array_equal(bins[order], ob)
    assert np.array_equal(sph[order], osph)
    cd, co = dev.counts(), ora.counts()
    assert cd.nBinSphereTouches == co.nBinSphereTouches
    assert cd.nActiveBins == co.nActiveBins
    assert cd.maxSpheresInBin == co.maxSpheresInBin


@pytest.mark.parametrize("make", SCENES)
def test_contacts_and_history_map_bit_exact(make):
    sc, dev, ora = systems(make)
    total = 0
    for k in range(6):
        ora.set_state(dev.get_state())
        dev.cd_run()
        ora.cd_run()
        da, db, dt, dm = dev.contacts()
        oa, ob, ot, om = ora.contacts()
        assert np.array_equal(da, oa), f"idA differs at update {k}"
        assert np.array_equal(db, ob), f"idB differs at update {k}"
        assert np.array_equal(dt, ot), f"contact types differ at update {k}"
        assert np.array_equal(dm, om), f"history map differs at update {k}"
        total += len(da)
        for _ in range(3):
            dev.force()
            dev.integrate()
    assert total > 0


@pytest.mark.parametrize("make", SCENES)
def test_staged_bin_path_matches_cell_path_and_oracle(make):
    """The per-bin stages (count, scan+fill, per-bin pairs, sort+remap) and the cell-list detection give the same
    canonical contact list, equal to the oracle's."""
    sc, dev, ora = systems(make, pre_steps=10)
    state = dev.get_state()
    ora.set_state(state)
    ora.cd_run()
    oa, ob, ot, _ = ora.contacts()
    dev.cd_bin_count()
    dev.cd_bin_fill()
    dev.cd_pairs()
    dev.cd_sort_remap()
    sa, sb, st, _ = dev.contacts()
    dev.set_state(state)
    dev.cd_run()
    fa, fb, ft, _ = dev.contacts()
    for x, y in ((sa, oa), (sb, ob), (st, ot), (fa, oa), (fb, ob), (ft, ot)):
        assert np.array_equal(x, y)


def aligned_force_pass(make):
    sc, dev, ora = systems(make)
    ora.step(15, 1)
    dev.set_state(ora.get_state())
    ora.cd_run()
    dev.cd_run()
    da, db, dt, _ = dev.contacts()
    oa, ob, ot, _ = ora.contacts()
    assert np.array_equal(da, oa) and np.array_equal(db, ob) and np.array_equal(dt, ot)
    dev.set_wildcards(ora.wildcards())
    F64, T64 = ora.forces_fp64()
    ora.force()
    dev.force()
    return sc, dev, ora, F64, T64


@pytest.mark.parametrize("make", SCENES)
def test_force_matches_fp64_and_fp32_oracle(make):
    sc, dev, ora, F64, T64 = aligned_force_pass(make)
    F, T, cA, cB = dev.forces()
    Fo, To, cAo, cBo = ora.forces()
    scale = max(np.abs(F64).max(), 1e-30)
    assert np.abs(F - F64).max() <= 1e-5 * scale
    assert np.abs(F - Fo).max() <= 1e-6 * scale
    # rolling resistance: direction of v_rot = rotVelCPB - rotVelCPA, a normalised difference of nearly equal float
    # vectors, so the reference's own FP32 math is only ~1e-5 from FP64 there; held to the FP32 oracle tightly
    tscale = max(np.abs(T64).max(), 1e-30)
    assert np.abs(T - T64).max() <= 1e-4 * tscale + 1e-12
    assert np.abs(T - To).max() <= 1e-6 * tscale + 1e-12
    assert np.abs(cA - cAo).max() <= 1e-6 * max(np.abs(cAo).max(), 1e-30)
    assert np.abs(cB[:, :] - cBo).max() <= 1e-6 * max(np.abs(cBo).max(), 1e-30)
    w, wo = dev.wildcards(), ora.wildcards()
    assert np.abs(w[3] - wo[3]).max() <= 1e-7 * max(np.abs(wo[3]).max(), 1e-30)
    assert np.abs(w[:3] - wo[:3]).max() <= 1e-5 * max(np.abs(wo[:3]).max(), 1e-30)


@pytest.mark.parametrize("make", SCENES)
def test_owner_reduction_matches_oracle(make):
    sc, dev, ora, _, _ = aligned_force_pass(make)
    dev.reduce_owner()
    ora.reduce_owner()
    sd, so = dev.get_state(), ora.get_state()
    for k in ("aX", "aY", "aZ", "alphaX", "alphaY", "alphaZ"):
        d, o = sd[k].astype(np.float64), so[k].astype(np.float64)
        scale = max(np.abs(o).max(), 1e-30)
        assert np.abs(d - o).max() <= 1e-5 * scale, k


@pytest.mark.parametrize("make", SCENES)
def test_integration_matches_oracle(make):
    sc, dev, ora, _, _ = aligned_force_pass(make)
    dev.integrate()
    ora.reduce_owner()
    ora.integrate()
    sd, so = dev.get_state(), ora.get_state()
    assert np.array_equal(sd["family"], so["family"])
    from deme_amd.abi import decode_positions

    pd = decode_positions(sc.params, sd["voxelID"], sd["locX"], sd["locY"], sd["locZ"]).astype(np.float64)
    po = decode_positions(sc.params, so["voxelID"], so["locX"], so["locY"], so["locZ"]).astype(np.float64)
    assert np.abs(pd - po).max() <= 1e-7
    # most owners agree to the last sub-voxel
    same = (sd["voxelID"] == so["voxelID"]) & (sd["locX"] == so["locX"]) & (sd["locY"] == so["locY"]) & (
        sd["locZ"] == so["locZ"])
    assert same.mean() > 0.99
    for k in ("vX", "vY", "vZ", "omgX", "omgY", "omgZ", "oriQw", "oriQx", "oriQy", "oriQz"):
        scale = max(np.abs(so[k]).max(), 1e-30)
        assert np.abs(sd[k].astype(np.float64) - so[k]).max() <= 1e-5 * scale, k


@pytest.mark.parametrize("make,freq", [(packed_spheres, 1), (loose_box, 1), (expanded_clumps, 5), (polydisperse, 1),
                                       (expanded_polydisperse, 5)])
def test_trajectory_parity(make, freq):
    import pyoracle

    sc = make()
    dev = DeviceSystem(sc)
    ora = pyoracle.OracleSystem(sc)
    n = 20
    dev.step(n, freq)
    ora.step(n, freq)
    sd, so = dev.get_state(), ora.get_state()
    from deme_amd.abi import decode_positions

    pd = decode_positions(sc.params, sd["voxelID"], sd["locX"], sd["locY"], sd["locZ"]).astype(np.float64)
    po = decode_positions(sc.params, so["voxelID"], so["locX"], so["locY"], so["locZ"]).astype(np.float64)
    assert np.abs(pd - po).max() <= 1e-7
    vscale = max(np.abs(so["vZ"]).max(), 1e-30)
    assert np.abs(sd["vZ"] - so["vZ"]).max() <= 1e-4 * vscale
    assert dev.counts().nContacts == ora.counts().nContacts
    assert dev.counts().nSteps == n


def test_bin_overflow_is_reported_not_trapped():
    """More than 256 spheres in one bin: the staged per-bin path reports DEME_ERR_BIN_OVERFLOW like the reference
    (DEMContactKernels_Blockwise.cu:42-45, minus the trap); the cell-list path has no per-bin capacity and returns
    the brute-force contact set."""
    import pyoracle
    from deme_amd.abi import build_sim_params

    sc = scenes.sphere_settling(n=3000, spacing=1.98)
    p = sc.params
    sc.params = build_sim_params((0.3, 0.3, 0.6), tuple(p.LBF), 0.005, p.h, bin_size=0.2)
    dev = DeviceSystem(sc)
    dev.cd_bin_count()
    with pytest.raises(DemeError) as e:
        dev.cd_bin_fill()
    assert e.value.code == DEME_ERR_BIN_OVERFLOW
    # the context stays usable after the error
    dev.cd_run()
    a, b, t, _ = dev.contacts()
    ora = pyoracle.OracleSystem(sc)
    ba, bb = ora.brute_force_pairs()
    got = set(zip(a[t == 1].tolist(), b[t == 1].tolist()))
    assert got == set(zip(ba.tolist(), bb.tolist()))
    assert len(got) > 0


def test_prescribed_plate_moves_at_its_velocity():
    from deme_amd.abi import decode_positions

    sc = polydisperse()
    dev = DeviceSystem(sc)
    p0 = decode_positions(sc.params, *(dev.get_state()[k] for k in ("voxelID", "locX", "locY", "locZ")))[-1]
    n = 50
    dev.step(n, 1)
    p1 = decode_positions(sc.params, *(dev.get_state()[k] for k in ("voxelID", "locX", "locY", "locZ")))[-1]
    dz = float(p1[2]) - float(p0[2])
    assert abs(dz - (-0.05 * n * float(np.float32(sc.params.h)))) < 1e-7
    assert abs(float(p1[0]) - float(p0[0])) < 1e-9


def test_masked_family_pairs_never_touch():
    sc, dev, ora = systems(masked_families, pre_steps=5)
    dev.cd_run()
    a, b, t, _ = dev.contacts()
    fam = sc.state["family"]
    sph = t == 1
    fa, fb = fam[a[sph]], fam[b[sph]]
    assert np.all(fa == fb) and sph.sum() > 500  # only same-family sphere pairs remain
    ora.cd_run()
    oa, ob, ot, _ = ora.contacts()
    assert np.array_equal(a, oa) and np.array_equal(b, ob) and np.array_equal(t, ot)


def test_empty_contact_list_steps_and_differentiates():
    import pyoracle

    # 27 spheres far apart, no walls: every step runs with zero contacts
    g = np.arange(3) * 0.05 - 0.05
    xyz = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(2)
    sc = scenes.spheres_scene(xyz, 0.003, (0.4, 0.4, 0.4), planes=False, vel=rng.normal(0, 0.1, (27, 3)), h=1e-5)
    dev = DeviceSystem(sc)
    ora = pyoracle.OracleSystem(sc)
    dev.step(12, 1)
    ora.step(12, 1)
    assert dev.counts().nContacts == 0
    dev.force()
    F, T, _, _ = dev.forces()
    assert F.shape == (0, 3) and T.shape == (0, 3)
    sd, so = dev.get_state(), ora.get_state()
    for k in ("voxelID", "locX", "locY", "locZ", "vX", "vY", "vZ"):
        assert np.array_equal(sd[k], so[k]), k
    # adjoint of free flight: dX_T/dv_0 = T h
    dev.clear_contacts()
    dev.adjoint_forward(10, 5)
    w = rng.normal(size=(27, 3))
    gV, gM = dev.adjoint_backward(w)
    assert np.allclose(gV[:27], w * 10 * np.float32(sc.params.h), rtol=1e-5, atol=0)
    assert np.all(gM == 0.0)


def test_integration_of_uploaded_accelerations_is_bit_exact():
    """integrateOwners on identical states and accelerations (random, uploaded to both sides): voxel ids, sub-voxel
    positions and velocities equal the oracle's bit for bit (DEMIntegrationKernels.cu:46-167: float velocity update,
    double position update and re-encoding); orientations within one float ulp (the rotation increment's cos / sin in
    double come from different libraries)."""
    import pyoracle

    sc = scenes.clump_settling(n_clumps=60000, spacing=(0.0075, 0.005, 0.005), velocity_sigma=0.3, omega_sigma=30.0,
                               random_orientation=True)
    rng = np.random.default_rng(21)
    st = {k: v.copy() for k, v in sc.state.items()}
    n = sc.n_owners
    for k, s in (("aX", 50.0), ("aY", 50.0), ("aZ", 50.0), ("alphaX", 5e3), ("alphaY", 5e3), ("alphaZ", 5e3)):
        st[k] = (rng.normal(size=n) * s).astype(np.float32)
    dev = DeviceSystem(sc)
    dev.set_state(st)
    ora = pyoracle.OracleSystem(sc)
    ora.set_state(st)
    for _ in range(3):
        dev.integrate_owners()
        ora.integrate()
        sd, so = dev.get_state(), ora.get_state()
        for k in ("voxelID", "locX", "locY", "locZ", "vX", "vY", "vZ", "omgX", "omgY", "omgZ"):
            assert np.array_equal(sd[k], so[k]), k
        for k in ("oriQw", "oriQx", "oriQy", "oriQz"):
            assert np.abs(sd[k].astype(np.float64) - so[k]).max() <= 1.2e-7, k
        ora.set_state(sd)  # the next step from the same orientations
    moved = sd["voxelID"] != sc.state["voxelID"]
    assert moved.mean() > 0.01 or np.any(sd["locZ"] != sc.state["locZ"])
