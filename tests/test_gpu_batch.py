"""Batches of independent scenes in one context (deme_upload_scenes): parity with the oracle, and every scene of a batch
follows bit for bit the trajectory it follows when run alone in the same world."""
import numpy as np
import pytest

from deme_amd import scenes
from deme_amd.abi import DeviceSystem

pytestmark = pytest.mark.gpu


def small_batch(update_freq=1, beta=0.0):
    # uneven scenes (20 to ~950 spheres), per-scene materials and walls, overlapping lattice so contacts exist at once
    return scenes.batched_columns(n_scenes=6, total_scale=0.01, pitch_scale=0.75, update_freq=update_freq, beta=beta)


def test_batch_contacts_match_oracle_and_stay_within_scenes():
    import pyoracle

    sc = small_batch()
    dev = DeviceSystem(sc)
    dev.step(5, 1)
    ora = pyoracle.OracleSystem(sc)
    ora.set_state(dev.get_state())
    dev.cd_run()
    ora.cd_run()
    da, db, dt, _ = dev.contacts()
    oa, ob, ot, _ = ora.contacts()
    assert np.array_equal(da, oa) and np.array_equal(db, ob) and np.array_equal(dt, ot)
    n = sc.n_spheres
    sph = dt == 1
    assert np.all(sc.owner_scene[da[sph]] == sc.owner_scene[db[sph]])
    pl = dt == 3
    assert pl.sum() > 0
    assert np.all(sc.owner_scene[da[pl]] == sc.anal_scene[db[pl]])
    assert sph.sum() > 0 and da.max() < n


@pytest.mark.parametrize("freq", [1, 4])
def test_batch_equals_scenes_run_alone(freq):
    beta = 0.0 if freq == 1 else 2e-4
    sc = small_batch(freq, beta)
    dev = DeviceSystem(sc)
    n_steps = 12
    dev.step(n_steps, freq)
    sb = dev.get_state()
    for k, part in enumerate(scenes.split_batch(sc)):
        d = DeviceSystem(part)
        d.step(n_steps, freq)
        s = d.get_state()
        idx = np.nonzero(sc.owner_scene[:sc.n_spheres] == k)[0]
        for key in ("voxelID", "locX", "locY", "locZ", "oriQw", "vX", "vY", "vZ", "omgX", "omgY", "omgZ"):
            assert np.array_equal(s[key][:-1], sb[key][idx]), f"scene {k}: {key}"
        d.close()


def test_batch_trajectory_matches_oracle():
    import pyoracle
    from deme_amd.abi import decode_positions

    sc = small_batch()
    dev = DeviceSystem(sc)
    ora = pyoracle.OracleSystem(sc)
    dev.step(20, 1)
    ora.step(20, 1)
    sd, so = dev.get_state(), ora.get_state()
    pd = decode_positions(sc.params, sd["voxelID"], sd["locX"], sd["locY"], sd["locZ"]).astype(np.float64)
    po = decode_positions(sc.params, so["voxelID"], so["locX"], so["locY"], so["locZ"]).astype(np.float64)
    assert np.abs(pd - po).max() <= 1e-7
    assert dev.counts().nContacts == ora.counts().nContacts


def test_autotune_picks_a_candidate_and_keeps_it_configured():
    import pyoracle

    sc = small_batch()
    dev = DeviceSystem(sc)
    f, table = dev.autotune(freqs=(1, 3), cell_scales=(1.0, 1.5), trial_steps=6, max_vel=1.0, safety=1.2)
    assert f in (1, 3)
    assert table.shape == (2, 2) and np.all(table > 0)
    # the configured margin is the one of the chosen interval; detection with it still matches the oracle
    f32 = np.float32
    beta = (f32(1.0) * f32(sc.params.h) * f32(f) * f32(1.2)) if f > 1 else f32(0.0)  # float, as the library forms it
    sc.params.beta = float(beta)
    ora = pyoracle.OracleSystem(sc)
    ora.set_state(dev.get_state())
    dev.cd_run()
    ora.cd_run()
    da, db, _, _ = dev.contacts()
    oa, ob, _, _ = ora.contacts()
    assert np.array_equal(da, oa) and np.array_equal(db, ob)
