"""The pair sweep's launch configurations (deme_set_pair_sweep: exact test, float pre-reject, float sweep with the
exact test in the shared second phase; workgroups of 64, 128 and 256 lanes) give the oracle's contact list and
persistent-contact map bit for bit, on clumps (near-uniform radii, same-owner pairs) and on a polydisperse bed with a
drift margin (many candidates far from touching), as do candidate lists built by the sweep; the launch-configuration
autotuner keeps a configuration that still gives the same lists."""
import numpy as np
import pytest

from deme_amd import scenes
from deme_amd.abi import DeviceSystem

pytestmark = pytest.mark.gpu

CONFIGS = [(m, b) for m in (0, 1, 2) for b in (64, 128, 256)]


def _scene(kind):
    if kind == "clumps":
        return scenes.clump_settling(n_clumps=4000, box=(0.1, 0.1, 0.2), update_freq=1, beta_frac=0.0), 1
    # a drift margin of 0.2 r_min: most candidates do not touch
    return scenes.polydisperse_bed(n=30_000, seed=1031, beta=1e-4), 1


@pytest.mark.parametrize("kind", ["clumps", "bed"])
def test_every_sweep_configuration_gives_the_oracle_list(kind):
    import pyoracle

    sc, f = _scene(kind)
    dev = DeviceSystem(sc)
    if kind == "bed":
        scenes.settle_bed(dev, sc, 50)
    dev.step(10, f)
    ora = pyoracle.OracleSystem(sc)
    ora.set_state(dev.get_state())
    dev.cd_run()
    ora.cd_run()
    ref = ora.contacts()
    assert len(ref[0]) > 1000
    # the cell sweep every pass in each configuration, then candidate lists (automatic and a fixed skin) built by
    # the sweep's build mode at each workgroup size
    runs = [(0.0, mode, block) for mode, block in CONFIGS] + [(skin, 2, block) for skin in (-1.0, 3e-5)
                                                               for block in (64, 128, 256)]
    for skin, mode, block in runs:
        print(f"{kind}: skin {skin}, sweep mode {mode}, block {block}", flush=True)
        dev.set_skin(skin)
        dev.set_pair_sweep(mode, block)
        dev.cd_run()  # the previous list is the same each time, so the maps must agree too
        got = dev.contacts()
        for x, y in zip(got[:3], ref[:3]):
            assert np.array_equal(x, y), (skin, mode, block)
        ora.cd_run()
        assert np.array_equal(got[3], ora.contacts()[3]), (skin, mode, block)


def test_sweep_autotuner_keeps_an_equivalent_configuration():
    sc, f = _scene("clumps")
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    a.set_skin(0.0)  # the cell sweep every pass, where the mode and workgroup size apply
    b.set_skin(0.0)
    a.step(10, f)
    b.step(10, f)
    mode, block, table = a.autotune_sweep(trial_steps=4, update_freq=f)
    assert (mode, block) in CONFIGS and table.shape == (3, 3) and np.all(table > 0)
    b.step(int(3 * 3 * (4 + 1)), f)  # the tuner's trials advanced a: a pass plus 4 steps per configuration
    sa, sb = a.get_state(), b.get_state()
    for k in ("voxelID", "locX", "locY", "locZ", "vX", "vY", "vZ"):
        assert np.array_equal(sa[k], sb[k]), k


def test_detection_autotuner_over_skins_keeps_an_equivalent_configuration():
    sc, f = _scene("clumps")
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    a.step(10, f)
    b.step(10, f)
    skin, mode, block, table = a.autotune_sweep(skins=(-1.0, 0.0), modes=(0, 2), blocks=(128, 256), trial_steps=4,
                                                update_freq=f)
    assert skin in (-1.0, 0.0) and mode in (0, 2) and block in (128, 256)
    assert table.shape == (2, 2, 2) and np.all(table > 0)
    # with candidate lists the mode is not tried (one timing stands for both): 2 trials; the cell sweep: 4
    assert np.array_equal(table[0, 0], table[0, 1])
    b.step(int((2 + 4) * (4 + 1)), f)
    sa, sb = a.get_state(), b.get_state()
    for k in ("voxelID", "locX", "locY", "locZ", "vX", "vY", "vZ"):
        assert np.array_equal(sa[k], sb[k]), k
    assert a.candidate_stats()[0] > 0.0 if skin < 0 else a.candidate_stats()[0] == 0.0


KEYS = ("voxelID", "locX", "locY", "locZ", "oriQw", "oriQx", "oriQy", "oriQz", "vX", "vY", "vZ", "omgX", "omgY", "omgZ")


@pytest.mark.parametrize("skin,freq,graphs", [(None, 1, False), (2e-5, 1, False), (None, 3, True), (2e-5, 2, True)])
def test_candidate_lists_give_the_cell_sweeps_trajectory(skin, freq, graphs):
    """Detection from candidate lists (default skin, and a skin so thin that passes find the lists stale and rebuild
    them) against the full cell sweep every pass: identical contact lists, maps and trajectories, bit for bit, also
    with whole intervals replayed from graphs."""
    sc = scenes.sphere_settling(n=6000, spacing=2.02, velocity_sigma=0.5, omega_sigma=3.0, seed=1032,
                                beta=(2e-4 if freq > 1 else 0.0))
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    if skin is not None:
        a.set_skin(skin)
    b.set_skin(0.0)
    a.set_step_graphs(graphs)
    for chunk in (freq * 10, 7, freq * 15):
        a.step(chunk, freq)
        b.step(chunk, freq)
        sa, sb = a.get_state(), b.get_state()
        for k in KEYS:
            assert np.array_equal(sa[k], sb[k]), (chunk, k)
        for x, y in zip(a.contacts(), b.contacts()):
            assert np.array_equal(x, y)
    sk, n_cand, builds = a.candidate_stats()
    print(f"skin {sk:.3e} m, {n_cand} candidates at the last build, {builds} builds, {a.counts().nContacts} contacts",
          flush=True)
    assert sk > 0 and builds >= 1 and n_cand >= a.counts().nContacts - a.counts().nSphereGeoContacts
    if skin is not None:
        assert builds >= 3  # the thin skin went stale and was rebuilt


@pytest.mark.parametrize("graphs", [False, True])
def test_automatic_skin_gives_way_to_the_cell_sweep_and_back(graphs):
    """A sparse gas of spheres fast enough (tens of m/s, 1e-4 s between passes) that candidate lists die within a
    pass even at the widest automatic skin: the automatic mode then runs the cell sweep for a while and tries the
    lists again. Every switch keeps the contact lists, maps and trajectory of the cell sweep every pass, bit for
    bit."""
    sc = scenes.sphere_settling(n=4000, spacing=4.0, velocity_sigma=20.0, omega_sigma=3.0, seed=1033, beta=2e-4)
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    b.set_skin(0.0)
    a.set_step_graphs(graphs)
    modes = []
    for k in range(180):  # one detection pass each
        a.step(10, 10)
        b.step(10, 10)
        modes.append(a.candidate_stats()[0] > 0)
        if k % 20 == 19:
            sa, sb = a.get_state(), b.get_state()
            for key in KEYS:
                assert np.array_equal(sa[key], sb[key]), (k, key)
            for x, y in zip(a.contacts(), b.contacts()):
                assert np.array_equal(x, y), k
    print("candidate lists in force per pass:", "".join("L" if m else "c" for m in modes), flush=True)
    assert modes[0] and not all(modes)
    first_cell = modes.index(False)
    assert any(modes[first_cell:]), "the lists were not tried again"


def test_skin_below_the_build_position_resolution_runs_the_cell_sweep():
    """Build positions are kept in fixed point (a unit near 2^-31 of twice the world's extent): a skin whose half does
    not exceed 1 nm plus four units could never hold its lists, so the cell sweep runs instead -- no builds, a reported
    skin of 0 -- with the cell sweep's trajectory and contacts, bit for bit."""
    sc = scenes.sphere_settling(n=3000, spacing=2.02, velocity_sigma=0.5, seed=1036)
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    a.set_skin(2e-9)
    b.set_skin(0.0)
    a.step(20, 1)
    b.step(20, 1)
    sa, sb = a.get_state(), b.get_state()
    for k in KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    for x, y in zip(a.contacts(), b.contacts()):
        assert np.array_equal(x, y)
    skin, n_cand, builds = a.candidate_stats()
    assert skin == 0.0 and builds == 0
    # a skin well above the resolution keeps its lists
    a.set_skin(2e-5)
    a.step(5, 1)
    assert a.candidate_stats()[0] == 2e-5 and a.candidate_stats()[2] >= 1


@pytest.mark.parametrize("graphs", [False, True])
def test_lists_survive_a_common_translation(graphs):
    """Every sphere drifting and falling together (0.5 m/s sideways plus gravity, clear of each other and of the
    walls): displacements are measured against a common translation, so the candidate lists built at the start stay
    valid over 30 passes (1.5 mm of travel against a 1 mm skin, which would otherwise be rebuilt every ~7 passes) --
    and the contacts and trajectory stay the cell sweep's, bit for bit."""
    sc = scenes.sphere_settling(n=4000, spacing=2.2, jitter=0.05, seed=1035)
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    st = a.get_state()
    st["vX"] = np.full_like(st["vX"], 0.5)
    for d in (a, b):
        d.set_state(st)
        d.set_step_graphs(graphs)
    b.set_skin(0.0)
    a.step(300, 10)
    b.step(300, 10)
    sa, sb = a.get_state(), b.get_state()
    for k in KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    for x, y in zip(a.contacts(), b.contacts()):
        assert np.array_equal(x, y)
    skin, n_cand, builds = a.candidate_stats()
    print(f"skin {skin:.3e} m, {builds} builds over 30 passes", flush=True)
    assert skin > 0 and builds <= 2
    if graphs:
        captures, replays, fallbacks = a.graph_stats()
        assert replays >= 20 and fallbacks == 0, (captures, replays, fallbacks)
