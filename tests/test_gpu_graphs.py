"""Steps replayed from captured HIP graphs give the same states, bit for bit, as the same steps launched directly:
whole detection intervals (the detection pass and the steps up to the next, one graph launch each) with and without a
drift margin, calls that start and end inside intervals (per-step graphs), a batch of scenes, re-captures after a
change of parameters, and intervals whose lists outgrow the captured capacity (the replay changes nothing and the
interval runs directly)."""
import numpy as np
import pytest

from deme_amd import scenes
from deme_amd.abi import DeviceSystem

pytestmark = pytest.mark.gpu

KEYS = ("voxelID", "locX", "locY", "locZ", "oriQw", "oriQx", "oriQy", "oriQz", "vX", "vY", "vZ", "omgX", "omgY", "omgZ",
        "aX", "aY", "aZ", "alphaX", "alphaY", "alphaZ")


def _pair(sc):
    a, b = DeviceSystem(sc), DeviceSystem(sc)
    a.set_step_graphs(True)
    b.set_step_graphs(False)
    return a, b


def _same(a, b):
    sa, sb = a.get_state(), b.get_state()
    for k in KEYS:
        assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("freq", [2, 5])
def test_graph_replay_equals_direct_launches(freq):
    sc = scenes.sphere_settling(n=4000, spacing=2.02, velocity_sigma=0.2, omega_sigma=3.0, beta=2e-4)
    a, b = _pair(sc)
    for chunk in (7, 1, 13):  # calls that start and end inside detection intervals
        a.step(chunk, freq)
        b.step(chunk, freq)
        _same(a, b)
    assert a.counts().nContacts == b.counts().nContacts > 500


def test_graph_replay_after_parameter_and_state_changes():
    sc = scenes.sphere_settling(n=3000, spacing=2.02, velocity_sigma=0.2, beta=2e-4)
    a, b = _pair(sc)
    a.step(9, 3)
    b.step(9, 3)
    for d in (a, b):
        d.set_beta(3e-4)  # kernel parameters change: the graph is captured again
    a.step(10, 3)
    b.step(10, 3)
    _same(a, b)
    st = a.get_state()
    st["vZ"] = st["vZ"] + np.float32(0.01)
    for d in (a, b):
        d.set_state(st)
    a.step(10, 3)
    b.step(10, 3)
    _same(a, b)


def test_graph_replay_in_a_batch():
    sc = scenes.batched_columns(n_scenes=6, total_scale=0.01, pitch_scale=0.75, update_freq=4, beta=2e-4)
    a, b = _pair(sc)
    a.step(17, 4)
    b.step(17, 4)
    _same(a, b)


@pytest.mark.parametrize("freq,beta", [(1, 0.0), (5, 2e-4)])
def test_whole_interval_graphs_equal_direct_launches(freq, beta):
    sc = scenes.sphere_settling(n=4000, spacing=2.02, velocity_sigma=0.2, omega_sigma=3.0, beta=beta, seed=1013)
    a, b = _pair(sc)
    for chunk in (freq * 6, 3, freq * 4 + 2):
        a.step(chunk, freq)
        b.step(chunk, freq)
        _same(a, b)
        da, db = a.contacts(), b.contacts()
        for x, y in zip(da, db):
            assert np.array_equal(x, y)
    captures, replays, fallbacks = a.graph_stats()
    assert replays >= 6 and captures >= 1
    assert a.counts().nContacts == b.counts().nContacts > 500
    assert a.counts().nSteps == b.counts().nSteps


def test_whole_interval_graphs_fall_back_when_lists_outgrow_them():
    # no headroom: every pass whose list grew outgrows the capture, runs directly and is captured again
    sc = scenes.sphere_settling(n=4000, spacing=2.02, velocity_sigma=0.3, omega_sigma=3.0, beta=2e-4, seed=1014)
    a, b = _pair(sc)
    a.set_graph_headroom(1.0, 0)
    a.step(60, 3)
    b.step(60, 3)
    _same(a, b)
    captures, replays, fallbacks = a.graph_stats()
    assert fallbacks >= 1 and replays >= 1 and captures >= fallbacks
    a.set_graph_headroom(1.25, 1024)
    a.step(30, 3)
    b.step(30, 3)
    _same(a, b)
