"""Steps between detection passes replayed from a captured HIP graph give the same states, bit for bit, as the same
steps launched directly: several detection intervals (a capture per pass and replays between them, with the eager
detection work in between), a batch of scenes, a drift margin, and re-captures after a change of parameters."""
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
