"""Device adjoint (deme_adjoint_forward / deme_adjoint_backward through the C ABI) against the FP64 host adjoint on
the same initial state: the recorded forward run, the gradients with respect to initial velocities and materials, the
state left behind, determinism, and the torch.autograd calibration harness."""
import numpy as np
import pytest

import pyoracle
from deme_amd import scenes
from deme_amd.abi import DemeError, DeviceSystem, positions_f64

pytestmark = pytest.mark.gpu

N, STEPS, WINDOW = 400, 60, 20


def _weights(n):
    rng = np.random.default_rng(3)
    return rng.normal(size=(n, 3)), rng.normal(size=(n, 3)) * 1e-3


@pytest.fixture(scope="module")
def case():
    sc = scenes.adjoint_drop(n=N, steps=STEPS, window=WINDOW, seed=11)
    dev = DeviceSystem(sc)
    st0 = dev.get_state()
    ref = pyoracle.AdjointFP64(sc, STEPS, WINDOW, state=st0)
    wX, wV = _weights(N)
    dev.adjoint_forward(STEPS, WINDOW)
    s_fwd = dev.get_state()
    gV, gM = dev.adjoint_backward(wX, wV)
    s_after = dev.get_state()
    gVh, gMh, Xh, Vh = ref.gradient(wX, wV)
    return dict(sc=sc, dev=dev, wX=wX, wV=wV, s_fwd=s_fwd, s_after=s_after, gV=gV, gM=gM, gVh=gVh, gMh=gMh, Xh=Xh,
                Vh=Vh, ref=ref)


def test_recorded_forward_matches_fp64(case):
    sc, s = case["sc"], case["s_fwd"]
    X = positions_f64(sc.params, s["voxelID"], s["locX"], s["locY"], s["locZ"])[:N]
    V = np.stack([s["vX"], s["vY"], s["vZ"]], 1)[:N]
    assert np.abs(X - case["ref"].X0).max() > 1e-4
    assert np.abs(X - case["Xh"]).max() <= 1e-8
    assert np.abs(V - case["Vh"]).max() <= 1e-5
    assert case["dev"].counts().nContacts > N // 4


def test_backward_leaves_the_forward_end_state(case):
    for k in ("voxelID", "locX", "locY", "locZ", "vX", "vY", "vZ", "omgX", "omgY", "omgZ", "oriQw", "aX", "aY", "aZ",
              "alphaX", "alphaY", "alphaZ"):
        assert np.array_equal(case["s_fwd"][k], case["s_after"][k]), k


def test_velocity_gradients_match_fp64_adjoint(case):
    g, h = case["gV"][:N].astype(np.float64), case["gVh"]
    # float32 forward tapes against double: 1e-3 of the gradient's scale, entry by entry
    assert np.abs(g - h).max() <= 1e-3 * np.abs(h).max()
    assert np.linalg.norm(g - h) <= 1e-3 * np.linalg.norm(h)
    assert np.abs(case["gV"][N:]).max() == 0.0  # the wall owner


def test_material_gradients_match_fp64_adjoint(case):
    g, h = case["gM"].astype(np.float64), case["gMh"]
    for k in range(4):
        scale = np.abs(h[:, k]).max()
        assert scale > 0.0
        assert np.abs(g[:, k] - h[:, k]).max() <= 2e-3 * scale, (k, g[:, k], h[:, k])


def test_backward_is_deterministic(case):
    dev = case["dev"]
    dev.set_state(case["sc"].state)  # the backward pass left the system at the forward end
    dev.clear_contacts()
    dev.adjoint_forward(STEPS, WINDOW)
    gV, gM = dev.adjoint_backward(case["wX"], case["wV"])
    assert np.array_equal(gV, case["gV"]) and np.array_equal(gM, case["gM"])


def test_single_window_and_uneven_last_window():
    sc = scenes.adjoint_drop(n=120, steps=25, window=10, seed=5)
    dev = DeviceSystem(sc)
    ref = pyoracle.AdjointFP64(sc, 25, 10, state=dev.get_state())
    wX, wV = _weights(120)
    dev.adjoint_forward(25, 10)
    gV, gM = dev.adjoint_backward(wX, wV)
    gVh, gMh, _, _ = ref.gradient(wX, wV)
    assert np.abs(gV[:120] - gVh).max() <= 1e-3 * np.abs(gVh).max()
    assert np.all(np.abs(gM - gMh) <= 2e-3 * np.abs(gMh).max(axis=0) + 1e-30)
    # the system continues from the forward end (step 25, between detection passes at 20 and 30) exactly as an
    # uninterrupted run with the same schedule
    dev.step(12, 10)
    straight = DeviceSystem(sc)
    straight.step(37, 10)
    a, b = dev.get_state(), straight.get_state()
    for k in ("voxelID", "locX", "locY", "locZ", "vX", "vY", "vZ", "omgX", "omgY", "omgZ"):
        assert np.array_equal(a[k], b[k]), k


def test_backward_without_forward_and_unsupported_systems_fail_loudly():
    sc = scenes.adjoint_drop(n=30, steps=5, window=5, seed=2)
    dev = DeviceSystem(sc)
    with pytest.raises(DemeError):
        dev.adjoint_backward(np.zeros((30, 3)))
    dev.adjoint_forward(5, 5)
    dev.step(1, 1)  # the checkpoints no longer describe the system
    with pytest.raises(DemeError):
        dev.adjoint_backward(np.zeros((30, 3)))
    clumps = scenes.clump_settling(n_clumps=200, steps=5)
    with pytest.raises(DemeError):
        DeviceSystem(clumps).adjoint_forward(5, 5)


def test_torch_harness_gradients_are_the_adjoints():
    import torch
    from deme_amd import calibrate

    sc = scenes.adjoint_drop(n=150, steps=40, window=20, seed=9)
    runner = calibrate.TrajectoryRunner(sc, 40, 20)
    v0, E, nu, CoR, mu = runner.parameters()
    with torch.no_grad():
        target = calibrate.final_positions(runner, v0 * 1.01, E * 1.2, nu, CoR * 0.9, mu).numpy()
    loss = calibrate.mse_to_target(runner, target, v0, E, nu, CoR, mu)
    loss.backward()
    X = calibrate.final_positions(runner, v0.detach(), E.detach(), nu.detach(), CoR.detach(), mu.detach()).numpy()
    gX = 2.0 * (X - target) / 150
    ref = pyoracle.AdjointFP64(sc, 40, 20)
    gVh, gMh, _, _ = ref.gradient(gX, None)
    assert np.abs(v0.grad.numpy() - gVh).max() <= 2e-3 * np.abs(gVh).max()
    for k, t in enumerate((E, nu, CoR, mu)):
        scale = np.abs(gMh[:, k]).max()
        assert np.abs(t.grad.numpy() - gMh[:, k]).max() <= 5e-3 * scale + 1e-30
