"""Runtime force-model specialisation through the C ABI (deme_jit_build / deme_set_force_model): the built-in model
with the constants baked in is bit-identical to the shipped kernel, the frictionless model equals the frictional
one at mu = 0, a custom model's forces and wildcards match their definition computed on the host, compile errors
surface, and the built-in kernel comes back."""
import numpy as np
import pytest

from deme_amd import scenes
from deme_amd.abi import DemeError, DeviceSystem, builtin_force_model, positions_f64

pytestmark = pytest.mark.gpu

STATE_KEYS = ("voxelID", "locX", "locY", "locZ", "oriQw", "oriQx", "oriQy", "oriQz", "vX", "vY", "vZ", "omgX", "omgY",
              "omgZ")


def _scene():
    return scenes.sphere_settling(n=3000, spacing=2.02, velocity_sigma=0.3, omega_sigma=5.0, random_orientation=True)


def _run(sc, setup, steps=25):
    dev = DeviceSystem(sc)
    setup(dev)
    dev.step(steps, 1)
    dev.force()
    return dev, dev.get_state(), dev.forces(), dev.wildcards()


def test_baked_builtin_model_is_bit_identical_to_the_shipped_kernel():
    sc = _scene()
    _, sa, fa, wa = _run(sc, lambda d: None)
    dev, sb, fb, wb = _run(sc, lambda d: d.jit_build())
    for k in STATE_KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y)
    assert np.array_equal(wa, wb)
    assert dev.counts().nContacts > 1000
    # and back to the shipped kernel
    dev.set_force_model(None)
    dev.step(1, 1)


def test_frictionless_model_equals_frictional_at_zero_friction():
    sc = _scene()
    sc.materials["mu"] = 0.0
    sc.materials["Crr"] = 0.0
    _, sa, fa, _ = _run(sc, lambda d: None)
    _, sb, fb, wb = _run(sc, lambda d: d.set_force_model(builtin_force_model(1)))
    for k in STATE_KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(fa[0], fb[0])
    assert np.all(wb == 0.0)  # no wildcards: nothing kept per contact


def test_custom_model_forces_and_wildcards_follow_their_definition():
    sc = _scene()
    K = 3.0e4
    model = f"""
    force = (float)({K} * overlapDepth) * B2A;
    contact_age += ts;
    """
    dev = DeviceSystem(sc)
    dev.set_force_model(model, ("contact_age",))
    dev.step(6, 1)
    dev.force()
    a, b, t, _ = dev.contacts()
    F, T, _, _ = dev.forces()
    w = dev.wildcards()
    st = dev.get_state()
    X = positions_f64(sc.params, st["voxelID"], st["locX"], st["locY"], st["locZ"])
    r = sc.sph_radius.astype(np.float64)
    expect = np.zeros((len(a), 3))
    sph = t == 1
    d = X[a[sph]] - X[b[sph]]
    dist = np.linalg.norm(d, axis=1)
    delta = r[a[sph]] + r[b[sph]] - dist
    on = delta >= 0
    n = d / dist[:, None]
    expect[np.nonzero(sph)[0][on]] = (K * delta[on])[:, None] * n[on]
    pl = np.nonzero(t == 3)[0]
    anal = sc.anal
    # plane pose: its owner's position (identity orientation) plus the component offset
    P = X[anal["owner"][b[pl]]] + np.asarray(anal["relPos"][b[pl]], np.float64)
    N = np.asarray(anal["rot"][b[pl]], np.float64)
    distp = np.einsum("ij,ij->i", X[a[pl]] - P, N)
    dp = (r[a[pl]] - distp) / 2.0
    onp = dp >= 0
    expect[pl[onp]] = (K * dp[onp])[:, None] * N[onp]
    scale = np.abs(expect).max()
    assert scale > 0 and onp.any() and on.sum() > 100
    assert np.abs(F - expect).max() <= 2e-5 * scale
    assert np.all(T == 0.0)
    # the age of every active contact counts whole steps of h (carried over detection passes by the persistent map)
    h = np.float32(sc.params.h)
    active = np.linalg.norm(F, axis=1) > 0
    ages = w[0][active] / h
    assert np.all(np.abs(ages - np.round(ages)) < 1e-3) and ages.max() >= 6.5
    assert np.all(w[0][~active] == 0.0)


def test_compile_errors_surface_and_the_adjoint_refuses_custom_models():
    sc = scenes.adjoint_drop(n=50, steps=5, window=5)
    dev = DeviceSystem(sc)
    with pytest.raises(DemeError) as e:
        dev.set_force_model("force = make_float3(overlapDepth, ;")
    assert e.value.code == 6 and "error" in str(e.value)  # DEME_ERR_JIT with the compiler log
    dev.set_force_model("force = (float)(1e4 * overlapDepth) * B2A;")
    with pytest.raises(DemeError):
        dev.adjoint_forward(5, 5)
    dev.set_force_model(None)
    dev.adjoint_forward(5, 5)  # the built-in model differentiates again
