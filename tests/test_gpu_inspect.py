"""Inspectors through the C ABI (deme_inspect) against the restated inspectors (oracle/inspect_ref.py) on the same
device state: maxima and minima bit for bit, sums within one float ulp of the exact sum; regions compiled by hipRTC
select exactly the entities the same predicate selects on the CPU; repeated calls return the same bits; unknown
quantities, malformed regions and regions that fail to compile are reported as errors."""
import numpy as np
import pytest

import inspect_ref
from deme_amd import scenes
from deme_amd.abi import DEME_ERR_INVALID_ARGUMENT, DemeError, DeviceSystem

pytestmark = pytest.mark.gpu
f32 = np.float32

REGIONS = {
    "all": (None, None),
    "bottom": ("return Z < -0.49f;", lambda X, Y, Z: Z < f32(-0.49)),  # the bed spans -0.5 .. -0.482
    "cylinder": ("float r2 = X * X + Y * Y; return r2 < 0.01f && Z > -0.488f;",
                 lambda X, Y, Z: ((X * X + Y * Y) < f32(0.01)) & (Z > f32(-0.488))),
    "nowhere": ("return X > 10.0f && Z < 0.0f;", lambda X, Y, Z: (X > f32(10)) & (Z < f32(0))),
}


@pytest.fixture(scope="module")
def system():
    sc = scenes.clump_settling(n_clumps=20000, spacing=(0.0070, 0.0046, 0.0046), velocity_sigma=0.05,
                               omega_sigma=5.0, random_orientation=True)
    print("inspect scene: clump_settling 20000 clumps, 30 steps", flush=True)
    ds = DeviceSystem(sc)
    ds.step(30)
    vol = np.random.default_rng(2).uniform(1e-9, 5e-8, sc.n_clumps).astype(np.float32)
    ds.set_owner_volumes(vol)
    return sc, ds, vol


@pytest.mark.parametrize("region", list(REGIONS))
@pytest.mark.parametrize("quantity", ["clump_max_z", "clump_min_z", "clump_max_absv", "clump_volume"])
def test_inspector_matches_the_oracle(system, quantity, region):
    sc, ds, vol = system
    code, pred = REGIONS[region]
    print(f"deme_inspect {quantity} region={region}", flush=True)
    got = ds.inspect(quantity, code)
    st = ds.get_state()
    want = inspect_ref.inspect(quantity, sc.params, st, sc.sph_owner, sc.sph_radius, sc.sph_relpos,
                               n_clumps=sc.n_clumps, volume=vol, region=pred)
    if quantity == "clump_volume" and region != "nowhere":
        assert abs(float(got) - float(want)) <= float(np.spacing(want)), (got, want)
    else:
        assert got.view(np.uint32) == f32(want).view(np.uint32), (got, want)
    assert ds.inspect(quantity, code).view(np.uint32) == got.view(np.uint32)  # repeats bit for bit


def test_regions_select_a_proper_subset(system):
    sc, ds, _ = system
    full = ds.inspect("clump_max_z")
    low = ds.inspect("clump_max_z", REGIONS["bottom"][0])
    assert -0.5 < low < full


def test_inspector_errors(system):
    _, ds, _ = system
    with pytest.raises(DemeError) as e:
        ds.inspect("clump_max_x")
    assert e.value.code == DEME_ERR_INVALID_ARGUMENT
    with pytest.raises(DemeError) as e:
        ds.inspect("clump_max_z", "Z < 0.0f;")  # no return
    assert e.value.code == DEME_ERR_INVALID_ARGUMENT
    with pytest.raises(DemeError) as e:
        ds.inspect("clump_max_z", "return Z < undefined_name;")
    assert "undefined_name" in str(e.value)
    assert np.isfinite(ds.inspect("clump_max_z"))  # the context still works
