"""The DEMSolver facade end to end: a C++ program written against the reference's user API (host/facade_demo.cpp,
built with the library) drops spheres onto the floor of a box; each sphere must come to rest where its force model
balances gravity. The built-in frictional and frictionless Hertz models rest at the Hertz depth (halved plane
overlap), a custom linear model given as a model string (compiled by hipRTC) rests at its spring depth."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dem-engine-sbel_amd", "deme_amd",
                    "facade_demo")


def run(model, n=9, seconds=0.15):
    print(f"facade_demo {model} {n} {seconds}", flush=True)
    out = subprocess.run([DEMO, model, str(n), str(seconds)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("model", ["hertz", "frictionless"])
def test_hertz_models_rest_at_the_hertz_depth(model):
    d = run(model)
    m, r, E, nu = d["mass"], d["r"], d["E"], d["nu"]
    E_eff = 1.0 / (2 * (1 - nu * nu) / E)  # sphere and floor of one material; the floor's radius is infinite
    delta = (3 * m * 9.81 / (4 * E_eff * np.sqrt(r))) ** (2.0 / 3.0)  # F = 4/3 E* sqrt(R) delta^1.5 = m g
    depth = r - d["height"]  # the reference halves the plane overlap: geometric depth = 2 delta
    assert d["vmax"] < 1e-3
    assert abs(depth - 2 * delta) <= 0.02 * 2 * delta, (depth, 2 * delta)
    assert d["contacts"] >= 9


def test_custom_linear_model_rests_at_its_spring_depth():
    d = run("linear")
    depth = d["r"] - d["height"]
    expect = 2 * d["mass"] * 9.81 / d["k"]  # k * (geometric depth / 2) = m g
    assert d["vmax"] < 1e-3
    assert abs(depth - expect) <= 0.05 * expect, (depth, expect)


def test_expression_prescription_and_family_change_through_the_facade():
    d = run("prescribed", n=9, seconds=0.02)
    # sphere 0: vx = 0.1 t of the last step (t as the step saw it, float), vy = vz = 0
    h = 1e-5
    steps = round(d["time"] / h)
    t_last = np.float32(sum([h] * (steps - 1)))
    assert d["v0"][0] == pytest.approx(float(np.float32(0.1) * t_last), rel=1e-6)
    assert d["v0"][1] == 0.0 and d["v0"][2] == 0.0
    # the free spheres fell faster than 0.05 m/s within the first 6 ms and moved to family 6; sphere 0 stayed in 5
    assert d["families"][0] == 5 and all(f == 6 for f in d["families"][1:])
