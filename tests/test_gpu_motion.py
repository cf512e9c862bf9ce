"""Family motion rules through the owner update kernel specialised by hipRTC (deme_jit_build with the reference's
prescription-strategy and family-change placeholders) and ChangeFamily: prescribed velocities and positions given as
expressions in t follow their closed forms step by step, time- and position-dependent family changes happen at the
step their condition first holds, and the rest of the system is untouched."""
import numpy as np
import pytest

from deme_amd import scenes
from deme_amd.abi import DemeError, DeviceSystem, positions_f64

pytestmark = pytest.mark.gpu


def isolated(n=8, fam=None):
    # spheres far apart, no walls: each owner evolves alone
    xyz = np.stack([np.linspace(-0.15, 0.15, n), np.zeros(n), np.zeros(n)], 1)
    sc = scenes.spheres_scene(xyz, 0.003, (0.4, 0.4, 0.4), planes=False, vel=np.full((n, 3), 0.02), h=1e-5)
    if fam is not None:
        sc.state["family"][:n] = fam
    return sc


def test_expression_velocity_follows_its_formula():
    fam = np.zeros(8, np.uint8)
    fam[3] = 7
    sc = isolated(fam=fam)
    dev, ref = DeviceSystem(sc), DeviceSystem(sc)
    dev.set_motion_rules(
        vel_strategy="case 7: { vX = 0.f; vY = 0.f; vZ = -0.05f * t; LinPrescribed = 1; RotPrescribed = 1; break; }",
        pos_strategy="case 7: { LinPrescribed = 0; RotPrescribed = 0; break; }")
    n = 25
    dev.step(n, 1)
    ref.step(n, 1)
    s, r = dev.get_state(), ref.get_state()
    h = float(sc.params.h)
    t_last = np.float32(sum([h] * (n - 1)))  # the time the last step saw (accumulated in double)
    assert s["vZ"][3] == np.float32(-0.05) * t_last and s["vX"][3] == 0.0
    # position: extended Taylor with the velocity dictated moves by h times the velocity entering each step
    z = positions_f64(sc.params, s["voxelID"], s["locX"], s["locY"], s["locZ"])[3, 2]
    z0 = positions_f64(sc.params, sc.state["voxelID"], sc.state["locX"], sc.state["locY"], sc.state["locZ"])[3, 2]
    vz = [np.float32(0.02)] + [np.float32(-0.05) * np.float32(sum([h] * k)) for k in range(n - 1)]
    assert abs(z - (z0 + h * sum(float(v) for v in vz))) < 1e-9
    others = np.arange(8) != 3
    for k in ("voxelID", "locX", "locY", "locZ", "vX", "vY", "vZ"):
        assert np.array_equal(s[k][others], r[k][others]), k


def test_expression_position_is_applied_each_step():
    fam = np.zeros(8, np.uint8)
    fam[5] = 8
    sc = isolated(fam=fam)
    dev = DeviceSystem(sc)
    dev.set_motion_rules(vel_strategy="case 8: { LinPrescribed = 0; RotPrescribed = 0; break; }",
                         pos_strategy="case 8: { X = 0.1 + 10.0 * t; LinPrescribed = 1; RotPrescribed = 0; break; }")
    n = 10
    dev.step(n, 1)
    s = dev.get_state()
    x = positions_f64(sc.params, s["voxelID"], s["locX"], s["locY"], s["locZ"])[5, 0]
    t_last = float(np.float32(sum([float(sc.params.h)] * (n - 1))))
    assert abs(x - (0.1 + 10.0 * t_last)) <= 2 * sc.params.l


def test_time_dependent_family_change_happens_at_its_step():
    sc = isolated(fam=np.ones(8, np.uint8))
    dev = DeviceSystem(sc)
    dev.set_motion_rules(family_change_rules="if (family_code == 1) { if (t > 4.5e-5f) newFamily = 2; }")
    dev.step(5, 1)  # steps see t = 0 .. 4e-5
    assert np.all(dev.get_state()["family"][:8] == 1)
    dev.step(1, 1)  # t = 5e-5
    assert np.all(dev.get_state()["family"][:8] == 2)
    dev.change_family(2, 3)
    assert np.all(dev.get_state()["family"][:8] == 3)
    # an owner whose new family prescribes its velocity moves by that prescription from the step of the change
    with pytest.raises(DemeError):
        dev.adjoint_forward(5, 5)  # no adjoint through motion rules
    dev.set_motion_rules()
    dev.step(1, 1)


def test_position_dependent_family_change_under_gravity():
    sc = isolated(fam=np.ones(8, np.uint8))
    sc.state["vZ"][:8] = np.linspace(-0.5, 0.5, 8).astype(np.float32)
    dev = DeviceSystem(sc)
    X0 = positions_f64(sc.params, sc.state["voxelID"], sc.state["locX"], sc.state["locY"], sc.state["locZ"])
    zc = float(X0[0, 2])
    dev.set_motion_rules(family_change_rules=f"if (family_code == 1 && pos.z < {zc!r} && vel.z < 0.f) newFamily = 4;")
    dev.step(20, 1)
    s = dev.get_state()
    z = positions_f64(sc.params, s["voxelID"], s["locX"], s["locY"], s["locZ"])[:8, 2]
    fam = s["family"][:8]
    # owners moving down crossed below their start height at their first step and changed; the others did not
    assert np.all(fam[sc.state["vZ"][:8] < 0] == 4) and np.all(fam[sc.state["vZ"][:8] > 0] == 1)
    assert np.all(z[fam == 4] < zc)
