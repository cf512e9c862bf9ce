"""Parity of the HIP path (through the C ABI) with the oracle on every benchmark configuration at its full size.

Configurations (the workloads the benchmark line quotes):
  1  20,000 spheres r = 5 mm settling in a 0.3 x 0.3 x 0.6 m box, 1,000 steps, detection every step;
  2  1,000,000 three-sphere clumps in a 0.5 x 0.5 x 1.0 m box, detection every step (run A) and every 5 steps with a
     drift margin of 0.2 r_min (run B);
  3  8,000,000 lognormal spheres (4 materials) packed at ~0.6 solid fraction under a prescribed top plate;
  4  256 independent columns (11.6M spheres) batched in one context, on lattices of 0.75 x the largest diameter so
     that they carry contacts;
  5  50,000 spheres dropped onto a plane, 200 steps with the contact list rebuilt every 20: adjoint gradients.

Bars (the same on every configuration):
  - contact lists (idA, idB, type) bit-exact; persistent-contact maps bit-exact when both sides hold the same
    previous list;
  - per contact, within 1e-6 |Fo| + 1e-7 max|F| of the FP32 oracle (the reference's arithmetic, same operations in
    the same order; in practice bit-identical); against the FP64 evaluation, |F - F64| <= 1e-5 |F64| for 99.9% of the
    touching contacts and <= 1e-4 |F64| + 1e-6 max|F64| for all (the FP32 arithmetic itself, which the device
    reproduces bit for bit, strays up to ~2e-5 from FP64 on contacts whose force is a near-cancellation of elastic,
    damping and tangential terms);
  - owner sums bit-exact when the oracle reduces the device's own per-contact records (same order, same arithmetic);
    the per-contact records themselves meet the force bars below;
  - integration from identical states and accelerations: voxel and sub-voxel positions and velocities bit-exact,
    orientations within one float ulp (the rotation increment's cos / sin come from the device's and the host's
    libraries); after five full steps >= 99% of owners at identical sub-voxels, all within 8;
  - configuration 1 after 1,000 steps: positions within 1e-6 m (2e-4 r) of the oracle's trajectory;
  - configuration 5: the float32 forward within 1e-6 m (5e-4 r_min) of the FP64 simulator after 200 steps; velocity
    gradients within 2e-3 of the FP64 adjoint's scale for >= 99.9% of the spheres (the rest sit at kinks of the model,
    where float and double linearize different pieces); material gradients within 2e-3 of scale plus twice the spread
    of the FP64 gradient itself under a 4-ulp change of the initial velocities (its conditioning at float precision).
"""
import numpy as np
import pytest

from deme_amd import scenes
from deme_amd.abi import DeviceSystem, decode_positions, positions_f64

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def same_contacts(dev, ora, maps=False):
    da, db, dt, dm = dev.contacts()
    oa, ob, ot, om = ora.contacts()
    assert len(da) == len(oa)
    assert np.array_equal(da, oa) and np.array_equal(db, ob) and np.array_equal(dt, ot)
    if maps:
        assert np.array_equal(dm, om)
    return len(da)


def check_forces(dev, ora, label):
    """Device forces against the FP64 evaluation and the FP32 oracle, per contact; returns the worst relative errors."""
    ora.set_wildcards(dev.wildcards())
    F64, T64 = ora.forces_fp64()
    ora.force()
    dev.force()
    F, T, cA, cB = dev.forces()
    Fo, To, _, _ = ora.forces()
    mag64 = np.linalg.norm(F64, axis=1)
    fmax = mag64.max()
    err64 = np.linalg.norm(F - F64, axis=1)
    err32 = np.linalg.norm(F - Fo, axis=1)
    touching = mag64 > 0
    rel = err64[touching] / mag64[touching]
    say(f"{label}: {len(F)} contacts, max|F| {fmax:.3e}, per-contact rel err vs FP64: median {np.median(rel):.2e} "
        f"p99.9 {np.quantile(rel, 0.999):.2e} max {rel.max():.2e}; vs FP32 oracle max {err32.max() / fmax:.2e}")
    assert np.all(err32 <= 1e-6 * np.linalg.norm(Fo, axis=1) + 1e-7 * fmax)
    assert np.quantile(rel, 0.999) <= 1e-5
    assert np.all(err64 <= 1e-4 * mag64 + 1e-6 * fmax)
    tmax = max(np.linalg.norm(T64, axis=1).max(), 1e-30)
    assert np.all(np.linalg.norm(T - T64, axis=1) <= 1e-4 * np.linalg.norm(T64, axis=1) + 1e-5 * tmax)
    return rel


def check_owner_sums(dev, ora, n_clumps):
    """Owner sums on identical inputs: the oracle reduces the device's own per-contact records (after check_forces),
    in the same order with the same arithmetic, so every clump's accelerations must agree bit for bit."""
    F, T, cA, cB = dev.forces()
    ora.set_forces(F, T, cA, cB)
    dev.reduce_owner()
    ora.reduce_owner()
    sd, so = dev.get_state(), ora.get_state()
    for k in ("aX", "aY", "aZ", "alphaX", "alphaY", "alphaZ"):
        d, o = sd[k][:n_clumps], so[k][:n_clumps].astype(np.float32)
        bad = np.flatnonzero(d != o)
        assert len(bad) == 0, (k, len(bad), bad[:5], d[bad[:5]], o[bad[:5]])
    say(f"owner sums of {n_clumps} clumps bit-exact on identical per-contact records")


def check_positions(sc, sd, so, frac=0.999, max_subvoxels=2):
    same = ((sd["voxelID"] == so["voxelID"]) & (sd["locX"] == so["locX"]) & (sd["locY"] == so["locY"])
            & (sd["locZ"] == so["locZ"]))
    pd = positions_f64(sc.params, sd["voxelID"], sd["locX"], sd["locY"], sd["locZ"])
    po = positions_f64(sc.params, so["voxelID"], so["locX"], so["locY"], so["locZ"])
    worst = np.abs(pd - po).max() / sc.params.l
    say(f"positions: {same.mean() * 100:.4f}% of owners identical, worst {worst:.1f} sub-voxels")
    assert same.mean() >= frac and worst <= max_subvoxels


def same_integration(sd, so):
    for k in ("voxelID", "locX", "locY", "locZ", "vX", "vY", "vZ", "omgX", "omgY", "omgZ", "family"):
        assert np.array_equal(sd[k], so[k]), k
    for k in ("oriQw", "oriQx", "oriQy", "oriQz"):
        # the rotation increment's cos / sin in double (device and host libraries) may differ in the last bit
        assert np.abs(sd[k].astype(np.float64) - so[k]).max() <= 2.4e-7, k


def test_cfg1_20k_spheres_1000_steps():
    import pyoracle

    sc = scenes.sphere_settling()  # 20,000 spheres, r = 5 mm, E = 1e7, CoR 0.5, mu 0.5, h = 1e-5
    assert sc.n_spheres == 20000
    dev, ora = DeviceSystem(sc), pyoracle.OracleSystem(sc)
    worst = 0.0
    for k in range(10):
        dev.step(100, 1)
        ora.step(100, 1)
        sd, so = dev.get_state(), ora.get_state()
        pd = positions_f64(sc.params, sd["voxelID"], sd["locX"], sd["locY"], sd["locZ"])
        po = positions_f64(sc.params, so["voxelID"], so["locX"], so["locY"], so["locZ"])
        worst = max(worst, np.abs(pd - po).max())
        same = np.mean((sd["voxelID"] == so["voxelID"]) & (sd["locX"] == so["locX"]) & (sd["locY"] == so["locY"])
                       & (sd["locZ"] == so["locZ"]))
        say(f"cfg1 step {100 * (k + 1)}: contacts {dev.counts().nContacts} / {ora.counts().nContacts}, "
            f"max |dX| {np.abs(pd - po).max():.3e} m, identical {same * 100:.3f}%")
    assert worst <= 1e-6
    assert dev.counts().nSteps == 1000
    # after 0.01 s of fall the bottom layer rests on the floor (about 380 contacts); the lists agree at the end
    dev.cd_run()
    ora.set_state(dev.get_state())
    ora.cd_run()
    assert same_contacts(dev, ora) > 300


def clump_systems(update_freq, beta_frac, warm=20):
    import pyoracle

    sc = scenes.clump_settling(update_freq=update_freq, beta_frac=beta_frac)  # 1,000,000 clumps
    assert sc.n_clumps == 1_000_000
    dev = DeviceSystem(sc)
    dev.step(warm, update_freq)  # contacts with tangential history
    ora = pyoracle.OracleSystem(sc)
    ora.set_state(dev.get_state())
    return sc, dev, ora


def test_cfg2_1M_clumps_detection_every_step():
    sc, dev, ora = clump_systems(1, 0.0)
    dev.cd_run()
    ora.cd_run()
    # a second pass on both sides: the persistent maps link the same lists
    dev.cd_run()
    ora.cd_run()
    n = same_contacts(dev, ora, maps=True)
    say(f"cfg2 run A: {n} contacts")
    assert n > 2_000_000
    check_forces(dev, ora, "cfg2 run A")
    check_owner_sums(dev, ora, sc.n_clumps)
    # integration from identical states and accelerations (the device's, given to the oracle): positions, sub-voxels
    # and velocities bit-exact
    ora.set_state(dev.get_state())
    dev.integrate_owners()
    ora.integrate()
    same_integration(dev.get_state(), ora.get_state())
    # five full steps from one state (each starts with a detection pass that carries the history over)
    ora.set_state(dev.get_state())
    ora.set_wildcards(dev.wildcards())
    dev.step(5, 1)
    ora.step(5, 1)
    check_positions(sc, dev.get_state(), ora.get_state(), frac=0.99, max_subvoxels=8)
    assert dev.counts().nContacts == ora.counts().nContacts


def test_cfg2_1M_clumps_drift_margin_every_5_steps():
    sc, dev, ora = clump_systems(5, 0.2)
    assert sc.params.beta > 0
    dev.cd_run()
    ora.cd_run()
    n = same_contacts(dev, ora)
    say(f"cfg2 run B: {n} contacts in the list (drift margin {sc.params.beta:.3e} m)")
    assert n > 5_000_000
    check_forces(dev, ora, "cfg2 run B")


def test_cfg3_8M_polydisperse_bed():
    import pyoracle

    say("cfg3: packing 8M spheres")
    sc = scenes.polydisperse_bed()  # 8,000,000 spheres
    say(f"cfg3 bed: solid fraction {sc.info['solid_fraction']:.3f}, box {sc.info['box']}")
    assert sc.n_spheres == 8_000_000 and 0.55 <= sc.info["solid_fraction"] <= 0.65
    dev = DeviceSystem(sc)
    scenes.settle_bed(dev, sc, steps=200)
    dev.step(5, 1)
    ora = pyoracle.OracleSystem(sc)
    ora.set_state(dev.get_state())
    dev.cd_run()
    ora.cd_run()
    n = same_contacts(dev, ora)
    say(f"cfg3: {n} contacts")
    assert n > 8_000_000
    check_forces(dev, ora, "cfg3")
    check_owner_sums(dev, ora, sc.n_clumps)


def test_cfg4_256_scene_batch():
    import pyoracle

    sc = scenes.batched_columns(pitch_scale=0.75)  # 256 columns, neighbours closer than their radii sum
    assert sc.n_scenes == 256
    dev = DeviceSystem(sc)
    dev.step(10, 1)
    ora = pyoracle.OracleSystem(sc)
    ora.set_state(dev.get_state())
    dev.cd_run()
    ora.cd_run()
    n = same_contacts(dev, ora)
    say(f"cfg4: {sc.n_spheres} spheres, {n} contacts")
    assert n > 1_000_000
    check_forces(dev, ora, "cfg4")
    check_owner_sums(dev, ora, sc.n_clumps)


def test_cfg5_adjoint_50k_spheres_200_steps():
    import pyoracle

    n, steps, window = 50_000, 200, 20
    sc = scenes.adjoint_drop(n=n, steps=steps, window=window)
    dev = DeviceSystem(sc)
    s0 = dev.get_state()
    ref = pyoracle.AdjointFP64(sc, steps, window, state=s0)
    rng = np.random.default_rng(3)
    wX, wV = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)) * 1e-3
    dev.adjoint_forward(steps, window)
    s = dev.get_state()
    gV, gM = dev.adjoint_backward(wX, wV)
    say("cfg5: FP64 adjoint on the host")
    gVh, gMh, Xh, Vh = ref.gradient(wX, wV)
    X = positions_f64(sc.params, s["voxelID"], s["locX"], s["locY"], s["locZ"])[:n]
    say(f"cfg5: forward max |dX| {np.abs(X - Xh).max():.3e} m, contacts {dev.counts().nContacts}")
    # the float32 forward against the same model in double after 200 steps: a few float ulps of a 0.3 m coordinate
    assert np.abs(X - Xh).max() <= 1e-6
    g, h = gV[:n].astype(np.float64), gVh
    # How well-defined the gradient is at float precision: the FP64 adjoint again from initial velocities moved by four
    # float ulps (up, then down), which moves the final positions by ~1e-6 m, as much as float32 rounding moves the
    # device's. Over 200 steps of 50,000 colliding spheres some contacts sit at the model's kinks (overlaps near zero,
    # where the damping term's derivative ~ overlap^-3/4 is unbounded; forces at the Coulomb cap, where sticking and
    # sliding give different derivatives) and the gradients of their spheres change wholesale under such changes.
    spread_v, spread_m = np.zeros(n), np.zeros(4)
    for direction in (np.inf, -np.inf):
        st1 = {k: v.copy() for k, v in s0.items()}
        for k in ("vX", "vY", "vZ"):
            for _ in range(4):
                st1[k][:n] = np.nextafter(st1[k][:n], np.float32(direction))
        gv1, gm1, _, _ = pyoracle.AdjointFP64(sc, steps, window, state=st1).gradient(wX, wV)
        spread_v = np.maximum(spread_v, np.abs(gv1 - h).max(1))
        spread_m = np.maximum(spread_m, np.abs(gm1 - gMh).max(0))
    scale = np.abs(h).max()
    stable = spread_v <= 2e-3 * scale
    err = np.abs(g - h).max(1)
    off = err > 2e-3 * scale
    say(f"cfg5: velocity gradients: {off.sum()} of {n} spheres beyond 2e-3 of scale ({(off & ~stable).sum()} of them "
        f"among the {n - stable.sum()} whose FP64 gradient itself moves that much under the 4-ulp change); median "
        f"error {np.median(err) / scale:.1e} of scale")
    assert off.sum() <= n // 1000
    for k in range(4):
        ms = np.abs(gMh[:, k]).max()
        e = np.abs(gM[:, k] - gMh[:, k]).max()
        say(f"cfg5: material parameter {k}: device error {e / ms:.2e} of scale, FP64 4-ulp spread "
            f"{spread_m[k] / ms:.2e}")
        assert e <= 2e-3 * ms + 2.0 * spread_m[k]
