"""Pins of the CPU oracle against facts the reference's code does not itself state: geometry, closed-form contact
mechanics, conservation laws and exact kinematics. The reference ships no tests or golden vectors for this path, so
these are what make the oracle trustworthy as the parity checker of the HIP path.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a wrong sign or index, a transposed
operand, a missing halving of the plane depth) fails it. CPU only.
"""
import numpy as np
import pytest

from deme_amd import scenes
from deme_amd.abi import NUM_MASK_ENTRIES, RULE_DTYPE


def oracle(sc, threads=2):
    import pyoracle

    return pyoracle.OracleSystem(sc, threads=threads)


def world_positions(sc, state=None):
    """Sphere centres relative to the world's left-bottom-front corner, from the voxel codec's definition
    X = voxel index * voxelSize + sub-voxel index * l per axis, plus the rotated template offset (float64)."""
    p = sc.params
    st = sc.state if state is None else state
    vox = st["voxelID"].astype(np.uint64)
    nx, ny = int(p.nvXp2), int(p.nvYp2)
    ix = (vox & np.uint64((1 << nx) - 1)).astype(np.float64)
    iy = ((vox >> np.uint64(nx)) & np.uint64((1 << ny) - 1)).astype(np.float64)
    iz = (vox >> np.uint64(nx + ny)).astype(np.float64)
    own = np.stack([ix * p.voxelSize + st["locX"].astype(np.float64) * p.l,
                    iy * p.voxelSize + st["locY"].astype(np.float64) * p.l,
                    iz * p.voxelSize + st["locZ"].astype(np.float64) * p.l], 1)
    q = np.stack([st["oriQw"], st["oriQx"], st["oriQy"], st["oriQz"]], 1).astype(np.float64)
    o = sc.sph_owner
    v = sc.sph_relpos.astype(np.float64)
    w, u = q[o, :1], q[o, 1:]
    t = 2.0 * np.cross(u, v)
    return own[o] + v + w * t + np.cross(u, t)


# --------------------------------------------------------------------------------------------------------------
# a3 / a7 / a9 helpers
# --------------------------------------------------------------------------------------------------------------
def test_mask_pair_index_is_a_bijection():
    import pyoracle

    lib = pyoracle.load()
    seen = set()
    for i in range(256):
        for j in range(i, 256):
            k = lib.ora_locate_mask_pair(i, j)
            assert k == lib.ora_locate_mask_pair(j, i)
            seen.add(k)
    assert seen == set(range(NUM_MASK_ENTRIES))


def test_recover_cnt_pair_enumerates_unordered_pairs():
    import ctypes as C

    import pyoracle

    lib = pyoracle.load()
    for n in (2, 3, 17, 64, 256):
        got = set()
        for ind in range(n * (n - 1) // 2):
            a, b = C.c_uint32(), C.c_uint32()
            lib.ora_recover_cnt_pair(ind, n, C.byref(a), C.byref(b))
            assert a.value < b.value < n
            got.add((a.value, b.value))
        assert len(got) == n * (n - 1) // 2


# --------------------------------------------------------------------------------------------------------------
# a6 / a8: bins touched by each sphere = bins its AABB intersects
# --------------------------------------------------------------------------------------------------------------
BIN_SCENES = [
    lambda: scenes.sphere_settling(n=600, spacing=2.05, velocity_sigma=0.0),
    lambda: scenes.sphere_settling(n=400, beta=0.001),
    lambda: scenes.clump_settling(n_clumps=300, spacing=(0.0070, 0.0046, 0.0046), random_orientation=True),
]


@pytest.mark.parametrize("make", BIN_SCENES)
def test_bin_table_is_the_aabb_bin_enumeration(make):
    sc = make()
    p = sc.params
    X = world_positions(sc)
    rcd = sc.sph_radius.astype(np.float64) + np.float64(np.float32(p.beta))
    ora = oracle(sc)
    ora.cd_bins()
    nb, _ = ora.bin_counts()
    bins, sph = ora.bin_pairs()
    b = p.binSize
    expected = []
    ambiguous = set()
    for s in range(sc.n_spheres):
        lo = np.floor((X[s] - rcd[s]) / b).astype(np.int64)
        hi = np.floor((X[s] + rcd[s]) / b).astype(np.int64)
        # an AABB face within 1e-9 of a bin face may land on either side in the reference's arithmetic
        if np.any(np.abs((X[s] - rcd[s]) / b - np.round((X[s] - rcd[s]) / b)) < 1e-9) or \
                np.any(np.abs((X[s] + rcd[s]) / b - np.round((X[s] + rcd[s]) / b)) < 1e-9):
            ambiguous.add(s)
            continue
        lo = np.maximum(lo, 0)
        cnt = np.prod(hi - lo + 1)
        assert nb[s] == cnt, f"sphere {s}"
        for k in range(lo[2], hi[2] + 1):
            for j in range(lo[1], hi[1] + 1):
                for i in range(lo[0], hi[0] + 1):
                    expected.append((i + j * p.nbX + k * p.nbX * p.nbY, s))
    got = [(int(x), int(y)) for x, y in zip(bins, sph) if int(y) not in ambiguous]
    assert sorted(expected) == got  # the oracle's table is sorted by (bin, sphere)
    assert len(ambiguous) <= sc.n_spheres // 100


# --------------------------------------------------------------------------------------------------------------
# a9 / a10: contact set = physical overlap of expanded spheres / plane distance within the envelope
# --------------------------------------------------------------------------------------------------------------
PAIR_SCENES = [
    lambda: scenes.sphere_settling(n=800, spacing=1.99, velocity_sigma=0.0),
    lambda: scenes.sphere_settling(n=500, spacing=2.02, beta=0.0006),
    lambda: scenes.clump_settling(n_clumps=400, spacing=(0.0069, 0.0045, 0.0045), random_orientation=True),
    lambda: scenes.clump_settling(n_clumps=300, spacing=(0.0072, 0.0049, 0.0049), beta_frac=0.2),
]


@pytest.mark.parametrize("make", PAIR_SCENES)
def test_pair_set_is_the_overlap_predicate(make):
    sc = make()
    p = sc.params
    X = world_positions(sc)
    rexp = (sc.sph_radius + np.float32(p.beta)).astype(np.float64)  # float(r + beta), as detection forms it
    ora = oracle(sc)
    ora.cd_run()
    a, b, t, _ = ora.contacts()
    got = set(zip(a[t == 1].tolist(), b[t == 1].tolist()))
    d = np.sqrt(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1))
    cut = rexp[:, None] + rexp[None, :]
    iu = np.triu_indices(sc.n_spheres, 1)
    same_owner = sc.sph_owner[iu[0]] == sc.sph_owner[iu[1]]
    rel = (d[iu] - cut[iu]) / cut[iu]
    inside = (rel <= 0) & ~same_owner
    clear = np.abs(rel) > 1e-6
    want = set(zip(iu[0][inside & clear].tolist(), iu[1][inside & clear].tolist()))
    band = set(zip(iu[0][~clear].tolist(), iu[1][~clear].tolist()))
    assert want <= got
    assert got - want <= band
    assert len(want) > 0
    # each pair appears once, with the lower id as A
    assert len(got) == int((t == 1).sum())
    assert np.all(a[t == 1] < b[t == 1])


def near_walls(beta):
    rng = np.random.default_rng(7)
    n, r = 400, 0.004
    xyz = rng.uniform(-0.05 + 0.5 * r, 0.05 - 0.5 * r, size=(n, 3))
    return scenes.spheres_scene(xyz, r, (0.1, 0.1, 0.1), beta=beta)


@pytest.mark.parametrize("make", [lambda: near_walls(0.0), lambda: near_walls(0.0005), PAIR_SCENES[1]])
def test_plane_contacts_are_within_the_envelope(make):
    sc = make()
    p = sc.params
    X = world_positions(sc)
    lbf = np.array(p.LBF, np.float64)
    rcd = sc.sph_radius.astype(np.float64) + np.float64(np.float32(p.beta))
    ora = oracle(sc)
    ora.cd_run()
    a, b, t, _ = ora.contacts()
    got = set(zip(a[t == 3].tolist(), b[t == 3].tolist()))
    want, band = set(), set()
    for k, pl in enumerate(sc.anal):
        P = pl["relPos"].astype(np.float64) - lbf
        n = pl["rot"].astype(np.float64)
        dist = (X - P) @ n
        env = rcd + np.float64(np.float32(p.beta))
        for s in np.nonzero(dist <= env + 1e-9)[0]:
            (want if dist[s] <= env[s] - 1e-9 else band).add((int(s), k))
    assert want <= got and got - want <= band
    assert len(want) > 0


# --------------------------------------------------------------------------------------------------------------
# a11 / a12: canonical order and the persistent-contact map
# --------------------------------------------------------------------------------------------------------------
def test_canonical_order_and_history_map_join():
    sc = scenes.sphere_settling(n=1500, spacing=1.99, velocity_sigma=0.3)
    ora = oracle(sc)
    ora.cd_run()
    prev = ora.contacts()
    for _ in range(6):
        ora.force()
        ora.reduce_owner()
        ora.integrate()
    ora.cd_run()
    a, b, t, m = ora.contacts()
    # by idA; planes (type 3) before spheres (type 1) within a row; partners ascending
    key = np.lexsort((b, (t == 1).astype(np.int64), a))
    assert np.array_equal(key, np.arange(len(a)))
    old = {(int(x), int(y), int(z)): i for i, (x, y, z) in enumerate(zip(*prev[:3]))}
    expect = np.array([old.get((int(x), int(y), int(z)), 0xFFFFFFFF) for x, y, z in zip(a, b, t)], np.uint32)
    assert np.array_equal(m, expect)
    assert (m != 0xFFFFFFFF).sum() > 0 and (m == 0xFFFFFFFF).sum() > 0


# --------------------------------------------------------------------------------------------------------------
# a15: contact mechanics
# --------------------------------------------------------------------------------------------------------------
E, NU = 1e7, 0.3
E_STAR = E / (2.0 * (1.0 - NU ** 2))  # two bodies of one material


@pytest.mark.parametrize("delta", [2e-6, 3e-5, 1e-4])
def test_head_on_hertz_force(delta):
    """F = 4/3 E* sqrt(R*) delta^1.5 along the line of centres, pointing from B to A; no damping at rest."""
    r = 0.005
    sc = scenes.spheres_scene([[0.001, 0.002, 0.003], [0.001 + 2 * r - delta, 0.002, 0.003]], r, (0.1, 0.1, 0.1),
                              planes=False, g=(0, 0, 0), E=E, nu=NU)
    ora = oracle(sc)
    ora.cd_run()
    ora.force()
    F, T, cA, cB = ora.forces()
    X = world_positions(sc)
    d = np.linalg.norm(X[1] - X[0])
    r32 = float(np.float32(r))  # radii are stored in float
    dq = 2 * r32 - d  # the encoded (quantised) overlap
    expect = 4.0 / 3.0 * E_STAR * np.sqrt(r32 / 2) * dq ** 1.5
    assert F.shape == (1, 3)
    assert F[0, 0] < 0 and abs(F[0, 1]) < 1e-6 * expect and abs(F[0, 2]) < 1e-6 * expect
    assert abs(np.linalg.norm(F[0]) - expect) <= 2e-5 * expect
    assert np.all(T == 0)
    # contact point halfway into the overlap, in each body's frame
    assert abs(cA[0, 0] - (r - dq / 2)) < 1e-7 and abs(cB[0, 0] + (r - dq / 2)) < 1e-7


def test_resting_sphere_equilibrium_depth_on_a_plane():
    """The reference halves the sphere-plane penetration: at rest, F_n = m g when the geometric penetration is
    2 (3 m g / (4 E* sqrt(r)))^(2/3). A slip dropping the halving is off by 2^1.5."""
    r, g = 0.005, 9.81
    box = (0.1, 0.1, 0.1)
    m = 2600.0 * 4.0 / 3.0 * np.pi * r ** 3
    pen = 2.0 * (3.0 * m * g / (4.0 * E_STAR * np.sqrt(r))) ** (2.0 / 3.0)
    zc = -0.05 + r - pen
    sc = scenes.spheres_scene([[0.0, 0.0, zc]], r, box, E=E, nu=NU, g=(0, 0, -g))
    ora = oracle(sc)
    ora.cd_run()
    ora.force()
    F, _, _, _ = ora.forces()
    a, b, t, _ = ora.contacts()
    assert list(t) == [3] and b[0] == 0  # floor plane only
    X = world_positions(sc)
    pen_q = r - (X[0, 2] - 0.0)  # floor at world z = 0 relative to the corner
    expect = 4.0 / 3.0 * E_STAR * np.sqrt(r) * (pen_q / 2.0) ** 1.5
    assert abs(F[0, 2] - expect) <= 1e-4 * expect
    assert abs(F[0, 2] - m * g) <= 1e-3 * m * g
    # and it stays put: 200 steps of the full schedule move it by far less than the penetration
    ora.step(200, 1)
    X1 = world_positions(sc, ora.get_state())
    assert abs(X1[0, 2] - X[0, 2]) < 1e-3 * pen


def test_coulomb_cap_and_newtons_third_law():
    """Sliding contact: |F_t| <= mu |F_n| once the spring saturates; the reduction's net force is zero."""
    r, mu = 0.005, 0.3
    sc = scenes.spheres_scene([[0.0, 0.0, 0.0], [2 * r - 5e-5, 0.0, 0.0]], r, (0.1, 0.1, 0.1), planes=False,
                              g=(0, 0, 0), E=E, nu=NU, mu=mu, vel=[[0.0, 0.5, 0.0], [0.0, -0.5, 0.0]], h=1e-5)
    ora = oracle(sc)
    ora.cd_run()
    saturated = False
    for _ in range(40):
        ora.force()
        F, _, _, _ = ora.forces()
        n = np.array([-1.0, 0.0, 0.0])
        fn = abs(F[0] @ n)
        ft = np.linalg.norm(F[0] - (F[0] @ n) * n)
        assert ft <= mu * fn * (1 + 1e-5) + 1e-12
        saturated |= ft >= mu * fn * (1 - 1e-4)
    assert saturated
    ora.reduce_owner()
    st = ora.get_state()
    mass = sc.owner_mass[:2].astype(np.float64)
    net = sum(mass[i] * np.array([st["aX"][i], st["aY"][i], st["aZ"][i]], np.float64) for i in range(2))
    scale = mass[0] * np.linalg.norm([st["aX"][0], st["aY"][0], st["aZ"][0]])
    assert np.linalg.norm(net) <= 1e-6 * scale


def test_zero_net_force_in_a_packed_bed_without_planes():
    sc = scenes.clump_settling(n_clumps=600, spacing=(0.0069, 0.0045, 0.0045), random_orientation=True,
                               velocity_sigma=0.05)
    sc.anal = sc.anal[:0]
    ora = oracle(sc)
    ora.cd_run()
    ora.force()
    ora.reduce_owner()
    st = ora.get_state()
    m = sc.owner_mass[:sc.n_clumps].astype(np.float64)
    f = np.stack([st[k][:sc.n_clumps] for k in ("aX", "aY", "aZ")], 1).astype(np.float64) * m[:, None]
    assert np.abs(f).sum() > 0
    assert np.linalg.norm(f.sum(0)) <= 1e-5 * np.abs(f).sum()


# --------------------------------------------------------------------------------------------------------------
# a17: kinematics
# --------------------------------------------------------------------------------------------------------------
def test_free_fall_is_exact_for_constant_gravity():
    g, h, n = 9.81, 1e-5, 300
    sc = scenes.spheres_scene([[0.0, 0.0, 0.02]], 0.005, (0.1, 0.1, 0.1), planes=False, g=(0, 0, -g), h=h)
    X0 = world_positions(sc)[0]
    ora = oracle(sc)
    ora.step(n, 1)
    st = ora.get_state()
    X1 = world_positions(sc, st)[0]
    t = n * np.float32(h)
    assert abs((X1[2] - X0[2]) - (-0.5 * g * t * t)) < 1e-8
    assert abs(X1[0] - X0[0]) < 1e-12 and abs(X1[1] - X0[1]) < 1e-12
    assert abs(st["vZ"][0] - (-g * t)) < 1e-5 * g * t


def test_constant_spin_composes_to_the_total_angle():
    w, h, n = 40.0, 1e-5, 500
    axis = np.array([1.0, 2.0, 2.0]) / 3.0
    sc = scenes.spheres_scene([[0.0, 0.0, 0.0]], 0.005, (0.1, 0.1, 0.1), planes=False, g=(0, 0, 0), h=h,
                              omg=[w * axis])
    ora = oracle(sc)
    ora.step(n, 1)
    st = ora.get_state()
    q = np.array([st["oriQw"][0], st["oriQx"][0], st["oriQy"][0], st["oriQz"][0]], np.float64)
    theta = w * n * h
    expect = np.concatenate([[np.cos(theta / 2)], np.sin(theta / 2) * axis])
    assert np.abs(q - expect).max() < 2e-5


def test_fixed_and_prescribed_families():
    rules = np.zeros(256, RULE_DTYPE)
    rules[255]["used"] = 1
    rules[255]["linVelPrescribed"] = rules[255]["rotVelPrescribed"] = 1
    rules[255]["linPosPrescribed"] = rules[255]["rotPosPrescribed"] = 1
    rules[255]["velSet"] = 0x3F
    rules[7]["used"] = 1
    rules[7]["linVelPrescribed"] = 1
    rules[7]["velSet"] = 0x07
    rules[7]["vel"][:3] = (0.1, -0.05, 0.0)
    mask = np.zeros(NUM_MASK_ENTRIES, np.uint8)
    h, n = 1e-5, 200
    # The extended-Taylor pass-on moves a body with the mean of its old and new velocity, and a prescription only
    # replaces the new one (DEMIntegrationKernels.cu:57-89, IntegrationVelPassOnExtendedTaylor.cu): start the
    # prescribed body at its prescribed velocity so that every step moves it by v h.
    sc = scenes.spheres_scene([[0.0, 0.0, 0.0], [0.02, 0.0, 0.0]], 0.005, (0.1, 0.1, 0.1), planes=False, h=h,
                              families=np.array([255, 7], np.uint8), rules=rules, mask=mask,
                              vel=[[0.0, 0.0, 0.0], [0.1, -0.05, 0.0]])
    X0 = world_positions(sc)
    ora = oracle(sc)
    ora.step(n, 1)
    st = ora.get_state()
    X1 = world_positions(sc, st)
    assert np.array_equal(X1[0], X0[0])  # fixed: not even gravity moves it
    t = n * np.float32(h)
    assert np.abs((X1[1] - X0[1]) - np.array([0.1, -0.05, 0.0]) * t).max() < 1e-9


# --------------------------------------------------------------------------------------------------------------
# a15: the rest of the Hertz-Mindlin model against its closed forms (FullHertzianForceModel.cu:37-120,
# matProxy2ContactParam at DEMHelperKernels.cu:316-339). Two spheres on the x axis (A at the origin, B at +x, so the
# unit normal B2A is -x and the force returned is the force on A); expected values are formed in double from the
# model's definitions. Both the FP32 oracle (1e-4 relative: float arithmetic) and the FP64 restatement the GPU parity
# tests use as their reference (1e-9) are held to them.
# --------------------------------------------------------------------------------------------------------------
def mixed(m1, m2):
    """E*, G*, CoR, mu, Crr of a pair of materials (E, nu, CoR, mu, Crr) by the model's mixing rule."""
    E1, n1, e1, u1, c1 = m1
    E2, n2, e2, u2, c2 = m2
    E_eff = 1.0 / ((1 - n1 * n1) / E1 + (1 - n2 * n2) / E2)
    G_eff = 1.0 / (2 * (2 - n1) * (1 + n1) / E1 + 2 * (2 - n2) * (1 + n2) / E2)
    return E_eff, G_eff, min(e1, e2), max(u1, u2), max(c1, c2)


def f32(x):
    return float(np.float32(x))


def pair(delta, rA=0.005, rB=0.005, vA=(0, 0, 0), vB=(0, 0, 0), wA=(0, 0, 0), wB=(0, 0, 0),
         matA=(E, NU, 0.5, 0.5, 0.0), matB=None, hist=None, h=1e-5):
    """One force pass of the oracle on a two-sphere contact of overlap ~delta, with preset contact history
    hist = (delta_tan_x, delta_tan_y, delta_tan_z, delta_time); returns the closed-form ingredients with the FP32
    and FP64 results."""
    from deme_amd.abi import MATERIAL_DTYPE

    matB = matA if matB is None else matB
    sc = scenes.spheres_scene([[0.0, 0.0, 0.0], [rA + rB - delta, 0.0, 0.0]], np.array([rA, rB]), (0.1, 0.1, 0.1),
                              planes=False, g=(0, 0, 0), vel=[vA, vB], omg=[wA, wB], h=h)
    sc.materials = np.zeros(2, MATERIAL_DTYPE)
    sc.materials[0] = matA
    sc.materials[1] = matB
    sc.sph_mat = np.array([0, 1], np.uint8)
    ora = oracle(sc)
    ora.cd_run()
    a, b, t, _ = ora.contacts()
    assert list(t) == [1] and a[0] == 0 and b[0] == 1
    if hist is not None:
        ora.set_wildcards(np.asarray(hist, np.float32).reshape(4, 1))
    F64, T64 = ora.forces_fp64()
    ora.force()
    F, T, cA, cB = ora.forces()
    X = world_positions(sc)
    assert X[0, 1] == X[1, 1] and X[0, 2] == X[1, 2]  # the normal is exactly -x
    ra, rb = f32(rA), f32(rB)
    d = ra + rb - (X[1, 0] - X[0, 0])  # the encoded overlap
    mA, mB = (float(m) for m in sc.owner_mass[:2])
    mats = [tuple(float(sc.materials[k][f]) for f in ("E", "nu", "CoR", "mu", "Crr")) for k in (0, 1)]
    E_eff, G_eff, e, mu, crr = mixed(*mats)
    R = ra * rb / (ra + rb)
    m = mA * mB / (mA + mB)
    Sn = 2.0 * E_eff * np.sqrt(R * d)
    loge = np.log(e)
    beta = loge / np.sqrt(loge * loge + np.pi ** 2)
    cf = dict(d=d, R=R, m=m, E=E_eff, G=G_eff, e=e, mu=mu, crr=crr, beta=beta, kn=2.0 / 3.0 * Sn,
              gn=2.0 * np.sqrt(5.0 / 6.0) * beta * np.sqrt(Sn * m), kt=8.0 * G_eff * np.sqrt(R * d), h=f32(h))
    cf["gt"] = -2.0 * np.sqrt(5.0 / 6.0) * beta * np.sqrt(m * cf["kt"])
    return cf, F[0].astype(np.float64), T[0].astype(np.float64), F64[0], T64[0], ora.wildcards()[:, 0]


def close(got, want, rel):
    return abs(got - want) <= rel * max(abs(want), 1e-30)


@pytest.mark.parametrize("speed,cor", [(0.3, 0.5), (-0.2, 0.5), (0.3, 0.2), (0.3, 0.9)])
def test_normal_damping(speed, cor):
    """F_n = k_n delta + gamma_n v_n with k_n = 2/3 S_n, S_n = 2 E* sqrt(R* delta), gamma_n = 2 sqrt(5/6) beta
    sqrt(S_n m*), beta = ln e / sqrt(ln^2 e + pi^2) and v_n = (v_A - v_B) . B2A (speed > 0: approaching)."""
    cf, F, T, F64, _, w = pair(3e-5, vA=(speed / 2, 0, 0), vB=(-speed / 2, 0, 0), matA=(E, NU, cor, 0.5, 0.0))
    vn = -(f32(speed / 2) + f32(speed / 2))  # velocities are stored in float
    fn = cf["kn"] * cf["d"] + cf["gn"] * vn
    assert abs(cf["gn"] * vn) > 0.25 * cf["kn"] * cf["d"]  # the damping term is not a rounding detail here
    assert close(-F[0], fn, 1e-4) and abs(F[1]) < 1e-7 * abs(fn) and abs(F[2]) < 1e-7 * abs(fn)
    assert close(-F64[0], fn, 1e-9)
    assert np.all(T == 0) and w[0] == w[1] == w[2] == 0.0 and w[3] == np.float32(cf["h"])


def test_tangential_spring_stiffness():
    """Preset tangential displacement, no relative motion: F_t = -k_t delta_t with k_t = 8 G* sqrt(R* delta), and
    G* = 1 / (2 (2 - nu)(1 + nu) / E + ...) for the pair; the history is kept (below the Coulomb cap)."""
    d0 = 4e-6
    cf, F, T, F64, _, w = pair(3e-5, hist=(0.0, d0, 0.0, 0.0))
    ft = -cf["kt"] * f32(d0)
    assert abs(ft) < cf["mu"] * cf["kn"] * cf["d"]
    assert close(F[1], ft, 1e-4) and close(F64[1], ft, 1e-9)
    assert close(-F[0], cf["kn"] * cf["d"], 1e-4) and F[2] == 0.0
    assert w[1] == np.float32(d0) and w[0] == 0.0 and w[2] == 0.0


def test_tangential_damping_on_the_first_sliding_step():
    """From zero history, the first step of a sliding contact: delta_t = h v_t and
    F_t = -k_t delta_t - gamma_t v_t with gamma_t = -2 sqrt(5/6) beta sqrt(m* k_t)."""
    u = 0.005
    cf, F, T, F64, _, w = pair(3e-5, vA=(0, u, 0), vB=(0, -u, 0))
    vt = 2 * f32(u)
    ft = -(cf["kt"] * cf["h"] * vt) - cf["gt"] * vt
    assert cf["gt"] * vt > 5 * cf["kt"] * cf["h"] * vt  # the damping term dominates the first step
    assert abs(ft) < cf["mu"] * cf["kn"] * cf["d"]
    assert close(F[1], ft, 1e-4) and close(F64[1], ft, 1e-9)
    assert close(w[1], cf["h"] * vt, 1e-6) and w[3] == np.float32(cf["h"])


@pytest.mark.parametrize("u", [0.0, 0.004])
def test_coulomb_cap_back_solves_the_history(u):
    """A tangential force beyond mu |F_n| is scaled onto the cap and the history is back-solved from it:
    delta_t = (F_t + gamma_t v_t) / (-k_t), with or without sliding."""
    d0 = 1e-4
    cf, F, T, F64, _, w = pair(3e-5, hist=(0.0, d0, 0.0, 0.0), matA=(E, NU, 0.5, 0.3, 0.0), vA=(0, u, 0),
                               vB=(0, -u, 0))
    fn = cf["kn"] * cf["d"]
    vt = 2 * f32(u)
    assert cf["kt"] * d0 > 2 * cf["mu"] * fn
    assert close(F[1], -cf["mu"] * fn, 1e-4) and close(F64[1], -cf["mu"] * fn, 1e-9)
    want = (cf["mu"] * fn - cf["gt"] * vt) / cf["kt"]
    if u:
        assert abs(cf["gt"] * vt) > 0.2 * cf["mu"] * fn  # the damping share of the back-solve matters
    assert close(w[1], want, 1e-4)


def t_collision(cf):
    """The rolling-resistance switch: t_c = pi sqrt(m* / (k (1 - d^2))) with k = 4/3 E* R_e^(1/2), R_e = sqrt(R*),
    d = g / (2 sqrt(k m*)), g = -2 sqrt(5/3 m* E*) beta R_e^(1/4) (FullHertzianForceModel.cu:70-82)."""
    Re = np.sqrt(cf["R"])
    k = 4.0 / 3.0 * cf["E"] * np.sqrt(Re)
    g = -2.0 * np.sqrt(5.0 / 3.0 * cf["m"] * cf["E"]) * cf["beta"] * Re ** 0.25
    dc = g / (2.0 * np.sqrt(k * cf["m"]))
    assert dc < 1.0
    return np.pi * np.sqrt(cf["m"] / (k * (1.0 - dc * dc)))


@pytest.mark.parametrize("frac,spin", [(0.5, 40.0), (0.99, 40.0), (1.01, 40.0), (2.0, 40.0), (2.0, -25.0)])
def test_rolling_resistance(frac, spin):
    """Crr > 0: once the contact has lasted longer than t_c, a torque-only force Crr |F| along the rolling velocity
    of B relative to A at the contact point (v_rot = w_B x a_B - w_A x a_A, world frame); none before. Both spheres
    spin about z, so v_rot = -2 w a y-hat; mu = 0 keeps the tangential part out of |F|."""
    crr, h = 0.02, 1e-5
    cf0, *_ = pair(3e-5, matA=(E, NU, 0.5, 0.0, crr))
    tc = t_collision(cf0)
    dt0 = f32(frac * tc - h)  # the pass adds h first
    cf, F, T, F64, T64, w = pair(3e-5, wA=(0, 0, spin), wB=(0, 0, spin), matA=(E, NU, 0.5, 0.0, crr),
                                 hist=(0.0, 0.0, 0.0, dt0), h=h)
    fn = cf["kn"] * cf["d"]
    assert close(-F[0], fn, 1e-4) and F[1] == 0.0 and F[2] == 0.0
    if frac < 1.0:
        assert np.all(T == 0) and np.all(T64 == 0)
    else:
        want = np.array([0.0, -np.sign(spin) * cf["crr"] * fn, 0.0])  # Crr as stored (float)
        assert np.abs(T - want).max() <= 1e-4 * crr * fn
        assert np.abs(T64 - want).max() <= 1e-9 * crr * fn


@pytest.mark.parametrize("swap", [False, True])
def test_two_material_mixing(swap):
    """E* and G* by the compliance sums of both materials, CoR the smaller, mu and Crr the larger (an exchange of min
    and max, or of the two materials, changes the numbers); symmetric in the order of the materials."""
    m1 = (1e7, 0.30, 0.3, 0.2, 0.01)
    m2 = (4e7, 0.20, 0.8, 0.6, 0.03)
    mA, mB = (m2, m1) if swap else (m1, m2)
    rA, rB = 0.004, 0.006  # unequal radii and masses: R* and m* are the harmonic forms
    # normal, with damping from the smaller CoR
    cf, F, T, F64, _, _ = pair(3e-5, rA=rA, rB=rB, vA=(0.1, 0, 0), matA=mA, matB=mB)
    E_eff = 1.0 / ((1 - 0.09) / 1e7 + (1 - 0.04) / 4e7)
    assert close(cf["E"], E_eff, 1e-6) and cf["e"] == f32(0.3) and cf["mu"] == f32(0.6) and cf["crr"] == f32(0.03)
    assert close(cf["R"], f32(rA) * f32(rB) / (f32(rA) + f32(rB)), 1e-12)
    fn = cf["kn"] * cf["d"] + cf["gn"] * (-f32(0.1))
    assert close(-F[0], fn, 1e-4) and close(-F64[0], fn, 1e-9)
    # spring stiffness from G*
    cf, F, T, F64, _, _ = pair(3e-5, rA=rA, rB=rB, matA=mA, matB=mB, hist=(0.0, 2e-6, 0.0, 0.0))
    G_eff = 1.0 / (2 * 1.7 * 1.3 / 1e7 + 2 * 1.8 * 1.2 / 4e7)
    assert close(cf["G"], G_eff, 1e-6)
    assert close(F[1], -8.0 * G_eff * np.sqrt(cf["R"] * cf["d"]) * f32(2e-6), 1e-4)
    # the cap from the larger mu
    cf, F, T, F64, _, _ = pair(3e-5, rA=rA, rB=rB, matA=mA, matB=mB, hist=(0.0, 1e-4, 0.0, 0.0))
    assert close(F[1], -f32(0.6) * cf["kn"] * cf["d"], 1e-4)
    # rolling resistance from the larger Crr, past the collision time of the mixed contact
    tc = t_collision(cf)
    cf, F, T, F64, _, _ = pair(3e-5, rA=rA, rB=rB, matA=mA, matB=mB, wA=(0, 0, 30.0), wB=(0, 0, 30.0),
                               hist=(0.0, 0.0, 0.0, f32(2 * tc)))
    # |F| when the torque-only force is formed is the normal force's (the tangential part comes after it)
    assert close(np.linalg.norm(T), f32(0.03) * cf["kn"] * cf["d"], 1e-4)
