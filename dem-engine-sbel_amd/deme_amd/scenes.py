"""Synthetic workloads of the DEM hot path, built the way DEMSolver::Initialize lays data out.

sphere_settling   monodisperse spheres settling in a plane box (20,000 spheres r = 5 mm at full size)
clump_settling    three-sphere clumps settling in a plane box (1,000,000 clumps at full size)

All randomness is seeded (numpy PCG64, seed = 1000 + workload index) so every run, the oracle's and the device's,
starts from identical integer positions.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .abi import (ANAL_DTYPE, MATERIAL_DTYPE, NUM_FAMILIES, NUM_MASK_ENTRIES, RULE_DTYPE, SimParams,
                  build_sim_params, empty_state, encode_positions)

RHO = 2600.0


@dataclass
class Scene:
    name: str
    params: SimParams
    n_spheres: int
    n_owners: int
    n_clumps: int
    sph_owner: np.ndarray
    sph_radius: np.ndarray
    sph_relpos: np.ndarray
    sph_mat: np.ndarray
    owner_mass: np.ndarray
    owner_moi: np.ndarray
    materials: np.ndarray
    anal: np.ndarray
    mask: np.ndarray
    rules: np.ndarray
    state: dict
    update_freq: int = 1
    steps: int = 1
    info: dict = field(default_factory=dict)


def f32(x):
    return np.float32(x)


def default_families(extra_rules=None):
    """Family mask (all contacts allowed) and rules with family 255 fixed (validateUserInputs, APIPrivate.cpp:906)."""
    mask = np.zeros(NUM_MASK_ENTRIES, np.uint8)
    rules = np.zeros(NUM_FAMILIES, RULE_DTYPE)
    r = rules[255]
    r["used"] = 1
    r["linVelPrescribed"] = r["rotVelPrescribed"] = r["linPosPrescribed"] = r["rotPosPrescribed"] = 1
    r["velSet"] = 0x3F
    rules[255] = r
    for fam, rule in (extra_rules or {}).items():
        rules[fam] = rule
    return mask, rules


def box_planes(lbf, box, owner, material, top=True):
    """The bounding-box planes of InstructBoxDomainBoundingBC("all"/"top_open") (APIPrivate.cpp:660-687)."""
    L = [float(f32(v)) for v in lbf]
    S = [float(f32(v)) for v in box]
    pos = [(L[0] + S[0] / 2., L[1] + S[1] / 2., L[2]),
           (L[0], L[1] + S[1] / 2., L[2] + S[2] / 2.),
           (L[0] + S[0], L[1] + S[1] / 2., L[2] + S[2] / 2.),
           (L[0] + S[0] / 2., L[1], L[2] + S[2] / 2.),
           (L[0] + S[0] / 2., L[1] + S[1], L[2] + S[2] / 2.)]
    nrm = [(0, 0, 1), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)]
    if top:
        pos.append((L[0] + S[0] / 2., L[1] + S[1] / 2., L[2] + S[2]))
        nrm.append((0, 0, -1))
    a = np.zeros(len(pos), ANAL_DTYPE)
    for i, (p, n) in enumerate(zip(pos, nrm)):
        a[i]["owner"] = owner
        a[i]["type"] = 0
        a[i]["material"] = material
        a[i]["relPos"] = np.array(p, np.float32)
        a[i]["rot"] = np.array(n, np.float32)
        a[i]["mass"] = f32(1e15)
    return a


def random_motion(rng, n, velocity_sigma=0.0, omega_sigma=0.0, random_orientation=False):
    """Seeded initial velocities N(0, sigma), angular velocities and uniformly random unit quaternions."""
    vel = rng.normal(0.0, velocity_sigma, size=(n, 3)).astype(np.float32) if velocity_sigma > 0 else None
    omg = rng.normal(0.0, omega_sigma, size=(n, 3)).astype(np.float32) if omega_sigma > 0 else None
    q = None
    if random_orientation:
        q = rng.normal(size=(n, 4))
        q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    return vel, omg, q


def _finish(name, params, xyz_clumps, sph_owner, sph_radius, sph_relpos, sph_mat, clump_mass, clump_moi, materials,
            anal, n_ext, update_freq, steps, vel=None, info=None, mask=None, rules=None, families=None, omg=None,
            quat=None):
    n_clumps = xyz_clumps.shape[0]
    n_owners = n_clumps + n_ext
    owner_mass = np.concatenate([clump_mass.astype(np.float32), np.full(n_ext, 1e15, np.float32)])
    owner_moi = np.concatenate([clump_moi.astype(np.float32), np.full((n_ext, 3), 1e15, np.float32)])
    ext_pos = np.zeros((n_ext, 3), np.float32)
    vox, lx, ly, lz = encode_positions(params, np.concatenate([xyz_clumps.astype(np.float32), ext_pos]))
    st = empty_state(n_owners)
    st.update(voxelID=vox, locX=lx, locY=ly, locZ=lz)
    st["oriQw"][:] = 1.0
    if vel is not None:
        st["vX"][:n_clumps], st["vY"][:n_clumps], st["vZ"][:n_clumps] = vel[:, 0], vel[:, 1], vel[:, 2]
    if omg is not None:
        st["omgX"][:n_clumps], st["omgY"][:n_clumps], st["omgZ"][:n_clumps] = omg[:, 0], omg[:, 1], omg[:, 2]
    if quat is not None:
        st["oriQw"][:n_clumps], st["oriQx"][:n_clumps] = quat[:, 0], quat[:, 1]
        st["oriQy"][:n_clumps], st["oriQz"][:n_clumps] = quat[:, 2], quat[:, 3]
    if families is not None:
        st["family"][:n_clumps] = families
    st["family"][n_clumps:] = 255
    if mask is None or rules is None:
        mask, rules = default_families()
    return Scene(name=name, params=params, n_spheres=len(sph_owner), n_owners=n_owners, n_clumps=n_clumps,
                 sph_owner=np.ascontiguousarray(sph_owner, np.uint32),
                 sph_radius=np.ascontiguousarray(sph_radius, np.float32),
                 sph_relpos=np.ascontiguousarray(sph_relpos, np.float32),
                 sph_mat=np.ascontiguousarray(sph_mat, np.uint8), owner_mass=owner_mass,
                 owner_moi=np.ascontiguousarray(owner_moi), materials=materials, anal=anal, mask=mask, rules=rules,
                 state=st, update_freq=update_freq, steps=steps, info=info or {})


def sphere_settling(n=20000, seed=1001, r=0.005, box=(0.3, 0.3, 0.6), h=1e-5, steps=1000, update_freq=1,
                    E=1e7, nu=0.3, CoR=0.5, mu=0.5, Crr=0.0, jitter=0.1, spacing=2.2, velocity_sigma=0.0,
                    omega_sigma=0.0, random_orientation=False, beta=0.0, n_big=0, r_big=None):
    """Monodisperse FP32 spheres seeded on a cubic lattice (spacing 2.2 r, jitter +-0.1 r), settling under gravity
    in a box of analytical planes. Optionally n_big spheres of radius r_big sit on a row above the lattice (clear of
    it), which widens the detection cells so that many small spheres share one cell."""
    rng = np.random.default_rng(seed)
    box = tuple(float(f32(b)) for b in box)
    lbf = tuple(float(f32(-f32(b) / f32(2.0))) for b in box)
    s = spacing * r
    nx = int(np.floor((box[0] - 2 * r) / s)) + 1
    ny = int(np.floor((box[1] - 2 * r) / s)) + 1
    per_layer = nx * ny
    nz = int(np.ceil(n / per_layer))
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    order = np.lexsort((ii.ravel(), jj.ravel(), kk.ravel()))  # x fastest, then y, then z
    i, j, k = ii.ravel()[order][:n], jj.ravel()[order][:n], kk.ravel()[order][:n]
    x0 = lbf[0] + (box[0] - (nx - 1) * s) / 2
    y0 = lbf[1] + (box[1] - (ny - 1) * s) / 2
    z0 = lbf[2] + 1.1 * r
    xyz = np.stack([x0 + i * s, y0 + j * s, z0 + k * s], 1)
    xyz += rng.uniform(-jitter * r, jitter * r, size=xyz.shape)
    radius = np.full(n, r, np.float32)
    if n_big:
        rb = float(f32(r_big))
        zb = float(xyz[:, 2].max()) + r + 1.5 * rb
        xb = lbf[0] + rb * 1.5 + np.arange(n_big) * 2.5 * rb
        if xb[-1] + rb > lbf[0] + box[0] or zb + rb > lbf[2] + box[2]:
            raise ValueError("large spheres do not fit above the lattice")
        big = np.stack([xb, np.full(n_big, lbf[1] + box[1] / 2), np.full(n_big, zb)], 1)
        xyz = np.concatenate([xyz, big])
        radius = np.concatenate([radius, np.full(n_big, rb, np.float32)])
    m = n + n_big
    mass = (RHO * 4.0 / 3.0 * np.pi * radius.astype(np.float64) ** 3).astype(np.float32)
    moi = (2.0 / 5.0 * mass.astype(np.float64) * radius.astype(np.float64) ** 2).astype(np.float32)
    params = build_sim_params(box, lbf, float(f32(r)), h, beta=beta)
    mats = np.zeros(1, MATERIAL_DTYPE)
    mats[0] = (E, nu, CoR, mu, Crr)
    anal = box_planes(lbf, box, owner=m, material=0)
    vel, omg, q = random_motion(rng, m, velocity_sigma, omega_sigma, random_orientation)
    return _finish("sphere_settling", params, xyz, np.arange(m), radius, np.zeros((m, 3)), np.zeros(m), mass,
                   np.repeat(moi[:, None], 3, 1), mats, anal, 1, update_freq, steps, vel=vel, omg=omg, quat=q,
                   info=dict(r=r, box=box))


def clump_template(seed=1002, r0=0.002, offset_frac=0.6, samples=2_000_000):
    """Three overlapping spheres on the x axis (centres 0.6 r0 apart, radii r0 * U[0.8, 1.2]); mass, centroid and
    principal moments of the union by seeded Monte Carlo at density 2600."""
    rng = np.random.default_rng(seed)
    radii = (r0 * rng.uniform(0.8, 1.2, 3)).astype(np.float32)
    cx = np.array([-offset_frac * r0, 0.0, offset_frac * r0])
    lo = np.array([cx.min() - radii.max(), -radii.max(), -radii.max()])
    hi = np.array([cx.max() + radii.max(), radii.max(), radii.max()])
    pts = rng.uniform(lo, hi, size=(samples, 3))
    inside = np.zeros(samples, bool)
    for c, rr in zip(cx, radii):
        inside |= ((pts[:, 0] - c) ** 2 + pts[:, 1] ** 2 + pts[:, 2] ** 2) <= float(rr) ** 2
    vol = np.prod(hi - lo) * inside.mean()
    p = pts[inside]
    # the union is symmetric about the x axis: centroid on the axis, Iyy = Izz
    com = np.array([p[:, 0].mean(), 0.0, 0.0])
    q = p - com
    mass = RHO * vol
    Ixx = mass * np.mean(q[:, 1] ** 2 + q[:, 2] ** 2)
    Iyz = mass * np.mean(q[:, 0] ** 2 + 0.5 * (q[:, 1] ** 2 + q[:, 2] ** 2))
    rel = np.stack([cx - com[0], np.zeros(3), np.zeros(3)], 1).astype(np.float32)
    return dict(radii=radii, relpos=rel, mass=np.float32(mass), moi=np.array([Ixx, Iyz, Iyz], np.float32))


def clump_settling(n_clumps=1_000_000, seed=1002, box=(0.5, 0.5, 1.0), h=5e-6, steps=2000, update_freq=1,
                   beta_frac=0.0, E=5e7, nu=0.3, CoR=0.6, mu=0.4, Crr=0.0, spacing=None, compaction=0.998,
                   jitter=0.1, velocity_sigma=0.0, omega_sigma=0.0, random_orientation=False):
    """Three-sphere clumps on an orthorhombic lattice settling in a plane box. update_freq > 1 runs with a drift
    margin beta = beta_frac * smallest radius.

    Default lattice: a compacted bed -- neighbour spacing = compaction x the clumps' contact distance along each
    axis, bottom layer resting on the floor -- so the contact path is loaded from the first step (a loose lattice
    would free-fall for the whole run)."""
    tpl = clump_template(seed=seed)
    rng = np.random.default_rng(seed + 17)
    box = tuple(float(f32(b)) for b in box)
    lbf = tuple(float(f32(-f32(b) / f32(2.0))) for b in box)
    xs = tpl["relpos"][:, 0]
    x_lo = float((xs - tpl["radii"]).min())
    x_hi = float((xs + tpl["radii"]).max())
    half_x = max(-x_lo, x_hi)
    half_yz = float(tpl["radii"].max())
    if spacing is None:
        spacing = ((x_hi - x_lo) * compaction, 2 * half_yz * compaction, 2 * half_yz * compaction)
    sx, sy, sz = spacing
    margin = max(half_x, half_yz) + 0.0002
    z_floor = half_yz * compaction
    nx = int(np.floor((box[0] - 2 * margin) / sx)) + 1
    ny = int(np.floor((box[1] - 2 * margin) / sy)) + 1
    nz = int(np.ceil(n_clumps / (nx * ny)))
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    order = np.lexsort((ii.ravel(), jj.ravel(), kk.ravel()))
    i, j, k = ii.ravel()[order][:n_clumps], jj.ravel()[order][:n_clumps], kk.ravel()[order][:n_clumps]
    x0 = lbf[0] + (box[0] - (nx - 1) * sx) / 2
    y0 = lbf[1] + (box[1] - (ny - 1) * sy) / 2
    z0 = lbf[2] + z_floor
    xyz = np.stack([x0 + i * sx, y0 + j * sy, z0 + k * sz], 1)
    r_mean = float(tpl["radii"].mean())
    xyz += rng.uniform(-jitter * 0.02 * r_mean, jitter * 0.02 * r_mean, size=xyz.shape)
    if xyz[:, 2].max() + half_yz > lbf[2] + box[2]:
        raise ValueError("lattice does not fit the box")
    rmin = float(tpl["radii"].min())
    beta = float(f32(beta_frac * rmin))
    params = build_sim_params(box, lbf, rmin, h, beta=beta)
    mats = np.zeros(1, MATERIAL_DTYPE)
    mats[0] = (E, nu, CoR, mu, Crr)
    anal = box_planes(lbf, box, owner=n_clumps, material=0)
    n_s = 3 * n_clumps
    sph_owner = np.repeat(np.arange(n_clumps, dtype=np.uint32), 3)
    sph_radius = np.tile(tpl["radii"], n_clumps)
    sph_relpos = np.tile(tpl["relpos"], (n_clumps, 1))
    vel, omg, q = random_motion(rng, n_clumps, velocity_sigma, omega_sigma, random_orientation)
    return _finish("clump_settling", params, xyz, sph_owner, sph_radius, sph_relpos, np.zeros(n_s),
                   np.full(n_clumps, tpl["mass"]), np.tile(tpl["moi"], (n_clumps, 1)), mats, anal, 1, update_freq,
                   steps, vel=vel, omg=omg, quat=q, info=dict(template=tpl, box=box, beta=beta))


def spheres_scene(xyz, radius, box, lbf=None, h=1e-5, beta=0.0, g=(0.0, 0.0, -9.81), E=1e7, nu=0.3, CoR=0.5,
                  mu=0.5, Crr=0.0, planes=True, vel=None, omg=None, quat=None, families=None, rules=None, mask=None,
                  integrator="extended_taylor", name="spheres"):
    """Single-sphere owners at given centres (world coordinates) in a box whose left-bottom-front corner defaults to
    (-box/2); planes=True adds the six bounding planes (family 255), planes=False none."""
    xyz = np.atleast_2d(np.asarray(xyz, np.float64))
    n = xyz.shape[0]
    radius = np.broadcast_to(np.asarray(radius, np.float32), (n,)).copy()
    box = tuple(float(f32(b)) for b in box)
    if lbf is None:
        lbf = tuple(float(f32(-f32(b) / f32(2.0))) for b in box)
    params = build_sim_params(box, lbf, float(radius.min()), h, beta=beta, g=g, integrator=integrator)
    mats = np.zeros(1, MATERIAL_DTYPE)
    mats[0] = (E, nu, CoR, mu, Crr)
    mass = (RHO * 4.0 / 3.0 * np.pi * radius.astype(np.float64) ** 3).astype(np.float32)
    moi = (2.0 / 5.0 * mass.astype(np.float64) * radius.astype(np.float64) ** 2).astype(np.float32)
    if planes:
        anal = box_planes(lbf, box, owner=n, material=0)
        n_ext = 1
    else:
        anal = np.zeros(0, ANAL_DTYPE)
        n_ext = 0
    if mask is None or rules is None:
        m0, r0 = default_families()
        mask = m0 if mask is None else mask
        rules = r0 if rules is None else rules
    to32 = (lambda a: None if a is None else np.atleast_2d(np.asarray(a, np.float32)))
    return _finish(name, params, xyz, np.arange(n), radius, np.zeros((n, 3)), np.zeros(n), mass,
                   np.repeat(moi[:, None], 3, 1), mats, anal, n_ext, 1, 1, vel=to32(vel), omg=to32(omg),
                   quat=to32(quat), families=families, mask=mask, rules=rules, info=dict(box=box))
