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
                  build_sim_params, decode_positions, empty_state, encode_positions)

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
    owner_scene: np.ndarray = None  # batch: scene index of every owner (None: one scene)
    anal_scene: np.ndarray = None   # batch: scene index of every analytical component (None: its owner's)
    n_scenes: int = 1


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


def fcc_neighbours(n, mx, my, mz):
    """The first n sites of an FCC lattice of mx x my x mz cubic cells (four sites per cell), in z-major order, as
    integer half-cell coordinates, with each site's 12 nearest neighbours (-1 where the lattice ends)."""
    base = np.array([[0, 0, 0], [1, 1, 0], [1, 0, 1], [0, 1, 1]], np.int64)
    ii, jj, kk = np.meshgrid(np.arange(mx), np.arange(my), np.arange(mz), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1)
    hc = (2 * cells[:, None, :] + base[None]).reshape(-1, 3)
    hc = hc[np.lexsort((hc[:, 0], hc[:, 1], hc[:, 2]))][:n]
    grid = np.full((2 * mx, 2 * my, 2 * mz), -1, np.int64)
    grid[hc[:, 0], hc[:, 1], hc[:, 2]] = np.arange(n)
    offs = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
            if abs(dx) + abs(dy) + abs(dz) == 2]
    nb = np.full((n, 12), -1, np.int64)
    dims = np.array([2 * mx, 2 * my, 2 * mz])
    for k, o in enumerate(offs):
        q = hc + np.array(o)
        ok = np.all((q >= 0) & (q < dims), 1)
        nb[ok, k] = grid[q[ok, 0], q[ok, 1], q[ok, 2]]
    return hc, nb


def relax_overlaps(x, r, nb, iters, step=0.4):
    """Pushes lattice neighbours apart: each iteration moves every sphere by step x half the overlap with each of its
    neighbours, along the line of centres (a Jacobi sweep over the fixed neighbour table; gathers only, so the result
    does not depend on summation order). Runs on the CPU through torch's threaded kernels when torch is present:
    never on the GPU, since torch's wheel carries a HIP runtime of its own and a process must not drive the device
    through two runtimes (the library's and torch's)."""
    try:
        import torch

        dev = "cpu"
        X = torch.from_numpy(np.ascontiguousarray(x, np.float64)).to(dev)
        R = torch.from_numpy(np.ascontiguousarray(r, np.float64)).to(dev)
        N = torch.from_numpy(nb).to(dev)
        valid = N >= 0
        Nc = torch.where(valid, N, torch.zeros_like(N))
        for _ in range(iters):
            d = X[:, None, :] - X[Nc]                       # n x 12 x 3, from each neighbour to the sphere
            dist = torch.sqrt((d * d).sum(-1))
            o = R[:, None] + R[Nc] - dist
            w = torch.where(valid & (o > 0), step * 0.5 * o / dist.clamp_min(1e-12), torch.zeros_like(o))
            X = X + (d * w[..., None]).sum(1)
        out = X.numpy()
        del X, R, N, Nc
        return out
    except ImportError:
        X = np.array(x, np.float64)
        valid = nb >= 0
        Nc = np.where(valid, nb, 0)
        for _ in range(iters):
            d = X[:, None, :] - X[Nc]
            dist = np.sqrt((d * d).sum(-1))
            o = r[:, None] + r[Nc] - dist
            w = np.where(valid & (o > 0), step * 0.5 * o / np.maximum(dist, 1e-12), 0.0)
            X = X + (d * w[..., None]).sum(1)
        return X


def polydisperse_bed(n=8_000_000, seed=1003, median_r=0.001, sigma=0.2, r_clip=(0.0005, 0.002), n_mat=4,
                     h=2e-6, plate_speed=0.05, beta=0.0, velocity_sigma=0.0, steps=500, solid_fraction=0.6,
                     relax_iters=12):
    """Lognormal spheres (median 1 mm, sigma 0.2, clipped to [0.5, 2] mm) with four materials assigned at random,
    packed at `solid_fraction` in an open-top box under a top plate (its own owner, a prescribed family moving down
    at plate_speed) one largest diameter above the bed. At 8M spheres the bed is 0.40 x 0.40 x 0.41 m.

    Packing: sphere i takes site i of a face-centred cubic lattice (sites numbered x fastest, then y, then z, as a grid
    sampler emits them; radii drawn independently per site; a jitter of 2% of the spacing) whose spacing gives the
    requested solid fraction; a few sweeps pushing lattice neighbours apart then spread the overlaps of the larger
    spheres. Residual overlaps remain (about 1.5 overlapping pairs per sphere,
    mean depth ~5% of a radius); settle_bed() relaxes them on the device."""
    rng = np.random.default_rng(seed)
    radius = np.clip(median_r * np.exp(sigma * rng.standard_normal(n)), *r_clip).astype(np.float32)
    r64 = radius.astype(np.float64)
    # FCC: volume per site a^3 / sqrt(2) for nearest-neighbour spacing a
    a = float((4.0 / 3.0 * np.pi * np.sqrt(2.0) * np.mean(r64 ** 3) / solid_fraction) ** (1.0 / 3.0))
    L = a * np.sqrt(2.0)
    mx = max(2, int(np.ceil((n / (4 * 1.04)) ** (1.0 / 3.0))))  # a bed about as tall as it is wide
    mz = int(np.ceil(n / (4 * mx * mx)))
    hc, nb = fcc_neighbours(n, mx, mx, mz)
    dmax = 2.0 * float(radius.max())
    margin = dmax
    bx = mx * L + 2 * margin
    bed_h = mz * L
    bz = bed_h + 3 * dmax + margin
    box = (float(f32(bx)), float(f32(bx)), float(f32(bz)))
    lbf = tuple(float(f32(-f32(b) / f32(2.0))) for b in box)
    xyz = (hc * 0.5 + 0.25) * L + rng.uniform(-0.02 * a, 0.02 * a, size=(n, 3))
    xyz = relax_overlaps(xyz, r64, nb, relax_iters)
    # inside the walls: no sphere deeper into a wall than its radius allows
    lo = np.array([margin, margin, 0.0])
    xyz = np.maximum(xyz + lo, r64[:, None] * 0.98)
    xyz[:, :2] = np.minimum(xyz[:, :2], box[0] - r64[:, None] * 0.98)
    xyz = xyz + np.array(lbf)
    mat = rng.integers(0, n_mat, n).astype(np.uint8)
    mats = np.zeros(n_mat, MATERIAL_DTYPE)
    for m in range(n_mat):
        mats[m] = (1e7 * (1 + m), 0.25 + 0.02 * m, 0.4 + 0.1 * m, 0.3 + 0.1 * m, 0.0)
    mass = (RHO * 4.0 / 3.0 * np.pi * r64 ** 3).astype(np.float32)
    moi = (2.0 / 5.0 * mass.astype(np.float64) * r64 ** 2).astype(np.float32)
    params = build_sim_params(box, lbf, float(radius.min()), h, beta=beta)
    walls = box_planes(lbf, box, owner=n, material=0, top=False)
    plate = np.zeros(1, ANAL_DTYPE)
    plate[0]["owner"] = n + 1
    plate[0]["type"] = 0
    plate[0]["material"] = n_mat - 1
    plate[0]["relPos"] = np.array([0.0, 0.0, 0.0], np.float32)
    plate[0]["rot"] = np.array([0.0, 0.0, -1.0], np.float32)
    plate[0]["mass"] = f32(1e15)
    anal = np.concatenate([walls, plate])
    rule = np.zeros(1, RULE_DTYPE)[0]
    rule["used"] = 1
    rule["linVelPrescribed"] = rule["rotVelPrescribed"] = 1
    rule["velSet"] = 0x3F
    rule["vel"][:] = (0.0, 0.0, -plate_speed, 0.0, 0.0, 0.0)
    mask, rules = default_families({10: rule})
    vel, omg, q = random_motion(rng, n, velocity_sigma)
    top = float((xyz[:, 2] + r64).max())
    sphere_volume = float((4.0 / 3.0 * np.pi * r64 ** 3).sum())
    fill = sphere_volume / ((mx * L) * (mx * L) * (top - lbf[2]))
    sc = _finish("polydisperse_bed", params, xyz, np.arange(n), radius, np.zeros((n, 3)), mat, mass,
                 np.repeat(moi[:, None], 3, 1), mats, anal, 2, 1, steps, vel=vel, mask=mask, rules=rules,
                 info=dict(box=box, spacing=a, solid_fraction=fill, bed_top=top))
    # the plate owner: centred above the bed, in the prescribed family
    plate_xyz = np.array([[lbf[0] + box[0] / 2, lbf[1] + box[1] / 2, top + dmax]], np.float32)
    v, lx, ly, lz = encode_positions(params, plate_xyz)
    st = sc.state
    st["voxelID"][n + 1], st["locX"][n + 1], st["locY"][n + 1], st["locZ"][n + 1] = v[0], lx[0], ly[0], lz[0]
    st["family"][n + 1] = 10
    st["vZ"][n + 1] = -plate_speed
    return sc


def settle_bed(dev, sc, steps=400, cor=0.05):
    """Relaxes a packed bed's residual overlaps on the device: `steps` steps with every material's restitution
    lowered to `cor` (strongly damped contacts), then the scene's materials back and all sphere velocities zeroed.
    The plate keeps its prescribed motion (it moves steps x h x plate speed, far less than its clearance)."""
    damped = sc.materials.copy()
    damped["CoR"] = np.float32(cor)
    dev.upload_materials(damped)
    dev.step(steps, 1)
    dev.upload_materials(sc.materials)
    st = dev.get_state()
    for k in ("vX", "vY", "vZ", "omgX", "omgY", "omgZ"):
        st[k][:sc.n_clumps] = 0.0
    dev.set_state(st)
    dev.clear_contacts()


def adjoint_drop(n=50_000, seed=1005, r_range=(0.002, 0.004), n_mat=4, velocity_sigma=0.5, h=1e-5, steps=200,
                 window=20, beta=None, gap=0.02, E0=2e7):
    """Spheres of radii uniform in r_range with n_mat materials assigned at random, on a cubic lattice of pitch
    2 r_max (1 + gap) whose bottom layer rests on a floor plane, with seeded initial velocities N(0, velocity_sigma):
    the differentiation workload (steps at h, contact list rebuilt every `window` steps with drift margin beta,
    default 1.2 x a 4.5-sigma speed over one window)."""
    rng = np.random.default_rng(seed)
    radius = rng.uniform(r_range[0], r_range[1], n).astype(np.float32)
    pitch = 2.0 * r_range[1] * (1.0 + gap)
    side = max(2, int(np.ceil((2.0 * n) ** (1.0 / 3.0))))
    layers = int(np.ceil(n / (side * side)))
    bx = (side + 2) * pitch
    bz = max((layers + 4) * pitch, bx)  # a cube: the voxel split of the reference sizing takes near-cubic boxes
    box = (float(f32(bx)), float(f32(bx)), float(f32(bz)))
    lbf = tuple(float(f32(-f32(b) / f32(2.0))) for b in box)
    ii, jj, kk = np.meshgrid(np.arange(side), np.arange(side), np.arange(layers), indexing="ij")
    order = np.lexsort((ii.ravel(), jj.ravel(), kk.ravel()))
    i, j, k = ii.ravel()[order][:n], jj.ravel()[order][:n], kk.ravel()[order][:n]
    floor = lbf[2] + pitch
    xyz = np.stack([lbf[0] + 1.5 * pitch + i * pitch, lbf[1] + 1.5 * pitch + j * pitch,
                    floor + radius.astype(np.float64) * (1.0 - 1e-3) * (k == 0) + (k > 0) * (0.5 * pitch + k * pitch)],
                   1)  # the bottom layer presses 0.1% of its radius into the floor: no contact starts at zero overlap
    mat = rng.integers(0, n_mat, n).astype(np.uint8)
    mats = np.zeros(n_mat, MATERIAL_DTYPE)
    for m in range(n_mat):
        mats[m] = (E0 * (1 + m), 0.25 + 0.03 * m, 0.5 + 0.1 * m, 0.3 + 0.1 * m, 0.0)
    mass = (RHO * 4.0 / 3.0 * np.pi * radius.astype(np.float64) ** 3).astype(np.float32)
    moi = (2.0 / 5.0 * mass.astype(np.float64) * radius.astype(np.float64) ** 2).astype(np.float32)
    if beta is None:
        beta = 1.2 * 4.5 * velocity_sigma * h * window
    params = build_sim_params(box, lbf, float(radius.min()), h, beta=float(f32(beta)))
    plane = np.zeros(1, ANAL_DTYPE)
    plane[0]["owner"] = n
    plane[0]["type"] = 0
    plane[0]["material"] = 0
    plane[0]["relPos"] = np.array([0.0, 0.0, floor], np.float32)
    plane[0]["rot"] = np.array([0.0, 0.0, 1.0], np.float32)
    plane[0]["mass"] = f32(1e15)
    vel, _, _ = random_motion(rng, n, velocity_sigma)
    return _finish("adjoint_drop", params, xyz, np.arange(n), radius, np.zeros((n, 3)), mat, mass,
                   np.repeat(moi[:, None], 3, 1), mats, plane, 1, window, steps, vel=vel,
                   info=dict(box=box, pitch=pitch, window=window))


def voxel_split(box):
    """Bits per axis (summing to 64) and sub-voxel unit l covering `box` with the smallest l."""
    best = None
    for a in range(8, 41):
        for b in range(8, 41):
            c = 64 - a - b
            if c < 8 or c > 40:
                continue
            l = max(box[0] / 2.0 ** (a + 16), box[1] / 2.0 ** (b + 16), box[2] / 2.0 ** (c + 16))
            if best is None or l < best[1]:
                best = ((a, b, c), l)
    nv, l = best
    return nv, float(np.nextafter(f32(l), f32(np.inf)))


def batched_columns(n_scenes=256, seed=1004, s_range=(2_000, 200_000), r_range=(0.001, 0.003), height=0.3,
                    h=5e-6, steps=1000, update_freq=1, beta=0.0, total_scale=1.0, pitch_scale=1.05):
    """Independent settling columns batched into one system. Scene k has S_k spheres (log-uniform in s_range, times
    total_scale), radii uniform in r_range, its own material (E log-uniform in [1e6, 1e8], mu in [0.2, 0.8], CoR in
    [0.3, 0.9], nu 0.3) and its own open-top box of five planes. All columns are about `height` tall; their bases grow
    with S_k, and they sit on shelves in the xy plane, two lattice pitches apart, so no pair of spheres from different
    scenes can ever be in range. The walls of every scene belong to one fixed owner; each plane is tagged with its
    scene."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(s_range[0]), np.log(s_range[1])
    counts = np.maximum(1, np.round(np.exp(rng.uniform(lo, hi, n_scenes)) * total_scale)).astype(np.int64)
    pitch = 2.0 * r_range[1] * pitch_scale
    layers = max(1, int(height / pitch))
    sides = np.ceil(np.sqrt(np.ceil(counts / layers))).astype(np.int64)
    widths = sides * pitch + pitch  # half a pitch of wall clearance on each side
    gap = 2.0 * pitch
    # shelf packing: widest first, rows of bounded length
    order = np.argsort(-widths, kind="stable")
    row_len = float(np.sqrt((widths + gap) ** 2 @ np.ones(n_scenes))) * 1.05
    origin = np.zeros((n_scenes, 2))
    x = y = row_h = 0.0
    for k in order:
        if x > 0 and x + widths[k] > row_len:
            x = 0.0
            y += row_h + gap
            row_h = 0.0
        origin[k] = (x, y)
        x += widths[k] + gap
        row_h = max(row_h, widths[k])
    world = (float(origin[:, 0].max() + widths.max() + gap), float(y + row_h + gap), float(height + 4 * pitch))
    box = tuple(float(f32(b)) for b in world)
    lbf = (0.0, 0.0, 0.0)
    xyz, rad, mat, owner_scene = [], [], [], []
    anal = []
    anal_scene = []
    mats = np.zeros(n_scenes, MATERIAL_DTYPE)
    n_total = int(counts.sum())
    for k in range(n_scenes):
        E = float(np.exp(rng.uniform(np.log(1e6), np.log(1e8))))
        mats[k] = (E, 0.3, rng.uniform(0.3, 0.9), rng.uniform(0.2, 0.8), 0.0)
        n = int(counts[k])
        side = int(sides[k])
        i = np.arange(n) % side
        j = (np.arange(n) // side) % side
        l = np.arange(n) // (side * side)
        p = np.stack([origin[k, 0] + pitch * (i + 1.0), origin[k, 1] + pitch * (j + 1.0), pitch * (l + 0.55)], 1)
        p += rng.uniform(-0.05 * pitch, 0.05 * pitch, size=p.shape)
        xyz.append(p)
        rad.append(rng.uniform(r_range[0], r_range[1], n))
        mat.append(np.full(n, k))
        owner_scene.append(np.full(n, k))
        w = float(widths[k])
        walls = box_planes(lbf, (w, w, world[2]), owner=n_total, material=k, top=False)
        walls["relPos"][:, 0] += np.float32(origin[k, 0])
        walls["relPos"][:, 1] += np.float32(origin[k, 1])
        anal.append(walls)
        anal_scene.append(np.full(len(walls), k))
    xyz = np.concatenate(xyz)
    radius = np.concatenate(rad).astype(np.float32)
    mass = (RHO * 4.0 / 3.0 * np.pi * radius.astype(np.float64) ** 3).astype(np.float32)
    moi = (2.0 / 5.0 * mass.astype(np.float64) * radius.astype(np.float64) ** 2).astype(np.float32)
    # explicit voxel split (InstructBoxDomainNumVoxel): a flat world gets more bits along x and y
    nv, l = voxel_split(box)
    params = build_sim_params(box, lbf, float(radius.min()), h, beta=beta, nv=nv, l=l)
    sc = _finish("batched_columns", params, xyz, np.arange(n_total), radius, np.zeros((n_total, 3)),
                 np.concatenate(mat).astype(np.uint8), mass, np.repeat(moi[:, None], 3, 1), mats,
                 np.concatenate(anal), 1, update_freq, steps, info=dict(counts=counts, box=box, pitch=pitch))
    sc.owner_scene = np.concatenate(owner_scene + [np.zeros(1, np.int64)]).astype(np.uint32)
    sc.anal_scene = np.concatenate(anal_scene).astype(np.uint32)
    sc.n_scenes = n_scenes
    return sc


def split_batch(batch, own_world=False):
    """The scenes of a batch as separate systems. By default each keeps the batch's world (same voxel grid, materials
    and planes), so a scene run alone follows exactly the trajectory it follows inside the batch; own_world=True gives
    each scene a world of its own column's size instead (as a scene set up on its own would have)."""
    out = []
    n_sph = batch.n_spheres
    if own_world:
        world_pos = decode_positions(batch.params, batch.state["voxelID"], batch.state["locX"], batch.state["locY"],
                                     batch.state["locZ"]).astype(np.float64)
    for k in range(batch.n_scenes):
        sph = np.nonzero(batch.owner_scene[:n_sph] == k)[0]
        n = len(sph)
        walls = batch.anal[batch.anal_scene == k].copy()
        walls["owner"] = n
        owners = np.concatenate([batch.sph_owner[sph], [batch.n_owners - 1]])  # single-sphere owners, then the walls
        st = {key: v[owners].copy() for key, v in batch.state.items()}
        params = batch.params
        if own_world:
            # the column's own world: its walls' extent, origin at the column's corner
            x0 = float(walls["relPos"][:, 0].min())
            y0 = float(walls["relPos"][:, 1].min())
            w = float(walls["relPos"][:, 0].max()) - x0
            box = (float(f32(w)), float(f32(w)), float(batch.info["box"][2]))
            nv, l = voxel_split(box)
            params = build_sim_params(box, (0.0, 0.0, 0.0), float(batch.sph_radius[sph].min()), batch.params.h,
                                      beta=batch.params.beta, nv=nv, l=l)
            xyz = world_pos[owners].copy()
            xyz[:, 0] -= x0
            xyz[:, 1] -= y0
            xyz[-1] = 0.0
            v, lx, ly, lz = encode_positions(params, xyz.astype(np.float32))
            st.update(voxelID=v, locX=lx, locY=ly, locZ=lz)
            walls["relPos"][:, 0] -= np.float32(x0)
            walls["relPos"][:, 1] -= np.float32(y0)
        sc = Scene(name=f"{batch.name}[{k}]", params=params, n_spheres=n, n_owners=n + 1, n_clumps=n,
                   sph_owner=np.arange(n, dtype=np.uint32), sph_radius=batch.sph_radius[sph].copy(),
                   sph_relpos=batch.sph_relpos[sph].copy(), sph_mat=batch.sph_mat[sph].copy(),
                   owner_mass=batch.owner_mass[owners].copy(), owner_moi=batch.owner_moi[owners].copy(),
                   materials=batch.materials, anal=walls, mask=batch.mask, rules=batch.rules, state=st,
                   update_freq=batch.update_freq, steps=batch.steps, info=dict(scene=k))
        out.append(sc)
    return out
