"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the checker or the
reported CPU baseline. Its interface mirrors deme_amd.abi.DeviceSystem so a test runs the same sequence on both.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "dem-engine-sbel_amd"))
from deme_amd.abi import Counts, OwnerState, SimParams, empty_state, make_state_struct, ptr  # noqa: E402

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise OSError(f"oracle library missing at {LIB}; run `make -C oracle`")
    lib = C.CDLL(LIB)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    lib.ora_create.restype = vp
    lib.ora_create.argtypes = [C.POINTER(SimParams), u32, vp, vp, vp, vp, u32, vp, vp, u32, vp, u32, vp, vp, vp]
    lib.ora_destroy.argtypes = [vp]
    lib.ora_set_threads.argtypes = [i32]
    lib.ora_set_state.argtypes = [vp, C.POINTER(OwnerState)]
    lib.ora_set_scenes.argtypes = [vp, u32, vp, vp]
    lib.ora_get_state.argtypes = [vp, C.POINTER(OwnerState)]
    lib.ora_get_counts.argtypes = [vp, C.POINTER(Counts)]
    for n in ("ora_cd_bins", "ora_cd_pairs", "ora_cd", "ora_force", "ora_collect", "ora_integrate"):
        getattr(lib, n).argtypes = [vp]
        getattr(lib, n).restype = i32
    lib.ora_step.argtypes = [vp, i32, i32]
    lib.ora_step.restype = i32
    lib.ora_get_bin_counts.argtypes = [vp, vp, vp]
    lib.ora_get_bin_pairs.argtypes = [vp, vp, vp]
    lib.ora_get_active_bins.argtypes = [vp, vp, vp]
    lib.ora_get_contacts.argtypes = [vp, vp, vp, vp, vp]
    lib.ora_set_contacts.argtypes = [vp, u64, vp, vp, vp]
    lib.ora_set_contacts.restype = i32
    lib.ora_brute_force_pairs.argtypes = [vp, vp, vp, u64]
    lib.ora_brute_force_pairs.restype = u64
    lib.ora_get_contact_forces.argtypes = [vp, vp, vp, vp, vp]
    lib.ora_get_wildcards.argtypes = [vp, vp]
    lib.ora_set_wildcards.argtypes = [vp, vp]
    lib.ora_set_contact_forces.argtypes = [vp, vp, vp, vp, vp]
    lib.ora_force_fp64.argtypes = [vp, vp, vp]
    lib.ora_get_time.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    lib.ora_set_time.argtypes = [vp, C.c_double]
    lib.ora_phase_times.argtypes = [vp, vp, i32]
    lib.ora_figure_out_nv.argtypes = [C.c_float] * 3 + [C.POINTER(C.c_int)] * 3 + [C.POINTER(C.c_float)]
    lib.ora_literal17.argtypes = [C.c_double]
    lib.ora_literal17.restype = C.c_double
    lib.ora_locate_mask_pair.argtypes = [u32, u32]
    lib.ora_locate_mask_pair.restype = u32
    lib.ora_recover_cnt_pair.argtypes = [u32, u32, C.POINTER(u32), C.POINTER(u32)]
    _lib = lib
    return lib


class OracleError(RuntimeError):
    pass


class OracleSystem:
    """The oracle restatement of the hot path holding one scene (same call sequence as DeviceSystem)."""

    def __init__(self, scene, threads: int = 0):
        self.lib = load()
        if threads:
            self.lib.ora_set_threads(threads)
        self.scene = sc = scene
        self._keep = [sc.sph_owner, sc.sph_radius, sc.sph_relpos, sc.sph_mat, sc.owner_mass, sc.owner_moi,
                      sc.anal, sc.materials, sc.mask, sc.rules]
        self.h = self.lib.ora_create(C.byref(sc.params), sc.n_spheres, ptr(sc.sph_owner), ptr(sc.sph_radius),
                                     ptr(sc.sph_relpos), ptr(sc.sph_mat), sc.n_owners, ptr(sc.owner_mass),
                                     ptr(sc.owner_moi), len(sc.anal), ptr(sc.anal), len(sc.materials),
                                     ptr(sc.materials), ptr(sc.mask), ptr(sc.rules))
        if getattr(sc, "owner_scene", None) is not None:
            self._keep += [sc.owner_scene, getattr(sc, "anal_scene", None)]
            self.lib.ora_set_scenes(self.h, int(sc.n_scenes), ptr(sc.owner_scene), ptr(getattr(sc, "anal_scene", None)))
        self.set_state(sc.state)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st:
            raise OracleError(f"oracle status {st}")

    def set_state(self, state):
        from deme_amd.abi import STATE_FIELDS
        arrs = {k: (np.ascontiguousarray(v, STATE_FIELDS[k]) if v is not None else None) for k, v in state.items()}
        self.lib.ora_set_state(self.h, C.byref(make_state_struct(arrs)))

    def get_state(self):
        arrs = empty_state(self.scene.n_owners)
        self.lib.ora_get_state(self.h, C.byref(make_state_struct(arrs)))
        return arrs

    def counts(self):
        c = Counts()
        self.lib.ora_get_counts(self.h, C.byref(c))
        return c

    def cd_bins(self): self._chk(self.lib.ora_cd_bins(self.h))
    def cd_pairs(self): self._chk(self.lib.ora_cd_pairs(self.h))
    def cd_run(self): self._chk(self.lib.ora_cd(self.h))
    def force(self): self._chk(self.lib.ora_force(self.h))
    def reduce_owner(self): self._chk(self.lib.ora_collect(self.h))
    def integrate(self): self._chk(self.lib.ora_integrate(self.h))
    def step(self, n, update_freq=1): self._chk(self.lib.ora_step(self.h, int(n), int(update_freq)))
    def set_time(self, t): self.lib.ora_set_time(self.h, float(t))

    PHASES = ("Discretize domain", "Find contact pairs", "Build history map", "Calculate contact forces",
              "Collect contact forces", "Integration")

    def phase_times(self, reset=False):
        """Wall seconds per phase since the last reset, under the reference's timer names."""
        out = np.zeros(6, np.float64)
        self.lib.ora_phase_times(self.h, ptr(out), 1 if reset else 0)
        return dict(zip(self.PHASES, out.tolist()))

    def bin_counts(self):
        n = self.scene.n_spheres
        nb = np.zeros(n, np.uint16)
        na = np.zeros(n, np.uint8)
        self.lib.ora_get_bin_counts(self.h, ptr(nb), ptr(na))
        return nb, na

    def bin_pairs(self):
        P = self.counts().nBinSphereTouches
        b = np.zeros(P, np.uint32)
        s = np.zeros(P, np.uint32)
        self.lib.ora_get_bin_pairs(self.h, ptr(b), ptr(s))
        return b, s

    def active_bins(self):
        n = self.counts().nActiveBins
        b = np.zeros(n, np.uint32)
        c = np.zeros(n, np.uint16)
        self.lib.ora_get_active_bins(self.h, ptr(b), ptr(c))
        return b, c

    def contacts(self):
        n = self.counts().nContacts
        a, b, m = (np.zeros(n, np.uint32) for _ in range(3))
        t = np.zeros(n, np.uint8)
        self.lib.ora_get_contacts(self.h, ptr(a), ptr(b), ptr(t), ptr(m))
        return a, b, t, m

    def set_contacts(self, a, b, t):
        a, b = (np.ascontiguousarray(x, np.uint32) for x in (a, b))
        t = np.ascontiguousarray(t, np.uint8)
        self._chk(self.lib.ora_set_contacts(self.h, len(a), ptr(a), ptr(b), ptr(t)))

    def brute_force_pairs(self, cap=10_000_000):
        a = np.zeros(cap, np.uint32)
        b = np.zeros(cap, np.uint32)
        n = self.lib.ora_brute_force_pairs(self.h, ptr(a), ptr(b), cap)
        return a[:n], b[:n]

    def forces(self):
        n = self.counts().nContacts
        F, T, cA, cB = (np.zeros((n, 3), np.float32) for _ in range(4))
        self.lib.ora_get_contact_forces(self.h, ptr(F), ptr(T), ptr(cA), ptr(cB))
        return F, T, cA, cB

    def forces_fp64(self):
        n = self.counts().nContacts
        F = np.zeros((n, 3), np.float64)
        T = np.zeros((n, 3), np.float64)
        self.lib.ora_force_fp64(self.h, ptr(F), ptr(T))
        return F, T

    def wildcards(self):
        n = self.counts().nContacts
        w = np.zeros((4, n), np.float32)
        self.lib.ora_get_wildcards(self.h, ptr(w))
        return w

    def set_forces(self, F, T, cA, cB):
        """Per-contact records of the current list (n x 3 float32 each), e.g. the device's."""
        a = [np.ascontiguousarray(x, np.float32) for x in (F, T, cA, cB)]
        n = self.counts().nContacts
        assert all(x.shape == (n, 3) for x in a)
        self.lib.ora_set_contact_forces(self.h, *(ptr(x) for x in a))

    def set_wildcards(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self.lib.ora_set_wildcards(self.h, ptr(w))


def figure_out_nv(bx, by, bz):
    lib = load()
    x, y, z = C.c_int(), C.c_int(), C.c_int()
    l = C.c_float()
    lib.ora_figure_out_nv(bx, by, bz, C.byref(x), C.byref(y), C.byref(z), C.byref(l))
    return (x.value, y.value, z.value), l.value


def literal17(v):
    return load().ora_literal17(v)


# ---------------------------------------------------------------------------------------------------------------------
# FP64 host adjoint (oracle/adjoint_ref.cpp): sphere owners, fixed planes, contact set frozen per window.
# ---------------------------------------------------------------------------------------------------------------------
class AdjProblem(C.Structure):
    _fields_ = [("n", C.c_uint32), ("X0", C.c_void_p), ("V0", C.c_void_p), ("W0", C.c_void_p), ("r", C.c_void_p),
                ("m", C.c_void_p), ("I", C.c_void_p), ("mat", C.c_void_p), ("nP", C.c_uint32),
                ("planeP", C.c_void_p), ("planeN", C.c_void_p), ("planeMat", C.c_void_p), ("planeMass", C.c_void_p),
                ("nMat", C.c_uint32), ("E", C.c_void_p), ("nu", C.c_void_p), ("CoR", C.c_void_p), ("mu", C.c_void_p),
                ("h", C.c_double), ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
                ("beta", C.c_double), ("nSteps", C.c_int), ("window", C.c_int)]


def positions_fp64(params, vox, lx, ly, lz):
    """Owner positions in double in the frame of the voxel grid (offset LBF), as ownerPosition computes them."""
    vox = np.asarray(vox, np.uint64)
    nx, ny = int(params.nvXp2), int(params.nvYp2)
    ix = (vox & np.uint64((1 << nx) - 1)).astype(np.float64)
    iy = ((vox >> np.uint64(nx)) & np.uint64((1 << ny) - 1)).astype(np.float64)
    iz = (vox >> np.uint64(nx + ny)).astype(np.float64)
    vs, l = params.voxelSize, params.l
    return np.stack([ix * vs + np.asarray(lx, np.float64) * l, iy * vs + np.asarray(ly, np.float64) * l,
                     iz * vs + np.asarray(lz, np.float64) * l], 1)


def _rotate(q, v):
    """Rotate rows of v by unit quaternions q (w, x, y, z) in double."""
    w, x, y, z = (q[:, k].astype(np.float64) for k in range(4))
    u = np.stack([x, y, z], 1)
    t = 2.0 * np.cross(u, v)
    return v + w[:, None] * t + np.cross(u, t)


class AdjointFP64:
    """The FP64 forward simulator and reverse sweep of a scene of single-sphere owners (each sphere its own owner,
    centred) with fixed planes, starting from `state` (default: the scene's)."""

    def __init__(self, scene, n_steps, window, state=None):
        self.lib = load()
        lib = self.lib
        lib.ora_adj_forward.argtypes = [C.POINTER(AdjProblem)] + [C.c_void_p] * 6 + [C.POINTER(C.c_uint64)]
        lib.ora_adj_forward.restype = C.c_int
        lib.ora_adj_gradient.argtypes = [C.POINTER(AdjProblem)] + [C.c_void_p] * 6
        lib.ora_adj_gradient.restype = C.c_int
        st = scene.state if state is None else state
        n = scene.n_spheres
        assert scene.n_clumps == n and np.array_equal(scene.sph_owner, np.arange(n)), "single-sphere owners only"
        p = scene.params
        X = positions_fp64(p, st["voxelID"], st["locX"], st["locY"], st["locZ"])
        q = np.stack([st["oriQw"], st["oriQx"], st["oriQy"], st["oriQz"]], 1)
        self.X0 = np.ascontiguousarray(X[:n])
        self.V0 = np.ascontiguousarray(np.stack([st["vX"], st["vY"], st["vZ"]], 1)[:n].astype(np.float64))
        wb = np.stack([st["omgX"], st["omgY"], st["omgZ"]], 1).astype(np.float64)
        self.W0 = np.ascontiguousarray(_rotate(q[:n], wb[:n]))
        self.r = np.ascontiguousarray(scene.sph_radius.astype(np.float64))
        self.m = np.ascontiguousarray(scene.owner_mass[:n].astype(np.float64))
        self.I = np.ascontiguousarray(scene.owner_moi[:n, 0].astype(np.float64))
        self.mat = np.ascontiguousarray(scene.sph_mat.astype(np.uint32))
        planes = [a for a in scene.anal if a["type"] == 0]
        assert len(planes) == len(scene.anal), "planes only"
        P, N = [], []
        for a in planes:
            o = int(a["owner"])
            qo = q[o:o + 1]
            P.append(X[o] + _rotate(qo, a["relPos"][None].astype(np.float64))[0])
            N.append(_rotate(qo, a["rot"][None].astype(np.float64))[0])
        self.planeP = np.ascontiguousarray(np.array(P, np.float64).reshape(-1, 3))
        self.planeN = np.ascontiguousarray(np.array(N, np.float64).reshape(-1, 3))
        self.planeMat = np.ascontiguousarray(np.array([a["material"] for a in planes], np.uint32))
        self.planeMass = np.ascontiguousarray(np.array([a["mass"] for a in planes], np.float64))
        mats = scene.materials
        self.mats = {k: np.ascontiguousarray(mats[k].astype(np.float64)) for k in ("E", "nu", "CoR", "mu")}
        pr = AdjProblem()
        pr.n, pr.nP, pr.nMat = n, len(planes), len(mats)
        for k in ("X0", "V0", "W0", "r", "m", "I", "mat", "planeP", "planeN", "planeMat", "planeMass"):
            setattr(pr, k, getattr(self, k).ctypes.data)
        for k in ("E", "nu", "CoR", "mu"):
            setattr(pr, k, self.mats[k].ctypes.data)
        pr.h = float(p.h)
        pr.gx, pr.gy, pr.gz = (float(g) for g in p.G)
        pr.beta = float(p.beta)
        pr.nSteps, pr.window = int(n_steps), int(window)
        self.pr, self.n, self.n_mat = pr, n, len(mats)

    def forward(self, E=None, nu=None, CoR=None, mu=None, V0=None):
        """Final positions and velocities (n x 3 each) with optionally perturbed materials or initial velocities."""
        keep = []

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64)
            keep.append(a)
            return a.ctypes.data

        pr = self.pr
        if V0 is not None:
            pr = AdjProblem.from_buffer_copy(self.pr)
            v0 = np.ascontiguousarray(V0, np.float64)
            keep.append(v0)
            pr.V0 = v0.ctypes.data
        X = np.zeros((self.n, 3))
        V = np.zeros((self.n, 3))
        nc = C.c_uint64()
        st = self.lib.ora_adj_forward(C.byref(pr), arr(E), arr(nu), arr(CoR), arr(mu), ptr(X), ptr(V), C.byref(nc))
        if st:
            raise OracleError(f"ora_adj_forward: {st}")
        return X, V

    def gradient(self, dLdX=None, dLdV=None):
        """Gradients with respect to the initial velocities (n x 3) and the materials (nMat x 4: E, nu, CoR, mu) of the
        loss whose derivatives on the final positions and velocities are dLdX, dLdV; also the final state."""
        gV = np.zeros((self.n, 3))
        gM = np.zeros((self.n_mat, 4))
        X = np.zeros((self.n, 3))
        V = np.zeros((self.n, 3))
        a = None if dLdX is None else np.ascontiguousarray(dLdX, np.float64)
        b = None if dLdV is None else np.ascontiguousarray(dLdV, np.float64)
        st = self.lib.ora_adj_gradient(C.byref(self.pr), None if a is None else ptr(a), None if b is None else ptr(b),
                                       ptr(gV), ptr(gM), ptr(X), ptr(V))
        if st:
            raise OracleError(f"ora_adj_gradient: {st}")
        return gV, gM, X, V
