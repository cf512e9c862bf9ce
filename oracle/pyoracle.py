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
    lib.ora_force_fp64.argtypes = [vp, vp, vp]
    lib.ora_get_time.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    lib.ora_set_time.argtypes = [vp, C.c_double]
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
