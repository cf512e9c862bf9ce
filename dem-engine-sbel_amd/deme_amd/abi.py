"""ctypes declarations of the C ABI in include/deme/deme.h (libdeme.so) and a numpy-facing device system.

The product path is libdeme.so; nothing here computes physics. Loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdeme.so")

# status codes / constants (include/deme/deme_types.h)
DEME_OK = 0
DEME_ERR_INVALID_ARGUMENT = 1
DEME_ERR_HIP = 2
DEME_ERR_BIN_OVERFLOW = 3
DEME_ERR_CAPACITY = 4
DEME_ERR_NOT_READY = 5
DEME_ERR_UNSUPPORTED = 7
DEME_ERR_NO_DEVICE = 8
SPHERE_SPHERE = 1
SPHERE_PLANE = 3
NULL_MAP = 0xFFFFFFFF
NUM_MASK_ENTRIES = 256 * 257 // 2
NUM_FAMILIES = 256
# contacts per workgroup tile of the plane-owner sums (ANAL_TILE in csrc/deme_internal.h): the sums write one set of
# partials per tile
ANAL_TILE = 1024
INTEGRATOR = {"forward_euler": 0, "centered_difference": 1, "extended_taylor": 2}


class SimParams(C.Structure):
    _fields_ = [
        ("nvXp2", C.c_uint8), ("nvYp2", C.c_uint8), ("nvZp2", C.c_uint8), ("integrator", C.c_uint8),
        ("nbX", C.c_uint32), ("nbY", C.c_uint32), ("nbZ", C.c_uint32), ("pad0", C.c_uint32),
        ("l", C.c_double), ("voxelSize", C.c_double), ("l_exact", C.c_double), ("voxelSize_exact", C.c_double),
        ("binSize", C.c_double),
        ("LBF", C.c_float * 3), ("G", C.c_float * 3), ("h", C.c_float), ("beta", C.c_float),
        ("boxSize", C.c_float * 3), ("pad1", C.c_float),
    ]


class Counts(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "nSpheres", "nOwners", "nAnal", "nBinSphereTouches", "nActiveBins", "nContacts", "nPrevContacts",
        "nSphereGeoContacts", "maxSpheresInBin", "nSteps")] + [("time", C.c_double), ("nCells", C.c_uint64)]


class OwnerState(C.Structure):
    _fields_ = [("voxelID", C.c_void_p)] + [(n, C.c_void_p) for n in (
        "locX", "locY", "locZ", "oriQw", "oriQx", "oriQy", "oriQz", "vX", "vY", "vZ", "omgX", "omgY", "omgZ",
        "aX", "aY", "aZ", "alphaX", "alphaY", "alphaZ", "family")]


# numpy mirrors of the POD structs
MATERIAL_DTYPE = np.dtype([("E", "<f4"), ("nu", "<f4"), ("CoR", "<f4"), ("mu", "<f4"), ("Crr", "<f4")])
ANAL_DTYPE = np.dtype([("owner", "<u4"), ("type", "u1"), ("material", "u1"), ("normal", "u1"), ("pad", "u1"),
                       ("relPos", "<f4", (3,)), ("rot", "<f4", (3,)), ("size", "<f4", (3,)), ("mass", "<f4")])
RULE_DTYPE = np.dtype([("used", "u1"), ("linVelPrescribed", "u1"), ("rotVelPrescribed", "u1"),
                       ("linPosPrescribed", "u1"), ("rotPosPrescribed", "u1"), ("velSet", "u1"), ("posSet", "u1"),
                       ("pad", "u1"), ("vel", "<f4", (6,)), ("pos", "<f8", (3,))])
assert MATERIAL_DTYPE.itemsize == 20 and ANAL_DTYPE.itemsize == 48 and RULE_DTYPE.itemsize == 56

STATE_FIELDS = {
    "voxelID": np.uint64, "locX": np.uint16, "locY": np.uint16, "locZ": np.uint16,
    "oriQw": np.float32, "oriQx": np.float32, "oriQy": np.float32, "oriQz": np.float32,
    "vX": np.float32, "vY": np.float32, "vZ": np.float32, "omgX": np.float32, "omgY": np.float32,
    "omgZ": np.float32, "aX": np.float32, "aY": np.float32, "aZ": np.float32, "alphaX": np.float32,
    "alphaY": np.float32, "alphaZ": np.float32, "family": np.uint8,
}

EXPORTED = [
    "deme_ctx_create", "deme_ctx_destroy", "deme_last_error", "deme_device_count", "deme_get_stream", "deme_sync",
    "deme_figure_out_nv", "deme_kernel_literal", "deme_build_sim_params", "deme_encode_positions",
    "deme_decode_positions", "deme_set_params", "deme_upload_templates", "deme_upload_materials",
    "deme_upload_analytical", "deme_upload_families", "deme_upload_scenes", "deme_upload_state", "deme_download_state",
    "deme_cd_bin_count", "deme_cd_bin_fill", "deme_cd_pairs", "deme_cd_sort_remap", "deme_cd_run",
    "deme_download_bin_counts", "deme_download_bins", "deme_download_contacts", "deme_force", "deme_reduce_owner",
    "deme_integrate", "deme_step", "deme_download_forces", "deme_download_wildcards", "deme_upload_wildcards",
    "deme_get_counts", "deme_set_time", "deme_enable_kernel_timing", "deme_reset_kernel_timing",
    "deme_kernel_timing", "deme_set_beta", "deme_set_cell_scale", "deme_autotune", "deme_adjoint_forward",
    "deme_adjoint_backward", "deme_clear_contacts", "deme_set_force_model", "deme_jit_build",
    "deme_builtin_force_model", "deme_jit_render", "deme_jit_compile", "deme_set_step_graphs",
    "deme_set_keep_contact_records", "deme_set_owner_volumes", "deme_inspect", "deme_append_clumps",
    "deme_change_clump_sizes",
    "deme_clear_motion_rules", "deme_change_family", "deme_jit_render_owner", "deme_jit_render_inspect",
    "deme_prefix_scan_u32", "deme_integrate_owners", "deme_prefix_scan_u32_noalloc", "deme_set_graph_headroom",
    "deme_graph_stats", "deme_set_pair_sweep", "deme_autotune_sweep", "deme_set_skin", "deme_candidate_stats",
]

_lib = None


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def load(path: str = LIB_PATH):
    """Load libdeme.so (raises OSError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"libdeme.so not found at {path}; run __graft_entry__.build() first")
    lib = C.CDLL(path)
    vp, i32, u32, u64, f32, f64 = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_float, C.c_double
    sig = {
        "deme_ctx_create": ([i32, vp, C.POINTER(vp)], i32),
        "deme_ctx_destroy": ([vp], i32),
        "deme_last_error": ([vp], C.c_char_p),
        "deme_device_count": ([C.POINTER(i32)], i32),
        "deme_get_stream": ([vp, C.POINTER(vp)], i32),
        "deme_sync": ([vp], i32),
        "deme_figure_out_nv": ([f32, f32, f32, i32, vp, C.POINTER(f32)], i32),
        "deme_kernel_literal": ([f64], f64),
        "deme_build_sim_params": ([f32, f32, f32, i32, vp, f32, vp, f32, f64, vp, f32, f32, i32,
                                   C.POINTER(SimParams)], i32),
        "deme_encode_positions": ([C.POINTER(SimParams), u64, vp, vp, vp, vp, vp], i32),
        "deme_decode_positions": ([C.POINTER(SimParams), u64, vp, vp, vp, vp, vp], i32),
        "deme_set_params": ([vp, C.POINTER(SimParams)], i32),
        "deme_upload_templates": ([vp, u32, vp, vp, vp, vp, u32, u32, vp, vp], i32),
        "deme_upload_materials": ([vp, u32, vp], i32),
        "deme_upload_analytical": ([vp, u32, vp], i32),
        "deme_upload_families": ([vp, vp, vp], i32),
        "deme_upload_scenes": ([vp, u32, vp, vp], i32),
        "deme_upload_state": ([vp, C.POINTER(OwnerState)], i32),
        "deme_download_state": ([vp, C.POINTER(OwnerState)], i32),
        "deme_cd_bin_count": ([vp], i32), "deme_cd_bin_fill": ([vp], i32), "deme_cd_pairs": ([vp], i32),
        "deme_cd_sort_remap": ([vp], i32), "deme_cd_run": ([vp], i32),
        "deme_download_bin_counts": ([vp, vp, vp], i32),
        "deme_download_bins": ([vp, vp, vp], i32),
        "deme_download_contacts": ([vp, vp, vp, vp, vp], i32),
        "deme_prefix_scan_u32": ([vp, vp, vp, u64, C.POINTER(u64)], i32),
        "deme_prefix_scan_u32_noalloc": ([vp, vp, vp, u64, vp], i32),
        "deme_set_graph_headroom": ([vp, C.c_double, u64], i32),
        "deme_graph_stats": ([vp, vp], i32),
        "deme_set_pair_sweep": ([vp, i32, i32], i32),
        "deme_set_skin": ([vp, C.c_double], i32),
        "deme_candidate_stats": ([vp, vp, vp, vp], i32),
        "deme_autotune_sweep": ([vp, vp, i32, vp, i32, vp, i32, i32, i32, C.POINTER(C.c_double), C.POINTER(i32),
                                 C.POINTER(i32), vp], i32),
        "deme_force": ([vp], i32), "deme_reduce_owner": ([vp], i32), "deme_integrate": ([vp], i32),
        "deme_integrate_owners": ([vp], i32),
        "deme_step": ([vp, i32, i32], i32),
        "deme_download_forces": ([vp, vp, vp, vp, vp], i32),
        "deme_download_wildcards": ([vp, vp], i32), "deme_upload_wildcards": ([vp, vp], i32),
        "deme_get_counts": ([vp, C.POINTER(Counts)], i32),
        "deme_set_time": ([vp, f64], i32),
        "deme_enable_kernel_timing": ([vp, i32], i32), "deme_reset_kernel_timing": ([vp], i32),
        "deme_kernel_timing": ([vp, i32, vp, vp, vp, C.POINTER(i32)], i32),
        "deme_set_beta": ([vp, f32], i32),
        "deme_set_cell_scale": ([vp, f64], i32),
        "deme_autotune": ([vp, vp, i32, vp, i32, i32, f32, f32, C.POINTER(i32), vp], i32),
        "deme_adjoint_forward": ([vp, i32, i32], i32),
        "deme_clear_contacts": ([vp], i32),
        "deme_set_step_graphs": ([vp, i32], i32),
        "deme_set_keep_contact_records": ([vp, i32], i32),
        "deme_set_owner_volumes": ([vp, vp], i32),
        "deme_inspect": ([vp, C.c_char_p, C.c_char_p, vp], i32),
        "deme_append_clumps": ([vp, u32, vp, vp, vp, vp, u32, vp, vp, vp], i32),
        "deme_change_clump_sizes": ([vp, u32, vp, vp], i32),
        "deme_clear_motion_rules": ([vp], i32),
        "deme_change_family": ([vp, i32, i32], i32),
        "deme_jit_render_owner": ([C.c_char_p, C.c_char_p, C.c_char_p, vp, u64, C.POINTER(u64)], i32),
        "deme_jit_render_inspect": ([C.c_char_p, vp, u64, C.POINTER(u64)], i32),
        "deme_set_force_model": ([vp, C.c_char_p, vp, i32], i32),
        "deme_jit_build": ([vp, vp, vp, i32], i32),
        "deme_builtin_force_model": ([i32], C.c_char_p),
        "deme_jit_render": ([C.c_char_p, vp, i32, vp, vp, u64, C.POINTER(u64)], i32),
        "deme_jit_compile": ([C.c_char_p, vp, u64], i32),
        "deme_adjoint_backward": ([vp, vp, vp, vp, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class DemeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"DEME status {code}: {msg}")
        self.code = code


def build_sim_params(box, lbf, min_radius, h, beta=0.0, g=(0.0, 0.0, -9.81), bin_size=0.0,
                     integrator="extended_taylor", exact_dir=-1, nv=None, l=0.0) -> SimParams:
    lib = load()
    out = SimParams()
    # keep every array bound to a name until the call returns (ctypes does not own the memory)
    nv_arr = None if nv is None else np.asarray(nv, np.uint8)
    lbf_arr = np.asarray(lbf, np.float32)
    g_arr = np.asarray(g, np.float32)
    st = lib.deme_build_sim_params(C.c_float(box[0]), C.c_float(box[1]), C.c_float(box[2]), exact_dir, ptr(nv_arr),
                                   C.c_float(l), ptr(lbf_arr), C.c_float(min_radius), C.c_double(bin_size),
                                   ptr(g_arr), C.c_float(h), C.c_float(beta), INTEGRATOR[integrator], C.byref(out))
    if st != DEME_OK:
        raise DemeError(st, "deme_build_sim_params failed")
    return out


def encode_positions(params: SimParams, xyz: np.ndarray):
    lib = load()
    xyz = np.ascontiguousarray(xyz, np.float32)
    n = xyz.shape[0]
    vox = np.zeros(n, np.uint64)
    lx, ly, lz = (np.zeros(n, np.uint16) for _ in range(3))
    lib.deme_encode_positions(C.byref(params), n, ptr(xyz), ptr(vox), ptr(lx), ptr(ly), ptr(lz))
    return vox, lx, ly, lz


def decode_positions(params: SimParams, vox, lx, ly, lz) -> np.ndarray:
    lib = load()
    n = len(vox)
    out = np.zeros((n, 3), np.float32)
    v = np.ascontiguousarray(vox, np.uint64)
    x, y, z = (np.ascontiguousarray(a, np.uint16) for a in (lx, ly, lz))
    lib.deme_decode_positions(C.byref(params), n, ptr(v), ptr(x), ptr(y), ptr(z), ptr(out))
    return out


def positions_f64(params: SimParams, vox, lx, ly, lz) -> np.ndarray:
    """Owner positions in double in the voxel-grid frame (the world frame offset by LBF): voxel index times the voxel
    size plus sub-voxel count times l per axis (ownerPosition, DEMHelperKernels.cu)."""
    vox = np.asarray(vox, np.uint64)
    nx, ny = int(params.nvXp2), int(params.nvYp2)
    idx = (vox & np.uint64((1 << nx) - 1), (vox >> np.uint64(nx)) & np.uint64((1 << ny) - 1),
           vox >> np.uint64(nx + ny))
    loc = (lx, ly, lz)
    return np.stack([idx[k].astype(np.float64) * params.voxelSize + np.asarray(loc[k], np.float64) * params.l
                     for k in range(3)], 1)


def _cstr_array(items):
    if not items:
        return None
    return (C.c_char_p * len(items))(*[x.encode() for x in items])


def builtin_force_model(which=0) -> str:
    """Source of a built-in force model body: 0 frictional Hertz-Mindlin, 1 frictionless Hertz."""
    return load().deme_builtin_force_model(int(which)).decode()


def jit_render(model=None, wildcards=(), params=None) -> str:
    """The force kernel source specialised for a model (None: the built-in one) and constants (None: runtime)."""
    lib = load()
    arr = _cstr_array(list(wildcards))
    m = None if model is None else model.encode()
    need = C.c_uint64()
    args = (m, C.cast(arr, C.c_void_p) if arr else None, len(wildcards), C.byref(params) if params else None)
    st = lib.deme_jit_render(*args, None, 0, C.byref(need))
    if st:
        raise DemeError(st, "deme_jit_render failed")
    buf = C.create_string_buffer(need.value)
    lib.deme_jit_render(*args, buf, need.value, C.byref(need))
    return buf.value.decode()


def jit_render_inspect(region=None) -> str:
    """The inspection kernels' source for a region (None: the whole domain)."""
    lib = load()
    arg = None if region is None else region.encode()
    need = C.c_uint64()
    lib.deme_jit_render_inspect(arg, None, 0, C.byref(need))
    buf = C.create_string_buffer(need.value)
    lib.deme_jit_render_inspect(arg, buf, need.value, C.byref(need))
    return buf.value.decode()


def jit_render_owner(vel_strategy=None, pos_strategy=None, family_change_rules=None) -> str:
    """The owner update kernel source for the given strategies and rules."""
    lib = load()
    args = tuple(None if x is None else x.encode() for x in (vel_strategy, pos_strategy, family_change_rules))
    need = C.c_uint64()
    lib.deme_jit_render_owner(*args, None, 0, C.byref(need))
    buf = C.create_string_buffer(need.value)
    lib.deme_jit_render_owner(*args, buf, need.value, C.byref(need))
    return buf.value.decode()


def jit_compile(source: str):
    """hipRTC compile of a force kernel source for gfx950: (status, log)."""
    log = C.create_string_buffer(1 << 16)
    st = load().deme_jit_compile(source.encode(), log, len(log))
    return st, log.value.decode(errors="replace")


def make_state_struct(arrays: dict) -> OwnerState:
    s = OwnerState()
    for k in STATE_FIELDS:
        a = arrays.get(k)
        setattr(s, k, None if a is None else a.ctypes.data)
    return s


def empty_state(n: int) -> dict:
    return {k: np.zeros(n, t) for k, t in STATE_FIELDS.items()}


class DeviceSystem:
    """A scene resident on the GPU behind the C ABI. Every method maps to one ABI call."""

    def __init__(self, scene, device: int = 0):
        self.lib = load()
        self.scene = scene
        h = C.c_void_p()
        st = self.lib.deme_ctx_create(device, None, C.byref(h))
        if st != DEME_OK:
            raise DemeError(st, "deme_ctx_create failed (no HIP device?)")
        self.ctx = h
        sc = scene
        self._chk(self.lib.deme_set_params(self.ctx, C.byref(sc.params)))
        self._chk(self.lib.deme_upload_templates(
            self.ctx, sc.n_spheres, ptr(sc.sph_owner), ptr(sc.sph_radius), ptr(sc.sph_relpos), ptr(sc.sph_mat),
            sc.n_owners, sc.n_clumps, ptr(sc.owner_mass), ptr(sc.owner_moi)))
        self._chk(self.lib.deme_upload_materials(self.ctx, len(sc.materials), ptr(sc.materials)))
        self._chk(self.lib.deme_upload_analytical(self.ctx, len(sc.anal), ptr(sc.anal)))
        self._chk(self.lib.deme_upload_families(self.ctx, ptr(sc.mask), ptr(sc.rules)))
        if getattr(sc, "owner_scene", None) is not None:
            self._chk(self.lib.deme_upload_scenes(self.ctx, int(sc.n_scenes), ptr(sc.owner_scene),
                                                  ptr(getattr(sc, "anal_scene", None))))
        self.set_state(sc.state)

    def _chk(self, st):
        if st != DEME_OK:
            raise DemeError(st, self.lib.deme_last_error(self.ctx).decode())
        return st

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.deme_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, state: dict):
        arrs = {k: (np.ascontiguousarray(v, STATE_FIELDS[k]) if v is not None else None) for k, v in state.items()}
        s = make_state_struct(arrs)
        self._chk(self.lib.deme_upload_state(self.ctx, C.byref(s)))

    def set_force_model(self, model, wildcards=()):
        """Force-model body (reference model-string names) compiled by hipRTC; None returns to the built-in kernel."""
        if model is None:
            self._chk(self.lib.deme_set_force_model(self.ctx, None, None, 0))
            return
        arr = _cstr_array(wildcards)
        self._chk(self.lib.deme_set_force_model(self.ctx, model.encode(), C.cast(arr, C.c_void_p) if arr else None,
                                                len(wildcards)))

    def jit_build(self, substitutions=None):
        """Specialise the force kernel at run time with placeholder values {key: value} (none: the constants only)."""
        subs = dict(substitutions or {})
        keys, vals = _cstr_array(list(subs)), _cstr_array(list(subs.values()))
        self._chk(self.lib.deme_jit_build(self.ctx, C.cast(keys, C.c_void_p) if keys else None,
                                          C.cast(vals, C.c_void_p) if vals else None, len(subs)))

    def set_motion_rules(self, vel_strategy=None, pos_strategy=None, family_change_rules=None):
        """Owner update kernel specialised (hipRTC) with prescription strategies / family change rules."""
        subs = {k: v for k, v in (("_velPrescriptionStrategy_", vel_strategy), ("_posPrescriptionStrategy_",
                                   pos_strategy), ("_familyChangeRules_", family_change_rules)) if v is not None}
        if not subs:
            self._chk(self.lib.deme_clear_motion_rules(self.ctx))
            return
        self.jit_build(subs)

    def change_family(self, src, dst):
        self._chk(self.lib.deme_change_family(self.ctx, int(src), int(dst)))

    def set_step_graphs(self, on=True):
        """Replay whole detection intervals (and the steps outside them) from captured HIP graphs, or launch every
        kernel directly (default)."""
        self._chk(self.lib.deme_set_step_graphs(self.ctx, 1 if on else 0))

    def set_graph_headroom(self, factor, slack):
        self._chk(self.lib.deme_set_graph_headroom(self.ctx, float(factor), int(slack)))

    def set_skin(self, skin):
        """Candidate-list skin in metres (0: the cell sweep every pass; -1: automatic, 0.2 x the mean radius)."""
        self._chk(self.lib.deme_set_skin(self.ctx, float(skin)))

    def candidate_stats(self):
        """(skin in force, candidate pairs of the last build, builds so far)"""
        sk = C.c_double()
        n, b = C.c_uint64(), C.c_uint64()
        self._chk(self.lib.deme_candidate_stats(self.ctx, C.byref(sk), C.byref(n), C.byref(b)))
        return sk.value, n.value, b.value

    def set_pair_sweep(self, mode, block=0):
        """Pair sweep of the cell-list detection: mode 0 exact, 1 float pre-reject, 2 float sweep, -1 automatic;
        block: workgroup size 64, 128 or 256 (0 keeps it)."""
        self._chk(self.lib.deme_set_pair_sweep(self.ctx, int(mode), int(block)))

    def autotune_sweep(self, modes=(0, 1, 2), blocks=(64, 128, 256), trial_steps=10, update_freq=1, skins=None):
        """Times the detection configurations (skins x pair sweep modes x workgroup sizes; skins None: the skin in
        force) on the live system and keeps the fastest; returns (skin, mode, block, table of ms per step)."""
        m = np.asarray(modes, np.int32)
        b = np.asarray(blocks, np.int32)
        sk = None if skins is None else np.asarray(skins, np.float64)
        ns = 1 if sk is None else len(sk)
        table = np.zeros((ns, len(m), len(b)), np.float64)
        bs, bm, bb = C.c_double(), C.c_int32(), C.c_int32()
        self._chk(self.lib.deme_autotune_sweep(self.ctx, None if sk is None else C.c_void_p(sk.ctypes.data), ns,
                                               C.c_void_p(m.ctypes.data), len(m), C.c_void_p(b.ctypes.data), len(b),
                                               int(trial_steps), int(update_freq), C.byref(bs), C.byref(bm),
                                               C.byref(bb), C.c_void_p(table.ctypes.data)))
        if sk is None:
            return bm.value, bb.value, table[0]
        return bs.value, bm.value, bb.value, table

    def graph_stats(self):
        """(captures, replays that advanced the system, replays that fell back to direct launches)"""
        out = np.zeros(3, np.uint64)
        self._chk(self.lib.deme_graph_stats(self.ctx, C.c_void_p(out.ctypes.data)))
        return tuple(int(x) for x in out)

    def set_owner_volumes(self, volume):
        v = np.ascontiguousarray(volume, dtype=np.float32)
        self._chk(self.lib.deme_set_owner_volumes(self.ctx, ptr(v)))

    def inspect(self, quantity, region=None):
        """Value of an inspector: quantity in clump_max_z, clump_min_z, clump_max_absv, clump_volume; region a function
        body over X, Y, Z returning whether an entity counts (None: the whole domain)."""
        out = np.zeros(1, np.float32)
        reg = None if region is None else region.encode()
        self._chk(self.lib.deme_inspect(self.ctx, quantity.encode(), reg, ptr(out)))
        return np.float32(out[0])

    def append_clumps(self, sph_owner, sph_radius, sph_relpos, sph_mat, mass, moi, state):
        """UpdateClumps: append clumps (sph_owner counts from the first new clump) with their mass, MOI and state (a
        dict of owner-state arrays). The scene description follows: new spheres after the old ones, new clumps after
        the old clumps, external-object owners moved up."""
        own = np.ascontiguousarray(sph_owner, np.uint32)
        rad = np.ascontiguousarray(sph_radius, np.float32)
        rel = np.ascontiguousarray(np.asarray(sph_relpos, np.float32).reshape(-1, 3))
        mat = np.ascontiguousarray(sph_mat, np.uint8)
        mass = np.ascontiguousarray(mass, np.float32)
        moi = np.ascontiguousarray(np.asarray(moi, np.float32).reshape(-1, 3))
        n_new = len(mass)
        st = {k: np.ascontiguousarray(v, STATE_FIELDS[k]) for k, v in state.items() if v is not None}
        s = make_state_struct(st)
        self._chk(self.lib.deme_append_clumps(self.ctx, len(own), ptr(own), ptr(rad), ptr(rel), ptr(mat), n_new,
                                              ptr(mass), ptr(moi), C.byref(s)))
        sc = self.scene
        c0 = sc.n_clumps
        anal = sc.anal.copy()
        anal["owner"] += n_new
        init = {}
        for k, v in sc.state.items():
            add = st.get(k, np.zeros(n_new, STATE_FIELDS[k]) if k not in ("oriQw",) else np.ones(n_new, np.float32))
            init[k] = np.concatenate([v[:c0], add.astype(v.dtype), v[c0:]])
        self.scene = dataclasses.replace(
            sc, n_spheres=sc.n_spheres + len(own), n_owners=sc.n_owners + n_new, n_clumps=c0 + n_new,
            sph_owner=np.concatenate([sc.sph_owner, own + np.uint32(c0)]).astype(np.uint32),
            sph_radius=np.concatenate([sc.sph_radius, rad]).astype(np.float32),
            sph_relpos=np.concatenate([sc.sph_relpos, rel]).astype(np.float32),
            sph_mat=np.concatenate([sc.sph_mat, mat]).astype(sc.sph_mat.dtype),
            owner_mass=np.concatenate([sc.owner_mass[:c0], mass, sc.owner_mass[c0:]]).astype(np.float32),
            owner_moi=np.concatenate([sc.owner_moi[:c0], moi, sc.owner_moi[c0:]]).astype(np.float32),
            anal=anal, state=init)

    def change_clump_sizes(self, ids, factors):
        """ChangeClumpSizes: scale the component offsets and radii of the given clumps."""
        ids = np.ascontiguousarray(ids, np.uint32)
        f = np.ascontiguousarray(factors, np.float32)
        self._chk(self.lib.deme_change_clump_sizes(self.ctx, len(ids), ptr(ids), ptr(f)))
        sc = self.scene
        per = np.ones(sc.n_clumps, np.float32)
        per[ids] = f
        k = per[sc.sph_owner]
        self.scene = dataclasses.replace(sc, sph_radius=(sc.sph_radius * k).astype(np.float32),
                                         sph_relpos=(sc.sph_relpos * k[:, None]).astype(np.float32))

    def keep_contact_records(self, on=True):
        """Store the per-contact records (download_forces) on the last step of every step() call."""
        self._chk(self.lib.deme_set_keep_contact_records(self.ctx, 1 if on else 0))

    def clear_contacts(self):
        """Forget the contact list and histories; the step count and time restart at zero."""
        self._chk(self.lib.deme_clear_contacts(self.ctx))

    def get_state(self) -> dict:
        arrs = empty_state(self.scene.n_owners)
        s = make_state_struct(arrs)
        self._chk(self.lib.deme_download_state(self.ctx, C.byref(s)))
        return arrs

    def counts(self) -> Counts:
        c = Counts()
        self._chk(self.lib.deme_get_counts(self.ctx, C.byref(c)))
        return c

    def cd_bin_count(self): self._chk(self.lib.deme_cd_bin_count(self.ctx))
    def cd_bin_fill(self): self._chk(self.lib.deme_cd_bin_fill(self.ctx))
    def cd_pairs(self): self._chk(self.lib.deme_cd_pairs(self.ctx))
    def cd_sort_remap(self): self._chk(self.lib.deme_cd_sort_remap(self.ctx))
    def cd_run(self): self._chk(self.lib.deme_cd_run(self.ctx))
    def force(self): self._chk(self.lib.deme_force(self.ctx))
    def reduce_owner(self): self._chk(self.lib.deme_reduce_owner(self.ctx))
    def integrate(self): self._chk(self.lib.deme_integrate(self.ctx))
    def integrate_owners(self): self._chk(self.lib.deme_integrate_owners(self.ctx))
    def step(self, n, update_freq=1): self._chk(self.lib.deme_step(self.ctx, int(n), int(update_freq)))
    def sync(self): self._chk(self.lib.deme_sync(self.ctx))
    def set_time(self, t): self._chk(self.lib.deme_set_time(self.ctx, float(t)))

    def bin_counts(self):
        n = self.scene.n_spheres
        nb = np.zeros(n, np.uint16)
        na = np.zeros(n, np.uint8)
        self._chk(self.lib.deme_download_bin_counts(self.ctx, ptr(nb), ptr(na)))
        return nb, na

    def bins(self):
        p = self.scene.params
        nbins = p.nbX * p.nbY * p.nbZ
        start = np.zeros(nbins + 1, np.uint32)
        P = self.counts().nBinSphereTouches
        sph = np.zeros(P, np.uint32)
        self._chk(self.lib.deme_download_bins(self.ctx, ptr(start), ptr(sph)))
        return start, sph

    def contacts(self):
        n = self.counts().nContacts
        a, b, m = (np.zeros(n, np.uint32) for _ in range(3))
        t = np.zeros(n, np.uint8)
        self._chk(self.lib.deme_download_contacts(self.ctx, ptr(a), ptr(b), ptr(t), ptr(m)))
        return a, b, t, m

    def forces(self):
        n = self.counts().nContacts
        F, T, cA, cB = (np.zeros((n, 3), np.float32) for _ in range(4))
        self._chk(self.lib.deme_download_forces(self.ctx, ptr(F), ptr(T), ptr(cA), ptr(cB)))
        return F, T, cA, cB

    def wildcards(self):
        n = self.counts().nContacts
        w = np.zeros((4, n), np.float32)
        self._chk(self.lib.deme_download_wildcards(self.ctx, ptr(w)))
        return w

    def set_wildcards(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._chk(self.lib.deme_upload_wildcards(self.ctx, ptr(w)))

    def stream_handle(self) -> int:
        s = C.c_void_p()
        self._chk(self.lib.deme_get_stream(self.ctx, C.byref(s)))
        return s.value or 0

    def enable_timing(self, on=True): self._chk(self.lib.deme_enable_kernel_timing(self.ctx, 1 if on else 0))
    def reset_timing(self): self._chk(self.lib.deme_reset_kernel_timing(self.ctx))

    def set_beta(self, beta): self._chk(self.lib.deme_set_beta(self.ctx, float(beta)))

    def upload_materials(self, mats):
        mats = np.ascontiguousarray(mats, MATERIAL_DTYPE)
        self._chk(self.lib.deme_upload_materials(self.ctx, len(mats), ptr(mats)))

    def set_cell_scale(self, scale): self._chk(self.lib.deme_set_cell_scale(self.ctx, float(scale)))

    def autotune(self, freqs=(1, 2, 5, 10), cell_scales=(1.0, 1.25), trial_steps=20, max_vel=0.0, safety=1.5):
        """Times each (update interval, cell scale) on the live system; keeps the fastest. Returns (interval, table)."""
        f = np.asarray(freqs, np.int32)
        cs = np.asarray(cell_scales, np.float64)
        ms = np.zeros(len(f) * len(cs), np.float64)
        best = C.c_int()
        self._chk(self.lib.deme_autotune(self.ctx, ptr(f), len(f), ptr(cs), len(cs), int(trial_steps),
                                         float(max_vel), float(safety), C.byref(best), ptr(ms)))
        return best.value, ms.reshape(len(f), len(cs))

    def adjoint_forward(self, n_steps, window):
        """Forward run of n_steps with the contact list rebuilt every `window` steps, checkpointed for the adjoint."""
        self._chk(self.lib.deme_adjoint_forward(self.ctx, int(n_steps), int(window)))

    def adjoint_backward(self, dLdX=None, dLdV=None):
        """Gradients of a loss with final-state derivatives dLdX, dLdV (n_owners x 3, float) with respect to the initial
        velocities (n_owners x 3) and the materials (n_materials x 4: E, nu, CoR, mu)."""
        no, nm = self.scene.n_owners, len(self.scene.materials)

        def soa(a):
            if a is None:
                return None
            a = np.asarray(a, np.float32).reshape(-1, 3)
            full = np.zeros((no, 3), np.float32)
            full[:a.shape[0]] = a
            return np.ascontiguousarray(full.T)

        x, v = soa(dLdX), soa(dLdV)
        gV = np.zeros((3, no), np.float32)
        gM = np.zeros((nm, 4), np.float32)
        self._chk(self.lib.deme_adjoint_backward(self.ctx, ptr(x), ptr(v), ptr(gV), ptr(gM)))
        return gV.T.copy(), gM

    def kernel_timing(self) -> dict:
        mx = 64
        names = (C.c_char_p * mx)()
        ms = np.zeros(mx, np.float64)
        cnt = np.zeros(mx, np.uint64)
        n = C.c_int()
        self._chk(self.lib.deme_kernel_timing(self.ctx, mx, C.cast(names, C.c_void_p), ptr(ms), ptr(cnt),
                                              C.byref(n)))
        return {names[i].decode(): (float(ms[i]), int(cnt[i])) for i in range(min(n.value, mx))}


def device_count() -> int:
    lib = load()
    n = C.c_int(0)
    lib.deme_device_count(C.byref(n))
    return n.value
