"""The C ABI library loads without a GPU, exports every entry point include/deme/deme.h declares, and its host-side
sizing agrees with the oracle's restatement and with closed forms."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "deme", "deme.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(deme_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from deme_amd.abi import EXPORTED

    assert sorted(EXPORTED) == declared_functions()


def test_device_count_reports_without_gpu(lib):
    n = C.c_int(-1)
    st = lib.deme_device_count(C.byref(n))
    assert n.value >= 0
    assert (st == 0) == (n.value > 0)


def test_context_creation_fails_cleanly_without_device(lib):
    n = C.c_int(0)
    lib.deme_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert lib.deme_ctx_create(0, None, C.byref(h)) == 8  # DEME_ERR_NO_DEVICE
    assert not h.value


@pytest.mark.parametrize("box,bits", [((0.3, 0.3, 0.6), (21, 21, 22)), ((0.5, 0.5, 1.0), (21, 21, 22)),
                                      ((0.4, 0.4, 0.45), (21, 21, 22)), ((20, 20, 30), (21, 21, 22)),
                                      ((1.0, 1.0, 1.0), (21, 21, 22)), ((0.1, 1.0, 8.0), (20, 22, 25))])
def test_figure_out_nv_matches_oracle_and_closed_form(lib, box, bits):
    # (0.1, 1, 8): the reference's split gives 64 + (extra bits of the middle axis) = 67 bits (APIPrivate.cpp:268-272
    # add the first surplus to both larger axes); the restatements keep that, the sim-params builder refuses it.
    import pyoracle

    nv = np.zeros(3, np.uint8)
    l = C.c_float()
    assert lib.deme_figure_out_nv(C.c_float(box[0]), C.c_float(box[1]), C.c_float(box[2]), -1,
                                  nv.ctypes.data_as(C.c_void_p), C.byref(l)) == 0
    ora_bits, ora_l = pyoracle.figure_out_nv(*box)
    assert tuple(nv) == ora_bits == bits
    assert l.value == ora_l
    # closed form: l is the largest size / 2^(16 + bits) over the axes, rounded to float
    sizes = np.float32(box).astype(np.float64)
    expect = np.float32(max(sizes[k] / 2.0 ** (16 + bits[k]) for k in range(3)))
    assert np.float32(l.value) == expect


def test_sim_params_refuse_more_than_64_voxel_bits():
    from deme_amd.abi import DEME_ERR_UNSUPPORTED, DemeError, build_sim_params

    with pytest.raises(DemeError) as e:
        build_sim_params((0.1, 1.0, 8.0), (0, 0, 0), 0.01, 1e-5)
    assert e.value.code == DEME_ERR_UNSUPPORTED


def test_kernel_literal_is_17_decimal_fixed(lib):
    import pyoracle

    l = float(np.float32(0.3) / np.float32(2.0 ** 37))
    lit = lib.deme_kernel_literal(C.c_double(l))
    assert lit == float("%.17f" % l) == pyoracle.literal17(l)
    assert lit == 2.18279e-12  # "0.00000000000218279": only six significant digits survive


def test_sim_params_for_sphere_settling():
    from deme_amd import scenes

    sc = scenes.sphere_settling(n=100)
    p = sc.params
    assert (p.nvXp2, p.nvYp2, p.nvZp2) == (21, 21, 22)
    assert p.binSize == 2.0 * float(np.float32(0.005))
    world = p.voxelSize_exact * 2 ** 21
    assert p.nbX == int(world / p.binSize) + 1 == 31 and p.nbZ == 61
    assert list(p.LBF) == [np.float32(-0.15), np.float32(-0.15), np.float32(-0.3)]


def test_encode_decode_round_trip():
    from deme_amd import scenes
    from deme_amd.abi import decode_positions, encode_positions

    sc = scenes.sphere_settling(n=100)
    rng = np.random.default_rng(5)
    xyz = rng.uniform(-0.14, 0.14, size=(1000, 3)).astype(np.float32)
    v, x, y, z = encode_positions(sc.params, xyz)
    back = decode_positions(sc.params, v, x, y, z)
    # decode(encode(p)) is p up to float rounding of the world coordinate (l ~ 2e-12 m is far below a float ulp)
    assert np.abs(back.astype(np.float64) - xyz).max() <= 3e-8
    # voxel indices stay inside their bit fields for in-world positions
    p = sc.params
    vx = v & np.uint64((1 << p.nvXp2) - 1)
    vz = v >> np.uint64(p.nvXp2 + p.nvYp2)
    assert vx.max() < 2 ** p.nvXp2 and vz.max() < 2 ** p.nvZp2


def test_plane_sum_tile_matches_the_library():
    # bench.py sizes the plane-owner partials' byte model with the Python constant; it must be the kernels' tile
    from deme_amd.abi import ANAL_TILE

    text = open(os.path.join(ROOT, "dem-engine-sbel_amd", "csrc", "deme_internal.h")).read()
    items = int(re.search(r"constexpr uint32_t ANAL_ITEMS = (\d+);", text).group(1))
    assert re.search(r"constexpr uint32_t ANAL_TILE = 256 \* ANAL_ITEMS;", text)
    assert ANAL_TILE == 256 * items
