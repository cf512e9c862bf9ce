"""The CSV output files and clump-file readers on the CPU.

The restated writers (oracle/csv_ref.py) are pinned on a hand-built state whose files are written out below by hand:
binary-fraction positions, the identity and a half-turn quaternion, a 3-4-5 speed, contact points and normals that
follow from the geometry, family filtering and every content flag. The restated number reader is pinned on exactly
representable decimals, on its distance from correctly rounded conversion, and on the error and saturation rules of
the reference's reader. The facade's C++ readers (DEMSolver::ReadClumpXyzFromCsv / ReadClumpQuatFromCsv, run through
host/facade_demo.cpp's readcsv mode, which makes no GPU call) must then return the oracle's floats bit for bit and
fail on the same files."""
import json
import os
import subprocess

import numpy as np
import pytest

import csv_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "dem-engine-sbel_amd", "deme_amd", "facade_demo")

L = 2.0 ** -26  # sub-voxel unit; the voxel is 2^16 of them = 2^-10


def hand_state():
    # owners: 0 a ball at (0, 0.25, 2^-12); 1 a ball of family 3 (output disabled); 2 a two-sphere clump at
    # (0.5, -0.25, 0.25) turned half a turn about x; 3 the wall owning plane 0
    def vox(ix, iy, iz):
        return ix + (iy << 21) + (iz << 42)

    return {
        "nvXp2": 21, "nvYp2": 21, "voxelSize": 2.0 ** -10, "l": L, "LBF": [-0.5, -0.5, -0.5],
        "voxel": [str(vox(512, 768, 512)), str(vox(600, 600, 600)), str(vox(1024, 256, 768)), str(vox(0, 0, 0))],
        "locX": [0, 0, 0, 0], "locY": [0, 0, 0, 0], "locZ": [1 << 14, 0, 0, 0],
        "qw": [1, 1, 0, 1], "qx": [0, 0, 1, 0], "qy": [0, 0, 0, 0], "qz": [0, 0, 0, 0],
        "vx": [3, 0, 0, 0], "vy": [4, 0, 0, 0], "vz": [0, 0, -1.5, 0],
        "family": [0, 3, 2, 255],
        "clump_type": ["ball", "ball", "tri"], "n_clumps": 3,
        "sph_owner": [0, 1, 2, 2], "sph_r": [0.125, 0.125, 0.0625, 0.0625],
        "sph_rel": [0, 0, 0, 0, 0, 0, 0.25, 0.125, 0.5, -0.25, 0, 0],
        "anal_owner": [3],
        "idA": [0, 3, 0, 1], "idB": [2, 0, 3, 1], "type": [1, 3, 1, 0],
        "force": [1.5, -2, 0.25, 0, 0, 9.5, 0, 0, 0, 0, 0, 0],
        "tof": [0, 0, 1e-3, 0, 0, 0, 0, 0, 0, 0, 0, 0],
        "cpA": [0.0625, 0, 0, -0.25, 0, 0.125, 0, 0, 0, 0, 0, 0],
    }


def test_clump_file_of_a_hand_built_state():
    st = hand_state()
    assert csv_ref.clumps_csv(st, 1 | 2 | 4 | 64, no_output={3}) == (
        "X,Y,Z,Qw,Qx,Qy,Qz,clump_type,absv,v_x,v_y,v_z,family\n"
        "0,0.25,0.000244141,1,0,0,0,ball,5,3,4,0,0\n"
        "0.5,-0.25,0.25,0,1,0,0,tri,1.5,0,0,-1.5,2\n")
    assert csv_ref.clumps_csv(st, 1 | 2) == (
        "X,Y,Z,Qw,Qx,Qy,Qz,clump_type,absv\n"
        "0,0.25,0.000244141,1,0,0,0,ball,5\n"
        "0.0859375,0.0859375,0.0859375,1,0,0,0,ball,0\n"
        "0.5,-0.25,0.25,0,1,0,0,tri,1.5\n")


def test_sphere_file_of_a_hand_built_state():
    st = hand_state()
    # the half turn about x maps (x, y, z) to (x, -y, -z)
    assert csv_ref.spheres_csv(st, 0, no_output={3}) == (
        "X,Y,Z,r\n"
        "0,0.25,0.000244141,0.125\n"
        "0.75,-0.375,-0.25,0.0625\n"
        "0.25,-0.25,0.25,0.0625\n")
    assert csv_ref.spheres_csv(st, 4 | 64, no_output={0, 2}) == (
        "X,Y,Z,r,v_x,v_y,v_z,family\n"
        "0.0859375,0.0859375,0.0859375,0.125,0,0,0,3\n")


def test_contact_file_of_a_hand_built_state():
    st = hand_state()
    # contact 0: ball 0 meets the clump, point 1/16 along +x from its centre; contact 1: the clump's second sphere
    # (centre (0.25, -0.25, 0.25)) on plane 0 of owner 3, point (-0.25, 0, 0.125) in the clump frame = (0.25, -0.25,
    # 0.125) in the world, normal -z; contact 2: a pair that does not touch (zero point, zero normal); 3: no contact
    assert csv_ref.contacts_csv(st, 1 | 2 | 8 | 16) == (
        "A,B,contact_type,f_x,f_y,f_z,X,Y,Z,n_x,n_y,n_z,tof_x,tof_y,tof_z\n"
        "0,2,SS,1.5,-2,0.25,0.0625,0.25,0.000244141,1,0,0,0,0,0.001\n"
        "2,3,SA,0,0,9.5,0.25,-0.25,0.125,0,0,-1,0,0,0\n"
        "0,2,SS,0,0,0,0,0.25,0.000244141,0,0,0,0,0,0\n")
    assert csv_ref.contacts_csv(st, 0) == "A,B,contact_type\n0,2,SS\n2,3,SA\n0,2,SS\n"


@pytest.mark.parametrize("text,value", [("0", 0.0), ("7", 7.0), ("+7", 7.0), ("-3", -3.0), ("", 0.0),
                                        ("123456", 123456.0), ("1e3", 1000.0), ("25E+2", 2500.0), ("-2e1", -20.0),
                                        ("1e", 1.0), ("0.1", float(np.float32(0.1)))])
def test_number_reader_exact_cases(text, value):
    assert csv_ref.csv_float(text) == np.float32(value)


def test_number_reader_signs_and_saturation():
    assert np.signbit(csv_ref.csv_float("-0"))
    assert csv_ref.csv_float("1e99999999999") == np.inf  # the exponent saturates at INT_MAX
    z = csv_ref.csv_float("-1e-99999999999")  # ... and at INT_MIN
    assert z == 0 and np.signbit(z)


@pytest.mark.parametrize("text", ["1x", "abc", "1e5x", "1.2.3", "--1", "1e+-2", " 1"])
def test_number_reader_rejects(text):
    with pytest.raises(ValueError):
        csv_ref.csv_float(text)


def _random_numbers(rng, n):
    out = []
    for _ in range(n):
        digits = "".join(str(d) for d in rng.integers(0, 10, size=int(rng.integers(1, 8))))
        k = int(rng.integers(0, len(digits) + 1))
        s = digits[:k] + ("." + digits[k:] if k < len(digits) else "")
        if rng.random() < 0.5:
            s += ("e" if rng.random() < 0.5 else "E") + str(int(rng.integers(-8, 9)))
        if rng.random() < 0.3:
            s = "-" + s
        out.append(s)
    return out


def test_number_reader_is_within_a_few_ulps_of_correct_rounding():
    rng = np.random.default_rng(3)
    for s in _random_numbers(rng, 3000):
        v = csv_ref.csv_float(s)
        exact = float(s) if s.lstrip("-").strip(".") else 0.0
        ref = np.float32(exact)
        ulp = np.spacing(np.abs(ref)) if ref != 0 else np.float32(1e-45)
        assert abs(float(v) - exact) <= 12 * float(ulp) + 1e-38, (s, v, exact)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def test_reader_structure(tmp_path):
    text = ("\n  \t\n Y , clump_type,X ,extra,Z\r\n"  # blank lines before the header; trimmed names; any order
            "1.5,a,2,zzz,-3\r\n"
            "\n"
            "4,b,5,,6\n"
            "7,a,8,q,9")  # last line without a newline
    p = _write(tmp_path, "s.csv", text)
    got = csv_ref.read_clumps(p, "xyz")
    assert {k: [[float(x) for x in r] for r in v] for k, v in got.items()} == {
        "a": [[2.0, 1.5, -3.0], [8.0, 7.0, 9.0]], "b": [[5.0, 4.0, 6.0]]}


@pytest.mark.parametrize("text,reason", [
    ("X,Y,clump_type\n1,2,a\n", "missing"),
    ("X,Y,Z,X,clump_type\n1,2,3,4,a\n", "twice"),
    ("X,Y,Z,clump_type\n1,2,a\n", "too few"),
    ("X,Y,Z,clump_type\n1,2,3,a,5\n", "too many"),
    ("X,Y,Z,clump_type\n1,2,3e,a\n", None),  # "3e" is 3: an empty exponent reads as zero
    ("X,Y,Z,clump_type\n1,2,3q,a\n", "not a number"),
    ("\n \n", "header"),
])
def test_reader_errors(tmp_path, text, reason):
    p = _write(tmp_path, "e.csv", text)
    if reason is None:
        assert float(csv_ref.read_clumps(p, "xyz")["a"][0][2]) == 3.0
    else:
        with pytest.raises(ValueError):
            csv_ref.read_clumps(p, "xyz")


# ---- the facade's C++ readers against the oracle --------------------------------------------------------------
def _demo_read(path):
    assert os.path.exists(DEMO), f"{DEMO} is not built (run build())"
    out = subprocess.run([DEMO, "readcsv", path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout.strip().splitlines()[-1])


def _bits(rows):
    return [[int(np.float32(x).view(np.uint32)) for x in r] for r in rows]


def test_cpp_reader_matches_the_oracle_bit_for_bit(tmp_path):
    rng = np.random.default_rng(11)
    nums = _random_numbers(rng, 8 * 700)
    names = ["ball", "0001", "tri", "x y"]
    lines = ["X,Y,Z,Qw,Qx,Qy,Qz,clump_type,absv"]
    for r in range(700):
        v = nums[8 * r: 8 * r + 8]
        pad = [" ", "\t", "", ""][r % 4]
        lines.append(",".join(pad + s + pad for s in v[:7]) + "," + names[r % 4] + "," + v[7])
        if r % 50 == 0:
            lines.append("   ")
    p = _write(tmp_path, "big.csv", "\r\n".join(lines) + "\r\n")
    got = _demo_read(p)
    xyz, q = csv_ref.read_clumps(p, "xyz"), csv_ref.read_clumps(p, "quat")
    assert got["xyz"] == {k: _bits(v) for k, v in xyz.items()}
    assert got["quat"] == {k: _bits(v) for k, v in q.items()}


@pytest.mark.parametrize("text,reason", [
    ("X,Y,clump_type,Qw,Qx,Qy,Qz\n1,2,a,1,0,0,0\n", "missing"),
    ("X,Y,Z,X,clump_type,Qw,Qx,Qy,Qz\n1,2,3,4,a,1,0,0,0\n", "twice"),
    ("X,Y,Z,clump_type,Qw,Qx,Qy,Qz\n1,2,a,1,0,0,0\n", "too few"),
    ("X,Y,Z,clump_type,Qw,Qx,Qy,Qz\n1,2,3,a,1,0,0,0,9\n", "too many"),
    ("X,Y,Z,clump_type,Qw,Qx,Qy,Qz\n1,2,3q,a,1,0,0,0\n", "not a number"),
    ("\n \n", "header"),
])
def test_cpp_reader_fails_where_the_oracle_fails(tmp_path, text, reason):
    p = _write(tmp_path, "e.csv", text)
    with pytest.raises(ValueError):
        csv_ref.read_clumps(p, "xyz")
    got = _demo_read(p)
    assert reason in got.get("error", ""), got
