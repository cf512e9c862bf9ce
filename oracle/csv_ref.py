"""CPU restatement of the reference's CSV output files and clump-file readers -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of the DEMSolver facade's WriteSphereFile / WriteClumpFile /
WriteContactFile and ReadClumpXyzFromCsv / ReadClumpQuatFromCsv. Every number is computed in float32 with the
operation order of the reference's host code and printed as std::ostream prints a float by default (printf "%g").

  spheres_csv   dT.cpp:896-986      writeSpheresAsCsv
  clumps_csv    dT.cpp:990-1052     writeClumpsAsCsv
  contacts_csv  dT.cpp:1054-1180    writeContactsAsCsv (orientation of owner A, see contacts_csv)
  csv_float     csv.hpp:937-997     parse_float<float> of the vendored fast-cpp-csv-parser
  read_clumps   API.h:326-371       ReadClumpXyzFromCsv / ReadClumpQuatFromCsv over io::CSVReader (csv.hpp:1010-1160)

Pins (tests/test_csv_io.py): hand-built states whose files are known exactly (binary-fraction positions, the identity
and a half-turn quaternion, a 3-4-5 speed, a known contact point and normal, family filtering, every content flag);
csv_float against exactly representable decimals, against correctly rounded conversion within a few float ulps for
random decimals, and against the reference's error and saturation rules.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32

ABSV, VEL, FAMILY = 2, 4, 64  # OUTPUT_CONTENT (Defines.h:118-131)
C_FORCE, C_POINT, C_NORMAL, C_TOF = 1, 2, 8, 16  # CNT_OUTPUT_CONTENT (Defines.h:135-143)
SS, SM, SPLANE, SPLATE = 1, 2, 3, 4  # contact types


def g(v) -> str:
    """A float32 as std::ostream << float writes it by default."""
    return "%g" % float(v)


def com(st, o):
    """Owner centre of mass: hostVoxelIDToPosition<float> (HostSideHelpers.hpp:307-324) + the domain corner."""
    vid = int(st["voxel"][o])
    nx, ny = st["nvXp2"], st["nvYp2"]
    vx, vy, vz = vid & ((1 << nx) - 1), (vid >> nx) & ((1 << ny) - 1), vid >> (nx + ny)
    vs, l = f32(st["voxelSize"]), f32(st["l"])
    X = f32(vx) * vs + f32(st["locX"][o]) * l
    Y = f32(vy) * vs + f32(st["locY"][o]) * l
    Z = f32(vz) * vs + f32(st["locZ"][o]) * l
    L = st["LBF"]
    return X + f32(L[0]), Y + f32(L[1]), Z + f32(L[2])


def rotate(v, q):
    """hostApplyOriQToVector3<float, float> (HostSideHelpers.hpp:383-394)."""
    x, y, z = (f32(c) for c in v)
    w, a, b, c = (f32(t) for t in q)
    one, two = f32(1), f32(2)
    X = (two * (w * w + a * a) - one) * x + (two * (a * b - w * c)) * y + (two * (a * c + w * b)) * z
    Y = (two * (a * b + w * c)) * x + (two * (w * w + b * b) - one) * y + (two * (b * c - w * a)) * z
    Z = (two * (a * c - w * b)) * x + (two * (b * c + w * a)) * y + (two * (w * w + c * c) - one) * z
    return X, Y, Z


def _quat(st, o):
    return st["qw"][o], st["qx"][o], st["qy"][o], st["qz"][o]


def _vel_cols(st, o, content):
    vx, vy, vz = f32(st["vx"][o]), f32(st["vy"][o]), f32(st["vz"][o])
    out = ""
    if content & ABSV:
        out += "," + g(np.sqrt(vx * vx + vy * vy + vz * vz))
    if content & VEL:
        out += "," + g(vx) + "," + g(vy) + "," + g(vz)
    if content & FAMILY:
        out += "," + str(int(st["family"][o]))
    return out


def _extra_header(content):
    h = ""
    if content & ABSV:
        h += ",absv"
    if content & VEL:
        h += ",v_x,v_y,v_z"
    if content & FAMILY:
        h += ",family"
    return h


def spheres_csv(st, content, no_output=()):
    lines = ["X,Y,Z,r" + _extra_header(content)]
    rel = st["sph_rel"]
    for i, o in enumerate(st["sph_owner"]):
        if int(st["family"][o]) in no_output:
            continue
        c = com(st, o)
        d = rotate((rel[3 * i], rel[3 * i + 1], rel[3 * i + 2]), _quat(st, o))
        p = [c[k] + d[k] for k in range(3)]
        lines.append(",".join(g(v) for v in p) + "," + g(f32(st["sph_r"][i])) + _vel_cols(st, o, content))
    return "\n".join(lines) + "\n"


def clumps_csv(st, content, no_output=()):
    lines = ["X,Y,Z,Qw,Qx,Qy,Qz,clump_type" + _extra_header(content)]
    for o in range(st["n_clumps"]):
        if int(st["family"][o]) in no_output:
            continue
        c = com(st, o)
        q = [f32(v) for v in _quat(st, o)]
        lines.append(",".join(g(v) for v in c) + "," + ",".join(g(v) for v in q) + "," + st["clump_type"][o] +
                     _vel_cols(st, o, content))
    return "\n".join(lines) + "\n"


def contacts_csv(st, content):
    """The contact file. The point of contact i is kept in owner A's frame and placed with owner A's orientation
    (the reference reads the orientation at the contact's own index there, dT.cpp:1131-1134, which is owner A's only
    when the two numbers coincide). A pair of the list that does not touch has a zero normal here."""
    h = "A,B,contact_type"
    if content & C_FORCE:
        h += ",f_x,f_y,f_z"
    if content & C_POINT:
        h += ",X,Y,Z"
    if content & C_NORMAL:
        h += ",n_x,n_y,n_z"
    if content & C_TOF:
        h += ",tof_x,tof_y,tof_z"
    lines = [h]
    F, T, P, rel = st["force"], st["tof"], st["cpA"], st["sph_rel"]
    for i, t in enumerate(st["type"]):
        if t == 0:
            continue
        a, b = st["idA"][i], st["idB"][i]
        oA = st["sph_owner"][a]
        if t == SS:
            oB, name = st["sph_owner"][b], "SS"
        elif t == SM:
            raise ValueError("sphere-mesh contacts do not occur")
        else:
            oB, name = st["anal_owner"][b], "SA"
        row = f"{oA},{oB},{name}"
        if content & C_FORCE:
            row += "," + ",".join(g(f32(F[3 * i + k])) for k in range(3))
        if content & (C_POINT | C_NORMAL):
            q = _quat(st, oA)
            c = com(st, oA)
            r = rotate((P[3 * i], P[3 * i + 1], P[3 * i + 2]), q)
            p = [r[k] + c[k] for k in range(3)]
            if content & C_POINT:
                row += "," + ",".join(g(v) for v in p)
            if content & C_NORMAL:
                d = rotate((rel[3 * a], rel[3 * a + 1], rel[3 * a + 2]), q)
                n = [p[k] - (c[k] + d[k]) for k in range(3)]
                dd = n[0] * n[0] + n[1] * n[1] + n[2] * n[2]
                if dd > 0:
                    inv = f32(1) / np.sqrt(dd)
                    n = [v * inv for v in n]
                else:
                    n = [f32(0)] * 3
                row += "," + ",".join(g(v) for v in n)
        if content & C_TOF:
            row += "," + ",".join(g(f32(T[3 * i + k])) for k in range(3))
        lines.append(row)
    return "\n".join(lines) + "\n"


# ---- reading ----------------------------------------------------------------------------------------------------
INT_MIN, INT_MAX = -(2 ** 31), 2 ** 31 - 1


def csv_float(text: str) -> np.float32:
    """parse_float<float> (csv.hpp:937-997): digit accumulation in float32; raises ValueError where the reference
    throws io::error::no_digit."""
    s, i = text, 0
    neg = False
    if i < len(s) and s[i] == "-":
        neg, i = True, i + 1
    elif i < len(s) and s[i] == "+":
        i += 1
    x = f32(0)
    while i < len(s) and "0" <= s[i] <= "9":
        x = x * f32(10)
        x = x + f32(ord(s[i]) - 48)
        i += 1
    if i < len(s) and s[i] in ".,":
        i += 1
        p = f32(1)
        while i < len(s) and "0" <= s[i] <= "9":
            p = p / f32(10)
            x = x + f32(ord(s[i]) - 48) * p
            i += 1
    if i < len(s) and s[i] in "eE":
        i += 1
        e = 0
        if i < len(s) and s[i] == "-":
            i += 1
            while i < len(s):
                if not "0" <= s[i] <= "9":
                    raise ValueError(text)
                d = ord(s[i]) - 48
                if e < -((-(INT_MIN + d)) // 10):  # C++ integer division truncates toward zero
                    e = INT_MIN
                    break
                e = 10 * e - d
                i += 1
        else:
            if i < len(s) and s[i] == "+":
                i += 1
            while i < len(s):
                if not "0" <= s[i] <= "9":
                    raise ValueError(text)
                d = ord(s[i]) - 48
                if e > (INT_MAX - d) // 10:
                    e = INT_MAX
                    break
                e = 10 * e + d
                i += 1
        if e != 0:
            base = f32(0.1) if e < 0 else f32(10)
            m = -e if e < 0 else e
            with np.errstate(over="ignore", under="ignore"):
                while m != 1:
                    if m % 2 == 0:
                        base = base * base
                        m //= 2
                    else:
                        x = x * base
                        m -= 1
                x = x * base
    elif i < len(s):
        raise ValueError(text)
    return -x if neg else x


def _fields(line):
    return [f.strip(" \t") for f in line.split(",")]


def read_clumps(path, kind, clump_header="clump_type", names=None):
    """ReadClumpXyzFromCsv (kind "xyz") or ReadClumpQuatFromCsv (kind "quat"): {type name: [[float32, ...], ...]}.
    Raises ValueError where the reference's reader throws."""
    if names is None:
        names = ["X", "Y", "Z"] if kind == "xyz" else ["Qw", "Qx", "Qy", "Qz"]
    with open(path, "rb") as fh:
        raw = fh.read().decode()
    lines = raw.split("\n")
    if lines and lines[-1] == "":
        lines.pop()  # the text after the last newline is a line only when it is not empty
    lines = [ln[:-1] if ln.endswith("\r") else ln for ln in lines]
    lines = [ln for ln in lines if ln.strip(" \t") != ""]
    if not lines:
        raise ValueError("header missing")
    head = _fields(lines[0])
    want = [clump_header] + list(names)
    col = [-1] * len(want)
    for i, f in enumerate(head):
        for k, w in enumerate(want):
            if f == w:
                if col[k] != -1:
                    raise ValueError(f"duplicated column {w}")
                col[k] = i
                break
    for k, w in enumerate(want):
        if col[k] == -1:
            raise ValueError(f"missing column {w}")
    out: dict[str, list] = {}
    for ln in lines[1:]:
        f = _fields(ln)
        if len(f) != len(head):
            raise ValueError("too few columns" if len(f) < len(head) else "too many columns")
        vals = [csv_float(f[c]) for c in col[1:]]
        out.setdefault(f[col[0]], []).append(vals)
    return out
