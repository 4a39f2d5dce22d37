"""Readers of NekCEM's case files: the .rea parameter/mesh/boundary file and
the binary .re2 mesh, into the arrays the reference's own readers fill
(rdparam / rdmesh / rdbdry, bin_rd1_mesh / bin_rd1_bc,
src/nek5_connect2.F): param[1..], xc/yc/zc (8 corners per element), and per
face a boundary code cbc and 5 parameters bc.  ``MaxwellSolver.from_rea``
hands them to nkc_create_hex.

Formats:
* .rea: "N PARAMETERS FOLLOW" then one value per line; "**MESH DATA**"
  then "NEL NDIM NELV" (NEL < 0: the mesh is in the .re2) and per element a
  header line plus 6 lines of 4 numbers (x, y, z of corners 1-4, then of
  5-8); "CURVED SIDE DATA" (must be empty here); "FLUID BOUNDARY
  CONDITIONS" with one line per element face: code (columns 2-4), element,
  face, 5 parameters.
* .re2 (v002): an 80-byte header "#v002 nelgt ndim nelgv", a float32
  6.54321 byte-order mark, per element 25 float64 (group, x[8], y[8],
  z[8]), then ncurve (float64) and its 64-byte records, then nbc (float64)
  and 64-byte records (element, face, 5 parameters as float64, code as 8
  characters).  Faces the .re2 does not list are interior ('E').
"""
from __future__ import annotations

import struct

import numpy as np


def _floats(line):
    return [float(v.replace("D", "E").replace("d", "e")) for v in line.split()]


def read_rea(path):
    """dict(param (1-based list), ndim, nel, xc, yc, zc, cbc, bc, in_re2)."""
    with open(path) as f:
        lines = f.read().splitlines()
    out = {}
    i = 0
    while "PARAMETERS FOLLOW" not in lines[i]:
        i += 1
    npar = int(lines[i].split()[0])
    out["param"] = [0.0] + [float(lines[i + 1 + k].split()[0].replace("D", "E").replace("d", "e"))
                            for k in range(npar)]
    while "MESH DATA" not in lines[i]:
        i += 1
    nel, ndim, _ = (int(v) for v in lines[i + 1].split()[:3])
    if ndim != 3:
        raise ValueError(f"{path}: {ndim}-D mesh; the HIP path is 3-D")
    out["ndim"], out["nel"], out["in_re2"] = ndim, abs(nel), nel < 0
    i += 2
    if nel > 0:
        xc, yc, zc = (np.zeros((nel, 8)) for _ in range(3))
        for e in range(nel):
            rows = [_floats(lines[i + 1 + r]) for r in range(6)]
            xc[e, :4], yc[e, :4], zc[e, :4] = rows[0], rows[1], rows[2]
            xc[e, 4:], yc[e, 4:], zc[e, 4:] = rows[3], rows[4], rows[5]
            i += 7
        out["xc"], out["yc"], out["zc"] = xc, yc, zc
        while "CURVED SIDE DATA" not in lines[i]:
            i += 1
        ncurve = int(lines[i + 1].split()[0])
        if ncurve:
            raise ValueError(f"{path}: {ncurve} curved sides; the HIP path takes straight-sided boxes")
        while "FLUID" not in lines[i] or "BOUNDARY" not in lines[i]:
            i += 1
        i += 1
        cbc = np.full((nel, 6), "E  ", dtype="<U3")
        bc = np.zeros((nel, 6, 5))
        for _ in range(6 * nel):
            line = lines[i]
            code = line[1:4]
            rest = _floats(line[4:])
            e, f = int(rest[0]), int(rest[1])
            cbc[e - 1, f - 1] = code
            bc[e - 1, f - 1] = rest[2:7]
            i += 1
        out["cbc"], out["bc"] = cbc, bc
    return out


def read_re2(path):
    """dict(nel, xc, yc, zc, cbc, bc) of a v002 .re2 file."""
    b = open(path, "rb").read()
    hdr = b[:80].decode("ascii", "replace")
    if not hdr.startswith("#v002"):
        raise ValueError(f"{path}: header {hdr[:5]!r}; only the v002 (float64) .re2 is read")
    nel, ndim = int(hdr[5:14]), int(hdr[14:17])
    if ndim != 3:
        raise ValueError(f"{path}: {ndim}-D mesh; the HIP path is 3-D")
    (mark,) = struct.unpack("<f", b[80:84])
    end = "<" if abs(mark - 6.54321) < 1e-5 else ">"
    off = 84
    m = np.frombuffer(b, dtype=end + "f8", count=25 * nel, offset=off).reshape(nel, 25)
    off += 200 * nel
    xc, yc, zc = m[:, 1:9].copy(), m[:, 9:17].copy(), m[:, 17:25].copy()
    (ncurve,) = struct.unpack(end + "d", b[off:off + 8])
    off += 8
    if int(ncurve):
        raise ValueError(f"{path}: {int(ncurve)} curved sides; the HIP path takes straight-sided boxes")
    (nbc,) = struct.unpack(end + "d", b[off:off + 8])
    off += 8
    cbc = np.full((nel, 6), "E  ", dtype="<U3")
    bc = np.zeros((nel, 6, 5))
    for _ in range(int(nbc)):
        v = struct.unpack(end + "7d", b[off:off + 56])
        code = b[off + 56:off + 64].decode("ascii", "replace")[:3]
        e, f = int(v[0]), int(v[1])
        cbc[e - 1, f - 1] = code
        bc[e - 1, f - 1] = v[2:7]
        off += 64
    return dict(nel=nel, xc=xc, yc=yc, zc=zc, cbc=cbc, bc=bc)


def read_case(rea_path, re2_path=None):
    """The mesh of a case: from the .rea, or from the .re2 beside it when the
    .rea says so (NEL < 0); plus the .rea parameters."""
    r = read_rea(rea_path)
    if r["in_re2"]:
        re2 = re2_path or rea_path[:-4] + ".re2"
        m = read_re2(re2)
        if m["nel"] != r["nel"]:
            raise ValueError(f"{re2}: {m['nel']} elements, the .rea says {r['nel']}")
        r.update(m)
    return r
