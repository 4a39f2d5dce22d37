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
  5-8); "CURVED SIDE DATA": a count, then per curved edge or face
  "IEDGE IEL CURVE(1..5) CCURVE" (edge 1..8 or face 1..6, 1-based element,
  5 parameters, a 1-character type such as 'C', a circular arc of radius
  CURVE(1)); "FLUID BOUNDARY
  CONDITIONS" with one line per element face: code (columns 2-4), element,
  face, 5 parameters.
* .box (genbox2 / getbox / makebox, src/nek5_genbox.F:2-555; used when
  the .rea's NDIM is negative): lines starting with '#' are comments; a
  dummy line, NDIM, the number of fields, then a box: a line starting
  'Box', "nelx nely nelz" (negative: x0 x1 ratio on the next line, a
  geometric spacing; positive: the nel+1 vertex coordinates), the same for
  y and z, and a line of six comma-separated 3-character codes for the x-,
  x+, y-, y+, z-, z+ walls.  Elements are numbered x fastest; interior
  faces are 'E' with their neighbour, periodic walls 'P' with theirs.
* .re2 (v002): an 80-byte header "#v002 nelgt ndim nelgv", a float32
  6.54321 byte-order mark, per element 25 float64 (group, x[8], y[8],
  z[8]), then ncurve (float64) and its 64-byte records (element, side,
  5 parameters as float64, type as 8 characters), then nbc (float64)
  and 64-byte records (element, face, 5 parameters as float64, code as 8
  characters).  Faces the .re2 does not list are interior ('E').
"""
from __future__ import annotations

import struct

import numpy as np


def _floats(line):
    return [float(v.replace("D", "E").replace("d", "e")) for v in line.split()]


def read_rea(path):
    """dict(param (1-based list), ndim, nel, xc, yc, zc, cbc, bc, curves,
    in_re2, in_box); curves is a list of (side 1..8, element 0-based,
    parameters (5,), type)."""
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
    out["in_box"] = ndim < 0          # ifgtp: the mesh is the .box beside it
    if out["in_box"]:
        out["ndim"], out["nel"], out["in_re2"] = None, None, False
        return out
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
        curves = []
        for k in range(ncurve):
            # rdcurve (src/nek5_connect2.F): fixed columns for nel < 1000
            line = lines[i + 2 + k]
            side, iel = int(line[0:3]), int(line[3:6])
            curves.append((side, iel - 1, np.array(_floats(line[6:76])[:5]), line[76:].strip()[:1] or line.split()[-1]))
        out["curves"] = curves
        i += 2 + ncurve
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
    curves = []
    for _ in range(int(ncurve)):
        v = struct.unpack(end + "7d", b[off:off + 56])
        kind = b[off + 56:off + 64].decode("ascii", "replace").strip()[:1]
        curves.append((int(v[1]), int(v[0]) - 1, np.array(v[2:7]), kind))
        off += 64
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
    return dict(nel=nel, xc=xc, yc=yc, zc=zc, cbc=cbc, bc=bc, curves=curves)


def _box_vertices(tokens, n):
    """One axis of getbox: n < 0 -> x0 x1 ratio (geometric spacing scaled
    onto [x0, x1]); n > 0 -> the n + 1 vertices."""
    if n < 0:
        n = -n
        x0, x1, ratio = (float(tokens.pop(0)) for _ in range(3))
        if ratio <= 0:
            ratio = 1.0
        x = np.zeros(n + 1)
        dx = (x1 - x0) / n
        for k in range(1, n + 1):
            x[k] = x[k - 1] + dx
            dx = dx * ratio
        return x0 + (x1 - x0) / x[n] * x
    return np.array([float(tokens.pop(0)) for _ in range(n + 1)])


def read_box(path):
    """dict(ndim, nel, nelxyz, xc, yc, zc, cbc, bc) of a one-box .box file,
    elements and faces as makebox lays them out."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip() and not ln.lstrip().startswith("#")]
    ndim = abs(int(lines[1].split()[0]))
    if ndim != 3:
        raise ValueError(f"{path}: {ndim}-D box; the HIP path is 3-D")
    k = 3
    if not lines[k].lstrip()[:1] in ("B", "b"):
        raise ValueError(f"{path}: only a plain 'Box' is read (got {lines[k]!r})")
    k += 1
    nels = [int(v) for v in lines[k].split()[:3]]
    k += 1
    tokens = []
    verts = []
    for n in nels:
        need = 3 if n < 0 else abs(n) + 1
        while len(tokens) < need:
            tokens += [t.replace("D", "E").replace("d", "e") for t in lines[k].replace(",", " ").split()
                       if not t.startswith("(")][:need - len(tokens)]
            k += 1
        verts.append(_box_vertices(tokens, n))
        tokens = []
    codes = [c.strip()[:3] for c in lines[k].split(",")[:6]]
    codes = [(c + "  ")[:3] for c in codes]
    nx, ny, nz = (abs(n) for n in nels)
    nelt = nx * ny * nz
    e = np.arange(nelt)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    X, Y, Z = verts
    xc = np.stack([X[ex], X[ex + 1], X[ex + 1], X[ex]] * 2, axis=1)
    yc = np.stack([Y[ey], Y[ey], Y[ey + 1], Y[ey + 1]] * 2, axis=1)
    zc = np.stack([Z[ez]] * 4 + [Z[ez + 1]] * 4, axis=1)
    cbc = np.full((nelt, 6), "E  ", dtype="<U3")
    bc = np.zeros((nelt, 6, 5))
    gid = lambda a, b, c: 1 + a + nx * (b + ny * c)
    # (wall code index, reference face, axis, side, neighbour's face)
    for wall, f, ax, side, nf in ((0, 4, 0, 0, 2), (1, 2, 0, 1, 4), (2, 1, 1, 0, 3), (3, 3, 1, 1, 1),
                                  (4, 5, 2, 0, 6), (5, 6, 2, 1, 5)):
        idx = [ex, ey, ez]
        nn = (nx, ny, nz)[ax]
        nb = [ex.copy(), ey.copy(), ez.copy()]
        nb[ax] = (idx[ax] + (1 if side else -1)) % nn
        on_wall = idx[ax] == (nn - 1 if side else 0)
        cbc[:, f - 1] = np.where(on_wall, codes[wall], "E  ")
        link = (~on_wall) | (codes[wall].strip() == "P")
        bc[:, f - 1, 0] = np.where(link, gid(*nb), 0.0)
        bc[:, f - 1, 1] = np.where(link, nf, 0.0)
    return dict(ndim=3, nel=nelt, nelxyz=(nx, ny, nz), xc=xc, yc=yc, zc=zc, cbc=cbc, bc=bc)


def read_case(rea_path, re2_path=None):
    """The mesh of a case: from the .rea, from the .re2 beside it when the
    .rea says so (NEL < 0), or from the .box beside it (NDIM < 0); plus the
    .rea parameters."""
    r = read_rea(rea_path)
    if r["in_box"]:
        m = read_box(rea_path[:-4] + ".box")
        r.update(m)
        return r
    if r["in_re2"]:
        re2 = re2_path or rea_path[:-4] + ".re2"
        m = read_re2(re2)
        if m["nel"] != r["nel"]:
            raise ValueError(f"{re2}: {m['nel']} elements, the .rea says {r['nel']}")
        r.update(m)
    return r
