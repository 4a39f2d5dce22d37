"""Readers of the reference's case files (CPU): the .rea of 3dboxpec (mesh
and boundary conditions in the text file) and the .rea + binary .re2 of
3dboxper (mesh and periodic links in the .re2)."""
import os

import numpy as np

from nekcem_amd import rea

CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")


def _check_boxes(r):
    xc, yc, zc = r["xc"], r["yc"], r["zc"]
    # corners 1-4 on z-, counter-clockwise from (x-, y-); 5-8 above them
    for c, lo_idx in ((xc, [0, 3, 4, 7]), (yc, [0, 1, 4, 5]), (zc, [0, 1, 2, 3])):
        hi_idx = [q for q in range(8) if q not in lo_idx]
        assert np.allclose(c[:, lo_idx], c[:, lo_idx[:1]])
        assert np.allclose(c[:, hi_idx], c[:, hi_idx[:1]])
        assert np.all(c[:, hi_idx[0]] > c[:, lo_idx[0]])


def test_rea_3dboxpec():
    r = rea.read_case(os.path.join(CASES, "3dboxpec.rea"))
    assert r["nel"] == 27 and not r["in_re2"]
    assert r["param"][11] == 1000 and r["param"][12] == -0.005
    _check_boxes(r)
    assert set(np.unique(r["cbc"])) == {"E  ", "PEC"}
    # every interior link is mutual and to the opposite face
    opp = {1: 3, 3: 1, 2: 4, 4: 2, 5: 6, 6: 5}
    for e in range(27):
        for f in range(6):
            if r["cbc"][e, f] == "E  ":
                ne, nf = int(r["bc"][e, f, 0]), int(r["bc"][e, f, 1])
                assert nf == opp[f + 1]
                assert int(r["bc"][ne - 1, nf - 1, 0]) == e + 1
    # a 3x3x3 box: 9 PEC faces on each of its 6 sides
    assert np.sum(r["cbc"] == "PEC") == 54


def test_rea_re2_3dboxper():
    r = rea.read_case(os.path.join(CASES, "3dboxper.rea"))
    assert r["nel"] == 128 and r["in_re2"]
    assert r["param"][11] == 50 and r["param"][12] == -0.0002
    _check_boxes(r)
    assert (r["xc"].min(), r["xc"].max(), r["zc"].min(), r["zc"].max()) == (-1.0, 1.0, -1.0, 1.0)
    per = r["cbc"] == "P  "
    # 4x4x8: 2 (4*8 + 4*8 + 4*4) periodic faces; the others are interior
    assert per.sum() == 160 and np.all(r["cbc"][~per] == "E  ")
    for e, f in zip(*np.nonzero(per)):
        ne, nf = int(r["bc"][e, f, 0]), int(r["bc"][e, f, 1])
        assert r["cbc"][ne - 1, nf - 1] == "P  " and int(r["bc"][ne - 1, nf - 1, 0]) == e + 1
