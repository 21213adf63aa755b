"""CPU: the oracle's SetCAD (InitSt 'IN', initial/SetCAD.f:84-418).

SetCAD builds the lattice from a geometry file of one id per grid point (10
a wall, 1 an atom on a diamond site), then runs the same coordination and
probability passes as the shape builders.  The regression decks do not use
it, so it is pinned through the builders they do use:

* the geometry of SetSiFINFET.f, written as a geometry file from its
  description (walls z <= 5 and at the top, wall columns beside the fin
  between Len2 and LenZ-7, walls above Len3, Si below Len1), gives through
  SetCAD the very state SetSiFINFET gives (pinned by the finfet fixture);
* the box of SetNC.f and the sphere of SetNCSp.f (Si on sublattice A, C on
  B), written the same way, give SetNC's and SetNCSp's states (pinned by
  the cube and sphere fixtures) away from SetCAD's own wall planes;
* known answers for what only SetCAD does: an isolated atom is removed, a
  vacancy inside a filled block joins the void list with coordination 4, a
  wall counts toward an atom's coordination but not an empty site's.
"""
import os

import numpy as np

from oracle.oracle import Oracle
from tests.helpers import SMALL, list_map, oracle_state, small_deck

WALL_ID, ATOM_ID = 10, 1


def dia_type(box):
    """1 (sublattice A) / 2 (B) / 0 for every grid point z < LenZ, by the
    diamond-site rule of the builders (SetCAD.f:116-126)."""
    lx, ly, lz = box
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    on = (x % 3 == 0) & (y % 3 == 0) & (z % 3 == 0)
    x1, y1, z1 = x // 3, y // 3, z // 3
    on &= (x1 % 2 == y1 % 2) & (x1 % 2 == z1 % 2)
    m = (x1 + y1 + z1) % 4
    return np.where(on & (m == 0), 1, np.where(on & (m == 3), 2, 0)).astype(np.int8)


def grids(box):
    lx, ly, lz = box
    return np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")


def state(o):
    lat, mc, bulk, void = oracle_state(o)
    return lat, list_map(mc), list_map(bulk), list_map(void), mc


def same_states(a, b, z_ok=None):
    lat_a, mc_a, bulk_a, void_a, mcr_a = a
    lat_b, mc_b, bulk_b, void_b, mcr_b = b
    if z_ok is None:
        np.testing.assert_array_equal(lat_a, lat_b)
    else:
        np.testing.assert_array_equal(lat_a[z_ok], lat_b[z_ok])
    assert mc_a == mc_b and bulk_a == bulk_b and void_a == void_b
    # the same MC particles with the same rates (the species pre-picked for a
    # deposition consumes random numbers, so Ind_Event may differ)
    ra = {tuple(s): p for s, p in zip(mcr_a["xyz"], mcr_a["prob"])}
    rb = {tuple(s): p for s, p in zip(mcr_b["xyz"], mcr_b["prob"])}
    assert ra == rb


def test_cad_of_the_finfet_geometry_is_setsifinfet(tmp_path):
    base, box, lens = SMALL["finfet"]
    path, _ = small_deck(tmp_path, base, lens)
    ref = Oracle(*box)
    ref.load_deck(path)
    lx, ly, lz = box
    ln, lww, lwwtop, lap = (v - v % 24 for v in lens)
    # SetSiFINFET.f:25-312 as a geometry: walls and the Si slab
    z, y, x = grids(box)
    xa, xb = int(np.float32(0.5) * np.float32(lx - lap)), int(np.float32(0.5) * np.float32(lx + lap))
    wall = (z <= 5) | (z >= lz - 6) | (z >= lwwtop) | ((z >= lww) & ((x < xa) | (x >= xb)))
    ids = np.where(wall, WALL_ID, np.where((z >= 6) & (z <= ln - 1), ATOM_ID, 0)).astype(np.int8)
    cad = Oracle(*box)
    cad.load_deck(path)
    cad.build_cad(ids)
    a, b = state(ref), state(cad)
    assert len(a[1]) > 100 and len(a[2]) > 1000
    same_states(a, b)
    ref.close()
    cad.close()


def test_cad_of_the_box_and_the_sphere_is_setnc(tmp_path):
    for name in ("cube", "sphere"):
        base, box, lens = SMALL[name]
        path, _ = small_deck(tmp_path, base, lens)
        ref = Oracle(*box)
        ref.load_deck(path)
        lx, ly, lz = box
        cx, cy, cz = lx // 2, ly // 2, lz // 2
        z, y, x = grids(box)
        if name == "cube":
            l1, l2, l3 = lens
            inside = ((x >= cx - l1 // 2) & (x <= cx + l1 // 2 - 1) & (y >= cy - l2 // 2) & (y <= cy + l2 // 2 - 1) &
                      (z >= cz - l3 // 2) & (z <= cz + l3 // 2 - 1))
        else:
            l1 = lens[0]
            inside = ((x >= cx - l1 // 2) & (x <= cx + l1 // 2 - 1) & (y >= cy - l1 // 2) & (y <= cy + l1 // 2 - 1) &
                      (z >= cz - l1 // 2) & (z <= cz + l1 // 2 - 1) &
                      ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= l1 * l1 // 4))
        ids = np.where(inside, ATOM_ID, 0).astype(np.int8)
        cad = Oracle(*box)
        cad.load_deck(path)
        cad.build_cad(ids, species2=dia_type(box) == 2)
        zz = np.arange(lz + 1)
        z_ok = (zz >= 4) & (zz <= lz - 5)  # SetCAD's own wall planes aside
        a, b = state(ref), state(cad)
        assert len(a[1]) > 100, name
        same_states(a, b, z_ok)
        ref.close()
        cad.close()


def test_cad_known_answers(tmp_path):
    base, box, lens = SMALL["cube"]
    path, _ = small_deck(tmp_path, base, lens)
    o = Oracle(*box)
    o.load_deck(path)
    lx, ly, lz = box
    t = dia_type(box)
    z, y, x = grids(box)
    ids = np.zeros((lz, ly, lx), np.int8)
    # a filled block with one empty diamond site inside: a vacancy
    block = (x >= 30) & (x < 90) & (y >= 30) & (y < 90) & (z >= 30) & (z < 90)
    ids[block] = ATOM_ID
    vac = (60, 60, 60)
    assert t[vac[2], vac[1], vac[0]] == 1
    ids[vac[2], vac[1], vac[0]] = 0
    # an isolated atom far from everything
    iso = (111, 111, 111)
    assert t[iso[2], iso[1], iso[0]] != 0
    ids[iso[2], iso[1], iso[0]] = ATOM_ID
    # a diamond site with walls on its four neighbours, one an atom, one empty
    ids[z == 120] = WALL_ID
    o.build_cad(ids)
    lat, mc, bulk, void, mcr = state(o)
    occ = lambda p: (lat[p[2], p[1], p[0]] >> 23) & 1  # noqa: E731
    coor = lambda p: (lat[p[2], p[1], p[0]] >> 24) & 7  # noqa: E731
    assert vac in void and coor(vac) == 4 and not occ(vac)
    assert occ(iso) == 0 and lat[iso[2], iso[1], iso[0]] == 0 and iso not in mc
    # the vacancy's four neighbours are atoms of coordination 3 (MC
    # particles); every other atom of the block's interior is bulk
    nbrs = [tuple(int(v) for v in n) for n in np.asarray(void[vac]).reshape(4, 3)]
    assert all(occ(p) and coor(p) == 3 and p in mc for p in nbrs)
    inner = [(int(a), int(b), int(c)) for c, b, a in zip(*np.nonzero(t != 0)) if 36 <= a < 84 and 36 <= b < 84
             and 36 <= c < 84]
    inner = [p for p in inner if p != vac and p not in nbrs]
    assert len(inner) > 300 and all(p in bulk for p in inner)
    # on the wall plane z = 120: an empty site just below with a wall neighbour
    # above counts only its atom neighbours; an atom there counts the wall
    below = [(xx, yy, 117) for xx in range(0, lx, 3) for yy in range(0, ly, 3) if t[117, yy, xx]]
    assert all(coor(p) == 0 and p not in mc for p in below)
    ids2 = ids.copy()
    for p in below[:4]:
        ids2[p[2], p[1], p[0]] = ATOM_ID
    o.build_cad(ids2)
    lat, mc, bulk, void, mcr = state(o)
    for p in below[:4]:
        # neighbours above are walls (two of the four bonds point up)
        assert occ(p) and coor(p) == 2 and p in mc, (p, coor(p))
    o.close()
