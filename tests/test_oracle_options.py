"""The oracle's two engine options that have no reference counterpart,
pinned without the oracle's own formulas.

* The second-neighbour term multiplies an evaporating atom's rate by
  f2[m], m the occupied non-wall points among its 12 fcc-shell positions
  (+-6, +-6, 0) and permutations (x, y periodic).  On a grown surface,
  the rates with the term divided by the rates recomputed without it must
  equal f2[m] with m counted here by array indexing on LattCoo.
* A frozen substrate takes the rate of evaporating atoms with z below the
  bound (and only theirs) to the never-picked 1e-300, and leaves every
  other rate as it was.
"""
import os

import numpy as np

from mulskips_amd import deck as deckmod
from mulskips_amd import models
from oracle.oracle import Oracle

BOX = (96, 96, 144)
SLAB = 48
SHELL = np.array([(a, b, 0) for a in (6, -6) for b in (6, -6)] + [(a, 0, b) for a in (6, -6) for b in (6, -6)] +
                 [(0, a, b) for a in (6, -6) for b in (6, -6)])


def model():
    """The SiC pair energies at 1100 K, where the film grows, with 0.2 eV per
    occupied shell point and cubic stacking."""
    tab = models.sic_c2_tables(temperature=1100.0, e2=0.2)
    tab["ptrans_zig"] = 1.0
    return tab


def grown(tmp_path, f2, frozen=SLAB - 12, events=20000):
    tab = model()
    d = deckmod.deck_from_tables(tab, "F", [SLAB, 120, 120], idum=4242)
    path = os.path.join(str(tmp_path), "start-sic.dat")
    deckmod.write_start_dat(d, path)
    o = Oracle(*BOX)
    o.load_deck(path)
    if frozen:
        o.set_frozen(frozen)
    o.set_second_factor(f2)
    o.refresh_all()
    o.step(events, record=False)
    return o, tab


def occupied_shell(lat, xyz):
    lz1, ly, lx = lat.shape
    m = np.zeros(len(xyz), np.int64)
    for v in SHELL:
        x = (xyz[:, 0] + v[0]) % lx
        y = (xyz[:, 1] + v[1]) % ly
        z = xyz[:, 2] + v[2]
        ok = (z >= 0) & (z < lz1)
        w = lat[np.clip(z, 0, lz1 - 1), y, x]
        m += ok & (((w >> 23) & 1) == 1) & ((w & 255) != 113)
    return m


def stale(lat, s):
    """An empty point of coordination 1: the reference does not re-evaluate
    its rate (placeholder or adsorption, Get_Prob.f:155-214) when its
    neighbour's coordination changes, so a full refresh may change it."""
    w = int(lat[s[2], s[1], s[0]])
    return ((w >> 23) & 1) == 0 and ((w >> 24) & 7) == 1


def rates(o):
    mc = o.lists()[0]
    return {tuple(int(c) for c in s): (int(e), float(p)) for s, e, p in zip(mc["xyz"], mc["event"], mc["prob"])}


def test_second_neighbour_factor(tmp_path):
    f2 = model()["second_factor"]
    o, _ = grown(tmp_path, f2)
    with_term = rates(o)
    lat = o.lattcoo()
    o.set_second_factor(np.ones(13))
    o.refresh_all()
    without = rates(o)
    o.close()
    assert with_term.keys() == without.keys()
    sites = np.array([s for s, (e, p) in with_term.items() if e == 4 and p > 1e-299])
    coor = (lat[sites[:, 2], sites[:, 1], sites[:, 0]] >> 24) & 7
    sites = sites[coor > 0]  # isolated atoms take 1e5 R, R itself changes with the term
    m = occupied_shell(lat, sites)
    ratio = np.array([with_term[tuple(s)][1] / without[tuple(s)][1] for s in sites])
    np.testing.assert_allclose(ratio, f2[m], rtol=1e-14)
    assert len(set(m.tolist())) >= 4  # several shell occupancies occur on the grown surface
    # other rates do not change (a refresh redraws an adsorption site's species,
    # Get_Prob.f:164,189, but not its total rate)
    others = [s for s, (e, p) in with_term.items() if e != 4 and not stale(lat, s)]
    assert all(with_term[s][1] == without[s][1] for s in others)


def test_frozen_substrate(tmp_path):
    o, _ = grown(tmp_path, model()["second_factor"])
    free = rates(o)
    lat = o.lattcoo()
    z_evap = [s[2] for s, (e, p) in free.items() if e == 4 and p > 1e-299]
    z0 = int(np.median(z_evap)) + 1  # raise the bound through the evaporating atoms
    o.set_frozen(z0)
    o.refresh_all()
    frozen = rates(o)
    o.close()
    assert free.keys() == frozen.keys()
    n_frozen = n_free = 0
    for s, (e, p) in free.items():
        coor = (int(lat[s[2], s[1], s[0]]) >> 24) & 7
        if e == 4 and s[2] < z0:
            assert frozen[s] == (4, 1e-300), s
            n_frozen += p > 1e-299
        elif not (e == 4 and coor == 0) and not stale(lat, s):  # isolated atoms follow R
            assert frozen[s][1] == p, s
            n_free += e == 4
    assert n_frozen > 20 and n_free > 20
