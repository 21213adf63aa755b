"""CPU: the oracle's restatement of the laser-annealing build's rate form.

mulskips-source-LaserAnnealing/mcevent/Get_Prob.f:140-153 gives each species
its own P0, Tm, sigma and Tflex, and scales solidification by the liquid
fraction CountLiquid(k) / SUM(CountLiquid), CountLiquid = InitCountCrystal -
CountCrystal refreshed at the top of each iteration (its main.f:637); its
main.f:676-681 swaps the deposited SiGe species whose liquid is used up.  No
reference fixture covers that build (its LA fixture is stale), so the
restatement is pinned by cases the formula reduces to:

* equal per-species constants and equal species energies: melting rates
  equal the main build's LA form (pinned by the closed form in
  tests/test_gpu_la.py), and since the liquid fractions sum to one, a
  site's solidification rate is half the main form's two-species sum;
* scaling one species' P0 scales that species' melting and solidification
  rates by the same factor and leaves the other's;
* a species with no liquid left never solidifies;
* the SiGe swap (applied here to the SiC deck's two species) keeps
  CountCrystal(1) from passing InitCountCrystal(1), where the same run
  without it overshoots through species drawn before the liquid ran out.
"""
import numpy as np
import pytest

from mulskips_amd import deck as deckmod
from oracle.oracle import Oracle
from tests.helpers import SMALL, small_deck
from tests.test_oracle_la_coupling import to_socket

LA = dict(tm=1687.0, p0=1e13, sigma=50.0, tflex=1650.0)
E_PAIR = ((0.6, 0.8), (0.6, 0.8))  # eV per bond of an atom of species i to species 1 / 2


def la_energy_deck(tmp_path, key, e_pair=E_PAIR, e_dep=(0.55, 0.55)):
    """A regression deck with energies in its tables: melting the sum over
    bonds of e_pair[species][neighbour species], solidification
    c * e_dep[species]."""
    base, box, lens = SMALL[key]
    _, d = small_deck(tmp_path, base, lens)
    d.evap = [(sp, digits, sum(int(ch) * e_pair[sp - 1][i] for i, ch in enumerate(digits)))
              for sp, digits, _ in d.evap]
    d.depo = [(sp, c, c * e_dep[sp - 1]) for sp, c, _ in d.depo]
    path = str(tmp_path / f"start-la-{key}.dat")
    deckmod.write_start_dat(d, path)
    return path, d, box


def field(box):
    lx, ly, lz = box
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    return to_socket((1560 + 2 * z + (x % 24) + 3 * (y % 12)).astype(np.int32))


def species_counts(o):
    lat = o.lattcoo()
    occ = ((lat >> 23) & 1) == 1
    sic = (lat >> 27) & 7
    wall = (lat & 255) == 113
    return [int((occ & ~wall & (sic == k)).sum()) for k in (1, 2)]


def rates(o):
    mc, _, _ = o.lists()
    return mc["xyz"], np.array([o.pure_rate(s)[1] for s in mc["xyz"]]), mc["event"]


def loaded(tmp_path, key="surface", **kw):
    path, d, box = la_energy_deck(tmp_path, key, **kw)
    o = Oracle(*box)
    o.load_deck(path)
    o.set_temperature_field(field(box), "socket")
    return o, box


def set_variant(o, p0=(1e13, 1e13), tm=(1687.0, 1687.0), sigma=(50.0, 50.0), tflex=(1650.0, 1650.0),
                liquid=(4000, 4000), swap=False):
    cc = species_counts(o)
    o.set_la_variant(p0, tm, sigma, tflex, [cc[0] + liquid[0], cc[1] + liquid[1]], cc, swap)


def test_reduces_to_the_main_form(tmp_path):
    o, box = loaded(tmp_path, e_pair=((0.7, 0.7), (0.7, 0.7)), e_dep=(0.55, 0.55))
    o.set_la(**LA)
    _, r_main, ev = rates(o)
    set_variant(o, liquid=(1234, 5678))
    _, r_var, _ = rates(o)
    melt = ev == 4
    assert melt.sum() > 100 and (~melt).sum() > 100
    np.testing.assert_allclose(r_var[melt], r_main[melt], rtol=1e-13)
    # the fractions sum to one: one species' worth of the main form's
    # two-species sum
    np.testing.assert_allclose(r_var[~melt], r_main[~melt] / 2, rtol=1e-13)


def test_p0_of_one_species_scales_its_rates(tmp_path):
    o, box = loaded(tmp_path)
    set_variant(o, liquid=(3000, 1000))
    o.step(3000, record=False)  # both species at the surface
    set_variant(o, liquid=(3000, 1000))
    xyz, r1, ev = rates(o)
    set_variant(o, p0=(1e13, 3e13), liquid=(3000, 1000))
    _, r2, _ = rates(o)
    lat = o.lattcoo()
    sp = np.array([(lat[z, y, x] >> 27) & 7 for x, y, z in xyz])
    melt = ev == 4
    np.testing.assert_allclose(r2[melt & (sp == 1)], r1[melt & (sp == 1)], rtol=1e-13)
    np.testing.assert_allclose(r2[melt & (sp == 2)], 3 * r1[melt & (sp == 2)], rtol=1e-13)
    assert (melt & (sp == 1)).sum() > 10 and (melt & (sp == 2)).sum() > 10
    # solidification: PD = PD_1 + PD_2 with PD_2 tripled, and PD_1 : PD_2 fixed
    # by the liquid (3:1) and the equal constants (same energies for both)
    sol = (ev != 4) & (r1 != float(np.float32(1e-8)))  # placeholders keep their fixed rate
    assert sol.sum() > 10
    np.testing.assert_allclose(r2[sol], r1[sol] * (3 + 3 * 1) / (3 + 1), rtol=1e-12)


def test_no_liquid_no_solidification(tmp_path):
    """Until an atom of species 2 melts (CountLiquid(2) > 0 again), no site
    draws species 2: every deposition is of species 1."""
    o, box = loaded(tmp_path, e_pair=((0.6, 0.8), (2.5, 2.5)))  # species 2 hardly melts
    set_variant(o, liquid=(2000, 0))
    o.refresh_all()  # every site's species drawn with these counts
    c2 = o.la_counts()[0][1]
    deps = []
    for _ in range(20000):
        ev = o.step(1)
        if o.la_counts()[0][1] < c2:
            break
        if ev["itrans"][0] <= 3:
            deps.append(int(ev["itrans"][0]))
    assert len(deps) > 50
    assert set(deps) == {1}


def test_sige_swap_holds_the_count(tmp_path):
    """Species drawn with liquid of both species left, then species 1's
    liquid used up (no redraw): with the rule those depositions become
    species 2 and CountLiquid(1) never drops below zero; without it they
    overshoot: CountLiquid(1) goes negative, or the reference's tree update
    stops on a solidification rate that cancels to zero (UpdateTree.f:37-41)."""
    results = {}
    for swap in (False, True):
        (tmp_path / str(swap)).mkdir()
        o, box = loaded(tmp_path / str(swap))
        set_variant(o, liquid=(5000, 5000), swap=swap)
        o.refresh_all()                                # about half the sites draw species 1
        set_variant(o, liquid=(0, 100000), swap=swap)  # its liquid is gone
        lows, stopped = [0], False
        for _ in range(600):
            try:
                o.step(1, record=False)
            except RuntimeError as exc:
                assert "Update" in str(exc)
                stopped = True
                break
            lows.append(int(o.la_counts()[1][0]))
        results[swap] = (min(lows), stopped)
    assert results[False][0] < 0 or results[False][1]
    assert results[True] == (0, False)
