"""GPU: laser-annealing rates (Get_Prob.f:127-149) from a per-site
temperature field (the T bits of LattCoo, SetT.f:25-50).

In LA mode the tables hold energies and the rates are
  evaporation  P0 exp(-E / (kB T)) [E != 0]
  deposition   0.5 P0 exp(-E / (kB Tm)) (1 + erf((T - Tflex) / sigma)) [E != 0]
per site, with the reference's kB = 8.61733035e-5 eV/K (Get_Prob.f:70).  No
regression fixture of the reference covers LA mode (its LA fixture is stale
against its source), so the rates are pinned by these closed forms,
evaluated in FP64 in the test, for a temperature field that varies across
the slab; a short run then checks that LA mode executes and keeps every
site's temperature.
"""
import math

import numpy as np
import pytest

from mulskips_amd import models
from mulskips_amd.engine import Engine

pytestmark = pytest.mark.gpu
KB = 0.00008617330350
LA = dict(tm=1687.0, p0=1e13, sigma=50.0, tflex=1650.0)


def la_tables():
    # energies (eV) in the rate-table slots: evaporation n * 0.6 eV, deposition 0.3 eV
    t = models.pair_energy_tables([[0.6]], [1.0], 1.0, 1.0, (14,))
    pe = np.zeros(3 * 4 ** 6)
    for n in (1, 2, 3):
        pe[models.pe_index(1, n, 0, 0)] = 0.6 * n
    t["ptrans_e"] = pe
    pd = np.zeros(9)
    pd[0::3] = 0.3  # species 1 at coordination 1..3
    t["ptrans_d"] = pd
    return t


def temperature_field(box):
    lx, ly, lz = box
    z, y, x = np.meshgrid(np.arange(lz + 1), np.arange(ly), np.arange(lx), indexing="ij")
    return (1500 + 2 * z + (x % 24) + 3 * (y % 12)).astype(np.int32)


def test_la_rates_closed_form():
    box = (96, 96, 144)
    eng = Engine(box, la_tables(), seed=5, la=LA)
    eng.build_slab(0, 72)
    T = temperature_field(box)
    eng.set_temperature(T)
    eng.refresh_rates()
    lat, mc, _, _ = eng.store_state()
    cls, rate = eng.site_rates(mc["xyz"])
    n_evap = n_dep = 0
    for (x, y, z), c, r in zip(mc["xyz"], cls, rate):
        t = float(T[z, y, x])
        assert ((lat[z, y, x] >> 9) & 8191) == T[z, y, x]  # LattCoo's T bits carry the field
        coor = (lat[z, y, x] >> 24) & 7
        if c == 4:
            want = LA["p0"] * math.exp(-0.6 * coor / (KB * t))
            n_evap += 1
        elif c == 1:
            want = 0.5 * LA["p0"] * math.exp(-0.3 / (KB * LA["tm"])) * (1 + math.erf((t - LA["tflex"]) / LA["sigma"]))
            n_dep += 1
        else:
            continue
        assert r == pytest.approx(want, rel=1e-6), (x, y, z, c, t)
    assert n_evap > 100 and n_dep > 100
    # LA mode runs, and evaporation/deposition keep the temperature bits
    n, clock, stop = eng.run(20000)
    assert n >= 20000
    lat2 = eng.store_state()[0]
    valid = ((lat2 & 255) != 113)
    np.testing.assert_array_equal(((lat2 >> 9) & 8191)[valid & (((lat >> 9) & 8191) > 0)],
                                  ((lat >> 9) & 8191)[valid & (((lat >> 9) & 8191) > 0)])
    eng.close()
