"""CPU: the oracle's state hand-over (orc_load_state).

A state taken from one oracle and handed to another (LattCoo and the three
lists, the layout of the engine's mskg_store_state) continues exactly like
the original: both apply the same event log and end with the same lattice
and lists; after a refresh the rates equal the original's Get_Prob.  This is
what lets the GPU tests load an engine-built state (a device alloy slab)
into the oracle and replay the engine's runs there.
"""
import numpy as np

from oracle.oracle import Oracle
from tests.helpers import SMALL, list_map, oracle_state, small_deck


def test_handed_over_state_continues_like_the_original(tmp_path):
    base, box, lens = SMALL["cube-zigzag"]
    path, d = small_deck(tmp_path, base, lens)
    a = Oracle(*box)
    a.load_deck(path)
    a.step(3000, record=False)
    lat, mc, bulk, void = oracle_state(a)
    b = Oracle(*box)
    b.load_deck(path)
    b.load_state(lat, mc, bulk, void, kmc_time=a.time())
    b.refresh_all()
    mc_b = b.lists()[0]
    ma, mb = list_map(mc), list_map(mc_b)
    assert ma == mb
    # Get_Prob of the same state in FP64: evaporation rates of atoms with
    # neighbours (an adsorption site's rate follows its pre-drawn species,
    # an isolated atom's the total rate at its evaluation)
    coor = (lat >> 24) & 7

    def evap(m):
        x, y, z = m["xyz"].T
        keep = (m["event"] == 4) & (coor[z, y, x] > 0)
        return dict(zip(map(tuple, m["xyz"][keep].tolist()), m["prob"][keep]))
    pa, pb = evap(mc), evap(mc_b)
    assert len(pa) > 100 and pa == pb
    log = a.step(4000)
    b.apply_log(log)
    np.testing.assert_array_equal(a.lattcoo(), b.lattcoo())
    for la, lb in zip(a.lists(), b.lists()):
        assert list_map(la) == list_map(lb)
    a.close()
    b.close()


def test_frozen_atoms_take_the_vanishing_rate(tmp_path):
    """An MC particle handed over without a rate (a frozen-substrate atom in
    the engine's lists) is listed with the rate the frozen option gives it."""
    base, box, lens = SMALL["Si-SL"]
    path, d = small_deck(tmp_path, base, lens)
    a = Oracle(*box)
    a.load_deck(path)
    lat, mc, bulk, void = oracle_state(a)
    mc = dict(mc)
    mc["prob"] = mc["prob"].copy()
    mc["prob"][:5] = 0.0
    b = Oracle(*box)
    b.load_deck(path)
    b.load_state(lat, mc, bulk, void)
    assert (b.lists()[0]["prob"][:5] == 1e-300).all()
    a.close()
    b.close()
