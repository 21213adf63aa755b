"""GPU: the rate-update sweep over the tile layers that hold MC particles
(mskg_refresh_active) against the sweep over every stored point
(mskg_refresh_rates), from the same state after a run:

* the MC lists (membership, neighbour lists, ProbTrans) and LattCoo equal
  bit for bit, and so do the classes and rates of every MC particle;
* the active sweep reads the points of the layers only, the full sweep all
  of them (mskg_last_timing);
* with lists that are not current (a fresh builder state) the active sweep
  reads every point;
* a second active sweep in a row still reads the layers only (the sweep
  keeps the engine's tile lists current).
"""
import numpy as np
import pytest

from mulskips_amd.engine import Engine
from oracle.oracle import Oracle
from tests.helpers import list_map, oracle_state, small_deck

pytestmark = pytest.mark.gpu

F2 = np.exp(-np.arange(13) * 0.05 / (8.617333262e-5 * 1600.0))

CASES = {  # name: (deck, builder mode (None: the deck's own start), box, lens, second factor)
    "SiC-f2": ("surface", 1, (192, 192, 240), (72, 72, 72), True),
    "CVD": ("CVD-Si-SiH2", None, (96, 96, 192), (72, 120, 120), False),
}


def make(tmp_path, case, seed=17):
    deck, mode, box, lens, f2 = CASES[case]
    path, d = small_deck(tmp_path, deck, lens)
    tab = d.tables()
    if f2:
        tab["second_factor"] = F2
    eng = Engine(box, tab, seed=seed, tile=24, events_per_pass=4.0)
    if mode is None:
        o = Oracle(*box)
        o.load_deck(path)
        eng.load_state(*oracle_state(o))
        o.close()
    else:
        eng.build_slab(mode, lens[0] - lens[0] % 24)
    return eng, box


def state(eng):
    lat, mc, bulk, void = eng.store_state()
    m = list_map(mc)
    prob = dict(zip(map(tuple, mc["xyz"].tolist()), mc["prob"].tolist()))
    cls, rate = eng.site_rates(mc["xyz"])
    return lat, m, prob, cls, rate


@pytest.mark.parametrize("case", list(CASES))
def test_active_sweep_equals_full_sweep(tmp_path, case):
    eng, box = make(tmp_path, case)
    npts = (box[0] // 2) * (box[1] // 2) * (box[2] + 1)
    eng.refresh_rates(active=True)  # builder / upload state: lists not current
    assert eng.last_timing()["sweep_points"] == npts
    n, t, stop = eng.run(40000)
    assert n >= 40000 or stop
    ref = Engine(box, eng_tables(tmp_path, case), seed=17, tile=24, events_per_pass=4.0)
    ref.copy_state_from(eng)
    eng.refresh_rates(active=True)
    swept = eng.last_timing()["sweep_points"]
    ref.refresh_rates()
    assert ref.last_timing()["sweep_points"] == npts
    print(f"[{case}] active sweep read {swept} of {npts} points", flush=True)
    assert 0 < swept < npts
    lat_a, m_a, p_a, c_a, r_a = state(eng)
    lat_b, m_b, p_b, c_b, r_b = state(ref)
    assert np.array_equal(lat_a, lat_b)
    assert m_a == m_b, "MC lists differ"
    assert p_a == p_b, "ProbTrans differs"
    np.testing.assert_array_equal(c_a, c_b)
    np.testing.assert_array_equal(r_a, r_b)
    # again: the lists stay current, so the second active sweep reads the layers only
    eng.refresh_rates(active=True)
    assert eng.last_timing()["sweep_points"] == swept
    ref.close()
    eng.close()


def eng_tables(tmp_path, case):
    deck, mode, box, lens, f2 = CASES[case]
    _, d = small_deck(tmp_path, deck, lens)
    tab = d.tables()
    if f2:
        tab["second_factor"] = F2
    return tab
