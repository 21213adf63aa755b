"""GPU: the parallel synchronous-sublattice engine and the device builder.

* the device-built slab (mskg_build_slab) equals the reference builders
  SetSi / SetSiC3C as restated by the oracle, site by site;
* mskg_run's own event log, replayed through the oracle (reference update
  rules, serial), reproduces the engine's final lattice bit for bit: the
  parallel execution of concurrent tiles is a valid serial KMC history;
* the rates the engine keeps after a run equal a fresh Get_Prob sweep except
  for the reference's stale-rate cases (placeholder / isolated atom).
"""
import numpy as np
import pytest

from mulskips_amd.autotune import tile_capacity
from mulskips_amd.engine import Engine
from oracle.oracle import Oracle
from tests.helpers import list_map, oracle_state, parity_mask, small_deck

pytestmark = pytest.mark.gpu

CASES = {  # name: (deck, mode, box, lens)
    "SiC": ("surface", 1, (96, 96, 144), (72, 72, 72)),
    "Si": ("Si-SL", 0, (96, 96, 144), (72, 120, 120)),
}


def compare(o, eng, box, classes_only_isolated=True):
    lat_o, mc_o, bulk_o, void_o = oracle_state(o)
    lat_e, mc_e, bulk_e, void_e = eng.store_state(lattcoo=lat_o.copy())
    m = parity_mask(box)
    bad = np.nonzero((lat_o != lat_e) & m)
    assert bad[0].size == 0, f"LattCoo differs at {bad[0].size} points"
    for name, a, b in (("MC", mc_o, mc_e), ("bulk", bulk_o, bulk_e), ("void", void_o, void_e)):
        ma, mb = list_map(a), list_map(b)
        assert ma.keys() == mb.keys(), f"{name} membership differs at {len(ma.keys() ^ mb.keys())} sites"
        assert all(ma[k] == mb[k] for k in ma), f"{name} neighbour lists differ"
    return mc_o


@pytest.mark.parametrize("case", list(CASES))
def test_device_slab_equals_reference_builder(tmp_path, case):
    deck, mode, box, lens = CASES[case]
    path, d = small_deck(tmp_path, deck, lens)
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables())
    eng.build_slab(mode, lens[0] - lens[0] % 24)
    mc_o = compare(o, eng, box)
    cls, rate = eng.site_rates(mc_o["xyz"])
    np.testing.assert_allclose(rate, mc_o["prob"], rtol=1e-6)
    eng.close()


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("tile,epp,wide", [(24, 1.0, False), (16, 4.0, False), (48, 4.0, False), (32, 4.0, True)])
def test_parallel_run_is_a_serial_history(tmp_path, case, tile, epp, wide):
    """Tile edges 16, 24 and 48 on the 96-wide box; 32 (an even number of
    tiles needs a width that is a multiple of 64, and the 3C cell one of
    12) on a 192 x 192 x 144 box."""
    deck, mode, box, lens = CASES[case]
    if wide:
        box = (192, 192, box[2])
    path, d = small_deck(tmp_path, deck, lens)
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables(), seed=11, tile=tile, tile_capacity=tile_capacity(tile), events_per_pass=epp)
    eng.build_slab(mode, lens[0] - lens[0] % 24)
    eng.set_event_log(400000)
    n, t, stop = eng.run(100000)
    log = eng.event_log()
    print(f"[{case} T={tile}] {n} events, {len(log)} logged, stop={stop}", flush=True)
    assert len(log) == n and n > 1000 and stop in (0, 1)
    o.apply_log(log)
    assert o.anomalies()["oob_reads"] == 0, "the film reached the top of the box"
    mc_o = compare(o, eng, box)
    cls, rate = eng.site_rates(mc_o["xyz"])
    keep = cls != 6
    np.testing.assert_allclose(rate[keep], mc_o["prob"][keep], rtol=1e-6)
    counts, clock, lists = eng.counters()
    assert counts.sum() == n
    assert lists == o.counts()
    eng.close()


def test_rates_after_run_match_fresh_sweep(tmp_path):
    deck, mode, box, lens = CASES["SiC"]
    path, d = small_deck(tmp_path, deck, lens)
    eng = Engine(box, d.tables(), seed=5)
    eng.build_slab(mode, lens[0] - lens[0] % 24)
    eng.run(50000)
    _, mc, _, _ = eng.store_state()
    cls0, r0 = eng.site_rates(mc["xyz"])
    eng.refresh_rates()
    cls1, r1 = eng.site_rates(mc["xyz"])
    stale = np.isin(cls0, (2, 6)) | np.isin(cls1, (2, 6))
    np.testing.assert_array_equal(cls0[~stale], cls1[~stale])
    np.testing.assert_allclose(r0[~stale], r1[~stale], rtol=1e-6)
    eng.close()


# ---- second-neighbour energy term (engine option, restated by the oracle) ----
F2 = np.exp(-np.arange(13) * 0.05 / (8.617333262e-5 * 1600.0))


def n2_tables(d):
    t = d.tables()
    t["second_factor"] = F2
    return t


def test_second_neighbour_slab_rates(tmp_path):
    deck, mode, box, lens = CASES["SiC"]
    path, d = small_deck(tmp_path, deck, lens)
    o = Oracle(*box)
    o.load_deck(path)
    o.set_second_factor(F2)
    eng = Engine(box, n2_tables(d), tile=24)
    eng.build_slab(mode, lens[0] - lens[0] % 24)
    mc_o = compare(o, eng, box)
    cls, rate = eng.site_rates(mc_o["xyz"])
    np.testing.assert_allclose(rate, mc_o["prob"], rtol=1e-6)
    # the factor is really applied: surface atoms of a flat slab see 8 occupied second neighbours
    atoms = cls == 4
    assert atoms.any()
    plain = d.tables()["ptrans_e"]
    assert np.all(rate[atoms] < plain.max() * F2[7])
    eng.close()


def test_second_neighbour_replay_bit_exact(tmp_path):
    deck, mode, box, lens = CASES["SiC"]
    path, d = small_deck(tmp_path, deck, lens)
    o = Oracle(*box)
    o.load_deck(path)
    o.set_second_factor(F2)
    lat0, mc0, bulk0, void0 = oracle_state(o)
    eng = Engine(box, n2_tables(d), tile=24)
    eng.load_state(lat0, mc0, bulk0, void0)
    log = o.step(20000)
    eng.replay(log)
    mc_o = compare(o, eng, box)
    cls, rate = eng.site_rates(mc_o["xyz"])
    keep = cls != 6
    np.testing.assert_allclose(rate[keep], mc_o["prob"][keep], rtol=1e-6)
    counts, clock, _ = eng.counters(lists=False)
    assert clock == o.time()
    eng.close()


@pytest.mark.parametrize("tile", [24, 48])
def test_second_neighbour_parallel_run_is_a_serial_history(tmp_path, tile):
    deck, mode, box, lens = CASES["SiC"]
    box = (96, 96, 288)  # the film grows faster with the second-neighbour term: keep it off the box top
    path, d = small_deck(tmp_path, deck, lens)
    o = Oracle(*box)
    o.load_deck(path)
    o.set_second_factor(F2)
    eng = Engine(box, n2_tables(d), seed=23, tile=tile)
    eng.build_slab(mode, lens[0] - lens[0] % 24)
    eng.set_event_log(400000)
    n, t, stop = eng.run(100000)
    log = eng.event_log()
    print(f"[n2 T={tile}] {n} events, stop={stop}", flush=True)
    assert len(log) == n and n > 1000 and stop in (0, 1)
    o.apply_log(log)
    assert o.anomalies()["oob_reads"] == 0, "the film reached the top of the box"
    mc_o = compare(o, eng, box)
    cls, rate = eng.site_rates(mc_o["xyz"])
    keep = cls != 6
    np.testing.assert_allclose(rate[keep], mc_o["prob"][keep], rtol=1e-6)
    eng.close()


def test_second_neighbour_needs_tile_24(tmp_path):
    deck, mode, box, lens = CASES["SiC"]
    path, d = small_deck(tmp_path, deck, lens)
    from mulskips_amd.engine import EngineError
    with pytest.raises(EngineError):
        Engine(box, n2_tables(d), tile=16)


@pytest.mark.parametrize("case", ["SiC", "n2"])
def test_rerun_with_the_same_seed_is_identical(tmp_path, case):
    """Two engines with the same seed, run from the same state, give the
    same lattice, clock and event count: the window comes from fixed-order
    sums and each tile's list order from its contents, not from timing."""
    deck, mode, box, lens = CASES["SiC"]
    path, d = small_deck(tmp_path, deck, lens)
    tab = d.tables() if case == "SiC" else n2_tables(d)
    if case == "n2":
        box = (96, 96, 288)
    runs = []
    for _ in range(2):
        eng = Engine(box, tab, seed=31, tile=24, events_per_pass=4.0, tau_policy=1)
        eng.build_slab(mode, lens[0] - lens[0] % 24)
        n, t, stop = eng.run(60000)
        lat, mc, _, _ = eng.store_state()
        runs.append((n, t, lat, list_map(mc)))
        eng.close()
    (n0, t0, lat0, mc0), (n1, t1, lat1, mc1) = runs
    assert n0 == n1 and n0 > 10000
    assert t0 == t1
    np.testing.assert_array_equal(lat0, lat1)
    assert mc0 == mc1
