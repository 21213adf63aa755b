"""GPU: the batched synchronous-sublattice ensemble and the stored-planes box.

* mskg_run_ensemble advances replicas of different boxes, rate models, tile
  edges, window policies and limits in one run; every replica ends exactly
  as mskg_run leaves the same engine run alone (lattice, lists, counters,
  clock, events and stop reason, bit for bit).
* A replica's event log from the ensemble, replayed serially by the oracle
  from the replica's initial state, reproduces its lattice bit for bit: the
  ensemble's concurrent tiles of many lattices form a valid serial history
  of each lattice.
* An engine that stores only the bottom planes of its box (stored_len_z)
  runs, replays and reports exactly as the engine storing every plane, and
  a film that grows towards the stored top stops the run with an error
  before any event reaches past it.
"""
import os

import numpy as np
import pytest

from mulskips_amd import deck as deckmod
from mulskips_amd import models
from mulskips_amd.engine import Engine, EngineError, run_ensemble, run_serial
from oracle.oracle import Oracle
from tests.helpers import list_map, oracle_state, parity_mask

pytestmark = pytest.mark.gpu

SLAB, FROZEN = 48, 36

# name: box, stored planes (None = all), rate model, engine options, max events, max time
REPLICAS = {
    "sige-96-windowed": ((96, 96, 768), 120, ("sige", 1200.0, 1.0, 0.2), dict(tile=24, events_per_pass=4.0,
                                                                             tau_policy=1), 30000, None),
    "sige-192": ((192, 192, 144), None, ("sige", 1000.0, 5.0, 0.4), dict(tile=16), 40000, None),
    "sige-384-windowed": ((384, 384, 1536), 144, ("sige", 1400.0, 0.3, 0.1), dict(tile=24, events_per_pass=4.0,
                                                                                 tau_policy=1), 10 ** 9, 2e-6),
    "sic-n2": ((96, 96, 144), None, ("sic",), dict(tile=24, events_per_pass=4.0, tau_policy=1), 50000, None),
    "si-144-windowed": ((144, 144, 288), 96, ("si",), dict(tile=24, events_per_pass=1.0), 20000, None),
    "sige-48-short": ((48, 48, 192), 96, ("sige", 900.0, 10.0, 0.5), dict(tile=24), 500, None),
}


def tables_of(model):
    if model[0] == "sige":
        return models.sige_alloy_tables(*model[1:])
    if model[0] == "sic":
        return models.sic_c2_tables()
    return models.si_c1_tables()


def build(name, seed, log_cap=0):
    box, stored, model, opts, _, _ = REPLICAS[name]
    e = Engine(box, tables_of(model), seed=seed, frozen_below_z=FROZEN, stored_planes=stored, **opts)
    if model[0] == "sige":
        e.build_alloy_slab(SLAB, model[3])
    else:
        e.build_slab(1 if model[0] == "sic" else 0, SLAB)
    e.refresh_rates()
    if log_cap:
        e.set_event_log(log_cap)
    return e


def state_of(e):
    lat, mc, bulk, void = e.store_state()
    counts, clock, lists = e.counters()
    return lat, list_map(mc), list_map(bulk), list_map(void), counts, clock, lists


def test_ensemble_equals_runs_alone():
    names = list(REPLICAS)
    seeds = [101 + 7 * i for i in range(len(names))]
    limits_e = [REPLICAS[n][4] for n in names]
    limits_t = [REPLICAS[n][5] or float("inf") for n in names]
    engs = [build(n, s) for n, s in zip(names, seeds)]
    done, clock, stop = run_ensemble(engs, limits_e, limits_t)
    print(f"[ensemble] events {done.tolist()} stop {stop.tolist()} clock {clock.tolist()}", flush=True)
    after = [state_of(e) for e in engs]
    for e in engs:
        e.close()
    for i, (n, s) in enumerate(zip(names, seeds)):
        e = build(n, s)
        k, t, st = e.run(limits_e[i], limits_t[i])
        alone = state_of(e)
        e.close()
        print(f"[alone] {n}: {k} events, clock {t:.6g}, stop {st}", flush=True)
        assert k == done[i] and t == clock[i] and st == stop[i], n
        assert k > 0, n
        np.testing.assert_array_equal(alone[0], after[i][0], err_msg=n)
        for j in (1, 2, 3):
            assert alone[j] == after[i][j], (n, j)
        np.testing.assert_array_equal(alone[4], after[i][4], err_msg=n)
        assert alone[5] == after[i][5] and alone[6] == after[i][6], n
    # the time-limited replica ran to its limit, the short one to its event count
    assert clock[names.index("sige-384-windowed")] == REPLICAS["sige-384-windowed"][5]
    assert done[names.index("sige-48-short")] >= 500


def oracle_with_state(tmp_path, name, eng):
    """An oracle of the replica's whole box holding the engine's state: the
    builder's walls from the oracle's own builder, the lattice points of the
    stored planes from the engine."""
    box, stored, model, _, _, _ = REPLICAS[name]
    tab = tables_of(model)
    d = deckmod.deck_from_tables(tab, "FS", [SLAB, 120, 120], idum=5)
    path = os.path.join(str(tmp_path), f"start-{name}.dat")
    deckmod.write_start_dat(d, path)
    o = Oracle(*box)
    o.load_deck(path)
    lat0 = o.lattcoo().copy()
    lat_s, mc_s, bulk_s, void_s = eng.store_state()
    m = parity_mask(box)[:lat_s.shape[0]]
    part = lat0[:lat_s.shape[0]]
    part[m] = lat_s[m]
    o.load_state(lat0, mc_s, bulk_s, void_s)
    o.set_frozen(FROZEN)
    if tab.get("second_factor") is not None:
        o.set_second_factor(tab["second_factor"])
    o.refresh_all()
    return o


@pytest.mark.parametrize("name", ["sige-96-windowed", "sic-n2"])
def test_ensemble_log_is_a_serial_history(tmp_path, name):
    others = ["sige-192", "si-144-windowed"]
    engs = [build(name, 77, log_cap=400000)] + [build(n, 5 + i) for i, n in enumerate(others)]
    o = oracle_with_state(tmp_path, name, engs[0])
    done, clock, stop = run_ensemble(engs, [REPLICAS[name][4]] + [REPLICAS[n][4] for n in others])
    log = engs[0].event_log()
    print(f"[{name}] {done[0]} events in the ensemble, {len(log)} logged", flush=True)
    assert len(log) == done[0] > 1000
    o.apply_log(log)
    assert o.anomalies()["oob_reads"] == 0
    lat_o, mc_o, bulk_o, void_o = oracle_state(o)
    lat_e, mc_e, bulk_e, void_e = engs[0].store_state()
    k = lat_e.shape[0]
    m = parity_mask(REPLICAS[name][0])[:k]
    assert np.count_nonzero((lat_o[:k] != lat_e) & m) == 0
    for a, b in ((mc_o, mc_e), (bulk_o, bulk_e), (void_o, void_e)):
        assert list_map(a) == list_map(b)
    cls, rate = engs[0].site_rates(mc_o["xyz"])
    keep = (cls != 6) & (mc_o["xyz"][:, 2] >= FROZEN)
    np.testing.assert_allclose(rate[keep], mc_o["prob"][keep], rtol=1e-6)
    o.close()
    for e in engs:
        e.close()


def test_stored_planes_equal_the_whole_box():
    box = (96, 96, 384)
    tab = models.sige_alloy_tables(1100.0, 2.0, 0.3)
    res = []
    for stored in (None, 120):
        e = Engine(box, tab, seed=9, tile=24, events_per_pass=4.0, tau_policy=1, frozen_below_z=FROZEN,
                   stored_planes=stored)
        e.build_alloy_slab(SLAB, 0.3)
        e.refresh_rates()
        e.set_event_log(200000)
        n1, t1, s1 = e.run(40000)
        n2, t2, s2 = e.run_serial(20000)
        lat, mc, bulk, void = e.store_state()
        frame = e.frame()
        log = e.event_log()
        res.append(((n1, t1, s1, n2, t2, s2), lat[:120], list_map(mc), list_map(bulk), list_map(void),
                    e.counters(), frame, log))
        e.close()
    full, win = res
    assert full[0] == win[0] and full[0][0] >= 40000 and full[0][3] == 20000
    np.testing.assert_array_equal(full[1], win[1])
    assert full[2] == win[2] and full[3] == win[3] and full[4] == win[4]
    np.testing.assert_array_equal(full[5][0], win[5][0])
    assert full[5][1:] == win[5][1:]
    for a, b in zip(full[6], win[6]):
        np.testing.assert_array_equal(a, b)
    # the serial part of the log is ordered: the same events in the same order
    for f in ("site", "itrans", "arm", "time"):
        np.testing.assert_array_equal(full[7][f][-20000:], win[7][f][-20000:])


def test_film_reaching_the_stored_top_stops_the_run():
    """Si at 400 K grows about one monolayer (3 planes) per second of KMC
    time: the film reaches the guard 24 planes below the stored top and the
    run stops with an error instead of reading or writing past it."""
    box = (96, 96, 768)
    tab = models.si_c1_tables(temperature=400.0, flux=1.0)
    stored = SLAB + 72
    e = Engine(box, tab, seed=3, tile=24, stored_planes=stored)
    e.build_slab(0, SLAB)
    with pytest.raises(EngineError) as err:
        e.run(10 ** 7, max_time=100.0)
    assert "stored planes" in str(err.value)
    lat, mc, _, _ = e.store_state()
    assert lat.shape[0] == stored
    # nothing was written into the top 18 stored planes
    assert not np.any(lat[stored - 18:] & (1 << 23))
    e.close()
    with pytest.raises(EngineError):
        Engine(box, tab, stored_planes=20)


def test_stored_planes_refuse_the_whole_grid_maps():
    box = (96, 96, 384)
    tab, _ = models.la_sige_tables(2026)
    la = models.LA_CONSTANTS
    e = Engine(box, tab, la=la, stored_planes=120)
    e.build_alloy_slab(SLAB, 0.3)
    with pytest.raises(EngineError):
        e.set_temperature_field(models.la_front_field(box, 60.0), "socket")
    with pytest.raises(EngineError):
        e.phases()
    e.close()


def test_ensemble_rejects_mixed_kernels():
    a = build("sige-192", 1)
    tab, _ = models.la_sige_tables(2026)
    b = Engine((96, 96, 144), tab, la=models.LA_CONSTANTS)
    b.build_alloy_slab(SLAB, 0.3)
    with pytest.raises(EngineError):
        run_ensemble([a, b], [100, 100])
    with pytest.raises(EngineError):
        run_ensemble([a, a], [100, 100])
    a.close()
    b.close()


def test_ensemble_of_coverage_and_laser_annealing_replicas(tmp_path):
    """The other kernel instances of the ensemble's pass: replicas with
    coverage species (the CVD-Si-SiH2 deck, H absorption and desorption)
    and replicas in laser-annealing mode (SiGe under the tanh front) each
    end as mskg_run leaves them alone.  The melting front carries MC
    particles into new tile layers within a sweep, where the passes of the
    later colours must visit them: the case that needs each colour's plan
    made just before its pass (a plan per sweep parted from mskg_run after
    ~2e4 events here)."""
    from tests.helpers import SMALL, small_deck
    base, box, lens = SMALL["CVD-Si-SiH2"]
    path, d = small_deck(tmp_path, base, lens)
    o = Oracle(*box)
    o.load_deck(path)
    start = oracle_state(o)
    o.close()

    def cvd(seed):
        e = Engine(box, d.tables(), seed=seed, tile=24, events_per_pass=4.0, tau_policy=1)
        e.load_state(*start)
        return e

    la_box = (96, 96, 240)
    tab, _ = models.la_sige_tables(2026)

    def la(seed, zm):
        e = Engine(la_box, tab, seed=seed, tile=24, events_per_pass=4.0, tau_policy=1, la=models.LA_CONSTANTS)
        e.build_alloy_slab(120, 0.3)
        e.set_temperature_field(models.la_front_field(la_box, zm), "socket")
        e.refresh_rates()
        return e

    for make, args, limit in ((cvd, [(3,), (4,)], 40000), (la, [(5, 120.0), (6, 126.0)], 30000)):
        engs = [make(*a) for a in args]
        done, clock, stop = run_ensemble(engs, [limit] * len(engs))
        after = [state_of(e) for e in engs]
        for e in engs:
            e.close()
        for i, a in enumerate(args):
            e = make(*a)
            k, t, s = e.run(limit)
            alone = state_of(e)
            e.close()
            print(f"[{make.__name__} {a}] ensemble {done[i]} events, alone {k}", flush=True)
            assert (k, t, s) == (done[i], clock[i], stop[i]) and k >= limit
            np.testing.assert_array_equal(alone[0], after[i][0])
            assert alone[1:4] == after[i][1:4]
            np.testing.assert_array_equal(alone[4], after[i][4])
            assert alone[5:] == after[i][5:]
