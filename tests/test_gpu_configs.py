"""GPU: the batched-ensemble and laser-annealing configurations at their
configured sizes, checked against the oracle.

* Batched ensemble: Si(1-x)Ge(x) replicas with the ensemble's draws of T,
  flux and Ge fraction, one replica per configured lateral size (32, 64,
  128 and 256 cells) at configured heights (64-256 cells), built on the
  device as random-alloy slabs and run together in one mskg_run_serial
  launch.  For the replicas whose grid the oracle holds in host memory
  (lateral 32-128 cells), the oracle takes the engine's initial state and
  replays the replica's serial event log: lattice, lists and rates equal
  the engine's.  The 256-cell replica (a 3072 x 3072 x 768 grid, beyond
  the oracle's host memory) gives the same log and frame records run alone
  as in the batch, and its event counts agree with the change of its atoms.
* The same replicas through the batched ensemble (mskg_run_ensemble), each
  storing its bottom planes, run to 20 ML deposited: the oracle replays the
  logs of those it holds, and the 256-cell replica ends as mskg_run alone.
* Laser annealing: Si0.7Ge0.3 on the 264 x 264 x 384 grid in LA mode, the
  tanh front T(z) = 1700 + 300 tanh((z - zm)/5 nm) exchanged twice in the
  FEM socket layout, 1e5 events of the synchronous-sublattice engine per
  exchange: the oracle, holding the engine's post-build state and the
  same temperature maps, replays the engine's logs bit for bit and its LA
  rates (Get_Prob.f:127-149) agree within 1e-6 after each exchange.
"""
import os

import numpy as np
import pytest

from mulskips_amd import deck as deckmod
from mulskips_amd import models
from mulskips_amd.engine import Engine, run_ensemble, run_serial
from oracle.oracle import Oracle
from tests.helpers import list_map, oracle_state, parity_mask

pytestmark = pytest.mark.gpu

SLAB, FROZEN = 48, 36  # the ensemble's starting slab and frozen substrate (grid units)


def oracle_for(tmp_path, box, tables, init_st, slab, name):
    """An oracle of this box whose builder lays the same walls as the device
    builder (SetSi: wall planes on every grid point)."""
    d = deckmod.deck_from_tables(tables, init_st, [slab, 120, 120], idum=5)
    path = os.path.join(str(tmp_path), f"start-{name}.dat")
    deckmod.write_start_dat(d, path)
    o = Oracle(*box)
    o.load_deck(path)
    return o


def same_state(o, eng, box, tag):
    lat_o, mc_o, bulk_o, void_o = oracle_state(o)
    lat_e, mc_e, bulk_e, void_e = eng.store_state(lattcoo=lat_o.copy())
    bad = np.count_nonzero((lat_o != lat_e) & parity_mask(box))
    assert bad == 0, f"{tag}: LattCoo differs at {bad} points"
    for a, b in ((mc_o, mc_e), (bulk_o, bulk_e), (void_o, void_e)):
        assert list_map(a) == list_map(b), f"{tag}: lists differ"
    return mc_o


def rates_agree(eng, mc, tag, z_min=0):
    cls, rate = eng.site_rates(mc["xyz"])
    keep = (cls != 6) & (mc["xyz"][:, 2] >= z_min)
    assert keep.sum() > 100, tag
    np.testing.assert_allclose(rate[keep], mc["prob"][keep], rtol=1e-6, err_msg=tag)


# (lateral cells, height cells): each configured lateral size once, at configured heights
C3_BOXES = [(32, 256), (64, 128), (128, 64), (256, 64)]
ORACLE_MAX_LATERAL = 128
C3_EVENTS = 200_000


def c3_replicas():
    draws = models.ensemble_replicas(len(C3_BOXES))
    reps = []
    for (lc, hc), r in zip(C3_BOXES, draws):
        reps.append(dict(r, box=(12 * lc, 12 * lc, 12 * hc)))
    return reps


def c3_engine(r, log_cap=0):
    e = Engine(r["box"], r["tables"], seed=r["seed"], tile=24, frozen_below_z=FROZEN)
    e.build_alloy_slab(SLAB, r["x"])
    e.refresh_rates()
    if log_cap:
        e.set_event_log(log_cap)
    return e


@pytest.mark.timeout(900)
def test_c3_replicas_at_configured_sizes(tmp_path):
    reps = c3_replicas()
    for r in reps:
        print(f"[c3] replica box {r['box']} T={r['temperature']:.0f} K flux={r['flux']:.3g} ML/s "
              f"x={r['x']:.3f}", flush=True)
    engs = [c3_engine(r, C3_EVENTS + 1000) for r in reps]
    starts = []
    for r, e in zip(reps, engs):
        if r["box"][0] // 12 <= ORACLE_MAX_LATERAL:
            starts.append(e.store_state())
        else:
            starts.append(None)
    done, clock, stop = run_serial(engs, [C3_EVENTS] * len(engs))
    print(f"[c3] batched: events {done.tolist()}, stop {stop.tolist()}", flush=True)
    assert (done == C3_EVENTS).all() and (stop == 0).all()
    for r, e, st in zip(reps, engs, starts):
        lc = r["box"][0] // 12
        log = e.event_log()
        assert len(log) == C3_EVENTS
        if st is None:
            continue
        o = oracle_for(tmp_path, r["box"], r["tables"], "FS", SLAB, f"c3-{lc}")
        lat0 = o.lattcoo().copy()
        lat_s, mc_s, bulk_s, void_s = st
        # the lattice points from the engine, the off-lattice wall planes from the oracle's builder
        m = parity_mask(r["box"])
        lat0[m] = lat_s[m]
        o.load_state(lat0, mc_s, bulk_s, void_s)
        o.set_frozen(FROZEN)
        o.refresh_all()
        print(f"[c3] oracle replays {len(log)} events of the {lc}-cell replica", flush=True)
        o.apply_log(log)
        mc = same_state(o, e, r["box"], f"c3 lateral {lc}")
        rates_agree(e, mc, f"c3 lateral {lc}", z_min=FROZEN)
        o.close()
        del lat0, st
    # the 256-cell replica alone gives the same trajectory as in the batch
    big = [i for i, r in enumerate(reps) if r["box"][0] // 12 > ORACLE_MAX_LATERAL]
    for i in big:
        r = reps[i]
        log_b = engs[i].event_log()
        counts = engs[i].counters(lists=False)[0]
        evap, _ = engs[i].species_counts()
        # every deposition adds an atom of its species and every evaporation removes one
        assert int(counts[3]) == int(evap.sum())
        # the lattice's fingerprint without a 29 GB host copy: the frame
        # records (every atom of coordination <= 3, vacancy and
        # non-epitaxial atom, gathered on the device)
        frame_b = engs[i].frame()
        engs[i].close()
        engs[i] = None
        alone = c3_engine(r, C3_EVENTS + 1000)
        n, t, s = alone.run_serial(C3_EVENTS)
        log_a = alone.event_log()
        for f in ("site", "itrans", "arm", "coor", "time"):
            np.testing.assert_array_equal(log_a[f], log_b[f])
        frame_a = alone.frame()
        assert len(frame_a[0]) > 1000
        for a, b in zip(frame_a, frame_b):
            np.testing.assert_array_equal(a, b)
        alone.close()
        print(f"[c3] {r['box'][0] // 12}-cell replica: batch and single runs agree", flush=True)
    for e in engs:
        if e is not None:
            e.close()


ENS_STORED = 192  # the planes the batched ensemble stores per replica (its film stays far below)


@pytest.mark.timeout(900)
def test_c3_ensemble_at_configured_sizes(tmp_path):
    """The batched ensemble path on one replica per configured lateral
    size at configured heights, each storing its bottom 192 planes, run to
    20 ML deposited in one mskg_run_ensemble: the oracle replays the logs of
    the replicas whose whole box it holds (lateral 32-128 cells) bit for
    bit, and the 256-cell replica ends exactly as mskg_run leaves it alone."""
    reps = c3_replicas()
    engs, starts = [], []
    for r in reps:
        lc = r["box"][0] // 12
        e = Engine(r["box"], r["tables"], seed=r["seed"], tile=24, events_per_pass=4.0, tau_policy=1,
                   frozen_below_z=FROZEN, stored_planes=ENS_STORED)
        e.build_alloy_slab(SLAB, r["x"])
        e.refresh_rates()
        if lc <= ORACLE_MAX_LATERAL:
            e.set_event_log(4_000_000)
            starts.append(e.store_state())
        else:
            starts.append(None)
        engs.append(e)
    done, clock, stop = run_ensemble(engs, [10 ** 12] * len(engs), [r["t_end"] for r in reps])
    print(f"[c3 ensemble] events {done.tolist()}, stop {stop.tolist()}", flush=True)
    assert (stop == 0).all() and all(c == r["t_end"] for c, r in zip(clock, reps))
    for r, e, st, n in zip(reps, engs, starts, done):
        lc = r["box"][0] // 12
        # 20 ML over the lateral cells: at least 2 deposited atoms per cell and ML
        assert n >= 40 * lc * lc, (lc, n)
        if st is None:
            continue
        log = e.event_log()
        assert len(log) == n
        o = oracle_for(tmp_path, r["box"], r["tables"], "FS", SLAB, f"c3e-{lc}")
        lat0 = o.lattcoo().copy()
        lat_s, mc_s, bulk_s, void_s = st
        k = lat_s.shape[0]
        m = parity_mask(r["box"])[:k]
        part = lat0[:k]
        part[m] = lat_s[m]
        o.load_state(lat0, mc_s, bulk_s, void_s)
        o.set_frozen(FROZEN)
        o.refresh_all()
        print(f"[c3 ensemble] oracle replays {len(log)} events of the {lc}-cell replica", flush=True)
        o.apply_log(log)
        lat_o, mc_o, bulk_o, void_o = oracle_state(o)
        lat_e, mc_e, bulk_e, void_e = e.store_state()
        assert np.count_nonzero((lat_o[:k] != lat_e) & m) == 0, f"lateral {lc}: LattCoo differs"
        for a, b in ((mc_o, mc_e), (bulk_o, bulk_e), (void_o, void_e)):
            assert list_map(a) == list_map(b), f"lateral {lc}: lists differ"
        rates_agree(e, mc_o, f"c3 ensemble lateral {lc}", z_min=FROZEN)
        o.close()
        del lat0, st, lat_o
    for i, r in enumerate(reps):
        if r["box"][0] // 12 <= ORACLE_MAX_LATERAL:
            continue
        lat_b, mc_b, _, _ = engs[i].store_state()
        cnt_b = engs[i].counters()
        engs[i].close()
        engs[i] = None
        alone = Engine(r["box"], r["tables"], seed=r["seed"], tile=24, events_per_pass=4.0, tau_policy=1,
                       frozen_below_z=FROZEN, stored_planes=ENS_STORED)
        alone.build_alloy_slab(SLAB, r["x"])
        alone.refresh_rates()
        n, t, s = alone.run(10 ** 12, r["t_end"])
        assert n == done[i] and t == clock[i] and s == stop[i]
        lat_a, mc_a, _, _ = alone.store_state()
        np.testing.assert_array_equal(lat_a, lat_b)
        assert list_map(mc_a) == list_map(mc_b)
        cnt_a = alone.counters()
        np.testing.assert_array_equal(cnt_a[0], cnt_b[0])
        assert cnt_a[1:] == cnt_b[1:]
        alone.close()
        print(f"[c3 ensemble] {r['box'][0] // 12}-cell replica: ensemble and mskg_run agree", flush=True)
    for e in engs:
        if e is not None:
            e.close()


C4_BOX = (264, 264, 384)
C4_SLAB = 240
C4_EVENTS = 100_000


@pytest.mark.timeout(600)
def test_c4_laser_annealing_sige_is_a_serial_history(tmp_path):
    tab, barriers = models.la_sige_tables(2026)
    la = models.LA_CONSTANTS
    print(f"[c4] barriers {barriers}", flush=True)
    o = oracle_for(tmp_path, C4_BOX, tab, "FS", C4_SLAB, "c4")
    o.set_la(la["tm"], la["p0"], la["sigma"], la["tflex"])
    eng = Engine(C4_BOX, tab, seed=2026, tau_policy=1, events_per_pass=4.0, la=la)
    eng.build_alloy_slab(C4_SLAB, 0.3)
    lat, mc, bulk, void = eng.store_state(lattcoo=o.lattcoo().copy())
    sp = (lat >> 27) & 7
    atoms = ((lat >> 23) & 1).astype(bool) & ((lat & 255) != 113)
    x_ge = float(np.mean(sp[atoms] == 2))
    assert abs(x_ge - 0.3) < 0.01, x_ge
    o.load_state(lat, mc, bulk, void)
    eng.set_event_log(3 * C4_EVENTS)
    seen = 0
    zm = float(C4_SLAB)
    for cycle in range(2):
        f = models.la_front_field(C4_BOX, zm)
        for side in (o, eng):
            side.set_temperature_field(f, "socket")
        o.refresh_all()
        eng.refresh_rates()
        n, t, stop = eng.run(C4_EVENTS)
        log = eng.event_log()
        new = log[seen:]
        seen = len(log)
        kinds = np.bincount(new["itrans"], minlength=5)[1:5]
        print(f"[c4] exchange {cycle}: front at z={zm:.0f}, {n} events (deposits Si/Ge, melts: "
              f"{kinds[0]}/{kinds[1]}, {kinds[3]}), stop {stop}", flush=True)
        assert n >= C4_EVENTS and len(new) == n and stop == 0
        o.apply_log(new)
        mc_o = same_state(o, eng, C4_BOX, f"c4 exchange {cycle}")
        rates_agree(eng, mc_o, f"c4 exchange {cycle}")
        assert eng.diagnostics()["guard_exits"] == 0
        zm += 3.0
    eng.close()
    o.close()
