"""GPU: the laser-annealing coupling maps on the device against the oracle.

mskg_set_temperature_field restates SetT (mulskips-source-LaserAnnealing/
initial/SetT.f:38-48, :56-66) and mskg_phases the phase maps of main.f:796-818
of that build; the oracle's restatements are pinned in
tests/test_oracle_la_coupling.py.  On a regression deck grown by replayed
oracle events, the engine's maps equal the oracle's at every grid point and
its LattCoo (T bits included) equals the oracle's at every lattice point,
bit for bit, for both map layouts and for host and device buffers.
"""
import numpy as np
import pytest
import torch

from mulskips_amd import deck as deckmod
from mulskips_amd.engine import Engine, EngineError
from oracle.oracle import Oracle
from tests.helpers import SMALL, oracle_state, parity_mask, small_deck
from tests.test_oracle_la_coupling import self_naming_field, to_socket

pytestmark = pytest.mark.gpu


def grown(tmp_path, key, n=3000):
    base, box, lens = SMALL[key]
    path, d = small_deck(tmp_path, base, lens)
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables())
    eng.load_state(*oracle_state(o))
    log = o.step(n)
    print(f"[{key}] replaying {len(log)} events", flush=True)
    eng.replay(log)
    return o, eng, box


def same_lattcoo(o, eng, box):
    lat_o = o.lattcoo()
    lat_e = eng.store_state(lattcoo=lat_o.copy())[0]
    m = parity_mask(box)
    bad = np.nonzero((lat_o != lat_e) & m)
    assert bad[0].size == 0, f"LattCoo differs at {bad[0].size} points, first (z,y,x)={[int(a[0]) for a in bad]}"


@pytest.mark.parametrize("key", ["surface", "finfet", "CVD-Si-SiH2"])
def test_phases_match_the_oracle(tmp_path, key):
    o, eng, box = grown(tmp_path, key)
    for layout in ("socket", "planes"):
        ce, de = eng.phases(layout)
        co, do = o.phases(layout)
        np.testing.assert_array_equal(ce, co)
        np.testing.assert_array_equal(de, do)
    cd, dd = eng.phases("socket", device=True)
    np.testing.assert_array_equal(cd.cpu().numpy(), o.phases("socket")[0])
    np.testing.assert_array_equal(dd.cpu().numpy(), o.phases("socket")[1])
    eng.close()


@pytest.mark.parametrize("deck,mode,box,lens", [("Si-SL", 0, (96, 96, 144), (72, 120, 120)),
                                                 ("surface", 1, (96, 96, 144), (72, 72, 72))])
def test_phases_of_the_device_builders(tmp_path, deck, mode, box, lens):
    """SetSi fills its wall planes at every grid point, on and off the
    diamond super-lattice; SetSiC3C has no walls.  The device-built slab's
    phase maps equal the oracle builder's everywhere, and survive a state
    copy into another engine."""
    path, d = small_deck(tmp_path, deck, lens)
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables())
    eng.build_slab(mode, lens[0] - lens[0] % 24)
    co, do = o.phases("socket")
    assert ((co == 10).sum() > 0) == (mode == 0)
    ce, de = eng.phases("socket")
    np.testing.assert_array_equal(ce, co)
    np.testing.assert_array_equal(de, do)
    other = Engine(box, d.tables())
    other.copy_state_from(eng)
    np.testing.assert_array_equal(other.phases("socket")[0], co)
    other.close()
    eng.close()


@pytest.mark.parametrize("key", ["surface", "finfet"])
def test_temperature_field_matches_the_oracle(tmp_path, key):
    o, eng, box = grown(tmp_path, key)
    lx, ly, lz = box
    t = self_naming_field(box) + 1000
    t[3, 5, :] = 10           # wall marker: the point keeps its temperature
    t[4, 6, 8] = 8192 + 300   # low 13 bits
    t[6, 2, 2] = -1
    for layout, f in (("planes", t), ("socket", to_socket(t[::-1].copy()))):
        o.set_temperature_field(f, layout)
        eng.set_temperature_field(f, layout)
        same_lattcoo(o, eng, box)
    # a device map, read in place
    t2 = np.ascontiguousarray(t[:, ::-1, :]) + 7
    o.set_temperature_field(t2, "planes")
    eng.set_temperature_field(torch.from_numpy(t2).cuda(), "planes")
    same_lattcoo(o, eng, box)
    # the maps do not change what the phase maps read
    np.testing.assert_array_equal(eng.phases("planes")[0], o.phases("planes")[0])
    eng.close()


LA = dict(tm=1687.0, p0=1e13, sigma=50.0, tflex=1650.0)


def la_deck(tmp_path):
    """The Si-SL regression deck, small box, with energies in its tables
    (as an LA start.dat holds them): melting 0.6 eV per bond, solidification
    0.55 eV per bond of the site."""
    base, box, lens = SMALL["Si-SL"]
    _, d = small_deck(tmp_path, base, lens)
    d.evap = [(sp, digits, 0.6 * int(digits)) for sp, digits, _ in d.evap]
    d.depo = [(sp, c, 0.55 * c) for sp, c, _ in d.depo]
    path = str(tmp_path / "start-la.dat")
    deckmod.write_start_dat(d, path)
    return path, d, box


def la_field(box, shift):
    lx, ly, lz = box
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    return (1500 + 2 * z + (x % 24) + 3 * ((y + shift) % 12)).astype(np.int32)


def test_la_cycle_replays_bit_exact(tmp_path):
    """One laser-annealing cycle twice over: temperature map in (socket
    layout), rates re-evaluated, 6000 oracle events replayed by the engine.
    The lattice (T bits included) equals the oracle's bit for bit and every
    MC particle's LA rate (Get_Prob.f:127-149) is within 1e-6 of the
    oracle's FP64 rate, after each cycle."""
    path, d, box = la_deck(tmp_path)
    o = Oracle(*box)
    o.load_deck(path)
    o.set_la(LA["tm"], LA["p0"], LA["sigma"], LA["tflex"])
    eng = Engine(box, d.tables(), la=LA)
    eng.load_state(*oracle_state(o))
    for cycle in range(2):
        f = to_socket(la_field(box, cycle))
        o.set_temperature_field(f, "socket")
        o.refresh_all()
        eng.set_temperature_field(f, "socket")
        eng.refresh_rates()
        log = o.step(6000)
        print(f"[LA cycle {cycle}] replaying {len(log)} events", flush=True)
        eng.replay(log)
        same_lattcoo(o, eng, box)
        mc, _, _ = o.lists()
        cls, rate = eng.site_rates(mc["xyz"])
        ev = mc["event"]
        np.testing.assert_array_equal(np.where(cls == 6, 4, np.where(cls == 2, 1, cls)),
                                      np.where(ev == 4, 4, np.where(ev == 5, 5, 1)))
        keep = cls != 6
        rel = np.abs(rate[keep].astype(np.float64) - mc["prob"][keep]) / mc["prob"][keep]
        assert rel.max() <= 1e-6, rel.max()
        c_e, d_e = eng.phases("socket")
        c_o, d_o = o.phases("socket")
        np.testing.assert_array_equal(c_e, c_o)
        np.testing.assert_array_equal(d_e, d_o)
    eng.close()


VARIANT = dict(p0=(1e13, 2e13), tm=(1687.0, 1600.0), sigma=(50.0, 70.0), tflex=(1650.0, 1580.0))


def variant_pair(tmp_path):
    """Oracle and engine on the SiC deck with energy tables, in the
    laser-annealing build's form (per-species constants, liquid fraction)."""
    from tests.test_oracle_la_variant import field as vfield, la_energy_deck, species_counts
    path, d, box = la_energy_deck(tmp_path, "surface")
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables(), la=LA)
    eng.load_state(*oracle_state(o))
    f = vfield(box)
    o.set_temperature_field(f, "socket")
    eng.set_temperature_field(f, "socket")
    cc = species_counts(o)
    init = [cc[0] + 50000, cc[1] + 30000]  # enough liquid that neither runs out
    o.set_la_variant(VARIANT["p0"], VARIANT["tm"], VARIANT["sigma"], VARIANT["tflex"], init, cc, False)
    eng.set_la_variant(VARIANT["p0"], VARIANT["tm"], VARIANT["sigma"], VARIANT["tflex"], init, cc)
    o.refresh_all()
    eng.refresh_rates()
    return o, eng, box


def test_la_variant_replays_bit_exact(tmp_path):
    """The laser-annealing build's rate form (mulskips-source-LaserAnnealing/
    mcevent/Get_Prob.f:140-153): 8000 oracle events replayed by the engine;
    lattice bit for bit, every rate within 1e-6, and the same CountCrystal /
    CountLiquid (refreshed per event, as that build's main.f:637 does)."""
    o, eng, box = variant_pair(tmp_path)
    log = o.step(8000)
    print(f"[LA variant] replaying {len(log)} events, deposits by species "
          f"{np.bincount(log['itrans'], minlength=5)[1:3]}", flush=True)
    eng.replay(log)
    same_lattcoo(o, eng, box)
    mc, _, _ = o.lists()
    cls, rate = eng.site_rates(mc["xyz"])
    keep = cls != 6
    rel = np.abs(rate[keep].astype(np.float64) - mc["prob"][keep]) / mc["prob"][keep]
    assert rel.max() <= 1e-6, rel.max()
    c_o, q_o = o.la_counts()
    c_e, q_e = eng.la_counts()
    np.testing.assert_array_equal(c_e[:2], c_o[:2])
    np.testing.assert_array_equal(q_e[:2], q_o[:2])
    eng.close()


def test_la_variant_parallel_run_counts(tmp_path):
    """mskg_run in the same form: the crystal counts move by the deposits
    and evaporations the run counts, and CountLiquid stays consistent."""
    o, eng, box = variant_pair(tmp_path)
    c0, q0 = eng.la_counts()
    n, t, stop = eng.run(20000)
    counts = eng.counters(lists=False)[0]
    c1, q1 = eng.la_counts()
    print(f"[LA variant run] {n} events, stop {stop}, counts {c0} -> {c1}", flush=True)
    assert n >= 20000
    assert int((c1 - c0).sum()) == int(counts[0] + counts[1] + counts[2] - counts[3])
    np.testing.assert_array_equal(c1 + q1, c0 + q0)  # InitCountCrystal is fixed
    assert (q1[:2] >= 0).all()
    eng.close()


def test_la_variant_needs_la_mode(tmp_path):
    base, box, lens = SMALL["surface"]
    _, d = small_deck(tmp_path, base, lens)
    eng = Engine(box, d.tables())
    eng.build_slab(1, 72)
    with pytest.raises(EngineError, match="la_mode"):
        eng.set_la_variant(*VARIANT.values(), [1, 1], [0, 0])
    eng.close()


def test_st_tables_replay_bit_exact(tmp_path):
    """InitSt 'ST' (InterpProbFromField.f:22-57 at each site's temperature):
    reference tables at 1400 / 1500 / 1600 K, a temperature field spread
    over 1410-1590 K; 6000 oracle events replayed by the engine bit for bit,
    every rate within 1e-6 of the oracle's, then a parallel run in the same
    mode."""
    from tests.test_oracle_st import TEMPS, scaled
    base, box, lens = SMALL["Si-SL"]
    path, d = small_deck(tmp_path, base, lens)
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables())
    eng.load_state(*oracle_state(o))
    lx, ly, lz = box
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    f = (1410 + (z * 7 + x * 3 + y) % 181).astype(np.int32)
    tab = d.tables()
    refs = [scaled(tab, 0.5), tab, scaled(tab, 2.0)]
    o.set_temperature_field(f, "planes")
    eng.set_temperature_field(f, "planes")
    o.set_st_tables(TEMPS, refs)
    eng.set_st_tables(TEMPS, refs)
    o.refresh_all()
    eng.refresh_rates()
    log = o.step(6000)
    print(f"[ST] replaying {len(log)} events", flush=True)
    eng.replay(log)
    same_lattcoo(o, eng, box)
    mc, _, _ = o.lists()
    cls, rate = eng.site_rates(mc["xyz"])
    keep = cls != 6
    rel = np.abs(rate[keep].astype(np.float64) - mc["prob"][keep]) / mc["prob"][keep]
    assert rel.max() <= 1e-6, rel.max()
    n, t, stop = eng.run(20000)
    print(f"[ST run] {n} events, stop {stop}", flush=True)
    # the deck evaporates at these temperatures: the reference's few-MC stop may come first
    assert n > 5000 and stop in (0, 1)
    eng.close()


@pytest.mark.parametrize("mode", ["parallel", "serial"])
@pytest.mark.parametrize("model", ["st", "st-cvd", "la"])
def test_runs_are_serial_histories(tmp_path, model, mode):
    """InitSt 'ST' tables (on the Si-SL deck, and on the CVD deck, whose
    adsorption and desorption tables are interpolated too), and the main
    build's laser-annealing rates from a temperature map, in the
    synchronous-sublattice engine (mskg_run) and in
    the serial kernel (mskg_run_serial): the engine's own event log, applied
    by the oracle with the same tables and temperature field, reproduces the
    engine's lattice bit for bit and its rates within 1e-6."""
    if model.startswith("st"):
        from tests.test_oracle_st import TEMPS, scaled
        base, box, lens = SMALL["CVD-Si-SiH2" if model == "st-cvd" else "Si-SL"]
        path, d = small_deck(tmp_path, base, lens)
        la = None
    else:
        path, d, box = la_deck(tmp_path)
        la = LA
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables(), seed=7, tile=24, events_per_pass=1.0, la=la)
    eng.load_state(*oracle_state(o))
    if model.startswith("st"):
        lx, ly, lz = box
        z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
        f = (1460 + (z * 5 + x + 2 * y) % 121).astype(np.int32)
        tab = d.tables()
        refs = [scaled(tab, 0.5), tab, scaled(tab, 2.0)]
        for side in (o, eng):
            side.set_temperature_field(f, "planes")
            side.set_st_tables(TEMPS, refs)
    else:
        o.set_la(LA["tm"], LA["p0"], LA["sigma"], LA["tflex"])
        f = to_socket(la_field(box, 0))
        for side in (o, eng):
            side.set_temperature_field(f, "socket")
    o.refresh_all()
    eng.refresh_rates()
    eng.set_event_log(200000)
    n, t, stop = eng.run(15000) if mode == "parallel" else eng.run_serial(15000)
    log = eng.event_log()
    print(f"[{model} {mode}] {n} events, {len(log)} logged, stop {stop}", flush=True)
    assert len(log) == n and n > 3000 and stop in (0, 1)
    o.apply_log(log)
    same_lattcoo(o, eng, box)
    mc, _, _ = o.lists()
    cls, rate = eng.site_rates(mc["xyz"])
    keep = cls != 6
    np.testing.assert_allclose(rate[keep], mc["prob"][keep], rtol=1e-6)
    eng.close()


def test_st_tables_set_after_a_run(tmp_path):
    """A run captures the sweep; InitSt 'ST' tables set after it switch the
    next run to the interpolated rates: the oracle applies the first run's
    events with the plain tables and the second's with the same ST tables,
    and ends with the engine's lattice and rates."""
    from tests.test_oracle_st import TEMPS, scaled
    base, box, lens = SMALL["Si-SL"]
    path, d = small_deck(tmp_path, base, lens)
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables(), seed=5)
    eng.load_state(*oracle_state(o))
    eng.set_event_log(200000)
    n1, _, stop1 = eng.run(3000)
    log1 = eng.event_log()
    assert len(log1) == n1 and n1 >= 3000 and stop1 in (0, 1)
    o.apply_log(log1)
    lx, ly, lz = box
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    f = (1460 + (z * 5 + x + 2 * y) % 121).astype(np.int32)
    tab = d.tables()
    refs = [scaled(tab, 0.5), tab, scaled(tab, 2.0)]
    for side in (o, eng):
        side.set_temperature_field(f, "planes")
        side.set_st_tables(TEMPS, refs)
    o.refresh_all()
    eng.refresh_rates()
    n2, _, stop2 = eng.run(10000)
    log = eng.event_log()
    print(f"[ST after a run] {n1} + {n2} events, stop {stop2}", flush=True)
    assert len(log) == n1 + n2 and n2 > 3000 and stop2 in (0, 1)
    np.testing.assert_array_equal(log[:n1], log1)
    o.apply_log(log[n1:])
    same_lattcoo(o, eng, box)
    mc, _, _ = o.lists()
    cls, rate = eng.site_rates(mc["xyz"])
    keep = cls != 6
    np.testing.assert_allclose(rate[keep], mc["prob"][keep], rtol=1e-6)
    eng.close()


def test_la_variant_set_after_a_run(tmp_path):
    """A run in the main build's laser-annealing form captures the sweep; the
    laser-annealing build's form set after it, with no liquid of species 2
    (which hardly melts: 2.5 eV bonds), keeps every deposition of the next
    run to species 1 (the main form would deposit both)."""
    from types import SimpleNamespace
    from tests.test_oracle_la_variant import field as vfield, la_energy_deck, species_counts
    path, d, box = la_energy_deck(tmp_path, "surface", e_pair=((0.6, 0.8), (2.5, 2.5)))
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables(), la=LA)
    eng.load_state(*oracle_state(o))
    eng.set_temperature_field(vfield(box), "socket")
    eng.refresh_rates()
    n1, _, _ = eng.run(3000)
    dep1 = eng.counters(lists=False)[0][:3].copy()
    assert dep1[0] > 0 and dep1[1] > 0, dep1
    cc = species_counts(SimpleNamespace(lattcoo=lambda: eng.store_state()[0]))
    eng.set_la_variant(VARIANT["p0"], VARIANT["tm"], VARIANT["sigma"], VARIANT["tflex"],
                       [cc[0] + 20000, cc[1]], cc)
    eng.refresh_rates()
    n2, _, stop = eng.run(10000)
    dep2 = eng.counters(lists=False)[0][:3] - dep1
    evap, _ = eng.species_counts()
    print(f"[LA variant after a run] {n1} + {n2} events, stop {stop}, deposits {dep2}, "
          f"evaporated {evap}", flush=True)
    assert n2 >= 10000 and dep2[0] > 1000
    assert dep2[1] <= evap[1]
    eng.close()


def test_strict_temperature_rejects_and_keeps(tmp_path):
    o, eng, box = grown(tmp_path, "surface", n=500)
    lx, ly, lz = box
    full = np.full((lz + 1, ly, lx), 1700, np.int32)
    eng.set_temperature(full)
    before = eng.store_state()[0]
    bad = full.copy()
    bad[10, 4, 4] = 9000  # a lattice point (all even): beyond 13 bits
    with pytest.raises(EngineError, match="13-bit"):
        eng.set_temperature(bad)
    np.testing.assert_array_equal(eng.store_state()[0], before)
    eng.close()
