"""GPU: the two orders of a sweep's colour passes (mskg_set_schedule).

* MSKG_SCHEDULE_TILES (one launch per sweep, each tile after its lower-
  colour neighbours) and MSKG_SCHEDULE_COLOURS (a barrier after each colour)
  give the same lattice, lists, counters and clock bit for bit from the same
  seed: the SiC model with the second-neighbour term at tile 24 and k = 4
  (the benchmark's model), Si PVD at tiles 16 and 48, and CVD with coverage
  species;
* the tile schedule's event log, replayed serially through the oracle,
  reproduces the engine's lattice (its order is a valid serial history);
* switching the schedule between runs of one engine changes nothing.
"""
import numpy as np
import pytest

from mulskips_amd.autotune import tile_capacity
from mulskips_amd.engine import Engine
from oracle.oracle import Oracle
from tests.helpers import list_map, oracle_state, parity_mask, small_deck

pytestmark = pytest.mark.gpu

F2 = np.exp(-np.arange(13) * 0.05 / (8.617333262e-5 * 1600.0))

CASES = {  # name: (deck, builder mode (None: the deck's own start), box, lens, tile, k, second factor)
    "SiC-f2": ("surface", 1, (192, 192, 144), (72, 72, 72), 24, 4.0, True),
    "Si-T16": ("Si-SL", 0, (96, 96, 144), (72, 120, 120), 16, 4.0, False),
    "Si-T48": ("Si-SL", 0, (96, 96, 144), (72, 120, 120), 48, 1.0, False),
    "CVD": ("CVD-Si-SiH2", None, (96, 96, 192), (72, 120, 120), 24, 4.0, False),
}


def make(tmp_path, case, schedule, seed=29):
    deck, mode, box, lens, tile, k, f2 = CASES[case]
    path, d = small_deck(tmp_path, deck, lens)
    tab = d.tables()
    if f2:
        tab["second_factor"] = F2
    eng = Engine(box, tab, seed=seed, tile=tile, tile_capacity=tile_capacity(tile), events_per_pass=k)
    eng.set_schedule(schedule)
    if mode is None:
        o = Oracle(*box)
        o.load_deck(path)
        if f2:
            o.set_second_factor(F2)
        eng.load_state(*oracle_state(o))
        o.close()
    else:
        eng.build_slab(mode, lens[0] - lens[0] % 24)
    return eng, path, box, f2


def snapshot(eng):
    lat, mc, bulk, void = eng.store_state()
    counts, clock, lists = eng.counters()
    return lat, list_map(mc), counts, clock, lists


@pytest.mark.parametrize("case", list(CASES))
def test_schedules_agree_bit_for_bit(tmp_path, case):
    out = {}
    for sch in ("colours", "tiles", "mixed"):
        eng, _, box, _ = make(tmp_path, case, "colours" if sch == "colours" else "tiles")
        r1 = eng.run(60000)
        if sch == "mixed":  # the second run with the barrier schedule
            eng.set_schedule("colours")
        r2 = eng.run(30000)
        print(f"[{case} {sch}] runs {r1} {r2}", flush=True)
        out[sch] = (snapshot(eng), [r1, r2], eng.diagnostics())
        eng.close()
    (la, ma, ca, ta, na), ra, da = out["colours"]
    # the Si deck's film can run down to the reference's few-MC stop
    assert ra[0][0] >= 10000 and ra[0][2] in (0, 1)
    for sch in ("tiles", "mixed"):
        (lb, mb, cb, tb, nb), rb, db = out[sch]
        assert ra == rb, sch
        assert np.array_equal(la, lb), f"LattCoo differs ({sch})"
        assert ma == mb, f"MC lists differ ({sch})"
        assert np.array_equal(ca, cb) and ta == tb and na == nb, sch
        assert da == db, sch


def test_tile_schedule_log_is_a_serial_history(tmp_path):
    """The deck's own start is the device-built slab (the builders agree
    site by site), so the oracle replays the log from its loaded deck."""
    eng, path, box, f2 = make(tmp_path, "SiC-f2", "tiles", seed=3)
    o = Oracle(*box)
    o.load_deck(path)
    o.set_second_factor(F2)
    eng.set_event_log(400000)
    n, t, stop = eng.run(80000)
    log = eng.event_log()
    print(f"[tiles log] {n} events, {len(log)} logged, stop={stop}", flush=True)
    assert len(log) == n and n >= 80000
    o.apply_log(log)
    assert o.anomalies()["oob_reads"] == 0
    lat_o, mc_o, _, _ = oracle_state(o)
    lat_e, mc_e, _, _ = eng.store_state(lattcoo=lat_o.copy())
    m = parity_mask(box)
    assert not np.any((lat_o != lat_e) & m), "LattCoo differs from the serial replay"
    assert list_map(mc_o) == list_map(mc_e)
    o.close()
    eng.close()


def test_schedules_agree_in_laser_annealing_mode():
    """The laser-annealing kernels (the bench's melt-front leg runs the tile
    schedule): SiGe x=0.3 on the 264 x 264 x 384 grid under the tanh front,
    two temperature exchanges, the same state and clock from both orders."""
    from mulskips_amd import models
    tab, _ = models.la_sige_tables(2026)
    box = (264, 264, 384)
    out = {}
    for sch in ("colours", "tiles"):
        eng = Engine(box, tab, seed=2026, tau_policy=1, events_per_pass=4.0, la=models.LA_CONSTANTS)
        eng.build_alloy_slab(240, 0.3)
        eng.set_schedule(sch)
        runs = []
        for zm in (240.0, 243.0):
            eng.set_temperature_field(models.la_front_field(box, zm), "socket")
            eng.refresh_rates()
            runs.append(eng.run(50000))
        print(f"[LA {sch}] runs {runs}", flush=True)
        out[sch] = (snapshot(eng), runs)
        eng.close()
    (la, ma, ca, ta, na), ra = out["colours"]
    (lb, mb, cb, tb, nb), rb = out["tiles"]
    assert ra == rb and ra[0][0] >= 50000
    assert np.array_equal(la, lb) and ma == mb
    assert np.array_equal(ca, cb) and ta == tb and na == nb
