"""GPU: batched serial KMC over independent lattices (mskg_run_serial) and the
random-alloy device builder.

* batching does not change a replica's trajectory: replicas of different
  boxes, models and seeds give identical event logs and lattices when run
  alone or together in one launch;
* the serial kernel is the reference algorithm: its event log, replayed by
  the oracle from the same device-built slab, reaches the same lattice
  (also covered for the SiC model by test_adjoint);
* the alloy slab has the SetSi structure with species 2 on a fraction of
  the sites within 5 sigma of the requested one.
"""
import numpy as np
import pytest

from mulskips_amd import models
from mulskips_amd.engine import Engine, run_serial
from oracle.oracle import Oracle
from tests.helpers import list_map, oracle_state, parity_mask, small_deck

pytestmark = pytest.mark.gpu


def replicas():
    si = models.si_c1_tables(temperature=800.0, e_bond=0.9)
    sic = models.sic_c2_tables()
    sige = models.pair_energy_tables([[0.55, 0.5], [0.5, 0.45]], [0.7, 0.3], 1100.0, 1e13, (14, 32))
    return [((96, 96, 144), si, 0, 48, 3, 0.0), ((144, 144, 144), sic, 1, 48, 5, 0.0),
            ((48, 48, 96), sige, 2, 30, 9, 0.3)]


def make(box, tab, mode, length, seed, frac):
    e = Engine(box, tab, seed=seed, tile=24, frozen_below_z=length - 12)
    if mode == 2:
        e.build_alloy_slab(length, frac)
    else:
        e.build_slab(mode, length)
    e.refresh_rates()
    e.set_event_log(30000)
    return e


def test_batch_equals_single_runs():
    reps = replicas()
    single = []
    for r in reps:
        e = make(*r)
        n, t, stop = e.run_serial(20000, max_time=1e300)
        single.append((e.event_log(), e.store_state()[0]))
        e.close()
    engs = [make(*r) for r in reps]
    n, t, stop = run_serial(engs, [20000] * len(engs))
    assert np.all(n == 20000)
    for (log, lat), e in zip(single, engs):
        log_b = e.event_log()
        assert len(log_b) == len(log)
        for f in ("site", "itrans", "arm", "coor"):
            np.testing.assert_array_equal(log_b[f], log[f])
        np.testing.assert_array_equal(log_b["time"], log["time"])
        np.testing.assert_array_equal(e.store_state()[0], lat)
        e.close()


def test_serial_run_is_the_reference_algorithm(tmp_path):
    """A serial run on a device-built Si slab, replayed by the oracle."""
    path, d = small_deck(tmp_path, "Si-SL", (72, 120, 120))
    box = (96, 96, 144)
    tab = d.tables()
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, tab, seed=31, tile=24)
    eng.build_slab(0, 72)
    eng.set_event_log(50000)
    n, t, stop = eng.run_serial(20000)
    log = eng.event_log()
    assert len(log) == n
    o.apply_log(log)
    lat_o, mc_o, bulk_o, void_o = oracle_state(o)
    lat_e, mc_e, bulk_e, void_e = eng.store_state(lattcoo=lat_o.copy())
    assert not np.any((lat_o != lat_e) & parity_mask(box))
    for a, b in ((mc_o, mc_e), (bulk_o, bulk_e), (void_o, void_e)):
        assert list_map(a) == list_map(b)
    eng.close()


def test_alloy_slab():
    box = (96, 96, 144)
    tab = models.pair_energy_tables([[0.55, 0.5], [0.5, 0.45]], [0.7, 0.3], 1100.0, 1e13, (14, 32))
    plain = Engine(box, tab, seed=3)
    plain.build_slab(0, 72)
    lat0 = plain.store_state()[0]
    plain.close()
    for frac in (0.0, 0.3, 1.0):
        e = Engine(box, tab, seed=3, frozen_below_z=36)  # the alloy sublimates at 1100 K: keep a substrate
        e.build_alloy_slab(72, frac)
        lat = e.store_state()[0]
        occ = ((lat >> 23) & 1) == 1
        wall = (lat & 255) == 113
        np.testing.assert_array_equal(occ, ((lat0 >> 23) & 1) == 1)
        np.testing.assert_array_equal((lat >> 24) & 7, (lat0 >> 24) & 7)  # coordination
        sp = (lat >> 27) & 7
        atoms = occ & ~wall
        f = np.mean(sp[atoms] == 2)
        n = atoms.sum()
        assert abs(f - frac) <= 5 * np.sqrt(frac * (1 - frac) / n) + 1e-12, (frac, f)
        e.refresh_rates()
        m, t, stop = e.run_serial(5000)
        assert m == 5000
        e.close()
