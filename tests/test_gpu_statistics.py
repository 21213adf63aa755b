"""Statistical parity of the synchronous-sublattice engine with serial KMC.

The parallel engine is not a replay of the serial algorithm: tiles of one
colour advance concurrently over a window tau, with Philox draws.  Its
trajectories must agree in distribution with the reference's serial BKL
(the oracle, RunType 'R', ran3).  For each case 32 independent seeds of both
are run to the same KMC time from the same initial lattice, and the means of
the film observables must agree within 2 standard errors of their
difference:
* growth (atoms), RMS roughness and mean height: Si PVD on a flat (001) slab
  with a frozen substrate, E = 1.0 eV per bond at 700 K, PtransZig 0.5;
* growth and stacking-fault (non-cubic site) fraction: the cube-zigzag
  regression deck (PtransZig 0.8) on a 144^3 grid;
* growth and coverage: the CVD-Si-SiH2 regression deck (H coverage).
Both window sizes the engine is run with are covered: about 1 and 4 events
per tile and colour pass (tau policy 1).  The seeds are fixed, so the test
is deterministic.
"""
import multiprocessing as mp
import os

import numpy as np
import pytest

from mulskips_amd import deck as deckmod
from mulskips_amd import models, observables
from mulskips_amd.engine import Engine
from oracle.oracle import Oracle
from tests.helpers import SMALL, oracle_observables, oracle_state, small_deck

pytestmark = pytest.mark.gpu
SEEDS = 32


def si_pvd_deck(tmp_path):
    tab = models.pair_energy_tables([[1.0]], [1.0], 700.0, 1e13, (14,))
    tab["ptrans_zig"] = 0.5
    d = deckmod.deck_from_tables(tab, "FS", [48, 120, 120])
    path = os.path.join(str(tmp_path), "start-si-pvd.dat")
    deckmod.write_start_dat(d, path)
    return path, d


def case(tmp_path, name):
    """(deck path, deck, box, z0, frozen, t_end, observables)"""
    if name == "si-pvd":
        path, d = si_pvd_deck(tmp_path)
        return path, d, (96, 96, 240), 36, 36, 60.0, ("atoms", "height", "roughness")
    base, box, lens = SMALL[name]
    path, d = small_deck(tmp_path, base, lens)
    if name == "cube-zigzag":
        return path, d, box, 0, 0, 3.0e5, ("atoms", "defects")
    if name == "CVD-Si-SiH2":
        return path, d, box, 0, 0, 5.0, ("atoms", "coverage", "roughness")
    raise KeyError(name)


@pytest.mark.parametrize("epp", [1.0, 4.0])
@pytest.mark.parametrize("name", ["si-pvd", "cube-zigzag", "CVD-Si-SiH2"])
def test_sublattice_matches_serial_statistics(tmp_path, name, epp):
    path, d, box, z0, frozen, t_end, keys = case(tmp_path, name)
    # serial reference: the oracle, one process per seed
    ctx = mp.get_context("spawn")
    with ctx.Pool(min(8, SEEDS)) as pool:
        serial = pool.map(oracle_observables, [(path, box, 1001 + 2 * s, t_end, z0, frozen) for s in range(SEEDS)])
    # the engine from the same initial lattice
    o = Oracle(*box)
    o.load_deck(path)
    if frozen:
        o.set_frozen(frozen)
    lat0, mc0, bulk0, void0 = oracle_state(o)
    o.close()
    parallel = []
    for s in range(SEEDS):
        eng = Engine(box, d.tables(), seed=7000 + s, tile=24, tau_policy=1, events_per_pass=epp,
                     frozen_below_z=frozen)
        eng.load_state(lat0, mc0, bulk0, void0)
        n, t, stop = eng.run(10 ** 12, max_time=t_end)
        assert stop == 0 and t >= t_end
        lat, _, _, _ = eng.store_state()
        r = observables.film(lat, box, z0)
        r["events"] = n
        parallel.append(r)
        eng.close()
    res = observables.compare(parallel, serial, keys + ("events",), nsigma=2.0)
    print(f"\n[{name}, {epp} events per tile and pass] t_end={t_end}, {SEEDS} seeds each (parallel vs serial):",
          flush=True)
    for k, (ma, mb, dd, bound, ok) in res.items():
        print(f"  {k:10s} {ma:14.5g} {mb:14.5g}  diff {dd:+.4g}  2se {bound:.4g}  {'ok' if ok else 'OUT'}", flush=True)
    for k in keys:
        assert res[k][4], f"{name}: {k} differs by more than 2 standard errors: {res[k]}"
    # the comparison is not vacuous: the films changed
    assert np.mean([r["atoms"] for r in serial]) != observables.film(lat0, box, z0)["atoms"]
