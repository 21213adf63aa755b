"""Statistical parity of the synchronous-sublattice engine with serial KMC.

The parallel engine is not a replay of the serial algorithm: tiles of one
colour advance concurrently over a window tau, with Philox draws.  Its
trajectories must agree in distribution with the reference's serial BKL
(the oracle, RunType 'R', ran3).  For each case 32 independent seeds of both
are run to the same KMC time from the same initial lattice, and the means of
the film observables must agree within 2 standard errors of their
difference:
* growth (atoms), RMS roughness, mean height and pymulskips' surface height
  (pymulskips/analyze.py:124-207, whose change over the run is its growth
  rate): Si PVD on a flat (001) slab with a frozen substrate, E = 1.0 eV per
  bond at 700 K, PtransZig 0.5;
* growth and stacking-fault (non-cubic site) fraction: the cube-zigzag
  regression deck (PtransZig 0.8) on a 144^3 grid;
* growth and coverage: the CVD-Si-SiH2 regression deck (H coverage, also
  as pymulskips' get_coverage, analyze.py:418-483, and its surface height);
* the benchmark's SiC model with its 0.05 eV second-neighbour term, at the
  benchmark's tile edge 24 and 4 events per tile and pass: at 1600 K on a
  frozen SiC substrate (the benchmark's own steady state: adatom population
  and deposits by species) and at 900 K on the SiC parallelepiped of the
  cube-zigzag deck (growth, roughness, stacking).
Both window sizes the engine is run with are covered: about 1 and 4 events
per tile and colour pass (tau policy 1).  The event count is asserted like
the observables.  The engine's last window ends at the time limit, so both
sides stop at the same KMC time.  The seeds are fixed and both sides repeat
a run bit for bit, so the test is deterministic.

A case compares several observables, and the eight cases hold about 35
comparisons, so a correct engine puts one of them outside 2 standard
errors in most realisations.  The test is therefore sequential: an
observable outside 2 standard errors at 32 seeds is measured again on 128
fresh seeds of both sides and must lie within 2 standard errors there.  A
real bias that reaches 2 standard errors at 32 seeds reaches about 4 at 128.
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


def sic2_deck(tmp_path, name):
    """The SiC model of the large-replica benchmark (pair energies Si-Si 0.6 /
    Si-C 1.1 / C-C 0.9 eV, 0.05 eV per occupied second-neighbour position,
    Si and C fluxes 2 and 1.8 ML/s):
    sic2-bench: at its 1600 K on a flat SiC slab over a frozen substrate --
      every surface class evaporates faster than the flux arrives, so the
      run is the flicker of adatoms on the frozen surface that the benchmark
      times: its deposits of each species and its evaporations;
    sic2-growth: at 900 K, where it grows, on the cube-zigzag deck's SiC
      parallelepiped (PtransZig 0.8): growth, roughness and stacking."""
    if name == "sic2-bench":
        tab = models.sic_c2_tables()
        d = deckmod.deck_from_tables(tab, "F", [72, 120, 120])
    else:
        tab = models.sic_c2_tables(temperature=900.0)
        base, box, lens = SMALL["cube-zigzag"]
        d = deckmod.deck_from_tables(tab, "C", list(lens), sf=0.38583333333333333)
        d.ptrans_zig = 0.8
    tab["ptrans_zig"] = d.ptrans_zig
    path = os.path.join(str(tmp_path), f"start-{name}.dat")
    deckmod.write_start_dat(d, path)
    return path, d, tab


def case(tmp_path, name):
    """(deck path, engine tables, box, z0, frozen, t_end, super-lattice step,
    observables)"""
    if name == "si-pvd":
        path, d = si_pvd_deck(tmp_path)
        return path, d.tables(), (96, 96, 240), 36, 36, 60.0, d.sf, ("atoms", "height", "roughness", "surface")
    if name == "sic2-bench":
        path, d, tab = sic2_deck(tmp_path, name)
        return path, tab, (144, 144, 240), 60, 61, 4.0, d.sf, ("dep1", "dep2", "evap")
    if name == "sic2-growth":
        path, d, tab = sic2_deck(tmp_path, name)
        return path, tab, SMALL["cube-zigzag"][1], 0, 0, 4.0, d.sf, ("atoms", "species1", "height", "roughness",
                                                                      "defects")
    base, box, lens = SMALL[name]
    path, d = small_deck(tmp_path, base, lens)
    if name == "cube-zigzag":
        return path, d.tables(), box, 0, 0, 3.0e5, d.sf, ("atoms", "defects")
    if name == "CVD-Si-SiH2":
        return path, d.tables(), box, 0, 0, 5.0, d.sf, ("atoms", "coverage", "roughness", "surface", "cov_h")
    raise KeyError(name)


def serial_runs(path, box, t_end, z0, frozen, seeds, f2=None, sf=None):
    ctx = mp.get_context("spawn")
    with ctx.Pool(min(8, len(seeds))) as pool:
        return pool.map(oracle_observables, [(path, box, s, t_end, z0, frozen, f2, sf) for s in seeds])


def parallel_runs(path, tab, box, t_end, z0, frozen, epp, seeds, sf=None):
    o = Oracle(*box)
    o.load_deck(path)
    if frozen:
        o.set_frozen(frozen)
    lat0, mc0, bulk0, void0 = oracle_state(o)
    o.close()
    out = []
    for s in seeds:
        eng = Engine(box, tab, seed=s, tile=24, tau_policy=1, events_per_pass=epp, frozen_below_z=frozen)
        eng.load_state(lat0, mc0, bulk0, void0)
        if tab.get("second_factor") is not None:
            eng.refresh_rates()  # the listed atoms' rates with the second-neighbour factor
        n, t, stop = eng.run(10 ** 12, max_time=t_end)
        # the last window ends at the time limit: the engine stops exactly there
        assert stop == 0 and t == t_end, (stop, t)
        assert eng.diagnostics()["guard_exits"] == 0
        lat, _, _, _ = eng.store_state()
        r = observables.film(lat, box, z0, sf=sf)
        r["events"] = n
        counts = eng.counters(lists=False)[0]
        r["dep1"], r["dep2"], r["evap"] = int(counts[0]), int(counts[1]), int(counts[3])
        out.append(r)
        eng.close()
        if len(out) % 32 == 0:
            print(f"  engine runs: {len(out)} of {len(seeds)}", flush=True)
    return out, lat0


def report(title, res):
    print(title, flush=True)
    for k, (ma, mb, dd, bound, ok) in res.items():
        print(f"  {k:10s} {ma:14.5g} {mb:14.5g}  diff {dd:+.4g}  2se {bound:.4g}  {'ok' if ok else 'OUT'}", flush=True)


def check_statistics(tmp_path, name, epp):
    path, tab, box, z0, frozen, t_end, sf, keys = case(tmp_path, name)
    keys = keys + ("events",)
    f2 = tab.get("second_factor")
    # serial reference: the oracle, one process per seed; the engine from the same initial lattice
    serial = serial_runs(path, box, t_end, z0, frozen, [1001 + 2 * s for s in range(SEEDS)], f2, sf)
    parallel, lat0 = parallel_runs(path, tab, box, t_end, z0, frozen, epp, [7000 + s for s in range(SEEDS)], sf)
    res = observables.compare(parallel, serial, keys, nsigma=2.0)
    report(f"\n[{name}, {epp} events per tile and pass] t_end={t_end}, {SEEDS} seeds each (parallel vs serial):", res)
    out = [k for k in keys if not res[k][4]]
    if out:  # second stage on fresh seeds
        n2 = 4 * SEEDS
        serial2 = serial_runs(path, box, t_end, z0, frozen, [50001 + 2 * s for s in range(n2)], f2, sf)
        parallel2, _ = parallel_runs(path, tab, box, t_end, z0, frozen, epp, [17000 + s for s in range(n2)], sf)
        res2 = observables.compare(parallel2, serial2, tuple(out), nsigma=2.0)
        report(f"[{name}, {epp}] second stage, {n2} fresh seeds each, for {', '.join(out)}:", res2)
        for k in out:
            assert res2[k][4], f"{name}: {k} differs by more than 2 standard errors at 32 and at {n2} seeds: {res2[k]}"
    # the comparison is not vacuous: the films changed
    assert np.mean([r["atoms"] for r in serial]) != observables.film(lat0, box, z0)["atoms"]
    assert np.mean([r["events"] for r in serial]) > 1000


@pytest.mark.parametrize("epp", [1.0, 4.0])
@pytest.mark.parametrize("name", ["si-pvd", "cube-zigzag", "CVD-Si-SiH2"])
def test_sublattice_matches_serial_statistics(tmp_path, name, epp):
    check_statistics(tmp_path, name, epp)


@pytest.mark.parametrize("name", ["sic2-bench", "sic2-growth"])
def test_benchmark_sic_model_matches_serial_statistics(tmp_path, name):
    """The benchmark's SiC model with its second-neighbour term, at the
    benchmark's tile edge 24 and window of 4 events per tile and pass."""
    check_statistics(tmp_path, name, 4.0)
