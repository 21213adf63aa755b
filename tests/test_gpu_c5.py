"""GPU: the calibration configuration at the benchmark's size.

A 10^6-event trajectory of the SiC second-neighbour model on the 144^3 grid
(frozen substrate, 1600 K), recorded by the engine's serial kernel with a
lattice checkpoint every 10^5 events and replayed segment by segment by the
adjoint kernel in one launch, with class-population checkpoints every 10^4
events, against the oracle replaying the whole log from the same start:

* S (selections per class) equal exactly, A (time integrals of the class
  populations) to 1e-10 relative, the 100 class-population checkpoints
  exactly;
* the 8-parameter log-likelihood gradients agree to 1e-7 relative (terms of
  ~1e4 cancel to components of ~1, so A's 1e-10 grows there to ~5e-9);
* the oracle's replay reaches the engine's final lattice (the serial
  kernel's trajectory is a valid history).
"""
import os

import numpy as np
import pytest

from mulskips_amd import adjoint, deck as deckmod

pytestmark = pytest.mark.gpu

BOX = (144, 144, 144)
SLAB = 48
FROZEN = 36
TEMP = 1600.0
N_EVENTS = 1_000_000
SEGMENT = 100_000
EVERY = 10_000


def test_c5_trajectory_against_oracle(tmp_path):
    from mulskips_amd.engine import Engine
    from oracle.oracle import Oracle
    from tests.helpers import list_map, oracle_state, parity_mask
    theta0 = adjoint.theta_sic()
    tab = adjoint.tables_from_theta(theta0, TEMP)

    def make(seed):
        e = Engine(BOX, tab, seed=seed if seed is not None else 1, tile=24, frozen_below_z=FROZEN)
        if seed is not None:
            e.build_slab(1, SLAB)
            e.refresh_rates()
        return e

    logs, dts, segments = adjoint.record_segmented(make, [501], N_EVENTS, SEGMENT)
    log, dt = logs[0], dts[0]
    assert len(log) == N_EVENTS and len(segments[0]) == N_EVENTS // SEGMENT
    print(f"[C5] recorded {len(log)} events, KMC time {float(np.sum(dt)):.4g} s", flush=True)
    final = segments[0][-1][0]  # the last segment's engine ends in the trajectory's final state
    per_traj, ck = adjoint.segmented_stats(segments, every=EVERY)
    S, A, sc = per_traj[0]
    hist = ck[0]
    assert hist.shape == (N_EVENTS // EVERY, adjoint.KEY_N)
    # the oracle replays the whole log from the same device-built start
    d = deckmod.deck_from_tables(tab, "F", [SLAB, 120, 120], idum=4242)
    path = os.path.join(str(tmp_path), "start.dat")
    deckmod.write_start_dat(d, path)
    o = Oracle(*BOX)
    o.load_deck(path)
    o.set_frozen(FROZEN)
    o.set_second_factor(tab["second_factor"])
    o.refresh_all()
    S_o, A_o, sc_o, hist_o = o.adjoint_stats(log, dt, every=EVERY)
    print(f"[C5] oracle replayed the log; S {int(S_o.sum())} selections", flush=True)
    np.testing.assert_array_equal(S, S_o)
    np.testing.assert_allclose(A, A_o, rtol=1e-10, atol=1e-300)
    np.testing.assert_array_equal(hist, hist_o)
    assert sc[3] == N_EVENTS
    g_o = adjoint.gradient(theta0, adjoint.TrajectoryStats(S_o, A_o, sc_o), TEMP)[1]
    g_e = adjoint.gradient(theta0, adjoint.TrajectoryStats(S, A, sc), TEMP)[1]
    np.testing.assert_allclose(g_e.numpy(), g_o.numpy(), rtol=1e-7)
    # the statistics are not vacuous: both deposited species, evaporations
    assert S[adjoint.KEY_N:adjoint.KEY_N + 3].sum() > 0 and S[adjoint.KEY_N + 3:adjoint.KEY_N + 6].sum() > 0
    assert S[:adjoint.KEY_ADS].sum() > 0
    # the oracle's replay reaches the engine's final lattice
    lat_o, mc_o, _, _ = oracle_state(o)
    lat_e, mc_e, _, _ = final.store_state(lattcoo=lat_o.copy())
    m = parity_mask(BOX)
    assert not np.any((lat_o != lat_e) & m)
    assert list_map(mc_o).keys() == list_map(mc_e).keys()
    o.close()
    for traj in segments:
        for c, _, _ in traj:
            c.close()
