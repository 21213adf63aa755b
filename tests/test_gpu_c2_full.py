"""The SiC benchmark configuration at its full size: 512^3 diamond sites
(a 3072^3 grid, 7.25e9 stored points), a device-built 3C-SiC slab with a
40-monolayer frozen substrate, the second-neighbour model at 1600 K.

The oracle cannot hold this lattice (its grid alone is 232 GB), so the test
checks properties that need no reference run: two runs with the same seed
from the same state give the same event count, clock, counters and the
same frame records (every surface atom, vacancy and non-epitaxial atom,
compacted on the device), and the counters account for every event."""
import hashlib

import pytest

from mulskips_amd import models
from mulskips_amd.engine import Engine

pytestmark = pytest.mark.gpu
BOX = (3072, 3072, 3072)
SUBSTRATE = 120  # 40 monolayers


def run_once():
    eng = Engine(BOX, models.sic_c2_tables(), seed=20261202, tile=24, events_per_pass=4.0, tau_policy=1,
                 frozen_below_z=SUBSTRATE + 1)
    eng.build_slab(1, SUBSTRATE)
    n, t, stop = eng.run(2_000_000)
    counts, t2, _ = eng.counters(lists=False)
    main, voids, noepi, nat, ncov = eng.frame()
    h = hashlib.sha256()
    for a in (main, voids, noepi):
        h.update(a.tobytes())
    eng.close()
    print(f"[sic 512^3] {n} events, t={t:.6g}, {len(main)} frame records, {nat} undercoordinated atoms", flush=True)
    return {"n": n, "t": t, "stop": stop, "counts": counts.tolist(), "t2": t2, "nat": nat, "ncov": ncov,
            "records": (len(main), len(voids), len(noepi)), "digest": h.hexdigest()}


def test_sic_full_size_rerun_is_identical():
    a = run_once()
    assert a["stop"] == 0 and a["n"] >= 2_000_000 and a["t"] == a["t2"]
    assert sum(a["counts"]) == a["n"]  # every event is counted under its type
    assert a["nat"] > 131_072  # at least the slab's top monolayer (256^2 cells x 2 atoms)
    b = run_once()
    assert a == b
