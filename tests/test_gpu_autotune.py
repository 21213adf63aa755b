"""GPU: the launch-shape autotuner times real runs of the SiC benchmark model
(second-neighbour term, frozen substrate) on a 96 x 96 x 192 box, keeps the
fastest candidate and answers a second call from its cache."""
import pytest

from mulskips_amd import models
from mulskips_amd.autotune import tune

pytestmark = pytest.mark.gpu


def test_tune_sic_small(tmp_path):
    tab = models.sic_c2_tables()
    cache = str(tmp_path / "tune.json")
    kw = dict(tau_policy=1, frozen_below_z=49, seed=3)
    out = tune((96, 96, 192), tab, lambda e: e.build_slab(1, 48), events_per_pass=(1.0, 4.0), warmup=2000,
               events=20000, cache=cache, **kw)
    print(f"[tune] {out['trials']}", flush=True)
    rates = [t["events_per_s"] for t in out["trials"] if "events_per_s" in t]
    assert len(out["trials"]) == 4 and len(rates) >= 2 and all(r > 0 for r in rates)
    assert out["tile"] in (24, 48) and out["events_per_s"] == max(rates) and not out["cached"]

    def no_run(*a):
        raise AssertionError("a cached choice ran again")
    again = tune((96, 96, 192), tab, lambda e: e.build_slab(1, 48), events_per_pass=(1.0, 4.0), cache=cache,
                 runner=no_run, **kw)
    assert again["cached"] and (again["tile"], again["events_per_pass"]) == (out["tile"], out["events_per_pass"])
