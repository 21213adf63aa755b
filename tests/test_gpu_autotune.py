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


def test_tune_ensemble_small(tmp_path):
    """The ensemble tuner runs the sample replicas together once per tile
    edge, keeps the fastest and answers a second call from its cache."""
    from mulskips_amd.autotune import tune_ensemble
    from mulskips_amd.engine import Engine
    reps = [dict(r, box=(192, 192, 768)) for r in models.ensemble_replicas(3)]
    built = []

    def make(r, tile):
        built.append(tile)
        e = Engine(r["box"], r["tables"], seed=r["seed"], tile=tile, events_per_pass=4.0, tau_policy=1,
                   frozen_below_z=36, stored_planes=120)
        e.build_alloy_slab(48, r["x"])
        e.refresh_rates()
        return e
    cache = str(tmp_path / "tune.json")
    out = tune_ensemble(make, reps, (16, 24), [r["t_end"] for r in reps], cache=cache, setup_tag="3 small")
    print(f"[tune_ensemble] {out['trials']}", flush=True)
    assert built == [16] * 3 + [24] * 3
    rates = [t["events_per_s"] for t in out["trials"]]
    assert len(rates) == 2 and min(rates) > 0 and out["events_per_s"] == max(rates) and not out["cached"]
    again = tune_ensemble(make, reps, (24, 16), [r["t_end"] for r in reps], cache=cache, setup_tag="3 small")
    assert again["cached"] and again["tile"] == out["tile"] and len(built) == 6
