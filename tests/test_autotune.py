"""CPU: the launch-shape autotuner's candidates, choice and cache (the timed
runs replaced by a stand-in runner; the GPU test times real ones)."""
import json

import pytest

from mulskips_amd import models
from mulskips_amd.autotune import candidate_tiles, tile_capacity, tune
from mulskips_amd.engine import EngineError


def test_candidate_tiles():
    assert candidate_tiles((3072, 3072, 3072), second_neighbour=True) == [24, 32, 48]
    assert candidate_tiles((96, 96, 144)) == [16, 24, 48]
    assert candidate_tiles((96, 96, 144), second_neighbour=True) == [24, 48]
    assert candidate_tiles((264, 264, 384)) == [22, 44]
    assert candidate_tiles((100, 100, 100)) == []  # 20 leaves 5 tiles, 50 is above the cap


def test_tile_capacity():
    assert [tile_capacity(t) for t in (16, 24, 32, 48, 96)] == [1024, 1024, 1856, 4096, 4096]


class Runner:
    def __init__(self, rates, fail=()):
        self.rates, self.fail, self.calls = rates, set(fail), []

    def __call__(self, box, tables, setup, tile, k, warmup, events, engine_kw):
        self.calls.append((tile, k))
        if (tile, k) in self.fail:
            raise EngineError(5, "a tile's candidate list overflowed tile_capacity")
        return self.rates[(tile, k)]


def test_choice_and_cache(tmp_path):
    tab = models.sic_c2_tables()
    cache = str(tmp_path / "tune.json")
    rates = {(24, 1.0): 3.0, (24, 4.0): 5.0, (48, 1.0): 9.0, (48, 4.0): 2.0}
    r = Runner(rates, fail=[(48, 1.0)])
    out = tune((96, 96, 192), tab, setup=None, events_per_pass=(1.0, 4.0), cache=cache, device="dev", runner=r,
               tau_policy=1)
    assert (out["tile"], out["events_per_pass"], out["cached"]) == (24, 4.0, False)
    assert len(r.calls) == 4 and sum("error" in t for t in out["trials"]) == 1
    with open(cache) as f:
        assert out["key"] in json.load(f)
    again = Runner({}, fail=rates)  # must not be called
    out2 = tune((96, 96, 192), tab, setup=None, events_per_pass=(1.0, 4.0), cache=cache, device="dev",
                runner=again, tau_policy=1, seed=99)  # the seed is not part of the key
    assert out2["cached"] and again.calls == [] and out2["tile"] == 24
    # another model, option or device is another key
    for kw in ({"tau_policy": 0}, {"tau_policy": 1, "frozen_below_z": 49}):
        r3 = Runner(rates)
        assert not tune((96, 96, 192), tab, setup=None, events_per_pass=(1.0, 4.0), cache=cache, device="dev",
                        runner=r3, **kw)["cached"]
        assert len(r3.calls) == 4
    tab2 = dict(tab, ptrans_d=tab["ptrans_d"] * 2)
    assert not tune((96, 96, 192), tab2, setup=None, events_per_pass=(1.0,), cache=cache, device="dev",
                    runner=Runner(rates), tau_policy=1)["cached"]


def test_unvalidated_window_refused():
    tab = models.sic_c2_tables()
    with pytest.raises(ValueError, match="validated"):
        tune((96, 96, 192), tab, setup=None, events_per_pass=(4.0, 8.0), device="dev", runner=Runner({}))
    out = tune((96, 96, 192), tab, setup=None, events_per_pass=(8.0,), device="dev", allow_unvalidated=True,
               runner=Runner({(24, 8.0): 1.0, (48, 8.0): 2.0}))
    assert (out["tile"], out["events_per_pass"]) == (48, 8.0)


def test_all_failed_and_no_tile():
    tab = models.sic_c2_tables()
    with pytest.raises(RuntimeError, match="every candidate failed"):
        tune((96, 96, 192), tab, setup=None, events_per_pass=(1.0,), device="dev",
             runner=Runner({}, fail=[(24, 1.0), (48, 1.0)]))
    with pytest.raises(ValueError, match="no valid tile"):
        tune((100, 100, 100), tab, setup=None, device="dev", runner=Runner({}))


def test_capacity_beyond_lds_is_refused():
    """mskg_create rejects a list capacity whose LDS would not fit one wave's
    64 KB, before it looks for a device."""
    from mulskips_amd.engine import Engine
    with pytest.raises(EngineError, match="tile_capacity 7000"):
        Engine((96, 96, 192), models.sic_c2_tables(), tile=24, tile_capacity=7000)


def test_cache_key_holds_the_request(tmp_path):
    """A choice made under allow_unvalidated (k = 8) or on a restricted tile
    list is not returned for a default request; the result names the list
    capacity the chosen trial ran with."""
    tab = models.sic_c2_tables()
    cache = str(tmp_path / "tune.json")
    fast8 = Runner({(24, 4.0): 1.0, (24, 8.0): 5.0, (48, 4.0): 2.0, (48, 8.0): 3.0})
    out = tune((96, 96, 192), tab, setup=None, events_per_pass=(4.0, 8.0), cache=cache, device="dev",
               allow_unvalidated=True, runner=fast8)
    assert (out["tile"], out["events_per_pass"], out["tile_capacity"]) == (24, 8.0, 1024)
    assert all(t["lds_workgroups_per_cu"] == min(12, 160 * 1024 // (9 * t["tile_capacity"] + 2048))
               for t in out["trials"])
    r = Runner({(24, 4.0): 1.0, (48, 4.0): 2.0})
    out2 = tune((96, 96, 192), tab, setup=None, events_per_pass=(4.0, 8.0)[:1], cache=cache, device="dev", runner=r)
    assert not out2["cached"] and out2["events_per_pass"] == 4.0 and out2["tile"] == 48 and len(r.calls) == 2
    assert out2["tile_capacity"] == tile_capacity(48)
    r24 = Runner({(24, 4.0): 1.0})
    out3 = tune((96, 96, 192), tab, setup=None, tiles=[24], events_per_pass=(4.0,), cache=cache, device="dev",
                runner=r24)
    assert not out3["cached"] and out3["tile"] == 24 and len(r24.calls) == 1
    # the same request again comes from the cache; another starting state does not
    assert tune((96, 96, 192), tab, setup=None, tiles=[24], events_per_pass=(4.0,), cache=cache, device="dev",
                runner=Runner({}))["cached"]
    r4 = Runner({(24, 4.0): 1.0})
    assert not tune((96, 96, 192), tab, setup=None, tiles=[24], events_per_pass=(4.0,), cache=cache, device="dev",
                    setup_tag="thicker slab", runner=r4)["cached"]
    # an entry edited to an unvalidated window is refused for a default request
    with open(cache) as f:
        store = json.load(f)
    k3 = out3["key"]
    store[k3]["events_per_pass"] = 8.0
    with open(cache, "w") as f:
        json.dump(store, f)
    r5 = Runner({(24, 4.0): 1.0})
    assert not tune((96, 96, 192), tab, setup=None, tiles=[24], events_per_pass=(4.0,), cache=cache, device="dev",
                    runner=r5)["cached"]


def test_ensemble_key_names_the_question():
    """The ensemble tuner's cache key: the same for a reordered candidate
    list, different for another sample size, limit or starting state."""
    from mulskips_amd.autotune import ensemble_key
    k = ensemble_key("dev", (16, 24, 32), "s", [1.0, 2.0])
    assert k == ensemble_key("dev", (32, 16, 24), "s", [1.0, 2.0])
    others = [ensemble_key("dev", (16, 24), "s", [1.0, 2.0]),
              ensemble_key("dev", (16, 24, 32), "t", [1.0, 2.0]),
              ensemble_key("dev", (16, 24, 32), "s", [1.0, 2.5]),
              ensemble_key("dev", (16, 24, 32), "s", [1.0, 2.0, 3.0]),
              ensemble_key("dev", (16, 24, 32), "s", [1.0, 2.0], [10, 10]),
              ensemble_key("other", (16, 24, 32), "s", [1.0, 2.0])]
    assert len(set(others)) == len(others) and k not in others
