"""Launch-shape autotuner of the synchronous-sublattice engine (mskg_run).

The engine's free launch shape is the tile edge T (one wave works one tile
in each colour pass) and the number of events per tile and pass k, which
sets the window tau. tune() times short runs of each candidate (T, k) from
the same starting state and keeps the fastest, in events per device second.
The choice is cached in a JSON file, keyed by the box, a digest of the rate
tables and engine options, the device name and the request (candidate
tiles and windows, whether unvalidated windows were allowed, a tag of the
starting state). A later call with the same key runs nothing.

No reference counterpart: the reference runs one serial tree per binary.
"""
import hashlib
import json
import os

import numpy as np

from .engine import Engine, EngineError

TABLE_KEYS = ("ptrans_e", "ptrans_d", "ptrans_abs", "ptrans_des")


def candidate_tiles(box, second_neighbour=False, max_tile=48):
    """Tile edges the engine accepts for this box (its choose_tile rule): even,
    at least 16 (24 with the second-neighbour term, whose reads reach twice
    as far), dividing LenX and LenY into an even number of tiles each."""
    lx, ly, _ = (int(v) for v in box)
    lo = 24 if second_neighbour else 16
    hi = min(lx // 2, ly // 2, 254, int(max_tile))
    return [t for t in range(lo, hi + 1, 2)
            if lx % t == 0 and ly % t == 0 and (lx // t) % 2 == 0 and (ly // t) % 2 == 0]


# Window sizes whose statistics match serial KMC: growth, roughness,
# coverage and stacking faults within 2 standard errors over 32 seeds at
# k = 1 and k = 4. At k = 8 the growth rate drifts out (-3.5 SE on the
# cube-zigzag deck): the larger window trades accuracy for speed.
MAX_VALIDATED_EVENTS_PER_PASS = 4.0

MAX_CAPACITY = 4096  # 9 B of LDS per entry in the pass kernel: 37 KB per wave
LDS_PER_CU = 160 * 1024
PASS_STATIC_LDS = 2048   # the pass kernel's fixed LDS (touched points, counters, event scratch)
PASS_WAVES_PER_CU = 12   # its launch bounds: 3 waves per SIMD, 4 SIMDs per CU


def tile_capacity(tile):
    """List capacity per tile: 1024 entries at T = 24, scaled with the tile's
    cross-section (the candidates of a growing film lie near one surface),
    at most MAX_CAPACITY. A tile that overflows it stops its trial with an
    error, which tune() records."""
    return min(MAX_CAPACITY, max(1024, -(-1024 * tile * tile // (24 * 24) // 64) * 64))


def lds_occupancy(capacity):
    """One-wave pass workgroups per CU that the LDS allows at this list
    capacity (at most the launch bounds' 12): a larger capacity reserves
    more LDS per wave and can hold fewer tiles in flight."""
    return min(PASS_WAVES_PER_CU, LDS_PER_CU // (9 * int(capacity) + PASS_STATIC_LDS))


def tune_key(box, tables, engine_kw, device, tiles=None, events_per_pass=None, allow_unvalidated=False,
             setup_tag=None):
    """Cache key: the device, the box, a digest of the rate tables and engine
    options, and the request itself -- the candidate tiles and windows, whether
    unvalidated windows were allowed and the starting state's tag -- so that a
    choice made under one request is never handed to another."""
    h = hashlib.sha256()
    for k in TABLE_KEYS:
        h.update(np.ascontiguousarray(tables[k], np.float64).tobytes())
    if tables.get("second_factor") is not None:
        h.update(np.ascontiguousarray(tables["second_factor"], np.float64).tobytes())
    scal = {k: (list(map(int, tables[k])) if k.startswith("list_") else float(tables[k]))
            for k in ("n_crystal", "n_cov", "list_crystal", "list_cov", "ptrans_zig")}
    opts = {k: v for k, v in sorted(engine_kw.items()) if k != "seed"}
    req = {"tiles": sorted(int(t) for t in tiles) if tiles is not None else None,
           "events_per_pass": sorted(float(k) for k in events_per_pass) if events_per_pass is not None else None,
           "allow_unvalidated": bool(allow_unvalidated), "setup": setup_tag}
    h.update(json.dumps([scal, opts, req], sort_keys=True, default=str).encode())
    return f"{device}|{box[0]}x{box[1]}x{box[2]}|{h.hexdigest()[:16]}"


def device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # noqa: BLE001 - no device: the key still names the box and model
        return "unknown"


def time_candidate(box, tables, setup, tile, events_per_pass, warmup, events, engine_kw):
    """Events per device second of one (tile, events_per_pass) on a fresh
    engine brought to the starting state by setup(engine)."""
    eng = Engine(box, tables, tile=tile, tile_capacity=tile_capacity(tile), events_per_pass=events_per_pass,
                 **engine_kw)
    try:
        setup(eng)
        if warmup:
            eng.run(warmup)
        n, _, _ = eng.run(events)
        ms = eng.last_timing()["run_ms"]
        return n / (ms * 1e-3) if ms > 0 else 0.0
    finally:
        eng.close()


def tune(box, tables, setup, tiles=None, events_per_pass=(1.0, 2.0, 4.0), warmup=20000, events=200000,
         cache=None, device=None, runner=time_candidate, allow_unvalidated=False, setup_tag=None, **engine_kw):
    """Pick the fastest (tile, events_per_pass) for this box and model.

    setup(engine) brings a new engine to the starting state (for example
    ``lambda e: e.build_slab(1, 120)``). engine_kw are the other Engine
    options (tau_policy, frozen_below_z, la, seed). A candidate whose run
    fails (say, a tile list over capacity) is recorded with its error and
    skipped. Returns a dict with tile, events_per_pass, events_per_s, the
    trials and the cache key; cached=True when it came from the cache file.
    Candidates above MAX_VALIDATED_EVENTS_PER_PASS are refused unless
    allow_unvalidated is set: the tuner measures speed, not accuracy.
    The result's tile_capacity is the list capacity the chosen trial ran with
    (tile_capacity(tile)); build the deployed engine with it, so that it is
    the engine that was timed.  setup_tag names the starting state setup()
    builds, for the cache key.  Each trial records its capacity and the
    pass workgroups per CU its LDS reservation allows.
    """
    bad = [k for k in events_per_pass if k > MAX_VALIDATED_EVENTS_PER_PASS]
    if bad and not allow_unvalidated:
        raise ValueError(f"events_per_pass {bad} above {MAX_VALIDATED_EVENTS_PER_PASS}: the window's statistics "
                         "are validated up to there (allow_unvalidated=True to time them anyway)")
    box = tuple(int(v) for v in box)
    device = device or device_name()
    if tiles is None:
        tiles = candidate_tiles(box, second_neighbour=tables.get("second_factor") is not None)
    if not tiles:
        raise ValueError(f"no valid tile edge for box {box}")
    key = tune_key(box, tables, engine_kw, device, tiles, events_per_pass, allow_unvalidated, setup_tag)
    store = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            store = json.load(f)
        hit = store.get(key)
        # a cached entry must answer this request: a window within the
        # validated ones unless allowed, a tile among the candidates
        if hit and (allow_unvalidated or hit["events_per_pass"] <= MAX_VALIDATED_EVENTS_PER_PASS) \
                and int(hit["tile"]) in [int(t) for t in tiles]:
            return dict(hit, cached=True)
    trials = []
    for t in tiles:
        for k in events_per_pass:
            cap = tile_capacity(int(t))
            tr = {"tile": int(t), "events_per_pass": float(k), "tile_capacity": cap,
                  "lds_workgroups_per_cu": lds_occupancy(cap)}
            try:
                tr["events_per_s"] = float(runner(box, tables, setup, int(t), float(k), int(warmup), int(events),
                                                  engine_kw))
            except EngineError as exc:
                tr["error"] = str(exc)
            trials.append(tr)
    ok = [tr for tr in trials if "events_per_s" in tr]
    if not ok:
        raise RuntimeError(f"every candidate failed: {trials}")
    best = max(ok, key=lambda tr: tr["events_per_s"])
    out = {"tile": best["tile"], "events_per_pass": best["events_per_pass"], "events_per_s": best["events_per_s"],
           "tile_capacity": best["tile_capacity"], "trials": trials, "key": key}
    if cache:
        store[key] = out
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(store, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return dict(out, cached=False)


def ensemble_key(device, tiles, setup_tag, max_time, max_events=None):
    """Cache key of tune_ensemble: the device, the candidate tiles (in any
    order), the sample (setup_tag, its size) and the timed runs' per-replica
    limits, so that a choice is reused only for the same question."""
    tiles = sorted(int(t) for t in tiles)
    limits = {"max_time": [float(t) for t in max_time],
              "max_events": None if max_events is None else [int(k) for k in max_events]}
    body = {"tiles": tiles, "setup": setup_tag, "replicas": len(limits["max_time"]), "limits": limits}
    return f"{device}|ensemble|{json.dumps(body, sort_keys=True)}"


def tune_ensemble(make, replicas, tiles, max_time, max_events=None, cache=None, device=None, setup_tag=None):
    """Pick the tile edge of a batched ensemble (mskg_run_ensemble): the
    sample replicas run together once per candidate tile, each engine from
    make(replica, tile) in its starting state, and the tile with the most
    events per device second wins.  Every replica of an ensemble takes the
    chosen edge (it must divide every box, as candidate_tiles() checks).
    max_time / max_events: per-replica limits of the timed runs.  Cached
    like tune(), keyed by the device, the candidates and setup_tag (which
    must name the sample and its starting states) and the limits
    (ensemble_key)."""
    from .engine import run_ensemble
    device = device or device_name()
    tiles = sorted(int(t) for t in tiles)
    if len(max_time) != len(replicas) or (max_events is not None and len(max_events) != len(replicas)):
        raise ValueError("one max_time (and max_events, if given) per sample replica")
    key = ensemble_key(device, tiles, setup_tag, max_time, max_events)
    store = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            store = json.load(f)
        if key in store:
            return dict(store[key], cached=True)
    n = len(replicas)
    me = [10 ** 12] * n if max_events is None else list(max_events)
    trials = []
    for t in tiles:
        tr = {"tile": t}
        engs = []
        try:
            engs = [make(r, t) for r in replicas]
            done, _, _ = run_ensemble(engs, me, list(max_time))
            ms = engs[0].last_timing()["run_ms"]
            tr["events"] = int(done.sum())
            tr["events_per_s"] = float(done.sum()) / (ms * 1e-3) if ms > 0 else 0.0
        except EngineError as exc:
            tr["error"] = str(exc)
        finally:
            for e in engs:
                e.close()
        trials.append(tr)
    ok = [tr for tr in trials if "events_per_s" in tr]
    if not ok:
        raise RuntimeError(f"every candidate failed: {trials}")
    best = max(ok, key=lambda tr: tr["events_per_s"])
    out = {"tile": best["tile"], "events_per_s": best["events_per_s"], "trials": trials, "key": key}
    if cache:
        store[key] = out
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(store, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return dict(out, cached=False)
