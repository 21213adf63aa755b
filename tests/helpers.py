"""Shared test helpers: small decks, oracle <-> engine state plumbing."""
import os

import numpy as np

from mulskips_amd import deck as deckmod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REGRESSION = os.path.join(GOLDEN, "regression")
DECKS = ["cube", "cube-zigzag", "sphere", "surface", "inverted-pyramid", "finfet", "Si-SL", "CVD-Si-SiH2"]


def regression_deck(name):
    return os.path.join(REGRESSION, f"start-{name}.dat")


def small_deck(tmpdir, base, lens, run_type="R", idum=9117116, max_iter=10 ** 9, ptrans_zig=None):
    """A regression deck with its shape lengths changed so it fits a small box."""
    d = deckmod.read_start_dat(regression_deck(base))
    d.lens = list(lens)
    d.run_type = run_type
    d.idum = idum
    if d.exit_strategy == "Time":
        d.exit_strategy, d.max_iter, d.out_molmol = "Iter", max_iter, max_iter
    else:
        d.max_iter = d.out_molmol = max_iter
    if ptrans_zig is not None:
        d.ptrans_zig = ptrans_zig
    path = os.path.join(str(tmpdir), f"start-{base}-small.dat")
    deckmod.write_start_dat(d, path)
    return path, d


# (base deck, box, lens): small versions of the regression shapes
SMALL = {
    "cube": ("cube", (144, 144, 144), (72, 72, 72)),
    "cube-zigzag": ("cube-zigzag", (144, 144, 144), (72, 72, 72)),
    "sphere": ("sphere", (144, 144, 144), (96, 96, 96)),
    "surface": ("surface", (96, 96, 192), (72, 72, 72)),
    "inverted-pyramid": ("inverted-pyramid", (144, 144, 144), (48, 60, 60)),
    "Si-SL": ("Si-SL", (96, 96, 192), (96, 120, 120)),
    "finfet": ("finfet", (96, 96, 240), (72, 96, 168, 48)),
    "CVD-Si-SiH2": ("CVD-Si-SiH2", (96, 96, 192), (72, 120, 120)),
}


def oracle_state(o):
    mc, bulk, void = o.lists()
    return o.lattcoo(), mc, bulk, void


def parity_mask(box):
    """The lattice points of the grid (x = y = z mod 2), as a bool array of
    LattCoo's shape (LenZ+1, LenY, LenX), built by broadcasting so that a
    large grid costs one byte per point."""
    lx, ly, lz = box
    z = (np.arange(lz + 1, dtype=np.int8) & 1)[:, None, None]
    y = (np.arange(ly, dtype=np.int8) & 1)[None, :, None]
    x = (np.arange(lx, dtype=np.int8) & 1)[None, None, :]
    return (x == y) & (x == z)


def list_map(d):
    """{site tuple: neighbour tuple} of a list dict."""
    return {tuple(int(v) for v in s): tuple(int(v) for v in n.reshape(-1)) for s, n in zip(d["xyz"], d["nn"])}


def oracle_run_to_time(o, t_end):
    """Run the oracle's serial loop to the KMC time t_end (or until it
    stops): the state at t_end, every event up to t_end executed and none
    after it, as the engine stops its last window at t_end.  Returns the
    number of events."""
    n, _ = o.run(10 ** 15, t_end, at_time=True)
    return n


def oracle_observables(args):
    """Worker (separate process): run one deck through the oracle to t_end
    with RunType 'R' and seed idum, return the film observables, the events
    and the executed events by type.  args: (path, box, idum, t_end, z0,
    frozen[, second_factor[, sf]]); with the super-lattice step sf the
    observables include pymulskips' surface height and H coverage."""
    import os as _os
    import tempfile as _tf

    from mulskips_amd import deck as _deck
    from mulskips_amd import observables as _ob
    from oracle.oracle import Oracle as _Oracle
    path, box, idum, t_end, z0, frozen = args[:6]
    f2 = args[6] if len(args) > 6 else None
    sf = args[7] if len(args) > 7 else None
    d = _deck.read_start_dat(path)
    d.run_type, d.idum = "R", int(idum)
    with _tf.TemporaryDirectory() as tmp:
        p = _os.path.join(tmp, "start.dat")
        _deck.write_start_dat(d, p)
        o = _Oracle(*box)
        o.load_deck(p)
        if frozen:
            o.set_frozen(frozen)
        if f2 is not None:
            o.set_second_factor(f2)
        n = oracle_run_to_time(o, t_end)
        r = _ob.film(o.lattcoo(), box, z0, sf=sf)
        r["events"] = n
        r["time"] = o.time()
        ex = o.executed()
        r["dep1"], r["dep2"], r["evap"] = int(ex[0]), int(ex[1]), int(ex[3])

        o.close()
    return r
