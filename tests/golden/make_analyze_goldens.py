"""Write the goldens of pymulskips' own film analyses (run only in a
container that holds the reference; the outputs are committed).

pymulskips/analyze.py (get_surface_height :124-207, analyze_growth_rate
:209-416, get_coverage :418-483) is imported from the reference and applied
to
* the reference's own regression frames (tests/golden/regression/
  I0000000[01]-*.xyz.gz): surface heights at two bin / roughness settings
  and the H / Cl coverage;
* a growth run: a Si (001) slab growing at 700 K (1.0 eV bonds, 1 ML/s),
  written frame by frame by the oracle (whose writer reproduces the
  reference's frames byte for byte) into tests/golden/analysis/growth/,
  and its growth rate by every method of analyze_growth_rate.
The values go to tests/golden/analysis/analyze.json, for the frames in the
order analyze_growth_rate's glob found them and in name (time) order.

    python tests/golden/make_analyze_goldens.py
"""
import glob
import gzip
import json
import os
import shutil
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFERENCE = "/root/reference"
OUT = os.path.join(HERE, "analysis")
GROWTH = os.path.join(OUT, "growth")
sys.path.insert(0, ROOT)

SETTINGS = [(5.0, 20.0), (3.0, 10.0)]
METHODS = ["finitediff", "spline", "polyfit", "finiteANDfit"]


class Miller100:
    miller = 100


def growth_frames(outdir):
    """Frames of a growing Si (001) slab, one every 2000 events, from the
    oracle's main loop (RunType 'R', seed 77)."""
    from mulskips_amd import deck as deckmod
    from mulskips_amd import models
    from oracle.oracle import Oracle
    tab = models.pair_energy_tables([[1.0]], [1.0], 700.0, 1e13, (14,))
    d = deckmod.deck_from_tables(tab, "FS", [48, 120, 120], max_iter=26000, idum=77)
    d.out_molmol = 2000
    d.sf = 0.4525
    box = (96, 96, 192)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "start.dat")
        deckmod.write_start_dat(d, path)
        o = Oracle(*box)
        o.load_deck(path)
        o.run_main(outdir=tmp)
        o.close()
        os.makedirs(outdir, exist_ok=True)
        for f in sorted(glob.glob(os.path.join(tmp, "I*[0-9].xyz"))):
            with open(f, "rb") as src, gzip.open(os.path.join(outdir, os.path.basename(f) + ".gz"), "wb") as dst:
                shutil.copyfileobj(src, dst)


def main():
    sys.path.insert(0, os.path.join(REFERENCE, "pymulskips"))
    import analyze  # noqa: E402  (pymulskips/analyze.py)
    os.makedirs(OUT, exist_ok=True)
    out = {"frames": {}, "growth": {}}
    with tempfile.TemporaryDirectory() as tmp:
        for gz in sorted(glob.glob(os.path.join(HERE, "regression", "I*.xyz.gz"))):
            name = os.path.basename(gz)[:-3]
            f = os.path.join(tmp, name)
            with gzip.open(gz, "rb") as src, open(f, "wb") as dst:
                shutil.copyfileobj(src, dst)
            rec = {"height": [analyze.get_surface_height(f, b, r) for b, r in SETTINGS]}
            cov = analyze.get_coverage(f, Miller100())
            rec["coverage"] = {k: [float(x) for x in v] for k, v in cov.items()}
            out["frames"][name] = rec
        growth_frames(GROWTH)
        run = os.path.join(tmp, "growth")
        os.makedirs(run)
        for gz in glob.glob(os.path.join(GROWTH, "*.gz")):
            with gzip.open(gz, "rb") as src, open(os.path.join(run, os.path.basename(gz)[:-3]), "wb") as dst:
                shutil.copyfileobj(src, dst)
        # analyze_growth_rate takes the frames in the order glob gives them,
        # which need not be time order (its spline then refuses the data):
        # run it on that order and on the frames sorted by name
        found = analyze.read_output_files
        for key, sort in (("glob", False), ("sorted", True)):
            analyze.read_output_files = (lambda d, w="undercoordinated", f=found, s=sort:
                                         sorted(f(d, w)) if s else f(d, w))
            order = [os.path.basename(f) for f in analyze.read_output_files(run, "undercoordinated")]
            res = {"order": order}
            for m in METHODS:
                try:
                    ave, heights = analyze.analyze_growth_rate(run, method=m, plotting=False,
                                                               return_surf_height=True)
                    res[m] = {"average": [float(v) for v in ave] if isinstance(ave, tuple) else float(ave),
                              "heights": [float(v) for v in heights]}
                except ValueError as exc:
                    res[m] = {"error": type(exc).__name__}
            out["growth"][key] = res
        analyze.read_output_files = found
    with open(os.path.join(OUT, "analyze.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {len(out['frames'])} frame analyses and {len(METHODS)} growth-rate methods to {OUT}")


if __name__ == "__main__":
    main()
