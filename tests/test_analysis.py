"""CPU: pymulskips' film analyses (pymulskips/analyze.py:124-483) restated in
mulskips_amd.analysis, against the values pymulskips' own functions gave
(tests/golden/analysis, written by tests/golden/make_analyze_goldens.py):

* the surface height of each of the reference's 16 regression frames at two
  bin / roughness settings, and their H / Cl coverage, equal exactly;
* a growth run's frames (a growing Si (001) slab): the per-frame heights
  and the growth rate of every method equal pymulskips', for the frames in
  the order its glob found them (where its spline refuses the unsorted
  times, so does the restatement) and in time order;
* the frame's z values computed from compacted records (frame_z) equal the
  ones read back from the written frame, so the engine's records give the
  same heights without a file.
"""
import gzip
import json
import os

import numpy as np
import pytest

from mulskips_amd import analysis, frames
from tests.helpers import GOLDEN

DATA = os.path.join(GOLDEN, "analysis")
REF = json.load(open(os.path.join(DATA, "analyze.json")))
SETTINGS = [(5.0, 20.0), (3.0, 10.0)]


def unzip(gz, tmp_path):
    out = tmp_path / os.path.basename(gz)[:-3]
    with gzip.open(gz, "rb") as f:
        out.write_bytes(f.read())
    return str(out)


@pytest.mark.parametrize("name", sorted(REF["frames"]))
def test_frame_height_and_coverage(name, tmp_path):
    f = unzip(os.path.join(GOLDEN, "regression", name + ".gz"), tmp_path)
    sp, z, box = analysis.read_frame(f)
    for (b, r), want in zip(SETTINGS, REF["frames"][name]["height"]):
        assert analysis.surface_height(z, box[2], b, r) == want
    assert analysis.coverage(sp) == REF["frames"][name]["coverage"]


def frame_time(path):
    with open(path) as f:
        return float(f.readline().split()[3])


@pytest.mark.parametrize("order", ["glob", "sorted"])
def test_growth_rate(order, tmp_path):
    ref = REF["growth"][order]
    files = [unzip(os.path.join(DATA, "growth", n + ".gz"), tmp_path) for n in ref["order"]]
    heights = []
    for f in files:
        _, z, box = analysis.read_frame(f)
        heights.append(analysis.surface_height(z, box[2]))
    times = [frame_time(f) for f in files]
    for method in ("finitediff", "spline", "polyfit", "finiteANDfit"):
        want = ref[method]
        if "error" in want:
            with pytest.raises(ValueError):
                analysis.growth_rate(times, heights, method)
            continue
        assert heights == want["heights"]
        _, ave, _ = analysis.growth_rate(times, heights, method)
        if isinstance(ave, tuple):
            assert [float(v) for v in ave] == want["average"]
        else:
            assert float(ave) == want["average"]
    if order == "sorted":
        assert REF["growth"][order]["polyfit"]["average"] > 0.1  # the slab grows


def test_growth_rate_closed_form():
    """Heights rising linearly at v Angstrom/s give v * 0.36 micron/hour by
    every method."""
    t = np.linspace(0.0, 10.0, 11)
    h = 3.0 + 2.5 * t
    for m in ("finitediff", "spline", "polyfit"):
        _, ave, _ = analysis.growth_rate(t, h, m)
        assert abs(float(ave) - 2.5 * 0.36) < 1e-9, m


def test_record_z_equals_written_frame(tmp_path):
    from tests.test_frames import frame_case, host_records
    o, d, box, lat, void = frame_case(tmp_path, "CVD-Si-SiH2")
    lcov = [int(v) for v in d.tables()["list_cov"]]
    main, noepi, nat, ncov = host_records(lat, box, lcov)
    voids = np.concatenate([void["xyz"], np.zeros((len(void["xyz"]), 1), np.int32)], axis=1)
    f = str(tmp_path / "f.xyz")
    frames.write_frame(f, (main, voids, noepi, nat, ncov), box, d.sf, o.time(), 1,
                       [int(v) for v in d.tables()["list_crystal"]], lcov)
    sp, z, fbox = analysis.read_frame(f)
    zr = analysis.frame_z(np.concatenate([main, voids]), d.sf)
    np.testing.assert_array_equal(zr, z)
    assert analysis.surface_height(zr, d.sf * box[2]) == analysis.surface_height(z, fbox[2])
    o.close()


def test_film_observables_use_the_frame_analyses(tmp_path):
    """observables.film's pymulskips observables (surface height, H
    coverage), computed from LattCoo, equal the analyses of the frame file
    the oracle writes for the same state."""
    from mulskips_amd import observables
    from oracle.oracle import Oracle
    from tests.helpers import SMALL, small_deck
    for key in ("CVD-Si-SiH2", "cube-zigzag"):
        base, box, lens = SMALL[key]
        path, d = small_deck(tmp_path, base, lens, idum=777)
        o = Oracle(*box)
        o.load_deck(path)
        o.step(20000, record=False)
        r = observables.film(o.lattcoo(), box, 0, sf=d.sf)
        f = str(tmp_path / f"{key}.xyz")
        o.write_frame(f, None, 1)
        sp, z, fb = analysis.read_frame(f)
        assert r["surface"] == analysis.surface_height(z, fb[2])
        assert r["cov_h"] == analysis.coverage(sp)["H"][0]
        o.close()
