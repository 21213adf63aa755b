"""The CPU oracle against the reference's own regression fixtures.

regression-tests/input-and-references-files holds, for 8 decks, the start.dat,
the first and last xyz frames (I00000000.xyz, I00000001.xyz) and the run log
of the reference Fortran program.  The oracle runs each deck the way
mulskips.e does (main.f) on the 480^3 box the fixtures were produced with
(modules/defsystem.f:27) and must reproduce both frames byte for byte and the
final clock of the log to the last printed digit.  RunType 'T' decks replay
the fixed uniform table of PickTreeEvent.f:72-1079.

The 8 decks run concurrently in subprocesses (about 45 s of wall time on 8
cores, 0.9 GB of host memory each).
"""
import gzip
import json
import os
import re
import subprocess
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import DECKS, GOLDEN, REGRESSION, ROOT  # noqa: E402

CLI = os.path.join(ROOT, "oracle", "kmc_oracle_cli")
BOX = ("480", "480", "480")


def _log_final(name):
    """(Iter, Time) of the last 'Iter, Time, Site' line of the reference log."""
    last = None
    with open(os.path.join(REGRESSION, f"log-{name}.txt")) as f:
        for line in f:
            if "Iter, Time, Site" in line:
                last = line
    toks = re.split(r"\s+", last.split("Site", 1)[1].strip())
    return int(toks[0]), float(toks[1]), toks[1]


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    from oracle import oracle
    oracle.build()
    procs = {}
    for name in DECKS:
        d = tmp_path_factory.mktemp(f"deck-{name}")
        with open(os.path.join(REGRESSION, f"start-{name}.dat")) as src, open(d / "start.dat", "w") as dst:
            dst.write(src.read())
        procs[name] = (d, subprocess.Popen([CLI, "start.dat", *BOX, ".", os.path.join(GOLDEN, "numberslist.txt")],
                                           cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    out = {}
    for name, (d, p) in procs.items():
        so, se = p.communicate(timeout=600)
        out[name] = (d, p.returncode, so, se)
    return out


@pytest.mark.parametrize("name", DECKS)
def test_frames_byte_identical(runs, name):
    d, rc, so, se = runs[name]
    assert rc == 0, se
    for frame in ("I00000000", "I00000001"):
        with gzip.open(os.path.join(REGRESSION, f"{frame}-{name}.xyz.gz"), "rb") as f:
            want = f.read()
        got = (d / f"{frame}.xyz").read_bytes()
        assert got == want, f"{name} {frame}.xyz differs"


@pytest.mark.parametrize("name", DECKS)
def test_final_clock_matches_log(runs, name):
    d, rc, so, se = runs[name]
    assert rc == 0, se
    res = json.loads(so.strip().splitlines()[-1])
    it, t, t_txt = _log_final(name)
    assert res["iter"] == it + 1
    # the log prints 17 significant digits: the FP64 clock must agree with them exactly
    assert res["time"] == t, (res["time"], t_txt)


def test_frames_change_the_lattice_plain_text(runs):
    """Guard against a vacuous comparison: the two frames of a growth deck differ."""
    d, rc, so, se = runs["cube"]
    assert (d / "I00000000.xyz").read_bytes() != (d / "I00000001.xyz").read_bytes()
