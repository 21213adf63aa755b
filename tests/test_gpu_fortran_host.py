"""GPU: the engine driven from Fortran through its iso_c_binding interfaces.

mulskips_amd/fortran/mskg_host_demo.f90 is the Fortran host's side in
miniature (the binding style of mulskips-source-LaserAnnealing/fsockets.f:
44-65): it creates the engine, loads a state handed over as the host's
module arrays (LattCoo and the three lists with NextNXYZ), replays a
recorded event history with mskg_replay, takes LattCoo back with
mskg_store_state, then runs the KMC loop as one logged mskg_run and takes
LattCoo back again.  The oracle, from the same state, applies the same
history and then the engine's logged run: both lattices equal the ones the
Fortran program returned, bit for bit.  Needs ROCm's amdflang; skipped
without it.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mulskips_amd import _lib
from oracle.oracle import Oracle
from tests.helpers import ROOT, SMALL, oracle_state, parity_mask, small_deck

pytestmark = pytest.mark.gpu
FLANG = shutil.which("amdflang") or (os.path.exists("/opt/rocm/bin/amdflang") and "/opt/rocm/bin/amdflang")
FORTRAN = os.path.join(ROOT, "mulskips_amd", "fortran")
LIB = os.path.join(ROOT, "mulskips_amd", "libmskg.so")


def build(tmp_path):
    exe = str(tmp_path / "mskg_host_demo")
    subprocess.run([FLANG, "-O2", "-J", str(tmp_path), "-o", exe, os.path.join(FORTRAN, "mskg_bindings.f90"),
                    os.path.join(FORTRAN, "mskg_host_demo.f90"), LIB, f"-Wl,-rpath,{os.path.dirname(LIB)}"],
                   check=True, cwd=tmp_path, timeout=300)
    return exe


def write_input(path, box, tab, lat, mc, bulk, void, log, run_events):
    lx, ly, lz = box
    hdr = np.array([lx, ly, lz, tab["n_crystal"], tab["n_cov"], *tab["list_crystal"][:3], *tab["list_cov"][:3],
                    len(mc["xyz"]), len(bulk["xyz"]), len(void["xyz"])], np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([tab["ptrans_zig"]], np.float64).tobytes())
        for k in ("ptrans_e", "ptrans_d", "ptrans_abs", "ptrans_des"):
            f.write(np.ascontiguousarray(tab[k], np.float64).tobytes())
        f.write(np.array([len(log), run_events], np.int64).tobytes())
        f.write(np.ascontiguousarray(lat, np.int32).tobytes())
        if len(mc["xyz"]):
            for k, dt in (("xyz", np.int32), ("nn", np.int32), ("event", np.int32), ("prob", np.float64)):
                f.write(np.ascontiguousarray(mc[k], dt).tobytes())
        for d in (bulk, void):
            if len(d["xyz"]):
                f.write(np.ascontiguousarray(d["xyz"], np.int32).tobytes())
                f.write(np.ascontiguousarray(d["nn"], np.int32).tobytes())
        f.write(np.ascontiguousarray(log, _lib.EVENT_DTYPE).tobytes())


@pytest.mark.skipif(not FLANG, reason="needs ROCm's amdflang")
@pytest.mark.timeout(600)
@pytest.mark.parametrize("key", ["cube-zigzag", "CVD-Si-SiH2"])
def test_fortran_host_drives_the_engine(tmp_path, key):
    exe = build(tmp_path)
    base, box, lens = SMALL[key]
    path, d = small_deck(tmp_path, base, lens)
    tab = d.tables()
    o = Oracle(*box)
    o.load_deck(path)
    lat0, mc0, bulk0, void0 = oracle_state(o)
    log = o.step(5000)
    run_events = 5000
    fin, fout = str(tmp_path / "state.bin"), str(tmp_path / "out.bin")
    write_input(fin, box, tab, lat0, mc0, bulk0, void0, log, run_events)
    res = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=300)
    print(f"[fortran {key}] {res.stdout.strip()} {res.stderr.strip()}", flush=True)
    assert res.returncode == 0, res.stdout + res.stderr
    lx, ly, lz = box
    npts = lx * ly * (lz + 1)
    raw = open(fout, "rb").read()
    off = 4 * npts
    lat_replay = np.frombuffer(raw[:off], np.int32).reshape(lz + 1, ly, lx)
    n_log, stop = np.frombuffer(raw[off:off + 8], np.int64)[0], np.frombuffer(raw[off + 8:off + 12], np.int32)[0]
    off += 8 + 4 + 8
    run_log = np.frombuffer(raw[off:off + n_log * _lib.EVENT_DTYPE.itemsize], _lib.EVENT_DTYPE)
    off += n_log * _lib.EVENT_DTYPE.itemsize
    lat_run = np.frombuffer(raw[off:off + 4 * npts], np.int32).reshape(lz + 1, ly, lx)
    m = parity_mask(box)
    lat_o = o.lattcoo()
    assert np.array_equal(lat_o, lat_replay), "the replayed lattice differs from the oracle's"
    assert n_log >= run_events and stop in (0, 1)
    o.apply_log(run_log)
    lat_o = o.lattcoo()
    bad = np.count_nonzero((lat_o != lat_run) & m)
    assert bad == 0, f"the run's lattice differs from the oracle's at {bad} points"
    o.close()
