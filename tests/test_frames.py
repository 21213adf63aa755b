"""xyz frames (inpout/WriteMolMolXYZFile.f) from compacted records.

CPU: the host writer (mulskips_amd/frames.py), fed with the records of a
state taken from the oracle, reproduces the oracle's frame files byte for
byte (the oracle's writer reproduces the reference's fixtures byte for
byte).  The records are derived on the host here with the kernel's rules,
and the vacancies are given in the oracle's ListVoid order.
GPU: the device compaction (mskg_frame) returns exactly those records
(vacancies in grid order).
"""
import os

import numpy as np
import pytest

from mulskips_amd import frames
from oracle.oracle import Oracle
from tests.helpers import SMALL, oracle_state, parity_mask, small_deck

CASES = ["cube", "CVD-Si-SiH2", "surface", "inverted-pyramid"]


def host_records(lat, box, list_cov):
    """The records mskg_frame defines, from a LattCoo array (grid order)."""
    lx, ly, lz = box
    z, y, x = np.meshgrid(np.arange(lz + 1), np.arange(ly), np.arange(lx), indexing="ij")
    cov = lat & 255
    occ = (lat >> 23) & 1
    coor = (lat >> 24) & 7
    sp = (lat >> 27) & 7
    ok = parity_mask(box) & (z <= lz - 1) & (cov != 113)
    atom = ok & (occ == 1) & (sp >= 1)
    covered = ok & (occ == 0) & (cov != 0)
    slot = np.zeros_like(cov)
    for k, c in enumerate(list_cov):
        slot[cov == c] = k + 1
    main_mask = (atom & (coor <= 3)) | covered
    code = np.where(atom, sp, -slot)
    order = np.argsort((z * ly + y) * lx + x, axis=None, kind="stable")

    def pick(mask, codes):
        m = mask.reshape(-1)[order]
        return np.stack([x.reshape(-1)[order][m], y.reshape(-1)[order][m], z.reshape(-1)[order][m],
                         codes.reshape(-1)[order][m]], axis=1).astype(np.int32)

    xs, ys, zs = x, y, z
    dia = (xs % 3 == 0) & (ys % 3 == 0) & (zs % 3 == 0)
    x1, y1, z1 = xs // 3, ys // 3, zs // 3
    dia &= ((x1 % 2) == (y1 % 2)) & ((x1 % 2) == (z1 % 2)) & (((x1 + y1 + z1) % 4 == 0) | ((x1 + y1 + z1) % 4 == 3))
    noepi = atom & (coor == 4) & ~dia
    inner = (z >= 1)
    nat = int((atom & (coor <= 3) & inner).sum())
    ncov = int((covered & inner).sum())
    return pick(main_mask, code), pick(noepi, sp), nat, ncov


def frame_case(tmp_path, key, n_events=30000):
    base, box, lens = SMALL[key]
    path, d = small_deck(tmp_path, base, lens, idum=777)
    o = Oracle(*box)
    o.load_deck(path)
    o.step(n_events, record=False)
    lat, mc, bulk, void = oracle_state(o)
    return o, d, box, lat, void


@pytest.mark.parametrize("key", CASES)
def test_host_writer_reproduces_oracle_frames(tmp_path, key):
    o, d, box, lat, void = frame_case(tmp_path, key)
    lc = [int(v) for v in d.tables()["list_crystal"]]
    lcov = [int(v) for v in d.tables()["list_cov"]]
    main, noepi, nat, ncov = host_records(lat, box, lcov)
    voids = np.concatenate([void["xyz"], np.zeros((len(void["xyz"]), 1), np.int32)], axis=1)
    it = 1234567
    ox, od = str(tmp_path / "o.xyz"), str(tmp_path / "o_d.xyz")
    fx, fd = str(tmp_path / "f.xyz"), str(tmp_path / "f_d.xyz")
    frames.write_frame(fx, (main, voids, noepi, nat, ncov), box, d.sf, o.time(), it, lc, lcov, d_path=fd)
    o.write_frame(ox, od, it)
    assert open(fx, "rb").read() == open(ox, "rb").read()
    assert open(fd, "rb").read() == open(od, "rb").read()
    assert len(main) > 100
    o.close()


def test_fortran_e_format():
    assert frames.fort_e(0.0) == "  0.00000000E+00"[1:]
    assert frames.fort_e(2876829.4401611313) == " 0.28768294E+07"
    assert frames.fort_e(-1.5e-3) == "-0.15000000E-02"
    assert frames.fort_e(9.999999999) == " 0.10000000E+02"


@pytest.mark.gpu
@pytest.mark.parametrize("key", CASES)
def test_device_frame_records(tmp_path, key):
    from mulskips_amd.engine import Engine
    o, d, box, lat, void = frame_case(tmp_path, key)
    lcov = [int(v) for v in d.tables()["list_cov"]]
    main, noepi, nat, ncov = host_records(lat, box, lcov)
    _, mc, bulk, _ = oracle_state(o)
    eng = Engine(box, d.tables())
    eng.load_state(lat, mc, bulk, void)
    m, v, ne, nat_e, ncov_e = eng.frame()
    np.testing.assert_array_equal(m, main)
    np.testing.assert_array_equal(ne, noepi)
    assert (nat_e, ncov_e) == (nat, ncov)
    got = sorted(map(tuple, v[:, :3].tolist()))
    want = sorted(map(tuple, void["xyz"].tolist()))
    assert got == want
    # the written frame equals the oracle's but for the order of the vacancy lines
    fx, ox = str(tmp_path / "e.xyz"), str(tmp_path / "o.xyz")
    frames.write_frame(fx, (m, v, ne, nat_e, ncov_e), box, d.sf, o.time(), 7, [int(c) for c in d.tables()["list_crystal"]],
                       lcov)
    o.write_frame(ox, None, 7)
    a, b = open(fx).read().splitlines(), open(ox).read().splitlines()
    assert len(a) == len(b)
    n_o = len(v)
    assert a[:len(a) - n_o] == b[:len(b) - n_o]
    assert sorted(a[len(a) - n_o:]) == sorted(b[len(b) - n_o:])
    eng.close()
    o.close()
