"""xyz frames (inpout/WriteMolMolXYZFile.f) from compacted records.

CPU: the host writer (mulskips_amd/frames.py), fed with the records of a
state taken from the oracle, reproduces the oracle's frame files byte for
byte (the oracle's writer reproduces the reference's fixtures byte for
byte).  The records are derived on the host here with the kernel's rules,
and the vacancies are given in the oracle's ListVoid order.  The same holds
for the SiC defect frames (_v, _w), whose records the test derives by the
reference's own loops (sequential wrong-site flips in grid order) on a
grown SiC state with synthetic vacancies and swapped species; that pins the
oracle's restatement of those frames too.
GPU: the device compaction (mskg_frame, mskg_frame_defects) returns exactly
those records (vacancies in grid order).
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


# ---- the SiC defect frames (WriteMolMolXYZFile.f:222-398) ----------------------
def defect_state(tmp_path, seed=5):
    """A grown SiC state (cube-zigzag deck, PtransZig 0.8) with synthetic
    defects: interior bulk atoms turned into vacancies (empty, coordination
    4, listed with their neighbour lists) and the species of some bulk atoms
    swapped, so the vacancies see every neighbour-species sum and wrong
    occupations of all kinds occur.  The frames only read the state, so it
    need not be one the events could reach; the oracle takes it over with
    its lists and LattInd rebuilt."""
    o, d, box, lat, void = frame_case(tmp_path, "cube-zigzag")
    _, mc, bulk, _ = oracle_state(o)
    rng = np.random.default_rng(seed)
    lat = lat.copy()
    xyz = bulk["xyz"]
    inner = np.nonzero((xyz[:, 2] > 30) & (xyz[:, 2] < 110) & (xyz[:, 0] > 30) & (xyz[:, 0] < 110))[0]
    pick = rng.choice(inner, 60, replace=False)
    vac, swap = pick[:30], pick[30:]
    for i in swap:  # Si <-> C
        x, y, z = xyz[i]
        sp = (lat[z, y, x] >> 27) & 7
        lat[z, y, x] = (lat[z, y, x] & ~(7 << 27)) | ((3 - sp) << 27)
    for i in vac:
        x, y, z = xyz[i]
        lat[z, y, x] = 4 << 24  # empty, coordination 4
    keep = np.setdiff1d(np.arange(len(xyz)), vac)
    bulk2 = {"xyz": xyz[keep], "nn": bulk["nn"][keep]}
    void2 = {"xyz": np.concatenate([void["xyz"], xyz[vac]]), "nn": np.concatenate([void["nn"], bulk["nn"][vac]])}
    o.load_state(lat, mc, bulk2, void2, kmc_time=o.time())
    return o, d, box, lat, mc, bulk2, void2


def species_bits(word):
    return (word >> 27) & 7


def host_defect_records(lat, box, bulk, void, ncrystal, list_crystal):
    """The _v classification and the _w records by the reference's loops:
    vacancies in the given list order with the species sum of their
    neighbours; then, for two crystal species with C, every bulk atom with
    no neighbour of the other species flipped in grid order (later atoms see
    the flips), and the bulk atoms with three such neighbours listed."""
    lx, ly, lz = box
    vac = []
    for (x, y, z), nn in zip(void["xyz"], void["nn"]):
        if z <= lz - 1:
            vac.append((x, y, z, int(sum(species_bits(lat[c[2], c[1], c[0]]) for c in nn))))
    wrong = []
    if ncrystal == 2 and 6 in list_crystal:
        sp = species_bits(lat).copy()
        nbr = {tuple(int(v) for v in s): nn for s, nn in zip(bulk["xyz"], bulk["nn"])}
        occ = (lat >> 23) & 1
        coor = (lat >> 24) & 7
        wall = (lat & 255) == 113

        def corr(x, y, z):
            s = sp[z, y, x]
            nn = nbr.get((x, y, z))
            if nn is None:
                return 0
            other = [(int(sp[c[2], c[1], c[0]]) >> 1) & 1 for c in nn] if s == 1 else \
                [int(sp[c[2], c[1], c[0]]) & 1 for c in nn] if s == 2 else [0]
            return sum(other)
        zz, yy, xx = np.nonzero((occ == 1) & (coor == 4))
        order = np.lexsort((xx, yy, zz))
        cells = [(int(xx[i]), int(yy[i]), int(zz[i])) for i in order if 1 <= zz[i] <= lz - 1]
        for x, y, z in cells:
            if not wall[z, y, x] and corr(x, y, z) == 0:
                sp[z, y, x] = 2 if sp[z, y, x] == 1 else 1
        for x, y, z in cells:
            if corr(x, y, z) == 3:
                wrong.append((x, y, z, int(sp[z, y, x])))
    return np.array(vac, np.int32).reshape(-1, 4), np.array(wrong, np.int32).reshape(-1, 4)


def test_host_writer_reproduces_oracle_defect_frames(tmp_path):
    o, d, box, lat, mc, bulk, void = defect_state(tmp_path)
    lc = [int(v) for v in d.tables()["list_crystal"]]
    vac, wrong = host_defect_records(lat, box, bulk, void, d.n_crystal, lc)
    codes = set(vac[:, 3].tolist())
    assert {4, 8}.issubset(codes) and len(codes) >= 3, codes  # C and Si vacancies and antisite classes
    assert len(wrong) > 10
    fv, fw = str(tmp_path / "f_v.xyz"), str(tmp_path / "f_w.xyz")
    ov, ow = str(tmp_path / "o_v.xyz"), str(tmp_path / "o_w.xyz")
    frames.write_defect_frames(fv, fw, (vac, wrong), box, d.sf, lc, d.n_cov)
    o.write_frames(v_path=ov, w_path=ow, it=3)
    assert open(fv, "rb").read() == open(ov, "rb").read()
    assert open(fw, "rb").read() == open(ow, "rb").read()
    o.close()


def test_vacancy_frame_of_other_models(tmp_path):
    """Outside SiC PVD the vacancy frame lists 'O' lines, and there is no
    wrong-occupation frame."""
    base, box, lens = SMALL["CVD-Si-SiH2"]
    path, d = small_deck(tmp_path, base, lens, idum=777)
    o = Oracle(*box)
    o.load_deck(path)
    o.step(5000, record=False)
    lat, mc, bulk, void = oracle_state(o)
    vac, wrong = host_defect_records(lat, box, bulk, void, d.n_crystal, [int(v) for v in d.list_crystal])
    assert len(wrong) == 0
    fv, ov = str(tmp_path / "f_v.xyz"), str(tmp_path / "o_v.xyz")
    fw, ow = str(tmp_path / "f_w.xyz"), str(tmp_path / "o_w.xyz")
    frames.write_defect_frames(fv, fw, (vac, wrong), box, d.sf, [int(v) for v in d.tables()["list_crystal"]],
                               d.n_cov)
    o.write_frames(v_path=ov, w_path=ow, it=3)
    assert open(fv, "rb").read() == open(ov, "rb").read()
    assert not os.path.exists(fw) and not os.path.exists(ow)
    o.close()


@pytest.mark.gpu
def test_device_defect_records(tmp_path):
    """mskg_frame_defects returns the host-derived records (vacancies in
    grid order) and leaves the lattice as it found it."""
    from mulskips_amd.engine import Engine
    o, d, box, lat, mc, bulk, void = defect_state(tmp_path)
    lc = [int(v) for v in d.tables()["list_crystal"]]
    vac, wrong = host_defect_records(lat, box, bulk, void, d.n_crystal, lc)
    eng = Engine(box, d.tables())
    eng.load_state(lat, mc, bulk, void)
    before = eng.store_state()[0]
    v, w = eng.frame_defects()
    np.testing.assert_array_equal(w, wrong)
    lx, ly, _ = box
    order = np.lexsort((vac[:, 0], vac[:, 1], vac[:, 2]))
    np.testing.assert_array_equal(v, vac[order])
    np.testing.assert_array_equal(eng.store_state()[0], before)
    # the written frames equal the oracle's but for the order of the vacancy lines
    fv, fw = str(tmp_path / "e_v.xyz"), str(tmp_path / "e_w.xyz")
    ov, ow = str(tmp_path / "o_v.xyz"), str(tmp_path / "o_w.xyz")
    frames.write_defect_frames(fv, fw, (v, w), box, d.sf, lc, d.n_cov)
    o.write_frames(v_path=ov, w_path=ow, it=3)
    a, b = open(fv).read().splitlines(), open(ov).read().splitlines()
    assert a[:2] == b[:2] and sorted(a[2:]) == sorted(b[2:])
    assert open(fw, "rb").read() == open(ow, "rb").read()
    eng.close()
    o.close()
