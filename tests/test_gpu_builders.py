"""GPU: the shaped initial states built on the device equal the reference's
builders (main.f:224-294) as the oracle restates them.

* mskg_build_shape for the InitSt of the regression decks' shapes -- 'C'
  (cube, cube-zigzag), 'S' (sphere), 'I' (inverted pyramids), 'FF' (FinFET)
  -- against the oracle's SetNC / SetNCSp / SetInvPOff / SetSiFINFET
  (pinned by those decks' fixtures): every lattice point's LattCoo word, the
  three lists with their NextNXYZ, the MC particles' rates within 1e-6 and
  the phase maps of every grid point (walls off the lattice included).
* mskg_build_cad against the oracle's SetCAD on the FinFET geometry written
  as a geometry file and on a random geometry with walls, atoms, vacancies
  and isolated atoms; with an alloy fraction, the same structure, each
  atom's species the engine's Philox draw (restated here in numpy) and the
  species-2 share within 4 sigma of the fraction.
* A run from a device-built shape is a serial history: the oracle replays
  the engine's log from the built state to the engine's lattice.
"""
import numpy as np
import pytest

from mulskips_amd.engine import Engine
from oracle.oracle import Oracle
from tests.helpers import SMALL, list_map, oracle_state, parity_mask, small_deck
from tests.test_oracle_cad import ATOM_ID, WALL_ID, dia_type, grids

pytestmark = pytest.mark.gpu


def compare(o, eng, box, tag):
    lat_o, mc_o, bulk_o, void_o = oracle_state(o)
    lat_e, mc_e, bulk_e, void_e = eng.store_state(lattcoo=lat_o.copy())
    bad = np.count_nonzero((lat_o != lat_e) & parity_mask(box))
    assert bad == 0, f"{tag}: LattCoo differs at {bad} lattice points"
    for name, a, b in (("MC", mc_o, mc_e), ("bulk", bulk_o, bulk_e), ("void", void_o, void_e)):
        ma, mb = list_map(a), list_map(b)
        assert ma == mb, f"{tag}: {name} lists differ ({len(ma)} vs {len(mb)})"
    assert len(mc_o["xyz"]) > 50, tag
    cls, rate = eng.site_rates(mc_o["xyz"])
    np.testing.assert_allclose(rate, mc_o["prob"], rtol=1e-6, err_msg=tag)
    for layout in ("planes",):
        co, so = o.phases(layout)
        ce, se = eng.phases(layout)
        np.testing.assert_array_equal(co, ce, err_msg=f"{tag}: coordination map")
        np.testing.assert_array_equal(so, se, err_msg=f"{tag}: species map")
    return mc_o


@pytest.mark.parametrize("name", ["cube", "cube-zigzag", "sphere", "inverted-pyramid", "finfet"])
def test_device_shape_equals_reference_builder(tmp_path, name):
    base, box, lens = SMALL[name]
    path, d = small_deck(tmp_path, base, lens)
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables())
    eng.build_shape(d.init_st, d.lens)
    print(f"[{name}] InitSt {d.init_st} lens {d.lens}: {o.counts()}", flush=True)
    compare(o, eng, box, name)
    o.close()
    eng.close()


def test_run_from_a_device_shape_is_a_serial_history(tmp_path):
    base, box, lens = SMALL["sphere"]
    path, d = small_deck(tmp_path, base, lens)
    o = Oracle(*box)
    o.load_deck(path)
    eng = Engine(box, d.tables(), seed=17, tile=24, events_per_pass=4.0, tau_policy=1)
    eng.build_shape(d.init_st, d.lens)
    eng.set_event_log(200000)
    n, t, stop = eng.run(50000)
    log = eng.event_log()
    assert len(log) == n > 10000
    o.apply_log(log)
    lat_o, mc_o, _, _ = oracle_state(o)
    lat_e, mc_e, _, _ = eng.store_state(lattcoo=lat_o.copy())
    assert np.count_nonzero((lat_o != lat_e) & parity_mask(box)) == 0
    assert list_map(mc_o) == list_map(mc_e)
    o.close()
    eng.close()


def finfet_ids(box, lens):
    lx, ly, lz = box
    ln, lww, lwwtop, lap = (v - v % 24 for v in lens)
    z, y, x = grids(box)
    xa, xb = int(np.float32(0.5) * np.float32(lx - lap)), int(np.float32(0.5) * np.float32(lx + lap))
    wall = (z <= 5) | (z >= lz - 6) | (z >= lwwtop) | ((z >= lww) & ((x < xa) | (x >= xb)))
    return np.where(wall, WALL_ID, np.where((z >= 6) & (z <= ln - 1), ATOM_ID, 0)).astype(np.int8)


def random_ids(box, seed=7):
    """Walls, a rough slab with vacancies, scattered atoms (some isolated)."""
    lx, ly, lz = box
    rng = np.random.default_rng(seed)
    z, y, x = grids(box)
    top = 40 + (6 * np.sin(x / 7.0) * np.cos(y / 5.0)).astype(int)
    ids = np.where(z <= top, ATOM_ID, 0).astype(np.int8)
    ids[(rng.random(ids.shape) < 0.02) & (z <= top)] = 0              # vacancies
    ids[(rng.random(ids.shape) < 0.01) & (z > top) & (z < 90)] = ATOM_ID  # scattered atoms
    ids[(z >= 60) & (z < 66) & (x < 24)] = WALL_ID                    # a wall ledge
    return ids


@pytest.mark.parametrize("geometry", ["finfet", "random"])
def test_device_cad_equals_reference_setcad(tmp_path, geometry):
    base, box, lens = SMALL["finfet"]
    path, d = small_deck(tmp_path, base, lens)
    ids = finfet_ids(box, lens) if geometry == "finfet" else random_ids(box)
    o = Oracle(*box)
    o.load_deck(path)
    o.build_cad(ids)
    eng = Engine(box, d.tables())
    eng.build_cad(ids)
    print(f"[cad {geometry}] lists {o.counts()}", flush=True)
    compare(o, eng, box, f"cad {geometry}")
    if geometry == "random":
        assert o.counts()[2] > 10  # vacancies became voids
    o.close()
    eng.close()


def test_device_cad_alloy(tmp_path):
    base, box, lens = SMALL["finfet"]
    path, d = small_deck(tmp_path, base, lens)
    tab = d.tables()
    tab["n_crystal"] = 2
    tab["list_crystal"] = np.array([14, 32, 0], np.int32)
    ids = random_ids(box, seed=3)
    a = Engine(box, tab, seed=5)
    a.build_cad(ids)
    b = Engine(box, tab, seed=5)
    b.build_cad(ids, alloy_fraction=0.3)
    lat_a, mc_a, bulk_a, void_a = a.store_state()
    lat_b, mc_b, bulk_b, void_b = b.store_state()
    occ_a = (lat_a >> 23) & 1
    np.testing.assert_array_equal(occ_a, (lat_b >> 23) & 1)
    np.testing.assert_array_equal((lat_a >> 24) & 7, (lat_b >> 24) & 7)
    assert list_map(mc_a) == list_map(mc_b) and list_map(bulk_a) == list_map(bulk_b)
    atoms = occ_a.astype(bool) & ((lat_a & 255) != 113)
    sp2 = ((lat_b >> 27) & 7) == 2
    # the draw of each atom: Philox-4x32-10 (Salmon et al., SC'11) keyed by
    # the seed, counter (point index, 0, 0x5e6e0002, 0), its first 53 bits
    # against the fraction
    zz, yy, xx = np.nonzero(atoms)
    lx, ly, _ = box
    pidx = (zz.astype(np.int64) * (ly // 2) + yy // 2) * (lx // 2) + xx // 2
    w0, w1, _, _ = philox4x32_10(pidx & 0xFFFFFFFF, pidx >> 32, 0x5e6e0002, 0, 5, 0)
    u = ((((w0 << np.uint64(32)) | w1) >> np.uint64(11)).astype(np.float64) + 0.5) / 2.0 ** 53
    np.testing.assert_array_equal(sp2[atoms], u < 0.3)
    frac = float(np.mean(sp2[atoms]))
    assert abs(frac - 0.3) < 4 * np.sqrt(0.21 / atoms.sum()), (frac, atoms.sum())
    a.close()
    b.close()


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox-4x32-10 on numpy uint64 lanes holding 32-bit words."""
    m = np.uint64(0xFFFFFFFF)
    c = [np.asarray(v, np.uint64) & m for v in np.broadcast_arrays(c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0), np.uint64(k1)
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & m, p1 & m, ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & m, p0 & m]
        k0, k1 = (k0 + np.uint64(0x9E3779B9)) & m, (k1 + np.uint64(0xBB67AE85)) & m
    return c
