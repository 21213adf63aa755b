"""Host-side mirror of the reference's KMC loop, driving the HIP engine.

The reference's Fortran host (mulskips-source/main.f) parses start.dat,
builds the lattice, then loops PickTreeEvent / Deposition / Evaporation /
absorption / desorption (main.f:304-371).  Here the lattice is handed to the
GPU engine with `load_state`, advanced with `run` (or `replay` of a recorded
event log) and fetched back with `store_state`, through the C ABI of
include/mskg.h.  Errors the reference turns into STOP become exceptions.
"""
import ctypes as C
import warnings

import numpy as np

from . import _lib
from ._lib import Config, EVENT_DTYPE, MSKG_FIELD_PLANES, MSKG_FIELD_SOCKET, ptr

CLASS_NAMES = {0: "none", 1: "adsorb", 2: "placeholder", 4: "evaporate", 5: "desorb", 6: "isolated"}


def _layout(name):
    codes = {"socket": MSKG_FIELD_SOCKET, "planes": MSKG_FIELD_PLANES}
    if name not in codes:
        raise ValueError(f"layout {name!r}: 'socket' or 'planes'")
    return codes[name]


def _torch():
    import torch  # device tensors only
    return torch


def _device_tensor(a):
    """a, if it is a torch tensor in GPU memory, else None."""
    if type(a).__module__.startswith("torch") and getattr(a, "is_cuda", False):
        return a
    return None


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mskg error {code}: {msg}")
        self.code = code


class Engine:
    """One lattice resident on the current GPU.

    tables: dict with ptrans_e (12288, Fortran order of PTransE), ptrans_d (9),
    ptrans_abs (12), ptrans_des (12), ptrans_zig, n_crystal, n_cov,
    list_crystal, list_cov -- the quantities SetProbability.f reads from
    start.dat.

    stored_planes: hold only the grid planes z < stored_planes in device
    memory (mskg_config.stored_len_z; None: all LenZ+1).  The planes above
    are vacuum no event reaches; a film that grows to within 24 planes of
    the stored top stops the run with an error.  store_state() then returns
    the stored planes of LattCoo.
    """

    def __init__(self, box, tables, seed=1, tile=0, tile_capacity=0, events_per_pass=0.0, tau_policy=0,
                 frozen_below_z=0, la=None, stored_planes=None):
        self.L = _lib.lib()
        lx, ly, lz = box
        self.box = (int(lx), int(ly), int(lz))
        self.stored_planes = int(lz) + 1 if stored_planes is None else min(int(stored_planes), int(lz) + 1)
        self._arrays = {k: np.ascontiguousarray(tables[k], np.float64)
                        for k in ("ptrans_e", "ptrans_d", "ptrans_abs", "ptrans_des")}
        cfg = Config()
        cfg.len_x, cfg.len_y, cfg.len_z = self.box
        cfg.n_crystal = int(tables["n_crystal"])
        cfg.n_cov = int(tables["n_cov"])
        for i in range(3):
            cfg.list_crystal[i] = int(tables["list_crystal"][i])
            cfg.list_cov[i] = int(tables["list_cov"][i])
        cfg.ptrans_e = ptr(self._arrays["ptrans_e"])
        cfg.ptrans_d = ptr(self._arrays["ptrans_d"])
        cfg.ptrans_abs = ptr(self._arrays["ptrans_abs"])
        cfg.ptrans_des = ptr(self._arrays["ptrans_des"])
        cfg.ptrans_zig = float(tables["ptrans_zig"])
        if la:
            cfg.la_mode = 1
            cfg.tm, cfg.p0, cfg.sigma, cfg.tflex = (float(la[k]) for k in ("tm", "p0", "sigma", "tflex"))
        cfg.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        cfg.tile = int(tile)
        cfg.tile_capacity = int(tile_capacity)
        cfg.events_per_pass = float(events_per_pass)
        cfg.tau_policy = int(tau_policy)
        cfg.frozen_below_z = int(frozen_below_z)
        cfg.stored_len_z = 0 if stored_planes is None else int(stored_planes)
        if tables.get("second_factor") is not None:
            self._arrays["second_factor"] = np.ascontiguousarray(tables["second_factor"], np.float64)
            if self._arrays["second_factor"].shape != (13,):
                raise ValueError("second_factor needs 13 values (0..12 occupied second neighbours)")
            cfg.second_factor = self._arrays["second_factor"].ctypes.data
        self._cfg = cfg
        self._guard_seen = 0
        h = C.c_void_p()
        rc = self.L.mskg_create(C.byref(cfg), C.byref(h))
        self.h = h
        if rc:
            msg = self.L.mskg_last_error(h).decode() if h else "allocation failed"
            self.close()
            raise EngineError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self.L.mskg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise EngineError(rc, self.L.mskg_last_error(self.h).decode())

    def load_state(self, lattcoo, mc, bulk, void, kmc_time=0.0):
        """lattcoo: int32 array (LenZ+1, LenY, LenX); mc/bulk/void: dicts with
        'xyz' (n,3) and 'nn' (n,4,3) int32 arrays (NextNXYZ), mc also
        'event' (Ind_Event) and 'prob' (ProbTrans)."""
        lat = np.ascontiguousarray(lattcoo, np.int32)
        lx, ly, lz = self.box
        assert lat.shape == (lz + 1, ly, lx), lat.shape

        def arr(d, k, dt):
            return np.ascontiguousarray(d[k], dt) if len(d["xyz"]) else np.zeros((1,) + np.shape(d[k])[1:], dt)

        keep = []
        args = []
        for d, extra in ((mc, True), (bulk, False), (void, False)):
            n = len(d["xyz"])
            xyz, nn = arr(d, "xyz", np.int32), arr(d, "nn", np.int32)
            keep += [xyz, nn]
            args += [n, ptr(xyz), ptr(nn)]
            if extra:
                ev, pr = arr(d, "event", np.int32), arr(d, "prob", np.float64)
                keep += [ev, pr]
                args += [ptr(ev), ptr(pr)]
        self._check(self.L.mskg_load_state(self.h, ptr(lat), *args, float(kmc_time)))

    def build_slab(self, mode, length):
        """Flat (001) slab built on the device: mode 0 = SetSi (Si, walls),
        1 = SetSiC3C (SiC)."""
        self._check(self.L.mskg_build_slab(self.h, int(mode), int(length)))

    def build_alloy_slab(self, length, fraction):
        """The SetSi slab as a random alloy: species 2 with probability fraction."""
        self._check(self.L.mskg_build_alloy_slab(self.h, int(length), float(fraction)))

    def build_shape(self, init_st, lens):
        """A shaped initial state on the device (mskg_build_shape): InitSt
        'C', 'S', 'I', 'FF', 'F' or 'FS' with start.dat's Len1..Len4."""
        ln = np.zeros(4, np.int32)
        ln[:len(lens)] = lens
        self._check(self.L.mskg_build_shape(self.h, str(init_st).encode(), ptr(ln)))

    def build_cad(self, ids, alloy_fraction=0.0):
        """SetCAD (InitSt 'IN') on the device: ids int8 (LenZ, LenY, LenX)
        geometry ids (10 wall, 1 atom on a diamond site), a numpy array or an
        int8 torch tensor on the engine's device."""
        lx, ly, lz = self.box
        dev = _device_tensor(ids)
        if dev is not None:
            assert tuple(dev.shape) == (lz, ly, lx) and dev.dtype == _torch().int8 and dev.is_contiguous()
            _torch().cuda.current_stream(dev.device).synchronize()
            self._check(self.L.mskg_build_cad(self.h, C.c_void_p(dev.data_ptr()), 1, float(alloy_fraction)))
            return
        a = np.ascontiguousarray(ids, np.int8)
        assert a.shape == (lz, ly, lx), a.shape
        self._check(self.L.mskg_build_cad(self.h, ptr(a), 0, float(alloy_fraction)))

    def frame(self):
        """Device frame compaction: (main, voids, non_epitaxial) record arrays
        of (x, y, z, code) rows in grid order and the header counts
        (atoms of coordination <= 3, covered points; both for 1 <= z < LenZ)."""
        c = np.zeros(5, np.int64)
        z = None
        self._check(self.L.mskg_frame(self.h, z, 0, z, 0, z, 0, ptr(c)))
        bufs = [np.zeros((max(int(c[k]), 1), 4), np.int32) for k in range(3)]
        self._check(self.L.mskg_frame(self.h, ptr(bufs[0]), int(c[0]), ptr(bufs[1]), int(c[1]), ptr(bufs[2]),
                                      int(c[2]), ptr(c)))
        return bufs[0][:c[0]], bufs[1][:c[1]], bufs[2][:c[2]], int(c[3]), int(c[4])

    def frame_defects(self):
        """The SiC defect frames' records (mskg_frame_defects): (vacancies
        with the species sum of their neighbours, wrong occupations with the
        flipped species), (x, y, z, code) rows in grid order."""
        c = np.zeros(2, np.int64)
        self._check(self.L.mskg_frame_defects(self.h, None, 0, None, 0, ptr(c)))
        bufs = [np.zeros((max(int(c[k]), 1), 4), np.int32) for k in range(2)]
        self._check(self.L.mskg_frame_defects(self.h, ptr(bufs[0]), int(c[0]), ptr(bufs[1]), int(c[1]), ptr(c)))
        return bufs[0][:c[0]], bufs[1][:c[1]]

    def track_classes(self, on=True):
        """Keep the rate-class keys current through runs (likelihood segments)."""
        self._check(self.L.mskg_track_classes(self.h, 1 if on else 0))

    def copy_state_from(self, src):
        """This engine takes src's lattice, rates, class keys and clock."""
        self._check(self.L.mskg_copy_state(self.h, src.h))

    def set_event_log(self, capacity):
        self._check(self.L.mskg_set_event_log(self.h, int(capacity)))

    def event_log(self):
        n = C.c_int64()
        self._check(self.L.mskg_get_event_log(self.h, None, 0, C.byref(n)))
        log = np.zeros(max(n.value, 1), EVENT_DTYPE)
        self._check(self.L.mskg_get_event_log(self.h, ptr(log), n.value, C.byref(n)))
        return log[:n.value]

    def set_temperature(self, temp_k):
        t = np.ascontiguousarray(temp_k, np.int32)
        self._check(self.L.mskg_set_temperature(self.h, ptr(t)))

    def set_temperature_field(self, field, layout="socket"):
        """SetT of the laser-annealing build on the device: the FEM map in its
        socket layout (Fortran b(LenZ, LenX*LenY): numpy shape (LenX*LenY,
        LenZ), z fastest) or its file layout (numpy shape (LenZ, LenY,
        LenX)); value 10 leaves a point as it is, other values
        give their low 13 bits.  field: a numpy array, or an int32 torch
        tensor on the engine's device (read in place).  Rates are left as
        they are (refresh_rates())."""
        code = _layout(layout)
        shape = self._map_shape(code)
        dev = _device_tensor(field)
        if dev is not None:
            assert tuple(dev.shape) == shape and dev.dtype == _torch().int32 and dev.is_contiguous(), dev.shape
            _torch().cuda.current_stream(dev.device).synchronize()  # the engine reads it on its own stream
            self._check(self.L.mskg_set_temperature_field(self.h, C.c_void_p(dev.data_ptr()), code, 1))
            return
        f = np.ascontiguousarray(field, np.int32)
        assert f.shape == shape, (f.shape, shape)
        self._check(self.L.mskg_set_temperature_field(self.h, ptr(f), code, 0))

    def phases(self, layout="socket", device=False):
        """The phase maps the laser-annealing build returns to its FEM driver:
        (coor, species) int32 of numpy shape (LenX*LenY, LenZ) for the socket
        layout (Fortran c(LenZ, LenX*LenY)) or (LenZ, LenY, LenX) for the file
        layout; coordination and species of
        occupied points, 10 on walls, 0 elsewhere.  device=True returns torch
        tensors on the engine's device."""
        code = _layout(layout)
        shape = self._map_shape(code)
        if device:
            torch = _torch()
            out = [torch.empty(shape, dtype=torch.int32, device="cuda") for _ in range(2)]
            self._check(self.L.mskg_phases(self.h, C.c_void_p(out[0].data_ptr()), C.c_void_p(out[1].data_ptr()),
                                           code, 1))
            return out[0], out[1]
        # page-locked host buffers (torch's pinned allocator caches them):
        # the device-to-host copies run at PCIe speed instead of staging
        # through pageable memory; the arrays keep their buffers alive
        torch = _torch()
        out = [torch.empty(shape, dtype=torch.int32, pin_memory=True).numpy() for _ in range(2)]
        self._check(self.L.mskg_phases(self.h, ptr(out[0]), ptr(out[1]), code, 0))
        return out[0], out[1]

    def set_la_variant(self, p0, tm, sigma, tflex, init_count, count):
        """The laser-annealing build's rate form: per-species (up to 3)
        constants, InitCountCrystal and the current CountCrystal; see
        mskg_set_la_variant.  Call after load_state / build_slab."""
        def d3(v):
            return np.ascontiguousarray((list(v) + [0.0, 0.0, 0.0])[:3], np.float64)

        def i3(v):
            return np.ascontiguousarray((list(v) + [0, 0, 0])[:3], np.int64)
        arrs = [d3(p0), d3(tm), d3(sigma), d3(tflex), i3(init_count), i3(count)]
        self._check(self.L.mskg_set_la_variant(self.h, *[ptr(a) for a in arrs]))

    def set_st_tables(self, temps, tables):
        """InitSt 'ST': reference temperatures (increasing) and one table dict
        per temperature (ptrans_e, ptrans_d, ptrans_abs, ptrans_des as for
        the constructor); rates interpolate them at each site's temperature.
        An empty list switches it off."""
        n = len(temps)
        t = np.ascontiguousarray(temps, np.int32) if n else np.zeros(1, np.int32)

        def stack(key, size):
            if not n:
                return np.zeros(size, np.float64)
            return np.ascontiguousarray(np.concatenate([np.asarray(tb[key], np.float64).reshape(-1) for tb in tables]))
        arrs = [stack("ptrans_e", 12288), stack("ptrans_d", 9), stack("ptrans_abs", 12), stack("ptrans_des", 12)]
        self._check(self.L.mskg_set_st_tables(self.h, n, ptr(t), *[ptr(a) for a in arrs]))

    def species_counts(self):
        """(evaporations by species, desorptions by coverage species) since
        the state was loaded."""
        ev = np.zeros(3, np.int64)
        de = np.zeros(3, np.int64)
        self._check(self.L.mskg_species_counts(self.h, ptr(ev), ptr(de)))
        return ev, de

    def la_counts(self):
        """(CountCrystal, CountLiquid) of the laser-annealing build's form."""
        c = np.zeros(3, np.int64)
        q = np.zeros(3, np.int64)
        self._check(self.L.mskg_la_counts(self.h, ptr(c), ptr(q)))
        return c, q

    def _map_shape(self, code):
        lx, ly, lz = self.box
        return (lx * ly, lz) if code == MSKG_FIELD_SOCKET else (lz, ly, lx)

    def refresh_rates(self, active=False):
        """Get_Prob for every MC particle: over every stored point, or with
        active=True over the tile layers that hold MC particles (the whole
        lattice when the engine's tile lists are not current)."""
        self._check(self.L.mskg_refresh_active(self.h) if active else self.L.mskg_refresh_rates(self.h))

    def run(self, max_events, max_time=float("inf")):
        """Synchronous-sublattice KMC until max_events events or the clock
        reaches max_time (the last window is shortened to end there).
        Returns (events, clock, stop reason).  Warns when a tile window hit
        the per-window event cap (see diagnostics())."""
        n = C.c_int64()
        t = C.c_double()
        s = C.c_int32()
        mt = float(max_time) if np.isfinite(max_time) else 1e300
        self._check(self.L.mskg_run(self.h, int(max_events), mt, C.byref(n), C.byref(t), C.byref(s)))
        g = self.diagnostics()["guard_exits"]
        if g < self._guard_seen:  # a new state since: the count restarted
            self._guard_seen = 0
        if g > self._guard_seen:
            warnings.warn(f"{g - self._guard_seen} tile windows ended at the per-window event cap: the window "
                          "policy lets single tiles run far ahead (lower events_per_pass)", RuntimeWarning)
            self._guard_seen = g
        return n.value, t.value, s.value

    def set_schedule(self, schedule):
        """Order of a sweep's colour passes (mskg_set_schedule): "colours" (a
        barrier after each colour; the default) or "tiles" (each tile as soon
        as its lower-colour neighbours are done).  Same results bit for bit."""
        code = {"tiles": _lib.MSKG_SCHEDULE_TILES, "colours": _lib.MSKG_SCHEDULE_COLOURS}[schedule]
        self._check(self.L.mskg_set_schedule(self.h, code))

    def diagnostics(self):
        """Window diagnostics since the state was loaded: guard_exits (tile
        windows cut at the per-window event cap), sweeps, the next window
        tau and the MC particles at the last sweep."""
        g, k, m = C.c_int64(), C.c_int64(), C.c_int64()
        tau = C.c_double()
        self._check(self.L.mskg_run_diagnostics(self.h, C.byref(g), C.byref(k), C.byref(tau), C.byref(m)))
        return {"guard_exits": g.value, "sweeps": k.value, "tau": tau.value, "n_mc": m.value}

    def run_serial(self, max_events, max_time=float("inf")):
        """Serial BKL (one workgroup): see run_serial()."""
        n, t, s = run_serial([self], [max_events], max_time)
        return n[0], t[0], s[0]

    def replay(self, log, clock_mode=0):
        log = np.ascontiguousarray(log, EVENT_DTYPE)
        self._check(self.L.mskg_replay(self.h, ptr(log), len(log), int(clock_mode)))

    def store_state(self, lattcoo=None, capacity=None):
        """(LattCoo, MC list, bulk list, void list).  LattCoo is (LenZ+1,
        LenY, LenX), or its stored planes for an engine holding only the
        bottom ones; the lattice points are written, other grid points (and
        planes above the stored ones) keep what lattcoo held."""
        lx, ly, lz = self.box
        lat = (np.zeros((self.stored_planes, ly, lx), np.int32) if lattcoo is None
               else np.ascontiguousarray(lattcoo, np.int32))
        assert lat.shape[0] >= self.stored_planes and lat.shape[1:] == (ly, lx), lat.shape
        na, nb, nv = C.c_int32(), C.c_int32(), C.c_int32()
        z = None
        self.L.mskg_store_state(self.h, None, 0, C.byref(na), z, z, z, z, 0, C.byref(nb), z, z, 0, C.byref(nv), z, z)
        cap = capacity or (na.value, nb.value, nv.value)
        mc = {"xyz": np.zeros((max(cap[0], 1), 3), np.int32), "nn": np.zeros((max(cap[0], 1), 4, 3), np.int32),
              "event": np.zeros(max(cap[0], 1), np.int32), "prob": np.zeros(max(cap[0], 1), np.float64)}
        bulk = {"xyz": np.zeros((max(cap[1], 1), 3), np.int32), "nn": np.zeros((max(cap[1], 1), 4, 3), np.int32)}
        void = {"xyz": np.zeros((max(cap[2], 1), 3), np.int32), "nn": np.zeros((max(cap[2], 1), 4, 3), np.int32)}
        self._check(self.L.mskg_store_state(
            self.h, ptr(lat), cap[0], C.byref(na), ptr(mc["xyz"]), ptr(mc["nn"]), ptr(mc["event"]), ptr(mc["prob"]),
            cap[1], C.byref(nb), ptr(bulk["xyz"]), ptr(bulk["nn"]), cap[2], C.byref(nv), ptr(void["xyz"]),
            ptr(void["nn"])))
        for d, n in ((mc, na.value), (bulk, nb.value), (void, nv.value)):
            for k in list(d):
                d[k] = d[k][:n]
        return lat, mc, bulk, void

    def site_rates(self, xyz):
        a = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        cls = np.zeros(len(a), np.int32)
        rate = np.zeros(len(a), np.float32)
        self._check(self.L.mskg_site_rates(self.h, len(a), ptr(a), ptr(cls), ptr(rate)))
        return cls, rate

    def counters(self, lists=True):
        c = np.zeros(8, np.int64)
        t = C.c_double()
        na, nb, nv = C.c_int64(), C.c_int64(), C.c_int64()
        if lists:
            self._check(self.L.mskg_counters(self.h, ptr(c), C.byref(t), C.byref(na), C.byref(nb), C.byref(nv)))
            return c, t.value, (na.value, nb.value, nv.value)
        self._check(self.L.mskg_counters(self.h, ptr(c), C.byref(t), None, None, None))
        return c, t.value, None

    def last_timing(self):
        r, s = C.c_double(), C.c_double()
        n, k = C.c_int64(), C.c_int64()
        self._check(self.L.mskg_last_timing(self.h, C.byref(r), C.byref(s), C.byref(n), C.byref(k)))
        return {"run_ms": r.value, "sweep_ms": s.value, "sweep_points": n.value, "sweeps": k.value}


def find_nn(s1, s2, sc, lx, ly):
    """The engine's FIND_NN geometry evaluated on the host (no GPU needed)."""
    L = _lib.lib()
    a, b, c = (np.ascontiguousarray(v, np.int32) for v in (s1, s2, sc))
    out = np.zeros(6, np.int32)
    L.mskg_geometry_find_nn(ptr(a), ptr(b), ptr(c), int(lx), int(ly), ptr(out))
    return out.reshape(2, 3)


def find_sn(site, sitec, sn, lx, ly):
    """The engine's FIND_SN geometry evaluated on the host (no GPU needed)."""
    L = _lib.lib()
    a, c = np.ascontiguousarray(site, np.int32), np.ascontiguousarray(sitec, np.int32)
    n = np.ascontiguousarray(sn, np.int32).reshape(9)
    zig = np.zeros(9, np.int32)
    arm = np.zeros(9, np.int32)
    L.mskg_geometry_find_sn(ptr(a), ptr(c), ptr(n), int(lx), int(ly), ptr(zig), ptr(arm))
    return zig.reshape(3, 3), arm.reshape(3, 3)


def _engine_array(engines):
    arr = (C.c_void_p * len(engines))(*[e.h for e in engines])
    return arr


def run_serial(engines, max_events, max_time=float("inf")):
    """Run the serial BKL algorithm on several engines (independent lattices)
    in one launch, one workgroup each.  Returns per-engine (events, clock,
    stop reason) arrays."""
    L = _lib.lib()
    n = len(engines)
    me = np.ascontiguousarray(np.broadcast_to(np.asarray(max_events, np.int64), (n,)))
    mt = np.ascontiguousarray(np.broadcast_to(np.asarray(max_time, np.float64), (n,)))
    mt = np.where(np.isfinite(mt), mt, 1e300)
    done = np.zeros(n, np.int64)
    clock = np.zeros(n, np.float64)
    stop = np.zeros(n, np.int32)
    rc = L.mskg_run_serial(_engine_array(engines), n, ptr(me), ptr(mt), ptr(done), ptr(clock), ptr(stop))
    if rc:
        bad = [e for e in engines if L.mskg_last_error(e.h)]
        msg = "; ".join(L.mskg_last_error(e.h).decode() for e in bad) or "mskg_run_serial failed"
        raise EngineError(rc, msg)
    return done, clock, stop


def run_ensemble(engines, max_events, max_time=float("inf")):
    """The batched ensemble (mskg_run_ensemble): the synchronous-sublattice
    run of every engine at once, each colour pass one launch over the tiles
    of all running replicas.  Each engine ends as mskg_run alone would leave
    it.  Returns per-engine (events, clock, stop reason) arrays."""
    L = _lib.lib()
    n = len(engines)
    me = np.ascontiguousarray(np.broadcast_to(np.asarray(max_events, np.int64), (n,)))
    mt = np.ascontiguousarray(np.broadcast_to(np.asarray(max_time, np.float64), (n,)))
    mt = np.where(np.isfinite(mt), mt, 1e300)
    done = np.zeros(n, np.int64)
    clock = np.zeros(n, np.float64)
    stop = np.zeros(n, np.int32)
    rc = L.mskg_run_ensemble(_engine_array(engines), n, ptr(me), ptr(mt), ptr(done), ptr(clock), ptr(stop))
    if rc:
        raise EngineError(rc, L.mskg_last_error(engines[0].h).decode() or "mskg_run_ensemble failed")
    return done, clock, stop


def adjoint_stats(engines, logs, dts, every=0):
    """Likelihood statistics of recorded trajectories (mskg_adjoint): engine i
    holds the initial state of trajectory i, logs[i] its events and dts[i]
    the waiting time before each.  Returns S (n, ADJ_NS), A (n, KEY_N) and
    scalars (n, 4) = (sum dt, isolated-rate integral, isolated log-rate sum,
    events replayed); with every > 0 also the checkpoints
    (mskg_adjoint_checkpointed): for each trajectory an int32 array
    (len(log) // every, KEY_N) of the class populations after every
    `every` events.  The engines end in the final states."""
    L = _lib.lib()
    n = len(engines)
    logs = [np.ascontiguousarray(lg, EVENT_DTYPE) for lg in logs]
    dts = [np.ascontiguousarray(d, np.float64) for d in dts]
    for lg, d in zip(logs, dts):
        if len(lg) != len(d):
            raise ValueError("each log needs one waiting time per event")
    lp = (C.c_void_p * n)(*[lg.ctypes.data for lg in logs])
    dp = (C.c_void_p * n)(*[d.ctypes.data for d in dts])
    ne = np.array([len(lg) for lg in logs], np.int64)
    S = np.zeros((n, _lib.ADJ_NS))
    A = np.zeros((n, _lib.KEY_N))
    sc = np.zeros((n, 4))
    if every > 0:
        hist = [np.zeros((len(lg) // every, _lib.KEY_N), np.int32) for lg in logs]
        hp = (C.c_void_p * n)(*[h.ctypes.data for h in hist])
        rc = L.mskg_adjoint_checkpointed(_engine_array(engines), n, lp, dp, ptr(ne), ptr(S), ptr(A), ptr(sc),
                                         int(every), hp)
    else:
        rc = L.mskg_adjoint(_engine_array(engines), n, lp, dp, ptr(ne), ptr(S), ptr(A), ptr(sc))
    if rc:
        msg = "; ".join(L.mskg_last_error(e.h).decode() for e in engines if L.mskg_last_error(e.h))
        raise EngineError(rc, msg or "mskg_adjoint failed")
    if every > 0:
        return S, A, sc, hist
    return S, A, sc
