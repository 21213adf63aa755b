"""ctypes wrapper of the C oracle (kmc_oracle.c) -- test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libkmc_oracle.so")
CLI = os.path.join(HERE, "kmc_oracle_cli")
NUMBERSLIST = os.path.join(os.path.dirname(HERE), "tests", "golden", "numberslist.txt")

EVENT_DTYPE = np.dtype([("iter", "<i8"), ("site", "<i4", (3,)), ("itrans", "<i4"), ("arm", "<i4"),
                        ("coor", "<i4"), ("rtot", "<f8"), ("time", "<f8")])
assert EVENT_DTYPE.itemsize == 48


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def numberslist():
    return np.loadtxt(NUMBERSLIST, dtype=np.float64)


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I, D, I64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        L.orc_create.restype = P
        L.orc_create.argtypes = [I, I, I]
        L.orc_destroy.argtypes = [P]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_error.argtypes = [P]
        L.orc_set_numberslist.argtypes = [P, P, I]
        L.orc_load_deck.argtypes = [P, C.c_char_p]
        L.orc_build_cad.argtypes = [P, P, P]
        L.orc_main.argtypes = [P, C.c_char_p, P, I64, P, P]
        L.orc_step.argtypes = [P, P]
        L.orc_apply.argtypes = [P, P, I, I]
        L.orc_set_frozen.argtypes = [P, I]
        L.orc_load_state.argtypes = [P, P, I, P, P, P, P, I, P, P, I, P, P, D]
        L.orc_set_second_factor.argtypes = [P, P]
        L.orc_refresh_all.argtypes = [P]
        L.orc_set_la.argtypes = [P, D, D, D, D]
        L.orc_set_la_variant.argtypes = [P, P, P, P, P, P, P, I]
        L.orc_la_counts.argtypes = [P, P, P]
        L.orc_set_st_tables.argtypes = [P, I, P, P, P, P, P]
        L.orc_run.argtypes = [P, I64, D, P]
        L.orc_run_to.argtypes = [P, I64, D, P]
        L.orc_executed.argtypes = [P, P]
        L.orc_adjoint_stats.argtypes = [P, P, I64, P, P, P, P]
        L.orc_adjoint_checkpoints.argtypes = [P, P, I64, P, P, P, P, I64, P]
        L.orc_write_frame.argtypes = [P, C.c_char_p, C.c_char_p, I64]
        L.orc_write_frames.argtypes = [P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, I64]
        L.orc_set_temperature_field.argtypes = [P, P, I]
        L.orc_phases.argtypes = [P, P, P, I]
        L.orc_lattcoo.restype = C.POINTER(C.c_int32)
        L.orc_lattcoo.argtypes = [P]
        L.orc_dims.argtypes = [P, P, P, P]
        L.orc_counts.argtypes = [P, P, P, P]
        L.orc_get_list.argtypes = [P, I, P, P, P, P]
        L.orc_time.restype = D
        L.orc_time.argtypes = [P]
        L.orc_total_rate.restype = D
        L.orc_total_rate.argtypes = [P]
        L.orc_iter.restype = I64
        L.orc_iter.argtypes = [P]
        L.orc_count_crystal.argtypes = [P, P, P]
        L.orc_ncrystal.argtypes = [P, P]
        L.orc_tables.argtypes = [P, P, P, P, P, P, P, P]
        L.orc_pure_rate.argtypes = [P, P, P, P]
        L.orc_find_nn.argtypes = [P, P, P, P, I, I]
        L.orc_find_sn.argtypes = [P, P, P, P, P, I, I]
        L.orc_anomalies.argtypes = [P, P]
        L.orc_ran3_next.restype = D
        L.orc_ran3_next.argtypes = [P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One reference-faithful serial KMC simulation on the CPU."""

    def __init__(self, lx, ly, lz):
        self.L = _load()
        self.h = self.L.orc_create(lx, ly, lz)
        if not self.h:
            raise MemoryError("orc_create failed")
        self.dims = (lx, ly, lz)
        nl = numberslist()
        self.L.orc_set_numberslist(self.h, _p(nl), len(nl))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what}: {self.L.orc_last_error(self.h).decode()}")

    def load_deck(self, path):
        self._check(self.L.orc_load_deck(self.h, os.fsencode(path)), "orc_load_deck")

    def load_state(self, lattcoo, mc, bulk, void, kmc_time=0.0):
        """Replace the state by a handed-over one (the engine's store_state
        layout): LattCoo (LenZ+1, LenY, LenX) and the three list dicts. The
        tables and options stay; refresh_all() re-evaluates the rates."""
        lat = np.ascontiguousarray(lattcoo, np.int32)
        assert lat.shape == (self.dims[2] + 1, self.dims[1], self.dims[0]), lat.shape

        def arr(d, k, dt, shape):
            a = np.ascontiguousarray(d[k], dt)
            return a if len(a) else np.zeros(shape, dt)
        keep = [arr(mc, "xyz", np.int32, (1, 3)), arr(mc, "nn", np.int32, (1, 4, 3)),
                arr(mc, "event", np.int32, (1,)), arr(mc, "prob", np.float64, (1,)),
                arr(bulk, "xyz", np.int32, (1, 3)), arr(bulk, "nn", np.int32, (1, 4, 3)),
                arr(void, "xyz", np.int32, (1, 3)), arr(void, "nn", np.int32, (1, 4, 3))]
        rc = self.L.orc_load_state(self.h, _p(lat), len(mc["xyz"]), *[_p(a) for a in keep[:4]],
                                   len(bulk["xyz"]), _p(keep[4]), _p(keep[5]), len(void["xyz"]), _p(keep[6]),
                                   _p(keep[7]), float(kmc_time))
        self._check(rc, "orc_load_state")

    def build_cad(self, ids, species2=None):
        """SetCAD (InitSt 'IN'): ids int8 (LenZ, LenY, LenX) geometry ids
        (10 wall, 1 atom on a diamond site); species2: None or a bool array
        of the same shape, species 2 where set."""
        lx, ly, lz = self.dims
        a = np.ascontiguousarray(ids, np.int8)
        assert a.shape == (lz, ly, lx), a.shape
        b = None if species2 is None else np.ascontiguousarray(species2, np.int8)
        assert b is None or b.shape == a.shape
        self._check(self.L.orc_build_cad(self.h, _p(a), None if b is None else _p(b)), "orc_build_cad")

    def run_main(self, outdir=None, log_cap=0):
        log = np.zeros(max(log_cap, 1), EVENT_DTYPE)
        n = C.c_int64(0)
        stop = C.c_int(0)
        rc = self.L.orc_main(self.h, os.fsencode(outdir) if outdir else None, _p(log) if log_cap else None,
                             log_cap, C.byref(n), C.byref(stop))
        self._check(rc, "orc_main")
        return log[:n.value] if log_cap else None, stop.value

    def step(self, n, record=True):
        """Run n iterations of the KMC loop; returns the event log."""
        log = np.zeros(n, EVENT_DTYPE)
        done = 0
        for k in range(n):
            rc = self.L.orc_step(self.h, _p(log[k:k + 1]) if record else None)
            if rc == 1:
                break
            self._check(rc, "orc_step")
            done += 1
        return log[:done]

    def set_frozen(self, z):
        self._check(self.L.orc_set_frozen(self.h, int(z)), "orc_set_frozen")

    def set_second_factor(self, f2):
        """Second-neighbour evaporation factors f2[0..12] (None = off)."""
        if f2 is None:
            rc = self.L.orc_set_second_factor(self.h, None)
        else:
            self._f2 = np.ascontiguousarray(f2, np.float64)
            assert self._f2.shape == (13,)
            rc = self.L.orc_set_second_factor(self.h, _p(self._f2))
        self._check(rc, "orc_set_second_factor")

    def run(self, max_events, max_time=float("inf"), at_time=False):
        """Serial loop until max_events or the clock reaches max_time;
        returns (events, stop) with stop 0 limit, 1 few MC particles.
        at_time=True stops before the first event past max_time instead of
        after it (the state at the KMC time max_time)."""
        done = C.c_int64(0)
        f = self.L.orc_run_to if at_time else self.L.orc_run
        rc = f(self.h, int(max_events), float(max_time) if np.isfinite(max_time) else 1e300, C.byref(done))
        if rc == 2:
            self._check(rc, "orc_run")
        return done.value, rc

    def executed(self):
        """Events executed by type (the engine's counters layout)."""
        out = np.zeros(8, np.int64)
        self.L.orc_executed(self.h, _p(out))
        return out

    def set_la(self, tm, p0, sigma, tflex):
        """Laser-annealing rates from the T bits; the tables hold energies."""
        self._check(self.L.orc_set_la(self.h, float(tm), float(p0), float(sigma), float(tflex)), "orc_set_la")

    def set_la_variant(self, p0, tm, sigma, tflex, init_count, count, swap):
        """The laser-annealing build's rate form (per-species constants,
        liquid fraction on solidification, SiGe species swap if swap)."""
        d3 = [np.ascontiguousarray((list(v) + [0.0] * 3)[:3], np.float64) for v in (p0, tm, sigma, tflex)]
        i3 = [np.ascontiguousarray((list(v) + [0] * 3)[:3], np.int32) for v in (init_count, count)]
        self._keep_la = d3 + i3
        self._check(self.L.orc_set_la_variant(self.h, *[_p(a) for a in d3 + i3], 1 if swap else 0),
                    "orc_set_la_variant")

    def set_st_tables(self, temps, tables):
        """InitSt 'ST': reference temperatures and one table dict each."""
        n = len(temps)
        self._keep_st = [np.ascontiguousarray(temps, np.int32)] + [
            np.ascontiguousarray(np.concatenate([np.asarray(tb[k], np.float64).reshape(-1) for tb in tables]))
            for k in ("ptrans_e", "ptrans_d", "ptrans_abs", "ptrans_des")]
        self._check(self.L.orc_set_st_tables(self.h, n, *[_p(a) for a in self._keep_st]), "orc_set_st_tables")

    def la_counts(self):
        c = np.zeros(3, np.int32)
        q = np.zeros(3, np.int32)
        self.L.orc_la_counts(self.h, _p(c), _p(q))
        return c, q

    def refresh_all(self):
        """Get_Prob for every MC particle (fresh rates and class keys)."""
        self._check(self.L.orc_refresh_all(self.h), "orc_refresh_all")

    def adjoint_stats(self, log, dt, every=0):
        """Brute-force likelihood statistics of a trajectory replayed from the
        current state: S (2510,), A (2501,), scal (4,) = (sum dt, isolated-rate
        integral, isolated log-rate sum, direct log-likelihood); with every > 0
        also the class populations after every `every` events, (len // every,
        2501) int32."""
        log = np.ascontiguousarray(log, EVENT_DTYPE)
        dt = np.ascontiguousarray(dt, np.float64)
        assert len(dt) == len(log)
        S = np.zeros(2510)
        A = np.zeros(2501)
        sc = np.zeros(4)
        if every > 0:
            hist = np.zeros((len(log) // every, 2501), np.int32)
            self._check(self.L.orc_adjoint_checkpoints(self.h, _p(log), len(log), _p(dt), _p(S), _p(A), _p(sc),
                                                       int(every), _p(hist)), "orc_adjoint_checkpoints")
            return S, A, sc, hist
        self._check(self.L.orc_adjoint_stats(self.h, _p(log), len(log), _p(dt), _p(S), _p(A), _p(sc)),
                    "orc_adjoint_stats")
        return S, A, sc

    def apply_log(self, log):
        """Execute a logged event sequence (e.g. the engine's) in order."""
        for k, ev in enumerate(log):
            site = np.ascontiguousarray(ev["site"], np.int32)
            rc = self.L.orc_apply(self.h, _p(site), int(ev["itrans"]), int(ev["arm"]))
            if rc:
                raise RuntimeError(f"event {k}: {self.L.orc_last_error(self.h).decode()}")

    def set_temperature_field(self, field, layout="socket"):
        """SetT of the laser-annealing build (layout 'socket': numpy shape
        (LenX*LenY, LenZ), z fastest; 'planes': (LenZ, LenY, LenX))."""
        code = {"socket": 0, "planes": 1}[layout]
        f = np.ascontiguousarray(field, np.int32)
        lx, ly, lz = self.dims
        assert f.size == lx * ly * lz
        self._check(self.L.orc_set_temperature_field(self.h, _p(f), code), "orc_set_temperature_field")

    def phases(self, layout="socket"):
        """The phase maps (coor, species) of the laser-annealing build."""
        code = {"socket": 0, "planes": 1}[layout]
        lx, ly, lz = self.dims
        shape = (lx * ly, lz) if code == 0 else (lz, ly, lx)
        c = np.empty(shape, np.int32)
        d = np.empty(shape, np.int32)
        self._check(self.L.orc_phases(self.h, _p(c), _p(d), code), "orc_phases")
        return c, d

    def lattcoo(self):
        lx, ly, lz = self.dims
        ptr = self.L.orc_lattcoo(self.h)
        return np.ctypeslib.as_array(ptr, shape=(lz + 1, ly, lx)).copy()

    def counts(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self.L.orc_counts(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def lists(self):
        """(mc, bulk, void) as dicts of numpy arrays, in list order."""
        na, nb, nv = self.counts()
        out = []
        for which, n in ((0, na), (1, nb), (2, nv)):
            xyz = np.zeros((max(n, 1), 3), np.int32)
            nn = np.zeros((max(n, 1), 4, 3), np.int32)
            ev = np.zeros(max(n, 1), np.int32)
            pr = np.zeros(max(n, 1), np.float64)
            self.L.orc_get_list(self.h, which, _p(xyz), _p(nn), _p(ev), _p(pr))
            d = {"xyz": xyz[:n], "nn": nn[:n]}
            if which == 0:
                d["event"] = ev[:n]
                d["prob"] = pr[:n]
            out.append(d)
        return out

    def anomalies(self):
        a = np.zeros(3, np.int64)
        self.L.orc_anomalies(self.h, _p(a))
        return {"nonbond_arm": int(a[0]), "nonbond_nn": int(a[1]), "oob_reads": int(a[2])}

    def time(self):
        return self.L.orc_time(self.h)

    def total_rate(self):
        return self.L.orc_total_rate(self.h)

    def iteration(self):
        return self.L.orc_iter(self.h)

    def count_crystal(self):
        cc = np.zeros(3, np.int32)
        cv = np.zeros(3, np.int32)
        self.L.orc_count_crystal(self.h, _p(cc), _p(cv))
        return cc, cv

    def tables(self):
        pe = np.zeros(3 * 4096)
        pd = np.zeros(9)
        pabs = np.zeros(12)
        pdes = np.zeros(12)
        pz = C.c_double()
        lc = np.zeros(3, np.int32)
        lcov = np.zeros(3, np.int32)
        self.L.orc_tables(self.h, _p(pe), _p(pd), _p(pabs), _p(pdes), C.byref(pz), _p(lc), _p(lcov))
        ncov = C.c_int()
        ncr = self.L.orc_ncrystal(self.h, C.byref(ncov))
        return dict(ptrans_e=pe, ptrans_d=pd, ptrans_abs=pabs, ptrans_des=pdes, ptrans_zig=pz.value,
                    list_crystal=lc, list_cov=lcov, n_crystal=ncr, n_cov=ncov.value)

    def pure_rate(self, xyz):
        cat = C.c_int32()
        r = C.c_double()
        a = np.asarray(xyz, np.int32)
        self._check(self.L.orc_pure_rate(self.h, _p(a), C.byref(cat), C.byref(r)), "orc_pure_rate")
        return cat.value, r.value

    def write_frame(self, xyz_path, d_path=None, it=0):
        self._check(self.L.orc_write_frame(self.h, os.fsencode(xyz_path), os.fsencode(d_path) if d_path else None,
                                           it), "orc_write_frame")

    def write_frames(self, xyz_path=None, d_path=None, v_path=None, w_path=None, it=0):
        """The main, _d, _v and _w frames of WriteMolMolXYZFile.f (any path may be None)."""
        enc = lambda p: os.fsencode(p) if p else None  # noqa: E731
        self._check(self.L.orc_write_frames(self.h, enc(xyz_path), enc(d_path), enc(v_path), enc(w_path), it),
                    "orc_write_frames")


def find_nn(s1, s2, sc, lx, ly):
    L = _load()
    a, b, c = (np.asarray(v, np.int32) for v in (s1, s2, sc))
    out = np.zeros((2, 3), np.int32)
    L.orc_find_nn(_p(a), _p(b), _p(c), _p(out), lx, ly)
    return out


def find_sn(site, sitec, sn, lx, ly):
    L = _load()
    a, b = np.asarray(site, np.int32), np.asarray(sitec, np.int32)
    s = np.ascontiguousarray(sn, np.int32)
    zig = np.zeros((3, 3), np.int32)
    arm = np.zeros((3, 3), np.int32)
    L.orc_find_sn(_p(a), _p(b), _p(s), _p(zig), _p(arm), lx, ly)
    return zig, arm


def ran3_sequence(idum, n):
    L = _load()
    st = np.zeros(59, np.int32)  # orc_ran3: ma[56], inext, inextp, iff
    d = C.c_int32(idum)
    return np.array([L.orc_ran3_next(_p(st), C.byref(d)) for _ in range(n)])
