"""start.dat input decks: parser and writer.

Restates the reference's input grammar: the positional, list-directed header
read by mulskips-source/main.f:83-212 and the rate blocks read by
mcevent/SetProbability.f:21-143, in the form pymulskips writes them
(pymulskips/setuprun.py:292-389).  The engine needs only the rate tables
and species lists (mskg_config); the rest of the header drives the host.
"""
import dataclasses
import re

import numpy as np

# Number of evaporation lines per (NCrystal, NCov), SetProbability.f:47-75.
MAX_ELINES = {(1, 0): 3, (2, 0): 18, (3, 0): 57, (1, 1): 9, (2, 1): 50, (3, 1): 150,
              (1, 2): 19, (2, 2): 100, (3, 2): 288, (1, 3): 34, (2, 3): 172, (3, 3): 480}


@dataclasses.dataclass
class Deck:
    n_crystal: int
    n_cov: int
    list_crystal: list
    list_cov: list
    init_st: str
    lens: list                 # Len1..Len4 (or the LA block for 'LA')
    ptrans_zig: float
    sf: float
    exit_strategy: str         # 'Iter' or 'Time'
    max_iter: int = 0
    out_molmol: int = 0
    max_time: float = 0.0
    out_time: float = 0.0
    run_type: str = "R"
    idum: int = 1
    save_final_state: str = "F"
    save_coo: str = "F"
    alloy_fraction: float = 0.0
    evap: list = dataclasses.field(default_factory=list)  # (indsp, digits-string, value)
    depo: list = dataclasses.field(default_factory=list)  # (indsp, coor, value)
    absn: list = dataclasses.field(default_factory=list)  # (indsp, coor, value)
    deso: list = dataclasses.field(default_factory=list)
    extra_lines: dict = dataclasses.field(default_factory=dict)

    # ---- rate tables as the engine takes them (Fortran order) -----------
    def tables(self):
        pe = np.zeros(3 * 4 ** 6)
        nd = self.n_crystal + self.n_cov
        for indsp, digits, value in self.evap:
            elind = int(digits)
            ind = [0] * 6
            for j in range(1, self.n_crystal + 1):
                ind[j - 1] = (elind // 10 ** (nd - j)) % 10
            for j in range(self.n_crystal + 1, nd + 1):
                ind[3 + j - self.n_crystal - 1] = (elind // 10 ** (nd - j)) % 10
            flat = (indsp - 1) + 3 * (ind[0] + 4 * (ind[1] + 4 * (ind[2] + 4 * (ind[3] + 4 * (ind[4] + 4 * ind[5])))))
            pe[flat] = value
        pd = np.zeros(9)
        for indsp, coor, value in self.depo:
            pd[(indsp - 1) + 3 * (coor - 1)] = value
        pabs = np.zeros(12)
        for indsp, coor, value in self.absn:
            pabs[(indsp - self.n_crystal - 1) + 3 * coor] = value
        pdes = np.zeros(12)
        for indsp, coor, value in self.deso:
            pdes[(indsp - self.n_crystal - 1) + 3 * coor] = value
        lc = (list(self.list_crystal) + [0, 0, 0])[:3]
        lcov = (list(self.list_cov) + [0, 0, 0])[:3]
        return dict(ptrans_e=pe, ptrans_d=pd, ptrans_abs=pabs, ptrans_des=pdes, ptrans_zig=self.ptrans_zig,
                    n_crystal=self.n_crystal, n_cov=self.n_cov, list_crystal=np.array(lc, np.int32),
                    list_cov=np.array(lcov, np.int32))


def _fields(line):
    """Tokens of a list-directed record (the reference reads only the leading
    items it needs; everything after them, comments included, is skipped)."""
    return re.split(r"[\s,]+", line.strip())


def _real(tok):
    return float(tok.replace("d", "e").replace("D", "e"))


def read_start_dat(path):
    with open(path) as f:
        lines = f.read().splitlines()
    it = iter(lines)
    nxt = lambda: next(it)  # noqa: E731
    t = _fields(nxt())
    ncr, ncov = int(t[0]), int(t[1])
    if not (1 <= ncr <= 3) or not (0 <= ncov <= 3):
        raise ValueError("NCrystal must be 1..3 and NCov 0..3")
    lc = [int(v) for v in _fields(nxt())[:ncr]]
    lcov = [int(v) for v in _fields(nxt())[:ncov]] if ncov else []
    init_st = _fields(nxt())[0][:2]
    extra = {}
    if init_st == "FF":
        lens = [int(v) for v in _fields(nxt())[:4]]
    elif init_st == "ST":
        t = _fields(nxt())
        lens = [int(t[0]), int(t[1]), int(t[2]), _real(t[3]), _real(t[4]), _real(t[5])]
    elif init_st in ("IN", "SH"):
        extra["cadfilename"] = nxt()
        lens = []
    elif init_st == "LA":
        extra["cadfilename"] = nxt()
        extra["tempfilename"] = nxt()
        lens = [_real(_fields(nxt())[0]) for _ in range(6)]
        extra["homogeneous"] = _fields(nxt())[0]
    else:
        lens = [int(v) for v in _fields(nxt())[:3]]
    alloy = 0.0
    if lc[:2] == [14, 32] and (ncr == 2 or (ncr == 3 and lc[2] == 5)):
        alloy = _real(_fields(nxt())[0])  # main.f:142-148
    zig = _real(_fields(nxt())[0])
    sf = _real(_fields(nxt())[0])
    exit_s = _fields(nxt())[0][:4]
    d = Deck(ncr, ncov, lc, lcov, init_st, lens, zig, sf, exit_s, alloy_fraction=alloy, extra_lines=extra)
    if exit_s == "Iter":
        d.max_iter = int(_fields(nxt())[0])
        d.out_molmol = int(_fields(nxt())[0])
    elif exit_s == "Time":
        d.max_time = _real(_fields(nxt())[0])
        d.out_time = _real(_fields(nxt())[0])
    d.run_type = _fields(nxt())[0][:2]
    d.idum = int(_fields(nxt())[0])
    d.save_final_state = _fields(nxt())[0][:2]
    if d.save_final_state == "T":
        extra["restartfilename"] = nxt()
    d.save_coo = _fields(nxt())[0][:2]
    if d.save_coo == "T":
        extra["coofilename"] = nxt()
    for _ in range(MAX_ELINES[(ncr, ncov)]):
        t = _fields(nxt())
        indsp = int(t[0])
        if indsp > ncr:
            raise ValueError("1st index in evaporation block must be <= NCrystal")
        d.evap.append((indsp, t[1], _real(t[2])))
    for _ in range(3 * ncr):
        t = _fields(nxt())
        d.depo.append((int(t[0]), int(t[1]), _real(t[2])))
    for blk in (d.absn, d.deso):
        for _ in range(3 * ncov):
            t = _fields(nxt())
            blk.append((int(t[0]), int(t[1]), _real(t[2])))
    return d


def write_start_dat(d, path):
    """Write a deck in the grammar read_start_dat (and main.f) accepts."""
    out = [f"{d.n_crystal} {d.n_cov} ! NCrystal, NCov",
           " ".join(str(v) for v in d.list_crystal) + " ! crystal species"]
    if d.n_cov:
        out.append(" ".join(str(v) for v in d.list_cov) + " ! coverage species")
    out.append(f"{d.init_st} ! InitSt")
    if d.init_st in ("IN", "SH"):
        out.append(d.extra_lines["cadfilename"])
    elif d.init_st == "LA":
        out += [d.extra_lines["cadfilename"], d.extra_lines["tempfilename"]]
        out += [repr(float(v)) for v in d.lens[:6]]
        out.append(d.extra_lines.get("homogeneous", "F"))
    else:
        out.append(" ".join(str(v) for v in d.lens) + " ! Len1 Len2 Len3 [Len4]")
    if d.list_crystal[:2] == [14, 32]:
        out.append(repr(float(d.alloy_fraction)) + " ! alloy fraction")
    out.append(repr(float(d.ptrans_zig)) + " ! PtransZig")
    out.append(repr(float(d.sf)) + " ! KMC_sf")
    out.append(f"{d.exit_strategy} ! ExitStrategy")
    if d.exit_strategy == "Iter":
        out += [f"{d.max_iter} ! MaxIter", f"{d.out_molmol} ! OutMolMol"]
    else:
        out += [f"{repr(float(d.max_time))} ! MaxTime", f"{repr(float(d.out_time))} ! OutTime"]
    out += [f"{d.run_type} ! RunType", f"{d.idum} ! IDUM", f"{d.save_final_state} ! SaveFinalState"]
    if d.save_final_state == "T":
        out.append(d.extra_lines["restartfilename"])
    out.append(f"{d.save_coo} ! SaveCoo")
    if d.save_coo == "T":
        out.append(d.extra_lines["coofilename"])
    out += [f"{a} {b} {repr(float(v))}" for a, b, v in d.evap]
    out += [f"{a} {b} {repr(float(v))}" for a, b, v in d.depo]
    out += [f"{a} {b} {repr(float(v))}" for a, b, v in d.absn]
    out += [f"{a} {b} {repr(float(v))}" for a, b, v in d.deso]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def deck_from_tables(tables, init_st, lens, max_iter=10 ** 9, run_type="R", idum=1, sf=0.38583333333333333):
    """A start.dat deck carrying the given rate tables (the inverse of
    Deck.tables(): one evaporation line per species and neighbour composition
    with 1..3 neighbours, as SetProbability.f:79-100 expects)."""
    ncr, ncov = int(tables["n_crystal"]), int(tables["n_cov"])
    if ncov:
        raise ValueError("coverage species are not supported by deck_from_tables")
    pe = np.asarray(tables["ptrans_e"])
    evap = []
    for s in range(1, ncr + 1):
        ranges = [range(4)] * ncr
        import itertools
        for comb in itertools.product(*ranges):
            if not 1 <= sum(comb) <= 3:
                continue
            n = list(comb) + [0] * (3 - ncr)
            flat = (s - 1) + 3 * (n[0] + 4 * (n[1] + 4 * n[2]))
            evap.append((s, "".join(str(c) for c in comb), float(pe[flat])))
    if len(evap) != ncr * MAX_ELINES[(ncr, 0)] // ncr:
        raise ValueError("unexpected number of evaporation lines")
    pd = np.asarray(tables["ptrans_d"])
    depo = [(s, c, float(pd[(s - 1) + 3 * (c - 1)])) for s in range(1, ncr + 1) for c in (1, 2, 3)]
    lc = [int(v) for v in tables["list_crystal"][:ncr]]
    return Deck(ncr, 0, lc, [], init_st, list(lens), float(tables["ptrans_zig"]), sf, "Iter", max_iter=max_iter,
                out_molmol=max_iter, run_type=run_type, idum=idum, evap=evap, depo=depo)


# ---- start.dat as pymulskips writes it --------------------------------------
# The header comments and line formats of setuprun.py:292-389, kept
# character for character: a deck written here is byte-identical to the one
# pymulskips writes for the same run, so either can drive the host.
INITSTAT_NOTE = (" ! Initstat: S Sphere, C Parellelepipid, F Flat SiC-3C(100) surface, FS Flat Si(100) surface, "
                 "A APB, I inverted pyramid, D inverted pyramid of C, Z inverted pyramid of Si, J inverted pyramid "
                 "with APB, Fw Flat SiC(100) surface with walls, Fa Flat SiC(100) surface with aperture, SL Si(100) "
                 "liquid-solid interface, FF FinFET Si (100) liquid-solid, SS Sphere Si liquid-solid")
PROB_FLOOR = 5e-12  # setuprun.py:355: every rate written is at least this


def _fmt15(v):
    return "{:.15f}".format(v)


def pymulskips_start_dat(run):
    """The text of the start.dat pymulskips' run_mulskips writes
    (setuprun.py:292-389) for the run described by `run`, a dict of its
    arguments and process data: Simulation, listcryZ, listcovZ, KMC_sf,
    Seed_box, PtransZig, ExitStrategy with IterMax / OutMolMol or TotTime /
    OutTime, RunType, IDUM, SaveFinalState (+ restartfilename), SaveCoo (+
    coofilename), path (the working directory the file names are joined to),
    the rate arrays PtransE (species, n_1, ..) / PtransD (species, 3) and for
    coverage species PtransAbs / PtransDes (cov, 3) with the evaporation index
    list `indices`; for 'LA' also driver, cadfilename, tempfilename, Tm, P0,
    A, Tflex, LenVac, LenNuc, initialState, LenSiGe; for SiGe X0 (the Ge
    fraction); for 'ST' useProbTable / ProbTable; for 'IN' / 'SH'
    cadfilename.  Every rate is written floored at 5e-12 with Python's str()
    of the float."""
    sim = run["Simulation"]
    cz, vz = list(run["listcryZ"]), list(run.get("listcovZ", []))
    ncry, ncov = len(cz), len(vz)
    path = run.get("path", ".")
    driver = run.get("driver")
    o = [f"{ncry} {ncov} ! NCrystal, NCov \n",
         "{} ! Atomic numbers of crystal species \n".format(" ".join(str(x) for x in cz))]
    if ncov:
        o.append("{} ! Atomic numbers of coverage species \n".format(" ".join(str(x) for x in vz)))
    o.append(sim + INITSTAT_NOTE + "\n")
    if sim in ("IN", "SH"):
        o.append(path + "/" + run["cadfilename"] + "\n")
    elif sim == "LA":
        if driver is None:
            o += [path + "/" + run["cadfilename"] + "\n", path + "/" + run["tempfilename"] + "\n"]
        else:
            o += ["None\n", "None\n"]
        o.append("{} ! Melting temperature in Kelvin \n".format(" ".join(str(x) for x in run["Tm"])))
        o.append("{} ! prefactor P0 \n".format(" ".join(str(x) for x in run["P0"])))
        o.append("{} ! errf param A \n".format(" ".join(str(x) for x in run["A"])))
        o.append("{} ! errf param Tflex \n".format(" ".join(str(x) for x in run["Tflex"])))
        o.append(_fmt15(run["LenVac"]) + " ! LenVac -> Top air thickness in Ang\n")
        o.append(_fmt15(run["LenNuc"]) + " ! LenNuc -> Nucleus radius in Ang\n")
        flag = {"homogeneous": "T", "inhomogeneous": "F", "tripleSF": "D", "stripe": "R"}[run["initialState"]]
        o.append(flag + " ! Initial Nucleation State [homogeneous, inhomogeneous, tripleSF, stripe] \n")
        if run["initialState"] == "tripleSF":
            o.append(" ".join(str(v) for v in run["Seed_box"]) + " ! Len1 Len2 Len3 Len4 Len5 for configuring a triple SF\n")
    else:
        o.append(" ".join(str(v) for v in run["Seed_box"]) + " ! Len1 [Len2 Len3 Len4 Len5]\n")
    if 14 in cz and 32 in cz:
        o.append(str(run["X0"]) + " ! Ge fraction in substrate \n")
        if sim == "LA":
            o.append(_fmt15(run["LenSiGe"]) + " ! LenSiGe -> SiGe thickness in Ang\n")
    o.append(str(run["PtransZig"]) + " ! PtransZig\n")
    o.append(_fmt15(run["KMC_sf"]) + " ! KMC_sf -> KMC Super-Lattice parameter in Ang\n")
    ex = run["ExitStrategy"]
    o.append(ex + " ! ExitStrategy. Iter or Time. Decides whether to use iterations or time to check simulation "
                  "end and output frequency\n")
    if ex == "Iter":
        o.append(str(run["IterMax"]) + " ! IterMax -> Max number of iterations\n")
        o.append(str(run["OutMolMol"]) + " ! OutMolMol -> output frequency\n")
    if ex == "Time":
        o.append(_fmt15(run["TotTime"]) + " ! TotTime -> Max time of simulation in seconds\n")
        o.append(_fmt15(run["OutTime"]) + " ! OutTime -> Delta Time between consecutive print, in seconds\n")
    o.append(run["RunType"] + " ! RunType. R = truly random . T = not so random (use this for regression tests). "
                              "C = continuation run\n")
    o.append(str(run["IDUM"]) + " ! Random seed\n")
    o.append(("T" if run.get("SaveFinalState") else "F") + " ! Save final state\n")
    if run.get("SaveFinalState"):
        o.append(path + "/" + run["restartfilename"] + "\n")
    o.append(("T" if run.get("SaveCoo") else "F") + " ! Save final Coor\n")
    if run.get("SaveCoo"):
        o.append((path + "/" + run["coofilename"] if driver is None else "None") + "\n")
    if sim == "ST" and run.get("useProbTable"):
        o.append(run["ProbTable"] + "\n")
    else:
        pe = np.asarray(run["PtransE"], np.float64)
        pd = np.asarray(run["PtransD"], np.float64)
        for s in range(ncry):
            for ind in np.asarray(run["indices"]):
                digits = "".join(str(v) for v in ind)
                prob = np.amax([PROB_FLOOR, pe[(s,) + tuple(int(v) for v in ind)]])
                o.append("{} {} {} ! PtransE[{},{}] \n".format(s + 1, digits, prob, s + 1, ind))
        for s in range(ncry):
            for c in range(3):
                o.append("{} {} {} ! PtransD[{},{}] \n".format(s + 1, c + 1, np.amax([PROB_FLOOR, pd[s, c]]), s + 1,
                                                               c + 1))
        if ncov:
            for name in ("PtransAbs", "PtransDes"):
                t = np.asarray(run[name], np.float64)
                for s in range(ncry, ncry + ncov):
                    for c in range(3):
                        o.append("{} {} {} ! {}[{},{}] \n".format(s + 1, c + 1, np.amax([PROB_FLOOR, t[s - ncry, c]]),
                                                                  name, s + 1, c + 1))
    return "".join(o)
