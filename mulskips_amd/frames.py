"""xyz frames in the format of inpout/WriteMolMolXYZFile.f, written from the
engine's device-compacted records (Engine.frame).

Main frame (I????????.xyz, :123-176): a header with the number of listed
points and the KMC time in Fortran E15.8, the periodic box in Angstrom, one
line per coverage point (H, Cl) and per atom of coordination <= 3 in grid
order, then one 'O' line per vacancy.  The _d frame (:77-121) lists the bulk
atoms that sit off the cubic diamond sites under the reference's own
symbols (Ge for species 1, B for 2, Cl for 3).  The SiC defect frames
(:222-398) list the vacancies by class (_v) and the wrong occupations (_w).
"""
SYMBOLS = {6: "C ", 32: "Ge ", 5: "B "}
D_SYMBOLS = {1: "Ge ", 2: "B ", 3: "Cl "}
# _v frame of SiC PVD (:63-67): the species sum of a vacancy's neighbours
VACANCY_LINES = {4: ("Mg ", " # CV (C vacancy)"), 8: ("He ", " # SV (Si vacancy)"),
                 7: ("Ne ", " # SAV (Si antisite vac)"), 5: ("Be ", " # CAV (C antisite vac)"),
                 6: ("Ar ", " # XV (near 2C and 2Si)")}
WRONG_SYMBOLS = {1: "Si ", 2: "C "}


def fort_e(v, w=15, d=8):
    """Fortran Ew.d output: [-]0.ddddddddE+xx right-justified in w columns."""
    if v == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        mant, ex = f"{abs(v):.{d - 1}E}".split("E")
        e = int(ex) + 1
        s = ("-" if v < 0 else "") + "0." + mant.replace(".", "") + "E" + ("-" if e < 0 else "+") + f"{abs(e):02d}"
    return s.rjust(w)


def _line(sym, sf, x, y, z, tail="  "):
    return f"{sym}{sf * x:10.5f} {sf * y:10.5f} {sf * z:10.5f}{tail}\n"


def write_frame(path, records, box, sf, kmc_time, iteration, list_crystal, list_cov, d_path=None):
    """records: the tuple returned by Engine.frame()."""
    main, voids, noepi, nat, ncov = records
    lx, ly, lz = box
    head = f" {{:8d}} angstroem KMC-time: {fort_e(kmc_time)} Iter: {int(iteration):12d}\n"
    cell = f" periodic {fort_e(sf * lx)} {fort_e(sf * ly)} {fort_e(sf * lz)}\n"
    with open(path, "w") as f:
        f.write(head.format(nat + len(voids) + ncov))
        f.write(cell)
        for x, y, z, code in main:
            if code < 0:
                cv = list_cov[-code - 1]
                if cv == 1:
                    f.write(_line("H ", sf, x, y, z))
                elif cv == 17:
                    f.write(_line("Cl ", sf, x, y, z))
            elif code == 1:
                f.write(_line("Si ", sf, x, y, z))
            else:
                sym = SYMBOLS.get(int(list_crystal[code - 1]))
                if sym is None:
                    raise ValueError("the second and third crystal species can only be C, Ge or B")
                f.write(_line(sym, sf, x, y, z))
        for x, y, z, _ in voids:
            f.write(_line("O ", sf, x, y, z))
    if d_path:
        with open(d_path, "w") as f:
            f.write(head.format(len(noepi)))
            f.write(cell)
            for x, y, z, code in noepi:
                f.write(_line(D_SYMBOLS[int(code)], sf, x, y, z))


def write_defect_frames(v_path, w_path, records, box, sf, list_crystal, n_cov):
    """The SiC defect frames (WriteMolMolXYZFile.f:222-398) from the records
    of Engine.frame_defects(): v_path the vacancy frame (every vacancy, as
    its SiC PVD class -- NCov 0, two crystal species, the second C -- or as
    an 'O' line otherwise), w_path the wrong-occupation frame, written only
    for two crystal species one of which is C.  Either path may be None."""
    vac, wrong = records
    lx, ly, lz = box
    cell = f" periodic {fort_e(sf * lx)} {fort_e(sf * ly)} {fort_e(sf * lz)}\n"
    ncr = sum(1 for v in list_crystal if int(v))
    sic_pvd = n_cov == 0 and ncr == 2 and int(list_crystal[1]) == 6
    if v_path:
        with open(v_path, "w") as f:
            f.write(f" {len(vac):8d} angstroem\n")
            f.write(cell)
            for x, y, z, code in vac:
                if not sic_pvd:
                    f.write(_line("O ", sf, x, y, z))
                elif int(code) in VACANCY_LINES:
                    sym, tail = VACANCY_LINES[int(code)]
                    f.write(_line(sym, sf, x, y, z, tail))
                else:
                    raise ValueError(f"vacancy at ({x},{y},{z}) with neighbour species sum {code}: "
                                     "a defect type the reference stops on")
    if w_path and ncr == 2 and 6 in [int(v) for v in list_crystal]:
        with open(w_path, "w") as f:
            f.write(f" {len(wrong):8d} angstroem\n")
            f.write(cell)
            for x, y, z, code in wrong:
                f.write(_line(WRONG_SYMBOLS[int(code)], sf, x, y, z))
