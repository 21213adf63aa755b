"""xyz frames in the format of inpout/WriteMolMolXYZFile.f, written from the
engine's device-compacted records (Engine.frame).

Main frame (I????????.xyz, :123-176): a header with the number of listed
points and the KMC time in Fortran E15.8, the periodic box in Angstrom, one
line per coverage point (H, Cl) and per atom of coordination <= 3 in grid
order, then one 'O' line per vacancy.  The _d frame (:77-121) lists the bulk
atoms that sit off the cubic diamond sites under the reference's own
symbols (Ge for species 1, B for 2, Cl for 3).
"""
SYMBOLS = {6: "C ", 32: "Ge ", 5: "B "}
D_SYMBOLS = {1: "Ge ", 2: "B ", 3: "Cl "}


def fort_e(v, w=15, d=8):
    """Fortran Ew.d output: [-]0.ddddddddE+xx right-justified in w columns."""
    if v == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        mant, ex = f"{abs(v):.{d - 1}E}".split("E")
        e = int(ex) + 1
        s = ("-" if v < 0 else "") + "0." + mant.replace(".", "") + "E" + ("-" if e < 0 else "+") + f"{abs(e):02d}"
    return s.rjust(w)


def _line(sym, sf, x, y, z):
    return f"{sym}{sf * x:10.5f} {sf * y:10.5f} {sf * z:10.5f}  \n"


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
