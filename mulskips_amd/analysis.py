"""Film observables as pymulskips defines them (pymulskips/analyze.py:124-483),
computed from an xyz frame file or straight from the engine's frame records.

* surface_height: the surface height of a frame of undercoordinated atoms
  (analyze.py:124-207): the z values go into a histogram of bin_size
  Angstrom bins over [0, zl); the fullest bin (its next one when the bin
  lies wholly below the lowest atom) gives a first average, and the atoms
  within surface_roughness of it give the height.  A frame whose z range is
  null or narrower than one bin gives the plain mean.
* growth_rate: d height / d t in micron/hour from a sequence of frames
  (analyze.py:209-416): central finite differences (one-sided at the ends)
  averaged without the Nexclude first and last, or the derivative of a
  smoothing spline, or the slope of a straight-line fit, both over the
  frames without the Nexclude first and last.
* coverage: H and Cl points per surface atom of a frame (analyze.py:
  418-483), the surface atoms being the undercoordinated crystal atoms less
  four per vacancy.

Frames carry coordinates as F10.5 text; frame_z() rounds the engine's
records the same way, so a value computed from the records equals the one
computed from the written file.
"""
import numpy as np


def _per_hour(v):
    """Angstrom/s to micron/hour as analyze.py:302 scales, (v * 3600) * 1e-4."""
    return v * 3600 * 1e-4


def read_frame(path):
    """(species list, z array, box (xl, yl, zl)) of an xyz frame file."""
    with open(path) as f:
        lines = f.read().splitlines()
    nat = int(lines[0].split()[0])
    cell = lines[1].split()
    box = tuple(float(v) for v in cell[1:4])
    sp, z = [], []
    for line in lines[2:]:
        t = line.split()
        sp.append(t[0])
        z.append(float(t[3]))
    if len(z) != nat:
        raise ValueError(f"{path}: header says {nat} points, the frame lists {len(z)}")
    return sp, np.asarray(z, np.float64), box


def frame_z(records, sf):
    """The z values a frame written from these (x, y, z, code) records
    lists, rounded as its F10.5 text is: concatenate the main records and
    the vacancy records for the whole main frame."""
    return np.array([float(f"{sf * float(z):10.5f}") for z in np.asarray(records).reshape(-1, 4)[:, 2]])


def surface_height(z, zl, bin_size=5.0, surface_roughness=20.0):
    """Surface height of a frame (analyze.py:124-207); z in Angstrom, zl the
    box height in Angstrom."""
    z = np.asarray(z, np.float64)
    zs = list(z)
    if zs.count(zs[0]) == len(zs) or (max(zs) - min(zs)) < bin_size:
        return sum(zs) / len(zs)
    counts, _ = np.histogram(zs, bins=int(zl / bin_size), range=(0.0, zl))
    top = int(np.flatnonzero(counts == counts.max())[0])
    lo = float(bin_size * top)
    hi = lo + bin_size
    if hi < min(zs):
        lo, hi = lo + bin_size, hi + bin_size
    n = s = 0.0
    for v in zs:
        if lo < v < hi:
            n += 1
            s = s + v
    first = s / n
    n = s = 0.0
    for v in zs:
        if first - surface_roughness < v < first + surface_roughness:
            n += 1
            s = s + v
    return s / n


def growth_rate(times, heights, method="finitediff", n_exclude=2):
    """(per-frame rates, average rate, fitted heights) in micron/hour
    (analyze.py:293-370) for frames at KMC times `times` with surface heights
    `heights` (Angstrom), in the order given."""
    t = [float(v) for v in times]
    h = [float(v) for v in heights]
    if method in ("finitediff", "finiteANDfit"):
        gr = [_per_hour((h[1] - h[0]) / (t[1] - t[0]))]
        for i in range(1, len(t) - 1):
            gr.append(_per_hour((h[i + 1] - h[i - 1]) / (t[i + 1] - t[i - 1])))
        gr.append(_per_hour((h[-1] - h[-2]) / (t[-1] - t[-2])))
        ave = np.mean(np.array(gr[n_exclude:-n_exclude]))
        if method == "finitediff":
            return gr, ave, None
        c = np.polynomial.polynomial.polyfit(np.asarray(t[n_exclude:-n_exclude]),
                                             np.asarray(h[n_exclude:-n_exclude]), 1)
        return gr, (ave, _per_hour(c[1])), c[0] + c[1] * np.asarray(t[n_exclude:-n_exclude])
    if method == "spline":
        from scipy.interpolate import UnivariateSpline
        x, y = np.asarray(t), np.asarray(h)
        spl = UnivariateSpline(x, y)
        gr = _per_hour(spl.derivative()(x))
        return gr, np.mean(np.array(gr[n_exclude:-n_exclude])), spl(x)[n_exclude:-n_exclude]
    if method == "polyfit":
        x = np.asarray(t[n_exclude:-n_exclude])
        c = np.polynomial.polynomial.polyfit(x, np.asarray(h[n_exclude:-n_exclude]), 1)
        ave = _per_hour(c[1])
        return np.ones(len(t)) * ave, ave, c[0] + c[1] * x
    raise ValueError(f"method {method!r}")


def coverage(species):
    """{'H': [..], 'Cl': [..]}: coverage points per surface atom of a main
    frame's species column (analyze.py:418-483); the surface atoms are the
    crystal atoms listed less four per vacancy ('O')."""
    h = sum(1 for s in species if s == "H")
    cl = sum(1 for s in species if s == "Cl")
    vac = sum(1 for s in species if s == "O")
    under = len(species) - h - cl - vac
    surface = (under - vac * 4) * 1.0
    return {"H": [h / surface], "Cl": [cl / surface]}


def main_frame(latt, box, sf):
    """(species, z) of the main xyz frame of a LattCoo state (int32 (LenZ+1,
    LenY, LenX), WriteMolMolXYZFile.f:123-176): coverage points as H / Cl and
    atoms of coordination <= 3 in grid order, then the vacancies as 'O' (the
    reference lists those in ListVoid order), for the planes z <= LenZ-1,
    with z as the frame's F10.5 text gives it.  Crystal atoms are named 'X':
    the analyses only tell H, Cl and O apart."""
    lx, ly, lz = box
    lat = np.asarray(latt)[:lz]
    cov = lat & 255
    occ = (lat >> 23) & 1
    coor = (lat >> 24) & 7
    wall = cov == 113
    atom = (occ == 1) & ~wall & (coor <= 3)
    vac = (occ == 0) & (coor == 4) & ~wall
    main = atom | (((cov == 1) | (cov == 17)) & ~wall)
    zm, ym, xm = np.nonzero(main)  # grid order, as the frame lists them
    kind = cov[zm, ym, xm]
    species = ["H" if c == 1 else "Cl" if c == 17 else "X" for c in kind.tolist()]
    zv = np.nonzero(vac)[0]  # the vacancies follow (here in grid order)
    species += ["O"] * len(zv)
    zs = np.concatenate([zm, zv]).astype(np.float64)
    return species, np.array([float(f"{sf * v:10.5f}") for v in zs])
