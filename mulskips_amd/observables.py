"""Film observables of a lattice state, for comparing runs statistically.

All functions take LattCoo as the engine and the oracle hand it out: int32
(LenZ+1, LenY, LenX) with the bit layout of mulskips-source/modules/
defsystem.f:87-91 (CovInd b0-7, occupancy b23, coordination b24-26,
species b27-29).
"""
import numpy as np

WALL = 113


def occupied(latt):
    """Atoms (occupied points that are not wall)."""
    return (((latt >> 23) & 1) == 1) & ((latt & 255) != WALL)


def covered(latt):
    """Empty points carrying a coverage species (H, Cl, ...)."""
    cov = latt & 255
    return (((latt >> 23) & 1) == 0) & (cov != 0) & (cov != WALL)


def atoms_per_layer(box):
    """Atoms of one (001) atomic layer (two per (a/2)^2 with a = 12 grid units)."""
    lx, ly, _ = box
    return lx * ly // 72


def cubic_site(x, y, z):
    """True where a point is a site of the cubic diamond lattice the flat
    builders start from (initial/SetSi.f); atoms elsewhere sit on
    hexagonal-stacking (stacking-fault) positions of the super-lattice."""
    ok = (x % 3 == 0) & (y % 3 == 0) & (z % 3 == 0)
    x1, y1, z1 = x // 3, y // 3, z // 3
    ok &= ((x1 % 2) == (y1 % 2)) & ((x1 % 2) == (z1 % 2))
    m = (x1 + y1 + z1) % 4
    return ok & ((m == 0) | (m == 3))


def film(latt, box, z0, sf=None):
    """Observables of the film above height z0 (grid units):
    atoms: atoms with z >= z0;
    height: mean of the column heights (highest atom in each 6x6 column);
    roughness: RMS deviation of the column heights;
    defects: fraction of the film's atoms on non-cubic (stacking-fault) sites;
    coverage: covered points per surface site (atoms_per_layer);
    species1, species2: the film's atoms of species 1 and 2;
    and with the super-lattice step sf (Angstrom) the observables pymulskips
    reads off the main xyz frame (pymulskips/analyze.py:124-207, :418-483):
    surface: its surface height in Angstrom; cov_h: its H coverage per
    surface atom (0 without surface atoms)."""
    lx, ly, lz = box
    occ = occupied(latt)
    z, y, x = np.nonzero(occ)
    keep = z >= z0
    z, y, x = z[keep], y[keep], x[keep]
    sp = (latt[z, y, x] >> 27) & 7
    cols = np.full((ly // 6, lx // 6), z0 - 3, np.int64)
    np.maximum.at(cols, (y // 6, x // 6), z)
    defects = float(np.mean(~cubic_site(x, y, z))) if len(z) else 0.0
    out = {
        "atoms": int(len(z)),
        "height": float(cols.mean()),
        "roughness": float(cols.std()),
        "defects": defects,
        "coverage": float(covered(latt).sum()) / atoms_per_layer(box),
        "species1": int(np.count_nonzero(sp == 1)),
        "species2": int(np.count_nonzero(sp == 2)),
    }
    if sf is not None:
        from . import analysis
        names, zf = analysis.main_frame(latt, box, sf)
        out["surface"] = analysis.surface_height(zf, sf * lz) if len(zf) else 0.0
        under = sum(1 for n in names if n == "X") - 4 * sum(1 for n in names if n == "O")
        out["cov_h"] = analysis.coverage(names)["H"][0] if under > 0 else 0.0
    return out


def compare(a, b, keys, nsigma=2.0):
    """Two-sample comparison of per-run observables: for every key, the
    difference of the means against nsigma standard errors of that
    difference.  Returns {key: (mean_a, mean_b, diff, bound, ok)}."""
    out = {}
    for k in keys:
        va = np.array([r[k] for r in a], float)
        vb = np.array([r[k] for r in b], float)
        se = np.sqrt(va.var(ddof=1) / len(va) + vb.var(ddof=1) / len(vb))
        d = va.mean() - vb.mean()
        out[k] = (va.mean(), vb.mean(), d, nsigma * se, bool(abs(d) <= nsigma * se))
    return out
