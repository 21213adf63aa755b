"""Melting/solidification rate tables of pymulskips' PhaseChange process as
differentiable torch functions (pymulskips/process.py:854-1152).

PhaseChange turns calibration parameters into the start.dat rate tables of
a Si, Ge or Si(1-x)Ge(x) melt/solid interface:

* evaporation (melting) energetics E[species, n_Si, n_Ge] by the number of
  first neighbours of each species (process.py:1006-1061): n * Ed2 for one
  species; for 'SiGe_X' the mixed-neighbour barriers are interpolated and
  scaled by the alpha / beta corrections;
* PtransE = P0[species] * exp(-E / (kB T)) (process.py:1063-1086);
* deposition (solidification) energetics 2 (Ed2 - delta), 2 Ed2,
  2 (Ed2 + delta) for coordination 1, 2, 3 (process.py:1088-1121);
* PtransD = P0 X exp(-E / (kB Tm)) * (1 + erf((T - Tflex) / A)) / 2
  (process.py:1123-1152).

Everything is float64 torch, so a gradient of anything built on the tables
(the trajectory likelihood of mulskips_amd.adjoint, say) flows back to the
calibration parameters.  tables() writes them into the engine's table form
with the start.dat writer's floor of 5e-12 (pymulskips/setuprun.py:355-386).
"""
import itertools

import numpy as np
import torch

KB_EV = 8.617333262145e-5          # process.py:867
SIGE_ALPHA, SIGE_BETA = 0.063, 0.085  # process.py:1029-1030
MELTING_K = {"Si": (1688.0,), "Ge": (1210.0,), "SiGe_X": (1688.0, 1210.0)}  # process.py:879, 903, 926
SPECIES_Z = {"Si": (14,), "Ge": (32,), "SiGe_X": (14, 32)}
SUPERLATTICE_A = {"Si": 5.43 / 12.0, "Ge": 5.66 / 12.0, "SiGe_X": 5.43 / 12.0}  # KMC_sf
PROB_FLOOR = 5e-12                 # setuprun.py:355 (mulskipsProbThreshold)


def _t(v):
    return v if torch.is_tensor(v) else torch.as_tensor(v, dtype=torch.float64)


def _vec(params, key, n):
    v = _t(params[key]).to(torch.float64).reshape(-1)
    if v.numel() < n:
        raise ValueError(f"calibration parameter {key!r} needs {n} values (one per species)")
    return v[:n]


def indices(ncry, ncov=0):
    """Neighbour compositions that get an evaporation line (process.py:977-
    999): every (n_1 .. n_{ncry+ncov}) in 0..3 with at most 4 neighbours, at
    least one and at most three of them crystal atoms, in itertools.product
    order."""
    out = []
    for p in itertools.product(range(4), repeat=ncry + ncov):
        if sum(p) > 4 or sum(p) == 0:
            continue
        c = sum(p[:ncry])
        if c == 0 or c > 3:
            continue
        out.append(p)
    return np.array(out, dtype=int).reshape(-1, ncry + ncov)


def evap_energetics(substrate, params):
    """E[species, n_1, .., n_ncry] in eV (process.py:1006-1061): n * Ed2 for
    Si or Ge; the SiGe_X table of mixed neighbours, scaled by (1 - alpha) for
    Si and (1 + k beta) for Ge."""
    if substrate in ("Si", "Ge"):
        ed = _vec(params, "Ed2", 1)[0]
        return torch.stack([torch.zeros((), dtype=torch.float64), ed, 2 * ed, 3 * ed]).reshape(1, 4)
    if substrate != "SiGe_X":
        raise ValueError(f"substrate {substrate!r}: 'Si', 'Ge' or 'SiGe_X'")
    e = _vec(params, "Ed2", 2)
    es, eg = e[0], e[1]
    a, b = SIGE_ALPHA, SIGE_BETA
    z = torch.zeros((), dtype=torch.float64)
    si = {(1, 0): es, (0, 1): (1 - a) * (es + eg) / 2, (1, 1): (1 - a) * 2 * (3 * es + eg) / 4,
          (2, 0): 2 * es, (0, 2): (1 - a) * 2 * (es + eg) / 2, (2, 1): (1 - a) * 3 * (5 * es + eg) / 6,
          (1, 2): (1 - a) * 3 * (2 * es + eg) / 3, (3, 0): 3 * es, (0, 3): (1 - a) * 3 * (es + eg) / 2}
    ge = {(1, 0): (1 + 2 * b) * (eg + es) / 2, (0, 1): eg, (1, 1): (1 + b) * 2 * (3 * eg + es) / 4,
          (2, 0): (1 + 2.5 * b) * 2 * (eg + es) / 2, (0, 2): 2 * eg, (2, 1): (1 + b) * 3 * (2 * eg + es) / 3,
          (1, 2): (1 + b) * 3 * (5 * eg + es) / 6, (3, 0): (1 + 4 * b) * 3 * (eg + es) / 2, (0, 3): 3 * eg}
    rows = []
    for table in (si, ge):
        rows.append(torch.stack([torch.stack([table.get((i, j), z) for j in range(4)]) for i in range(4)]))
    return torch.stack(rows)


def depo_energetics(substrate, params):
    """E[species, coordination-1] in eV (process.py:1088-1121)."""
    n = len(SPECIES_Z[substrate])
    ed, dl = _vec(params, "Ed2", n), _vec(params, "delta", n)
    return torch.stack([2 * (ed - dl), 2 * ed, 2 * (ed + dl)], dim=1)


def damping(substrate, params, temp):
    """(1 + erf((T - Tflex) / A)) / 2 per species, for each of the three
    coordinations (process.py:1123-1131)."""
    n = len(SPECIES_Z[substrate])
    f = 0.5 * (1.0 + torch.erf((_t(temp) - _vec(params, "Tflex", n)) / _vec(params, "A", n)))
    return f.reshape(n, 1).expand(n, 3)


def ptrans_e(substrate, params, temp):
    """P0[species] exp(-E / (kB T)) over the evaporation energetics
    (process.py:1063-1086)."""
    ene = evap_energetics(substrate, params)
    p0 = _vec(params, "P0", ene.shape[0]).reshape((-1,) + (1,) * (ene.dim() - 1))
    return p0 * torch.exp(-ene / (KB_EV * _t(temp)))


def ptrans_d(substrate, params, temp):
    """P0 X exp(-E / (kB Tm)) times the damping factor at T (process.py:
    1133-1152), Tm the species' melting point."""
    ene = depo_energetics(substrate, params)
    n = ene.shape[0]
    p0x = (_vec(params, "P0", n) * _vec(params, "X", n)).reshape(n, 1)
    tm = torch.as_tensor(MELTING_K[substrate], dtype=torch.float64).reshape(n, 1)
    return p0x * torch.exp(-ene / (KB_EV * tm)) * damping(substrate, params, temp)


def start_dat_lines(substrate, pe, pd):
    """The evaporation and deposition lines pymulskips writes for these
    tables (setuprun.py:352-372): (species, digits, value) with every value
    floored at 5e-12."""
    n = len(SPECIES_Z[substrate])
    pe_np = pe.detach().cpu().numpy()
    pd_np = pd.detach().cpu().numpy()
    evap = [(s + 1, "".join(str(v) for v in ind), float(max(PROB_FLOOR, pe_np[(s,) + tuple(ind)])))
            for s in range(n) for ind in indices(n)]
    depo = [(s + 1, c + 1, float(max(PROB_FLOOR, pd_np[s, c]))) for s in range(n) for c in range(3)]
    return evap, depo


def tables(substrate, params, temp, ptrans_zig=1.0):
    """The engine's rate tables (mskg_config layout: PTransE in Fortran order,
    PTransD) for a PhaseChange process at temperature temp, through the
    start.dat values pymulskips would write (floored at 5e-12)."""
    from .deck import Deck
    pe, pd = ptrans_e(substrate, params, temp), ptrans_d(substrate, params, temp)
    evap, depo = start_dat_lines(substrate, pe, pd)
    lc = list(SPECIES_Z[substrate])
    d = Deck(len(lc), 0, lc, [], "FS", [120, 120, 120], float(ptrans_zig), SUPERLATTICE_A[substrate], "Iter",
             evap=evap, depo=depo)
    return d.tables()


def torch_pe_flat(substrate, params, temp):
    """PTransE as the engine's 12288-entry Fortran-order table, a
    differentiable torch tensor (entries pymulskips writes, before the
    5e-12 floor; zero elsewhere)."""
    n = len(SPECIES_Z[substrate])
    pe = ptrans_e(substrate, params, temp)
    out = torch.zeros(3 * 4 ** 6, dtype=torch.float64)
    for s in range(n):
        for ind in indices(n):
            n3 = list(ind) + [0] * (3 - n)
            out = out.index_put((torch.tensor([s + 3 * (n3[0] + 4 * (n3[1] + 4 * n3[2]))]),),
                                pe[(s,) + tuple(int(v) for v in ind)].reshape(1))
    return out


def melting_point(substrate):
    return MELTING_K[substrate]

