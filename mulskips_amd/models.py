"""Rate tables for the engine's workloads, in the reference's table form.

The reference reads FP64 rate tables from start.dat; pymulskips computes them
from physical parameters with closed forms such as P0*exp(-E/kT)
(pymulskips/process.py:1082,1146).  The functions here build the same
PTransE / PTransD arrays (Fortran order, SetProbability.f:79-140) for the
pair-bond-energy model of the large SiC and Si/SiGe benchmark configurations.
"""
import numpy as np

KB_EV = 8.617333262e-5  # Boltzmann constant, eV/K


def pe_index(species, n1, n2, n3, c1=0, c2=0, c3=0):
    """Flat index of PTransE(species, n1, n2, n3, c1, c2, c3) in Fortran order."""
    return (species - 1) + 3 * (n1 + 4 * (n2 + 4 * (n3 + 4 * (c1 + 4 * (c2 + 4 * c3)))))


def pair_energy_tables(pair, fluxes, temperature, prefactor=1e13, list_crystal=(14, 6, 0)):
    """Evaporation rate prefactor*exp(-E/kT) with E the sum of the pair
    energies to the (reciprocal) first neighbours, for every species and
    neighbour composition with 1..3 neighbours; deposition rate of species s
    equal to its flux in ML/s (the per-site normalisation of
    pymulskips PVD tables).

    pair: symmetric (n x n) array of bond energies in eV, n = number of species.
    fluxes: per-species deposition rate (ML/s).
    """
    pair = np.asarray(pair, float)
    ns = pair.shape[0]
    kt = KB_EV * temperature
    pe = np.zeros(3 * 4 ** 6)
    for s in range(1, ns + 1):
        for n1 in range(4):
            for n2 in range(4 if ns >= 2 else 1):
                for n3 in range(4 if ns >= 3 else 1):
                    tot = n1 + n2 + n3
                    if tot < 1 or tot > 3:
                        continue
                    counts = [n1, n2, n3][:ns]
                    e = sum(c * pair[s - 1, j] for j, c in enumerate(counts))
                    pe[pe_index(s, n1, n2, n3)] = prefactor * np.exp(-e / kt)
    pd = np.zeros(9)
    for s in range(1, ns + 1):
        for coor in (1, 2, 3):
            pd[(s - 1) + 3 * (coor - 1)] = fluxes[s - 1]
    lc = np.zeros(3, np.int32)
    lc[:len(list_crystal)] = list_crystal[:3]
    return dict(ptrans_e=pe, ptrans_d=pd, ptrans_abs=np.zeros(12), ptrans_des=np.zeros(12), ptrans_zig=1.0,
                n_crystal=ns, n_cov=0, list_crystal=lc, list_cov=np.zeros(3, np.int32))


def second_neighbour_factors(e2, temperature):
    """exp(-n2*E2/kT) for n2 = 0..12 occupied second-neighbour positions: the
    engine's second-neighbour energy term (E2 in eV per occupied position)."""
    return np.exp(-np.arange(13) * e2 / (KB_EV * temperature))


def sic_c2_tables(temperature=1600.0, e_sisi=0.6, e_sic=1.1, e_cc=0.9, flux_si=2.0, flux_c=1.8, prefactor=1e13,
                  e2=0.05):
    """3C-SiC CVD benchmark model: Si-Si 0.6, Si-C 1.1, C-C 0.9 eV pair
    energies over the first neighbours, 0.05 eV per occupied
    second-neighbour position, Si and C fluxes 2 and 1.8 ML/s, T = 1600 K.
    e2=None leaves the second-neighbour term out (the reference's model
    class)."""
    pair = [[e_sisi, e_sic], [e_sic, e_cc]]
    t = pair_energy_tables(pair, [flux_si, flux_c], temperature, prefactor, (14, 6))
    t["second_factor"] = None if e2 is None else second_neighbour_factors(e2, temperature)
    return t


def si_c1_tables(temperature=1100.0, e_bond=0.55, flux=1.0, prefactor=1e13):
    """Si PVD reference configuration: evaporation 1e13*exp(-n*0.55 eV/kT),
    deposition 1 ML/s, T = 1100 K."""
    return pair_energy_tables([[e_bond]], [flux], temperature, prefactor, (14,))


def sige_alloy_tables(temperature, flux, x_ge, prefactor=1e13):
    """The Si(1-x)Ge(x) model of the batched-ensemble configuration: pair
    energies Si-Si 0.55 / Si-Ge 0.5 / Ge-Ge 0.45 eV, total flux split by the
    Ge fraction."""
    return pair_energy_tables([[0.55, 0.5], [0.5, 0.45]], [flux * (1 - x_ge), flux * x_ge], temperature,
                              prefactor, (14, 32))


def ensemble_replicas(n, seed=2026, laterals=(32, 64, 128, 256), heights=(64, 128, 256)):
    """The batched-ensemble configuration: n Si(1-x)Ge(x) replicas of uneven
    box sizes (lateral edge and height in conventional 12-grid-unit cells,
    drawn uniformly from `laterals` and `heights`), T uniform in [900, 1500]
    K, flux log-uniform in [0.1, 10] ML/s, Ge fraction uniform in [0, 0.5]
    (random substitutional alloy, also the Ge share of the flux), each run
    for the KMC time in which 20 ML are deposited (20 / flux); a replica's
    own Philox seed.  Draws come from numpy's default_rng(seed) in a fixed
    order, so the mix is reproducible."""
    rng = np.random.default_rng(seed)
    reps = []
    for i in range(n):
        lc = int(rng.choice(laterals))
        hc = int(rng.choice(heights))
        temp = float(rng.uniform(900.0, 1500.0))
        flux = float(10 ** rng.uniform(-1.0, 1.0))
        x = float(rng.uniform(0.0, 0.5))
        reps.append(dict(box=(12 * lc, 12 * lc, 12 * hc), tables=sige_alloy_tables(temp, flux, x), x=x,
                         temperature=temp, flux=flux, t_end=20.0 / flux, seed=10_000 + i))
    return reps


LA_CONSTANTS = dict(tm=1687.0, p0=1e13, sigma=50.0, tflex=1650.0)


def la_sige_tables(seed=2026):
    """Energy tables of the laser-annealing configuration (Si0.7Ge0.3 melt and
    regrowth, rates computed on the device from each site's temperature):
    pair barriers Si-Si / Si-Ge / Ge-Ge drawn uniformly in [0.4, 1.2] eV from
    default_rng(seed); an atom's melting barrier is the sum over its bonds,
    the solidification barrier of a site of coordination c is c times the
    species' own pair barrier.  Returns (tables, barriers)."""
    rng = np.random.default_rng(seed)
    e_ss, e_sg, e_gg = (float(v) for v in rng.uniform(0.4, 1.2, 3))
    pair = {(1, 1): e_ss, (1, 2): e_sg, (2, 1): e_sg, (2, 2): e_gg}
    tab = pair_energy_tables([[0.55, 0.5], [0.5, 0.45]], [1.0, 1.0], 1000.0, 1e13, (14, 32))
    pe = np.zeros_like(np.asarray(tab["ptrans_e"], np.float64))
    for sp in (1, 2):
        for a in range(4):
            for b in range(4 - a):
                if a + b:
                    pe[pe_index(sp, a, b, 0)] = a * pair[(sp, 1)] + b * pair[(sp, 2)]
    pd = np.zeros(9)
    for sp in (1, 2):
        for c in (1, 2, 3):
            pd[(sp - 1) + 3 * (c - 1)] = c * pair[(sp, sp)]
    tab["ptrans_e"], tab["ptrans_d"] = pe, pd
    return tab, {"Si-Si": e_ss, "Si-Ge": e_sg, "Ge-Ge": e_gg}


def la_front_field(box, zm, grid_nm=0.5431 / 12, width_nm=5.0):
    """The laser-annealing temperature map T(z) = 1700 K + 300 K tanh((z -
    zm) / 5 nm) (z in grid units of a(Si)/12) in the socket layout of the
    FEM exchange: numpy shape (LenX*LenY, LenZ), z fastest, int32 K."""
    lx, ly, lz = box
    zs = np.arange(lz, dtype=np.float64)
    t = np.rint(1700.0 + 300.0 * np.tanh((zs - zm) * grid_nm / width_nm)).astype(np.int32)
    return np.ascontiguousarray(np.broadcast_to(t, (lx * ly, lz)))
