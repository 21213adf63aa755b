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
