"""CPU: start.dat as pymulskips writes it, and the PhaseChange rate forms.

* mulskips_amd.deck.pymulskips_start_dat writes, byte for byte, the decks
  pymulskips' own writer wrote for five runs (tests/golden/start_dat,
  written by tests/golden/make_setuprun_goldens.py from
  pymulskips/setuprun.py:292-389): Si and SiGe PhaseChange decks (one with
  melting rates under the writer's 5e-12 floor), the socket-driven
  laser-annealing deck of energetics, and a coverage deck.  The decks read
  back through read_start_dat into the tables the lines hold.
* mulskips_amd.phasechange restates pymulskips/process.py:977-1152 in torch;
  pymulskips.process cannot be imported here (it imports cantera), so its
  forms are pinned by their closed forms and the line counts of the
  reference's reader: the evaporation index list has the SetProbability.f
  line counts for every species / coverage mix, a Si melting barrier is n
  Ed2, the SiGe corrections reduce to the single-species barriers where
  only one species neighbours, the damping factor is 1/2 at Tflex and
  antisymmetric about it, and the torch gradients equal finite differences.
"""
import json
import math
import os

import numpy as np
import pytest
import torch

from mulskips_amd import deck as deckmod
from mulskips_amd import phasechange as pc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "start_dat")
RUNS = json.load(open(os.path.join(GOLDEN, "runs.json")))


@pytest.mark.parametrize("name", sorted(RUNS))
def test_writer_matches_pymulskips(name, tmp_path):
    run = dict(RUNS[name])
    text = deckmod.pymulskips_start_dat(run)
    with open(os.path.join(GOLDEN, name + ".dat")) as f:
        want = f.read()
    assert text == want
    if run["Simulation"] != "LA":  # the reader takes the rates back
        p = tmp_path / "start.dat"
        p.write_text(text)
        d = deckmod.read_start_dat(str(p))
        pe = np.asarray(run["PtransE"])
        for (s, digits, v) in d.evap:
            ind = tuple(int(c) for c in digits)
            assert v == max(5e-12, float(pe[(s - 1,) + ind]))
        assert len(d.evap) == deckmod.MAX_ELINES[(d.n_crystal, d.n_cov)]


def test_floor_applies():
    run = dict(RUNS["si-fs-cold"])
    assert min(run["PtransE"][0][1:]) < 5e-12
    assert "5e-12" in deckmod.pymulskips_start_dat(run)


@pytest.mark.parametrize("ncry,ncov", [(c, v) for c in (1, 2, 3) for v in (0, 1, 2, 3)])
def test_index_list_has_the_reader_line_counts(ncry, ncov):
    ind = pc.indices(ncry, ncov)
    # SetProbability.f:47-75 reads exactly one line per species and composition
    assert ncry * len(ind) == deckmod.MAX_ELINES[(ncry, ncov)]
    assert len({tuple(r) for r in ind}) == len(ind)
    assert [tuple(r) for r in ind] == sorted(tuple(r) for r in ind)  # itertools.product order


SI = {"Ed2": [0.96], "delta": [0.03], "P0": [1.33e17], "A": [280.0], "Tflex": [1080.0], "X": [1.0]}
SIGE = {"Ed2": [0.96, 0.65], "delta": [0.03, 0.0195], "P0": [1.33e17, 2.596e16], "A": [280.0, 210.0],
        "Tflex": [1080.0, 900.0], "X": [0.7, 0.3]}


def test_single_species_closed_forms():
    T = 1450.0
    pe = pc.ptrans_e("Si", SI, T)
    pd = pc.ptrans_d("Si", SI, T)
    kt = 8.617333262145e-5 * T
    for n in (1, 2, 3):
        assert math.isclose(float(pe[0, n]), 1.33e17 * math.exp(-n * 0.96 / kt), rel_tol=1e-14)
    damp = 0.5 * (1 + math.erf((T - 1080.0) / 280.0))
    ktm = 8.617333262145e-5 * 1688.0
    for c, e in ((0, 2 * (0.96 - 0.03)), (1, 2 * 0.96), (2, 2 * (0.96 + 0.03))):
        assert math.isclose(float(pd[0, c]), 1.33e17 * math.exp(-e / ktm) * damp, rel_tol=1e-14)
    # Ge has the same forms with its own melting point
    pd_ge = pc.ptrans_d("Ge", dict(SI, Ed2=[0.65], delta=[0.0195]), T)
    assert math.isclose(float(pd_ge[0, 1]), 1.33e17 * math.exp(-1.3 / (8.617333262145e-5 * 1210.0)) * damp,
                        rel_tol=1e-14)


def test_damping_is_an_erf_step():
    p = dict(SI)
    assert float(pc.damping("Si", p, 1080.0)[0, 0]) == 0.5
    for x in (10.0, 137.0, 600.0):
        lo, hi = (float(pc.damping("Si", p, 1080.0 + s * x)[0, 0]) for s in (-1, 1))
        assert math.isclose(lo + hi, 1.0, rel_tol=1e-15)
    assert float(pc.damping("Si", p, 1e5)[0, 2]) == 1.0


def test_sige_barriers_reduce_to_pure_neighbours():
    """With Ge neighbours only a Ge atom has n Ed2(Ge), with Si neighbours
    only a Si atom n Ed2(Si); the mixed barriers carry the (1 - alpha) and
    (1 + k beta) corrections of the interpolated pair energies."""
    E = pc.evap_energetics("SiGe_X", SIGE).numpy()
    es, eg = 0.96, 0.65
    a, b = 0.063, 0.085
    for n in (1, 2, 3):
        assert math.isclose(E[0, n, 0], n * es) and math.isclose(E[1, 0, n], n * eg)
    assert math.isclose(E[0, 0, 1], (1 - a) * (es + eg) / 2)
    assert math.isclose(E[1, 3, 0], (1 + 4 * b) * 3 * (es + eg) / 2)
    assert math.isclose(E[1, 2, 0], (1 + 2.5 * b) * 2 * (es + eg) / 2)
    assert E[0, 0, 0] == 0 and E[0, 2, 2] == 0  # compositions without a line
    pd = pc.ptrans_d("SiGe_X", SIGE, 1300.0).numpy()
    assert math.isclose(pd[1, 1], 2.596e16 * 0.3 * math.exp(-1.3 / (8.617333262145e-5 * 1210.0)) *
                        0.5 * (1 + math.erf((1300.0 - 900.0) / 210.0)), rel_tol=1e-14)


def test_gradients_match_finite_differences():
    ed = torch.tensor([0.96, 0.65], dtype=torch.float64, requires_grad=True)
    tflex = torch.tensor([1080.0, 900.0], dtype=torch.float64, requires_grad=True)

    def f(ed_, tf_):
        p = dict(SIGE, Ed2=ed_, Tflex=tf_)
        return torch.log(pc.ptrans_e("SiGe_X", p, 1300.0)[1, 2, 1]) + torch.log(pc.ptrans_d("SiGe_X", p, 1300.0)[0, 2])
    y = f(ed, tflex)
    y.backward()
    for var, grad, h in ((0, ed.grad, 1e-6), (1, tflex.grad, 1e-3)):
        for i in range(2):
            args = [ed.detach().clone(), tflex.detach().clone()]
            up, dn = [a.clone() for a in args], [a.clone() for a in args]
            up[var][i] += h
            dn[var][i] -= h
            fd = (float(f(*up)) - float(f(*dn))) / (2 * h)
            assert math.isclose(float(grad[i]), fd, rel_tol=1e-6, abs_tol=1e-9), (var, i, float(grad[i]), fd)


def test_tables_route_through_the_deck_lines():
    t = pc.tables("SiGe_X", SIGE, 1350.0)
    pe = pc.ptrans_e("SiGe_X", SIGE, 1350.0).numpy()
    from mulskips_amd.models import pe_index
    assert t["ptrans_e"][pe_index(2, 1, 2, 0)] == max(5e-12, pe[1, 1, 2])
    assert t["ptrans_e"][pe_index(1, 0, 3, 0)] == max(5e-12, pe[0, 0, 3])
    flat = pc.torch_pe_flat("SiGe_X", SIGE, 1350.0).numpy()
    assert flat[pe_index(2, 1, 2, 0)] == pe[1, 1, 2]
    assert list(t["list_crystal"][:2]) == [14, 32]
