"""Unit pins of oracle pieces that the regression decks exercise only in bulk:
ran3 (random.f:22-69), the RunType 'T' table (PickTreeEvent.f:72-1079) and
the benchmark rate-table models (closed forms quoted for the Si PVD and SiC
configurations)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import GOLDEN  # noqa: E402

from mulskips_amd import models  # noqa: E402
from oracle import oracle  # noqa: E402

FAC = float(np.float32(1e-9))  # ran3 returns mj*fac with fac = 1./Mbig a REAL(4) constant (random.f:30)


def _ints(x):
    y = np.rint(np.asarray(x) / FAC).astype(np.int64)
    assert np.array_equal(y * FAC, np.asarray(x)), "outputs are not integer multiples of the REAL(4) 1e-9"
    return y


@pytest.mark.parametrize("idum", [-1, -9117116, -123456789])
def test_ran3_is_knuth_subtractive(idum):
    """Past the 55 warm-up outputs, ran3 is the lagged recurrence
    y[n] = y[n-55] - y[n-24] (mod 1e9): inext and inextp=inext+31 walk the
    55-entry table, and each output overwrites the slot it read."""
    x = oracle.ran3_sequence(idum, 2000)
    assert np.all((x >= 0) & (x < 1))
    y = _ints(x)
    n = np.arange(55, len(y))
    assert np.array_equal(y[n], (y[n - 55] - y[n - 24]) % 1000000000)


def test_ran3_reseeds_and_differs_by_seed():
    a = oracle.ran3_sequence(-7, 100)
    b = oracle.ran3_sequence(-7, 100)
    c = oracle.ran3_sequence(-8, 100)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    # uniformity of the first 20000 outputs (a gross-error check, 5 sigma)
    u = oracle.ran3_sequence(-3, 20000)
    hist = np.histogram(u, bins=10, range=(0, 1))[0]
    assert np.all(np.abs(hist - 2000) < 5 * np.sqrt(2000))


def test_numberslist_table():
    """The fixed uniforms of RunType 'T': 1000 values in (0,1), REAL(4)
    literals widened to FP64 (PickTreeEvent.f:72-1079)."""
    v = np.loadtxt(os.path.join(GOLDEN, "numberslist.txt"))
    assert v.shape == (1000,)
    assert np.all((v > 0) & (v < 1))
    assert np.array_equal(v.astype(np.float32).astype(np.float64), v)
    assert len(np.unique(v)) > 990


def test_si_pvd_rates_closed_form():
    """Si PVD: 1e13*exp(-n*0.55 eV/kT) at 1100 K = 3.0207e10, 9.1248e7,
    2.7563e5 s^-1 for n = 1, 2, 3; deposition 1 ML/s at every coordination."""
    t = models.si_c1_tables()
    want = [3.0207e10, 9.1248e7, 2.7563e5]
    for n, w in zip((1, 2, 3), want):
        assert t["ptrans_e"][models.pe_index(1, n, 0, 0)] == pytest.approx(w, rel=1e-4)
    assert np.count_nonzero(t["ptrans_e"]) == 3
    pd = t["ptrans_d"].reshape(3, 3)  # PTransD(species, coor) in Fortran order: row = coor, column = species
    assert np.all(pd[:, 0] == 1.0) and np.all(pd[:, 1:] == 0)


def test_sic_rates_closed_form():
    """SiC: E = sum of pair energies (Si-Si 0.6, Si-C 1.1, C-C 0.9 eV);
    fastest class Si on one Si at 1600 K is 1.29e11 s^-1; C bonded to three
    Si is 1e13*exp(-3.3 eV/kT)."""
    t = models.sic_c2_tables()
    kt = models.KB_EV * 1600.0
    pe = t["ptrans_e"]
    assert pe[models.pe_index(1, 1, 0, 0)] == pytest.approx(1.288e11, rel=2e-3)
    assert pe[models.pe_index(2, 3, 0, 0)] == pytest.approx(1e13 * np.exp(-3.3 / kt), rel=1e-12)
    assert pe[models.pe_index(1, 1, 2, 0)] == pytest.approx(1e13 * np.exp(-(0.6 + 2.2) / kt), rel=1e-12)
    # 2 species x 9 compositions with 1..3 neighbours
    assert np.count_nonzero(pe) == 18
    pd = t["ptrans_d"].reshape(3, 3)
    assert np.all(pd[:, 0] == 2.0) and np.all(pd[:, 1] == 1.8) and np.all(pd[:, 2] == 0)
