"""CPU: the oracle's InitSt 'ST' interpolation (mcevent/InterpProbFromField.f:22-57,
called by Get_Prob.f:77-90 at each site's temperature), pinned by known answers.

No reference fixture runs an 'ST' deck (its initialiser, SetTrenchSi, is a
host builder), so the interpolation is pinned by what its rule fixes:

* at a reference temperature every rate equals the plain deck's with that
  table;
* between references 0.5 x B at 1400 K and B at 1500 K, the integer weights
  wl = |1500 - T| on the lower and wu = |1400 - T| on the upper table give
  0.75 x B's rates at 1450 K and 0.625 x at 1425 K (swapped weights would
  give 0.875 x);
* a temperature outside the references fails (the reference indexes out of
  its arrays there).

Each holds on the Si-SL deck and on the CVD deck, whose adsorption
(PTransAbs) and desorption (PTransDes) tables are interpolated alike.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.helpers import SMALL, small_deck

TEMPS = [1400, 1500, 1600]


def scaled(tab, f):
    out = dict(tab)
    for k in ("ptrans_e", "ptrans_d", "ptrans_abs", "ptrans_des"):
        out[k] = np.asarray(tab[k], np.float64) * f
    return out


def deck_oracle(tmp_path, key="Si-SL"):
    """The deck; the CVD deck then grown by 3000 events with its plain
    tables (the same events every time), so that coverage is present."""
    base, box, lens = SMALL[key]
    path, d = small_deck(tmp_path, base, lens)
    o = Oracle(*box)
    o.load_deck(path)
    if d.tables()["ptrans_des"].max() > 0:
        o.step(3000, record=False)
    return o, d, box


def uniform_t(box, t):
    lx, ly, lz = box
    return np.full((lz, ly, lx), t, np.int32)


def scalable_rates(o, cats=False):
    """Rates of the MC particles whose rate is a sum of table entries
    (deposition and adsorption, evaporation, desorption; no placeholder, no
    isolated atom)."""
    mc, _, _ = o.lists()
    out, cs = [], []
    for s in mc["xyz"]:
        cat, r = o.pure_rate(s)
        out.append(r if cat in (1, 4, 5) else np.nan)
        cs.append(cat)
    return (np.array(out), np.array(cs)) if cats else np.array(out)


@pytest.fixture(params=["Si-SL", "CVD-Si-SiH2"])
def key(request):
    return request.param


@pytest.fixture
def plain(tmp_path, key):
    (tmp_path / "plain").mkdir()
    o, d, box = deck_oracle(tmp_path / "plain", key)
    return scalable_rates(o)


def test_cvd_deck_has_coverage_rates(tmp_path):
    """The CVD deck's MC particles include desorption and adsorption ones, so
    the interpolation of PTransAbs / PTransDes is exercised above."""
    o, d, box = deck_oracle(tmp_path, "CVD-Si-SiH2")
    r, cat = scalable_rates(o, cats=True)
    assert (cat == 5).sum() > 10 and (cat == 1).sum() > 10
    assert np.any(np.asarray(d.tables()["ptrans_abs"]) > 0)


@pytest.fixture
def st_oracle_of(key):
    return lambda tmp_path, t: st_oracle(tmp_path, t, key)


def st_oracle(tmp_path, t, key="Si-SL"):
    o, d, box = deck_oracle(tmp_path, key)
    tab = d.tables()
    o.set_temperature_field(uniform_t(box, t), "planes")
    o.set_st_tables(TEMPS, [scaled(tab, 0.5), tab, scaled(tab, 2.0)])
    return o


def test_reference_temperature_takes_its_table(tmp_path, plain, st_oracle_of):
    r = scalable_rates(st_oracle_of(tmp_path, 1500))
    ok = ~np.isnan(plain)
    assert ok.sum() > 100
    np.testing.assert_array_equal(r[ok], plain[ok])
    r2 = scalable_rates(st_oracle_of(tmp_path, 1600))
    np.testing.assert_allclose(r2[ok], 2.0 * plain[ok], rtol=1e-15)


@pytest.mark.parametrize("t,factor", [(1450, 0.75), (1425, 0.625), (1550, 1.5)])
def test_integer_weights(tmp_path, plain, st_oracle_of, t, factor):
    r = scalable_rates(st_oracle_of(tmp_path, t))
    ok = ~np.isnan(plain)
    np.testing.assert_allclose(r[ok], factor * plain[ok], rtol=1e-14)


@pytest.mark.parametrize("t", [1300, 1700])
def test_outside_the_references_fails(tmp_path, t):
    o = st_oracle(tmp_path, t)
    with pytest.raises(RuntimeError, match="outside"):
        o.refresh_all()
