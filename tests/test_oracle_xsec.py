"""Pins of the oracle's LXCat reader and cross-section interpolation.

References restated by oracle/xsec.py: cross_section.py:31-69 (the parse itself
is done upstream by the third-party lxcat_data_parser, not installed, no version
pinned; its published LXCat block layout is restated) and collisions.py:64-79
(interp1d, linear, flat outside the table).  Checked here against
  * scipy.interpolate.interp1d with the same options;
  * the reference's own fixture lxcat_data/eAr_crs.synthetic.3sp2r (copied as data
    into tests/golden/), read a second way (plain line scan between the dashed
    lines) for record count, parameters and every table value;
  * the product reader (boltzmann_amd.cross_section) on both data files.
"""
import os

import numpy as np
import pytest
import scipy.interpolate

from conftest import XSEC_FILE
from oracle import xsec

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "eAr_crs.synthetic.3sp2r")


def test_interp_matches_scipy_interp1d():
    rng = np.random.default_rng(3)
    e = np.sort(np.concatenate([[0.0], rng.uniform(0, 100, 40)]))
    s = rng.uniform(0, 1e-19, len(e))
    x = np.concatenate([rng.uniform(-5, 120, 500), e, [e[0], e[-1]]])
    f = scipy.interpolate.interp1d(e, s, kind="linear", bounds_error=False, fill_value=(s[0], s[-1]))
    np.testing.assert_allclose(xsec.interp_linear_flat(e, s, x), f(x), rtol=1e-15, atol=1e-35)


def blocks_by_line_scan(path):
    """Independent reading: each table is the numeric lines between a pair of dashed
    lines; the three lines after a type keyword are species / parameter."""
    lines = open(path).read().splitlines()
    out = []
    i = 0
    while i < len(lines):
        if lines[i].strip() in ("ELASTIC", "EFFECTIVE", "EXCITATION", "IONIZATION", "ATTACHMENT"):
            kind, param = lines[i].strip(), float(lines[i + 2].split()[0])
            a = next(j for j in range(i, len(lines)) if lines[j].startswith("-----"))
            b = next(j for j in range(a + 1, len(lines)) if lines[j].startswith("-----"))
            tab = np.loadtxt(lines[a + 1:b])
            out.append((kind, param, tab))
            i = b
        i += 1
    return out


def test_reference_fixture_parses():
    recs, species = xsec.read_cross_section_data(FIXTURE)
    scan = blocks_by_line_scan(FIXTURE)
    assert len(recs) == len(scan) == 2
    assert species == ["Ar"]
    assert [r["type"] for r in recs] == ["ELASTIC", "IONIZATION"]
    assert recs[0]["mass_ratio"] == pytest.approx(1.36e-5, rel=1e-15)
    assert recs[1]["threshold"] == pytest.approx(15.75961, rel=1e-15)
    assert recs[0]["process"] == "E + Ar -> E + Ar"
    assert recs[1]["process"] == "E + Ar -> E + E + Ar+"
    for r, (kind, param, tab) in zip(recs, scan):
        np.testing.assert_array_equal(r["energy"], tab[:, 0])
        np.testing.assert_array_equal(r["cross section"], tab[:, 1])
    # a physical sanity check of the columns: ionization vanishes at threshold
    ion = recs[1]
    below = ion["energy"] <= ion["threshold"]
    assert below.any() and (ion["cross section"][below] == 0).all()


@pytest.mark.parametrize("path", [FIXTURE, XSEC_FILE])
def test_product_reader_agrees_with_oracle(path):
    from boltzmann_amd import cross_section as pcs
    recs, _ = xsec.read_cross_section_data(path)
    prod = pcs.read_cross_section_data(path)
    assert list(prod.keys()) == [r["process"] for r in recs]
    for r in recs:
        p = prod[r["process"]]
        assert p["type"] == r["type"]
        np.testing.assert_array_equal(np.asarray(p["energy"]), r["energy"])
        np.testing.assert_array_equal(np.asarray(p["cross section"]), r["cross section"])
        assert p["threshold"] == r["threshold"]
    assert pcs.read_available_species(path) == xsec.read_cross_section_data(path)[1]
