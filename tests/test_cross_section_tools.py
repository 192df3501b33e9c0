"""Cross-section preprocessing of the product (boltzmann_amd/cross_section.py): Maxwellian
rate coefficients (cross_section.py:71-110) and the lumping of lump_crs.py, checked against
closed forms and invariants on the CPU."""
import math

import numpy as np
import pytest
import scipy.constants as sc

from boltzmann_amd import cross_section as cs


def test_maxwellian_rates_closed_forms():
    """<sigma v> of a Maxwellian at kT: sigma0 <v> = sigma0 sqrt(8 kT / (pi m)) for a constant
    sigma0, and a <eps v> = 2 a kT <v> for sigma = a eps (from <v^3> = (4 / sqrt(pi))
    (2 kT / m)^(3/2)).  The trapezoid rule on 1000 points of [0, 30] vth has the
    Euler-Maclaurin error h^4 f3(0) / 720 for f = v^3 exp(-v^2) (third derivative f3(0) = 6,
    h = 30 / 999): 1.36e-8 of the integral 1/2; for v^5 exp(-v^2) the first term is O(h^6)."""
    Te = np.array([0.5, 2.0, 7.5])
    s0, a = 3e-20, 2e-21
    rates = cs.compute_mw_reaction_rates(Te, [lambda e: np.full_like(e, s0), lambda e: a * e])
    vbar = np.sqrt(8 * Te * sc.electron_volt / (math.pi * sc.electron_mass))
    h = 30.0 / 999
    np.testing.assert_allclose(rates[:, 0], s0 * vbar * (1 + h ** 4 * 6 / 720 / 0.5), rtol=1e-11)
    np.testing.assert_allclose(rates[:, 1], 2 * a * Te * vbar, rtol=1e-10)


def _rec(species, e, s, ptype="EXCITATION"):
    return {"type": ptype, "species": species, "energy": np.asarray(e, float), "cross section": np.asarray(s, float),
            "threshold": float(e[0]), "mass_ratio": None, "attachment_energy": 0.0}


def test_lumping_weights_grid_and_round_trip(tmp_path):
    """Boltzmann weights g exp(-E/Tg) normalised to one; identical members lump to themselves;
    the grid spans the common range (0 prepended from 1e-4); the written LXCat block reads
    back to the same table."""
    e1 = np.array([11.5, 12.0, 20.0, 100.0])
    s1 = np.array([0.0, 1e-21, 3e-21, 1e-21])
    e2 = np.array([11.6, 13.0, 25.0, 80.0])
    s2 = np.array([0.0, 2e-21, 1e-21, 5e-22])
    data = {"E + Ar -> E + ArA": _rec("ArA", e1, s1), "E + Ar -> E + ArB": _rec("ArB", e2, s2),
            "E + Ar -> E + ArC": _rec("ArC", e1, s1)}
    energy = {"ArA": 11.5, "ArB": 11.6, "ArC": 11.5}
    degen = {"ArA": 5, "ArB": 3, "ArC": 1}
    Tg = 1.0
    out = cs.lump_cross_sections(data, {"E + Ar -> E + Ar*": ["E + Ar -> E + ArA", "E + Ar -> E + ArB"],
                                        "E + Ar -> E + Ar**": ["E + Ar -> E + ArA", "E + Ar -> E + ArC"]},
                                 energy, degen, Tg, num_cs_pts=64)
    a = out["E + Ar -> E + Ar*"]
    w = np.array([5 * math.exp(-11.5), 3 * math.exp(-11.6)])
    np.testing.assert_allclose(a["weights"], w / w.sum(), rtol=1e-14)
    assert a["energy"][0] == pytest.approx(11.5) and a["energy"][-1] == pytest.approx(80.0)
    assert len(a["energy"]) == 64 and np.all(np.diff(a["energy"]) > 0)
    ref = w[0] / w.sum() * np.interp(a["energy"], e1, s1) + w[1] / w.sum() * np.interp(a["energy"], e2, s2)
    np.testing.assert_allclose(a["cross section"], ref, rtol=1e-14, atol=0)
    b = out["E + Ar -> E + Ar**"]                 # identical tables: the lump is the table itself
    np.testing.assert_allclose(b["cross section"], np.interp(b["energy"], e1, s1), rtol=1e-14)
    # a range starting at zero: 0 then log-spaced from 1e-4
    z = cs.lump_cross_sections({"E + Ar -> E + Ar": _rec("Ar", [0.0, 1.0, 10.0], [1e-20, 2e-20, 5e-21],
                                                          ptype="ELASTIC") | {"mass_ratio": 1.36e-5}},
                               {"E + Ar -> E + Ar": ["E + Ar -> E + Ar"]}, {"Ar": 0.0}, {"Ar": 1}, Tg, 32)
    ze = z["E + Ar -> E + Ar"]["energy"]
    assert ze[0] == 0.0 and ze[1] == pytest.approx(1e-4) and ze[-1] == pytest.approx(10.0)
    f = tmp_path / "lumped.txt"
    cs.write_lxcat(str(f), {**out, **z})
    back = cs.read_cross_section_data(str(f))
    for k, r in {**out, **z}.items():
        np.testing.assert_allclose(back[k]["energy"], r["energy"], rtol=1e-12)
        np.testing.assert_allclose(back[k]["cross section"], r["cross section"], rtol=1e-12)
        assert back[k]["type"] == r["type"]
    assert back["E + Ar -> E + Ar"]["mass_ratio"] == pytest.approx(1.36e-5)
    with pytest.raises(cs.LXCatError):
        cs.lump_cross_sections(data | {"E + Ar -> E + Ar": _rec("Ar", e1, s1, ptype="ELASTIC")},
                               {"x": ["E + Ar -> E + ArA", "E + Ar -> E + Ar"]}, energy | {"Ar": 0.0},
                               degen | {"Ar": 1}, Tg)
