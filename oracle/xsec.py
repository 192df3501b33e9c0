"""Oracle: LXCat text parsing and cross-section interpolation (FP64).

Restates
  * cross_section.read_available_species / read_cross_section_data
    (BESolver/python/cross_section.py:31-69): dict keyed by the PROCESS name
    (text before the first comma) with type, species, energy, cross section,
    threshold and mass_ratio.  The reference delegates the parse to the
    third-party ``lxcat_data_parser`` (not installed here, no version pinned);
    this restates the published LXCat block layout: a type line, a species
    line, a parameter line (threshold in eV for EXCITATION/IONIZATION, m/M
    for ELASTIC/EFFECTIVE, energy for ATTACHMENT), key lines, and the data
    table between two dashed lines.
  * Collisions.load_cross_section / total_cross_section default mode
    (collisions.py:64-79, 235-262): scipy interp1d(kind='linear',
    bounds_error=False, fill_value=(sigma[0], sigma[-1])), i.e. piecewise
    linear inside the table and flat outside.
Pinned by tests/test_oracle_xsec.py against scipy interp1d and the reference's
own fixture file lxcat_data/eAr_crs.synthetic.3sp2r (values re-read as text).
"""
import numpy as np

PROCESS_TYPES = ("ELASTIC", "EFFECTIVE", "EXCITATION", "IONIZATION", "ATTACHMENT")


def parse_lxcat(text):
    """Return list of dicts (in file order) with keys
    type, species, process, threshold, mass_ratio, energy, cross section, info."""
    lines = text.splitlines()
    out = []
    i = 0
    n = len(lines)
    while i < n:
        tok = lines[i].strip().upper()
        if tok in PROCESS_TYPES:
            ptype = tok
            species_line = lines[i + 1].strip()
            param = float(lines[i + 2].split()[0])
            info = {}
            j = i + 3
            while j < n and not lines[j].strip().startswith("-----"):
                if ":" in lines[j]:
                    key, val = lines[j].split(":", 1)
                    info[key.strip()] = val.strip()
                j += 1
            j += 1
            data = []
            while j < n and not lines[j].strip().startswith("-----"):
                parts = lines[j].split()
                if len(parts) >= 2:
                    data.append((float(parts[0]), float(parts[1])))
                j += 1
            data = np.array(data, dtype=np.float64)
            proc = info.get("PROCESS", species_line).split(",")[0].strip()
            sp = info.get("SPECIES", "e / " + species_line).split("/")[-1].strip()
            rec = {
                "type": ptype, "species": sp, "process": proc, "info": info,
                "energy": data[:, 0].copy(), "cross section": data[:, 1].copy(),
                "threshold": param if ptype in ("EXCITATION", "IONIZATION") else 0.0,
                "mass_ratio": param if ptype in ("ELASTIC", "EFFECTIVE") else None,
                "attachment_energy": param if ptype == "ATTACHMENT" else 0.0,
            }
            out.append(rec)
            i = j + 1
        else:
            i += 1
    return out


def read_cross_section_data(path):
    with open(path, "r") as f:
        recs = parse_lxcat(f.read())
    species = []
    for r in recs:
        if r["species"] not in species:
            species.append(r["species"])
    return recs, species


def interp_linear_flat(energy_tab, sigma_tab, ev):
    """collisions.py:79 interp1d(kind='linear', fill_value=(s[0], s[-1])).

    Inside [e0, e_last]: y_lo + (x - x_lo) * (y_hi - y_lo) / (x_hi - x_lo)
    with the interval found by searchsorted(side='left') - 1, clipped, which
    is scipy's choice for points equal to a table node."""
    x = np.asarray(ev, dtype=np.float64)
    e = energy_tab
    s = sigma_tab
    idx = np.searchsorted(e, x, side="left")
    idx = np.clip(idx, 1, len(e) - 1)
    lo = idx - 1
    hi = idx
    slope = (s[hi] - s[lo]) / (e[hi] - e[lo])
    y = slope * (x - e[lo]) + s[lo]
    y = np.where(x < e[0], s[0], y)
    y = np.where(x > e[-1], s[-1], y)
    return y
