"""LXCat cross-section files (host-side parse, once per grid setup).

Mirrors the reference interface cross_section.read_available_species /
read_cross_section_data (BESolver/python/cross_section.py:31-69): a dict keyed
by the PROCESS name (text before the first comma) with 'type', 'species',
'energy', 'cross section', 'threshold', 'mass_ratio', 'info'.  The reference
delegates the parse to the third-party lxcat_data_parser; this parser reads
the LXCat block layout directly: a process-type line, a target line, a
parameter line (threshold for EXCITATION/IONIZATION, m/M for
ELASTIC/EFFECTIVE, energy for ATTACHMENT), 'KEY: value' lines, and the data
table between two dashed lines.
"""
import numpy as np

from .vspace import XSection

PROCESS_TYPES = ("ELASTIC", "EFFECTIVE", "EXCITATION", "IONIZATION", "ATTACHMENT")


class LXCatError(ValueError):
    pass


def _parse(text):
    lines = text.splitlines()
    recs = []
    i, n = 0, len(lines)
    while i < n:
        tok = lines[i].strip().upper()
        if tok not in PROCESS_TYPES:
            i += 1
            continue
        if i + 2 >= n:
            raise LXCatError("truncated process block")
        target = lines[i + 1].strip()
        try:
            param = float(lines[i + 2].split()[0])
        except (IndexError, ValueError) as e:
            raise LXCatError(f"bad parameter line after {tok}: {lines[i + 2]!r}") from e
        info = {}
        j = i + 3
        while j < n and not lines[j].strip().startswith("-----"):
            if ":" in lines[j]:
                k, v = lines[j].split(":", 1)
                info[k.strip()] = v.strip()
            j += 1
        if j >= n:
            raise LXCatError("missing data table")
        j += 1
        rows = []
        while j < n and not lines[j].strip().startswith("-----"):
            parts = lines[j].split()
            if len(parts) >= 2:
                rows.append((float(parts[0]), float(parts[1])))
            j += 1
        if not rows:
            raise LXCatError("empty data table")
        data = np.array(rows, dtype=np.float64)
        process = info.get("PROCESS", target).split(",")[0].strip()
        species = info.get("SPECIES", "e / " + target).split("/")[-1].strip()
        recs.append({
            "process": process, "type": tok, "species": species, "info": info,
            "energy": data[:, 0].copy(), "cross section": data[:, 1].copy(),
            "threshold": param if tok in ("EXCITATION", "IONIZATION") else 0.0,
            "mass_ratio": param if tok in ("ELASTIC", "EFFECTIVE") else None,
            "attachment_energy": param if tok == "ATTACHMENT" else 0.0,
        })
        i = j + 1
    return recs


def read_available_species(file):
    """cross_section.py:31-40."""
    with open(file) as f:
        species = [line.split(":")[1].split("/")[1].strip() for line in f if "SPECIES:" in line]
    return list(sorted(set(species), key=species.index))


def read_cross_section_data(file):
    """cross_section.py:42-69: dict keyed by process name, in file order."""
    with open(file) as f:
        recs = _parse(f.read())
    out = {}
    for r in recs:
        out[r["process"]] = r
    return out


def xsections_from_file(file, device=None):
    """Device cross sections of every process; m/M of each process is the elastic one
    (collisions.py:75, bte_0d3v_batched.py:186-214)."""
    data = read_cross_section_data(file)
    mbym = None
    for r in data.values():
        if r["type"] in ("ELASTIC", "EFFECTIVE"):
            mbym = r["mass_ratio"]
            break
    if mbym is None:
        raise LXCatError("no ELASTIC process: the mass ratio is undefined")
    return [XSection(r["type"], r["energy"], r["cross section"], r["threshold"], mbym, r["attachment_energy"],
                     r["species"], name, device=device) for name, r in data.items()]


# ----------------------------------------------------------------------------- preprocessing

def compute_mw_reaction_rates(Te, cs_data, normalized_v_domain=(0, 30), num_vr_pts=1000):
    """cross_section.compute_mw_reaction_rates (cross_section.py:71-110): the rate
    coefficients k = <sigma v> of Maxwellian EEDFs at the electron temperatures Te (eV) for
    the cross sections cs_data (callables sigma(eV)), by the composite trapezoid rule on
    num_vr_pts points of v / vth over normalized_v_domain.  Returns (len(Te), len(cs_data)).
    Host-side set-up work, as in the reference."""
    import scipy.constants as sc
    Te = np.asarray(Te, dtype=np.float64).reshape(-1)
    qe, me, kB = sc.electron_volt, sc.electron_mass, sc.Boltzmann
    c_gamma = np.sqrt(2 * qe / me)
    vr = np.linspace(normalized_v_domain[0], normalized_v_domain[1], num_vr_pts)
    w = np.full(num_vr_pts, vr[1] - vr[0])
    w[0] *= 0.5
    w[-1] *= 0.5
    vth = np.sqrt(Te) * c_gamma
    f0 = 4 * np.pi * (1.0 / (vth[:, None] * np.sqrt(np.pi)) ** 3) * np.exp(-vr[None, :] ** 2)
    mass = (vth ** 3 * ((vr ** 2 * f0) @ w)).reshape(-1, 1)
    f0n = f0 / mass
    temp = (2.0 / 3.0 / kB) * (0.5 * me) * vth ** 5 * ((vr ** 4 * f0n) @ w) / (qe / kB)
    if np.linalg.norm(Te - temp) / np.linalg.norm(Te) >= 1e-10:
        raise ValueError("the quadrature does not resolve the Maxwellians: widen normalized_v_domain or add points")
    rf = np.zeros((Te.shape[0], len(cs_data)))
    for i, cs in enumerate(cs_data):
        for j in range(Te.shape[0]):
            rf[j, i] = vth[j] ** 4 * ((vr ** 3 * cs((vth[j] * vr / c_gamma) ** 2) * f0n[j]) @ w)
    return rf


def lump_cross_sections(crs_data, lump_mechanism, species_energy, species_degeneracy, Tg, num_cs_pts=256):
    """The cross-section lumping of lump_crs.py (lump_crs.py:103-151): every lumped process is
    the mix sum_r w_r sigma_r(eps) of its member processes (all of one type), with Boltzmann
    weights w_r ~ g_r exp(-E_r / Tg) of the members' initial states (degeneracy g, energy E in
    eV, Tg in eV), normalised to one.  The energy grid spans [max(min member start, 1e-4),
    min member end]: num_cs_pts log-spaced points, with 0 prepended in place of the first
    point when the range starts at 1e-4; members are interpolated linearly and held
    constant beyond their tables.  Returns {lumped name: record} with the keys of
    read_cross_section_data; the threshold of a lumped process is the start of its grid.
    An elastic process is passed through, never lumped with another."""
    out = {}
    for name, members in lump_mechanism.items():
        recs = [crs_data[r] for r in members]
        types = {r["type"] for r in recs}
        if len(types) != 1:
            raise LXCatError(f"{name}: members of different types {sorted(types)}")
        ptype = types.pop()
        if ptype in ("ELASTIC", "EFFECTIVE") and len(members) != 1:
            raise LXCatError(f"{name}: elastic collisions cannot be lumped")
        wts = np.array([species_degeneracy[r["species"]] * np.exp(-species_energy[r["species"]] / Tg) for r in recs])
        wts = wts / wts.sum()
        ev_min = min(r["energy"][0] for r in recs)
        ev_max = min(r["energy"][-1] for r in recs)
        lo = max(ev_min, 1e-4)
        if lo == 1e-4:
            grid = np.append([0.0], np.logspace(np.log10(lo), np.log10(ev_max), num_cs_pts - 1))
        else:
            grid = np.logspace(np.log10(lo), np.log10(ev_max), num_cs_pts)
        sig = np.zeros_like(grid)
        for wt, r in zip(wts, recs):
            e, s = r["energy"], r["cross section"]
            sig += wt * np.interp(grid, e, s, left=s[0], right=s[-1])
        species = name.split("->")[0].split("+")[1].strip()
        out[name] = {"process": name, "type": ptype, "species": species, "energy": grid, "cross section": sig,
                     "threshold": lo if ptype in ("EXCITATION", "IONIZATION") else 0.0,
                     "mass_ratio": recs[0]["mass_ratio"] if ptype in ("ELASTIC", "EFFECTIVE") else None,
                     "attachment_energy": recs[0].get("attachment_energy", 0.0), "weights": wts,
                     "info": {"COMMENT": f"lumped at a Maxwellian of {Tg:.2E} eV"}}
    return out


def write_lxcat(file, records, comment=""):
    """Writes process records (read_cross_section_data / lump_cross_sections layout) as LXCat
    blocks: type, target, threshold (or m/M for elastic), SPECIES, PROCESS, PARAM., COMMENT,
    COLUMNS and the data table between dashed lines (the block lump_crs.py:138-151 writes)."""
    with open(file, "w") as f:
        for name, r in records.items():
            t = r["type"].upper()
            f.write(t + "\n" + r["species"] + "\n")
            if t in ("ELASTIC", "EFFECTIVE"):
                f.write("%.6E\n" % r["mass_ratio"])
            elif t == "ATTACHMENT":
                f.write("%.6E\n" % r.get("attachment_energy", 0.0))
            else:
                f.write("%.6E\n" % r["threshold"])
            f.write("SPECIES: e / %s\n" % r["species"])
            f.write("PROCESS: %s, %s\n" % (name, t.capitalize()))
            if t in ("ELASTIC", "EFFECTIVE"):
                f.write("PARAM.: m/M = %.6E, complete set\n" % r["mass_ratio"])
            f.write("COMMENT: %s\n" % (comment or r.get("info", {}).get("COMMENT", "")))
            f.write("COLUMNS: Energy (eV) | Cross section (m2)\n")
            f.write("-----------------------------\n")
            for e, s in zip(r["energy"], r["cross section"]):
                f.write("%.12E\t%.12E\n" % (e, s))
            f.write("-----------------------------\n\n")
