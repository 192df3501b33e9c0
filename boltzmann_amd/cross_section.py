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
