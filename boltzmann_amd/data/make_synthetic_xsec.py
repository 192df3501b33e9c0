"""Write the synthetic argon-like 3-process LXCat file used by the benchmarks.

Analytic forms (energies in eV, cross sections in m^2):
  elastic (momentum transfer, m/M = 1.36e-5), Ramsauer minimum near 0.25 eV:
      s_el(e) = s0 / (1 + (e/e1)^1.5) * (1 - D exp(-(ln(e/ed))^2 / (2 w^2)))
      s0 = 7.5e-20, e1 = 15, D = 0.9, ed = 0.25, w = 0.5
  excitation, threshold 11.5 eV, and ionization, threshold 15.76 eV:
      s(e) = A * E * (1 - exp(-(e - eth)/delta))^2 * ln(e/eth) / (e/eth),  e > eth
      (E = Euler's number so A is the peak of the ln(y)/y factor), delta = 2,
      A_ex = 1.2e-20, A_ion = 3.0e-20.
Each process is written under its own species name (Ar, Ar_x, Ar_i) so that a
per-process density factor can be carried in the batched solver's ``ns_by_n0``
array, one row per process.

Usage: python make_synthetic_xsec.py [out_path]
"""
import os
import sys
import numpy as np

MASS_RATIO = 1.36e-5
E_EX = 11.5
E_ION = 15.76


def sigma_elastic(e):
    e = np.asarray(e, dtype=np.float64)
    s0, e1, D, ed, w = 7.5e-20, 15.0, 0.9, 0.25, 0.5
    with np.errstate(divide="ignore"):
        lg = np.where(e > 0, np.log(np.maximum(e, 1e-300) / ed), -np.inf)
    dip = 1.0 - D * np.exp(-(lg ** 2) / (2 * w * w))
    return s0 / (1.0 + (e / e1) ** 1.5) * dip


def _inelastic(e, eth, amp, delta=2.0):
    e = np.asarray(e, dtype=np.float64)
    y = np.maximum(e, eth) / eth
    rise = (1.0 - np.exp(-(np.maximum(e, eth) - eth) / delta)) ** 2
    s = amp * np.e * rise * np.log(y) / y
    return np.where(e > eth, s, 0.0)


def sigma_excitation(e):
    return _inelastic(e, E_EX, 1.2e-20)


def sigma_ionization(e):
    return _inelastic(e, E_ION, 3.0e-20)


def _block(ptype, species, param, process, param_str, energy, sigma):
    rows = "\n".join("%.9E\t%.9E" % (a, b) for a, b in zip(energy, sigma))
    return (f"{ptype}\n{species}\n{param:.9E}\n"
            f"SPECIES: e / {species}\n"
            f"PROCESS: {process}, {ptype.capitalize()}\n"
            f"PARAM.:  {param_str}\n"
            f"COMMENT: synthetic argon-like cross section\n"
            f"COLUMNS: Energy (eV) | Cross section (m2)\n"
            f"-----------------------------\n{rows}\n-----------------------------\n\n")


def build_text():
    e_el = np.concatenate([[0.0], np.logspace(-4, 4, 801)])
    e_ex = E_EX + np.concatenate([[0.0], np.logspace(-3, np.log10(1e4 - E_EX), 600)])
    e_io = E_ION + np.concatenate([[0.0], np.logspace(-3, np.log10(1e4 - E_ION), 600)])
    txt = _block("ELASTIC", "Ar", MASS_RATIO, "E + Ar -> E + Ar", "m/M = %.8E" % MASS_RATIO,
                 e_el, sigma_elastic(e_el))
    txt += _block("EXCITATION", "Ar_x", E_EX, "E + Ar_x -> E + Ar*", "E = %.8E eV" % E_EX,
                  e_ex, sigma_excitation(e_ex))
    txt += _block("IONIZATION", "Ar_i", E_ION, "E + Ar_i -> E + E + Ar+", "E = %.8E eV" % E_ION,
                  e_io, sigma_ionization(e_io))
    return txt


DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eAr_synthetic_3proc.lxcat")

if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_PATH
    with open(out, "w") as f:
        f.write(build_text())
    print("wrote", out)
