"""Parameters of the 1D glow discharge: the reference's glow1d_utils.parameters
(BESolver/python/glow1d_utils.py:26-130).

Same attributes and nondimensionalisation: L is half the electrode gap, tau the RF period,
n0 the gas density in units of np0 = 8e16 m^-3, the ion and electron transport
coefficients scaled by tau / L^2 (diffusivities) and V0 tau / L^2 (mobilities), and alpha =
np0 L^2 qe / (eps0 V0) the Poisson coupling.  With args.par_file set, the [glow_1d] table
of that TOML file replaces the default 2.54 cm / 100 V / 13.56 MHz / 1 Torr / 300 K argon
case.  The reference's default branch never sets the boundary temperature Teb, so its
Maxwellian flux ks is defined only from a parameter file here; the glow step does not use
either.
"""
import numpy as np
import scipy.constants as sc


def _load_toml(path):
    try:
        import tomllib as tl
    except ImportError:
        import tomli as tl
    with open(path, "rb") as f:
        return tl.load(f)


class parameters:
    def __init__(self, args):
        par_file = getattr(args, "par_file", "") or ""
        if par_file != "":
            tp = _load_toml(par_file)["glow_1d"]
            self.L = 0.5 * tp["L"]
            self.V0 = tp["V0"]
            self.f = tp["freq"]
            self.Teb = tp.get("Teb")
            self.gamma = tp.get("gamma")
            self.Tg = tp["Tg"]
            self.p0 = tp["p0"]
            De, mu_e, Di, mu_i = tp["De"], tp["mu_e"], tp["Di"], tp["mu_i"]
        else:
            self.L = 0.5 * 2.54e-2
            self.V0 = 1e2
            self.f = 13.56e6
            self.Teb = None
            self.gamma = None
            self.Tg = 300.0
            self.p0 = 1.0
            De, mu_e, Di, mu_i = 3.86e22 * 1e2, 9.66e21 * 1e2, 2.07e18 * 1e2, 4.65e19 * 1e2
        self.tau = 1.0 / self.f
        self.qe = sc.e
        self.eps0 = sc.epsilon_0
        self.kB = sc.Boltzmann
        self.ev_to_K = sc.electron_volt / sc.Boltzmann
        self.me = sc.electron_mass
        self.n0 = self.p0 * sc.torr / (sc.Boltzmann * self.Tg) if self.Tg != 0 else 3.22e22
        self.np0 = 8e16
        self._De = De / self.n0
        self._mu_e = mu_e / self.n0
        self.Di = Di / self.n0
        self.mu_i = mu_i / self.n0

        self.mw_flux = lambda Te: (self.tau / self.L) * (self.kB * Te * self.ev_to_K / (2 * np.pi * self.me)) ** 0.5
        self.n0 /= self.np0
        self._De *= self.tau / self.L ** 2
        self.Di *= self.tau / self.L ** 2
        self._mu_e *= self.V0 * self.tau / self.L ** 2
        self.mu_i *= self.V0 * self.tau / self.L ** 2
        self.Hi = 15.76
        if self.Teb is not None:
            self.ks = self.mw_flux(self.Teb)
            self.ks0 = self.ks1 = self.ks
            self.Teb0 = self.Teb1 = self.Teb
        self.alpha = self.np0 * self.L ** 2 * self.qe / (self.eps0 * self.V0)
