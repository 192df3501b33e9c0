"""Drop-in single-instance 0D solver with the reference's interface, on the MI355X.

Mirrors bte_0d3v (BESolver/python/bte_0d3v_solver.py:75-1276) for the B-spline x
spherical-harmonic discretisation: ``bte_0d3v(args)``, ``setup()``,
``initialize()``, ``steady_state_solver(h_init=None)`` and
``compute_QoIs(hh, tgrid, effective_mobility=True)``; results are numpy arrays,
as in the reference.

The single path differs from the batched one in three places, all carried here:
  * grid: k_domain = (0, sqrt(ev_max) c_gamma / vth), no 1.2x extension and no
    threshold snapping (bte_0d3v_solver.py:161-190, use_dg = 0);
  * collisions: every process is assembled with tgK = Tg and scaled by n0 x_s
    (attachment by (n0 x_s)^2, bte_0d3v_solver.py:225-233);
  * growth rate: W = u^T M^-1 (C + E A), advection included (:787), so the fixed
    point is the dominant generalized eigenpair of (C + E A, M).
The solve runs in the fused Newton kernel (bz_steady_solve_0d) with one instance:
the reference Jacobian (QT L Q - (W.h) I, :775-778) by default, step halving
while |r| does not decrease (Armijo with c = 0, :52-56), rtol = 1e-13 on |r|/|r0|,
atol = 1e-8 on |r|, at most 300 iterations (:806).

Not available here: the BOLSIG+ run that sets the reference's vth and eps_max
(bte_0d3v_solver.py:124-158) - ``args.electron_volt`` and ``args.ev_max`` must be
given, and ``h_bolsig`` is None; the FVM v-space (vstype 1); DG splines
(use_dg 1); e-e collisions (ee_collisions 1).
"""
import copy
import math

import numpy as np
import torch

from . import _lib
from .cross_section import read_available_species, read_cross_section_data, xsections_from_file
from .solver0d import Batched0D
from .vspace import C_GAMMA, VGrid, make_knots, vth_from_Te


class bte_0d3v:

    def __init__(self, args) -> None:
        self._args = args
        self._n0 = args.n0
        self._tgK = args.Tg
        if getattr(args, "vstype", 0) != 0:
            raise NotImplementedError("only the B-spline x spherical-harmonic v-space (vstype 0)")
        self._dev = _lib.require_device()
        self._avail_species = read_available_species(args.collisions)
        self._cross_section_data = read_cross_section_data(args.collisions)
        self._coll_list = xsections_from_file(args.collisions, device=self._dev)
        self._collision_names = ["C%d" % (i + 1) for i in range(len(self._coll_list))]
        self.num_collisions = len(self._coll_list)
        self._solver = None

    def get_collision_list(self):
        return self._coll_list

    def get_collision_names(self):
        return self._collision_names

    def get_cross_section_data(self):
        return self._cross_section_data

    def run_bolsig_solver(self):
        raise NotImplementedError("BOLSIG+ is not part of this build: set args.electron_volt and args.ev_max")

    def setup(self):
        """bte_0d3v_solver.py:462-505 (B-spline branch), on the GPU."""
        args = self._args
        if getattr(args, "use_dg", 0) != 0:
            raise NotImplementedError("DG splines (use_dg = 1)")
        if getattr(args, "ee_collisions", 0) != 0:
            raise NotImplementedError("e-e collisions on the single path")
        self._vth = float(vth_from_Te(args.electron_volt))
        self._c_gamma = C_GAMMA
        kd = (0.0, math.sqrt(args.ev_max) * C_GAMMA / self._vth)
        t = make_knots(kd, args.sp_order, args.Nr + 1, None, extend_domain=False)
        self._grid = VGrid(t, args.Nr + 1, args.l_max, args.spline_qpts, self._vth, device=self._dev)
        # one operator group per (species, density power): attachment scales with (n0 x_s)^2
        species = read_available_species(args.collisions)
        xs_by_group = []
        self._group_coef = []
        ns_by_n0 = list(getattr(args, "ns_by_n0", [1.0] * len(species)))
        for x in self._coll_list:
            sidx = species.index(x.species) if x.species in species else 0
            xsn = ns_by_n0[sidx] * self._n0
            if x.ptype == 3:                       # attachment: its own group, coefficient (n0 x_s)^2
                x = copy.copy(x)
                x.species = x.species + "#att"
                coef = xsn ** 2
            else:
                coef = xsn
            xs_by_group.append(x)
            if x.species not in [g for g, _ in self._group_coef]:
                self._group_coef.append((x.species, coef))
        self._solver = Batched0D(self._grid, xs_by_group, growth="all")
        assert [g for g, _ in self._group_coef] == self._solver.species
        el = [c for (g, c), s in zip(self._group_coef, self._solver.species)
              if any(x.ptype == 0 and x.species == s for x in xs_by_group)]
        # the elastic kappa term is assembled with tgK = Tg inside n0 x_s C_s: the kernel's
        # gas-temperature coefficient n0 * Tg * tg_scale reproduces it with tg_scale = x_s
        self._tg_scale = (el[0] / self._n0) if el else 0.0
        self._mw = lambda v: math.exp(-v ** 2) / (math.sqrt(math.pi) * self._vth) ** 3
        return

    def initialize(self, init_type="maxwellian"):
        """bte_0d3v_solver.py:507-557 (B-spline branch, Maxwellian at vth)."""
        if init_type != "maxwellian":
            raise NotImplementedError(init_type)
        return self._solver.h0.cpu().numpy()

    def steady_state_solver(self, h_init=None, jacobian="reference", rtol=1e-13, atol=1e-8, max_iter=300):
        """bte_0d3v_solver.py:741-859 (no e-e).  Returns {'sol','h_bolsig','atol','rtol','tgrid'}
        with sol[0] the normalised initial data and sol[1] the normalised steady state."""
        s = self._solver
        dev = self._dev
        if h_init is None:
            x0 = s.h0.reshape(-1, 1).clone()
        else:
            x0 = torch.as_tensor(np.asarray(h_init, dtype=np.float64), device=dev).reshape(-1, 1).clone()
        x = x0.clone().contiguous()
        one = lambda v: torch.tensor([float(v)], dtype=torch.float64, device=dev)
        ns = torch.tensor([[c / self._n0] for _, c in self._group_coef], dtype=torch.float64, device=dev)
        sc = s.solve_raw(x, one(self._n0), one(self._tgK), one(self._args.E_field), ns.contiguous(), atol=atol,
                         rtol=rtol, max_iter=max_iter, alpha_min=1e-30, alpha_rho=0.5, ls_c=0.0,
                         mode=1 if jacobian == "exact" else 0, tg_scale=self._tg_scale)
        h0, _ = s.qois(x0.contiguous(), one(self._n0), one(self._args.E_field))
        h1, _ = s.qois(x, one(self._n0), one(self._args.E_field))
        sol = np.zeros((2, self._grid.n))
        sol[0] = h0[:, 0].cpu().numpy()
        sol[1] = h1[:, 0].cpu().numpy()
        self.last_info = {k: sc[k].cpu().numpy()[0] for k in ("iters", "status", "res_abs", "res0")}
        return {"sol": sol, "h_bolsig": None, "atol": float(self.last_info["res_abs"]),
                "rtol": float(self.last_info["res_abs"] / self.last_info["res0"]), "tgrid": None}

    def compute_QoIs(self, hh, tgrid, effective_mobility=True):
        """bte_0d3v_solver.py:1234-1276 for normalised distributions hh (rows)."""
        args = self._args
        s = self._solver
        dev = self._dev
        h = torch.as_tensor(np.asarray(hh, dtype=np.float64).reshape(-1, self._grid.n), device=dev).t().contiguous()
        B = h.shape[1]
        E_max = args.E_field
        period = getattr(args, "efield_period", 0)
        if period > 0 and tgrid is not None:
            e_t = E_max * np.sin(np.pi * 2 * np.asarray(tgrid, dtype=np.float64) / period)
            e_t[np.abs(e_t) < 1e-2] = 0
            E = torch.as_tensor(np.broadcast_to(e_t, (B,)).copy(), device=dev)
        else:
            E = torch.full((B,), float(E_max), dtype=torch.float64, device=dev)
        n0 = torch.full((B,), float(self._n0), dtype=torch.float64, device=dev)
        # the kernel normalises first; the reference's hh is already normalised, and the
        # normalisation of a normalised distribution is the identity
        _, q = s.qois(h, n0, E, effective_mobility)
        rates = q["rates"].cpu().numpy()
        return {"energy": q["energy"].cpu().numpy(), "mobility": q["mobility"].cpu().numpy(),
                "diffusion": q["diffusion"].cpu().numpy(), "rates": [rates[i] for i in range(rates.shape[0])]}
