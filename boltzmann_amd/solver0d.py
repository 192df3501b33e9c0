"""Batched 0D steady-state solver on one grid (device-resident operators + fused Newton kernel).

``Batched0D`` assembles, on the GPU, the per-species collision operators (tgK = 0),
the gas-temperature operator, the field operator and the mass matrix of one
velocity grid (bte_0d3v_batched.assemble_operators, bte_0d3v_batched.py:310-479,
without the M^-1 premultiplication and the Q basis, which the kernel does not
need), builds the solver tables, and solves any batch of instances with one
launch of bz_steady_solve_0d followed by bz_qoi_0d.
"""
import math

import numpy as np
import torch

from . import _lib
from .plan0d import VEC_ROWS, SolverTables
from .vspace import C_GAMMA, QE, ME, TEMP_EV_FAC, VGrid, XSection


# ptype of the processes that change the number of electrons (ionisation 2, attachment 3)
NONCONSERVATIVE = (2, 3)


class Batched0D:
    def __init__(self, vgrid: VGrid, xsecs, growth="collisions", ee_pattern=False):
        """growth: 'collisions' (batched path: W_s = u M^-1 C_s, bte_0d3v_batched.py:636) or
        'all' (single path: W = u M^-1 (C + E A), bte_0d3v_solver.py:787)."""
        self.g = vgrid
        self.xsecs = list(xsecs)
        species = []
        for x in self.xsecs:
            if x.species not in species:
                species.append(x.species)
        self.species = species
        self.S = len(species)
        g = vgrid
        n = g.n
        dev = g.dev
        C = [torch.zeros((n, n), dtype=torch.float64, device=dev) for _ in species]
        C_grow = [torch.zeros((n, n), dtype=torch.float64, device=dev) for _ in species]
        Cg = torch.zeros((n, n), dtype=torch.float64, device=dev)
        for x in self.xsecs:
            Cx = g.collision_matrix(x, tgK=0.0)
            C[species.index(x.species)] += Cx
            if x.ptype in NONCONSERVATIVE:
                C_grow[species.index(x.species)] += Cx
            if x.ptype == 0:
                Cg += g.gas_temperature_matrix(x)
        self.elastic_species = [species.index(x.species) for x in self.xsecs if x.ptype == 0]
        self.mass_radial = g.mass_radial()
        M = g.mass_matrix()
        A = g.advection_matrix(scale=(1.0 / g.vth) * (QE / ME))
        self.C, self.Cg, self.A, self.M = C, Cg, A, M
        self.C_grow = C_grow
        # QoI vectors (original order)
        self.mass = g.mass_op()
        self.temp = g.temp_op() * TEMP_EV_FAC
        self.mobility = g.mobility_op()
        self.diffusion = g.diffusion_op(self.xsecs)
        self.rates = torch.stack([g.rate_op(x) for x in self.xsecs]).contiguous()
        self.mw0 = 1.0 / (math.sqrt(math.pi) * g.vth) ** 3
        self.u = (self.mass * (self.mw0 * g.vth ** 3)).contiguous()
        # growth vectors W_m = Op_m^T M^-1 u from the processes that change the electron number
        # (ionisation, attachment) and, on the single path, the field term; for particle-
        # conserving operators (elastic, excitation, gas temperature) the vector is zero, which
        # the reference's product only reproduces up to roundoff
        gm = [True] * self.S + ([False, False] if growth == "collisions" else [False, True])
        self.ee_pattern = ee_pattern
        extra = None
        if ee_pattern:
            # e-e couples (p, q) to (k, l) for |k - p| <= 3 and every pair of harmonics
            pr = torch.arange(n, device=dev) // g.num_sh
            extra = (pr[:, None] - pr[None, :]).abs() <= 3
        self.tables = SolverTables(g, C + [Cg, A, M], gm, self.u, self.mass_radial, extra_mask=extra,
                                   growth_ops=C_grow + [Cg, A])
        self.h0 = g.maxwellian_init()
        self._scratch = {}

    def initial(self, B):
        return self.h0.reshape(-1, 1).repeat(1, B).contiguous()

    def new_scratch(self, B):
        """Scratch of B instances owned by the caller (for batches in flight concurrently)."""
        t = self.tables
        dev = self.g.dev
        return dict(
            kprof=torch.empty((B, t.P), dtype=torch.float64, device=dev),
            urows=torch.empty((B, t.n, t.WP), dtype=torch.float64, device=dev),
            vec=torch.empty((B, VEC_ROWS, t.n), dtype=torch.float64, device=dev),
            iters=torch.empty(B, dtype=torch.int32, device=dev),
            status=torch.empty(B, dtype=torch.int32, device=dev),
            res_abs=torch.empty(B, dtype=torch.float64, device=dev),
            res0=torch.empty(B, dtype=torch.float64, device=dev))

    def _get_scratch(self, B):
        if B not in self._scratch:
            if len(self._scratch) >= 3:                # keep the three most recent batch sizes
                self._scratch.pop(next(iter(self._scratch)))
            self._scratch[B] = self.new_scratch(B)
        return self._scratch[B]

    def batch_struct(self, x, n0, Tg, E, ns_by_n0, atol=0.0, rtol=0.0, max_iter=0, alpha_min=1e-16, alpha_rho=0.5,
                     ls_c=1e-4, mode=1, tg_scale=1.0, tol_abs=None, scratch=None):
        """bz_batch0d_t over x (n, B) and the per-instance parameters, with this solver's
        scratch for B instances (or the given one, from new_scratch); returns (struct, scratch)."""
        B = x.shape[1]
        assert x.shape[0] == self.g.n and x.dtype == torch.float64 and x.is_contiguous()
        for v in (n0, Tg, E):
            assert v.shape == (B,) and v.dtype == torch.float64 and v.is_contiguous()
        assert ns_by_n0.shape == (self.S, B) and ns_by_n0.is_contiguous()
        if tol_abs is not None:
            assert tol_abs.shape == (B,) and tol_abs.dtype == torch.float64 and tol_abs.is_contiguous()
        sc = self._get_scratch(B) if scratch is None else scratch
        assert sc["kprof"].shape == (B, self.tables.P)
        bt = _lib.Batch0D(B=B, n0=n0.data_ptr(), Tg=Tg.data_ptr(), E=E.data_ptr(), ns_by_n0=ns_by_n0.data_ptr(),
                          x=x.data_ptr(), kprof=sc["kprof"].data_ptr(), urows=sc["urows"].data_ptr(),
                          vec=sc["vec"].data_ptr(), iters=sc["iters"].data_ptr(), status=sc["status"].data_ptr(),
                          res_abs=sc["res_abs"].data_ptr(), res0=sc["res0"].data_ptr(), atol=float(atol),
                          rtol=float(rtol), alpha_min=float(alpha_min), alpha_rho=float(alpha_rho),
                          ls_c=float(ls_c), max_iter=int(max_iter), mode=int(mode), tg_scale=float(tg_scale),
                          tol_abs=None if tol_abs is None else tol_abs.data_ptr())
        self._last_batch = bt
        return bt, sc

    def solve_raw(self, x, n0, Tg, E, ns_by_n0, atol=1e-10, rtol=1e-10, max_iter=30, alpha_min=1e-16,
                  alpha_rho=0.5, ls_c=1e-4, mode=1, tg_scale=1.0, stream=None, tol_abs=None, ee=None, scratch=None):
        """In-place Newton on x (n, B); returns the scratch dict (iters, status, res_abs, res0).
        tol_abs: optional (B,) per-instance absolute tolerance; ee: optional _lib.EE0D with the
        frozen e-e operator of every instance (bz_steady_solve_0d_ee)."""
        bt, sc = self.batch_struct(x, n0, Tg, E, ns_by_n0, atol=atol, rtol=rtol, max_iter=max_iter,
                                   alpha_min=alpha_min, alpha_rho=alpha_rho, ls_c=ls_c, mode=mode,
                                   tg_scale=tg_scale, tol_abs=tol_abs, scratch=scratch)
        if ee is None:
            _lib.call("bz_steady_solve_0d", self.tables.struct, bt, stream=stream)
        else:
            if not self.ee_pattern:
                raise _lib.BzError("the solver tables were planned without the e-e pattern (ee_pattern=True)")
            _lib.call("bz_steady_solve_0d_ee", self.tables.struct, bt, ee, stream=stream)
        return sc

    def transient_step(self, x, n0, Tg, E, ns_by_n0, dt_tau, atol=1e-10, rtol=1e-10, max_iter=30, ustart=None,
                       avg=None, tg_scale=1.0, stream=None, store_factor=True):
        """One backward-Euler step of size dt_tau for every instance (bz_transient_step_0d):
        x (n, B) holds the initial iterate on entry and the new state on exit; ustart is the
        state at the start of the step (None: x itself); avg += ustart + x_new if given.
        store_factor: keep the L factor of the frozen matrix (B x P doubles of scratch) so
        that it is factorised once per step instead of once per Newton iteration.
        Returns the scratch dict (iters, status, res_abs, res0)."""
        B = x.shape[1]
        assert x.shape[0] == self.g.n and x.dtype == torch.float64 and x.is_contiguous()
        for v in (n0, Tg, E):
            assert v.shape == (B,) and v.dtype == torch.float64 and v.is_contiguous()
        assert ns_by_n0.shape == (self.S, B) and ns_by_n0.is_contiguous()
        for v in (ustart, avg):
            assert v is None or (v.shape == x.shape and v.is_contiguous() and v.dtype == torch.float64)
        sc = self._get_scratch(B)
        bt = _lib.Batch0D(B=B, n0=n0.data_ptr(), Tg=Tg.data_ptr(), E=E.data_ptr(), ns_by_n0=ns_by_n0.data_ptr(),
                          x=x.data_ptr(), kprof=sc["kprof"].data_ptr(), urows=sc["urows"].data_ptr(),
                          vec=sc["vec"].data_ptr(), iters=sc["iters"].data_ptr(), status=sc["status"].data_ptr(),
                          res_abs=sc["res_abs"].data_ptr(), res0=sc["res0"].data_ptr(), atol=float(atol),
                          rtol=float(rtol), alpha_min=1e-16, alpha_rho=0.5, ls_c=0.0, max_iter=int(max_iter), mode=0,
                          tg_scale=float(tg_scale))
        if store_factor:
            if sc.get("lprof") is None:
                sc["lprof"] = torch.empty((B, self.tables.P), dtype=torch.float64, device=self.g.dev)
            bt.lprof = sc["lprof"].data_ptr()
        _lib.call("bz_transient_step_0d", self.tables.struct, bt, float(dt_tau), ustart, avg, stream=stream)
        return sc

    def transient(self, f0, n0, Tg, eRe, eIm, ns_by_n0, dt, cycles, Efreq, atol=1e-10, rtol=1e-10, max_iter=30):
        """bte_0d3v_batched.solve(..., "transient") (bte_0d3v_batched.py:1040-1161): backward
        Euler in RF periods (step dt * tau, tau = 1/Efreq), E(t) = eRe cos 2 pi t + eIm sin 2 pi t,
        each step's Newton started from the previous step's increment, trapezoidal average over
        each cycle; returns (u, h_avg, qoi of the last full cycle's average, newton iterations)."""
        tau = 1e-7 if Efreq == 0 else 1.0 / Efreq
        io_cycle = 1.0
        io_freq = int(io_cycle / dt)
        steps_total = int(cycles / dt)
        u = f0 / (self.u @ f0)[None, :]
        u = u.contiguous()
        u_prev = u.clone()
        avg = torch.zeros_like(u)
        avg_v = torch.zeros_like(u)
        E = torch.empty_like(n0)
        x = torch.empty_like(u)
        tt = 0.0
        iters = []
        for ts_idx in range(steps_total + 1):
            if ts_idx > 0 and ts_idx % io_freq == 0:
                avg.mul_(0.5 * dt / io_cycle)
                avg_v.copy_(avg)
                avg.zero_()
                if ts_idx == steps_total:
                    break
            ph = 2.0 * math.pi * (tt + dt)
            if Efreq > 0:
                torch.add(eRe * math.cos(ph), eIm, alpha=math.sin(ph), out=E)
            else:
                E.copy_(eRe)
            # initial iterate u + (u - u_prev): the previous step's increment
            torch.sub(2.0 * u, u_prev, out=x)
            sc = self.transient_step(x, n0, Tg, E, ns_by_n0, dt * tau, atol, rtol, max_iter, ustart=u, avg=avg)
            iters.append(sc["iters"].clone())
            u_prev, u = u, x.clone()
            tt += dt
        h, q = self.qois(avg_v.contiguous(), n0, torch.ones_like(n0), effective_mobility=False)
        return u, h, q, torch.stack(iters) if iters else None

    def qois(self, h, n0, E, effective_mobility=True, stream=None):
        """Normalised distribution and QoIs (bte_0d3v_batched.py:1163-1217)."""
        g = self.g
        n, B = h.shape
        G = len(self.xsecs)
        hout = torch.empty_like(h)
        out = torch.empty((B, 3 + G), dtype=torch.float64, device=g.dev)
        mm_fac = (1.0 / (2.0 * math.sqrt(math.pi))) * 4.0 * math.pi
        mm_scale = 2.0 * (g.vth / C_GAMMA) ** 3 / mm_fac
        _lib.call("bz_qoi_0d", n, g.num_sh, B, G, h, self.mass, self.temp, self.mobility, self.diffusion,
                  self.rates, E, n0, mm_scale, self.mw0 * g.vth ** 3, self.mw0 * g.vth ** 5, C_GAMMA,
                  1 if effective_mobility else 0, hout, out, stream=stream)
        q = {"energy": out[:, 0], "mobility": out[:, 1], "diffusion": out[:, 2], "rates": out[:, 3:].t()}
        return hout, q

    def solve(self, n0, Tg, E, ns_by_n0, x0=None, mode=1, **kw):
        B = n0.shape[0]
        x = (self.initial(B) if x0 is None else x0.clone()).contiguous()
        sc = self.solve_raw(x, n0, Tg, E, ns_by_n0, mode=mode, **kw)
        h, q = self.qois(x, n0, E)
        info = {k: sc[k] for k in ("iters", "status", "res_abs", "res0")}
        return h, q, info
