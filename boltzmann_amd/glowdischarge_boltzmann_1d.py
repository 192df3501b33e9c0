"""The 1D3V glow-discharge solver: the reference's glowdischarge_boltzmann_1d.glow1d_boltzmann
(BESolver/python/glowdischarge_boltzmann_1d.py:79-2402) on the MI355X.

The electron Boltzmann equation in (x, v) with a B-spline x spherical-harmonic v-space on
Np Chebyshev points, coupled to an ion fluid and the Poisson field through Strang splitting
(heavies_freeze_E, operator-split scheme 0):

    field from the Poisson problem -> ion half step (backward Euler)   bz_glow_fluid_step
    -> field of the electron step                                       (one launch)
    -> x half step -> v step -> x half step                             glow1d.XAdvection,
                                                                        glow1d.VStep
    -> moments ne, Te, ionisation rate                                  bz_glow_moments
    -> field -> ion half step                                           bz_glow_fluid_step

Every per-step operation runs on the GPU; the host only launches and reads the status of
the v-step GMRES (which it also reads once per restart cycle).  Set-up (operator assembly on
the GPU, the x-advection eigenbasis and the Chebyshev tables on the host) happens once.

State: u (Np, 3) device tensor, columns ne, ni, Te (densities in units of np0, Te in eV),
and v (num_p * Nvt, Np) device tensor, the ordinate representation of the distribution,
as in the reference.  Differences from the reference:
  * the ionisation rate is the rate of the file's IONIZATION process; the reference takes
    the second process of its two-process argon files, which is that process;
  * plots are not produced; checkpoints are the reference's .npy files (u, v and the
    cycle averages), and store_checkpoint / restore_checkpoint (bte_1d3v_solver.py:
    2065-2103) write and read the same datasets in a .npz, h5py not being available;
  * ts_type is "BE" for the fluid and the v-step (the reference's default; RK2/RK4/IMEX
    fluid steps and the fully implicit split scheme 1 are not provided).
"""
import argparse
import math
import os

import numpy as np
import torch

from . import _lib
from .collocation_op import cheb_collocation_1d
from .cross_section import xsections_from_file
from .glow1d import Ordinates, VStep, XAdvection, ordinates
from .glow1d_utils import parameters
from .vspace import C_GAMMA, QE, ME, TEMP_EV_FAC, VGrid, make_knots, vth_from_Te

MOMENT_SPLITS = 64
TILE_LEN = 512 * 80           # BZ_GLOW_TILE_LEN


def args_parse(argv=None):
    """The reference's command line (glowdischarge_boltzmann_1d.py:3387-3535), including the
    TOML parameter file that overrides it."""
    p = argparse.ArgumentParser()
    a = p.add_argument
    a("-threads", "--threads", type=int, default=4)
    a("-out_fname", "--out_fname", type=str, default="bte_glow1d")
    a("-l_max", "--l_max", type=int, default=1)
    a("-c", "--collisions", type=str, default="lxcat_data/eAr_crs.Biagi.3sp2r")
    a("-ev_max", "--ev_max", type=float, default=50)
    a("-ev_extend", "--ev_extend", type=int, default=2)
    a("-sp_order", "--sp_order", type=int, default=3)
    a("-spline_qpts", "--spline_qpts", type=int, default=5)
    a("-steady", "--steady_state", type=int, default=1)
    a("-Te", "--Te", type=float, default=0.5)
    a("-Tg", "--Tg", type=float, default=0.0)
    a("-Nr", "--Nr", type=int, default=128)
    a("-Nvt", "--Nvt", type=int, default=3)
    a("-Ns", "--Ns", type=int, default=2)
    a("-NT", "--NT", type=int, default=1)
    a("-Np", "--Np", type=int, default=100)
    a("-cfl", "--cfl", type=float, default=1e-1)
    a("-cycles", "--cycles", type=float, default=10)
    a("-ts_type", "--ts_type", type=str, default="BE")
    a("-atol", "--atol", type=float, default=1e-10)
    a("-rtol", "--rtol", type=float, default=1e-10)
    a("-gmres_atol", "--gmres_atol", type=float, default=1e-20)
    a("-gmres_rtol", "--gmres_rtol", type=float, default=1e-1)
    a("-gmres_rsrt", "--gmres_rsrt", type=int, default=10)
    a("-max_iter", "--max_iter", type=int, default=1000)
    a("-profile", "--profile", type=int, default=0)
    a("-warm_up", "--warm_up", type=int, default=5)
    a("-runs", "--runs", type=int, default=10)
    a("-store_eedf", "--store_eedf", type=int, default=0)
    a("-use_gpu", "--use_gpu", type=int, default=0)
    a("-gpu_device_id", "--gpu_device_id", type=int, default=0)
    a("-store_csv", "--store_csv", type=int, default=1)
    a("-plot_data", "--plot_data", type=int, default=1)
    a("-ee_collisions", "--ee_collisions", type=float, default=0)
    a("-verbose", "--verbose", type=int, default=0)
    a("-restore", "--restore", type=int, default=0)
    a("-rs_idx", "--rs_idx", type=int, default=0)
    a("-ic_neTe", "--ic_neTe", type=int, default=0)
    a("-ca_macro_qoi", "--ca_macro_qoi", type=int, default=0)
    a("-fname", "--fname", type=str, default="1d_glow")
    a("-dir", "--dir", type=str, default="glow1d_dir")
    a("-glow_op_split_scheme", "--glow_op_split_scheme", type=int, default=0)
    a("-par_file", "--par_file", type=str, default="")
    a("-ic_file", "--ic_file", type=str, default="")
    a("-io_cycle_freq", "--io_cycle_freq", type=float, default=1e0)
    a("-cp_cycle_freq", "--cp_cycle_freq", type=float, default=1e1)
    a("-bte_with_E", "--bte_with_E", type=int, default=0)
    a("-vtDe", "--vtDe", type=float, default=0)
    a("-xadv_type", "--xadv_type", type=int, default=0)
    args = p.parse_args(argv)
    if args.par_file != "":
        from .glow1d_utils import _load_toml
        tp = _load_toml(args.par_file)
        b = tp["bte"]
        for k_arg, k_toml in (("Np", "Np"), ("l_max", "lmax"), ("collisions", "collisions"), ("ev_max", "ev_max"),
                              ("ev_extend", "ev_extend"), ("sp_order", "sp_order"), ("spline_qpts", "spline_qpts"),
                              ("Te", "Te"), ("Nr", "Nr"), ("Nvt", "Nvt"), ("cfl", "dt"), ("cycles", "cycles"),
                              ("vtDe", "vtDe"), ("xadv_type", "xadv_type")):
            setattr(args, k_arg, b[k_toml])
        args.Tg = tp["glow_1d"]["Tg"]
        s = tp["solver"]
        for k_arg, k_toml in (("atol", "atol"), ("rtol", "rtol"), ("gmres_atol", "gmres_atol"),
                              ("gmres_rtol", "gmres_rtol"), ("gmres_rsrt", "gmres_rsrt"), ("max_iter", "max_iter"),
                              ("use_gpu", "use_gpu"), ("gpu_device_id", "gpu_device_id"), ("restore", "restore"),
                              ("rs_idx", "rs_idx"), ("fname", "fname"), ("dir", "dir"), ("ic_file", "ic_file"),
                              ("io_cycle_freq", "io_cycle"), ("cp_cycle_freq", "cp_cycle"),
                              ("glow_op_split_scheme", "split_scheme")):
            if k_toml in s:
                setattr(args, k_arg, s[k_toml])
        c = tp["chemistry"]
        args.Ns, args.NT = c["Ns"], c["NT"]
    return args


def default_args(**kw):
    """args_parse([]) with keyword overrides."""
    args = args_parse([])
    for k, v in kw.items():
        setattr(args, k, v)
    return args


GLOW_MOMENTS, GLOW_STEP, GLOW_FIELD_AFTER = 1, 2, 4


class glow1d_boltzmann:
    def __init__(self, args):
        if args.Nvt % 2:
            raise ValueError("Nvt must be even (half-range Gauss ordinates)")
        if args.Ns != 2 or args.NT != 1:
            raise NotImplementedError("the coupled step has two species (electrons, ions) and one temperature")
        if getattr(args, "ee_collisions", 0):
            raise NotImplementedError("e-e collisions are not part of the 1D3V glow step here")
        self.dev = _lib.require_device()
        self.args = args
        self.param = parameters(args)
        self.ts_type_fluid = "BE"
        self.ts_type_bte_v = "BE"
        self.heavies_freeze_E = True
        self.weak_bc_ni = True
        self.Ns, self.NT = args.Ns, args.NT
        self.Nv = args.Ns + args.NT
        self.Nr = args.Nr + 1
        self.Nvt = args.Nvt
        self.dof_v = self.Nr * (args.l_max + 1)
        self.Np = args.Np
        self.deg = self.Np - 1
        self.ele_idx, self.ion_idx, self.Te_idx = 0, 1, args.Ns
        self.fluid_idx = [self.ion_idx]
        self.Zp = np.array([-1.0, 1.0])

        cheb = cheb_collocation_1d(self.Np)
        self.xp = cheb.xp
        self.Dp, self.Lp, self.LpD, self.LpD_inv = cheb.Dp, cheb.Lp, cheb.LpD, cheb.LpD_inv
        self.xp_vt, self.xp_vt_qw = ordinates(self.Nvt)
        self.mu = np.zeros((self.Np, self.Ns))
        self.D = np.zeros((self.Np, self.Ns))

        self.bs_Te = args.Te
        self.bs_vth = vth_from_Te(args.Te)
        self.bs_lm = [[l, 0] for l in range(args.l_max + 1)]
        self.ev_lim = (0.0, args.ev_max)
        self.coll_list = xsections_from_file(args.collisions, device=self.dev)
        ion = [i for i, x in enumerate(self.coll_list) if x.ptype == 2]
        if len(ion) != 1:
            raise ValueError("the glow step needs exactly one IONIZATION process")
        self.ion_proc = ion[0]
        self.initialize_boltzmann()
        self._dt = None

    # ------------------------------------------------------------------ set-up
    def initialize_boltzmann(self):
        """initialize_boltzmann (:272-539) on the GPU: the grid, M^-1-premultiplied
        collision, gas-temperature and field operators, QoI vectors and the ordinate
        transforms."""
        args = self.args
        vth = self.bs_vth
        sig = sorted(set(x.threshold for x in self.coll_list if x.threshold > 0))
        dg_nodes = np.sqrt(np.array(sig, dtype=np.float64)) * C_GAMMA / vth
        kd = (math.sqrt(self.ev_lim[0]) * C_GAMMA / vth, math.sqrt(self.ev_lim[1]) * C_GAMMA / vth)
        if args.ev_extend == 0:
            t = make_knots(kd, args.sp_order, self.Nr, dg_nodes, extend_domain=False)
        elif args.ev_extend == 1:
            t = make_knots(kd, args.sp_order, self.Nr, dg_nodes, extend_domain=True)
        elif args.ev_extend == 2:
            t = make_knots(kd, args.sp_order, self.Nr, dg_nodes, extend_domain_with_log=True)
        else:
            raise ValueError(f"ev_extend {args.ev_extend}")
        g = VGrid(t, self.Nr, args.l_max, args.spline_qpts, vth, device=self.dev)
        self.vgrid = g
        Minv = g.inverse_mass_matrix()
        FOp = torch.zeros((g.n, g.n), dtype=torch.float64, device=self.dev)
        FOp_g = torch.zeros_like(FOp)
        for x in self.coll_list:
            FOp += g.collision_matrix(x, tgK=0.0)
            if x.ptype == 0:
                FOp_g += g.gas_temperature_matrix(x)
        A = g.advection_matrix(scale=(1.0 / vth) * (QE / ME))
        self.op_col_en = (Minv @ FOp).contiguous()
        self.op_col_gT = (Minv @ FOp_g).contiguous()
        self.op_adv_v = (Minv @ A).contiguous()
        self.op_mass = g.mass_op()
        self.op_temp = g.temp_op() * (vth ** 2) * TEMP_EV_FAC
        self.op_rate = [g.rate_op(x) for x in self.coll_list]
        self.ords = Ordinates(self.Nr, args.l_max, self.Nvt, self.dev)
        self.op_po2sh, self.op_psh2o = self.ords.po2sh, self.ords.psh2o
        self.mm_fac = math.sqrt(4.0 * math.pi)
        ns = g.num_sh
        po0 = self.op_po2sh[0]
        W = torch.stack([self.op_mass[0::ns], self.op_temp[0::ns], self.op_rate[self.ion_proc]])   # (3, num_p)
        self._W = (W[:, :, None] * po0[None, None, :]).reshape(3, -1).contiguous()             # (3, num_p Nvt)
        torch.cuda.synchronize()

    def _setup_fluid(self):
        """Device tables and buffers of the ion-fluid launches."""
        p, n = self.param, self.Np
        d = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.dev)
        Dp_mod = self.Dp.copy()
        Dp_mod[0, :] = 0.0
        Dp_mod[-1, :] = 0.0
        G = -(self.Dp @ self.LpD_inv)
        self._DpT = d(self.Dp.T)
        self._DppT = d((self.Dp @ Dp_mod).T)
        self._GT = d(G.T)
        # the two matrices again in the order of the step kernel's register tiles
        self._DpTile = torch.zeros(TILE_LEN, dtype=torch.float64, device=self.dev)
        self._DppTile = torch.zeros(TILE_LEN, dtype=torch.float64, device=self.dev)
        _lib.call("bz_glow_tile_tables", n, self._DpT, self._DppT, self._DpTile, self._DppTile)
        self._ki = torch.zeros(n, dtype=torch.float64, device=self.dev)
        self._Ef = torch.zeros(n, dtype=torch.float64, device=self.dev)
        self.bs_E = torch.zeros(n, dtype=torch.float64, device=self.dev)
        self._partial = torch.zeros((MOMENT_SPLITS, 3, n), dtype=torch.float64, device=self.dev)
        # status and info in pinned host memory: the kernel stores them there, the host reads
        # them after the stream has passed the launch, with no copy
        self._status = torch.zeros(4, dtype=torch.int32).pin_memory()
        self._info = torch.zeros(3, dtype=torch.float64).pin_memory()
        c3 = 2.0 * (self.bs_vth / C_GAMMA) ** 3
        self._fl = _lib.GlowFluid(n=n, DpT=self._DpT.data_ptr(), DppT=self._DppT.data_ptr(), GT=self._GT.data_ptr(),
                               mu_i=float(p.mu_i), Di=float(p.Di), n0=float(p.n0), alpha=float(p.alpha),
                               field_scale=float(p.V0 / p.L), atol=float(self.args.atol), rtol=float(self.args.rtol),
                               max_iter=int(self.args.max_iter), u=0, ki=self._ki.data_ptr(),
                               E=self._Ef.data_ptr(), bsE=self.bs_E.data_ptr(), partial=self._partial.data_ptr(),
                               nsplit=MOMENT_SPLITS, rate_scale=self.mm_fac * p.np0 * p.tau / c3,
                               status=self._status.data_ptr(), info=self._info.data_ptr(),
                               DpTile=self._DpTile.data_ptr(), DppTile=self._DppTile.data_ptr())
        self.mu[:, self.ele_idx] = p._mu_e
        self.D[:, self.ele_idx] = p._De
        self.mu[:, self.ion_idx] = p.mu_i
        self.D[:, self.ion_idx] = p.Di

    def _fluid_launch(self, u, flags, h=0.0, t_before=0.0, t_after=0.0):
        assert u.shape == (self.Np, 3) and u.dtype == torch.float64 and u.is_contiguous() and u.device == self.dev
        self._fl.u = u.data_ptr()
        _lib.call("bz_glow_fluid_step", self._fl, int(flags), float(h), math.sin(2.0 * math.pi * t_before),
                  math.sin(2.0 * math.pi * t_after))

    def _moments(self, v):
        rows = self.Nr * self.Nvt
        assert v.shape == (rows, self.Np) and v.is_contiguous()
        _lib.call("bz_glow_moments", v, rows, self.Np, self._W, MOMENT_SPLITS, self._partial)

    def check_fluid_status(self):
        """Raises if an ion step has failed since the last reset (the reference exits)."""
        if int(self._status[2]) > 0:
            info = self._info.tolist()
            self._status[2] = 0
            raise RuntimeError("Fluid step non-linear solver step FAILED: ||res|| = %.6e, ||res||/||res0|| = %.6e "
                               "(alpha %.3g); try a smaller time step or more iterations" % tuple(info))

    # ------------------------------------------------------------------ reference methods
    def initialize(self, type=0):
        """initialize (:541-657): u = (ne, ni, Te) and v (Nr Nvt, Np) on the device.  The
        default state: ne = ni = 1e6 (1e7 + 1e9 s^2 (1 - s)^2) / np0, s = (xp + 1) / 2, Te =
        args.Te, and at every point a unit-mass Maxwellian at Te times ne.  With
        args.restore the state of checkpoint rs_idx is read back (Chebyshev interpolation
        onto this grid when the point counts differ); args.ic_file starts from a saved u."""
        args, p = self.args, self.param
        if args.restore == 1:
            u = np.load("%s_%04d_u.npy" % (self._fname(), args.rs_idx))
            v = np.load("%s_%04d_v.npy" % (self._fname(), args.rs_idx))
            if u.shape[0] != self.Np:
                npts = u.shape[0]
                xx1 = -np.cos(np.pi * np.linspace(0, npts - 1, npts) / (npts - 1))
                V0inv = np.linalg.solve(np.polynomial.chebyshev.chebvander(xx1, npts - 1), np.eye(npts))
                P1 = np.polynomial.chebyshev.chebvander(self.xp, npts - 1) @ V0inv
                u = P1 @ u
                v = (P1 @ v.T).T
            return self._d(u), self._d(v)
        if type != 0:
            raise NotImplementedError(type)
        xx = p.L * (self.xp + 1)
        U = np.zeros((self.Np, self.Nv))
        if args.ic_file != "":
            import scipy.interpolate as si
            u1 = np.load(args.ic_file)
            npts = u1.shape[0]
            xx1 = -np.cos(np.pi * np.linspace(0, npts - 1, npts) / (npts - 1))
            for k in (self.ele_idx, self.ion_idx, self.Te_idx):
                U[:, k] = si.interp1d(xx1, u1[:, k], kind="linear", bounds_error=True)(self.xp)
            if args.ic_neTe == 1:
                U[:, self.Te_idx] = U[:, self.Te_idx] / U[:, self.ele_idx]
        else:
            prof = 1e6 * (1e7 + 1e9 * (1 - 0.5 * xx / p.L) ** 2 * (0.5 * xx / p.L) ** 2) / p.np0
            U[:, self.ele_idx] = prof
            U[:, self.ion_idx] = prof
            U[:, self.Te_idx] = args.Te
        g = self.vgrid
        vth = self.bs_vth
        Te = U[:, self.Te_idx]
        cols, cache = [], {}
        for i in range(self.Np):
            v_ratio = C_GAMMA * math.sqrt(Te[i]) / vth
            if v_ratio not in cache:
                h = self._maxwellian(v_ratio)
                cache[v_ratio] = h / torch.dot(self.op_mass, h)
            cols.append(cache[v_ratio])
        H = torch.stack(cols, dim=1).contiguous()                      # (dof_v, Np)
        V = self.ords.sph_to_ords(H) * self._d(U[:, self.ele_idx])[None, :]
        return self._d(U), V.contiguous()

    def _maxwellian(self, v_ratio):
        """l = 0 coefficients of pi^-1.5 exp(-(v / v_ratio)^2) / v_ratio^3 (M^-1 of its
        moments against the radial basis), as initialize does per point.  Only the grid's
        own temperature (v_ratio = 1) is provided."""
        if abs(v_ratio - 1.0) > 1e-12:
            raise NotImplementedError("initial temperatures other than args.Te are not provided")
        return self.vgrid.maxwellian_init()

    def _d(self, a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.dev)

    def _fname(self):
        return os.path.join(self.args.dir, self.args.fname) if not str(self.args.fname).startswith(
            str(self.args.dir)) else self.args.fname

    def ords_to_sph(self, x):
        return self.ords.ords_to_sph(x)

    def sph_to_ords(self, x):
        return self.ords.sph_to_ords(x)

    def bte_eedf_normalization(self, v_lm):
        """:1021-1030: v_lm / (mass(v_lm) / sqrt(4 pi) * 2 (vth / c_gamma)^3) per point."""
        scale = (self.op_mass / self.mm_fac) @ v_lm * (2 * (self.bs_vth / C_GAMMA) ** 3)
        return v_lm / scale

    def solve_poisson(self, ne, ni, time):
        """:1005-1019: phi with phi'' = -alpha (ni - ne), phi(-1) = sin(2 pi t), phi(1) = 0."""
        r = -self.param.alpha * (ni - ne)
        r = r.clone() if isinstance(r, torch.Tensor) else self._d(r)
        r[0] = math.sin(2 * math.pi * time)
        r[-1] = 0.0
        if not hasattr(self, "_LpD_inv_d"):
            self._LpD_inv_d = self._d(self.LpD_inv)
        return self._LpD_inv_d @ r

    def step_init(self, Uin, Vin, dt):
        """step_init (:2073-2212): the v-step preconditioner table and x-advection operators
        for dt, then bte_to_fluid and fluid_to_bte at t = 0."""
        if not hasattr(self, "_fl"):
            self._setup_fluid()
        u = (Uin if isinstance(Uin, torch.Tensor) else self._d(Uin)).to(self.dev).clone().contiguous()
        v = (Vin if isinstance(Vin, torch.Tensor) else self._d(Vin)).to(self.dev).clone().contiguous()
        self.initialize_bte_adv_x(0.5 * dt)
        p = self.param
        self._vstep = VStep(self.op_col_en, self.op_col_gT, self.op_adv_v, p.tau, p.n0, p.np0, p.Tg, dt,
                            ordinates=self.ords, rtol=self.args.rtol, atol=self.args.atol,
                            gmres_rsrt=self.args.gmres_rsrt)
        self._dt = dt
        self.bte_to_fluid(u, v, 0.0, dt)
        self.fluid_to_bte(u, v, 0.0, dt)
        return u, v

    def initialize_bte_adv_x(self, dt):
        """initialize_bte_adv_x (:718-834, upwinded backward Euler): eigenbasis of
        M^-1 (vth tau / L) [int v^3 B_p B_k] and the shift factors for a step dt."""
        scale = self.bs_vth * self.param.tau / self.param.L
        self._xadv = XAdvection(self.vgrid, self.xp, self.Nvt, scale, dt)
        self.adv_setup_dt = dt

    def bte_to_fluid(self, u, v, time, dt):
        """:1050-1076 in place: u[:, ne], u[:, Te] and the ionisation rate from v."""
        self._moments(v)
        self._fluid_launch(u, GLOW_MOMENTS)

    def fluid_to_bte(self, u, v, time, dt):
        """:1078-1086: bs_E = -(V0 / L) Dp phi for the electron step."""
        self._fluid_launch(u, GLOW_FIELD_AFTER, t_after=time)

    def step_fluid(self, u, du, time, dt, ts_type, verbose=0):
        """:1318-1372 (ts_type "BE"): a new u with ni advanced by a backward-Euler step of dt,
        the field frozen at its value from the Poisson problem at `time`."""
        if ts_type != "BE":
            raise NotImplementedError(f"fluid ts_type {ts_type!r}: the ion step of this library is backward Euler")
        u1 = u.clone()
        self._fluid_launch(u1, GLOW_STEP, h=dt, t_before=time)
        torch.cuda.current_stream().synchronize()
        self.check_fluid_status()
        return u1

    def step_bte_x(self, v, time, dt, verbose=0):
        assert self.adv_setup_dt == dt, "x-advection operators were set up for another dt"
        return self._xadv.step(v, dt)

    def step_bte_v(self, v, dv, time, dt, ts_type, verbose=0):
        if ts_type != "BE":
            raise NotImplementedError(f"ts_type {ts_type!r}: the v-step of this library is backward Euler")
        assert dt == self._dt, "the v-step was set up for another dt"
        return self._vstep.step(v.contiguous(), self.bs_E)

    def step_bte(self, u, v, du, dv, time, dt, ts_type, verbose=0):
        """:2030-2071: x half step, v step under bs_E, x half step."""
        v = self.step_bte_x(v, time, 0.5 * dt)
        v = self.step_bte_v(v, None, time, dt, self.ts_type_bte_v, verbose)
        return self.step_bte_x(v, time + 0.5 * dt, 0.5 * dt)

    def _step_inplace(self, u, v, time, dt):
        """The Strang step (:2214-2241) with u updated in place; returns v."""
        h = 0.5 * dt
        self._fluid_launch(u, GLOW_STEP | GLOW_FIELD_AFTER, h=h, t_before=time, t_after=time)
        v = self.step_bte(u, v, None, None, time, dt, None)
        self._moments(v)
        self._fluid_launch(u, GLOW_MOMENTS | GLOW_STEP, h=h, t_before=time + h)
        return v

    def step(self, u, v, du, dv, time, dt, scheme="strang-splitting", verbose=0):
        """step (:2214-2269, Strang splitting): (u, v) after one step of dt."""
        if scheme != "strang-splitting":
            raise NotImplementedError(scheme)
        if dt != self._dt:
            raise AssertionError("step_init set the operators up for another dt")
        u = u.clone()
        v = self._step_inplace(u, v.contiguous(), time, dt)
        torch.cuda.current_stream().synchronize()
        self.check_fluid_status()
        return u, v

    def solve(self, Uin, Vin, output_cycle_averaged_qois=False):
        """solve (:2271-2402): args.cycles RF periods in steps of args.cfl with the
        reference's checkpoints (every cp_cycle_freq periods, .npy files under args.dir) and
        cycle-averaged u and v / ne."""
        args = self.args
        dt = args.cfl
        steps = int(max(1, round(args.cycles / dt)))
        u, v = self.step_init(Uin, Vin, dt)
        io_freq = int(round(args.io_cycle_freq / dt))
        cycle_freq = int(round(1.0 / dt))
        cp_freq = int(round(args.cp_cycle_freq / dt))
        ts_idx_b, tt = 0, 0.0
        if args.restore == 1:
            ts_idx_b = int(args.rs_idx * io_freq)
            tt = ts_idx_b * dt
        os.makedirs(args.dir, exist_ok=True)
        fname = self._fname()
        avg_u = torch.zeros_like(u) if output_cycle_averaged_qois else None
        avg_v = torch.zeros_like(v) if output_cycle_averaged_qois else None
        u0 = u.clone()
        self.cycle_residuals = []
        for ts_idx in range(ts_idx_b, steps):
            if ts_idx % cycle_freq == 0:
                self.check_fluid_status()
                a1 = float((u[:, 0] - u0[:, 0]).abs().max())
                a2 = a1 / float(u0[:, 0].abs().max())
                self.cycle_residuals.append((ts_idx, a1, a2))
                if args.verbose:
                    print("||u(t+T) - u(t)|| = %.8E and ||u(t+T) - u(t)||/||u(t)|| = %.8E" % (a1, a2))
                u0 = u.clone()
            if ts_idx % io_freq == 0:
                idx = ts_idx // io_freq
                if ts_idx % cp_freq == 0:
                    np.save("%s_%04d_u.npy" % (fname, idx), u.cpu().numpy())
                    np.save("%s_%04d_v.npy" % (fname, idx), v.cpu().numpy())
                if output_cycle_averaged_qois:
                    if ts_idx > ts_idx_b:
                        au = avg_u * (0.5 * dt / args.io_cycle_freq)
                        av = avg_v * (0.5 * dt / args.io_cycle_freq)
                        if ts_idx % cp_freq == 0:
                            np.save("%s_%04d_u_avg.npy" % (fname, idx), au.cpu().numpy())
                            np.save("%s_%04d_v_avg.npy" % (fname, idx), av.cpu().numpy())
                        avg_u.zero_()
                        avg_v.zero_()
                    elif ts_idx % cp_freq == 0:
                        np.save("%s_%04d_u_avg.npy" % (fname, idx), u.cpu().numpy())
                        np.save("%s_%04d_v_avg.npy" % (fname, idx), v.cpu().numpy())
            if output_cycle_averaged_qois:
                avg_u += u
                avg_v += v / u[:, self.ele_idx]
            v = self._step_inplace(u, v, tt, dt)
            if output_cycle_averaged_qois:
                avg_u += u
                avg_v += v / u[:, self.ele_idx]
            tt += dt
        torch.cuda.synchronize()
        self.check_fluid_status()
        return u, v

    # ------------------------------------------------------------------ checkpoints
    def store_checkpoint(self, v, time, dt, fprefix):
        """bte_1d3v_solver.store_checkpoint (bte_1d3v_solver.py:2065-2089): time, dt, the
        distribution and its density (m^-3) and temperature (eV), in fprefix.npz."""
        ns = self.vgrid.num_sh
        vsh = self.ords_to_sph(v)
        ne = self.op_mass @ vsh
        Te = (self.op_temp @ vsh) / ne
        np.savez("%s.npz" % fprefix, **{"time[T]": np.array([time]), "dt[T]": np.array([dt]),
                                          "edf": v.cpu().numpy(), "ne[m^-3]": self.param.np0 * ne.cpu().numpy(),
                                          "Te[eV]": Te.cpu().numpy()})

    def restore_checkpoint(self, fprefix):
        """bte_1d3v_solver.restore_checkpoint (:2091-2103): (time, dt, v) from fprefix.npz."""
        with np.load("%s.npz" % fprefix) as F:
            return float(F["time[T]"][0]), float(F["dt[T]"][0]), self._d(F["edf"])
