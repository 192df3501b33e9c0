"""Per-grid tables of the fused 0D Newton kernel (bz_steady_solve_0d).

Given the dense device operators of one grid (original radial-major order)
  ops = [C_0 .. C_{S-1}, C_Tg, A_E, M]
this builds, in the permuted order i' = (np_r-1-p)*ns + l (radial index from
high to low energy):
  * the row profiles of the union sparsity pattern (first column .. i'+bw),
  * the operator values gathered into profile storage (on the device),
  * the lane-slot schedule of the factorisation,
  * the mass-matrix Cholesky rows, u, uhat, M uhat and the growth vectors
    W_m = Op_m^T M^-1 u.
Only the sparsity pattern and the schedule are computed on the host (integer
planning work on an n x n boolean mask); every floating-point table is produced
by the HIP kernels.
"""
import heapq
import math

import numpy as np
import torch

from . import _lib

PREFETCH_WINDOWS = 2   # rows enter their slot 2W steps before their first column
VEC_ROWS = 12          # n-vectors of per-instance work space (bz_batch0d_t.vec)


def permutation(npr, ns):
    """perm[i'] = original index of permuted position i'."""
    q = np.arange(npr * ns) // ns
    l = np.arange(npr * ns) % ns
    return (npr - 1 - q) * ns + l


def plan_structure(mask_orig, npr, ns, window_fn):
    """Profile and slot schedule from the boolean union pattern (original order)."""
    n = npr * ns
    perm = permutation(npr, ns)
    P = mask_orig[np.ix_(perm, perm)]
    np.fill_diagonal(P, True)
    rows, cols = np.nonzero(P)
    bw = int(max(0, (cols - rows).max()))
    first = np.array([int(np.nonzero(P[i])[0].min()) for i in range(n)], dtype=np.int64)
    first = np.minimum(first, np.arange(n))
    last = np.minimum(np.arange(n) + bw, n - 1)
    lens = last - first + 1
    prof_off = np.zeros(n + 1, dtype=np.int64)
    prof_off[1:] = np.cumsum(lens)
    W = window_fn(bw)
    if W < 0:
        raise ValueError(f"upper bandwidth {bw} not supported by the 0D solver")
    # interval scheduling: row i occupies a slot from step first-2W to its pivot step i;
    # a slot freed at step i_prev can take a row with first - 2W >= i_prev.  Rows are
    # taken in order of their start and get the lowest free slot id, so slot ids >= 64
    # (register level 1+, slot s -> lane s % 64, level s // 64) are used only where
    # more than 64 rows are alive at once.
    start = first - PREFETCH_WINDOWS * W
    order = np.argsort(start, kind="stable")
    alive = []         # heap of (pivot_step, slot) of rows holding a slot
    free_ids = []      # heap of free slot ids
    nslots_used = 0
    slot_of = np.full(n, -1, dtype=np.int64)
    slot_rows = []
    for i in order:
        s0 = int(start[i])
        while alive and alive[0][0] <= s0:
            _, s_rel = heapq.heappop(alive)
            heapq.heappush(free_ids, s_rel)
        if free_ids:
            s = heapq.heappop(free_ids)
        else:
            s = nslots_used
            nslots_used += 1
            slot_rows.append([])
        slot_of[i] = s
        slot_rows[s].append(int(i))
        heapq.heappush(alive, (int(i), s))
    R = max(1, math.ceil(nslots_used / 64))
    nslots = 64 * R
    slot_first = np.full(nslots, -1, dtype=np.int64)
    sched_next = np.full(n, -1, dtype=np.int64)
    for s, rs in enumerate(slot_rows):
        slot_first[s] = rs[0]
        for a, b in zip(rs[:-1], rs[1:]):
            sched_next[a] = b
    # register levels each W-step chunk must run: a row of level r needs its level
    # executed from max(0, start) (first prefetch of its first column; earlier columns
    # are zero and its window was zeroed at entry or loaded at step 0) to its pivot
    # step, where it retires and its successor enters
    nchunk = (n + W - 1) // W
    levels = np.ones(nchunk, dtype=np.int64)
    for s, rs in enumerate(slot_rows):
        lev = s // 64 + 1
        if lev == 1:
            continue
        for i in rs:
            a = max(0, int(start[i]))
            levels[a // W: i // W + 1] = np.maximum(levels[a // W: i // W + 1], lev)
    # invariants the kernel relies on
    for s, rs in enumerate(slot_rows):
        for a, b in zip(rs[:-1], rs[1:]):
            assert a <= start[b], "slot overlap"
    assert (first <= np.arange(n)).all()
    # U stays in the band: no entry of the pattern above the band (no pivoting)
    assert ((cols - rows) <= bw).all()
    # (row << 16) | column of every profile entry, for the flat matvec
    ent_rc = np.concatenate([(np.uint32(i) << np.uint32(16)) | np.arange(first[i], last[i] + 1, dtype=np.uint32)
                             for i in range(n)]).astype(np.uint32)
    assert n < 32768
    return {
        "n": n, "bw": bw, "W": W, "R": R, "nslots": nslots, "perm": perm,
        "first_col": first, "prof_off": prof_off, "P": int(prof_off[-1]),
        "slot_first": slot_first, "sched_next": sched_next, "active_max": nslots_used,
        "chunk_levels": levels, "ent_rc": ent_rc,
    }


class SolverTables:
    """Device tables of one grid for bz_steady_solve_0d."""

    def __init__(self, vgrid, ops, growth_mask, u_orig, mass_radial, extra_mask=None, growth_ops=None):
        """vgrid: VGrid; ops: list of dense (n,n) device tensors [C_0..C_{S-1}, C_Tg, A_E, M];
        growth_mask[m] (m < nop-1): operator m contributes to the growth rate mu = w.x;
        u_orig: mass vector u (original order, device); mass_radial: radial mass block;
        extra_mask: (n, n) bool pattern of per-instance operators added later (e-e);
        growth_ops: operators (parallel to ops[:-1]) whose W_m = G_m^T M^-1 u form the growth
        vectors (default: ops themselves)."""
        dev = vgrid.dev
        self.dev = dev
        npr, ns, n = vgrid.npr, vgrid.num_sh, vgrid.n
        self.n, self.ns, self.npr = n, ns, npr
        self.nop = len(ops)
        self.S = self.nop - 3
        mask = torch.zeros((n, n), dtype=torch.bool, device=dev)
        for A in ops:
            mask |= (A != 0)
        if extra_mask is not None:
            mask |= extra_mask.to(dev)
        mask_np = mask.cpu().numpy()
        L = _lib.lib()
        st = plan_structure(mask_np, npr, ns, lambda bw: L.bz_solver0d_window(bw))
        self.st = st
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)
        self.first_col = i32(st["first_col"])
        self.prof_off = i32(st["prof_off"])
        self.slot_first = i32(st["slot_first"])
        self.sched_next = i32(st["sched_next"])
        self.chunk_levels = i32(st["chunk_levels"])
        self.ent_rc = torch.from_numpy(st["ent_rc"].view(np.int32).copy()).to(dev)
        P = st["P"]
        self.P = P
        self.opprof = torch.empty((self.nop, P), dtype=torch.float64, device=dev)
        for m, A in enumerate(ops):
            _lib.call("bz_gather_profile", A.contiguous(), n, ns, npr, self.first_col, self.prof_off, 1.0,
                      self.opprof[m])
        # mass Cholesky rows in the permuted order (reversed radial block)
        mrev = torch.flip(mass_radial, dims=(0, 1)).contiguous()
        lrev = torch.empty((npr, 4), dtype=torch.float64, device=dev)
        _lib.call("bz_band_cholesky", mrev, npr, 3, lrev)
        self.mchol = torch.empty((n, 4), dtype=torch.float64, device=dev)
        _lib.call("bz_mchol_perm", lrev, ns, npr, self.mchol)
        # u, uhat, M uhat, growth vectors
        Mdense = ops[-1]
        self.u = torch.empty(n, dtype=torch.float64, device=dev)
        _lib.call("bz_permute", u_orig.contiguous(), n, ns, npr, 0, self.u)
        uu = torch.empty(1, dtype=torch.float64, device=dev)
        _lib.call("bz_gemv", u_orig.contiguous(), 1, n, 0, u_orig.contiguous(), 1.0, uu)
        unorm = float(uu.item()) ** 0.5
        uhat_orig = torch.empty(n, dtype=torch.float64, device=dev)
        _lib.call("bz_gemv", u_orig.reshape(n, 1).contiguous(), n, 1, 0,
                  torch.ones(1, dtype=torch.float64, device=dev), 1.0 / unorm, uhat_orig)
        self.uhat = torch.empty(n, dtype=torch.float64, device=dev)
        _lib.call("bz_permute", uhat_orig, n, ns, npr, 0, self.uhat)
        muh = torch.empty(n, dtype=torch.float64, device=dev)
        _lib.call("bz_gemv", Mdense.contiguous(), n, n, 0, uhat_orig, 1.0, muh)
        self.muhat = torch.empty(n, dtype=torch.float64, device=dev)
        _lib.call("bz_permute", muh, n, ns, npr, 0, self.muhat)
        # v = M^-1 u: the radial block solved (refined) for the ns columns of u viewed as (npr, ns)
        lb = torch.empty((npr, 4), dtype=torch.float64, device=dev)
        _lib.call("bz_band_cholesky", mass_radial, npr, 3, lb)
        v = torch.empty(n, dtype=torch.float64, device=dev)
        work = torch.empty(2 * n, dtype=torch.float64, device=dev)
        _lib.call("bz_band_chol_solve_refined", mass_radial, lb, npr, 3, u_orig.contiguous(), ns, v, work, 2)
        self.wsp = torch.zeros((self.nop - 1, n), dtype=torch.float64, device=dev)
        tmp = torch.empty(n, dtype=torch.float64, device=dev)
        for m in range(self.nop - 1):
            if growth_mask[m]:
                G = ops[m] if growth_ops is None else growth_ops[m]
                _lib.call("bz_gemv", G.contiguous(), n, n, 1, v, 1.0, tmp)   # Op^T v
                _lib.call("bz_permute", tmp, n, ns, npr, 0, self.wsp[m])
        self.W = st["W"]
        self.WP = (self.W + 1) & ~1
        self.struct = _lib.Grid0D(
            n=n, ns=ns, npr=npr, bw=st["bw"], S=self.S, nop=self.nop, P=P, nslots=st["nslots"],
            first_col=self.first_col.data_ptr(), prof_off=self.prof_off.data_ptr(),
            opprof=self.opprof.data_ptr(), slot_first=self.slot_first.data_ptr(),
            sched_next=self.sched_next.data_ptr(), muhat=self.muhat.data_ptr(), uhat=self.uhat.data_ptr(),
            u=self.u.data_ptr(), wsp=self.wsp.data_ptr(), mchol=self.mchol.data_ptr(),
            chunk_levels=self.chunk_levels.data_ptr(), ent_rc=self.ent_rc.data_ptr(),
            max_row=int(np.diff(st["prof_off"]).max()))

    def scratch_bytes(self, B):
        return B * (self.P + self.n * self.WP + VEC_ROWS * self.n) * 8


class CPlannedGrid:
    """The tables of one grid planned by the C entry point bz_plan_grid0d (the path of a
    caller without this Python layer): same inputs as SolverTables, device arrays owned by
    the library and released with bz_free_grid0d when this object goes away."""

    def __init__(self, vgrid, ops, growth_ops, u_orig, mass_radial, ee_pattern=False):
        import ctypes
        self._keep = [o.contiguous() for o in ops] + [g.contiguous() for g in growth_ops if g is not None]
        arr = (ctypes.c_void_p * len(ops))(*[o.data_ptr() for o in ops])
        garr = (ctypes.c_void_p * (len(ops) - 1))(*[(g.data_ptr() if g is not None else None) for g in growth_ops])
        self.struct = _lib.Grid0D()
        self._handle = ctypes.c_void_p()
        _lib.call("bz_plan_grid0d", vgrid.npr, vgrid.num_sh, len(ops), ctypes.cast(arr, ctypes.c_void_p),
                  ctypes.cast(garr, ctypes.c_void_p), u_orig.contiguous(), mass_radial.contiguous(), int(ee_pattern),
                  self.struct, ctypes.cast(ctypes.pointer(self._handle), ctypes.c_void_p))
        self.P = self.struct.P
        self.n = self.struct.n
        W = _lib.lib().bz_solver0d_window(self.struct.bw)
        self.W = W
        self.WP = (W + 1) & ~1

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            _lib.lib().bz_free_grid0d(h)
            self._handle = None
