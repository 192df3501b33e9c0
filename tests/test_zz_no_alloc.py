"""bz_steady_solve_0d_noalloc (scratch in a static device array) against bz_steady_solve_0d
(caller-owned scratch): the same kernels, so the results must be bitwise equal.  Two calls
in a row, the second with a smaller batch, and a batch beyond the static scratch rejected
with BZ_ENOSPACE before anything is launched."""
import numpy as np
import pytest
import torch

from boltzmann_amd.plan0d import VEC_ROWS
from helpers import instance_params, oracle_batched, product_grid, product_xsecs

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


@pytest.fixture(scope="module")
def solver():
    from boltzmann_amd.solver0d import Batched0D
    print("case: no-alloc steady solve, Nr=32 lmax=2", flush=True)
    ops, g, procs = oracle_batched(32, 2)
    return Batched0D(product_grid(g), product_xsecs(), growth="collisions")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _solve_both(solver, B, seed):
    from boltzmann_amd import _lib
    prm = instance_params(B, seed=seed)
    n0, Tg, E, ns = _dev(prm["n0"]), _dev(prm["Tg"]), _dev(prm["E"]), _dev(prm["ns_by_n0"])
    kw = dict(atol=0.0, rtol=1e-12, max_iter=30, alpha_min=1e-16, alpha_rho=0.5, ls_c=1e-4, mode=1)
    x_ref = solver.initial(B)
    sc = solver.solve_raw(x_ref, n0, Tg, E, ns, **kw)
    torch.cuda.synchronize()
    ref = {k: sc[k].clone() for k in ("iters", "status", "res_abs", "res0")}
    x = solver.initial(B)
    out = dict(iters=torch.full((B,), -7, dtype=torch.int32, device="cuda"),
               status=torch.full((B,), -7, dtype=torch.int32, device="cuda"),
               res_abs=torch.full((B,), -1.0, dtype=torch.float64, device="cuda"),
               res0=torch.full((B,), -1.0, dtype=torch.float64, device="cuda"))
    bt = _lib.Batch0D(B=B, n0=n0.data_ptr(), Tg=Tg.data_ptr(), E=E.data_ptr(), ns_by_n0=ns.data_ptr(),
                      x=x.data_ptr(), kprof=None, urows=None, vec=None, iters=out["iters"].data_ptr(),
                      status=out["status"].data_ptr(), res_abs=out["res_abs"].data_ptr(),
                      res0=out["res0"].data_ptr(), atol=kw["atol"], rtol=kw["rtol"], alpha_min=kw["alpha_min"],
                      alpha_rho=kw["alpha_rho"], ls_c=kw["ls_c"], max_iter=kw["max_iter"], mode=kw["mode"],
                      tg_scale=1.0, tol_abs=None)
    _lib.call("bz_steady_solve_0d_noalloc", solver.tables.struct, bt)
    torch.cuda.synchronize()
    return x_ref, ref, x, out


def test_noalloc_equals_allocating_entry_point(solver):
    for B, seed in ((24, 5), (7, 9)):          # the second call is smaller
        print(f"no-alloc call B={B}", flush=True)
        x_ref, ref, x, out = _solve_both(solver, B, seed)
        assert torch.equal(x, x_ref)
        for k in ref:
            assert torch.equal(out[k], ref[k]), k
        assert (out["status"] == 0).all()


def test_noalloc_rejects_batch_beyond_scratch(solver):
    from boltzmann_amd import _lib
    t = solver.tables
    per = (t.P + t.n * t.WP + VEC_ROWS * t.n) * 8
    B = SCRATCH_BYTES // per + 1
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    p = dummy.data_ptr()
    bt = _lib.Batch0D(B=B, n0=p, Tg=p, E=p, ns_by_n0=p, x=p, kprof=None, urows=None, vec=None, iters=p, status=p,
                      res_abs=p, res0=p, atol=0.0, rtol=1e-12, alpha_min=1e-16, alpha_rho=0.5, ls_c=1e-4,
                      max_iter=30, mode=1, tg_scale=1.0, tol_abs=None)
    rc = _lib.lib().bz_steady_solve_0d_noalloc(_lib.ptr(t.struct), _lib.ptr(bt), _lib.stream_handle())
    assert rc == -3
    assert float(dummy[0]) == 0.0
