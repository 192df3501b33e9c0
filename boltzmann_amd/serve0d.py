"""Serving loop for mixed 0D EDF solve requests on one MI355X: continuous batching with
admission by an HBM budget.

Requests differ in grid (Nr, lmax), number of collision processes, E/N, gas parameters and
e-e (ionisation degree > 0). Each distinct grid signature gets its operators and solver
tables once (assembly on the device; the reference assembles per `bte_0d3v_batched` grid,
bte_0d3v_batched.py:310-520), cached for later requests. One host thread runs the loop:

* arrived requests go into a queue per grid signature;
* a batch is launched for a grid whose queue is not empty when the batch's device memory
  (solver scratch, state, parameters, QoIs; for e-e also the Newton-Krylov work arrays) fits
  in what is left of the budget. Each batch runs on its own HIP stream (default priority), so
  batches of different grids, and batches of later arrivals, overlap on the device with the
  ones already running;
* a batch whose completion event has fired is retired: its QoIs are copied to the host, its
  memory is returned to the budget, and every request in it gets its latency (arrival to
  QoIs on the host).

Non-e-e batches are asynchronous (one fused solver launch and one QoI launch). e-e batches
run the host-orchestrated Newton-Krylov solve (bte_ee.EEOperator.steady_state) after the
warm start without e-e, as the reference driver does (bte_0d3v_batched_driver.py:192-204);
they are started after the asynchronous launches of the same pass, so those keep the device
busy meanwhile.

Instances do not interact inside a batch (every kernel works on one instance per wavefront,
or elementwise), so a result does not depend on which other requests share its batch.
"""
import concurrent.futures
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .cross_section import xsections_from_file
from .solver0d import Batched0D
from .vspace import C_GAMMA, VGrid, make_knots, vth_from_Te


@dataclass
class Request:
    rid: int
    Nr: int
    lmax: int
    nproc: int                      # the first nproc processes of the cross-section file
    EbyN: float                     # Td
    ion_deg: float = 0.0            # > 0: e-e collisions with ne = ion_deg * n0
    n0: float = 3.22e22
    Tg: float = 300.0
    fac: tuple = (1.0, 1.0, 1.0)    # per-process density factors (ns_by_n0)
    arrival: float = 0.0            # seconds from the start of the run


@dataclass
class Result:
    rid: int
    energy: float
    mobility: float
    diffusion: float
    rates: np.ndarray
    status: int
    latency: float
    admitted: float
    done: float
    batch: int = field(default=-1)


class _Grid:
    """Operators and solver tables of one grid signature (Nr, lmax, nproc)."""

    def __init__(self, key, xs_all, Te, ev_max, q, dev):
        Nr, lmax, nproc = key
        xs = xs_all[:nproc]
        vth = vth_from_Te(Te)
        kd = (0.0, math.sqrt(ev_max) * C_GAMMA / vth)
        thr = sorted(set(x.threshold for x in xs if x.threshold > 0))
        t = make_knots(kd, 3, Nr + 1, np.sqrt(np.array(thr)) * C_GAMMA / vth if thr else None, extend_domain=True)
        self.g = VGrid(t, Nr + 1, lmax, q, vth, device=dev)
        self.solver = Batched0D(self.g, xs, growth="collisions")
        self.nproc = nproc
        self._ee = None

    def ee(self):
        """e-e solver (tables with the e-e pattern) and operator, built on first use."""
        if self._ee is None:
            from .bte_ee import EEOperator
            s = Batched0D(self.g, self.solver.xsecs, growth="collisions", ee_pattern=True)
            self._ee = (s, EEOperator(self.g, s))
        return self._ee

    def bytes_per_instance(self, ee, restart):
        t = self.solver.tables
        n = t.n
        b = t.scratch_bytes(1) + 4 * n * 8 + 64 * 8
        if ee:
            te = self.ee()[0].tables
            b += te.scratch_bytes(1)                                  # e-e solver scratch
            b += ((restart + 1) * n + 16 * n + te.P + restart * restart + 4 * restart) * 8
        return b


class ServingEngine:
    def __init__(self, xsec_file, hbm_budget, Te=4.0, ev_max=80.0, q=5, max_batch=4096, rtol=1e-14,
                 max_iter=30, ee_rtol=1e-15, ee_max_iter=60, restart=30, streams=4, ee_min_batch=256,
                 ee_max_wait=0.05, floor_rel=1e-11, ee_workers=4, device=None):
        self.dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.xs = xsections_from_file(xsec_file, device=self.dev)
        self.Te, self.ev_max, self.q = Te, ev_max, q
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.rtol, self.max_iter, self.floor_rel = rtol, max_iter, floor_rel
        self.ee_rtol, self.ee_max_iter, self.restart = ee_rtol, ee_max_iter, restart
        self.grids = {}
        self.stats = {}
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(max(1, int(streams)))]
        # e-e batches (host-orchestrated Newton-Krylov) run in worker threads, so the loop keeps
        # admitting and retiring the asynchronous batches meanwhile; up to ee_workers of them at
        # once, of different grid signatures (a grid's e-e work space serves one batch at a
        # time), each on an e-e stream of its own (default priority); they wait in their queue
        # for ee_min_batch requests or ee_max_wait seconds
        self.ee_workers = max(1, int(ee_workers))
        self._ee_pool = concurrent.futures.ThreadPoolExecutor(max_workers=self.ee_workers)
        self.ee_streams = [torch.cuda.Stream(device=self.dev) for _ in range(self.ee_workers)]
        self.ee_min_batch, self.ee_max_wait = int(ee_min_batch), float(ee_max_wait)

    def grid(self, key):
        if key not in self.grids:
            if key[2] < 1 or key[2] > len(self.xs):
                raise ValueError(f"nproc {key[2]} outside 1..{len(self.xs)}")
            self.grids[key] = _Grid(key, self.xs, self.Te, self.ev_max, self.q, self.dev)
        return self.grids[key]

    # ------------------------------------------------------------------ one batch
    def _params(self, reqs, S):
        f = lambda a: torch.tensor(a, dtype=torch.float64, device=self.dev)
        n0 = f([r.n0 for r in reqs])
        Tg = f([r.Tg for r in reqs])
        E = f([r.EbyN * r.n0 * 1e-21 for r in reqs])
        ns = f([[r.fac[s] for r in reqs] for s in range(S)])
        ne = f([r.ion_deg * r.n0 for r in reqs])
        return n0, Tg, E, ns, ne

    def _solve(self, key, reqs, ee, stream):
        gr = self.grid(key)
        s = gr.solver
        B = len(reqs)
        with torch.cuda.stream(stream):
            n0, Tg, E, ns, ne = self._params(reqs, s.S)
            x = s.initial(B)
            sc = s.new_scratch(B)
            s.solve_raw(x, n0, Tg, E, ns, atol=0.0, rtol=self.rtol, max_iter=self.max_iter, mode=1, scratch=sc)
            # an instance that ran out of iterations or stalled (no decrease of the true residual)
            # with |r| <= floor_rel |r0| sits at the FP64 floor of its residual: reported as
            # status 3, as the e-e solve does
            at_floor = ((sc["status"] == 1) | (sc["status"] == 2)) & (sc["res_abs"] <= self.floor_rel * sc["res0"])
            status = torch.where(at_floor, 3, sc["status"]).to(torch.int32)
            if ee:
                es, eop = gr.ee()
                info = eop.steady_state(x, n0, Tg, E, ns, ne, 0.0, self.ee_rtol, self.ee_max_iter,
                                        restart=self.restart)
                status = info["status"]
            _, q = s.qois(x, n0, E)
            ev = torch.cuda.Event()
            ev.record(stream)
        return {"ev": ev, "q": q, "status": status, "keep": (x, sc, n0, Tg, E, ns, ne)}

    def _launch(self, key, reqs, ee, stream):
        if not ee:
            return self._solve(key, reqs, False, stream)

        def work():
            out = self._solve(key, reqs, True, stream)
            out["ev"].synchronize()
            return {k: v.cpu().numpy() for k, v in out["q"].items()}, out["status"].cpu().numpy()
        return {"future": self._ee_pool.submit(work)}

    # ------------------------------------------------------------------ the loop
    def run(self, requests, realtime=True, poll=2e-4):
        """Serve the requests (arrival times in seconds from now); returns the results in
        request order and a statistics dict."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:                                  # assembly is not part of a request's latency
            g = self.grid((r.Nr, r.lmax, r.nproc))
            if r.ion_deg > 0:
                g.ee()
            if len(r.fac) < r.nproc:
                raise ValueError(f"request {r.rid}: {r.nproc} processes need {r.nproc} density factors")
        for key, g in self.grids.items():
            for ee in (False, True):
                if g.bytes_per_instance(ee, self.restart) > self.budget:
                    if any((r.Nr, r.lmax, r.nproc) == key and (r.ion_deg > 0) == ee for r in reqs):
                        raise ValueError(f"one instance of grid {key} (e-e {ee}) exceeds the HBM budget")
        torch.cuda.synchronize()
        queues = {}
        inflight = []
        results = {}
        used = 0
        peak = 0
        nb = 0
        i = 0
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        sim_now = 0.0
        while i < len(reqs) or any(queues.values()) or inflight:
            now = clock() if realtime else sim_now
            if not realtime and not inflight and not any(queues.values()) and i < len(reqs):
                sim_now = max(sim_now, reqs[i].arrival)
                now = sim_now
            while i < len(reqs) and reqs[i].arrival <= now:
                r = reqs[i]
                queues.setdefault((r.Nr, r.lmax, r.nproc, r.ion_deg > 0), []).append(r)
                i += 1
            progressed = False
            # retire finished batches
            for bt in list(inflight):
                if "future" in bt:
                    if not bt["future"].done():
                        continue
                    q, st = bt["future"].result()
                elif bt["ev"].query():
                    q = {k: v.cpu().numpy() for k, v in bt["q"].items()}
                    st = bt["status"].cpu().numpy()
                else:
                    continue
                done = time.perf_counter() - t0
                for j, r in enumerate(bt["reqs"]):
                    results[r.rid] = Result(r.rid, float(q["energy"][j]), float(q["mobility"][j]),
                                            float(q["diffusion"][j]), q["rates"][:, j].copy(), int(st[j]),
                                            done - r.arrival, bt["t"], done, bt["id"])
                used -= bt["bytes"]
                inflight.remove(bt)
                progressed = True
            # admit: asynchronous (no e-e) batches first, then e-e batches
            for want_ee in (False, True):
                for qk in sorted(queues, key=lambda k: queues[k][0].arrival if queues[k] else 1e300):
                    qq = queues[qk]
                    if not qq or qk[3] != want_ee:
                        continue
                    if want_ee:
                        busy = [b["key"] for b in inflight if "future" in b]
                        if len(busy) >= self.ee_workers or qk[:3] in busy:
                            continue                           # all e-e workers busy, or this grid's
                        if (len(qq) < self.ee_min_batch and now - qq[0].arrival < self.ee_max_wait
                                and i < len(reqs)):
                            continue                           # let the batch grow
                    gr = self.grids[qk[:3]]
                    per = gr.bytes_per_instance(qk[3], self.restart)
                    nfit = min(len(qq), self.max_batch, (self.budget - used) // per)
                    if nfit < 1:
                        continue
                    take, queues[qk] = qq[:nfit], qq[nfit:]
                    if want_ee:
                        used_s = {id(b["stream"]) for b in inflight if "future" in b}
                        stream = next(st_ for st_ in self.ee_streams if id(st_) not in used_s)
                    else:
                        stream = self.streams[nb % len(self.streams)]
                    bt = self._launch(qk[:3], take, qk[3], stream)
                    bt.update(reqs=take, bytes=nfit * per, t=time.perf_counter() - t0, id=nb,
                              stream=stream, key=qk[:3])
                    nb += 1
                    used += nfit * per
                    peak = max(peak, used)
                    inflight.append(bt)
                    progressed = True
            if not progressed:
                if realtime or any("future" in b for b in inflight):
                    time.sleep(poll)
                elif inflight:
                    torch.cuda.synchronize()
                else:
                    # offline and idle: move the simulated clock to the next event, the next
                    # arrival or the end of the wait of the oldest e-e queue still growing
                    events = [reqs[i].arrival] if i < len(reqs) else []
                    events += [qq[0].arrival + self.ee_max_wait for qk, qq in queues.items() if qq and qk[3]]
                    if not events:
                        raise RuntimeError("serving loop: queued requests but nothing can be admitted")
                    sim_now = max(sim_now, min(events))
        makespan = time.perf_counter() - t0
        lat = np.array([results[r.rid].latency for r in reqs])
        ok = np.array([results[r.rid].status in (0, 3) for r in reqs])
        self.stats = {"requests": len(reqs), "batches": nb, "makespan_s": makespan,
                      "solves_per_s": float(ok.sum()) / makespan if makespan > 0 else float("nan"),
                      "converged": int(ok.sum()), "p50_latency_s": float(np.percentile(lat, 50)),
                      "p99_latency_s": float(np.percentile(lat, 99)), "peak_bytes": int(peak),
                      "budget_bytes": self.budget, "grids": len(self.grids)}
        return [results[r.rid] for r in sorted(requests, key=lambda r: r.rid)], self.stats
