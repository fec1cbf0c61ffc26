"""Continuous-batching request loop for HDG solves.

Requests are independent HDG solves (Darcy or elasticity, 2D or 3D, orders
1..3, Cartesian meshes of different sizes).  The loop

  * admits waiting requests in arrival order while their device footprint
    (hdg_footprint_cartesian, computed on the host before anything is
    allocated) fits the free HBM budget;
  * condenses and assembles every request admitted in one round with one
    batch-graph launch (hdg_batch_launch: one condensation launch per class of
    equal local systems), asynchronously, on one of several contexts (HIP
    streams), so rounds admitted while others are on the GPU run beside them;
  * keeps each request's mesh, dof maps and factors resident until its
    recovery (hdg_batch_recover) has run, then copies the free values back and
    retires the request.

The skeleton solve of a request is either the host sparse direct solve
(``skeleton="spsolve"``, the reference's solve) or, for workloads where that
host solve would dominate, the interpolant of the manufactured trace
(``skeleton="interpolant"``), as in the 64^3 benchmark.

A finished round is released at once: its device blocks go back to its
context's pool (libhdg's per-context block pools), which needs no device
synchronisation, so the other context's round keeps the GPU busy meanwhile.
By default each context is driven by a host thread of its own, so that one
round's host work (mesh set-up, batch creation, result copies) overlaps the
other's.
"""
from __future__ import annotations

import ctypes as C
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import FP64, MIXED, HDGError, check, lib


def _out_of_memory(err):
    return err.code == _lib.HDG_ERR_ALLOC or "out of memory" in str(err)


@dataclass
class SolveRequest:
    rid: int
    arrival: float          # seconds after the start of the run
    physics: int            # 0 Darcy, 1 elasticity
    dim: int
    order: int
    n: int                  # cells per direction
    seed: int
    admitted: float = None
    done: float = None
    cells: int = 0
    footprint: int = 0
    result: np.ndarray = None
    sweeps: float = 0.0
    fp64_fallback: bool = False
    extra: dict = field(default_factory=dict)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def draw_requests(count, rate, seed, start=0.0):
    """Seeded Poisson stream of solves with the uneven-batch distribution:
    half Darcy (2D with n in [8,128] or 3D with n in [4,32]), half 3D
    elasticity (n in [4,32]); order uniform in {1,2,3}."""
    rng = np.random.default_rng(seed)
    t = start + np.cumsum(rng.exponential(1.0 / rate, count))
    out = []
    for i in range(count):
        phys = int(rng.integers(0, 2))
        D = int(rng.integers(2, 4)) if phys == 0 else 3
        n = int(rng.integers(8, 129)) if D == 2 else int(rng.integers(4, 33))
        k = int(rng.integers(1, 4))
        out.append(SolveRequest(i, float(t[i]), phys, D, k, n, int(rng.integers(0, 2 ** 31 - 1))))
    return out


def _problem(req):
    el = req.physics == 1
    src = _lib.SOURCE_ELAS_SIN if el else _lib.SOURCE_DARCY_SINPROD
    return _lib.Problem(int(req.physics), req.dim, _lib.SPACE_P if el else _lib.SPACE_Q, req.order,
                        req.order + 1 if el else req.order, req.order, 2 * (req.order + 1),
                        float(req.n) if el else 1.0, 1.0, src)


def footprint(req, precision):
    prob = _problem(req)
    cells = (C.c_int64 * 3)(*([req.n] * req.dim + [1] * (3 - req.dim)))
    b = C.c_int64(0)
    check(lib().hdg_footprint_cartesian(C.byref(prob), cells, int(precision), C.byref(b)), None,
          "footprint")
    return int(b.value)


def build_operator(gh, ctx, req):
    """HybridAffineFEOperator of a request (not yet condensed): kappa
    lognormal(0,1) per cell (Darcy) or nu uniform in [0.2,0.4] (elasticity)
    from the request's seed."""
    D, n = req.dim, req.n
    model = gh.CartesianDiscreteModel(tuple([0.0, 1.0] * D), tuple([n] * D))
    rng = np.random.default_rng(req.seed)
    nc = n ** D
    if req.physics == 0:
        X = gh.HDGSpaces.darcy_baseline(model, req.order)
        wf = gh.DarcyHDG(tau=1.0, kappa=np.exp(rng.standard_normal(nc)),
                         source=_lib.SOURCE_DARCY_SINPROD)
    else:
        X = gh.HDGSpaces.elasticity(model, req.order)
        wf = gh.ElasticityHDG(tau=float(n), nu=rng.uniform(0.2, 0.4, nc), E=1.0,
                              source=_lib.SOURCE_ELAS_SIN)
    return gh.HybridAffineFEOperator(wf, X, X, [1, 2], [3], ctx=ctx, assemble=False)


class _Round:
    """Requests admitted together: one batch graph on one context."""

    def __init__(self, server, ctx, reqs, now):
        import torch
        gh = server.gh
        self.server, self.ctx, self.reqs = server, ctx, reqs
        self.ops, self.batch = [], None
        t0 = time.perf_counter()
        try:
            for r in reqs:
                self.ops.append(build_operator(gh, ctx, r))
            t1 = time.perf_counter()
            self.batch = gh.HybridOperatorBatch(self.ops, autotune=False)
            t2 = time.perf_counter()
            self.batch.launch(server.precision)
            t3 = time.perf_counter()
        except Exception:
            self.release()  # (the caller decides whether to retry the requests)
            raise
        for r, op in zip(reqs, self.ops):
            r.admitted = now
            r.cells = op.info.ncells
        server.host_s["operators"] += t1 - t0
        server.host_s["batch_graph"] += t2 - t1
        server.host_s["launch"] += t3 - t2
        self.ev = torch.cuda.Event()
        self.ev.record(ctx.torch_stream)
        self.stage = "condense"
        self.bytes = sum(r.footprint for r in reqs)

    def advance(self, now):
        """Moves the round on when its GPU work is done; True when finished."""
        import torch
        if not self.ev.query():
            return False
        ctx = self.ctx
        t0 = time.perf_counter()
        if self.stage == "condense":
            lams, outs = [], []
            with ctx.stream():
                for op in self.ops:
                    if self.server.skeleton == "spsolve" and op.info.nfree:
                        import scipy.sparse.linalg as spla
                        K, rhs = op.skeleton_system()
                        lam = spla.spsolve(K.tocsc(), rhs)
                    else:
                        lam = op.interpolate_trace()
                    lams.append(torch.as_tensor(lam, dtype=torch.float64, device="cuda"))
                    outs.append(torch.empty(op.info.nfree_total, dtype=torch.float64,
                                            device="cuda"))
                self.batch.recover(lams, outs)
                self.outs = outs
                self.ev = torch.cuda.Event()
                self.ev.record(ctx.torch_stream)
            self.lams = lams
            self.stage = "recover"
            self.server.host_s["skeleton_and_recover"] += time.perf_counter() - t0
            return False
        for r, op, out in zip(self.reqs, self.ops, self.outs):
            r.result = out.cpu().numpy()
            st = op.refinement_stats()
            r.sweeps = st["mean_sweeps"]
            if st["unconverged"]:
                # the mixed refinement did not converge somewhere: redo in FP64
                check(lib().hdg_condense_assemble(op.mesh, int(FP64), None), ctx.h, "fallback")
                r.result = op.recover(self.lams[self.ops.index(op)]).cpu().numpy()
                r.fp64_fallback = True
            r.done = self.server.clock()
            if not self.server.keep_results:  # the solution reached the host; keep its norm
                r.extra["norm"] = float(np.linalg.norm(r.result))
                r.result = None
        self.server.host_s["results_to_host"] += time.perf_counter() - t0
        return True

    def release(self):
        t0 = time.perf_counter()
        if self.batch is not None:
            self.batch.close()
        for op in self.ops:
            op.close()
        self.server.host_s["release"] += time.perf_counter() - t0


class Server:
    """Continuous-batching HDG solve server on one GPU.  `max_batch` bounds the
    requests of one round: on MI355X with the uneven-batch stream (meshes of
    4k to 250k cells, each enough to fill the GPU), rounds of 2 served 28.6
    requests/s against 22 for rounds of up to 64, whose larger allocations
    miss the block pools more often; a host thread per context (`threaded`)
    raised rounds of 2 from 26.4 to 34.2 requests/s."""

    def __init__(self, gh, hbm_budget=200 << 30, precision=MIXED, streams=2, max_batch=2,
                 skeleton="interpolant", batching=True, keep_results=True, threaded=True):
        self.gh = gh
        self.keep_results = keep_results
        # threaded: one host thread per context builds, launches, recovers and
        # releases its own rounds, so one context's host work (mesh set-up,
        # batch creation, result copies: libhdg and torch release the GIL)
        # overlaps the other's; otherwise one thread drives every context
        self.threaded = threaded and streams > 1
        self.ctxs = [gh.Context(0) for _ in range(streams)]
        self.budget = int(hbm_budget)
        self.precision = precision
        self.max_batch = max_batch if batching else 1
        self.skeleton = skeleton
        self.t0 = None
        # host seconds per phase of the loop (the GPU works behind most of them)
        self.host_s = dict.fromkeys(("footprint", "operators", "batch_graph", "launch",
                                     "skeleton_and_recover", "results_to_host", "release"), 0.0)

    def clock(self):
        return time.perf_counter() - self.t0

    def run(self, requests, progress=None):
        """Serves `requests` (arrival times relative to the start) and returns
        them with latencies filled in; requests whose footprint exceeds the
        whole budget are rejected (done stays None)."""
        if self.threaded:
            return self._run_threaded(requests, progress)
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        inflight = []
        used = 0
        busy = set()
        blocked = False   # an admission ran out of device memory: wait for a round
        self.memory_retries = 0
        self.max_round = self.max_batch  # halved by an admission that ran out of memory alone
        self.rejected = []
        self.rounds = []
        self.t0 = time.perf_counter()
        last_report = 0.0
        while pending or waiting or inflight:
            now = self.clock()
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                tf = time.perf_counter()
                r.footprint = footprint(r, self.precision)
                self.host_s["footprint"] += time.perf_counter() - tf
                if r.footprint > self.budget:
                    self.rejected.append(r)
                else:
                    waiting.append(r)
            progressed = False
            for rd in list(inflight):
                if rd.advance(self.clock()):
                    # a finished round gives its device memory back at once: it
                    # returns to its context's pool without a device
                    # synchronisation, while the other context keeps running
                    inflight.remove(rd)
                    busy.discard(id(rd.ctx))
                    rd.release()
                    used -= rd.bytes
                    progressed = True
                    blocked = False
            free = [c for c in self.ctxs if id(c) not in busy]
            if blocked and not inflight:
                blocked = False
            if waiting and free and not blocked:
                take, need = [], 0
                while (waiting and len(take) < min(self.max_batch, self.max_round) and
                       used + need + waiting[0].footprint <= self.budget):
                    r = waiting.popleft()
                    take.append(r)
                    need += r.footprint
                if take:
                    ctx = free[0]
                    try:
                        rd = _Round(self, ctx, take, self.clock())
                    except HDGError as err:
                        # the device is fuller than the footprints say (the
                        # runtime's own memory is not in them): give the
                        # requests back and admit again once a round is done,
                        # or, with nothing on the device, as a smaller round
                        if not _out_of_memory(err) or (len(take) == 1 and not inflight):
                            raise
                        for r in reversed(take):
                            waiting.appendleft(r)
                        if not inflight:
                            self.max_round = max(1, len(take) // 2)
                        self.memory_retries += 1
                        blocked = bool(inflight)
                        continue
                    self.rounds.append(len(take))
                    inflight.append(rd)
                    busy.add(id(ctx))
                    used += rd.bytes
                    progressed = True
            if progress and self.clock() - last_report > 5.0:
                import torch
                last_report = self.clock()
                free, total = torch.cuda.mem_get_info()
                live = pooled = 0
                for cx in self.ctxs:
                    a, b, u = C.c_int64(), C.c_int64(), C.c_int64()
                    check(lib().hdg_ctx_memory(cx.h, C.byref(a), C.byref(b), C.byref(u)),
                          cx.h, "memory")
                    live, pooled = live + a.value, pooled + b.value
                progress(f"t={last_report:.1f}s done={sum(r.done is not None for r in requests)} "
                         f"waiting={len(waiting)} inflight={len(inflight)} "
                         f"admitted={used / 2**30:.1f}GiB live={live / 2**30:.1f}GiB "
                         f"pooled={pooled / 2**30:.1f}GiB unpooled={u.value / 2**30:.1f}GiB "
                         f"free={free / 2**30:.1f}GiB "
                         f"torch={torch.cuda.memory_reserved() / 2**30:.1f}GiB")
            if not progressed:
                wait = 2e-4
                if pending and not waiting and not inflight:
                    wait = max(0.0, min(0.05, pending[0].arrival - self.clock()))
                time.sleep(wait)
        for c in self.ctxs:
            c.synchronize()
        return requests

    def _report(self, requests, progress, used, inflight, waiting):
        import torch
        free, total = torch.cuda.mem_get_info()
        live = pooled = 0
        u = C.c_int64()
        for cx in self.ctxs:
            a, b = C.c_int64(), C.c_int64()
            check(lib().hdg_ctx_memory(cx.h, C.byref(a), C.byref(b), C.byref(u)), cx.h, "memory")
            live, pooled = live + a.value, pooled + b.value
        progress(f"t={self.clock():.1f}s done={sum(r.done is not None for r in requests)} "
                 f"waiting={waiting} inflight={inflight} "
                 f"admitted={used / 2**30:.1f}GiB live={live / 2**30:.1f}GiB "
                 f"pooled={pooled / 2**30:.1f}GiB unpooled={u.value / 2**30:.1f}GiB "
                 f"free={free / 2**30:.1f}GiB "
                 f"torch={torch.cuda.memory_reserved() / 2**30:.1f}GiB")

    def _run_threaded(self, requests, progress=None):
        """run() with one worker thread per context.  Admission (arrival order,
        footprint budget, round size) is shared under a lock; each worker
        admits a round, builds and launches it on its context, waits for its
        GPU work, recovers, copies the results and releases the round."""
        import threading
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        st = dict(used=0, inflight=0, blocked=False)
        lock = threading.Lock()
        errors = []
        self.memory_retries = 0
        self.max_round = self.max_batch
        self.rejected = []
        self.rounds = []
        self.t0 = time.perf_counter()

        def arrivals():  # (under the lock)
            now = self.clock()
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                tf = time.perf_counter()
                r.footprint = footprint(r, self.precision)
                self.host_s["footprint"] += time.perf_counter() - tf
                (self.rejected if r.footprint > self.budget else waiting).append(r)

        def worker(ctx):
            try:
                while True:
                    take, need, wait = [], 0, 2e-4
                    with lock:
                        if errors:
                            return
                        arrivals()
                        if not pending and not waiting:
                            return
                        if not st["blocked"]:
                            while (waiting and len(take) < min(self.max_batch, self.max_round)
                                   and st["used"] + need + waiting[0].footprint <= self.budget):
                                r = waiting.popleft()
                                take.append(r)
                                need += r.footprint
                        if take:
                            st["used"] += need
                            st["inflight"] += 1
                        elif pending and not waiting:
                            wait = max(0.0, min(0.05, pending[0].arrival - self.clock()))
                    if not take:
                        time.sleep(wait)
                        continue
                    try:
                        rd = _Round(self, ctx, take, self.clock())
                    except HDGError as err:
                        with lock:
                            st["used"] -= need
                            st["inflight"] -= 1
                            if not _out_of_memory(err) or (len(take) == 1 and not st["inflight"]):
                                raise
                            # give the requests back; admit again once a round
                            # is done, or, with nothing in flight, as a smaller round
                            for r in reversed(take):
                                waiting.appendleft(r)
                            if not st["inflight"]:
                                self.max_round = max(1, len(take) // 2)
                            self.memory_retries += 1
                            st["blocked"] = st["inflight"] > 0
                        continue
                    with lock:
                        self.rounds.append(len(take))
                    while not rd.advance(self.clock()):
                        time.sleep(1e-4)
                    rd.release()
                    with lock:
                        st["used"] -= rd.bytes
                        st["inflight"] -= 1
                        st["blocked"] = False
            except BaseException as e:  # handed to the caller's thread
                with lock:
                    errors.append(e)

        threads = [threading.Thread(target=worker, args=(c,), daemon=True) for c in self.ctxs]
        for t in threads:
            t.start()
        last_report = 0.0
        while any(t.is_alive() for t in threads):
            for t in threads:
                t.join(timeout=0.05)
            if progress and self.clock() - last_report > 5.0:
                last_report = self.clock()
                with lock:
                    used, inflight, nwait = st["used"], st["inflight"], len(waiting)
                self._report(requests, progress, used, inflight, nwait)
        for c in self.ctxs:
            c.synchronize()
        if errors:
            raise errors[0]
        return requests

    def close(self):
        for c in self.ctxs:
            c.close()


def summarize(requests, wall=None):
    """Latency percentiles, throughput (cells/s and requests/s over the span
    from the first arrival to the last completion, or `wall`)."""
    lat = np.array([r.latency for r in requests if r.latency is not None])
    cells = sum(r.cells for r in requests if r.done is not None)
    done = [r for r in requests if r.done is not None]
    span = wall if wall is not None else (max(r.done for r in done) - min(r.arrival for r in done))
    return dict(requests=len(done), rejected=len(requests) - len(done),
                requests_per_s=len(done) / span if span > 0 else None,
                p50_ms=float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
                p99_ms=float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
                mean_ms=float(lat.mean() * 1e3) if len(lat) else None,
                cells=int(cells), cells_per_s=cells / span if span > 0 else None,
                span_s=span,
                fp64_fallbacks=sum(r.fp64_fallback for r in done))
