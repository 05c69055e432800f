"""Request serving over HBM-resident factors with continuous batching.

A `SolverService` admits two kinds of requests for sparse SPD (or symmetric
indefinite, static LDL^T) matrices identified by a key:

* factor(key, A, options): symbolic analysis on a host worker, then admission
  once the factor's HBM footprint fits the budget (idle resident factors are
  evicted least-recently-used first), then DistributeMatrix + Factorize;
* solve(key, B): right-hand sides for a resident factor.  Solves queued for
  the same factor are merged into one multi-RHS solve (up to
  `max_batch_cols` columns) and split afterwards.

`slots` worker threads pick work as it becomes runnable: a request that
arrives while others run on the GPU is started as soon as a slot and the
memory are free.  `solve_slots` of them (default 1 when there are several)
take only solves, so a burst of large factorizations cannot hold every slot
while short solves on resident factors wait.  Every factor owns its own HIP stream, so the work of
different requests overlaps on the device.  One request runs on a factor at
a time; the results do not depend on what else is in flight.

Latency of every request (arrival to completion) is recorded; `report()`
returns p50/p99 per kind, solves/s and right-hand sides/s.

symPACK has no serving layer: this module drives the symPACKMatrix2D
interface (symPACKMatrix2D.hpp:2612-2757) the way a resident-factor service
would.
"""
import collections
import threading
import time
from concurrent.futures import Future

import numpy as np

from .matrix import DistSparseMatrix, symPACKMatrix2D, symPACKOptions


class _Entry:
    __slots__ = ("key", "A", "opts", "M", "state", "bytes", "last_used", "pending", "factor_futs",
                 "error", "resident")

    def __init__(self, key, A, opts):
        self.key = key
        self.A = A
        self.opts = opts
        self.M = None
        self.state = "new"          # new -> analyzing -> waiting -> factoring -> idle <-> solving
        self.bytes = 0
        self.last_used = 0.0
        self.pending = collections.deque()   # (B (n x k), future, t_arrival)
        self.factor_futs = []                # (future, t_arrival)
        self.error = None
        self.resident = False


class SolverService:
    def __init__(self, hbm_budget_bytes=64 << 30, slots=4, max_batch_cols=64,
                 matrix_factory=symPACKMatrix2D, solve_slots=None):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        if solve_slots is None:
            solve_slots = 1 if slots > 1 else 0
        if not 0 <= solve_slots < slots:
            raise ValueError("solve_slots must leave at least one slot for factorizations")
        self.budget = int(hbm_budget_bytes)
        self.max_batch_cols = int(max_batch_cols)
        self._factory = matrix_factory
        self._cv = threading.Condition()
        self._entries = {}
        self._analyze_q = collections.deque()
        self._admit_q = collections.deque()
        self._solve_q = collections.deque()
        self._resident_bytes = 0
        self._closing = False
        self.records = []   # (kind, key, n, ncols, t_arrival, t_start, t_done)
        self.evictions = 0
        self.batched_solves = 0
        self._workers = [threading.Thread(target=self._work, args=(i < solve_slots,), daemon=True,
                                          name="spk-slot-%d" % i)
                         for i in range(slots)]
        for w in self._workers:
            w.start()

    # ---------------------------------------------------------------- API
    def submit_factor(self, key, A, options=None):
        if not isinstance(A, DistSparseMatrix):
            raise TypeError("A must be a DistSparseMatrix")
        fut = Future()
        with self._cv:
            if self._closing:
                raise RuntimeError("service is closed")
            if key in self._entries:
                raise KeyError("key %r is already registered" % (key,))
            e = _Entry(key, A, options if options is not None else symPACKOptions())
            e.factor_futs.append((fut, time.perf_counter()))
            self._entries[key] = e
            self._analyze_q.append(e)
            self._cv.notify_all()
        return fut

    def submit_solve(self, key, B):
        fut = Future()
        t = time.perf_counter()
        with self._cv:
            e = self._entries.get(key)
            if e is None:
                raise KeyError("no factor registered under %r" % (key,))
            Bm = np.asarray(B, np.float64)
            Bm = Bm.reshape(e.A.size, -1)
            if e.error is not None:
                fut.set_exception(e.error)
                return fut
            e.pending.append((Bm, fut, t))
            if e.state == "idle" and e.resident:
                self._solve_q.append(e)
            elif e.state == "idle" and not e.resident:
                e.state = "waiting"        # evicted: factor again before solving
                self._admit_q.append(e)
            self._cv.notify_all()
        return fut

    def release(self, key):
        """Drop a factor once its queued work has finished."""
        with self._cv:
            e = self._entries.get(key)
            if e is None:
                return
            while e.state in ("analyzing", "factoring", "solving", "waiting", "new") or e.pending:
                self._cv.wait()
            self._evict(e)
            del self._entries[key]
            self._cv.notify_all()

    def close(self):
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        for w in self._workers:
            w.join()
        with self._cv:
            for e in list(self._entries.values()):
                self._evict(e)
            self._entries.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def drain(self):
        """Wait until every submitted request has completed."""
        with self._cv:
            while any(e.state not in ("idle", "failed") or e.pending for e in self._entries.values()):
                self._cv.wait()

    def report(self):
        out = {}
        for kind in ("factor", "solve"):
            lat = np.array([(r[6] - r[4]) * 1e3 for r in self.records if r[0] == kind])
            if lat.size:
                out[kind] = {"count": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
                             "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean())}
        if self.records:
            t0 = min(r[4] for r in self.records)
            t1 = max(r[6] for r in self.records)
            span = max(t1 - t0, 1e-9)
            ns = sum(1 for r in self.records if r[0] == "solve")
            cols = sum(r[3] for r in self.records if r[0] == "solve")
            out["makespan_s"] = span
            out["solves_per_s"] = ns / span
            out["rhs_per_s"] = cols / span
        out["evictions"] = self.evictions
        out["merged_solve_batches"] = self.batched_solves
        return out

    # ----------------------------------------------------------- internals
    def _evict(self, e):
        if e.resident:
            e.M.ReleaseDevice()
            e.resident = False
            self._resident_bytes -= e.bytes

    def _make_room(self, need):
        """Evict idle resident factors without queued solves, LRU first."""
        if self._resident_bytes + need <= self.budget:
            return True
        victims = sorted((x for x in self._entries.values()
                          if x.resident and x.state == "idle" and not x.pending),
                         key=lambda x: x.last_used)
        for v in victims:
            if self._resident_bytes + need <= self.budget:
                break
            self._evict(v)
            self.evictions += 1
        return self._resident_bytes + need <= self.budget

    def _next_job(self, solves_only=False):
        # 1. solves on resident factors (short, latency-critical)
        while self._solve_q:
            e = self._solve_q.popleft()
            if e.state == "idle" and e.resident and e.pending:
                batch, cols = [], 0
                while e.pending and (not batch or cols + e.pending[0][0].shape[1] <= self.max_batch_cols):
                    item = e.pending.popleft()
                    batch.append(item)
                    cols += item[0].shape[1]
                e.state = "solving"
                return ("solve", e, batch)
        if solves_only:
            return None
        # 2. factorizations admitted by memory (FIFO; a large one waits for room)
        if self._admit_q:
            e = self._admit_q[0]
            if self._make_room(e.bytes) or self._resident_bytes == 0:
                self._admit_q.popleft()
                self._resident_bytes += e.bytes
                e.resident = True
                e.state = "factoring"
                return ("factor", e, None)
        # 3. host-side analysis of new matrices
        if self._analyze_q:
            e = self._analyze_q.popleft()
            e.state = "analyzing"
            return ("analyze", e, None)
        return None

    def _work(self, solves_only):
        while True:
            with self._cv:
                job = None
                while not self._closing:
                    job = self._next_job(solves_only)
                    if job is not None:
                        break
                    self._cv.wait()
                if job is None:
                    return
            kind, e, batch = job
            try:
                if kind == "analyze":
                    M = self._factory()
                    M.Init(e.opts)
                    M.SymbolicFactorization(e.A)
                    nb = M.DeviceBytes(self.max_batch_cols)
                    with self._cv:
                        e.M, e.bytes, e.state = M, nb, "waiting"
                        self._admit_q.append(e)
                        self._cv.notify_all()
                elif kind == "factor":
                    t_start = time.perf_counter()
                    e.M.DistributeMatrix(e.A)
                    e.M.Factorize()
                    t_done = time.perf_counter()
                    with self._cv:
                        e.state = "idle"
                        e.last_used = t_done
                        for fut, ta in e.factor_futs:
                            self.records.append(("factor", e.key, e.A.size, 0, ta, t_start, t_done))
                            fut.set_result(e.M.stats())
                        e.factor_futs = []
                        if e.pending:
                            self._solve_q.append(e)
                        self._cv.notify_all()
                else:
                    t_start = time.perf_counter()
                    B = batch[0][0] if len(batch) == 1 else np.concatenate([b[0] for b in batch], axis=1)
                    k = B.shape[1]
                    e.M.Solve(B, k)
                    X = e.M.GetSolution(k).reshape(e.A.size, k)
                    t_done = time.perf_counter()
                    with self._cv:
                        if len(batch) > 1:
                            self.batched_solves += 1
                        c = 0
                        for Bm, fut, ta in batch:
                            w = Bm.shape[1]
                            fut.set_result(X[:, c:c + w].copy())
                            self.records.append(("solve", e.key, e.A.size, w, ta, t_start, t_done))
                            c += w
                        e.state = "idle"
                        e.last_used = t_done
                        if e.pending:
                            self._solve_q.append(e)
                        self._cv.notify_all()
            except BaseException as exc:  # noqa: BLE001 - reported through the futures
                with self._cv:
                    e.error = exc
                    e.state = "failed"
                    # whatever DistributeMatrix / Factorize allocated before
                    # the failure goes back before the budget counts it free
                    if e.M is not None:
                        try:
                            e.M.ReleaseDevice()
                        except Exception:  # noqa: BLE001 - the original error is the one reported
                            pass
                    if e.resident:
                        self._resident_bytes -= e.bytes
                        e.resident = False
                    for fut, _ in e.factor_futs:
                        if not fut.done():
                            fut.set_exception(exc)
                    e.factor_futs = []
                    for item in (batch or []):
                        if not item[1].done():
                            item[1].set_exception(exc)
                    while e.pending:
                        e.pending.popleft()[1].set_exception(exc)
                    self._cv.notify_all()


def make_request_stream(n_requests, rate_per_s, seed=0, solve_fraction=0.7, n_lo=2e3, n_hi=2e5,
                        max_nrhs=32):
    """Synthetic request mix: new factors of 2D 5-point or 3D 7-point
    random-coefficient Laplacians (coefficients 10**U(-2,2)) with n drawn
    log-uniform in [n_lo, n_hi] and snapped to a square / cube grid, and
    follow-up solves (nrhs uniform in 1..max_nrhs, entries N(0,1)) on factors
    submitted earlier; Poisson arrivals at `rate_per_s`.  Returns a list of
    (t_arrival, kind, key, payload) with payload (dim, k, seed) for factors
    and (nrhs, seed) for solves."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    keys = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate_per_s)
        if not keys or rng.random() >= solve_fraction:
            dim = 2 if rng.random() < 0.5 else 3
            n = 10.0 ** rng.uniform(np.log10(n_lo), np.log10(n_hi))
            k = max(4, int(round(n ** (1.0 / dim))))
            key = "m%d" % i
            keys.append(key)
            out.append((t, "factor", key, (dim, k, int(rng.integers(1 << 31)))))
        else:
            key = keys[int(rng.integers(len(keys)))]
            out.append((t, "solve", key, (int(rng.integers(1, max_nrhs + 1)), int(rng.integers(1 << 31)))))
    return out
