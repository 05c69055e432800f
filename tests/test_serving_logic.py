"""Scheduling logic of the serving layer (admission by memory budget, LRU
eviction, merged solves, error propagation, same result alone and in the mix)
on the CPU, with a dense stand-in for the factor object."""
import threading
import time

import numpy as np
import pytest
import scipy.linalg as sl

import sympack_amd as sp
from sympack_amd import matrices as M
from sympack_amd.serving import SolverService, make_request_stream


class DenseFactor:
    """Stand-in with the symPACKMatrix2D methods the service calls."""
    live = 0
    lock = threading.Lock()

    def Init(self, opts):
        self.opts = opts

    def SymbolicFactorization(self, A):
        self.n = A.size

    def DeviceBytes(self, nrhs):
        return self.n * self.n * 8

    def DistributeMatrix(self, A):
        self.A = A.to_scipy_full().toarray()
        with DenseFactor.lock:
            if not getattr(self, "held", False):
                DenseFactor.live += 1
            self.held = True

    def Factorize(self):
        if np.any(np.diag(self.A) <= 0):
            raise sp.NotHPDError(1)
        time.sleep(0.002)
        self.L = sl.cholesky(self.A, lower=True)

    def Solve(self, B, nrhs):
        time.sleep(0.001)
        self.X = sl.cho_solve((self.L, True), np.asarray(B).reshape(self.n, nrhs))

    def GetSolution(self, nrhs):
        return self.X

    def ReleaseDevice(self):
        # idempotent, as symPACKMatrix2D.ReleaseDevice
        self.L = None
        with DenseFactor.lock:
            if getattr(self, "held", False):
                DenseFactor.live -= 1
            self.held = False

    def stats(self):
        return {"n": self.n}


def _mat(i):
    n, cp, ri, v = M.random_laplacian("2d" if i % 2 else "3d", 6 + i % 3, seed=i)
    return sp.DistSparseMatrix(n, cp, ri, v)


def test_mix_equals_alone_and_batches():
    mats = {("m%d" % i): _mat(i) for i in range(6)}
    rng = np.random.default_rng(0)
    solves = [(k, rng.standard_normal((mats[k].size, int(rng.integers(1, 5)))))
              for k in rng.choice(sorted(mats), 40)]
    # alone: one slot, one request at a time
    alone = []
    with SolverService(slots=1, matrix_factory=DenseFactor) as s:
        for k, A in mats.items():
            s.submit_factor(k, A).result()
        for k, B in solves:
            alone.append(s.submit_solve(k, B).result())
    # mix: everything in flight at once on 4 slots
    with SolverService(slots=4, matrix_factory=DenseFactor) as s:
        ff = [s.submit_factor(k, A) for k, A in mats.items()]
        fs = [s.submit_solve(k, B) for k, B in solves]
        mix = [f.result() for f in fs]
        assert all(f.result()["n"] > 0 for f in ff)
        rep = s.report()
    for a, b in zip(alone, mix):
        assert a.shape == b.shape
        assert np.allclose(a, b, rtol=1e-12, atol=0)
    assert rep["solve"]["count"] == len(solves)
    assert rep["merged_solve_batches"] > 0
    for (k, B), X in zip(solves, mix):
        Af = mats[k].to_scipy_full().toarray()
        assert np.linalg.norm(Af @ X - B) <= 1e-10 * np.linalg.norm(B)


def test_budget_evicts_lru_and_refactors():
    mats = {("m%d" % i): _mat(i) for i in range(5)}
    biggest = max(A.size for A in mats.values()) ** 2 * 8
    with SolverService(hbm_budget_bytes=int(1.2 * biggest), slots=2, matrix_factory=DenseFactor) as s:
        for k, A in mats.items():
            s.submit_factor(k, A).result()
            assert s._resident_bytes <= s.budget
        # the first factors were evicted; a solve brings them back
        B = np.ones(mats["m0"].size)
        X = s.submit_solve("m0", B).result()
        Af = mats["m0"].to_scipy_full().toarray()
        assert np.allclose(Af @ X[:, 0], B)
        assert s.report()["evictions"] >= 3
        assert s._resident_bytes <= s.budget


def test_errors_reach_the_futures():
    n, cp, ri, v = M.laplacian_2d_5pt(5)
    v = v.copy()
    v[cp[3]] = -1.0
    bad = sp.DistSparseMatrix(n, cp, ri, v)
    live0 = DenseFactor.live
    with SolverService(slots=2, matrix_factory=DenseFactor) as s:
        f = s.submit_factor("bad", bad)
        g = s.submit_solve("bad", np.ones(n))
        with pytest.raises(sp.NotHPDError):
            f.result()
        with pytest.raises(sp.NotHPDError):
            g.result()
        with pytest.raises(KeyError):
            s.submit_solve("missing", np.ones(3))
        # the failed factorization gave its device memory back, and the
        # budget no longer counts it
        s.drain()
        assert DenseFactor.live == live0
        assert s._resident_bytes == 0


def test_request_stream_shape():
    st = make_request_stream(200, 100.0, seed=1)
    assert len(st) == 200
    ts = [r[0] for r in st]
    assert ts == sorted(ts)
    kinds = [r[1] for r in st]
    assert kinds[0] == "factor"
    frac = kinds.count("solve") / len(kinds)
    assert 0.6 < frac < 0.8
    seen = set()
    for t, kind, key, p in st:
        if kind == "factor":
            seen.add(key)
            dim, k, seed = p
            n = k ** dim
            assert 1000 <= n <= 3e5
        else:
            assert key in seen and 1 <= p[0] <= 32


class SlowFactor(DenseFactor):
    """Factorizations that take long enough to fill every general slot."""
    gate = threading.Event()

    def Factorize(self):
        if getattr(self.opts, "orderingStr", "") == "NATURAL":
            SlowFactor.gate.wait(10.0)
        super().Factorize()


def test_solve_slot_serves_while_factorizations_hold_the_other_slots():
    """With one of three slots reserved for solves, a solve on a resident
    factor completes while two factorizations occupy both general slots."""
    SlowFactor.gate.clear()
    mats = {("m%d" % i): _mat(i) for i in range(3)}
    with SolverService(slots=3, solve_slots=1, matrix_factory=SlowFactor) as s:
        s.submit_factor("m0", mats["m0"]).result(timeout=10)
        slow = [s.submit_factor(k, mats[k], sp.symPACKOptions(ordering="NATURAL")) for k in ("m1", "m2")]
        time.sleep(0.05)  # both slow factorizations are running (or queued behind analysis)
        B = np.ones((mats["m0"].size, 1))
        x = s.submit_solve("m0", B).result(timeout=5)
        assert not any(f.done() for f in slow)
        A = mats["m0"].to_scipy_full().toarray()
        assert np.allclose(A @ x, B)
        SlowFactor.gate.set()
        for f in slow:
            f.result(timeout=10)


def test_solve_slots_must_leave_a_general_slot():
    with pytest.raises(ValueError, match="solve_slots"):
        SolverService(slots=2, solve_slots=2, matrix_factory=DenseFactor)
