"""Serving on the GPU: results of requests processed in a mix of concurrent
factor and solve requests equal the results of the same requests processed
alone (FP64: 1e-12 relative; FP32+IR: both within backward error 1e-12, and
1e-8 relative to each other)."""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from sympack_amd.serving import SolverService
from oracle import numeric as onum

pytestmark = pytest.mark.gpu


def _opts(dim, prec):
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    return sp.symPACKOptions(ordering="NDGRID" if dim == 2 else "NDBOX", relax=rp, precision=prec)


@pytest.mark.parametrize("prec", ["FP64", "FP32_IR"])
def test_results_alone_equal_results_in_mix(prec):
    specs = [(2, 40, 1), (3, 12, 2), (2, 64, 3), (3, 16, 4), (2, 23, 5), (3, 9, 6)]
    mats = {}
    for dim, k, seed in specs:
        n, cp, ri, v = M.random_laplacian("2d" if dim == 2 else "3d", k, seed)
        mats["m%d" % seed] = (dim, sp.DistSparseMatrix(n, cp, ri, v))
    rng = np.random.default_rng(7)
    keys = sorted(mats)
    solves = [(k, rng.standard_normal((mats[k][1].size, int(rng.integers(1, 9)))))
              for k in rng.choice(keys, 30)]
    print("serving alone", prec, flush=True)
    alone = []
    with SolverService(slots=1) as s:
        for k in keys:
            dim, A = mats[k]
            s.submit_factor(k, A, _opts(dim, prec)).result()
        for k, B in solves:
            alone.append(s.submit_solve(k, B).result())
    print("serving mix", prec, flush=True)
    with SolverService(slots=4, max_batch_cols=16) as s:
        ff = [s.submit_factor(k, mats[k][1], _opts(mats[k][0], prec)) for k in keys]
        fs = [s.submit_solve(k, B) for k, B in solves]
        mix = [f.result(timeout=120) for f in fs]
        for f in ff:
            f.result(timeout=120)
        rep = s.report()
    assert rep["solve"]["count"] == len(solves)
    for (k, B), a, b in zip(solves, alone, mix):
        A = mats[k][1]
        rel = np.linalg.norm(a - b) / np.linalg.norm(a)
        if prec == "FP64":
            assert rel < 1e-12
        else:
            assert rel < 1e-8
            assert onum.backward_error(A, b, B) <= 1e-12
            assert onum.backward_error(A, a, B) <= 1e-12
