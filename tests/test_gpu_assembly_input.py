"""Assembly of caller CSC input that is not in canonical form.  The assembly
map has two paths: lower-triangle columns with strictly increasing rows (no
two entries can share a position; each entry looks its row up by itself,
on several host threads), and anything else (entries bucketed by permuted
column, earlier duplicates of a position dropped: the reference's
DistributeMatrix keeps the last value written).  The same matrix given in
both forms must give the same factor and solution (1e-12 relative; the
Schur sums arrive through atomics, so runs differ at rounding level), and
both must match the oracle (1e-10, the parity bar)."""
import numpy as np
import pytest

import sympack_amd as sp
from sympack_amd import matrices as M
from oracle import numeric as onum


def _messy(n, cp, ri, v):
    """Every column's rows reversed, preceded by a duplicate of its diagonal
    entry holding garbage (earlier, so it must lose)."""
    cp2, ri2, v2 = [0], [], []
    for c in range(n):
        rows, vals = ri[cp[c]:cp[c + 1]], v[cp[c]:cp[c + 1]]
        ri2.append(rows[0])
        v2.append(1e30)
        ri2.extend(rows[::-1])
        v2.extend(vals[::-1])
        cp2.append(len(ri2))
    return np.array(cp2), np.array(ri2), np.array(v2)


def _solve(A, ldl, B):
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(40)
    m = sp.symPACKMatrix2D()
    m.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL" if ldl else "LL", relax=rp))
    m.SymbolicFactorization(A)
    m.DistributeMatrix(A)
    m.Factorize()
    m.Solve(B, B.shape[1])
    return m, m.GetSolution(B.shape[1]).reshape(B.shape)


def test_messy_input_symbolic_matches_clean():
    n, cp, ri, v = M.diffusion_3d_7pt(10, seed=2)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    Am = sp.DistSparseMatrix(n, *_messy(n, cp, ri, v))
    for X in (A, Am):
        assert X.size == n
    ms = []
    for X in (A, Am):
        rp = sp.RelaxationParameters()
        rp.SetMaxSize(40)
        m = sp.symPACKMatrix2D()
        m.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp))
        m.SymbolicFactorization(X)
        ms.append(m)
    assert np.array_equal(ms[0].GetSupernodalPartition(), ms[1].GetSupernodalPartition())
    for x, y in zip(ms[0].GetStructure(), ms[1].GetStructure()):
        assert np.array_equal(x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("ldl", [False, True], ids=["LL", "LDL"])
def test_messy_input_assembles_like_clean(ldl):
    gen = (lambda: M.shifted_laplacian_3d_7pt(10, 0.3)) if ldl else (lambda: M.diffusion_3d_7pt(10, seed=2))
    n, cp, ri, v = gen()
    A = sp.DistSparseMatrix(n, cp, ri, v)
    Am = sp.DistSparseMatrix(n, *_messy(n, cp, ri, v))
    B = np.random.default_rng(5).uniform(-1, 1, (n, 2))
    print("case messy input 10^3 ldl=%s" % ldl, flush=True)
    m, X = _solve(A, ldl, B)
    _, Xm = _solve(Am, ldl, B)
    assert np.linalg.norm(Xm - X) / np.linalg.norm(X) < 1e-12
    S = onum.Structure.of(m)
    info, F = onum.factor(S, A, ldl)
    assert info == 0
    Xo = onum.solve(S, F, B, ldl)
    assert np.linalg.norm(Xm - Xo) / np.linalg.norm(Xo) < 1e-10
