"""Host-side basis construction of the engine and the C ABI surface (no GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    src = open(os.path.join(ROOT, "include", "evr_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(evr_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(evr):
    lib = ctypes.CDLL(evr._lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the python binding declares exactly the header's functions
    assert sorted(evr._lib.declared_symbols()) == names


@pytest.mark.parametrize("D,L,LB,law", [(6, 5, 5, (1, 2, 2, 2)), (3, 4, 2, (1, 2, 2, 2)),
                                        (6, 3, 3, (1, 2, 1, 2)), (4, 3, 3, (1, 1, 1, 1))])
def test_sg4_tables_match_oracle(evr, oracle_mod, D, L, LB, law):
    b = evr.Basis.sg4(D, L, LB, *law)
    s = oracle_mod.SG4(D, L, LB, *law)
    lv, w, _, _ = s.terms()
    assert b.nterms == s.nterms and b.nb_packed == s.nb_packed
    assert (b.term_levels() == lv).all()
    assert (b.weights() == w).all()
    assert (b.map() == s.map()).all()
    assert np.abs(b.grid() - s.grid()).max() < 4e-16 * 10


def test_1d_matrices_match_oracle(evr, oracle_mod):
    """The engine's 1D matrices (independent methods: Sturm bisection nodes,
    Cholesky pseudo-inverse) agree with the oracle's reference-style ones."""
    b = evr.Basis.sg4(2, 8)
    s = oracle_mod.SG4(2, 8)
    d = b.ptr.contents
    nlev = d.nlev
    offq = offB = offD = 0
    for l in range(nlev):
        nb, nq = d.nb[l], d.nq[l]
        o = s.one_d(0, l)
        x = np.ctypeslib.as_array(d.x, shape=(offq + nq,))[offq:]
        B = np.ctypeslib.as_array(d.B, shape=(offB + nb * nq,))[offB:].reshape(nb, nq).T
        W = np.ctypeslib.as_array(d.W, shape=(offB + nb * nq,))[offB:].reshape(nq, nb).T
        D1 = np.ctypeslib.as_array(d.D1, shape=(offD + nq * nq,))[offD:].reshape(nq, nq).T
        D2 = np.ctypeslib.as_array(d.D2, shape=(offD + nq * nq,))[offD:].reshape(nq, nq).T
        assert np.abs(x - o["x"]).max() < 1e-15 * 10
        assert np.abs(B - o["B"]).max() < 1e-14
        assert np.abs(W - o["W"]).max() < 1e-14
        assert np.abs(D1 - o["D1"]).max() < 1e-11 * max(1, np.abs(o["D1"]).max())
        assert np.abs(D2 - o["D2"]).max() < 1e-10 * max(1, np.abs(o["D2"]).max())
        offq += nq; offB += nb * nq; offD += nq * nq


def test_dp_desc(evr, oracle_mod):
    b = evr.Basis.dp([12, 12, 12], [14, 14, 14])
    d = oracle_mod.DP([12, 12, 12], [14, 14, 14])
    assert b.nb_packed == d.nb_tot == 1728 and b.sum_nq == d.nq_tot == 2744
    assert np.abs(b.grid() - d.grid()).max() < 4e-15
    m = b.map()
    assert sorted(m) == list(range(1728))
    assert m[1] == 144  # term-local first mode fastest -> DP last mode fastest


def test_bad_arguments_raise(evr):
    with pytest.raises(RuntimeError):
        evr.Basis.sg4(0, 3)
    with pytest.raises(RuntimeError):
        evr.Basis.sg4(3, 2, 4)
    with pytest.raises(RuntimeError):
        evr.Basis.dp([5], [4])


@pytest.mark.parametrize("r", [0.0, 0.3, 5.0, 57.3, 1234.5])
def test_bessel_coefficients(evr, r):
    """J_k(r) behind the Chebyshev coefficients (host code) against scipy and,
    where scipy's own error reaches 1e-14 (large r), 40-digit mpmath values."""
    import mpmath
    from scipy.special import jv
    n = int(r) + 60
    j = evr.bessel_j(n, r)
    assert np.abs(j - jv(np.arange(n + 1), r)).max() < 1e-12
    mpmath.mp.dps = 40
    for k in sorted({0, 1, 2, n // 3, n // 2, max(n - 61, 0), n - 30, n}):
        assert abs(j[k] - float(mpmath.besselj(k, r))) < 2e-16, k
    # J_0 + 2 sum J_2k = 1 and sum_k J_k^2 (k in Z) = 1 (tail beyond r + 30 r^(1/3) negligible)
    j = evr.bessel_j(int(r + 30 * r ** (1 / 3) + 60), r)
    assert abs(j[0] + 2 * j[2::2].sum() - 1.0) < 1e-14
    assert abs(j[0] ** 2 + 2 * (j[1:] ** 2).sum() - 1.0) < 1e-14


@pytest.mark.parametrize("args", [(2, 8), (2, 1, 1, 40, 1, 41, 1)])
def test_grid_to_basis_matrix_is_weighted_transpose(evr, args):
    """W = B^T diag(w) on every 1D slot, with w_q = W[0, q] / B[q, 0] (and
    B[q, 0] > 0): the identity behind the per-point quadrature weights of the
    control gradient dJ/dV and the transposed sequence of the hermitian grid
    forms (W diag(1/w) Op^T diag(w) B).  Also for slots of 41 functions."""
    b = evr.Basis.sg4(*args)
    d = b.ptr.contents
    offB = 0
    for l in range(d.nlev):
        nb, nq = d.nb[l], d.nq[l]
        B = np.ctypeslib.as_array(d.B, shape=(offB + nb * nq,))[offB:].reshape(nb, nq).T
        W = np.ctypeslib.as_array(d.W, shape=(offB + nb * nq,))[offB:].reshape(nq, nb).T
        assert (B[:, 0] > 0).all()
        w = W[0, :] / B[:, 0]
        assert np.abs(W - B.T * w[None, :]).max() < 1e-13 * np.abs(W).max()
        offB += nb * nq
