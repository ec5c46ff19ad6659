"""The oracle against the reference's 12D and 24D pyrazine propagation
benchmarks (two coupled electronic states,
Working_tests/MPI_tests/{12D,24D}_propagation_S0/benchmark): the per-mode SG4
basis, the type-1 and type-0 operators and the Chebyshev driver reproduce the
reference's autocorrelation."""
import functools

import numpy as np
import pytest
import scipy.linalg as sl

import propa
from pyrazine import P12, P24

MODELS = {"12d": P12, "24d": P24}


@functools.lru_cache(maxsize=4)
def _dense(name, L):
    import oracle
    P = MODELS[name]
    D, LG, LB, nbA, nbB, nqA, nqB = P.sg4_args(L)
    s = oracle.SG4(D, LG, LB, nbA=nbA, nbB=nbB, nqA=nqA, nqB=nqB)
    V11, V22, V12 = P.potentials(s.grid())
    I = np.eye(s.nb_packed)
    H11 = s.oppsi(oracle.Op(1, V=V11, F2=P.F2()), I)
    H22 = s.oppsi(oracle.Op(1, V=V22, F2=P.F2()), I)
    W = s.oppsi(oracle.Op(0, V=V12), I)
    return s, np.block([[H11, W], [W, H22]])


def coupled_dense(oracle_mod, P=P12, L=None):
    """(oracle SG4, dense coupled operator [[H11, V12], [V12, H22]])."""
    name = [k for k, v in MODELS.items() if v is P][0]
    return _dense(name, P.L if L is None else L)


def psi0(s):
    assert not s.packed()[0].any()          # packed function 0 is the ground tuple
    p = np.zeros(2 * s.nb_packed, complex)
    p[s.nb_packed] = 1.0                    # electronic state 2
    return p


def test_pyrazine_basis_sizes(oracle_mod):
    s, _ = coupled_dense(oracle_mod, P12)
    # 1 + 12 terms; packed: ground + (nb_d(1) - 1) functions per mode
    assert (s.nterms, s.nb_packed) == (13, 1 + sum(P12.NQB))
    lv, w, _, _ = s.terms()
    assert w.sum() == 1.0 and lv.sum(1).max() == 1
    s, _ = coupled_dense(oracle_mod, P24)
    # L = 2: ground, 2 B_d per mode (levels 1 and 2), B_d B_e per pair
    B = np.array(P24.NQB)
    assert s.nb_packed == 1 + 2 * B.sum() + (B.sum() ** 2 - (B * B).sum()) // 2
    assert s.nterms == 1 + 2 * 24 + 24 * 23 // 2


@pytest.mark.parametrize("name", ["12d", "24d"])
def test_pyrazine_exact_autocorr(oracle_mod, name):
    """exp(-i H t) of the oracle's coupled operator gives the benchmark's
    <psi0|psi(t)> (printed with 17 digits) to 1e-11."""
    P = MODELS[name]
    s, H = coupled_dense(oracle_mod, P)
    t, ref = P.autocorr_ref()
    assert np.allclose(t, P.DT * np.arange(P.NSTEPS + 1), rtol=1e-15, atol=1e-12)
    U = sl.expm(-1j * H * P.DT)
    p0 = psi0(s)
    p = p0.copy()
    got = [1.0]
    for _ in range(P.NSTEPS):
        p = U @ p
        got.append(np.vdot(p0, p))
    err = np.abs(np.array(got) - ref).max()
    print("pyrazine %s exact vs benchmark %.3e" % (name, err))
    assert err < 1e-11


@pytest.mark.parametrize("name", ["12d", "24d"])
def test_pyrazine_cheby_autocorr(oracle_mod, name):
    """The oracle's march_cheby at the reference's npoly/poly_tol reproduces the
    benchmark to 1e-10 (an error of the Bessel coefficients, the recurrence
    or the exit test would show at the first step)."""
    P = MODELS[name]
    s, H = coupled_dense(oracle_mod, P)
    ev = np.linalg.eigvalsh(0.5 * (H + H.T))   # bounds only: H is symmetric to 1e-15
    lo, hi = ev.min(), ev.max()
    t, ref = P.autocorr_ref()
    p0 = psi0(s)
    p = p0.reshape(-1, 1).copy()
    got = [1.0]
    for _ in range(P.NSTEPS):
        p, order, _ = propa.cheby_step(H, p, P.DT, lo - 1e-3 * (hi - lo), hi + 1e-3 * (hi - lo),
                                       P.NPOLY, P.POLY_TOL)
        got.append(np.vdot(p0, p[:, 0]))
    err = np.abs(np.array(got) - ref).max()
    print("pyrazine %s cheby vs benchmark %.3e, last order %d" % (name, err, order))
    assert err < 1e-10


def test_pyrazine_planted_errors_fail(oracle_mod):
    """A flipped sign of delta or a shifted diagonal block is seen by the check."""
    s, H = coupled_dense(oracle_mod, P12)
    n = s.nb_packed
    t, ref = P12.autocorr_ref()
    p0 = psi0(s)
    for Hb in (H - 2 * np.diag(np.r_[np.zeros(n), np.full(n, P12.fix["delta_eV"] / P12.AU)]),
               H + 1e-3 * np.diag(np.r_[np.ones(n), np.zeros(n)])):
        U = sl.expm(-1j * Hb * P12.DT * P12.NSTEPS)
        assert abs(np.vdot(p0, U @ p0) - ref[-1]) > 1e-6
