"""The oracle against the reference's 12D pyrazine propagation benchmark
(two coupled electronic states, Working_tests/MPI_tests/12D_propagation_S0/benchmark):
the per-mode SG4 basis, the type-1 and type-0 operators and the Chebyshev
driver reproduce the reference's autocorrelation."""
import numpy as np
import pytest
import scipy.linalg as sl

import pyrazine as P


def coupled_dense(oracle_mod, L=P.L):
    s = oracle_mod.SG4(P.D, L, L, nbA=P.NQA, nbB=P.NQB, nqA=P.NQA, nqB=P.NQB)
    V11, V22, V12 = P.potentials(s.grid())
    I = np.eye(s.nb_packed)
    H11 = s.oppsi(oracle_mod.Op(1, V=V11, F2=P.F2()), I)
    H22 = s.oppsi(oracle_mod.Op(1, V=V22, F2=P.F2()), I)
    W = s.oppsi(oracle_mod.Op(0, V=V12), I)
    return s, np.block([[H11, W], [W, H22]])


def psi0(s):
    assert not s.packed()[0].any()          # packed function 0 is the ground tuple
    p = np.zeros(2 * s.nb_packed, complex)
    p[s.nb_packed] = 1.0                    # electronic state 2
    return p


def test_pyrazine_basis_sizes(oracle_mod):
    s = oracle_mod.SG4(P.D, 1, 1, nbA=P.NQA, nbB=P.NQB, nqA=P.NQA, nqB=P.NQB)
    # 1 + 12 terms; packed: ground + (nb_d(1) - 1) functions per mode
    assert (s.nterms, s.nb_packed) == (13, 1 + sum(P.NQB))
    lv, w, _, _ = s.terms()
    assert w.sum() == 1.0 and lv.sum(1).max() == 1


def test_pyrazine_exact_autocorr(oracle_mod):
    """exp(-i H t) of the oracle's coupled operator gives the benchmark's
    <psi0|psi(t)> (printed with 17 digits) to 1e-12."""
    s, H = coupled_dense(oracle_mod)
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
    print("pyrazine exact vs benchmark", err)
    assert err < 1e-12


def test_pyrazine_cheby_autocorr(oracle_mod):
    """The oracle's march_cheby at the reference's npoly/poly_tol reproduces the
    benchmark to 1e-9 (an error of the Bessel coefficients, the recurrence
    or the exit test would show at the first step)."""
    s, H = coupled_dense(oracle_mod)
    ev = np.linalg.eigvals(H).real
    lo, hi = ev.min(), ev.max()
    import propa
    t, ref = P.autocorr_ref()
    p0 = psi0(s)
    p = p0.reshape(-1, 1).copy()
    got = [1.0]
    for _ in range(P.NSTEPS):
        p, order, _ = propa.cheby_step(H, p, P.DT, lo - 1e-3 * (hi - lo), hi + 1e-3 * (hi - lo),
                                       P.NPOLY, P.POLY_TOL)
        got.append(np.vdot(p0, p[:, 0]))
    err = np.abs(np.array(got) - ref).max()
    print("pyrazine cheby vs benchmark", err, "last order", order)
    assert err < 1e-9


def test_pyrazine_planted_sign_fails(oracle_mod):
    """A flipped sign of the coupling (V12 -> -V12 changes the spectrum of
    neither block but the interference) or of delta is seen by the check."""
    s, H = coupled_dense(oracle_mod)
    n = s.nb_packed
    t, ref = P.autocorr_ref()
    p0 = psi0(s)
    for Hb in (H - 2 * np.diag(np.r_[np.zeros(n), np.full(n, P.FIX["delta_eV"] / P.AU)]),
               H + 1e-3 * np.diag(np.r_[np.ones(n), np.zeros(n)])):
        U = sl.expm(-1j * Hb * P.DT * P.NSTEPS)
        assert abs(np.vdot(p0, U @ p0) - ref[-1]) > 1e-6
