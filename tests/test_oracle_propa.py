"""The propagation oracle (oracle/propa.py) pinned against the mathematics.

Chebyshev (march_cheby) and SIL (march_SIL) restatements must reproduce
exp(-i H dt) psi from an eigendecomposition, with and without a field term;
the Chebyshev coefficients must satisfy the Jacobi-Anger expansion
exp(-i x r) = J_0(r) + 2 sum_k (-i)^k J_k(r) T_k(x).  A dropped factor 2, a
wrong sign of -i, a missing phase exp(-i E0 dt) or an index shift in the
coefficients fails these checks.
"""
import numpy as np
import pytest

import propa


def rand_sym(n, seed, lo=0.0, hi=5.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.sort(rng.uniform(lo, hi, n))
    return (Q * ev) @ Q.T


def rand_psi(n, nwp, seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, nwp)) + 1j * rng.standard_normal((n, nwp))
    return p / np.linalg.norm(p, axis=0)


def test_jacobi_anger_expansion():
    x = np.linspace(-1, 1, 41)
    for r in [0.5, 7.0, 60.0]:
        c = propa.cheb_coefs(int(r) + 40, r)
        T = np.cos(np.arange(len(c))[:, None] * np.arccos(x)[None, :])
        s = (c[:, None] * (-1j) ** np.arange(len(c))[:, None] * T).sum(0)
        assert np.abs(s - np.exp(-1j * x * r)).max() < 1e-13


@pytest.mark.parametrize("dt", [0.3, 4.0])
def test_cheby_matches_exact(dt):
    H = rand_sym(60, 1)
    psi = rand_psi(60, 3, 2)
    out, order, _ = propa.cheby_step(H, psi, dt, -0.1, 5.1, 200, 1e-15)
    assert order < 199
    assert np.abs(out - propa.exact_step(H, psi, dt)).max() < 1e-12
    assert np.abs(np.linalg.norm(out, axis=0) - 1).max() < 1e-12


def test_cheby_hmax_retry():
    """An interval below the spectrum top diverges and is widened (MARCH:4348-4373)."""
    H = rand_sym(40, 3, 0.0, 8.0)
    psi = rand_psi(40, 1, 4)
    out, _, hmax = propa.cheby_step(H, psi, 2.0, -0.1, 4.0, 400, 1e-14)
    assert hmax > 4.0
    assert np.abs(out - propa.exact_step(H, psi, 2.0)).max() < 1e-10


def test_sil_matches_exact():
    H = rand_sym(60, 5)
    psi = rand_psi(60, 2, 6)
    out, n = propa.sil_step(H, psi, 0.2, 30, 1e-14)
    assert n < 30
    assert np.abs(out - propa.exact_step(H, psi, 0.2)).max() < 1e-11


def test_field_term():
    H = rand_sym(50, 7)
    M = np.diag(np.linspace(-1, 1, 50))
    psi = rand_psi(50, 2, 8)
    E = np.array([0.3, -0.7])
    ex = propa.exact_step(H, psi, 0.5, E, M)
    # exact_step with a field against the field-free formula on H - E M
    for w in range(2):
        assert np.abs(ex[:, w] - propa.exact_step(H - E[w] * M, psi[:, w:w + 1], 0.5)[:, 0]).max() < 1e-13
    out, _, _ = propa.cheby_step(H, psi, 0.5, -1.0, 6.0, 200, 1e-15, E, M)
    assert np.abs(out - ex).max() < 1e-12
    out2, _ = propa.sil_step(H, psi, 0.5, 40, 1e-14, E, M)
    assert np.abs(out2 - ex).max() < 1e-11


def test_cheby_on_symmetrised_sg4_operator(oracle_mod):
    from evrhip import hamiltonian as Hm
    D, L = 3, 3
    ham = Hm.smolyak_config(1, D)
    s = oracle_mod.SG4(D, L)
    M = s.oppsi(oracle_mod.Op(10, V=ham.potential(s.grid()), G0=ham.G0, Ga=ham.Ga_fortran()),
                np.eye(s.nb_packed))
    Hs = 0.5 * (M + M.T)
    ev = np.linalg.eigvalsh(Hs)
    psi = rand_psi(s.nb_packed, 2, 9)
    for dt in [2e-3, 2e-2]:   # energies in cm^-1: r = Esc dt = 13 and 134
        out, _, _ = propa.cheby_step(Hs, psi, dt, ev[0] - 1.0, ev[-1] + 1.0, 400, 1e-14)
        assert np.abs(out - propa.exact_step(Hs, psi, dt)).max() < 1e-11


def _sym(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return scale * (A + A.T) / 2


def test_field_derivs_closed_form():
    """E^(j)(t) of the control field against finite differences of E(t) and the
    direct product form sin^2(pi t/T) sum_k a_k sin(k pi t/T)."""
    import control
    rng = np.random.default_rng(1)
    amp = rng.uniform(-1, 1, (5, 2))
    T = 7.0
    E = lambda t: (np.sin(np.pi * t / T) ** 2) * (np.sin(np.outer([t], np.arange(1, 6)) * np.pi / T) @ amp)[0]
    for t in (0.3, 2.9, 5.5):
        d = control.field_derivs(amp, T, t, 3)
        assert np.allclose(d[0], E(t), atol=1e-14)
        h = 1e-4
        assert np.allclose(d[1], (E(t + h) - E(t - h)) / (2 * h), atol=1e-7)
        assert np.allclose(d[2], (E(t + h) - 2 * E(t) + E(t - h)) / h ** 2, atol=1e-5)
        assert np.allclose(d[3], (control.field_derivs(amp, T, t + h, 2)[2] -
                                  control.field_derivs(amp, T, t - h, 2)[2]) / (2 * h), atol=1e-5)


def test_taylor_field_converges_to_time_dependent_solution():
    """march_noD_field with the field derivatives reproduces the time-dependent
    solution (fine fourth-order Magnus) to 1e-11; with the derivatives dropped
    (the field frozen at t_n) it does not (a first-order error)."""
    import control
    n = 30
    H = _sym(n, 2) + np.diag(np.linspace(0, 3, n))
    M = _sym(n, 3, 0.3)
    rng = np.random.default_rng(4)
    psi = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    psi /= np.linalg.norm(psi, axis=0)
    amp = rng.uniform(-0.5, 0.5, (4, 2))
    T, dt, nst = 10.0, 0.25, 8
    E0 = 1.5
    got = psi.copy()
    frozen = psi.copy()
    for st in range(nst):
        d = control.field_derivs(amp, T, st * dt, 12)
        got, order = propa.taylor_field_step(H, got, dt, E0, 40, 1e-28, dnE=d, M=M)
        frozen, _ = propa.taylor_field_step(H, frozen, dt, E0, 40, 1e-28, dnE=d[:1], M=M)
    for w in range(2):
        f = lambda t: control.field_derivs(amp[:, w], T, t, 0)[0, 0]
        ref = propa.magnus4_propagate(H, psi[:, w:w + 1], 0.0, nst * dt, 400, f, M)[:, 0]
        assert np.abs(got[:, w] - ref).max() < 1e-11
        assert np.abs(frozen[:, w] - ref).max() > 1e-4


def test_rk4_field_order_four():
    """march_RK4_field: error against the time-dependent solution falls 16x
    per halving of dt (fourth order), and is small (1e-5) at dt = 0.02."""
    import control
    n = 20
    H = _sym(n, 5) + np.diag(np.linspace(0, 2, n))
    M = _sym(n, 6, 0.3)
    rng = np.random.default_rng(7)
    psi = rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1))
    psi /= np.linalg.norm(psi)
    amp = rng.uniform(-0.5, 0.5, (3, 1))
    T, t1, E0 = 6.0, 1.0, 0.7
    f = lambda t: control.field_derivs(amp, T, t, 0)[0, 0]
    ref = propa.magnus4_propagate(H, psi, 0.0, t1, 400, f, M)
    errs = []
    for nst in (25, 50):
        dt = t1 / nst
        x = psi.copy()
        for st in range(nst):
            ev = np.array([[f(st * dt)], [f(st * dt + dt / 2)], [f(st * dt + dt)]])
            x = propa.rk4_field_step(H, x, dt, E0, ev, M)
        errs.append(np.abs(x - ref).max())
    print(errs)
    assert errs[1] < 1e-5
    assert 12.0 < errs[0] / errs[1] < 20.0
