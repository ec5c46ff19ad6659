"""The control-gradient oracle (oracle/control.py) pinned against finite differences.

The exact gradient (Daleckii-Krein on every step) must agree with central
differences of the oracle's own objective; the grid projector P_g used for
dJ/dV_g must equal the oracle's type-0 operator with V = e_g.  A sign error in
the costate, a missing conjugate or a transposed F fails these checks.
"""
import numpy as np
import pytest

import control


@pytest.fixture(scope="module")
def small_system(oracle_mod):
    from evrhip import hamiltonian as Hm
    D, L = 3, 3
    ham = Hm.smolyak_config(1, D)
    s = oracle_mod.SG4(D, L)
    q = s.grid()
    M = s.oppsi(oracle_mod.Op(10, V=ham.potential(q), G0=ham.G0, Ga=ham.Ga_fortran()), np.eye(s.nb_packed))
    H = 0.5 * (M + M.T) / Hm.AU_TO_CM
    mu = 0.1 * q[:, 0] + 0.05 * q[:, 1] - 0.03 * q[:, 0] * q[:, 2]
    Mu = s.oppsi(oracle_mod.Op(0, V=mu), np.eye(s.nb_packed))
    Mu = 0.5 * (Mu + Mu.T)
    ev, V = np.linalg.eigh(H)
    return dict(s=s, H=H, Mu=Mu, mu=mu, psi0=V[:, 0].astype(complex), phi1=V[:, 1].astype(complex))


def test_projector_matches_type0_operator(small_system, oracle_mod):
    s = small_system["s"]
    Y = control.grid_basis(s)
    gw = control.grid_weights(s)
    for g in [0, 7, s.sum_nq // 2, s.sum_nq - 1]:
        e = np.zeros(s.sum_nq); e[g] = 1.0
        Pg = s.oppsi(oracle_mod.Op(0, V=e), np.eye(s.nb_packed))
        assert np.abs(Pg - gw[g] * np.outer(Y[g], Y[g])).max() < 1e-13


def test_field_gradient_fd(small_system):
    c = small_system
    T, N, dt = 400.0, 20, 20.0
    amp = np.array([[0.02, -0.01], [0.015, 0.02], [-0.01, 0.005], [0.01, -0.02]])
    E = control.field_values(amp, T, N)
    J, gE, _ = control.gradient(c["H"], c["Mu"], c["psi0"], c["phi1"], E, dt)
    assert np.allclose(J, control.objective(c["H"], c["Mu"], c["psi0"], c["phi1"], E, dt), atol=1e-14)
    assert J.min() > 1e-8
    h = 1e-6
    for n in [0, 7, N - 1]:
        Ep, Em = E.copy(), E.copy()
        Ep[n] += h
        Em[n] -= h
        fd = (control.objective(c["H"], c["Mu"], c["psi0"], c["phi1"], Ep, dt)
              - control.objective(c["H"], c["Mu"], c["psi0"], c["phi1"], Em, dt)) / (2 * h)
        assert np.abs(fd - gE[n]).max() < 1e-6 * np.abs(gE).max() + 1e-12
    # chain rule to the coefficients
    ga = control.field_basis(T, N, 4).T @ gE
    k = 1
    ap, am = amp.copy(), amp.copy()
    ap[k] += h
    am[k] -= h
    fd = (control.objective(c["H"], c["Mu"], c["psi0"], c["phi1"], control.field_values(ap, T, N), dt)
          - control.objective(c["H"], c["Mu"], c["psi0"], c["phi1"], control.field_values(am, T, N), dt)) / (2 * h)
    assert np.abs(fd - ga[k]).max() < 1e-6 * np.abs(ga).max()


def test_potential_gradient_fd(small_system):
    c = small_system
    s = c["s"]
    T, N, dt = 200.0, 10, 20.0
    E = control.field_values(np.array([[0.03], [-0.02]]), T, N)
    Y = control.grid_basis(s)
    gw = control.grid_weights(s)
    J, _, gV = control.gradient(c["H"], c["Mu"], c["psi0"], c["phi1"], E, dt, Y, gw)
    h = 1e-7
    for g in [3, s.sum_nq // 3, s.sum_nq - 2]:
        P = gw[g] * np.outer(Y[g], Y[g])
        fd = (control.objective(c["H"] + h * P, c["Mu"], c["psi0"], c["phi1"], E, dt)
              - control.objective(c["H"] - h * P, c["Mu"], c["psi0"], c["phi1"], E, dt)) / (2 * h)
        assert abs(fd[0] - gV[g, 0]) < 1e-5 * np.abs(gV).max()


@pytest.mark.parametrize("method,nsub", [(3, 6), (4, 3)])
def test_control_field_sub_matches_oracle(evr, method, nsub):
    """evr_control_field_sub (host code of the library) against the oracle's
    closed-form field derivatives: Taylor derivatives at t_n, RK4 values at
    t_n, t_n + dt/2, t_n + dt."""
    rng = np.random.default_rng(9)
    amp = rng.uniform(-1e-3, 1e-3, (8, 3))
    T, N = 2000.0, 20
    got = evr.control_field_sub(amp, T, N, method, nsub)
    dt = T / N
    for n in range(N):
        if method == 3:
            ref = control.field_derivs(amp, T, n * dt, nsub - 1)
        else:
            ref = np.stack([control.field_derivs(amp, T, n * dt + j * dt / 2, 0)[0] for j in range(3)])
        assert np.abs(got[n] - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_build_field_is_the_scaled_gradient():
    """The oracle's build_field (-(1/alpha) Im(<psi|chi><chi|mu|psi>) at the
    mid-steps, CTRL:902-971) equals (dJ/dE_n)/(2 alpha dt) of the oracle's
    exact gradient (pinned by finite differences above) up to the midpoint
    rule's error, which falls as dt^2; a sign error in either fails."""
    rng = np.random.default_rng(3)
    n = 12
    A = rng.standard_normal((n, n))
    H = (A + A.T) / 4 + np.diag(np.linspace(0, 1, n))
    B = rng.standard_normal((n, n))
    M = 0.2 * (B + B.T)
    psi0 = np.zeros(n, complex); psi0[0] = 1.0
    phi1 = np.zeros(n, complex); phi1[1] = 1.0
    T, alpha = 20.0, 0.7
    errs = []
    for N in (200, 400):
        dt = T / N
        E = (0.3 * np.sin(np.pi * (np.arange(N) + 0.5) * dt / T))[:, None]
        _, gE, _ = control.gradient(H, M, psi0, phi1, E, dt)
        bf = control.build_field(H, M, psi0, phi1, E[:, 0], dt, alpha)
        ref = gE[:, 0] / (2 * alpha * dt)
        errs.append(np.abs(bf - ref).max() / np.abs(ref).max())
        assert np.abs(bf + ref).max() > 1.0 * np.abs(ref).max()
    assert errs[1] < 2e-3
    assert 3.5 < errs[0] / errs[1] < 4.5          # the midpoint rule's dt^2
