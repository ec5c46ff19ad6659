"""Chebyshev and SIL propagation on the GPU against the oracle.

The operator is the symmetric (hermitian=1) form of the type-10 SG4 and DP
Hamiltonians; the oracle builds the same matrix densely ((M + M^T)/2 of the
oracle's operator) and propagates with exp(-i H dt) from an eigendecomposition
and with its restatement of march_cheby.  Energies are in cm^-1, so dt is in
units of 1/cm^-1 (dt = 2e-3 gives r = Esc dt of about 13 on the SG4 case).
Bar: psi(t) within 1e-10 (max-norm, psi normalised).
"""
import numpy as np
import pytest

import propa

pytestmark = pytest.mark.gpu


def rand_psi(n, nwp, seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, nwp)) + 1j * rng.standard_normal((n, nwp))
    return p / np.linalg.norm(p, axis=0)


def dipole(q):
    """A linear dipole surface mu(q) = sum_i (0.5 + 0.1 i) q_i (dimensionless)."""
    return (q * (0.5 + 0.1 * np.arange(q.shape[1]))[None, :]).sum(1)


@pytest.fixture(scope="module")
def sg4_case(evr, oracle_mod):
    from evrhip import hamiltonian as Hm
    D, L = 4, 3
    ham = Hm.smolyak_config(1, D)
    s = oracle_mod.SG4(D, L)
    V = ham.potential(s.grid())
    M = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), np.eye(s.nb_packed))
    Hs = 0.5 * (M + M.T)
    mu = dipole(s.grid())
    Mu = s.oppsi(oracle_mod.Op(0, V=mu), np.eye(s.nb_packed))
    b = evr.Basis.sg4(D, L)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=8, hermitian=True)
    mplan = evr.Plan(b, type=0, V=mu, max_nvec=8)
    ev = np.linalg.eigvalsh(Hs)
    return dict(Hs=Hs, Mu=0.5 * (Mu + Mu.T), plan=plan, mplan=mplan, ev=ev, nb=s.nb_packed)


def test_cheby_sg4(evr, sg4_case):
    c = sg4_case
    psi = rand_psi(c["nb"], 3, 1)
    dt, nst = 2e-3, 5
    hmin, hmax = c["ev"][0] - 1.0, c["ev"][-1] + 1.0
    out, ac, info = c["plan"].propagate(psi, dt, nst, method=evr.CHEBY, npoly=200, poly_tol=1e-14,
                                        Hmin=hmin, Hmax=hmax, autocorr=True)
    ex = psi.copy()
    ref_ac = []
    order = 0
    for _ in range(nst):
        _, k, _ = propa.cheby_step(c["Hs"], ex, dt, hmin, hmax, 200, 1e-14)
        order = max(order, k)
        ex = propa.exact_step(c["Hs"], ex, dt)
        ref_ac.append((psi.conj() * ex).sum(0))
    print("cheby orders", info.max_order, order, info.hpsi, np.abs(out - ex).max())
    assert np.abs(out - ex).max() < 1e-10
    assert np.abs(ac - np.array(ref_ac)).max() < 1e-10
    assert info.max_order == order          # same adaptive truncation as march_cheby
    assert info.retries == 0


def test_lanczos_bounds(evr, sg4_case):
    c = sg4_case
    lo, hi = c["plan"].bounds(60)
    print("bounds", lo, hi, c["ev"][0], c["ev"][-1])
    assert abs(lo - c["ev"][0]) < 1e-6 * abs(c["ev"][0])
    assert abs(hi - c["ev"][-1]) < 1e-2 * abs(c["ev"][-1])
    assert hi <= c["ev"][-1] * (1 + 1e-12)


def test_cheby_estimated_interval(evr, sg4_case):
    c = sg4_case
    psi = rand_psi(c["nb"], 2, 2)
    out, _, info = c["plan"].propagate(psi, 2e-3, 3, method=evr.CHEBY, npoly=300, poly_tol=1e-14)
    ex = psi
    for _ in range(3):
        ex = propa.exact_step(c["Hs"], ex, 2e-3)
    assert info.Hmin < c["ev"][0] and info.Hmax > c["ev"][-1]
    assert np.abs(out - ex).max() < 1e-10


def test_sil_sg4(evr, sg4_case):
    c = sg4_case
    psi = rand_psi(c["nb"], 2, 3)
    dt, nst = 2e-4, 4
    out, ac, info = c["plan"].propagate(psi, dt, nst, method=evr.SIL, npoly=40, poly_tol=1e-14,
                                        autocorr=True)
    ex = psi
    ors = 0
    for _ in range(nst):
        _, n = propa.sil_step(c["Hs"], ex, dt, 40, 1e-14)
        ors = max(ors, n)
        ex = propa.exact_step(c["Hs"], ex, dt)
    print("SIL", info.max_order, ors, np.abs(out - ex).max())
    assert np.abs(out - ex).max() < 1e-10
    assert np.abs(ac[-1] - (psi.conj() * ex).sum(0)).max() < 1e-10
    assert abs(info.max_order - ors) <= 1


@pytest.mark.parametrize("method", ["cheby", "sil"])
def test_field_coupling(evr, sg4_case, method):
    """H_n = H - E_n mu, field constant over each step, a different field per wavepacket."""
    c = sg4_case
    nwp, nst = 2, 4
    dt = 2e-3 if method == "cheby" else 2e-4
    psi = rand_psi(c["nb"], nwp, 4)
    rng = np.random.default_rng(5)
    field = rng.uniform(-300.0, 300.0, (nst, nwp))
    m = evr.CHEBY if method == "cheby" else evr.SIL
    out, _, info = c["plan"].propagate(psi, dt, nst, method=m, npoly=200 if m == evr.CHEBY else 40,
                                       poly_tol=1e-14, field=field, mu=c["mplan"])
    ex = psi
    for n in range(nst):
        ex = propa.exact_step(c["Hs"], ex, dt, field[n], c["Mu"])
    print(method, "field", info.max_order, np.abs(out - ex).max())
    assert np.abs(out - ex).max() < 1e-10


def test_cheby_c0_direct_product(evr, oracle_mod):
    from evrhip import hamiltonian as Hm
    ham = Hm.config0()
    nb, nq = [12] * 3, [14] * 3
    d = oracle_mod.DP(nb, nq)
    V = ham.potential(d.grid())
    M = d.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0), np.eye(1728))
    Hs = 0.5 * (M + M.T)
    ev = np.linalg.eigvalsh(Hs)
    plan = evr.Plan(evr.Basis.dp(nb, nq), type=10, V=V, G0=ham.G0, max_nvec=4, hermitian=True)
    psi = rand_psi(1728, 2, 6)
    dt = 1e-3
    out, _, info = plan.propagate(psi, dt, 3, npoly=400, poly_tol=1e-14, Hmin=ev[0] - 1.0,
                                  Hmax=ev[-1] + 1.0)
    ex = psi
    for _ in range(3):
        ex = propa.exact_step(Hs, ex, dt)
    print("C0 cheby", info.max_order, np.abs(out - ex).max())
    assert np.abs(out - ex).max() < 1e-10


@pytest.mark.parametrize("with_field", [False, True])
def test_cheby_graph(evr, sg4_case, with_field):
    """Steps after the first replayed as one captured HIP graph each (fixed
    order from the first step, exit test checked after every replay): the
    same propagation as the adaptive driver, to the exact propagator's bar."""
    c = sg4_case
    nwp, nst, dt = 3, 12, 2e-3
    psi = rand_psi(c["nb"], nwp, 8)
    hmin, hmax = c["ev"][0] - 1.0, c["ev"][-1] + 1.0
    field = np.random.default_rng(9).uniform(-300.0, 300.0, (nst, nwp)) if with_field else None
    mu = c["mplan"] if with_field else None
    kw = dict(method=evr.CHEBY, npoly=300, poly_tol=1e-14, Hmin=hmin, Hmax=hmax, field=field, mu=mu,
              autocorr=True)
    eager, ac_e, ie = c["plan"].propagate(psi, dt, nst, **kw)
    graph, ac_g, ig = c["plan"].propagate(psi, dt, nst, graph=True, **kw)
    ex = psi
    for n in range(nst):
        ex = propa.exact_step(c["Hs"], ex, dt, None if field is None else field[n],
                              None if field is None else c["Mu"])
    print("graph field=%s" % with_field, ig.max_order, ie.max_order, ig.hpsi, ie.hpsi,
          np.abs(graph - eager).max(), np.abs(graph - ex).max())
    assert np.abs(graph - ex).max() < 1e-10
    assert np.abs(graph - eager).max() < 1e-12
    assert np.abs(ac_g - ac_e).max() < 1e-12
    assert ig.max_order == ie.max_order


@pytest.mark.parametrize("method", ["taylor", "rk4"])
def test_time_dependent_field_integrators(evr, sg4_case, method):
    """march_noD_field (Taylor with the field derivatives) and
    march_RK4_field on the GPU with the control field E(t) = sin^2(pi t/T)
    sum_k a_k sin(k pi t/T) (samples from evr_control_field_sub), against the
    oracle's restatement of the same integrator (1e-10) and, for Taylor,
    against the time-dependent solution (fine Magnus, 1e-9)."""
    import control
    c = sg4_case
    nwp, nst = 2, 5
    psi = rand_psi(c["nb"], nwp, 11)
    rng = np.random.default_rng(12)
    amp = rng.uniform(-200.0, 200.0, (6, nwp))      # field in cm^-1 per unit of mu
    T = 0.01
    dt = T / 20 if method == "taylor" else T / 400
    m = evr.TAYLOR if method == "taylor" else evr.RK4
    nsub = 14 if method == "taylor" else 3
    fs = evr.control_field_sub(amp, T, int(round(T / dt)), m, nsub)[:nst]
    ev = np.linalg.eigvalsh(c["Hs"])
    out, _, info = c["plan"].propagate(psi, dt, nst, method=m, npoly=60, poly_tol=1e-26,
                                       Hmin=ev[0], Hmax=ev[-1], mu=c["mplan"], field_sub=fs)
    E0 = 0.5 * (ev[0] + ev[-1])
    ref = psi.copy()
    for n in range(nst):
        if m == evr.TAYLOR:
            ref, _ = propa.taylor_field_step(c["Hs"], ref, dt, E0, 60, 1e-26, dnE=fs[n], M=c["Mu"])
        else:
            ref = propa.rk4_field_step(c["Hs"], ref, dt, E0, fs[n], c["Mu"])
    err = np.abs(out - ref).max()
    print(method, "orders", info.max_order, "vs oracle", err)
    assert err < 1e-10
    if m == evr.TAYLOR:
        for w in range(nwp):
            f = lambda t: control.field_derivs(amp[:, w], T, t, 0)[0, 0]
            ex = propa.magnus4_propagate(c["Hs"], psi[:, w:w + 1], 0.0, nst * dt, 400, f, c["Mu"])[:, 0]
            assert np.abs(out[:, w] - ex).max() < 1e-9
