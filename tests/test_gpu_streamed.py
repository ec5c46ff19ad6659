"""Terms beyond one workgroup's LDS (more than 8192 basis functions, or
slabs that do not fit) run as streamed passes over global memory: the
reference's per-term basis -> grid, grid operator and grid -> basis sequence
(BGSG4:2385-2581, OPSG4:1448-1650), one mode at a time.  Checked against the
oracle on direct-product bases of 3 x 22 functions (10,648 > 8192; type 10
with q-linear G, type 1 with F2 and F1, FP64 and FP32), on an SG4 basis
whose four level-(1,1,1,0) terms have 13^3 x 5 = 10,985 functions beside
LDS-resident terms (H|psi>, basis <-> grid, the diagonal), on the hermitian
form, through block Davidson on a coupled harmonic direct-product operator
whose levels are known in closed form, through Chebyshev propagation
(eager and replayed as captured HIP graphs) against march_cheby on the CPU,
and through the optimal-control objective and its adjoint gradient."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _dp(evr, oracle_mod, n=22):
    nb, nq = [n] * 3, [n + 1] * 3
    return oracle_mod.DP(nb, nq), evr.Basis.dp(nb, nq)


def test_dp_streamed_type10(evr, oracle_mod):
    from evrhip import hamiltonian as H
    d, b = _dp(evr, oracle_mod)
    assert b.nb_packed == 22 ** 3 > 8192
    ham = H.smolyak_config(9, 3)
    V = ham.potential(d.grid())
    X = H.test_vectors(b.nb_packed, 3, 11)
    ref = d.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), X)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=3)
    e64 = relerr(plan.TabOpPsi(X), ref)
    e32 = relerr(plan.TabOpPsi(X, flags=evr.FP32), ref)
    e1 = relerr(plan.TabOpPsi(X[:, 1:2]), ref[:, 1:2])
    print("DP 3 x 22 type 10 streamed: FP64 %.2e, FP32 %.2e, one vector %.2e" % (e64, e32, e1))
    assert e64 < TOL64 and e1 < TOL64 and e32 < 2e-5


def test_dp_streamed_type1(evr, oracle_mod):
    from evrhip import hamiltonian as H
    d, b = _dp(evr, oracle_mod)
    q = d.grid()
    V = 0.5 * (q ** 2 * np.array([1.0, 1.3, 1.7])).sum(axis=1) + 0.02 * q[:, 0] * q[:, 1] * q[:, 2]
    F2 = -0.5 * np.diag([1.0, 1.3, 1.7]) + 0.03 * (np.ones((3, 3)) - np.eye(3))
    F2[0, 2] = 0.05
    F1 = np.array([0.02, -0.01, 0.015])
    X = H.test_vectors(b.nb_packed, 2, 12)
    ref = d.oppsi(oracle_mod.Op(1, V=V, F2=F2, F1=F1), X)
    plan = evr.Plan(b, type=1, V=V, F2=F2, F1=F1, max_nvec=2)
    e = relerr(plan.TabOpPsi(X), ref)
    print("DP 3 x 22 type 1 streamed: %.2e" % e)
    assert e < TOL64


def test_dp_streamed_hermitian(evr, oracle_mod):
    """hermitian = 1 evaluates (H + H^T) / 2: X^T H_sym X equals the
    symmetric part of X^T H X with the oracle's H, on 4 vectors."""
    from evrhip import hamiltonian as H
    d, b = _dp(evr, oracle_mod)
    ham = H.smolyak_config(9, 3)
    V = ham.potential(d.grid())
    X = H.test_vectors(b.nb_packed, 4, 13)
    M = X.T @ d.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), X)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=4, hermitian=True)
    Ms = X.T @ plan.TabOpPsi(X)
    e = np.abs(Ms - 0.5 * (M + M.T)).max() / np.abs(M).max()
    asym = np.abs(M - M.T).max() / np.abs(M).max()
    print("hermitian streamed: %.2e (asymmetry of the reference form %.2e)" % (e, asym))
    assert e < TOL64


def _sg4_big(evr, oracle_mod):
    # nb(l) = 5 + 8 l, nq(l) = 6 + 8 l: D=4 L=3 holds the level-(1,1,1,0)
    # terms of 13^3 x 5 = 10,985 functions beside LDS-resident terms
    args = dict(nbA=5, nbB=8, nqA=6, nqB=8)
    s = oracle_mod.SG4(4, 3, **args)
    b = evr.Basis.sg4(4, 3, **args)
    lv, _, onb, _ = s.terms()
    assert np.diff(onb).max() == 10985 and (np.diff(onb) > 8192).sum() == 4
    return s, b


def test_sg4_streamed_term_oppsi_and_transforms(evr, oracle_mod):
    from evrhip import hamiltonian as H
    s, b = _sg4_big(evr, oracle_mod)
    ham = H.smolyak_config(9, 4)
    V = ham.potential(s.grid())
    X = H.test_vectors(b.nb_packed, 3, 14)
    ref = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), X)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=64)
    e = relerr(plan.TabOpPsi(X), ref)
    x = X[:, 0]
    g = plan.BasisRep_TO_GridRep(x)
    eg = relerr(g, s.b2g(x))
    back = plan.GridRep_TO_BasisRep(g)
    eb = relerr(back, s.g2b(g))
    print("SG4 with a streamed term: H|psi> %.2e, B->G %.2e, G->B %.2e" % (e, eg, eb))
    assert e < TOL64 and eg < TOL64 and eb < TOL64
    assert relerr(back, x) < 1e-12
    # the plan diagonal (read off H on blocks of 64 unit vectors with
    # streamed terms)
    dg = plan.diagonal()
    idx = np.array([0, 7, 12345, b.nb_packed // 2, b.nb_packed - 1])
    E = np.zeros((b.nb_packed, len(idx)))
    E[idx, np.arange(len(idx))] = 1.0
    dref = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), E)[idx, np.arange(len(idx))]
    assert np.abs(dg[idx] - dref).max() < 1e-12 * np.abs(dref).max()


def test_streamed_plan_refusals(evr, oracle_mod):
    s, b = _sg4_big(evr, oracle_mod)
    plan = evr.Plan(b, type=10, V=np.ones(b.sum_nq), G0=np.eye(4), max_nvec=1)
    with pytest.raises(RuntimeError, match="streamed over global memory"):
        plan.assemble()
    with pytest.raises(RuntimeError, match="at most 8192 functions"):
        plan.set_keo_grid(np.broadcast_to(np.eye(4), (b.sum_nq, 4, 4)))


def test_dp_streamed_davidson_harmonic(evr):
    """Block Davidson through the streamed path: the 8 lowest levels of the
    coupled harmonic operator -1/2 sum_i w_i d^2/dq_i^2 + q^T K q / 2 on
    3 x 22 Hermite functions are sum_k W_k (n_k + 1/2), W^2 the eigenvalues
    of diag(w)^1/2 K diag(w)^1/2 (the same closed form holds to 1e-13 for
    the oracle's dense spectrum on 3 x 12 functions)."""
    w = np.array([1.0, 1.3, 1.7])
    K = np.diag(w)
    K[0, 1] = K[1, 0] = 0.15
    K[1, 2] = K[2, 1] = -0.1
    wh = np.diag(np.sqrt(w))
    om = np.sqrt(np.linalg.eigvalsh(wh @ K @ wh))
    levels = np.array(sorted(float(np.dot(om, np.array(n) + 0.5))
                             for n in itertools.product(range(8), repeat=3))[:8])
    b = evr.Basis.dp([22] * 3, [23] * 3)
    q = b.grid()
    V = 0.5 * np.einsum("pi,ij,pj->p", q, K, q)
    plan = evr.Plan(b, type=1, V=V, F2=-0.5 * np.diag(w), max_nvec=16)
    ev, _, info = plan.davidson(8, block=8, conv_resi=1e-9, conv_ene=1e-11)
    print("coupled harmonic DP Davidson", info.iterations, np.abs(ev - levels).max())
    assert info.converged == 1
    assert np.abs(ev - levels).max() < 1e-9


def test_dp_streamed_chebyshev_graph(evr, oracle_mod):
    """Chebyshev propagation through the streamed path (the hermitian form),
    eager and with the steps after the first replayed as captured HIP graphs,
    against march_cheby restated on the CPU with the oracle's H: 3 steps of
    two wavepackets on the coupled harmonic operator of 3 x 22 functions."""
    import propa
    w = np.array([1.0, 1.3, 1.7])
    K = np.diag(w)
    K[0, 1] = K[1, 0] = 0.15
    K[1, 2] = K[2, 1] = -0.1
    d, b = _dp(evr, oracle_mod)
    q = d.grid()
    V = 0.5 * np.einsum("pi,ij,pj->p", q, K, q)
    F2 = -0.5 * np.diag(w)
    plan = evr.Plan(b, type=1, V=V, F2=F2, max_nvec=4, hermitian=True)
    lo, hi = plan.bounds()
    hmin, hmax = lo - 1.0, hi + 1.0
    rng = np.random.default_rng(21)
    psi = rng.standard_normal((b.nb_packed, 2)) + 1j * rng.standard_normal((b.nb_packed, 2))
    psi /= np.linalg.norm(psi, axis=0)
    dt, nst = 0.05, 3
    kw = dict(npoly=400, poly_tol=1e-14, Hmin=hmin, Hmax=hmax)
    eager, _, ie = plan.propagate(psi, dt, nst, **kw)
    graph, _, ig = plan.propagate(psi, dt, nst, graph=True, **kw)
    op = oracle_mod.Op(1, V=V, F2=F2)

    class Hop:
        def __matmul__(self, x):
            return d.oppsi(op, np.ascontiguousarray(x.real)) + 1j * d.oppsi(op, np.ascontiguousarray(x.imag))

    ref = psi
    for _ in range(nst):
        ref, order, _ = propa.cheby_step(Hop(), ref, dt, hmin, hmax, 400, 1e-14)
    print("streamed Chebyshev: order %d, eager %.2e, graph %.2e"
          % (order, np.abs(eager - ref).max(), np.abs(graph - ref).max()))
    assert np.abs(eager - ref).max() < 1e-10
    assert np.abs(graph - ref).max() < 1e-10
    assert np.abs(np.linalg.norm(eager, axis=0) - 1.0).max() < 1e-10


def test_dp_streamed_control_gradient(evr, oracle_mod):
    """The optimal-control objective and adjoint gradient through a streamed
    one-term plan (DP 3 x 8, type 10, hermitian; the dipole a type-0 plan on
    the same basis): J and dJ/da against the oracle's exact propagation with
    the Daleckii-Krein derivative (1e-9), dJ/dV against central differences
    of the GPU objective at two grid points."""
    import control
    from evrhip import hamiltonian as Hm
    T, N, NK = 400.0, 20, 4
    d = oracle_mod.DP([8] * 3, [9] * 3)
    b = evr.Basis.dp([8] * 3, [9] * 3)
    ham = Hm.smolyak_config(1, 3)
    q = d.grid()
    V = ham.potential(q) / Hm.AU_TO_CM
    G0, Ga = ham.G0 / Hm.AU_TO_CM, ham.Ga_fortran() / Hm.AU_TO_CM
    M = d.oppsi(oracle_mod.Op(10, V=V, G0=G0, Ga=Ga), np.eye(d.nb_tot))
    H = 0.5 * (M + M.T)
    mu = 0.1 * q[:, 0] + 0.05 * q[:, 1] - 0.03 * q[:, 0] * q[:, 2]
    Mu = d.oppsi(oracle_mod.Op(0, V=mu), np.eye(d.nb_tot))
    Mu = 0.5 * (Mu + Mu.T)
    ev, Vec = np.linalg.eigh(H)
    psi0, phi1 = Vec[:, 0].astype(complex), Vec[:, 1].astype(complex)
    amp = np.random.default_rng(11).uniform(-0.02, 0.02, (NK, 2))

    def run(Vg, want):
        plan = evr.Plan(b, type=10, V=Vg, G0=G0, Ga=Ga, max_nvec=8, hermitian=True)
        mplan = evr.Plan(b, type=0, V=mu, max_nvec=8)
        return plan.control_gradient(mplan, amp, psi0, phi1, T, N, checkpoint=8, npoly=100,
                                     poly_tol=1e-15, want=want)

    J, g, _ = run(V, ("a", "V"))
    E = control.field_values(amp, T, N)
    Jr, gEr, _ = control.gradient(H, Mu, psi0, phi1, E, T / N)
    gar = control.field_basis(T, N, NK).T @ gEr
    h = 1e-6
    fd = []
    for p in (100, 364):
        Vp, Vm = V.copy(), V.copy()
        Vp[p] += h
        Vm[p] -= h
        fd.append((run(Vp, ())[0] - run(Vm, ())[0]) / (2 * h))
    fd = np.array(fd)
    print("streamed control: J %.2e, dJ/da %.2e, dJ/dV fd %.2e" % (
        np.abs(J - Jr).max(), np.abs(g["a"] - gar).max() / np.abs(gar).max(),
        np.abs(fd - g["V"][[100, 364]]).max() / np.abs(g["V"]).max()))
    assert np.abs(J - Jr).max() < 1e-12
    assert np.abs(g["a"] - gar).max() < 1e-9 * np.abs(gar).max()
    assert np.abs(fd - g["V"][[100, 364]]).max() < 1e-6 * np.abs(g["V"]).max()
