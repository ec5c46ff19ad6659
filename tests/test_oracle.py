"""The oracle pinned against the reference's own data and against mathematics.

Fixtures (tests/golden/*.json, written by tests/golden/make_golden.py from the
reference's data files):
  * Internal_data/HermQuadra/herm{nq}.txt: Gauss-Hermite nodes and w*exp(x^2);
  * Working_tests/MPI_tests/{6D,21D}_Davidson_S0/benchmark: Henon-Heiles SG4
    Davidson levels (Type_HamilOp=1, conv_resi=5e-4 au).
"""
import itertools
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, hh_potential


def test_quadrature_matches_reference_tables(oracle_mod):
    tab = json.load(open(os.path.join(GOLDEN, "herm_tables.json")))
    for nq in range(1, 41):
        x, w = oracle_mod.herm_quadrature(nq)
        xt = np.array([float(v) for v in tab[str(nq)]["x"]])
        wt = np.array([float(v) for v in tab[str(nq)]["wex2"]])
        assert np.abs(x - xt).max() <= 4e-16 * max(1.0, np.abs(xt).max()), nq
        assert (np.abs(w - wt) / wt).max() <= 4e-16, nq


@pytest.mark.parametrize("nb,nq", [(1, 2), (3, 4), (9, 10), (17, 18), (12, 14), (5, 5)])
def test_hermite_basis_orthonormal_and_derivatives(oracle_mod, nb, nq):
    x, w = oracle_mod.herm_quadrature(nq)
    B, B1, B2 = oracle_mod.herm_basis(nb, x)
    S = B.T @ np.diag(w) @ B                      # W B = I (exact quadrature)
    assert np.abs(S - np.eye(nb)).max() < 1e-13
    # closed forms of the Hermite functions: ladder relations
    from numpy.polynomial.hermite import hermval
    from math import factorial, pi, sqrt
    for n in range(nb):
        c = np.zeros(n + 1); c[n] = 1.0
        phi = hermval(x, c) * np.exp(-x * x / 2) / sqrt(2.0 ** n * factorial(n) * sqrt(pi))
        assert np.abs(B[:, n] - phi).max() < 1e-13
        assert np.abs(B2[:, n] - (x * x - 2 * n - 1) * phi).max() < 1e-12
    D0, D1, D2 = oracle_mod.herm_dngg(B, B1, B2)
    assert np.abs(D0 @ B - B).max() < 1e-12           # BAS:2399-2440 self-check
    assert np.abs(D1 @ B - B1).max() < 1e-12
    assert np.abs(D2 @ B - B2).max() < 1e-11


@pytest.mark.parametrize("D,L,sizes", [
    (6, 5, (462, 3653, 25954, 396032, 2048)),
    (9, 8, (24310, 598417, 14666470, 799884800, 131072)),
    (12, 6, (18564, 369305, 4195284, 2432102400, 262144)),
])
def test_smolyak_sizes(oracle_mod, D, L, sizes):
    s = oracle_mod.SG4(D, L)
    assert (s.nterms, s.nb_packed, s.sum_nb, s.sum_nq, s.max_nq) == sizes
    _, w, _, _ = s.terms()
    assert abs(w.sum() - 1.0) < 1e-12                 # Smolyak weights sum to 1


def test_smolyak_brute_force_small(oracle_mod):
    """Terms, weights, packed basis and map by brute force at D=3, L=3."""
    D, L = 3, 3
    s = oracle_mod.SG4(D, L)
    lv, w, onb, _ = s.terms()
    Lmin = max(0, L - D + 1)                         # SPB:994
    want = [l for l in itertools.product(range(L + 1), repeat=D) if Lmin <= sum(l) <= L]
    want.sort(key=lambda l: tuple(reversed(l)))       # first mode fastest
    assert [tuple(r) for r in lv] == want
    from math import comb
    for l, wt in zip(want, w):
        dl = L - sum(l)
        assert wt == ((-1) ** dl * comb(D - 1, dl) if dl <= D - 1 else 0)
    lev = lambda n: (n + 1) // 2                       # nb(l) = 1 + 2l
    packed = [t for t in itertools.product(range(2 * L + 1), repeat=D)
              if sum(lev(n) for n in t) <= L]          # last mode fastest
    assert [tuple(r) for r in s.packed()] == packed
    idx = {t: i for i, t in enumerate(packed)}
    m = s.map()
    for t, l in enumerate(want):
        nbs = [1 + 2 * li for li in l]
        for e, tup in enumerate(itertools.product(*[range(n) for n in reversed(nbs)])):
            assert m[onb[t] + e] == idx[tuple(reversed(tup))]


def test_harmonic_closed_form(oracle_mod):
    """Type 1 with F2 = -1/2 diag(w) and V = sum w_i q_i^2/2: with nq = nb + 1 the
    quadrature integrates every term matrix exactly, so the SG4 spectrum is
    sum_i w_i (n_i + 1/2) over the packed space, exactly."""
    D, L = 4, 4
    s = oracle_mod.SG4(D, L)
    w = np.array([900.0, 1300.0, 1700.0, 2500.0])
    V = 0.5 * (s.grid() ** 2) @ w
    H = s.oppsi(oracle_mod.Op(1, V=V, F2=-0.5 * np.diag(w)), np.eye(s.nb_packed))
    lev = lambda n: (n + 1) // 2
    P = s.packed()
    exact = np.array([sum(wi * (ni + 0.5) for wi, ni in zip(w, t)) for t in P])
    assert np.abs(H - np.diag(exact)).max() < 1e-9 * exact.max()
    assert len(P) == len([t for t in itertools.product(range(2 * L + 1), repeat=D)
                          if sum(lev(n) for n in t) <= L])


def test_type10_low_levels_near_harmonic(oracle_mod):
    """Type 10 (d G d with dnRGG%d1 = B' pinv(B)) is not exact at the top
    function of each term, but the low levels of a harmonic system stay close to
    sum_i w_i (n_i + 1/2)."""
    D, L = 4, 4
    s = oracle_mod.SG4(D, L)
    w = np.array([900.0, 1300.0, 1700.0, 2500.0])
    V = 0.5 * (s.grid() ** 2) @ w
    H = s.oppsi(oracle_mod.Op(10, V=V, G0=np.diag(w)), np.eye(s.nb_packed))
    e = np.sort(np.linalg.eigvals(H).real)
    assert abs(e[0] - 0.5 * w.sum()) < 1e-3
    assert abs(e[1] - (0.5 * w.sum() + w[0])) < 1e-2


def test_smolyak_matches_dense_kronecker(oracle_mod):
    """D=2: the SG4 type-10 operator equals the combination of per-term dense
    Kronecker operators built independently from the 1D matrices."""
    D, L = 2, 3
    s = oracle_mod.SG4(D, L)
    rng = np.random.default_rng(5)
    G0 = np.array([[1.0, 0.1], [0.1, 1.3]])
    Ga = rng.uniform(-0.02, 0.02, (D, D, D)); Ga = (Ga + Ga.transpose(1, 0, 2)) / 2
    q = s.grid()
    V = 0.5 * (q ** 2).sum(1) + 0.1 * q[:, 0] * q[:, 1] ** 2
    H = s.oppsi(oracle_mod.Op(10, V=V, G0=G0, Ga=Ga.ravel(order="F")), np.eye(s.nb_packed))
    lv, w, onb, onq = s.terms()
    m = s.map()
    Hd = np.zeros((s.nb_packed, s.nb_packed))
    for t in range(s.nterms):
        one = [s.one_d(d, lv[t, d]) for d in range(D)]
        # first mode fastest: kron(mode1, mode0)
        K = lambda A0, A1: np.kron(A1, A0)
        Bt = K(one[0]["B"], one[1]["B"]); Wt = K(one[0]["W"], one[1]["W"])
        I0, I1 = np.eye(one[0]["nq"]), np.eye(one[1]["nq"])
        Dm = [K(one[0]["D1"], I1), K(I0, one[1]["D1"])]
        qt = q[onq[t]:onq[t + 1]]
        op = np.diag(V[onq[t]:onq[t + 1]]).copy()
        for i in range(D):
            for j in range(D):
                g = G0[j, i] + qt @ Ga[j, i, :]
                op += -0.5 * Dm[i] @ np.diag(g) @ Dm[j]
        Ht = Wt @ op @ Bt
        rows = m[onb[t]:onb[t + 1]]
        Hd[np.ix_(rows, rows)] += w[t] * Ht
    assert np.abs(H - Hd).max() < 1e-11 * np.abs(Hd).max()


def test_smolyak_type10_d3_generator_matches_dense_kronecker(oracle_mod):
    """D=3 (the smallest D with three distinct modes in G_ij(q) = G0_ij +
    sum_k a_ijk q_k, the coupling the benchmark operator carries): the SG4
    type-10 operator with the generator potential evaluated per grid point
    (Vpoly) equals (i) the same operator given V explicitly and (ii) the
    combination of per-term dense Kronecker operators built from the 1D
    matrices, V from the host formula.  A transposed Ga index, a wrong
    potential coefficient or a dropped cross term fails (ii)."""
    from evrhip import hamiltonian as Hm
    D, L = 3, 3
    s = oracle_mod.SG4(D, L)
    ham = Hm.smolyak_config(1, D)
    q = s.grid()
    V = ham.potential(q)
    Ga = ham.Ga.copy()
    Ga[0, 1, 0] += 3.0           # break the symmetry: an (i, j) transposition shows
    Gaf = Ga.ravel(order="F")
    cf = oracle_mod.poly_coefs(ham.v2, ham.c3, ham.d22, ham.e4)
    X = np.eye(s.nb_packed)
    Hp = s.oppsi(oracle_mod.Op(10, G0=ham.G0, Ga=Gaf, Vpoly=cf), X)
    Hv = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=Gaf), X)
    assert np.abs(Hp - Hv).max() < 1e-12 * np.abs(Hv).max()
    lv, w, onb, onq = s.terms()
    m = s.map()
    Hd = np.zeros((s.nb_packed, s.nb_packed))
    Hx = np.zeros_like(Hd)
    for t in range(s.nterms):
        one = [s.one_d(d, lv[t, d]) for d in range(D)]
        K = lambda A0, A1, A2: np.kron(A2, np.kron(A1, A0))     # first mode fastest
        Bt = K(*[o["B"] for o in one]); Wt = K(*[o["W"] for o in one])
        I = [np.eye(o["nq"]) for o in one]
        Dm = [K(one[0]["D1"], I[1], I[2]), K(I[0], one[1]["D1"], I[2]), K(I[0], I[1], one[2]["D1"])]
        qt = q[onq[t]:onq[t + 1]]
        op = np.diag(V[onq[t]:onq[t + 1]]).copy()
        opx = op.copy()
        for i in range(D):
            for j in range(D):
                op += -0.5 * Dm[i] @ np.diag(ham.G0[j, i] + qt @ Ga[j, i, :]) @ Dm[j]
                opx += -0.5 * Dm[i] @ np.diag(ham.G0[i, j] + qt @ Ga[i, j, :]) @ Dm[j]
        rows = m[onb[t]:onb[t + 1]]
        Hd[np.ix_(rows, rows)] += w[t] * (Wt @ op @ Bt)
        Hx[np.ix_(rows, rows)] += w[t] * (Wt @ opx @ Bt)
    assert np.abs(Hp - Hd).max() < 1e-11 * np.abs(Hd).max()
    assert np.abs(Hp - Hx).max() > 1e-6 * np.abs(Hd).max()


def tabulated_keo(q, seed):
    """A per-point G (not symmetric in i, j, so a transposed index shows), a
    Jacobian and s = sqrt(rho/J) on the grid points q."""
    rng = np.random.default_rng(seed)
    npt, D = q.shape
    G0 = np.diag(rng.uniform(0.8, 1.4, D)) + 0.05 * rng.standard_normal((D, D))
    A = 0.03 * rng.standard_normal((D, D, D))
    Gq = G0[None] + np.einsum("pk,ijk->pij", q, A) + 0.01 * (q ** 2).sum(1)[:, None, None]
    J = 1.0 + 0.1 * np.tanh(q @ rng.standard_normal(D))
    s = 1.0 / np.sqrt(J) * (1.0 + 0.05 * np.cos(q[:, 0]))
    return Gq, J, s   # Gq[p, j, i] multiplies d_j in the i-th flux


def test_smolyak_tabulated_keo_matches_dense_kronecker(oracle_mod):
    """D=2, type 10 with G, J and s per grid point: -1/2 (1/(J s)) sum_i d_i
    [J sum_j G(:,j,i) d_j (s psi)] + V psi per term against dense Kronecker
    operators; the same with G(:,i,j) in place of G(:,j,i) differs."""
    D, L = 2, 3
    s = oracle_mod.SG4(D, L)
    q = s.grid()
    Gq, J, sq = tabulated_keo(q, 11)
    V = 0.5 * (q ** 2).sum(1)
    H = s.oppsi(oracle_mod.Op(10, V=V, Jac=J, s=sq, Gq=Gq), np.eye(s.nb_packed))
    lv, w, onb, onq = s.terms()
    m = s.map()
    Hd = np.zeros((s.nb_packed, s.nb_packed))
    Hx = np.zeros_like(Hd)
    for t in range(s.nterms):
        one = [s.one_d(d, lv[t, d]) for d in range(D)]
        K = lambda A0, A1: np.kron(A1, A0)
        Bt = K(one[0]["B"], one[1]["B"]); Wt = K(one[0]["W"], one[1]["W"])
        I0, I1 = np.eye(one[0]["nq"]), np.eye(one[1]["nq"])
        Dm = [K(one[0]["D1"], I1), K(I0, one[1]["D1"])]
        sl = slice(onq[t], onq[t + 1])
        Jt, st = J[sl], sq[sl]
        kin = sum(Dm[i] @ np.diag(Jt * Gq[sl, j, i]) @ Dm[j] for i in range(D) for j in range(D))
        kinx = sum(Dm[i] @ np.diag(Jt * Gq[sl, i, j]) @ Dm[j] for i in range(D) for j in range(D))
        op = np.diag(V[sl]) - 0.5 * np.diag(1.0 / (Jt * st)) @ kin @ np.diag(st)
        opx = np.diag(V[sl]) - 0.5 * np.diag(1.0 / (Jt * st)) @ kinx @ np.diag(st)
        rows = m[onb[t]:onb[t + 1]]
        Hd[np.ix_(rows, rows)] += w[t] * (Wt @ op @ Bt)
        Hx[np.ix_(rows, rows)] += w[t] * (Wt @ opx @ Bt)
    assert np.abs(H - Hd).max() < 1e-11 * np.abs(Hd).max()
    assert np.abs(H - Hx).max() > 1e-4 * np.abs(Hd).max()


def type1_grids(q, seed):
    """Per-point type-1 operator grids: F2[p, i, j] (not symmetric, so a
    transposed index shows) multiplying d_i d_j, and F1[p, i] multiplying d_i."""
    rng = np.random.default_rng(seed)
    npt, D = q.shape
    A = -0.5 * np.diag(rng.uniform(0.8, 1.3, D)) + 0.05 * rng.standard_normal((D, D))
    B = 0.02 * rng.standard_normal((D, D, D))
    F2 = A[None] + np.einsum("pk,ijk->pij", q, B)
    F1 = 0.1 * np.tanh(q @ rng.standard_normal((D, D)))
    return F2, F1


def test_smolyak_type1_grids_match_dense_kronecker(oracle_mod):
    """D=2, type 1 with the operator grids per point (get_OpGrid_type1,
    OPSG4:1490-1546): sum_ij F2(:,i,j) d_i d_j psi + sum_i F1(:,i) d_i psi + V psi
    per term, d_i d_i with dnRGG%d2 and d_i d_j with d1 on both modes, against
    dense Kronecker operators (d1 d1 in place of d2, or F1 on the wrong mode,
    differs); a constant table equals the constant-F2 path."""
    D, L = 2, 3
    s = oracle_mod.SG4(D, L)
    q = s.grid()
    F2, F1 = type1_grids(q, 21)
    V = 0.5 * (q ** 2).sum(1)
    I_ = np.eye(s.nb_packed)
    H = s.oppsi(oracle_mod.Op(1, V=V, F2q=F2, F1q=F1), I_)
    lv, w, onb, onq = s.terms()
    m = s.map()
    Hd = np.zeros((s.nb_packed, s.nb_packed))
    Hx = np.zeros_like(Hd)
    for t in range(s.nterms):
        one = [s.one_d(d, lv[t, d]) for d in range(D)]
        K = lambda A0, A1: np.kron(A1, A0)
        Bt = K(one[0]["B"], one[1]["B"]); Wt = K(one[0]["W"], one[1]["W"])
        I0, I1 = np.eye(one[0]["nq"]), np.eye(one[1]["nq"])
        D1 = [K(one[0]["D1"], I1), K(I0, one[1]["D1"])]
        D2 = [K(one[0]["D2"], I1), K(I0, one[1]["D2"])]
        sl = slice(onq[t], onq[t + 1])
        op = np.diag(V[sl]).copy()
        opx = op.copy()
        for i in range(D):
            op += np.diag(F1[sl, i]) @ D1[i]
            opx += np.diag(F1[sl, D - 1 - i]) @ D1[i]        # mistake: F1 mode index swapped
            for j in range(D):
                Dij = D2[i] if i == j else D1[i] @ D1[j]
                op += np.diag(F2[sl, i, j]) @ Dij
                opx += np.diag(F2[sl, i, j]) @ (D1[i] @ D1[j])  # mistake: d1 d1 for d2
        rows = m[onb[t]:onb[t + 1]]
        Hd[np.ix_(rows, rows)] += w[t] * (Wt @ op @ Bt)
        Hx[np.ix_(rows, rows)] += w[t] * (Wt @ opx @ Bt)
    assert np.abs(H - Hd).max() < 1e-11 * np.abs(Hd).max()
    assert np.abs(H - Hx).max() > 1e-4 * np.abs(Hd).max()
    F2c = np.array([[-0.6, 0.04], [0.03, -0.45]])
    Hc = s.oppsi(oracle_mod.Op(1, V=V, F2=F2c, F1=np.array([0.1, -0.2])), I_)
    Ht = s.oppsi(oracle_mod.Op(1, V=V, F2q=np.broadcast_to(F2c, (s.sum_nq, D, D)),
                               F1q=np.broadcast_to([0.1, -0.2], (s.sum_nq, D))), I_)
    assert np.abs(Hc - Ht).max() < 1e-13 * np.abs(Hc).max()


def test_tabulated_g_equals_polynomial(oracle_mod):
    """G0 + Ga q tabulated on the grid gives the polynomial operator to rounding."""
    D, L = 4, 3
    s = oracle_mod.SG4(D, L)
    q = s.grid()
    rng = np.random.default_rng(2)
    G0 = np.diag([1.0, 1.2, 0.9, 1.1])
    Ga = rng.uniform(-0.02, 0.02, (D, D, D))
    Gq = G0[None] + np.einsum("pk,ijk->pij", q, Ga)
    X = rng.standard_normal((s.nb_packed, 3))
    V = 0.5 * (q ** 2).sum(1)
    a = s.oppsi(oracle_mod.Op(10, V=V, G0=G0, Ga=Ga.ravel(order="F")), X)
    b = s.oppsi(oracle_mod.Op(10, V=V, Gq=Gq), X)
    assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()


def _hh_levels(oracle_mod, D, L):
    s = oracle_mod.SG4(D, L, L, nbA=1, nbB=2, nqA=1, nqB=2)
    V = hh_potential(s.grid())
    H = s.oppsi(oracle_mod.Op(1, V=V, F2=-0.5 * np.eye(D)), np.eye(s.nb_packed))
    return np.linalg.eigvalsh(0.5 * (H + H.T))


@pytest.mark.parametrize("name,D,L", [("6D_Davidson_S0", 6, 3), ("21D_Davidson_S0", 21, 2)])
def test_henon_heiles_reference_levels(oracle_mod, name, D, L):
    """The reference's Davidson levels are Ritz values converged to residual
    5e-4 au: they lie above the exact levels of the same discrete operator by at
    most ~r^2/gap.  A wrong term, weight, map or derivative matrix moves the
    levels by far more than that."""
    bench = np.array(json.load(open(os.path.join(GOLDEN, "hh_benchmarks.json")))[name])
    e = _hh_levels(oracle_mod, D, L)[:len(bench)]
    diff = bench - e
    assert diff.min() > -1e-9
    assert diff.max() < 2e-7


def test_dp_matches_dense(oracle_mod):
    """Direct product, D=2: the DP type-10 operator equals the dense form."""
    nb, nq = [4, 5], [6, 7]
    d = oracle_mod.DP(nb, nq)
    q = d.grid()
    G0 = np.array([[1.2, 0.2], [0.2, 0.8]])
    V = 0.5 * (q ** 2).sum(1) + 0.05 * q[:, 0] ** 3
    H = d.oppsi(oracle_mod.Op(10, V=V, G0=G0), np.eye(d.nb_tot))
    one = []
    for k in range(2):
        x, w = oracle_mod.herm_quadrature(nq[k])
        B, B1, B2 = oracle_mod.herm_basis(nb[k], x)
        _, D1, _ = oracle_mod.herm_dngg(B, B1, B2)
        one.append((B, np.diag(w) @ B, D1))
    # last mode fastest: kron(mode0, mode1)
    Bt = np.kron(one[0][0], one[1][0]); Wt = np.kron(one[0][1], one[1][1]).T
    Dm = [np.kron(one[0][2], np.eye(nq[1])), np.kron(np.eye(nq[0]), one[1][2])]
    op = np.diag(V) - 0.5 * sum(G0[j, i] * Dm[i] @ Dm[j] for i in range(2) for j in range(2))
    assert np.abs(H - Wt @ op @ Bt).max() < 1e-11 * np.abs(H).max()


def test_dp_ga_and_type1_match_dense(oracle_mod):
    """Direct product, D=3 (last mode fastest): the DP type-10 operator with
    G(q) = G0 + Ga q and the DP type-1 operator with F2 (off-diagonal pairs
    included) and F1 equal their dense Kronecker forms; these are the DP
    branches the streamed-term GPU tests compare against."""
    nb, nq = [3, 4, 3], [5, 5, 4]
    d = oracle_mod.DP(nb, nq)
    q = d.grid()
    D = 3
    rng = np.random.default_rng(5)
    G0 = np.diag([1.1, 0.9, 1.3]) + 0.05 * rng.standard_normal((D, D))
    Ga = 0.02 * rng.standard_normal((D, D, D))
    V = 0.5 * (q ** 2).sum(1) + 0.04 * q[:, 0] * q[:, 1] * q[:, 2]
    one = []
    for k in range(D):
        x, w = oracle_mod.herm_quadrature(nq[k])
        B, B1, B2 = oracle_mod.herm_basis(nb[k], x)
        _, D1, D2 = oracle_mod.herm_dngg(B, B1, B2)
        one.append((B, np.diag(w) @ B, D1, D2))

    def kron3(a, b, c):
        return np.kron(a, np.kron(b, c))

    Bt = kron3(one[0][0], one[1][0], one[2][0])
    Wt = kron3(one[0][1], one[1][1], one[2][1]).T
    I = [np.eye(n) for n in nq]
    d1 = [kron3(*(one[m][2] if m == k else I[m] for m in range(D))) for k in range(D)]
    d2 = [kron3(*(one[m][3] if m == k else I[m] for m in range(D))) for k in range(D)]
    # type 10: Ga in the Fortran layout Ga(i, j, m) = flat[i + D (j + D m)]
    Gf = Ga.reshape(-1, order="F")
    H = d.oppsi(oracle_mod.Op(10, V=V, G0=G0.reshape(-1, order="F"), Ga=Gf), np.eye(d.nb_tot))
    op = np.diag(V)
    for i in range(D):
        for j in range(D):
            gij = G0[i, j] + q @ Ga[i, j, :]
            op = op - 0.5 * d1[j] @ np.diag(gij) @ d1[i]
    assert np.abs(H - Wt @ op @ Bt).max() < 1e-11 * np.abs(H).max()
    # type 1: F2(i, j) d_i d_j (d2 on the diagonal) + F1(i) d_i
    F2 = -0.5 * np.diag([1.0, 1.3, 1.7]) + 0.03 * rng.standard_normal((D, D))
    F1 = np.array([0.02, -0.01, 0.015])
    H1 = d.oppsi(oracle_mod.Op(1, V=V, F2=F2, F1=F1), np.eye(d.nb_tot))
    op1 = np.diag(V)
    for i in range(D):
        op1 = op1 + F1[i] * d1[i]
        for j in range(D):
            op1 = op1 + F2[i, j] * (d2[i] if i == j else d1[i] @ d1[j])
    assert np.abs(H1 - Wt @ op1 @ Bt).max() < 1e-11 * np.abs(H1).max()


def test_grid_derivative_closed_form(oracle_mod):
    """orc_sg4_deriv / orc_dp_deriv on the grid image of a basis vector equal
    the closed-form derivatives of the Hermite functions: d/dq_i and
    d2/dq_i2 of phi_n(q_i) = H_n e^{-q^2/2}/sqrt(2^n n! sqrt(pi)), and the
    mixed d_i d_j (dnRGG%d1 B = B', dnRGG%d2 B = B'' exactly for functions
    of the basis).  A wrong mode index or d1 for d2 fails."""
    from numpy.polynomial.hermite import hermval
    from math import factorial, pi, sqrt

    def phi(n, x, der):
        c = np.zeros(n + 1); c[n] = 1.0
        f = hermval(x, c) * np.exp(-x * x / 2) / sqrt(2.0 ** n * factorial(n) * sqrt(pi))
        if der == 0:
            return f
        c1 = np.zeros(n + 2); c1[n + 1] = 1.0
        fp1 = hermval(x, c1) * np.exp(-x * x / 2) / sqrt(2.0 ** (n + 1) * factorial(n + 1) * sqrt(pi))
        fm1 = 0.0
        if n > 0:
            cm = np.zeros(n); cm[n - 1] = 1.0
            fm1 = hermval(x, cm) * np.exp(-x * x / 2) / sqrt(2.0 ** (n - 1) * factorial(n - 1) * sqrt(pi))
        if der == 1:                     # phi_n' = sqrt(n/2) phi_{n-1} - sqrt((n+1)/2) phi_{n+1}
            return sqrt(n / 2.0) * fm1 - sqrt((n + 1) / 2.0) * fp1
        return (x * x - 2 * n - 1) * f   # phi_n''

    D, L = 3, 3
    s = oracle_mod.SG4(D, L)
    q = s.grid()
    P = s.packed()
    k = int(np.argmax(P.sum(1)))         # a packed function with several nonzero indices
    e = np.zeros(s.nb_packed); e[k] = 1.0
    g = s.b2g(e)
    n = P[k]
    # the grid image is phi_n on the grids of the terms that hold function k
    _, _, onb, onq = s.terms()
    m = s.map()
    on = np.zeros(len(q), bool)
    for t in range(s.nterms):
        if (m[onb[t]:onb[t + 1]] == k).any():
            on[onq[t]:onq[t + 1]] = True
    assert on.sum() > 0 and np.abs(g[~on]).max() == 0.0
    for (i, j) in [(0, -1), (2, -1), (1, 1), (0, 2), (-1, -1)]:
        want = np.ones(len(q))
        for d in range(D):
            der = (d == i) + (d == j)
            want = want * phi(n[d], q[:, d], der)
        got = s.deriv(g, i, j)
        assert np.abs(got[on] - want[on]).max() < 1e-11, (i, j)
        assert np.abs(got[~on]).max() == 0.0
    assert np.abs(s.deriv(g, 1, -1) - s.deriv(g, 0, -1)).max() > 1e-3
    d = oracle_mod.DP([4, 5, 3], [6, 7, 5])
    qd = d.grid()
    e = np.zeros(d.nb_tot); e[7] = 1.0      # last mode fastest: (0, 2, 1)
    gd = d.b2g(e)
    want = phi(0, qd[:, 0], 0) * phi(2, qd[:, 1], 1) * phi(1, qd[:, 2], 1)
    assert np.abs(d.deriv(gd, 1, 2) - want).max() < 1e-11
