"""Type 10 with G, J and s = sqrt(rho/J) tabulated per grid point
(evr_plan_set_keo_grid: the reference's get_OpGrid_type10 tables and its grid
form OPSG4:1566-1632) against the oracle, FP64 (1e-12) and FP32 (2e-5), its
hermitian form (H + H^T) / 2, and against the polynomial G path when the
table is G0 + Ga q."""
import numpy as np
import pytest

from test_oracle import tabulated_keo

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("D,L,nvec", [(4, 4, 5), (6, 4, 3), (3, 6, 8)])
def test_keo_grid_vs_oracle(evr, oracle_mod, D, L, nvec):
    s = oracle_mod.SG4(D, L)
    b = evr.Basis.sg4(D, L)
    q = s.grid()
    Gq, J, sq = tabulated_keo(q, 20 + D)
    V = 0.5 * (q ** 2).sum(1) + 0.05 * q[:, 0] ** 3
    plan = evr.Plan(b, type=10, V=V, G0=np.eye(D), max_nvec=nvec)
    plan.set_keo_grid(Gq, J, sq)
    X = evr.hamiltonian.test_vectors(s.nb_packed, nvec, D)
    ref = s.oppsi(oracle_mod.Op(10, V=V, Jac=J, s=sq, Gq=Gq), X)
    err = relerr(plan.TabOpPsi(X), ref)
    err32 = relerr(plan.TabOpPsi(X, flags=evr.FP32), ref)
    print("G, J, s per point D=%d L=%d: rel err %.3e, FP32 %.3e" % (D, L, err, err32))
    assert err < 1e-12
    assert err32 < 2e-5


def test_keo_grid_equals_polynomial_g(evr):
    """G0 + Ga q tabulated, J = s = 1: the grid form equals the basis-space
    form of the polynomial G (two different algorithms for the same operator)."""
    D, L = 5, 4
    from evrhip import hamiltonian as H
    ham = H.smolyak_config(9, D)
    b = evr.Basis.sg4(D, L)
    q = b.grid()
    V = ham.potential(q)
    poly = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=4)
    Ga = ham.Ga_fortran().reshape(D, D, D, order="F")            # Ga[i, j, k]
    Gq = ham.G0[None] + np.einsum("pk,ijk->pji", q, Ga)          # Gq[p, j, i] = G_ji(q)
    grid = evr.Plan(b, type=10, V=V, G0=np.eye(D), max_nvec=4)
    grid.set_keo_grid(Gq)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 4, 3)
    err = relerr(grid.TabOpPsi(X), poly.TabOpPsi(X))
    print("grid form vs basis-space form: %.3e" % err)
    assert err < 1e-12


def test_keo_grid_refusals(evr):
    D, L = 4, 3
    b = evr.Basis.sg4(D, L)
    q = b.grid()
    Gq, J, sq = tabulated_keo(q, 1)
    p = evr.Plan(b, type=10, V=np.zeros(len(q)), G0=np.eye(D), max_nvec=2)
    p.set_keo_grid(Gq, J, sq)
    with pytest.raises(RuntimeError, match="not assembled"):
        p.assemble()
    p1 = evr.Plan(b, type=1, V=np.zeros(len(q)), F2=-0.5 * np.eye(D), max_nvec=2)
    with pytest.raises(RuntimeError, match="type-10"):
        p1.set_keo_grid(Gq)


def test_keo_grid_diagonal(evr, oracle_mod):
    """The diagonal of the grid form (H applied to blocks of unit vectors, the
    last block partial) against the oracle's dense operator."""
    D, L = 4, 4
    s = oracle_mod.SG4(D, L)
    b = evr.Basis.sg4(D, L)
    q = s.grid()
    Gq, J, sq = tabulated_keo(q, 31)
    V = 0.5 * (q ** 2).sum(1) + 0.05 * q[:, 0] ** 3
    plan = evr.Plan(b, type=10, V=V, G0=np.eye(D), max_nvec=4)
    plan.set_keo_grid(Gq, J, sq)
    n = s.nb_packed
    assert n % 4 != 0
    Hd = s.oppsi(oracle_mod.Op(10, V=V, Jac=J, s=sq, Gq=Gq), np.eye(n))
    err = relerr(plan.diagonal(), np.diag(Hd))
    print("G per point: diagonal rel err %.3e" % err)
    assert err < 1e-12


def test_keo_grid_hermitian(evr, oracle_mod):
    """hermitian = 1 with G, J, s per point: (H + H^T) / 2 of the reference's
    grid form, through the transposed sequence W diag(1/w) Op^T diag(w) B,
    against the oracle's dense operator (FP64 1e-12, FP32 2e-5), and its
    diagonal."""
    D, L = 3, 4
    s = oracle_mod.SG4(D, L)
    b = evr.Basis.sg4(D, L)
    q = s.grid()
    Gq, J, sq = tabulated_keo(q, 17)
    V = 0.5 * (q ** 2).sum(1) + 0.05 * q[:, 0] ** 3
    n = s.nb_packed
    M = s.oppsi(oracle_mod.Op(10, V=V, Jac=J, s=sq, Gq=Gq), np.eye(n))
    Hs = 0.5 * (M + M.T)
    plan = evr.Plan(b, type=10, V=V, G0=np.eye(D), max_nvec=6, hermitian=True)
    plan.set_keo_grid(Gq, J, sq)
    X = evr.hamiltonian.test_vectors(n, 6, 5)
    ref = Hs @ X
    err = relerr(plan.TabOpPsi(X), ref)
    err32 = relerr(plan.TabOpPsi(X, flags=evr.FP32), ref)
    errd = relerr(plan.diagonal(), np.diag(Hs))
    asym = np.abs(M - M.T).max() / np.abs(M).max()
    print("hermitian G per point: %.3e, FP32 %.3e, diagonal %.3e (asymmetry of H %.2e)" % (err, err32, errd, asym))
    assert asym > 1e-6
    assert err < 1e-12 and errd < 1e-12 and err32 < 2e-5
