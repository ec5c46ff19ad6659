"""Multi-vector BasisTOGrid / GridTOBasis and the grid derivative
(evr_basis_to_grid_n, evr_grid_to_basis_n, evr_grid_deriv) through the C ABI,
host and device pointers, SG4 and direct-product bases, against the oracle
(1e-12 relative max-norm); GridTOBasis(BasisTOGrid(x)) = x on the packed
space."""
import numpy as np
import pytest

from conftest import DeviceBuffer

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


DERS = [(0, 0), (1, 0), (0, 2), (2, 2), (1, 3), (3, 1)]


@pytest.mark.parametrize("D,L,ncol,max_nvec", [(3, 4, 3, 2), (5, 4, 4, 4)])
def test_sg4_transforms_vs_oracle(evr, oracle_mod, D, L, ncol, max_nvec):
    s = oracle_mod.SG4(D, L)
    b = evr.Basis.sg4(D, L)
    plan = evr.Plan(b, type=0, V=np.zeros(b.sum_nq), max_nvec=max_nvec)
    X = evr.hamiltonian.test_vectors(b.nb_packed, ncol, 40 + D)
    G = plan.BasisRep_TO_GridRep_n(X)
    Gref = np.stack([s.b2g(X[:, c]) for c in range(ncol)], axis=1)
    assert relerr(G, Gref) < 1e-12
    Y = plan.GridRep_TO_BasisRep_n(G)
    assert relerr(Y, X) < 1e-12                       # the packed-space identity
    rng = np.random.default_rng(D)
    Gr = rng.standard_normal((b.sum_nq, ncol))
    Yr = plan.GridRep_TO_BasisRep_n(Gr)
    assert relerr(Yr, np.stack([s.g2b(Gr[:, c]) for c in range(ncol)], axis=1)) < 1e-12
    for (i, j) in DERS:
        if max(i, j) > D:
            continue
        got = plan.DerivOp_TO_RVecG(Gr, (i, j))
        ref = np.stack([s.deriv(Gr[:, c], i - 1, j - 1) for c in range(ncol)], axis=1)
        assert relerr(got, ref) < 1e-12, (i, j)
    # device pointers
    dx = DeviceBuffer.of(X)
    dg = DeviceBuffer(8 * b.sum_nq * ncol)
    dd = DeviceBuffer(8 * b.sum_nq * ncol)
    dy = DeviceBuffer(8 * b.nb_packed * ncol)
    plan.transform_device("b2g", ncol, dx.ptr, dg.ptr)
    plan.transform_device("deriv", ncol, dg.ptr, dd.ptr, (1, 2))
    plan.transform_device("g2b", ncol, dg.ptr, dy.ptr)
    DeviceBuffer.sync(plan.stream())
    assert np.array_equal(dg.download((b.sum_nq, ncol)), G)
    assert relerr(dy.download((b.nb_packed, ncol)), X) < 1e-12
    assert np.array_equal(dd.download((b.sum_nq, ncol)), plan.DerivOp_TO_RVecG(G, (1, 2)))
    with pytest.raises(RuntimeError, match="tab_der"):
        plan.DerivOp_TO_RVecG(Gr, (D + 1, 0))
    plan.close()


def test_dp_transforms_vs_oracle(evr, oracle_mod):
    nb, nq = [4, 5, 3], [6, 7, 5]
    d = oracle_mod.DP(nb, nq)
    b = evr.Basis.dp(nb, nq)
    plan = evr.Plan(b, type=0, V=np.zeros(b.sum_nq), max_nvec=2)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 3, 7)
    G = plan.BasisRep_TO_GridRep_n(X)
    assert relerr(G, np.stack([d.b2g(X[:, c]) for c in range(3)], axis=1)) < 1e-12
    assert relerr(plan.GridRep_TO_BasisRep_n(G), X) < 1e-12
    for (i, j) in [(1, 0), (3, 3), (2, 3)]:
        got = plan.DerivOp_TO_RVecG(G, (i, j))
        ref = np.stack([d.deriv(G[:, c], i - 1, j - 1) for c in range(3)], axis=1)
        assert relerr(got, ref) < 1e-12, (i, j)
    plan.close()
