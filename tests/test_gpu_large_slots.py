"""1D slots beyond 20 points: SG4 levels up to nb = 31 functions (the most a
5-bit basis digit of the term kernels holds) and nq = 32 points, H|psi>
(sum-factorised and assembled, FP64) and the basis <-> grid transforms
against the oracle; slots of 32 to 64 functions run streamed (checked
against the oracle), a slot of 65 is refused with its message."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("D,L", [(2, 14), (3, 15)])
def test_large_slots_oppsi(evr, oracle_mod, D, L):
    from evrhip import hamiltonian as H
    ham = H.smolyak_config(9, D)
    s = oracle_mod.SG4(D, L)
    assert s.max_nq > 20
    V = ham.potential(s.grid())
    b = evr.Basis.sg4(D, L)
    X = H.test_vectors(b.nb_packed, 3, 5)
    ref = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), X)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=3)
    e_mf = relerr(plan.TabOpPsi(X, flags=evr.MATRIX_FREE), ref)
    plan.assemble()
    e_asm = relerr(plan.TabOpPsi(X), ref)
    print("D=%d L=%d (nq up to %d): matrix-free %.2e, assembled %.2e" % (D, L, 2 + 2 * L, e_mf, e_asm))
    assert e_mf < TOL64 and e_asm < TOL64


def test_large_slots_transforms(evr, oracle_mod):
    D, L = 3, 15
    s = oracle_mod.SG4(D, L)
    b = evr.Basis.sg4(D, L)
    plan = evr.Plan(b, type=0, V=np.ones(s.sum_nq), max_nvec=1)
    x = evr.hamiltonian.test_vectors(b.nb_packed, 1, 7)[:, 0]
    g = plan.BasisRep_TO_GridRep(x)
    assert relerr(g, s.b2g(x)) < TOL64
    back = plan.GridRep_TO_BasisRep(g)
    assert relerr(back, s.g2b(g)) < TOL64
    assert relerr(back, x) < 1e-12


def test_slots_above_31_functions_streamed(evr, oracle_mod):
    from evrhip import hamiltonian as H
    # nb(l) = 32 + 8 l, nq(l) = 33 + 8 l: every term has a slot beyond 31
    s = oracle_mod.SG4(2, 2, 2, 32, 8, 33, 8)
    b = evr.Basis.sg4(2, 2, 2, 32, 8, 33, 8)
    ham = H.smolyak_config(9, 2)
    V = ham.potential(s.grid())
    X = H.test_vectors(b.nb_packed, 2, 6)
    ref = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), X)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=2)
    e = relerr(plan.TabOpPsi(X), ref)
    g = plan.BasisRep_TO_GridRep(X[:, 0])
    print("slots of 32..48 functions (streamed): H|psi> %.2e, B->G %.2e" % (e, relerr(g, s.b2g(X[:, 0]))))
    assert e < TOL64 and relerr(g, s.b2g(X[:, 0])) < TOL64


def test_slot_above_64_functions_refused(evr):
    b = evr.Basis.sg4(2, 1, 1, 65, 1, 66, 1)   # nb(0) = 65 functions
    with pytest.raises(RuntimeError, match="nb <= nq <= 64"):
        evr.Plan(b, type=0, V=np.ones(b.sum_nq), max_nvec=1)
