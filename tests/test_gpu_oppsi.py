"""H|psi> parity of the HIP engine against the CPU oracle (through the C ABI)."""
import numpy as np
import pytest

from conftest import hh_potential

pytestmark = pytest.mark.gpu

TOL64 = 1e-12   # relative max-norm, FP64 path


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _c1(evr, oracle_mod, D=6, L=5):
    from evrhip import hamiltonian as H
    ham = H.smolyak_config(1, D)
    s = oracle_mod.SG4(D, L)
    V = ham.potential(s.grid())
    return ham, s, V


def test_c1_type10_fp64(evr, oracle_mod):
    ham, s, V = _c1(evr, oracle_mod)
    b = evr.Basis.sg4(6, 5)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=16)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 16, 1)
    ref = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), X)
    got = plan.TabOpPsi(X)
    print("C1 type10 rel err", relerr(got, ref))
    assert relerr(got, ref) < TOL64


def test_c1_gen_potential(evr, oracle_mod):
    ham, s, V = _c1(evr, oracle_mod)
    b = evr.Basis.sg4(6, 5)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=2)
    plan.gen_potential_poly(ham)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 2, 1)
    ref = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), X)
    assert relerr(plan.TabOpPsi(X), ref) < TOL64


def test_c1_type0_and_constant_G(evr, oracle_mod):
    ham, s, V = _c1(evr, oracle_mod)
    b = evr.Basis.sg4(6, 5)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 3, 1)
    p0 = evr.Plan(b, type=0, V=V, max_nvec=3)
    assert relerr(p0.TabOpPsi(X), s.oppsi(oracle_mod.Op(0, V=V), X)) < TOL64
    G0 = ham.G0 + 0.01 * np.ones((6, 6)) * ham.omega.min()
    p1 = evr.Plan(b, type=10, V=V, G0=G0, max_nvec=3)
    assert relerr(p1.TabOpPsi(X), s.oppsi(oracle_mod.Op(10, V=V, G0=G0), X)) < TOL64


def test_hh6d_type1(evr, oracle_mod):
    D = 6
    s = oracle_mod.SG4(D, 3, 3, nbA=1, nbB=2, nqA=1, nqB=2)
    b = evr.Basis.sg4(D, 3, 3, 1, 2, 1, 2)
    V = hh_potential(s.grid())
    F2 = -0.5 * np.eye(D)
    F1 = np.linspace(-0.1, 0.1, D)
    plan = evr.Plan(b, type=1, V=V, F2=F2, F1=F1, max_nvec=4)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 4, 7)
    ref = s.oppsi(oracle_mod.Op(1, V=V, F2=F2, F1=F1), X)
    assert relerr(plan.TabOpPsi(X), ref) < TOL64


def test_c0_direct_product(evr, oracle_mod):
    from evrhip import hamiltonian as H
    ham = H.config0()
    nb, nq = [12] * 3, [14] * 3
    d = oracle_mod.DP(nb, nq)
    V = ham.potential(d.grid())
    b = evr.Basis.dp(nb, nq)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, max_nvec=8)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 8, 0)
    ref = d.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0), X)
    assert relerr(plan.TabOpPsi(X), ref) < TOL64


def test_basis_grid_roundtrip(evr, oracle_mod):
    ham, s, V = _c1(evr, oracle_mod)
    b = evr.Basis.sg4(6, 5)
    plan = evr.Plan(b, type=0, V=V, max_nvec=1)
    x = evr.hamiltonian.test_vectors(b.nb_packed, 1, 3)[:, 0]
    g = plan.BasisRep_TO_GridRep(x)
    assert relerr(g, s.b2g(x)) < TOL64
    back = plan.GridRep_TO_BasisRep(g)
    assert relerr(back, s.g2b(g)) < TOL64
    assert relerr(back, x) < 1e-12   # B->G->B is the identity on the packed space


def test_c1_fp32(evr, oracle_mod):
    ham, s, V = _c1(evr, oracle_mod)
    b = evr.Basis.sg4(6, 5)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=4)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 4, 1)
    ref = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), X)
    got = plan.TabOpPsi(X, flags=evr.FP32)
    assert relerr(got, ref) < 2e-5   # FP32 contractions, FP32 accumulation
