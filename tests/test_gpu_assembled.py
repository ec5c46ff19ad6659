"""Assembled per-term operators (evr_plan_assemble + MFMA apply) against the
oracle and against the sum-factorised path of the same plan.

Bars: FP64 relative max-norm 1e-12 against the oracle, FP32 (FP64-built
matrices stored in FP32, FP32 MFMA) 2e-5; batch sizes that exercise one and
two 16-vector tiles and several 32-vector passes.
"""
import numpy as np
import pytest

from conftest import hh_potential

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def c1(evr, oracle_mod):
    from evrhip import hamiltonian as H
    ham = H.smolyak_config(1, 6)
    s = oracle_mod.SG4(6, 5)
    V = ham.potential(s.grid())
    return ham, s, V


@pytest.mark.parametrize("nvec", [1, 5, 16, 17, 32, 40])
def test_c1_type10_assembled(evr, oracle_mod, c1, nvec):
    ham, s, V = c1
    b = evr.Basis.sg4(6, 5)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=40)
    plan.assemble(evr.ASM_FP64 | evr.ASM_FP32)
    mask, nbytes = plan.assembled()
    assert mask == 3 and nbytes > 0
    X = evr.hamiltonian.test_vectors(b.nb_packed, nvec, 5)
    ref = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), X)
    assert relerr(plan.TabOpPsi(X), ref) < 1e-12
    assert relerr(plan.TabOpPsi(X, flags=evr.FP32), ref) < 2e-5
    assert relerr(plan.TabOpPsi(X, flags=evr.MATRIX_FREE), ref) < 1e-12
    # whichever path the cost model picks, the operator is the same
    assert relerr(plan.TabOpPsi(X, flags=evr.AUTO_PATH), ref) < 1e-12
    assert relerr(plan.TabOpPsi(X, flags=evr.AUTO_PATH | evr.FP32), ref) < 2e-5


def test_hermitian_assembled_matches_matrix_free(evr, c1):
    ham, s, V = c1
    b = evr.Basis.sg4(6, 5)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=16, hermitian=True)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 16, 6)
    mf = plan.TabOpPsi(X)
    plan.assemble()
    assert relerr(plan.TabOpPsi(X), mf) < 1e-12
    # a new potential drops the assembled operator
    plan.set_potential(V)
    assert plan.assembled()[0] == 0


def test_type1_and_direct_product_assembled(evr, oracle_mod):
    from evrhip import hamiltonian as H
    D = 6
    s = oracle_mod.SG4(D, 3, 3, 1, 2, 1, 2)
    b = evr.Basis.sg4(D, 3, 3, 1, 2, 1, 2)
    V = hh_potential(s.grid())
    plan = evr.Plan(b, type=1, V=V, F2=-0.5 * np.eye(D), max_nvec=8)
    plan.assemble()
    X = evr.hamiltonian.test_vectors(b.nb_packed, 8, 7)
    assert relerr(plan.TabOpPsi(X), s.oppsi(oracle_mod.Op(1, V=V, F2=-0.5 * np.eye(D)), X)) < 1e-12
    ham = H.config0()
    d = oracle_mod.DP([12] * 3, [14] * 3)
    Vd = ham.potential(d.grid())
    pd = evr.Plan(evr.Basis.dp([12] * 3, [14] * 3), type=10, V=Vd, G0=ham.G0, max_nvec=8)
    pd.assemble(evr.ASM_FP64)
    Xd = evr.hamiltonian.test_vectors(1728, 8, 0)
    assert relerr(pd.TabOpPsi(Xd), d.oppsi(oracle_mod.Op(10, V=Vd, G0=ham.G0), Xd)) < 1e-12


def test_davidson_on_assembled_operator(evr, c1):
    ham, s, V = c1
    b = evr.Basis.sg4(6, 5)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=16, hermitian=True)
    ev_mf, _, _ = plan.davidson(40, block=16, conv_resi=1e-8, conv_ene=1e-10)
    plan.assemble(evr.ASM_FP64 | evr.ASM_FP32)
    ev_a, _, info = plan.davidson(40, block=16, conv_resi=1e-8, conv_ene=1e-10, precision=evr.FP32,
                                  refine_resi=1e-1)
    assert info.converged == 1
    assert np.abs(ev_a - ev_mf).max() < 1e-7
