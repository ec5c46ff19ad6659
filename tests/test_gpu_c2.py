"""D=9 SG4 L=8 (24,310 Smolyak terms, 598,417 packed functions, 8.0e8 grid
points): parity at the full size of the H|psi> benchmark.

- the kinetic part (type 10, q-linear G, no potential) against the oracle on
  one vector, 1e-12 relative max-norm;
- a constant potential (type 0): H psi = c psi exactly (W_d B_d = I and the
  Smolyak weights sum to one), matrix-free and assembled, 1e-10;
- the full generator Hamiltonian: assembled operator against the
  sum-factorised kernels, 1e-11; the hermitian form is symmetric, 1e-12.

The assembled pool is 144 GB (FP64) at this size; the plans are closed
between the tests to release it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D, L = 9, 8


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def c2(evr):
    from evrhip import hamiltonian as H
    return H.smolyak_config(2, D), evr.Basis.sg4(D, L)


def test_c2_kinetic_vs_oracle(evr, oracle_mod, c2):
    ham, b = c2
    s = oracle_mod.SG4(D, L)
    x = evr.hamiltonian.test_vectors(b.nb_packed, 1, 4)
    ref = s.oppsi(oracle_mod.Op(10, G0=ham.G0, Ga=ham.Ga_fortran()), x, nthreads=16)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=1)
    got = plan.TabOpPsi(x)
    plan.close()
    print("D=9 L=8 kinetic part vs oracle", relerr(got, ref))
    assert relerr(got, ref) < 1e-12


def test_c2_constant_potential_identity(evr, c2):
    _, b = c2
    X = evr.hamiltonian.test_vectors(b.nb_packed, 4, 5)
    plan = evr.Plan(b, type=0, V=np.full(b.sum_nq, 2.5), max_nvec=4)
    mf = plan.TabOpPsi(X)
    plan.assemble(evr.ASM_FP64)
    asm = plan.TabOpPsi(X)
    plan.close()
    print("D=9 L=8 constant V: matrix-free", relerr(mf, 2.5 * X), "assembled", relerr(asm, 2.5 * X))
    assert relerr(mf, 2.5 * X) < 1e-10
    assert relerr(asm, 2.5 * X) < 1e-10


@pytest.mark.parametrize("hermitian", [False, True])
def test_c2_assembled_matches_matrix_free(evr, c2, hermitian):
    ham, b = c2
    X = evr.hamiltonian.test_vectors(b.nb_packed, 8, 6)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=8, hermitian=hermitian)
    plan.gen_potential_poly(ham)
    mf = plan.TabOpPsi(X)
    plan.assemble(evr.ASM_FP64)
    asm = plan.TabOpPsi(X)
    plan.close()
    print("D=9 L=8 hermitian=%s assembled vs matrix-free" % hermitian, relerr(asm, mf))
    assert relerr(asm, mf) < 1e-11
    if hermitian:
        for H_ in (mf, asm):
            S = X.T @ H_
            assert np.abs(S - S.T).max() < 1e-12 * np.abs(S).max()
