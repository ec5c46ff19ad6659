"""D=12 SG4 L=6 (18,564 Smolyak terms, 369,305 packed functions, 2.4e9 grid
points; twelve modes is the most the term kernel takes): the assembled
operator against the sum-factorised kernels with the generator Hamiltonian
(1e-11 relative max-norm) and the symmetry of the hermitian form (1e-12).
The kinetic part of both rests on the same per-mode factors as at D=6 L=5
and D=9 L=8, where the oracle pins it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_c4_assembled_matches_matrix_free(evr):
    from evrhip import hamiltonian as H
    D, L = 12, 6
    ham = H.smolyak_config(4, D)
    b = evr.Basis.sg4(D, L)
    assert (b.nterms, b.nb_packed, b.sum_nb, b.sum_nq) == (18564, 369305, 4195284, 2432102400)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 8, 7)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=8, hermitian=True)
    plan.gen_potential_poly(ham)
    mf = plan.TabOpPsi(X)
    plan.assemble(evr.ASM_FP64)
    asm = plan.TabOpPsi(X)
    plan.close()
    print("D=12 L=6 assembled vs matrix-free", relerr(asm, mf))
    assert relerr(asm, mf) < 1e-11
    for H_ in (mf, asm):
        S = X.T @ H_
        assert np.abs(S - S.T).max() < 1e-12 * np.abs(S).max()
