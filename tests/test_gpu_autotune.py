"""The per-term tile autotuner (evr_plan_autotune): every bucket of terms
tries the larger kernel classes, the layout kept gives the same H|psi> (the
same arithmetic per term, to rounding of the slab order), and is not slower
than the static rule by more than the timing noise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("D,L,nvec", [(6, 5, 16), (9, 5, 8)])
def test_autotune(evr, D, L, nvec):
    from evrhip import hamiltonian as H
    ham = H.smolyak_config(9, D)
    b = evr.Basis.sg4(D, L)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=nvec, hermitian=True)
    plan.gen_potential_poly(ham)
    X = evr.hamiltonian.test_vectors(b.nb_packed, nvec, 21)
    ref = plan.TabOpPsi(X, evr.MATRIX_FREE)
    chosen, ms0, ms1 = plan.autotune(nvec, evr.FP64, reps=3)
    print("D=%d L=%d autotune: classes %s, static %.3f ms, tuned %.3f ms" % (D, L, chosen, ms0, ms1))
    assert all(c >= k for k, c in enumerate(chosen))
    assert ms1 <= 1.05 * ms0
    got = plan.TabOpPsi(X, evr.MATRIX_FREE)
    assert relerr(got, ref) < 1e-13
    plan.assemble(evr.ASM_FP64)                 # the assembly runs the tuned layout too
    assert relerr(plan.TabOpPsi(X), ref) < 1e-12
    with pytest.raises(RuntimeError, match="max_nvec"):
        plan.autotune(nvec + 1)
