"""D=12 SG4 L=6 (18,564 Smolyak terms, 369,305 packed functions, 2.4e9 grid
points; twelve modes is the most the term kernel takes):

- the generator operator (type 10, q-linear G, potential per grid point,
  hermitian = 0) against the oracle on one vector, FP64 sum-factorised and
  assembled 1e-12 (the oracle on every host core, a few minutes); FP32
  assembled 2e-3 (Smolyak weights up to C(11, 5) = 462 amplify the FP32
  rounding of the term rows);
- the assembled operator against the sum-factorised kernels with the
  hermitian form (1e-11 relative max-norm) and its symmetry (1e-12).
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def host_threads():
    n = len(os.sched_getaffinity(0))
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def oracle_chunks(s, op, X, nchunk=16):
    """The oracle's H|psi> as the sum over consecutive term ranges (a progress
    line per range: the full operator takes minutes on the host cores)."""
    ref = np.zeros_like(X)
    edges = np.linspace(0, s.nterms, nchunk + 1).astype(int)
    t = time.time()
    for k in range(nchunk):
        ref += s.oppsi(op, X, nthreads=host_threads(), t0=int(edges[k]), t1=int(edges[k + 1]))
        print("  oracle terms %d..%d done (%.0f s)" % (edges[k], edges[k + 1], time.time() - t), flush=True)
    return ref


@pytest.mark.timeout(1200)
def test_c4_operator_vs_oracle(evr, oracle_mod):
    from evrhip import hamiltonian as H
    D, L = 12, 6
    ham = H.smolyak_config(4, D)
    b = evr.Basis.sg4(D, L)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 1, 9)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=1, hermitian=False)
    plan.gen_potential_poly(ham)
    mf = plan.TabOpPsi(X)
    plan.assemble(evr.ASM_FP64 | evr.ASM_FP32)
    asm = plan.TabOpPsi(X)
    asm32 = plan.TabOpPsi(X, evr.FP32)
    plan.close()
    print("D=12 L=6 operator: oracle on %d threads" % host_threads(), flush=True)
    t = time.time()
    s = oracle_mod.SG4(D, L)
    op = oracle_mod.Op(10, G0=ham.G0, Ga=ham.Ga_fortran(),
                       Vpoly=oracle_mod.poly_coefs(ham.v2, ham.c3, ham.d22, ham.e4))
    ref = oracle_chunks(s, op, X)
    print("oracle %.1f s" % (time.time() - t), flush=True)
    errs = dict(mf64=relerr(mf, ref), asm64=relerr(asm, ref), asm32=relerr(asm32, ref))
    print("D=12 L=6 operator vs oracle", errs)
    assert errs["mf64"] < 1e-12 and errs["asm64"] < 1e-12
    assert errs["asm32"] < 2e-3


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
