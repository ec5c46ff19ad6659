"""D=9 SG4 L=8 (24,310 Smolyak terms, 598,417 packed functions, 8.0e8 grid
points): parity at the full size of the H|psi> benchmark.

- the benchmark operator itself (type 10, q-linear G, the generator potential,
  hermitian = 0) against the oracle on two vectors: FP64 sum-factorised and
  assembled 1e-12 (relative max-norm); the oracle evaluates the potential per
  grid point as the reference fills its grid cache and runs on every host
  core.  FP32: 2e-4 sum-factorised, 1e-3 assembled (measured 6.8e-5 and
  5.0e-4).  The Smolyak sum cancels terms of weight up to C(8, 4) = 70 with
  alternating signs, so an FP32 rounding of a term's rows (about 6e-8 times
  sqrt of the contraction length: 17 per mode sum-factorised, up to 6,561
  for an assembled term row) is amplified up to 70 times in the packed
  result; the FP32 H|psi> serves the Davidson iterations, whose levels are
  brought to FP64 accuracy by the FP64 defect corrections;
- the kinetic part (type 10, q-linear G, no potential) against the oracle on
  one vector, 1e-12 relative max-norm;
- a constant potential (type 0): H psi = c psi exactly (W_d B_d = I and the
  Smolyak weights sum to one), matrix-free and assembled, 1e-10;
- the full generator Hamiltonian: assembled operator against the
  sum-factorised kernels, 1e-11; the hermitian form is symmetric, 1e-12.

The assembled pool is 144 GB (FP64) at this size; the plans are closed
between the tests to release it.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D, L = 9, 8


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


@pytest.mark.timeout(900)
def test_c2_benchmark_operator_vs_oracle(evr, oracle_mod, c2):
    ham, b = c2
    X = evr.hamiltonian.test_vectors(b.nb_packed, 2, 8)
    print("D=9 L=8 benchmark operator: oracle on %d threads" % host_threads(), flush=True)
    t = time.time()
    s = oracle_mod.SG4(D, L)
    op = oracle_mod.Op(10, G0=ham.G0, Ga=ham.Ga_fortran(),
                       Vpoly=oracle_mod.poly_coefs(ham.v2, ham.c3, ham.d22, ham.e4))
    ref = oracle_chunks(s, op, X)
    print("oracle %.1f s" % (time.time() - t), flush=True)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=2, hermitian=False)
    plan.gen_potential_poly(ham)
    mf = plan.TabOpPsi(X)
    mf32 = plan.TabOpPsi(X, evr.FP32)
    plan.assemble(evr.ASM_FP64 | evr.ASM_FP32)
    asm = plan.TabOpPsi(X)
    asm32 = plan.TabOpPsi(X, evr.FP32)
    plan.close()
    errs = dict(mf64=relerr(mf, ref), asm64=relerr(asm, ref), mf32=relerr(mf32, ref),
                asm32=relerr(asm32, ref))
    print("D=9 L=8 benchmark operator vs oracle", errs)
    assert errs["mf64"] < 1e-12 and errs["asm64"] < 1e-12
    assert errs["mf32"] < 2e-4 and errs["asm32"] < 1e-3


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
