"""evr_oppsi_noalloc (the allocation-free H|psi>) against evr_oppsi through
the C ABI: the same kernels on the same inputs, so the results are identical
to the last bit; two calls in a row, the second with fewer vectors (the
shared scratch keeps the first call's contents), FP64 and FP32, sum-factorised
and assembled; a plan whose work memory exceeds the scratch is refused."""
import numpy as np
import pytest

from conftest import DeviceBuffer

pytestmark = pytest.mark.gpu


def _run(evr, plan, X, nvec, flags, noalloc):
    x = DeviceBuffer.of(X[:, :nvec])
    y = DeviceBuffer(x.nbytes)
    y.fill_nan()
    f = plan.TabOpPsi_noalloc_device if noalloc else plan.TabOpPsi_device
    f(x.ptr, y.ptr, nvec, flags)
    DeviceBuffer.sync(plan.stream())
    out = y.download((X.shape[0], nvec))
    x.free()
    y.free()
    return out


@pytest.mark.parametrize("assembled", [False, True])
def test_noalloc_matches_oppsi(evr, assembled):
    from evrhip import hamiltonian as H
    D, L = 6, 5
    ham = H.smolyak_config(1, D)
    b = evr.Basis.sg4(D, L)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=16, hermitian=False)
    plan.gen_potential_poly(ham)
    if assembled:
        plan.assemble(evr.ASM_FP64 | evr.ASM_FP32)
    X = H.test_vectors(b.nb_packed, 16, 12)
    print("noalloc: D=%d L=%d assembled=%s" % (D, L, assembled), flush=True)
    for flags in (evr.FP64, evr.FP32):
        for nvec in (16, 5):          # second call smaller: stale scratch must not leak
            ref = _run(evr, plan, X, nvec, flags, False)
            got = _run(evr, plan, X, nvec, flags, True)
            assert np.isfinite(got).all()
            assert np.array_equal(got, ref), (flags, nvec, np.abs(got - ref).max())
    plan.close()


def test_noalloc_refuses_oversized(evr):
    from evrhip import hamiltonian as H
    D, L = 8, 6          # 8 x 32 x (40,081 packed + 439,019 term rows) bytes: 117 MiB of work
    ham = H.smolyak_config(9, D)
    b = evr.Basis.sg4(D, L)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=32, hermitian=True)
    plan.gen_potential_poly(ham)
    X = H.test_vectors(b.nb_packed, 32, 13)
    need = 8 * 32 * (b.nb_packed + b.sum_nb)     # about the work memory (pruned rows aside)
    assert need > 1.5 * evr.NOALLOC_SCRATCH_BYTES
    with pytest.raises(RuntimeError, match="exceed the"):
        _run(evr, plan, X, 32, evr.FP64, True)
    with pytest.raises(RuntimeError, match="device pointers only"):
        evr.check(evr._lib.load().evr_oppsi_noalloc(plan.h, 1, 1, 1, evr.FP64))
    # a call that fits still runs after the refusals
    got = _run(evr, plan, X, 2, evr.FP64, True)
    ref = _run(evr, plan, X, 2, evr.FP64, False)
    assert np.array_equal(got, ref)
    plan.close()
