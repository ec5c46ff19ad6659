"""Error paths of the C ABI on the GPU: every bad argument comes back as a
nonzero status with a message (the reference STOPs on the same conditions),
and the plan stays usable afterwards."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(evr):
    b = evr.Basis.sg4(3, 3)
    V = 0.5 * (b.grid() ** 2).sum(1)
    return b, V


def test_oppsi_vector_counts(evr, small):
    b, V = small
    p = evr.Plan(b, type=1, V=V, F2=-0.5 * np.eye(3), max_nvec=4)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 5, 1)
    with pytest.raises(RuntimeError, match="max_nvec"):
        p.TabOpPsi(X)                       # 5 > max_nvec
    with pytest.raises(RuntimeError, match="max_nvec"):
        p.TabOpPsi(np.zeros((b.nb_packed, 0)))
    ok = p.TabOpPsi(X[:, :4])               # still usable
    assert np.isfinite(ok).all()


def test_operator_descriptions(evr, small):
    b, V = small
    with pytest.raises(RuntimeError, match="type"):
        evr.Plan(b, type=2, V=V, max_nvec=1)
    with pytest.raises(RuntimeError, match="G0"):
        evr.Plan(b, type=10, V=V, max_nvec=1)
    with pytest.raises(RuntimeError, match="F2"):
        evr.Plan(b, type=1, V=V, max_nvec=1)
    with pytest.raises(RuntimeError, match="max_nvec"):
        evr.Plan(b, type=0, V=V, max_nvec=0)
    G0 = np.array([[1.0, 0.2, 0.0], [0.1, 1.0, 0.0], [0.0, 0.0, 1.0]])
    with pytest.raises(RuntimeError, match="symmetric G0"):
        evr.Plan(b, type=10, V=V, G0=G0, max_nvec=1, hermitian=True)


def test_coupled_block_checks(evr, small):
    b, V = small
    p = evr.Plan(b, type=0, V=V, max_nvec=2)
    other = evr.Plan(evr.Basis.sg4(3, 2), type=0, V=np.zeros(evr.Basis.sg4(3, 2).sum_nq), max_nvec=2)
    with pytest.raises(RuntimeError, match="different basis sizes"):
        evr.Coupled([[p, other], [other, p]])
    with pytest.raises(RuntimeError, match="state 1 has no block"):
        evr.Coupled([[p, None], [None, None]])


def test_davidson_and_propagation_arguments(evr, small):
    b, V = small
    p = evr.Plan(b, type=1, V=V, F2=-0.5 * np.eye(3), max_nvec=4)
    with pytest.raises(RuntimeError, match="nev"):
        p.davidson(b.nb_packed + 1)
    psi = np.zeros(b.nb_packed, complex)
    psi[0] = 1.0
    with pytest.raises(RuntimeError, match="bad dt"):
        p.propagate(psi, 0.0, 2)
    with pytest.raises(RuntimeError, match="SIL needs"):
        p.propagate(psi, 0.1, 1, method=evr.SIL, npoly=64)
    with pytest.raises(RuntimeError, match="npoly >= 3"):
        p.propagate(psi, 0.1, 1, npoly=2)
    out, _, info = p.propagate(psi, 0.1, 2)  # and the plan still propagates
    assert abs(np.linalg.norm(out) - 1.0) < 1e-10
