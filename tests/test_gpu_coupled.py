"""Coupled electronic states (nb0 = 2) through the C ABI, on the 12D and 24D
pyrazine models of the reference's propagation tests: H psi against the
oracle's coupled operator, and the Chebyshev autocorrelation against the
reference's benchmarks
(Working_tests/MPI_tests/{12D,24D}_propagation_S0/benchmark)."""
import numpy as np
import pytest
import scipy.linalg as sl

from pyrazine import P12, P24
from test_oracle_pyrazine import MODELS, coupled_dense, psi0

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _coupled(evr, P, L=None, max_nvec=4, assemble=False):
    D, LG, LB, nbA, nbB, nqA, nqB = P.sg4_args(L)
    b = evr.Basis.sg4_modes(D, LG, nbA, nbB, nqA, nqB, LB=LB)
    V11, V22, V12 = P.potentials(b.grid())
    p11 = evr.Plan(b, type=1, V=V11, F2=P.F2(), max_nvec=max_nvec)
    p22 = evr.Plan(b, type=1, V=V22, F2=P.F2(), max_nvec=max_nvec)
    w = evr.Plan(b, type=0, V=V12, max_nvec=max_nvec)
    if assemble:
        for p in (p11, p22, w):
            p.assemble()
    return b, evr.Coupled([[p11, w], [w, p22]])


@pytest.mark.parametrize("name,L", [("12d", 1), ("12d", 2), ("24d", 2)])
def test_coupled_oppsi_vs_oracle(evr, oracle_mod, name, L):
    P = MODELS[name]
    b, H = _coupled(evr, P, L)
    s, Hd = coupled_dense(oracle_mod, P, L)
    assert H.nb_packed == s.nb_packed and H.nb0 == 2
    X = evr.hamiltonian.test_vectors(2 * s.nb_packed, 6, 5)   # 6 columns: two calls of 4 + 2
    got = H.TabOpPsi(X)
    err = relerr(got, Hd @ X)
    print("pyrazine %s L=%d coupled H psi rel err %.3e" % (name, L, err))
    assert err < 1e-12
    H.close()


def test_coupled_assembled_matches_matrix_free(evr):
    _, Hm = _coupled(evr, P24)
    _, Ha = _coupled(evr, P24, assemble=True)
    X = evr.hamiltonian.test_vectors(2 * Hm.nb_packed, 4, 9)
    assert relerr(Ha.TabOpPsi(X), Hm.TabOpPsi(X)) < 1e-12


@pytest.mark.parametrize("name,graph", [("12d", False), ("12d", True), ("24d", False), ("24d", True)])
def test_pyrazine_benchmark_autocorr(evr, oracle_mod, name, graph):
    """march_cheby on the GPU at the reference's npoly = 300, poly_tol = 1e-10,
    60 steps of 0.1 fs: <psi0|psi(t)> within 1e-9 of the reference's output."""
    P = MODELS[name]
    b, H = _coupled(evr, P)
    s, _ = coupled_dense(oracle_mod, P)
    p0 = psi0(s)
    t, ref = P.autocorr_ref()
    psi, ac, info = H.propagate(p0, P.DT, P.NSTEPS, npoly=P.NPOLY, poly_tol=P.POLY_TOL, autocorr=True,
                                graph=graph)
    got = np.r_[1.0, ac[:, 0]]
    err = np.abs(got - ref).max()
    print("pyrazine %s GPU (graph=%s) vs benchmark %.3e, norm %.15f, max order %d"
          % (name, graph, err, np.linalg.norm(psi), info.max_order))
    assert err < 1e-9
    assert abs(np.linalg.norm(psi) - 1.0) < 1e-9


def test_pyrazine_L2_vs_exact(evr, oracle_mod):
    """A larger 12D basis (362 functions per state): 20 steps against
    exp(-i H t) of the oracle's coupled operator, two wavepackets at once."""
    b, H = _coupled(evr, P12, 2, max_nvec=4)
    s, Hd = coupled_dense(oracle_mod, P12, 2)
    n = s.nb_packed
    rng = np.random.default_rng(3)
    p = np.zeros((2 * n, 2), complex)
    p[n, 0] = 1.0
    p[:, 1] = rng.standard_normal(2 * n) + 1j * rng.standard_normal(2 * n)
    p[:, 1] /= np.linalg.norm(p[:, 1])
    nsteps = 20
    got, _, _ = H.propagate(p, P12.DT, nsteps, npoly=P12.NPOLY, poly_tol=1e-12)
    ref = sl.expm(-1j * Hd * P12.DT * nsteps) @ p
    err = np.abs(got - ref).max()
    print("pyrazine L=2 GPU vs exact %.3e" % err)
    assert err < 1e-9


@pytest.mark.parametrize("graph", [False, True])
def test_pyrazine_field_driven(evr, oracle_mod, graph):
    """Coupled states driven by a field through a dipole operator of the
    states (mu_11 = 0.1 Q_2, mu_12 = mu_21 = 0.3 + 0.05 Q_1, type-0 blocks), 12D
    model, 30 steps with a piecewise-constant field per wavepacket, two
    wavepackets, adaptive and with captured steps: against
    exp(-i (H - E_n mu) dt) of the oracle's operators."""
    P = P12
    b, H = _coupled(evr, P)
    q = b.grid()
    m11, m12 = 0.1 * q[:, 1], 0.3 + 0.05 * q[:, 0]
    p11 = evr.Plan(b, type=0, V=m11, max_nvec=4)
    p12 = evr.Plan(b, type=0, V=m12, max_nvec=4)
    Mu = evr.Coupled([[p11, p12], [p12, None]])
    s, Hd = coupled_dense(oracle_mod, P)
    n = s.nb_packed
    I = np.eye(n)
    M11 = s.oppsi(oracle_mod.Op(0, V=m11), I)
    M12 = s.oppsi(oracle_mod.Op(0, V=m12), I)
    Md = np.block([[M11, M12], [M12, np.zeros((n, n))]])
    nsteps, dt = 30, P.DT
    E = 0.02 * np.sin(0.3 * np.arange(nsteps)[:, None] + np.arange(2)[None, :])
    rng = np.random.default_rng(8)
    p = np.zeros((2 * n, 2), complex)
    p[n, 0] = 1.0
    p[:, 1] = rng.standard_normal(2 * n) + 1j * rng.standard_normal(2 * n)
    p[:, 1] /= np.linalg.norm(p[:, 1])
    got, _, info = H.propagate(p, dt, nsteps, npoly=P.NPOLY, poly_tol=1e-12, field=E, mu=Mu,
                               graph=graph)
    ref = p.copy()
    for k in range(nsteps):
        for w in range(2):
            ref[:, w] = sl.expm(-1j * (Hd - E[k, w] * Md) * dt) @ ref[:, w]
    err = np.abs(got - ref).max()
    print("field-driven coupled states (graph=%s) vs exact %.3e (max order %d)" % (graph, err, info.max_order))
    assert err < 1e-9
    with pytest.raises(RuntimeError, match="evr_coupled_propagate_field"):
        H.propagate(p, dt, 2, field=E[:2])          # a field needs the coupled dipole


def test_pyrazine_davidson_levels(evr, oracle_mod):
    """Lowest 8 vibronic levels of the 12D pyrazine model on the L=2 basis (724
    coupled functions) by the block Davidson with the coupled diagonal as
    preconditioner, against the dense eigenvalues of the oracle's operator."""
    b, H = _coupled(evr, P12, 2, max_nvec=8)
    s, Hd = coupled_dense(oracle_mod, P12, 2)
    n = s.nb_packed
    assert relerr(H.diagonal(), np.diag(Hd)) < 1e-12
    ref = np.linalg.eigvalsh(0.5 * (Hd + Hd.T))[:8]
    ev, vec, info = H.davidson(8, block=8, conv_resi=1e-9, want_vectors=True)
    err = np.abs(ev - ref).max() * 219474.6313705      # hartree -> cm-1
    print("pyrazine 12D L=2 Davidson: max |dE| %.3e cm-1 in %d iterations" % (err, info.iterations))
    assert err < 1e-6
    r = Hd @ vec - vec * ev[None, :]
    assert np.abs(r).max() < 1e-7 and vec.shape == (2 * n, 8)


def test_pyrazine_davidson_guess_and_collapse(evr, oracle_mod):
    """Coupled-state Davidson started from user guess vectors (the dense
    eigenvectors, perturbed) with a user diagonal and a subspace of nev +
    block (a collapse every iteration), against the dense levels; FP32 is
    refused with its message."""
    b, H = _coupled(evr, P12, 2, max_nvec=4)
    s, Hd = coupled_dense(oracle_mod, P12, 2)
    Hs = 0.5 * (Hd + Hd.T)
    w, U = np.linalg.eigh(Hs)
    rng = np.random.default_rng(3)
    guess = U[:, :6] + 1e-3 * rng.standard_normal((Hd.shape[0], 6))
    ev, vec, info = H.davidson(6, block=4, max_sub=10, conv_resi=1e-9, guess=guess, diag=np.diag(Hd),
                               want_vectors=True)
    err = np.abs(ev - w[:6]).max() * 219474.6313705
    print("pyrazine guess + collapse: max |dE| %.3e cm-1 in %d iterations" % (err, info.iterations))
    assert info.converged == 1 and err < 1e-6
    with pytest.raises(RuntimeError, match="FP64 only"):
        H.davidson(6, block=4, precision=evr.FP32)
