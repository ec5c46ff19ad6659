"""More than 12 modes: the plan merges each term's one-function, one-point
modes into one scalar position of the term kernel.  Checked on the
reference's 21D Henon-Heiles Davidson test
(Working_tests/MPI_tests/21D_Davidson_S0: SG4 LB = LG = 2, nb = nq = 1 + 2 l)
against the oracle, which has no mode limit, and against the reference's own
levels; and on the basis <-> grid transforms and the plan diagonal of the 24D
pyrazine basis."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, hh_potential
from pyrazine import P24

pytestmark = pytest.mark.gpu

D21, L21 = 21, 2


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def hh21(evr, oracle_mod):
    s = oracle_mod.SG4(D21, L21, L21, nbA=1, nbB=2, nqA=1, nqB=2)
    b = evr.Basis.sg4(D21, L21, L21, 1, 2, 1, 2)
    V = hh_potential(s.grid())
    op = oracle_mod.Op(1, V=V, F2=-0.5 * np.eye(D21))
    plan = evr.Plan(b, type=1, V=V, F2=-0.5 * np.eye(D21), max_nvec=16)
    return s, b, V, op, plan


def test_hh21_oppsi(evr, oracle_mod, hh21):
    s, b, V, op, plan = hh21
    assert plan.nb_packed == s.nb_packed == 1 + 4 * D21 + 4 * D21 * (D21 - 1) // 2
    X = evr.hamiltonian.test_vectors(s.nb_packed, 5, 2)
    err = relerr(plan.TabOpPsi(X), s.oppsi(op, X))
    print("HH 21D H psi rel err %.3e" % err)
    assert err < 1e-12


def test_hh21_assembled(evr, hh21):
    s, b, V, op, plan = hh21
    X = evr.hamiltonian.test_vectors(s.nb_packed, 3, 4)
    ref = plan.TabOpPsi(X)
    plan.assemble(evr.ASM_FP64)
    try:
        assert relerr(plan.TabOpPsi(X), ref) < 1e-12
    finally:
        plan.release_assembled()


def test_hh21_davidson_reference_levels(evr, oracle_mod, hh21):
    """The 22 levels of the reference's 21D benchmark: the GPU Davidson levels
    equal the exact levels of the oracle's operator (1e-9 au), and the
    reference's Ritz values (converged to residual 5e-4 au) lie 0..2e-7 au
    above them."""
    s, b, V, op, plan = hh21
    bench = np.array(json.load(open(os.path.join(GOLDEN, "hh_benchmarks.json")))["21D_Davidson_S0"])
    H = s.oppsi(op, np.eye(s.nb_packed))
    exact = np.linalg.eigvalsh(0.5 * (H + H.T))[:len(bench)]
    ev, _, info = plan.davidson(len(bench), block=16, conv_resi=1e-9, conv_ene=1e-11)
    print("HH 21D Davidson: %d iterations, max |dE| %.3e au" % (info.iterations, np.abs(ev - exact).max()))
    assert np.abs(ev - exact).max() < 1e-9
    diff = bench - ev
    assert diff.min() > -1e-9 and diff.max() < 2e-7


def test_hh21_diagonal(evr, hh21):
    s, b, V, op, plan = hh21
    H = s.oppsi(op, np.eye(s.nb_packed))
    assert relerr(plan.diagonal(), np.diag(H)) < 1e-12


def test_pyrazine24_basis_grid(evr, oracle_mod):
    D, LG, LB, nbA, nbB, nqA, nqB = P24.sg4_args()
    s = oracle_mod.SG4(D, LG, LB, nbA=nbA, nbB=nbB, nqA=nqA, nqB=nqB)
    b = evr.Basis.sg4_modes(D, LG, nbA, nbB, nqA, nqB, LB=LB)
    V11, _, _ = P24.potentials(b.grid())
    plan = evr.Plan(b, type=1, V=V11, F2=P24.F2(), max_nvec=1)
    x = evr.hamiltonian.test_vectors(s.nb_packed, 1, 6)[:, 0]
    g = plan.BasisRep_TO_GridRep(x)
    assert relerr(g, s.b2g(x)) < 1e-12
    assert relerr(plan.GridRep_TO_BasisRep(g), s.g2b(g)) < 1e-12


def test_compaction_refusals(evr):
    """What the merged positions cannot express fails loudly."""
    b = evr.Basis.sg4(D21, L21, L21, 1, 2, 1, 2)
    G0 = np.eye(D21)
    Ga = np.zeros((D21, D21, D21))
    with pytest.raises(RuntimeError, match="Ga"):
        evr.Plan(b, type=10, G0=G0, Ga=Ga, max_nvec=1)
    p = evr.Plan(b, type=10, G0=G0, max_nvec=1)
    from evrhip import hamiltonian
    with pytest.raises(RuntimeError, match="polynomial generator"):
        p.gen_potential_poly(hamiltonian.smolyak_config(1, D21))
    # nq = 2 at level 0: every mode keeps two grid points, 21 > 11 positions
    b2 = evr.Basis.sg4(D21, 1, 1, 1, 1, 2, 1)
    with pytest.raises(RuntimeError, match="at most 11 modes"):
        evr.Plan(b2, type=1, F2=-0.5 * np.eye(D21), max_nvec=1)
