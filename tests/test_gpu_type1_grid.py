"""Type 1 with the operator grids per grid point (evr_plan_set_type1_grid: the
reference's get_OpGrid_type1 tables, combined as OPSG4:1490-1546) against the
oracle, 1e-12 relative max-norm (FP32 2e-5), and its hermitian form (H + H^T) / 2;
a constant table equals the constant-F2
basis-space path; a Henon-Heiles operator written as per-point grids gives the
reference's 6D Davidson levels."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, hh_potential
from test_oracle import type1_grids

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("D,L,nvec,with_f1", [(3, 5, 4, True), (4, 4, 6, True), (6, 4, 3, False)])
def test_type1_grid_vs_oracle(evr, oracle_mod, D, L, nvec, with_f1):
    s = oracle_mod.SG4(D, L)
    b = evr.Basis.sg4(D, L)
    q = s.grid()
    F2, F1 = type1_grids(q, 40 + D)
    if not with_f1:
        F1 = None
    V = 0.5 * (q ** 2).sum(1) + 0.03 * q[:, 0] ** 3
    plan = evr.Plan(b, type=1, V=V, F2=-0.5 * np.eye(D), max_nvec=nvec)
    plan.set_type1_grid(F2, F1)
    X = evr.hamiltonian.test_vectors(s.nb_packed, nvec, 30 + D)
    ref = s.oppsi(oracle_mod.Op(1, V=V, F2q=F2, F1q=F1), X)
    err = relerr(plan.TabOpPsi(X), ref)
    err32 = relerr(plan.TabOpPsi(X, evr.FP32), ref)
    print("type 1 per-point grids D=%d L=%d: rel err %.3e, FP32 %.3e" % (D, L, err, err32))
    assert err < 1e-12
    assert err32 < 2e-5
    with pytest.raises(RuntimeError, match="not assembled"):
        plan.assemble(evr.ASM_FP64)
    plan.close()


def test_type1_constant_grid_equals_basis_space(evr):
    D, L = 5, 4
    b = evr.Basis.sg4(D, L)
    q = b.grid()
    rng = np.random.default_rng(4)
    F2c = -0.5 * np.diag(rng.uniform(0.8, 1.2, D)) + 0.02 * rng.standard_normal((D, D))
    F1c = 0.05 * rng.standard_normal(D)
    V = 0.5 * (q ** 2).sum(1)
    const = evr.Plan(b, type=1, V=V, F2=F2c, F1=F1c, max_nvec=4)
    grid = evr.Plan(b, type=1, V=V, F2=F2c, max_nvec=4)
    grid.set_type1_grid(np.broadcast_to(F2c, (len(q), D, D)), np.broadcast_to(F1c, (len(q), D)))
    X = evr.hamiltonian.test_vectors(b.nb_packed, 4, 5)
    err = relerr(grid.TabOpPsi(X), const.TabOpPsi(X))
    print("type 1 constant grid vs basis-space form: %.3e" % err)
    assert err < 1e-12


def test_type1_grid_henon_heiles_levels(evr):
    """The reference's 6D Henon-Heiles test (Type_HamilOp=1, G = I: F2 = -1/2
    on the diagonal at every point) through the per-point path: its Davidson
    levels (Working_tests/MPI_tests/6D_Davidson_S0/benchmark) sit 0..2e-7 au
    above ours, as Ritz values converged to 5e-4 au must."""
    bench = np.array(json.load(open(os.path.join(GOLDEN, "hh_benchmarks.json")))["6D_Davidson_S0"])
    D = 6
    b = evr.Basis.sg4(D, 3, 3, 1, 2, 1, 2)
    q = b.grid()
    F2 = np.zeros((len(q), D, D))
    F2[:, np.arange(D), np.arange(D)] = -0.5
    plan = evr.Plan(b, type=1, V=hh_potential(q), F2=-0.5 * np.eye(D), max_nvec=16)
    plan.set_type1_grid(F2)
    dg = plan.diagonal()
    ev, _, info = plan.davidson(28, block=16, conv_resi=1e-9, conv_ene=1e-11, diag=dg)
    assert info.converged == 1
    diff = bench - ev
    print("per-point type 1 HH 6D: levels - reference in [%.2e, %.2e]" % (diff.min(), diff.max()))
    assert diff.min() > -1e-9 and diff.max() < 2e-7


def test_type1_grid_hermitian(evr, oracle_mod):
    """hermitian = 1 with F2, F1 per point: (H + H^T) / 2 of the reference's
    type-1 grid form (the transpose as W diag(1/w) Op^T diag(w) B), against
    the oracle's dense operator (FP64 1e-12, FP32 2e-5) and its diagonal."""
    D, L = 3, 4
    s = oracle_mod.SG4(D, L)
    b = evr.Basis.sg4(D, L)
    q = s.grid()
    F2, F1 = type1_grids(q, 47)
    V = 0.5 * (q ** 2).sum(1) + 0.03 * q[:, 0] ** 3
    n = s.nb_packed
    M = s.oppsi(oracle_mod.Op(1, V=V, F2q=F2, F1q=F1), np.eye(n))
    Hs = 0.5 * (M + M.T)
    plan = evr.Plan(b, type=1, V=V, F2=-0.5 * np.eye(D), max_nvec=6, hermitian=True)
    plan.set_type1_grid(F2, F1)
    X = evr.hamiltonian.test_vectors(n, 6, 9)
    ref = Hs @ X
    err = relerr(plan.TabOpPsi(X), ref)
    err32 = relerr(plan.TabOpPsi(X, evr.FP32), ref)
    errd = relerr(plan.diagonal(), np.diag(Hs))
    asym = np.abs(M - M.T).max() / np.abs(M).max()
    print("hermitian type-1 grids: %.3e, FP32 %.3e, diagonal %.3e (asymmetry of H %.2e)" % (err, err32, errd, asym))
    assert asym > 1e-6
    assert err < 1e-12 and errd < 1e-12 and err32 < 2e-5
