"""Symmetric (hermitian) operator and block Davidson on the GPU, against the oracle."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, hh_potential

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_hermitian_form_is_symmetric_part(evr, oracle_mod):
    """hermitian=1 evaluates (H + H^T)/2 of the reference operator exactly."""
    from evrhip import hamiltonian as H
    D, L = 4, 4
    ham = H.smolyak_config(1, D)
    s = oracle_mod.SG4(D, L)
    V = ham.potential(s.grid())
    M = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), np.eye(s.nb_packed))
    Ms = 0.5 * (M + M.T)
    b = evr.Basis.sg4(D, L)
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=8, hermitian=True)
    X = evr.hamiltonian.test_vectors(b.nb_packed, 8, 4)
    assert relerr(plan.TabOpPsi(X), Ms @ X) < 1e-12


def test_plan_diagonal(evr, oracle_mod):
    from evrhip import hamiltonian as H
    D, L = 4, 3
    ham = H.smolyak_config(1, D)
    s = oracle_mod.SG4(D, L)
    V = ham.potential(s.grid())
    M = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), np.eye(s.nb_packed))
    plan = evr.Plan(evr.Basis.sg4(D, L), type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=1,
                    hermitian=True)
    # exact V part; kinetic part from factor diagonals (P, Q diagonals vanish)
    assert relerr(plan.diagonal(), np.diag(M)) < 1e-10


def test_davidson_c0_direct_product(evr):
    from evrhip import hamiltonian as H
    ref = np.array(json.load(open(os.path.join(GOLDEN, "oracle_levels.json")))["c0"])
    ham = H.config0()
    b = evr.Basis.dp([12] * 3, [14] * 3)
    plan = evr.Plan(b, type=10, G0=ham.G0, max_nvec=8, hermitian=True)
    plan.set_potential(ham.potential(b.grid()))
    ev, _, info = plan.davidson(20, block=8, conv_resi=1e-8, conv_ene=1e-9)
    print("C0 Davidson", info.iterations, info.hpsi_fp64, np.abs(ev - ref[:20]).max())
    assert info.converged == 1
    assert np.abs(ev - ref[:20]).max() < 1e-6   # cm^-1


def test_davidson_henon_heiles_6d(evr):
    """Type-1 SG4 Davidson reproduces the reference's 6D Henon-Heiles levels
    (Working_tests/MPI_tests/6D_Davidson_S0/benchmark; the reference converged
    its Ritz values to 5e-4 au residual, so they sit 0..2e-7 au above ours)."""
    bench = np.array(json.load(open(os.path.join(GOLDEN, "hh_benchmarks.json")))["6D_Davidson_S0"])
    D = 6
    b = evr.Basis.sg4(D, 3, 3, 1, 2, 1, 2)
    plan = evr.Plan(b, type=1, V=hh_potential(b.grid()), F2=-0.5 * np.eye(D), max_nvec=16)
    ev, _, info = plan.davidson(28, block=16, conv_resi=1e-9, conv_ene=1e-11)
    assert info.converged == 1
    diff = bench - ev
    assert diff.min() > -1e-9 and diff.max() < 2e-7


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_davidson_c1(evr, precision):
    """C1: D=6 SG4 L=5, 100 lowest levels, block 16, FP64 and FP32-refined."""
    from evrhip import hamiltonian as H
    ref = np.array(json.load(open(os.path.join(GOLDEN, "oracle_levels.json")))["c1"])[:100]
    ham = H.smolyak_config(1, 6)
    b = evr.Basis.sg4(6, 5)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=16, hermitian=True)
    plan.gen_potential_poly(ham)
    prec = evr.FP64 if precision == "fp64" else evr.FP32
    ev, _, info = plan.davidson(100, block=16, conv_resi=1e-6, conv_ene=1e-8, precision=prec,
                                refine_resi=1e-1)
    err = np.abs(ev - ref).max()
    print("C1 Davidson %s: it %d fp64 %d fp32 %d max|dE| %.2e cm-1 (%.2f s)" % (
        precision, info.iterations, info.hpsi_fp64, info.hpsi_fp32, err, info.seconds))
    assert info.converged == 1
    assert err < 1e-6
    if precision == "fp32":
        # the bulk of the H|psi> in FP32, FP64 only for the defect corrections
        assert info.hpsi_fp32 >= 0.6 * (info.hpsi_fp32 + info.hpsi_fp64)


def test_davidson_max_sub(evr, oracle_mod):
    """max_sub below nev + block is refused (a collapse must keep every wanted
    Ritz vector, DAV:516-606); at nev + block the subspace collapses every
    iteration and the levels still converge to the dense ones."""
    from evrhip import hamiltonian as H
    D, L = 4, 3
    ham = H.smolyak_config(1, D)
    s = oracle_mod.SG4(D, L)
    V = ham.potential(s.grid())
    M = s.oppsi(oracle_mod.Op(10, V=V, G0=ham.G0, Ga=ham.Ga_fortran()), np.eye(s.nb_packed))
    ref = np.linalg.eigvalsh(0.5 * (M + M.T))[:8]
    plan = evr.Plan(evr.Basis.sg4(D, L), type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=8,
                    hermitian=True)
    with pytest.raises(RuntimeError, match="max_sub must be at least nev"):
        plan.davidson(8, block=8, max_sub=9)
    for prec in (evr.FP64, evr.FP32):
        ev, _, info = plan.davidson(8, block=8, max_sub=16, conv_resi=1e-7, conv_ene=1e-9, precision=prec,
                                    max_it=400)
        print("max_sub 16 precision %d: it %d fp64 %d fp32 %d" % (prec, info.iterations, info.hpsi_fp64,
                                                                 info.hpsi_fp32))
        assert info.converged == 1
        assert np.abs(ev - ref).max() < 1e-8 * np.abs(ref).max()


def test_davidson_save_and_restart_files(evr, oracle_mod, tmp_path):
    """The Ritz vectors saved with sub_save_psi's format (save_file) restart a
    Davidson run (guess_file, read_WP of the reference): the second run
    converges in fewer iterations to the same levels."""
    from evrhip import hamiltonian as H
    D, L = 4, 3
    ham = H.smolyak_config(1, D)
    b = evr.Basis.sg4(D, L)
    s = oracle_mod.SG4(D, L)
    V = ham.potential(s.grid())
    plan = evr.Plan(b, type=10, V=V, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=8, hermitian=True)
    path = str(tmp_path / "dav_wp")
    ev1, _, info1 = plan.davidson(10, block=8, conv_resi=1e-3, conv_ene=1e-6, save_file=path)
    ev2, vec2, info2 = plan.davidson(10, block=8, conv_resi=1e-8, conv_ene=1e-10, guess_file=path,
                                     want_vectors=True)
    ev3, _, info3 = plan.davidson(10, block=8, conv_resi=1e-8, conv_ene=1e-10)
    print("restart: it %d (cold start %d)" % (info2.iterations, info3.iterations))
    assert info2.converged == 1
    assert np.abs(ev2 - ev3).max() < 1e-8 * np.abs(ev3).max()
    assert info2.iterations < info3.iterations
    got = b.read_psi(path, cplx=False, max_wp=20)
    assert got.shape[1] == 10


def test_davidson_graph_replays_match_eager(evr):
    """The captured H|psi> of the new-vector block (graph = 1) replayed every
    iteration, with the driver's own eager work in between, gives the same
    levels to the last bit as the eager driver, in FP64 and in the mixed
    precision (whose FP32 graph and FP64 corrections alternate)."""
    from evrhip import hamiltonian as H
    D, L = 5, 4
    ham = H.smolyak_config(9, D)
    b = evr.Basis.sg4(D, L)
    plan = evr.Plan(b, type=10, G0=ham.G0, Ga=ham.Ga_fortran(), max_nvec=8, hermitian=True)
    plan.gen_potential_poly(ham)
    for prec in (evr.FP64, evr.FP32):
        e0, _, i0 = plan.davidson(12, block=8, conv_resi=1e-6, conv_ene=1e-9, precision=prec, refine_resi=1e-1)
        e1, _, i1 = plan.davidson(12, block=8, conv_resi=1e-6, conv_ene=1e-9, precision=prec, refine_resi=1e-1,
                                  graph=True)
        print("graph: precision %d, %d iterations, %d / %d H|psi>" % (prec, i1.iterations, i1.hpsi_fp64,
                                                                      i1.hpsi_fp32))
        assert i0.converged == 1 and i1.converged == 1
        assert np.array_equal(e0, e1)
        assert (i0.iterations, i0.hpsi_fp64, i0.hpsi_fp32) == (i1.iterations, i1.hpsi_fp64, i1.hpsi_fp32)
