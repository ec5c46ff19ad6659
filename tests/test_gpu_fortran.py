"""The Fortran ISO_C_BINDING layer (elvibrot-tnumtana_amd/fortran) as a host.

evr_f_host builds a D=4 SG4 Henon-Heiles type-1 operator through evr_hip_f,
applies it to real and complex vectors, runs Davidson and 5 Chebyshev steps,
builds the assembled per-term operators and applies them again, transforms
the vectors to the grid and takes a grid derivative (the multi-vector and
DerivOp wrappers), and writes the results; the oracle recomputes all of them.
It also runs the control field update (build_field, Zhu-Rabitz and Krotov)
on a given dJ/dE through the Fortran interface.  evr_f_pyrazine
propagates the two-state 12D pyrazine model and writes the autocorrelation
file the reference's benchmark holds.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import propa
from conftest import ROOT, hh_potential

pytestmark = pytest.mark.gpu

HOST = os.path.join(ROOT, "elvibrot-tnumtana_amd", "fortran", "evr_f_host")


def test_fortran_host(oracle_mod, tmp_path):
    if not os.path.exists(HOST):
        if shutil.which("flang") is None and not os.path.exists("/opt/rocm/lib/llvm/bin/flang"):
            pytest.skip("no Fortran compiler: the Fortran host is not built")
        raise RuntimeError("evr_f_host is not built (run __graft_entry__.build())")
    out = tmp_path / "f_host.txt"
    r = subprocess.run([HOST, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    vals = open(out).read().split()
    nb, npts = int(vals[0]), int(vals[1])
    x = np.array([float(v) for v in vals[2:]])
    D = 4
    s = oracle_mod.SG4(D, 3, 3, 1, 2, 2, 2)
    assert nb == s.nb_packed and npts == s.sum_nq
    i = np.arange(1, nb + 1)[:, None]
    k = np.arange(1, 5)[None, :]
    psi = np.sin(0.37 * i + 1.3 * k) / (1.0 + 0.01 * i)
    op = oracle_mod.Op(1, V=hh_potential(s.grid()), F2=-0.5 * np.eye(D))
    ref = s.oppsi(op, psi)
    off = 0
    Hpsi = x[off:off + 4 * nb].reshape(4, nb).T
    off += 4 * nb
    assert np.abs(Hpsi - ref).max() < 1e-12 * np.abs(ref).max()
    zH = x[off:off + 4 * nb].reshape(2, nb, 2)
    off += 4 * nb
    zH = zH[..., 0] + 1j * zH[..., 1]
    zref = np.stack([ref[:, 0] + 1j * ref[:, 1], ref[:, 2] - 1j * ref[:, 3]])
    assert np.abs(zH - zref).max() < 1e-12 * np.abs(ref).max()
    ene = x[off:off + 8]
    off += 8
    M = s.oppsi(op, np.eye(nb))
    Hs = 0.5 * (M + M.T)
    ev = np.linalg.eigvalsh(Hs)
    assert np.abs(ene - ev[:8]).max() < 1e-9
    zp = x[off:off + 2 * nb].reshape(nb, 2)
    zp = zp[:, 0] + 1j * zp[:, 1]
    z0 = psi[:, 0] + 1j * psi[:, 1]
    z0 = z0 / np.linalg.norm(z0)
    ex = z0[:, None]
    for _ in range(5):
        ex = propa.exact_step(Hs, ex, 0.5)
    assert np.abs(zp - ex[:, 0]).max() < 1e-10
    off += 2 * nb
    Hasm = x[off:off + 4 * nb].reshape(4, nb).T
    assert np.abs(Hasm - ref).max() < 1e-12 * np.abs(ref).max()
    off += 4 * nb
    gd = x[off:off + 8 * npts].reshape(4, npts, 2)
    for c in range(4):
        g = s.b2g(psi[:, c])
        assert np.abs(gd[c, :, 0] - g).max() < 1e-12 * np.abs(g).max()
        dref = s.deriv(g, 0, 1)
        assert np.abs(gd[c, :, 1] - dref).max() < 1e-12 * np.abs(dref).max()
    # build_field through the Fortran interface: T = 100, 10 steps, alpha = 2,
    # envelope sin^2(pi t / 100) at the mid-steps; Krotov adds the old field
    f = np.loadtxt(str(out) + ".field").reshape(10, 2, 4)
    gE, Eold, Ezr, Ekr = (f[:, :, j] for j in range(4))
    dt = 10.0
    env = np.sin(np.pi * (np.arange(10) + 0.5) * dt / 100.0) ** 2
    assert np.abs(Ezr - env[:, None] * gE / (2 * 2.0 * dt)).max() < 1e-15
    assert np.abs(Ekr - (Eold + Ezr)).max() < 1e-15


PYRAZINE = os.path.join(ROOT, "elvibrot-tnumtana_amd", "fortran", "evr_f_pyrazine")


def test_fortran_pyrazine_benchmark(tmp_path):
    """evr_f_pyrazine (two coupled states through evr_coupled_* from Fortran)
    writes the reference's 12D pyrazine autocorrelation file; it must agree
    with the reference's benchmark to 1e-9."""
    from pyrazine import P12 as P
    if not os.path.exists(PYRAZINE):
        if shutil.which("flang") is None and not os.path.exists("/opt/rocm/lib/llvm/bin/flang"):
            pytest.skip("no Fortran compiler: the Fortran host is not built")
        raise RuntimeError("evr_f_pyrazine is not built (run __graft_entry__.build())")
    out = tmp_path / "pyrazine.txt"
    r = subprocess.run([PYRAZINE, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = np.loadtxt(out)
    t, ref = P.autocorr_ref()
    assert got.shape == (P.NSTEPS + 1, 4)
    assert np.abs(got[:, 0] - t).max() < 1e-12
    assert np.abs(got[:, 1] + 1j * got[:, 2] - ref).max() < 1e-9
    assert np.abs(got[:, 3] - np.abs(ref)).max() < 1e-9
    # the final wavepacket, saved from Fortran in the reference's version-0
    # format: unit norm and <psi0|psi(T)> = the last autocorrelation value
    import evrhip
    D, LG, LB, nbA, nbB, nqA, nqB = P.sg4_args()
    b = evrhip.Basis.sg4_modes(D, LG, nbA, nbB, nqA, nqB, LB=LB)
    psi = b.read_psi(str(out) + ".psi", nb0=2, cplx=True, max_wp=1)[:, 0]
    assert abs(np.linalg.norm(psi) - 1.0) < 1e-9
    assert abs(psi[b.nb_packed] - (got[-1, 1] + 1j * got[-1, 2])) < 1e-15
