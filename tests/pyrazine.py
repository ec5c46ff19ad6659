"""The 12-mode, two-state pyrazine model of the reference's propagation test
(Working_tests/MPI_tests/12D_propagation_S0), from the committed fixture
golden/pyrazine_12d.json (written by golden/make_pyrazine.py).

Model (sub_system_pyrazine.f:29-63, calc_f2_f1Q.f90:98-108), energies in eV
converted with 1/auTOeV:
  T   = sum_i -w_i/2 d^2/dQ_i^2            (type 1 with F2 = diag(-w/2))
  V11 = sum_i w_i Q_i^2/2 - delta + k1.Q
  V22 = sum_i w_i Q_i^2/2 + delta + k2.Q
  V12 = V21 = lambda Q_1
Basis: SG4 with LB = LG = 1, Hermite (Q0 = 0, scaleQ = 1) with
nb_d(l) = nq_d(l) = 1 + B_d l, B = (3, 3, 2, ..., 2).  Initial packet: the
ground vibrational function on electronic state 2.  Chebyshev propagation,
npoly 300, poly_tol 1e-10, 60 steps of 0.1 fs, autocorrelation <psi0|psi(t)>.
"""
import json
import os

import numpy as np

FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                  "pyrazine_12d.json")))
D = FIX["basis"]["D"]
L = FIX["basis"]["L"]
NQA = FIX["basis"]["nqA"]
NQB = FIX["basis"]["nqB"]
AU = FIX["auTOeV"]
DT = FIX["propa"]["dt_fs"] * FIX["fs_TO_au"]
NSTEPS = FIX["propa"]["nsteps"]
NPOLY = FIX["propa"]["npoly"]
POLY_TOL = FIX["propa"]["poly_tol"]


def autocorr_ref():
    """(times in au, complex <psi0|psi(t_n)>) for n = 0..nsteps."""
    return np.array(FIX["time_au"]), np.array(FIX["re"]) + 1j * np.array(FIX["im"])


def potentials(q):
    """V11, V22, V12 (au) on grid points q (npts x D)."""
    w = np.array(FIX["w_eV"])
    k1 = np.array(FIX["k1_eV"])
    k2 = np.array(FIX["k2_eV"])
    h = 0.5 * (q * q) @ w
    V11 = (h - FIX["delta_eV"] + q @ k1) / AU
    V22 = (h + FIX["delta_eV"] + q @ k2) / AU
    V12 = FIX["lambda_eV"] * q[:, 0] / AU
    return V11, V22, V12


def F2():
    return np.diag(-0.5 * np.array(FIX["w_eV"]) / AU)
