"""The two-state pyrazine models of the reference's propagation tests
(Working_tests/MPI_tests/12D_propagation_S0 and 24D_propagation_S0), from the
committed fixtures golden/pyrazine_12d.json and golden/pyrazine_24d.json
(written by golden/make_pyrazine.py).

Model (sub_system_pyrazine.f:29-63, calc_f2_f1Q.f90:98-108), energies in eV
converted with 1/auTOeV:
  T   = sum_i -w_i/2 d^2/dQ_i^2            (type 1 with F2 = diag(-w/2))
  V11 = sum_i w_i Q_i^2/2 - delta + k1.Q
  V22 = sum_i w_i Q_i^2/2 + delta + k2.Q
  V12 = V21 = lambda Q_1
Basis: SG4 with LB = LG = L, Hermite (Q0 = 0, scaleQ = 1) with
nb_d(l) = nq_d(l) = 1 + B_d l (12D: L = 1, B = (3, 3, 2, ..., 2); 24D: L = 2,
B = (4, 4, 3, 3, 2, ..., 2)).  Initial packet: the ground vibrational function
on electronic state 2.  Chebyshev propagation, npoly 300, poly_tol 1e-10, 60
steps of 0.1 fs, autocorrelation <psi0|psi(t)>.
"""
import json
import os

import numpy as np

_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Pyrazine:
    def __init__(self, fname):
        self.fix = f = json.load(open(os.path.join(_GOLDEN, fname)))
        self.D = f["basis"]["D"]
        self.L = f["basis"]["L"]
        self.NQA = f["basis"]["nqA"]
        self.NQB = f["basis"]["nqB"]
        self.AU = f["auTOeV"]
        self.DT = f["propa"]["dt_fs"] * f["fs_TO_au"]
        self.NSTEPS = f["propa"]["nsteps"]
        self.NPOLY = f["propa"]["npoly"]
        self.POLY_TOL = f["propa"]["poly_tol"]

    def autocorr_ref(self):
        """(times in au, complex <psi0|psi(t_n)>) for n = 0..nsteps."""
        f = self.fix
        return np.array(f["time_au"]), np.array(f["re"]) + 1j * np.array(f["im"])

    def potentials(self, q):
        """V11, V22, V12 (au) on grid points q (npts x D)."""
        f = self.fix
        w, k1, k2 = np.array(f["w_eV"]), np.array(f["k1_eV"]), np.array(f["k2_eV"])
        h = 0.5 * (q * q) @ w
        V11 = (h - f["delta_eV"] + q @ k1) / self.AU
        V22 = (h + f["delta_eV"] + q @ k2) / self.AU
        V12 = f["lambda_eV"] * q[:, 0] / self.AU
        return V11, V22, V12

    def F2(self):
        return np.diag(-0.5 * np.array(self.fix["w_eV"]) / self.AU)

    def sg4_args(self, L=None):
        """(D, LG, LB, nbA, nbB, nqA, nqB) of the oracle's SG4 (nb = nq)."""
        L = self.L if L is None else L
        return self.D, L, L, self.NQA, self.NQB, self.NQA, self.NQB


P12 = Pyrazine("pyrazine_12d.json")
P24 = Pyrazine("pyrazine_24d.json")
