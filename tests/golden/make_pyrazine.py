"""Writes pyrazine_12d.json and pyrazine_24d.json from the reference's 12- and
24-mode pyrazine propagation tests (run once where the reference is mounted;
the tests read only the JSON).  The MPI-scheme variants S1-S3 and the openMP
copies of each test hold the same benchmark file.

Source files (relative to the reference root; 24D_propagation_S0 alike, with
24 modes, L = 2 and B = (4, 4, 3, 3, 2, ..., 2)):
  Working_tests/MPI_tests/12D_propagation_S0/benchmark      autocorrelation
      <psi0|psi(t)> every 0.1 fs: time (au), Re, Im, |.|
  Working_tests/MPI_tests/12D_propagation_S0/sub_system_pyrazine.f:29-48
      the two-state vibronic model: frequencies w, gradients k1, k2, the
      half gap delta and the coupling lambda, all in eV; V is
      V11 = w.Q^2/2 - delta + k1.Q, V22 = w.Q^2/2 + delta + k2.Q,
      V12 = V21 = lambda Q_1 (lines 55-63), times 1/auTOeV
  Working_tests/MPI_tests/12D_propagation_S0/calc_f2_f1Q.f90:98-108
      kinetic part Tdef2(i,i) = -w_i/2 (eV -> au), no Tdef1
  Working_tests/MPI_tests/12D_propagation_S0/shell_run (+ run_jobs: L=1)
      SG4 LB = LG = 1, Hermite bases with L_TO_nq_A = 1, L_TO_nq_B =
      (3, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2), nb = nq; Chebyshev
      (type_WPpropa=1) npoly=300, poly_tol=1e-10, 60 steps of 0.1 fs
      (WPTmax 6 fs, nb_micro 20 per 2 fs); initial packet: the ground
      vibrational function on electronic state 2
  Source_PhysicalConstants/sub_module_constant.f90:786-805 (CODATA 2006) and
      :270-310: auTOeV = Eh/e and the atomic unit of time, computed here the
      same way.
"""
import json
import math
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

# CODATA 2006 as in the reference's constants module
c, mu0, h, e, me = 299792458.0, math.pi * 4e-7, 6.62606896e-34, 1.602176487e-19, 9.10938215e-31
eps0 = 1.0 / (mu0 * c * c)
hb = h / (2 * math.pi)
a0 = (hb / e) ** 2 / me * 4 * math.pi * eps0
Eh = (e * e / a0) / (4 * math.pi * eps0)


def fixture(test, out_name):
    T = os.path.join(REF, "Working_tests/MPI_tests", test)
    rows = [[float(x.replace("D", "E")) for x in l.split()] for l in open(os.path.join(T, "benchmark"))
            if l.strip()]
    src = open(os.path.join(T, "sub_system_pyrazine.f")).read()
    shell = open(os.path.join(T, "shell_run")).read()
    jobs = open(os.path.join(T, "run_jobs")).read()

    def block(name):
        """Values of the Fortran array constructor `name(ndim) = (/ ... /)`."""
        m = re.search(r"\b%s\(ndim\)\s*=\s*\(/(.*?)/\)" % name, src, re.S)
        return [0.0 if tok == "ZERO" else float(tok)
                for tok in re.findall(r"(ZERO|-?\d+\.\d*)(?:_Rkind)?", m.group(1))]

    def scalar(name):
        return float(re.search(r"\b%s\s*=\s*(-?\d+\.\d*)_Rkind" % name, src).group(1))

    def shell_int(name):
        return int(re.search(r"^%s=(\d+)" % name, shell, re.M).group(1))

    D = int(re.search(r"integer,\s*parameter\s*::\s*ndim=(\d+)", src).group(1))
    L = int(re.search(r"shell_run\s+\d+\s+(\d+)", jobs).group(1))
    B = [shell_int("B1"), shell_int("B2"), shell_int("B3"), shell_int("B4")] + [shell_int("B")] * (D - 4)
    tail = shell[shell.index("New_Read_WP0"):]
    wp0_line = re.search(r"^ *((?:\d+ +){%d}[-\d.]+ +[-\d.]+) *$" % (D + 2), tail, re.M).group(1)
    out = {
        "time_au": [r[0] for r in rows], "re": [r[1] for r in rows], "im": [r[2] for r in rows],
        "abs": [r[3] for r in rows],
        "w_eV": block("w"), "k1_eV": block("k1"), "k2_eV": block("k2"),
        "delta_eV": scalar("delta"), "lambda_eV": scalar("lambda"),
        "auTOeV": Eh / e, "fs_TO_au": 1e-15 / (hb / Eh),
        "basis": {"D": D, "L": L, "nqA": 1, "nqB": B, "nb_equals_nq": True},
        "propa": {"dt_fs": 0.1, "nsteps": 60, "npoly": 300, "poly_tol": 1e-10},
        "psi0": {"state": 2, "vibrational": "ground (all modes n=0)"},
        # the initial wavepacket as the input gives it (New_Read_WP0): D
        # 1-based function indices, i_h, i_e, Re, Im (lect_psiBasisRepnotall_nD)
        "wp0_line": wp0_line,
    }
    assert len(out["w_eV"]) == len(out["k1_eV"]) == len(out["k2_eV"]) == D
    json.dump(out, open(os.path.join(HERE, out_name), "w"), indent=0)
    print(out_name, ": D", D, "L", L, "B", B, len(rows), "autocorrelation values")


fixture("12D_propagation_S0", "pyrazine_12d.json")
fixture("24D_propagation_S0", "pyrazine_24d.json")
