"""Writes oracle_levels.json: lowest levels of the symmetric part (H + H^T)/2 of
the type-10 Hamiltonians of the benchmark generator, from dense matrices built
column by column with the CPU oracle (oracle/evr_oracle.c).  These are oracle
outputs cached for the GPU tests (building the dense C1 matrix takes a minute
or more); rerun this script to regenerate them.

  c0: direct product D=3, nb=12, nq=14 (configuration 0), 40 levels
  c1: SG4 D=6 L=5 (configuration 1), 120 levels
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "elvibrot-tnumtana_amd"))
import oracle  # noqa: E402
from evrhip import hamiltonian as H  # noqa: E402

out = {}
ham0 = H.config0()
d = oracle.DP([12, 12, 12], [14, 14, 14])
V0 = ham0.potential(d.grid())
M0 = d.oppsi(oracle.Op(10, V=V0, G0=ham0.G0), np.eye(d.nb_tot))
out["c0"] = np.linalg.eigvalsh(0.5 * (M0 + M0.T))[:40].tolist()

ham1 = H.smolyak_config(1, 6)
s = oracle.SG4(6, 5)
V1 = ham1.potential(s.grid())
M1 = s.oppsi(oracle.Op(10, V=V1, G0=ham1.G0, Ga=ham1.Ga_fortran()), np.eye(s.nb_packed))
out["c1"] = np.linalg.eigvalsh(0.5 * (M1 + M1.T))[:120].tolist()
json.dump(out, open(os.path.join(HERE, "oracle_levels.json"), "w"), indent=1)
print("wrote oracle_levels.json")
