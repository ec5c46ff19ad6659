"""Writes the golden fixtures of this directory from the reference's own data
files (run once in a container where the reference is mounted; the fixtures are
committed and the tests never read the reference).

  herm_tables.json   Internal_data/HermQuadra/herm{1..40}.txt: nodes and w*exp(x^2)
                     (column 4, the value sub_quadra_hermite stores in base%w)
  hh_benchmarks.json Working_tests/MPI_tests/{6D,21D}_Davidson_S0/benchmark levels
"""
import json
import os
import sys
from decimal import Decimal

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

tables = {}
for nq in range(1, 41):
    rows = [l.split() for l in open(os.path.join(REF, "Internal_data/HermQuadra/herm%d.txt" % nq))
            if l.strip()]
    tables[str(nq)] = {"x": [repr(float(Decimal(r[1]))) for r in rows],
                       "wex2": [repr(float(Decimal(r[3]))) for r in rows]}
json.dump(tables, open(os.path.join(HERE, "herm_tables.json"), "w"), indent=0)

bench = {}
for name in ("6D_Davidson_S0", "21D_Davidson_S0"):
    lv = [float(l.split()[0]) for l in open(os.path.join(REF, "Working_tests/MPI_tests", name, "benchmark"))
          if l.strip()]
    bench[name] = lv
bench["_source"] = {"6D_Davidson_S0": "HenonHeiles D=6, SG4 LB=LG=3, L_TO_nq_A=1, L_TO_nq_B=2, "
                                      "lambda=0.111803, Type_HamilOp=1, conv_resi=5e-4 au",
                    "21D_Davidson_S0": "HenonHeiles D=21, SG4 LB=LG=2, L_TO_nq_A=1, L_TO_nq_B=2, "
                                       "lambda=0.111803, Type_HamilOp=1, conv_resi=5e-4 au"}
json.dump(bench, open(os.path.join(HERE, "hh_benchmarks.json"), "w"), indent=1)
print("wrote fixtures")
