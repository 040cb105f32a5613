"""Writes tests/golden/ref_solvers.npz: outputs of the reference's own
src/eikon_solver_2d.f90 (FSM_UW_PS_lonlat_2d, FSM_O1_JSE_lonlat_2d, mask),
compiled as-is by oracle/Makefile into oracle/_ref/libsatt_ref.so, on the
deterministic cases of tests/golden/cases.py.

Run in the container that holds the reference tree:
    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)
import oracle  # noqa: E402
from cases import CASES, solver_case  # noqa: E402


def main():
    ref = oracle.RefSolver()
    out = {}
    for name in CASES:
        c = solver_case(name)
        T = ref.fwd(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], c["fun"], c["x0"], c["y0"])
        Ta = ref.adj(c["xx"], c["yy"], c["spha"], c["sphb"], c["sphc"], T, c["xr"], c["yr"], c["sa"])
        Tm = ref.mask(c["xx"], c["yy"], Ta, c["x0"], c["y0"])
        out[f"{name}/T"] = T
        out[f"{name}/Ta"] = Ta
        out[f"{name}/Ta_masked"] = Tm
        # inputs, to detect any drift of the case generator
        for k in ("spha", "sphb", "sphc", "fun", "xr", "yr", "sa"):
            out[f"{name}/in_{k}"] = c[k]
    np.savez_compressed(os.path.join(HERE, "ref_solvers.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_solvers.npz"))


if __name__ == "__main__":
    main()
