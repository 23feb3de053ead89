"""Writes the golden fixtures of tests/golden/ from the oracle.

CPS ships no vectors for this path and cannot be built here, so these
fixtures record the oracle's own answers (FP64, CPS layouts) for small seeded
inputs; they pin the oracle against later changes and let GPU tests check
the library without the oracle library present.  Inputs are regenerated
from their seeds by cps_veloc_amd.lattice (SplitMix64 / Box-Muller, the
same generator on every machine).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from cases import CASES, inputs  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

def main():
    for name, c in CASES.items():
        U, src = inputs(c)
        o = Oracle(c["dims"], c["ls"], c["action"], m5=1.8, b=1.5, c=0.5, gauge_canonical=U)
        out = {}
        if c["cg"]:
            sol, iters, tr, hist = o.invcg(src, c["mass"], stop_rsd=1e-10, max_iter=10000)
            out.update(iters=iters, true_res=tr, residual_history=hist[:iters],
                       solution_checksum=np.array([np.sum(sol), np.dot(sol, sol)]))
        else:
            mdm, dot = o.matpcdagmatpc(src, c["mass"])
            out.update(matpcdagmatpc=mdm, dot_prd=dot)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, {k: (v if np.ndim(v) == 0 else np.shape(v)) for k, v in out.items()})


if __name__ == "__main__":
    main()
