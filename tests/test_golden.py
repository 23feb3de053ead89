"""The committed fixtures of tests/golden/ (written by make_golden.py from the
oracle): the oracle must keep reproducing them (CPU), and the HIP library
must match them at the FP64 parity bar without the oracle (GPU)."""
import os

import numpy as np
import pytest

from golden.cases import CASES, inputs

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))


def site_rel(got, want):
    g = np.asarray(got).reshape(-1, 24)
    w = np.asarray(want).reshape(-1, 24)
    return float(np.max(np.linalg.norm(g - w, axis=1) / np.linalg.norm(w, axis=1)))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(name):
    from oracle.oracle import Oracle
    c = CASES[name]
    U, src = inputs(c)
    o = Oracle(c["dims"], c["ls"], c["action"], m5=1.8, b=1.5, c=0.5, gauge_canonical=U)
    f = load(name)
    if c["cg"]:
        sol, iters, tr, hist = o.invcg(src, c["mass"], stop_rsd=1e-10, max_iter=10000)
        assert iters == int(f["iters"])
        assert np.allclose(hist[:iters], f["residual_history"], rtol=1e-10, atol=0)
        assert abs(tr - float(f["true_res"])) < 1e-6 * tr
    else:
        mdm, dot = o.matpcdagmatpc(src, c["mass"])
        assert site_rel(mdm, f["matpcdagmatpc"]) < 1e-14
        assert abs(dot - float(f["dot_prd"])) < 1e-12 * dot


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_fixture(name):
    from cps_veloc_amd.dwf import CgArg, DiracOpDwfGpu, DiracOpMobiusGpu
    c = CASES[name]
    print(f"[gpu] golden fixture {name}", flush=True)
    U, src = inputs(c)
    if c["action"] == "dwf":
        g = DiracOpDwfGpu(U, c["dims"], c["ls"], m5=1.8)
    else:
        g = DiracOpMobiusGpu(U, c["dims"], c["ls"], m5=1.8, b=1.5, c=0.5)
    f = load(name)
    if c["cg"]:
        sol = np.zeros(g.n)
        iters, tr, st = g.InvCg(sol, src, CgArg(c["mass"], 10000, 1e-10))
        assert st == 0 and iters == int(f["iters"])
        assert abs(tr - float(f["true_res"])) < 0.05 * tr
    else:
        out = np.zeros(g.n)
        dot = g.MatPcDagMatPc(out, src, c["mass"])
        assert site_rel(out, f["matpcdagmatpc"]) < 1e-13
        assert abs(dot - float(f["dot_prd"])) < 1e-13 * dot
