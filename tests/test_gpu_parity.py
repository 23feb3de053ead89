"""GPU parity: the HIP library through its C ABI against the CPU oracle.

Bar (FP64 path): every site of every operator output within 1e-13 relative
(|gpu - oracle|_site / |oracle|_site), dot products within 1e-13, CG with the
same iteration count as the oracle's InvCgShift and a true residual below the
stop target.  FP32 storage: 2e-6 per site.
"""
import sys

import numpy as np
import pytest

from cps_veloc_amd import lattice as lat
from cps_veloc_amd.dwf import CgArg, DiracOpDwfGpu, DiracOpMobiusGpu
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

M5 = 1.8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def site_rel(got, want):
    g = np.asarray(got).reshape(-1, 24)
    w = np.asarray(want).reshape(-1, 24)
    nw = np.linalg.norm(w, axis=1)
    return float(np.max(np.linalg.norm(g - w, axis=1) / np.maximum(nw, 1e-300)))


def make(action, dims, ls, eps=0.3, seed=101, precision="f64"):
    U = lat.random_gauge(dims, eps, seed)
    o = Oracle(dims, ls, action, m5=M5, b=1.5, c=0.5, gauge_canonical=U)
    if action == "dwf":
        g = DiracOpDwfGpu(U, dims, ls, m5=M5, precision=precision)
    else:
        g = DiracOpMobiusGpu(U, dims, ls, m5=M5, b=1.5, c=0.5, precision=precision)
    return U, o, g


def src_vec(dims, ls, seed):
    return lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), seed))


CASES = [
    ("dwf", (4, 4, 4, 4), 4),
    ("dwf", (4, 2, 6, 4), 6),      # 48 cb sites: a partial work-group
    ("dwf", (8, 8, 8, 8), 8),      # the 8^3x8, Ls=8 lattice
    ("mobius", (4, 4, 4, 4), 4),
    ("mobius", (4, 2, 6, 4), 6),
    ("mobius", (8, 8, 4, 8), 12),
    ("mobius", (4, 4, 4, 4), 8),   # the M5^-1 column swept in registers: Ls 8,
    ("mobius", (4, 4, 2, 4), 16),  # 12 and 16 (other Ls take the LDS loop)
    ("mobius", (4, 4, 4, 4), 2),   # Ls 2: slices s-1 and s+1 are one slice
    ("mobius", (4, 4, 2, 4), 24),  # Ls 24: the s-columns need 8 sites per work-group
    ("dwf", (6, 2, 2, 2), 4),      # 24 cb sites: the second work-group is half
    ("mobius", (6, 2, 2, 2), 4),   # empty (inactive lanes), odd Lx/2
]


@pytest.mark.parametrize("action,dims,ls", CASES)
def test_matpcdagmatpc(action, dims, ls):
    say(f"[gpu] MatPcDagMatPc {action} {dims} ls={ls}")
    U, o, g = make(action, dims, ls)
    x = src_vec(dims, ls, 7)
    for mass in (0.04, 0.3):
        want, dwant = o.matpcdagmatpc(x, mass)
        got = np.zeros(g.n)
        d = g.MatPcDagMatPc(got, x, mass)
        assert site_rel(got, want) < 1e-13
        assert abs(d - dwant) / dwant < 1e-13


@pytest.mark.parametrize("action,dims,ls", CASES)
@pytest.mark.parametrize("dag", [0, 1])
def test_matpc(action, dims, ls, dag):
    say(f"[gpu] MatPc dag={dag} {action} {dims} ls={ls}")
    U, o, g = make(action, dims, ls)
    x = src_vec(dims, ls, 9)
    want = o.matpc(x, 0.05, dag)
    got = np.zeros(g.n)
    (g.MatPcDag if dag else g.MatPc)(got, x, 0.05)
    assert site_rel(got, want) < 1e-13


@pytest.mark.parametrize("action,dims,ls", CASES[:2] + CASES[3:5] + CASES[-2:])
@pytest.mark.parametrize("cb", [0, 1])
@pytest.mark.parametrize("dag", [0, 1])
def test_dslash(action, dims, ls, cb, dag):
    say(f"[gpu] Dslash cb={cb} dag={dag} {action} {dims} ls={ls}")
    U, o, g = make(action, dims, ls)
    x = src_vec(dims, ls, 11)
    want = o.dslash(x, 0.1, cb, dag)
    got = np.zeros(g.n)
    g.Dslash(got, x, 0.1, cb, dag)
    assert site_rel(got, want) < 1e-13


@pytest.mark.parametrize("action,dims,ls,mass", [
    ("dwf", (8, 8, 8, 8), 8, 0.04),        # 8^3x8, Ls=8, M5=1.8, mf=0.04
    ("mobius", (4, 4, 4, 8), 8, 0.04),
    ("mobius", (4, 2, 6, 4), 6, 0.1),
    ("dwf", (6, 2, 2, 2), 4, 0.1),
])
def test_invcg_matches_oracle(action, dims, ls, mass):
    say(f"[gpu] InvCg {action} {dims} ls={ls} mf={mass}")
    U, o, g = make(action, dims, ls)
    src = src_vec(dims, ls, 13)
    osol, oit, otr, _ = o.invcg(src, mass, stop_rsd=1e-10, max_iter=5000)
    gsol = np.zeros(g.n)
    git, gtr, st = g.InvCg(gsol, src, CgArg(mass, 5000, 1e-10))
    say(f"  oracle iters={oit} true_res={otr:.3e}  gpu iters={git} true_res={gtr:.3e}")
    assert st == 0
    assert git == oit
    assert gtr < 2e-10 and abs(gtr - otr) < 0.05 * otr + 1e-14
    assert np.linalg.norm(gsol - osol) / np.linalg.norm(osol) < 1e-8


def test_invcg_edges():
    say("[gpu] InvCg edge cases")
    dims, ls, mass = (4, 4, 4, 4), 4, 0.1
    U, o, g = make("dwf", dims, ls)
    src = src_vec(dims, ls, 17)
    sol = np.zeros(g.n)
    it, tr, st = g.InvCg(sol, src, CgArg(mass, 5, 1e-14))
    assert (it, st) == (5, 2)        # MAXITER, CPS count = max_num_iter
    _, oit, otr, _ = o.invcg(src, mass, stop_rsd=1e-14, max_iter=5)
    assert oit == 5 and abs(tr - otr) < 1e-6 * otr
    sol2 = np.zeros(g.n)
    g.InvCg(sol2, src, CgArg(mass, 1000, 1e-11))
    it0, tr0, st0 = g.InvCg(sol2.copy(), src, CgArg(mass, 1000, 1e-6))
    assert (it0, st0) == (1, 0)      # guess already converged


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_matinv_matches_oracle(action):
    say(f"[gpu] MatInv {action}")
    dims, ls, mass = (4, 4, 4, 4), 4, 0.1
    U, o, g = make(action, dims, ls)
    src = np.concatenate([src_vec(dims, ls, 19), src_vec(dims, ls, 20)])
    want, oit, otr = o.matinv(src, mass, stop_rsd=1e-11, max_iter=3000)
    got = np.zeros(2 * g.n)
    keep = src.copy()
    git, gtr, st = g.MatInv(got, src, CgArg(mass, 3000, 1e-11))
    assert np.array_equal(src, keep)
    assert git == oit
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-9


@pytest.mark.parametrize("action,dims,ls", [
    ("dwf", (4, 4, 4, 8), 8),
    ("mobius", (4, 4, 4, 8), 8),
    ("mobius", (6, 2, 2, 2), 4),   # a half-empty work-group in the slice exchange
    ("mobius", (4, 4, 4, 4), 2),
])
def test_fp32_storage(action, dims, ls):
    say(f"[gpu] FP32 storage {action} {dims} ls={ls}")
    U, o, g = make(action, dims, ls, precision="f32")
    x = src_vec(dims, ls, 23)
    want, dwant = o.matpcdagmatpc(x, 0.04)
    got = np.zeros(g.n)
    d = g.MatPcDagMatPc(got, x, 0.04)
    assert site_rel(got, want) < 2e-5
    assert abs(d - dwant) / dwant < 1e-5


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_invcg_fp32_handle(action):
    """InvCg on a handle with FP32 storage: the same CG (r-only hop epilogue,
    then x and p in one pass) converges to a stop target FP32 can reach, in
    about the oracle's FP64 iteration count for that target."""
    say(f"[gpu] InvCg FP32 handle {action}")
    dims, ls, mass = (4, 4, 4, 8), 8, 0.04
    U, o, g = make(action, dims, ls, precision="f32")
    src = src_vec(dims, ls, 31)
    osol, oit, otr, _ = o.invcg(src, mass, stop_rsd=1e-5, max_iter=2000)
    gsol = np.zeros(g.n)
    git, gtr, st = g.InvCg(gsol, src, CgArg(mass, 2000, 1e-5))
    say(f"  oracle iters={oit}  gpu iters={git} true_res={gtr:.3e}")
    assert st == 0
    assert abs(git - oit) <= max(2, oit // 10)
    assert gtr < 2e-5
    assert np.linalg.norm(gsol - osol) / np.linalg.norm(osol) < 1e-3


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_device_vectors(action):
    say(f"[gpu] device vectors {action}")
    dims, ls = (4, 4, 4, 4), 4
    U, o, g = make(action, dims, ls)
    x = src_vec(dims, ls, 29)
    vin, vout = g.vector(x), g.vector()
    assert np.array_equal(vin.download(), x)
    d = g.MatPcDagMatPc_dev(vout, vin, 0.04)
    host = np.zeros(g.n)
    dh = g.MatPcDagMatPc(host, x, 0.04)
    assert np.array_equal(vout.download(), host) and d == dh
    ms, msk, byk = g.time_MatPcDagMatPc(vout, vin, 0.04, 3)
    assert ms > 0 and len(msk) == 5
    assert np.array_equal(vout.download(), host)


@pytest.mark.parametrize("action,dims,ls", [("dwf", (4, 4, 4, 8), 8), ("mobius", (4, 2, 6, 4), 6)])
def test_minvcg_matches_oracle(action, dims, ls):
    say(f"[gpu] MInvCG {action} {dims} ls={ls}")
    U, o, g = make(action, dims, ls)
    chi = src_vec(dims, ls, 41)
    mass = 0.05
    shifts = np.array([0.0, 0.003, 0.02, 0.1, 0.6])
    gp, gk, grs, git, gst = g.MInvCG(chi, mass, shifts, 1e-10)
    op, ok, ors, oit = o.minvcg(chi, mass, shifts, 1e-10)
    say(f"  k gpu={gk} oracle={ok}; per-shift iters gpu={git.tolist()} oracle={oit.tolist()}")
    assert gst == 0 and gk == ok and np.array_equal(git, oit)
    assert np.allclose(grs, ors, rtol=1e-3)
    for s in range(len(shifts)):
        assert np.linalg.norm(gp[s] - op[s]) / np.linalg.norm(op[s]) < 1e-9
    # SINGLE: the alpha sum, added to the vector passed in
    alpha = np.array([0.5, -1.0, 2.0, 0.25, 1.5])
    start = src_vec(dims, ls, 42)
    sp, sk, _, _, sst = g.MInvCG(chi, mass, shifts, 1e-10, multi=False, alpha=alpha, psi=start)
    assert sst == 0 and sk == gk
    want = start + (alpha[:, None] * gp).sum(0)
    assert np.linalg.norm(sp - want) / np.linalg.norm(want) < 1e-10
