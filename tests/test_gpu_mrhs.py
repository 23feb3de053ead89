"""GPU: the batched many-RHS operator, CG and MatInv (DWF and Mobius).

Bars
  * batched FP64 MatPcDagMatPc of every RHS == the single-RHS FP64 operator
    bit for bit (same arithmetic per site), each RHS with its own mass;
    FP32 / int16 batched operator against the FP64 oracle within the
    reduced-precision bounds of tests/test_gpu_mixed.py;
  * batched reliable-update CG: every RHS reaches |src - A x| <= stop_rsd |src|
    (checked with the FP64 operator), its solution agrees with the oracle's
    solve to 1e-8, masses give different iteration counts (RHS retire on their
    own), its count is within 2 of the single-RHS solver's and within 3 of the
    oracle's, and the graph-replayed solve equals the eagerly launched one bit
    for bit, iteration counts included;
  * batched MatInv (DiracOpMobius::MatInv's vector form, and DWF): every RHS's
    full solution within 1e-8 of the single-RHS MatInv (FP64 CG) and the
    sources left unchanged.
"""
import sys

import numpy as np
import pytest

from cps_veloc_amd import lattice as lat
from cps_veloc_amd.dwf import BAD_ARG, CgArg, DiracOpDwfGpu, DiracOpMobiusGpu, DwfGpuError
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

M5 = 1.8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def src_vec(dims, ls, seed):
    return lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), seed))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def setup(dims, ls, nrhs, seed=3, action="dwf"):
    U = lat.random_gauge(dims, 0.3, 101)
    if action == "dwf":
        g = DiracOpDwfGpu(U, dims, ls, m5=M5)
    else:
        g = DiracOpMobiusGpu(U, dims, ls, m5=M5, b=1.5, c=0.5)
    o = Oracle(dims, ls, action, m5=M5, b=1.5, c=0.5, gauge_canonical=U)
    xs = [src_vec(dims, ls, seed + j) for j in range(nrhs)]
    return g, o, xs


@pytest.mark.parametrize("action", ["dwf", "mobius"])
@pytest.mark.parametrize("dims,ls,nrhs", [((4, 4, 4, 8), 8, 5), ((4, 2, 6, 4), 6, 3)])
def test_batched_operator(dims, ls, nrhs, action):
    say(f"[gpu] batched MatPcDagMatPc {action} {dims} ls={ls} nrhs={nrhs}")
    g, o, xs = setup(dims, ls, nrhs, action=action)
    masses = np.array([0.01, 0.04, 0.3, 0.005, 0.1][:nrhs])
    vin, vout = g.multi_vector(nrhs, xs), g.multi_vector(nrhs)
    g.MatPcDagMatPc_mrhs_dev(vout, vin, masses, "f64")
    for j in range(nrhs):
        single = np.zeros(g.n)
        g.MatPcDagMatPc(single, xs[j], masses[j])
        assert np.array_equal(vout.download(j), single), j
    for prec, tol in (("f32", 1e-5), ("i16", 2e-4)):
        g.MatPcDagMatPc_mrhs_dev(vout, vin, masses, prec)
        for j in range(nrhs):
            want, _ = o.matpcdagmatpc(xs[j], masses[j])
            assert rel(vout.download(j), want) < tol, (prec, j)


@pytest.mark.parametrize("action", ["dwf", "mobius"])
@pytest.mark.parametrize("sloppy", ["f32", "i16"])
def test_batched_cg(sloppy, action):
    dims, ls = (4, 4, 4, 8), 8
    masses = np.array([0.01, 0.04, 0.5, 0.04])
    nrhs = len(masses)
    say(f"[gpu] batched CG {action} sloppy={sloppy} nrhs={nrhs} masses={masses.tolist()}")
    g, o, xs = setup(dims, ls, nrhs, seed=11, action=action)
    arg = CgArg(0.0, 5000, 1e-10)
    res = {}
    for eager in (False, True):
        sol, src = g.multi_vector(nrhs), g.multi_vector(nrhs, xs)
        it, tr, st, nrel = g.InvCgMrhs_dev(sol, src, arg, masses, sloppy, 0.1, eager=eager)
        say(f"  eager={eager} iters={it.tolist()} reliable={nrel.tolist()} "
            f"true_res={np.array2string(tr, precision=2)} shape={g.mrhs_shape()}")
        res[eager] = (it.copy(), tr.copy(), [sol.download(j) for j in range(nrhs)])
        assert np.all(st == 0) and np.all(tr <= 1e-10) and np.all(nrel >= 1)
    # graph replay == eager launches, bit for bit
    assert np.array_equal(res[False][0], res[True][0])
    assert np.array_equal(res[False][1], res[True][1])
    for a, b in zip(res[False][2], res[True][2]):
        assert np.array_equal(a, b)
    it = res[False][0]
    assert it[2] < it[0]  # the heavy RHS retires first, the others go on
    for j in range(nrhs):
        x = res[False][2][j]
        ax = np.zeros(g.n)
        g.MatPcDagMatPc(ax, x, masses[j])
        assert rel(ax, xs[j]) <= 1.0001e-10
        osol, oit, _, _ = o.invcg(xs[j], masses[j], stop_rsd=1e-10, max_iter=5000)
        assert rel(x, osol) < 1e-8
        # the single-RHS reliable-update CG of the same RHS: within 2 steps
        s1 = np.zeros(g.n)
        i1, _, st1, _ = g.InvCgMixed(s1, xs[j], CgArg(masses[j], 5000, 1e-10), sloppy, 0.1)
        assert st1 == 0 and abs(int(it[j]) - i1) <= 2, (j, it[j], i1)
        assert it[j] <= oit + 3


def test_batched_cg_maxiter_and_args():
    say("[gpu] batched CG: MAXITER per RHS and argument checks")
    dims, ls = (4, 4, 4, 4), 4
    g, o, xs = setup(dims, ls, 2, seed=21)
    sol, src = g.multi_vector(2), g.multi_vector(2, xs)
    it, tr, st, _ = g.InvCgMrhs_dev(sol, src, CgArg(0.0, 6, 1e-14), np.array([0.1, 0.1]), "f32")
    assert np.all(st == 2) and np.all(it == 6)
    for j in range(2):
        ax = np.zeros(g.n)
        g.MatPcDagMatPc(ax, sol.download(j), 0.1)
        assert abs(rel(ax, xs[j]) - tr[j]) < 1e-6 * tr[j]
    with pytest.raises(DwfGpuError) as e:
        g.InvCgMrhs_dev(sol, src, CgArg(0.0, 10, 1e-8), np.array([0.1, 0.1]), "f32", delta=1.5)
    assert e.value.code == BAD_ARG
    # MatInv: the four vectors must share one RHS count
    a, b = g.multi_vector(2), g.multi_vector(1)
    with pytest.raises(DwfGpuError) as e:
        g.MatInvMrhs_dev(a, a, a, b, CgArg(0.0, 10, 1e-8), np.array([0.1, 0.1]), "f32")
    assert e.value.code == BAD_ARG


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_batched_matinv(action):
    """Full-vector solves of several RHS in one batched CG against the
    single-RHS MatInv (FP64 InvCg) of each."""
    dims, ls = (4, 4, 4, 8), 8
    masses = np.array([0.04, 0.04, 0.1])
    nrhs = len(masses)
    say(f"[gpu] batched MatInv {action} nrhs={nrhs}")
    g, o, _ = setup(dims, ls, nrhs, action=action)
    firsts = [src_vec(dims, ls, 40 + j) for j in range(nrhs)]
    seconds = [src_vec(dims, ls, 50 + j) for j in range(nrhs)]
    inf, ins = g.multi_vector(nrhs, firsts), g.multi_vector(nrhs, seconds)
    outf, outs = g.multi_vector(nrhs), g.multi_vector(nrhs)
    it, tr, st, nrel = g.MatInvMrhs_dev(outf, outs, inf, ins, CgArg(0.0, 5000, 1e-10), masses,
                                        "f32", 0.1)
    say(f"  batched iters {it.tolist()} true_res {np.array2string(tr, precision=2)}")
    assert np.all(st == 0) and np.all(tr <= 1e-10)
    v = [g.vector() for _ in range(4)]
    for j in range(nrhs):
        assert np.array_equal(inf.download(j), firsts[j])   # sources preserved
        assert np.array_equal(ins.download(j), seconds[j])
        v[0].upload(np.zeros(g.n))
        v[1].upload(np.zeros(g.n))
        v[2].upload(firsts[j])
        v[3].upload(seconds[j])
        i1, tr1, st1 = g.MatInv_dev(v[0], v[1], v[2], v[3], CgArg(masses[j], 5000, 1e-10))
        assert st1 == 0
        want = np.concatenate([v[0].download(), v[1].download()])
        got = np.concatenate([outf.download(j), outs.download(j)])
        err = rel(got, want)
        say(f"  RHS {j}: batched {it[j]} / single {i1} iterations, |x - x_single|/|x| {err:.2e}")
        assert err < 1e-8 and abs(int(it[j]) - i1) <= 3
