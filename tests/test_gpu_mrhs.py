"""GPU: the batched many-RHS DWF operator and CG.

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
    for bit, iteration counts included.
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


def setup(dims, ls, nrhs, seed=3):
    U = lat.random_gauge(dims, 0.3, 101)
    g = DiracOpDwfGpu(U, dims, ls, m5=M5)
    o = Oracle(dims, ls, "dwf", m5=M5, gauge_canonical=U)
    xs = [src_vec(dims, ls, seed + j) for j in range(nrhs)]
    return g, o, xs


@pytest.mark.parametrize("dims,ls,nrhs", [((4, 4, 4, 8), 8, 5), ((4, 2, 6, 4), 6, 3)])
def test_batched_operator(dims, ls, nrhs):
    say(f"[gpu] batched MatPcDagMatPc {dims} ls={ls} nrhs={nrhs}")
    g, o, xs = setup(dims, ls, nrhs)
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


@pytest.mark.parametrize("sloppy", ["f32", "i16"])
def test_batched_cg(sloppy):
    dims, ls = (4, 4, 4, 8), 8
    masses = np.array([0.01, 0.04, 0.5, 0.04])
    nrhs = len(masses)
    say(f"[gpu] batched CG sloppy={sloppy} nrhs={nrhs} masses={masses.tolist()}")
    g, o, xs = setup(dims, ls, nrhs, seed=11)
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
    U = lat.random_gauge(dims, 0.3, 1)
    mob = DiracOpMobiusGpu(U, dims, ls, m5=M5)
    a, b = mob.multi_vector(1), mob.multi_vector(1)
    with pytest.raises(DwfGpuError) as e:
        mob.MatPcDagMatPc_mrhs_dev(a, b, np.array([0.1]))
    assert e.value.code == BAD_ARG
