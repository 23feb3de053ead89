"""dwfgpu_matpcdagmatpc_dev_noalloc through the C ABI against
dwfgpu_matpcdagmatpc_dev (and, in FP64, against the oracle).

The allocation-free entry point runs the same kernels in the same order as
the allocating one, with its temporaries in the library's static device
array instead of the handle's buffers, so the two must agree bit for bit:
output vector and |MatPc in|^2.  Each case makes two calls in a row, the
second on a smaller lattice than the first, so the second reuses the array
the first left filled.  A lattice whose temporaries exceed the array is
refused with DWFGPU_BAD_ARG.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from cps_veloc_amd import dwf
from cps_veloc_amd import lattice as lat
from cps_veloc_amd.dwf import DiracOpDwfGpu, DiracOpMobiusGpu
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

M5 = 1.8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(action, dims, ls, prec, seed):
    U = lat.random_gauge(dims, 0.3, seed)
    if action == "dwf":
        op = DiracOpDwfGpu(U, dims, ls, m5=M5, precision=prec)
    else:
        op = DiracOpMobiusGpu(U, dims, ls, m5=M5, b=1.5, c=0.5, precision=prec)
    return U, op


def noalloc(op, out, inp, mass, want_dot=True):
    """One call of the entry point through the raw C ABI."""
    d = C.c_double(0.0)
    st = dwf.lib().dwfgpu_matpcdagmatpc_dev_noalloc(op.h, mass, out.h, inp.h,
                                                   C.byref(d) if want_dot else None)
    return st, d.value


def site_rel(got, want):
    g, w = got.reshape(-1, 24), want.reshape(-1, 24)
    return float(np.max(np.linalg.norm(g - w, axis=1) / np.linalg.norm(w, axis=1)))


def one_case(action, dims, ls, prec, mass, seed):
    say(f"[gpu] noalloc {action} {dims} ls={ls} {prec} mf={mass}")
    U, op = make(action, dims, ls, prec, seed)
    x = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), seed + 1))
    vin, va, vb = op.vector(x), op.vector(), op.vector()
    da = op.MatPcDagMatPc_dev(va, vin, mass)
    st, db = noalloc(op, vb, vin, mass)
    assert st == dwf.OK, dwf.lib().dwfgpu_last_error()
    a, b = va.download(), vb.download()
    assert np.array_equal(a, b), "noalloc output differs from dwfgpu_matpcdagmatpc_dev"
    assert db == da, (db, da)
    # without the norm (dot_prd NULL): the same output
    vb.upload(np.zeros(op.n))
    st, _ = noalloc(op, vb, vin, mass, want_dot=False)
    assert st == dwf.OK and np.array_equal(vb.download(), a)
    if prec == "f64":
        o = Oracle(dims, ls, action, m5=M5, b=1.5, c=0.5, gauge_canonical=U)
        want, dwant = o.matpcdagmatpc(x, mass)
        err = site_rel(b, want)
        say(f"  vs oracle: max site rel err {err:.2e}, dot rel err {abs(db - dwant) / dwant:.2e}")
        assert err < 1e-13 and abs(db - dwant) / dwant < 1e-13
    for v in (vin, va, vb):
        v.free()
    op.close()


@pytest.mark.parametrize("action", ["dwf", "mobius"])
@pytest.mark.parametrize("prec", ["f64", "i16"])
def test_noalloc_matches_allocating_entry(action, prec):
    # first call on an 8^3x8 Ls 8 lattice, the second on a smaller one
    one_case(action, (8, 8, 8, 8), 8, prec, 0.04, 21)
    one_case(action, (4, 4, 4, 8), 4, prec, 0.1, 33)


def test_noalloc_refuses_lattice_beyond_the_array():
    # 16^3x32, Ls 12: two FP64 half vectors alone are 2 x 151 MB > 256 MiB
    dims, ls = (16, 16, 16, 32), 12
    _, op = make("dwf", dims, ls, "f64", 5)
    vin, vout = op.vector(), op.vector()
    st, _ = noalloc(op, vout, vin, 0.01)
    msg = dwf.lib().dwfgpu_last_error().decode()
    say(f"  refused: status {st}, '{msg}'")
    assert st == dwf.BAD_ARG and "scratch" in msg
    vin.free()
    vout.free()
    op.close()
