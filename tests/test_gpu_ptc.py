"""GPU: the parallel-transport network (BF16 MFMA kernels) against its CPU
restatement (oracle/ptc_net.py), its training, and the flexible CG it
preconditions.

Bars (the oracle runs with the same bf16-rounded weights, so the remaining
difference is the bf16 rounding of activations, of the links in the
transport and of the transported fields, about 2^-9 = 2e-3 per rounding, a
handful of roundings per layer):
  * network output vs oracle, normwise: < 3e-2
  * loss vs oracle: < 3e-2 relative; weight gradient per layer, normwise:
    < 6e-2
  * training lowers the loss; the trained network used by the flexible CG
    reaches the FP64 true residual of the stop rule in fewer iterations than
    the same solver without the network.
"""
import sys

import numpy as np
import pytest

from cps_veloc_amd import lattice as lat
from cps_veloc_amd.dwf import CgArg, DiracOpDwfGpu, DiracOpMobiusGpu
from oracle import ptc_net as pn

pytestmark = pytest.mark.gpu

M5 = 1.8
HID = 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16_round(x):
    """round-to-nearest-even to bfloat16, as float64"""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split_layers(w, ls):
    shapes = [(9, HID, ls, 4, 4), (9, HID, HID, 4, 4), (9, ls, HID, 4, 4)]
    out, off = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(np.asarray(w[off:off + n]).reshape(sh))
        off += n
    return out


def make(action, dims, ls, seed=101):
    U = lat.random_gauge(dims, 0.3, seed)
    if action == "dwf":
        g = DiracOpDwfGpu(U, dims, ls, m5=M5)
    else:
        g = DiracOpMobiusGpu(U, dims, ls, m5=M5, b=1.5, c=0.5)
    return U, g


def rounded_weights(g, rng, ls, scale=0.3):
    n = g.ptc_num_weights()
    w = (scale * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    w = (bf16_round(w.real) + 1j * bf16_round(w.imag)).astype(np.complex64)
    g.ptc_set_weights(w)
    return [x.astype(np.complex128) for x in split_layers(w, ls)]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


# Ls 4: layers <4,16>, <16,16>, <16,4> (one K step; 4 output channels take
# the LDS output path); Ls 12: <12,16>, <16,16>, <16,12> (three K steps, 1.5
# items per thread, 12 output channels through LDS, 16 from registers)
@pytest.mark.parametrize("ls", [4, 12])
@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_network_forward_matches_oracle(action, ls):
    dims = (4, 4, 4, 8)
    say(f"[gpu] network forward {action} {dims} ls={ls}")
    U, g = make(action, dims, ls)
    g.ptc_init(1, 0.02)
    rng = np.random.default_rng(3)
    Ws = rounded_weights(g, rng, ls)
    r = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 9))
    got = g.ptc_apply(r)
    want = pn.apply_net(Ws, r, pn.links_with_bc(U, dims), dims, ls, action)
    err = rel(got, want)
    say(f"  normwise rel err {err:.3e}")
    assert err < 3e-2


@pytest.mark.parametrize("ls", [4, 12])
@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_network_gradients_match_oracle(action, ls):
    dims = (4, 4, 4, 8)
    say(f"[gpu] network gradients {action} ls={ls}")
    U, g = make(action, dims, ls)
    g.ptc_init(1, 0.02)
    rng = np.random.default_rng(4)
    Ws = rounded_weights(g, rng, ls, scale=0.2)
    vh = int(np.prod(dims)) // 2
    r = lat.to_reals(lat.gaussian_spinor((ls, vh), 11))
    e = lat.to_reals(lat.gaussian_spinor((ls, vh), 12))
    loss, grad = g.ptc_loss_grad(r, e)
    wl, wg = pn.loss_and_grads_cg(Ws, r, e, pn.links_with_bc(U, dims), dims, ls, action)
    say(f"  loss gpu {loss:.6e} oracle {wl:.6e}")
    assert abs(loss - wl) < 3e-2 * wl
    for k, (a, b) in enumerate(zip(split_layers(grad, ls), wg)):
        err = rel(a, b)
        say(f"  layer {k}: gradient normwise rel err {err:.3e}")
        assert err < 6e-2


def test_gradients_on_fewer_tiles_than_xcds():
    """4x4x2x4: four 32-site tiles, so the weight gradient runs four
    work-groups per direction in one range (fewer than the XCDs) and the
    forward layer four work-groups over eight XCDs."""
    dims, ls = (4, 4, 2, 4), 4
    say(f"[gpu] network gradients on {dims} ls={ls}")
    U, g = make("mobius", dims, ls)
    g.ptc_init(1, 0.02)
    rng = np.random.default_rng(5)
    Ws = rounded_weights(g, rng, ls, scale=0.2)
    vh = int(np.prod(dims)) // 2
    r = lat.to_reals(lat.gaussian_spinor((ls, vh), 13))
    e = lat.to_reals(lat.gaussian_spinor((ls, vh), 14))
    links = pn.links_with_bc(U, dims)
    err = rel(g.ptc_apply(r), pn.apply_net(Ws, r, links, dims, ls, "mobius"))
    say(f"  forward normwise rel err {err:.3e}")
    assert err < 3e-2
    loss, grad = g.ptc_loss_grad(r, e)
    wl, wg = pn.loss_and_grads_cg(Ws, r, e, links, dims, ls, "mobius")
    assert abs(loss - wl) < 3e-2 * wl
    for k, (a, b) in enumerate(zip(split_layers(grad, ls), wg)):
        err = rel(a, b)
        say(f"  layer {k}: gradient normwise rel err {err:.3e}")
        assert err < 6e-2


def test_training_and_preconditioned_solve():
    dims, ls, mass = (8, 8, 8, 8), 8, 0.04
    say(f"[gpu] train the network and solve with it, dwf {dims} ls={ls} mf={mass}")
    U, g = make("dwf", dims, ls)
    src = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 31))
    arg = CgArg(mass, 5000, 1e-10)
    # without the network: reliable-update CG through the same solver
    vs, vb = g.vector(), g.vector(src)
    it0, tr0, st0, _, ns0 = g.InvCgPtc_dev(vs, vb, arg, "f32")
    say(f"  no network: iters={it0} true_res={tr0:.3e}")
    assert st0 == 0 and tr0 <= 1e-10 and ns0 == 0
    g.ptc_init(5, 0.02)
    # high-mode pairs (r = A e, e Gaussian): the part of A^-1 a three-hop
    # network can represent
    l0, l1, npairs = g.ptc_train(mass, nsolves=6, pairs_per_solve=0, random_pairs=32,
                                 stride=3, epochs=3, batch=4, lr=2e-3, seed=17)
    say(f"  training: {npairs} pairs, loss {l0:.4f} -> {l1:.4f}")
    assert npairs > 0 and l1 < l0
    vs = g.vector()
    it1, tr1, st1, nrel, ns1 = g.InvCgPtc_dev(vs, vb, arg, "f32")
    say(f"  with network: iters={it1} (net used on {ns1}) true_res={tr1:.3e} reliable={nrel}")
    assert st1 == 0 and tr1 <= 1e-10
    x = vs.download()
    ax = np.zeros(g.n)
    g.MatPcDagMatPc(ax, x, mass)
    assert rel(ax, src) <= 1.0001e-10
    assert ns1 > 0 and it1 < it0
    # flexible GCR(8) with the same network: residual-minimising steps in FP64
    vs = g.vector()
    it2, tr2, st2, ns2, nc2 = g.InvGcrPtc_dev(vs, vb, arg, nkrylov=8)
    say(f"  flexible GCR(8): iters={it2} (net helped on {ns2}, {nc2} restarts) "
        f"true_res={tr2:.3e}")
    assert st2 == 0 and tr2 <= 1e-10
    ax = np.zeros(g.n)
    g.MatPcDagMatPc(ax, vs.download(), mass)
    assert rel(ax, src) <= 1.0001e-10
    assert ns2 > 0 and it2 < it0


def test_gcr_without_network_is_cg():
    """No network: the GCR entry hands the whole solve to InvCg (same
    iterations and true residual as dwfgpu_invcg_dev)."""
    dims, ls, mass = (4, 4, 4, 8), 4, 0.1
    say("[gpu] flexible GCR without a network")
    U, g = make("mobius", dims, ls)
    src = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 41))
    arg = CgArg(mass, 5000, 1e-10)
    vs, vb = g.vector(), g.vector(src)
    it, tr, st, ns, nc = g.InvGcrPtc_dev(vs, vb, arg, nkrylov=4)
    v2 = g.vector()
    it_cg, tr_cg, st_cg = g.InvCg_dev(v2, vb, arg)
    say(f"  gcr entry {it} iters, true res {tr:.3e}; InvCg {it_cg} iters")
    assert st == 0 and ns == 0 and it == it_cg and abs(tr - tr_cg) <= 1e-3 * tr_cg
