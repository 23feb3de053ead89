"""Pin the CPU restatement of the parallel-transport network (oracle/ptc_net.py).

* a single-channel layer with W_{+mu} = 1 - gamma_mu, W_{-mu} = 1 + gamma_mu is
  exactly the Wilson hopping term of the independent textbook operator
  (tests/textbook.py), boundary signs included;
* <g, layer(phi)> = <layer_adjoint(g), phi> for random fields and weights;
* gauge covariance: net(V phi, U^V) = V net(phi, U);
* the weight gradients agree with central finite differences of the loss;
* embed / extract invert each other on CG half vectors of both actions.
"""
import numpy as np
import pytest

import textbook as tb
from cps_veloc_amd import lattice as lat
from oracle import ptc_net as pn

DIMS = (4, 2, 2, 4)


def cfield(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rand_weights(rng, chans, scale=0.3):
    return [scale * cfield(rng, (pn.NDIR, co, ci, 4, 4)) for ci, co in zip(chans[:-1], chans[1:])]


def su3(rng):
    q, r = np.linalg.qr(cfield(rng, (3, 3)))
    q = q * (np.diag(r) / np.abs(np.diag(r)))
    return q / np.linalg.det(q) ** (1 / 3)


@pytest.mark.parametrize("bc", [(0, 0, 0, 1), (1, 0, 1, 0)])
def test_layer_is_wilson_hop(bc):
    Uc = lat.random_gauge(DIMS, 0.4, 5)
    U = pn.links_with_bc(Uc, DIMS, bc)
    rng = np.random.default_rng(1)
    lx, ly, lz, lt = DIMS
    psi = cfield(rng, (lt, lz, ly, lx, 4, 3))
    W = np.zeros((pn.NDIR, 1, 1, 4, 4), dtype=complex)
    for mu in range(4):
        W[1 + mu, 0, 0] = np.eye(4) - tb.GAMMA[mu]
        W[5 + mu, 0, 0] = np.eye(4) + tb.GAMMA[mu]
    got = pn.layer(W, psi[None], U)[0]
    Uraw = np.array(Uc).reshape(lt, lz, ly, lx, 4, 3, 3)
    want = tb.wilson_hop(Uraw, psi, bc=bc)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_layer_adjoint():
    rng = np.random.default_rng(2)
    U = pn.links_with_bc(lat.random_gauge(DIMS, 0.4, 6), DIMS)
    lx, ly, lz, lt = DIMS
    W = rand_weights(rng, [3, 2])[0]
    phi = cfield(rng, (3, lt, lz, ly, lx, 4, 3))
    g = cfield(rng, (2, lt, lz, ly, lx, 4, 3))
    lhs = np.vdot(g, pn.layer(W, phi, U))
    rhs = np.vdot(pn.layer_adjoint(W, g, U), phi)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_gauge_covariance():
    rng = np.random.default_rng(3)
    lx, ly, lz, lt = DIMS
    U = pn.links_with_bc(lat.random_gauge(DIMS, 0.4, 7), DIMS)
    V = np.array([su3(rng) for _ in range(lx * ly * lz * lt)]).reshape(lt, lz, ly, lx, 3, 3)
    UV = np.empty_like(U)
    for mu in range(4):
        Vn = np.roll(V, -1, axis=3 - mu)  # V(x + mu)
        UV[..., mu, :, :] = V @ U[..., mu, :, :] @ np.conj(np.swapaxes(Vn, -1, -2))
    Ws = rand_weights(rng, [2, 3, 3, 2])
    phi = cfield(rng, (2, lt, lz, ly, lx, 4, 3))
    rot = lambda f: np.einsum("tzyxab,ctzyxsb->ctzyxsa", V, f)
    a = pn.forward(Ws, rot(phi), UV)[-1]
    b = rot(pn.forward(Ws, phi, U)[-1])
    assert np.abs(a - b).max() < 1e-11 * np.abs(b).max()


def test_weight_gradients_finite_difference():
    rng = np.random.default_rng(4)
    lx, ly, lz, lt = DIMS
    U = pn.links_with_bc(lat.random_gauge(DIMS, 0.4, 8), DIMS)
    Ws = rand_weights(rng, [2, 3, 3, 2])
    phi = cfield(rng, (2, lt, lz, ly, lx, 4, 3))
    target = cfield(rng, (2, lt, lz, ly, lx, 4, 3))
    L0, grads = pn.loss_and_grads(Ws, phi, target, U)
    eps = 1e-6
    for k in range(3):
        for _ in range(3):
            D = cfield(rng, Ws[k].shape)
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[k] = Wp[k] + eps * D
            Wm[k] = Wm[k] - eps * D
            fd = (pn.loss_and_grads(Wp, phi, target, U)[0] -
                  pn.loss_and_grads(Wm, phi, target, U)[0]) / (2 * eps)
            an = np.real(np.vdot(grads[k], D))
            assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (k, fd, an)


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_embed_extract(action):
    ls = 4
    rng = np.random.default_rng(5)
    vh = int(np.prod(DIMS)) // 2
    h = rng.standard_normal(ls * vh * 24)
    f = pn.embed(h, DIMS, ls, action)
    assert np.array_equal(pn.extract(f, DIMS, ls, action), h)
    # the other parity of every slice stays empty
    assert np.count_nonzero(f) == np.count_nonzero(h.view(np.complex128))


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_cg_loss_gradients_finite_difference(action):
    ls = 2
    rng = np.random.default_rng(6)
    U = pn.links_with_bc(lat.random_gauge(DIMS, 0.4, 9), DIMS)
    Ws = rand_weights(rng, [ls, 2, 2, ls])
    vh = int(np.prod(DIMS)) // 2
    r = rng.standard_normal(ls * vh * 24)
    e = rng.standard_normal(ls * vh * 24)
    L0, grads = pn.loss_and_grads_cg(Ws, r, e, U, DIMS, ls, action)
    eps = 1e-6
    for k in range(3):
        D = cfield(rng, Ws[k].shape)
        Wp = [w.copy() for w in Ws]
        Wm = [w.copy() for w in Ws]
        Wp[k] = Wp[k] + eps * D
        Wm[k] = Wm[k] - eps * D
        fd = (pn.loss_and_grads_cg(Wp, r, e, U, DIMS, ls, action)[0] -
              pn.loss_and_grads_cg(Wm, r, e, U, DIMS, ls, action)[0]) / (2 * eps)
        an = np.real(np.vdot(grads[k], D))
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (k, fd, an)
