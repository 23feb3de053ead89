"""CPU restatement of the gauge-equivariant parallel-transport network used as
a preconditioner (TEST INFRASTRUCTURE ONLY: the checker of the HIP kernels in
cps_veloc_amd/csrc/ptc_kernels.hpp; only tests/ may import it).

The network has no counterpart in CPS; it acts on the CG vector of the
even-odd preconditioned operator (cps_pp/src/util/dirac_op/d_op_dwf/d_op_dwf.C
MatPcDagMatPc, d_op_mobius.C) and is pinned here by mathematics only:
finite-difference gradients, the adjoint identity of each layer, gauge
covariance and the explicit definition below.

Fields: phi[c, t, z, y, x, spin, colour] complex (c = channel) on the full 4D
lattice, x fastest as in the canonical site order.  Links U[t, z, y, x, mu]
(3x3 complex) with the anti-periodic boundary sign already applied (the
Lattice::BondCond convention, lattice_base.C:5336-5391).

Parallel transport along direction d (d = 0: identity; 1..4: +x,+y,+z,+t;
5..8: -x,-y,-z,-t):
    T_{+mu} phi(x) = U_mu(x) phi(x + mu)
    T_{-mu} phi(x) = U_mu(x - mu)^dag phi(x - mu)
A layer maps Cin channels to Cout channels:
    out_o(x) = sum_d sum_i W[d, o, i] (T_d phi_i)(x)
with W[d, o, i] a 4x4 complex spin matrix, so the map commutes with gauge
transformations phi(x) -> V(x) phi(x), U_mu(x) -> V(x) U_mu(x) V(x+mu)^dag.
The network is three layers, Ls -> H -> H -> Ls, without activation.
"""
from __future__ import annotations

import numpy as np

NDIR = 9
AXIS = {0: 4, 1: 3, 2: 2, 3: 1}  # mu -> axis of the [t, z, y, x] site block (after channel)


def links_with_bc(U_canonical, dims, bc=(0, 0, 0, 1)):
    """CANONICAL gauge [site][mu][3][3] -> U[t, z, y, x, mu, 3, 3] with the
    anti-periodic sign on links leaving x_mu = L_mu - 1."""
    lx, ly, lz, lt = dims
    U = np.array(U_canonical, dtype=np.complex128).reshape(lt, lz, ly, lx, 4, 3, 3).copy()
    for mu in range(4):
        if bc[mu]:
            idx = [slice(None)] * 4
            idx[3 - mu] = -1
            U[tuple(idx) + (mu,)] *= -1
    return U


def transport(phi, U, d):
    """T_d phi for phi[c, t, z, y, x, 4, 3]."""
    if d == 0:
        return phi.copy()
    mu = (d - 1) % 4
    ax = AXIS[mu]
    Umu = U[..., mu, :, :]  # [t, z, y, x, 3, 3]
    # out[c, x, s, a] = sum_b M[x, a, b] nb[c, x, s, b] = (nb @ M^T)[c, x, s, a]
    if d <= 4:
        nb = np.roll(phi, -1, axis=ax)  # phi(x + mu)
        return np.matmul(nb, np.swapaxes(Umu, -1, -2))  # M = U_mu(x)
    nb = np.roll(phi, 1, axis=ax)  # phi(x - mu)
    Ub = np.roll(Umu, 1, axis=ax - 1)  # U_mu(x - mu)
    return np.matmul(nb, Ub.conj())  # M = U_mu(x - mu)^dag, M^T = conj(U)


def _spin_mix(W, f):
    """out[o, x, a, c] = sum_{i, b} W[o, i, a, b] f[i, x, b, c] (a BLAS
    contraction over (i, b))."""
    return np.moveaxis(np.tensordot(W, f, axes=([1, 3], [0, 5])), 1, 5)


def layer(W, phi, U):
    """W[d, o, i, 4, 4] complex: sum_d W_d (T_d phi)."""
    out = _spin_mix(W[0], phi)
    for d in range(1, NDIR):
        out += _spin_mix(W[d], transport(phi, U, d))
    return out


def layer_adjoint(W, g, U):
    """The adjoint map: sum_d T_{-d} (W_d^dag g), so <g, layer(phi)> = <layer_adjoint(g), phi>."""
    out = None
    for d in range(NDIR):
        dm = 0 if d == 0 else (d + 4 if d <= 4 else d - 4)
        Wd = np.conj(np.transpose(W[d], (1, 0, 3, 2)))  # [i, o, b, a]
        t = transport(_spin_mix(Wd, g), U, dm)
        if out is None:
            out = t
        else:
            out += t
    return out


def weight_grad(g, phi, U):
    """G[d, o, i] = 2 sum_x,colour g_o(x) conj(T_d phi_i(x))^T, the gradient
    dL/dRe W + i dL/dIm W of a real loss whose Wirtinger derivative with
    respect to conj(out) is g."""
    G = []
    for d in range(NDIR):
        t = transport(phi, U, d)
        # sum over the sites and colour: [o, a, i, b] -> [o, i, a, b]
        G.append(2.0 * np.tensordot(g, np.conj(t), axes=([1, 2, 3, 4, 6], [1, 2, 3, 4, 6]))
                 .transpose(0, 2, 1, 3))
    return np.array(G)


def forward(Ws, phi, U):
    acts = [phi]
    for W in Ws:
        acts.append(layer(W, acts[-1], U))
    return acts


def loss_and_grads(Ws, phi, target, U):
    """L = |net(phi) - target|^2 / |target|^2 and dL/dW for every layer."""
    acts = forward(Ws, phi, U)
    diff = acts[-1] - target
    nt = np.vdot(target, target).real
    L = np.vdot(diff, diff).real / nt
    g = diff / nt  # dL / d conj(out)
    grads = [None] * len(Ws)
    for k in range(len(Ws) - 1, -1, -1):
        grads[k] = weight_grad(g, acts[k], U)
        g = layer_adjoint(Ws[k], g, U)
    return L, grads


# ---- the CG vector <-> network field ------------------------------------
def parity_of_slice(action, s):
    """4D parity held by slice s of the CG half vector (DWF: (s+1)%2,
    convert.C:704-760; Mobius: 1, convert.C:797-831)."""
    return (s + 1) % 2 if action == "dwf" else 1


def embed(half, dims, ls, action):
    """CG half vector (CPS order [s][cb site][4][3]) -> field[s, t, z, y, x, 4, 3]."""
    lx, ly, lz, lt = dims
    h = np.asarray(half).view(np.complex128).reshape(ls, -1, 4, 3)
    f = np.zeros((ls, lt, lz, ly, lx, 4, 3), dtype=np.complex128)
    t, z, y, x = np.meshgrid(np.arange(lt), np.arange(lz), np.arange(ly), np.arange(lx),
                             indexing="ij")
    par = (x + y + z + t) % 2
    cb = (x // 2) + (lx // 2) * (y + ly * (z + lz * t))
    for s in range(ls):
        m = par == parity_of_slice(action, s)
        f[s][m] = h[s][cb[m]]
    return f


def extract(f, dims, ls, action):
    lx, ly, lz, lt = dims
    vh = lx * ly * lz * lt // 2
    out = np.zeros((ls, vh, 4, 3), dtype=np.complex128)
    t, z, y, x = np.meshgrid(np.arange(lt), np.arange(lz), np.arange(ly), np.arange(lx),
                             indexing="ij")
    par = (x + y + z + t) % 2
    cb = (x // 2) + (lx // 2) * (y + ly * (z + lz * t))
    for s in range(ls):
        m = par == parity_of_slice(action, s)
        out[s][cb[m]] = f[s][m]
    return out.reshape(-1).view(np.float64)


def apply_net(Ws, half, U, dims, ls, action):
    """z = |r| net(r / |r|) on a CG half vector (FP64 reference of the
    preconditioner application)."""
    nr = np.linalg.norm(half)
    f = embed(np.asarray(half) / nr, dims, ls, action)
    return nr * extract(forward(Ws, f, U)[-1], dims, ls, action)


def loss_and_grads_cg(Ws, r_half, e_half, U, dims, ls, action):
    """Training loss of one (residual, correction) pair of CG half vectors:
    with r^ = r/|r| and e^ = e/|r|,
        L = |extract(net(embed(r^))) - e^|^2 / |e^|^2,
    and dL/dW per layer (the gradient flows only through the parity that
    extract keeps)."""
    nr = np.linalg.norm(r_half)
    rh = np.asarray(r_half) / nr
    eh = np.asarray(e_half) / nr
    acts = forward(Ws, embed(rh, dims, ls, action), U)
    out = extract(acts[-1], dims, ls, action)
    diff = out - eh
    nt = float(np.dot(eh, eh))
    L = float(np.dot(diff, diff)) / nt
    g = embed(diff / nt, dims, ls, action)
    grads = [None] * len(Ws)
    for k in range(len(Ws) - 1, -1, -1):
        grads[k] = weight_grad(g, acts[k], U)
        g = layer_adjoint(Ws[k], g, U)
    return L, grads
