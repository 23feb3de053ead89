"""Independent textbook operators used to pin the oracle.

Nothing here shares code with oracle/cps_oracle.c.  The Wilson hopping term is
written with dense 4x4 gamma matrices taken from the documentation block of
cps_pp/include/util/dirac_op.h:956-975, lattice shifts are ``np.roll`` on a
full 4D array, the anti-periodic time boundary is applied as an explicit sign
on the wrapped hop (not folded into the links), and the even-odd Schur
complements are formed from the full 5D operators:

* Shamir DWF (one s-node):
  D = (4 - M5 + 1/a5) - 1/2 D_hop - (1/a5) (P+ S_- + P- S_+),
  with S_- psi(s) = psi(s-1), S_+ psi(s) = psi(s+1) and the walls
  psi(-1) -> -m psi(ls-1), psi(ls) -> -m psi(0).
* Mobius (real b, c):
  D = (b D_W + 1/a5) + (c D_W - 1/a5) (P+ S_- + P- S_+),  D_W = (4-M5) - 1/2 D_hop.

P+- = (1 +- gamma5)/2 with gamma5 = gamma_x gamma_y gamma_z gamma_t.
The normalisations 2*kappa*D that CPS preconditions follow from these forms.
"""
from __future__ import annotations

import numpy as np

# documented gamma matrices (dirac_op.h:956-975); index 0..3 = x, y, z, t
GAMMA = np.zeros((4, 4, 4), dtype=np.complex128)
GAMMA[3] = [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]]
GAMMA[0] = [[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]]
GAMMA[1] = [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]]
GAMMA[2] = [[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]]
GAMMA5 = GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3]
PPLUS = 0.5 * (np.eye(4) + GAMMA5)
PMINUS = 0.5 * (np.eye(4) - GAMMA5)

# numpy axis of the 4D array (t,z,y,x,...) that carries direction mu
_AXIS = {0: 3, 1: 2, 2: 1, 3: 0}


def _shift(f, mu, step, bc):
    """f(x + step*mu) with sign -1 on wrapped hops where bc[mu] is set."""
    ax = _AXIS[mu]
    g = np.roll(f, -step, axis=ax)
    if bc[mu]:
        L = f.shape[ax]
        idx = [slice(None)] * f.ndim
        idx[ax] = L - 1 if step == 1 else 0
        g[tuple(idx)] *= -1.0
    return g


def spin_mul(m, f):
    """Apply a 4x4 spin matrix to f[..., spin, colour]."""
    return np.einsum("ab,...bc->...ac", m, f)


def colour_mul(u, f, dagger=False):
    """u[...,3,3] acting on f[..., spin, colour]."""
    if dagger:
        return np.einsum("...ba,...sb->...sa", np.conj(u), f)
    return np.einsum("...ab,...sb->...sa", u, f)


def wilson_hop(U, psi, dag=False, bc=(0, 0, 0, 1)):
    """sum_mu (1 - s g_mu) U_mu(x) psi(x+mu) + (1 + s g_mu) U_mu(x-mu)^+ psi(x-mu).

    U: (lt,lz,ly,lx,4,3,3) raw links; psi: (lt,lz,ly,lx,4,3).  s = +1 (dag
    False) or -1.
    """
    s = -1.0 if dag else 1.0
    out = np.zeros_like(psi)
    one = np.eye(4)
    for mu in range(4):
        fwd = colour_mul(U[..., mu, :, :], _shift(psi, mu, +1, bc))
        out += spin_mul(one - s * GAMMA[mu], fwd)
        ub = np.roll(U[..., mu, :, :], 1, axis=_AXIS[mu])  # U_mu(x - mu)
        bwd = colour_mul(ub, _shift(psi, mu, -1, bc), dagger=True)
        out += spin_mul(one + s * GAMMA[mu], bwd)
    return out


def hop5(psi5, mass):
    """(P+ S_- + P- S_+) psi with the -m walls; psi5: (ls, ..., 4, 3)."""
    ls = psi5.shape[0]
    minus = np.concatenate([-mass * psi5[ls - 1:], psi5[:ls - 1]])  # psi(s-1)
    plus = np.concatenate([psi5[1:], -mass * psi5[:1]])            # psi(s+1)
    return spin_mul(PPLUS, minus) + spin_mul(PMINUS, plus)


def dwf_full(U, psi5, mass, m5, a5inv=1.0, bc=(0, 0, 0, 1)):
    """Textbook Shamir operator D psi5, psi5: (ls,lt,lz,ly,lx,4,3)."""
    hop4 = np.stack([wilson_hop(U, p, bc=bc) for p in psi5])
    return (4 - m5 + a5inv) * psi5 - 0.5 * hop4 - a5inv * hop5(psi5, mass)


def mobius_full(U, psi5, mass, m5, b, c, a5inv=1.0, bc=(0, 0, 0, 1)):
    """Textbook Mobius operator (b D_W + 1/a5) + (c D_W - 1/a5) hop5."""
    def dw(f):
        return (4 - m5) * f - 0.5 * np.stack([wilson_hop(U, p, bc=bc) for p in f])
    h = hop5(psi5, mass)
    return b * dw(psi5) + a5inv * psi5 + c * dw(h) - a5inv * h


def parity5(dims, ls):
    lx, ly, lz, lt = dims
    s, t, z, y, x = np.meshgrid(np.arange(ls), np.arange(lt), np.arange(lz),
                                np.arange(ly), np.arange(lx), indexing="ij")
    return (x + y + z + t + s) % 2, (x + y + z + t) % 2


def dwf_schur_odd(U, psi5_odd, mass, m5, a5inv=1.0, bc=(0, 0, 0, 1)):
    """Schur complement of 2 kappa D on the 5D-odd sites, applied to a 5D field
    supported on 5D-odd sites.  Diagonal blocks of 2 kappa D are the identity."""
    kappa = 1.0 / (2 * (4 - m5 + a5inv))
    dims = psi5_odd.shape[4:0:-1]
    par5, _ = parity5(dims, psi5_odd.shape[0])

    def offdiag(f):  # 2 kappa D - 1
        return 2 * kappa * dwf_full(U, f, mass, m5, a5inv, bc) - f

    a = offdiag(psi5_odd)
    a[par5 == 1] = 0  # keep the even part
    c = offdiag(a)
    c[par5 == 0] = 0
    return psi5_odd - c


def mobius_m5_matrix(ls, mass, K):
    """(4 ls) x (4 ls) matrix of 1 + K hop5 in (s, spin) space."""
    m = np.eye(4 * ls, dtype=np.complex128)
    for s in range(ls):
        sm, cm = (s - 1, 1.0) if s > 0 else (ls - 1, -mass)
        sp, cp = (s + 1, 1.0) if s < ls - 1 else (0, -mass)
        m[4 * s:4 * s + 4, 4 * sm:4 * sm + 4] += K * cm * PPLUS
        m[4 * s:4 * s + 4, 4 * sp:4 * sp + 4] += K * cp * PMINUS
    return m


def apply_s_spin(mat, f):
    """Apply a (4 ls)x(4 ls) (s,spin) matrix to f: (ls, ..., 4, 3)."""
    ls = f.shape[0]
    g = np.moveaxis(f, 0, -3)                      # (..., ls, 4, 3)
    sh = g.shape
    g = g.reshape(sh[:-3] + (4 * ls, 3))
    g = np.einsum("ab,...bc->...ac", mat, g).reshape(sh)
    return np.moveaxis(g, -3, 0)


def mobius_schur_odd(U, psi5_odd, mass, m5, b, c, a5inv=1.0, bc=(0, 0, 0, 1)):
    """Schur complement of 2 kappa_b D_M on the 4D-odd sites."""
    kb = 1.0 / (2 * (b * (4 - m5) + a5inv))
    kc = 1.0 / (2 * (c * (4 - m5) - a5inv))
    ls = psi5_odd.shape[0]
    dims = psi5_odd.shape[4:0:-1]
    _, par4 = parity5(dims, ls)
    m5mat = mobius_m5_matrix(ls, mass, kb / kc)

    def full(f):
        return 2 * kb * mobius_full(U, f, mass, m5, b, c, a5inv, bc)

    def m5op(f):
        return apply_s_spin(m5mat, f)

    a = full(psi5_odd) - m5op(psi5_odd)   # M_eo psi
    a[par4 == 1] = 0
    binv = apply_s_spin(np.linalg.inv(m5mat), a)
    cc = full(binv) - m5op(binv)          # M_oe M5^-1 M_eo psi
    cc[par4 == 0] = 0
    return m5op(psi5_odd) - cc
