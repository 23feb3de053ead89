"""Synthetic lattice inputs and CPS storage orders (host side, numpy).

Gauge links follow the recipe ``U = reunitarize(exp(i*eps*H))`` with ``H`` a
seeded Gaussian traceless Hermitian 3x3 matrix per link, and reunitarize the
row-wise Gram-Schmidt of CPS's ``Matrix::Unitarize``
(cps_pp/src/util/vector/comsrc/unitarize.C:28-56): normalise row 0,
orthogonalise and normalise row 1, row 2 = conj(row0 x row1).

Storage orders (CPS's own):

* CANONICAL gauge: ``[t][z][y][x][mu][row][col]`` complex, i.e. site
  ``x + lx*(y + ly*(z + lz*t))``, then mu, then a row-major 3x3
  (cps_pp/include/util/lattice.h:327, convert_func.C:393-405).
* CANONICAL 5D fermion: ``[s][t][z][y][x][spin][colour]`` complex.
* checkerboard site index ``(x/2) + (lx/2)*(y + ly*(z + lz*t))``
  (d_op_wilson/noarch/wilson_dslash.C:288).
* DWF full vector ``[5D-odd half | 5D-even half]``; slice ``s`` of the
  5D-parity-``cb`` half holds 4D parity ``(s+cb)%2``
  (lattice/convert/convert.C:704-760, d_op_dwf/d_op_dwf.C:357-361).
* Mobius full vector ``[4D-even (all s) | 4D-odd (all s)]``
  (convert.C:797-831, d_op_mobius/d_op_mobius.C:454-457).
"""
from __future__ import annotations

import numpy as np

_GELL_MANN = np.zeros((8, 3, 3), dtype=np.complex128)
_GELL_MANN[0, 0, 1] = _GELL_MANN[0, 1, 0] = 1
_GELL_MANN[1, 0, 1], _GELL_MANN[1, 1, 0] = -1j, 1j
_GELL_MANN[2, 0, 0], _GELL_MANN[2, 1, 1] = 1, -1
_GELL_MANN[3, 0, 2] = _GELL_MANN[3, 2, 0] = 1
_GELL_MANN[4, 0, 2], _GELL_MANN[4, 2, 0] = -1j, 1j
_GELL_MANN[5, 1, 2] = _GELL_MANN[5, 2, 1] = 1
_GELL_MANN[6, 1, 2], _GELL_MANN[6, 2, 1] = -1j, 1j
_GELL_MANN[7] = np.diag([1, 1, -2]) / np.sqrt(3)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def expm_su3_algebra(x: np.ndarray) -> np.ndarray:
    """exp(x) for a stack of 3x3 anti-Hermitian matrices (scaling+squaring)."""
    norm = np.max(np.abs(x).sum(axis=-1), axis=-1)
    nsq = int(max(0, np.ceil(np.log2(max(float(norm.max(initial=0.0)), 1e-300)))) + 4)
    y = x / (2.0 ** nsq)
    out = np.broadcast_to(np.eye(3, dtype=np.complex128), x.shape).copy()
    term = out.copy()
    for k in range(1, 18):
        term = term @ y / k
        out += term
    for _ in range(nsq):
        out = out @ out
    return out


def reunitarize(u: np.ndarray) -> np.ndarray:
    """Matrix::Unitarize on a stack of 3x3 matrices (unitarize.C:28-56)."""
    u = u.copy()
    r0 = u[..., 0, :]
    r0 /= np.sqrt(np.sum(np.abs(r0) ** 2, axis=-1, keepdims=True))
    r1 = u[..., 1, :]
    r1 -= r0 * np.sum(np.conj(r0) * r1, axis=-1, keepdims=True)
    r1 /= np.sqrt(np.sum(np.abs(r1) ** 2, axis=-1, keepdims=True))
    u[..., 2, :] = np.conj(np.cross(r0, r1))
    return u


def random_gauge(dims, eps: float, seed: int) -> np.ndarray:
    """CANONICAL gauge field, complex array of shape (lt,lz,ly,lx,4,3,3)."""
    lx, ly, lz, lt = dims
    n = lx * ly * lz * lt * 4
    rng = _rng(seed)
    out = np.empty((n, 3, 3), dtype=np.complex128)
    chunk = 1 << 20
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        w = rng.standard_normal((hi - lo, 8))
        h = np.einsum("na,aij->nij", w, _GELL_MANN) * 0.5
        out[lo:hi] = reunitarize(expm_su3_algebra(1j * eps * h))
    return out.reshape(lt, lz, ly, lx, 4, 3, 3)


def unit_gauge(dims) -> np.ndarray:
    lx, ly, lz, lt = dims
    u = np.zeros((lt, lz, ly, lx, 4, 3, 3), dtype=np.complex128)
    u[..., :, :] = np.eye(3)
    return u


def gauge_reals(u: np.ndarray) -> np.ndarray:
    """Complex CANONICAL gauge -> flat float64 array [site][mu][row][col][re,im]."""
    return np.ascontiguousarray(u).view(np.float64).reshape(-1)


def gaussian_spinor(shape, seed: int) -> np.ndarray:
    """Seeded unit Gaussian complex spinor of the given leading shape (...,4,3)."""
    rng = _rng(seed)
    re = rng.standard_normal(tuple(shape) + (4, 3))
    im = rng.standard_normal(tuple(shape) + (4, 3))
    return re + 1j * im


# ---------------------------------------------------------------------------
# storage orders
# ---------------------------------------------------------------------------
def parity_mask(dims) -> np.ndarray:
    """4D parity (x+y+z+t)%2 in CANONICAL site order, shape (lt,lz,ly,lx)."""
    lx, ly, lz, lt = dims
    t, z, y, x = np.meshgrid(np.arange(lt), np.arange(lz), np.arange(ly),
                             np.arange(lx), indexing="ij")
    return (x + y + z + t) % 2


def cb_site_of_canonical(dims) -> np.ndarray:
    """Checkerboard index of each CANONICAL site, shape (lt,lz,ly,lx)."""
    lx, ly, lz, lt = dims
    t, z, y, x = np.meshgrid(np.arange(lt), np.arange(lz), np.arange(ly),
                             np.arange(lx), indexing="ij")
    return x // 2 + (lx // 2) * (y + ly * (z + lz * t))


def _half_gather(dims):
    """For parity p: canonical flat indices ordered by checkerboard index."""
    par = parity_mask(dims).reshape(-1)
    cbi = cb_site_of_canonical(dims).reshape(-1)
    out = []
    for p in (0, 1):
        sites = np.nonzero(par == p)[0]
        order = np.argsort(cbi[sites], kind="stable")
        out.append(sites[order])
    return out


def canonical_to_halves(psi5: np.ndarray, dims, kind: str):
    """CANONICAL 5D fermion (ls,lt,lz,ly,lx,4,3) -> two CPS half vectors.

    kind='dwf'    -> (odd5, even5), each complex (ls, V4/2, 4, 3)
    kind='mobius' -> (even4, odd4)
    """
    ls = psi5.shape[0]
    flat = psi5.reshape(ls, -1, 4, 3)
    g = _half_gather(dims)
    if kind == "dwf":
        odd = np.stack([flat[s][g[(s + 1) % 2]] for s in range(ls)])
        even = np.stack([flat[s][g[s % 2]] for s in range(ls)])
        return odd, even
    even = np.stack([flat[s][g[0]] for s in range(ls)])
    odd = np.stack([flat[s][g[1]] for s in range(ls)])
    return even, odd


def halves_to_canonical(first: np.ndarray, second: np.ndarray, dims, kind: str):
    """Inverse of :func:`canonical_to_halves` (same argument order)."""
    ls = first.shape[0]
    lx, ly, lz, lt = dims
    g = _half_gather(dims)
    out = np.zeros((ls, lx * ly * lz * lt, 4, 3), dtype=np.complex128)
    for s in range(ls):
        if kind == "dwf":
            out[s][g[(s + 1) % 2]] = first[s]
            out[s][g[s % 2]] = second[s]
        else:
            out[s][g[0]] = first[s]
            out[s][g[1]] = second[s]
    return out.reshape(ls, lt, lz, ly, lx, 4, 3)


def to_reals(v: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(v).view(np.float64).reshape(-1)


def from_reals(v: np.ndarray, ls: int, vh: int) -> np.ndarray:
    return np.ascontiguousarray(v).view(np.complex128).reshape(ls, vh, 4, 3)
