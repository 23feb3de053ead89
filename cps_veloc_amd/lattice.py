"""Synthetic lattice inputs and CPS storage orders (host side, numpy).

Gauge links follow the recipe ``U = reunitarize(exp(i*eps*H))`` with ``H`` a
seeded Gaussian traceless Hermitian 3x3 matrix per link (generated in C++, see
include/dwf_synth.h), reunitarize being CPS's ``Matrix::Unitarize``
(cps_pp/src/util/vector/comsrc/unitarize.C:28-56).

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

import ctypes

import numpy as np


def random_gauge(dims, eps: float, seed: int) -> np.ndarray:
    """CANONICAL gauge field, complex array of shape (lt,lz,ly,lx,4,3,3),
    from the library's counter-based generator (include/dwf_synth.h)."""
    from .dwf import lib
    lx, ly, lz, lt = dims
    out = np.empty(lx * ly * lz * lt * 4 * 18)
    lib().dwfsynth_gauge(out, (ctypes.c_int * 4)(*dims), float(eps), int(seed))
    return out.view(np.complex128).reshape(lt, lz, ly, lx, 4, 3, 3)


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
    from .dwf import lib
    n = int(np.prod(shape)) * 12
    out = np.empty(2 * n)
    lib().dwfsynth_gaussian(out, out.size, int(seed))
    return out.view(np.complex128).reshape(tuple(shape) + (4, 3))


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
