"""Whole-field formats on the CPU: the oracle's restatement of Fconvert
(CANONICAL <-> the MatInv full vector) and Ffour2five / Ffive2four, pinned
against independent constructions; the plaquette of oracle/oracle.py pinned
by closed forms; and the library's NERSC reader (host code, no GPU) against
files laid out as CPS writes them."""
import os

import numpy as np
import pytest

from cps_veloc_amd import dwf
from cps_veloc_amd import lattice as lat
from oracle.oracle import Oracle, extract_cb, plaquette_linktrace
import nersc_file

DIMS, LS = (4, 2, 2, 6), 4


def rand(n, seed):
    return np.random.default_rng(seed).standard_normal(n)


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_canonical_to_full_vs_slicewise_extract(action):
    o = Oracle(DIMS, LS, action)
    v4 = int(np.prod(DIMS))
    canon = rand(LS * v4 * 24, 1)
    full = o.canonical_to_full(canon)
    first, second = full[:o.n].reshape(LS, -1), full[o.n:].reshape(LS, -1)
    for s in range(LS):
        sl = canon[s * v4 * 24:(s + 1) * v4 * 24]
        pf = (s + 1) % 2 if action == "dwf" else 0   # 5D-odd / 4D-even
        assert np.array_equal(first[s], extract_cb(sl, DIMS, pf))
        assert np.array_equal(second[s], extract_cb(sl, DIMS, 1 - pf))
    # the package's numpy gather (an independent construction)
    c5 = canon.view(np.complex128).reshape((LS,) + DIMS[::-1] + (4, 3))
    a, b = lat.canonical_to_halves(c5, DIMS, action)
    assert np.array_equal(first.reshape(-1), lat.to_reals(a))
    assert np.array_equal(second.reshape(-1), lat.to_reals(b))
    assert np.array_equal(o.full_to_canonical(full), canon)


def test_dwf_first_half_is_5d_odd():
    o = Oracle(DIMS, LS, "dwf")
    lx, ly, lz, lt = DIMS
    s, t, z, y, x = np.meshgrid(np.arange(LS), np.arange(lt), np.arange(lz), np.arange(ly),
                                np.arange(lx), indexing="ij")
    par5 = ((x + y + z + t + s) % 2).astype(np.float64)
    canon = np.repeat(par5.reshape(-1), 24)
    full = o.canonical_to_full(canon)
    assert np.all(full[:o.n] == 1.0) and np.all(full[o.n:] == 0.0)


@pytest.mark.parametrize("s_u,s_l", [(0, LS - 1), (2, 2), (LS - 1, 1)])
def test_four2five_five2four(s_u, s_l):
    o = Oracle(DIMS, LS, "dwf")
    v4 = int(np.prod(DIMS))
    four = rand(v4 * 24, 2)
    five = o.four2five(four, s_u, s_l).reshape(LS, v4, 2, 12)
    # upper spin pair (gamma5 = +1) at s_u, lower pair (gamma5 = -1) at s_l
    f4 = four.reshape(v4, 2, 12)
    assert np.array_equal(five[s_u, :, 0], f4[:, 0])
    assert np.array_equal(five[s_l, :, 1], f4[:, 1])
    mask = np.ones((LS, 2), bool)
    mask[s_u, 0] = mask[s_l, 1] = False
    assert not np.any(five.transpose(0, 2, 1, 3)[mask])
    assert np.array_equal(o.five2four(five.reshape(-1), s_u, s_l), four)
    # five2four reads only the two chiral slices
    other = rand(LS * v4 * 24, 3).reshape(LS, v4, 2, 12)
    other[s_u, :, 0] = f4[:, 0]
    other[s_l, :, 1] = f4[:, 1]
    assert np.array_equal(o.five2four(other.reshape(-1), s_u, s_l), four)


def test_plaquette_closed_forms():
    dims = (4, 4, 2, 2)
    lx, ly, lz, lt = dims
    v4 = lx * ly * lz * lt
    unit = lat.gauge_reals(lat.unit_gauge(dims))
    assert np.allclose(plaquette_linktrace(unit, dims), (1.0, 1.0), atol=1e-15)
    # constant abelian flux n through every xy plaquette: theta = 2 pi n/(lx ly)
    n = 1
    th = 2 * np.pi * n / (lx * ly)
    u = lat.unit_gauge(dims)
    for y in range(ly):
        for x in range(lx):
            ph = th * y  # U_x(x, y) = exp(i th y) on the first two colours
            u[:, :, y, x, 0] = np.diag([np.exp(1j * ph), np.exp(-1j * ph), 1.0])
    for x in range(lx):
        ph = -th * ly * x  # U_y(x, ly-1) closes the flux across the y boundary
        u[:, :, ly - 1, x, 1] = np.diag([np.exp(1j * ph), np.exp(-1j * ph), 1.0])
    # U_x(x,y) U_y(x+1,y) U_x(x,y+1)^dag U_y(x,y)^dag = exp(-i th) on colour 0
    # for every xy plaquette; the other five planes give 3
    want = (6 * v4 * 3 - v4 * 3 + v4 * (2 * np.cos(th) + 1)) / (18.0 * v4)
    assert abs(plaquette_linktrace(lat.gauge_reals(u), dims)[0] - want) < 1e-14
    # gauge invariance of the plaquette, not of the link trace
    U = lat.random_gauge(dims, 0.4, 7)
    g = lat.random_gauge(dims, 0.9, 8)[..., 0, :, :]   # one SU(3) per site
    ax = {0: 3, 1: 2, 2: 1, 3: 0}
    V = U.copy()
    for mu in range(4):
        gp = np.roll(g, -1, axis=ax[mu])
        V[..., mu, :, :] = g @ U[..., mu, :, :] @ np.conj(np.swapaxes(gp, -1, -2))
    p0, t0 = plaquette_linktrace(lat.gauge_reals(U), dims)
    p1, t1 = plaquette_linktrace(lat.gauge_reals(V), dims)
    assert abs(p0 - p1) < 1e-13 and abs(t0 - t1) > 1e-3 and p0 < 1.0


# ---- NERSC reader (host code of the library) ------------------------------
@pytest.mark.parametrize("fp", ["IEEE64BIG", "IEEE64LITTLE", "IEEE32BIG", "IEEE32LITTLE"])
@pytest.mark.parametrize("two_rows", [True, False])
def test_nersc_read(tmp_path, fp, two_rows):
    dims = (4, 2, 2, 4)
    links = lat.gauge_reals(lat.random_gauge(dims, 0.4, 11))
    path = os.path.join(tmp_path, "cfg")
    plaq, tr = plaquette_linktrace(links, dims)
    cs = nersc_file.write(path, links, dims, fp, two_rows, plaq, tr)
    got, h = dwf.read_nersc(path)
    assert h.dims == dims and h.recon_row_3 == int(two_rows)
    assert h.fp_bytes == (8 if "64" in fp else 4) and h.big_endian == ("BIG" in fp)
    assert h.checksum == cs == h.checksum_data
    assert abs(h.plaquette - plaq) <= 1e-7 * plaq    # header read as float, as CPS
    want = links if "64" in fp else links.astype(np.float32).astype(np.float64)
    if two_rows:
        # stored rows exact, third row = conj(row0 x row1) of the stored rows
        g = got.reshape(-1, 18)
        w = want.reshape(-1, 18)
        assert np.array_equal(g[:, :12], w[:, :12])
        tol = 1e-15 if "64" in fp else 1e-6
        assert np.max(np.abs(g[:, 12:] - links.reshape(-1, 18)[:, 12:])) < tol
    else:
        assert np.array_equal(got, want)


def test_nersc_checksum_known_words(tmp_path):
    # one site of a 2^4 lattice: every stored real 1.0 (IEEE64BIG words
    # 0x3ff00000, 0x00000000) -> 16 * 4 * 12 * 0x3ff00000 mod 2^32
    dims = (2, 2, 2, 2)
    links = np.ones(16 * 4 * 18)
    path = os.path.join(tmp_path, "ones")
    cs = nersc_file.write(path, links, dims, "IEEE64BIG", True)
    assert cs == (16 * 4 * 12 * 0x3FF00000) % (1 << 32)
    _, h = dwf.read_nersc(path, with_links=False)
    assert h.checksum == cs


def test_nersc_errors(tmp_path):
    dims = (2, 2, 2, 2)
    links = lat.gauge_reals(lat.random_gauge(dims, 0.3, 5))
    p = os.path.join(tmp_path, "c")
    nersc_file.write(p, links, dims, checksum_override=0x1234)
    with pytest.raises(dwf.DwfGpuError) as e:
        dwf.read_nersc(p)
    assert e.value.code == dwf.CHECKSUM_ERROR
    nersc_file.write(p, links, dims, truncate=8)
    with pytest.raises(dwf.DwfGpuError) as e:
        dwf.read_nersc(p)
    assert e.value.code == dwf.IO_ERROR
    nersc_file.write(p, links, dims, end_header=False)
    with pytest.raises(dwf.DwfGpuError) as e:
        dwf.read_nersc(p)
    assert e.value.code == dwf.IO_ERROR
    with pytest.raises(dwf.DwfGpuError) as e:
        dwf.read_nersc(os.path.join(tmp_path, "missing"))
    assert e.value.code == dwf.IO_ERROR
