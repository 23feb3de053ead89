"""GPU: whole-field formats and the gauge observables through the C ABI.

Bar: the format kernels move FP64 values, so their results equal the
oracle's bit for bit; the MatInv chain (Ffour2five -> MatInv -> Ffive2four)
within 1e-9 relative of the oracle's chain (two CG solves to 1e-10 that run
the same iterations); the plaquette within 1e-13 of the numpy oracle."""
import os
import sys

import numpy as np
import pytest

from cps_veloc_amd import dwf
from cps_veloc_amd import lattice as lat
from cps_veloc_amd.dwf import CgArg, DiracOpDwfGpu, DiracOpMobiusGpu
from oracle.oracle import Oracle, plaquette_linktrace
import nersc_file

pytestmark = pytest.mark.gpu

CASES = [("dwf", (4, 2, 2, 6), 4), ("mobius", (2, 4, 2, 4), 6)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(action, dims, ls, seed=31):
    U = lat.random_gauge(dims, 0.3, seed)
    o = Oracle(dims, ls, action, m5=1.8, b=1.5, c=0.5, gauge_canonical=U)
    if action == "dwf":
        g = DiracOpDwfGpu(U, dims, ls, m5=1.8)
    else:
        g = DiracOpMobiusGpu(U, dims, ls, m5=1.8, b=1.5, c=0.5)
    return U, o, g


def rand(n, seed):
    return np.random.default_rng(seed).standard_normal(n)


@pytest.mark.parametrize("action,dims,ls", CASES)
def test_canonical_round_trip(action, dims, ls):
    say(f"[gpu] canonical <-> halves {action} {dims} ls={ls}")
    U, o, g = make(action, dims, ls)
    canon = rand(ls * int(np.prod(dims)) * 24, 3)
    a, b = g.fermion_from_canonical(canon)
    full = o.canonical_to_full(canon)
    assert np.array_equal(a.download(), full[:o.n])
    assert np.array_equal(b.download(), full[o.n:])
    assert np.array_equal(g.fermion_to_canonical(a, b), canon)


@pytest.mark.parametrize("action,dims,ls", CASES)
@pytest.mark.parametrize("s_u,s_l", [(0, -1), (1, 1)])
def test_four2five_five2four(action, dims, ls, s_u, s_l):
    s_l = s_l % ls
    say(f"[gpu] Ffour2five/Ffive2four {action} {dims} ls={ls} s_u={s_u} s_l={s_l}")
    U, o, g = make(action, dims, ls)
    four = rand(int(np.prod(dims)) * 24, 4)
    a, b = g.Ffour2five(four, s_u, s_l)
    full = o.canonical_to_full(o.four2five(four, s_u, s_l))
    assert np.array_equal(a.download(), full[:o.n])
    assert np.array_equal(b.download(), full[o.n:])
    assert np.array_equal(g.Ffive2four(a, b, s_u, s_l), four)
    with pytest.raises(dwf.DwfGpuError):
        g.Ffour2five(four, ls, 0)


@pytest.mark.parametrize("action,dims,ls", CASES)
def test_point_source_propagator_chain(action, dims, ls):
    """QPropW-style: 4D source -> Ffour2five(0, Ls-1) -> MatInv -> Ffive2four."""
    say(f"[gpu] four2five -> MatInv_dev -> five2four {action} {dims} ls={ls}")
    U, o, g = make(action, dims, ls)
    mass = 0.1
    four = np.zeros(int(np.prod(dims)) * 24)
    four[0 * 24 + 2 * 4] = 1.0          # spin 1, colour 1 at the origin
    four[5 * 24 + 2 * 9 + 1] = 0.5      # spin 3, colour 0, imaginary, site 5
    a, b = g.Ffour2five(four, 0, ls - 1)
    oa, ob = g.vector(), g.vector()
    it, tr, st = g.MatInv_dev(oa, ob, a, b, CgArg(mass, 5000, 1e-10))
    got4 = g.Ffive2four(oa, ob, ls - 1, 0)   # the 4D quark field q = P_L psi(0) + P_R psi(Ls-1)
    src_full = o.canonical_to_full(o.four2five(four, 0, ls - 1))
    sol_full, oit, otr = o.matinv(src_full, mass, stop_rsd=1e-10, max_iter=5000)
    want4 = o.five2four(o.full_to_canonical(sol_full), ls - 1, 0)
    say(f"  iters gpu={it} oracle={oit} true_res gpu={tr:.3e} oracle={otr:.3e}")
    assert st == 0 and it == oit
    assert np.linalg.norm(got4 - want4) / np.linalg.norm(want4) < 1e-9
    # the sources were not touched (PRESERVE_YES)
    assert np.array_equal(g.Ffive2four(a, b, 0, ls - 1), four)


def test_plaquette_matches_oracle():
    dims, ls = (4, 4, 2, 6), 4
    say(f"[gpu] plaquette {dims}")
    U, o, g = make("dwf", dims, ls, seed=77)
    plaq, tr = g.plaquette()
    wp, wt = plaquette_linktrace(lat.gauge_reals(U), dims)
    assert abs(plaq - wp) < 1e-13 and abs(tr - wt) < 1e-13


def test_nersc_load_and_check(tmp_path):
    dims, ls = (4, 2, 2, 4), 4
    say(f"[gpu] NERSC load {dims}")
    U = lat.gauge_reals(lat.random_gauge(dims, 0.4, 13))
    plaq, tr = plaquette_linktrace(U, dims)
    path = os.path.join(tmp_path, "cfg")
    nersc_file.write(path, U, dims, "IEEE64BIG", True, plaq, tr)
    links, h = dwf.read_nersc(path)
    g = DiracOpMobiusGpu(links, h.dims, ls, m5=1.8, b=1.5, c=0.5)
    p, t = dwf.check_plaquette(g, h, 1e-6)
    assert abs(p - plaq) < 1e-14
    bad = os.path.join(tmp_path, "bad")
    nersc_file.write(bad, U, dims, "IEEE64BIG", True, plaq * 1.001, tr)
    links, h = dwf.read_nersc(bad)
    g2 = DiracOpMobiusGpu(links, h.dims, ls, m5=1.8, b=1.5, c=0.5)
    with pytest.raises(ValueError):
        dwf.check_plaquette(g2, h, 1e-6)
