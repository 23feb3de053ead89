"""The implicitly restarted Lanczos restatement (oracle/lanczos.py) pinned
on the CPU by the dense spectrum of MatPcDagMatPc on a 2x2x2x4, Ls 4
lattice: the nt returned Rayleigh quotients are its nt lowest eigenvalues
and the vectors are eigenvectors to the stop residual (single precision
bounds both), for DWF and Mobius; the Chebyshev filter is pinned by the
scalar Chebyshev polynomial on eigenvectors."""
import numpy as np
import pytest

from cps_veloc_amd import lattice as lat
from oracle.lanczos import cheby, imp_res_lanczos
from oracle.oracle import Oracle


def setup(action, dims=(2, 2, 2, 4), ls=4, mass=0.05):
    U = lat.random_gauge(dims, 0.3, 3)
    o = Oracle(dims, ls, action, m5=1.8, b=1.5, c=0.5, gauge_canonical=U)
    return o, (lambda x: o.matpcdagmatpc(x, mass)[0])


def dense(o, A):
    cols = []
    for j in range(o.n // 2):
        e = np.zeros(o.n)
        e[2 * j] = 1.0
        cols.append(A(e).view(np.complex128))
    return np.array(cols).T


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_lanczos_finds_lowest_eigenpairs(action):
    o, A = setup(action)
    lam, vec = np.linalg.eigh(dense(o, A))
    v0 = np.random.default_rng(0).standard_normal(o.n)
    V, l, it = imp_res_lanczos(A, v0, nk=10, nt=4, np_=10, maxiters=50, stop_res=1e-6,
                               npol=8, p0=np.sqrt(1.05 * lam[-1]), p1=np.sqrt(lam[12]))
    assert np.all(np.abs(l - lam[:4]) < 1e-7 * lam[:4])
    for k in range(4):
        v = V[k]
        r = A(v) - l[k] * v
        assert np.linalg.norm(r) / np.linalg.norm(v) < 1e-5
        assert abs(np.linalg.norm(v) - 1.0) < 1e-6


def test_chebyshev_filter_on_eigenvectors():
    o, A = setup("dwf")
    lam, vec = np.linalg.eigh(dense(o, A))
    p0, p1 = np.sqrt(1.05 * lam[-1]), np.sqrt(lam[10])
    a, b = p0 * p0, p1 * p1
    for k in (0, 5, 40):
        v = np.ascontiguousarray(vec[:, k]).view(np.float64).copy()
        q = (2 * lam[k] - (a + b)) / (a - b)
        for n in (1, 2, 7):
            want = np.polynomial.chebyshev.chebval(q, [0] * n + [1])
            got = cheby(A, v, n, p0, p1)
            assert np.linalg.norm(got - want * v) < 1e-10 * max(1.0, abs(want))
