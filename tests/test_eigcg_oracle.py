"""The eigCG restatement (oracle/eigcg.py) pinned on the CPU:

* without deflation space its CG iterates are InvCg's (same solution, the
  loop count two below InvCg's itr+1 count);
* its lowest Ritz values approach the lowest eigenvalues of the dense
  MatPcDagMatPc of a tiny lattice;
* the init-CG guess is the exact solution on span(U), so a source inside an
  exact eigenspace needs no iteration, and U^dag r = 0 for any source;
* the eig_Run bookkeeping keeps H = U^dag A U and cuts the iterations of
  later right-hand sides."""
import numpy as np
import pytest

from cps_veloc_amd import lattice as lat
from oracle.eigcg import EigRun, cdot, inveigcg
from oracle.oracle import Oracle


def setup(action="dwf", dims=(2, 2, 2, 4), ls=4, mass=0.05, seed=3):
    U = lat.random_gauge(dims, 0.3, seed)
    o = Oracle(dims, ls, action, m5=1.8, b=1.5, c=0.5, gauge_canonical=U)
    return o, (lambda x: o.matpcdagmatpc(x, mass)), mass


def dense(o, A):
    n = o.n // 2
    cols = []
    for j in range(n):
        e = np.zeros(o.n)
        e[2 * j] = 1.0
        cols.append(A(e)[0].view(np.complex128))
    return np.array(cols).T


def src(o, seed):
    return np.random.default_rng(seed).standard_normal(o.n)


def test_eigcg_iterates_are_cg():
    o, A, mass = setup()
    b = src(o, 1)
    x, it, tr, V, M = inveigcg(A, b, 1e-10, 1000, nev=3, m=12)
    x0, it0, tr0, _ = o.invcg(b, mass, stop_rsd=1e-10, max_iter=1000)
    assert it == it0 - 2
    assert np.linalg.norm(x - x0) / np.linalg.norm(x0) < 1e-11
    assert tr < 1e-10


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_ritz_values_approach_lowest_eigenvalues(action):
    o, A, mass = setup(action)
    lam = np.linalg.eigvalsh(dense(o, A))
    x, it, tr, V, M = inveigcg(A, src(o, 2), 1e-12, 2000, nev=4, m=20)
    m = np.sort(M[M > 0])
    # the CG run is short (tens of iterations) and the low end is clustered:
    # the lowest Ritz value is close, none lies below its eigenvalue
    # (Ritz values of a subspace interlace: theta_k >= lambda_k)
    assert abs(m[0] - lam[0]) / lam[0] < 1e-3
    assert np.all(m[:6] >= lam[:6] * (1 - 1e-12))
    # every Ritz pair: unit vector whose Rayleigh quotient is its value
    for k in range(len(M)):
        if M[k] > 0:
            v = V[k]
            av = A(v)[0]
            assert abs(np.linalg.norm(v) - 1.0) < 1e-8
            assert abs(cdot(v, av).real - M[k]) < 1e-8 * M[k]
    k = int(np.argmin(np.where(M > 0, M, np.inf)))
    assert np.linalg.norm(A(V[k])[0] - M[k] * V[k]) < 0.05 * M[k]


def test_init_cg_guess_is_exact_on_span_u():
    o, A, mass = setup()
    Ad = dense(o, A)
    lam, vec = np.linalg.eigh(Ad)
    U = [np.ascontiguousarray(vec[:, k]).view(np.float64).copy() for k in range(6)]
    H = np.array([[cdot(U[i], A(U[j])[0]) for j in range(6)] for i in range(6)])
    invH = np.linalg.inv(H)
    coef = np.random.default_rng(4).standard_normal(6)
    b = sum(c * u for c, u in zip(coef, U))
    x, it, tr, _, _ = inveigcg(A, b, 1e-10, 100, U=U, invH=invH)
    assert it == 0 and tr < 1e-12
    # any source: the first residual is orthogonal to U
    b = src(o, 5)
    sol0 = np.zeros(o.n).view(np.complex128)
    ub = np.array([cdot(u, b) for u in U])
    for c, u in zip(invH @ ub, U):
        sol0 += c * u.view(np.complex128)
    r = b - A(sol0.view(np.float64))[0]
    assert max(abs(cdot(u, r)) for u in U) < 1e-12 * np.linalg.norm(b)


def rough(dims=(4, 2, 2, 4), ls=8, mass=0.01):
    """A rough gauge field (eps = 1): its low modes are separated enough for
    deflation to pay on a lattice this small."""
    U = lat.random_gauge(dims, 1.0, 3)
    o = Oracle(dims, ls, "dwf", m5=1.8, gauge_canonical=U)
    return o, (lambda x: o.matpcdagmatpc(x, mass)), mass


def test_eig_run_grows_space_and_cuts_iterations():
    o, A, mass = rough()
    run = EigRun(A, nev=6, m=30, max_def_len=36, max_eig=1.0)
    its = []
    for k in range(4):
        b = src(o, 10 + k)
        x, it, tr = run.solve(b, 1e-10, 2000)
        assert tr < 1e-10
        its.append(it)
    dl = len(run.U)
    assert dl == 36
    # H is U^dag A U of the stored (single-precision) vectors
    H = np.array([[cdot(run.U[i], A(run.U[j])[0]) for j in range(dl)] for i in range(dl)])
    assert np.allclose(run.H[:dl, :dl], H, rtol=0, atol=1e-12 * np.abs(H).max())
    assert all(np.array_equal(u, u.astype(np.float32).astype(np.float64)) for u in run.U)
    assert its[-1] < 0.8 * its[0]
