"""GPU ImpResLanczos against the dense spectrum of MatPcDagMatPc (2x2x2x4,
Ls 4) and the oracle restatement (oracle/lanczos.py) on the same start
vector.  Bar: the nt eigenvalues within 1e-7 relative of the dense ones (the
Lanczos vectors are single precision, as CPS keeps them), eigenvector
residuals below 1e-5, the restart count equal to the oracle's."""
import sys

import numpy as np
import pytest

from cps_veloc_amd import dwf
from cps_veloc_amd import lattice as lat
from cps_veloc_amd.dwf import DiracOpDwfGpu, DiracOpMobiusGpu
from oracle.lanczos import imp_res_lanczos
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(action, dims=(2, 2, 2, 4), ls=4):
    U = lat.random_gauge(dims, 0.3, 3)
    o = Oracle(dims, ls, action, m5=1.8, b=1.5, c=0.5, gauge_canonical=U)
    if action == "dwf":
        g = DiracOpDwfGpu(U, dims, ls, m5=1.8)
    else:
        g = DiracOpMobiusGpu(U, dims, ls, m5=1.8, b=1.5, c=0.5)
    return o, g


def dense(o, A):
    cols = []
    for j in range(o.n // 2):
        e = np.zeros(o.n)
        e[2 * j] = 1.0
        cols.append(A(e).view(np.complex128))
    return np.array(cols).T


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_lanczos_matches_dense_spectrum(action):
    say(f"[gpu] ImpResLanczos {action}")
    mass = 0.05
    o, g = make(action)
    A = lambda x: o.matpcdagmatpc(x, mass)[0]
    lam = np.linalg.eigvalsh(dense(o, A))
    v0 = np.random.default_rng(0).standard_normal(o.n)
    kw = dict(nk=10, nt=4, maxiters=50, npol=8, p0=np.sqrt(1.05 * lam[-1]), p1=np.sqrt(lam[12]))
    l, V, it = g.ImpResLanczos(v0, mass, kw["nk"], kw["nt"], 10, kw["maxiters"], 1e-6, kw["npol"],
                               kw["p0"], kw["p1"])
    Vo, lo, ito = imp_res_lanczos(A, v0, np_=10, stop_res=1e-6, **kw)
    say(f"  restarts gpu={it} oracle={ito}; lambda gpu={l} dense={lam[:4]}")
    assert it == ito
    assert np.all(np.abs(l - lam[:4]) < 1e-7 * lam[:4])
    for k in range(4):
        r = A(V[k]) - l[k] * V[k]
        assert np.linalg.norm(r) / np.linalg.norm(V[k]) < 1e-5
        assert np.array_equal(V[k], V[k].astype(np.float32).astype(np.float64))


def test_lanczos_rejects_bad_arguments():
    o, g = make("dwf")
    v0 = np.ones(o.n)
    with pytest.raises(dwf.DwfGpuError):
        g.ImpResLanczos(v0, 0.05, 4, 6, 4, 10, 1e-6, 8, 1.0, 0.3)   # nt > nk
    with pytest.raises(dwf.DwfGpuError):
        g.ImpResLanczos(v0, 0.05, 4, 2, 4, 10, 1e-6, 8, 1.0, 1.0)   # empty filter interval
