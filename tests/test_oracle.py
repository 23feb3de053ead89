"""Pin the CPU oracle against what the mathematics fixes.

CPS ships no golden vectors for this path and cannot be built here without
stand-ins for generated/QUDA headers, so the oracle is pinned by:
* an independent numpy textbook Wilson / Shamir / Mobius operator
  (tests/textbook.py) and dense even-odd Schur complements,
* the adjoint identities <x, A y> = <A^+ x, y>,
* M5 * M5^-1 = 1 and a dense inverse of the (s,spin) block,
* CG converging in exactly k iterations on a source spanned by k eigenvectors,
* full-system solves of MatInv checked against the textbook operator.
Tolerances: FP64 rounding only, 1e-12 relative unless stated.
"""
import numpy as np
import pytest

import textbook as tb
from cps_veloc_amd import lattice as lat
from oracle.oracle import Oracle

M5 = 1.8


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def cplx(v):
    return np.ascontiguousarray(v).view(np.complex128)


def setup(dims, ls, action, eps=0.3, seed=11, bc=(0, 0, 0, 1), b=1.5, c=0.5):
    U = lat.random_gauge(dims, eps, seed)
    o = Oracle(dims, ls, action, m5=M5, b=b, c=c, gauge_canonical=U, bc=bc)
    return U, o


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 2, 6, 4), (2, 4, 2, 6)])
@pytest.mark.parametrize("cb", [0, 1])
@pytest.mark.parametrize("dag", [0, 1])
def test_wilson_hop_vs_textbook(dims, cb, dag):
    U, o = setup(dims, 1, "dwf", bc=(0, 0, 0, 1))
    psi = lat.gaussian_spinor((1,) + tuple(reversed(dims)), seed=3)
    full = tb.wilson_hop(U, psi[0], dag=bool(dag))
    # oracle acts on the checkerboard cb of the input, producing parity 1-cb
    g = lat._half_gather(dims)
    flat_in = psi[0].reshape(-1, 4, 3)
    half_in = lat.to_reals(flat_in[g[cb]])
    out = o.wilson_dslash(half_in, cb, dag)
    want = full.reshape(-1, 4, 3)[g[1 - cb]]
    assert rel(cplx(out).reshape(want.shape), want) < 1e-13


@pytest.mark.parametrize("bc", [(0, 0, 0, 1), (1, 1, 1, 1), (0, 0, 0, 0)])
def test_boundary_conditions_vs_textbook(bc):
    dims = (4, 2, 2, 4)
    U, o = setup(dims, 1, "dwf", bc=bc)
    psi = lat.gaussian_spinor(tuple(reversed(dims)), seed=5)
    full = tb.wilson_hop(U, psi, bc=bc)
    g = lat._half_gather(dims)
    out = o.wilson_dslash(lat.to_reals(psi.reshape(-1, 4, 3)[g[1]]), 1, 0)
    assert rel(cplx(out).reshape(-1, 4, 3), full.reshape(-1, 4, 3)[g[0]]) < 1e-13


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_serial_hop_same_bits(action):
    """The CPS-faithful timing variant (4D hop in one thread, as
    wilson_dslash.C:280-477) computes exactly what the OpenMP one does."""
    from oracle import oracle as orc
    dims, ls = (4, 4, 4, 8), 4
    _, o = setup(dims, ls, action)
    x = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 8))
    want, dwant = o.matpcdagmatpc(x, 0.04)
    orc.set_serial_hop(True)
    try:
        got, dgot = o.matpcdagmatpc(x, 0.04)
    finally:
        orc.set_serial_hop(False)
    assert np.array_equal(got, want) and dgot == dwant


def test_free_field_plane_wave():
    """U = 1: a plane wave e^{ipx} is an eigenvector of the hopping term with
    multiplier sum_mu [(1-g_mu) e^{ip_mu} + (1+g_mu) e^{-ip_mu}] (closed form)."""
    dims = (4, 4, 4, 8)
    o = Oracle(dims, 1, "dwf", gauge_canonical=lat.unit_gauge(dims), bc=(0, 0, 0, 0))
    n = [1, 2, 3, 5]
    p = [2 * np.pi * n[m] / dims[m] for m in range(4)]
    mult = sum((np.eye(4) - tb.GAMMA[m]) * np.exp(1j * p[m]) +
               (np.eye(4) + tb.GAMMA[m]) * np.exp(-1j * p[m]) for m in range(4))
    chi0 = np.random.default_rng(0).standard_normal((4, 3)) + 0j
    lx, ly, lz, lt = dims
    t, z, y, x = np.meshgrid(*[np.arange(d) for d in (lt, lz, ly, lx)], indexing="ij")
    phase = np.exp(1j * (p[0] * x + p[1] * y + p[2] * z + p[3] * t))
    psi = phase[..., None, None] * chi0
    g = lat._half_gather(dims)
    out = cplx(o.wilson_dslash(lat.to_reals(psi.reshape(-1, 4, 3)[g[0]]), 0, 0))
    want = (phase[..., None, None] * (mult @ chi0)).reshape(-1, 4, 3)[g[1]]
    assert rel(out.reshape(want.shape), want) < 1e-13


def _embed(vec_half, dims, ls, kind, which):
    """CPS half vector (reals) -> canonical 5D field with zeros elsewhere."""
    vh = int(np.prod(dims)) // 2
    h = lat.from_reals(vec_half, ls, vh)
    z = np.zeros_like(h)
    a, b = (h, z) if which == 0 else (z, h)
    return lat.halves_to_canonical(a, b, dims, kind)


@pytest.mark.parametrize("dims,ls", [((4, 4, 2, 2), 4), ((2, 4, 4, 2), 6)])
@pytest.mark.parametrize("mass", [0.04, 0.3])
def test_dwf_matpc_vs_textbook_schur(dims, ls, mass):
    U, o = setup(dims, ls, "dwf")
    x = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 7))
    got = o.matpc(x, mass, 0)
    f = _embed(x, dims, ls, "dwf", 0)  # odd half is the first half
    want = tb.dwf_schur_odd(U, f, mass, M5)
    odd, even = lat.canonical_to_halves(want, dims, "dwf")
    assert np.abs(even).max() == 0.0
    assert rel(cplx(got).reshape(odd.shape), odd) < 1e-13


@pytest.mark.parametrize("dag", [0, 1])
def test_dwf_dslash_vs_textbook(dag):
    dims, ls, mass = (4, 2, 2, 4), 4, 0.1
    U, o = setup(dims, ls, "dwf")
    kappa = 1.0 / (2 * (4 - M5 + 1.0))
    x = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 9))
    got = o.dslash(x, mass, 1, dag)  # odd -> even
    f = _embed(x, dims, ls, "dwf", 0)
    if dag == 0:
        hop = -(2 * kappa * tb.dwf_full(U, f, mass, M5) - f) / kappa
        _, even = lat.canonical_to_halves(hop, dims, "dwf")
        assert rel(cplx(got).reshape(even.shape), even) < 1e-13
    else:
        # <y, D^dag x> = <D_{e->o} y, x> with the textbook D_{e->o}
        y = lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 10)
        fy = lat.halves_to_canonical(np.zeros_like(y), y, dims, "dwf")
        hy = -(2 * kappa * tb.dwf_full(U, fy, mass, M5) - fy) / kappa
        odd_hy, _ = lat.canonical_to_halves(hy, dims, "dwf")
        lhs = np.vdot(y.reshape(-1), cplx(got))
        rhs = np.vdot(odd_hy.reshape(-1), cplx(x))
        assert abs(lhs - rhs) / abs(rhs) < 1e-13


@pytest.mark.parametrize("dims,ls", [((4, 4, 2, 2), 4), ((2, 2, 4, 4), 6)])
@pytest.mark.parametrize("b,c", [(1.5, 0.5), (1.0, 0.0), (2.0, 1.0)])
def test_mobius_matpc_vs_textbook_schur(dims, ls, b, c):
    mass = 0.05
    U, o = setup(dims, ls, "mobius", b=b, c=c)
    x = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 13))
    got = o.matpc(x, mass, 0)
    f = _embed(x, dims, ls, "mobius", 1)  # odd is the second half
    want = tb.mobius_schur_odd(U, f, mass, M5, b, c)
    even, odd = lat.canonical_to_halves(want, dims, "mobius")
    assert np.abs(even).max() < 1e-14
    assert rel(cplx(got).reshape(odd.shape), odd) < 1e-13


def test_mobius_dslash4_vs_textbook():
    dims, ls, mass, b, c = (4, 2, 2, 4), 4, 0.07, 1.5, 0.5
    U, o = setup(dims, ls, "mobius", b=b, c=c)
    x = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 17))
    got = o.dslash(x, mass, 1, 0)  # odd -> even, D_hop (b + c hop5)
    f = _embed(x, dims, ls, "mobius", 1)
    t = b * f + c * tb.hop5(f, mass)
    want = np.stack([tb.wilson_hop(U, p) for p in t])
    even, _ = lat.canonical_to_halves(want, dims, "mobius")
    assert rel(cplx(got).reshape(even.shape), even) < 1e-13


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_matpc_adjoint_and_mdagm(action):
    dims, ls, mass = (4, 4, 2, 2), 4, 0.02
    U, o = setup(dims, ls, action)
    vh = int(np.prod(dims)) // 2
    x = lat.to_reals(lat.gaussian_spinor((ls, vh), 21))
    y = lat.to_reals(lat.gaussian_spinor((ls, vh), 22))
    mx = o.matpc(x, mass, 0)
    mdy = o.matpc(y, mass, 1)
    lhs = np.vdot(cplx(y), cplx(mx))
    rhs = np.vdot(cplx(mdy), cplx(x))
    assert abs(lhs - rhs) / abs(lhs) < 1e-13
    out, dot = o.matpcdagmatpc(x, mass)
    assert rel(out, o.matpc(mx, mass, 1)) < 1e-14
    assert abs(dot - np.dot(mx, mx)) / dot < 1e-14
    # MdagM is Hermitian positive: <x, A x> = |M x|^2 > 0
    assert abs(np.vdot(cplx(x), cplx(out)) - dot) / dot < 1e-12


@pytest.mark.parametrize("dag", [0, 1])
@pytest.mark.parametrize("ls", [4, 8, 12])
def test_m5inv_identity_and_dense(dag, ls):
    dims, mass, b, c = (2, 2, 2, 4), 0.01, 1.5, 0.5
    U, o = setup(dims, ls, "mobius", b=b, c=c)
    kb, kc = o.kappas()
    K = np.full(ls, kb / kc, dtype=np.complex128)
    vh = int(np.prod(dims)) // 2
    x = lat.to_reals(lat.gaussian_spinor((ls, vh), 31))
    y = o.m5inv(x, mass, dag, K)
    # M5 y = y + K D5 y   (D5 via the oracle's own 5D hop)
    back = o.kappa_dslash_5_plus(y, y, mass, dag, K)
    assert rel(back, x) < 1e-14
    # dense (s,spin) inverse from the textbook projectors
    mat = tb.mobius_m5_matrix(ls, mass, kb / kc)
    if dag:
        mat = mat.conj().T
    xs = lat.from_reals(x, ls, vh)
    want = tb.apply_s_spin(np.linalg.inv(mat), xs)
    assert rel(lat.from_reals(y, ls, vh), want) < 1e-13


def _dense_mdagm(o, mass):
    n = o.n
    A = np.zeros((n, n))
    e = np.zeros(n)
    for k in range(n):
        e[:] = 0.0
        e[k] = 1.0
        A[:, k] = o.matpcdagmatpc(e, mass)[0]
    return A


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_cg_iterations_equal_eigen_count(action):
    """src spanned by k eigenvectors of distinct eigenvalues: CG ends after k
    iterations and InvCgShift returns k+1 (inv_cg.C:446)."""
    dims, ls, mass = (2, 2, 2, 2), 4, 0.1
    U, o = setup(dims, ls, action)
    A = _dense_mdagm(o, mass)
    assert np.abs(A - A.T).max() < 1e-12
    w, V = np.linalg.eigh(A)
    for k in (1, 2, 3):
        idx = [5, 80, 200][:k]
        src = V[:, idx].sum(axis=1)
        _, it, tr, hist = o.invcg(src, mass, stop_rsd=1e-8, max_iter=50)
        assert it == k + 1, (k, it, hist)
        assert tr < 1e-8


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_cg_shift_eigen_count_and_closed_form(action):
    """InvCgShift with a shift s (inv_cg.C:248-251,304-310): (A + s) has A's
    eigenvectors, so a source spanned by k of them ends after k iterations
    with the solution sum c_i v_i / (lambda_i + s)."""
    dims, ls, mass, s = (2, 2, 2, 2), 4, 0.1, 0.37
    U, o = setup(dims, ls, action)
    A = _dense_mdagm(o, mass)
    w, V = np.linalg.eigh(A)
    for k in (1, 2, 3):
        idx = [7, 90, 210][:k]
        coef = np.array([1.0, -0.5, 2.0])[:k]
        src = V[:, idx] @ coef
        sol, it, tr, _ = o.invcg_shift(src, mass, shift=s, stop_rsd=1e-9, max_iter=50)
        assert it == k + 1, (k, it)
        want = V[:, idx] @ (coef / (w[idx] + s))
        assert rel(sol, want) < 1e-12
        assert tr < 1e-9
    # shift NULL and shift 0 are the unshifted CG
    src = V[:, [3, 50, 120, 300]].sum(axis=1)
    a = o.invcg(src, mass, stop_rsd=1e-10)
    b = o.invcg_shift(src, mass, shift=0.0, stop_rsd=1e-10)
    assert a[1] == b[1] and rel(b[0], a[0]) < 1e-14


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_cg_shift_solution_and_src_norm_sq(action):
    """The shifted solve against the dense solve of (A + s) x = b and the
    pinned multi-shift oracle; a caller-supplied src_norm_sq sets the stop
    rule stop_rsd^2 src_norm_sq and the true-residual denominator
    (inv_cg.C:166-174, 403-415)."""
    dims, ls, mass, s = (2, 2, 2, 4), 4, 0.05, 0.02
    U, o = setup(dims, ls, action)
    vh = int(np.prod(dims)) // 2
    b = lat.to_reals(lat.gaussian_spinor((ls, vh), 43))
    A = _dense_mdagm(o, mass)
    x, it, tr, _ = o.invcg_shift(b, mass, shift=s, stop_rsd=1e-12, max_iter=3000)
    want = np.linalg.solve(A + s * np.eye(o.n), b)
    assert rel(x, want) < 1e-10
    r = b - A @ x - s * x
    assert abs(tr - np.linalg.norm(r) / np.linalg.norm(b)) < 1e-14
    psi, _, _, its = o.minvcg(b, mass, [s], 1e-12)
    assert rel(x, psi[0]) < 1e-10 and abs(its[0] - (it - 1)) <= 1
    nb2 = float(np.dot(b, b))
    x1, it1, tr1, _ = o.invcg_shift(b, mass, src_norm_sq=4 * nb2, stop_rsd=1e-9)
    x2, it2, tr2, _ = o.invcg_shift(b, mass, src_norm_sq=0.0, stop_rsd=2e-9)
    assert it1 == it2 and np.array_equal(x1, x2)
    assert abs(tr1 - 0.5 * tr2) < 1e-15


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_cg_solution_and_edges(action):
    dims, ls, mass = (4, 4, 2, 2), 4, 0.05
    U, o = setup(dims, ls, action)
    vh = int(np.prod(dims)) // 2
    src = lat.to_reals(lat.gaussian_spinor((ls, vh), 41))
    sol, it, tr, hist = o.invcg(src, mass, stop_rsd=1e-11, max_iter=2000)
    r = src - o.matpcdagmatpc(sol, mass)[0]
    assert np.linalg.norm(r) / np.linalg.norm(src) < 2e-11
    assert abs(tr - np.linalg.norm(r) / np.linalg.norm(src)) < 1e-15
    assert hist[-1] <= (1e-11) ** 2 * np.dot(src, src) < hist[-2]
    assert it == len(hist)  # itr+1 counts the initial residual entry
    # initial guess = solution: |res|^2 <= stp_cnd at start -> 0 iterations
    _, it0, _, _ = o.invcg(src, mass, stop_rsd=1e-6, max_iter=100, guess=sol)
    assert it0 == 1
    # max iterations: returns max_num_iter
    _, itm, trm, _ = o.invcg(src, mass, stop_rsd=1e-14, max_iter=5)
    assert itm == 5 and trm > 1e-6


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_matinv_solves_full_textbook_system(action):
    dims, ls, mass = (4, 2, 2, 4), 4, 0.1
    U, o = setup(dims, ls, action)
    src5 = lat.gaussian_spinor((ls,) + tuple(reversed(dims)), 51)
    first, second = lat.canonical_to_halves(src5, dims, action)
    src_full = np.concatenate([lat.to_reals(first), lat.to_reals(second)])
    keep = src_full.copy()
    out, it, tr = o.matinv(src_full, mass, stop_rsd=1e-12, max_iter=5000)
    assert np.array_equal(src_full, keep)  # PRESERVE_YES
    vh = int(np.prod(dims)) // 2
    a = lat.from_reals(out[: o.n], ls, vh)
    b = lat.from_reals(out[o.n:], ls, vh)
    x5 = lat.halves_to_canonical(a, b, dims, action)
    if action == "dwf":
        kappa = 1.0 / (2 * (4 - M5 + 1.0))
        back = 2 * kappa * tb.dwf_full(U, x5, mass, M5)
    else:
        back = tb.mobius_full(U, x5, mass, M5, 1.5, 0.5)
    assert rel(back, src5) < 1e-9


def test_layout_roundtrip():
    dims, ls = (4, 2, 6, 2), 3
    f = lat.gaussian_spinor((ls,) + tuple(reversed(dims)), 61)
    for kind in ("dwf", "mobius"):
        a, b = lat.canonical_to_halves(f, dims, kind)
        assert np.array_equal(lat.halves_to_canonical(a, b, dims, kind), f)


def test_gauge_is_su3():
    U = lat.random_gauge((4, 4, 2, 2), 0.4, 3).reshape(-1, 3, 3)
    eye = np.eye(3)
    assert np.abs(U @ np.conj(np.swapaxes(U, -1, -2)) - eye).max() < 1e-14
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-14
    # deterministic
    assert np.array_equal(U, lat.random_gauge((4, 4, 2, 2), 0.4, 3).reshape(-1, 3, 3))


@pytest.mark.parametrize("action", ["dwf", "mobius"])
def test_minvcg_solves_every_shift(action):
    """MInvCG: each psi_s solves (A + shift_s) psi_s = chi to its stop
    residual (checked with the pinned MatPcDagMatPc), SINGLE is the alpha sum
    of MULTI, and with one zero shift it takes InvCg's number of updates."""
    dims, ls, mass = (4, 4, 2, 4), 4, 0.1
    U, o = setup(dims, ls, action)
    chi = lat.to_reals(lat.gaussian_spinor((ls, int(np.prod(dims)) // 2), 21))
    shifts = np.array([0.0, 0.01, 0.1, 0.5])
    psi, k, rs, its = o.minvcg(chi, mass, shifts, 1e-10, multi=True)
    nchi = np.linalg.norm(chi)
    for s, sh in enumerate(shifts):
        ax, _ = o.matpcdagmatpc(psi[s], mass)
        res = np.linalg.norm(chi - ax - sh * psi[s]) / nchi
        assert res < 1e-9, (s, res)
        assert 0 < its[s] < k and rs[s] < 1e-10 * nchi
    assert its[0] >= its[1] >= its[2] >= its[3]  # larger shift, faster
    alpha = np.array([0.3, -1.2, 0.7, 2.0])
    single, k1, _, _ = o.minvcg(chi, mass, shifts, 1e-10, multi=False, alpha=alpha)
    assert k1 == k
    assert rel(single, (alpha[:, None] * psi).sum(0)) < 1e-12
    # one zero shift: the same CG updates as InvCg (which counts one more)
    p1, k0, _, _ = o.minvcg(chi, mass, [0.0], 1e-10)
    sol, it, _, _ = o.invcg(chi, mass, stop_rsd=1e-10)
    assert k0 == it - 1
    assert rel(p1[0], sol) < 1e-12
