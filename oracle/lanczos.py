"""CPU restatement of DiracOp::ImpResLanczos (implicitly restarted Lanczos
with a Chebyshev filter) around the C oracle's MatPcDagMatPc.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Follows cps_pp/src/util/dirac_op/d_op_base/comsrc/lanczos.C:193-736 and the
Lanczos step :756-840, with the Chebyshev filter of DiracOpDwf::RitzMat
(d_op_dwf/d_op_dwf.C:852-940):
  Q = c1/2 (c0 A - 1), c0 = 2/(a+b), c1 = 2(a+b)/(a-b), a = p0^2,
  b = (p1 + |shift|)^2, out = T_Npol(Q) in by the three-term recurrence;
the m = nk + np step Lanczos factorisation with single-precision Lanczos
vectors and a full Gram-Schmidt of each new residual against them; each
restart applies np implicit QR steps whose shifts are the np smallest Ritz
values of T_m (the filter maps the wanted low modes of A to the top), keeps
nk vectors, and every conv_check restarts checks the nt lowest Ritz vectors
of A against stop_residual (|A v - <v,A v> v| / |v|).  On return V[0..nt)
are those Ritz vectors and lam[0..nt) their Rayleigh quotients.

Dense tridiagonal work uses numpy (eigh; explicit QR steps with Givens
rotations); CPS uses tqri / QRtrf (eigen_solver.C:206-305).  The restarted
Krylov space does not depend on how the shifted QR steps are carried out.
"""
from __future__ import annotations

import numpy as np


def cheby(A, x, npol, p0, p1, shift=0.0):
    a, b = p0 * p0, (p1 + abs(shift)) ** 2
    c1, c0 = 2.0 * (a + b) / (a - b), 2.0 / (a + b)
    t_prev = x.copy()
    t = 0.5 * c1 * (c0 * A(x) - x)
    if npol == 1:
        return t
    out = None
    for _ in range(2, npol + 1):
        out = c1 * (c0 * A(t) - t) - t_prev
        t_prev, t = t, out
    return out


def _f32(v):
    return v.astype(np.float32).astype(np.float64)


def _cdot(a, b):
    return np.vdot(a.view(np.complex128), b.view(np.complex128))


def _qr_step(alpha, beta, Q, mu):
    """One shifted QR step T - mu = QR, T <- RQ + mu on the tridiagonal
    (alpha, beta); Q <- Q G (returns new alpha, beta, Q)."""
    m = alpha.size
    T = np.diag(alpha) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
    q, r = np.linalg.qr(T - mu * np.eye(m))
    # sign convention: positive diagonal of R keeps q continuous
    s = np.sign(np.diag(r))
    s[s == 0] = 1.0
    q = q * s
    r = (r.T * s).T
    T2 = r @ q + mu * np.eye(m)
    a2 = np.diag(T2).copy()
    b2 = beta.copy()
    b2[:m - 1] = np.diag(T2, 1)
    return a2, b2, Q @ q


def lanczos_ext(Afilt, k0, m, r, V, alpha, beta):
    V[k0] = _f32(r / np.sqrt(np.dot(r, r)))
    for k in range(k0, m):
        v = V[k]
        r = Afilt(v)
        if k > 0:
            r = r - beta[k - 1] * V[k - 1]
        alpha[k] = float(np.dot(v, r))
        r = r - alpha[k] * v
        beta[k] = np.sqrt(float(np.dot(r, r)))
        if k > 0:
            for i in range(k + 1):       # Gram-Schmidt against V[0..k]
                rc = r.view(np.complex128)
                rc -= _cdot(V[i], r) * V[i].view(np.complex128)
        if k + 1 < m:
            V[k + 1] = _f32(r / beta[k])
    return r


def imp_res_lanczos(A, v0, nk, nt, np_, maxiters, stop_res, npol, p0, p1,
                    conv_check=1):
    """Returns (V[0..nt), lam[0..nt), iterations)."""
    m = nk + np_
    Afilt = lambda x: cheby(A, x, npol, p0, p1)
    V = [None] * m
    alpha = np.zeros(m)
    beta = np.zeros(m)
    r = lanczos_ext(Afilt, 0, m, v0.copy(), V, alpha, beta)
    it = 1
    while it <= maxiters:
        T = np.diag(alpha) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
        shifts = np.sort(np.linalg.eigvalsh(T))[::-1]
        Q = np.eye(m)
        for i in range(np_):
            alpha, beta, Q = _qr_step(alpha, beta, Q, shifts[nk + i])
        Vm = np.array(V)
        Vn = Q.T @ Vm
        V = [_f32(Vn[k]) for k in range(m)]
        r = r * Q[m - 1, nk - 1]
        if nk < m:
            r = r + beta[nk - 1] * V[nk]
        beta[nk - 1] = np.sqrt(float(np.dot(r, r)))
        if it % conv_check == 0:
            Tk = np.diag(alpha[:nk]) + np.diag(beta[:nk - 1], 1) + np.diag(beta[:nk - 1], -1)
            w, Z = np.linalg.eigh(Tk)
            order = np.argsort(w)[::-1]
            Z = Z[:, order]
            done = 0
            Vk = np.array(V[:nk])
            for k in range(nk):
                v = Z[:, k] @ Vk
                av = A(v)
                bt = np.sqrt(float(np.dot(v, v)))
                alp = float(np.dot(av, v)) / bt
                rn = np.sqrt(float(np.dot(av - alp * v, av - alp * v))) / bt
                if rn < stop_res and k < nt:
                    done += 1
            if done == nt:
                break
        r = lanczos_ext(Afilt, nk, m, r, V, alpha, beta)
        it += 1
    if it == maxiters + 1:
        it -= 1
    Tk = np.diag(alpha[:nk]) + np.diag(beta[:nk - 1], 1) + np.diag(beta[:nk - 1], -1)
    w, Z = np.linalg.eigh(Tk)
    Z = Z[:, np.argsort(w)[::-1]]
    Vk = np.array(V[:nk])
    Vout = [_f32(Z[:, k] @ Vk) for k in range(nt)]
    lam = np.array([float(np.dot(v, A(v))) for v in Vout])
    return Vout, lam, it
