"""CPU restatement of DiracOp::InvEigCg and of QPropW::eig_Run's deflation
space (numpy around the C oracle's MatPcDagMatPc).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

* inveigcg follows cps_pp/src/util/dirac_op/d_op_base/comsrc/inv_eig_cg.C:
  the init-CG guess U invH U^dag src (:202-225), CG on MatPcDagMatPc with
  pAp from the operator's fused norm (:387-392, :493-503), the eigCG window
  (:296-491: T(i,i) = 1/alpha + beta_old/alpha_old, T(i,i-1) = -sqrt(beta)/alpha,
  V[i] = r/|r|; when the window fills, the nev lowest eigenvectors of T(m) and
  of T(m-1) (last row zero), Gram-Schmidt with the 1e-16 rank cut, the Ritz
  pairs of Q^T T Q, V <- V Q Z, T <- diag(M) plus the row <A r, V>/|r|), the
  optional init-CG restarts (:504-559) and the true residual (:590-601).
* EigRun follows cps_pp/src/alg/alg_qpropw/eig_QPropW.C:253-330: invH of the
  leading block of H, the window/restart choice, then the new Ritz vectors
  (value in (1e-30, max_eig), lowest first, single precision) into U and H.

Dense eigenproblems use numpy.linalg.eigh (CPS: tred2 + tqli); the Ritz pairs
do not depend on the solver up to rounding and eigenvector signs.
Vectors are float64 CPS half vectors; <a, b> = sum conj(a) b over the complex
view.
"""
from __future__ import annotations

import numpy as np


def cdot(a, b):
    return np.vdot(a.view(np.complex128), b.view(np.complex128))


def inveigcg(A, src, stop_rsd, max_iter, nev=0, m=0, U=(), invH=None, restart=()):
    """A(x) -> (A x, |MatPc x|^2).  Returns (sol, iters, true_res, V, M)."""
    n = src.size
    src_norm_sq = float(np.dot(src, src))
    stp_cnd = src_norm_sq * stop_rsd * stop_rsd
    def_len = len(U)
    sol = np.zeros(n)
    if def_len:
        ub = np.array([cdot(u, src) for u in U])
        c = invH @ ub
        solc = sol.view(np.complex128)
        for i in range(def_len):
            solc += c[i] * U[i].view(np.complex128)
    mmp, _ = A(sol)
    r = src - mmp
    d = r.copy()
    rr = float(np.dot(r, r))
    V = [np.zeros(n) for _ in range(m)] if nev > 0 else []
    M = np.zeros(2 * nev)
    T = np.zeros((m, m))
    i_eig = 0
    alpha, beta, alpha_old, beta_old = 1.0, 0.0, 0.0, 0.0
    rcond = [x * x * src_norm_sq for x in restart]
    restarted = 0
    mmp_prev = None
    it = 0
    for it in range(max_iter):
        if nev > 0 and i_eig == m:
            mmp_prev = mmp.copy()
        rr_prv = rr
        mmp, pap = A(d)
        if pap == 0.0:
            break
        if nev > 0:
            if i_eig != 0:
                T[i_eig - 1, i_eig - 1] = 1.0 / alpha + beta_old / alpha_old
            if i_eig == m:
                nev2 = 2 * nev
                Q = np.zeros((m, nev2))
                _, y1 = np.linalg.eigh(T[:m - 1, :m - 1])
                Q[:m - 1, nev:] = y1[:, :nev]
                _, y0 = np.linalg.eigh(T)
                Q[:, :nev] = y0[:, :nev]
                rank = nev
                for i in range(nev, nev2):
                    q = Q[:, i].copy()
                    for j in range(rank):
                        q -= np.dot(q, Q[:, j]) * Q[:, j]
                    xx = float(np.dot(q, q))
                    if xx > 1e-16:
                        Q[:, rank] = q / np.sqrt(xx)
                        rank += 1
                Q = Q[:, :rank]
                w_h, Z = np.linalg.eigh(Q.T @ T @ Q)
                M[:] = 0.0
                M[:rank] = w_h
                QZ = Q @ Z
                Vm = np.array(V)
                Vn = QZ.T @ Vm
                for j in range(rank):
                    V[j] = Vn[j].copy()
                T[:] = 0.0
                T[np.arange(rank), np.arange(rank)] = w_h
                i_eig = rank
                w = mmp - beta * mmp_prev
                for ii in range(i_eig):
                    T[i_eig, ii] = T[ii, i_eig] = cdot(w, V[ii]).real / np.sqrt(rr)
            elif i_eig != 0:
                T[i_eig, i_eig - 1] = T[i_eig - 1, i_eig] = -np.sqrt(beta) / alpha
            V[i_eig] = r / np.sqrt(rr)
            i_eig += 1
        alpha_old = alpha
        alpha = -rr_prv / pap
        r = r + alpha * mmp
        rr = float(np.dot(r, r))
        alpha = -alpha
        beta_old = beta
        beta = rr / rr_prv
        sol = sol + alpha * d
        d = r + beta * d
        if restarted < len(rcond) and def_len and rr < rcond[restarted]:
            restarted += 1
            ub = np.array([cdot(u, r) for u in U])
            c = invH @ ub
            e = np.zeros(n)
            ec = e.view(np.complex128)
            for i in range(def_len):
                ec += c[i] * U[i].view(np.complex128)
            sol = sol + e
            ae, _ = A(e)
            r = r - ae
            rr = float(np.dot(r, r))
            d = r.copy()
            mmp = r.copy()
        if rr <= stp_cnd:
            break
    else:
        it = max_iter
    ax, _ = A(sol)
    res = src - ax
    true_res = np.sqrt(float(np.dot(res, res)) / src_norm_sq)
    return sol, it, true_res, V, M


class EigRun:
    """The deflation space of QPropW::eig_Run for a sequence of right-hand sides."""

    def __init__(self, A, nev, m, max_def_len, max_eig):
        self.A, self.nev, self.m = A, nev, m
        self.max_def_len, self.max_eig = max_def_len, max_eig
        self.U = []
        self.H = np.zeros((max_def_len, max_def_len), dtype=np.complex128)
        self.M = np.zeros(2 * nev)

    def solve(self, src, stop_rsd, max_iter, restart=(), always_restart=False):
        dl = len(self.U)
        invH = np.linalg.inv(self.H[:dl, :dl]) if dl else None
        nev, rs = self.nev, list(restart)
        if not always_restart and dl < self.max_def_len:
            rs = []
        elif dl == self.max_def_len:
            nev = 0
        sol, it, tr, V, M = inveigcg(self.A, src, stop_rsd, max_iter, nev,
                                     self.m if nev else 0, self.U, invH, rs)
        if nev:
            self.M = M.copy()
        if nev and dl < self.max_def_len:
            for k in np.argsort(M, kind="stable"):
                if not (1e-30 < M[k] < self.max_eig) or len(self.U) >= self.max_def_len:
                    continue
                v = V[k].astype(np.float32).astype(np.float64)
                au, _ = self.A(v)
                d = len(self.U)
                for j in range(d):
                    self.H[j, d] = cdot(self.U[j], au)
                    self.H[d, j] = np.conj(self.H[j, d])
                self.H[d, d] = float(np.dot(v, au))
                self.U.append(v)
        return sol, it, tr
