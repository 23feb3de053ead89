"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
All vectors are float64 numpy arrays in CPS's own orders (see cps_oracle.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("L", C.c_int * 4), ("ls", C.c_int), ("m5", C.c_double),
                ("a5inv", C.c_double), ("mob_b", C.c_double), ("mob_c", C.c_double),
                ("action", C.c_int)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER(Params)
        L.orc_half_size.restype = C.c_size_t
        L.orc_half_size.argtypes = [P]
        L.orc_gauge_canonical_to_wilson.argtypes = [_dp, _dp, C.c_int * 4, C.c_int * 4]
        L.orc_wilson_dslash.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int * 4]
        L.orc_dwf_dslash_5_plus.argtypes = [_dp, _dp, C.c_double, C.c_int, P]
        L.orc_zmobius_kappa_dslash_5_plus.argtypes = [_dp, _dp, C.c_double, C.c_int, _dp, P]
        L.orc_zmobius_m5inv.argtypes = [_dp, C.c_double, C.c_int, _dp, P]
        L.orc_matpcdagmatpc.argtypes = [_dp, _dp, _dp, C.POINTER(C.c_double), C.c_double, P]
        L.orc_matpc.argtypes = [_dp, _dp, _dp, C.c_double, C.c_int, P]
        L.orc_dslash.argtypes = [_dp, _dp, _dp, C.c_double, C.c_int, C.c_int, P]
        L.orc_invcg.restype = C.c_int
        L.orc_invcg.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                C.POINTER(C.c_double), _dp, P]
        L.orc_matinv.restype = C.c_int
        L.orc_matinv.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                 C.POINTER(C.c_double), P]
        _ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.orc_minvcg.restype = C.c_int
        L.orc_minvcg.argtypes = [_dp, _dp, C.c_double, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int,
                                 _ip, _dp, C.c_double, P]
        L.orc_norm_sq.restype = C.c_double
        L.orc_norm_sq.argtypes = [_dp, C.c_size_t]
        _lib = L
    return _lib


class Oracle:
    """One lattice + action; mirrors the DiracOpDwf/DiracOpMobius surface."""

    def __init__(self, dims, ls, action="dwf", m5=1.8, a5inv=1.0, b=1.5, c=0.5,
                 gauge_canonical=None, bc=(0, 0, 0, 1)):
        self.dims = tuple(int(d) for d in dims)
        self.ls = int(ls)
        self.p = Params((C.c_int * 4)(*self.dims), self.ls, m5, a5inv, b, c,
                        0 if action == "dwf" else 1)
        self.action = action
        self.n = int(lib().orc_half_size(C.byref(self.p)))
        v4 = int(np.prod(self.dims))
        self.u = np.zeros(v4 * 72)
        if gauge_canonical is not None:
            g = np.ascontiguousarray(gauge_canonical).view(np.float64).reshape(-1)
            lib().orc_gauge_canonical_to_wilson(self.u, np.ascontiguousarray(g),
                                                (C.c_int * 4)(*self.dims),
                                                (C.c_int * 4)(*bc))

    # -- operators on CPS half vectors (float64, n reals) -------------------
    def matpcdagmatpc(self, x, mass):
        out = np.zeros(self.n)
        d = C.c_double(0.0)
        lib().orc_matpcdagmatpc(out, self.u, np.ascontiguousarray(x, np.float64),
                                C.byref(d), mass, C.byref(self.p))
        return out, d.value

    def matpc(self, x, mass, dag=0):
        out = np.zeros(self.n)
        lib().orc_matpc(out, self.u, np.ascontiguousarray(x, np.float64), mass, dag,
                        C.byref(self.p))
        return out

    def dslash(self, x, mass, cb, dag=0):
        out = np.zeros(self.n)
        lib().orc_dslash(out, self.u, np.ascontiguousarray(x, np.float64), mass, cb,
                         dag, C.byref(self.p))
        return out

    def wilson_dslash(self, psi_half, cb, dag=0):
        out = np.zeros_like(psi_half, dtype=np.float64)
        lib().orc_wilson_dslash(out, self.u, np.ascontiguousarray(psi_half, np.float64),
                                cb, dag, (C.c_int * 4)(*self.dims))
        return out

    def dwf_dslash_5_plus(self, out, x, mass, dag=0):
        out = np.ascontiguousarray(out, np.float64).copy()
        lib().orc_dwf_dslash_5_plus(out, np.ascontiguousarray(x, np.float64), mass,
                                    dag, C.byref(self.p))
        return out

    def kappa_dslash_5_plus(self, out, x, mass, dag, fact):
        out = np.ascontiguousarray(out, np.float64).copy()
        f = np.ascontiguousarray(np.asarray(fact, np.complex128)).view(np.float64)
        lib().orc_zmobius_kappa_dslash_5_plus(out, np.ascontiguousarray(x, np.float64),
                                              mass, dag, f, C.byref(self.p))
        return out

    def m5inv(self, x, mass, dag, K):
        out = np.ascontiguousarray(x, np.float64).copy()
        k = np.ascontiguousarray(np.asarray(K, np.complex128)).view(np.float64)
        lib().orc_zmobius_m5inv(out, mass, dag, k, C.byref(self.p))
        return out

    def invcg(self, src, mass, stop_rsd=1e-10, max_iter=10000, guess=None):
        sol = np.zeros(self.n) if guess is None else np.array(guess, np.float64)
        tr = C.c_double(0.0)
        hist = np.zeros(max_iter + 2)
        it = lib().orc_invcg(sol, np.ascontiguousarray(src, np.float64), self.u, mass,
                             stop_rsd, max_iter, C.byref(tr), hist, C.byref(self.p))
        return sol, it, tr.value, hist[: max(it, 1)]

    def matinv(self, src_full, mass, stop_rsd=1e-10, max_iter=10000, guess=None):
        out = np.zeros(2 * self.n) if guess is None else np.array(guess, np.float64)
        src = np.array(src_full, np.float64)
        tr = C.c_double(0.0)
        it = lib().orc_matinv(out, src, self.u, mass, stop_rsd, max_iter,
                              C.byref(tr), C.byref(self.p))
        return out, it, tr.value

    def minvcg(self, chi, mass, shifts, rsd, multi=True, alpha=None, max_iter=10000):
        """DiracOp::MInvCG: returns (psi [nshift, n] or [n], k, |r_s|, iters_s)."""
        shifts = np.ascontiguousarray(shifts, np.float64)
        ns = shifts.size
        rs = np.ascontiguousarray(np.broadcast_to(rsd, (ns,)), np.float64).copy()
        al = np.ones(ns) if alpha is None else np.ascontiguousarray(alpha, np.float64)
        psi = np.zeros((ns if multi else 1) * self.n)
        its = np.zeros(ns, dtype=np.int32)
        chi = np.ascontiguousarray(chi, np.float64)
        chi_norm = float(lib().orc_norm_sq(chi, chi.size))
        k = lib().orc_minvcg(psi, chi, chi_norm, shifts, ns, rs, 1 if multi else 0, al,
                             max_iter, its, self.u, mass, C.byref(self.p))
        return psi.reshape(ns, self.n) if multi else psi, k, rs, its

    def kappas(self):
        p = self.p
        if self.action == "dwf":
            return 1.0 / (2 * (4 + p.a5inv - p.m5)), None
        kb = 1.0 / (2 * (p.mob_b * (4 - p.m5) + p.a5inv))
        kc = 1.0 / (2 * (p.mob_c * (4 - p.m5) - p.a5inv))
        return kb, kc
