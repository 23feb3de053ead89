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
        L.orc_invcg_shift.restype = C.c_int
        L.orc_invcg_shift.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double,
                                      C.POINTER(C.c_double), C.c_double, C.c_int,
                                      C.POINTER(C.c_double), _dp, P]
        _ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.orc_minvcg.restype = C.c_int
        L.orc_minvcg.argtypes = [_dp, _dp, C.c_double, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int,
                                 _ip, _dp, C.c_double, P]
        L.orc_fermion_extract_cb.argtypes = [_dp, _dp, C.c_int * 4, C.c_int]
        L.orc_fermion_insert_cb.argtypes = [_dp, _dp, C.c_int * 4, C.c_int]
        L.orc_fermion5d_canonical_to_full.argtypes = [_dp, _dp, P]
        L.orc_fermion5d_full_to_canonical.argtypes = [_dp, _dp, P]
        L.orc_four2five.argtypes = [_dp, _dp, C.c_int, C.c_int, P]
        L.orc_five2four.argtypes = [_dp, _dp, C.c_int, C.c_int, P]
        L.orc_set_serial_hop.argtypes = [C.c_int]
        L.orc_norm_sq.restype = C.c_double
        L.orc_norm_sq.argtypes = [_dp, C.c_size_t]
        _lib = L
    return _lib


def set_serial_hop(on: bool) -> None:
    """Run the 4D hop in one thread, as CPS's noarch wilson_dslash.C:280-477
    does (the rest keeps its OpenMP loops, as dwf_dslash_5_plus_omp.C and
    vector_util.C do); results are bit-identical either way."""
    lib().orc_set_serial_hop(1 if on else 0)


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

    def invcg_shift(self, src, mass, src_norm_sq=0.0, shift=None, stop_rsd=1e-10,
                    max_iter=10000, guess=None):
        """InvCgShift(out, in, src_norm_sq, true_res, shift); shift None = NULL."""
        sol = np.zeros(self.n) if guess is None else np.array(guess, np.float64)
        tr = C.c_double(0.0)
        hist = np.zeros(max_iter + 2)
        sh = None if shift is None else C.byref(C.c_double(shift))
        it = lib().orc_invcg_shift(sol, np.ascontiguousarray(src, np.float64), self.u, mass,
                                   src_norm_sq, sh, stop_rsd, max_iter, C.byref(tr), hist,
                                   C.byref(self.p))
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

    # -- f2: whole 5D fermion fields (CANONICAL [s][4D site][24]) ----------
    def canonical_to_full(self, canon):
        full = np.zeros(2 * self.n)
        lib().orc_fermion5d_canonical_to_full(full, np.ascontiguousarray(canon, np.float64),
                                              C.byref(self.p))
        return full

    def full_to_canonical(self, full):
        canon = np.zeros(2 * self.n)
        lib().orc_fermion5d_full_to_canonical(canon, np.ascontiguousarray(full, np.float64),
                                              C.byref(self.p))
        return canon

    def four2five(self, four, s_u, s_l):
        five = np.zeros(2 * self.n)
        lib().orc_four2five(five, np.ascontiguousarray(four, np.float64), s_u, s_l,
                            C.byref(self.p))
        return five

    def five2four(self, five, s_u, s_l):
        four = np.zeros(2 * self.n // self.ls)
        lib().orc_five2four(four, np.ascontiguousarray(five, np.float64), s_u, s_l,
                            C.byref(self.p))
        return four

    def kappas(self):
        p = self.p
        if self.action == "dwf":
            return 1.0 / (2 * (4 + p.a5inv - p.m5)), None
        kb = 1.0 / (2 * (p.mob_b * (4 - p.m5) + p.a5inv))
        kc = 1.0 / (2 * (p.mob_c * (4 - p.m5) - p.a5inv))
        return kb, kc


def extract_cb(canon4d, dims, parity):
    """orc_fermion_extract_cb: one CANONICAL 4D slice -> checkerboard half."""
    half = np.zeros(canon4d.size // 2)
    lib().orc_fermion_extract_cb(half, np.ascontiguousarray(canon4d, np.float64),
                                 (C.c_int * 4)(*dims), parity)
    return half


def plaquette_linktrace(links, dims):
    """Lattice::SumReTrPlaq / (18 V) and sum ReTr U / (12 V) of CANONICAL links
    (no boundary signs), as ReadLatticeParallel::CheckPlaqLinktrace forms them
    (ReadLatticePar.C:403, 420-428; lattice_base.C:2736-2778, 2878-2905).
    numpy, FP64."""
    lx, ly, lz, lt = dims
    u = np.asarray(links, np.float64).reshape(lt, lz, ly, lx, 4, 3, 3, 2)
    u = u[..., 0] + 1j * u[..., 1]                 # [t,z,y,x][mu][row][col]
    axis = {0: 3, 1: 2, 2: 1, 3: 0}                # mu -> array axis of that direction
    plaq = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            umu, unu = u[..., mu, :, :], u[..., nu, :, :]
            unu_xpmu = np.roll(unu, -1, axis=axis[mu])  # U_nu(x + mu)
            umu_xpnu = np.roll(umu, -1, axis=axis[nu])  # U_mu(x + nu)
            p = umu @ unu_xpmu @ np.conj(np.swapaxes(umu_xpnu, -1, -2)) @ \
                np.conj(np.swapaxes(unu, -1, -2))
            plaq += np.trace(p, axis1=-2, axis2=-1).real.sum()
    v4 = lx * ly * lz * lt
    tr = np.trace(u, axis1=-2, axis2=-1).real.sum()
    return plaq / (18.0 * v4), tr / (12.0 * v4)
