"""ctypes binding of libcps_dwf_gpu.so with the DiracOpDwf / DiracOpMobius
call surface of CPS (cps_pp/include/util/dirac_op.h):

    op = DiracOpDwfGpu(gauge, dims, ls, m5=1.8)       # DiracOpDwf ctor
    dot = op.MatPcDagMatPc(out, inp, mass)             # dirac_op.h:259
    op.MatPc(out, inp, mass); op.MatPcDag(out, inp, mass)
    op.Dslash(out, inp, mass, cb, dag)
    iters, true_res = op.InvCg(out, inp, CgArg(mass, ...))   # dirac_op.h:138
    iters, true_res = op.MatInv(out, inp, CgArg(...))
    iters, true_res, st, nrel = op.InvCgMixed_dev(sol, src, CgArg(...), "bf16")

Vectors are float64 numpy arrays in CPS half-vector order.  The HIP library
must be built (cps_veloc_amd/build.py); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CPS_DWF_GPU_LIB selects another build of the same library (A/B timing)
LIB_PATH = os.environ.get("CPS_DWF_GPU_LIB", os.path.join(_HERE, "libcps_dwf_gpu.so"))

OK, BREAKDOWN, MAXITER, BAD_ARG, HIP_ERROR, IO_ERROR, CHECKSUM_ERROR = 0, 1, 2, 3, 4, 5, 6
# bytes of the static device array behind dwfgpu_matpcdagmatpc_dev_noalloc
NOALLOC_SCRATCH_BYTES = 256 << 20
DWF, MOBIUS = 0, 1
F64, F32, BF16, I16 = 0, 1, 2, 3
_PREC = {"f64": F64, "f32": F32, "bf16": BF16, "i16": I16}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None

# every entry point of include/dwf_gpu.h
EXPORTS = (
    "dwfgpu_create", "dwfgpu_destroy", "dwfgpu_half_size", "dwfgpu_last_error",
    "dwfgpu_matpcdagmatpc", "dwfgpu_matpc", "dwfgpu_dslash", "dwfgpu_invcg",
    "dwfgpu_invcg_shift", "dwfgpu_matinv", "dwfgpu_vec_create", "dwfgpu_vec_destroy",
    "dwfgpu_vec_upload", "dwfgpu_vec_download", "dwfgpu_matpcdagmatpc_dev", "dwfgpu_invcg_dev",
    "dwfgpu_matpcdagmatpc_dev_noalloc",
    "dwfgpu_invcg_shift_dev",
    "dwfgpu_invcg_mixed", "dwfgpu_invcg_mixed_dev",
    "dwfgpu_mvec_create", "dwfgpu_mvec_destroy", "dwfgpu_mvec_upload", "dwfgpu_mvec_download",
    "dwfgpu_matpcdagmatpc_mrhs_dev", "dwfgpu_invcg_mrhs_dev", "dwfgpu_mrhs_shape",
    "dwfgpu_matinv_mrhs_dev",
    "dwfgpu_ptc_num_weights", "dwfgpu_ptc_init", "dwfgpu_ptc_set_weights",
    "dwfgpu_ptc_get_weights", "dwfgpu_ptc_apply", "dwfgpu_ptc_loss_grad", "dwfgpu_ptc_train",
    "dwfgpu_invcg_ptc_dev", "dwfgpu_invgcr_ptc_dev", "dwfgpu_minvcg", "dwfgpu_minvcg_dev",
    "dwfgpu_time_matpcdagmatpc", "dwfgpu_matpcdagmatpc_cost", "dwfgpu_time_copy",
    "dwfgpu_time_bandwidth",
    "dwfgpu_fermion_from_canonical", "dwfgpu_fermion_to_canonical", "dwfgpu_four2five",
    "dwfgpu_five2four", "dwfgpu_matinv_dev", "dwfgpu_nersc_read", "dwfgpu_plaquette",
    "dwfgpu_inveigcg_dev", "dwfgpu_eigcg_create", "dwfgpu_eigcg_solve", "dwfgpu_eigcg_def_len",
    "dwfgpu_eigcg_values", "dwfgpu_eigcg_destroy", "dwfgpu_lanczos",
)
# every entry point of include/dwf_synth.h
SYNTH_EXPORTS = ("dwfsynth_gauge", "dwfsynth_gaussian")


class DwfGpuError(RuntimeError):
    def __init__(self, code, where):
        msg = lib().dwfgpu_last_error().decode(errors="replace")
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


class PtcTrainArgC(C.Structure):
    _fields_ = [("mass", C.c_double), ("nsolves", C.c_int), ("pairs_per_solve", C.c_int),
                ("random_pairs", C.c_int), ("stride", C.c_int), ("epochs", C.c_int),
                ("batch", C.c_int),
                ("lr", C.c_double), ("seed", C.c_ulonglong)]


class NerscHeaderC(C.Structure):
    _fields_ = [("dims", C.c_int * 4), ("recon_row_3", C.c_int), ("fp_bytes", C.c_int),
                ("big_endian", C.c_int), ("checksum", C.c_uint), ("checksum_data", C.c_uint),
                ("plaquette", C.c_double), ("link_trace", C.c_double),
                ("data_start", C.c_longlong)]


class LanczosArgC(C.Structure):
    _fields_ = [("mass", C.c_double), ("nk", C.c_int), ("nt", C.c_int), ("np", C.c_int),
                ("maxiters", C.c_int), ("stop_residual", C.c_double), ("npol", C.c_int),
                ("p0", C.c_double), ("p1", C.c_double), ("shift", C.c_double),
                ("conv_check", C.c_int)]


class CgArgC(C.Structure):
    _fields_ = [("mass", C.c_double), ("max_num_iter", C.c_int),
                ("stop_rsd", C.c_double), ("true_rsd", C.c_double)]


@dataclass
class CgArg:
    """Mirror of CPS's CgArg (cps_pp/include/input/cg_arg.x:3-28)."""
    mass: float
    max_num_iter: int = 10000
    stop_rsd: float = 1e-10
    true_rsd: float = 1e-10

    def c(self):
        return CgArgC(self.mass, self.max_num_iter, self.stop_rsd, self.true_rsd)


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run cps_veloc_amd/build.py "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        H = C.c_void_p
        L.dwfgpu_last_error.restype = C.c_char_p
        L.dwfgpu_create.argtypes = [_dp, C.c_int * 4, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_int * 4, C.c_int,
                                    C.POINTER(H)]
        L.dwfgpu_destroy.argtypes = [H]
        L.dwfgpu_half_size.argtypes = [H]
        L.dwfgpu_half_size.restype = C.c_size_t
        L.dwfgpu_matpcdagmatpc.argtypes = [H, C.c_double, _dp, _dp, C.POINTER(C.c_double)]
        L.dwfgpu_matpc.argtypes = [H, C.c_double, _dp, _dp, C.c_int]
        L.dwfgpu_dslash.argtypes = [H, C.c_double, _dp, _dp, C.c_int, C.c_int]
        L.dwfgpu_invcg.argtypes = [H, C.POINTER(CgArgC), _dp, _dp, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int)]
        L.dwfgpu_invcg_shift.argtypes = [H, C.POINTER(CgArgC), _dp, _dp, C.c_double,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_int)]
        L.dwfgpu_invcg_shift_dev.argtypes = [H, C.POINTER(CgArgC), H, H, C.c_double,
                                             C.POINTER(C.c_double), C.POINTER(C.c_double),
                                             C.POINTER(C.c_int)]
        L.dwfgpu_matinv.argtypes = [H, C.POINTER(CgArgC), _dp, _dp, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int)]
        L.dwfgpu_vec_create.argtypes = [H, C.POINTER(H)]
        L.dwfgpu_vec_destroy.argtypes = [H]
        L.dwfgpu_vec_upload.argtypes = [H, _dp]
        L.dwfgpu_vec_download.argtypes = [H, _dp]
        L.dwfgpu_matpcdagmatpc_dev.argtypes = [H, C.c_double, H, H, C.POINTER(C.c_double)]
        L.dwfgpu_matpcdagmatpc_dev_noalloc.argtypes = [H, C.c_double, H, H,
                                                       C.POINTER(C.c_double)]
        L.dwfgpu_invcg_dev.argtypes = [H, C.POINTER(CgArgC), H, H, C.POINTER(C.c_double),
                                       C.POINTER(C.c_int)]
        L.dwfgpu_invcg_mixed.argtypes = [H, C.POINTER(CgArgC), C.c_int, C.c_double, _dp, _dp,
                                         C.POINTER(C.c_double), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
        L.dwfgpu_invcg_mixed_dev.argtypes = [H, C.POINTER(CgArgC), C.c_int, C.c_double, H, H,
                                             C.POINTER(C.c_double), C.POINTER(C.c_int),
                                             C.POINTER(C.c_int)]
        L.dwfgpu_mvec_create.argtypes = [H, C.c_int, C.POINTER(H)]
        L.dwfgpu_mvec_destroy.argtypes = [H]
        L.dwfgpu_mvec_upload.argtypes = [H, C.c_int, _dp]
        L.dwfgpu_mvec_download.argtypes = [H, C.c_int, _dp]
        L.dwfgpu_matpcdagmatpc_mrhs_dev.argtypes = [H, C.c_int, _dp, H, H]
        _ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.dwfgpu_invcg_mrhs_dev.argtypes = [H, C.POINTER(CgArgC), C.c_int, C.c_double, C.c_int,
                                            _dp, H, H, _dp, _ip, _ip, _ip]
        L.dwfgpu_mrhs_shape.argtypes = [H, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.dwfgpu_matinv_mrhs_dev.argtypes = [H, C.POINTER(CgArgC), C.c_int, C.c_double, C.c_int,
                                             _dp, H, H, H, H, _dp, _ip, _ip, _ip]
        _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.dwfgpu_ptc_num_weights.argtypes = [H]
        L.dwfgpu_ptc_num_weights.restype = C.c_size_t
        L.dwfgpu_ptc_init.argtypes = [H, C.c_ulonglong, C.c_double]
        L.dwfgpu_ptc_set_weights.argtypes = [H, _fp, C.c_size_t]
        L.dwfgpu_ptc_get_weights.argtypes = [H, _fp, C.c_size_t]
        L.dwfgpu_ptc_apply.argtypes = [H, _dp, _dp]
        L.dwfgpu_ptc_loss_grad.argtypes = [H, _dp, _dp, C.POINTER(C.c_double), _fp]
        L.dwfgpu_ptc_train.argtypes = [H, C.POINTER(PtcTrainArgC), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.dwfgpu_minvcg.argtypes = [H, C.c_double, _dp, C.c_int, _dp, _dp, C.c_double, _dp,
                                    C.c_int, C.c_void_p, C.c_int, _ip, C.POINTER(C.c_int)]
        L.dwfgpu_minvcg_dev.argtypes = [H, C.c_double, _dp, C.c_int, H, H, C.c_double, _dp,
                                        C.c_int, C.c_void_p, C.c_int, _ip, C.POINTER(C.c_int)]
        L.dwfgpu_invcg_ptc_dev.argtypes = [H, C.POINTER(CgArgC), C.c_int, C.c_double, H, H,
                                           C.POINTER(C.c_double), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.dwfgpu_invgcr_ptc_dev.argtypes = [H, C.POINTER(CgArgC), C.c_int, H, H,
                                            C.POINTER(C.c_double), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.dwfgpu_time_matpcdagmatpc.argtypes = [H, C.c_int, C.c_double, H, H, C.c_int,
                                                C.POINTER(C.c_double), _dp, _dp,
                                                C.POINTER(C.c_int)]
        L.dwfgpu_matpcdagmatpc_cost.argtypes = [H, C.c_int, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double)]
        L.dwfgpu_time_copy.argtypes = [H, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.dwfgpu_time_bandwidth.argtypes = [H, C.c_size_t, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.dwfgpu_fermion_from_canonical.argtypes = [H, H, H, _dp]
        L.dwfgpu_fermion_to_canonical.argtypes = [H, _dp, H, H]
        L.dwfgpu_four2five.argtypes = [H, H, H, _dp, C.c_int, C.c_int]
        L.dwfgpu_five2four.argtypes = [H, _dp, H, H, C.c_int, C.c_int]
        L.dwfgpu_matinv_dev.argtypes = [H, C.POINTER(CgArgC), H, H, H, H,
                                        C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.dwfgpu_nersc_read.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t,
                                        C.POINTER(NerscHeaderC)]
        L.dwfgpu_plaquette.argtypes = [H, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.dwfgpu_inveigcg_dev.argtypes = [H, C.POINTER(CgArgC), H, H, C.c_int, C.c_int, H,
                                          C.c_void_p, H, C.c_void_p, C.c_int, C.c_void_p,
                                          C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.dwfgpu_eigcg_create.argtypes = [H, C.c_int, C.c_int, C.c_int, C.c_double,
                                          C.POINTER(H)]
        L.dwfgpu_eigcg_solve.argtypes = [H, C.POINTER(CgArgC), H, H, C.c_void_p, C.c_int,
                                         C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.dwfgpu_eigcg_def_len.argtypes = [H]
        L.dwfgpu_eigcg_values.argtypes = [H, C.c_void_p, C.c_void_p]
        L.dwfgpu_eigcg_destroy.argtypes = [H]
        L.dwfgpu_lanczos.argtypes = [H, C.POINTER(LanczosArgC), _dp, C.c_void_p, _dp,
                                     C.POINTER(C.c_int)]
        L.dwfsynth_gauge.argtypes = [_dp, C.c_int * 4, C.c_double, C.c_ulonglong]
        L.dwfsynth_gauge.restype = None
        L.dwfsynth_gaussian.argtypes = [_dp, C.c_size_t, C.c_ulonglong]
        L.dwfsynth_gaussian.restype = None
        _lib = L
    return _lib


def _chk(code, where, allow=(OK,)):
    if code not in allow:
        raise DwfGpuError(code, where)
    return code


def _arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class DeviceVector:
    """A CG half vector resident in HBM (dwfgpu_vec_t)."""

    def __init__(self, op: "_DiracOpGpu", host=None):
        self.op = op
        self.h = C.c_void_p()
        _chk(lib().dwfgpu_vec_create(op.h, C.byref(self.h)), "dwfgpu_vec_create")
        if host is not None:
            self.upload(host)

    def upload(self, host):
        _chk(lib().dwfgpu_vec_upload(self.h, _arr(host)), "dwfgpu_vec_upload")

    def download(self):
        out = np.empty(self.op.n)
        _chk(lib().dwfgpu_vec_download(self.h, out), "dwfgpu_vec_download")
        return out

    def free(self):
        if self.h:
            lib().dwfgpu_vec_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MultiVector:
    """nrhs CG half vectors resident in HBM (dwfgpu_mvec_t, FP64 handle)."""

    def __init__(self, op: "_DiracOpGpu", nrhs, hosts=None):
        self.op = op
        self.nrhs = int(nrhs)
        self.h = C.c_void_p()
        _chk(lib().dwfgpu_mvec_create(op.h, self.nrhs, C.byref(self.h)), "dwfgpu_mvec_create")
        if hosts is not None:
            for j, x in enumerate(hosts):
                self.upload(j, x)

    def upload(self, j, host):
        _chk(lib().dwfgpu_mvec_upload(self.h, j, _arr(host)), "dwfgpu_mvec_upload")

    def download(self, j):
        out = np.empty(self.op.n)
        _chk(lib().dwfgpu_mvec_download(self.h, j, out), "dwfgpu_mvec_download")
        return out

    def free(self):
        if self.h:
            lib().dwfgpu_mvec_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _DiracOpGpu:
    action = DWF

    def __init__(self, gauge_canonical, dims, ls, m5=1.8, a5inv=1.0, b=1.0, c=0.0,
                 bc=(0, 0, 0, 1), precision="f64"):
        g = np.ascontiguousarray(gauge_canonical)
        g = g.view(np.float64).reshape(-1) if np.iscomplexobj(g) else g.reshape(-1)
        self.dims = tuple(int(d) for d in dims)
        self.ls = int(ls)
        self.precision = precision
        self.h = C.c_void_p()
        _chk(lib().dwfgpu_create(_arr(g), (C.c_int * 4)(*self.dims), self.ls, self.action,
                                 m5, a5inv, b, c, (C.c_int * 4)(*bc), _PREC[precision],
                                 C.byref(self.h)), "dwfgpu_create")
        self.n = int(lib().dwfgpu_half_size(self.h))

    def close(self):
        if self.h:
            lib().dwfgpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- DiracOp surface on host arrays -----------------------------------
    def MatPcDagMatPc(self, out, inp, mass):
        d = C.c_double()
        _chk(lib().dwfgpu_matpcdagmatpc(self.h, mass, out, _arr(inp), C.byref(d)),
             "MatPcDagMatPc")
        return d.value

    def MatPc(self, out, inp, mass):
        _chk(lib().dwfgpu_matpc(self.h, mass, out, _arr(inp), 0), "MatPc")

    def MatPcDag(self, out, inp, mass):
        _chk(lib().dwfgpu_matpc(self.h, mass, out, _arr(inp), 1), "MatPcDag")

    def Dslash(self, out, inp, mass, cb, dag=0):
        _chk(lib().dwfgpu_dslash(self.h, mass, out, _arr(inp), cb, dag), "Dslash")

    def InvCg(self, out, inp, cg_arg: CgArg):
        """out: initial guess in, solution out.  Returns (iters, true_res, status)."""
        tr, it = C.c_double(), C.c_int()
        a = cg_arg.c()
        st = _chk(lib().dwfgpu_invcg(self.h, C.byref(a), out, _arr(inp), C.byref(tr),
                                     C.byref(it)), "InvCg", allow=(OK, MAXITER))
        return it.value, tr.value, st

    def InvCgShift(self, out, inp, cg_arg: CgArg, src_norm_sq=0.0, shift=None):
        """DiracOp::InvCgShift(out, in, src_norm_sq, true_res, shift); shift None
        is CPS's NULL.  Returns (iters, true_res, status)."""
        tr, it = C.c_double(), C.c_int()
        a = cg_arg.c()
        sh = None if shift is None else C.byref(C.c_double(shift))
        st = _chk(lib().dwfgpu_invcg_shift(self.h, C.byref(a), out, _arr(inp), src_norm_sq,
                                           C.byref(tr), sh, C.byref(it)), "InvCgShift",
                  allow=(OK, MAXITER))
        return it.value, tr.value, st

    def InvCgShift_dev(self, sol: DeviceVector, src: DeviceVector, cg_arg: CgArg,
                       src_norm_sq=0.0, shift=None):
        tr, it = C.c_double(), C.c_int()
        a = cg_arg.c()
        sh = None if shift is None else C.byref(C.c_double(shift))
        st = _chk(lib().dwfgpu_invcg_shift_dev(self.h, C.byref(a), sol.h, src.h, src_norm_sq,
                                               C.byref(tr), sh, C.byref(it)), "InvCgShift_dev",
                  allow=(OK, MAXITER))
        return it.value, tr.value, st

    def MatInv(self, out, inp, cg_arg: CgArg):
        tr, it = C.c_double(), C.c_int()
        a = cg_arg.c()
        st = _chk(lib().dwfgpu_matinv(self.h, C.byref(a), out, _arr(inp), C.byref(tr),
                                      C.byref(it)), "MatInv", allow=(OK, MAXITER))
        return it.value, tr.value, st

    # --- device-resident fast path ----------------------------------------
    def vector(self, host=None):
        return DeviceVector(self, host)

    def MatPcDagMatPc_dev(self, out: DeviceVector, inp: DeviceVector, mass):
        d = C.c_double()
        _chk(lib().dwfgpu_matpcdagmatpc_dev(self.h, mass, out.h, inp.h, C.byref(d)),
             "MatPcDagMatPc_dev")
        return d.value

    def MatPcDagMatPc_dev_noalloc(self, out: DeviceVector, inp: DeviceVector, mass):
        """MatPcDagMatPc_dev with its temporaries in the library's static device
        array (no device allocation; lattices up to NOALLOC_SCRATCH_BYTES)."""
        d = C.c_double()
        _chk(lib().dwfgpu_matpcdagmatpc_dev_noalloc(self.h, mass, out.h, inp.h, C.byref(d)),
             "MatPcDagMatPc_dev_noalloc")
        return d.value

    def InvCg_dev(self, sol: DeviceVector, src: DeviceVector, cg_arg: CgArg):
        tr, it = C.c_double(), C.c_int()
        a = cg_arg.c()
        st = _chk(lib().dwfgpu_invcg_dev(self.h, C.byref(a), sol.h, src.h, C.byref(tr),
                                         C.byref(it)), "InvCg_dev", allow=(OK, MAXITER))
        return it.value, tr.value, st

    def InvCgMixed(self, out, inp, cg_arg: CgArg, sloppy="f32", delta=0.1):
        """FP64 reliable-update CG with a `sloppy` ("f64", "f32", "bf16", "i16")
        operator.  out: initial guess in, solution out.
        Returns (iters, true_res, status, reliable_updates)."""
        tr, it, nr = C.c_double(), C.c_int(), C.c_int()
        a = cg_arg.c()
        st = _chk(lib().dwfgpu_invcg_mixed(self.h, C.byref(a), _PREC[sloppy], delta, out,
                                           _arr(inp), C.byref(tr), C.byref(it), C.byref(nr)),
                  "InvCgMixed", allow=(OK, MAXITER))
        return it.value, tr.value, st, nr.value

    def InvCgMixed_dev(self, sol: DeviceVector, src: DeviceVector, cg_arg: CgArg,
                       sloppy="f32", delta=0.1):
        tr, it, nr = C.c_double(), C.c_int(), C.c_int()
        a = cg_arg.c()
        st = _chk(lib().dwfgpu_invcg_mixed_dev(self.h, C.byref(a), _PREC[sloppy], delta, sol.h,
                                               src.h, C.byref(tr), C.byref(it), C.byref(nr)),
                  "InvCgMixed_dev", allow=(OK, MAXITER))
        return it.value, tr.value, st, nr.value

    # --- many right-hand sides (DWF) ----------------------------------------
    def multi_vector(self, nrhs, hosts=None):
        return MultiVector(self, nrhs, hosts)

    def MatPcDagMatPc_mrhs_dev(self, out: MultiVector, inp: MultiVector, masses,
                               precision="f64"):
        m = _arr(masses)
        _chk(lib().dwfgpu_matpcdagmatpc_mrhs_dev(self.h, _PREC[precision], m, out.h, inp.h),
             "MatPcDagMatPc_mrhs_dev")

    def InvCgMrhs_dev(self, sol: MultiVector, src: MultiVector, cg_arg: CgArg, masses,
                      sloppy="f32", delta=0.1, eager=False):
        """Batched reliable-update CG; returns per-RHS arrays
        (iters, true_res, status, reliable_updates)."""
        n = sol.nrhs
        tr = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        nr = np.zeros(n, dtype=np.int32)
        stv = np.zeros(n, dtype=np.int32)
        a = cg_arg.c()
        _chk(lib().dwfgpu_invcg_mrhs_dev(self.h, C.byref(a), _PREC[sloppy], delta,
                                         1 if eager else 0, _arr(masses), sol.h, src.h, tr, it,
                                         nr, stv), "InvCgMrhs_dev")
        return it, tr, stv, nr

    # --- multi-shift CG (DiracOp::MInvCG) -----------------------------------
    def MInvCG(self, chi, mass, shifts, rsd, multi=True, alpha=None, max_iter=10000,
               psi=None):
        """Returns (psi [nshift, n] (MULTI) or [n] (SINGLE), k, |r_s|, iters_s, status)."""
        shifts = np.ascontiguousarray(shifts, np.float64)
        ns = shifts.size
        rs = np.ascontiguousarray(np.broadcast_to(rsd, (ns,)), np.float64).copy()
        al = None if alpha is None else np.ascontiguousarray(alpha, np.float64)
        if psi is None:
            psi = np.zeros((ns if multi else 1) * self.n)
        # the result goes into a copy: the caller's starting vector is kept
        psi = np.array(psi, dtype=np.float64, copy=True).reshape(-1)
        its = np.zeros(ns, dtype=np.int32)
        k = C.c_int()
        st = _chk(lib().dwfgpu_minvcg(self.h, mass, shifts, ns, psi, _arr(chi), -1.0, rs,
                                      1 if multi else 0,
                                      None if al is None else al.ctypes.data_as(C.c_void_p),
                                      max_iter, its, C.byref(k)), "MInvCG",
                  allow=(OK, MAXITER))
        return (psi.reshape(ns, self.n) if multi else psi), k.value, rs, its, st

    # --- learned preconditioner -------------------------------------------
    def ptc_init(self, seed=1, scale=0.02):
        _chk(lib().dwfgpu_ptc_init(self.h, seed, scale), "ptc_init")

    def ptc_num_weights(self):
        return int(lib().dwfgpu_ptc_num_weights(self.h))

    def ptc_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.complex64)
        f = w.view(np.float32)
        _chk(lib().dwfgpu_ptc_set_weights(self.h, f, w.size), "ptc_set_weights")

    def ptc_get_weights(self):
        f = np.zeros(2 * self.ptc_num_weights(), dtype=np.float32)
        _chk(lib().dwfgpu_ptc_get_weights(self.h, f, f.size // 2), "ptc_get_weights")
        return f.view(np.complex64)

    def ptc_apply(self, r):
        out = np.zeros(self.n)
        _chk(lib().dwfgpu_ptc_apply(self.h, out, _arr(r)), "ptc_apply")
        return out

    def ptc_loss_grad(self, r, e):
        loss = C.c_double()
        g = np.zeros(2 * self.ptc_num_weights(), dtype=np.float32)
        _chk(lib().dwfgpu_ptc_loss_grad(self.h, _arr(r), _arr(e), C.byref(loss), g),
             "ptc_loss_grad")
        return loss.value, g.view(np.complex64)

    def ptc_train(self, mass, nsolves=2, pairs_per_solve=16, random_pairs=16, stride=4,
                  epochs=4, batch=4, lr=1e-3, seed=7):
        a = PtcTrainArgC(mass, nsolves, pairs_per_solve, random_pairs, stride, epochs, batch, lr,
                         seed)
        l0, l1, n = C.c_double(), C.c_double(), C.c_int()
        _chk(lib().dwfgpu_ptc_train(self.h, C.byref(a), C.byref(l0), C.byref(l1), C.byref(n)),
             "ptc_train")
        return l0.value, l1.value, n.value

    def InvCgPtc_dev(self, sol: DeviceVector, src: DeviceVector, cg_arg: CgArg, sloppy="f32",
                     delta=0.1):
        """Flexible CG with the network as preconditioner; returns
        (iters, true_res, status, reliable_updates, net_steps)."""
        tr, it, nr, ns = C.c_double(), C.c_int(), C.c_int(), C.c_int()
        a = cg_arg.c()
        st = _chk(lib().dwfgpu_invcg_ptc_dev(self.h, C.byref(a), _PREC[sloppy], delta, sol.h,
                                             src.h, C.byref(tr), C.byref(it), C.byref(nr),
                                             C.byref(ns)), "InvCgPtc_dev", allow=(OK, MAXITER))
        return it.value, tr.value, st, nr.value, ns.value

    def InvGcrPtc_dev(self, sol: DeviceVector, src: DeviceVector, cg_arg: CgArg, nkrylov=8):
        """Flexible GCR(nkrylov) with the network as preconditioner (plain CG
        finishes if the network stops helping); returns
        (iters, true_res, status, net_steps, restarts)."""
        tr, it, ns, nc = C.c_double(), C.c_int(), C.c_int(), C.c_int()
        a = cg_arg.c()
        st = _chk(lib().dwfgpu_invgcr_ptc_dev(self.h, C.byref(a), nkrylov, sol.h, src.h,
                                              C.byref(tr), C.byref(it), C.byref(ns),
                                              C.byref(nc)), "InvGcrPtc_dev", allow=(OK, MAXITER))
        return it.value, tr.value, st, ns.value, nc.value

    def MatInvMrhs_dev(self, out_first: MultiVector, out_second: MultiVector,
                       in_first: MultiVector, in_second: MultiVector, cg_arg: CgArg, masses,
                       sloppy="f32", delta=0.1, eager=False):
        """DiracOpMobius::MatInv(vector) (and DWF): full vectors [first | second]
        of every RHS, one batched CG; returns per-RHS arrays
        (iters, true_res, status, reliable_updates)."""
        n = in_first.nrhs
        tr = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        nr = np.zeros(n, dtype=np.int32)
        stv = np.zeros(n, dtype=np.int32)
        a = cg_arg.c()
        _chk(lib().dwfgpu_matinv_mrhs_dev(self.h, C.byref(a), _PREC[sloppy], delta,
                                          1 if eager else 0, _arr(masses), out_first.h,
                                          out_second.h, in_first.h, in_second.h, tr, it, nr,
                                          stv), "MatInvMrhs_dev")
        return it, tr, stv, nr

    def mrhs_shape(self):
        ns, nr = C.c_int(), C.c_int()
        _chk(lib().dwfgpu_mrhs_shape(self.h, C.byref(ns), C.byref(nr)), "mrhs_shape")
        return ns.value, nr.value

    def time_MatPcDagMatPc(self, out: DeviceVector, inp: DeviceVector, mass, reps,
                           precision=None):
        """Per-kernel timing of `reps` applications; precision defaults to the
        handle's own, else a sloppy precision of an FP64 handle."""
        ms = C.c_double()
        msk = np.zeros(8)
        byk = np.zeros(8)
        nk = C.c_int()
        p = _PREC[precision or self.precision]
        _chk(lib().dwfgpu_time_matpcdagmatpc(self.h, p, mass, out.h, inp.h, reps, C.byref(ms),
                                             msk, byk, C.byref(nk)), "time_MatPcDagMatPc")
        return ms.value, msk[:nk.value], byk[:nk.value]

    def time_copy(self, nbytes, reps=10):
        g = C.c_double()
        _chk(lib().dwfgpu_time_copy(self.h, int(nbytes), reps, C.byref(g)), "time_copy")
        return g.value

    def time_bandwidth(self, nbytes, reps=10):
        """(copy, read-only, three-read-one-write) GB/s of the device's probes."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _chk(lib().dwfgpu_time_bandwidth(self.h, int(nbytes), reps, C.byref(a), C.byref(b),
                                         C.byref(c)), "time_bandwidth")
        return a.value, b.value, c.value

    def cost(self, precision=None):
        b, f = C.c_double(), C.c_double()
        p = _PREC[precision or self.precision]
        _chk(lib().dwfgpu_matpcdagmatpc_cost(self.h, p, C.byref(b), C.byref(f)), "cost")
        return b.value, f.value

    # --- host formats (convert.C:639-831, f_dwf_base.C:961-1168) ----------
    def _v4(self):
        return int(np.prod(self.dims))

    def fermion_from_canonical(self, canon):
        """CANONICAL 5D fermion [s][x,y,z,t][24] -> (first, second) device halves
        of the MatInv full vector (Fconvert CANONICAL -> WILSON / EOPREC_EE)."""
        canon = _arr(canon).reshape(-1)
        if canon.size != self.ls * self._v4() * 24:
            raise ValueError("canonical field must hold Ls*V4*24 reals")
        a, b = self.vector(), self.vector()
        _chk(lib().dwfgpu_fermion_from_canonical(self.h, a.h, b.h, canon),
             "fermion_from_canonical")
        return a, b

    def fermion_to_canonical(self, first: DeviceVector, second: DeviceVector):
        out = np.empty(self.ls * self._v4() * 24)
        _chk(lib().dwfgpu_fermion_to_canonical(self.h, out, first.h, second.h),
             "fermion_to_canonical")
        return out

    def Ffour2five(self, four, s_u, s_l):
        """FdwfBase::Ffour2five (Ncb 2) into device halves (first, second)."""
        four = _arr(four).reshape(-1)
        if four.size != self._v4() * 24:
            raise ValueError("4D field must hold V4*24 reals")
        a, b = self.vector(), self.vector()
        _chk(lib().dwfgpu_four2five(self.h, a.h, b.h, four, s_u, s_l), "Ffour2five")
        return a, b

    def Ffive2four(self, first: DeviceVector, second: DeviceVector, s_u, s_l):
        out = np.empty(self._v4() * 24)
        _chk(lib().dwfgpu_five2four(self.h, out, first.h, second.h, s_u, s_l), "Ffive2four")
        return out

    def MatInv_dev(self, out_first, out_second, in_first, in_second, cg_arg: CgArg):
        """MatInv on device halves; returns (iters, true_res, status)."""
        tr, it = C.c_double(), C.c_int()
        a = cg_arg.c()
        st = _chk(lib().dwfgpu_matinv_dev(self.h, C.byref(a), out_first.h, out_second.h,
                                          in_first.h, in_second.h, C.byref(tr), C.byref(it)),
                  "MatInv_dev", allow=(OK, MAXITER))
        return it.value, tr.value, st

    # --- eigCG (inv_eig_cg.C:129-631) -------------------------------------
    def InvEigCg_dev(self, sol: DeviceVector, src: DeviceVector, cg_arg: CgArg, nev=0, m=0,
                     V: MultiVector = None, U: MultiVector = None, invH=None, restart=()):
        """DiracOp::InvEigCg.  Returns (iters, true_res, status, M)."""
        tr, it = C.c_double(), C.c_int()
        a = cg_arg.c()
        M = np.zeros(max(2 * nev, 1))
        dl = 0 if invH is None else int(np.asarray(invH).shape[0])
        ih = None
        if dl:
            ih = np.ascontiguousarray(np.asarray(invH, np.complex128)).view(np.float64)
        rs = np.ascontiguousarray(restart, np.float64) if len(restart) else None
        st = _chk(lib().dwfgpu_inveigcg_dev(
            self.h, C.byref(a), sol.h, src.h, nev, m, V.h if V is not None else None,
            M.ctypes.data_as(C.c_void_p), U.h if U is not None else None,
            None if ih is None else ih.ctypes.data_as(C.c_void_p), dl,
            None if rs is None else rs.ctypes.data_as(C.c_void_p), 0 if rs is None else rs.size,
            C.byref(tr), C.byref(it)), "InvEigCg", allow=(OK, MAXITER))
        return it.value, tr.value, st, M[:2 * nev]

    def ImpResLanczos(self, v0, mass, nk, nt, np_, maxiters, stop_residual, npol, p0, p1,
                      shift=0.0, conv_check=1, vectors=True):
        """DiracOp::ImpResLanczos on the Chebyshev-filtered MatPcDagMatPc.
        Returns (lambda[nt], vectors [nt, n] or None, restarts)."""
        a = LanczosArgC(mass, nk, nt, np_, maxiters, stop_residual, npol, p0, p1, shift,
                        conv_check)
        lam = np.zeros(nt)
        ev = np.zeros(nt * self.n) if vectors else None
        it = C.c_int()
        _chk(lib().dwfgpu_lanczos(self.h, C.byref(a), _arr(v0),
                                  None if ev is None else ev.ctypes.data_as(C.c_void_p), lam,
                                  C.byref(it)), "ImpResLanczos")
        return lam, (None if ev is None else ev.reshape(nt, self.n)), it.value

    def plaquette(self):
        """(SumReTrPlaq / (18 V), sum ReTr U / (12 V)) of the handle's links."""
        p, t = C.c_double(), C.c_double()
        _chk(lib().dwfgpu_plaquette(self.h, C.byref(p), C.byref(t)), "plaquette")
        return p.value, t.value


class EigCG:
    """The deflation space of QPropW::eig_Run (eig_QPropW.C:85-330) around
    InvEigCg: one solve() per right-hand side grows U (up to max_def_len
    single-precision Ritz vectors) and H = U^dag A U on the device."""

    def __init__(self, op: "_DiracOpGpu", nev, m, max_def_len, max_eig):
        self.op, self.nev = op, nev
        self.h = C.c_void_p()
        _chk(lib().dwfgpu_eigcg_create(op.h, nev, m, max_def_len, max_eig, C.byref(self.h)),
             "eigcg_create")

    def solve(self, sol: DeviceVector, src: DeviceVector, cg_arg: CgArg, restart=(),
              always_restart=False):
        tr, it = C.c_double(), C.c_int()
        a = cg_arg.c()
        rs = np.ascontiguousarray(restart, np.float64) if len(restart) else None
        st = _chk(lib().dwfgpu_eigcg_solve(
            self.h, C.byref(a), sol.h, src.h,
            None if rs is None else rs.ctypes.data_as(C.c_void_p), 0 if rs is None else rs.size,
            1 if always_restart else 0, C.byref(tr), C.byref(it)), "eigcg_solve",
            allow=(OK, MAXITER))
        return it.value, tr.value, st

    @property
    def def_len(self):
        return int(lib().dwfgpu_eigcg_def_len(self.h))

    def values(self):
        """(M: the last window's 2 nev Ritz values, H: def_len x def_len complex)."""
        M = np.zeros(2 * self.nev)
        dl = self.def_len
        Hm = np.zeros(max(2 * dl * dl, 1))
        _chk(lib().dwfgpu_eigcg_values(self.h, M.ctypes.data_as(C.c_void_p),
                                       Hm.ctypes.data_as(C.c_void_p)), "eigcg_values")
        return M, Hm[:2 * dl * dl].view(np.complex128).reshape(dl, dl)

    def free(self):
        if self.h:
            lib().dwfgpu_eigcg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class NerscHeader:
    dims: tuple
    recon_row_3: int
    fp_bytes: int
    big_endian: int
    checksum: int
    checksum_data: int
    plaquette: float
    link_trace: float
    data_start: int


def read_nersc(path, with_links=True):
    """ReadLatticeParallel::read steps 0-2 (ReadLatticePar.C:26-175): returns
    (links CANONICAL [V4*4*18] or None, NerscHeader).  Raises DwfGpuError
    on a missing/short file or a checksum mismatch."""
    h = NerscHeaderC()
    _chk(lib().dwfgpu_nersc_read(os.fsencode(path), None, 0, C.byref(h)), "nersc_read")
    links = None
    if with_links:
        v4 = int(np.prod(list(h.dims)))
        links = np.empty(v4 * 4 * 18)
        _chk(lib().dwfgpu_nersc_read(os.fsencode(path), links.ctypes.data_as(C.c_void_p),
                                     links.size, C.byref(h)), "nersc_read")
    return links, NerscHeader(tuple(h.dims), h.recon_row_3, h.fp_bytes, h.big_endian,
                              h.checksum, h.checksum_data, h.plaquette, h.link_trace,
                              h.data_start)


def check_plaquette(op: "_DiracOpGpu", hdr: NerscHeader, chk_prec):
    """ReadLatticeParallel::CheckPlaqLinktrace (ReadLatticePar.C:396-450):
    the plaquette of the loaded links against the header's, relative
    deviation at most chk_prec (the link trace is reported, not checked)."""
    plaq, tr = op.plaquette()
    dev = abs((plaq - hdr.plaquette) / plaq)
    if dev > chk_prec:
        raise ValueError(f"plaquette {plaq:.8e} differs from header {hdr.plaquette:.8e} "
                         f"(rel. dev. {dev:.3e} > {chk_prec})")
    return plaq, tr


class DiracOpDwfGpu(_DiracOpGpu):
    """DiracOpDwf (cps_pp/src/util/dirac_op/d_op_dwf/d_op_dwf.C)."""
    action = DWF

    def __init__(self, gauge_canonical, dims, ls, m5=1.8, a5inv=1.0, bc=(0, 0, 0, 1),
                 precision="f64"):
        super().__init__(gauge_canonical, dims, ls, m5, a5inv, 1.0, 0.0, bc, precision)


class DiracOpMobiusGpu(_DiracOpGpu):
    """DiracOpMobius (cps_pp/src/util/dirac_op/d_op_mobius/d_op_mobius.C)."""
    action = MOBIUS

    def __init__(self, gauge_canonical, dims, ls, m5=1.8, b=1.5, c=0.5, a5inv=1.0,
                 bc=(0, 0, 0, 1), precision="f64"):
        super().__init__(gauge_canonical, dims, ls, m5, a5inv, b, c, bc, precision)
