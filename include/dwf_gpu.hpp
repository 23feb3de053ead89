// dwf_gpu.hpp -- header-only C++ facade over the C ABI of dwf_gpu.h with the
// member functions of CPS's DiracOpDwf / DiracOpMobius
// (include/util/dirac_op/d_op_dwf.h:26-140, include/util/dirac_op.h:1796-1900),
// so that FmatEvlInv-style host code swaps one type.  Vectors are CPS
// half (or, for MatInv, full) vectors of FP64 reals in CPS's own order;
// CPS's Vector* maps to double* when Float is double (its default).
//
// Errors: CPS aborts through ERR.General; the facade throws dwfgpu::Error
// carrying the C status and message.  MAXITER is not an error (CPS warns):
// the iteration count and true residual are returned as usual.
#ifndef DWF_GPU_HPP
#define DWF_GPU_HPP

#include <stdexcept>
#include <string>

#include "dwf_gpu.h"

namespace dwfgpu {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& where)
      : std::runtime_error(where + " failed with status " + std::to_string(code) + ": " +
                           dwfgpu_last_error()),
        code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline int check(int code, const char* where, bool maxiter_ok = false) {
  if (code == DWFGPU_OK || (maxiter_ok && code == DWFGPU_MAXITER)) return code;
  throw Error(code, where);
}

// Mirror of CgArg (include/input/cg_arg.x:3-28), the fields InvCg reads.
struct CgArg {
  double mass = 0.0;
  int max_num_iter = 10000;
  double stop_rsd = 1e-10;
  double true_rsd = 1e-10;
  dwfgpu_cg_arg c() const { return dwfgpu_cg_arg{mass, max_num_iter, stop_rsd, true_rsd}; }
};

// Lattice and action parameters CPS reads from GJP (GJP.NodeSites, SnodeSites,
// DwfHeight, DwfA5Inv, Mobius_b/c, Bc).
struct Params {
  int dims[4] = {4, 4, 4, 4};
  int ls = 4;
  double m5 = 1.8;
  double a5inv = 1.0;
  double mobius_b = 1.5;
  double mobius_c = 0.5;
  int bc[4] = {0, 0, 0, 1};  // 1: anti-periodic links at x_mu = L_mu - 1
  int precision = DWFGPU_F64;
};

class DiracOpGpu {
 public:
  DiracOpGpu(const DiracOpGpu&) = delete;
  DiracOpGpu& operator=(const DiracOpGpu&) = delete;
  virtual ~DiracOpGpu() {
    if (h_) dwfgpu_destroy(h_);
  }

  // number of reals in one half vector (CPS: RitzLatSize / f_size_cb)
  size_t HalfSize() const { return dwfgpu_half_size(h_); }
  dwfgpu_t handle() const { return h_; }

  // DiracOp::MatPcDagMatPc(out, in, dot_prd)     dirac_op.h:259
  void MatPcDagMatPc(double* out, const double* in, double* dot_prd = nullptr) {
    check(dwfgpu_matpcdagmatpc(h_, arg_.mass, out, in, dot_prd), "MatPcDagMatPc");
  }
  // DiracOp::MatPc / MatPcDag                     d_op_dwf.C:245-300
  void MatPc(double* out, const double* in) {
    check(dwfgpu_matpc(h_, arg_.mass, out, in, 0), "MatPc");
  }
  void MatPcDag(double* out, const double* in) {
    check(dwfgpu_matpc(h_, arg_.mass, out, in, 1), "MatPcDag");
  }
  // DiracOp::Dslash(out, in, cb, dag)             d_op_dwf.C:210-236
  void Dslash(double* out, const double* in, int cb, int dag) {
    check(dwfgpu_dslash(h_, arg_.mass, out, in, cb, dag), "Dslash");
  }
  // DiracOp::InvCg(out, in, true_res): returns CPS's iteration count
  int InvCg(double* out, const double* in, double* true_res = nullptr) {
    int iters = 0;
    const dwfgpu_cg_arg a = arg_.c();
    check(dwfgpu_invcg(h_, &a, out, in, true_res, &iters), "InvCg", true);
    return iters;
  }
  // DiracOp::InvCg(Float *true_res) on the vectors given at construction
  // (inv_cg.C:599-600; CPS's DWF solve calls it from FdwfBase::FmatEvlInv,
  // f_dwf_base.C:200): out = f_out, in = f_in of SetVectors
  int InvCg(double* true_res = nullptr) {
    if (!f_out_ || !f_in_) check(DWFGPU_BAD_ARG, "InvCg(true_res): no vectors set");
    return InvCg(f_out_, f_in_, true_res);
  }
  // DiracOp::InvCgShift(out, in, src_norm_sq, true_res, shift) (inv_cg.C:84-448)
  // and InvCg(out, in, src_norm_sq, true_res) (shift nullptr)
  int InvCgShift(double* out, const double* in, double src_norm_sq, double* true_res,
                 const double* shift) {
    int iters = 0;
    const dwfgpu_cg_arg a = arg_.c();
    check(dwfgpu_invcg_shift(h_, &a, out, in, src_norm_sq, true_res, shift, &iters),
          "InvCgShift", true);
    return iters;
  }
  int InvCg(double* out, const double* in, double src_norm_sq, double* true_res) {
    return InvCgShift(out, in, src_norm_sq, true_res, nullptr);
  }
  // the DiracOp constructor's f_field_out / f_field_in (dirac_op_base.C:95-110),
  // the vectors of InvCg(true_res)
  void SetVectors(double* f_out, const double* f_in) {
    f_out_ = f_out;
    f_in_ = f_in;
  }
  // DiracOpDwf::MatInv / DiracOpMobius::MatInv (PRESERVE_YES)
  int MatInv(double* out, const double* in, double* true_res = nullptr) {
    int iters = 0;
    const dwfgpu_cg_arg a = arg_.c();
    check(dwfgpu_matinv(h_, &a, out, in, true_res, &iters), "MatInv", true);
    return iters;
  }
  // InvCg's solve with a sloppy operator and FP64 reliable updates
  int InvCgMixed(double* out, const double* in, int sloppy, double delta = 0.1,
                 double* true_res = nullptr, int* reliable_updates = nullptr) {
    int iters = 0, nrel = 0;
    const dwfgpu_cg_arg a = arg_.c();
    check(dwfgpu_invcg_mixed(h_, &a, sloppy, delta, out, in, true_res, &iters, &nrel),
          "InvCgMixed", true);
    if (reliable_updates) *reliable_updates = nrel;
    return iters;
  }
  // DiracOp::MInvCG (d_op_base/noarch/minvcg.C:47-355): psi = nshift vectors
  // (MULTI) or one (SINGLE, with alpha); rsdcg in: relative stop residuals,
  // out: |r_s|; returns CPS's loop count k
  int MInvCG(double* psi, const double* chi, const double* shifts, int nshift, double* rsdcg,
             int type = DWFGPU_SHIFT_MULTI, const double* alpha = nullptr,
             int* iters = nullptr, double chi_norm = 0.0) {
    int k = 0;
    check(dwfgpu_minvcg(h_, arg_.mass, shifts, nshift, psi, chi, chi_norm > 0 ? chi_norm : -1.0,
                        rsdcg, type, alpha, arg_.max_num_iter, iters, &k),
          "MInvCG", true);
    return k;
  }
  // DiracOp::ImpResLanczos (d_op_base/comsrc/lanczos.C:193-840) on the
  // Chebyshev-filtered MatPcDagMatPc: lambda[nt], evecs nt half vectors (or
  // nullptr); returns the restart count
  int ImpResLanczos(const double* v0, double* lambda, double* evecs,
                    const dwfgpu_lanczos_arg& a) {
    int it = 0;
    dwfgpu_lanczos_arg arg = a;
    arg.mass = arg_.mass;
    check(dwfgpu_lanczos(h_, &arg, v0, evecs, lambda, &it), "ImpResLanczos");
    return it;
  }
  // QPropW's 4D propagator column (QPropW.C:1204-1236): Ffour2five(src4,
  // 0, Ls-1) -> MatInv -> Ffive2four(sol4, Ls-1, 0), all 5D work on the
  // device; src4 / sol4 are CANONICAL 4D fermions (V4*24 reals)
  int PropagatorColumn(double* sol4, const double* src4, int ls, double* true_res = nullptr) {
    dwfgpu_vec_t v[4] = {nullptr, nullptr, nullptr, nullptr};
    int st = DWFGPU_OK;
    for (auto& x : v)
      if (!st) st = dwfgpu_vec_create(h_, &x);
    int iters = 0;
    const dwfgpu_cg_arg a = arg_.c();
    if (!st) st = dwfgpu_four2five(h_, v[0], v[1], src4, 0, ls - 1);
    if (!st) st = dwfgpu_matinv_dev(h_, &a, v[2], v[3], v[0], v[1], true_res, &iters);
    if (st == DWFGPU_OK || st == DWFGPU_MAXITER) {
      const int r = dwfgpu_five2four(h_, sol4, v[2], v[3], ls - 1, 0);
      if (r) st = r;
    }
    for (auto& x : v)
      if (x) dwfgpu_vec_destroy(x);
    check(st, "PropagatorColumn", true);
    return iters;
  }
  // DiracOp::DiracArg(CgArg*): the mass and CG controls of later calls
  void DiracArg(const CgArg& arg) { arg_ = arg; }

 protected:
  DiracOpGpu(const double* gauge_canonical, const Params& p, const CgArg& arg, int action)
      : arg_(arg) {
    check(dwfgpu_create(gauge_canonical, p.dims, p.ls, action, p.m5, p.a5inv, p.mobius_b,
                        p.mobius_c, p.bc, p.precision, &h_),
          "dwfgpu_create");
  }

 private:
  dwfgpu_t h_ = nullptr;
  CgArg arg_;
  double* f_out_ = nullptr;
  const double* f_in_ = nullptr;
};

// DiracOpDwf (src/util/dirac_op/d_op_dwf/d_op_dwf.C)
class DiracOpDwf : public DiracOpGpu {
 public:
  DiracOpDwf(const double* gauge_canonical, const Params& p, const CgArg& arg)
      : DiracOpGpu(gauge_canonical, p, arg, DWFGPU_DWF) {}
};

// DiracOpMobius (src/util/dirac_op/d_op_mobius/d_op_mobius.C)
class DiracOpMobius : public DiracOpGpu {
 public:
  DiracOpMobius(const double* gauge_canonical, const Params& p, const CgArg& arg)
      : DiracOpGpu(gauge_canonical, p, arg, DWFGPU_MOBIUS) {}
};

}  // namespace dwfgpu

#endif
