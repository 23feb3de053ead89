// dwf_gpu.hip -- host side of the MI355X domain-wall / Mobius library: the
// C ABI of include/dwf_gpu.h, device memory, kernel launches and the CG loop.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dwf_gpu.h"
#include "dwf_kernels.hpp"
#include "roofline.h"

using dwfk::cx;
using dwfk::Geom;
using dwfk::M5Side;
using dwfk::MobArgs;

namespace {

thread_local std::string g_err;

void set_err(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_err("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));       \
      return DWFGPU_HIP_ERROR;                                                         \
    }                                                                                  \
  } while (0)

#define TRY(expr)                     \
  do {                                \
    int r_ = (expr);                  \
    if (r_ != DWFGPU_OK) return r_;   \
  } while (0)


}  // namespace

struct dwfgpu_s {
  int action = 0, prec = 0;
  int dims[4] = {0, 0, 0, 0};
  int ls = 0;
  double m5 = 0, a5inv = 1, b = 1, c = 0;
  Geom g{};
  size_t n = 0;        // reals per half vector
  size_t elem = 8;     // bytes per stored real
  hipStream_t st = nullptr;
  void* U = nullptr;
  void* te = nullptr;  // opposite-parity temporary
  void* t1 = nullptr;  // MatPc output inside MatPcDagMatPc
  void* cps_stage = nullptr;  // FP64 buffer in CPS order (host-array API)
  void* work[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double* partial = nullptr;
  int npartial = 0;
  double* dscal = nullptr;
  double* hscal = nullptr;  // pinned
  int grid_hop = 0;
  int ns = 16;   // sites per work-group of the hopping kernels (x); y = all slices
  int wpe = 2;   // minimum waves per SIMD the hopping kernels are compiled for
};

struct dwfgpu_vec_s {
  dwfgpu_s* h = nullptr;
  void* d = nullptr;
};

namespace {

// ---------------------------------------------------------------------------
// coefficients
// ---------------------------------------------------------------------------
double dwf_kappa(const dwfgpu_s* h) { return 1.0 / (2 * (4 + h->a5inv - h->m5)); }
double mob_kappa_b(const dwfgpu_s* h) { return 1.0 / (2 * (h->b * (4 - h->m5) + h->a5inv)); }
double mob_kappa_c(const dwfgpu_s* h) { return 1.0 / (2 * (h->c * (4 - h->m5) - h->a5inv)); }

// M5^-1 coefficients following the products of zmobius_m5inv.C (dag 0 :271-346,
// dag 1 :559-631) for a constant real K = kappa_b/kappa_c.
M5Side m5_side(const dwfgpu_s* h, double mass, int dag) {
  typedef std::complex<double> C;
  const int ls = h->ls;
  const C K(mob_kappa_b(h) / mob_kappa_c(h), 0.0);
  M5Side k;
  std::memset(&k, 0, sizeof k);
  if (!dag) {
    C prod = 1.0;
    for (int i = 0; i < ls; ++i) prod *= -K;
    prod *= mass;
    const C inv_d = 1.0 / (1.0 + prod);
    // P-: backward-corner (fout[ls-1] gathers, then backward sweep)
    C fL = K * mass * inv_d;
    for (int s = 0; s <= ls - 2; ++s) {
      k.cbk[s] = fL.real();
      fL *= -K;
    }
    for (int s = 0; s < ls; ++s) k.ab[s] = (-K).real();
    k.inv_b = inv_d.real();
    // P+: forward sweep then corner correction
    for (int s = 0; s < ls; ++s) k.af[s] = (-K).real();
    C q = 1.0;
    for (int i = 0; i < ls - 1; ++i) q *= -K;
    C fR = -q * mass * inv_d;
    for (int s = ls - 2; s >= 0; --s) {
      k.cf[s] = fR.real();
      fR /= -K;
    }
    k.inv_f = inv_d.real();
  } else {
    C prod = 1.0;
    for (int i = 0; i < ls; ++i) prod *= K;
    prod *= mass;
    const C inv_d = std::conj(1.0 / (1.0 + prod));
    // P+: backward-corner with conj coefficients
    C fR = std::conj(K) * mass * inv_d;
    for (int s = 0; s <= ls - 2; ++s) {
      k.cbk[s] = fR.real();
      fR *= -std::conj(K);
    }
    for (int s = 0; s < ls; ++s) k.ab[s] = (-std::conj(K)).real();
    k.inv_b = inv_d.real();
    // P-: forward sweep then corner correction
    for (int s = 0; s < ls; ++s) k.af[s] = (-std::conj(K)).real();
    C q = 1.0;
    for (int i = 0; i < ls - 2; ++i) q *= -K;
    q *= -K;
    C fL = -std::conj(q) * mass * inv_d;
    for (int s = ls - 2; s >= 0; --s) {
      k.cf[s] = fL.real();
      if (s > 0) fL /= -std::conj(K);
    }
    k.inv_f = inv_d.real();
  }
  return k;
}

MobArgs mob_args(const dwfgpu_s* h, double mass, int dag) {
  MobArgs a;
  const double kb = mob_kappa_b(h), kc = mob_kappa_c(h);
  a.b = h->b;
  a.c = h->c;
  a.mc = -mass * h->c;
  a.K = kb / kc;
  a.mK = -mass * a.K;
  a.kappa_b = kb;
  a.m5 = m5_side(h, mass, dag);
  return a;
}

// LDS of the s-column buffer (M5^-1 and (b + c D5)^dag epilogues)
size_t col_lds_bytes(const dwfgpu_s* h) { return (size_t)h->ns * (h->ls * 12 + 1) * 2 * h->elem; }
// LDS of the staged links: nsets parities x 8 directions x ns sites x 9
size_t link_lds_bytes(const dwfgpu_s* h, int nsets) {
  return (size_t)nsets * 8 * h->ns * 9 * 2 * h->elem;
}
// largest LDS request of any kernel of this handle
size_t max_lds_bytes(const dwfgpu_s* h) {
  const size_t dwf = link_lds_bytes(h, 2);
  const size_t mob = link_lds_bytes(h, 1) + col_lds_bytes(h);
  return h->action == DWFGPU_DWF ? dwf : mob;
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
int fetch_sum(dwfgpu_s* h, int nparts, double* out) {
  hipLaunchKernelGGL(dwfk::k_reduce_partials, dim3(1), dim3(1024), 0, h->st, h->partial, nparts,
                     h->dscal);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h->hscal, h->dscal, sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  *out = h->hscal[0];
  return DWFGPU_OK;
}

dim3 blas_grid() { return dim3(dwfk::BLAS_BLOCKS); }
dim3 blas_block() { return dim3(dwfk::BLAS_THREADS); }

// ---------------------------------------------------------------------------
// operator pieces, templated on the storage real R
// ---------------------------------------------------------------------------
template <typename R>
struct Ops {
  typedef cx<R> C;

  // DWF hop kernel: epi 0 -> out = D in ; epi 1 -> out = x0 + alpha D in
  static int dwf(dwfgpu_s* h, double mass, void* out, const void* in, int cb_in, int dag,
                 int epi, const void* x0, double alpha, bool norm) {
    const dim3 grid(h->grid_hop), block(h->ns, h->ls);
    const double two_a5 = 2.0 * h->a5inv, nm = -2.0 * mass * h->a5inv;
    C* o = (C*)out;
    const C* i = (const C*)in;
    const C* u = (const C*)h->U;
    const C* x = (const C*)x0;
    const size_t lds = link_lds_bytes(h, 2);
#define DWF_LAUNCH(DAG, EPI, NORM)                                                              \
  do {                                                                                          \
    if (h->wpe >= 4)                                                                            \
      hipLaunchKernelGGL((dwfk::k_dwf_dslash<R, DAG, EPI, NORM, 4>), grid, block, lds, h->st, o,  \
                         i, u, x, h->partial, h->g, cb_in, two_a5, nm, alpha);                  \
    else                                                                                        \
      hipLaunchKernelGGL((dwfk::k_dwf_dslash<R, DAG, EPI, NORM, 2>), grid, block, lds, h->st, o,  \
                         i, u, x, h->partial, h->g, cb_in, two_a5, nm, alpha);                  \
  } while (0)
    if (epi == 0) {
      if (dag) DWF_LAUNCH(1, 0, false); else DWF_LAUNCH(0, 0, false);
    } else if (norm) {
      if (dag) DWF_LAUNCH(1, 1, true); else DWF_LAUNCH(0, 1, true);
    } else {
      if (dag) DWF_LAUNCH(1, 1, false); else DWF_LAUNCH(0, 1, false);
    }
#undef DWF_LAUNCH
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  // Mobius hop kernel; fl: dwfk::FL_* flags
  static int mob(dwfgpu_s* h, double mass, void* out, const void* in, const void* psi, int p_out,
                 int dag, int epi, int fl) {
    const dim3 grid(h->grid_hop), block(h->ns, h->ls);
    const MobArgs a = mob_args(h, mass, dag);
    C* o = (C*)out;
    const C* i = (const C*)in;
    const C* p = (const C*)psi;
    const C* u = (const C*)h->U;
    const bool col = dag || epi == dwfk::EPI_M5INV;
    const size_t lds = link_lds_bytes(h, 1) + (col ? col_lds_bytes(h) : 0);
#define MOB_LAUNCH(DAG, EPI, FL)                                                               \
  do {                                                                                         \
    if (h->wpe >= 4)                                                                           \
      hipLaunchKernelGGL((dwfk::k_mobius<R, DAG, EPI, FL, 4>), grid, block, lds, h->st, o, i, p, \
                         u, h->partial, h->g, p_out, a);                                       \
    else                                                                                       \
      hipLaunchKernelGGL((dwfk::k_mobius<R, DAG, EPI, FL, 2>), grid, block, lds, h->st, o, i, p, \
                         u, h->partial, h->g, p_out, a);                                       \
  } while (0)
    using dwfk::FL_NORM;
    using dwfk::FL_POST_T;
    using dwfk::FL_PRE_T;
    if (dag == 0 && epi == dwfk::EPI_M5INV) {
      if (fl & FL_POST_T) MOB_LAUNCH(0, dwfk::EPI_M5INV, FL_POST_T);
      else MOB_LAUNCH(0, dwfk::EPI_M5INV, 0);
    } else if (dag == 0 && epi == dwfk::EPI_SCHUR) {
      if ((fl & FL_PRE_T) && (fl & FL_NORM)) MOB_LAUNCH(0, dwfk::EPI_SCHUR, FL_PRE_T | FL_NORM);
      else if (fl & FL_PRE_T) MOB_LAUNCH(0, dwfk::EPI_SCHUR, FL_PRE_T);
      else if (fl & FL_NORM) MOB_LAUNCH(0, dwfk::EPI_SCHUR, FL_NORM);
      else MOB_LAUNCH(0, dwfk::EPI_SCHUR, 0);
    } else if (dag == 0) {
      MOB_LAUNCH(0, dwfk::EPI_PLAIN, 0);
    } else if (epi == dwfk::EPI_M5INV) {
      MOB_LAUNCH(1, dwfk::EPI_M5INV, 0);
    } else if (epi == dwfk::EPI_SCHUR) {
      MOB_LAUNCH(1, dwfk::EPI_SCHUR, 0);
    } else {
      MOB_LAUNCH(1, dwfk::EPI_PLAIN, 0);
    }
#undef MOB_LAUNCH
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  // first / second half of MatPc (dag 0) or MatPcDag (dag 1): the first writes
  // h->te (for dag 0 already multiplied by (b + c D5), consumed by the second)
  static int mob_first(dwfgpu_s* h, double mass, void* te, const void* in, int dag) {
    return mob(h, mass, te, in, nullptr, 0, dag, dwfk::EPI_M5INV, dag ? 0 : dwfk::FL_POST_T);
  }
  static int mob_second(dwfgpu_s* h, double mass, void* out, const void* te, const void* in,
                        int dag, bool norm) {
    const int fl = (dag ? 0 : dwfk::FL_PRE_T) | (norm && !dag ? dwfk::FL_NORM : 0);
    return mob(h, mass, out, te, in, 1, dag, dwfk::EPI_SCHUR, fl);
  }

  static int m5inv(dwfgpu_s* h, double mass, void* out, const void* in, int dag) {
    const dim3 grid(h->grid_hop), block(h->ns, h->ls);
    const M5Side k = m5_side(h, mass, dag);
    if (dag)
      hipLaunchKernelGGL((dwfk::k_m5inv<R, 1>), grid, block, col_lds_bytes(h), h->st, (C*)out,
                         (const C*)in, h->g, k);
    else
      hipLaunchKernelGGL((dwfk::k_m5inv<R, 0>), grid, block, col_lds_bytes(h), h->st, (C*)out,
                         (const C*)in, h->g, k);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  // MatPc / MatPcDag: out = M in (out on the CG parity).  norm -> partials.
  static int matpc(dwfgpu_s* h, double mass, void* out, const void* in, int dag, bool norm) {
    if (h->action == DWFGPU_DWF) {
      const double mk2 = -dwf_kappa(h) * dwf_kappa(h);
      TRY(dwf(h, mass, h->te, in, 1, dag, 0, nullptr, 0.0, false));
      TRY(dwf(h, mass, out, h->te, 0, dag, 1, in, mk2, norm));
    } else {
      TRY(mob_first(h, mass, h->te, in, dag));
      TRY(mob_second(h, mass, out, h->te, in, dag, norm));
    }
    return DWFGPU_OK;
  }

  static int mdagm(dwfgpu_s* h, double mass, void* out, const void* in, double* dot) {
    TRY(matpc(h, mass, h->t1, in, 0, dot != nullptr));
    if (dot) TRY(fetch_sum(h, h->grid_hop, dot));
    TRY(matpc(h, mass, out, h->t1, 1, false));
    return DWFGPU_OK;
  }

  static int dslash(dwfgpu_s* h, double mass, void* out, const void* in, int cb, int dag) {
    if (h->action == DWFGPU_DWF) return dwf(h, mass, out, in, cb, dag, 0, nullptr, 0.0, false);
    return mob(h, mass, out, in, nullptr, 1 - cb, dag, dwfk::EPI_PLAIN, 0);
  }

  static int norm(dwfgpu_s* h, const void* x, double* out) {
    hipLaunchKernelGGL((dwfk::k_norm<R>), blas_grid(), blas_block(), 0, h->st, (const R*)x, h->n,
                       h->partial);
    HIPCHK(hipGetLastError());
    return fetch_sum(h, dwfk::BLAS_BLOCKS, out);
  }

  static int axpy_to(dwfgpu_s* h, void* z, const void* x, const void* y, double a) {
    hipLaunchKernelGGL((dwfk::k_axpy_to<R, false>), blas_grid(), blas_block(), 0, h->st, (R*)z,
                       (const R*)x, (const R*)y, a, h->n, h->partial);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  static int axpby(dwfgpu_s* h, void* y, const void* x, double a, double b) {
    hipLaunchKernelGGL((dwfk::k_axpby<R>), blas_grid(), blas_block(), 0, h->st, (R*)y,
                       (const R*)x, a, b, h->n);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  static int scale(dwfgpu_s* h, void* x, double a) {
    hipLaunchKernelGGL((dwfk::k_scale<R>), blas_grid(), blas_block(), 0, h->st, (R*)x, a, h->n);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  // DiracOp::InvCgShift, shift = NULL (inv_cg.C:84-448)
  static int invcg(dwfgpu_s* h, const dwfgpu_cg_arg* arg, void* sol, const void* src,
                   double* true_res, int* iters) {
    void* res = h->work[0];
    void* dir = h->work[1];
    void* mmp = h->work[2];
    const double mass = arg->mass;
    double src_norm_sq;
    TRY(norm(h, src, &src_norm_sq));
    const double stp_cnd = src_norm_sq * arg->stop_rsd * arg->stop_rsd;
    TRY(mdagm(h, mass, mmp, sol, nullptr));
    // res = src - mmp ; dir = res ; rr = |res|^2
    hipLaunchKernelGGL((dwfk::k_axpy_to<R, true>), blas_grid(), blas_block(), 0, h->st, (R*)res,
                       (const R*)mmp, (const R*)src, -1.0, h->n, h->partial);
    HIPCHK(hipGetLastError());
    double rr;
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
    HIPCHK(hipMemcpyAsync(dir, res, h->n * h->elem, hipMemcpyDeviceToDevice, h->st));
    int itr = 0;
    int max_itr = arg->max_num_iter - 1;
    if (rr <= stp_cnd) max_itr = 0;
    int status = DWFGPU_OK;
    for (int k = 0; k < max_itr; ++k) {
      itr = itr + 1;
      const double rr_prv = rr;
      double d;
      TRY(mdagm(h, mass, mmp, dir, &d));
      if (d == 0.0) {
        status = DWFGPU_BREAKDOWN;
        break;
      }
      const double a = rr_prv / d;
      hipLaunchKernelGGL((dwfk::k_cg_update<R>), blas_grid(), blas_block(), 0, h->st, (R*)sol,
                         (R*)res, (const R*)dir, (const R*)mmp, a, h->n, h->partial);
      HIPCHK(hipGetLastError());
      TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
      if (rr <= stp_cnd) break;
      const double b = rr / rr_prv;
      TRY(axpby(h, dir, res, 1.0, b));  // dir = res + b dir
    }
    // true residual |src - A sol| / |src|  (inv_cg.C:403-415)
    TRY(mdagm(h, mass, mmp, sol, nullptr));
    hipLaunchKernelGGL((dwfk::k_axpy_to<R, true>), blas_grid(), blas_block(), 0, h->st, (R*)res,
                       (const R*)mmp, (const R*)src, -1.0, h->n, h->partial);
    HIPCHK(hipGetLastError());
    double tr;
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &tr));
    if (true_res) *true_res = std::sqrt(tr / src_norm_sq);
    if (iters) *iters = itr + 1;
    if (status == DWFGPU_OK && itr == arg->max_num_iter - 1) status = DWFGPU_MAXITER;
    return status;
  }

  // DiracOpDwf::MatInv (d_op_dwf.C:316-507) / DiracOpMobius::MatInv ORIG
  // (d_op_mobius.C:407-696).  fin/fout: the two device halves in CPS order of
  // the full vector (first, second).
  static int matinv(dwfgpu_s* h, const dwfgpu_cg_arg* arg, void* out_first, void* out_second,
                    const void* in_first, const void* in_second, double* true_res, int* iters) {
    const double mass = arg->mass;
    void* temp = h->work[3];
    void* src = h->work[4];
    int st;
    if (h->action == DWFGPU_DWF) {
      const double kappa = dwf_kappa(h);
      // [odd | even]
      TRY(dwf(h, mass, temp, in_second, 0, 0, 1, in_first, kappa, false));
      TRY(matpc(h, mass, src, temp, 1, false));
      st = invcg(h, arg, out_first, src, true_res, iters);
      if (st != DWFGPU_OK && st != DWFGPU_MAXITER) return st;
      TRY(dwf(h, mass, out_second, out_first, 1, 0, 1, in_second, kappa, false));
    } else {
      // [even | odd]
      const double kb = mob_kappa_b(h);
      void* temp2 = h->work[5];
      TRY(m5inv(h, mass, temp, in_first, 0));
      TRY(mob(h, mass, temp2, temp, nullptr, 1, 0, dwfk::EPI_PLAIN, 0));
      TRY(axpy_to(h, temp, temp2, in_second, kb));
      TRY(matpc(h, mass, src, temp, 1, false));
      st = invcg(h, arg, out_second, src, true_res, iters);
      if (st != DWFGPU_OK && st != DWFGPU_MAXITER) return st;
      TRY(mob(h, mass, temp, out_second, nullptr, 0, 0, dwfk::EPI_PLAIN, 0));
      TRY(scale(h, temp, kb));
      TRY(m5inv(h, mass, out_first, temp, 0));
      TRY(m5inv(h, mass, temp, in_first, 0));
      TRY(axpby(h, out_first, temp, 1.0, 1.0));
      TRY(scale(h, out_first, 2.0 * kb));
      TRY(scale(h, out_second, 2.0 * kb));
    }
    return st;
  }

  static int to_device(dwfgpu_s* h, void* dst, const double* host) {
    HIPCHK(hipMemcpyAsync(h->cps_stage, host, h->n * sizeof(double), hipMemcpyHostToDevice,
                          h->st));
    hipLaunchKernelGGL((dwfk::k_cps_to_soa<R>), dim3(2048), dim3(256), 0, h->st, (C*)dst,
                       (const double*)h->cps_stage, h->g);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  static int to_host(dwfgpu_s* h, double* host, const void* src) {
    hipLaunchKernelGGL((dwfk::k_soa_to_cps<R>), dim3(2048), dim3(256), 0, h->st,
                       (double*)h->cps_stage, (const C*)src, h->g);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, h->cps_stage, h->n * sizeof(double), hipMemcpyDeviceToHost,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return DWFGPU_OK;
  }
};

#define DISPATCH(h, call) ((h)->prec == DWFGPU_F64 ? Ops<double>::call : Ops<float>::call)

int alloc(dwfgpu_s* h, void** p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes));
  HIPCHK(hipMemsetAsync(*p, 0, bytes, h->st));
  return DWFGPU_OK;
}

int ensure_work(dwfgpu_s* h) {
  for (int k = 0; k < 6; ++k)
    if (!h->work[k]) TRY(alloc(h, &h->work[k], h->n * h->elem));
  return DWFGPU_OK;
}

// CANONICAL host gauge -> device layout with BondCond signs
template <typename R>
int upload_gauge(dwfgpu_s* h, const double* uc, const int bc[4]) {
  const int lx = h->dims[0], ly = h->dims[1], lz = h->dims[2], lt = h->dims[3];
  const size_t vh = h->g.vh;
  std::vector<R> dev((size_t)4 * 2 * 9 * vh * 2);
  for (int t = 0; t < lt; ++t)
    for (int z = 0; z < lz; ++z)
      for (int y = 0; y < ly; ++y)
        for (int x = 0; x < lx; ++x) {
          const size_t site = x + (size_t)lx * (y + (size_t)ly * (z + (size_t)lz * t));
          const int par = (x + y + z + t) & 1;
          const size_t cbs = dwfk::site_index(h->g, x / 2, y, z, t);
          const int pos[4] = {x, y, z, t};
          for (int mu = 0; mu < 4; ++mu) {
            const double sg = (bc[mu] && pos[mu] == h->dims[mu] - 1) ? -1.0 : 1.0;
            const double* src = uc + (site * 4 + mu) * 18;
            for (int e = 0; e < 9; ++e) {
              const size_t o = (((size_t)mu * 2 + par) * 9 + e) * vh + cbs;
              dev[2 * o] = (R)(sg * src[2 * e]);
              dev[2 * o + 1] = (R)(sg * src[2 * e + 1]);
            }
          }
        }
  HIPCHK(hipMalloc(&h->U, dev.size() * sizeof(R)));
  HIPCHK(hipMemcpy(h->U, dev.data(), dev.size() * sizeof(R), hipMemcpyHostToDevice));
  return DWFGPU_OK;
}

// Events around each kernel of MatPcDagMatPc (the 4 hopping passes and the
// norm's final reduction); they are recorded on the stream the kernels run on.
template <typename R>
int time_mdagm(dwfgpu_s* h, double mass, void* out, const void* in, int reps, double* ms_app,
               double* ms_kernel, double* bytes_kernel, int* nkernel) {
  typedef Ops<R> O;
  const int nk = 5;
  std::vector<hipEvent_t> ev((size_t)reps * nk + 1, nullptr);
  int st = DWFGPU_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) st = DWFGPU_HIP_ERROR;
  const double half = (double)h->n * h->elem;
  const double links = 72.0 * 2.0 * h->g.vh * h->elem;
  const double kb[nk] = {2 * half + links, 3 * half + links, 2 * half + links, 3 * half + links, 0};
  const bool dwf = h->action == DWFGPU_DWF;
  const double mk2 = dwf ? -dwf_kappa(h) * dwf_kappa(h) : 0.0;
  auto rec = [&](size_t k) {
    if (!st && hipEventRecord(ev[k], h->st) != hipSuccess) st = DWFGPU_HIP_ERROR;
  };
  for (int r = 0; r < reps && !st; ++r) {
    const size_t b = (size_t)r * nk;
    rec(b + 0);
    if (!st) st = dwf ? O::dwf(h, mass, h->te, in, 1, 0, 0, nullptr, 0, false)
                      : O::mob_first(h, mass, h->te, in, 0);
    rec(b + 1);
    if (!st) st = dwf ? O::dwf(h, mass, h->t1, h->te, 0, 0, 1, in, mk2, true)
                      : O::mob_second(h, mass, h->t1, h->te, in, 0, true);
    rec(b + 2);
    if (!st) st = dwf ? O::dwf(h, mass, h->te, h->t1, 1, 1, 0, nullptr, 0, false)
                      : O::mob_first(h, mass, h->te, h->t1, 1);
    rec(b + 3);
    if (!st) st = dwf ? O::dwf(h, mass, out, h->te, 0, 1, 1, h->t1, mk2, false)
                      : O::mob_second(h, mass, out, h->te, h->t1, 1, false);
    rec(b + 4);
    if (!st) {
      hipLaunchKernelGGL(dwfk::k_reduce_partials, dim3(1), dim3(1024), 0, h->st, h->partial,
                         h->grid_hop, h->dscal);
      if (hipGetLastError() != hipSuccess) st = DWFGPU_HIP_ERROR;
    }
  }
  rec((size_t)reps * nk);
  if (!st && hipEventSynchronize(ev[(size_t)reps * nk]) != hipSuccess) st = DWFGPU_HIP_ERROR;
  if (!st) {
    double sum[nk] = {0, 0, 0, 0, 0};
    for (int r = 0; r < reps && !st; ++r)
      for (int k = 0; k < nk; ++k) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev[(size_t)r * nk + k], ev[(size_t)r * nk + k + 1]) !=
            hipSuccess)
          st = DWFGPU_HIP_ERROR;
        sum[k] += ms;
      }
    float tot = 0;
    if (hipEventElapsedTime(&tot, ev[0], ev[(size_t)reps * nk]) != hipSuccess) st = DWFGPU_HIP_ERROR;
    if (ms_app) *ms_app = tot / reps;
    for (int k = 0; k < nk; ++k) {
      if (ms_kernel) ms_kernel[k] = sum[k] / reps;
      if (bytes_kernel) bytes_kernel[k] = kb[k];
    }
    if (nkernel) *nkernel = nk;
  }
  if (st) set_err("dwfgpu_time_matpcdagmatpc: HIP event failure");
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return st;
}

void free_all(dwfgpu_s* h) {
  void* ptrs[] = {h->U, h->te, h->t1, h->cps_stage, h->partial, h->dscal,
                  h->work[0], h->work[1], h->work[2], h->work[3], h->work[4], h->work[5]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->hscal) (void)hipHostFree(h->hscal);
  if (h->st) (void)hipStreamDestroy(h->st);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* dwfgpu_last_error(void) { return g_err.c_str(); }

int dwfgpu_create(const double* gauge, const int dims[4], int ls, int action, double m5,
                  double a5inv, double mob_b, double mob_c, const int bc[4], int precision,
                  dwfgpu_t* out) {
  if (!gauge || !dims || !bc || !out) {
    set_err("dwfgpu_create: null pointer argument");
    return DWFGPU_BAD_ARG;
  }
  for (int mu = 0; mu < 4; ++mu)
    if (dims[mu] < 2 || dims[mu] % 2) {
      set_err("dwfgpu_create: every extent must be even and >= 2 (dims[%d]=%d)", mu, dims[mu]);
      return DWFGPU_BAD_ARG;
    }
  if (ls < 2 || ls > dwfk::MAXLS || ls % 2) {
    set_err("dwfgpu_create: Ls must be even and in [2, %d] (got %d)", dwfk::MAXLS, ls);
    return DWFGPU_BAD_ARG;
  }
  if (action != DWFGPU_DWF && action != DWFGPU_MOBIUS) {
    set_err("dwfgpu_create: unknown action %d", action);
    return DWFGPU_BAD_ARG;
  }
  if (precision != DWFGPU_F64 && precision != DWFGPU_F32) {
    set_err("dwfgpu_create: unknown precision %d", precision);
    return DWFGPU_BAD_ARG;
  }
  dwfgpu_s* h = new dwfgpu_s;
  h->action = action;
  h->prec = precision;
  h->elem = precision == DWFGPU_F64 ? 8 : 4;
  for (int mu = 0; mu < 4; ++mu) h->dims[mu] = dims[mu];
  h->ls = ls;
  h->m5 = m5;
  h->a5inv = a5inv;
  h->b = mob_b;
  h->c = mob_c;
  const long v4 = (long)dims[0] * dims[1] * dims[2] * dims[3];
  if (v4 / 2 > (1L << 30)) {
    delete h;
    set_err("dwfgpu_create: local volume too large");
    return DWFGPU_BAD_ARG;
  }
  // t tiles of 4 slices when possible; 8 XCDs on gfx950 (a different count
  // only changes the work-group order, not the result)
  const int bt = dims[3] % 4 == 0 ? 4 : 2;
  const int bz = dims[2] % 4 == 0 ? 4 : 2;
  h->g = Geom{dims[0], dims[1], dims[2], dims[3], dims[0] / 2, (int)(v4 / 2), ls, bt, bz, 8};
  h->n = (size_t)24 * (v4 / 2) * ls;
  int dev = 0, max_lds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess) {
    set_err("dwfgpu_create: no HIP device");
    delete h;
    return DWFGPU_HIP_ERROR;
  }
  // launch shape of the hopping kernels; DWFGPU_NS / DWFGPU_WPE override
  // the defaults (sites per work-group, minimum waves per SIMD).  Without an
  // override, fewer sites per work-group are taken when the LDS stage of a
  // long fifth dimension would not fit.
  const char* env_ns = getenv("DWFGPU_NS");
  if (env_ns) h->ns = atoi(env_ns);
  if (const char* e = getenv("DWFGPU_WPE")) h->wpe = atoi(e);
  while (!env_ns && h->ns > 4 && max_lds_bytes(h) > (size_t)max_lds) h->ns /= 2;
  if (h->ns < 1 || h->ns * ls > 512 || (h->wpe != 2 && h->wpe != 4)) {
    set_err("dwfgpu_create: bad launch shape ns=%d (ns*Ls <= 512) wpe=%d (2 or 4)", h->ns, h->wpe);
    delete h;
    return DWFGPU_BAD_ARG;
  }
  if (max_lds_bytes(h) > (size_t)max_lds) {
    set_err("dwfgpu_create: Ls=%d needs %zu bytes of LDS per work-group (limit %d)", ls,
            max_lds_bytes(h), max_lds);
    delete h;
    return DWFGPU_BAD_ARG;
  }
  h->grid_hop = (h->g.vh + h->ns - 1) / h->ns;
  h->npartial = h->grid_hop > dwfk::BLAS_BLOCKS ? h->grid_hop : dwfk::BLAS_BLOCKS;
  int r = DWFGPU_OK;
  auto fail = [&](int code) {
    free_all(h);
    delete h;
    return code;
  };
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
    set_err("dwfgpu_create: hipStreamCreate failed");
    delete h;
    return DWFGPU_HIP_ERROR;
  }
  r = precision == DWFGPU_F64 ? upload_gauge<double>(h, gauge, bc) : upload_gauge<float>(h, gauge, bc);
  if (r) return fail(r);
  if ((r = alloc(h, &h->te, h->n * h->elem))) return fail(r);
  if ((r = alloc(h, &h->t1, h->n * h->elem))) return fail(r);
  if ((r = alloc(h, &h->cps_stage, h->n * sizeof(double)))) return fail(r);
  if ((r = alloc(h, (void**)&h->partial, h->npartial * sizeof(double)))) return fail(r);
  if ((r = alloc(h, (void**)&h->dscal, 8 * sizeof(double)))) return fail(r);
  if (hipHostMalloc((void**)&h->hscal, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    set_err("dwfgpu_create: hipHostMalloc failed");
    return fail(DWFGPU_HIP_ERROR);
  }
  if (hipStreamSynchronize(h->st) != hipSuccess) {
    set_err("dwfgpu_create: stream synchronize failed");
    return fail(DWFGPU_HIP_ERROR);
  }
  *out = h;
  return DWFGPU_OK;
}

int dwfgpu_destroy(dwfgpu_t h) {
  if (!h) return DWFGPU_BAD_ARG;
  (void)hipStreamSynchronize(h->st);
  free_all(h);
  delete h;
  return DWFGPU_OK;
}

size_t dwfgpu_half_size(dwfgpu_t h) { return h ? h->n : 0; }

int dwfgpu_matpcdagmatpc(dwfgpu_t h, double mass, double* out, const double* in,
                         double* dot_prd) {
  if (!h || !out || !in) return DWFGPU_BAD_ARG;
  TRY(ensure_work(h));
  TRY(DISPATCH(h, to_device(h, h->work[0], in)));
  TRY(DISPATCH(h, mdagm(h, mass, h->work[1], h->work[0], dot_prd)));
  return DISPATCH(h, to_host(h, out, h->work[1]));
}

int dwfgpu_matpc(dwfgpu_t h, double mass, double* out, const double* in, int dag) {
  if (!h || !out || !in || (dag != 0 && dag != 1)) return DWFGPU_BAD_ARG;
  TRY(ensure_work(h));
  TRY(DISPATCH(h, to_device(h, h->work[0], in)));
  TRY(DISPATCH(h, matpc(h, mass, h->work[1], h->work[0], dag, false)));
  return DISPATCH(h, to_host(h, out, h->work[1]));
}

int dwfgpu_dslash(dwfgpu_t h, double mass, double* out, const double* in, int cb, int dag) {
  if (!h || !out || !in || (dag != 0 && dag != 1) || (cb != 0 && cb != 1)) return DWFGPU_BAD_ARG;
  TRY(ensure_work(h));
  TRY(DISPATCH(h, to_device(h, h->work[0], in)));
  TRY(DISPATCH(h, dslash(h, mass, h->work[1], h->work[0], cb, dag)));
  return DISPATCH(h, to_host(h, out, h->work[1]));
}

int dwfgpu_invcg(dwfgpu_t h, const dwfgpu_cg_arg* arg, double* sol, const double* src,
                 double* true_res, int* iters) {
  if (!h || !arg || !sol || !src) return DWFGPU_BAD_ARG;
  if (arg->max_num_iter < 1) {
    set_err("dwfgpu_invcg: max_num_iter must be >= 1");
    return DWFGPU_BAD_ARG;
  }
  TRY(ensure_work(h));
  void* dsol = h->work[3];
  void* dsrc = h->work[4];
  TRY(DISPATCH(h, to_device(h, dsol, sol)));
  TRY(DISPATCH(h, to_device(h, dsrc, src)));
  const int st = DISPATCH(h, invcg(h, arg, dsol, dsrc, true_res, iters));
  if (st != DWFGPU_OK && st != DWFGPU_MAXITER) return st;
  TRY(DISPATCH(h, to_host(h, sol, dsol)));
  return st;
}

int dwfgpu_matinv(dwfgpu_t h, const dwfgpu_cg_arg* arg, double* out, const double* in,
                  double* true_res, int* iters) {
  if (!h || !arg || !out || !in) return DWFGPU_BAD_ARG;
  if (arg->max_num_iter < 1) return DWFGPU_BAD_ARG;
  TRY(ensure_work(h));
  void *i1, *i2, *o1, *o2;
  const size_t bytes = h->n * h->elem;
  HIPCHK(hipMalloc(&i1, bytes));
  HIPCHK(hipMalloc(&i2, bytes));
  HIPCHK(hipMalloc(&o1, bytes));
  HIPCHK(hipMalloc(&o2, bytes));
  int st = DISPATCH(h, to_device(h, i1, in));
  if (!st) st = DISPATCH(h, to_device(h, i2, in + h->n));
  if (!st) st = DISPATCH(h, to_device(h, o1, out));
  if (!st) st = DISPATCH(h, to_device(h, o2, out + h->n));
  // the CG works in work[0..2]; MatInv temporaries are work[3..5]
  if (!st) st = DISPATCH(h, matinv(h, arg, o1, o2, i1, i2, true_res, iters));
  if (st == DWFGPU_OK || st == DWFGPU_MAXITER) {
    int r = DISPATCH(h, to_host(h, out, o1));
    if (!r) r = DISPATCH(h, to_host(h, out + h->n, o2));
    if (r) st = r;
  }
  (void)hipFree(i1);
  (void)hipFree(i2);
  (void)hipFree(o1);
  (void)hipFree(o2);
  return st;
}

int dwfgpu_vec_create(dwfgpu_t h, dwfgpu_vec_t* v) {
  if (!h || !v) return DWFGPU_BAD_ARG;
  dwfgpu_vec_s* p = new dwfgpu_vec_s;
  p->h = h;
  if (hipMalloc(&p->d, h->n * h->elem) != hipSuccess) {
    delete p;
    set_err("dwfgpu_vec_create: hipMalloc of %zu bytes failed", h->n * h->elem);
    return DWFGPU_HIP_ERROR;
  }
  HIPCHK(hipMemsetAsync(p->d, 0, h->n * h->elem, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  *v = p;
  return DWFGPU_OK;
}

int dwfgpu_vec_destroy(dwfgpu_vec_t v) {
  if (!v) return DWFGPU_BAD_ARG;
  (void)hipStreamSynchronize(v->h->st);
  (void)hipFree(v->d);
  delete v;
  return DWFGPU_OK;
}

int dwfgpu_vec_upload(dwfgpu_vec_t v, const double* host) {
  if (!v || !host) return DWFGPU_BAD_ARG;
  TRY(DISPATCH(v->h, to_device(v->h, v->d, host)));
  HIPCHK(hipStreamSynchronize(v->h->st));
  return DWFGPU_OK;
}

int dwfgpu_vec_download(dwfgpu_vec_t v, double* host) {
  if (!v || !host) return DWFGPU_BAD_ARG;
  return DISPATCH(v->h, to_host(v->h, host, v->d));
}

int dwfgpu_matpcdagmatpc_dev(dwfgpu_t h, double mass, dwfgpu_vec_t out, dwfgpu_vec_t in,
                             double* dot_prd) {
  if (!h || !out || !in || out->h != h || in->h != h) return DWFGPU_BAD_ARG;
  TRY(DISPATCH(h, mdagm(h, mass, out->d, in->d, dot_prd)));
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

int dwfgpu_invcg_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, dwfgpu_vec_t sol, dwfgpu_vec_t src,
                     double* true_res, int* iters) {
  if (!h || !arg || !sol || !src || sol->h != h || src->h != h) return DWFGPU_BAD_ARG;
  if (arg->max_num_iter < 1) return DWFGPU_BAD_ARG;
  TRY(ensure_work(h));
  return DISPATCH(h, invcg(h, arg, sol->d, src->d, true_res, iters));
}

int dwfgpu_matpcdagmatpc_cost(dwfgpu_t h, double* bytes, double* flops) {
  if (!h) return DWFGPU_BAD_ARG;
  const double nh = (double)h->g.vh * h->ls;
  const double v4 = 2.0 * h->g.vh;
  if (bytes)
    *bytes = DWF_MDAGM_SPINOR_REALS_PER_NH * nh * h->elem +
             DWF_MDAGM_LINK_REALS_PER_V4 * v4 * h->elem;
  if (flops)
    *flops = nh * (h->action == DWFGPU_DWF ? DWF_MDAGM_FLOPS_PER_NH_DWF
                                           : DWF_MDAGM_FLOPS_PER_NH_MOBIUS);
  return DWFGPU_OK;
}

int dwfgpu_time_copy(dwfgpu_t h, size_t bytes, int reps, double* gbs) {
  if (!h || reps < 1 || bytes < 16 || !gbs) return DWFGPU_BAD_ARG;
  const size_t n16 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, n16 * 16));
  if (hipMalloc(&b, n16 * 16) != hipSuccess) {
    (void)hipFree(a);
    set_err("dwfgpu_time_copy: hipMalloc failed");
    return DWFGPU_HIP_ERROR;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = DWFGPU_OK;
  if (hipMemsetAsync(a, 1, n16 * 16, h->st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    st = DWFGPU_HIP_ERROR;
  const dim3 grid(8192), block(256);
  if (!st) {  // one untimed pass to fault in the pages
    hipLaunchKernelGGL(dwfk::k_copy16, grid, block, 0, h->st, (uint4*)b, (const uint4*)a, n16);
    if (hipEventRecord(e0, h->st) != hipSuccess) st = DWFGPU_HIP_ERROR;
  }
  for (int r = 0; r < reps && !st; ++r)
    hipLaunchKernelGGL(dwfk::k_copy16, grid, block, 0, h->st, (uint4*)b, (const uint4*)a, n16);
  if (!st && (hipGetLastError() != hipSuccess || hipEventRecord(e1, h->st) != hipSuccess ||
              hipEventSynchronize(e1) != hipSuccess))
    st = DWFGPU_HIP_ERROR;
  float ms = 0;
  if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = DWFGPU_HIP_ERROR;
  if (!st) *gbs = 2.0 * n16 * 16 * reps / (ms * 1e-3) / 1e9;
  if (st) set_err("dwfgpu_time_copy: HIP failure");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return st;
}

int dwfgpu_time_matpcdagmatpc(dwfgpu_t h, double mass, dwfgpu_vec_t out, dwfgpu_vec_t in,
                              int reps, double* ms_app, double* ms_kernel, double* bytes_kernel,
                              int* nkernel) {
  if (!h || !out || !in || reps < 1 || out->h != h || in->h != h) return DWFGPU_BAD_ARG;
  return h->prec == DWFGPU_F64 ? time_mdagm<double>(h, mass, out->d, in->d, reps, ms_app, ms_kernel,
                                                    bytes_kernel, nkernel)
                               : time_mdagm<float>(h, mass, out->d, in->d, reps, ms_app, ms_kernel,
                                                   bytes_kernel, nkernel);
}

}  // extern "C"
