// dwf_gpu.hip -- host side of the MI355X domain-wall / Mobius library: the
// C ABI of include/dwf_gpu.h, device memory, kernel launches, the CG loop and
// the reliable-update mixed-precision CG.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <random>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "dwf_gpu.h"
#include "dwf_synth.h"
#include "dwf_kernels.hpp"
#include "ptc_kernels.hpp"
#include "roofline.h"

using dwfk::cx;
using dwfk::Geom;
using dwfk::M5Side;
using dwfk::MobArgs;
using dwfk::P32;
using dwfk::P64;
using dwfk::PBF16;
using dwfk::PI16;

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

constexpr int NPREC = 4;  // DWFGPU_F64, F32, BF16, I16

// a spinor half vector on the device in some storage precision
struct DVec {
  void* d = nullptr;
  float* sc = nullptr;  // per-(site, slice) scales of the 16-bit formats
};

// the operator's fields in one storage precision
struct PrecFields {
  bool ready = false;
  void* U = nullptr;  // links (18 reals in FP64, first two rows otherwise)
  DVec te, t1;        // opposite-parity temporary, MatPc output inside MdagM
  DVec work[6];       // solver vectors, allocated on first use
  std::vector<DVec> ring;  // CG directions of the last steps (mixed CG)
};

}  // namespace

struct dwfgpu_s {
  int action = 0, prec = 0;  // prec: storage precision of the API vectors
  int dev = 0;               // HIP device the handle was created on
  int dims[4] = {0, 0, 0, 0};
  int bc[4] = {0, 0, 0, 0};
  int ls = 0;
  double m5 = 0, a5inv = 1, b = 1, c = 0;
  Geom g{};
  size_t n = 0;  // reals per half vector
  hipStream_t st = nullptr;
  PrecFields f[NPREC];
  void* cps_stage = nullptr;  // FP64 buffer in CPS order (host-array API)
  double* stage4d = nullptr;  // one 4D CANONICAL fermion slice (format calls)
  int* nbt = nullptr;         // neighbour table of the hop kernels (g.nbt)
  double* partial = nullptr;
  int npartial = 0;
  double* dscal = nullptr;
  double* hscal = nullptr;   // pinned
  double* dalpha = nullptr;  // alphas of the pending sloppy CG steps (device)
  double* hring = nullptr;   // pinned copies of the sloppy CG scalars, per step
  int grid_hop = 0;
  int ns = 16;  // sites per work-group of the hopping kernels (x); y = all slices
  int max_lds = 65536;
  // launch shape of the many-RHS operator chosen by the autotuner, per
  // sloppy precision and RHS count (0: not tuned yet)
  int mr_ns[NPREC] = {0, 0, 0, 0}, mr_nr[NPREC] = {0, 0, 0, 0}, mr_nrhs[NPREC] = {0, 0, 0, 0};
  int mr_last_ns = 0, mr_last_nr = 0;
  struct PtcNet* net = nullptr;  // learned preconditioner (dwfgpu_ptc_*)
};

struct dwfgpu_vec_s {
  dwfgpu_s* h = nullptr;
  DVec v;
};

// nrhs FP64 half vectors in one allocation, RHS j at chunk offset j * n/2
struct dwfgpu_mvec_s {
  dwfgpu_s* h = nullptr;
  int nrhs = 0;
  DVec v;
};

namespace {

template <class P>
constexpr int pidx();
template <>
constexpr int pidx<P64>() { return DWFGPU_F64; }
template <>
constexpr int pidx<P32>() { return DWFGPU_F32; }
template <>
constexpr int pidx<PBF16>() { return DWFGPU_BF16; }
template <>
constexpr int pidx<PI16>() { return DWFGPU_I16; }

template <class P>
dwfk::Fld<P> fld(const DVec& v) {
  return dwfk::Fld<P>{(typename P::C*)v.d, v.sc};
}
template <class P>
dwfk::CFld<P> cfld(const DVec& v) {
  return dwfk::CFld<P>{(const typename P::C*)v.d, v.sc};
}

// bytes of one half vector / of all links (once) in policy P
template <class P>
double half_bytes(const dwfgpu_s* h) {
  return (double)h->n / 2 * sizeof(typename P::S) + (P::kScaled ? 4.0 * h->g.vh * h->ls : 0.0);
}
template <class P>
double link_bytes(const dwfgpu_s* h) {
  return 8.0 * h->g.vh * P::kLinkElems * sizeof(typename P::S);
}

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
  for (int s = 0; s < dwfk::MAXLS; ++s) {
    k.faf[s] = (float)k.af[s];
    k.fcf[s] = (float)k.cf[s];
    k.fab[s] = (float)k.ab[s];
    k.fcbk[s] = (float)k.cbk[s];
  }
  k.finv_f = (float)k.inv_f;
  k.finv_b = (float)k.inv_b;
  return k;
}

MobArgs mob_args(const dwfgpu_s* h, double mass, int dag) {
  MobArgs a;
  a.active = nullptr;
  const double kb = mob_kappa_b(h), kc = mob_kappa_c(h);
  a.b = h->b;
  a.c = h->c;
  a.mc = -mass * h->c;
  a.K = kb / kc;
  a.mK = -mass * a.K;
  a.kappa_b = kb;
  a.fb = (float)a.b;
  a.fc = (float)a.c;
  a.fmc = (float)a.mc;
  a.fK = (float)a.K;
  a.fmK = (float)a.mK;
  a.fkappa_b = (float)a.kappa_b;
  a.m5 = m5_side(h, mass, dag);
  return a;
}

// LDS of the s-column buffer (M5^-1 and (b + c D5)^dag epilogues)
size_t col_lds_bytes(const dwfgpu_s* h, size_t creal) {
  return (size_t)h->ns * (h->ls * 12 + 1) * 2 * creal;
}
// LDS of the staged links and neighbour indices: nsets parities x 8
// directions x ns sites x (9 complex + 1 int)
size_t link_lds_bytes(const dwfgpu_s* h, int nsets, size_t creal) {
  return (size_t)nsets * 8 * h->ns * (9 * 2 * creal + 4);
}
// largest LDS request of any kernel of this handle (FP64 compute is the widest)
size_t max_lds_bytes(const dwfgpu_s* h) {
  return h->action == DWFGPU_DWF ? link_lds_bytes(h, 2, 8)
                                 : std::max(link_lds_bytes(h, 1, 8), col_lds_bytes(h, 8));
}

// ---------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------
int alloc(dwfgpu_s* h, void** p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes));
  HIPCHK(hipMemsetAsync(*p, 0, bytes, h->st));
  return DWFGPU_OK;
}

template <class P>
int alloc_vec(dwfgpu_s* h, DVec* v) {
  TRY(alloc(h, &v->d, h->n / 2 * sizeof(typename P::S)));
  if (P::kScaled) TRY(alloc(h, (void**)&v->sc, (size_t)h->g.vh * h->ls * sizeof(float)));
  return DWFGPU_OK;
}

void free_vec(DVec* v) {
  if (v->d) (void)hipFree(v->d);
  if (v->sc) (void)hipFree(v->sc);
  v->d = nullptr;
  v->sc = nullptr;
}

dim3 blas_grid() { return dim3(dwfk::BLAS_BLOCKS); }
dim3 blas_block() { return dim3(dwfk::BLAS_THREADS); }

// the fields of precision P: links derived from the FP64 links on the device
template <class P>
int ensure_fields(dwfgpu_s* h) {
  PrecFields& F = h->f[pidx<P>()];
  if (F.ready) return DWFGPU_OK;
  if (pidx<P>() != DWFGPU_F64) {
    TRY(alloc(h, &F.U, (size_t)link_bytes<P>(h)));
    hipLaunchKernelGGL((dwfk::k_compress_links<P>), dim3(2048), dim3(256), 0, h->st,
                       (typename P::S*)F.U, (const cx<double>*)h->f[DWFGPU_F64].U, h->g);
    HIPCHK(hipGetLastError());
  }
  TRY(alloc_vec<P>(h, &F.te));
  TRY(alloc_vec<P>(h, &F.t1));
  F.ready = true;
  return DWFGPU_OK;
}

template <class P>
int ensure_work(dwfgpu_s* h, int count) {
  TRY(ensure_fields<P>(h));
  PrecFields& F = h->f[pidx<P>()];
  for (int k = 0; k < count; ++k)
    if (!F.work[k].d) TRY(alloc_vec<P>(h, &F.work[k]));
  return DWFGPU_OK;
}

template <class P>
int ensure_ring(dwfgpu_s* h, int count) {
  TRY(ensure_fields<P>(h));
  std::vector<DVec>& ring = h->f[pidx<P>()].ring;
  if ((int)ring.size() < count) ring.resize(count);
  for (int k = 0; k < count; ++k)
    if (!ring[k].d) TRY(alloc_vec<P>(h, &ring[k]));
  return DWFGPU_OK;
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// sum of nparts partials into the device scalar dscal[slot] (no host sync)
int reduce_to(dwfgpu_s* h, int nparts, int slot) {
  hipLaunchKernelGGL(dwfk::k_reduce_partials, dim3(1), dim3(1024), 0, h->st, h->partial, nparts,
                     h->dscal + slot);
  HIPCHK(hipGetLastError());
  return DWFGPU_OK;
}

// copy dscal[0..n) to the host and wait
int fetch_scalars(dwfgpu_s* h, int n, double* out) {
  HIPCHK(hipMemcpyAsync(h->hscal, h->dscal, n * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  for (int k = 0; k < n; ++k) out[k] = h->hscal[k];
  return DWFGPU_OK;
}

// sum of nparts partials in `partial` through the device scalar `dscal`, to
// the host (waits)
int fetch_sum_to(dwfgpu_s* h, const double* partial, int nparts, double* dscal, double* out) {
  hipLaunchKernelGGL(dwfk::k_reduce_partials, dim3(1), dim3(1024), 0, h->st, partial, nparts,
                     dscal);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h->hscal, dscal, sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  *out = h->hscal[0];
  return DWFGPU_OK;
}
int fetch_sum(dwfgpu_s* h, int nparts, double* out) {
  return fetch_sum_to(h, h->partial, nparts, h->dscal, out);
}

// the temporaries of one MatPcDagMatPc: the opposite-parity field te, the
// MatPc output t1, the per-work-group partial sums and one device scalar
struct Scratch {
  DVec te, t1;
  double* partial;
  double* dscal;
};

// ---------------------------------------------------------------------------
// operator pieces for storage policy P
// ---------------------------------------------------------------------------
template <class P>
struct Ops {
  typedef typename P::R R;
  typedef typename P::S S;
  static PrecFields& F(dwfgpu_s* h) { return h->f[pidx<P>()]; }

  // DWF hop kernel: epi 0 -> out = D in ; epi 1 -> out = x0 + alpha D in
  static int dwf(dwfgpu_s* h, double mass, const DVec& out, const DVec& in, int cb_in, int dag,
                 int epi, const DVec& x0, double alpha, bool norm, double* partial = nullptr) {
    if (!partial) partial = h->partial;
    const dim3 grid(h->grid_hop), block(h->ns, h->ls);
    const double two_a5 = 2.0 * h->a5inv, nm = -2.0 * mass * h->a5inv;
    const S* u = (const S*)F(h).U;
    const size_t lds = link_lds_bytes(h, 2, sizeof(R));
    Geom gm = h->g;
    gm.rev = epi;  // the x0 + alpha D passes sweep backwards, as the Schur ones in mob()
#define DWF_LAUNCH(DAG, EPI, NORM)                                                           \
  hipLaunchKernelGGL((dwfk::k_dwf_dslash<P, DAG, EPI, NORM>), grid, block, lds, h->st,       \
                     fld<P>(out), cfld<P>(in), u, cfld<P>(x0), partial, gm, cb_in, two_a5,    \
                     nm, alpha, dwfk::CgEpi<P, P>{})
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
  static int mob(dwfgpu_s* h, double mass, const DVec& out, const DVec& in, const DVec& psi,
                 int p_out, int dag, int epi, int fl, double* partial = nullptr,
                 const int* active = nullptr) {
    if (!partial) partial = h->partial;
    const dim3 grid(h->grid_hop), block(h->ns, h->ls);
    MobArgs a = mob_args(h, mass, dag);
    a.active = active;
    // the Schur passes sweep backwards: they start on the sites the pass
    // before them wrote last, which the last-level cache still holds
    Geom gm = h->g;
    gm.rev = epi == dwfk::EPI_SCHUR ? 1 : 0;
    const S* u = (const S*)F(h).U;
    const bool col = dag || epi == dwfk::EPI_M5INV || epi == dwfk::EPI_SCHUR;
    // exchange slots after the links; the M5^-1 / Schur s-columns reuse the
    // LDS from offset 0 once the hop is done
    const size_t split_lds =
        std::max(col ? col_lds_bytes(h, sizeof(R)) : (size_t)0,
                 dwfk::split_col_offset<R>(h->ns) +
                     (size_t)h->ls * h->ns * dwfk::TX_SLOT * 2 * sizeof(R));
    const bool split = dwfk::MobSplitT<P>::value && !dag && !(fl & dwfk::FL_PRE_T) &&
                       split_lds <= (size_t)h->max_lds;
    const size_t lds = split ? split_lds
                             : std::max(link_lds_bytes(h, 1, sizeof(R)),
                                        col ? col_lds_bytes(h, sizeof(R)) : (size_t)0);
#define MOB_LAUNCH(DAG, EPI, FL)                                                              \
  hipLaunchKernelGGL((dwfk::k_mobius<P, DAG, EPI, FL>), grid, block, lds, h->st, fld<P>(out), \
                     cfld<P>(in), cfld<P>(psi), u, partial, gm, p_out, a, dwfk::CgEpi<P, P>{})
    using dwfk::FL_NORM;
    using dwfk::FL_POST_T;
    using dwfk::FL_PRE_T;
    if (split) {
      if constexpr (dwfk::MobSplitT<P>::value) {
        using dwfk::FL_SPLIT_T;
        if (epi == dwfk::EPI_M5INV) {
          if (fl & FL_POST_T) MOB_LAUNCH(0, dwfk::EPI_M5INV, FL_POST_T | FL_SPLIT_T);
          else MOB_LAUNCH(0, dwfk::EPI_M5INV, FL_SPLIT_T);
        } else if (epi == dwfk::EPI_SCHUR) {
          if (fl & FL_NORM) MOB_LAUNCH(0, dwfk::EPI_SCHUR, FL_NORM | FL_SPLIT_T);
          else MOB_LAUNCH(0, dwfk::EPI_SCHUR, FL_SPLIT_T);
        } else {
          MOB_LAUNCH(0, dwfk::EPI_PLAIN, FL_SPLIT_T);
        }
      }
    } else if (dag == 0 && epi == dwfk::EPI_M5INV) {
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
  // te (for dag 0 already multiplied by (b + c D5), consumed by the second)
  static int mob_first(dwfgpu_s* h, double mass, const DVec& te, const DVec& in, int dag,
                       const int* active = nullptr) {
    return mob(h, mass, te, in, DVec(), 0, dag, dwfk::EPI_M5INV, dag ? 0 : dwfk::FL_POST_T,
               nullptr, active);
  }
  static int mob_second(dwfgpu_s* h, double mass, const DVec& out, const DVec& te,
                        const DVec& in, int dag, bool norm, double* partial = nullptr,
                        const int* active = nullptr) {
    const int fl = (dag ? 0 : dwfk::FL_PRE_T) | (norm && !dag ? dwfk::FL_NORM : 0);
    return mob(h, mass, out, te, in, 1, dag, dwfk::EPI_SCHUR, fl, partial, active);
  }
  // the last pass of a Mobius CG step: dagger hop of te, (b + c D5^dag) and
  // the Schur term with t1, then r -= alpha A p and |r|^2 partials (and
  // x += alpha p unless PX = NoX), alpha from scal[0] / scal[1]
  template <class PX>
  static int mob_cg_last(dwfgpu_s* h, double mass, const DVec& te, const DVec& t1,
                         const DVec& r, const DVec& p, const DVec& x, const double* scal,
                         double* partial, const int* active) {
    const dim3 grid(h->grid_hop), block(h->ns, h->ls);
    const dwfk::CgEpi<P, PX> cg{fld<P>(r), cfld<P>(p), fld<PX>(x), scal};
    MobArgs a = mob_args(h, mass, 1);
    a.active = active;
    const size_t lds = std::max(link_lds_bytes(h, 1, sizeof(R)), col_lds_bytes(h, sizeof(R)));
    Geom gm = h->g;
    gm.rev = 1;  // a Schur pass: backwards, as in mob()
    hipLaunchKernelGGL((dwfk::k_mobius<P, 1, dwfk::EPI_SCHUR, dwfk::FL_CG, PX>), grid, block,
                       lds, h->st, fld<P>(te), cfld<P>(te), cfld<P>(t1), (const S*)F(h).U,
                       partial, gm, 1, a, cg);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  static int m5inv(dwfgpu_s* h, double mass, const DVec& out, const DVec& in, int dag) {
    const dim3 grid(h->grid_hop), block(h->ns, h->ls);
    const M5Side k = m5_side(h, mass, dag);
    const size_t lds = col_lds_bytes(h, sizeof(R));
    if (dag)
      hipLaunchKernelGGL((dwfk::k_m5inv<P, 1>), grid, block, lds, h->st, fld<P>(out), cfld<P>(in),
                         h->g, k);
    else
      hipLaunchKernelGGL((dwfk::k_m5inv<P, 0>), grid, block, lds, h->st, fld<P>(out), cfld<P>(in),
                         h->g, k);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  // the handle's own temporaries of this precision
  static Scratch own(dwfgpu_s* h) { return Scratch{F(h).te, F(h).t1, h->partial, h->dscal}; }

  // MatPc / MatPcDag: out = M in (out on the CG parity).  norm -> partials.
  static int matpc_w(dwfgpu_s* h, double mass, const DVec& out, const DVec& in, int dag,
                     bool norm, const Scratch& w) {
    if (h->action == DWFGPU_DWF) {
      const double mk2 = -dwf_kappa(h) * dwf_kappa(h);
      TRY(dwf(h, mass, w.te, in, 1, dag, 0, DVec(), 0.0, false, w.partial));
      TRY(dwf(h, mass, out, w.te, 0, dag, 1, in, mk2, norm, w.partial));
    } else {
      TRY(mob_first(h, mass, w.te, in, dag));
      TRY(mob_second(h, mass, out, w.te, in, dag, norm, w.partial));
    }
    return DWFGPU_OK;
  }
  static int matpc(dwfgpu_s* h, double mass, const DVec& out, const DVec& in, int dag,
                   bool norm) {
    return matpc_w(h, mass, out, in, dag, norm, own(h));
  }

  // MatPcDagMatPc; *dot (if requested) = |MatPc in|^2 (dwf_mdagm.C:49-55)
  static int mdagm_w(dwfgpu_s* h, double mass, const DVec& out, const DVec& in, double* dot,
                     const Scratch& w) {
    TRY(matpc_w(h, mass, w.t1, in, 0, dot != nullptr, w));
    if (dot) TRY(fetch_sum_to(h, w.partial, h->grid_hop, w.dscal, dot));
    TRY(matpc_w(h, mass, out, w.t1, 1, false, w));
    return DWFGPU_OK;
  }
  static int mdagm(dwfgpu_s* h, double mass, const DVec& out, const DVec& in, double* dot) {
    return mdagm_w(h, mass, out, in, dot, own(h));
  }

  // One CG step around MatPcDagMatPc (inv_cg.C:318-349) with the vector
  // updates fused into the last kernel: with rr = dscal[0] on entry,
  // d = <p, A p> -> dscal[1], a = rr / d, x += a p, r -= a A p, and
  // |r|^2 -> dscal[0] (mdagm_cg; *rr_new, *d receive the two scalars).
  // mdagm_cg_launch issues the kernels only: the |r|^2 partials are left in
  // h->partial (grid_hop of them) and <p, A p> in dscal[1].
  template <class PX>
  static int mdagm_cg_launch(dwfgpu_s* h, double mass, const DVec& p, const DVec& r,
                             const DVec& x) {
    const DVec& te = F(h).te;
    const DVec& t1 = F(h).t1;
    const dwfk::CgEpi<P, PX> cg{fld<P>(r), cfld<P>(p), fld<PX>(x), h->dscal};
    const dim3 grid(h->grid_hop), block(h->ns, h->ls);
    const S* u = (const S*)F(h).U;
    if (h->action == DWFGPU_DWF) {
      const double mk2 = -dwf_kappa(h) * dwf_kappa(h);
      TRY(dwf(h, mass, te, p, 1, 0, 0, DVec(), 0.0, false));
      TRY(dwf(h, mass, t1, te, 0, 0, 1, p, mk2, true));
      TRY(reduce_to(h, h->grid_hop, 1));
      TRY(dwf(h, mass, te, t1, 1, 1, 0, DVec(), 0.0, false));
      Geom gm = h->g;
      gm.rev = 1;  // an x0 + alpha D pass: backwards, as in dwf()
      hipLaunchKernelGGL((dwfk::k_dwf_dslash<P, 1, 1, false, true, PX>), grid, block,
                         link_lds_bytes(h, 2, sizeof(R)), h->st, fld<P>(te), cfld<P>(te), u,
                         cfld<P>(t1), h->partial, gm, 0, 2.0 * h->a5inv,
                         -2.0 * mass * h->a5inv, mk2, cg);
    } else {
      TRY(mob_first(h, mass, te, p, 0));
      TRY(mob_second(h, mass, t1, te, p, 0, true));
      TRY(reduce_to(h, h->grid_hop, 1));
      TRY(mob_first(h, mass, te, t1, 1));
      TRY(mob_cg_last<PX>(h, mass, te, t1, r, p, x, h->dscal, h->partial, nullptr));
    }
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  template <class PX = P>
  static int mdagm_cg(dwfgpu_s* h, double mass, const DVec& p, const DVec& r, const DVec& x,
                      double* rr_new, double* d) {
    TRY(mdagm_cg_launch<PX>(h, mass, p, r, x));
    TRY(reduce_to(h, h->grid_hop, 0));
    double sc[2];
    TRY(fetch_scalars(h, 2, sc));
    *rr_new = sc[0];
    *d = sc[1];
    return DWFGPU_OK;
  }

  static int dslash(dwfgpu_s* h, double mass, const DVec& out, const DVec& in, int cb, int dag) {
    if (h->action == DWFGPU_DWF) return dwf(h, mass, out, in, cb, dag, 0, DVec(), 0.0, false);
    return mob(h, mass, out, in, DVec(), 1 - cb, dag, dwfk::EPI_PLAIN, 0);
  }

  static int norm(dwfgpu_s* h, const DVec& x, double* out) {
    hipLaunchKernelGGL((dwfk::k_norm<P>), blas_grid(), blas_block(), 0, h->st, cfld<P>(x), h->g,
                       h->partial);
    HIPCHK(hipGetLastError());
    return fetch_sum(h, dwfk::BLAS_BLOCKS, out);
  }

  // z = a x + y (with |z|^2 partials when norm)
  static int axpy_to(dwfgpu_s* h, const DVec& z, const DVec& x, const DVec& y, double a,
                     bool norm) {
    if (norm)
      hipLaunchKernelGGL((dwfk::k_axpy_to<P, true>), blas_grid(), blas_block(), 0, h->st, fld<P>(z),
                         cfld<P>(x), cfld<P>(y), a, h->g, h->partial);
    else
      hipLaunchKernelGGL((dwfk::k_axpy_to<P, false>), blas_grid(), blas_block(), 0, h->st,
                         fld<P>(z), cfld<P>(x), cfld<P>(y), a, h->g, h->partial);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  // y = a x + b y
  static int axpby(dwfgpu_s* h, const DVec& y, const DVec& x, double a, double b) {
    hipLaunchKernelGGL((dwfk::k_axpby<P>), blas_grid(), blas_block(), 0, h->st, fld<P>(y),
                       cfld<P>(x), a, b, h->g);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  static int scale(dwfgpu_s* h, const DVec& x, double a) {
    hipLaunchKernelGGL((dwfk::k_scale<P>), blas_grid(), blas_block(), 0, h->st, fld<P>(x), a, h->g);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  static int copy(dwfgpu_s* h, const DVec& dst, const DVec& src) {
    HIPCHK(hipMemcpyAsync(dst.d, src.d, h->n / 2 * sizeof(S), hipMemcpyDeviceToDevice, h->st));
    if (P::kScaled)
      HIPCHK(hipMemcpyAsync(dst.sc, src.sc, (size_t)h->g.vh * h->ls * sizeof(float),
                            hipMemcpyDeviceToDevice, h->st));
    return DWFGPU_OK;
  }

  // DiracOp::InvCgShift(out, in, src_norm_sq, true_res, shift)
  // (inv_cg.C:84-448).  src_norm_sq == 0: |src|^2 is computed (:166-169).
  // shift != NULL: the solve is (MatPcDagMatPc + *shift) sol = src, with
  // A p + shift p formed after the operator and d = <p, A p> + shift |p|^2
  // (:248-251, 304-310), so its steps run unfused (invcg_shifted_step).
  static int invcg(dwfgpu_s* h, const dwfgpu_cg_arg* arg, const DVec& sol, const DVec& src,
                   double* true_res, int* iters, double src_norm_sq = 0.0,
                   const double* shift = nullptr) {
    TRY(ensure_work<P>(h, 3));
    const DVec& res = F(h).work[0];
    const DVec& dir = F(h).work[1];
    const DVec& mmp = F(h).work[2];
    const double mass = arg->mass;
    if (src_norm_sq == 0.0) TRY(norm(h, src, &src_norm_sq));
    const double stp_cnd = src_norm_sq * arg->stop_rsd * arg->stop_rsd;
    TRY(mdagm(h, mass, mmp, sol, nullptr));
    if (shift) TRY(axpby(h, mmp, sol, *shift, 1.0));  // mmp = shift sol + mmp
    TRY(axpy_to(h, res, mmp, src, -1.0, true));  // res = src - mmp
    double rr;
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
    TRY(copy(h, dir, res));
    int itr = 0;
    int max_itr = arg->max_num_iter - 1;
    if (rr <= stp_cnd) max_itr = 0;
    int status = DWFGPU_OK;
    // dscal[0] holds rr (fetch_sum above); each step replaces it by the new
    // rr.  The hop epilogue updates res only; sol += a dir and dir = res + b
    // dir follow in one pass (k_cg_xp), which reads dir once for both
    for (int k = 0; k < max_itr; ++k) {
      itr = itr + 1;
      const double rr_prv = rr;
      double d;
      if (shift) {
        bool stop = false;
        TRY(invcg_shifted_step(h, mass, *shift, sol, dir, res, mmp, stp_cnd, &rr, &d, &stop));
        if (d == 0.0) {
          status = DWFGPU_BREAKDOWN;
          break;
        }
        if (stop) break;
        continue;
      }
      TRY(mdagm_cg<dwfk::NoX>(h, mass, dir, res, DVec(), &rr, &d));
      if (d == 0.0) {  // a = 0 was used: res is unchanged
        status = DWFGPU_BREAKDOWN;
        break;
      }
      const double a = rr_prv / d;  // the epilogue's alpha (cg_alpha)
      if (rr <= stp_cnd) {
        hipLaunchKernelGGL((dwfk::k_cg_xp<P, P, false>), blas_grid(), blas_block(), 0, h->st,
                           fld<P>(sol), fld<P>(dir), cfld<P>(res), a, 0.0, h->g);
        HIPCHK(hipGetLastError());
        break;
      }
      const double b = rr / rr_prv;
      hipLaunchKernelGGL((dwfk::k_cg_xp<P, P, true>), blas_grid(), blas_block(), 0, h->st,
                         fld<P>(sol), fld<P>(dir), cfld<P>(res), a, b, h->g);
      HIPCHK(hipGetLastError());
    }
    // true residual |src - A sol| / |src|  (inv_cg.C:403-415)
    TRY(mdagm(h, mass, mmp, sol, nullptr));
    if (shift) TRY(axpby(h, mmp, sol, *shift, 1.0));
    TRY(axpy_to(h, res, mmp, src, -1.0, true));
    double tr;
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &tr));
    if (true_res) *true_res = std::sqrt(tr / src_norm_sq);
    if (iters) *iters = itr + 1;
    if (status == DWFGPU_OK && itr == arg->max_num_iter - 1) status = DWFGPU_MAXITER;
    return status;
  }

  // one step of the shifted CG in CPS's order (inv_cg.C:293-363): mmp = A dir
  // with d = |MatPc dir|^2, mmp += shift dir, d += shift |dir|^2, then (d != 0)
  // a = rr / d, sol += a dir, res -= a mmp, rr = |res|^2 and, unless rr meets
  // the stop rule (*stop), dir = b dir + res.
  static int invcg_shifted_step(dwfgpu_s* h, double mass, double shift, const DVec& sol,
                                const DVec& dir, const DVec& res, const DVec& mmp,
                                double stp_cnd, double* rr, double* d, bool* stop) {
    const double rr_prv = *rr;
    TRY(mdagm(h, mass, mmp, dir, d));
    TRY(axpby(h, mmp, dir, shift, 1.0));
    double dir_sq;
    TRY(norm(h, dir, &dir_sq));
    *d += shift * dir_sq;
    if (*d == 0.0) return DWFGPU_OK;
    const double a = rr_prv / *d;
    TRY(axpby(h, sol, dir, a, 1.0));
    TRY(axpy_to(h, res, mmp, res, -a, true));
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, rr));
    if (*rr <= stp_cnd) {
      *stop = true;
      return DWFGPU_OK;
    }
    return axpby(h, dir, res, 1.0, *rr / rr_prv);
  }

  // DiracOpDwf::MatInv (d_op_dwf.C:316-507) / DiracOpMobius::MatInv ORIG
  // (d_op_mobius.C:407-696) on the two device halves (first, second) of the
  // full vectors in CPS order.
  // MatInv's source of the CG (d_op_dwf.C:366-419; d_op_mobius.C:476-486 and
  // :546, ORIG pc) from the full vector [in_first | in_second] into src; uses
  // work[3] and work[5] (ensure_work(h, 6) first)
  static int matinv_src(dwfgpu_s* h, double mass, const DVec& src, const DVec& in_first,
                        const DVec& in_second) {
    const DVec& temp = F(h).work[3];
    if (h->action == DWFGPU_DWF) {
      // [odd | even]: MatPcDag(in_odd + kappa D in_even)
      TRY(dwf(h, mass, temp, in_second, 0, 0, 1, in_first, dwf_kappa(h), false));
    } else {
      // [even | odd]: MatPcDag(in_odd + kappa_b D4 M5^-1 in_even)
      const DVec& temp2 = F(h).work[5];
      TRY(m5inv(h, mass, temp, in_first, 0));
      TRY(mob(h, mass, temp2, temp, DVec(), 1, 0, dwfk::EPI_PLAIN, 0));
      TRY(axpy_to(h, temp, temp2, in_second, mob_kappa_b(h), false));
    }
    return matpc(h, mass, src, temp, 1, false);
  }
  // the half of the solution the CG did not give (DWF: even from odd;
  // Mobius: even from odd, then both halves times 2 kappa_b), d_op_dwf.C:
  // 489-494, d_op_mobius.C:648-677; uses work[3]
  static int matinv_sol(dwfgpu_s* h, double mass, const DVec& out_first,
                        const DVec& out_second, const DVec& in_first, const DVec& in_second) {
    const DVec& temp = F(h).work[3];
    if (h->action == DWFGPU_DWF)
      return dwf(h, mass, out_second, out_first, 1, 0, 1, in_second, dwf_kappa(h), false);
    const double kb = mob_kappa_b(h);
    TRY(mob(h, mass, temp, out_second, DVec(), 0, 0, dwfk::EPI_PLAIN, 0));
    TRY(scale(h, temp, kb));
    TRY(m5inv(h, mass, out_first, temp, 0));
    TRY(m5inv(h, mass, temp, in_first, 0));
    TRY(axpby(h, out_first, temp, 1.0, 1.0));
    TRY(scale(h, out_first, 2.0 * kb));
    return scale(h, out_second, 2.0 * kb);
  }
  // the CG half of a full vector: DWF [odd | even] -> first, Mobius
  // [even | odd] -> second
  static const DVec& cg_half(const dwfgpu_s* h, const DVec& first, const DVec& second) {
    return h->action == DWFGPU_DWF ? first : second;
  }

  static int matinv(dwfgpu_s* h, const dwfgpu_cg_arg* arg, const DVec& out_first,
                    const DVec& out_second, const DVec& in_first, const DVec& in_second,
                    double* true_res, int* iters) {
    TRY(ensure_work<P>(h, 6));
    const DVec& src = F(h).work[4];
    TRY(matinv_src(h, arg->mass, src, in_first, in_second));
    const int st = invcg(h, arg, cg_half(h, out_first, out_second), src, true_res, iters);
    if (st != DWFGPU_OK && st != DWFGPU_MAXITER) return st;
    TRY(matinv_sol(h, arg->mass, out_first, out_second, in_first, in_second));
    return st;
  }

  static int to_device(dwfgpu_s* h, const DVec& dst, const double* host) {
    HIPCHK(hipMemcpyAsync(h->cps_stage, host, h->n * sizeof(double), hipMemcpyHostToDevice,
                          h->st));
    hipLaunchKernelGGL((dwfk::k_cps_to_field<P>), dim3(2048), dim3(256), 0, h->st, fld<P>(dst),
                       (const double*)h->cps_stage, h->g);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }

  static int to_host(dwfgpu_s* h, double* host, const DVec& src) {
    hipLaunchKernelGGL((dwfk::k_field_to_cps<P>), dim3(2048), dim3(256), 0, h->st,
                       (double*)h->cps_stage, cfld<P>(src), h->g);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, h->cps_stage, h->n * sizeof(double), hipMemcpyDeviceToHost,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return DWFGPU_OK;
  }

  // one 4D CANONICAL slice on the device <-> slice s of the halves of a full
  // vector (up / lo: which spin pairs are taken; canon_to_halves zeroes the rest)
  static int canon_to_halves(dwfgpu_s* h, const DVec& first, const DVec& second,
                             const double* canon, int s, int up, int lo) {
    hipLaunchKernelGGL((dwfk::k_canon_to_halves<P>), dim3(1024), dim3(256), 0, h->st,
                       fld<P>(first), fld<P>(second), canon, h->g, s,
                       h->action == DWFGPU_DWF ? 1 : 0, up, lo);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }
  static int halves_to_canon(dwfgpu_s* h, double* canon, const DVec& first, const DVec& second,
                             int s, int up, int lo) {
    hipLaunchKernelGGL((dwfk::k_halves_to_canon<P>), dim3(1024), dim3(256), 0, h->st, canon,
                       cfld<P>(first), cfld<P>(second), h->g, s,
                       h->action == DWFGPU_DWF ? 1 : 0, up, lo);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  }
};

// FP64-reliable-update CG with the bulk of MatPcDagMatPc in sloppy policy PS.
// High precision (FP64): sol, src, the true residual r and the solution
// increments; sloppy: r_s, p_s, A p_s; the running correction x_s is FP32.
// A reliable update (x += x_s, r = src - A x in FP64, r_s = r) happens when
// |r_s| falls below delta times |r| at the previous update, and whenever the
// sloppy residual meets the stop condition; the solve ends only when the
// FP64 true residual meets it (stop rule of inv_cg.C:174,354).
// The correction x_s is kept in FP32 for every sloppy policy: an int16 x_s
// (per-site scale) rounds each step's increment to about 2^-15 of the site's
// accumulated correction, which on the 24^3x64 Mobius system raised the
// iteration count from 525 to 661.
template <class PS>
int invcg_mixed(dwfgpu_s* h, const dwfgpu_cg_arg* arg, double delta, const DVec& sol,
                const DVec& src, double* true_res, int* iters, int* nreliable) {
  typedef Ops<P64> H;
  typedef Ops<PS> L;
  typedef P32 XS;
  // the directions of the last KRING steps stay in a ring of sloppy vectors;
  // x_s += a_j p_j for all of them is one pass over x_s (k_x_flush) instead
  // of one per step: the FP32 x_s moves through HBM once every KRING steps
  constexpr int KRING = 16;
  TRY(ensure_work<P64>(h, 2));
  TRY(ensure_work<PS>(h, 1));
  TRY(ensure_work<P32>(h, 4));
  TRY(ensure_ring<PS>(h, KRING));
  const DVec& r = h->f[DWFGPU_F64].work[0];
  const DVec& ax = h->f[DWFGPU_F64].work[1];
  const DVec& rs = h->f[pidx<PS>()].work[0];
  const std::vector<DVec>& ring = h->f[pidx<PS>()].ring;
  const DVec& xs = h->f[DWFGPU_F32].work[3];
  const double mass = arg->mass;
  double src_norm_sq;
  TRY(H::norm(h, src, &src_norm_sq));
  const double stp_cnd = src_norm_sq * arg->stop_rsd * arg->stop_rsd;
  // r = src - A sol ; r_s = p_s = r ; x_s = 0
  TRY(H::mdagm(h, mass, ax, sol, nullptr));
  TRY(H::axpy_to(h, r, ax, src, -1.0, true));
  double rr;
  TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
  hipLaunchKernelGGL((dwfk::k_convert<PS, P64>), blas_grid(), blas_block(), 0, h->st, fld<PS>(rs),
                     cfld<P64>(r), h->g);
  HIPCHK(hipGetLastError());
  int slot = 0;
  TRY(L::copy(h, ring[0], rs));
  HIPCHK(hipMemsetAsync(xs.d, 0, h->n / 2 * sizeof(typename XS::S), h->st));
  dwfk::XFlush<PS> xf;
  xf.cnt = 0;
  xf.alpha = h->dalpha;
  auto flush = [&]() -> int {
    if (xf.cnt == 0) return DWFGPU_OK;
    hipLaunchKernelGGL((dwfk::k_x_flush<PS, XS>), blas_grid(), blas_block(), 0, h->st,
                       fld<XS>(xs), xf, h->g);
    HIPCHK(hipGetLastError());
    xf.cnt = 0;
    return DWFGPU_OK;
  };
  // The host examines the scalars of step k while the GPU runs steps k+1 and
  // k+2 (alpha, beta and the x_s updates take their scalars on the device),
  // so the GPU never waits for the host between steps.  A reliable update is
  // therefore taken LAG steps after its condition first holds, with the
  // state of the current step.  BF16 keeps the test of every step (LAG 0):
  // its recurrence loses conjugacy between updates, and two-step-late
  // updates more than doubled its iteration count at 24^3x64 (3178 -> 7019).
  constexpr int LAG = std::is_same<PS, PBF16>::value ? 0 : 2, NEV = 4;
  struct Events {
    hipEvent_t e[NEV] = {};
    ~Events() {
      for (auto& x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  for (auto& x : ev.e) HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  double rr_update = rr;  // |r|^2 at the last reliable update
  int itr = 0, nrel = 0, status = DWFGPU_OK;
  int fresh_from = 1;  // steps before this one were covered by a reliable update
  const int max_itr = rr <= stp_cnd ? 0 : arg->max_num_iter - 1;
  bool converged = rr <= stp_cnd;
  while (!converged && itr < max_itr) {
    itr = itr + 1;
    const DVec& ps = ring[slot];
    // dscal[0] holds |r_s|^2; the fused step updates r_s, and
    // k_reduce_cg_step the scalars and the step's alpha (x_s += a p_s pending)
    TRY(L::template mdagm_cg_launch<dwfk::NoX>(h, mass, ps, rs, DVec()));
    hipLaunchKernelGGL(dwfk::k_reduce_cg_step, dim3(1), dim3(1024), 0, h->st, h->partial,
                       h->grid_hop, h->dscal, h->dalpha + slot);
    HIPCHK(hipGetLastError());
    double* hs = h->hring + 4 * (itr % NEV);
    HIPCHK(hipMemcpyAsync(hs, h->dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipEventRecord(ev.e[itr % NEV], h->st));
    xf.p[xf.cnt] = (const typename PS::C*)ps.d;
    xf.sc[xf.cnt] = ps.sc;
    xf.ai[xf.cnt] = slot;
    xf.cnt++;
    // the scalars of step itr - LAG (or of this step once max_itr is reached)
    const int k = itr == max_itr ? itr : itr - LAG;
    bool update = false;
    if (k >= fresh_from) {
      HIPCHK(hipEventSynchronize(ev.e[k % NEV]));
      const double rr_k = h->hring[4 * (k % NEV)], d_k = h->hring[4 * (k % NEV) + 1];
      if (d_k == 0.0) {
        status = DWFGPU_BREAKDOWN;
        break;
      }
      update = rr_k <= stp_cnd || rr_k < delta * delta * rr_update;
    }
    if (update) {
      // reliable update in FP64 (after step itr)
      TRY(flush());
      hipLaunchKernelGGL((dwfk::k_accumulate<XS>), blas_grid(), blas_block(), 0, h->st,
                         fld<P64>(sol), fld<XS>(xs), h->g);
      HIPCHK(hipGetLastError());
      TRY(H::mdagm(h, mass, ax, sol, nullptr));
      TRY(H::axpy_to(h, r, ax, src, -1.0, true));
      double rr_new;
      TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr_new));  // into dscal[0]: beta uses it
      hipLaunchKernelGGL((dwfk::k_convert<PS, P64>), blas_grid(), blas_block(), 0, h->st,
                         fld<PS>(rs), cfld<P64>(r), h->g);
      HIPCHK(hipGetLastError());
      rr_update = rr_new;
      ++nrel;
      fresh_from = itr + 1;
      if (rr_new <= stp_cnd) {
        rr = rr_new;
        converged = true;
        break;
      }
    }
    // p_s(next) = r_s + b p_s into the next ring slot (b from the device
    // scalars); when every slot holds a pending direction the oldest is
    // about to be overwritten: flush first
    const int next = (slot + 1) % KRING;
    if (xf.cnt == KRING) TRY(flush());
    hipLaunchKernelGGL((dwfk::k_pnext<PS>), blas_grid(), blas_block(), 0, h->st, fld<PS>(ring[next]),
                       cfld<PS>(rs), cfld<PS>(ps), h->dscal, h->g);
    HIPCHK(hipGetLastError());
    slot = next;
  }
  if (!converged && status == DWFGPU_OK) {
    // fold the pending correction in and measure the true residual
    TRY(flush());
    hipLaunchKernelGGL((dwfk::k_accumulate<XS>), blas_grid(), blas_block(), 0, h->st,
                       fld<P64>(sol), fld<XS>(xs), h->g);
    HIPCHK(hipGetLastError());
    TRY(H::mdagm(h, mass, ax, sol, nullptr));
    TRY(H::axpy_to(h, r, ax, src, -1.0, true));
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
    status = DWFGPU_MAXITER;
  }
  if (true_res) *true_res = std::sqrt(rr / src_norm_sq);
  if (iters) *iters = itr + 1;
  if (nreliable) *nreliable = nrel;
  return status;
}

// CANONICAL host gauge -> FP64 device layout with BondCond signs
int upload_gauge(dwfgpu_s* h, const double* uc) {
  const int lx = h->dims[0], ly = h->dims[1], lz = h->dims[2], lt = h->dims[3];
  const size_t vh = h->g.vh;
  std::vector<double> dev((size_t)4 * 2 * 9 * vh * 2);
  for (int t = 0; t < lt; ++t)
    for (int z = 0; z < lz; ++z)
      for (int y = 0; y < ly; ++y)
        for (int x = 0; x < lx; ++x) {
          const size_t site = x + (size_t)lx * (y + (size_t)ly * (z + (size_t)lz * t));
          const int par = (x + y + z + t) & 1;
          const size_t cbs = dwfk::site_index(h->g, x / 2, y, z, t);
          const int pos[4] = {x, y, z, t};
          for (int mu = 0; mu < 4; ++mu) {
            const double sg = (h->bc[mu] && pos[mu] == h->dims[mu] - 1) ? -1.0 : 1.0;
            const double* src = uc + (site * 4 + mu) * 18;
            for (int e = 0; e < 9; ++e) {
              const size_t o = (((size_t)mu * 2 + par) * 9 + e) * vh + cbs;
              dev[2 * o] = sg * src[2 * e];
              dev[2 * o + 1] = sg * src[2 * e + 1];
            }
          }
        }
  void** U = &h->f[DWFGPU_F64].U;
  HIPCHK(hipMalloc(U, dev.size() * sizeof(double)));
  HIPCHK(hipMemcpy(*U, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice));
  return DWFGPU_OK;
}

// Events around each kernel of MatPcDagMatPc (the 4 hopping passes and the
// norm's final reduction); they are recorded on the stream the kernels run on.
template <class P>
int time_mdagm(dwfgpu_s* h, double mass, const DVec& out, const DVec& in, int reps,
               double* ms_app, double* ms_kernel, double* bytes_kernel, int* nkernel) {
  typedef Ops<P> O;
  TRY(ensure_fields<P>(h));
  const DVec& te = h->f[pidx<P>()].te;
  const DVec& t1 = h->f[pidx<P>()].t1;
  const int nk = 5;
  std::vector<hipEvent_t> ev((size_t)reps * nk + 1, nullptr);
  int st = DWFGPU_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) st = DWFGPU_HIP_ERROR;
  const double half = half_bytes<P>(h);
  const double links = link_bytes<P>(h);
  const double kb[nk] = {2 * half + links, 3 * half + links, 2 * half + links, 3 * half + links, 0};
  const bool dwf = h->action == DWFGPU_DWF;
  const double mk2 = dwf ? -dwf_kappa(h) * dwf_kappa(h) : 0.0;
  auto rec = [&](size_t k) {
    if (!st && hipEventRecord(ev[k], h->st) != hipSuccess) st = DWFGPU_HIP_ERROR;
  };
  for (int r = 0; r < reps && !st; ++r) {
    const size_t b = (size_t)r * nk;
    rec(b + 0);
    if (!st) st = dwf ? O::dwf(h, mass, te, in, 1, 0, 0, DVec(), 0, false)
                      : O::mob_first(h, mass, te, in, 0);
    rec(b + 1);
    if (!st) st = dwf ? O::dwf(h, mass, t1, te, 0, 0, 1, in, mk2, true)
                      : O::mob_second(h, mass, t1, te, in, 0, true);
    rec(b + 2);
    if (!st) st = dwf ? O::dwf(h, mass, te, t1, 1, 1, 0, DVec(), 0, false)
                      : O::mob_first(h, mass, te, t1, 1);
    rec(b + 3);
    if (!st) st = dwf ? O::dwf(h, mass, out, te, 0, 1, 1, t1, mk2, false)
                      : O::mob_second(h, mass, out, te, t1, 1, false);
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

void ptc_free(PtcNet* n);
void free_all(dwfgpu_s* h) {
  if (h->net) ptc_free(h->net);
  h->net = nullptr;
  for (int p = 0; p < NPREC; ++p) {
    PrecFields& F = h->f[p];
    if (F.U) (void)hipFree(F.U);
    F.U = nullptr;
    free_vec(&F.te);
    free_vec(&F.t1);
    for (auto& w : F.work) free_vec(&w);
    for (auto& w : F.ring) free_vec(&w);
    F.ring.clear();
    F.ready = false;
  }
  void* ptrs[] = {h->cps_stage, h->stage4d, h->partial, h->dscal, h->nbt, h->dalpha};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->hscal) (void)hipHostFree(h->hscal);
  if (h->hring) (void)hipHostFree(h->hring);
  if (h->st) (void)hipStreamDestroy(h->st);
}

// dispatch on the API precision of the handle
#define DISPATCH(h, call)                                \
  ((h)->prec == DWFGPU_F64    ? Ops<P64>::call           \
   : (h)->prec == DWFGPU_F32  ? Ops<P32>::call           \
   : (h)->prec == DWFGPU_BF16 ? Ops<PBF16>::call         \
                              : Ops<PI16>::call)

int ensure_api(dwfgpu_s* h, int nwork) {
  switch (h->prec) {
    case DWFGPU_F64: return ensure_work<P64>(h, nwork);
    case DWFGPU_F32: return ensure_work<P32>(h, nwork);
    case DWFGPU_BF16: return ensure_work<PBF16>(h, nwork);
    default: return ensure_work<PI16>(h, nwork);
  }
}
int alloc_api_vec(dwfgpu_s* h, DVec* v) {
  switch (h->prec) {
    case DWFGPU_F64: return alloc_vec<P64>(h, v);
    case DWFGPU_F32: return alloc_vec<P32>(h, v);
    case DWFGPU_BF16: return alloc_vec<PBF16>(h, v);
    default: return alloc_vec<PI16>(h, v);
  }
}
const DVec& api_work(dwfgpu_s* h, int k) { return h->f[h->prec].work[k]; }

}  // namespace



// ===========================================================================
// Learned preconditioner: the parallel-transport network (ptc_kernels.hpp),
// its training on the GPU and the flexible CG that applies it
// ===========================================================================
struct PtcNet {
  int ls = 0, hid = 16;
  std::vector<std::complex<float>> W[3];  // [9][cout][cin][4][4]
  std::vector<double> m[3], v[3];         // Adam moments of (re, im)
  long adam_t = 0;
  uint16_t* A[3] = {nullptr, nullptr, nullptr};     // bf16 real forms, [9][M'][K']
  uint16_t* Aadj[3] = {nullptr, nullptr, nullptr};  // of the adjoint layers
  uint32_t* act[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t* gact[2] = {nullptr, nullptr};
  float* gwpart = nullptr;
  int nwg[3] = {0, 0, 0};  // work-groups per direction of each layer's weight gradient
  cx<float>* G[3] = {nullptr, nullptr, nullptr};
  int cin(int k) const { return k == 0 ? ls : hid; }
  int cout(int k) const { return k == 2 ? ls : hid; }
  size_t nw(int k) const { return (size_t)ptck::NDIR * cout(k) * cin(k) * 16; }
};

namespace {

void ptc_free(PtcNet* n) {
  if (!n) return;
  for (int k = 0; k < 3; ++k) {
    if (n->A[k]) (void)hipFree(n->A[k]);
    if (n->Aadj[k]) (void)hipFree(n->Aadj[k]);
    if (n->G[k]) (void)hipFree(n->G[k]);
  }
  for (auto p : n->act)
    if (p) (void)hipFree(p);
  for (auto p : n->gact)
    if (p) (void)hipFree(p);
  if (n->gwpart) (void)hipFree(n->gwpart);
  delete n;
}

uint16_t bf16_host(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// A_d = [[Re W_d, -Im W_d], [Im W_d, Re W_d]] (rows (part, o, a), columns
// (part, i, b)) for every d, from W[d][o][i][a][b]; adj: the real form of
// the adjoint layer, W'_{d'}[i][o][b][a] = conj(W_d[o][i][a][b]), d' = -d
void ptc_real_form(const std::vector<std::complex<float>>& W, int cin, int cout, bool adj,
                   std::vector<uint16_t>& out) {
  const int ci = adj ? cout : cin, co = adj ? cin : cout;
  const int MP = 8 * co, KP = 8 * ci;
  out.assign((size_t)ptck::NDIR * MP * KP, 0);
  for (int d = 0; d < ptck::NDIR; ++d) {
    const int ds = adj ? (d == 0 ? 0 : (d <= 4 ? d + 4 : d - 4)) : d;  // source direction
    for (int o = 0; o < co; ++o)
      for (int i = 0; i < ci; ++i)
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) {
            std::complex<float> w;
            if (!adj)
              w = W[(((size_t)(ds * cout + o) * cin + i) * 4 + a) * 4 + b];
            else  // W'[o'=i][i'=o][a'=b][b'=a] = conj(W[o][i][a][b])
              w = std::conj(W[(((size_t)(ds * cout + i) * cin + o) * 4 + b) * 4 + a]);
            const int mre = o * 4 + a, mim = co * 4 + o * 4 + a;
            const int kre = i * 4 + b, kim = ci * 4 + i * 4 + b;
            uint16_t* Ad = out.data() + (size_t)d * MP * KP;
            Ad[mre * KP + kre] = bf16_host(w.real());
            Ad[mre * KP + kim] = bf16_host(-w.imag());
            Ad[mim * KP + kre] = bf16_host(w.imag());
            Ad[mim * KP + kim] = bf16_host(w.real());
          }
  }
}

// [9][M'][K'] row-major -> the MFMA fragment order of k_ptc_fwd: 8 bf16
// A[d][16*blk + l%16][32*kb + 8*(l/16) + e] at (((d*MB + blk)*K32 + kb)*64 + l)*8 + e
void ptc_fragment_order(std::vector<uint16_t>& buf, int mp, int kp) {
  const int mb = mp / 16, k32 = kp / 32;
  std::vector<uint16_t> rm(buf);
  for (int d = 0; d < ptck::NDIR; ++d)
    for (int blk = 0; blk < mb; ++blk)
      for (int kb = 0; kb < k32; ++kb)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 8; ++e)
            buf[((((size_t)d * mb + blk) * k32 + kb) * 64 + l) * 8 + e] =
                rm[(size_t)d * mp * kp + (size_t)(16 * blk + l % 16) * kp + 32 * kb + 8 * (l / 16) + e];
}

int ptc_upload(dwfgpu_s* h) {
  PtcNet* n = h->net;
  std::vector<uint16_t> buf;
  for (int k = 0; k < 3; ++k) {
    ptc_real_form(n->W[k], n->cin(k), n->cout(k), false, buf);
    ptc_fragment_order(buf, 8 * n->cout(k), 8 * n->cin(k));
    HIPCHK(hipMemcpyAsync(n->A[k], buf.data(), buf.size() * 2, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    ptc_real_form(n->W[k], n->cin(k), n->cout(k), true, buf);
    ptc_fragment_order(buf, 8 * n->cin(k), 8 * n->cout(k));
    HIPCHK(hipMemcpyAsync(n->Aadj[k], buf.data(), buf.size() * 2, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
  }
  return DWFGPU_OK;
}

template <int CIN, int COUT>
int ptc_fwd_t(dwfgpu_s* h, uint32_t* out, const uint32_t* in, const uint16_t* A) {
  constexpr size_t lds = ptck::fwd_lds<CIN, COUT>();
  if (lds > (size_t)h->max_lds) {
    set_err("ptc: layer %dx%d needs %zu bytes of LDS", CIN, COUT, lds);
    return DWFGPU_BAD_ARG;
  }
  hipLaunchKernelGGL((ptck::k_ptc_fwd<CIN, COUT>), dim3(2 * h->g.vh / ptck::NT), dim3(ptck::NTHR),
                     lds, h->st, out, in, reinterpret_cast<const ptck::u32x4*>(A),
                     (const typename P32::S*)h->f[DWFGPU_F32].U, h->g);
  HIPCHK(hipGetLastError());
  return DWFGPU_OK;
}

// Work-groups per direction of the weight-gradient kernel: as many as fill
// every CU once (all NDIR directions resident together), so the loop over
// tiles ends at the same time everywhere (576 work-groups over 512 resident
// slots had left a second round of 64 with the device almost idle), rounded
// down to a multiple of the XCD count so that each XCD holds all directions
// of its work-groups.
template <int CIN, int COUT>
int ptc_gw_grid(dwfgpu_s* h, int* nwg) {
  constexpr size_t lds = ptck::gw_lds<CIN, COUT>();
  int ncu = 0, per_cu = 0;
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->dev));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ptck::k_ptc_gw<CIN, COUT>,
                                                      ptck::NTHR, lds));
  int n = std::max(1, std::min(ncu * std::max(per_cu, 1) / ptck::NDIR,
                               (int)(2 * h->g.vh / ptck::NT)));
  if (n >= h->g.nxcd) n -= n % h->g.nxcd;
  *nwg = n;
  return DWFGPU_OK;
}

template <int CIN, int COUT>
int ptc_gw_t(dwfgpu_s* h, float* part, int nwg, const uint32_t* gout, const uint32_t* in) {
  constexpr size_t lds = ptck::gw_lds<CIN, COUT>();
  if (lds > (size_t)h->max_lds) {
    set_err("ptc: gradient %dx%d needs %zu bytes of LDS", CIN, COUT, lds);
    return DWFGPU_BAD_ARG;
  }
  hipLaunchKernelGGL((ptck::k_ptc_gw<CIN, COUT>), dim3(ptck::NDIR * nwg), dim3(ptck::NTHR), lds,
                     h->st, part, nwg, gout, in, (const typename P32::S*)h->f[DWFGPU_F32].U, h->g);
  HIPCHK(hipGetLastError());
  return DWFGPU_OK;
}

// the channel pairs the kernels are built for: Ls in {4, 8, 12, 16}, H = 16
#define PTC_PAIRS(X) X(4, 16) X(16, 4) X(8, 16) X(16, 8) X(12, 16) X(16, 12) X(16, 16)

int ptc_fwd(dwfgpu_s* h, int cin, int cout, uint32_t* out, const uint32_t* in, const uint16_t* A) {
#define PTC_F(CI, CO) \
  if (cin == CI && cout == CO) return ptc_fwd_t<CI, CO>(h, out, in, A);
  PTC_PAIRS(PTC_F)
#undef PTC_F
  set_err("ptc: no layer kernel for %d -> %d channels", cin, cout);
  return DWFGPU_BAD_ARG;
}

int ptc_gw(dwfgpu_s* h, int cin, int cout, float* part, int nwg, const uint32_t* gout,
           const uint32_t* in) {
#define PTC_G(CI, CO) \
  if (cin == CI && cout == CO) return ptc_gw_t<CI, CO>(h, part, nwg, gout, in);
  PTC_PAIRS(PTC_G)
#undef PTC_G
  set_err("ptc: no gradient kernel for %d -> %d channels", cin, cout);
  return DWFGPU_BAD_ARG;
}

int ptc_gw_nwg(dwfgpu_s* h, int cin, int cout, int* nwg) {
#define PTC_N(CI, CO) \
  if (cin == CI && cout == CO) return ptc_gw_grid<CI, CO>(h, nwg);
  PTC_PAIRS(PTC_N)
#undef PTC_N
  set_err("ptc: no gradient kernel for %d -> %d channels", cin, cout);
  return DWFGPU_BAD_ARG;
}

// z = |r| net(r/|r|) on sloppy vectors (z must differ from r)
template <class P>
int ptc_apply(dwfgpu_s* h, const DVec& z, const DVec& r, double* rnorm_out = nullptr) {
  PtcNet* n = h->net;
  double rr;
  TRY(Ops<P>::norm(h, r, &rr));
  const double rn = std::sqrt(rr);
  if (rnorm_out) *rnorm_out = rn;
  if (rn == 0.0) return Ops<P>::copy(h, z, r);  // r = 0: z = 0
  const int action = h->action == DWFGPU_DWF ? 0 : 1;
  hipLaunchKernelGGL((ptck::k_ptc_embed<P>), dim3(2048), dim3(256), 0, h->st, n->act[0],
                     cfld<P>(r), h->g, action, (float)(1.0 / rn));
  HIPCHK(hipGetLastError());
  for (int k = 0; k < 3; ++k) TRY(ptc_fwd(h, n->cin(k), n->cout(k), n->act[k + 1], n->act[k], n->A[k]));
  hipLaunchKernelGGL((ptck::k_ptc_extract<P>), dim3(2048), dim3(256), 0, h->st, fld<P>(z),
                     n->act[3], h->g, action, (float)rn);
  HIPCHK(hipGetLastError());
  return DWFGPU_OK;
}

// loss of one (r, e) pair (FP32 vectors) and G[k] (+)= dL/dW[k]
int ptc_grad_pair(dwfgpu_s* h, const DVec& r, const DVec& e, bool accumulate, double* loss) {
  PtcNet* n = h->net;
  const int action = h->action == DWFGPU_DWF ? 0 : 1;
  double rr, ee;
  TRY(Ops<P32>::norm(h, r, &rr));
  TRY(Ops<P32>::norm(h, e, &ee));
  const double rn = std::sqrt(rr);
  if (rn == 0.0 || ee == 0.0) {
    set_err("ptc: empty training pair");
    return DWFGPU_BAD_ARG;
  }
  const double nt = ee / rr;  // |e/|r||^2
  hipLaunchKernelGGL((ptck::k_ptc_embed<P32>), dim3(2048), dim3(256), 0, h->st, n->act[0],
                     cfld<P32>(r), h->g, action, (float)(1.0 / rn));
  HIPCHK(hipGetLastError());
  for (int k = 0; k < 3; ++k) TRY(ptc_fwd(h, n->cin(k), n->cout(k), n->act[k + 1], n->act[k], n->A[k]));
  hipLaunchKernelGGL(ptck::k_ptc_loss, dim3(dwfk::BLAS_BLOCKS), dim3(256), 0, h->st, n->gact[0],
                     n->act[3], cfld<P32>(e), (float)(1.0 / rn), h->g, action, (float)(1.0 / nt),
                     h->partial);
  HIPCHK(hipGetLastError());
  double dd;
  TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &dd));
  if (loss) *loss = dd / nt;
  // backward: layer 2, 1, 0 (gact ping-pong)
  int cur = 0;
  for (int k = 2; k >= 0; --k) {
    TRY(ptc_gw(h, n->cin(k), n->cout(k), n->gwpart, n->nwg[k], n->gact[cur], n->act[k]));
    hipLaunchKernelGGL(ptck::k_ptc_gw_reduce, dim3(128), dim3(256), 0, h->st, n->G[k], n->gwpart,
                       n->nwg[k], n->cin(k), n->cout(k), accumulate ? 1 : 0);
    HIPCHK(hipGetLastError());
    if (k > 0) {
      TRY(ptc_fwd(h, n->cout(k), n->cin(k), n->gact[1 - cur], n->gact[cur], n->Aadj[k]));
      cur = 1 - cur;
    }
  }
  return DWFGPU_OK;
}

// one Adam step with the gradients accumulated over `count` pairs
int ptc_adam(dwfgpu_s* h, int count, double lr) {
  PtcNet* n = h->net;
  const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
  ++n->adam_t;
  const double c1 = 1.0 - std::pow(b1, (double)n->adam_t), c2 = 1.0 - std::pow(b2, (double)n->adam_t);
  for (int k = 0; k < 3; ++k) {
    std::vector<std::complex<float>> G(n->nw(k));
    HIPCHK(hipMemcpyAsync(G.data(), n->G[k], G.size() * sizeof(cx<float>), hipMemcpyDeviceToHost,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (size_t q = 0; q < G.size(); ++q) {
      const double g2[2] = {G[q].real() / count, G[q].imag() / count};
      double w2[2] = {n->W[k][q].real(), n->W[k][q].imag()};
      for (int c = 0; c < 2; ++c) {
        double& m = n->m[k][2 * q + c];
        double& v = n->v[k][2 * q + c];
        m = b1 * m + (1.0 - b1) * g2[c];
        v = b2 * v + (1.0 - b2) * g2[c] * g2[c];
        w2[c] -= lr * (m / c1) / (std::sqrt(v / c2) + eps);
      }
      n->W[k][q] = std::complex<float>((float)w2[0], (float)w2[1]);
    }
  }
  return ptc_upload(h);
}

int ptc_ensure(dwfgpu_s* h) {
  if (h->net) return DWFGPU_OK;
  const int ls = h->ls;
  if (!(ls == 4 || ls == 8 || ls == 12 || ls == 16)) {
    set_err("ptc: the network is built for Ls in {4, 8, 12, 16} (got %d)", ls);
    return DWFGPU_BAD_ARG;
  }
  if (h->g.vh % ptck::NT) {
    set_err("ptc: V4/2 must be a multiple of %d", ptck::NT);
    return DWFGPU_BAD_ARG;
  }
  TRY(ensure_fields<P32>(h));  // compressed FP32 links
  PtcNet* n = new PtcNet;
  n->ls = ls;
  h->net = n;
  int st = DWFGPU_OK;
  const size_t V = 2 * (size_t)h->g.vh;
  const int cmax = std::max(ls, n->hid);
  for (int k = 0; k < 3 && !st; ++k) {
    n->W[k].assign(n->nw(k), std::complex<float>(0.f, 0.f));
    n->m[k].assign(2 * n->nw(k), 0.0);
    n->v[k].assign(2 * n->nw(k), 0.0);
    const size_t abytes = (size_t)ptck::NDIR * 64 * n->cin(k) * n->cout(k) * 2;
    st = alloc(h, (void**)&n->A[k], abytes);
    if (!st) st = alloc(h, (void**)&n->Aadj[k], abytes);
    if (!st) st = alloc(h, (void**)&n->G[k], n->nw(k) * sizeof(cx<float>));
  }
  const int chans[4] = {ls, n->hid, n->hid, ls};
  for (int k = 0; k < 4 && !st; ++k) st = alloc(h, (void**)&n->act[k], V * chans[k] * 12 * 4);
  for (int k = 0; k < 2 && !st; ++k) st = alloc(h, (void**)&n->gact[k], V * cmax * 12 * 4);
  int nwg_max = 0;
  for (int k = 0; k < 3 && !st; ++k) {
    st = ptc_gw_nwg(h, n->cin(k), n->cout(k), &n->nwg[k]);
    nwg_max = std::max(nwg_max, n->nwg[k]);
  }
  if (!st)
    st = alloc(h, (void**)&n->gwpart,
               (size_t)ptck::NDIR * nwg_max * 64 * n->hid * n->hid * sizeof(float));
  if (st) {
    ptc_free(n);
    h->net = nullptr;
  }
  return st;
}

// default weights: a damped local copy plus small random hops, so that the
// untrained network is a scaled identity (z = 0.5 r) and training starts
// from a valid preconditioner
void ptc_init_weights(PtcNet* n, unsigned long long seed, double scale) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int k = 0; k < 3; ++k) {
    const int ci = n->cin(k), co = n->cout(k);
    for (size_t q = 0; q < n->nw(k); ++q)
      n->W[k][q] = std::complex<float>((float)(scale * nd(rng)), (float)(scale * nd(rng)));
    // d = 0, o = i (channels that exist on both sides), a = b: a real diagonal
    const double diag = k == 2 ? 0.5 : 1.0;
    for (int o = 0; o < std::min(ci, co); ++o)
      for (int a = 0; a < 4; ++a)
        n->W[k][(((size_t)(0 * co + o) * ci + o) * 4 + a) * 4 + a] += std::complex<float>((float)diag, 0.f);
  }
  for (int k = 0; k < 3; ++k) {
    std::fill(n->m[k].begin(), n->m[k].end(), 0.0);
    std::fill(n->v[k].begin(), n->v[k].end(), 0.0);
  }
  n->adam_t = 0;
}

}  // namespace


namespace {

template <class P>
int dot(dwfgpu_s* h, const DVec& x, const DVec& y, double* out) {
  hipLaunchKernelGGL((ptck::k_dot<P>), blas_grid(), blas_block(), 0, h->st, cfld<P>(x), cfld<P>(y),
                     h->g, h->partial);
  HIPCHK(hipGetLastError());
  return fetch_sum(h, dwfk::BLAS_BLOCKS, out);
}

// Training pairs from one solve of A x = b: FP32 CG from x = 0 keeps x_k at
// iterations stride, 2 stride, ...; with x* from a reliable-update solve,
// e_k = x* - x_k and r_k = A e_k (FP32 operator), so A e_k = r_k holds for
// the stored vectors exactly up to FP32 rounding.  Late CG residuals are
// rich in the low modes a preconditioner has to capture.
int ptc_pairs(dwfgpu_s* h, double mass, const DVec& b64, int stride, int want,
              std::vector<DVec>& R, std::vector<DVec>& E, int* made) {
  // pairs R/E[0..want) from the CG snapshots (filled below)
  typedef Ops<P32> L;
  *made = 0;
  DVec x64, x32, r32, p32;
  int st = alloc_vec<P64>(h, &x64);
  if (!st) st = alloc_vec<P32>(h, &x32);
  if (!st) st = alloc_vec<P32>(h, &r32);
  if (!st) st = alloc_vec<P32>(h, &p32);
  auto done = [&](int code) {
    free_vec(&x64);
    free_vec(&x32);
    free_vec(&r32);
    free_vec(&p32);
    return code;
  };
  if (st) return done(st);
  const dwfgpu_cg_arg a{mass, 20000, 1e-9, 1e-9};
  double tr;
  int it, nrel;
  st = invcg_mixed<P32>(h, &a, 0.1, x64, b64, &tr, &it, &nrel);
  if (st != DWFGPU_OK && st != DWFGPU_MAXITER) return done(st);
  // FP32 CG with snapshots
  hipLaunchKernelGGL((dwfk::k_convert<P32, P64>), blas_grid(), blas_block(), 0, h->st,
                     fld<P32>(r32), cfld<P64>(b64), h->g);
  if (hipGetLastError() != hipSuccess) return done(DWFGPU_HIP_ERROR);
  if ((st = L::copy(h, p32, r32))) return done(st);
  if ((st = L::scale(h, x32, 0.0))) return done(st);
  double rr;
  if ((st = L::norm(h, r32, &rr))) return done(st);
  // dscal[0] = rr for the fused step
  HIPCHK(hipMemcpyAsync(h->dscal, &rr, sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  const double stp = rr * 1e-18;
  int k = 0;
  for (int itr = 1; itr <= stride * want && k < want; ++itr) {
    double rrn, d;
    if ((st = L::template mdagm_cg<P32>(h, mass, p32, r32, x32, &rrn, &d))) return done(st);
    if (d == 0.0 || rrn <= stp) break;
    if (itr % stride == 0) {
      if ((st = L::copy(h, E[k], x32))) return done(st);
      ++k;
    }
    if ((st = L::axpby(h, p32, r32, 1.0, rrn / rr))) return done(st);
    rr = rrn;
  }
  // e_k = x* - x_k, r_k = A e_k
  hipLaunchKernelGGL((dwfk::k_convert<P32, P64>), blas_grid(), blas_block(), 0, h->st,
                     fld<P32>(x32), cfld<P64>(x64), h->g);
  if (hipGetLastError() != hipSuccess) return done(DWFGPU_HIP_ERROR);
  for (int j = 0; j < k; ++j) {
    if ((st = L::axpby(h, E[j], x32, 1.0, -1.0))) return done(st);
    if ((st = L::mdagm(h, mass, R[j], E[j], nullptr))) return done(st);
  }
  *made = k;
  return done(DWFGPU_OK);
}

// Flexible CG (truncated to one previous direction) on A = MatPcDagMatPc
// with z = net(r) as preconditioner, sloppy vectors in PS and FP64
// reliable updates; the search direction restarts at every reliable update.
// The net is switched off (plain CG from then on) when (r, net(r)) <= 0 or
// is not finite, or when the residual stalls for `stall` steps.
template <class PS>
int invfcg_ptc(dwfgpu_s* h, const dwfgpu_cg_arg* arg, double delta, const DVec& sol,
               const DVec& src, double* true_res, int* iters, int* nrel_out, int* net_steps) {
  typedef Ops<P64> H;
  typedef Ops<PS> L;
  TRY(ensure_work<P64>(h, 2));
  TRY(ensure_work<PS>(h, 4));
  TRY(ensure_work<P32>(h, 5));
  const DVec& r = h->f[DWFGPU_F64].work[0];
  const DVec& ax = h->f[DWFGPU_F64].work[1];
  const DVec& rs = h->f[pidx<PS>()].work[0];
  const DVec& ps = h->f[pidx<PS>()].work[1];
  const DVec& qs = h->f[pidx<PS>()].work[2];
  const DVec& z = h->f[pidx<PS>()].work[3];
  const DVec& xs = h->f[DWFGPU_F32].work[4];
  const double mass = arg->mass;
  double sn;
  TRY(H::norm(h, src, &sn));
  const double stp = sn * arg->stop_rsd * arg->stop_rsd;
  TRY(H::mdagm(h, mass, ax, sol, nullptr));
  TRY(H::axpy_to(h, r, ax, src, -1.0, true));
  double rr;
  TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
  hipLaunchKernelGGL((dwfk::k_convert<PS, P64>), blas_grid(), blas_block(), 0, h->st, fld<PS>(rs),
                     cfld<P64>(r), h->g);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(xs.d, 0, h->n / 2 * sizeof(cx<float>), h->st));
  bool use_net = h->net != nullptr;
  int nsteps = 0, nbad = 0;
  // z = net(rs); a z nearly orthogonal to rs (cos < 1e-2) or not finite is
  // replaced by rs for this step, and after 10 such steps the net is dropped
  auto precondition = [&](double rrs) -> int {
    bool ok = false;
    if (use_net) {
      TRY(ptc_apply<PS>(h, z, rs));
      double rz, zz;
      TRY(dot<PS>(h, rs, z, &rz));
      TRY(L::norm(h, z, &zz));
      ok = std::isfinite(rz) && std::isfinite(zz) && zz > 0.0 &&
           rz > 1e-2 * std::sqrt(rrs * zz);
      if (ok) ++nsteps;
      else if (++nbad > 10) use_net = false;
    }
    if (!ok) TRY(L::copy(h, z, rs));
    return DWFGPU_OK;
  };
  TRY(precondition(rr));
  TRY(L::copy(h, ps, z));
  double rr_upd = rr;
  int itr = 0, nrel = 0, status = DWFGPU_OK, since = 0;
  const int max_itr = arg->max_num_iter - 1;
  const int stall = 400;
  bool converged = rr <= stp;
  while (!converged && itr < max_itr) {
    ++itr;
    ++since;
    double d, rp;
    TRY(L::mdagm(h, mass, qs, ps, &d));
    if (d == 0.0) {
      status = DWFGPU_BREAKDOWN;
      break;
    }
    TRY(dot<PS>(h, rs, ps, &rp));
    const double a = rp / d;  // exact line minimum along p
    hipLaunchKernelGGL((dwfk::k_cg_update<PS, P32>), blas_grid(), blas_block(), 0, h->st,
                       fld<P32>(xs), fld<PS>(rs), cfld<PS>(ps), cfld<PS>(qs), a, h->g, h->partial);
    HIPCHK(hipGetLastError());
    double rrn;
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rrn));
    bool restart = false;
    if (rrn <= stp || rrn < delta * delta * rr_upd || (use_net && since > stall)) {
      if (use_net && since > stall && !(rrn < delta * delta * rr_upd)) use_net = false;
      hipLaunchKernelGGL((dwfk::k_accumulate<P32>), blas_grid(), blas_block(), 0, h->st,
                         fld<P64>(sol), fld<P32>(xs), h->g);
      HIPCHK(hipGetLastError());
      TRY(H::mdagm(h, mass, ax, sol, nullptr));
      TRY(H::axpy_to(h, r, ax, src, -1.0, true));
      TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rrn));
      hipLaunchKernelGGL((dwfk::k_convert<PS, P64>), blas_grid(), blas_block(), 0, h->st,
                         fld<PS>(rs), cfld<P64>(r), h->g);
      HIPCHK(hipGetLastError());
      rr_upd = rrn;
      ++nrel;
      since = 0;
      restart = true;
      if (rrn <= stp) {
        rr = rrn;
        converged = true;
        break;
      }
    }
    TRY(precondition(rrn));
    double beta = 0.0;
    if (!restart) {
      double zq;
      TRY(dot<PS>(h, z, qs, &zq));
      beta = -zq / d;
    }
    TRY(L::axpby(h, ps, z, 1.0, beta));  // p = z + beta p
    rr = rrn;
  }
  if (!converged && status == DWFGPU_OK) {
    hipLaunchKernelGGL((dwfk::k_accumulate<P32>), blas_grid(), blas_block(), 0, h->st,
                       fld<P64>(sol), fld<P32>(xs), h->g);
    HIPCHK(hipGetLastError());
    TRY(H::mdagm(h, mass, ax, sol, nullptr));
    TRY(H::axpy_to(h, r, ax, src, -1.0, true));
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
    status = DWFGPU_MAXITER;
  }
  if (true_res) *true_res = std::sqrt(rr / sn);
  if (iters) *iters = itr + 1;
  if (nrel_out) *nrel_out = nrel;
  if (net_steps) *net_steps = nsteps;
  return status;
}

}  // namespace


// ===========================================================================
// Many right-hand sides (DWF): batched operator and the batched
// reliable-update CG with device-side retirement and a graph-captured step
// ===========================================================================
namespace {

// a slice of an nrhs-vector allocation of policy P
template <class P>
DVec mslice(const dwfgpu_s* h, const DVec& v, int j) {
  const size_t vs = h->n / 2 / P::kCPC;  // chunks per vector
  DVec r;
  r.d = (char*)v.d + (size_t)j * vs * sizeof(typename P::C);
  r.sc = v.sc ? v.sc + (size_t)j * h->g.vh * h->ls : nullptr;
  return r;
}

template <class P>
int alloc_mvec(dwfgpu_s* h, int nrhs, DVec* v) {
  TRY(alloc(h, &v->d, (size_t)nrhs * (h->n / 2) * sizeof(typename P::S)));
  if (P::kScaled)
    TRY(alloc(h, (void**)&v->sc, (size_t)nrhs * h->g.vh * h->ls * sizeof(float)));
  return DWFGPU_OK;
}

struct MrShape {
  int ns, nr;
};

template <class PS>
struct MrState {
  int nrhs = 0;
  DVec te, t1, rs, ps, xs;  // xs in P32
  double* partial = nullptr;
  double* scal = nullptr;
  double* mass = nullptr;
  int* ibuf = nullptr;  // active | flag | iters
  int gx = 0;
  ~MrState() {
    free_vec(&te);
    free_vec(&t1);
    free_vec(&rs);
    free_vec(&ps);
    free_vec(&xs);
    void* p[] = {partial, scal, mass, ibuf};
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
  int* active() const { return ibuf; }
  int* flag() const { return ibuf + nrhs; }
  int* iters() const { return ibuf + 2 * nrhs; }
};

size_t mr_lds(const dwfgpu_s* h, const MrShape& t, size_t creal) {
  return (size_t)2 * 8 * t.ns * (9 * 2 * creal + 4) + (size_t)t.ns * h->ls * t.nr * 8;
}

// the batched kernels write one partial sum per (RHS, work-group): the
// buffers hold ceil(vh / MR_MIN_NS) work-groups per RHS, so no shape may use
// fewer sites per work-group than that
constexpr int MR_MIN_NS = 8;
int mr_partials_per_rhs(const dwfgpu_s* h) { return (h->g.vh + MR_MIN_NS - 1) / MR_MIN_NS; }
// partials per RHS of the batched buffers: the DWF shapes' bound, or the
// Mobius passes' grid (the handle's hop shape, which may hold fewer than
// MR_MIN_NS sites per work-group at long Ls)
int mr_partials_stride(const dwfgpu_s* h) {
  return std::max(mr_partials_per_rhs(h), h->grid_hop);
}

bool mr_shape_ok(const dwfgpu_s* h, const MrShape& t, int nrhs, size_t creal) {
  return t.ns >= MR_MIN_NS && t.nr >= 1 && t.nr <= nrhs && t.ns * h->ls * t.nr <= 512 &&
         mr_lds(h, t, creal) <= (size_t)h->max_lds;
}

// kind 0: out = D p (cb 1)          kind 1: out = x0 + alpha D in, |out|^2
// kind 2: out = D^dag in (cb 1)      kind 3: x0 + alpha D^dag in -> CG epilogue
// kind 4: out = x0 + alpha D^dag in (plain batched MatPcDagMatPc)
template <class P>
int mr_launch(dwfgpu_s* h, const MrShape& t, int nrhs, dwfk::MrArgs<P, P32> a, int kind) {
  const dim3 grid((h->g.vh + t.ns - 1) / t.ns, (nrhs + t.nr - 1) / t.nr);
  const dim3 block(t.ns, h->ls, t.nr);
  const size_t lds = mr_lds(h, t, sizeof(typename P::R));
  a.cb_in = (kind == 0 || kind == 2) ? 1 : 0;
  switch (kind) {
    case 0:
      hipLaunchKernelGGL((dwfk::k_dwf_mrhs<P, 0, 0, false, false, P32>), grid, block, lds, h->st, a);
      break;
    case 1:
      hipLaunchKernelGGL((dwfk::k_dwf_mrhs<P, 0, 1, true, false, P32>), grid, block, lds, h->st, a);
      break;
    case 2:
      hipLaunchKernelGGL((dwfk::k_dwf_mrhs<P, 1, 0, false, false, P32>), grid, block, lds, h->st, a);
      break;
    case 3:
      hipLaunchKernelGGL((dwfk::k_dwf_mrhs<P, 1, 1, false, true, P32>), grid, block, lds, h->st, a);
      break;
    default:
      hipLaunchKernelGGL((dwfk::k_dwf_mrhs<P, 1, 1, false, false, P32>), grid, block, lds, h->st, a);
      break;
  }
  HIPCHK(hipGetLastError());
  return DWFGPU_OK;
}

// batched MatPcDagMatPc (cg: the CG step of every active RHS instead of `out`)
template <class P>
int mr_mdagm(dwfgpu_s* h, const MrShape& t, int nrhs, const DVec& out, const DVec& in,
             const DVec& te, const DVec& t1, const double* mass, const int* active,
             double* partial, double* scal, const DVec* r, const DVec* x, bool cg) {
  const double kap = dwf_kappa(h);
  dwfk::MrArgs<P, P32> a{};
  a.U = (const typename P::S*)h->f[pidx<P>()].U;
  a.partial = partial;
  a.g = h->g;
  a.two_a5 = 2.0 * h->a5inv;
  a.mass = mass;
  a.active = active;
  a.vs = h->n / 2 / P::kCPC;
  a.nrhs = nrhs;
  a.scal = scal;
  const int gx = (h->g.vh + t.ns - 1) / t.ns;
  // te = D in
  a.out = fld<P>(te);
  a.in = cfld<P>(in);
  TRY(mr_launch<P>(h, t, nrhs, a, 0));
  // t1 = in - kappa^2 D te, |t1|^2 -> scal[D]
  a.out = fld<P>(t1);
  a.in = cfld<P>(te);
  a.x0 = cfld<P>(in);
  a.alpha = -kap * kap;
  TRY(mr_launch<P>(h, t, nrhs, a, 1));
  if (scal) {
    hipLaunchKernelGGL(dwfk::k_reduce_mrhs, dim3(nrhs), dim3(1024), 0, h->st, partial, gx, scal,
                       (int)dwfk::MR_D, active);
    HIPCHK(hipGetLastError());
  }
  // te = D^dag t1
  a.out = fld<P>(te);
  a.in = cfld<P>(t1);
  a.x0 = dwfk::CFld<P>{};
  a.alpha = 0.0;
  TRY(mr_launch<P>(h, t, nrhs, a, 2));
  // t1 - kappa^2 D^dag te: into out, or into the CG step
  a.in = cfld<P>(te);
  a.x0 = cfld<P>(t1);
  a.alpha = -kap * kap;
  if (cg) {
    a.r = fld<P>(*r);
    a.p = cfld<P>(in);
    a.x = fld<P32>(*x);
    TRY(mr_launch<P>(h, t, nrhs, a, 3));
    hipLaunchKernelGGL(dwfk::k_reduce_mrhs, dim3(nrhs), dim3(1024), 0, h->st, partial, gx, scal,
                       (int)dwfk::MR_RRNEW, active);
    HIPCHK(hipGetLastError());
  } else {
    a.out = fld<P>(out);
    TRY(mr_launch<P>(h, t, nrhs, a, 4));
  }
  return DWFGPU_OK;
}

// batched Mobius MatPcDagMatPc: the four hop passes of each RHS with its own
// mass, launched per RHS (the M5^-1 and (b + c D5) epilogues keep every s-
// column of a site in one work-group's LDS, so RHS share a launch shape but
// not a work-group).  A paused RHS (active[j] == 0) costs four launches whose
// work-groups leave at once.  Partials: grid_hop per RHS, RHS j at j*grid_hop.
// cg: the CG step of every active RHS (r -= alpha A p, x += alpha p in the
// last pass, alpha and |r|^2 on the device) instead of `out`.
template <class P>
int mr_mdagm_mob(dwfgpu_s* h, int nrhs, const DVec& out, const DVec& in, const DVec& te,
                 const DVec& t1, const double* masses, const int* active, double* partial,
                 double* scal, const DVec* r, const DVec* x, bool cg) {
  typedef Ops<P> O;
  const int gx = h->grid_hop;
  for (int j = 0; j < nrhs; ++j) {
    const DVec inj = mslice<P>(h, in, j), tej = mslice<P>(h, te, j), t1j = mslice<P>(h, t1, j);
    TRY(O::mob_first(h, masses[j], tej, inj, 0, active + j));
    TRY(O::mob_second(h, masses[j], t1j, tej, inj, 0, scal != nullptr, partial + (size_t)j * gx,
                      active + j));
  }
  if (scal) {
    hipLaunchKernelGGL(dwfk::k_reduce_mrhs, dim3(nrhs), dim3(1024), 0, h->st, partial, gx, scal,
                       (int)dwfk::MR_D, active);
    HIPCHK(hipGetLastError());
  }
  for (int j = 0; j < nrhs; ++j) {
    const DVec tej = mslice<P>(h, te, j), t1j = mslice<P>(h, t1, j);
    TRY(O::mob_first(h, masses[j], tej, t1j, 1, active + j));
    if (cg)
      TRY(O::template mob_cg_last<P32>(h, masses[j], tej, t1j, mslice<P>(h, *r, j),
                                       mslice<P>(h, in, j), mslice<P32>(h, *x, j),
                                       scal + (size_t)j * dwfk::MR_NSCAL,
                                       partial + (size_t)j * gx, active + j));
    else
      TRY(O::mob_second(h, masses[j], mslice<P>(h, out, j), tej, t1j, 1, false, nullptr,
                        active + j));
  }
  if (cg) {
    hipLaunchKernelGGL(dwfk::k_reduce_mrhs, dim3(nrhs), dim3(1024), 0, h->st, partial, gx, scal,
                       (int)dwfk::MR_RRNEW, active);
    HIPCHK(hipGetLastError());
  }
  return DWFGPU_OK;
}

// launch shapes the autotuner tries: ns sites x Ls slices x nr RHS threads
std::vector<MrShape> mr_candidates(const dwfgpu_s* h, int nrhs, size_t creal) {
  std::vector<MrShape> c;
  for (int ns : {8, 16, 32})
    for (int nr : {1, 2, 4}) {
      const MrShape t{ns, nr};
      if (ns * h->ls * nr < 64 || !mr_shape_ok(h, t, nrhs, creal)) continue;
      c.push_back(t);
    }
  if (c.empty() && mr_shape_ok(h, MrShape{MR_MIN_NS, 1}, nrhs, creal))
    c.push_back(MrShape{MR_MIN_NS, 1});
  return c;
}

// Project-written autotuner: time one batched MatPcDagMatPc (all RHS active)
// for every candidate shape, twice, and keep the fastest; cached per handle,
// sloppy precision and RHS count.
template <class P>
int mr_tune(dwfgpu_s* h, int nrhs, MrState<P>& m, const int* all_active, MrShape* best) {
  const int pi = pidx<P>();
  if (h->mr_ns[pi] && h->mr_nrhs[pi] == nrhs) {
    *best = MrShape{h->mr_ns[pi], h->mr_nr[pi]};
    return DWFGPU_OK;
  }
  const char* env = getenv("DWFGPU_MRHS_SHAPE");  // "ns,nr" pins the shape
  if (env) {
    int ns = 0, nr = 0;
    if (sscanf(env, "%d,%d", &ns, &nr) != 2 ||
        !mr_shape_ok(h, MrShape{ns, nr}, nrhs, sizeof(typename P::R))) {
      set_err("DWFGPU_MRHS_SHAPE=%s: need ns >= %d, 1 <= nr <= nrhs, ns*Ls*nr <= 512 and the "
              "LDS of the shape within the device limit", env, MR_MIN_NS);
      return DWFGPU_BAD_ARG;
    }
    *best = MrShape{ns, nr};
    h->mr_ns[pi] = ns;
    h->mr_nr[pi] = nr;
    h->mr_nrhs[pi] = nrhs;
    return DWFGPU_OK;
  }
  const auto cand = mr_candidates(h, nrhs, sizeof(typename P::R));
  if (cand.empty()) {
    set_err("mr_tune: no launch shape of the batched operator fits Ls %d", h->ls);
    return DWFGPU_BAD_ARG;
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_err("mr_tune: hipEventCreate failed");
    return DWFGPU_HIP_ERROR;
  }
  float best_ms = 1e30f;
  int st = DWFGPU_OK;
  for (const MrShape& t : cand) {
    float tot = 0;
    for (int rep = 0; rep < 3 && !st; ++rep) {
      if (hipEventRecord(e0, h->st) != hipSuccess) st = DWFGPU_HIP_ERROR;
      if (!st)
        st = mr_mdagm<P>(h, t, nrhs, m.rs, m.ps, m.te, m.t1, m.mass, all_active, m.partial, nullptr,
                         nullptr, nullptr, false);
      if (!st && (hipEventRecord(e1, h->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        st = DWFGPU_HIP_ERROR;
      float ms = 0;
      if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = DWFGPU_HIP_ERROR;
      if (rep > 0) tot += ms;  // the first run only warms up
    }
    if (st) break;
    if (tot < best_ms) {
      best_ms = tot;
      *best = t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) {
    set_err("mr_tune: timing failed");
    return st;
  }
  h->mr_ns[pi] = best->ns;
  h->mr_nr[pi] = best->nr;
  h->mr_nrhs[pi] = nrhs;
  return DWFGPU_OK;
}

template <class PS>
int invcg_mrhs(dwfgpu_s* h, const dwfgpu_cg_arg* arg, int nrhs, const double* masses,
               double delta, int flags, const DVec& sol, const DVec& src, double* true_res,
               int* iters, int* nrel, int* status) {
  typedef Ops<P64> H;
  typedef Ops<PS> L;
  const Geom& g = h->g;
  TRY(ensure_work<P64>(h, 2));
  TRY(ensure_fields<PS>(h));
  const DVec& r = h->f[DWFGPU_F64].work[0];
  const DVec& ax = h->f[DWFGPU_F64].work[1];
  MrState<PS> m;
  m.nrhs = nrhs;
  m.gx = mr_partials_stride(h);
  TRY(alloc_mvec<PS>(h, nrhs, &m.te));
  TRY(alloc_mvec<PS>(h, nrhs, &m.t1));
  TRY(alloc_mvec<PS>(h, nrhs, &m.rs));
  TRY(alloc_mvec<PS>(h, nrhs, &m.ps));
  TRY(alloc_mvec<P32>(h, nrhs, &m.xs));
  TRY(alloc(h, (void**)&m.partial, (size_t)nrhs * m.gx * sizeof(double)));
  TRY(alloc(h, (void**)&m.scal, (size_t)nrhs * dwfk::MR_NSCAL * sizeof(double)));
  TRY(alloc(h, (void**)&m.mass, (size_t)nrhs * sizeof(double)));
  TRY(alloc(h, (void**)&m.ibuf, (size_t)3 * nrhs * sizeof(int)));
  HIPCHK(hipMemcpyAsync(m.mass, masses, nrhs * sizeof(double), hipMemcpyHostToDevice, h->st));

  std::vector<double> sc((size_t)nrhs * dwfk::MR_NSCAL, 0.0), src_nsq(nrhs, 0.0);
  std::vector<int> act(nrhs, 0), flg(nrhs, 0), its(nrhs, 0), done(nrhs, 0);
  const int max_itr = arg->max_num_iter - 1;
  for (int j = 0; j < nrhs; ++j) {
    nrel[j] = 0;
    status[j] = DWFGPU_OK;
  }

  // autotune with every RHS marked active (operands: the zeroed sloppy fields)
  {
    std::vector<int> ones(nrhs, 1);
    HIPCHK(hipMemcpyAsync(m.ibuf, ones.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, h->st));
  }
  MrShape t{h->ns, 1};  // Mobius: the handle's own hop shape, one RHS per work-group
  if (h->action == DWFGPU_DWF) TRY(mr_tune<PS>(h, nrhs, m, m.active(), &t));
  h->mr_last_ns = t.ns;
  h->mr_last_nr = t.nr;

  // initial residuals in FP64, one RHS at a time
  for (int j = 0; j < nrhs; ++j) {
    const DVec xj = mslice<P64>(h, sol, j), bj = mslice<P64>(h, src, j);
    TRY(H::norm(h, bj, &src_nsq[j]));
    const double stp = src_nsq[j] * arg->stop_rsd * arg->stop_rsd;
    TRY(H::mdagm(h, masses[j], ax, xj, nullptr));
    TRY(H::axpy_to(h, r, ax, bj, -1.0, true));
    double rr;
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
    double* sj = &sc[(size_t)j * dwfk::MR_NSCAL];
    sj[dwfk::MR_RR] = rr;
    sj[dwfk::MR_RRUPD] = rr;
    sj[dwfk::MR_STOP] = stp;
    sj[dwfk::MR_DELTA2] = delta * delta;
    sj[dwfk::MR_MAXIT] = max_itr;
    if (rr <= stp || max_itr <= 0) {
      done[j] = 1;
      true_res[j] = std::sqrt(rr / src_nsq[j]);
      if (rr > stp) status[j] = DWFGPU_MAXITER;
      continue;
    }
    const DVec rsj = mslice<PS>(h, m.rs, j), psj = mslice<PS>(h, m.ps, j);
    hipLaunchKernelGGL((dwfk::k_convert<PS, P64>), blas_grid(), blas_block(), 0, h->st, fld<PS>(rsj),
                       cfld<P64>(r), g);
    HIPCHK(hipGetLastError());
    TRY(L::copy(h, psj, rsj));
    act[j] = 1;
  }
  HIPCHK(hipMemsetAsync(m.xs.d, 0, (size_t)nrhs * (h->n / 2) * sizeof(cx<float>), h->st));
  HIPCHK(hipMemcpyAsync(m.scal, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(m.ibuf, act.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(m.flag(), flg.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(m.iters(), its.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, h->st));

  const size_t vs = h->n / 2 / PS::kCPC;
  auto step = [&]() -> int {
    if (h->action == DWFGPU_DWF)
      TRY(mr_mdagm<PS>(h, t, nrhs, DVec(), m.ps, m.te, m.t1, m.mass, m.active(), m.partial,
                       m.scal, &m.rs, &m.xs, true));
    else
      TRY(mr_mdagm_mob<PS>(h, nrhs, DVec(), m.ps, m.te, m.t1, masses, m.active(), m.partial,
                           m.scal, &m.rs, &m.xs, true));
    hipLaunchKernelGGL(dwfk::k_cg_scalars_mrhs, dim3(1), dim3(256), 0, h->st, m.scal, m.active(),
                       m.flag(), m.iters(), nrhs);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL((dwfk::k_xpby_mrhs<PS>), dim3(dwfk::BLAS_BLOCKS / 4, nrhs), blas_block(), 0,
                       h->st, fld<PS>(m.ps), cfld<PS>(m.rs), m.scal, m.active(), vs, g);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  };

  // K CG steps per host round, captured once into a graph (no memset or
  // copy inside it) unless DWFGPU_MRHS_EAGER asks for plain launches
  const int K = 4;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int st = DWFGPU_OK;
  if (!(flags & DWFGPU_MRHS_EAGER)) {
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < K && !st; ++k) st = step();
    hipGraph_t gcap = nullptr;
    const hipError_t ec = hipStreamEndCapture(h->st, &gcap);
    if (st || ec != hipSuccess) {
      if (gcap) (void)hipGraphDestroy(gcap);
      set_err("invcg_mrhs: graph capture failed");
      return st ? st : DWFGPU_HIP_ERROR;
    }
    graph = gcap;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      set_err("invcg_mrhs: graph instantiate failed");
      return DWFGPU_HIP_ERROR;
    }
  }
  auto cleanup = [&]() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  };
  // every RHS either retires or gains at least one iteration per round, and
  // its iterations are capped at max_itr, so the loop ends
  while (!st) {
    int nact = 0;
    for (int j = 0; j < nrhs; ++j) nact += act[j];
    if (nact == 0) break;
    if (exec) {
      if (hipGraphLaunch(exec, h->st) != hipSuccess) st = DWFGPU_HIP_ERROR;
    } else {
      for (int k = 0; k < K && !st; ++k) st = step();
    }
    if (st) break;
    std::vector<int> ib(3 * nrhs);
    if (hipMemcpyAsync(ib.data(), m.ibuf, ib.size() * sizeof(int), hipMemcpyDeviceToHost, h->st) !=
            hipSuccess ||
        hipMemcpyAsync(sc.data(), m.scal, sc.size() * sizeof(double), hipMemcpyDeviceToHost,
                       h->st) != hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess) {
      st = DWFGPU_HIP_ERROR;
      break;
    }
    bool changed = false;
    for (int j = 0; j < nrhs && !st; ++j) {
      act[j] = ib[j];
      flg[j] = ib[nrhs + j];
      its[j] = ib[2 * nrhs + j];
      if (flg[j] == 0) continue;
      changed = true;
      double* sj = &sc[(size_t)j * dwfk::MR_NSCAL];
      const DVec xj = mslice<P64>(h, sol, j), bj = mslice<P64>(h, src, j);
      const DVec rsj = mslice<PS>(h, m.rs, j), psj = mslice<PS>(h, m.ps, j);
      const DVec xsj = mslice<P32>(h, m.xs, j);
      if (flg[j] == 2) {  // <p, A p> == 0: CPS's breakdown
        status[j] = DWFGPU_BREAKDOWN;
        done[j] = 1;
        act[j] = 0;
        flg[j] = 0;
        continue;
      }
      // reliable update: x += xs, r = b - A x in FP64
      hipLaunchKernelGGL((dwfk::k_accumulate<P32>), blas_grid(), blas_block(), 0, h->st,
                         fld<P64>(xj), fld<P32>(xsj), g);
      if (hipGetLastError() != hipSuccess) st = DWFGPU_HIP_ERROR;
      if (!st) st = H::mdagm(h, masses[j], ax, xj, nullptr);
      if (!st) st = H::axpy_to(h, r, ax, bj, -1.0, true);
      double rr = 0;
      if (!st) st = fetch_sum(h, dwfk::BLAS_BLOCKS, &rr);
      if (st) break;
      ++nrel[j];
      if (rr <= sj[dwfk::MR_STOP] || its[j] >= max_itr || flg[j] == 3) {
        done[j] = 1;
        act[j] = 0;
        true_res[j] = std::sqrt(rr / src_nsq[j]);
        if (rr > sj[dwfk::MR_STOP]) status[j] = DWFGPU_MAXITER;
        flg[j] = 0;
        continue;
      }
      // continue: r_s = r, p_s = r_s + beta' p_s with beta' = rr / rr_prev
      hipLaunchKernelGGL((dwfk::k_convert<PS, P64>), blas_grid(), blas_block(), 0, h->st,
                         fld<PS>(rsj), cfld<P64>(r), g);
      if (hipGetLastError() != hipSuccess) st = DWFGPU_HIP_ERROR;
      const double beta = sj[dwfk::MR_BETA] * rr / sj[dwfk::MR_RR];
      if (!st) st = L::axpby(h, psj, rsj, 1.0, beta);
      sj[dwfk::MR_RR] = rr;
      sj[dwfk::MR_RRUPD] = rr;
      act[j] = 1;
      flg[j] = 0;
    }
    if (st || !changed) continue;
    for (int j = 0; j < nrhs; ++j) {
      ib[j] = act[j];
      ib[nrhs + j] = flg[j];
    }
    if (hipMemcpyAsync(m.scal, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice,
                       h->st) != hipSuccess ||
        hipMemcpyAsync(m.ibuf, ib.data(), 2 * nrhs * sizeof(int), hipMemcpyHostToDevice, h->st) !=
            hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess)
      st = DWFGPU_HIP_ERROR;
  }
  cleanup();
  if (st) {
    if (g_err.empty()) set_err("invcg_mrhs: HIP failure");
    return st;
  }
  for (int j = 0; j < nrhs; ++j) iters[j] = its[j] + 1;
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

}  // namespace

// ===========================================================================
// Multi-shift CG (DiracOp::MInvCG, d_op_base/noarch/minvcg.C:47-355), FP64
// ===========================================================================
namespace {

// psi: MULTI nshift vectors (mvec layout), SINGLE one vector; scalars and
// stop tests exactly as minvcg.C, the vector updates of all shifts fused
// into one pass per iteration (k_minv_update).
int minvcg_run(dwfgpu_s* h, double mass, const double* shift, int nshift, const DVec& psi,
               int multi, const double* alpha, const DVec& chi, double chi_norm, double* rsdcg,
               int max_num_iter, int* iters, int* k_out) {
  typedef Ops<P64> H;
  TRY(ensure_work<P64>(h, 2));
  const DVec& r = h->f[DWFGPU_F64].work[0];
  const DVec& ap = h->f[DWFGPU_F64].work[1];
  DVec pv;
  TRY(alloc_mvec<P64>(h, nshift, &pv));
  struct Free {
    DVec* v;
    ~Free() { free_vec(v); }
  } fr{&pv};
  const size_t vbytes = h->n / 2 * sizeof(cx<double>);
  dwfk::ShiftArgs<P64> A;
  std::memset(&A, 0, sizeof A);
  A.nshift = nshift;
  A.single = multi ? 0 : 1;
  for (int s = 0; s < nshift; ++s) {
    A.p[s] = (cx<double>*)mslice<P64>(h, pv, s).d;
    A.psi[s] = (cx<double>*)(multi ? mslice<P64>(h, psi, s).d : psi.d);
  }
  auto launch = [&](unsigned psi_mask, unsigned p_mask) -> int {
    A.psi_mask = psi_mask;
    A.p_mask = p_mask;
    hipLaunchKernelGGL((dwfk::k_minv_update<P64>), blas_grid(), blas_block(), 0, h->st, A,
                       cfld<P64>(r), h->g);
    HIPCHK(hipGetLastError());
    return DWFGPU_OK;
  };
  // d = <p0, (A + shift0) p0>
  DVec p0 = mslice<P64>(h, pv, 0);
  auto shifted_a = [&](double* d) -> int {
    if (shift[0] > 0) {
      TRY(H::mdagm(h, mass, ap, p0, nullptr));
      TRY(H::axpby(h, ap, p0, shift[0], 1.0));
      return dot<P64>(h, p0, ap, d);
    }
    return H::mdagm(h, mass, ap, p0, d);
  };
  if (multi) HIPCHK(hipMemsetAsync(psi.d, 0, (size_t)nshift * vbytes, h->st));  // :118-121
  HIPCHK(hipMemcpyAsync(r.d, chi.d, vbytes, hipMemcpyDeviceToDevice, h->st));
  for (int s = 0; s < nshift; ++s)
    HIPCHK(hipMemcpyAsync(A.p[s], chi.d, vbytes, hipMemcpyDeviceToDevice, h->st));
  double cp = chi_norm;
  std::vector<double> rsd_sq(nshift), bs(nshift), z0(nshift), z1(nshift), fin(nshift);
  double* z[2] = {z0.data(), z1.data()};
  std::vector<int> convs(nshift, 0), conv_at(nshift, 0);
  for (int s = 0; s < nshift; ++s) rsd_sq[s] = cp * rsdcg[s] * rsdcg[s];
  double d;
  TRY(shifted_a(&d));
  double b = -cp / d;
  z[0][0] = 1.0;
  z[1][0] = 1.0;
  bs[0] = b;
  int iz = 1;
  for (int s = 1; s < nshift; ++s) {
    z[1 - iz][s] = 1.0;
    z[iz][s] = 1.0 / (1.0 - b * (shift[s] - shift[0]));
    bs[s] = b * z[iz][s];
  }
  TRY(H::axpy_to(h, r, ap, r, b, true));  // r += b A p0
  double c;
  TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &c));
  // the solution update of the first step is pending (psi_s -= bs_s chi with
  // p_s = chi): it is applied with the next direction update or at the end
  unsigned pending = 0;
  for (int s = 0; s < nshift; ++s) {
    A.fpsi[s] = multi ? -bs[s] : -bs[s] * alpha[s];
    pending |= 1u << s;
  }
  int convp = c < rsd_sq[0] ? 1 : 0;
  double a = c / cp;
  int k;
  for (k = 1; k <= max_num_iter && !convp; k++) {
    a = c / cp;
    unsigned pmask = 0;
    for (int s = 0; s < nshift; ++s) {
      if (convs[s]) continue;
      pmask |= 1u << s;
      if (s == 0) {
        A.zr[0] = 1.0;
        A.ap[0] = a;
      } else {
        A.zr[s] = z[iz][s];
        A.ap[s] = a * (z[iz][s] * bs[s]) / (z[1 - iz][s] * b);
      }
    }
    TRY(launch(pending, pmask));
    pending = 0;
    cp = c;
    TRY(shifted_a(&d));
    const double bp = b;
    b = -cp / d;
    bs[0] = b;
    iz = 1 - iz;
    for (int s = 1; s < nshift; ++s) {
      if (convs[s]) continue;
      const double ztmp = z[1 - iz][s] * z[iz][s] * bp /
                          (b * a * (z[iz][s] - z[1 - iz][s]) +
                           z[iz][s] * bp * (1 - b * (shift[s] - shift[0])));
      bs[s] = b * ztmp / z[1 - iz][s];
      z[iz][s] = ztmp;
    }
    TRY(H::axpy_to(h, r, ap, r, b, true));
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &c));
    for (int s = 0; s < nshift; ++s) {
      if (convs[s]) continue;
      A.fpsi[s] = multi ? -bs[s] : -bs[s] * alpha[s];
      pending |= 1u << s;
    }
    for (int s = 0; s < nshift; ++s) {
      if (convs[s]) continue;
      const double css = c * z[iz][s] * z[iz][s];
      convs[s] = css < rsd_sq[s] ? 1 : 0;
      if (convs[s]) {
        fin[s] = css;
        conv_at[s] = k;
      }
    }
    convp = convs[0];
    if (convp)
      for (int s = 0; s < nshift; ++s)
        if (!convs[s]) convp = 0;
  }
  if (pending) TRY(launch(pending, 0u));
  HIPCHK(hipStreamSynchronize(h->st));
  for (int s = nshift - 1; s >= 0; s--) rsdcg[s] = std::sqrt(convs[s] ? fin[s] : c * z[iz][s] * z[iz][s]);
  if (iters)
    for (int s = 0; s < nshift; ++s) iters[s] = conv_at[s];
  if (k_out) *k_out = k;
  for (int s = 0; s < nshift; ++s)
    if (!convs[s]) return DWFGPU_MAXITER;
  return DWFGPU_OK;
}

// ---------------------------------------------------------------------------
// eigCG: DiracOp::InvEigCg (d_op_base/comsrc/inv_eig_cg.C:129-631) on an FP64
// handle, and the deflation-space bookkeeping of QPropW::eig_Run
// (alg/alg_qpropw/eig_QPropW.C:85-330).  The vectors (solution, residual,
// the eigCG window V, the deflation space U) stay in HBM; the small dense
// problems (the window's tridiagonal T, its projection H, invH) are solved on
// the host, where they are a few hundred numbers.
// ---------------------------------------------------------------------------
// Symmetric eigenproblem of the n x n matrix a (row-major, overwritten) by
// cyclic Jacobi rotations: w ascending, z[i * n + k] = component i of the
// k-th eigenvector.
void sym_eig(std::vector<double> a, int n, std::vector<double>& w, std::vector<double>& z) {
  z.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double v = a[(size_t)i * n + j] * a[(size_t)i * n + j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - sn * akq;
          a[(size_t)k * n + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - sn * aqk;
          a[(size_t)q * n + k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double zkp = z[(size_t)k * n + p], zkq = z[(size_t)k * n + q];
          z[(size_t)k * n + p] = c * zkp - sn * zkq;
          z[(size_t)k * n + q] = sn * zkp + c * zkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(),
            [&](int x, int y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
  std::vector<double> zs((size_t)n * n);
  w.resize(n);
  for (int k = 0; k < n; ++k) {
    w[k] = a[(size_t)idx[k] * n + idx[k]];
    for (int i = 0; i < n; ++i) zs[(size_t)i * n + k] = z[(size_t)i * n + idx[k]];
  }
  z.swap(zs);
}

// inverse of the complex n x n matrix a (row-major) by Gauss-Jordan with
// partial pivoting (the role of invert_H_matrix, eig_QPropW.C:33-83)
bool cinverse(std::vector<std::complex<double>>& a, int n) {
  std::vector<std::complex<double>> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::abs(a[(size_t)r * n + c]) > std::abs(a[(size_t)piv * n + c])) piv = r;
    if (std::abs(a[(size_t)piv * n + c]) == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
        std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
      }
    const std::complex<double> d = 1.0 / a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] *= d;
      inv[(size_t)c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const std::complex<double> f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  a.swap(inv);
  return true;
}

struct EigWork {
  double* buf = nullptr;   // device: qz / coefficients
  double* red = nullptr;   // device: reduced dot products
  double* part = nullptr;  // device: per-block partial dot products
  ~EigWork() {
    if (buf) (void)hipFree(buf);
    if (red) (void)hipFree(red);
    if (part) (void)hipFree(part);
  }
  int init(int m) {
    const size_t nbuf = std::max((size_t)m * dwfk::EIG_MAXRANK, (size_t)2 * dwfk::EIG_DOTK);
    HIPCHK(hipMalloc(&buf, nbuf * sizeof(double)));
    HIPCHK(hipMalloc(&red, 2 * dwfk::EIG_DOTK * sizeof(double)));
    HIPCHK(hipMalloc(&part, (size_t)2 * dwfk::EIG_DOTK * dwfk::EIG_BLOCKS * sizeof(double)));
    return DWFGPU_OK;
  }
};

// out[k] = <W_k, x> (complex) for k < K; W: K vectors `stride` reals apart
int eig_mdot(dwfgpu_s* h, EigWork& w, const double* W, size_t stride, int K, const double* x,
             std::vector<std::complex<double>>& out) {
  out.assign(K, 0.0);
  const int nb = dwfk::EIG_BLOCKS;
  std::vector<double> hv(2 * dwfk::EIG_DOTK);
  for (int k0 = 0; k0 < K; k0 += dwfk::EIG_DOTK) {
    const int kk = std::min(dwfk::EIG_DOTK, K - k0);
    hipLaunchKernelGGL(dwfk::k_eig_mdot<double>, dim3(nb), dim3(256), 0, h->st, W + (size_t)k0 * stride,
                       stride, kk, x, h->n / 2, w.part);
    hipLaunchKernelGGL(dwfk::k_eig_reduce_rows, dim3(2 * kk), dim3(256), 0, h->st, w.part, nb,
                       w.red);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hv.data(), w.red, 2 * kk * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (int k = 0; k < kk; ++k) out[k0 + k] = {hv[2 * k], hv[2 * k + 1]};
  }
  return DWFGPU_OK;
}

// out += sum_k c_k W_k
int eig_mcombo(dwfgpu_s* h, EigWork& w, double* out, const double* W, size_t stride, int K,
               const std::vector<std::complex<double>>& c) {
  for (int k0 = 0; k0 < K; k0 += dwfk::EIG_DOTK) {
    const int kk = std::min(dwfk::EIG_DOTK, K - k0);
    std::vector<double> hv(2 * kk);
    for (int k = 0; k < kk; ++k) {
      hv[2 * k] = c[k0 + k].real();
      hv[2 * k + 1] = c[k0 + k].imag();
    }
    HIPCHK(hipMemcpyAsync(w.buf, hv.data(), 2 * kk * sizeof(double), hipMemcpyHostToDevice, h->st));
    hipLaunchKernelGGL(dwfk::k_eig_mcombo<double>, blas_grid(), blas_block(), 0, h->st, out,
                       W + (size_t)k0 * stride, stride, kk, w.buf, h->n / 2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));  // w.buf is reused by the next chunk
  }
  return DWFGPU_OK;
}

// e = U invH U^dag v  (the init-CG guess of inv_eig_cg.C:202-225, 514-537)
int eig_deflate(dwfgpu_s* h, EigWork& w, const DVec& e, const double* U, int def_len,
                const std::vector<std::complex<double>>& invH, const DVec& v) {
  std::vector<std::complex<double>> ub, c(def_len, 0.0);
  TRY(eig_mdot(h, w, U, h->n, def_len, (const double*)v.d, ub));
  for (int i = 0; i < def_len; ++i)
    for (int j = 0; j < def_len; ++j) c[i] += invH[(size_t)i * def_len + j] * ub[j];
  HIPCHK(hipMemsetAsync(e.d, 0, h->n * sizeof(double), h->st));
  return eig_mcombo(h, w, (double*)e.d, U, h->n, def_len, c);
}

// Flexible GCR(m) with the network as preconditioner (no CPS counterpart;
// InvCg's stop rule on the FP64 true residual decides when it is done).  Per step: z = N(r),
// w = A z made orthogonal to the cycle's earlier w_j by classical
// Gram-Schmidt with complex coefficients c_j = <w_j, w> / |w_j|^2 (the
// network is not Hermitian), p = z updated with the same coefficients, then
// x += alpha p, r -= alpha w with alpha = <w, r> / |w|^2, which minimises
// |r| over the cycle's space; the cycle restarts after m directions.  All of
// it in FP64 with the FP64 operator.  A step that cuts |r|^2 by less than
// 0.1 % counts against the network; after 10 such steps, after a cycle that
// does not halve |r|^2, or without a network, plain CG (InvCg) finishes from
// the current solution.  The end is always the true residual |src - A sol|.
int invgcr_ptc(dwfgpu_s* h, const dwfgpu_cg_arg* arg, int m, const DVec& sol, const DVec& src,
               double* true_res, int* iters, int* net_steps, int* restarts) {
  typedef Ops<P64> H;
  TRY(ensure_work<P64>(h, 6));  // work[0..2] stay free for the CG that may finish
  const DVec& r = h->f[DWFGPU_F64].work[3];
  const DVec& z = h->f[DWFGPU_F64].work[4];
  const DVec& w = h->f[DWFGPU_F64].work[5];
  const size_t n = h->n;
  const double mass = arg->mass;
  struct Dirs {  // p_j and w_j = A p_j of the cycle, m vectors each
    double* p = nullptr;
    double* w = nullptr;
    ~Dirs() {
      if (p) (void)hipFree(p);
      if (w) (void)hipFree(w);
    }
  } B;
  HIPCHK(hipMalloc(&B.p, (size_t)m * n * sizeof(double)));
  HIPCHK(hipMalloc(&B.w, (size_t)m * n * sizeof(double)));
  EigWork ew;
  TRY(ew.init(1));
  double sn;
  TRY(H::norm(h, src, &sn));
  const double stp = sn * arg->stop_rsd * arg->stop_rsd;
  auto true_residual = [&](double* rr) -> int {  // r = src - A sol, |r|^2
    TRY(H::mdagm(h, mass, w, sol, nullptr));
    TRY(H::axpy_to(h, r, w, src, -1.0, true));
    return fetch_sum(h, dwfk::BLAS_BLOCKS, rr);
  };
  double rr;
  TRY(true_residual(&rr));
  std::vector<double> wn(m, 0.0);
  double rr_cycle = rr;  // |r|^2 at the start of the cycle
  int itr = 0, k = 0, nsteps = 0, nbad = 0, ncycle = 0, status = DWFGPU_OK;
  const int max_itr = arg->max_num_iter - 1;
  bool use_net = h->net != nullptr;
  while (use_net && rr > stp && itr < max_itr) {
    ++itr;
    TRY(ptc_apply<P64>(h, z, r));
    double zz;
    TRY(H::norm(h, z, &zz));
    if (!std::isfinite(zz) || zz == 0.0) TRY(H::copy(h, z, r));
    TRY(H::mdagm(h, mass, w, z, nullptr));
    double* pk = B.p + (size_t)k * n;
    double* wk = B.w + (size_t)k * n;
    HIPCHK(hipMemcpyAsync(pk, z.d, n * sizeof(double), hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(wk, w.d, n * sizeof(double), hipMemcpyDeviceToDevice, h->st));
    if (k > 0) {
      std::vector<std::complex<double>> c;
      TRY(eig_mdot(h, ew, B.w, n, k, wk, c));
      for (int j = 0; j < k; ++j) c[j] = -c[j] / wn[j];
      TRY(eig_mcombo(h, ew, wk, B.w, n, k, c));
      TRY(eig_mcombo(h, ew, pk, B.p, n, k, c));
    }
    DVec wkv;
    wkv.d = wk;
    TRY(H::norm(h, wkv, &wn[k]));
    if (!(wn[k] > 0.0) || !std::isfinite(wn[k])) {
      status = DWFGPU_BREAKDOWN;
      break;
    }
    std::vector<std::complex<double>> a;
    TRY(eig_mdot(h, ew, wk, n, 1, (const double*)r.d, a));
    const std::complex<double> alpha = a[0] / wn[k];
    TRY(eig_mcombo(h, ew, (double*)sol.d, pk, n, 1, {alpha}));
    TRY(eig_mcombo(h, ew, (double*)r.d, wk, n, 1, {-alpha}));
    double rrn;
    TRY(H::norm(h, r, &rrn));
    if (rrn > (1.0 - 1e-3) * rr) {
      if (++nbad >= 10) use_net = false;
    } else {
      ++nsteps;
    }
    rr = rrn;
    if (++k == m) {
      k = 0;
      ++ncycle;
      // a cycle of m network directions that does not halve |r|^2 loses to
      // CG (about 8 % per step on the 24^3x64 Ls 12 Mobius system, mf 0.005)
      if (rr > 0.5 * rr_cycle) use_net = false;
      rr_cycle = rr;
    }
  }
  if (status == DWFGPU_OK) {
    TRY(true_residual(&rr));
    if (rr > stp && itr < max_itr) {
      // plain CG from the current solution (the fallback)
      dwfgpu_cg_arg a2 = *arg;
      a2.max_num_iter = arg->max_num_iter - itr;
      double tr2 = 0.0;
      int it2 = 0;
      status = H::invcg(h, &a2, sol, src, &tr2, &it2);
      if (status != DWFGPU_OK && status != DWFGPU_MAXITER) return status;
      itr += it2 - 1;
      rr = tr2 * tr2 * sn;
    } else if (rr > stp) {
      status = DWFGPU_MAXITER;
    }
  }
  if (true_res) *true_res = std::sqrt(rr / sn);
  if (iters) *iters = itr + 1;
  if (net_steps) *net_steps = nsteps;
  if (restarts) *restarts = ncycle;
  return status;
}

int inveigcg_run(dwfgpu_s* h, const dwfgpu_cg_arg* arg, const DVec& sol, const DVec& src,
                 int nev, int m, double* V, double* M, const double* U,
                 const std::vector<std::complex<double>>& invH, int def_len,
                 const double* restart, int restart_len, double* true_res, int* iters) {
  typedef Ops<P64> H;
  TRY(ensure_work<P64>(h, 6));
  const DVec& r = h->f[DWFGPU_F64].work[3];
  const DVec& dir = h->f[DWFGPU_F64].work[4];
  const DVec& mmp = h->f[DWFGPU_F64].work[5];
  DVec mmp_prev;
  struct Free {
    DVec* v;
    ~Free() { free_vec(v); }
  } fr{&mmp_prev};
  if (nev > 0) TRY(alloc_vec<P64>(h, &mmp_prev));
  EigWork w;
  TRY(w.init(m));
  const size_t n = h->n;
  const double mass = arg->mass;
  double src_norm_sq;
  TRY(H::norm(h, src, &src_norm_sq));
  const double stp_cnd = src_norm_sq * arg->stop_rsd * arg->stop_rsd;
  // initial guess: 0, or U invH U^dag src
  if (def_len > 0)
    TRY(eig_deflate(h, w, sol, U, def_len, invH, src));
  else
    HIPCHK(hipMemsetAsync(sol.d, 0, n * sizeof(double), h->st));
  TRY(H::mdagm(h, mass, mmp, sol, nullptr));
  TRY(H::axpy_to(h, r, mmp, src, -1.0, true));  // r = src - A sol
  double rr;
  TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
  TRY(H::copy(h, dir, r));
  // eigCG state: T (m x m, symmetric, full storage here), the window count
  const int nev2 = 2 * nev;
  std::vector<double> T((size_t)m * m, 0.0);
  if (nev > 0)
    for (int i = 0; i < nev2; ++i) M[i] = 0.0;
  int i_eig = 0;
  double alpha = 1.0, beta = 0.0, alpha_old = 0.0, beta_old = 0.0;
  auto Tij = [&](int i, int j) -> double& { return T[(size_t)i * m + j]; };
  auto vptr = [&](int k) { return V + (size_t)k * n; };
  std::vector<double> restartcond(restart_len);
  for (int i = 0; i < restart_len; ++i) restartcond[i] = restart[i] * restart[i] * src_norm_sq;
  int restarted = 0;
  int iter;
  bool breakdown = false;
  const int max_iter = arg->max_num_iter;
  for (iter = 0; iter < max_iter; iter++) {
    if (nev > 0 && i_eig == m) TRY(H::copy(h, mmp_prev, mmp));
    const double rr_prv = rr;
    double pap;
    TRY(H::mdagm(h, mass, mmp, dir, &pap));
    if (pap == 0.0) {
      breakdown = true;
      break;
    }
    if (nev > 0) {
      if (i_eig != 0) Tij(i_eig - 1, i_eig - 1) = 1.0 / alpha + beta_old / alpha_old;
      if (i_eig == m) {
        // the window is full: keep the nev lowest Ritz vectors of T(m) and of
        // T(m-1) (inv_eig_cg.C:318-411), orthonormalised, and the Ritz pairs
        // of T in their span
        std::vector<double> Q((size_t)m * nev2, 0.0), ev, Y;
        std::vector<double> Tm1((size_t)(m - 1) * (m - 1));
        for (int i = 0; i < m - 1; ++i)
          for (int j = 0; j < m - 1; ++j) Tm1[(size_t)i * (m - 1) + j] = Tij(i, j);
        sym_eig(Tm1, m - 1, ev, Y);
        for (int c = 0; c < nev; ++c)
          for (int j = 0; j < m - 1; ++j) Q[(size_t)j * nev2 + nev + c] = Y[(size_t)j * (m - 1) + c];
        sym_eig(T, m, ev, Y);
        for (int c = 0; c < nev; ++c)
          for (int j = 0; j < m; ++j) Q[(size_t)j * nev2 + c] = Y[(size_t)j * m + c];
        int rank = nev;
        for (int i = nev; i < nev2; ++i) {
          for (int j = 0; j < rank; ++j) {
            double xy = 0.0;
            for (int k = 0; k < m; ++k) xy += Q[(size_t)k * nev2 + i] * Q[(size_t)k * nev2 + j];
            for (int k = 0; k < m; ++k) Q[(size_t)k * nev2 + i] -= xy * Q[(size_t)k * nev2 + j];
          }
          double xx = 0.0;
          for (int k = 0; k < m; ++k) xx += Q[(size_t)k * nev2 + i] * Q[(size_t)k * nev2 + i];
          if (xx > 1e-16) {
            xx = 1.0 / std::sqrt(xx);
            for (int k = 0; k < m; ++k) Q[(size_t)k * nev2 + rank] = xx * Q[(size_t)k * nev2 + i];
            rank++;
          }
        }
        // Hq = Q^T T Q (rank x rank), its eigenpairs, QZ = Q Z (m x rank)
        std::vector<double> TQ((size_t)m * rank, 0.0), Hq((size_t)rank * rank, 0.0);
        for (int l = 0; l < m; ++l)
          for (int j = 0; j < rank; ++j) {
            double a = 0.0;
            for (int k = 0; k < m; ++k) a += Tij(l, k) * Q[(size_t)k * nev2 + j];
            TQ[(size_t)l * rank + j] = a;
          }
        for (int i = 0; i < rank; ++i)
          for (int j = 0; j < rank; ++j) {
            double a = 0.0;
            for (int l = 0; l < m; ++l) a += Q[(size_t)l * nev2 + i] * TQ[(size_t)l * rank + j];
            Hq[(size_t)i * rank + j] = a;
          }
        std::vector<double> mz, Z;
        sym_eig(Hq, rank, mz, Z);
        for (int i = 0; i < nev2; ++i) M[i] = i < rank ? mz[i] : 0.0;
        std::vector<double> qz((size_t)m * rank, 0.0);  // qz[i + m j]
        for (int j = 0; j < rank; ++j)
          for (int i = 0; i < m; ++i) {
            double a = 0.0;
            for (int k = 0; k < rank; ++k) a += Q[(size_t)i * nev2 + k] * Z[(size_t)k * rank + j];
            qz[i + (size_t)m * j] = a;
          }
        HIPCHK(hipMemcpyAsync(w.buf, qz.data(), qz.size() * sizeof(double), hipMemcpyHostToDevice,
                              h->st));
        hipLaunchKernelGGL(dwfk::k_eig_vmul, blas_grid(), blas_block(),
                           (size_t)m * rank * sizeof(double), h->st, V, n, n, m, rank, w.buf);
        HIPCHK(hipGetLastError());
        std::fill(T.begin(), T.end(), 0.0);
        for (int i = 0; i < rank; ++i) Tij(i, i) = M[i];
        i_eig = rank;
        // w = A r = mmp - beta mmp_prev; T(rank, ii) = <w, V_ii> / |r|
        TRY(H::axpby(h, mmp_prev, mmp, 1.0, -beta));
        std::vector<std::complex<double>> tw;
        TRY(eig_mdot(h, w, V, n, rank, (const double*)mmp_prev.d, tw));
        for (int ii = 0; ii < i_eig; ++ii) {
          Tij(i_eig, ii) = tw[ii].real() / std::sqrt(rr);
          Tij(ii, i_eig) = Tij(i_eig, ii);
        }
      } else if (i_eig != 0) {
        Tij(i_eig, i_eig - 1) = -std::sqrt(beta) / alpha;
        Tij(i_eig - 1, i_eig) = Tij(i_eig, i_eig - 1);
      }
      // V[i_eig] = r / |r|
      DVec vi;
      vi.d = vptr(i_eig);
      TRY(H::copy(h, vi, r));
      TRY(H::scale(h, vi, 1.0 / std::sqrt(rr)));
      i_eig++;
    }
    alpha_old = alpha;
    alpha = -rr_prv / pap;
    TRY(H::axpy_to(h, r, mmp, r, alpha, true));  // r += alpha A p, |r|^2
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
    alpha = -alpha;
    beta_old = beta;
    beta = rr / rr_prv;
    // sol += alpha dir ; dir = r + beta dir
    hipLaunchKernelGGL((dwfk::k_cg_xp<P64, P64, true>), blas_grid(), blas_block(), 0, h->st,
                       fld<P64>(sol), fld<P64>(dir), cfld<P64>(r), alpha, beta, h->g);
    HIPCHK(hipGetLastError());
    if (restarted < restart_len && def_len > 0 && rr < restartcond[restarted]) {
      // init-CG again on the current residual (inv_eig_cg.C:504-559)
      restarted++;
      TRY(eig_deflate(h, w, dir, U, def_len, invH, r));
      TRY(H::axpby(h, sol, dir, 1.0, 1.0));
      TRY(H::mdagm(h, mass, mmp, dir, nullptr));
      TRY(H::axpy_to(h, r, mmp, r, -1.0, true));
      TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &rr));
      TRY(H::copy(h, dir, r));
      TRY(H::copy(h, mmp, r));  // as CPS, whose mmp holds the residual here
    }
    if (rr <= stp_cnd) break;
  }
  // true residual (inv_eig_cg.C:590-601)
  TRY(H::mdagm(h, mass, mmp, sol, nullptr));
  TRY(H::axpy_to(h, r, mmp, src, -1.0, true));
  double tr;
  TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &tr));
  if (true_res) *true_res = std::sqrt(tr / src_norm_sq);
  if (iters) *iters = iter;
  if (breakdown) return DWFGPU_BREAKDOWN;
  return iter == max_iter ? DWFGPU_MAXITER : DWFGPU_OK;
}

// ---------------------------------------------------------------------------
// DiracOp::ImpResLanczos (d_op_base/comsrc/lanczos.C:193-840) with the
// Chebyshev filter of DiracOpDwf::RitzMat (d_op_dwf/d_op_dwf.C:852-940) on
// MatPcDagMatPc.  Lanczos vectors V: m = nk + np single-precision device
// vectors; r, the filter's temporaries and the Ritz vectors of the checks
// are FP64.  The tridiagonal algebra runs on the host.
// ---------------------------------------------------------------------------
// out = T_Npol(Q) in, Q = c1/2 (c0 A - 1)
int lz_filter(dwfgpu_s* h, const dwfgpu_lanczos_arg* a, const DVec& out, const DVec& in,
              const DVec& t1, const DVec& t2) {
  typedef Ops<P64> H;
  const double al = a->p0 * a->p0, be = (a->p1 + std::fabs(a->shift)) * (a->p1 + std::fabs(a->shift));
  const double c1 = 2.0 * (al + be) / (al - be), c0 = 2.0 / (al + be);
  DVec tmp = t1, tmp2 = t2;
  TRY(H::copy(h, tmp2, in));
  TRY(H::mdagm(h, a->mass, tmp, in, nullptr));
  TRY(H::axpby(h, tmp, in, -0.5 * c1, 0.5 * c1 * c0));  // tmp = 0.5 c1 (c0 A in - in)
  if (a->npol == 1) return H::copy(h, out, tmp);
  for (int i = 2; i <= a->npol; ++i) {
    TRY(H::mdagm(h, a->mass, out, tmp, nullptr));
    TRY(H::axpby(h, out, tmp, -c1, c1 * c0));   // out = c1 (c0 A tmp - tmp)
    TRY(H::axpby(h, out, tmp2, -1.0, 1.0));     // out -= tmp2
    if (i != a->npol) {
      std::swap(tmp, tmp2);
      TRY(H::copy(h, tmp, out));
    }
  }
  return DWFGPU_OK;
}

// One shifted QR step of the symmetric tridiagonal (al, be): T - mu = QR,
// T <- RQ + mu, Qacc <- Qacc Q (Givens rotations on the dense m x m form).
void lz_qr_step(std::vector<double>& al, std::vector<double>& be, std::vector<double>& Qacc,
                int m, double mu) {
  std::vector<double> T((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) {
    T[(size_t)i * m + i] = al[i] - mu;
    if (i + 1 < m) T[(size_t)i * m + i + 1] = T[(size_t)(i + 1) * m + i] = be[i];
  }
  std::vector<double> cs(m), sn(m);
  // R = G_{m-2}^T ... G_0^T (T - mu)
  for (int k = 0; k + 1 < m; ++k) {
    const double a = T[(size_t)k * m + k], b = T[(size_t)(k + 1) * m + k];
    const double rr = std::hypot(a, b);
    const double c = rr == 0.0 ? 1.0 : a / rr, s = rr == 0.0 ? 0.0 : b / rr;
    cs[k] = c;
    sn[k] = s;
    for (int j = 0; j < m; ++j) {
      const double x = T[(size_t)k * m + j], y = T[(size_t)(k + 1) * m + j];
      T[(size_t)k * m + j] = c * x + s * y;
      T[(size_t)(k + 1) * m + j] = -s * x + c * y;
    }
  }
  // T = R Q + mu, Q = G_0 ... G_{m-2}; Qacc <- Qacc Q
  for (int k = 0; k + 1 < m; ++k) {
    const double c = cs[k], s = sn[k];
    for (int i = 0; i < m; ++i) {
      const double x = T[(size_t)i * m + k], y = T[(size_t)i * m + k + 1];
      T[(size_t)i * m + k] = c * x + s * y;
      T[(size_t)i * m + k + 1] = -s * x + c * y;
      const double qx = Qacc[(size_t)i * m + k], qy = Qacc[(size_t)i * m + k + 1];
      Qacc[(size_t)i * m + k] = c * qx + s * qy;
      Qacc[(size_t)i * m + k + 1] = -s * qx + c * qy;
    }
  }
  for (int i = 0; i < m; ++i) {
    al[i] = T[(size_t)i * m + i] + mu;
    if (i + 1 < m) be[i] = T[(size_t)i * m + i + 1];
  }
}

struct LzState {
  dwfgpu_s* h;
  const dwfgpu_lanczos_arg* a;
  float* V;  // m vectors, n floats apart
  DVec r, f1, f2, t1, t2;
  std::vector<double> al, be;
  EigWork w;
  size_t n;
};

// V[k0] = r / |r|, then Lanczos steps k0 .. m-1 (lanczos.C:756-840)
int lz_extend(LzState& S, int k0, int m) {
  dwfgpu_s* h = S.h;
  typedef Ops<P64> H;
  const size_t n = S.n;
  double rr;
  TRY(H::norm(h, S.r, &rr));
  hipLaunchKernelGGL(dwfk::k_lz_to_f32, blas_grid(), blas_block(), 0, h->st, S.V + (size_t)k0 * n,
                     (const double*)S.r.d, 1.0 / std::sqrt(rr), n);
  for (int k = k0; k < m; ++k) {
    // f1 = V[k] (FP64), r = filter(f1) - be[k-1] V[k-1]
    hipLaunchKernelGGL(dwfk::k_lz_to_f64, blas_grid(), blas_block(), 0, h->st, (double*)S.f1.d,
                       S.V + (size_t)k * n, n);
    HIPCHK(hipGetLastError());
    TRY(lz_filter(h, S.a, S.r, S.f1, S.t1, S.t2));
    if (k > 0) {
      hipLaunchKernelGGL(dwfk::k_lz_to_f64, blas_grid(), blas_block(), 0, h->st, (double*)S.f2.d,
                         S.V + (size_t)(k - 1) * n, n);
      TRY(H::axpby(h, S.r, S.f2, -S.be[k - 1], 1.0));
    }
    double ak;
    TRY(dot<P64>(h, S.f1, S.r, &ak));
    S.al[k] = ak;
    TRY(H::axpy_to(h, S.r, S.f1, S.r, -ak, true));
    double bk;
    TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &bk));
    S.be[k] = std::sqrt(bk);
    if (k > 0) {
      // Gram-Schmidt of r against V[0..k], one vector at a time; the complex
      // coefficient stays on the device
      for (int i = 0; i <= k; ++i) {
        hipLaunchKernelGGL(dwfk::k_eig_mdot<float>, dim3(dwfk::EIG_BLOCKS), dim3(256), 0, h->st,
                           S.V + (size_t)i * n, n, 1, (const double*)S.r.d, n / 2, S.w.part);
        hipLaunchKernelGGL(dwfk::k_eig_reduce_rows, dim3(2), dim3(256), 0, h->st, S.w.part,
                           dwfk::EIG_BLOCKS, S.w.red);
        hipLaunchKernelGGL(dwfk::k_lz_gs_axpy, blas_grid(), blas_block(), 0, h->st,
                           (double*)S.r.d, S.V + (size_t)i * n, S.w.red, n / 2);
      }
      HIPCHK(hipGetLastError());
    }
    if (k + 1 < m)
      hipLaunchKernelGGL(dwfk::k_lz_to_f32, blas_grid(), blas_block(), 0, h->st,
                         S.V + (size_t)(k + 1) * n, (const double*)S.r.d, 1.0 / S.be[k], n);
  }
  HIPCHK(hipGetLastError());
  return DWFGPU_OK;
}

// Ritz vectors of the leading nk x nk block, descending in the filtered
// value: z[j * nk + k] = component j of vector k
void lz_ritz(const std::vector<double>& al, const std::vector<double>& be, int nk,
             std::vector<double>& w, std::vector<double>& z) {
  std::vector<double> T((size_t)nk * nk, 0.0);
  for (int i = 0; i < nk; ++i) {
    T[(size_t)i * nk + i] = al[i];
    if (i + 1 < nk) T[(size_t)i * nk + i + 1] = T[(size_t)(i + 1) * nk + i] = be[i];
  }
  std::vector<double> wa, za;
  sym_eig(T, nk, wa, za);  // ascending
  w.resize(nk);
  z.assign((size_t)nk * nk, 0.0);
  for (int k = 0; k < nk; ++k) {
    w[k] = wa[nk - 1 - k];
    for (int j = 0; j < nk; ++j) z[(size_t)j * nk + k] = za[(size_t)j * nk + (nk - 1 - k)];
  }
}

// out (FP64) = sum_{j<nk} z[j * nk + k] V[j]
int lz_combine(LzState& S, const DVec& out, const std::vector<double>& z, int nk, int k) {
  std::vector<std::complex<double>> c(nk);
  for (int j = 0; j < nk; ++j) c[j] = z[(size_t)j * nk + k];
  HIPCHK(hipMemsetAsync(out.d, 0, S.n * sizeof(double), S.h->st));
  for (int k0 = 0; k0 < nk; k0 += dwfk::EIG_DOTK) {
    const int kk = std::min(dwfk::EIG_DOTK, nk - k0);
    std::vector<double> hv(2 * kk);
    for (int q = 0; q < kk; ++q) {
      hv[2 * q] = c[k0 + q].real();
      hv[2 * q + 1] = 0.0;
    }
    HIPCHK(hipMemcpyAsync(S.w.buf, hv.data(), 2 * kk * sizeof(double), hipMemcpyHostToDevice,
                          S.h->st));
    hipLaunchKernelGGL(dwfk::k_eig_mcombo<float>, blas_grid(), blas_block(), 0, S.h->st,
                       (double*)out.d, S.V + (size_t)k0 * S.n, S.n, kk, S.w.buf, S.n / 2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(S.h->st));
  }
  return DWFGPU_OK;
}

int lanczos_run(dwfgpu_s* h, const dwfgpu_lanczos_arg* a, const DVec& v0, double* evec_host,
                const DVec* evec_dev, double* lambda, int* iters) {
  typedef Ops<P64> H;
  const int nk = a->nk, np = a->np, nt = a->nt, m = nk + np;
  LzState S;
  S.h = h;
  S.a = a;
  S.n = h->n;
  S.al.assign(m, 0.0);
  S.be.assign(m, 0.0);
  TRY(S.w.init(std::max(m, 1)));
  DVec vbuf;
  struct Free {
    std::vector<DVec*> v;
    ~Free() {
      for (DVec* p : v) free_vec(p);
    }
  } fr;
  fr.v = {&vbuf, &S.r, &S.f1, &S.f2, &S.t1, &S.t2};
  TRY(alloc(h, &vbuf.d, (size_t)m * S.n * sizeof(float)));
  S.V = (float*)vbuf.d;
  for (DVec* p : {&S.r, &S.f1, &S.f2, &S.t1, &S.t2}) TRY(alloc_vec<P64>(h, p));
  double* dq = nullptr;
  HIPCHK(hipMalloc(&dq, (size_t)m * m * sizeof(double)));
  struct FreeQ {
    double* p;
    ~FreeQ() {
      if (p) (void)hipFree(p);
    }
  } fq{dq};
  TRY(H::copy(h, S.r, v0));
  TRY(lz_extend(S, 0, m));
  int it = 1;
  std::vector<double> w, z;
  while (it <= a->maxiters) {
    // shifts: the Ritz values of T_m, descending; the np smallest go out
    std::vector<double> T((size_t)m * m, 0.0), ev, Y;
    for (int i = 0; i < m; ++i) {
      T[(size_t)i * m + i] = S.al[i];
      if (i + 1 < m) T[(size_t)i * m + i + 1] = T[(size_t)(i + 1) * m + i] = S.be[i];
    }
    sym_eig(T, m, ev, Y);  // ascending: ev[0..np) are the np smallest
    std::vector<double> Q((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Q[(size_t)i * m + i] = 1.0;
    // CPS applies shifts[nk + i], i < np, of the descending list
    for (int i = 0; i < np; ++i) lz_qr_step(S.al, S.be, Q, m, ev[np - 1 - i]);
    HIPCHK(hipMemcpyAsync(dq, Q.data(), Q.size() * sizeof(double), hipMemcpyHostToDevice, h->st));
    const int ncol = nk < m ? nk + 1 : nk;
    hipLaunchKernelGGL(dwfk::k_lz_vq, dim3(1024), dim3(dwfk::LZ_TILE),
                       (size_t)m * dwfk::LZ_TILE * sizeof(float), h->st, S.V, S.n, S.n, m, ncol, dq);
    HIPCHK(hipGetLastError());
    // r <- Q[m-1][nk-1] r + be[nk-1] V[nk]; be[nk-1] = |r|
    TRY(H::scale(h, S.r, Q[(size_t)(m - 1) * m + nk - 1]));
    if (nk < m) {
      hipLaunchKernelGGL(dwfk::k_lz_to_f64, blas_grid(), blas_block(), 0, h->st, (double*)S.f1.d,
                         S.V + (size_t)nk * S.n, S.n);
      TRY(H::axpby(h, S.r, S.f1, S.be[nk - 1], 1.0));
    }
    double rr;
    TRY(H::norm(h, S.r, &rr));
    S.be[nk - 1] = std::sqrt(rr);
    if (it % std::max(a->conv_check, 1) == 0) {
      lz_ritz(S.al, S.be, nk, w, z);
      int done = 0;
      for (int k = 0; k < nk; ++k) {
        TRY(lz_combine(S, S.f1, z, nk, k));
        TRY(H::mdagm(h, a->mass, S.f2, S.f1, nullptr));
        double v2, avv;
        TRY(H::norm(h, S.f1, &v2));
        TRY(dot<P64>(h, S.f2, S.f1, &avv));
        const double bt = std::sqrt(v2), alp = avv / bt;
        TRY(H::axpy_to(h, S.f2, S.f1, S.f2, -alp, true));
        double res2;
        TRY(fetch_sum(h, dwfk::BLAS_BLOCKS, &res2));
        if (std::sqrt(res2) / bt < a->stop_residual && k < nt) done++;
      }
      if (done == nt) break;
    }
    TRY(lz_extend(S, nk, m));
    it++;
  }
  if (it == a->maxiters + 1) it--;
  // the nt Ritz vectors, rounded to single precision as CPS stores them, and
  // their Rayleigh quotients (lanczos.C:588-700)
  lz_ritz(S.al, S.be, nk, w, z);
  for (int k = 0; k < nt; ++k) {
    TRY(lz_combine(S, S.f1, z, nk, k));
    hipLaunchKernelGGL(dwfk::k_eig_round_f32, blas_grid(), blas_block(), 0, h->st,
                       (double*)S.f1.d, (double*)S.f1.d, S.n);
    TRY(H::mdagm(h, a->mass, S.f2, S.f1, nullptr));
    double lam;
    TRY(dot<P64>(h, S.f1, S.f2, &lam));
    lambda[k] = lam;
    if (evec_dev) {
      DVec o;
      o.d = (char*)evec_dev->d + (size_t)k * S.n * sizeof(double);
      TRY(H::copy(h, o, S.f1));
    }
    if (evec_host) TRY(H::to_host(h, evec_host + (size_t)k * S.n, S.f1));
  }
  HIPCHK(hipStreamSynchronize(h->st));
  if (iters) *iters = it;
  return DWFGPU_OK;
}

}  // namespace

// the deflation space of QPropW::eig_Run: U (float-valued FP64 vectors on the
// device), H = U^dag A U, the eigCG window V and its Ritz values M
struct dwfgpu_eigcg_s {
  dwfgpu_s* h = nullptr;
  int nev = 0, m = 0, max_def_len = 0, def_len = 0;
  double max_eig = 0.0;
  DVec U, V;
  std::vector<std::complex<double>> H;  // max_def_len^2, row-major
  std::vector<double> M;
};

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* dwfgpu_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// allocation-free MatPcDagMatPc (dwfgpu_matpcdagmatpc_dev_noalloc): its
// temporaries live in one static device array, which every GPU the library
// runs on holds from module load to the end of the process.  In host code
// `scratch` only names the symbol; its device address (one per GPU) comes from
// hipGetSymbolAddress, once per device, when the first handle is created there.
// ---------------------------------------------------------------------------
extern "C++" {
__device__ __align__(256) unsigned char scratch[DWFGPU_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxDevices] = {};

// with `dev` current: look up this device's copy of `scratch`
int scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_err("scratch_setup: device %d beyond the %d the library tracks", dev, kMaxDevices);
    return DWFGPU_BAD_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev]) return DWFGPU_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_err("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    return DWFGPU_HIP_ERROR;
  }
  g_scratch_dev[dev] = (unsigned char*)ptr;
  return DWFGPU_OK;
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// carve te, t1 (with their scales in the 16-bit formats), grid_hop partials
// and one device scalar out of base (nullptr: only count); returns the bytes
template <class P>
size_t carve_scratch(const dwfgpu_s* h, unsigned char* base, Scratch* w) {
  const size_t vb = round256((size_t)(h->n / 2) * sizeof(typename P::S));
  const size_t sb = P::kScaled ? round256((size_t)h->g.vh * h->ls * sizeof(float)) : 0;
  size_t off = 0;
  auto take = [&](size_t bytes) -> unsigned char* {
    unsigned char* p = base ? base + off : nullptr;
    off += bytes;
    return p;
  };
  DVec te, t1;
  te.d = take(vb);
  te.sc = P::kScaled ? (float*)take(sb) : nullptr;
  t1.d = take(vb);
  t1.sc = P::kScaled ? (float*)take(sb) : nullptr;
  double* partial = (double*)take(round256((size_t)h->grid_hop * sizeof(double)));
  double* dscal = (double*)take(256);
  if (w) *w = Scratch{te, t1, partial, dscal};
  return off;
}
size_t scratch_need(const dwfgpu_s* h) {
  switch (h->prec) {
    case DWFGPU_F64: return carve_scratch<P64>(h, nullptr, nullptr);
    case DWFGPU_F32: return carve_scratch<P32>(h, nullptr, nullptr);
    case DWFGPU_BF16: return carve_scratch<PBF16>(h, nullptr, nullptr);
    default: return carve_scratch<PI16>(h, nullptr, nullptr);
  }
}
}  // namespace
}  // extern "C++"

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
  if (precision < DWFGPU_F64 || precision > DWFGPU_I16) {
    set_err("dwfgpu_create: unknown precision %d", precision);
    return DWFGPU_BAD_ARG;
  }
  dwfgpu_s* h = new dwfgpu_s;
  h->action = action;
  h->prec = precision;
  for (int mu = 0; mu < 4; ++mu) {
    h->dims[mu] = dims[mu];
    h->bc[mu] = bc[mu] ? 1 : 0;
  }
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
  // tiles of 8 t-slices x 4 z-slices when the extents allow (the fastest of
  // 14 shapes timed on 24^3x64 Ls 12 Mobius: 4.12 ms against 4.36 ms for
  // 4 x 4); 8 XCDs on gfx950 (a different count only changes the
  // work-group order, not the result)
  int bt = dims[3] % 8 == 0 ? 8 : dims[3] % 4 == 0 ? 4 : 2;
  int bz = dims[2] % 4 == 0 ? 4 : 2;
  // DWFGPU_TILE="bt,bz" overrides the tile (each must divide its extent)
  if (const char* env = getenv("DWFGPU_TILE")) {
    int et = 0, ez = 0;
    if (sscanf(env, "%d,%d", &et, &ez) == 2 && et > 0 && ez > 0 && dims[3] % et == 0 &&
        dims[2] % ez == 0) {
      bt = et;
      bz = ez;
    }
  }
  const int bcmask = h->bc[0] | h->bc[1] << 1 | h->bc[2] << 2 | h->bc[3] << 3;
  h->g = Geom{dims[0], dims[1], dims[2], dims[3], dims[0] / 2, (int)(v4 / 2), ls, bt, bz, 8,
              bcmask};
  h->n = (size_t)24 * (v4 / 2) * ls;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&h->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess) {
    set_err("dwfgpu_create: no HIP device");
    delete h;
    return DWFGPU_HIP_ERROR;
  }
  h->dev = dev;
  if (int rs = scratch_setup(dev)) {
    delete h;
    return rs;
  }
  // sites per work-group of the hopping kernels; DWFGPU_NS overrides the
  // default.  Without an override, fewer sites per work-group are taken when
  // the LDS stage of a long fifth dimension would not fit.
  const char* env_ns = getenv("DWFGPU_NS");
  if (env_ns) h->ns = atoi(env_ns);
  while (!env_ns && h->ns > 4 && max_lds_bytes(h) > (size_t)h->max_lds) h->ns /= 2;
  if (h->ns < 1 || h->ns * ls > 512) {
    set_err("dwfgpu_create: bad launch shape ns=%d (ns*Ls <= 512)", h->ns);
    delete h;
    return DWFGPU_BAD_ARG;
  }
  if (max_lds_bytes(h) > (size_t)h->max_lds) {
    set_err("dwfgpu_create: Ls=%d needs %zu bytes of LDS per work-group (limit %d)", ls,
            max_lds_bytes(h), h->max_lds);
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
  if ((r = alloc(h, (void**)&h->nbt, (size_t)2 * 8 * h->g.vh * sizeof(int)))) return fail(r);
  hipLaunchKernelGGL(dwfk::k_build_nbt, dim3(2048), dim3(256), 0, h->st, h->nbt, h->g);
  if (hipGetLastError() != hipSuccess) {
    set_err("dwfgpu_create: neighbour table kernel failed to launch");
    return fail(DWFGPU_HIP_ERROR);
  }
  h->g.nbt = h->nbt;
  if ((r = upload_gauge(h, gauge))) return fail(r);
  if ((r = ensure_api(h, 0))) return fail(r);
  if ((r = alloc(h, &h->cps_stage, h->n * sizeof(double)))) return fail(r);
  if ((r = alloc(h, (void**)&h->partial, h->npartial * sizeof(double)))) return fail(r);
  if ((r = alloc(h, (void**)&h->dscal, 8 * sizeof(double)))) return fail(r);
  if ((r = alloc(h, (void**)&h->dalpha, 32 * sizeof(double)))) return fail(r);
  if (hipHostMalloc((void**)&h->hring, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    set_err("dwfgpu_create: hipHostMalloc failed");
    return fail(DWFGPU_HIP_ERROR);
  }
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
  TRY(ensure_api(h, 2));
  TRY(DISPATCH(h, to_device(h, api_work(h, 0), in)));
  TRY(DISPATCH(h, mdagm(h, mass, api_work(h, 1), api_work(h, 0), dot_prd)));
  return DISPATCH(h, to_host(h, out, api_work(h, 1)));
}

int dwfgpu_matpc(dwfgpu_t h, double mass, double* out, const double* in, int dag) {
  if (!h || !out || !in || (dag != 0 && dag != 1)) return DWFGPU_BAD_ARG;
  TRY(ensure_api(h, 2));
  TRY(DISPATCH(h, to_device(h, api_work(h, 0), in)));
  TRY(DISPATCH(h, matpc(h, mass, api_work(h, 1), api_work(h, 0), dag, false)));
  return DISPATCH(h, to_host(h, out, api_work(h, 1)));
}

int dwfgpu_dslash(dwfgpu_t h, double mass, double* out, const double* in, int cb, int dag) {
  if (!h || !out || !in || (dag != 0 && dag != 1) || (cb != 0 && cb != 1)) return DWFGPU_BAD_ARG;
  TRY(ensure_api(h, 2));
  TRY(DISPATCH(h, to_device(h, api_work(h, 0), in)));
  TRY(DISPATCH(h, dslash(h, mass, api_work(h, 1), api_work(h, 0), cb, dag)));
  return DISPATCH(h, to_host(h, out, api_work(h, 1)));
}

int dwfgpu_invcg_shift(dwfgpu_t h, const dwfgpu_cg_arg* arg, double* sol, const double* src,
                       double src_norm_sq, double* true_res, const double* shift, int* iters) {
  if (!h || !arg || !sol || !src) return DWFGPU_BAD_ARG;
  if (arg->max_num_iter < 1 || src_norm_sq < 0.0) {
    set_err("dwfgpu_invcg: need max_num_iter >= 1 and src_norm_sq >= 0");
    return DWFGPU_BAD_ARG;
  }
  TRY(ensure_api(h, 5));
  const DVec& dsol = api_work(h, 3);
  const DVec& dsrc = api_work(h, 4);
  TRY(DISPATCH(h, to_device(h, dsol, sol)));
  TRY(DISPATCH(h, to_device(h, dsrc, src)));
  const int st = DISPATCH(h, invcg(h, arg, dsol, dsrc, true_res, iters, src_norm_sq, shift));
  if (st != DWFGPU_OK && st != DWFGPU_MAXITER) return st;
  TRY(DISPATCH(h, to_host(h, sol, dsol)));
  return st;
}

int dwfgpu_invcg(dwfgpu_t h, const dwfgpu_cg_arg* arg, double* sol, const double* src,
                 double* true_res, int* iters) {
  return dwfgpu_invcg_shift(h, arg, sol, src, 0.0, true_res, nullptr, iters);
}

int dwfgpu_matinv(dwfgpu_t h, const dwfgpu_cg_arg* arg, double* out, const double* in,
                  double* true_res, int* iters) {
  if (!h || !arg || !out || !in) return DWFGPU_BAD_ARG;
  if (arg->max_num_iter < 1) return DWFGPU_BAD_ARG;
  TRY(ensure_api(h, 6));
  DVec v[4];
  int st = DWFGPU_OK;
  for (int k = 0; k < 4 && !st; ++k)
    st = alloc_api_vec(h, &v[k]);
  if (!st) st = DISPATCH(h, to_device(h, v[0], in));
  if (!st) st = DISPATCH(h, to_device(h, v[1], in + h->n));
  if (!st) st = DISPATCH(h, to_device(h, v[2], out));
  if (!st) st = DISPATCH(h, to_device(h, v[3], out + h->n));
  // the CG works in work[0..2]; MatInv temporaries are work[3..5]
  if (!st) st = DISPATCH(h, matinv(h, arg, v[2], v[3], v[0], v[1], true_res, iters));
  if (st == DWFGPU_OK || st == DWFGPU_MAXITER) {
    int r = DISPATCH(h, to_host(h, out, v[2]));
    if (!r) r = DISPATCH(h, to_host(h, out + h->n, v[3]));
    if (r) st = r;
  }
  (void)hipStreamSynchronize(h->st);
  for (auto& x : v) free_vec(&x);
  return st;
}

int dwfgpu_invcg_mixed(dwfgpu_t h, const dwfgpu_cg_arg* arg, int sloppy, double delta,
                       double* sol, const double* src, double* true_res, int* iters,
                       int* reliable_updates) {
  if (!h || !arg || !sol || !src) return DWFGPU_BAD_ARG;
  dwfgpu_vec_t vs = nullptr, vb = nullptr;
  TRY(dwfgpu_vec_create(h, &vs));
  int st = dwfgpu_vec_create(h, &vb);
  if (!st) st = dwfgpu_vec_upload(vs, sol);
  if (!st) st = dwfgpu_vec_upload(vb, src);
  if (!st) st = dwfgpu_invcg_mixed_dev(h, arg, sloppy, delta, vs, vb, true_res, iters,
                                       reliable_updates);
  if (st == DWFGPU_OK || st == DWFGPU_MAXITER) {
    const int r = dwfgpu_vec_download(vs, sol);
    if (r) st = r;
  }
  if (vs) dwfgpu_vec_destroy(vs);
  if (vb) dwfgpu_vec_destroy(vb);
  return st;
}

int dwfgpu_vec_create(dwfgpu_t h, dwfgpu_vec_t* v) {
  if (!h || !v) return DWFGPU_BAD_ARG;
  dwfgpu_vec_s* p = new dwfgpu_vec_s;
  p->h = h;
  const int st = alloc_api_vec(h, &p->v);
  if (st) {
    free_vec(&p->v);
    delete p;
    return st;
  }
  HIPCHK(hipStreamSynchronize(h->st));
  *v = p;
  return DWFGPU_OK;
}

int dwfgpu_vec_destroy(dwfgpu_vec_t v) {
  if (!v) return DWFGPU_BAD_ARG;
  (void)hipStreamSynchronize(v->h->st);
  free_vec(&v->v);
  delete v;
  return DWFGPU_OK;
}

int dwfgpu_vec_upload(dwfgpu_vec_t v, const double* host) {
  if (!v || !host) return DWFGPU_BAD_ARG;
  TRY(DISPATCH(v->h, to_device(v->h, v->v, host)));
  HIPCHK(hipStreamSynchronize(v->h->st));
  return DWFGPU_OK;
}

int dwfgpu_vec_download(dwfgpu_vec_t v, double* host) {
  if (!v || !host) return DWFGPU_BAD_ARG;
  return DISPATCH(v->h, to_host(v->h, host, v->v));
}

int dwfgpu_matpcdagmatpc_dev(dwfgpu_t h, double mass, dwfgpu_vec_t out, dwfgpu_vec_t in,
                             double* dot_prd) {
  if (!h || !out || !in || out->h != h || in->h != h) return DWFGPU_BAD_ARG;
  TRY(DISPATCH(h, mdagm(h, mass, out->v, in->v, dot_prd)));
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

int dwfgpu_matpcdagmatpc_dev_noalloc(dwfgpu_t h, double mass, dwfgpu_vec_t out,
                                     dwfgpu_vec_t in, double* dot_prd) {
  if (!h || !out || !in || out->h != h || in->h != h) return DWFGPU_BAD_ARG;
  const size_t need = scratch_need(h);
  if (need > DWFGPU_NOALLOC_SCRATCH_BYTES) {
    set_err("dwfgpu_matpcdagmatpc_dev_noalloc: this lattice needs %zu bytes of scratch, the "
            "static array holds %zu (use dwfgpu_matpcdagmatpc_dev)",
            need, (size_t)DWFGPU_NOALLOC_SCRATCH_BYTES);
    return DWFGPU_BAD_ARG;
  }
  unsigned char* base = g_scratch_dev[h->dev];
  if (!base) {
    set_err("dwfgpu_matpcdagmatpc_dev_noalloc: no scratch address for device %d", h->dev);
    return DWFGPU_HIP_ERROR;
  }
  // every kernel writes all of te, t1 and the partials it later reads, so
  // nothing left by an earlier call is consumed and nothing needs clearing
  switch (h->prec) {
    case DWFGPU_F64: {
      Scratch w;
      carve_scratch<P64>(h, base, &w);
      TRY(Ops<P64>::mdagm_w(h, mass, out->v, in->v, dot_prd, w));
      break;
    }
    case DWFGPU_F32: {
      Scratch w;
      carve_scratch<P32>(h, base, &w);
      TRY(Ops<P32>::mdagm_w(h, mass, out->v, in->v, dot_prd, w));
      break;
    }
    case DWFGPU_BF16: {
      Scratch w;
      carve_scratch<PBF16>(h, base, &w);
      TRY(Ops<PBF16>::mdagm_w(h, mass, out->v, in->v, dot_prd, w));
      break;
    }
    default: {
      Scratch w;
      carve_scratch<PI16>(h, base, &w);
      TRY(Ops<PI16>::mdagm_w(h, mass, out->v, in->v, dot_prd, w));
      break;
    }
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

int dwfgpu_invcg_shift_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, dwfgpu_vec_t sol,
                           dwfgpu_vec_t src, double src_norm_sq, double* true_res,
                           const double* shift, int* iters) {
  if (!h || !arg || !sol || !src || sol->h != h || src->h != h) return DWFGPU_BAD_ARG;
  if (arg->max_num_iter < 1 || src_norm_sq < 0.0) {
    set_err("dwfgpu_invcg: need max_num_iter >= 1 and src_norm_sq >= 0");
    return DWFGPU_BAD_ARG;
  }
  return DISPATCH(h, invcg(h, arg, sol->v, src->v, true_res, iters, src_norm_sq, shift));
}

int dwfgpu_invcg_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, dwfgpu_vec_t sol, dwfgpu_vec_t src,
                     double* true_res, int* iters) {
  return dwfgpu_invcg_shift_dev(h, arg, sol, src, 0.0, true_res, nullptr, iters);
}

int dwfgpu_invcg_mixed_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, int sloppy, double delta,
                           dwfgpu_vec_t sol, dwfgpu_vec_t src, double* true_res, int* iters,
                           int* reliable_updates) {
  if (!h || !arg || !sol || !src || sol->h != h || src->h != h) return DWFGPU_BAD_ARG;
  if (arg->max_num_iter < 1 || !(delta > 0.0 && delta < 1.0)) {
    set_err("dwfgpu_invcg_mixed: need max_num_iter >= 1 and 0 < delta < 1");
    return DWFGPU_BAD_ARG;
  }
  if (h->prec != DWFGPU_F64) {
    set_err("dwfgpu_invcg_mixed: the handle's API precision must be DWFGPU_F64");
    return DWFGPU_BAD_ARG;
  }
  switch (sloppy) {
    case DWFGPU_F64:
      return Ops<P64>::invcg(h, arg, sol->v, src->v, true_res, iters);
    case DWFGPU_F32:
      return invcg_mixed<P32>(h, arg, delta, sol->v, src->v, true_res, iters, reliable_updates);
    case DWFGPU_BF16:
      return invcg_mixed<PBF16>(h, arg, delta, sol->v, src->v, true_res, iters, reliable_updates);
    case DWFGPU_I16:
      return invcg_mixed<PI16>(h, arg, delta, sol->v, src->v, true_res, iters, reliable_updates);
    default:
      set_err("dwfgpu_invcg_mixed: unknown sloppy precision %d", sloppy);
      return DWFGPU_BAD_ARG;
  }
}

int dwfgpu_matpcdagmatpc_cost(dwfgpu_t h, int precision, double* bytes, double* flops) {
  if (!h) return DWFGPU_BAD_ARG;
  double half, links;
  switch (precision) {
    case DWFGPU_F64: half = half_bytes<P64>(h); links = link_bytes<P64>(h); break;
    case DWFGPU_F32: half = half_bytes<P32>(h); links = link_bytes<P32>(h); break;
    case DWFGPU_BF16: half = half_bytes<PBF16>(h); links = link_bytes<PBF16>(h); break;
    case DWFGPU_I16: half = half_bytes<PI16>(h); links = link_bytes<PI16>(h); break;
    default: return DWFGPU_BAD_ARG;
  }
  // 10 half-vector passes and 4 link passes (include/roofline.h)
  if (bytes) *bytes = 10.0 * half + 4.0 * links;
  const double nh = (double)h->g.vh * h->ls;
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

int dwfgpu_time_bandwidth(dwfgpu_t h, size_t bytes, int reps, double* copy_gbs,
                          double* read_gbs, double* r3w1_gbs) {
  if (!h || reps < 1 || bytes < 16 || !copy_gbs || !read_gbs || !r3w1_gbs)
    return DWFGPU_BAD_ARG;
  typedef dwfk::u32x4_t W;
  const size_t n16 = bytes / 16;
  const int grids[3] = {2048, 4096, 8192};
  W* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  W* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = DWFGPU_OK;
  for (int k = 0; k < 4 && !st; ++k)
    if (hipMalloc(&buf[k], n16 * 16) != hipSuccess ||
        hipMemsetAsync(buf[k], k + 1, n16 * 16, h->st) != hipSuccess)
      st = DWFGPU_HIP_ERROR;
  if (!st && (hipMalloc(&sink, (size_t)grids[2] * 256 * 16) != hipSuccess ||
              hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = DWFGPU_HIP_ERROR;
  // bytes moved per launch: copy 2, read 1, three-read-one-write 4 buffers
  const double moved[3] = {2.0, 1.0, 4.0};
  double best[3] = {0.0, 0.0, 0.0};
  for (int mode = 0; mode < 3 && !st; ++mode)
    for (int gi = 0; gi < 3 && !st; ++gi) {
      auto launch = [&]() {
        const dim3 grid(grids[gi]), block(256);
        if (mode == 0)
          hipLaunchKernelGGL(dwfk::k_bw_probe<0>, grid, block, 0, h->st, buf[3], buf[0], buf[1],
                             buf[2], n16, sink);
        else if (mode == 1)
          hipLaunchKernelGGL(dwfk::k_bw_probe<1>, grid, block, 0, h->st, buf[3], buf[0], buf[1],
                             buf[2], n16, sink);
        else
          hipLaunchKernelGGL(dwfk::k_bw_probe<2>, grid, block, 0, h->st, buf[3], buf[0], buf[1],
                             buf[2], n16, sink);
      };
      launch();  // untimed: first touch
      if (hipEventRecord(e0, h->st) != hipSuccess) st = DWFGPU_HIP_ERROR;
      for (int r = 0; r < reps && !st; ++r) launch();
      float ms = 0;
      if (!st && (hipGetLastError() != hipSuccess || hipEventRecord(e1, h->st) != hipSuccess ||
                  hipEventSynchronize(e1) != hipSuccess ||
                  hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = DWFGPU_HIP_ERROR;
      if (!st) best[mode] = std::max(best[mode], moved[mode] * n16 * 16 * reps / (ms * 1e-3) / 1e9);
    }
  if (!st) {
    *copy_gbs = best[0];
    *read_gbs = best[1];
    *r3w1_gbs = best[2];
  } else {
    set_err("dwfgpu_time_bandwidth: HIP failure");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (W* p : buf)
    if (p) (void)hipFree(p);
  if (sink) (void)hipFree(sink);
  return st;
}

int dwfgpu_time_matpcdagmatpc(dwfgpu_t h, int precision, double mass, dwfgpu_vec_t out,
                              dwfgpu_vec_t in, int reps, double* ms_app, double* ms_kernel,
                              double* bytes_kernel, int* nkernel) {
  if (!h || !out || !in || reps < 1 || out->h != h || in->h != h) return DWFGPU_BAD_ARG;
  if (precision == h->prec) {
    switch (h->prec) {
      case DWFGPU_F64:
        return time_mdagm<P64>(h, mass, out->v, in->v, reps, ms_app, ms_kernel, bytes_kernel,
                               nkernel);
      case DWFGPU_F32:
        return time_mdagm<P32>(h, mass, out->v, in->v, reps, ms_app, ms_kernel, bytes_kernel,
                               nkernel);
      case DWFGPU_BF16:
        return time_mdagm<PBF16>(h, mass, out->v, in->v, reps, ms_app, ms_kernel, bytes_kernel,
                                 nkernel);
      default:
        return time_mdagm<PI16>(h, mass, out->v, in->v, reps, ms_app, ms_kernel, bytes_kernel,
                                nkernel);
    }
  }
  if (h->prec != DWFGPU_F64) {
    set_err("dwfgpu_time_matpcdagmatpc: a precision other than the handle's needs an FP64 handle");
    return DWFGPU_BAD_ARG;
  }
  // time a sloppy-precision copy of `in` (converted on the device)
  auto run = [&](auto tag) -> int {
    typedef decltype(tag) P;
    TRY(ensure_work<P>(h, 2));
    const DVec& a = h->f[pidx<P>()].work[0];
    const DVec& b = h->f[pidx<P>()].work[1];
    hipLaunchKernelGGL((dwfk::k_convert<P, P64>), blas_grid(), blas_block(), 0, h->st, fld<P>(a),
                       cfld<P64>(in->v), h->g);
    HIPCHK(hipGetLastError());
    return time_mdagm<P>(h, mass, b, a, reps, ms_app, ms_kernel, bytes_kernel, nkernel);
  };
  switch (precision) {
    case DWFGPU_F32: return run(P32());
    case DWFGPU_BF16: return run(PBF16());
    case DWFGPU_I16: return run(PI16());
    default: return DWFGPU_BAD_ARG;
  }
}

int dwfgpu_mvec_create(dwfgpu_t h, int nrhs, dwfgpu_mvec_t* out) {
  if (!h || !out || nrhs < 1) return DWFGPU_BAD_ARG;
  if (h->prec != DWFGPU_F64) {
    set_err("dwfgpu_mvec_create: needs an FP64 handle");
    return DWFGPU_BAD_ARG;
  }
  dwfgpu_mvec_s* v = new dwfgpu_mvec_s;
  v->h = h;
  v->nrhs = nrhs;
  int st = alloc_mvec<P64>(h, nrhs, &v->v);
  if (!st && hipStreamSynchronize(h->st) != hipSuccess) st = DWFGPU_HIP_ERROR;
  if (st) {
    free_vec(&v->v);
    delete v;
    return st;
  }
  *out = v;
  return DWFGPU_OK;
}

int dwfgpu_mvec_destroy(dwfgpu_mvec_t v) {
  if (!v) return DWFGPU_BAD_ARG;
  (void)hipStreamSynchronize(v->h->st);
  free_vec(&v->v);
  delete v;
  return DWFGPU_OK;
}

int dwfgpu_mvec_upload(dwfgpu_mvec_t v, int j, const double* host) {
  if (!v || !host || j < 0 || j >= v->nrhs) return DWFGPU_BAD_ARG;
  TRY(Ops<P64>::to_device(v->h, mslice<P64>(v->h, v->v, j), host));
  HIPCHK(hipStreamSynchronize(v->h->st));
  return DWFGPU_OK;
}

int dwfgpu_mvec_download(dwfgpu_mvec_t v, int j, double* host) {
  if (!v || !host || j < 0 || j >= v->nrhs) return DWFGPU_BAD_ARG;
  return Ops<P64>::to_host(v->h, host, mslice<P64>(v->h, v->v, j));
}

int dwfgpu_matpcdagmatpc_mrhs_dev(dwfgpu_t h, int precision, const double* masses,
                                  dwfgpu_mvec_t out, dwfgpu_mvec_t in) {
  if (!h || !masses || !out || !in || out->h != h || in->h != h || out->nrhs != in->nrhs)
    return DWFGPU_BAD_ARG;
  const int nrhs = in->nrhs;
  auto run = [&](auto tag) -> int {
    typedef decltype(tag) P;
    TRY(ensure_fields<P>(h));
    MrState<P> m;
    m.nrhs = nrhs;
    m.gx = mr_partials_stride(h);
    TRY(alloc_mvec<P>(h, nrhs, &m.te));
    TRY(alloc_mvec<P>(h, nrhs, &m.t1));
    TRY(alloc_mvec<P>(h, nrhs, &m.rs));
    TRY(alloc_mvec<P>(h, nrhs, &m.ps));
    TRY(alloc(h, (void**)&m.partial, (size_t)nrhs * m.gx * sizeof(double)));
    TRY(alloc(h, (void**)&m.mass, (size_t)nrhs * sizeof(double)));
    TRY(alloc(h, (void**)&m.ibuf, (size_t)3 * nrhs * sizeof(int)));
    std::vector<int> ones(nrhs, 1);
    HIPCHK(hipMemcpyAsync(m.ibuf, ones.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(m.mass, masses, nrhs * sizeof(double), hipMemcpyHostToDevice, h->st));
    for (int j = 0; j < nrhs; ++j)
      hipLaunchKernelGGL((dwfk::k_convert<P, P64>), blas_grid(), blas_block(), 0, h->st,
                         fld<P>(mslice<P>(h, m.ps, j)), cfld<P64>(mslice<P64>(h, in->v, j)), h->g);
    HIPCHK(hipGetLastError());
    MrShape t{h->ns, 1};
    if (h->action == DWFGPU_DWF) {
      TRY(mr_tune<P>(h, nrhs, m, m.active(), &t));
      TRY(mr_mdagm<P>(h, t, nrhs, m.rs, m.ps, m.te, m.t1, m.mass, m.active(), m.partial,
                      nullptr, nullptr, nullptr, false));
    } else {
      TRY(mr_mdagm_mob<P>(h, nrhs, m.rs, m.ps, m.te, m.t1, masses, m.active(), m.partial,
                          nullptr, nullptr, nullptr, false));
    }
    h->mr_last_ns = t.ns;
    h->mr_last_nr = t.nr;
    for (int j = 0; j < nrhs; ++j)
      hipLaunchKernelGGL((dwfk::k_convert<P64, P>), blas_grid(), blas_block(), 0, h->st,
                         fld<P64>(mslice<P64>(h, out->v, j)), cfld<P>(mslice<P>(h, m.rs, j)), h->g);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    return DWFGPU_OK;
  };
  switch (precision) {
    case DWFGPU_F64: return run(P64());
    case DWFGPU_F32: return run(P32());
    case DWFGPU_BF16: return run(PBF16());
    case DWFGPU_I16: return run(PI16());
    default: return DWFGPU_BAD_ARG;
  }
}

int dwfgpu_invcg_mrhs_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, int sloppy, double delta,
                          int flags, const double* masses, dwfgpu_mvec_t sol, dwfgpu_mvec_t src,
                          double* true_res, int* iters, int* reliable_updates, int* status) {
  if (!h || !arg || !masses || !sol || !src || !true_res || !iters || !reliable_updates ||
      !status || sol->h != h || src->h != h || sol->nrhs != src->nrhs)
    return DWFGPU_BAD_ARG;

  if (arg->max_num_iter < 1 || !(delta > 0.0 && delta < 1.0)) {
    set_err("dwfgpu_invcg_mrhs_dev: need max_num_iter >= 1 and 0 < delta < 1");
    return DWFGPU_BAD_ARG;
  }
  const int nrhs = sol->nrhs;
  switch (sloppy) {
    case DWFGPU_F32:
      return invcg_mrhs<P32>(h, arg, nrhs, masses, delta, flags, sol->v, src->v, true_res, iters,
                             reliable_updates, status);
    case DWFGPU_BF16:
      return invcg_mrhs<PBF16>(h, arg, nrhs, masses, delta, flags, sol->v, src->v, true_res,
                               iters, reliable_updates, status);
    case DWFGPU_I16:
      return invcg_mrhs<PI16>(h, arg, nrhs, masses, delta, flags, sol->v, src->v, true_res, iters,
                              reliable_updates, status);
    default:
      set_err("dwfgpu_invcg_mrhs_dev: sloppy precision must be F32, BF16 or I16");
      return DWFGPU_BAD_ARG;
  }
}

int dwfgpu_matinv_mrhs_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, int sloppy, double delta,
                           int flags, const double* masses, dwfgpu_mvec_t out_first,
                           dwfgpu_mvec_t out_second, dwfgpu_mvec_t in_first,
                           dwfgpu_mvec_t in_second, double* true_res, int* iters,
                           int* reliable_updates, int* status) {
  if (!h || !arg || !masses || !out_first || !out_second || !in_first || !in_second ||
      !true_res || !iters || !reliable_updates || !status)
    return DWFGPU_BAD_ARG;
  const int nrhs = in_first->nrhs;
  for (dwfgpu_mvec_t v : {out_first, out_second, in_first, in_second})
    if (v->h != h || v->nrhs != nrhs) {
      set_err("dwfgpu_matinv_mrhs_dev: the four vectors need this handle and one RHS count");
      return DWFGPU_BAD_ARG;
    }
  if (arg->max_num_iter < 1 || !(delta > 0.0 && delta < 1.0)) {
    set_err("dwfgpu_matinv_mrhs_dev: need max_num_iter >= 1 and 0 < delta < 1");
    return DWFGPU_BAD_ARG;
  }
  if (sloppy != DWFGPU_F32 && sloppy != DWFGPU_BF16 && sloppy != DWFGPU_I16) {
    set_err("dwfgpu_matinv_mrhs_dev: sloppy precision must be F32, BF16 or I16");
    return DWFGPU_BAD_ARG;
  }
  typedef Ops<P64> O;
  TRY(ensure_work<P64>(h, 6));
  struct Src {  // the CG sources of all RHS, freed on every return
    DVec v;
    ~Src() { free_vec(&v); }
  } src;
  TRY(alloc_mvec<P64>(h, nrhs, &src.v));
  // per RHS: the CG source from [in_first | in_second]
  for (int j = 0; j < nrhs; ++j)
    TRY(O::matinv_src(h, masses[j], mslice<P64>(h, src.v, j), mslice<P64>(h, in_first->v, j),
                      mslice<P64>(h, in_second->v, j)));
  const DVec& sol = O::cg_half(h, out_first->v, out_second->v);
  int st;
  switch (sloppy) {
    case DWFGPU_F32:
      st = invcg_mrhs<P32>(h, arg, nrhs, masses, delta, flags, sol, src.v, true_res, iters,
                           reliable_updates, status);
      break;
    case DWFGPU_BF16:
      st = invcg_mrhs<PBF16>(h, arg, nrhs, masses, delta, flags, sol, src.v, true_res, iters,
                             reliable_updates, status);
      break;
    default:
      st = invcg_mrhs<PI16>(h, arg, nrhs, masses, delta, flags, sol, src.v, true_res, iters,
                            reliable_updates, status);
      break;
  }
  if (st) {
    (void)hipStreamSynchronize(h->st);
    return st;
  }
  // per RHS: the other half of the solution
  for (int j = 0; j < nrhs; ++j)
    TRY(O::matinv_sol(h, masses[j], mslice<P64>(h, out_first->v, j),
                      mslice<P64>(h, out_second->v, j), mslice<P64>(h, in_first->v, j),
                      mslice<P64>(h, in_second->v, j)));
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

int dwfgpu_mrhs_shape(dwfgpu_t h, int* ns, int* nr) {
  if (!h || !ns || !nr) return DWFGPU_BAD_ARG;
  *ns = h->mr_last_ns;
  *nr = h->mr_last_nr;
  return DWFGPU_OK;
}

size_t dwfgpu_ptc_num_weights(dwfgpu_t h) {
  if (!h) return 0;
  const size_t ls = h->ls, hid = 16;
  return (size_t)ptck::NDIR * 16 * (hid * ls + hid * hid + ls * hid);
}

int dwfgpu_ptc_init(dwfgpu_t h, unsigned long long seed, double scale) {
  if (!h || h->prec != DWFGPU_F64) return DWFGPU_BAD_ARG;
  TRY(ptc_ensure(h));
  ptc_init_weights(h->net, seed, scale);
  return ptc_upload(h);
}

int dwfgpu_ptc_set_weights(dwfgpu_t h, const float* w, size_t n) {
  if (!h || !w || !h->net || n != dwfgpu_ptc_num_weights(h)) return DWFGPU_BAD_ARG;
  size_t q = 0;
  for (int k = 0; k < 3; ++k)
    for (auto& x : h->net->W[k]) {
      x = std::complex<float>(w[2 * q], w[2 * q + 1]);
      ++q;
    }
  return ptc_upload(h);
}

int dwfgpu_ptc_get_weights(dwfgpu_t h, float* w, size_t n) {
  if (!h || !w || !h->net || n != dwfgpu_ptc_num_weights(h)) return DWFGPU_BAD_ARG;
  size_t q = 0;
  for (int k = 0; k < 3; ++k)
    for (auto& x : h->net->W[k]) {
      w[2 * q] = x.real();
      w[2 * q + 1] = x.imag();
      ++q;
    }
  return DWFGPU_OK;
}

int dwfgpu_ptc_apply(dwfgpu_t h, double* out, const double* in) {
  if (!h || !out || !in || !h->net) return DWFGPU_BAD_ARG;
  TRY(ensure_work<P32>(h, 6));
  const DVec& a = h->f[DWFGPU_F32].work[4];
  const DVec& b = h->f[DWFGPU_F32].work[5];
  TRY(Ops<P32>::to_device(h, a, in));
  TRY(ptc_apply<P32>(h, b, a));
  return Ops<P32>::to_host(h, out, b);
}

int dwfgpu_ptc_loss_grad(dwfgpu_t h, const double* r, const double* e, double* loss,
                         float* grads) {
  if (!h || !r || !e || !h->net) return DWFGPU_BAD_ARG;
  TRY(ensure_work<P32>(h, 6));
  const DVec& a = h->f[DWFGPU_F32].work[4];
  const DVec& b = h->f[DWFGPU_F32].work[5];
  TRY(Ops<P32>::to_device(h, a, r));
  TRY(Ops<P32>::to_device(h, b, e));
  TRY(ptc_grad_pair(h, a, b, false, loss));
  if (grads) {
    size_t off = 0;
    for (int k = 0; k < 3; ++k) {
      HIPCHK(hipMemcpyAsync(grads + 2 * off, h->net->G[k], h->net->nw(k) * sizeof(cx<float>),
                            hipMemcpyDeviceToHost, h->st));
      off += h->net->nw(k);
    }
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

int dwfgpu_ptc_train(dwfgpu_t h, const dwfgpu_ptc_train_arg* a, double* loss_first,
                     double* loss_last, int* pairs_used) {
  if (!h || !a || !h->net || a->nsolves < 1 || a->pairs_per_solve < 0 || a->random_pairs < 0 ||
      a->pairs_per_solve + a->random_pairs < 1 || a->stride < 1 || a->epochs < 1 ||
      a->batch < 1 || !(a->lr > 0.0))
    return DWFGPU_BAD_ARG;
  const int P = a->pairs_per_solve + a->random_pairs;
  std::vector<DVec> R(P), E(P);
  DVec b64;
  int st = alloc_vec<P64>(h, &b64);
  for (int j = 0; j < P && !st; ++j) {
    st = alloc_vec<P32>(h, &R[j]);
    if (!st) st = alloc_vec<P32>(h, &E[j]);
  }
  std::vector<double> host(a->pairs_per_solve > 0 ? h->n : 0);
  double first = -1.0, last = 0.0;
  int used = 0;
  for (int sidx = 0; sidx < a->nsolves && !st; ++sidx) {
    if (a->pairs_per_solve > 0) {  // the source of the solve whose iterates give pairs
      dwfsynth_gaussian(host.data(), h->n, a->seed + 7919ull * sidx);
      st = Ops<P64>::to_device(h, b64, host.data());
    }
    int made = 0;
    if (!st && a->pairs_per_solve > 0)
      st = ptc_pairs(h, a->mass, b64, a->stride, a->pairs_per_solve, R, E, &made);
    // pairs whose correction is a Gaussian vector e (drawn on the device) and
    // residual r = A e: they carry the high modes the CG residuals lack
    for (int j = 0; j < a->random_pairs && !st; ++j) {
      hipLaunchKernelGGL(ptck::k_gauss_f32, blas_grid(), blas_block(), 0, h->st,
                         reinterpret_cast<float*>(E[made].d), h->n,
                         (uint64_t)(a->seed + 104729ull * (sidx * 977 + j + 1)));
      if (hipGetLastError() != hipSuccess) st = DWFGPU_HIP_ERROR;
      if (!st) st = Ops<P32>::mdagm(h, a->mass, R[made], E[made], nullptr);
      if (!st) ++made;
    }
    if (st) break;
    // interleave the two kinds in the mini-batches
    if (a->random_pairs > 0) {
      std::vector<int> order(made);
      for (int j = 0; j < made; ++j) order[j] = j;
      std::mt19937_64 rng(a->seed + sidx);
      std::shuffle(order.begin(), order.end(), rng);
      std::vector<DVec> R2(made), E2(made);
      for (int j = 0; j < made; ++j) {
        R2[j] = R[order[j]];
        E2[j] = E[order[j]];
      }
      for (int j = 0; j < made; ++j) {
        R[j] = R2[j];
        E[j] = E2[j];
      }
    }
    used += made;
    for (int ep = 0; ep < a->epochs && !st; ++ep) {
      double sum = 0.0;
      for (int j0 = 0; j0 < made && !st; j0 += a->batch) {
        const int j1 = std::min(made, j0 + a->batch);
        for (int j = j0; j < j1 && !st; ++j) {
          double l = 0.0;
          st = ptc_grad_pair(h, R[j], E[j], j > j0, &l);
          sum += l;
          if (first < 0.0) first = l;
        }
        if (!st) st = ptc_adam(h, j1 - j0, a->lr);
      }
      if (made > 0) last = sum / made;
    }
  }
  (void)hipStreamSynchronize(h->st);
  free_vec(&b64);
  for (int j = 0; j < P; ++j) {
    free_vec(&R[j]);
    free_vec(&E[j]);
  }
  if (loss_first) *loss_first = first;
  if (loss_last) *loss_last = last;
  if (pairs_used) *pairs_used = used;
  return st;
}

int dwfgpu_invcg_ptc_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, int sloppy, double delta,
                         dwfgpu_vec_t sol, dwfgpu_vec_t src, double* true_res, int* iters,
                         int* reliable_updates, int* net_steps) {
  if (!h || !arg || !sol || !src || sol->h != h || src->h != h) return DWFGPU_BAD_ARG;
  if (h->prec != DWFGPU_F64 || arg->max_num_iter < 1 || !(delta > 0.0 && delta < 1.0)) {
    set_err("dwfgpu_invcg_ptc_dev: FP64 handle, max_num_iter >= 1 and 0 < delta < 1 needed");
    return DWFGPU_BAD_ARG;
  }
  switch (sloppy) {
    case DWFGPU_F32:
      return invfcg_ptc<P32>(h, arg, delta, sol->v, src->v, true_res, iters, reliable_updates,
                             net_steps);
    case DWFGPU_I16:
      return invfcg_ptc<PI16>(h, arg, delta, sol->v, src->v, true_res, iters, reliable_updates,
                              net_steps);
    case DWFGPU_BF16:
      return invfcg_ptc<PBF16>(h, arg, delta, sol->v, src->v, true_res, iters, reliable_updates,
                               net_steps);
    default:
      set_err("dwfgpu_invcg_ptc_dev: sloppy precision must be F32, BF16 or I16");
      return DWFGPU_BAD_ARG;
  }
}

int dwfgpu_invgcr_ptc_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, int nkrylov, dwfgpu_vec_t sol,
                          dwfgpu_vec_t src, double* true_res, int* iters, int* net_steps,
                          int* restarts) {
  if (!h || !arg || !sol || !src || sol->h != h || src->h != h) return DWFGPU_BAD_ARG;
  if (h->prec != DWFGPU_F64 || arg->max_num_iter < 1 || nkrylov < 1 || nkrylov > 64) {
    set_err("dwfgpu_invgcr_ptc_dev: FP64 handle, max_num_iter >= 1 and 1 <= nkrylov <= 64 "
            "needed");
    return DWFGPU_BAD_ARG;
  }
  const int st = invgcr_ptc(h, arg, nkrylov, sol->v, src->v, true_res, iters, net_steps, restarts);
  (void)hipStreamSynchronize(h->st);
  return st;
}

int dwfgpu_minvcg_dev(dwfgpu_t h, double mass, const double* shifts, int nshift,
                      dwfgpu_mvec_t psi, dwfgpu_vec_t chi, double chi_norm, double* rsdcg,
                      int type, const double* alpha, int max_num_iter, int* iters, int* k) {
  if (!h || !shifts || !psi || !chi || !rsdcg || psi->h != h || chi->h != h) return DWFGPU_BAD_ARG;
  if (h->prec != DWFGPU_F64 || nshift < 1 || nshift > dwfk::MAXSHIFT ||
      (type != DWFGPU_SHIFT_SINGLE && type != DWFGPU_SHIFT_MULTI) ||
      (type == DWFGPU_SHIFT_SINGLE && !alpha) ||
      psi->nrhs != (type == DWFGPU_SHIFT_MULTI ? nshift : 1) || max_num_iter < 1) {
    set_err("dwfgpu_minvcg: FP64 handle, 1..%d shifts, type SINGLE (with alpha, one "
            "vector) or MULTI (nshift vectors) and max_num_iter >= 1 needed", dwfk::MAXSHIFT);
    return DWFGPU_BAD_ARG;
  }
  if (chi_norm <= 0.0) TRY(Ops<P64>::norm(h, chi->v, &chi_norm));
  return minvcg_run(h, mass, shifts, nshift, psi->v, type == DWFGPU_SHIFT_MULTI, alpha, chi->v,
                    chi_norm, rsdcg, max_num_iter, iters, k);
}

int dwfgpu_minvcg(dwfgpu_t h, double mass, const double* shifts, int nshift, double* psi,
                  const double* chi, double chi_norm, double* rsdcg, int type,
                  const double* alpha, int max_num_iter, int* iters, int* k) {
  if (!h || !psi || !chi) return DWFGPU_BAD_ARG;
  const int nv = type == DWFGPU_SHIFT_MULTI ? nshift : 1;
  if (nv < 1) return DWFGPU_BAD_ARG;
  dwfgpu_mvec_t mv = nullptr;
  dwfgpu_vec_t cv = nullptr;
  TRY(dwfgpu_mvec_create(h, nv, &mv));
  int st = dwfgpu_vec_create(h, &cv);
  if (!st) st = dwfgpu_vec_upload(cv, chi);
  if (!st && type == DWFGPU_SHIFT_SINGLE) st = dwfgpu_mvec_upload(mv, 0, psi);
  if (!st)
    st = dwfgpu_minvcg_dev(h, mass, shifts, nshift, mv, cv, chi_norm, rsdcg, type, alpha,
                           max_num_iter, iters, k);
  if (st == DWFGPU_OK || st == DWFGPU_MAXITER)
    for (int j = 0; j < nv; ++j) {
      const int r = dwfgpu_mvec_download(mv, j, psi + (size_t)j * h->n);
      if (r) st = r;
    }
  if (cv) dwfgpu_vec_destroy(cv);
  if (mv) dwfgpu_mvec_destroy(mv);
  return st;
}

// ---- host formats of whole fermion fields (convert.C:639-831,
// f_dwf_base.C:961-1168) ----------------------------------------------------
namespace {
int ensure_stage4d(dwfgpu_s* h) {
  if (!h->stage4d) TRY(alloc(h, (void**)&h->stage4d, (size_t)2 * h->g.vh * 24 * sizeof(double)));
  return DWFGPU_OK;
}
bool own_pair(dwfgpu_t h, dwfgpu_vec_t a, dwfgpu_vec_t b) {
  return h && a && b && a->h == h && b->h == h && a != b;
}
}  // namespace

int dwfgpu_fermion_from_canonical(dwfgpu_t h, dwfgpu_vec_t first, dwfgpu_vec_t second,
                                  const double* canon) {
  if (!own_pair(h, first, second) || !canon) return DWFGPU_BAD_ARG;
  TRY(ensure_stage4d(h));
  const size_t v4r = (size_t)2 * h->g.vh * 24;
  for (int s = 0; s < h->ls; ++s) {
    HIPCHK(hipMemcpyAsync(h->stage4d, canon + (size_t)s * v4r, v4r * sizeof(double),
                          hipMemcpyHostToDevice, h->st));
    TRY(DISPATCH(h, canon_to_halves(h, first->v, second->v, h->stage4d, s, 1, 1)));
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

int dwfgpu_fermion_to_canonical(dwfgpu_t h, double* canon, dwfgpu_vec_t first,
                                dwfgpu_vec_t second) {
  if (!own_pair(h, first, second) || !canon) return DWFGPU_BAD_ARG;
  TRY(ensure_stage4d(h));
  const size_t v4r = (size_t)2 * h->g.vh * 24;
  for (int s = 0; s < h->ls; ++s) {
    TRY(DISPATCH(h, halves_to_canon(h, h->stage4d, first->v, second->v, s, 1, 1)));
    HIPCHK(hipMemcpyAsync(canon + (size_t)s * v4r, h->stage4d, v4r * sizeof(double),
                          hipMemcpyDeviceToHost, h->st));
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

int dwfgpu_four2five(dwfgpu_t h, dwfgpu_vec_t first, dwfgpu_vec_t second, const double* four,
                     int s_u, int s_l) {
  if (!own_pair(h, first, second) || !four) return DWFGPU_BAD_ARG;
  if (s_u < 0 || s_u >= h->ls || s_l < 0 || s_l >= h->ls) return DWFGPU_BAD_ARG;
  TRY(ensure_stage4d(h));
  const size_t v4r = (size_t)2 * h->g.vh * 24;
  HIPCHK(hipMemcpyAsync(h->stage4d, four, v4r * sizeof(double), hipMemcpyHostToDevice, h->st));
  for (int s = 0; s < h->ls; ++s)
    TRY(DISPATCH(h, canon_to_halves(h, first->v, second->v, h->stage4d, s, s == s_u ? 1 : 0,
                                    s == s_l ? 1 : 0)));
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

int dwfgpu_five2four(dwfgpu_t h, double* four, dwfgpu_vec_t first, dwfgpu_vec_t second,
                     int s_u, int s_l) {
  if (!own_pair(h, first, second) || !four) return DWFGPU_BAD_ARG;
  if (s_u < 0 || s_u >= h->ls || s_l < 0 || s_l >= h->ls) return DWFGPU_BAD_ARG;
  TRY(ensure_stage4d(h));
  const size_t v4r = (size_t)2 * h->g.vh * 24;
  HIPCHK(hipMemsetAsync(h->stage4d, 0, v4r * sizeof(double), h->st));
  TRY(DISPATCH(h, halves_to_canon(h, h->stage4d, first->v, second->v, s_u, 1, 0)));
  TRY(DISPATCH(h, halves_to_canon(h, h->stage4d, first->v, second->v, s_l, 0, 1)));
  HIPCHK(hipMemcpyAsync(four, h->stage4d, v4r * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return DWFGPU_OK;
}

int dwfgpu_matinv_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, dwfgpu_vec_t out_first,
                      dwfgpu_vec_t out_second, dwfgpu_vec_t in_first, dwfgpu_vec_t in_second,
                      double* true_res, int* iters) {
  if (!arg || !own_pair(h, out_first, out_second) || !own_pair(h, in_first, in_second))
    return DWFGPU_BAD_ARG;
  if (out_first == in_first || out_first == in_second || out_second == in_first ||
      out_second == in_second)
    return DWFGPU_BAD_ARG;
  if (arg->max_num_iter < 1) return DWFGPU_BAD_ARG;
  TRY(ensure_api(h, 6));
  const int st = DISPATCH(h, matinv(h, arg, out_first->v, out_second->v, in_first->v,
                                    in_second->v, true_res, iters));
  HIPCHK(hipStreamSynchronize(h->st));
  return st;
}

int dwfgpu_plaquette(dwfgpu_t h, double* plaq, double* link_trace) {
  if (!h || (!plaq && !link_trace)) return DWFGPU_BAD_ARG;
  const int nb = h->npartial / 2 < 512 ? h->npartial / 2 : 512;
  hipLaunchKernelGGL(dwfk::k_plaquette, dim3(nb), dim3(256), 0, h->st,
                     (const cx<double>*)h->f[DWFGPU_F64].U, h->g, h->partial);
  HIPCHK(hipGetLastError());
  TRY(reduce_to(h, nb, 0));
  hipLaunchKernelGGL(dwfk::k_reduce_partials, dim3(1), dim3(1024), 0, h->st, h->partial + nb, nb,
                     h->dscal + 1);
  HIPCHK(hipGetLastError());
  double sc[2];
  TRY(fetch_scalars(h, 2, sc));
  const double v4 = 2.0 * h->g.vh;
  // Lattice::SumReTrPlaq / (18 V), sum ReTr U / (12 V) (ReadLatticePar.C:403,428)
  if (plaq) *plaq = sc[0] / (18.0 * v4);
  if (link_trace) *link_trace = sc[1] / (12.0 * v4);
  return DWFGPU_OK;
}

// ---- (f4) eigCG deflation (inv_eig_cg.C:129-631, eig_QPropW.C:85-330) -----
int dwfgpu_inveigcg_dev(dwfgpu_t h, const dwfgpu_cg_arg* arg, dwfgpu_vec_t sol, dwfgpu_vec_t src,
                        int nev, int m, dwfgpu_mvec_t V, double* M, dwfgpu_mvec_t U,
                        const double* invH, int def_len, const double* restart, int restart_len,
                        double* true_res, int* iters) {
  if (!own_pair(h, sol, src) || !arg || arg->max_num_iter < 1 || nev < 0 || def_len < 0 ||
      restart_len < 0 || (restart_len > 0 && !restart)) {
    set_err("dwfgpu_inveigcg_dev: bad argument");
    return DWFGPU_BAD_ARG;
  }
  if (h->prec != DWFGPU_F64) {
    set_err("dwfgpu_inveigcg_dev: needs an FP64 handle");
    return DWFGPU_BAD_ARG;
  }
  if (nev > 0 && (m <= 2 * nev || m > 256 || 2 * nev > dwfk::EIG_MAXRANK || !V || V->h != h ||
                  V->nrhs < m || !M)) {
    set_err("dwfgpu_inveigcg_dev: 2 nev < m <= 256, 2 nev <= %d and an mvec of m vectors needed",
            dwfk::EIG_MAXRANK);
    return DWFGPU_BAD_ARG;
  }
  if (def_len > 0 && (!U || U->h != h || U->nrhs < def_len || !invH)) {
    set_err("dwfgpu_inveigcg_dev: U with def_len vectors and invH needed");
    return DWFGPU_BAD_ARG;
  }
  std::vector<std::complex<double>> ih((size_t)def_len * def_len);
  for (size_t k = 0; k < ih.size(); ++k) ih[k] = {invH[2 * k], invH[2 * k + 1]};
  return inveigcg_run(h, arg, sol->v, src->v, nev, nev > 0 ? m : 0,
                      nev > 0 ? (double*)V->v.d : nullptr, M,
                      def_len > 0 ? (const double*)U->v.d : nullptr, ih, def_len, restart,
                      restart_len, true_res, iters);
}

int dwfgpu_eigcg_create(dwfgpu_t h, int nev, int m, int max_def_len, double max_eig,
                        dwfgpu_eigcg_t* out) {
  if (!h || !out || nev < 1 || m <= 2 * nev || m > 256 || 2 * nev > dwfk::EIG_MAXRANK ||
      max_def_len < 1) {
    set_err("dwfgpu_eigcg_create: nev >= 1, 2 nev < m <= 256, 2 nev <= %d, max_def_len >= 1",
            dwfk::EIG_MAXRANK);
    return DWFGPU_BAD_ARG;
  }
  if (h->prec != DWFGPU_F64) {
    set_err("dwfgpu_eigcg_create: needs an FP64 handle");
    return DWFGPU_BAD_ARG;
  }
  dwfgpu_eigcg_s* e = new dwfgpu_eigcg_s;
  e->h = h;
  e->nev = nev;
  e->m = m;
  e->max_def_len = max_def_len;
  e->max_eig = max_eig;
  e->H.assign((size_t)max_def_len * max_def_len, 0.0);
  e->M.assign(2 * nev, 0.0);
  int st = alloc_mvec<P64>(h, m, &e->V);
  if (!st) st = alloc_mvec<P64>(h, max_def_len, &e->U);
  if (st) {
    free_vec(&e->V);
    free_vec(&e->U);
    delete e;
    return st;
  }
  *out = e;
  return DWFGPU_OK;
}

int dwfgpu_eigcg_destroy(dwfgpu_eigcg_t e) {
  if (!e) return DWFGPU_BAD_ARG;
  free_vec(&e->V);
  free_vec(&e->U);
  delete e;
  return DWFGPU_OK;
}

int dwfgpu_eigcg_def_len(dwfgpu_eigcg_t e) { return e ? e->def_len : -1; }

int dwfgpu_eigcg_solve(dwfgpu_eigcg_t e, const dwfgpu_cg_arg* arg, dwfgpu_vec_t sol,
                       dwfgpu_vec_t src, const double* restart, int restart_len,
                       int always_restart, double* true_res, int* iters) {
  if (!e || !arg || !own_pair(e->h, sol, src) || restart_len < 0 ||
      (restart_len > 0 && !restart) || arg->max_num_iter < 1)
    return DWFGPU_BAD_ARG;
  dwfgpu_s* h = e->h;
  const size_t n = h->n;
  const int dl = e->def_len;
  // invH from the leading def_len x def_len block of H (eig_QPropW.C:253-282)
  std::vector<std::complex<double>> ih((size_t)dl * dl);
  for (int i = 0; i < dl; ++i)
    for (int j = 0; j < dl; ++j) ih[(size_t)i * dl + j] = e->H[(size_t)i * e->max_def_len + j];
  if (dl > 0 && !cinverse(ih, dl)) {
    set_err("dwfgpu_eigcg_solve: U^dag A U is singular");
    return DWFGPU_BAD_ARG;
  }
  // window and restarts as eig_Run chooses them (eig_QPropW.C:283-290)
  int nev = e->nev, rl = restart_len;
  if (!always_restart && dl < e->max_def_len)
    rl = 0;
  else if (dl == e->max_def_len)
    nev = 0;
  int st = inveigcg_run(h, arg, sol->v, src->v, nev, nev > 0 ? e->m : 0, (double*)e->V.d,
                        e->M.data(), (const double*)e->U.d, ih, dl, restart, rl, true_res, iters);
  if (st != DWFGPU_OK && st != DWFGPU_MAXITER) return st;
  if (nev == 0 || e->def_len >= e->max_def_len) return st;
  // add the new low Ritz vectors, lowest first, to U and H (eig_QPropW.C:292-330)
  const int nev2 = 2 * e->nev;
  std::vector<int> low(nev2);
  for (int i = 0; i < nev2; ++i) low[i] = i;
  std::stable_sort(low.begin(), low.end(), [&](int a, int b) { return e->M[a] < e->M[b]; });
  const DVec& au = h->f[DWFGPU_F64].work[5];
  EigWork w;
  TRY(w.init(1));
  for (int i = 0; i < nev2; ++i) {
    const int k = low[i];
    if (!(e->M[k] < e->max_eig && e->M[k] > 1e-30 && e->def_len < e->max_def_len)) continue;
    const int d = e->def_len;
    DVec vk, ud;
    vk.d = (double*)e->V.d + (size_t)k * n;
    ud.d = (double*)e->U.d + (size_t)d * n;
    hipLaunchKernelGGL(dwfk::k_eig_round_f32, blas_grid(), blas_block(), 0, h->st,
                       (double*)ud.d, (double*)vk.d, n);
    HIPCHK(hipGetLastError());
    TRY(Ops<P64>::mdagm(h, arg->mass, au, vk, nullptr));
    if (d > 0) {
      std::vector<std::complex<double>> c;
      TRY(eig_mdot(h, w, (const double*)e->U.d, n, d, (const double*)au.d, c));
      for (int j = 0; j < d; ++j) {
        e->H[(size_t)j * e->max_def_len + d] = c[j];
        e->H[(size_t)d * e->max_def_len + j] = std::conj(c[j]);
      }
    }
    double hd;
    TRY(dot<P64>(h, vk, au, &hd));
    e->H[(size_t)d * e->max_def_len + d] = hd;
    e->def_len++;
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return st;
}

int dwfgpu_eigcg_values(dwfgpu_eigcg_t e, double* M, double* H) {
  if (!e) return DWFGPU_BAD_ARG;
  if (M) std::memcpy(M, e->M.data(), e->M.size() * sizeof(double));
  if (H)
    for (int i = 0; i < e->def_len; ++i)
      for (int j = 0; j < e->def_len; ++j) {
        const std::complex<double> v = e->H[(size_t)i * e->max_def_len + j];
        H[2 * ((size_t)i * e->def_len + j)] = v.real();
        H[2 * ((size_t)i * e->def_len + j) + 1] = v.imag();
      }
  return DWFGPU_OK;
}

int dwfgpu_lanczos(dwfgpu_t h, const dwfgpu_lanczos_arg* arg, const double* v0, double* evecs,
                   double* lambda, int* iters) {
  if (!h || !arg || !v0 || !lambda || arg->nk < 1 || arg->np < 1 || arg->nt < 1 ||
      arg->nt > arg->nk || arg->nk + arg->np > dwfk::LZ_MAXM || arg->npol < 1 ||
      arg->maxiters < 1 || !(arg->p0 * arg->p0 != (arg->p1 + std::fabs(arg->shift)) *
                                                      (arg->p1 + std::fabs(arg->shift)))) {
    set_err("dwfgpu_lanczos: 1 <= nt <= nk, np >= 1, nk + np <= %d, npol >= 1, maxiters >= 1 "
            "and a non-empty filter interval needed", dwfk::LZ_MAXM);
    return DWFGPU_BAD_ARG;
  }
  if (h->prec != DWFGPU_F64) {
    set_err("dwfgpu_lanczos: needs an FP64 handle");
    return DWFGPU_BAD_ARG;
  }
  TRY(ensure_fields<P64>(h));
  DVec v;
  TRY(alloc_vec<P64>(h, &v));
  int st = Ops<P64>::to_device(h, v, v0);
  if (!st) st = lanczos_run(h, arg, v, evecs, nullptr, lambda, iters);
  (void)hipStreamSynchronize(h->st);
  free_vec(&v);
  return st;
}

}  // extern "C"
