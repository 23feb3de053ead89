/*
 * cps_oracle.c -- CPU restatement (FP64) of CPS's DWF / Mobius MatPcDagMatPc
 * and DiracOp::InvCgShift, used as the parity oracle of the HIP library.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py; never by the product library.
 *
 * Each function cites the CPS file:line it restates (paths relative to
 * cps_pp/src/util/).  The loop order and per-element arithmetic follow CPS so
 * that results agree with CPS's own noarch build to rounding; the OpenMP
 * pragmas parallelise only loops whose iterations write disjoint outputs, so
 * results do not depend on the thread count.  Norms are serial sums (CPS's
 * OpenMP reduction order is thread-count dependent, vector_util.C:727-735).
 */
#include "cps_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* helpers                                                                   */
/* ------------------------------------------------------------------------ */
#define SPINOR 24
#define GAUGE_SITE 72
/* UELEM of wilson_dslash.C:54 -- element (row,col) of link mu at a site */
#define UEL(u, r, row, col, mu) (u)[(r) + 2 * ((row) + 3 * ((col) + 3 * (mu)))]
/* FERM of wilson_dslash.C:60 */
#define FRM(f, r, c, s) (f)[(r) + 2 * ((c) + 3 * (s))]

static inline int vol4(const int L[4]) { return L[0] * L[1] * L[2] * L[3]; }

size_t orc_half_size(const orc_params *p) {
  return (size_t)SPINOR * (size_t)vol4(p->L) * (size_t)p->ls / 2;
}

double orc_norm_sq(const double *a, size_t n) {
  /* Vector::NormSqNode -> dotProduct(v,v) (vector.h:936, vector_util.C:727) */
  double sum = 0.0;
  for (size_t i = 0; i < n; ++i) sum += a[i] * a[i];
  return sum;
}

/* fTimesV1PlusV2: a = b*c + d   (vector_util.C:827-834) */
static void f_times_v1_plus_v2(double *a, double b, const double *c, const double *d,
                               size_t n) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) a[i] = b * c[i] + d[i];
}

/* complex helpers with the textbook product (a+ib)(c+id) = (ac-bd, ad+bc) */
static inline void cmul(double *o_re, double *o_im, double ar, double ai, double br,
                        double bi) {
  *o_re = ar * br - ai * bi;
  *o_im = ar * bi + ai * br;
}

/* ------------------------------------------------------------------------ */
/* a13: layouts and boundary conditions                                      */
/* ------------------------------------------------------------------------ */
/* Lattice::BondCond (lattice_base.C:5336-5391) applies the sign to the links
 * U_mu(x) with x_mu = L_mu - 1 in CANONICAL order; CanonToAnything for WILSON
 * (convert_func.C:372-374,402-405) moves site x to
 * (x/2 + lx/2*(y+ly*(z+lz*t))) + vh*parity and transposes each 3x3 matrix.   */
void orc_gauge_canonical_to_wilson(double *uw, const double *uc, const int L[4],
                                   const int bc[4]) {
  const int lx = L[0], ly = L[1], lz = L[2], lt = L[3];
  const int vh = vol4(L) / 2;
  for (int t = 0; t < lt; ++t)
    for (int z = 0; z < lz; ++z)
      for (int y = 0; y < ly; ++y)
        for (int x = 0; x < lx; ++x) {
          const int site = x + lx * (y + ly * (z + lz * t));
          const int pos[4] = {x, y, z, t};
          const int par = (x + y + z + t) % 2;
          const int cbs = x / 2 + (lx / 2) * (y + ly * (z + lz * t));
          double *dst = uw + (size_t)(cbs + vh * par) * GAUGE_SITE;
          for (int mu = 0; mu < 4; ++mu) {
            const double sign = (bc[mu] && pos[mu] == L[mu] - 1) ? -1.0 : 1.0;
            const double *src = uc + (size_t)site * GAUGE_SITE + 18 * mu;
            for (int row = 0; row < 3; ++row)
              for (int col = 0; col < 3; ++col)
                for (int r = 0; r < 2; ++r)
                  UEL(dst, r, row, col, mu) = sign * src[r + 2 * (col + 3 * row)];
          }
        }
}

/* FcanonToWilson on one 4D slice (convert_func.C:429-494): half p of a slice */
void orc_fermion_extract_cb(double *half, const double *canon, const int L[4], int p) {
  const int lx = L[0], ly = L[1], lz = L[2], lt = L[3];
  for (int t = 0; t < lt; ++t)
    for (int z = 0; z < lz; ++z)
      for (int y = 0; y < ly; ++y)
        for (int x = 0; x < lx; ++x) {
          if ((x + y + z + t) % 2 != p) continue;
          const int site = x + lx * (y + ly * (z + lz * t));
          const int cbs = x / 2 + (lx / 2) * (y + ly * (z + lz * t));
          memcpy(half + (size_t)cbs * SPINOR, canon + (size_t)site * SPINOR,
                 SPINOR * sizeof(double));
        }
}

void orc_fermion_insert_cb(double *canon, const double *half, const int L[4], int p) {
  const int lx = L[0], ly = L[1], lz = L[2], lt = L[3];
  for (int t = 0; t < lt; ++t)
    for (int z = 0; z < lz; ++z)
      for (int y = 0; y < ly; ++y)
        for (int x = 0; x < lx; ++x) {
          if ((x + y + z + t) % 2 != p) continue;
          const int site = x + lx * (y + ly * (z + lz * t));
          const int cbs = x / 2 + (lx / 2) * (y + ly * (z + lz * t));
          memcpy(canon + (size_t)site * SPINOR, half + (size_t)cbs * SPINOR,
                 SPINOR * sizeof(double));
        }
}

/* ------------------------------------------------------------------------ */
/* f2: whole 5D fermion fields, CANONICAL <-> the full vector of MatInv      */
/* ------------------------------------------------------------------------ */
/* One CANONICAL 4D slice -> WILSON order, odd sites first
 * (convert_func.C:452-462: offset vol/2 * ((x+y+z+t+1)%2) + cb index).     */
static void canon_slice_to_wilson(double *wil, const double *canon, const int L[4]) {
  const int lx = L[0], ly = L[1], lz = L[2], lt = L[3];
  const size_t vh = (size_t)lx * ly * lz * lt / 2;
  for (int t = 0; t < lt; ++t)
    for (int z = 0; z < lz; ++z)
      for (int y = 0; y < ly; ++y)
        for (int x = 0; x < lx; ++x) {
          const size_t site = x + (size_t)lx * (y + (size_t)ly * (z + (size_t)lz * t));
          const size_t cbs = x / 2 + (size_t)(lx / 2) * (y + (size_t)ly * (z + (size_t)lz * t));
          const size_t w = vh * (size_t)((x + y + z + t + 1) % 2) + cbs;
          memcpy(wil + w * SPINOR, canon + site * SPINOR, SPINOR * sizeof(double));
        }
}

/* FdwfBase::Fconvert CANONICAL -> WILSON (DWF, convert.C:704-757) or
 * -> DWF_4D_EOPREC_EE (Mobius, convert.C:797-831): every slice to WILSON
 * order, then the halves of the slices gathered into [first | second].     */
void orc_fermion5d_canonical_to_full(double *full, const double *canon, const orc_params *p) {
  const size_t v4 = (size_t)p->L[0] * p->L[1] * p->L[2] * p->L[3];
  const size_t hs = v4 / 2 * SPINOR, stride = 2 * hs;
  const int ls = p->ls;
  double *tmp = (double *)malloc((size_t)ls * stride * sizeof(double));
  for (int s = 0; s < ls; ++s) canon_slice_to_wilson(tmp + s * stride, canon + s * stride, p->L);
  double *out = full;
  if (p->action == 0) {
    /* odd part: slice i takes the half of 4D parity (i+1)%2; the source
     * pointer advances by (2 parity + 1) halves (convert.C:724-739)        */
    const double *src = tmp;
    for (int i = 0; i < ls; ++i) {
      const int parity = (i + 1) % 2;
      memcpy(out, src, hs * sizeof(double));
      out += hs;
      src += (size_t)(2 * parity + 1) * hs;
    }
    /* even part, starting one half in (convert.C:741-756)                  */
    src = tmp + hs;
    for (int i = 0; i < ls; ++i) {
      const int parity = i % 2;
      memcpy(out, src, hs * sizeof(double));
      out += hs;
      src += (size_t)(2 * parity + 1) * hs;
    }
  } else {
    /* even part (second WILSON half of each slice), then odd part          */
    for (int i = 0; i < ls; ++i, out += hs)
      memcpy(out, tmp + i * stride + hs, hs * sizeof(double));
    for (int i = 0; i < ls; ++i, out += hs) memcpy(out, tmp + i * stride, hs * sizeof(double));
  }
  free(tmp);
}

/* the inverse permutation: [first | second] -> CANONICAL                    */
void orc_fermion5d_full_to_canonical(double *canon, const double *full, const orc_params *p) {
  const size_t v4 = (size_t)p->L[0] * p->L[1] * p->L[2] * p->L[3];
  const size_t hs = v4 / 2 * SPINOR, stride = 2 * hs;
  const int ls = p->ls;
  const double *first = full, *second = full + (size_t)ls * hs;
  for (int s = 0; s < ls; ++s) {
    const double *even, *odd;
    if (p->action == 0) {
      /* 5D-odd half holds 4D parity (s+1)%2 at slice s, 5D-even parity s%2 */
      odd = (s % 2 == 0) ? first + (size_t)s * hs : second + (size_t)s * hs;
      even = (s % 2 == 0) ? second + (size_t)s * hs : first + (size_t)s * hs;
    } else {
      even = first + (size_t)s * hs;
      odd = second + (size_t)s * hs;
    }
    orc_fermion_insert_cb(canon + s * stride, even, p->L, 0);
    orc_fermion_insert_cb(canon + s * stride, odd, p->L, 1);
  }
}

/* FdwfBase::Ffour2five(five, four, s_u, s_l, Ncb = 2), one node in s
 * (f_dwf_base.C:961-1043): zero, then the upper spin pair of slice s_u and
 * the lower pair of slice s_l from the 4D field.  CANONICAL order.          */
void orc_four2five(double *five, const double *four, int s_u, int s_l, const orc_params *p) {
  const size_t v4 = (size_t)p->L[0] * p->L[1] * p->L[2] * p->L[3];
  const size_t ls_stride = SPINOR * v4;
  memset(five, 0, (size_t)p->ls * ls_stride * sizeof(double));
  for (size_t x = 0; x < v4; ++x)
    for (int i = 0; i < 12; ++i) five[(size_t)s_u * ls_stride + x * SPINOR + i] = four[x * SPINOR + i];
  for (size_t x = 0; x < v4; ++x)
    for (int i = 12; i < 24; ++i) five[(size_t)s_l * ls_stride + x * SPINOR + i] = four[x * SPINOR + i];
}

/* FdwfBase::Ffive2four(four, five, s_u, s_l, Ncb = 2) (f_dwf_base.C:1064-1168) */
void orc_five2four(double *four, const double *five, int s_u, int s_l, const orc_params *p) {
  const size_t v4 = (size_t)p->L[0] * p->L[1] * p->L[2] * p->L[3];
  const size_t ls_stride = SPINOR * v4;
  memset(four, 0, v4 * SPINOR * sizeof(double));
  for (size_t x = 0; x < v4; ++x)
    for (int i = 0; i < 12; ++i) four[x * SPINOR + i] = five[(size_t)s_u * ls_stride + x * SPINOR + i];
  for (size_t x = 0; x < v4; ++x)
    for (int i = 12; i < 24; ++i) four[x * SPINOR + i] = five[(size_t)s_l * ls_stride + x * SPINOR + i];
}

/* ------------------------------------------------------------------------ */
/* a1: wilson_dslash  (d_op_wilson/noarch/wilson_dslash.C:64-167, 173-517)    */
/* ------------------------------------------------------------------------ */
/* (1 -/+ gamma_mu) psi, unprojected 4-spinor, sign = -1 for "minus" (mp=0),
 * +1 for "plus" (mp=1); gamma matrices as in OneMinusPlusGammaTimesPsi
 * (wilson_dslash.C:64-128).                                                 */
static void one_minus_plus_gamma(double *into, const double *psi, int sdag, int mu,
                                 int mp) {
  const double sg = (mp == 1 ? 1.0 : -1.0) * (double)sdag;
  for (int c = 0; c < 3; ++c) {
    switch (mu) {
      case 0:
        FRM(into, 0, c, 0) = FRM(psi, 0, c, 0) + sg * (-FRM(psi, 1, c, 3));
        FRM(into, 1, c, 0) = FRM(psi, 1, c, 0) + sg * (FRM(psi, 0, c, 3));
        FRM(into, 0, c, 1) = FRM(psi, 0, c, 1) + sg * (-FRM(psi, 1, c, 2));
        FRM(into, 1, c, 1) = FRM(psi, 1, c, 1) + sg * (FRM(psi, 0, c, 2));
        FRM(into, 0, c, 2) = FRM(psi, 0, c, 2) + sg * (FRM(psi, 1, c, 1));
        FRM(into, 1, c, 2) = FRM(psi, 1, c, 2) + sg * (-FRM(psi, 0, c, 1));
        FRM(into, 0, c, 3) = FRM(psi, 0, c, 3) + sg * (FRM(psi, 1, c, 0));
        FRM(into, 1, c, 3) = FRM(psi, 1, c, 3) + sg * (-FRM(psi, 0, c, 0));
        break;
      case 1:
        FRM(into, 0, c, 0) = FRM(psi, 0, c, 0) + sg * (-FRM(psi, 0, c, 3));
        FRM(into, 1, c, 0) = FRM(psi, 1, c, 0) + sg * (-FRM(psi, 1, c, 3));
        FRM(into, 0, c, 1) = FRM(psi, 0, c, 1) + sg * (FRM(psi, 0, c, 2));
        FRM(into, 1, c, 1) = FRM(psi, 1, c, 1) + sg * (FRM(psi, 1, c, 2));
        FRM(into, 0, c, 2) = FRM(psi, 0, c, 2) + sg * (FRM(psi, 0, c, 1));
        FRM(into, 1, c, 2) = FRM(psi, 1, c, 2) + sg * (FRM(psi, 1, c, 1));
        FRM(into, 0, c, 3) = FRM(psi, 0, c, 3) + sg * (-FRM(psi, 0, c, 0));
        FRM(into, 1, c, 3) = FRM(psi, 1, c, 3) + sg * (-FRM(psi, 1, c, 0));
        break;
      case 2:
        FRM(into, 0, c, 0) = FRM(psi, 0, c, 0) + sg * (-FRM(psi, 1, c, 2));
        FRM(into, 1, c, 0) = FRM(psi, 1, c, 0) + sg * (FRM(psi, 0, c, 2));
        FRM(into, 0, c, 1) = FRM(psi, 0, c, 1) + sg * (FRM(psi, 1, c, 3));
        FRM(into, 1, c, 1) = FRM(psi, 1, c, 1) + sg * (-FRM(psi, 0, c, 3));
        FRM(into, 0, c, 2) = FRM(psi, 0, c, 2) + sg * (FRM(psi, 1, c, 0));
        FRM(into, 1, c, 2) = FRM(psi, 1, c, 2) + sg * (-FRM(psi, 0, c, 0));
        FRM(into, 0, c, 3) = FRM(psi, 0, c, 3) + sg * (-FRM(psi, 1, c, 1));
        FRM(into, 1, c, 3) = FRM(psi, 1, c, 3) + sg * (FRM(psi, 0, c, 1));
        break;
      default:
        FRM(into, 0, c, 0) = FRM(psi, 0, c, 0) + sg * (FRM(psi, 0, c, 2));
        FRM(into, 1, c, 0) = FRM(psi, 1, c, 0) + sg * (FRM(psi, 1, c, 2));
        FRM(into, 0, c, 1) = FRM(psi, 0, c, 1) + sg * (FRM(psi, 0, c, 3));
        FRM(into, 1, c, 1) = FRM(psi, 1, c, 1) + sg * (FRM(psi, 1, c, 3));
        FRM(into, 0, c, 2) = FRM(psi, 0, c, 2) + sg * (FRM(psi, 0, c, 0));
        FRM(into, 1, c, 2) = FRM(psi, 1, c, 2) + sg * (FRM(psi, 1, c, 0));
        FRM(into, 0, c, 3) = FRM(psi, 0, c, 3) + sg * (FRM(psi, 0, c, 1));
        FRM(into, 1, c, 3) = FRM(psi, 1, c, 3) + sg * (FRM(psi, 1, c, 1));
        break;
    }
  }
}

/* U_mu psi (rowcol=0) or U_mu^dagger psi (rowcol=1), UmuTimes :131-167 */
static void umu_times(double *into, const double *u, const double *psi, int mu,
                      int rowcol) {
  for (int s = 0; s < 4; ++s)
    for (int c = 0; c < 3; ++c) {
      if (rowcol == 0) {
        FRM(into, 0, c, s) = (UEL(u, 0, c, 0, mu) * FRM(psi, 0, 0, s) +
                              UEL(u, 0, c, 1, mu) * FRM(psi, 0, 1, s) +
                              UEL(u, 0, c, 2, mu) * FRM(psi, 0, 2, s) -
                              UEL(u, 1, c, 0, mu) * FRM(psi, 1, 0, s) -
                              UEL(u, 1, c, 1, mu) * FRM(psi, 1, 1, s) -
                              UEL(u, 1, c, 2, mu) * FRM(psi, 1, 2, s));
        FRM(into, 1, c, s) = (UEL(u, 0, c, 0, mu) * FRM(psi, 1, 0, s) +
                              UEL(u, 0, c, 1, mu) * FRM(psi, 1, 1, s) +
                              UEL(u, 0, c, 2, mu) * FRM(psi, 1, 2, s) +
                              UEL(u, 1, c, 0, mu) * FRM(psi, 0, 0, s) +
                              UEL(u, 1, c, 1, mu) * FRM(psi, 0, 1, s) +
                              UEL(u, 1, c, 2, mu) * FRM(psi, 0, 2, s));
      } else {
        FRM(into, 0, c, s) = (UEL(u, 0, 0, c, mu) * FRM(psi, 0, 0, s) +
                              UEL(u, 0, 1, c, mu) * FRM(psi, 0, 1, s) +
                              UEL(u, 0, 2, c, mu) * FRM(psi, 0, 2, s) +
                              UEL(u, 1, 0, c, mu) * FRM(psi, 1, 0, s) +
                              UEL(u, 1, 1, c, mu) * FRM(psi, 1, 1, s) +
                              UEL(u, 1, 2, c, mu) * FRM(psi, 1, 2, s));
        FRM(into, 1, c, s) = (UEL(u, 0, 0, c, mu) * FRM(psi, 1, 0, s) +
                              UEL(u, 0, 1, c, mu) * FRM(psi, 1, 1, s) +
                              UEL(u, 0, 2, c, mu) * FRM(psi, 1, 2, s) -
                              UEL(u, 1, 0, c, mu) * FRM(psi, 0, 0, s) -
                              UEL(u, 1, 1, c, mu) * FRM(psi, 0, 1, s) -
                              UEL(u, 1, 2, c, mu) * FRM(psi, 0, 2, s));
      }
    }
}

/* chi (sites of parity 1-cb) = sum_mu (1 - s gamma_mu) U_mu(x) psi(x+mu)
 *                             + (1 + s gamma_mu) U_mu(x-mu)^dag psi(x-mu),
 * s = +1 (dag=0) or -1 (dag=1).  Periodic wrap; boundary signs live in U.   */
/* CPS's noarch 4D hop walks the sites in one thread (wilson_dslash.C:280-477);
 * the restatement spreads them over OpenMP threads unless this is set.  Each
 * site is written once from its own inputs, so both give the same bits.    */
static int g_serial_hop = 0;
void orc_set_serial_hop(int on) { g_serial_hop = on != 0; }

void orc_wilson_dslash(double *chi_p, const double *u_p, const double *psi_p, int cb,
                       int dag, const int L[4]) {
  const int lx = L[0], ly = L[1], lz = L[2], lt = L[3];
  const int vol = vol4(L) / 2; /* checkerboard volume, wilson_p->vol[0] */
  const int sdag = dag ? -1 : 1;
  const int cbn = cb ? 0 : 1;
#pragma omp parallel for collapse(2) schedule(static) if (!g_serial_hop)
  for (int x = 0; x < lx; ++x)
    for (int y = 0; y < ly; ++y)
      for (int z = 0; z < lz; ++z)
        for (int t = 0; t < lt; ++t) {
          if ((x + y + z + t) % 2 != cbn) continue;
          const int xyzt = (x / 2) + (lx / 2) * (y + ly * (z + lz * t));
          const int pos[4] = {x, y, z, t};
          double tmp[SPINOR], tp[4][2][SPINOR];
          for (int mu = 0; mu < 4; ++mu) {
            int pp[4] = {x, y, z, t};
            pp[mu] = (pp[mu] + 1) % L[mu];
            const int pxyzt = (pp[0] / 2) + (lx / 2) * (pp[1] + ly * (pp[2] + lz * pp[3]));
            const double *u = u_p + (size_t)GAUGE_SITE * (xyzt + vol * cbn);
            const double *psi = psi_p + (size_t)SPINOR * pxyzt;
            one_minus_plus_gamma(tmp, psi, sdag, mu, 0);
            umu_times(tp[mu][0], u, tmp, mu, 0);
          }
          for (int mu = 0; mu < 4; ++mu) {
            int pm[4] = {x, y, z, t};
            pm[mu] = pm[mu] - 1 + ((L[mu] - pm[mu]) / L[mu]) * L[mu];
            const int mxyzt = (pm[0] / 2) + (lx / 2) * (pm[1] + ly * (pm[2] + lz * pm[3]));
            const double *u = u_p + (size_t)GAUGE_SITE * (mxyzt + vol * cb);
            const double *psi = psi_p + (size_t)SPINOR * mxyzt;
            one_minus_plus_gamma(tmp, psi, sdag, mu, 1);
            umu_times(tp[mu][1], u, tmp, mu, 1);
          }
          (void)pos;
          double *chi = chi_p + (size_t)SPINOR * xyzt;
          /* summation order tmp1..tmp8 of wilson_dslash.C:437-450 */
          for (int k = 0; k < SPINOR; ++k)
            chi[k] = (tp[0][0][k] + tp[1][0][k] + tp[2][0][k] + tp[3][0][k] +
                      tp[0][1][k] + tp[1][1][k] + tp[2][1][k] + tp[3][1][k]);
        }
}

/* ------------------------------------------------------------------------ */
/* a2: dwf_dslash_5_plus  (d_op_dwf/noarch/dwf_dslash_5_plus_omp.C:115-270)   */
/* ------------------------------------------------------------------------ */
/* out(s) += 2/a5 [first 12 reals] of in(s-1)  (wall: -2 m/a5 in(ls-1))
 * out(s) += 2/a5 [last 12 reals]  of in(s+1)  (wall: -2 m/a5 in(0))
 * dag=1 swaps which 12 reals go with which direction.  One s-node.        */
void orc_dwf_dslash_5_plus(double *out, const double *in, double mass, int dag,
                           const orc_params *p) {
  const int ls = p->ls;
  const long vcb = vol4(p->L) / 2;
  const long ls_stride = SPINOR * vcb;
  const double two_a5 = 2.0 * p->a5inv;
  const double nm_two_a5 = -2.0 * mass * p->a5inv;
  const int off_m = dag ? 12 : 0; /* reals taken from s-1 */
  const int off_p = dag ? 0 : 12; /* reals taken from s+1 */
  /* out[s] += two_a5 * in[s-1], s = 1..ls-1 */
#pragma omp parallel for schedule(static)
  for (long x = 0; x < vcb * (ls - 1); ++x) {
    const double *fi = in + off_m + SPINOR * x;
    double *fo = out + ls_stride + off_m + SPINOR * x;
    for (int k = 0; k < 12; ++k) fo[k] = two_a5 * fi[k] + fo[k];
  }
  /* out[0] += -m two_a5 * in[ls-1] */
#pragma omp parallel for schedule(static)
  for (long x = 0; x < vcb; ++x) {
    const double *fi = in + (ls - 1) * ls_stride + off_m + SPINOR * x;
    double *fo = out + off_m + SPINOR * x;
    for (int k = 0; k < 12; ++k) fo[k] = nm_two_a5 * fi[k] + fo[k];
  }
  /* out[s] += two_a5 * in[s+1], s = 0..ls-2 */
#pragma omp parallel for schedule(static)
  for (long x = 0; x < vcb * (ls - 1); ++x) {
    const double *fi = in + ls_stride + off_p + SPINOR * x;
    double *fo = out + off_p + SPINOR * x;
    for (int k = 0; k < 12; ++k) fo[k] = two_a5 * fi[k] + fo[k];
  }
  /* out[ls-1] += -m two_a5 * in[0] */
#pragma omp parallel for schedule(static)
  for (long x = 0; x < vcb; ++x) {
    const double *fi = in + off_p + SPINOR * x;
    double *fo = out + (ls - 1) * ls_stride + off_p + SPINOR * x;
    for (int k = 0; k < 12; ++k) fo[k] = nm_two_a5 * fi[k] + fo[k];
  }
}

/* ------------------------------------------------------------------------ */
/* a3: dwf_dslash = dwf_dslash_4 + dwf_dslash_5_plus                        */
/* (d_op_dwf/noarch/dwf_dslash.C:32-51, dwf_dslash_4.C:68-154)               */
/* ------------------------------------------------------------------------ */
void orc_dwf_dslash(double *out, const double *u, const double *in, double mass,
                    int cb, int dag, const orc_params *p) {
  const size_t slice = (size_t)SPINOR * vol4(p->L) / 2;
  for (int s = 0; s < p->ls; ++s) {
    const int parity = (s + cb) % 2;
    orc_wilson_dslash(out + s * slice, u, in + s * slice, parity, dag, p->L);
  }
  orc_dwf_dslash_5_plus(out, in, mass, dag, p);
}

/* ------------------------------------------------------------------------ */
/* a4: dwf_m / dwf_mdag / dwf_mdagm                                          */
/* (d_op_dwf/noarch/dwf_m.C:31-63, dwf_mdag.C:32-63, dwf_mdagm.C:32-62)      */
/* kappa = 1/(2(4 + 1/a5 - M5))  (dwf_init.C:97-98, d_op_dwf.C:186-187)      */
/* ------------------------------------------------------------------------ */
static double dwf_kappa(const orc_params *p) {
  return 1.0 / (2 * (4 + p->a5inv - p->m5));
}

void orc_dwf_m(double *out, const double *u, const double *in, double mass,
               const orc_params *p, double *tmp2) {
  const size_t n = orc_half_size(p);
  const double mk2 = -dwf_kappa(p) * dwf_kappa(p);
  orc_dwf_dslash(tmp2, u, in, mass, 1, 0, p);
  orc_dwf_dslash(out, u, tmp2, mass, 0, 0, p);
  f_times_v1_plus_v2(out, mk2, out, in, n);
}

void orc_dwf_mdag(double *out, const double *u, const double *in, double mass,
                  const orc_params *p, double *tmp2) {
  const size_t n = orc_half_size(p);
  const double mk2 = -dwf_kappa(p) * dwf_kappa(p);
  orc_dwf_dslash(tmp2, u, in, mass, 1, 1, p);
  orc_dwf_dslash(out, u, tmp2, mass, 0, 1, p);
  f_times_v1_plus_v2(out, mk2, out, in, n);
}

void orc_dwf_mdagm(double *out, const double *u, const double *in, double *dot_prd,
                   double mass, const orc_params *p, double *tmp1, double *tmp2) {
  orc_dwf_m(tmp1, u, in, mass, p, tmp2);
  if (dot_prd) *dot_prd = orc_norm_sq(tmp1, orc_half_size(p));
  orc_dwf_mdag(out, u, tmp1, mass, p, tmp2);
}

/* ------------------------------------------------------------------------ */
/* a7: zmobius_kappa_dslash_5_plus_cmplx                                     */
/* (d_op_zmobius/noarch/zmobius_dslash_5_plus.C:125-374, one s-node)         */
/* ------------------------------------------------------------------------ */
/* a[i] = b*c[i] + d[i] over 6 complex numbers (zmobius.h:257-263) */
static inline void zaxpy6(double *a, double br, double bi, const double *c,
                          const double *d) {
  for (int i = 0; i < 6; ++i) {
    double re, im;
    cmul(&re, &im, br, bi, c[2 * i], c[2 * i + 1]);
    a[2 * i] = re + d[2 * i];
    a[2 * i + 1] = im + d[2 * i + 1];
  }
}

void orc_zmobius_kappa_dslash_5_plus(double *out, const double *in, double mass,
                                     int dag, const double *fact_in,
                                     const orc_params *p) {
  const int ls = p->ls;
  const long vcb = vol4(p->L) / 2;
  const long ls_stride = SPINOR * vcb;
  double *fact = (double *)malloc(sizeof(double) * 2 * ls);
  for (int i = 0; i < ls; ++i) {
    fact[2 * i] = fact_in[2 * i];
    fact[2 * i + 1] = dag ? -fact_in[2 * i + 1] : fact_in[2 * i + 1];
  }
  const double nm0r = -mass * fact[0], nm0i = -mass * fact[1];
  const double nmlr = -mass * fact[2 * (ls - 1)], nmli = -mass * fact[2 * (ls - 1) + 1];
  const int off_m = dag ? 12 : 0, off_p = dag ? 0 : 12;
  /* from s-1, s = 1..ls-1: dag0 fact[s], dag1 conj(fact[s-1]) */
  for (int s = 1; s < ls; ++s) {
    const int fs = dag ? s - 1 : s;
#pragma omp parallel for schedule(static)
    for (long x = 0; x < vcb; ++x) {
      const double *fi = in + off_m + SPINOR * (x + (s - 1) * vcb);
      double *fo = out + ls_stride + off_m + SPINOR * (x + (s - 1) * vcb);
      zaxpy6(fo, fact[2 * fs], fact[2 * fs + 1], fi, fo);
    }
  }
  /* wall s=0 from ls-1: dag0 -m fact[0], dag1 -m conj(fact[ls-1]) */
  {
    const double wr = dag ? nmlr : nm0r, wi = dag ? nmli : nm0i;
#pragma omp parallel for schedule(static)
    for (long x = 0; x < vcb; ++x) {
      const double *fi = in + (ls - 1) * ls_stride + off_m + SPINOR * x;
      double *fo = out + off_m + SPINOR * x;
      zaxpy6(fo, wr, wi, fi, fo);
    }
  }
  /* from s+1, s = 0..ls-2: dag0 fact[s], dag1 conj(fact[s+1]) */
  for (int s = 0; s < ls - 1; ++s) {
    const int fs = dag ? s + 1 : s;
#pragma omp parallel for schedule(static)
    for (long x = 0; x < vcb; ++x) {
      const double *fi = in + ls_stride + off_p + SPINOR * (x + s * vcb);
      double *fo = out + off_p + SPINOR * (x + s * vcb);
      zaxpy6(fo, fact[2 * fs], fact[2 * fs + 1], fi, fo);
    }
  }
  /* wall s=ls-1 from 0: dag0 -m fact[ls-1], dag1 -m conj(fact[0]) */
  {
    const double wr = dag ? nm0r : nmlr, wi = dag ? nm0i : nmli;
#pragma omp parallel for schedule(static)
    for (long x = 0; x < vcb; ++x) {
      const double *fi = in + off_p + SPINOR * x;
      double *fo = out + (ls - 1) * ls_stride + off_p + SPINOR * x;
      zaxpy6(fo, wr, wi, fi, fo);
    }
  }
  free(fact);
}

/* ------------------------------------------------------------------------ */
/* a6: zmobius_dslash_4  (d_op_zmobius/noarch/zmobius_dslash_4.C:34-221)      */
/* ------------------------------------------------------------------------ */
/* a = c*b over complex pairs (vecEqualsVecTimesEquComplex, zmobius.h:236-243) */
static void vec_eq_vec_times_cplx(double *a, const double *b, double cr, double ci,
                                  size_t nreal) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)nreal / 2; ++i) {
    double re, im;
    cmul(&re, &im, cr, ci, b[2 * i], b[2 * i + 1]);
    a[2 * i] = re;
    a[2 * i + 1] = im;
  }
}

void orc_zmobius_dslash_4(double *out, const double *u, const double *in, int cb,
                          int dag, double mass, const double *zb, const double *zc,
                          const orc_params *p, double *tmp3) {
  const int ls = p->ls;
  const size_t f_size = (size_t)SPINOR * vol4(p->L) / 2;
  if (!dag) {
    /* tmp3(s) = b(s) in(s); tmp3 += D5(c) in; out(s) = D_W tmp3(s) */
    for (int i = 0; i < ls; ++i)
      vec_eq_vec_times_cplx(tmp3 + i * f_size, in + i * f_size, zb[2 * i],
                            zb[2 * i + 1], f_size);
    orc_zmobius_kappa_dslash_5_plus(tmp3, in, mass, 0, zc, p);
    for (int i = 0; i < ls; ++i)
      orc_wilson_dslash(out + i * f_size, u, tmp3 + i * f_size, cb, 0, p->L);
  } else {
    /* tmp3(s) = D_W^dag in(s); out(s) = conj(b(s)) tmp3(s); out += D5^dag(c) tmp3 */
    for (int i = 0; i < ls; ++i)
      orc_wilson_dslash(tmp3 + i * f_size, u, in + i * f_size, cb, 1, p->L);
    for (int i = 0; i < ls; ++i)
      vec_eq_vec_times_cplx(out + i * f_size, tmp3 + i * f_size, zb[2 * i],
                            -zb[2 * i + 1], f_size);
    orc_zmobius_kappa_dslash_5_plus(out, tmp3, mass, 1, zc, p);
  }
}

/* ------------------------------------------------------------------------ */
/* a8: zmobius_m5inv  (d_op_zmobius/noarch/zmobius_m5inv.C:247-396 dag0,      */
/*     :534-679 dag1, dispatch :807-818).  Two s-sweeps per chirality.       */
/* ------------------------------------------------------------------------ */
static inline void c_scale6(double *a, double br, double bi) {
  for (int i = 0; i < 6; ++i) {
    double re, im;
    cmul(&re, &im, br, bi, a[2 * i], a[2 * i + 1]);
    a[2 * i] = re;
    a[2 * i + 1] = im;
  }
}
static inline void c_div(double *o_re, double *o_im, double ar, double ai, double br,
                         double bi) {
  /* (a)/(b) = a * conj(b) / |b|^2 */
  const double den = br * br + bi * bi;
  *o_re = (ar * br + ai * bi) / den;
  *o_im = (ai * br - ar * bi) / den;
}

void orc_zmobius_m5inv(double *f, double mass, int dag, const double *K,
                       const orc_params *p) {
  const int ls = p->ls;
  const long vcb = vol4(p->L) / 2;
#define AT(x, s) (f + SPINOR * ((x) + vcb * (s)))
  if (!dag) {
    double pr = 1.0, pi = 0.0;
    for (int i = 0; i < ls; ++i) {
      double re, im;
      cmul(&re, &im, pr, pi, -K[2 * i], -K[2 * i + 1]);
      pr = re;
      pi = im;
    }
    pr *= mass;
    pi *= mass;
    double idr, idi;
    c_div(&idr, &idi, 1.0, 0.0, 1.0 + pr, pi);
    /* lower part of s=ls-1 scaled first */
#pragma omp parallel for schedule(static)
    for (long x = 0; x < vcb; ++x) c_scale6(AT(x, ls - 1) + 12, idr, idi);
    double fLr, fLi;
    {
      double re, im;
      cmul(&re, &im, K[2 * (ls - 1)], K[2 * (ls - 1) + 1], mass, 0.0);
      cmul(&fLr, &fLi, re, im, idr, idi);
    }
    for (int s = 0; s <= ls - 2; ++s) {
      const double mkr = -K[2 * (s + 1)], mki = -K[2 * (s + 1) + 1];
#pragma omp parallel for schedule(static)
      for (long x = 0; x < vcb; ++x) {
        zaxpy6(AT(x, s + 1), mkr, mki, AT(x, s), AT(x, s + 1));
        zaxpy6(AT(x, ls - 1) + 12, fLr, fLi, AT(x, s) + 12, AT(x, ls - 1) + 12);
      }
      double re, im;
      cmul(&re, &im, fLr, fLi, -K[2 * s], -K[2 * s + 1]);
      fLr = re;
      fLi = im;
    }
    double fRr, fRi;
    {
      double qr = 1.0, qi = 0.0;
      for (int i = 0; i < ls - 1; ++i) {
        double re, im;
        cmul(&re, &im, qr, qi, -K[2 * i], -K[2 * i + 1]);
        qr = re;
        qi = im;
      }
      double re, im;
      cmul(&re, &im, qr * mass, qi * mass, idr, idi);
      fRr = -re;
      fRi = -im;
    }
    for (int s = ls - 2; s >= 0; --s) {
      const double mkr = -K[2 * s], mki = -K[2 * s + 1];
#pragma omp parallel for schedule(static)
      for (long x = 0; x < vcb; ++x) {
        zaxpy6(AT(x, s), fRr, fRi, AT(x, ls - 1), AT(x, s));
        zaxpy6(AT(x, s) + 12, mkr, mki, AT(x, s + 1) + 12, AT(x, s) + 12);
      }
      double re, im;
      c_div(&re, &im, fRr, fRi, -K[2 * s], -K[2 * s + 1]);
      fRr = re;
      fRi = im;
    }
#pragma omp parallel for schedule(static)
    for (long x = 0; x < vcb; ++x) c_scale6(AT(x, ls - 1), idr, idi);
  } else {
    double pr = 1.0, pi = 0.0;
    for (int i = 0; i < ls; ++i) {
      double re, im;
      cmul(&re, &im, pr, pi, K[2 * i], K[2 * i + 1]);
      pr = re;
      pi = im;
    }
    pr *= mass;
    pi *= mass;
    double idr, idi;
    c_div(&idr, &idi, 1.0, 0.0, 1.0 + pr, pi);
    idi = -idi; /* conj */
#pragma omp parallel for schedule(static)
    for (long x = 0; x < vcb; ++x) c_scale6(AT(x, ls - 1), idr, idi);
    double fRr, fRi;
    cmul(&fRr, &fRi, K[0] * mass, -K[1] * mass, idr, idi);
    for (int s = 0; s <= ls - 2; ++s) {
      const double mkr = -K[2 * s], mki = K[2 * s + 1]; /* -conj(K[s]) */
#pragma omp parallel for schedule(static)
      for (long x = 0; x < vcb; ++x) {
        zaxpy6(AT(x, ls - 1), fRr, fRi, AT(x, s), AT(x, ls - 1));
        zaxpy6(AT(x, s + 1) + 12, mkr, mki, AT(x, s) + 12, AT(x, s + 1) + 12);
      }
      double re, im;
      cmul(&re, &im, fRr, fRi, -K[2 * (s + 1)], K[2 * (s + 1) + 1]);
      fRr = re;
      fRi = im;
    }
    double fLr, fLi;
    {
      double qr = 1.0, qi = 0.0;
      for (int i = 0; i < ls - 2; ++i) {
        double re, im;
        cmul(&re, &im, qr, qi, -K[2 * i], -K[2 * i + 1]);
        qr = re;
        qi = im;
      }
      double re, im;
      cmul(&re, &im, qr, qi, -K[2 * (ls - 1)], -K[2 * (ls - 1) + 1]);
      /* factL = -conj(prod) * mass * inv_d_last */
      cmul(&fLr, &fLi, re * mass, -im * mass, idr, idi);
      fLr = -fLr;
      fLi = -fLi;
    }
    for (int s = ls - 2; s >= 0; --s) {
      const double mkr = -K[2 * (s + 1)], mki = K[2 * (s + 1) + 1];
#pragma omp parallel for schedule(static)
      for (long x = 0; x < vcb; ++x) {
        zaxpy6(AT(x, s), mkr, mki, AT(x, s + 1), AT(x, s));
        zaxpy6(AT(x, s) + 12, fLr, fLi, AT(x, ls - 1) + 12, AT(x, s) + 12);
      }
      if (s > 0) {
        double re, im;
        c_div(&re, &im, fLr, fLi, -K[2 * (s - 1)], K[2 * (s - 1) + 1]);
        fLr = re;
        fLi = im;
      }
    }
#pragma omp parallel for schedule(static)
    for (long x = 0; x < vcb; ++x) c_scale6(AT(x, ls - 1) + 12, idr, idi);
  }
#undef AT
}

/* ------------------------------------------------------------------------ */
/* a9: zmobius_m_orig / zmobius_mdag_orig / mobius_mdagm                     */
/* (zmobius_m-orig.h:3-126, zmobius_mdag-orig.h:4-119, mobius_mdagm.C:33-60)  */
/* ------------------------------------------------------------------------ */
void orc_mobius_coefficients(const orc_params *p, double *zb, double *zc,
                             double *zkb, double *zkr) {
  /* d_op_mobius.C:105-125 */
  const double kb = 1.0 / (2 * (p->mob_b * (4 - p->m5) + p->a5inv));
  const double kc = 1.0 / (2 * (p->mob_c * (4 - p->m5) - p->a5inv));
  for (int i = 0; i < p->ls; ++i) {
    zb[2 * i] = p->mob_b;
    zb[2 * i + 1] = 0.0;
    zc[2 * i] = p->mob_c;
    zc[2 * i + 1] = 0.0;
    zkb[2 * i] = kb;
    zkb[2 * i + 1] = 0.0;
    zkr[2 * i] = kb / kc; /* complex division with zero imaginary parts */
    zkr[2 * i + 1] = 0.0;
  }
}

typedef struct {
  double *zb, *zc, *zkb, *zkr;
} mob_coef;

static mob_coef mob_coef_new(const orc_params *p) {
  mob_coef c;
  c.zb = (double *)malloc(sizeof(double) * 8 * p->ls);
  c.zc = c.zb + 2 * p->ls;
  c.zkb = c.zc + 2 * p->ls;
  c.zkr = c.zkb + 2 * p->ls;
  orc_mobius_coefficients(p, c.zb, c.zc, c.zkb, c.zkr);
  return c;
}

/* vecTimesEquComplex on one s-slice (zmobius.h:265-271) */
static void slice_times_cplx(double *a, double br, double bi, size_t nreal) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)nreal / 2; ++i) {
    double re, im;
    cmul(&re, &im, a[2 * i], a[2 * i + 1], br, bi);
    a[2 * i] = re;
    a[2 * i + 1] = im;
  }
}

void orc_mobius_m(double *out, const double *u, const double *in, double mass,
                  const orc_params *p, double *tmp2, double *tmp3) {
  const int ls = p->ls;
  const size_t slice = (size_t)SPINOR * vol4(p->L) / 2;
  const size_t f_size = slice * ls;
  mob_coef c = mob_coef_new(p);
  orc_zmobius_dslash_4(out, u, in, 1, 0, mass, c.zb, c.zc, p, tmp3);
  for (int s = 0; s < ls; ++s)
    slice_times_cplx(out + s * slice, c.zkb[2 * s], c.zkb[2 * s + 1], slice);
  orc_zmobius_m5inv(out, mass, 0, c.zkr, p);
  orc_zmobius_dslash_4(tmp2, u, out, 0, 0, mass, c.zb, c.zc, p, tmp3);
  for (int s = 0; s < ls; ++s)
    slice_times_cplx(tmp2 + s * slice, -c.zkb[2 * s], -c.zkb[2 * s + 1], slice);
  memcpy(out, in, f_size * sizeof(double));
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)f_size; ++i) out[i] += tmp2[i];
  orc_zmobius_kappa_dslash_5_plus(out, in, mass, 0, c.zkr, p);
  free(c.zb);
}

void orc_mobius_mdag(double *out, const double *u, const double *in, double mass,
                     const orc_params *p, double *tmp2, double *tmp3) {
  const int ls = p->ls;
  const size_t slice = (size_t)SPINOR * vol4(p->L) / 2;
  const size_t f_size = slice * ls;
  mob_coef c = mob_coef_new(p);
  memcpy(out, in, f_size * sizeof(double));
  for (int s = 0; s < ls; ++s) /* -conj(kappa_b) */
    slice_times_cplx(out + s * slice, -c.zkb[2 * s], c.zkb[2 * s + 1], slice);
  orc_zmobius_dslash_4(tmp2, u, out, 1, 1, mass, c.zb, c.zc, p, tmp3);
  orc_zmobius_m5inv(tmp2, mass, 1, c.zkr, p);
  for (int s = 0; s < ls; ++s) /* conj(kappa_b) */
    slice_times_cplx(tmp2 + s * slice, c.zkb[2 * s], -c.zkb[2 * s + 1], slice);
  orc_zmobius_dslash_4(out, u, tmp2, 0, 1, mass, c.zb, c.zc, p, tmp3);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)f_size; ++i) out[i] += in[i];
  orc_zmobius_kappa_dslash_5_plus(out, in, mass, 1, c.zkr, p);
  free(c.zb);
}

void orc_mobius_mdagm(double *out, const double *u, const double *in, double *dot_prd,
                      double mass, const orc_params *p, double *tmp1, double *tmp2,
                      double *tmp3) {
  orc_mobius_m(tmp1, u, in, mass, p, tmp2, tmp3);
  if (dot_prd) *dot_prd = orc_norm_sq(tmp1, orc_half_size(p));
  orc_mobius_mdag(out, u, tmp1, mass, p, tmp2, tmp3);
}

/* ------------------------------------------------------------------------ */
/* a5 / a10: DiracOp front ends                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
  double *t1, *t2, *t3;
} work3;
static work3 work_new(const orc_params *p) {
  work3 w;
  const size_t n = orc_half_size(p);
  w.t1 = (double *)calloc(3 * n, sizeof(double));
  w.t2 = w.t1 + n;
  w.t3 = w.t2 + n;
  return w;
}

void orc_matpcdagmatpc(double *out, const double *u, const double *in, double *dot_prd,
                       double mass, const orc_params *p) {
  work3 w = work_new(p);
  if (p->action == 0)
    orc_dwf_mdagm(out, u, in, dot_prd, mass, p, w.t1, w.t2);
  else
    orc_mobius_mdagm(out, u, in, dot_prd, mass, p, w.t1, w.t2, w.t3);
  free(w.t1);
}

void orc_matpc(double *out, const double *u, const double *in, double mass, int dag,
               const orc_params *p) {
  work3 w = work_new(p);
  if (p->action == 0) {
    if (dag)
      orc_dwf_mdag(out, u, in, mass, p, w.t2);
    else
      orc_dwf_m(out, u, in, mass, p, w.t2);
  } else {
    if (dag)
      orc_mobius_mdag(out, u, in, mass, p, w.t2, w.t3);
    else
      orc_mobius_m(out, u, in, mass, p, w.t2, w.t3);
  }
  free(w.t1);
}

void orc_dslash(double *out, const double *u, const double *in, double mass, int cb,
                int dag, const orc_params *p) {
  if (p->action == 0) {
    orc_dwf_dslash(out, u, in, mass, cb, dag, p);
  } else {
    work3 w = work_new(p);
    mob_coef c = mob_coef_new(p);
    orc_zmobius_dslash_4(out, u, in, cb, dag, mass, c.zb, c.zc, p, w.t3);
    free(c.zb);
    free(w.t1);
  }
}

/* ------------------------------------------------------------------------ */
/* a11: DiracOp::InvCgShift with shift = NULL (d_op_base/comsrc/inv_cg.C:84-448) */
/* ------------------------------------------------------------------------ */
int orc_invcg(double *sol, const double *src, const double *u, double mass,
              double stop_rsd, int max_num_iter, double *true_res, double *hist,
              const orc_params *p) {
  return orc_invcg_shift(sol, src, u, mass, 0.0, NULL, stop_rsd, max_num_iter, true_res, hist,
                         p);
}

int orc_invcg_shift(double *sol, const double *src, const double *u, double mass,
                    double src_norm_sq, const double *shift, double stop_rsd,
                    int max_num_iter, double *true_res, double *hist, const orc_params *p) {
  const size_t n = orc_half_size(p);
  double *res = (double *)malloc(n * sizeof(double));
  double *dir = (double *)malloc(n * sizeof(double));
  double *mmp = (double *)malloc(n * sizeof(double));
  work3 w = work_new(p);
#define MDAGM(o, i, d)                                             \
  do {                                                             \
    if (p->action == 0)                                            \
      orc_dwf_mdagm(o, u, i, d, mass, p, w.t1, w.t2);              \
    else                                                           \
      orc_mobius_mdagm(o, u, i, d, mass, p, w.t1, w.t2, w.t3);     \
  } while (0)

  if (src_norm_sq == 0) src_norm_sq = orc_norm_sq(src, n);  /* :166-169 */
  const double stp_cnd = src_norm_sq * stop_rsd * stop_rsd; /* :174 */
  MDAGM(mmp, sol, NULL);                                    /* :248 */
  if (shift) f_times_v1_plus_v2(mmp, *shift, sol, mmp, n);  /* :249-251 */
  memcpy(res, src, n * sizeof(double));                     /* :255 */
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) res[i] -= mmp[i];      /* :259 */
  memcpy(dir, res, n * sizeof(double));                     /* :263 */
  double rr = orc_norm_sq(res, n);                          /* :267 */
  if (hist) hist[0] = rr;
  int itr = 0;
  int max_itr = max_num_iter - 1;
  if (rr <= stp_cnd) max_itr = 0;
  for (int i = 0; i < max_itr; ++i) {                       /* :293-364 */
    itr = itr + 1;
    const double rr_prv = rr;
    double d;
    MDAGM(mmp, dir, &d);
    if (shift) {                                            /* :304-310 */
      f_times_v1_plus_v2(mmp, *shift, dir, mmp, n);
      d += (*shift) * orc_norm_sq(dir, n);
    }
    if (d == 0.0) {                                         /* :326-331 */
      itr = -itr;
      break;
    }
    const double a = rr_prv / d;
    f_times_v1_plus_v2(sol, a, dir, sol, n);
    f_times_v1_plus_v2(res, -a, mmp, res, n);
    rr = orc_norm_sq(res, n);
    if (hist) hist[itr] = rr;
    if (rr <= stp_cnd) break;
    const double b = rr / rr_prv;
    f_times_v1_plus_v2(dir, b, dir, res, n);
  }
  if (itr < 0) { /* CPS exits(5) on d == 0; the oracle reports it */
    free(res); free(dir); free(mmp); free(w.t1);
    return 0;
  }
  MDAGM(mmp, sol, NULL);                                    /* :403-415 */
  if (shift) f_times_v1_plus_v2(mmp, *shift, sol, mmp, n);
  memcpy(res, src, n * sizeof(double));
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; ++i) res[i] -= mmp[i];
  const double tr = sqrt(orc_norm_sq(res, n) / src_norm_sq);
  if (true_res) *true_res = tr;
#undef MDAGM
  free(res);
  free(dir);
  free(mmp);
  free(w.t1);
  return itr + 1;                                           /* :446 */
}

/* ------------------------------------------------------------------------ */
/* (f1): DiracOpDwf::MatInv (d_op_dwf.C:316-507), DiracOpMobius::MatInv       */
/* (d_op_mobius.C:407-696, ORIG pc, one right-hand side), PRESERVE_YES.       */
/* ------------------------------------------------------------------------ */
int orc_matinv(double *out, double *in, const double *u, double mass, double stop_rsd,
               int max_iter, double *true_res, const orc_params *p) {
  const size_t n = orc_half_size(p);
  double *temp = (double *)calloc(n, sizeof(double));
  double *save = (double *)malloc(2 * n * sizeof(double));
  memcpy(save, in, 2 * n * sizeof(double));
  int iter;
  if (p->action == 0) {
    const double kappa = dwf_kappa(p);
    double *even_in = in + n, *even_out = out + n;
    orc_dwf_dslash(temp, u, even_in, mass, 0, 0, p);          /* :366 */
    f_times_v1_plus_v2(temp, kappa, temp, in, n);            /* :395 */
    orc_matpc(in, u, temp, mass, 1, p);                      /* :419 */
    iter = orc_invcg(out, in, u, mass, stop_rsd, max_iter, true_res, NULL, p);
    memcpy(in, save, 2 * n * sizeof(double));                /* :485 */
    orc_dwf_dslash(temp, u, out, mass, 1, 0, p);             /* :489 */
    f_times_v1_plus_v2(even_out, kappa, temp, even_in, n);   /* :491 */
  } else {
    const int ls = p->ls;
    const size_t slice = n / ls;
    work3 w = work_new(p);
    mob_coef c = mob_coef_new(p);
    double *even_in = in, *odd_in = in + n, *even_out = out, *odd_out = out + n;
    double *temp2 = w.t1;
    memcpy(temp, even_in, n * sizeof(double));               /* :483 m5inv(out,in) */
    orc_zmobius_m5inv(temp, mass, 0, c.zkr, p);
    orc_zmobius_dslash_4(temp2, u, temp, 0, 0, mass, c.zb, c.zc, p, w.t3);
    for (int s = 0; s < ls; ++s) {                           /* zvectTimesV1PlusV2 */
      const double kr = c.zkb[2 * s], ki = c.zkb[2 * s + 1];
      for (size_t k = 0; k < slice / 2; ++k) {
        double re, im;
        cmul(&re, &im, kr, ki, temp2[s * slice + 2 * k], temp2[s * slice + 2 * k + 1]);
        temp[s * slice + 2 * k] = re + odd_in[s * slice + 2 * k];
        temp[s * slice + 2 * k + 1] = im + odd_in[s * slice + 2 * k + 1];
      }
    }
    orc_matpc(odd_in, u, temp, mass, 1, p);                  /* :546 */
    iter = orc_invcg(odd_out, odd_in, u, mass, stop_rsd, max_iter, true_res, NULL, p);
    memcpy(in, save, 2 * n * sizeof(double));
    orc_zmobius_dslash_4(temp, u, odd_out, 1, 0, mass, c.zb, c.zc, p, w.t3); /* :650 */
    for (int s = 0; s < ls; ++s)
      slice_times_cplx(temp + s * slice, c.zkb[2 * s], c.zkb[2 * s + 1], slice);
    memcpy(even_out, temp, n * sizeof(double));
    orc_zmobius_m5inv(even_out, mass, 0, c.zkr, p);
    memcpy(temp, even_in, n * sizeof(double));
    orc_zmobius_m5inv(temp, mass, 0, c.zkr, p);
    for (size_t k = 0; k < n; ++k) even_out[k] += temp[k];
    const double two_kb = 2.0 * c.zkb[0];                    /* :668-678 */
    for (size_t k = 0; k < 2 * n; ++k) out[k] *= two_kb;
    free(c.zb);
    free(w.t1);
  }
  free(temp);
  free(save);
  return iter;
}

/* ---- (f3): DiracOp::MInvCG, multi-shift CG (d_op_base/noarch/minvcg.C:47-355)
 * Solves (A + shift[s]) psi_s = chi for every shift, A = MatPcDagMatPc, with
 * one A application per iteration; shift[0] is the base system.  type 0
 * (SINGLE): psi[0] accumulates sum_s alpha[s] psi_s; type 1 (MULTI): psi_s
 * separately (zeroed first).  rsdcg[s] holds the relative stop residual on
 * entry and the final |r_s| on return (sqrt of c z_s^2, :336-345); iters[s]
 * (may be NULL) the iteration at which shift s converged (0 if it did not).
 * Returns CPS's k (:345-350): the loop count at exit.                        */
int orc_minvcg(double *psi, const double *chi, double chi_norm, const double *shift,
               int nshift, double *rsdcg, int type, const double *alpha, int max_num_iter,
               int *iters, const double *u, double mass, const orc_params *p) {
  const size_t n = orc_half_size(p);
  double *r = (double *)malloc(n * sizeof(double));
  double *ap = (double *)malloc(n * sizeof(double));
  double *pv = (double *)malloc((size_t)nshift * n * sizeof(double));
  int *convs = (int *)calloc((size_t)nshift, sizeof(int));
  int *conv_at = (int *)calloc((size_t)nshift, sizeof(int));
  double *bs = (double *)malloc(nshift * sizeof(double));
  double *z[2];
  z[0] = (double *)malloc(nshift * sizeof(double));
  z[1] = (double *)malloc(nshift * sizeof(double));
  double *rsd_sq = (double *)malloc(nshift * sizeof(double));
  double *fin = (double *)malloc(nshift * sizeof(double));
  work3 w = work_new(p);
#define MDAGM(o, i, d)                                             \
  do {                                                             \
    if (p->action == 0)                                            \
      orc_dwf_mdagm(o, u, i, d, mass, p, w.t1, w.t2);              \
    else                                                           \
      orc_mobius_mdagm(o, u, i, d, mass, p, w.t1, w.t2, w.t3);     \
  } while (0)
  /* d = <p0, (A + shift[0]) p0> (:147-156, :229-238) */
#define SHIFTED_A(dout)                                            \
  do {                                                             \
    if (shift[0] > 0) {                                            \
      MDAGM(ap, pv, NULL);                                         \
      f_times_v1_plus_v2(ap, shift[0], pv, ap, n);                 \
      double dd_ = 0.0;                                            \
      for (size_t i_ = 0; i_ < n; ++i_) dd_ += pv[i_] * ap[i_];    \
      dout = dd_;                                                  \
    } else {                                                       \
      MDAGM(ap, pv, &dout);                                        \
    }                                                              \
  } while (0)

  const int nvec = type == 0 ? 1 : nshift;
  if (type == 1) memset(psi, 0, (size_t)nshift * n * sizeof(double)); /* :118-121 */
  memcpy(r, chi, n * sizeof(double));                                   /* :123-125 */
  double cp = chi_norm;
  for (int s = 0; s < nshift; ++s) memcpy(pv + (size_t)s * n, r, n * sizeof(double));
  for (int s = 0; s < nshift; ++s) {
    rsd_sq[s] = cp * rsdcg[s] * rsdcg[s]; /* :136-140 */
  }
  double d;
  SHIFTED_A(d);
  double b = -cp / d; /* :160 */
  z[0][0] = 1.0;
  z[1][0] = 1.0;
  bs[0] = b;
  int iz = 1;
  for (int s = 1; s < nshift; ++s) { /* :166-172 */
    z[1 - iz][s] = 1.0;
    z[iz][s] = 1.0 / (1.0 - b * (shift[s] - shift[0]));
    bs[s] = b * z[iz][s];
  }
  f_times_v1_plus_v2(r, b, ap, r, n); /* :175 */
  double c = orc_norm_sq(r, n);
  if (type == 0) {
    for (int s = 0; s < nshift; ++s) f_times_v1_plus_v2(psi, -bs[s] * alpha[s], chi, psi, n);
  } else {
    for (int s = 0; s < nshift; ++s)
      f_times_v1_plus_v2(psi + (size_t)s * n, -bs[s], chi, psi + (size_t)s * n, n);
  }
  int convp = c < rsd_sq[0] ? 1 : 0; /* :195 */
  double a = c / cp;
  double css = 0.0;
  int k;
  for (k = 1; k <= max_num_iter && !convp; k++) { /* :202 */
    a = c / cp;
    for (int s = 0; s < nshift; ++s) { /* :210-222 */
      if (convs[s]) continue;
      double *ps = pv + (size_t)s * n;
      if (s == 0) {
        f_times_v1_plus_v2(ps, a, ps, r, n);
      } else {
        const double as = a * (z[iz][s] * bs[s]) / (z[1 - iz][s] * b);
        for (size_t i = 0; i < n; ++i) ps[i] *= as;
        f_times_v1_plus_v2(ps, z[iz][s], r, ps, n);
      }
    }
    cp = c;
    SHIFTED_A(d);
    const double bp = b;
    b = -cp / d;
    bs[0] = b;
    iz = 1 - iz;
    for (int s = 1; s < nshift; ++s) { /* :246-254 */
      if (convs[s]) continue;
      const double ztmp = z[1 - iz][s] * z[iz][s] * bp /
                          (b * a * (z[iz][s] - z[1 - iz][s]) +
                           z[iz][s] * bp * (1 - b * (shift[s] - shift[0])));
      bs[s] = b * ztmp / z[1 - iz][s];
      z[iz][s] = ztmp;
    }
    f_times_v1_plus_v2(r, b, ap, r, n); /* :257 */
    c = orc_norm_sq(r, n);
    for (int s = 0; s < nshift; ++s) { /* :266-285 */
      if (convs[s]) continue;
      double *dst = type == 0 ? psi : psi + (size_t)s * n;
      const double f = type == 0 ? -bs[s] * alpha[s] : -bs[s];
      f_times_v1_plus_v2(dst, f, pv + (size_t)s * n, dst, n);
    }
    for (int s = 0; s < nshift; ++s) { /* :289-298 */
      if (convs[s]) continue;
      css = c * z[iz][s] * z[iz][s];
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
  for (int s = nshift - 1; s >= 0; s--) /* :330-345 */
    rsdcg[s] = sqrt(convs[s] ? fin[s] : c * z[iz][s] * z[iz][s]);
  if (iters)
    for (int s = 0; s < nshift; ++s) iters[s] = conv_at[s];
  (void)nvec;
#undef SHIFTED_A
#undef MDAGM
  free(r); free(ap); free(pv); free(convs); free(conv_at); free(bs);
  free(z[0]); free(z[1]); free(rsd_sq); free(fin); free(w.t1);
  return k;
}
