/*
 * cps_oracle.h -- CPU restatement of CPS's even-odd preconditioned DWF / Mobius
 * operator (MatPcDagMatPc) and the DiracOp::InvCgShift solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the parity checker for the HIP
 * library in cps_veloc_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls it.
 *
 * All arithmetic is FP64.  Every layout is CPS's own:
 *   gauge   (WILSON order)  u[(cbsite + vh*parity)*72 + mu*18 + re + 2*(row + 3*col)]
 *           even sites first; element (row,col) is canonical U(row,col)
 *           (cps_pp/src/util/lattice/convert/convert_func.C:372-374,403-405).
 *   spinor  [spin 4][colour 3][re,im] = 24 reals
 *           (cps_pp/src/util/dirac_op/d_op_wilson/noarch/wilson_dslash.C:60).
 *   cb site (x/2) + (lx/2)*(y + ly*(z + lz*t))          (wilson_dslash.C:288).
 *   DWF half vector    [s][cbsite][24], slice s holds 4D parity (s+cb)%2
 *                      (d_op_dwf/noarch/dwf_dslash_4.C:107-137).
 *   Mobius half vector [s][cbsite][24], every slice holds the same 4D parity
 *                      (d_op_zmobius/noarch/zmobius_dslash_4.C:109-121).
 *
 * CPS cannot be compiled here from its own sources without a configure-
 * generated config.h and stand-ins for the QUDA header dirac_op.h includes,
 * and the reference holds no vectors for this path, so each function is
 * pinned by mathematics (tests/test_oracle.py against the independent numpy
 * operators of tests/textbook.py):
 *   orc_gauge_canonical_to_wilson, orc_fermion_extract_cb/insert_cb
 *       round trips; BondCond signs through test_boundary_conditions_vs_textbook
 *   orc_wilson_dslash       textbook hop on 3 lattice shapes x cb x dag, three
 *                           boundary conditions, free-field plane-wave eigenvalue
 *   orc_dwf_dslash_5_plus, orc_dwf_dslash
 *                           dense textbook Shamir operator, both dag
 *   orc_dwf_m/mdag/mdagm, orc_matpc, orc_matpcdagmatpc (DWF)
 *                           dense even-odd Schur complement of the textbook
 *                           operator; adjoint identity; MdagM = Mdag M
 *   orc_zmobius_kappa_dslash_5_plus, orc_zmobius_dslash_4, orc_mobius_coefficients,
 *   orc_mobius_m/mdag/mdagm (Mobius)
 *                           dense Schur complement of the textbook Mobius
 *                           operator for (b,c) = (1.5,0.5), (1,0), (2,1);
 *                           textbook (b + c D5) D4; adjoint identity
 *   orc_zmobius_m5inv       M5 M5^-1 = 1 and the dense inverse, Ls 4/8/12, dag 0/1
 *   orc_invcg, orc_norm_sq  exactly k+1 iterations on a source spanned by k
 *                           eigenvectors; solution vs dense solve; MAXITER and
 *                           converged-guess edges
 *   orc_invcg_shift         shifted k-eigenvector source: k+1 iterations and
 *                           the closed-form solution sum c_i v_i/(lambda_i+s);
 *                           solution vs the dense solve of (A + s) x = b and
 *                           vs the pinned multi-shift oracle; src_norm_sq =
 *                           4|b|^2 at stop_rsd r reproduces |b|^2 at 2r with
 *                           half the true residual
 *   orc_matinv              full textbook system solved for both actions
 *   orc_fermion5d_canonical_to_full / _full_to_canonical
 *                           against the slice-wise composition of
 *                           orc_fermion_extract_cb with the parity rule of
 *                           each action; round trip; every slice of DWF's
 *                           5D-odd half holds sites with x+y+z+t+s odd
 *   orc_four2five, orc_five2four
 *                           five2four(four2five(f)) = f for s_u != s_l and
 *                           s_u == s_l; zero outside the two chiral slices;
 *                           chirality of the slices (gamma5 = +1 at s_u)
 */
#ifndef CPS_ORACLE_H
#define CPS_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Lattice / action description shared by all oracle calls. */
typedef struct {
  int L[4];          /* local 4D extents lx,ly,lz,lt (all even)          */
  int ls;            /* fifth-dimension extent                           */
  double m5;         /* domain-wall height M5  (GJP.DwfHeight)           */
  double a5inv;      /* 1/a5                   (GJP.DwfA5Inv)            */
  double mob_b;      /* Mobius b (GJP.Mobius_b); unused for DWF          */
  double mob_c;      /* Mobius c (GJP.Mobius_c); unused for DWF          */
  int action;        /* 0 = DWF (DiracOpDwf), 1 = Mobius (DiracOpMobius) */
} orc_params;

/* ---- layouts / boundary conditions (a13) ---------------------------- */
/* CANONICAL gauge [site][mu][row][col][re,im] -> WILSON order, applying
 * the anti-periodic sign of Lattice::BondCond for every mu with bc[mu]!=0. */
void orc_gauge_canonical_to_wilson(double *u_wilson, const double *u_canon,
                                   const int L[4], const int bc[4]);
/* One CANONICAL 4D fermion slice [site][24] -> half-lattice of parity p. */
void orc_fermion_extract_cb(double *half, const double *canon, const int L[4], int p);
void orc_fermion_insert_cb(double *canon, const double *half, const int L[4], int p);

/* ---- f2: whole 5D fermion fields (CANONICAL: [s][4D site][24]) -------- */
/* Fconvert CANONICAL -> the MatInv full vector [first | second]: DWF
 * [5D-odd | 5D-even] (convert.C:704-757), Mobius [4D-even | 4D-odd]
 * (convert.C:797-831); and back.                                            */
void orc_fermion5d_canonical_to_full(double *full, const double *canon, const orc_params *p);
void orc_fermion5d_full_to_canonical(double *canon, const double *full, const orc_params *p);
/* Ffour2five / Ffive2four with Ncb = 2 (f_dwf_base.C:961-1168).             */
void orc_four2five(double *five, const double *four, int s_u, int s_l, const orc_params *p);
void orc_five2four(double *four, const double *five, int s_u, int s_l, const orc_params *p);

/* ---- a1: Wilson 4D hopping term ------------------------------------- */
void orc_wilson_dslash(double *chi, const double *u, const double *psi,
                       int cb, int dag, const int L[4]);
/* 1: the 4D hop walks its sites in one thread, as CPS's noarch
 * wilson_dslash.C:280-477 does (the timing of the CPS-faithful CPU baseline);
 * 0 (default): OpenMP over sites.  Same bits either way.                  */
void orc_set_serial_hop(int on);

/* ---- a2..a4: DWF ----------------------------------------------------- */
void orc_dwf_dslash_5_plus(double *out, const double *in, double mass, int dag,
                           const orc_params *p);
void orc_dwf_dslash(double *out, const double *u, const double *in, double mass,
                    int cb, int dag, const orc_params *p);
void orc_dwf_m(double *out, const double *u, const double *in, double mass,
               const orc_params *p, double *tmp2);
void orc_dwf_mdag(double *out, const double *u, const double *in, double mass,
                  const orc_params *p, double *tmp2);
void orc_dwf_mdagm(double *out, const double *u, const double *in, double *dot_prd,
                   double mass, const orc_params *p, double *tmp1, double *tmp2);

/* ---- a6..a9: Mobius (zmobius arithmetic, complex per-s coefficients) --- */
/* coefficient arrays are complex [ls][2]                                   */
void orc_zmobius_kappa_dslash_5_plus(double *out, const double *in, double mass,
                                     int dag, const double *fact, const orc_params *p);
void orc_zmobius_dslash_4(double *out, const double *u, const double *in, int cb,
                          int dag, double mass, const double *zb, const double *zc,
                          const orc_params *p, double *tmp3);
void orc_zmobius_m5inv(double *inout, double mass, int dag, const double *K,
                       const orc_params *p);
void orc_mobius_m(double *out, const double *u, const double *in, double mass,
                  const orc_params *p, double *tmp2, double *tmp3);
void orc_mobius_mdag(double *out, const double *u, const double *in, double mass,
                     const orc_params *p, double *tmp2, double *tmp3);
void orc_mobius_mdagm(double *out, const double *u, const double *in, double *dot_prd,
                      double mass, const orc_params *p, double *tmp1, double *tmp2,
                      double *tmp3);
/* Mobius per-s coefficient arrays as DiracOpMobius builds them
 * (d_op_mobius.C:105-125): zb=b, zc=c, kappa_b, kappa_ratio=kappa_b/kappa_c. */
void orc_mobius_coefficients(const orc_params *p, double *zb, double *zc,
                             double *zkappa_b, double *zkappa_ratio);

/* ---- a5/a10: DiracOp front ends --------------------------------------- */
/* Number of reals in one CG (half) vector: 24*V4*ls/2. */
size_t orc_half_size(const orc_params *p);
void orc_matpcdagmatpc(double *out, const double *u, const double *in, double *dot_prd,
                       double mass, const orc_params *p);
void orc_matpc(double *out, const double *u, const double *in, double mass,
               int dag, const orc_params *p);
/* DWF:    out (5D parity 1-cb) = dwf_dslash(in, cb, dag)
 * Mobius: out (4D parity 1-cb) = zmobius_dslash_4(in, cb, dag)            */
void orc_dslash(double *out, const double *u, const double *in, double mass,
                int cb, int dag, const orc_params *p);

/* ---- a11: DiracOp::InvCgShift (d_op_base/comsrc/inv_cg.C:84-448) -------- */
/* Returns CPS's count itr+1.  hist (may be NULL, length >= max_iter+1)
 * receives |res|^2 after each iteration, hist[0] = initial |res|^2.
 * src_norm_sq == 0: |src|^2 is computed (inv_cg.C:166-169); otherwise it sets
 * the stop rule and the true residual's denominator.  shift (may be NULL):
 * the solve is (MatPcDagMatPc + *shift) sol = src (inv_cg.C:248-251,304-310).
 * orc_invcg is InvCgShift(out, in, 0, true_res, NULL).                      */
int orc_invcg(double *sol, const double *src, const double *u, double mass,
              double stop_rsd, int max_iter, double *true_res, double *hist,
              const orc_params *p);
int orc_invcg_shift(double *sol, const double *src, const double *u, double mass,
                    double src_norm_sq, const double *shift, double stop_rsd, int max_iter,
                    double *true_res, double *hist, const orc_params *p);

/* ---- a12: BLAS-1 ------------------------------------------------------- */
double orc_norm_sq(const double *a, size_t n);   /* Vector::NormSqNode, serial */

/* ---- (f1): DiracOpDwf::MatInv / DiracOpMobius::MatInv ----------------- */
/* full vectors: DWF [5D-odd | 5D-even], Mobius [4D-even | 4D-odd]        */
int orc_matinv(double *out_full, double *in_full, const double *u, double mass,
               double stop_rsd, int max_iter, double *true_res, const orc_params *p);

/* ---- (f3): DiracOp::MInvCG (d_op_base/noarch/minvcg.C:47-355) ----------- */
/* Multi-shift CG: (A + shift[s]) psi_s = chi.  type 0 SINGLE (psi[0] =
 * sum alpha_s psi_s), 1 MULTI (psi = nshift vectors).  rsdcg in: relative
 * stop residuals; out: |r_s|.  iters[s]: iteration of convergence.  Returns
 * CPS's loop count k.                                                        */
int orc_minvcg(double *psi, const double *chi, double chi_norm, const double *shift,
               int nshift, double *rsdcg, int type, const double *alpha, int max_num_iter,
               int *iters, const double *u, double mass, const orc_params *p);

#ifdef __cplusplus
}
#endif
#endif
