/*
 * dwf_gpu.h -- C ABI of the MI355X (gfx950) domain-wall / Mobius operator and
 * CG solver.  Every entry point replaces one piece of CPS's DiracOp call
 * surface (paths relative to cps_pp/); plain pointers and sizes only.
 *
 * Vectors passed as host arrays are FP64 CPS half vectors, 24*V4*Ls/2 reals,
 * in CPS's own order [s][checkerboard site][spin][colour][re,im]:
 *   DWF    : the 5D-odd half, slice s holds 4D parity (s+1)%2
 *            (src/util/lattice/convert/convert.C:704-760)
 *   Mobius : the 4D-odd half (every slice 4D-odd)
 *            (src/util/lattice/convert/convert.C:797-831).
 * Full vectors (MatInv) are [5D-odd | 5D-even] for DWF and [4D-even | 4D-odd]
 * for Mobius, as CPS stores them.
 *
 * Handles are not thread-safe; every call is synchronous with respect to the
 * host and runs on the handle's own HIP stream.
 */
#ifndef DWF_GPU_H
#define DWF_GPU_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (CPS aborts through ERR.General instead) */
#define DWFGPU_OK 0
#define DWFGPU_BREAKDOWN 1 /* CG found <p, A p> == 0 (CPS: exit(5), inv_cg.C:326-331) */
#define DWFGPU_MAXITER 2   /* CG hit max_num_iter (CPS: VRB.Warn, inv_cg.C:392-396)  */
#define DWFGPU_BAD_ARG 3
#define DWFGPU_HIP_ERROR 4
#define DWFGPU_IO_ERROR 5       /* file missing, short, or header unreadable (ERR.FileR) */
#define DWFGPU_CHECKSUM_ERROR 6 /* NERSC data checksum differs from the header's        */

/* actions */
#define DWFGPU_DWF 0    /* DiracOpDwf    (src/util/dirac_op/d_op_dwf/d_op_dwf.C)       */
#define DWFGPU_MOBIUS 1 /* DiracOpMobius (src/util/dirac_op/d_op_mobius/d_op_mobius.C) */

/* storage precision of the operator's fields on the device (the handle's
 * precision in dwfgpu_create, the sloppy precision of dwfgpu_invcg_mixed).
 * F32, BF16 and I16 store links as their first two rows (the third is
 * rebuilt as the conjugate cross product times the boundary sign); BF16 and I16
 * store each (site, slice) spinor as 24 16-bit numbers with one FP32 scale
 * (BF16: value = bf16 * scale; I16: value = int16 * scale, |int16| <= 16383
 * with the site maximum at 16383), and compute in FP32.                     */
#define DWFGPU_F64 0
#define DWFGPU_F32 1
#define DWFGPU_BF16 2
#define DWFGPU_I16 3

typedef struct dwfgpu_s *dwfgpu_t;
typedef struct dwfgpu_vec_s *dwfgpu_vec_t;

/* Mirrors CgArg (include/input/cg_arg.x:3-28): the fields InvCg reads. */
typedef struct {
  double mass;
  int max_num_iter;
  double stop_rsd;
  double true_rsd; /* unused by InvCg, kept for layout parity with CgArg */
} dwfgpu_cg_arg;

/* Construct the operator: the DiracOp ctor's BondCond + Convert(WILSON) +
 * dwf_init / DiracOpMobius coefficient set-up
 * (d_op_base/comsrc/dirac_op_base.C:95-170, d_op_dwf/d_op_dwf.C:67-123,
 *  d_op_mobius/d_op_mobius.C:57-140).
 * gauge_canonical: 4*V4*18 FP64 reals, CANONICAL order
 *   [site x+lx*(y+ly*(z+lz*t))][mu][row][col][re,im] (include/util/lattice.h:327).
 * bc[mu] = 1 multiplies the links U_mu(x) with x_mu = L_mu-1 by -1
 *   (Lattice::BondCond, src/util/lattice/lattice_base/lattice_base.C:5336-5391).
 * mob_b, mob_c are ignored for DWF.  precision: the storage precision of the
 * handle's vectors and operator (DWFGPU_F64 ... DWFGPU_I16); host arrays are
 * FP64 in every case and are converted on the device.                       */
int dwfgpu_create(const double *gauge_canonical, const int dims[4], int ls, int action,
                  double m5, double a5inv, double mob_b, double mob_c, const int bc[4],
                  int precision, dwfgpu_t *out);
int dwfgpu_destroy(dwfgpu_t h);
/* number of reals in one half vector: 24*V4*Ls/2 */
size_t dwfgpu_half_size(dwfgpu_t h);
/* message describing the last failure on any handle of this thread */
const char *dwfgpu_last_error(void);

/* ---- host-array entry points (CPS half-vector order) ------------------ */
/* DiracOp::MatPcDagMatPc(out, in, dot_prd)       include/util/dirac_op.h:259
 *   (d_op_dwf.C:174-200 -> dwf_mdagm, d_op_mobius.C:200-218 -> mobius_mdagm)
 * dot_prd (may be NULL) receives |MatPc in|^2.                              */
int dwfgpu_matpcdagmatpc(dwfgpu_t h, double mass, double *out, const double *in,
                         double *dot_prd);
/* DiracOp::MatPc / MatPcDag  (d_op_dwf.C:245-300, d_op_mobius.C:320-372) */
int dwfgpu_matpc(dwfgpu_t h, double mass, double *out, const double *in, int dag);
/* DiracOp::Dslash(out, in, cb, dag)  (d_op_dwf.C:210-236 -> dwf_dslash;
 * d_op_mobius.C:275-291 -> mobius_dslash_4 = zmobius_dslash_4).
 * in: half of parity cb, out: half of parity 1-cb.                          */
int dwfgpu_dslash(dwfgpu_t h, double mass, double *out, const double *in, int cb,
                  int dag);
/* DiracOp::InvCg(out, in, true_res) (include/util/dirac_op.h:138) through
 * InvCgShift(out, in, 0, true_res, NULL) (d_op_base/comsrc/inv_cg.C:84-448):
 * sol holds the initial guess on entry.  *iters receives CPS's count
 * (loop iterations + 1, inv_cg.C:446).  Returns DWFGPU_MAXITER when the
 * iteration cap was hit (solution and true_res still filled).               */
int dwfgpu_invcg(dwfgpu_t h, const dwfgpu_cg_arg *arg, double *sol, const double *src,
                 double *true_res, int *iters);
/* DiracOp::InvCgShift(out, in, src_norm_sq, true_res, shift)
 * (include/util/dirac_op.h:112-116, d_op_base/comsrc/inv_cg.C:84-448), which
 * also serves InvCg(out, in, src_norm_sq, true_res) (dirac_op.h:118, shift
 * NULL).  src_norm_sq == 0: |src|^2 is computed (inv_cg.C:166-169); otherwise
 * the stop rule is |r|^2 <= stop_rsd^2 src_norm_sq and *true_res =
 * |src - A sol| / sqrt(src_norm_sq).  shift (may be NULL): solve
 * (MatPcDagMatPc + *shift) sol = src with d = <p, A p> + shift |p|^2
 * (inv_cg.C:248-251, 304-310).  dwfgpu_invcg is (src_norm_sq 0, shift NULL).*/
int dwfgpu_invcg_shift(dwfgpu_t h, const dwfgpu_cg_arg *arg, double *sol, const double *src,
                       double src_norm_sq, double *true_res, const double *shift, int *iters);
/* DiracOpDwf::MatInv (d_op_dwf.C:316-507) / DiracOpMobius::MatInv
 * (d_op_mobius.C:388-696, ORIG pc) on full vectors; in is preserved
 * (PRESERVE_YES), out holds the initial guess of its CG half on entry.      */
int dwfgpu_matinv(dwfgpu_t h, const dwfgpu_cg_arg *arg, double *out, const double *in,
                  double *true_res, int *iters);

/* ---- device-resident vectors (no PCIe traffic per call) --------------- */
int dwfgpu_vec_create(dwfgpu_t h, dwfgpu_vec_t *v);
int dwfgpu_vec_destroy(dwfgpu_vec_t v);
int dwfgpu_vec_upload(dwfgpu_vec_t v, const double *host_cps_order);
int dwfgpu_vec_download(dwfgpu_vec_t v, double *host_cps_order);
int dwfgpu_matpcdagmatpc_dev(dwfgpu_t h, double mass, dwfgpu_vec_t out, dwfgpu_vec_t in,
                             double *dot_prd);
/* No counterpart in CPS.  dwfgpu_matpcdagmatpc_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): the same kernels in the same order, with
 * their temporaries (te and t1 half vectors in the handle's precision, the
 * per-work-group partial sums, one device scalar) in one static device array
 * of DWFGPU_NOALLOC_SCRATCH_BYTES per GPU, held from module load to process
 * end.  It calls no hipMalloc / hipFree.  A lattice whose temporaries do not
 * fit (2 FP64 half vectors: up to about 690k CG sites, e.g. 16^3x16 Ls 10)
 * gets DWFGPU_BAD_ARG.  Every call on one GPU shares the array: calls that
 * overlap on one GPU (two host threads, two handles' streams) race.        */
#define DWFGPU_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int dwfgpu_matpcdagmatpc_dev_noalloc(dwfgpu_t h, double mass, dwfgpu_vec_t out,
                                     dwfgpu_vec_t in, double *dot_prd);
int dwfgpu_invcg_dev(dwfgpu_t h, const dwfgpu_cg_arg *arg, dwfgpu_vec_t sol,
                     dwfgpu_vec_t src, double *true_res, int *iters);
int dwfgpu_invcg_shift_dev(dwfgpu_t h, const dwfgpu_cg_arg *arg, dwfgpu_vec_t sol,
                           dwfgpu_vec_t src, double src_norm_sq, double *true_res,
                           const double *shift, int *iters);

/* ---- mixed precision ---------------------------------------------------- */
/* The solve of DiracOp::InvCg (inv_cg.C:84-448: same system, same stop rule
 * |src - A sol|^2 <= stop_rsd^2 |src|^2 on the FP64 true residual) with the
 * bulk of the MatPcDagMatPc applications in `sloppy` precision and FP64
 * reliable updates.  CPS hands this scheme to QUDA under USE_QUDA
 * (QudaInvert in DiracOpDwf::MatInv, d_op_dwf.C:425-426; sloppy gauge and
 * spinor precisions and reliable_delta, quda/wilson_quda.C:84,134-137).
 * A reliable update (sol += sloppy correction, r = src - A sol in FP64)
 * happens when the sloppy residual norm drops below delta (0 < delta < 1)
 * times the norm at the previous update, or meets the stop rule.  For
 * FP32 and I16 the host tests each sloppy step's residual two steps later
 * (the GPU runs ahead), so updates and the stop come two steps after their
 * condition holds; BF16 tests every step as it ends.
 * sloppy = DWFGPU_F64 runs the plain CG of dwfgpu_invcg (per-step test).
 * The handle's API precision must be DWFGPU_F64.
 * *iters: sloppy iterations + 1; *reliable_updates: their count.           */
int dwfgpu_invcg_mixed(dwfgpu_t h, const dwfgpu_cg_arg *arg, int sloppy, double delta,
                       double *sol, const double *src, double *true_res, int *iters,
                       int *reliable_updates);
int dwfgpu_invcg_mixed_dev(dwfgpu_t h, const dwfgpu_cg_arg *arg, int sloppy, double delta,
                           dwfgpu_vec_t sol, dwfgpu_vec_t src, double *true_res, int *iters,
                           int *reliable_updates);

/* ---- many right-hand sides (DWF) ---------------------------------------- */
/* nrhs FP64 half vectors on the device (an FP64 handle's layout), filled and
 * read one RHS at a time in CPS half-vector order.                           */
typedef struct dwfgpu_mvec_s *dwfgpu_mvec_t;
int dwfgpu_mvec_create(dwfgpu_t h, int nrhs, dwfgpu_mvec_t *out);
int dwfgpu_mvec_destroy(dwfgpu_mvec_t v);
int dwfgpu_mvec_upload(dwfgpu_mvec_t v, int j, const double *host_cps_order);
int dwfgpu_mvec_download(dwfgpu_mvec_t v, int j, double *host_cps_order);
/* MatPcDagMatPc of every RHS j with mass masses[j], applied in `precision`.
 * DWF: the batched operator (one launch per hopping pass for all RHS, links
 * staged once per work-group for all of them).  Mobius: the four passes of
 * each RHS (their M5^-1 / (b + c D5) epilogues keep a site's s-column in one
 * work-group) issued back to back.  The per-RHS loop over
 * DiracOpDwf/DiracOpMobius::MatPcDagMatPc (d_op_dwf.C:174-200,
 * d_op_mobius.C:200-218) that CPS callers such as QPropW (QPropW.C:1236) run. */
int dwfgpu_matpcdagmatpc_mrhs_dev(dwfgpu_t h, int precision, const double *masses,
                                  dwfgpu_mvec_t out, dwfgpu_mvec_t in);
/* InvCg's solve (inv_cg.C:84-448 stop rule per RHS) for every RHS j with its
 * own mass, in one batched CG: `sloppy` (F32, BF16 or I16) steps with FP64
 * reliable updates as in dwfgpu_invcg_mixed; an RHS that converges retires on
 * the device while the others go on; the CG step is captured in a HIP graph
 * and replayed (flags & DWFGPU_MRHS_EAGER: plain launches instead).  Per RHS:
 * true_res[j], iters[j] (sloppy steps + 1), reliable_updates[j], status[j]
 * (DWFGPU_OK, DWFGPU_MAXITER or DWFGPU_BREAKDOWN).  The launch shape of the
 * batched operator is chosen on first use by timing the candidates.         */
#define DWFGPU_MRHS_EAGER 1
int dwfgpu_invcg_mrhs_dev(dwfgpu_t h, const dwfgpu_cg_arg *arg, int sloppy, double delta,
                          int flags, const double *masses, dwfgpu_mvec_t sol, dwfgpu_mvec_t src,
                          double *true_res, int *iters, int *reliable_updates, int *status);
/* the launch shape (sites, RHS per work-group) the last batched call used */
int dwfgpu_mrhs_shape(dwfgpu_t h, int *ns, int *nr);
/* DiracOpMobius::MatInv(std::vector<Vector*> out, std::vector<Vector*> in,
 * true_res, iters) (d_op_mobius.C:407-696, ORIG pc; the per-RHS InvCg of
 * :552-554 becomes one batched solve) and the same for DWF handles (the
 * DiracOpDwf::MatInv of d_op_dwf.C:316-507 per RHS): for every RHS j, full
 * vectors [in_first | in_second] -> [out_first | out_second] as
 * dwfgpu_matinv_dev (DWF [5D-odd | 5D-even], Mobius [4D-even | 4D-odd]),
 * with mass masses[j] (CPS's vector form uses one mass for all).  The CG
 * sources are prepared and the solutions completed on the device; the solve
 * is dwfgpu_invcg_mrhs_dev (sloppy, delta, flags and the per-RHS outputs as
 * there; the CG half of out holds the initial guess on entry).             */
int dwfgpu_matinv_mrhs_dev(dwfgpu_t h, const dwfgpu_cg_arg *arg, int sloppy, double delta,
                           int flags, const double *masses, dwfgpu_mvec_t out_first,
                           dwfgpu_mvec_t out_second, dwfgpu_mvec_t in_first,
                           dwfgpu_mvec_t in_second, double *true_res, int *iters,
                           int *reliable_updates, int *status);

/* ---- multi-shift CG ------------------------------------------------------ */
/* DiracOp::MInvCG(psi, chi, chi_norm, shift, Nshift, isz, RsdCG, type, alpha)
 * (include/util/dirac_op.h:222, d_op_base/noarch/minvcg.C:47-355): solves
 * (MatPcDagMatPc + shifts[s]) psi_s = chi for all shifts with one operator
 * application per iteration, in FP64 with CPS's recurrences and stop tests.
 * type DWFGPU_SHIFT_MULTI: psi holds nshift vectors (zeroed first);
 * DWFGPU_SHIFT_SINGLE: psi is one vector that receives sum_s alpha[s] psi_s
 * (added to its entry value, as CPS does).  GENERAL is not supported.
 * chi_norm <= 0: |chi|^2 is computed.  rsdcg[s]: relative stop residual on
 * entry, |r_s| on return (minvcg.C:330-345); iters[s] (may be NULL): the
 * iteration at which shift s converged (0: it did not); *k: CPS's return
 * value.  Returns DWFGPU_MAXITER when some shift did not converge.
 * At most 32 shifts; FP64 handles only.                                      */
#define DWFGPU_SHIFT_SINGLE 0 /* MultiShiftSolveType SINGLE (include/alg/enum_int.h:233) */
#define DWFGPU_SHIFT_MULTI 1  /* MULTI */
int dwfgpu_minvcg(dwfgpu_t h, double mass, const double *shifts, int nshift, double *psi,
                  const double *chi, double chi_norm, double *rsdcg, int type,
                  const double *alpha, int max_num_iter, int *iters, int *k);
int dwfgpu_minvcg_dev(dwfgpu_t h, double mass, const double *shifts, int nshift,
                      dwfgpu_mvec_t psi, dwfgpu_vec_t chi, double chi_norm, double *rsdcg,
                      int type, const double *alpha, int max_num_iter, int *iters, int *k);

/* ---- learned preconditioner (no CPS counterpart) ------------------------ */
/* A gauge-equivariant parallel-transport network z = |r| N(r/|r|) acting on
 * CG vectors: three layers Ls -> 16 -> 16 -> Ls channels (Ls in {4,8,12,16}),
 *   out_o(x) = sum_{d=0..8} sum_i W[d][o][i] (T_d in_i)(x),
 * T_0 = 1, T_{+mu} f(x) = U_mu(x) f(x+mu), T_{-mu} f(x) = U_mu(x-mu)^dag f(x-mu)
 * (d = 1..4: +x..+t, 5..8: -x..-t), W[d][o][i] a complex 4x4 spin matrix.
 * Weights: complex numbers as (re, im) FP32 pairs, layers concatenated, each
 * [9][cout][cin][4][4]; dwfgpu_ptc_num_weights gives their count.  BF16 MFMA
 * kernels with FP32 accumulation run the layers and their gradients; the
 * handle's API precision must be DWFGPU_F64.                                  */
size_t dwfgpu_ptc_num_weights(dwfgpu_t h);
/* create the network (once) and set random weights: scale * N(0,1) on every
 * weight plus a unit diagonal on the local spin term, so that the initial
 * network is close to z = 0.5 r                                               */
int dwfgpu_ptc_init(dwfgpu_t h, unsigned long long seed, double scale);
int dwfgpu_ptc_set_weights(dwfgpu_t h, const float *w, size_t n);
int dwfgpu_ptc_get_weights(dwfgpu_t h, float *w, size_t n);
/* z = N(r) on host CPS half vectors (computed in FP32 / BF16) */
int dwfgpu_ptc_apply(dwfgpu_t h, double *out, const double *in);
/* loss |N(r)/|r| - e/|r||^2 / |e/|r||^2 of one pair and its gradient
 * dL/dRe W + i dL/dIm W (same layout as the weights)                         */
int dwfgpu_ptc_loss_grad(dwfgpu_t h, const double *r, const double *e, double *loss,
                         float *grads);
/* Training on the GPU: for each of nsolves Gaussian sources (seeded), solve
 * A x* = b, run FP32 CG from 0 keeping x_k every `stride` iterations
 * (pairs_per_solve of them), form pairs (r_k = A e_k, e_k = x* - x_k), add
 * random_pairs pairs (r = A e, e Gaussian), and run `epochs` Adam passes over
 * them in mini-batches of `batch`.                                           */
typedef struct {
  double mass;
  int nsolves, pairs_per_solve, random_pairs, stride, epochs, batch;
  double lr;
  unsigned long long seed;
} dwfgpu_ptc_train_arg;
int dwfgpu_ptc_train(dwfgpu_t h, const dwfgpu_ptc_train_arg *a, double *loss_first,
                     double *loss_last, int *pairs_used);
/* InvCg's solve (same stop rule on the FP64 true residual) by flexible CG with
 * the network as preconditioner, `sloppy` vectors and FP64 reliable updates;
 * plain CG takes over when the network stops helping (*net_steps counts the
 * steps that used it).  Without a network it is plain reliable-update CG.   */
int dwfgpu_invcg_ptc_dev(dwfgpu_t h, const dwfgpu_cg_arg *arg, int sloppy, double delta,
                         dwfgpu_vec_t sol, dwfgpu_vec_t src, double *true_res, int *iters,
                         int *reliable_updates, int *net_steps);
/* InvCg's solve (same stop rule on the FP64 true residual) by flexible GCR
 * with the network as preconditioner, restarted every nkrylov (1..64)
 * directions: each step z = N(r), w = A z made orthogonal to the cycle's
 * earlier w_j (complex Gram-Schmidt; the network is not Hermitian), x and r
 * updated to minimise |r| over the cycle's space; FP64 vectors and operator.
 * After 10 steps that cut |r|^2 by less than 0.1 %, after a cycle that does
 * not halve |r|^2, or without a network, plain CG (dwfgpu_invcg) finishes
 * from the current solution; *iters counts
 * both (CPS's convention, steps + 1), *net_steps the steps the network
 * helped, *restarts the completed cycles.  No CPS counterpart.            */
int dwfgpu_invgcr_ptc_dev(dwfgpu_t h, const dwfgpu_cg_arg *arg, int nkrylov, dwfgpu_vec_t sol,
                          dwfgpu_vec_t src, double *true_res, int *iters, int *net_steps,
                          int *restarts);

/* ---- instrumentation (the role of CPS's print_flops / CGflops) --------- */
/* Times `reps` MatPcDagMatPc applications on the handle's stream with HIP
 * events around each of its kernel launches.  precision: the handle's API
 * precision (in and out are used directly), or, on an FP64 handle, a sloppy
 * precision (in is converted first; out is untouched).  ms_app: mean per
 * application; ms_kernel[k] (k < *nkernel on return, at most 8): mean per
 * launch of the k-th kernel of one application; bytes_kernel[k]: its
 * algorithmic HBM bytes (inputs read once, outputs written once).           */
int dwfgpu_time_matpcdagmatpc(dwfgpu_t h, int precision, double mass, dwfgpu_vec_t out,
                              dwfgpu_vec_t in, int reps, double *ms_app, double *ms_kernel,
                              double *bytes_kernel, int *nkernel);
/* Algorithmic bytes and flops of one MatPcDagMatPc application in the given
 * storage precision (see include/roofline.h).                               */
int dwfgpu_matpcdagmatpc_cost(dwfgpu_t h, int precision, double *bytes, double *flops);
/* Reference copy kernel on the handle's stream: copies `bytes` (rounded
 * down to 16) between two fresh device buffers `reps` times; *gbs receives
 * (read + write bytes) / mean kernel time.                                   */
int dwfgpu_time_copy(dwfgpu_t h, size_t bytes, int reps, double *gbs);
/* Memory-bandwidth ceilings of this device, beside which the operator's
 * GB/s are read: the best over three grid sizes of a copy (*copy_gbs, read +
 * write bytes), a read-only pass (*read_gbs) and a pass reading three buffers
 * and writing one (*r3w1_gbs, the read:write mix of the hop kernels), each
 * over `bytes` per buffer with 4 independent 16-byte non-temporal loads in
 * flight per thread.  Allocates and frees 4 buffers of `bytes`.  No
 * counterpart in CPS (measurement only).                                     */
int dwfgpu_time_bandwidth(dwfgpu_t h, size_t bytes, int reps, double *copy_gbs,
                          double *read_gbs, double *r3w1_gbs);

/* ---- (f2) host formats of whole fermion fields and gauge files --------- */
/* Full 5D fermion in CANONICAL order, FdwfBase::Fconvert(CANONICAL <->
 * WILSON / DWF_4D_EOPREC_EE) (src/util/lattice/convert/convert.C:639-831):
 * canon = Ls slices of V4 sites, site x + lx (y + ly (z + lz t)), 24 reals
 * each (Ls*V4*24 reals).  first / second are the device halves of the full
 * vector in MatInv order: DWF [5D-odd | 5D-even], Mobius [4D-even | 4D-odd].
 * Vectors are the handle's own (dwfgpu_vec_create).                        */
int dwfgpu_fermion_from_canonical(dwfgpu_t h, dwfgpu_vec_t first, dwfgpu_vec_t second,
                                  const double *canon);
int dwfgpu_fermion_to_canonical(dwfgpu_t h, double *canon, dwfgpu_vec_t first,
                                dwfgpu_vec_t second);
/* FdwfBase::Ffour2five(five, four, s_u, s_l, Ncb = 2)
 * (src/util/lattice/f_dwf_base/f_dwf_base.C:961-1043) fused with the
 * conversion above: four = one CANONICAL 4D fermion (V4*24 reals); the 5D
 * field is zero except the upper spin pair (reals 0..11) at slice s_u and the
 * lower pair (12..23) at slice s_l.  0 <= s_u, s_l < Ls (one node holds
 * every slice), DWFGPU_BAD_ARG otherwise.                                   */
int dwfgpu_four2five(dwfgpu_t h, dwfgpu_vec_t first, dwfgpu_vec_t second, const double *four,
                     int s_u, int s_l);
/* FdwfBase::Ffive2four(four, five, s_u, s_l, Ncb = 2) (f_dwf_base.C:1064-1168):
 * four = upper pair of slice s_u and lower pair of slice s_l, CANONICAL.     */
int dwfgpu_five2four(dwfgpu_t h, double *four, dwfgpu_vec_t first, dwfgpu_vec_t second,
                     int s_u, int s_l);
/* DiracOpDwf::MatInv / DiracOpMobius::MatInv (d_op_dwf.C:316-507,
 * d_op_mobius.C:407-696) on device halves; the inputs are preserved
 * (PRESERVE_YES).  out and in must be distinct vectors.                     */
int dwfgpu_matinv_dev(dwfgpu_t h, const dwfgpu_cg_arg *arg, dwfgpu_vec_t out_first,
                      dwfgpu_vec_t out_second, dwfgpu_vec_t in_first, dwfgpu_vec_t in_second,
                      double *true_res, int *iters);

/* Header of a NERSC gauge file (LatticeHeader::read, src/util/qcdio/comsrc/
 * latheader.C:267-312).                                                      */
typedef struct {
  int dims[4];                /* DIMENSION_1..4                                 */
  int recon_row_3;            /* DATATYPE 4D_SU3_GAUGE: two rows stored         */
  int fp_bytes;               /* 4 or 8 (FLOATING_POINT)                        */
  int big_endian;             /* byte order of FLOATING_POINT                   */
  unsigned int checksum;      /* CHECKSUM of the header                         */
  unsigned int checksum_data; /* the same 32-bit sum over the data read         */
  double plaquette;           /* PLAQUETTE (read as single precision, as CPS)   */
  double link_trace;          /* LINK_TRACE                                     */
  long long data_start;       /* byte offset of the data                        */
} dwfgpu_nersc_header;
/* ReadLatticeParallel::read steps 0-2 (src/util/qcdio/comsrc/
 * ReadLatticePar.C:26-175): header, data in file order t,z,y,x / mu / row-
 * major 3x3 (the CANONICAL order), the FPConv 32-bit checksum, and the third
 * row conj(row0 x row1) when two rows are stored.  gauge = 4*V4*18 reals
 * (n_reals), or NULL to read the header only.  DWFGPU_IO_ERROR: no file,
 * no END_HEADER, bad dimensions, unknown FLOATING_POINT or short data;
 * DWFGPU_CHECKSUM_ERROR: the links are filled but the sums differ.  Host
 * code (file I/O); the plaquette check is dwfgpu_plaquette.                 */
int dwfgpu_nersc_read(const char *path, double *gauge, size_t n_reals,
                      dwfgpu_nersc_header *hdr);
/* Lattice::SumReTrPlaq / (18 V) and sum ReTr U / (12 V) of the handle's
 * links without the boundary signs, as ReadLatticeParallel::
 * CheckPlaqLinktrace compares them with the header (ReadLatticePar.C:396-450).
 * Either pointer may be NULL.                                                */
int dwfgpu_plaquette(dwfgpu_t h, double *plaq, double *link_trace);

/* ---- (f4) eigCG deflation ---------------------------------------------- */
/* DiracOp::InvEigCg (src/util/dirac_op/d_op_base/comsrc/inv_eig_cg.C:129-631),
 * FP64 handle.  CG on MatPcDagMatPc from sol = U invH U^dag src (def_len > 0;
 * else 0).  With nev > 0 an eigCG window of m > 2 nev normalised residuals is
 * kept in V (>= m vectors); each time it fills, the Ritz pairs of the window's
 * Lanczos matrix in the span of its nev lowest eigenvectors and those of its
 * leading (m-1) block replace it: V[0..rank) and M[0..rank), M[rank..2nev) = 0
 * (2 nev <= 32, m <= 256; M is zeroed on entry).  U: >= def_len vectors; invH: def_len x
 * def_len complex, row-major, re/im interleaved.  restart[0..restart_len):
 * relative residuals at which the init-CG correction U invH U^dag r is
 * applied again.  *iters = CPS's iter (loop count); DWFGPU_MAXITER, or
 * DWFGPU_BREAKDOWN when <p, A p> = 0 (CPS breaks out of the loop).          */
int dwfgpu_inveigcg_dev(dwfgpu_t h, const dwfgpu_cg_arg *arg, dwfgpu_vec_t sol, dwfgpu_vec_t src,
                        int nev, int m, dwfgpu_mvec_t V, double *M, dwfgpu_mvec_t U,
                        const double *invH, int def_len, const double *restart, int restart_len,
                        double *true_res, int *iters);
/* The deflation-space bookkeeping of QPropW::eig_Run (src/alg/alg_qpropw/
 * eig_QPropW.C:85-330) around InvEigCg, one call per right-hand side:
 * invH = (U^dag A U)^-1 of the current space, the solve (an eigCG window
 * while the space is not full; restarts at restart[] when always_restart or
 * once it is full), then the new Ritz vectors with value in (1e-30, max_eig),
 * lowest first, rounded to single precision, appended to U with their row
 * and column of H = U^dag A U, up to max_def_len vectors.                   */
typedef struct dwfgpu_eigcg_s *dwfgpu_eigcg_t;
int dwfgpu_eigcg_create(dwfgpu_t h, int nev, int m, int max_def_len, double max_eig,
                        dwfgpu_eigcg_t *out);
int dwfgpu_eigcg_solve(dwfgpu_eigcg_t e, const dwfgpu_cg_arg *arg, dwfgpu_vec_t sol,
                       dwfgpu_vec_t src, const double *restart, int restart_len,
                       int always_restart, double *true_res, int *iters);
int dwfgpu_eigcg_def_len(dwfgpu_eigcg_t e);
/* M: the 2 nev Ritz values of the last window; H: def_len^2 complex (re/im). */
int dwfgpu_eigcg_values(dwfgpu_eigcg_t e, double *M, double *H);
int dwfgpu_eigcg_destroy(dwfgpu_eigcg_t e);

/* Mirrors LanczosArg (include/input/eig_arg.x) for DiracOp::ImpResLanczos
 * with RitzMat_lanczos = MATPCDAG_MATPC through the Chebyshev filter of
 * DiracOpDwf::RitzMat (d_op_dwf.C:852-940) and RitzMat_convcheck =
 * MATPCDAG_MATPC.                                                            */
typedef struct {
  double mass;
  int nk, nt, np;       /* kept, wanted, extra Lanczos vectors (nk + np <= 256) */
  int maxiters;         /* restarts                                           */
  double stop_residual; /* |A v - <v, A v> v| / |v| of the nt wanted vectors   */
  int npol;             /* Chebyshev degree (>= 1)                            */
  double p0, p1, shift; /* filter: a = p0^2, b = (p1 + |shift|)^2             */
  int conv_check;       /* check every conv_check restarts                    */
} dwfgpu_lanczos_arg;
/* DiracOp::ImpResLanczos (src/util/dirac_op/d_op_base/comsrc/lanczos.C:
 * 193-840), FP64 handle: implicitly restarted Lanczos on the filtered
 * MatPcDagMatPc with single-precision Lanczos vectors, exact shifts (the np
 * smallest filtered Ritz values), a Gram-Schmidt of every new residual
 * against all Lanczos vectors.  v0: the start vector (the CPS half vector
 * passed in V[0]).  lambda[0..nt): Rayleigh quotients of the nt returned
 * vectors (lowest modes of A for an even npol); evecs (host, nt*n reals, or
 * NULL) receive the vectors, rounded to single precision as CPS keeps them.
 * *iters = CPS's return value (restarts).                                   */
int dwfgpu_lanczos(dwfgpu_t h, const dwfgpu_lanczos_arg *arg, const double *v0,
                   double *evecs, double *lambda, int *iters);

#ifdef __cplusplus
}
#endif
#endif
