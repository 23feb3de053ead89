/*
 * roofline.h -- algorithmic cost model of one MatPcDagMatPc application.
 *
 * Minimum-traffic model: four hopping passes, each reading its input half
 * vector once and writing one half vector; the two passes that finish MatPc
 * and MatPcDag also read that operator's input once more (the Schur "1 +"
 * term); norms are fused.  All 4*V4 links are read once per pass.
 *   bytes = 240 * Nh * r + 288 * V4 * r_l
 * with Nh = V4*Ls/2, r = bytes per spinor real, r_l = bytes per stored link
 * real times (stored reals / 18).
 * Flops per Nh site use the nominal Wilson hop of 1320 flops:
 *   DWF    4*(1320+48) + 2*48 + 48             = 5616
 *   Mobius 4*(1320+72) + 2*96 + 2*(48+72) + 48 = 6048 (ORIG preconditioning)
 */
#ifndef DWF_ROOFLINE_H
#define DWF_ROOFLINE_H

#define DWF_MDAGM_SPINOR_REALS_PER_NH 240 /* 10 half-vector passes * 24 reals */
#define DWF_MDAGM_LINK_REALS_PER_V4 288   /* 4 passes * 4 links * 18 reals     */
#define DWF_MDAGM_FLOPS_PER_NH_DWF 5616
#define DWF_MDAGM_FLOPS_PER_NH_MOBIUS 6048

/* MI355X vendor peaks used as roofline denominators */
#define MI355X_HBM_PEAK_GBS 8000.0
#define MI355X_FP64_VECTOR_PEAK_TFLOPS 78.6

#endif
