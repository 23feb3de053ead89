/*
 * dwf_synth.h -- seeded synthetic inputs (host side, FP64), the role CPS's
 * disordered start configuration and Lattice::RandGaussVector play
 * (cps_pp/src/util/lattice/lattice_base/lattice_base.C:195-202, :4380-4460).
 *
 * A counter-based generator makes every value a function of (seed, index)
 * only, so results do not depend on the number of host threads.
 */
#ifndef DWF_SYNTH_H
#define DWF_SYNTH_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CANONICAL gauge field [site][mu][row][col][re,im] (4*V4*18 reals):
 * U = reunitarize(exp(i*eps*H)), H = sum_a w_a lambda_a / 2 with w_a seeded
 * unit Gaussians (lambda_a the Gell-Mann matrices); reunitarize follows
 * Matrix::Unitarize (cps_pp/src/util/vector/comsrc/unitarize.C:28-56).     */
void dwfsynth_gauge(double *out, const int dims[4], double eps, unsigned long long seed);

/* n seeded unit Gaussian reals (real and imaginary parts independent). */
void dwfsynth_gaussian(double *out, size_t n, unsigned long long seed);

#ifdef __cplusplus
}
#endif
#endif
