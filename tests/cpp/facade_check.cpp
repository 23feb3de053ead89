// Drives the C++ facade (include/dwf_gpu.hpp) the way CPS host code drives
// DiracOpDwf / DiracOpMobius, and checks it against the C ABI:
//   MatPcDag(MatPc(x)) == MatPcDagMatPc(x) bit for bit,
//   InvCg through the facade == dwfgpu_invcg (same iterations, same bits),
//   InvCg(true_res) on the constructor's vectors == InvCg(out, in, true_res),
//   a caller-supplied src_norm_sq and a shifted solve (InvCgShift),
//   |src - A sol| / |src| recomputed here == the returned true residual,
//   MAXITER is reported as a count, a bad argument as dwfgpu::Error;
//   MInvCG solves every shift, ImpResLanczos returns eigenpairs, the
//   Ffour2five -> MatInv -> Ffive2four column converges.
// Prints "facade_check OK" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dwf_gpu.hpp"
#include "dwf_synth.h"

#define REQUIRE(c)                                                      \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "facade_check: %s failed (line %d)\n", #c, __LINE__); \
      return 1;                                                         \
    }                                                                   \
  } while (0)

static double norm2(const std::vector<double>& v) {
  double s = 0;
  for (double x : v) s += x * x;
  return s;
}

template <class Op>
static int run(const char* name) {
  dwfgpu::Params p;
  p.dims[0] = p.dims[1] = p.dims[2] = 4;
  p.dims[3] = 8;
  p.ls = 8;
  const size_t v4 = 4 * 4 * 4 * 8;
  std::vector<double> gauge(v4 * 4 * 18);
  dwfsynth_gauge(gauge.data(), p.dims, 0.3, 11);
  dwfgpu::CgArg arg;
  arg.mass = 0.04;
  arg.max_num_iter = 5000;
  arg.stop_rsd = 1e-10;
  Op op(gauge.data(), p, arg);
  const size_t n = op.HalfSize();
  REQUIRE(n == 24 * v4 / 2 * (size_t)p.ls);
  std::vector<double> x(n), a(n), b(n), c(n);
  dwfsynth_gaussian(x.data(), n, 5);

  double dot = 0;
  op.MatPcDagMatPc(a.data(), x.data(), &dot);
  op.MatPc(b.data(), x.data());
  op.MatPcDag(c.data(), b.data());
  REQUIRE(std::memcmp(a.data(), c.data(), n * sizeof(double)) == 0);
  REQUIRE(std::fabs(dot - norm2(b)) <= 1e-12 * dot);

  std::vector<double> s1(n, 0.0), s2(n, 0.0);
  double tr1 = 0, tr2 = 0;
  const int it1 = op.InvCg(s1.data(), x.data(), &tr1);
  int it2 = 0;
  const dwfgpu_cg_arg ca = arg.c();
  REQUIRE(dwfgpu_invcg(op.handle(), &ca, s2.data(), x.data(), &tr2, &it2) == DWFGPU_OK);
  REQUIRE(it1 == it2 && tr1 == tr2);
  REQUIRE(std::memcmp(s1.data(), s2.data(), n * sizeof(double)) == 0);
  op.MatPcDagMatPc(a.data(), s1.data());
  for (size_t k = 0; k < n; ++k) a[k] = x[k] - a[k];
  const double tr = std::sqrt(norm2(a) / norm2(x));
  REQUIRE(tr <= 1e-10 && std::fabs(tr - tr1) <= 1e-6 * tr1);

  // InvCg(Float *true_res) on the constructor's vectors, the form CPS's DWF
  // solve calls (FdwfBase::FmatEvlInv, f_dwf_base.C:200): the same solve
  std::vector<double> s4(n, 0.0);
  op.SetVectors(s4.data(), x.data());
  double tr4 = 0;
  REQUIRE(op.InvCg(&tr4) == it1 && tr4 == tr1);
  REQUIRE(std::memcmp(s4.data(), s1.data(), n * sizeof(double)) == 0);
  // InvCg(out, in, src_norm_sq, true_res): src_norm_sq = 4 |x|^2 at stop_rsd
  // r is the computed-norm solve at 2r, with half its true residual
  dwfgpu::CgArg wide = arg;
  wide.stop_rsd = 2e-8;
  op.DiracArg(wide);
  std::vector<double> s5(n, 0.0), s6(n, 0.0);
  double tr5 = 0, tr6 = 0;
  const int it6 = op.InvCg(s6.data(), x.data(), &tr6);
  wide.stop_rsd = 1e-8;
  op.DiracArg(wide);
  REQUIRE(op.InvCg(s5.data(), x.data(), 4.0 * norm2(x), &tr5) == it6);
  REQUIRE(std::fabs(tr5 - 0.5 * tr6) <= 1e-12 * tr6);
  op.DiracArg(arg);
  // InvCgShift: (A + sigma) sol = x
  const double sigma = 0.05;
  std::vector<double> s7(n, 0.0);
  double tr7 = 0;
  const int it7 = op.InvCgShift(s7.data(), x.data(), 0.0, &tr7, &sigma);
  op.MatPcDagMatPc(a.data(), s7.data());
  for (size_t q = 0; q < n; ++q) a[q] = x[q] - a[q] - sigma * s7[q];
  REQUIRE(it7 > 1 && it7 < it1 && tr7 <= 1e-10);
  REQUIRE(std::fabs(std::sqrt(norm2(a) / norm2(x)) - tr7) <= 1e-6 * tr7);

  dwfgpu::CgArg few = arg;
  few.max_num_iter = 3;
  few.stop_rsd = 1e-14;
  op.DiracArg(few);
  std::vector<double> s3(n, 0.0);
  REQUIRE(op.InvCg(s3.data(), x.data()) == 3);

  bool threw = false;
  try {
    op.Dslash(a.data(), x.data(), 2, 0);
  } catch (const dwfgpu::Error& e) {
    threw = e.code() == DWFGPU_BAD_ARG;
  }
  REQUIRE(threw);
  op.DiracArg(arg);

  // MInvCG (MULTI): every shift solved, (A + sigma) psi_s = chi
  const double shifts[3] = {0.0, 0.01, 0.1};
  double rsd[3] = {1e-10, 1e-10, 1e-10};
  std::vector<double> psi(3 * n, 0.0);
  int its[3] = {0, 0, 0};
  const int k = op.MInvCG(psi.data(), x.data(), shifts, 3, rsd, DWFGPU_SHIFT_MULTI, nullptr, its);
  REQUIRE(k > 0 && its[0] >= its[1] && its[1] >= its[2]);
  for (int j = 0; j < 3; ++j) {
    op.MatPcDagMatPc(a.data(), psi.data() + j * n);
    for (size_t q = 0; q < n; ++q) a[q] = x[q] - a[q] - shifts[j] * psi[j * n + q];
    REQUIRE(std::sqrt(norm2(a) / norm2(x)) < 1e-9);
  }

  // ImpResLanczos: the lowest pairs are eigenpairs of MatPcDagMatPc
  dwfgpu_lanczos_arg la{};
  la.nk = 8;
  la.nt = 2;
  la.np = 8;
  la.maxiters = 100;
  la.stop_residual = 1e-6;
  la.npol = 6;
  la.p0 = 2.2;  // above the spectrum of MatPcDagMatPc here
  la.p1 = 0.4;
  la.conv_check = 1;
  double lam[2];
  std::vector<double> ev(2 * n);
  REQUIRE(op.ImpResLanczos(x.data(), lam, ev.data(), la) >= 1);
  for (int j = 0; j < 2; ++j) {
    op.MatPcDagMatPc(a.data(), ev.data() + j * n);
    std::vector<double> v(ev.begin() + j * n, ev.begin() + (j + 1) * n);
    for (size_t q = 0; q < n; ++q) a[q] -= lam[j] * v[q];
    REQUIRE(lam[j] > 0 && std::sqrt(norm2(a) / norm2(v)) < 1e-4 * lam[j] + 1e-6);
  }

  // a 4D point source through Ffour2five -> MatInv -> Ffive2four
  std::vector<double> src4(v4 * 24, 0.0), sol4(v4 * 24, 0.0);
  src4[2 * 4] = 1.0;
  double ptr = 1;
  const int pit = op.PropagatorColumn(sol4.data(), src4.data(), p.ls, &ptr);
  REQUIRE(pit > 0 && ptr <= 1e-10 && norm2(sol4) > 0);
  std::printf("%s: %d CG iterations, true residual %.3e; MInvCG k=%d; lowest eigenvalues "
              "%.6e %.6e; propagator column %d iterations\n",
              name, it1, tr1, k, lam[0], lam[1], pit);
  return 0;
}

int main() {
  if (run<dwfgpu::DiracOpDwf>("DiracOpDwf")) return 1;
  if (run<dwfgpu::DiracOpMobius>("DiracOpMobius")) return 1;
  std::printf("facade_check OK\n");
  return 0;
}
