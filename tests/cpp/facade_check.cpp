// Drives the C++ facade (include/dwf_gpu.hpp) the way CPS host code drives
// DiracOpDwf / DiracOpMobius, and checks it against the C ABI:
//   MatPcDag(MatPc(x)) == MatPcDagMatPc(x) bit for bit,
//   InvCg through the facade == dwfgpu_invcg (same iterations, same bits),
//   |src - A sol| / |src| recomputed here == the returned true residual,
//   MAXITER is reported as a count, a bad argument as dwfgpu::Error.
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
  std::printf("%s: %d CG iterations, true residual %.3e\n", name, it1, tr1);
  return 0;
}

int main() {
  if (run<dwfgpu::DiracOpDwf>("DiracOpDwf")) return 1;
  if (run<dwfgpu::DiracOpMobius>("DiracOpMobius")) return 1;
  std::printf("facade_check OK\n");
  return 0;
}
