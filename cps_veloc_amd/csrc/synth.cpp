// synth.cpp -- seeded synthetic lattice inputs (host C++, include/dwf_synth.h).
#include "dwf_synth.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

typedef std::complex<double> cd;

// SplitMix64 finaliser: a bijective mix of a 64-bit counter
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// uniform in (0, 1] from 53 random bits
inline double unit(uint64_t bits) { return ((bits >> 11) + 1) * (1.0 / 9007199254740992.0); }

// two independent unit Gaussians for counter k (Box-Muller)
inline void gauss2(uint64_t seed, uint64_t k, double* a, double* b) {
  const uint64_t key = mix64(seed ^ 0xD1B54A32D192ED03ULL);
  const double u1 = unit(mix64(key + 2 * k));
  const double u2 = unit(mix64(key + 2 * k + 1));
  const double r = std::sqrt(-2.0 * std::log(u1));
  const double ph = 6.283185307179586476925286766559 * u2;
  *a = r * std::cos(ph);
  *b = r * std::sin(ph);
}

template <class F>
void parallel_for(size_t n, F f) {
  unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  const size_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const size_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=] {
      for (size_t i = lo; i < hi; ++i) f(i);
    });
  }
  for (auto& x : th) x.join();
}

struct M3 {
  cd a[3][3];
};

inline M3 mul(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][0] * y.a[0][j] + x.a[i][1] * y.a[1][j] + x.a[i][2] * y.a[2][j];
  return r;
}

// exp(X): scale to norm <= 1/16, degree-10 Taylor, square back
M3 expm(const M3& x) {
  double nrm = 0;
  for (int i = 0; i < 3; ++i) {
    double row = 0;
    for (int j = 0; j < 3; ++j) row += std::abs(x.a[i][j]);
    nrm = std::max(nrm, row);
  }
  int nsq = nrm > 0 ? std::max(0, (int)std::ceil(std::log2(nrm)) + 4) : 0;
  const double sc = std::ldexp(1.0, -nsq);
  M3 y, out, term;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      y.a[i][j] = x.a[i][j] * sc;
      out.a[i][j] = term.a[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int k = 1; k <= 10; ++k) {
    term = mul(term, y);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        term.a[i][j] /= (double)k;
        out.a[i][j] += term.a[i][j];
      }
  }
  for (int k = 0; k < nsq; ++k) out = mul(out, out);
  return out;
}

// Matrix::Unitarize: normalise row 0, orthogonalise + normalise row 1,
// row 2 = conj(row0 x row1)
void unitarize(M3& u) {
  double n0 = 0;
  for (int j = 0; j < 3; ++j) n0 += std::norm(u.a[0][j]);
  n0 = 1.0 / std::sqrt(n0);
  for (int j = 0; j < 3; ++j) u.a[0][j] *= n0;
  cd dot = 0;
  for (int j = 0; j < 3; ++j) dot += std::conj(u.a[0][j]) * u.a[1][j];
  for (int j = 0; j < 3; ++j) u.a[1][j] -= u.a[0][j] * dot;
  double n1 = 0;
  for (int j = 0; j < 3; ++j) n1 += std::norm(u.a[1][j]);
  n1 = 1.0 / std::sqrt(n1);
  for (int j = 0; j < 3; ++j) u.a[1][j] *= n1;
  u.a[2][0] = std::conj(u.a[0][1] * u.a[1][2] - u.a[0][2] * u.a[1][1]);
  u.a[2][1] = std::conj(u.a[0][2] * u.a[1][0] - u.a[0][0] * u.a[1][2]);
  u.a[2][2] = std::conj(u.a[0][0] * u.a[1][1] - u.a[0][1] * u.a[1][0]);
}

}  // namespace

extern "C" {

void dwfsynth_gaussian(double* out, size_t n, unsigned long long seed) {
  const size_t npair = (n + 1) / 2;
  parallel_for(npair, [&](size_t k) {
    double a, b;
    gauss2(seed, k, &a, &b);
    out[2 * k] = a;
    if (2 * k + 1 < n) out[2 * k + 1] = b;
  });
}

void dwfsynth_gauge(double* out, const int dims[4], double eps, unsigned long long seed) {
  const size_t nlink = (size_t)4 * dims[0] * dims[1] * dims[2] * dims[3];
  const double s3 = 1.0 / std::sqrt(3.0);
  parallel_for(nlink, [&](size_t l) {
    double w[8];
    for (int q = 0; q < 4; ++q) gauss2(seed, 4 * (uint64_t)l + q, &w[2 * q], &w[2 * q + 1]);
    // i*eps*H with H = sum_a w_a lambda_a / 2
    cd H[3][3];
    H[0][1] = cd(w[0], -w[1]);
    H[1][0] = cd(w[0], w[1]);
    H[0][2] = cd(w[3], -w[4]);
    H[2][0] = cd(w[3], w[4]);
    H[1][2] = cd(w[5], -w[6]);
    H[2][1] = cd(w[5], w[6]);
    H[0][0] = w[2] + w[7] * s3;
    H[1][1] = -w[2] + w[7] * s3;
    H[2][2] = -2.0 * w[7] * s3;
    M3 x;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) x.a[i][j] = cd(0.0, 0.5 * eps) * H[i][j];
    M3 u = expm(x);
    unitarize(u);
    double* o = out + l * 18;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        o[2 * (3 * i + j)] = u.a[i][j].real();
        o[2 * (3 * i + j) + 1] = u.a[i][j].imag();
      }
  });
}

}  // extern "C"
