// dwf_kernels.hpp -- gfx950 device code for the even-odd preconditioned
// domain-wall / Mobius operator.
//
// Device layout (structure of arrays, one complex number per 8/16-byte word):
//   half vector  v[((s*12 + c)*vh + i)]     c = spin*3 + colour, i = site index
//   gauge        U[((mu*2 + parity)*9 + row*3 + col)*vh + i]   (canonical U(row,col),
//                boundary signs already folded in)
// Site index of a checkerboard site (z- and t-tiled, fastest first):
//   i = x/2 + lx/2*(t%bt + bt*(z%bz + bz*(y + ly*(z/bz + lz/bz*(t/bt)))))
// so y-, z- and t-hops inside a bz x bt tile stay within a few hundred sites
// and the neighbours of a work-group stay within one XCD's L2.
// A thread owns one output site (i, s); a work-group holds NS consecutive
// sites times all Ls slices, so the 8 links of a site are fetched once per
// group and served to all slices from L1/L2, and a whole s-column of a site
// sits in one group (needed by the Mobius M5^-1 recurrence, done in LDS).
// Work-groups are renumbered so each XCD sweeps one contiguous range.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dwfk {

constexpr int MAXLS = 32;

template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return {s * a.re, s * a.im}; }
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// conj(a) * b
template <typename R>
__device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b) {
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// sg * i * a  (sg = +-1)
template <int SG, typename R>
__device__ __forceinline__ cx<R> sgi(cx<R> a) {
  return SG > 0 ? cx<R>{-a.im, a.re} : cx<R>{a.im, -a.re};
}
template <int SG, typename R>
__device__ __forceinline__ cx<R> sgr(cx<R> a) {
  return SG > 0 ? a : cx<R>{-a.re, -a.im};
}

struct Geom {
  int lx, ly, lz, lt, lxh, vh, ls;
  int bt, bz;  // time / z slices per tile of the site index (divide lt / lz)
  int nxcd;    // XCDs the work-groups are spread over
};

// site index of (x/2, y, z, t): order (fastest first) x/2, t%bt, z%bz, y,
// z/bz, t/bt
__host__ __device__ __forceinline__ int site_index(const Geom& g, int xh, int y, int z, int t) {
  const int tb = t / g.bt, ti = t - tb * g.bt;
  const int zb = z / g.bz, zi = z - zb * g.bz;
  return xh + g.lxh * (ti + g.bt * (zi + g.bz * (y + g.ly * (zb + (g.lz / g.bz) * tb))));
}

// site index of a CPS checkerboard index (x/2) + (lx/2)(y + ly(z + lz t))
__host__ __device__ __forceinline__ int site_of_cps(const Geom& g, int c) {
  const int xh = c % g.lxh;
  int r = c / g.lxh;
  const int y = r % g.ly;
  r /= g.ly;
  const int z = r % g.lz;
  const int t = r / g.lz;
  return site_index(g, xh, y, z, t);
}

struct Nbr {
  int fwd[4], bwd[4];
};

__host__ __device__ __forceinline__ Nbr neighbours(const Geom& g, int i, int parity) {
  const int xh = i % g.lxh;
  int r = i / g.lxh;
  const int ti = r % g.bt;
  r /= g.bt;
  const int zi = r % g.bz;
  r /= g.bz;
  const int y = r % g.ly;
  r /= g.ly;
  const int nzb = g.lz / g.bz;
  const int z = (r % nzb) * g.bz + zi;
  const int t = (r / nzb) * g.bt + ti;
  const int x = 2 * xh + ((parity + y + z + t) & 1);
  const int xp = (x + 1 == g.lx) ? 0 : x + 1, xm = (x == 0) ? g.lx - 1 : x - 1;
  const int yp = (y + 1 == g.ly) ? 0 : y + 1, ym = (y == 0) ? g.ly - 1 : y - 1;
  const int zp = (z + 1 == g.lz) ? 0 : z + 1, zm = (z == 0) ? g.lz - 1 : z - 1;
  const int tp = (t + 1 == g.lt) ? 0 : t + 1, tm = (t == 0) ? g.lt - 1 : t - 1;
  const int row = i - xh;  // index of (0, y, z, t)
  Nbr n;
  n.fwd[0] = (xp >> 1) + row;
  n.bwd[0] = (xm >> 1) + row;
  n.fwd[1] = site_index(g, xh, yp, z, t);
  n.bwd[1] = site_index(g, xh, ym, z, t);
  n.fwd[2] = site_index(g, xh, y, zp, t);
  n.bwd[2] = site_index(g, xh, y, zm, t);
  n.fwd[3] = site_index(g, xh, y, z, tp);
  n.bwd[3] = site_index(g, xh, y, z, tm);
  return n;
}

// Work-groups are dispatched round-robin over the XCDs (hardware id b runs on
// XCD b % nxcd).  Renumber them so that XCD k receives one contiguous range of
// logical ids; neighbouring sites then share that XCD's L2.
__device__ __forceinline__ int xcd_block(int nxcd) {
  const int b = blockIdx.x, nb = gridDim.x;
  const int q = nb / nxcd, r = nb - q * nxcd;
  const int k = b % nxcd, j = b / nxcd;
  return k * q + (k < r ? k : r) + j;
}

// (1 + SG gamma_mu) psi -> two-spinor h (spins 0,1), gamma matrices of the
// chiral basis CPS uses (cps_pp/include/util/dirac_op.h:956-975).
template <int MU, int SG, typename R>
__device__ __forceinline__ void project(const cx<R>* p, cx<R>* h) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (MU == 0) {
      h[c] = p[c] + sgi<SG>(p[9 + c]);
      h[3 + c] = p[3 + c] + sgi<SG>(p[6 + c]);
    } else if (MU == 1) {
      h[c] = p[c] - sgr<SG>(p[9 + c]);
      h[3 + c] = p[3 + c] + sgr<SG>(p[6 + c]);
    } else if (MU == 2) {
      h[c] = p[c] + sgi<SG>(p[6 + c]);
      h[3 + c] = p[3 + c] - sgi<SG>(p[9 + c]);
    } else {
      h[c] = p[c] + sgr<SG>(p[6 + c]);
      h[3 + c] = p[3 + c] + sgr<SG>(p[9 + c]);
    }
  }
}

// accumulate the 4-spinor rebuilt from the transported two-spinor r
template <int MU, int SG, typename R>
__device__ __forceinline__ void reconstruct(cx<R>* a, const cx<R>* r) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    a[c] = a[c] + r[c];
    a[3 + c] = a[3 + c] + r[3 + c];
    if (MU == 0) {
      a[6 + c] = a[6 + c] - sgi<SG>(r[3 + c]);
      a[9 + c] = a[9 + c] - sgi<SG>(r[c]);
    } else if (MU == 1) {
      a[6 + c] = a[6 + c] + sgr<SG>(r[3 + c]);
      a[9 + c] = a[9 + c] - sgr<SG>(r[c]);
    } else if (MU == 2) {
      a[6 + c] = a[6 + c] - sgi<SG>(r[c]);
      a[9 + c] = a[9 + c] + sgi<SG>(r[3 + c]);
    } else {
      a[6 + c] = a[6 + c] + sgr<SG>(r[c]);
      a[9 + c] = a[9 + c] + sgr<SG>(r[3 + c]);
    }
  }
}

// plain load of the input spinor at checkerboard site n of one slice
template <typename R>
struct PlainLoader {
  const cx<R>* __restrict__ base;
  int vh;
  __device__ __forceinline__ void operator()(int n, cx<R>* p) const {
#pragma unroll
    for (int c = 0; c < 12; ++c) p[c] = base[c * vh + n];
  }
};

// on-the-fly (b + c D5) of zmobius_dslash_4_dag0 (zmobius_dslash_4.C:90-105):
// t(s) = b in(s) + c P+ in(s-1) + c P- in(s+1), walls carry -m c.
template <typename R>
struct MobTLoader {
  const cx<R>* __restrict__ cur;
  const cx<R>* __restrict__ prv;
  const cx<R>* __restrict__ nxt;
  int vh;
  R b, cm, cp;
  __device__ __forceinline__ void operator()(int n, cx<R>* p) const {
#pragma unroll
    for (int c = 0; c < 6; ++c) p[c] = b * cur[c * vh + n] + cm * prv[c * vh + n];
#pragma unroll
    for (int c = 6; c < 12; ++c) p[c] = b * cur[c * vh + n] + cp * nxt[c * vh + n];
  }
};

// Links of the work-group's sites, staged once in LDS and shared by all Ls
// slices: lu[((set*8 + d)*ns + site)*9 + e], d < 4 the forward link U_d(x),
// d >= 4 the backward link U_{d-4}(x - d) of the output site x; `set`
// distinguishes the two output parities of a DWF work-group.
template <typename R>
__device__ __forceinline__ void stage_links(cx<R>* __restrict__ lu, const cx<R>* __restrict__ U,
                                            const Geom& g, int blk, int ns, int nsets,
                                            int p_first) {
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int nthr = blockDim.x * blockDim.y;
  for (int w = tid; w < nsets * 8 * ns; w += nthr) {
    const int il = w % ns, d = (w / ns) % 8, set = w / (8 * ns);
    const int i = blk * ns + il;
    if (i >= g.vh) continue;
    const int p = (p_first + set) & 1;
    int site, par;
    if (d < 4) {
      site = i;
      par = p;
    } else {
      site = neighbours(g, i, p).bwd[d - 4];
      par = 1 - p;
    }
    const int mu = d & 3;
    const cx<R>* src = U + (size_t)((mu * 2 + par) * 9) * g.vh + site;
    cx<R>* dst = lu + (size_t)w * 9;
#pragma unroll
    for (int e = 0; e < 9; ++e) dst[e] = src[(size_t)e * g.vh];
  }
}

// lf / lb: this site's forward / backward link of direction MU (9 elements)
template <int MU, int DAG, typename R, class LD>
__device__ __forceinline__ void hop_dir(cx<R>* acc, const LD& ld, const cx<R>* lf,
                                        const cx<R>* lb, const Nbr& nb) {
  constexpr int SF = DAG ? 1 : -1;  // forward: 1 - s gamma, s = +1 (dag 0) / -1 (dag 1)
  cx<R> psi[12], h[6], r[6];
  {
    ld(nb.fwd[MU], psi);
    project<MU, SF>(psi, h);
    cx<R> m[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m[e] = lf[e];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int row = 0; row < 3; ++row)
        r[3 * k + row] = cmul(m[row * 3 + 0], h[3 * k + 0]) + cmul(m[row * 3 + 1], h[3 * k + 1]) +
                         cmul(m[row * 3 + 2], h[3 * k + 2]);
    reconstruct<MU, SF>(acc, r);
  }
  {
    ld(nb.bwd[MU], psi);
    project<MU, -SF>(psi, h);
    cx<R> m[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m[e] = lb[e];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int row = 0; row < 3; ++row)
        r[3 * k + row] = cmulc(m[0 * 3 + row], h[3 * k + 0]) + cmulc(m[1 * 3 + row], h[3 * k + 1]) +
                         cmulc(m[2 * 3 + row], h[3 * k + 2]);
    reconstruct<MU, -SF>(acc, r);
  }
}

// acc = D_hop^(dag) psi at output site i of parity p_out (wilson_dslash.C:173-517);
// lu points at this site's 8 staged links, consecutive directions `ds` apart.
template <int DAG, typename R, class LD>
__device__ __forceinline__ void hop4(cx<R>* acc, const LD& ld, const cx<R>* lu, int ds,
                                     const Geom& g, int i, int p_out) {
  const Nbr nb = neighbours(g, i, p_out);
  hop_dir<0, DAG>(acc, ld, lu + 0 * ds, lu + 4 * ds, nb);
  hop_dir<1, DAG>(acc, ld, lu + 1 * ds, lu + 5 * ds, nb);
  hop_dir<2, DAG>(acc, ld, lu + 2 * ds, lu + 6 * ds, nb);
  hop_dir<3, DAG>(acc, ld, lu + 3 * ds, lu + 7 * ds, nb);
}

// sum over the work-group; result valid in thread 0.  Fixed order, so the
// total is reproducible run to run.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int nthr = blockDim.x * blockDim.y;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (tid == 0)
    for (int w = 0; w < (nthr + 63) / 64; ++w) r += sh[w];
  return r;
}

// ---------------------------------------------------------------------------
// DWF: dwf_dslash = D4 on every slice with 4D parity (s+cb)%2 + 5th-dim hop
// (dwf_dslash_4.C:107-137, dwf_dslash_5_plus_omp.C:115-270).
// EPI 0: out = D in.   EPI 1: out = x0 + alpha D in.  NORM: |out|^2 partials.
// ---------------------------------------------------------------------------
template <typename R, int DAG, int EPI, bool NORM, int WPE>
__global__ void __launch_bounds__(512, WPE)
k_dwf_dslash(cx<R>* __restrict__ out, const cx<R>* __restrict__ in, const cx<R>* __restrict__ U,
             const cx<R>* __restrict__ x0, double* __restrict__ partial, Geom g, int cb_in,
             double two_a5, double nm_two_a5, double alpha) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double sh[16];
  cx<R>* lu = reinterpret_cast<cx<R>*>(smem_raw);  // [2 parities][8][ns][9]
  const int ns = blockDim.x, il = threadIdx.x, s = threadIdx.y;
  const int blk = xcd_block(g.nxcd);
  const int i = blk * ns + il;
  // slice s holds output parity (s + 1 + cb_in) % 2: set s % 2 of the stage
  stage_links(lu, U, g, blk, ns, 2, (1 + cb_in) & 1);
  __syncthreads();
  double nrm = 0.0;
  if (i < g.vh) {
    const int p_out = (s + 1 + cb_in) & 1;
    const size_t slice = (size_t)12 * g.vh;
    cx<R> acc[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) acc[c] = cx<R>{0, 0};
    hop4<DAG>(acc, PlainLoader<R>{in + s * slice, g.vh}, lu + ((s & 1) * 8 * ns + il) * 9, ns * 9,
              g, i, p_out);
    // fifth dimension: the same checkerboard index i in slices s-1 / s+1
    const cx<R>* sm = in + (s > 0 ? (s - 1) : (g.ls - 1)) * slice + i;
    const cx<R>* sp = in + (s < g.ls - 1 ? (s + 1) : 0) * slice + i;
    const R cm = (R)(s > 0 ? two_a5 : nm_two_a5);
    const R cp = (R)(s < g.ls - 1 ? two_a5 : nm_two_a5);
    const int lo_m = DAG ? 6 : 0;  // components fed from s-1
    const int lo_p = DAG ? 0 : 6;  // components fed from s+1
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      acc[lo_m + c] = acc[lo_m + c] + cm * sm[(lo_m + c) * g.vh];
      acc[lo_p + c] = acc[lo_p + c] + cp * sp[(lo_p + c) * g.vh];
    }
    cx<R>* o = out + s * slice + i;
    if (EPI == 0) {
#pragma unroll
      for (int c = 0; c < 12; ++c) o[c * g.vh] = acc[c];
    } else {
      const cx<R>* xs = x0 + s * slice + i;
      const R a = (R)alpha;
#pragma unroll
      for (int c = 0; c < 12; ++c) {
        const cx<R> v = xs[c * g.vh] + a * acc[c];
        o[c * g.vh] = v;
        if (NORM) nrm += (double)v.re * v.re + (double)v.im * v.im;
      }
    }
  }
  if (NORM) {
    const double t = block_sum(nrm, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.x] = t;
  }
}

// ---------------------------------------------------------------------------
// Mobius (zmobius ORIG arithmetic with real b, c):
//   EPI_M5INV : first half of MatPc / MatPcDag -> even sites
//       dag 0: out = M5^-1 kappa_b D_hop (b + c D5) in
//       dag 1: out = kappa_b M5^-dag (-kappa_b) (b + c D5^dag) D_hop^dag in
//   EPI_SCHUR : second half -> odd sites
//       dag 0: out = psi + K D5 psi - kappa_b D_hop (b + c D5) in      (+ |out|^2)
//       dag 1: out = (b + c D5^dag) D_hop^dag in + psi + K D5^dag psi
//   EPI_PLAIN : zmobius_dslash_4 alone (parity 1-cb)
// (zmobius_m-orig.h:3-126, zmobius_mdag-orig.h:4-119, zmobius_dslash_4.C:34-221)
// ---------------------------------------------------------------------------
enum { EPI_M5INV = 0, EPI_SCHUR = 1, EPI_PLAIN = 2 };

// One chirality's M5^-1 (or M5^-dag) along s: coefficients of the two-sweep
// recurrences of zmobius_m5inv.C:247-396 / :534-679, real for real b, c.
struct M5Side {
  double af[MAXLS];   // fwd-corner sweep: v[s] += af[s] v[s-1]
  double cf[MAXLS];   // fwd-corner correction: v[s] += cf[s] v[ls-1]
  double ab[MAXLS];   // bwd-corner sweep: v[s] += ab[s] v[s+1]
  double cbk[MAXLS];  // bwd-corner gather: v[ls-1] += cbk[s] v[s]
  double inv_f, inv_b;
};

struct MobArgs {
  double b, c, mc;     // Mobius b, c and -mass*c (wall coefficient)
  double K, mK;        // kappa_b/kappa_c and -mass*K
  double kappa_b;
  M5Side m5;
};

template <typename R>
__device__ __forceinline__ void m5_fwd_corner(cx<R>* v, int ls, const M5Side& k) {
  for (int s = 1; s < ls; ++s) v[s * 12] = v[s * 12] + (R)k.af[s] * v[(s - 1) * 12];
  const cx<R> zl = v[(ls - 1) * 12];
  for (int s = 0; s < ls - 1; ++s) v[s * 12] = v[s * 12] + (R)k.cf[s] * zl;
  v[(ls - 1) * 12] = (R)k.inv_f * zl;
}

template <typename R>
__device__ __forceinline__ void m5_bwd_corner(cx<R>* v, int ls, const M5Side& k) {
  cx<R> last = (R)k.inv_b * v[(ls - 1) * 12];
  for (int s = 0; s < ls - 1; ++s) last = last + (R)k.cbk[s] * v[s * 12];
  v[(ls - 1) * 12] = last;
  for (int s = ls - 2; s >= 0; --s) v[s * 12] = v[s * 12] + (R)k.ab[s] * v[(s + 1) * 12];
}

// M5^-1 (dag 0: P+ forward-corner, P- backward-corner; dag 1 swapped) on the
// s-columns staged in LDS; one thread per (site, complex component).
template <int DAG, typename R>
__device__ __forceinline__ void m5inv_columns(cx<R>* col, int sst, int ns, int nvalid, int ls,
                                              const M5Side& k) {
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int nthr = blockDim.x * blockDim.y;
  for (int w = tid; w < ns * 12; w += nthr) {
    const int sl = w / 12, cc = w % 12;
    if (sl >= nvalid) continue;
    cx<R>* v = col + sl * sst + cc;
    const bool upper = cc < 6;
    if (upper == (DAG == 0))
      m5_fwd_corner(v, ls, k);
    else
      m5_bwd_corner(v, ls, k);
  }
}

// FL flags of k_mobius
enum { FL_NORM = 1, FL_PRE_T = 2, FL_POST_T = 4 };
// FL_NORM   : |out|^2 partials (dag 0 EPI_SCHUR)
// FL_PRE_T  : dag 0, `in` already holds (b + c D5) applied to the operand, so
//             the gather is a plain load
// FL_POST_T : dag 0 EPI_M5INV, write (b + c D5) M5^-1 (...) so that the next
//             pass can use FL_PRE_T
template <typename R>
__device__ __forceinline__ void apply_t_col(cx<R>* tv, const cx<R>* col_site, int s, int ls,
                                            const MobArgs& a, int dag) {
  // (b + c D5) on one s-column held in LDS: dag 0 takes P+ from s-1 and P-
  // from s+1, dag 1 the reverse (zmobius_dslash_4.C:90-105, :189-202)
  const cx<R>* h0 = col_site + s * 12;
  const cx<R>* hm = col_site + (s > 0 ? s - 1 : ls - 1) * 12;
  const cx<R>* hp = col_site + (s < ls - 1 ? s + 1 : 0) * 12;
  const R cm = (R)(s > 0 ? a.c : a.mc), cp = (R)(s < ls - 1 ? a.c : a.mc);
  const R b = (R)a.b;
#pragma unroll
  for (int c = 0; c < 12; ++c) {
    const bool from_minus = (c < 6) == (dag == 0);
    tv[c] = b * h0[c] + (from_minus ? cm * hm[c] : cp * hp[c]);
  }
}

template <typename R, int DAG, int EPI, int FL, int WPE>
__global__ void __launch_bounds__(512, WPE)
k_mobius(cx<R>* __restrict__ out, const cx<R>* __restrict__ in, const cx<R>* __restrict__ psi,
         const cx<R>* __restrict__ U, double* __restrict__ partial, Geom g, int p_out, MobArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double sh[16];
  const int ns = blockDim.x, il = threadIdx.x, s = threadIdx.y;
  cx<R>* lu = reinterpret_cast<cx<R>*>(smem_raw);  // staged links [8][ns][9]
  cx<R>* col = lu + 8 * ns * 9;                     // s-columns [ns][sst]
  const int sst = g.ls * 12 + 1;                    // padded site stride in LDS
  const int blk = xcd_block(g.nxcd);
  const int i = blk * ns + il;
  const bool act = i < g.vh;
  const int nvalid = min(ns, g.vh - blk * ns);
  const size_t slice = (size_t)12 * g.vh;
  const int ls = g.ls;

  stage_links(lu, U, g, blk, ns, 1, p_out);
  __syncthreads();
  cx<R> acc[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) acc[c] = cx<R>{0, 0};
  if (act) {
    if (DAG == 0 && !(FL & FL_PRE_T)) {
      MobTLoader<R> ld;
      ld.cur = in + s * slice;
      ld.prv = in + (s > 0 ? s - 1 : ls - 1) * slice;
      ld.nxt = in + (s < ls - 1 ? s + 1 : 0) * slice;
      ld.vh = g.vh;
      ld.b = (R)a.b;
      ld.cm = (R)(s > 0 ? a.c : a.mc);
      ld.cp = (R)(s < ls - 1 ? a.c : a.mc);
      hop4<0>(acc, ld, lu + il * 9, ns * 9, g, i, p_out);
    } else {
      hop4<DAG>(acc, PlainLoader<R>{in + s * slice, g.vh}, lu + il * 9, ns * 9, g, i, p_out);
    }
  }
  cx<R>* o = out + s * slice + i;

  if (DAG == 0 && EPI == EPI_PLAIN) {
    if (act)
#pragma unroll
      for (int c = 0; c < 12; ++c) o[c * g.vh] = acc[c];
    return;
  }

  if (DAG == 0 && EPI == EPI_SCHUR) {
    double nrm = 0.0;
    if (act) {
      const cx<R>* p0 = psi + s * slice + i;
      const cx<R>* pm = psi + (s > 0 ? s - 1 : ls - 1) * slice + i;
      const cx<R>* pp = psi + (s < ls - 1 ? s + 1 : 0) * slice + i;
      const R km = (R)(s > 0 ? a.K : a.mK), kp = (R)(s < ls - 1 ? a.K : a.mK);
      const R nkb = (R)(-a.kappa_b);
#pragma unroll
      for (int c = 0; c < 12; ++c) {
        const cx<R> d5 = (c < 6) ? km * pm[c * g.vh] : kp * pp[c * g.vh];
        const cx<R> v = p0[c * g.vh] + d5 + nkb * acc[c];
        o[c * g.vh] = v;
        if (FL & FL_NORM) nrm += (double)v.re * v.re + (double)v.im * v.im;
      }
    }
    if (FL & FL_NORM) {
      const double t = block_sum(nrm, sh);
      if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.x] = t;
    }
    return;
  }

  cx<R>* mycol = col + il * sst;
  if (DAG == 1) {
    // (b + c D5^dag) on the hop output: needs s-1 / s+1 of the same site
#pragma unroll
    for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = acc[c];
    __syncthreads();
    cx<R> tv[12];
    apply_t_col(tv, mycol, s, ls, a, 1);
    if (EPI == EPI_PLAIN) {
      if (act)
#pragma unroll
        for (int c = 0; c < 12; ++c) o[c * g.vh] = tv[c];
      return;
    }
    if (EPI == EPI_SCHUR) {
      if (act) {
        const cx<R>* p0 = psi + s * slice + i;
        const cx<R>* pm = psi + (s > 0 ? s - 1 : ls - 1) * slice + i;
        const cx<R>* pp = psi + (s < ls - 1 ? s + 1 : 0) * slice + i;
        const R km = (R)(s > 0 ? a.K : a.mK), kp = (R)(s < ls - 1 ? a.K : a.mK);
#pragma unroll
        for (int c = 0; c < 12; ++c) {
          const cx<R> d5 = (c < 6) ? kp * pp[c * g.vh] : km * pm[c * g.vh];
          o[c * g.vh] = tv[c] + p0[c * g.vh] + d5;
        }
      }
      return;
    }
    __syncthreads();
    const R nkb = (R)(-a.kappa_b);
#pragma unroll
    for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = nkb * tv[c];
  } else {
    const R kb = (R)a.kappa_b;
#pragma unroll
    for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = kb * acc[c];
  }
  __syncthreads();
  m5inv_columns<DAG>(col, sst, ns, nvalid, ls, a.m5);
  __syncthreads();
  if (act) {
    if (DAG == 0 && (FL & FL_POST_T)) {
      cx<R> tv[12];
      apply_t_col(tv, mycol, s, ls, a, 0);
#pragma unroll
      for (int c = 0; c < 12; ++c) o[c * g.vh] = tv[c];
    } else {
      const R sc = (R)(DAG ? a.kappa_b : 1.0);
#pragma unroll
      for (int c = 0; c < 12; ++c) o[c * g.vh] = sc * mycol[s * 12 + c];
    }
  }
}

// stand-alone M5^-1 (dag) on a half vector: zmobius_m5inv(out, in, ...)
// (zmobius_m5inv.C:822-841)
template <typename R, int DAG>
__global__ void __launch_bounds__(512)
k_m5inv(cx<R>* __restrict__ out, const cx<R>* __restrict__ in, Geom g, M5Side k) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<R>* col = reinterpret_cast<cx<R>*>(smem_raw);
  const int ns = blockDim.x, il = threadIdx.x, s = threadIdx.y;
  const int sst = g.ls * 12 + 1;
  const int blk = xcd_block(g.nxcd);
  const int i = blk * ns + il;
  const bool act = i < g.vh;
  const int nvalid = min(ns, g.vh - blk * ns);
  const size_t slice = (size_t)12 * g.vh;
  cx<R>* mycol = col + il * sst;
  if (act)
#pragma unroll
    for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = in[s * slice + c * g.vh + i];
  __syncthreads();
  m5inv_columns<DAG>(col, sst, ns, nvalid, g.ls, k);
  __syncthreads();
  if (act)
#pragma unroll
    for (int c = 0; c < 12; ++c) out[s * slice + c * g.vh + i] = mycol[s * 12 + c];
}

// ---------------------------------------------------------------------------
// BLAS-1 on flat arrays of n reals (vector_util.C:727-900), FP64 reductions
// written as per-block partial sums (fixed grid => reproducible).
// ---------------------------------------------------------------------------
constexpr int BLAS_BLOCKS = 1024;
constexpr int BLAS_THREADS = 256;

template <typename R>
__global__ void __launch_bounds__(BLAS_THREADS)
k_norm(const R* __restrict__ x, size_t n, double* __restrict__ partial) {
  __shared__ double sh[16];
  double acc = 0.0;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    acc += (double)x[k] * x[k];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// sol += a dir ; res -= a mmp ; partial |res|^2   (inv_cg.C:340-349)
template <typename R>
__global__ void __launch_bounds__(BLAS_THREADS)
k_cg_update(R* __restrict__ sol, R* __restrict__ res, const R* __restrict__ dir,
            const R* __restrict__ mmp, double a, size_t n, double* __restrict__ partial) {
  __shared__ double sh[16];
  double acc = 0.0;
  const R ar = (R)a, mar = (R)(-a);
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    sol[k] = ar * dir[k] + sol[k];
    const R r = mar * mmp[k] + res[k];
    res[k] = r;
    acc += (double)r * r;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// y = a x + b y   (fTimesV1PlusV2 and friends)
template <typename R>
__global__ void __launch_bounds__(BLAS_THREADS)
k_axpby(R* __restrict__ y, const R* __restrict__ x, double a, double b, size_t n) {
  const R ar = (R)a, br = (R)b;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    y[k] = ar * x[k] + br * y[k];
}

// z = a x + y, partial |z|^2 optional
template <typename R, bool NORM>
__global__ void __launch_bounds__(BLAS_THREADS)
k_axpy_to(R* __restrict__ z, const R* __restrict__ x, const R* __restrict__ y, double a,
          size_t n, double* __restrict__ partial) {
  __shared__ double sh[16];
  double acc = 0.0;
  const R ar = (R)a;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const R v = ar * x[k] + y[k];
    z[k] = v;
    if (NORM) acc += (double)v * v;
  }
  if (NORM) {
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// per-slice complex-free scale of a half vector by a real factor
template <typename R>
__global__ void __launch_bounds__(BLAS_THREADS) k_scale(R* __restrict__ x, double a, size_t n) {
  const R ar = (R)a;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    x[k] = ar * x[k];
}

// reference streaming copy (16-byte words)
__global__ void __launch_bounds__(256)
k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n16;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

// deterministic final sum of per-block partials (one work-group of 1024
// threads, four independent accumulators per thread, fixed order)
__global__ void __launch_bounds__(1024)
k_reduce_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
  __shared__ double sh[16];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int k = threadIdx.x;
  for (; k + 3 * 1024 < n; k += 4 * 1024) {
    a0 += partial[k];
    a1 += partial[k + 1024];
    a2 += partial[k + 2 * 1024];
    a3 += partial[k + 3 * 1024];
  }
  for (; k < n; k += 1024) a0 += partial[k];
  const double t = block_sum((a0 + a1) + (a2 + a3), sh);
  if (threadIdx.x == 0) out[0] = t;
}

// CPS half vector [s][cb site][24 reals] (FP64, host order) <-> device SoA
template <typename R>
__global__ void __launch_bounds__(256)
k_cps_to_soa(cx<R>* __restrict__ soa, const double* __restrict__ cps, Geom g) {
  const int vh = g.vh, ls = g.ls;
  const size_t nsite = (size_t)vh * ls;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nsite * 12;
       k += (size_t)gridDim.x * blockDim.x) {
    const size_t site = k / 12;  // s*vh + i
    const int c = (int)(k % 12);
    const size_t s = site / vh, i = site_of_cps(g, (int)(site % vh));
    soa[(s * 12 + c) * vh + i] = cx<R>{(R)cps[k * 2], (R)cps[k * 2 + 1]};
  }
}

template <typename R>
__global__ void __launch_bounds__(256)
k_soa_to_cps(double* __restrict__ cps, const cx<R>* __restrict__ soa, Geom g) {
  const int vh = g.vh, ls = g.ls;
  const size_t nsite = (size_t)vh * ls;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nsite * 12;
       k += (size_t)gridDim.x * blockDim.x) {
    const size_t site = k / 12;
    const int c = (int)(k % 12);
    const size_t s = site / vh, i = site_of_cps(g, (int)(site % vh));
    const cx<R> v = soa[(s * 12 + c) * vh + i];
    cps[k * 2] = (double)v.re;
    cps[k * 2 + 1] = (double)v.im;
  }
}

}  // namespace dwfk
