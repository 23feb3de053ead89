// dwf_kernels.hpp -- gfx950 device code for the even-odd preconditioned
// domain-wall / Mobius operator and its CG vector algebra.
//
// Fields are stored in one of four policies (compute type R, stored type S):
//   P64   : FP64 spinors, FP64 links (18 reals)
//   P32   : FP32 spinors, FP32 links compressed to their first two rows
//   PBF16 : bfloat16 spinors with one FP32 scale per (site, slice), bf16
//           compressed links
//   PI16  : int16 fixed-point spinors with one FP32 scale per (site, slice),
//           int16 compressed links (|U_ij| <= 1)
// Compressed links are rebuilt as row2 = conj(row0 x row1) times the
// anti-periodic boundary sign when they are staged in LDS.
//
// Device layout (structure of arrays of 16-byte chunks):
//   spinor data   d[(s*NCH + k)*vh + i]  chunk k holds the complex components
//                 c = spin*3 + colour in [k*CPC, (k+1)*CPC): CPC = 1 (P64),
//                 2 (P32), 4 (16-bit), NCH = 12/CPC; one 16-byte load each
//   spinor scale  sc[s*vh + i]           (scaled policies only)
//   links         U[((mu*2 + parity)*NE + e)*vh + i]   NE = 9 (P64) or 6
//                 e = row*3 + col of canonical U(row, col), boundary signs folded in
// Site index of a checkerboard site (z- and t-tiled, fastest first):
//   i = x/2 + lx/2*(t%bt + bt*(z%bz + bz*(y + ly*(z/bz + lz/bz*(t/bt)))))
// so y-, z- and t-hops inside a bz x bt tile stay within a few hundred sites.
// A thread owns one output site (i, s); a work-group holds ns consecutive
// sites times all Ls slices.  The group stages the 8 links of each of its
// sites in LDS once (all slices share them), and the Mobius M5^-1 and
// (b + c D5) epilogues run on the s-columns of its sites in LDS.
// Work-groups are renumbered so each XCD sweeps one contiguous site range.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
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
// One row of U h and of U^dag h (3 complex products summed).  FP32 (VALU-
// bound kernels): chains of fused multiply-adds, the fewest instructions.
// FP64 (HBM-bound, two waves per SIMD): three independent products then a
// sum, which leaves more independent work to hide FP64 latency.
template <typename R>
__device__ __forceinline__ cx<R> row_dot(cx<R> a0, cx<R> a1, cx<R> a2, cx<R> b0, cx<R> b1,
                                         cx<R> b2) {
  if constexpr (sizeof(R) == 8) return cmul(a0, b0) + cmul(a1, b1) + cmul(a2, b2);
  R re = a0.re * b0.re;
  re = fma(-a0.im, b0.im, re);
  re = fma(a1.re, b1.re, re);
  re = fma(-a1.im, b1.im, re);
  re = fma(a2.re, b2.re, re);
  re = fma(-a2.im, b2.im, re);
  R im = a0.re * b0.im;
  im = fma(a0.im, b0.re, im);
  im = fma(a1.re, b1.im, im);
  im = fma(a1.im, b1.re, im);
  im = fma(a2.re, b2.im, im);
  im = fma(a2.im, b2.re, im);
  return {re, im};
}
// conj(a0) b0 + conj(a1) b1 + conj(a2) b2
template <typename R>
__device__ __forceinline__ cx<R> row_dotc(cx<R> a0, cx<R> a1, cx<R> a2, cx<R> b0, cx<R> b1,
                                          cx<R> b2) {
  if constexpr (sizeof(R) == 8) return cmulc(a0, b0) + cmulc(a1, b1) + cmulc(a2, b2);
  R re = a0.re * b0.re;
  re = fma(a0.im, b0.im, re);
  re = fma(a1.re, b1.re, re);
  re = fma(a1.im, b1.im, re);
  re = fma(a2.re, b2.re, re);
  re = fma(a2.im, b2.im, re);
  R im = a0.re * b0.im;
  im = fma(-a0.im, b0.re, im);
  im = fma(a1.re, b1.im, im);
  im = fma(-a1.im, b1.re, im);
  im = fma(a2.re, b2.im, im);
  im = fma(-a2.im, b2.re, im);
  return {re, im};
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

// ---------------------------------------------------------------------------
// storage policies
// ---------------------------------------------------------------------------
// R: compute type; S: one stored complex number (links use S);
// C: the 16-byte chunk of CPC complex numbers spinors are stored in
struct P64 {
  typedef double R;
  typedef cx<double> S;
  typedef cx<double> C;
  static constexpr bool kScaled = false;
  static constexpr int kLinkElems = 9;
  static constexpr int kCPC = 1;
};
struct P32 {
  typedef float R;
  typedef cx<float> S;
  typedef float4 C;
  static constexpr bool kScaled = false;
  static constexpr int kLinkElems = 6;
  static constexpr int kCPC = 2;
};
struct PBF16 {
  typedef float R;
  typedef uint32_t S;  // re in the low half, im in the high half
  typedef uint4 C;
  static constexpr bool kScaled = true;
  static constexpr int kLinkElems = 6;
  static constexpr int kCPC = 4;
};
struct PI16 {
  typedef float R;
  typedef uint32_t S;
  typedef uint4 C;
  static constexpr bool kScaled = true;
  static constexpr int kLinkElems = 6;
  static constexpr int kCPC = 4;
};
template <class P>
constexpr int nch() { return 12 / P::kCPC; }

// scale stored per (site, slice) from the largest |component| m of the site
template <class P>
__device__ __forceinline__ float scale_of(float) { return 1.0f; }
template <>
__device__ __forceinline__ float scale_of<PBF16>(float m) { return m; }
// int16 components are stored in [-16383, 16383] (site maximum -> 16383), so
// that the hop can add two of them in int16 before converting
template <>
__device__ __forceinline__ float scale_of<PI16>(float m) { return m * (1.0f / 16383.0f); }
// factor of 1 / m handed to the packer: pknorm maps [-1, 1] to +-32767
template <class P>
__device__ __forceinline__ float pack_norm() { return 1.0f; }
template <>
__device__ __forceinline__ float pack_norm<PI16>() { return 16383.0f / 32767.0f; }

// FP32 -> bf16 bits, rounded to nearest even, in integer arithmetic.  The
// operator's BF16 packing keeps this form: with the hardware conversion
// below the BF16 Schur passes were given twice the registers (94 -> 189
// VGPRs) and ran 53 % slower.
__device__ __forceinline__ uint32_t bf16_rne(float x) {
  uint32_t b = __float_as_uint(x);
  b += 0x7fffu + ((b >> 16) & 1u);
  return b >> 16;
}
// (lo, hi) -> bf16 bits lo | hi << 16, rounded to nearest even by gfx950's
// two-value conversion (v_cvt_pk_bf16_f32); used by the network kernels
__device__ __forceinline__ uint32_t bf16_pair(float lo, float hi) {
  typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, bf16x2_t{(__bf16)lo, (__bf16)hi});
}

template <class P>
struct Pack;
template <>
struct Pack<P64> {
  __device__ static cx<double> unpack(cx<double> v, float) { return v; }
  __device__ static cx<double> pack(cx<double> v, float) { return v; }
};
template <>
struct Pack<P32> {
  __device__ static cx<float> unpack(cx<float> v, float) { return v; }
  __device__ static cx<float> pack(cx<float> v, float) { return v; }
};
template <>
struct Pack<PBF16> {
  __device__ static cx<float> unpack(uint32_t v, float sc) {
    return {__uint_as_float(v << 16) * sc, __uint_as_float(v & 0xffff0000u) * sc};
  }
  __device__ static uint32_t pack(cx<float> v, float inv) {
    return bf16_rne(v.re * inv) | (bf16_rne(v.im * inv) << 16);
  }
};
template <>
struct Pack<PI16> {
  __device__ static cx<float> unpack(uint32_t v, float sc) {
    return {(float)(int16_t)(v & 0xffffu) * sc, (float)(int16_t)(v >> 16) * sc};
  }
  // v * inv in [-1, 1] (inv from st_site, or 1 for links): the
  // signed-normalised conversion clamps to [-1, 1], scales by 32767 and
  // rounds to nearest, two components per instruction
  __device__ static uint32_t pack(cx<float> v, float inv) {
    typedef short short2_t __attribute__((ext_vector_type(2)));
    const short2_t q = __builtin_amdgcn_cvt_pknorm_i16(v.re * inv, v.im * inv);
    return ((uint32_t)(uint16_t)q.x) | ((uint32_t)(uint16_t)q.y << 16);
  }
};
// fixed scale of compressed 16-bit links
template <class P>
__device__ __forceinline__ float link_scale() { return 1.0f; }
template <>
__device__ __forceinline__ float link_scale<PI16>() { return 1.0f / 32767.0f; }

// a spinor field: data + per-(site,slice) scales (scaled policies only)
template <class P>
struct Fld {
  typename P::C* d;
  float* sc;
};
template <class P>
struct CFld {
  const typename P::C* d;
  const float* sc;
};

// chunk <-> CPC complex numbers
template <class P>
__device__ __forceinline__ void chunk_unpack(const typename P::C& w, float sc,
                                             cx<typename P::R>* v) {
  if constexpr (P::kCPC == 1) {
    v[0] = w;
  } else if constexpr (P::kCPC == 2) {
    v[0] = cx<float>{w.x, w.y};
    v[1] = cx<float>{w.z, w.w};
  } else {
    v[0] = Pack<P>::unpack(w.x, sc);
    v[1] = Pack<P>::unpack(w.y, sc);
    v[2] = Pack<P>::unpack(w.z, sc);
    v[3] = Pack<P>::unpack(w.w, sc);
  }
}
template <class P>
__device__ __forceinline__ typename P::C chunk_pack(const cx<typename P::R>* v, float inv) {
  if constexpr (P::kCPC == 1) {
    return v[0];
  } else if constexpr (P::kCPC == 2) {
    return float4{v[0].re, v[0].im, v[1].re, v[1].im};
  } else {
    return uint4{Pack<P>::pack(v[0], inv), Pack<P>::pack(v[1], inv), Pack<P>::pack(v[2], inv),
                 Pack<P>::pack(v[3], inv)};
  }
}

template <class P>
__device__ __forceinline__ float ld_scale(const float* __restrict__ sc, int n) {
  return P::kScaled ? sc[n] : 1.0f;
}

// the 12 complex components of site n of one slice (d, sc: the slice's base).
// Chunks whose components the caller does not use are dead loads and vanish.
template <class P>
__device__ __forceinline__ void ld_spinor(const typename P::C* __restrict__ d,
                                          const float* __restrict__ sc, int vh, int n,
                                          cx<typename P::R>* v) {
  const float s = ld_scale<P>(sc, n);
#pragma unroll
  for (int k = 0; k < nch<P>(); ++k) chunk_unpack<P>(d[(size_t)k * vh + n], s, v + k * P::kCPC);
}

// the same without the per-site scale (16-bit policies: the raw integers or
// raw bf16; the caller applies the scale)
template <class P>
__device__ __forceinline__ void ld_spinor_raw(const typename P::C* __restrict__ d, int vh, int n,
                                              cx<typename P::R>* v) {
#pragma unroll
  for (int k = 0; k < nch<P>(); ++k) chunk_unpack<P>(d[(size_t)k * vh + n], 1.0f, v + k * P::kCPC);
}

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
// ld_spinor for data read once (psi of an epilogue, CG vectors): in FP64 and
// FP32 the loads are marked non-temporal so that they do not push the
// gather's neighbour spinors out of L2 (MatPcDagMatPc 3-5 % faster with
// st_stream below).  The 16-bit fields keep ordinary accesses: at 16^3x32
// they fit the last-level cache, which non-temporal accesses bypass (8-10 %
// slower there)
template <class P>
__device__ __forceinline__ void ld_spinor_once(const typename P::C* __restrict__ d,
                                               const float* __restrict__ sc, int vh, int n,
                                               cx<typename P::R>* v) {
  static_assert(sizeof(typename P::C) == 16, "16-byte chunks");
  if constexpr (P::kScaled) {
    ld_spinor<P>(d, sc, vh, n, v);
    return;
  }
  const float s = ld_scale<P>(sc, n);
#pragma unroll
  for (int k = 0; k < nch<P>(); ++k) {
    const u32x4_t w =
        __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(d + (size_t)k * vh + n));
    typename P::C c;
    __builtin_memcpy(&c, &w, 16);
    chunk_unpack<P>(c, s, v + k * P::kCPC);
  }
}

// one stored link element; non-temporal in FP64: a link is staged by two
// work-groups (as U_mu(x) and as U_mu(x - mu)), mostly far apart in time,
// and the spinor gather needs L2 more (FP64 MatPcDagMatPc 2 % faster; the
// FP32 links of a 16^3x32 lattice stay in the last-level cache between
// passes, and bypassing it made that operator 6 % slower)
template <class P>
__device__ __forceinline__ typename P::S ld_link_elem(const typename P::S* __restrict__ p) {
  typename P::S x;
  if constexpr (std::is_same<P, P64>::value) {
    static_assert(sizeof(x) == 16, "FP64 link elements are 16 bytes");
    const u32x4_t w = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
    __builtin_memcpy(&x, &w, 16);
  } else {
    x = *p;
  }
  return x;
}

// one 16-byte chunk of policy P stored non-temporally in FP64 / FP32 (see
// ld_spinor_once): the outputs of a pass are not re-read by it, so they
// should not evict its neighbours from L2
template <class P, class T>
__device__ __forceinline__ void st_stream(T* p, const T& v) {
  static_assert(sizeof(T) == 16, "16-byte chunks");
  if constexpr (P::kScaled) {
    *p = v;
    return;
  }
  u32x4_t w;
  __builtin_memcpy(&w, &v, 16);
  __builtin_nontemporal_store(w, reinterpret_cast<u32x4_t*>(p));
}

// store the 12 complex components of site n of a slice
template <class P>
__device__ __forceinline__ void st_site(typename P::C* __restrict__ d, float* __restrict__ sc,
                                        int vh, int n, const cx<typename P::R>* v) {
  float inv = 1.0f;
  if constexpr (P::kScaled) {
    float m = 0.0f;
#pragma unroll
    for (int c = 0; c < 12; ++c) m = fmaxf(m, fmaxf(fabsf((float)v[c].re), fabsf((float)v[c].im)));
    // the packers take 1 / m (bf16: value / scale; i16: times pack_norm,
    // so that the site maximum lands on 16383)
    inv = m > 0.0f ? pack_norm<P>() / m : 0.0f;
    sc[n] = scale_of<P>(m);
  }
#pragma unroll
  for (int k = 0; k < nch<P>(); ++k)
    st_stream<P>(d + (size_t)k * vh + n, chunk_pack<P>(v + k * P::kCPC, inv));
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
struct Geom {
  int lx, ly, lz, lt, lxh, vh, ls;
  int bt, bz;  // time / z slices per tile of the site index (divide lt / lz)
  int nxcd;    // XCDs the work-groups are spread over
  int bcmask;  // bit mu set: anti-periodic sign on the links leaving x_mu = L_mu - 1
  // neighbour table of the staged hop, nbt[(p * 8 + d) * vh + i] for output
  // site i of parity p: the site index of the neighbour x + d (d < 4) or
  // x - (d - 4), bit 31 set when the link between them carries the
  // anti-periodic sign (built once by k_build_nbt)
  const int* nbt;
  int rev;  // 1: each XCD sweeps its range of work-groups backwards (xcd_block)
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

// coordinates of site index i of parity `parity`
__host__ __device__ __forceinline__ void site_coords(const Geom& g, int i, int parity, int* c) {
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
  c[0] = 2 * xh + ((parity + y + z + t) & 1);
  c[1] = y;
  c[2] = z;
  c[3] = t;
}

struct Nbr {
  int fwd[4], bwd[4];
};

__host__ __device__ __forceinline__ Nbr neighbours(const Geom& g, int i, int parity) {
  int c[4];
  site_coords(g, i, parity, c);
  const int x = c[0], y = c[1], z = c[2], t = c[3], xh = x >> 1;
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
__device__ __forceinline__ int xcd_block(int nxcd, int rev = 0) {
  const int b = blockIdx.x, nb = gridDim.x;
  const int q = nb / nxcd, r = nb - q * nxcd;
  const int k = b % nxcd, j = b / nxcd;
  const int cnt = q + (k < r ? 1 : 0);  // work-groups of XCD k; j < cnt
  return k * q + (k < r ? k : r) + (rev ? cnt - 1 - j : j);
}

// ---------------------------------------------------------------------------
// Wilson hopping term
// ---------------------------------------------------------------------------
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
__device__ __forceinline__ void reconstruct(cx<R>* a, const cx<R>* r, R w) {
  // w: per-site scale of a 16-bit input (1 otherwise); each w * r term
  // contracts with its accumulation into one fused multiply-add
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const cx<R> ru = w * r[c], rl = w * r[3 + c];
    a[c] = a[c] + ru;
    a[3 + c] = a[3 + c] + rl;
    if (MU == 0) {
      a[6 + c] = a[6 + c] - sgi<SG>(rl);
      a[9 + c] = a[9 + c] - sgi<SG>(ru);
    } else if (MU == 1) {
      a[6 + c] = a[6 + c] + sgr<SG>(rl);
      a[9 + c] = a[9 + c] - sgr<SG>(ru);
    } else if (MU == 2) {
      a[6 + c] = a[6 + c] - sgi<SG>(ru);
      a[9 + c] = a[9 + c] + sgi<SG>(rl);
    } else {
      a[6 + c] = a[6 + c] + sgr<SG>(ru);
      a[9 + c] = a[9 + c] + sgr<SG>(rl);
    }
  }
}

// (1 + SG gamma_mu) on int16 components packed (re | im << 16), two lanes
// per instruction.  Stored int16 magnitudes are at most 16383, so sums and
// differences stay inside int16.
typedef short i16x2 __attribute__((ext_vector_type(2)));
template <int SG>
__device__ __forceinline__ i16x2 pk_addi(i16x2 a, i16x2 b) {  // a + SG i b
  // (re, im) of i b = (-b.im, b.re): b with its halves swapped (op_sel)
  // times (-1, 1), plus a, in one packed multiply-add (the low 16 bits of
  // a product do not depend on signedness)
  const i16x2 k = SG > 0 ? i16x2{-1, 1} : i16x2{1, -1};
  i16x2 r;
  asm("v_pk_mad_u16 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[0,1,1]"
      : "=v"(r)
      : "v"(b), "v"(k), "v"(a));
  return r;
}
template <int SG>
__device__ __forceinline__ i16x2 pk_addr(i16x2 a, i16x2 b) {  // a + SG b
  return SG > 0 ? a + b : a - b;
}
// the six projected components, still packed (re | im << 16)
template <int MU, int SG>
__device__ __forceinline__ void project_words(const uint32_t* w, i16x2* q) {
  i16x2 p[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) p[c] = __builtin_bit_cast(i16x2, w[c]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (MU == 0) {
      q[c] = pk_addi<SG>(p[c], p[9 + c]);
      q[3 + c] = pk_addi<SG>(p[3 + c], p[6 + c]);
    } else if (MU == 1) {
      q[c] = pk_addr<-SG>(p[c], p[9 + c]);
      q[3 + c] = pk_addr<SG>(p[3 + c], p[6 + c]);
    } else if (MU == 2) {
      q[c] = pk_addi<SG>(p[c], p[6 + c]);
      q[3 + c] = pk_addi<-SG>(p[3 + c], p[9 + c]);
    } else {
      q[c] = pk_addr<SG>(p[c], p[6 + c]);
      q[3 + c] = pk_addr<SG>(p[3 + c], p[9 + c]);
    }
  }
}
template <int MU, int SG>
__device__ __forceinline__ void project_pk(const uint32_t* w, cx<float>* h) {
  i16x2 q[6];
  project_words<MU, SG>(w, q);
#pragma unroll
  for (int c = 0; c < 6; ++c) h[c] = cx<float>{(float)q[c].x, (float)q[c].y};
}

// the 12 packed words of an int16 spinor (complex c in word c)
__device__ __forceinline__ void ld_words_i16(const uint4* __restrict__ d, int vh, int n,
                                             uint32_t* w) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint4 u = d[(size_t)k * vh + n];
    w[4 * k] = u.x;
    w[4 * k + 1] = u.y;
    w[4 * k + 2] = u.z;
    w[4 * k + 3] = u.w;
  }
}

// A loader gives the hop the projected two-spinor (1 + SG gamma_mu) psi of
// the input at site n; for 16-bit policies it leaves out the per-site scale
// and returns it, so the caller folds it into the accumulation of the hop's
// product (no separate multiplications).
template <class P>
struct PlainLoader {
  typedef typename P::R R;
  const typename P::C* __restrict__ d;
  const float* __restrict__ sc;
  int vh;
  // the 12 raw components (scale returned, not applied)
  __device__ __forceinline__ float load(int n, cx<R>* p) const {
#pragma unroll
    for (int k = 0; k < nch<P>(); ++k)
      chunk_unpack<P>(d[(size_t)k * vh + n], 1.0f, p + k * P::kCPC);
    return ld_scale<P>(sc, n);
  }
  // int16 only: the projected two-spinor as packed int16 words (scale returned)
  template <int MU, int SG>
  __device__ __forceinline__ float proj_words(int n, i16x2* q) const {
    static_assert(std::is_same<P, PI16>::value, "packed words are the int16 format");
    uint32_t w[12];
    ld_words_i16(d, vh, n, w);
    project_words<MU, SG>(w, q);
    return ld_scale<P>(sc, n);
  }
  template <int MU, int SG>
  __device__ __forceinline__ float proj(int n, cx<R>* h) const {
    if constexpr (std::is_same<P, PI16>::value) {
      // project the raw integers, then convert 6 complex instead of 12
      uint32_t w[12];
      ld_words_i16(d, vh, n, w);
      project_pk<MU, SG>(w, h);
    } else {
      cx<R> p[12];
#pragma unroll
      for (int k = 0; k < nch<P>(); ++k)
        chunk_unpack<P>(d[(size_t)k * vh + n], 1.0f, p + k * P::kCPC);
      project<MU, SG>(p, h);
    }
    return ld_scale<P>(sc, n);
  }
};

// on-the-fly (b + c D5) of zmobius_dslash_4_dag0 (zmobius_dslash_4.C:90-105):
// t(s) = b in(s) + c P+ in(s-1) + c P- in(s+1), walls carry -m c.
template <class P>
struct MobTLoader {
  typedef typename P::R R;
  const typename P::C* __restrict__ cur;
  const typename P::C* __restrict__ prv;
  const typename P::C* __restrict__ nxt;
  const float* __restrict__ sc_cur;
  const float* __restrict__ sc_prv;
  const float* __restrict__ sc_nxt;
  int vh;
  R b, cm, cp;
  template <int MU, int SG>
  __device__ __forceinline__ float proj(int n, cx<R>* h) const {
    cx<R> vm[12], vp[12], q[12];
    // raw values: the three per-site scales go into the coefficients
    // (3 products per site instead of one per component)
    if constexpr (std::is_same<P, PI16>::value) {
      // b proj(cur) + proj(c P+ prv + c P- nxt): cur projected in int16
      // (half the conversions); the second term takes the upper spins of prv
      // and the lower spins of nxt, one product each
      uint32_t w[12];
      cx<R> hq[6];
      ld_words_i16(cur, vh, n, w);
      project_pk<MU, SG>(w, h);
      ld_spinor_raw<P>(prv, vh, n, vm);
      ld_spinor_raw<P>(nxt, vh, n, vp);
      const R bs = b * sc_cur[n], cms = cm * sc_prv[n], cps = cp * sc_nxt[n];
#pragma unroll
      for (int c = 0; c < 6; ++c) q[c] = cms * vm[c];
#pragma unroll
      for (int c = 6; c < 12; ++c) q[c] = cps * vp[c];
      project<MU, SG>(q, hq);
#pragma unroll
      for (int c = 0; c < 6; ++c) h[c] = bs * h[c] + hq[c];
    } else {
      // t first, then its projection (the FP32 / FP64 kernels keep their
      // register budget this way)
      ld_spinor_raw<P>(cur, vh, n, q);
      ld_spinor_raw<P>(prv, vh, n, vm);
      ld_spinor_raw<P>(nxt, vh, n, vp);
      R bs = b, cms = cm, cps = cp;
      if constexpr (P::kScaled) {
        bs = b * sc_cur[n];
        cms = cm * sc_prv[n];
        cps = cp * sc_nxt[n];
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) q[c] = bs * q[c] + cms * vm[c];
#pragma unroll
      for (int c = 6; c < 12; ++c) q[c] = bs * q[c] + cps * vp[c];
      project<MU, SG>(q, h);
    }
    return 1.0f;
  }
};

// For the site with coordinates c: the index of its neighbour x + MU (fwd)
// or x - MU (bwd) on the other parity, whether the link U_MU(x) (fwd) or
// U_MU(x - MU) (bwd) carries the anti-periodic boundary sign, and the index of
// that link's site (x itself, or the backward neighbour).
template <int MU>
__device__ __forceinline__ void link_site(const Geom& g, const int* c, int i, bool bwd,
                                          int* site, bool* neg, int* nbr) {
  const int L = MU == 0 ? g.lx : MU == 1 ? g.ly : MU == 2 ? g.lz : g.lt;
  const int x = c[MU];
  const bool bc = (g.bcmask >> MU) & 1;
  *neg = bc && (bwd ? x == 0 : x == L - 1);
  const int m = bwd ? (x == 0 ? L - 1 : x - 1) : (x == L - 1 ? 0 : x + 1);
  *nbr = site_index(g, (MU == 0 ? m : c[0]) >> 1, MU == 1 ? m : c[1], MU == 2 ? m : c[2],
                    MU == 3 ? m : c[3]);
  *site = bwd ? *nbr : i;
}

// Links of the work-group's sites, staged once in LDS (compute type) and
// shared by all Ls slices: lu[((set*8 + d)*ns + site)*9 + e], d < 4 the
// forward link U_d(x), d >= 4 the backward link U_{d-4}(x - d) of the output
// site x; `set` distinguishes the two output parities of a DWF work-group.
// lnb[(set*8 + d)*ns + site] receives the index of the neighbour x + d
// (d < 4) or x - (d-4) that the hop reads, so no slice recomputes it.
// one entry of the neighbour table (layout: Geom::nbt)
__global__ void __launch_bounds__(256) k_build_nbt(int* __restrict__ nbt, Geom g) {
  const size_t n = (size_t)2 * 8 * g.vh;
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < n;
       w += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(w % g.vh), d = (int)((w / g.vh) % 8), p = (int)(w / (8 * (size_t)g.vh));
    int c[4];
    site_coords(g, i, p, c);
    int site = i, nbr = i;
    bool neg = false;
    switch (d & 3) {
      case 0: link_site<0>(g, c, i, d >= 4, &site, &neg, &nbr); break;
      case 1: link_site<1>(g, c, i, d >= 4, &site, &neg, &nbr); break;
      case 2: link_site<2>(g, c, i, d >= 4, &site, &neg, &nbr); break;
      default: link_site<3>(g, c, i, d >= 4, &site, &neg, &nbr); break;
    }
    nbt[w] = (int)((unsigned)nbr | (neg ? 0x80000000u : 0u));
  }
}

template <class P>
__device__ __forceinline__ void stage_links(cx<typename P::R>* __restrict__ lu,
                                            int* __restrict__ lnb,
                                            const typename P::S* __restrict__ U, const Geom& g,
                                            int blk, int ns, int nsets, int p_first,
                                            int only_dir = -1, int dst_base = 0) {
  typedef typename P::R R;
  constexpr int NE = P::kLinkElems;
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nthr = blockDim.x * blockDim.y * blockDim.z;
  for (int w = tid; w < nsets * 8 * ns; w += nthr) {
    const int il = w % ns, d = (w / ns) % 8, set = w / (8 * ns);
    const int i = blk * ns + il;
    if (i >= g.vh || (only_dir >= 0 && d != only_dir)) continue;
    const int p = (p_first + set) & 1;
    const int mu = d & 3;
    const int e = g.nbt[(size_t)(p * 8 + d) * g.vh + i];
    const int nbr = e & 0x7fffffff;
    const bool neg = e < 0;
    const int site = d >= 4 ? nbr : i;
    lnb[w - dst_base] = nbr;
    const int par = d >= 4 ? 1 - p : p;  // U_mu(x - mu) lives on the other parity
    const typename P::S* src = U + (size_t)((mu * 2 + par) * NE) * g.vh + site;
    cx<R>* dst = lu + (size_t)(w - dst_base) * 9;
    const float ls = link_scale<P>();
#pragma unroll
    for (int e = 0; e < NE; ++e) dst[e] = Pack<P>::unpack(ld_link_elem<P>(src + (size_t)e * g.vh), ls);
    if (NE == 6) {
      // row 2 = conj(row0 x row1) (det U = 1), times the boundary sign
      const cx<R> a0 = dst[0], a1 = dst[1], a2 = dst[2], b0 = dst[3], b1 = dst[4], b2 = dst[5];
      cx<R> r0 = cmul(a1, b2) - cmul(a2, b1);
      cx<R> r1 = cmul(a2, b0) - cmul(a0, b2);
      cx<R> r2 = cmul(a0, b1) - cmul(a1, b0);
      const R sg = neg ? (R)-1 : (R)1;
      dst[6] = cx<R>{sg * r0.re, -sg * r0.im};
      dst[7] = cx<R>{sg * r1.re, -sg * r1.im};
      dst[8] = cx<R>{sg * r2.re, -sg * r2.im};
    }
  }
}

// lf / lb: this site's forward / backward link of direction MU (9 elements)
template <int MU, int DAG, typename R, class LD>
__device__ __forceinline__ void hop_dir(cx<R>* acc, const LD& ld, const cx<R>* lf,
                                        const cx<R>* lb, const Nbr& nb) {
  constexpr int SF = DAG ? 1 : -1;  // forward: 1 - s gamma, s = +1 (dag 0) / -1 (dag 1)
  cx<R> h[6], r[6];
  {
    const R sc = (R)ld.template proj<MU, SF>(nb.fwd[MU], h);
    cx<R> m[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m[e] = lf[e];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int row = 0; row < 3; ++row)
        r[3 * k + row] = row_dot(m[row * 3 + 0], m[row * 3 + 1], m[row * 3 + 2], h[3 * k + 0],
                                 h[3 * k + 1], h[3 * k + 2]);
    reconstruct<MU, SF>(acc, r, sc);
  }
  {
    const R sc = (R)ld.template proj<MU, -SF>(nb.bwd[MU], h);
    cx<R> m[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m[e] = lb[e];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int row = 0; row < 3; ++row)
        r[3 * k + row] = row_dotc(m[0 * 3 + row], m[1 * 3 + row], m[2 * 3 + row], h[3 * k + 0],
                                  h[3 * k + 1], h[3 * k + 2]);
    reconstruct<MU, -SF>(acc, r, sc);
  }
}

// acc += the reconstructed (1 -/+ gamma_MU) U h of one projected neighbour:
// U h (FWD) or U^dag h, scaled by sc
template <int MU, int SG, bool FWD, typename R>
__device__ __forceinline__ void hop_half(cx<R>* acc, const cx<R>* h, const cx<R>* l, R sc) {
  cx<R> r[6];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int row = 0; row < 3; ++row)
      r[3 * k + row] = FWD ? row_dot(l[row * 3 + 0], l[row * 3 + 1], l[row * 3 + 2], h[3 * k + 0],
                                     h[3 * k + 1], h[3 * k + 2])
                           : row_dotc(l[0 * 3 + row], l[1 * 3 + row], l[2 * 3 + row],
                                      h[3 * k + 0], h[3 * k + 1], h[3 * k + 2]);
  reconstruct<MU, SG>(acc, r, sc);
}

// keeps the accumulated spinor computed where it stands: without it the
// compiler sinks the arithmetic of the exchange gather below into the
// (act) store after its last barrier and keeps every loaded operand alive
// until then
template <class R>
__device__ __forceinline__ void pin12(cx<R>* v) {
#pragma unroll
  for (int c = 0; c < 12; ++c) asm volatile("" : "+v"(v[c].re), "+v"(v[c].im));
}

// One direction of hop((b + c D5) in) at slice s, dag 0, reading every
// neighbour spinor once.  (b + c D5) in (s) = b in(s) + c P+ in(s-1) +
// c P- in(s+1) (zmobius_dslash_4.C:90-105), P+ in being the upper two spins
// in the chiral basis.  Each thread loads slice s of the neighbour; the upper
// half of slice s-1 and the lower half of slice s+1 come from the threads of
// those slices through the exchange slots `mine` / `prv` / `nxt` (6 complex
// a thread).  Four barriers: every thread of the work-group must call it.
template <int MU, int SG, bool FWD, typename R, class LD>
__device__ __forceinline__ void hop_dir_tx(cx<R>* acc, LD ld, int n, const cx<R>* l,
                                           cx<R>* mine, const cx<R>* prv, const cx<R>* nxt,
                                           R b, R cm, R cp) {
  cx<R> p[12], q[6], h[6];
  const R sc = (R)ld.load(n, p);
#pragma unroll
  for (int c = 0; c < 12; ++c) p[c] = sc * p[c];
  __syncthreads();  // the previous direction's readers are done with the slots
#pragma unroll
  for (int c = 0; c < 6; ++c) mine[c] = p[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 6; ++c) q[c] = prv[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 6; ++c) mine[c] = p[6 + c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 6; ++c) p[c] = b * p[c] + cm * q[c];
#pragma unroll
  for (int c = 0; c < 6; ++c) p[6 + c] = b * p[6 + c] + cp * nxt[c];
  project<MU, SG>(p, h);
  hop_half<MU, SG, FWD>(acc, h, l, (R)1);
  pin12(acc);
}

// acc = hop((b + c D5) in) at slice s (dag 0) through hop_dir_tx; inactive
// lanes read site 0 and take the barriers with the others
template <typename R, class LD>
__device__ __forceinline__ void hop4_tx(cx<R>* acc, LD ld, const cx<R>* lu, int ds,
                                        const int* nbi, int ns, bool act, cx<R>* mine,
                                        const cx<R>* prv, const cx<R>* nxt, R b, R cm, R cp) {
  Nbr nb;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    nb.fwd[mu] = act ? nbi[mu * ns] : 0;
    nb.bwd[mu] = act ? nbi[(4 + mu) * ns] : 0;
  }
  hop_dir_tx<0, -1, true>(acc, ld, nb.fwd[0], lu + 0 * ds, mine, prv, nxt, b, cm, cp);
  hop_dir_tx<0, 1, false>(acc, ld, nb.bwd[0], lu + 4 * ds, mine, prv, nxt, b, cm, cp);
  hop_dir_tx<1, -1, true>(acc, ld, nb.fwd[1], lu + 1 * ds, mine, prv, nxt, b, cm, cp);
  hop_dir_tx<1, 1, false>(acc, ld, nb.bwd[1], lu + 5 * ds, mine, prv, nxt, b, cm, cp);
  hop_dir_tx<2, -1, true>(acc, ld, nb.fwd[2], lu + 2 * ds, mine, prv, nxt, b, cm, cp);
  hop_dir_tx<2, 1, false>(acc, ld, nb.bwd[2], lu + 6 * ds, mine, prv, nxt, b, cm, cp);
  hop_dir_tx<3, -1, true>(acc, ld, nb.fwd[3], lu + 3 * ds, mine, prv, nxt, b, cm, cp);
  hop_dir_tx<3, 1, false>(acc, ld, nb.bwd[3], lu + 7 * ds, mine, prv, nxt, b, cm, cp);
}

// ---------------------------------------------------------------------------
// int16 hop on integer dot products
// ---------------------------------------------------------------------------
// The int16 kernels whose input needs no (b + c D5) on the fly multiply the
// projected two-spinor by the link in exact integer arithmetic: each complex
// product is two v_dot2_i32_i16 (re and im of U h from packed int16 words)
// instead of four FP32 multiply-adds, and the six products of a row sum in
// int32.  Links are staged as packed words p = (Ur, Ui) and q = (Ur, -Ui) of U
// scaled by 16383.5 (value = word * kLinkPQScale), so every partial sum of a
// row times a projected colour vector h (|h| <= sqrt(6) * 32766, int16 spinors
// stored in [-16383, 16383]) is below 16384 * 80260 < 1.32e9 < 2^31: no
// intermediate can overflow, the sums are exact, and one conversion per
// result component takes them to FP32.
struct LinkPQ {
  uint32_t p, q;
};
constexpr float kLinkPQScale = 2.0f / 32767.0f;

template <class P>
__device__ __forceinline__ LinkPQ link_pq(cx<float> u) {
  // cvt_pknorm_i16 rounds clamp(x, -1, 1) * 32767; x = u / 2
  const i16x2 pk = __builtin_amdgcn_cvt_pknorm_i16(0.5f * u.re, 0.5f * u.im);
  const i16x2 qk = pk * i16x2{1, -1};
  return LinkPQ{__builtin_bit_cast(uint32_t, pk), __builtin_bit_cast(uint32_t, qk)};
}

// stage_links for the integer hop: the same links and neighbour indices, the
// elements as LinkPQ (8 bytes each, the footprint of the FP32 staging)
__device__ __forceinline__ void stage_links_pq(LinkPQ* __restrict__ lu, int* __restrict__ lnb,
                                               const uint32_t* __restrict__ U, const Geom& g,
                                               int blk, int ns, int nsets, int p_first) {
  constexpr int NE = PI16::kLinkElems;
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nthr = blockDim.x * blockDim.y * blockDim.z;
  for (int w = tid; w < nsets * 8 * ns; w += nthr) {
    const int il = w % ns, d = (w / ns) % 8, set = w / (8 * ns);
    const int i = blk * ns + il;
    if (i >= g.vh) continue;
    const int p = (p_first + set) & 1;
    const int mu = d & 3;
    const int e = g.nbt[(size_t)(p * 8 + d) * g.vh + i];
    const int nbr = e & 0x7fffffff;
    const bool neg = e < 0;
    const int site = d >= 4 ? nbr : i;
    lnb[w] = nbr;
    const int par = d >= 4 ? 1 - p : p;
    const uint32_t* src = U + (size_t)((mu * 2 + par) * NE) * g.vh + site;
    cx<float> u[9];
#pragma unroll
    for (int k = 0; k < NE; ++k)
      u[k] = Pack<PI16>::unpack(src[(size_t)k * g.vh], link_scale<PI16>());
    // row 2 = conj(row0 x row1) times the boundary sign
    const cx<float> r0 = cmul(u[1], u[5]) - cmul(u[2], u[4]);
    const cx<float> r1 = cmul(u[2], u[3]) - cmul(u[0], u[5]);
    const cx<float> r2 = cmul(u[0], u[4]) - cmul(u[1], u[3]);
    const float sg = neg ? -1.0f : 1.0f;
    u[6] = cx<float>{sg * r0.re, -sg * r0.im};
    u[7] = cx<float>{sg * r1.re, -sg * r1.im};
    u[8] = cx<float>{sg * r2.re, -sg * r2.im};
    LinkPQ* dst = lu + (size_t)w * 9;
#pragma unroll
    for (int k = 0; k < 9; ++k) dst[k] = link_pq<PI16>(u[k]);
  }
}

__device__ __forceinline__ i16x2 swap16(i16x2 a) { return __builtin_shufflevector(a, a, 1, 0); }
// c + a.lo * b.lo + a.hi * b.hi in int32.  The sums here never leave the
// int32 range (bound above), so the clamped form gives the same bits; asking
// for it keeps the three-operand encoding, which takes an inline zero as the
// first accumulator and the half swap of b as an operand selector, where the
// accumulate-in-place form needs a move and a shift per chain.
__device__ __forceinline__ int dot2(i16x2 a, i16x2 b, int c) {
  return __builtin_amdgcn_sdot2(a, b, c, true);
}

// one direction of the hop with integer products; lf / lb as in hop_dir
template <int MU, int DAG>
__device__ __forceinline__ void hop_dir_pq(cx<float>* acc, const PlainLoader<PI16>& ld,
                                           const LinkPQ* lf, const LinkPQ* lb, const Nbr& nb) {
  constexpr int SF = DAG ? 1 : -1;
  i16x2 h[6];
  cx<float> r[6];
  {
    // (U h)_row: re = sum_j q_rj . h_j, im = sum_j p_rj . swap(h_j)
    const float sc = ld.template proj_words<MU, SF>(nb.fwd[MU], h);
    LinkPQ m[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m[e] = lf[e];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int row = 0; row < 3; ++row) {
        int re = 0, im = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const i16x2 hv = h[3 * k + j];
          re = dot2(__builtin_bit_cast(i16x2, m[row * 3 + j].q), hv, re);
          im = dot2(__builtin_bit_cast(i16x2, m[row * 3 + j].p), swap16(hv), im);
        }
        r[3 * k + row] = cx<float>{(float)re, (float)im};
      }
    reconstruct<MU, SF>(acc, r, sc * kLinkPQScale);
  }
  {
    // (U^dag h)_row: re = sum_j p_j,row . h_j, im = sum_j q_j,row . swap(h_j)
    const float sc = ld.template proj_words<MU, -SF>(nb.bwd[MU], h);
    LinkPQ m[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m[e] = lb[e];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int row = 0; row < 3; ++row) {
        int re = 0, im = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const i16x2 hv = h[3 * k + j];
          re = dot2(__builtin_bit_cast(i16x2, m[j * 3 + row].p), hv, re);
          im = dot2(__builtin_bit_cast(i16x2, m[j * 3 + row].q), swap16(hv), im);
        }
        r[3 * k + row] = cx<float>{(float)re, (float)im};
      }
    reconstruct<MU, -SF>(acc, r, sc * kLinkPQScale);
  }
}

template <int DAG>
__device__ __forceinline__ void hop4_pq(cx<float>* acc, const PlainLoader<PI16>& ld,
                                        const LinkPQ* lu, int ds, const int* nbi, int ns) {
  Nbr nb;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    nb.fwd[mu] = nbi[mu * ns];
    nb.bwd[mu] = nbi[(4 + mu) * ns];
  }
  hop_dir_pq<0, DAG>(acc, ld, lu + 0 * ds, lu + 4 * ds, nb);
  hop_dir_pq<1, DAG>(acc, ld, lu + 1 * ds, lu + 5 * ds, nb);
  hop_dir_pq<2, DAG>(acc, ld, lu + 2 * ds, lu + 6 * ds, nb);
  hop_dir_pq<3, DAG>(acc, ld, lu + 3 * ds, lu + 7 * ds, nb);
}

// acc = D_hop^(dag) psi at output site i of parity p_out (wilson_dslash.C:173-517);
// lu points at this site's 8 staged links, consecutive directions `ds` apart,
// nbi at its 8 staged neighbour indices, `ns` apart.
template <int DAG, typename R, class LD>
__device__ __forceinline__ void hop4(cx<R>* acc, const LD& ld, const cx<R>* lu, int ds,
                                     const int* nbi, int ns) {
  Nbr nb;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    nb.fwd[mu] = nbi[mu * ns];
    nb.bwd[mu] = nbi[(4 + mu) * ns];
  }
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

template <typename R>
__device__ __forceinline__ double norm12(const cx<R>* v) {
  double n = 0.0;
#pragma unroll
  for (int c = 0; c < 12; ++c) n += (double)v[c].re * v[c].re + (double)v[c].im * v[c].im;
  return n;
}

// ---------------------------------------------------------------------------
// CG update fused into the last kernel of MatPcDagMatPc (inv_cg.C:340-349):
// with alpha = scal[0] / scal[1] (|r|^2 and <p, A p> on the device; alpha = 0
// when <p, A p> == 0 so that the host can report the breakdown on unchanged
// vectors), r -= alpha A p and x += alpha p at the thread's (site, slice),
// returning |r|^2 of the site for the partial sums.  A p itself is not stored.
// ---------------------------------------------------------------------------
template <class P, class PX>
struct CgEpi {
  Fld<P> r;
  CFld<P> p;
  Fld<PX> x;
  const double* scal;
};

// PX = NoX: the epilogue updates r only; x += alpha p is left to the pass
// that forms the next direction (k_cg_xp), which reads p anyway.  This keeps
// the x traffic and its arithmetic out of the hop kernel.
struct NoX {
  typedef float R;
  typedef float S;
  typedef cx<float> C;
  static constexpr bool kScaled = false;
  static constexpr int kCPC = 1;
  static constexpr int kLinkElems = 9;
};

// alpha = rr / <p, A p> from the device scalars.  The kernels read it at
// their start: loaded and divided next to the epilogue, the compiler issued
// every neighbour load of the hop ahead of it and spilled them to scratch.
__device__ __forceinline__ double cg_alpha(const double* scal) {
  const double rr = scal[0], d = scal[1];
  return d != 0.0 ? rr / d : 0.0;
}

template <class P, class PX>
__device__ __forceinline__ double cg_epilogue(const CgEpi<P, PX>& cg, const Geom& g, int s,
                                              int i, const cx<typename P::R>* ap, double a) {
  typedef typename P::R R;
  typedef typename PX::R RX;
  const size_t sl = (size_t)s * nch<P>() * g.vh, slx = (size_t)s * nch<PX>() * g.vh;
  const size_t ss = (size_t)s * g.vh;
  cx<R> r[12];
  ld_spinor_once<P>(cg.r.d + sl, cg.r.sc + ss, g.vh, i, r);
  if constexpr (!std::is_same<PX, NoX>::value) {
    cx<R> p[12];
    cx<RX> x[12];
    ld_spinor_once<P>(cg.p.d + sl, cg.p.sc + ss, g.vh, i, p);
    ld_spinor_once<PX>(cg.x.d + slx, cg.x.sc + ss, g.vh, i, x);
#pragma unroll
    for (int c = 0; c < 12; ++c)
      x[c] = cx<RX>{(RX)(a * p[c].re) + x[c].re, (RX)(a * p[c].im) + x[c].im};
    st_site<PX>(cg.x.d + slx, cg.x.sc + ss, g.vh, i, x);
  }
#pragma unroll
  for (int c = 0; c < 12; ++c)
    r[c] = cx<R>{(R)(-a * ap[c].re) + r[c].re, (R)(-a * ap[c].im) + r[c].im};
  st_site<P>(cg.r.d + sl, cg.r.sc + ss, g.vh, i, r);
  return norm12(r);
}

// ---------------------------------------------------------------------------
// DWF: dwf_dslash = D4 on every slice with 4D parity (s+cb)%2 + 5th-dim hop
// (dwf_dslash_4.C:107-137, dwf_dslash_5_plus_omp.C:115-270).
// EPI 0: out = D in.   EPI 1: out = x0 + alpha D in.  NORM: |out|^2 partials.
// CG (EPI 1): out is not written; the result feeds cg_epilogue instead.
// ---------------------------------------------------------------------------
template <class P, int DAG, int EPI, bool NORM, bool CG = false, class PX = P>
__global__ void __launch_bounds__(512)
k_dwf_dslash(Fld<P> out, CFld<P> in, const typename P::S* __restrict__ U, CFld<P> x0,
             double* __restrict__ partial, Geom g, int cb_in, double two_a5, double nm_two_a5,
             double alpha, CgEpi<P, PX> cg) {
  typedef typename P::R R;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double sh[16];
  const int ns = blockDim.x, il = threadIdx.x, s = threadIdx.y;
  int* lnb = reinterpret_cast<int*>(smem_raw);                   // [2][8][ns]
  cx<R>* lu = reinterpret_cast<cx<R>*>(smem_raw + 2 * 8 * ns * 4);  // [2][8][ns][9]
  const double a_cg = CG ? cg_alpha(cg.scal) : 0.0;
  const int blk = xcd_block(g.nxcd, g.rev);
  const int i = blk * ns + il;
  // slice s holds output parity (s + 1 + cb_in) % 2: set s % 2 of the stage
  constexpr bool kPQ = std::is_same<P, PI16>::value;  // integer hop
  if constexpr (kPQ)
    stage_links_pq(reinterpret_cast<LinkPQ*>(lu), lnb, U, g, blk, ns, 2, (1 + cb_in) & 1);
  else
    stage_links<P>(lu, lnb, U, g, blk, ns, 2, (1 + cb_in) & 1);
  __syncthreads();
  double nrm = 0.0;
  if (i < g.vh) {
    const size_t slice = (size_t)nch<P>() * g.vh;
    cx<R> acc[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) acc[c] = cx<R>{0, 0};
    const PlainLoader<P> ld{in.d + s * slice, in.sc + (size_t)s * g.vh, g.vh};
    if constexpr (kPQ)
      hop4_pq<DAG>(acc, ld, reinterpret_cast<const LinkPQ*>(lu) + ((s & 1) * 8 * ns + il) * 9,
                   ns * 9, lnb + (s & 1) * 8 * ns + il, ns);
    else
      hop4<DAG>(acc, ld, lu + ((s & 1) * 8 * ns + il) * 9, ns * 9, lnb + (s & 1) * 8 * ns + il,
                ns);
    // fifth dimension: the same site index i in slices s-1 / s+1
    const int sm = s > 0 ? s - 1 : g.ls - 1, spl = s < g.ls - 1 ? s + 1 : 0;
    const R cm = (R)(s > 0 ? two_a5 : nm_two_a5);
    const R cp = (R)(s < g.ls - 1 ? two_a5 : nm_two_a5);
    const int lo_m = DAG ? 6 : 0;  // components fed from s-1
    const int lo_p = DAG ? 0 : 6;  // components fed from s+1
    {
      cx<R> vm[12], vp[12];
      ld_spinor<P>(in.d + sm * slice, in.sc + (size_t)sm * g.vh, g.vh, i, vm);
      ld_spinor<P>(in.d + spl * slice, in.sc + (size_t)spl * g.vh, g.vh, i, vp);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        acc[lo_m + c] = acc[lo_m + c] + cm * vm[lo_m + c];
        acc[lo_p + c] = acc[lo_p + c] + cp * vp[lo_p + c];
      }
    }
    if (EPI == 1) {
      cx<R> xv[12];
      ld_spinor<P>(x0.d + s * slice, x0.sc + (size_t)s * g.vh, g.vh, i, xv);
      const R a = (R)alpha;
#pragma unroll
      for (int c = 0; c < 12; ++c) acc[c] = xv[c] + a * acc[c];
      if (NORM) nrm = norm12(acc);
      if (CG) nrm = cg_epilogue<P, PX>(cg, g, s, i, acc, a_cg);
    }
    if (!CG) st_site<P>(out.d + s * slice, out.sc + (size_t)s * g.vh, g.vh, i, acc);
  }
  if (NORM || CG) {
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
// Coefficients are kept in FP64 and FP32 so that neither compute type
// converts them per use.
struct M5Side {
  double af[MAXLS];   // fwd-corner sweep: v[s] += af[s] v[s-1]
  double cf[MAXLS];   // fwd-corner correction: v[s] += cf[s] v[ls-1]
  double ab[MAXLS];   // bwd-corner sweep: v[s] += ab[s] v[s+1]
  double cbk[MAXLS];  // bwd-corner gather: v[ls-1] += cbk[s] v[s]
  double inv_f, inv_b;
  float faf[MAXLS], fcf[MAXLS], fab[MAXLS], fcbk[MAXLS];
  float finv_f, finv_b;
};

struct MobArgs {
  double b, c, mc;  // Mobius b, c and -mass*c (wall coefficient)
  double K, mK;     // kappa_b/kappa_c and -mass*K
  double kappa_b;
  float fb, fc, fmc, fK, fmK, fkappa_b;
  M5Side m5;
  // batched solves: the active flag of this launch's right-hand side
  // (nullptr: always run); a paused one's work-groups leave at once
  const int* active;
};

// the coefficient in the compute type R
template <typename R>
__device__ __forceinline__ R pick(double d, float f) {
  if constexpr (sizeof(R) == 4)
    return f;
  else
    return d;
}

template <typename R>
__device__ __forceinline__ void m5_fwd_corner(cx<R>* v, int ls, const M5Side& k) {
  for (int s = 1; s < ls; ++s)
    v[s * 12] = v[s * 12] + pick<R>(k.af[s], k.faf[s]) * v[(s - 1) * 12];
  const cx<R> zl = v[(ls - 1) * 12];
  for (int s = 0; s < ls - 1; ++s) v[s * 12] = v[s * 12] + pick<R>(k.cf[s], k.fcf[s]) * zl;
  v[(ls - 1) * 12] = pick<R>(k.inv_f, k.finv_f) * zl;
}

template <typename R>
__device__ __forceinline__ void m5_bwd_corner(cx<R>* v, int ls, const M5Side& k) {
  cx<R> last = pick<R>(k.inv_b, k.finv_b) * v[(ls - 1) * 12];
  for (int s = 0; s < ls - 1; ++s) last = last + pick<R>(k.cbk[s], k.fcbk[s]) * v[s * 12];
  v[(ls - 1) * 12] = last;
  for (int s = ls - 2; s >= 0; --s)
    v[s * 12] = v[s * 12] + pick<R>(k.ab[s], k.fab[s]) * v[(s + 1) * 12];
}

// the same on a column held in registers (LS known at compile time)
template <int LS, int DAG, typename R>
__device__ __forceinline__ void m5inv_column_reg(cx<R>* v, bool upper, const M5Side& k) {
  cx<R> r[LS];
#pragma unroll
  for (int s = 0; s < LS; ++s) r[s] = v[s * 12];
  if (upper == (DAG == 0)) {
#pragma unroll
    for (int s = 1; s < LS; ++s) r[s] = r[s] + pick<R>(k.af[s], k.faf[s]) * r[s - 1];
    const cx<R> zl = r[LS - 1];
#pragma unroll
    for (int s = 0; s < LS - 1; ++s) r[s] = r[s] + pick<R>(k.cf[s], k.fcf[s]) * zl;
    r[LS - 1] = pick<R>(k.inv_f, k.finv_f) * zl;
  } else {
    cx<R> last = pick<R>(k.inv_b, k.finv_b) * r[LS - 1];
#pragma unroll
    for (int s = 0; s < LS - 1; ++s) last = last + pick<R>(k.cbk[s], k.fcbk[s]) * r[s];
    r[LS - 1] = last;
#pragma unroll
    for (int s = LS - 2; s >= 0; --s) r[s] = r[s] + pick<R>(k.ab[s], k.fab[s]) * r[s + 1];
  }
#pragma unroll
  for (int s = 0; s < LS; ++s) v[s * 12] = r[s];
}

// M5^-1 (dag 0: P+ forward-corner, P- backward-corner; dag 1 swapped) on the
// s-columns staged in LDS; one thread per (site, complex component).  Columns
// are numbered component-major, so that whole wavefronts take the same
// corner; for the common Ls the column is swept in registers.
template <int DAG, typename R>
__device__ __forceinline__ void m5inv_columns(cx<R>* col, int sst, int ns, int nvalid, int ls,
                                              const M5Side& k) {
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int nthr = blockDim.x * blockDim.y;
  for (int w = tid; w < ns * 12; w += nthr) {
    const int cc = w / ns, sl = w - cc * ns;
    if (sl >= nvalid) continue;
    cx<R>* v = col + sl * sst + cc;
    const bool upper = cc < 6;
    if (ls == 12)
      m5inv_column_reg<12, DAG>(v, upper, k);
    else if (ls == 16)
      m5inv_column_reg<16, DAG>(v, upper, k);
    else if (ls == 8)
      m5inv_column_reg<8, DAG>(v, upper, k);
    else if (upper == (DAG == 0))
      m5_fwd_corner(v, ls, k);
    else
      m5_bwd_corner(v, ls, k);
  }
}

// FL flags of k_mobius
enum { FL_NORM = 1, FL_PRE_T = 2, FL_POST_T = 4, FL_CG = 8, FL_SPLIT_T = 16 };
// FL_NORM   : |out|^2 partials (dag 0 EPI_SCHUR)
// FL_PRE_T  : dag 0, `in` already holds (b + c D5) applied to the operand, so
//             the gather is a plain load
// FL_POST_T : dag 0 EPI_M5INV, write (b + c D5) M5^-1 (...) so that the next
//             pass can use FL_PRE_T
// FL_CG     : dag 1 EPI_SCHUR, feed the result to cg_epilogue instead of `out`
// FL_SPLIT_T: dag 0 without FL_PRE_T, gather (b + c D5) in reading each
//             neighbour once (hop4_tx; exchange slots after the links:
//             split_col_offset)
template <typename R>
__device__ __forceinline__ void apply_t_col(cx<R>* tv, const cx<R>* col_site, int s, int ls,
                                            const MobArgs& a, int dag) {
  // (b + c D5) on one s-column held in LDS: dag 0 takes P+ from s-1 and P-
  // from s+1, dag 1 the reverse (zmobius_dslash_4.C:90-105, :189-202)
  const cx<R>* h0 = col_site + s * 12;
  const cx<R>* hm = col_site + (s > 0 ? s - 1 : ls - 1) * 12;
  const cx<R>* hp = col_site + (s < ls - 1 ? s + 1 : 0) * 12;
  const R cm = s > 0 ? pick<R>(a.c, a.fc) : pick<R>(a.mc, a.fmc);
  const R cp = s < ls - 1 ? pick<R>(a.c, a.fc) : pick<R>(a.mc, a.fmc);
  const R b = pick<R>(a.b, a.fb);
#pragma unroll
  for (int c = 0; c < 12; ++c) {
    const bool from_minus = (c < 6) == (dag == 0);
    tv[c] = b * h0[c] + (from_minus ? cm * hm[c] : cp * hp[c]);
  }
}

// psi(s) + K D5^(dag) psi (s), the Schur "M5" term (zmobius_m-orig.h:96-121),
// with psi(s) already in v and the s-column of psi in LDS, so that psi is
// read from memory once per (site, slice) instead of three times
template <typename R>
__device__ __forceinline__ void m5_term_col(cx<R>* v, const cx<R>* col_site, int s, int ls,
                                            const MobArgs& a, int dag) {
  const cx<R>* hm = col_site + (s > 0 ? s - 1 : ls - 1) * 12;
  const cx<R>* hp = col_site + (s < ls - 1 ? s + 1 : 0) * 12;
  const R km = s > 0 ? pick<R>(a.K, a.fK) : pick<R>(a.mK, a.fmK);
  const R kp = s < ls - 1 ? pick<R>(a.K, a.fK) : pick<R>(a.mK, a.fmK);
#pragma unroll
  for (int c = 0; c < 12; ++c) {
    const bool from_minus = (c < 6) == (dag == 0);
    v[c] = v[c] + (from_minus ? km * hm[c] : kp * hp[c]);
  }
}

// whether the dag-0 gather of (b + c D5) in reads every neighbour once and
// trades the halves of slices s-1 / s+1 through LDS (hop4_tx) rather than
// reading the slices s-1, s, s+1 of every neighbour: FP64 and FP32, whose
// kernels wait on memory.  The 16-bit kernels are bound by instruction issue
// and keep the three-slice gather, which projects before converting (the
// exchange made them 7-8 % slower)
template <class P>
struct MobSplitT {
  static constexpr bool value = std::is_same<P, P64>::value || std::is_same<P, P32>::value;
};
// byte offset of the exchange slots of hop4_tx ([ls][ns] slots of
// TX_SLOT complex): after the staged neighbour indices and links, 16-byte
// aligned
constexpr int TX_SLOT = 7;  // 6 complex, padded against LDS bank conflicts
template <typename R>
__host__ __device__ constexpr size_t split_col_offset(int ns) {
  return ((size_t)8 * ns * 4 + (size_t)8 * ns * 9 * 2 * sizeof(R) + 15) / 16 * 16;
}

template <class P, int DAG, int EPI, int FL, class PX = P>
__global__ void __launch_bounds__(512)
k_mobius(Fld<P> out, CFld<P> in, CFld<P> psi, const typename P::S* __restrict__ U,
         double* __restrict__ partial, Geom g, int p_out, MobArgs a, CgEpi<P, PX> cg) {
  typedef typename P::R R;
  if (a.active && *a.active == 0) return;  // uniform: before any barrier
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double sh[16];
  const int ns = blockDim.x, il = threadIdx.x, s = threadIdx.y;
  int* lnb = reinterpret_cast<int*>(smem_raw);                   // neighbours [8][ns]
  cx<R>* lu = reinterpret_cast<cx<R>*>(smem_raw + 8 * ns * 4);  // staged links [8][ns][9]
  // s-columns [ns][sst] reuse the LDS of the links and neighbour indices,
  // which the hop no longer needs once every thread has passed the barrier
  // before the first column write: the smaller footprint lets more
  // work-groups share a CU
  cx<R>* col = reinterpret_cast<cx<R>*>(smem_raw);
  const int sst = g.ls * 12 + 1;                    // padded site stride in LDS
  const double a_cg = (FL & FL_CG) ? cg_alpha(cg.scal) : 0.0;
  const int blk = xcd_block(g.nxcd, g.rev);
  const int i = blk * ns + il;
  const bool act = i < g.vh;
  const int nvalid = min(ns, g.vh - blk * ns);
  const size_t slice = (size_t)nch<P>() * g.vh;
  const int ls = g.ls;
  typename P::C* od = out.d + s * slice;
  float* osc = out.sc + (size_t)s * g.vh;

  // int16 passes whose input is used as it is (no (b + c D5) on the fly):
  // integer hop on packed links
  constexpr bool kPQ = std::is_same<P, PI16>::value && !(DAG == 0 && !(FL & FL_PRE_T));
  if constexpr (kPQ)
    stage_links_pq(reinterpret_cast<LinkPQ*>(lu), lnb, U, g, blk, ns, 1, p_out);
  else
    stage_links<P>(lu, lnb, U, g, blk, ns, 1, p_out);
  __syncthreads();
  cx<R> acc[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) acc[c] = cx<R>{0, 0};
  constexpr bool split_t = DAG == 0 && !(FL & FL_PRE_T) && (FL & FL_SPLIT_T);
  cx<R>* mycol = col + il * sst;
  if constexpr (split_t) {
    const PlainLoader<P> ld{in.d + s * slice, in.sc + (size_t)s * g.vh, g.vh};
    cx<R>* tx = reinterpret_cast<cx<R>*>(smem_raw + split_col_offset<R>(ns));
    const int sm = s > 0 ? s - 1 : ls - 1, sp = s < ls - 1 ? s + 1 : 0;
    const R b = pick<R>(a.b, a.fb);
    const R cm = s > 0 ? pick<R>(a.c, a.fc) : pick<R>(a.mc, a.fmc);
    const R cp = s < ls - 1 ? pick<R>(a.c, a.fc) : pick<R>(a.mc, a.fmc);
    hop4_tx(acc, ld, lu + il * 9, ns * 9, lnb + il, ns, act, tx + (s * ns + il) * TX_SLOT,
            tx + (sm * ns + il) * TX_SLOT, tx + (sp * ns + il) * TX_SLOT, b, cm, cp);
  } else if (act) {
    if (DAG == 0 && !(FL & FL_PRE_T)) {
      MobTLoader<P> ld;
      const int sm = s > 0 ? s - 1 : ls - 1, sp = s < ls - 1 ? s + 1 : 0;
      ld.cur = in.d + s * slice;
      ld.prv = in.d + sm * slice;
      ld.nxt = in.d + sp * slice;
      ld.sc_cur = in.sc + (size_t)s * g.vh;
      ld.sc_prv = in.sc + (size_t)sm * g.vh;
      ld.sc_nxt = in.sc + (size_t)sp * g.vh;
      ld.vh = g.vh;
      ld.b = pick<R>(a.b, a.fb);
      ld.cm = s > 0 ? pick<R>(a.c, a.fc) : pick<R>(a.mc, a.fmc);
      ld.cp = s < ls - 1 ? pick<R>(a.c, a.fc) : pick<R>(a.mc, a.fmc);
      hop4<0>(acc, ld, lu + il * 9, ns * 9, lnb + il, ns);
    } else {
      const PlainLoader<P> ld{in.d + s * slice, in.sc + (size_t)s * g.vh, g.vh};
      if constexpr (kPQ)
        hop4_pq<DAG>(acc, ld, reinterpret_cast<const LinkPQ*>(lu) + il * 9, ns * 9, lnb + il, ns);
      else
        hop4<DAG>(acc, ld, lu + il * 9, ns * 9, lnb + il, ns);
    }
  }

  if (DAG == 0 && EPI == EPI_PLAIN) {
    if (act) st_site<P>(od, osc, g.vh, i, acc);
    return;
  }

  if (DAG == 0 && EPI == EPI_SCHUR) {
    // psi(s) through the s-column (the links' LDS is free after the barrier)
    cx<R> v[12];
    if (act) ld_spinor_once<P>(psi.d + s * slice, psi.sc + (size_t)s * g.vh, g.vh, i, v);
    __syncthreads();
    if (act)
#pragma unroll
      for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = v[c];
    __syncthreads();
    double nrm = 0.0;
    if (act) {
      m5_term_col(v, mycol, s, ls, a, 0);
      const R nkb = -pick<R>(a.kappa_b, a.fkappa_b);
#pragma unroll
      for (int c = 0; c < 12; ++c) v[c] = v[c] + nkb * acc[c];
      if (FL & FL_NORM) nrm = norm12(v);
      st_site<P>(od, osc, g.vh, i, v);
    }
    if (FL & FL_NORM) {
      const double t = block_sum(nrm, sh);
      if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.x] = t;
    }
    return;
  }

  __syncthreads();  // every thread is done with the staged links
  if (DAG == 1) {
    // (b + c D5^dag) on the hop output: needs s-1 / s+1 of the same site
#pragma unroll
    for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = acc[c];
    __syncthreads();
    cx<R> tv[12];
    apply_t_col(tv, mycol, s, ls, a, 1);
    if (EPI == EPI_PLAIN) {
      if (act) st_site<P>(od, osc, g.vh, i, tv);
      return;
    }
    if (EPI == EPI_SCHUR) {
      // psi(s) through the s-column, once every thread has read the hop's
      cx<R> v[12];
      if (act) ld_spinor_once<P>(psi.d + s * slice, psi.sc + (size_t)s * g.vh, g.vh, i, v);
      __syncthreads();
      if (act)
#pragma unroll
        for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = v[c];
      __syncthreads();
      double nrm = 0.0;
      if (act) {
        m5_term_col(v, mycol, s, ls, a, 1);
#pragma unroll
        for (int c = 0; c < 12; ++c) v[c] = v[c] + tv[c];
        if (FL & FL_CG)
          nrm = cg_epilogue<P, PX>(cg, g, s, i, v, a_cg);
        else
          st_site<P>(od, osc, g.vh, i, v);
      }
      if (FL & FL_CG) {
        const double t = block_sum(nrm, sh);
        if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.x] = t;
      }
      return;
    }
    __syncthreads();
    const R nkb = -pick<R>(a.kappa_b, a.fkappa_b);
#pragma unroll
    for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = nkb * tv[c];
  } else {
    const R kb = pick<R>(a.kappa_b, a.fkappa_b);
#pragma unroll
    for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = kb * acc[c];
  }
  __syncthreads();
  m5inv_columns<DAG>(col, sst, ns, nvalid, ls, a.m5);
  __syncthreads();
  if (act) {
    cx<R> v[12];
    if (DAG == 0 && (FL & FL_POST_T)) {
      apply_t_col(v, mycol, s, ls, a, 0);
    } else {
      const R sc = DAG ? pick<R>(a.kappa_b, a.fkappa_b) : (R)1;
#pragma unroll
      for (int c = 0; c < 12; ++c) v[c] = sc * mycol[s * 12 + c];
    }
    st_site<P>(od, osc, g.vh, i, v);
  }
}

// stand-alone M5^-1 (dag) on a half vector: zmobius_m5inv(out, in, ...)
// (zmobius_m5inv.C:822-841)
template <class P, int DAG>
__global__ void __launch_bounds__(512)
k_m5inv(Fld<P> out, CFld<P> in, Geom g, M5Side k) {
  typedef typename P::R R;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<R>* col = reinterpret_cast<cx<R>*>(smem_raw);
  const int ns = blockDim.x, il = threadIdx.x, s = threadIdx.y;
  const int sst = g.ls * 12 + 1;
  const int blk = xcd_block(g.nxcd, g.rev);
  const int i = blk * ns + il;
  const bool act = i < g.vh;
  const int nvalid = min(ns, g.vh - blk * ns);
  const size_t slice = (size_t)nch<P>() * g.vh;
  cx<R>* mycol = col + il * sst;
  if (act) {
    cx<R> v[12];
    ld_spinor_once<P>(in.d + s * slice, in.sc + (size_t)s * g.vh, g.vh, i, v);
#pragma unroll
    for (int c = 0; c < 12; ++c) mycol[s * 12 + c] = v[c];
  }
  __syncthreads();
  m5inv_columns<DAG>(col, sst, ns, nvalid, g.ls, k);
  __syncthreads();
  if (act) st_site<P>(out.d + s * slice, out.sc + (size_t)s * g.vh, g.vh, i, mycol + s * 12);
}

// ---------------------------------------------------------------------------
// Vector algebra (vector_util.C:727-900) on whole half vectors, one thread per
// (slice, site) with a grid-stride loop; FP64 reductions as per-block
// partial sums over a fixed grid, so totals are reproducible.
// ---------------------------------------------------------------------------
constexpr int BLAS_BLOCKS = 1024;
constexpr int BLAS_THREADS = 256;

template <class P>
__device__ __forceinline__ void ld_site(cx<typename P::R>* v, CFld<P> f, const Geom& g, size_t k) {
  // k = s*vh + i
  const size_t s = k / g.vh, i = k - s * g.vh;
  ld_spinor_once<P>(f.d + s * nch<P>() * (size_t)g.vh, f.sc + s * (size_t)g.vh, g.vh, (int)i, v);
}

template <class P>
__device__ __forceinline__ void st_site_k(Fld<P> f, const Geom& g, size_t k,
                                          const cx<typename P::R>* v) {
  const size_t s = k / g.vh, i = k - s * g.vh;
  st_site<P>(f.d + s * nch<P>() * (size_t)g.vh, f.sc + s * (size_t)g.vh, g.vh, (int)i, v);
}

#define DWFK_SITE_LOOP(k, g) \
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < (size_t)(g).vh * (g).ls; \
       k += (size_t)gridDim.x * blockDim.x)

template <class P>
__global__ void __launch_bounds__(BLAS_THREADS)
k_norm(CFld<P> x, Geom g, double* __restrict__ partial) {
  __shared__ double sh[16];
  double acc = 0.0;
  DWFK_SITE_LOOP(k, g) {
    cx<typename P::R> v[12];
    ld_site<P>(v, x, g, k);
    acc += norm12(v);
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// sol += a dir ; res -= a mmp ; partial |res|^2   (inv_cg.C:340-349)
template <class P, class PX>
__global__ void __launch_bounds__(BLAS_THREADS)
k_cg_update(Fld<PX> sol, Fld<P> res, CFld<P> dir, CFld<P> mmp, double a, Geom g,
            double* __restrict__ partial) {
  typedef typename P::R R;
  typedef typename PX::R RX;
  __shared__ double sh[16];
  double acc = 0.0;
  DWFK_SITE_LOOP(k, g) {
    cx<R> d[12], m[12], r[12];
    cx<RX> x[12];
    ld_site<P>(d, dir, g, k);
    ld_site<P>(m, mmp, g, k);
    ld_site<P>(r, CFld<P>{res.d, res.sc}, g, k);
    ld_site<PX>(x, CFld<PX>{sol.d, sol.sc}, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      x[c] = cx<RX>{(RX)(a * d[c].re) + x[c].re, (RX)(a * d[c].im) + x[c].im};
      r[c] = cx<R>{(R)(-a * m[c].re) + r[c].re, (R)(-a * m[c].im) + r[c].im};
    }
    st_site_k<PX>(sol, g, k, x);
    st_site_k<P>(res, g, k, r);
    acc += norm12(r);
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// y = a x + b y   (fTimesV1PlusV2 and friends)
template <class P>
__global__ void __launch_bounds__(BLAS_THREADS)
k_axpby(Fld<P> y, CFld<P> x, double a, double b, Geom g) {
  typedef typename P::R R;
  const R ar = (R)a, br = (R)b;
  DWFK_SITE_LOOP(k, g) {
    cx<R> xv[12], yv[12];
    ld_site<P>(xv, x, g, k);
    ld_site<P>(yv, CFld<P>{y.d, y.sc}, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) yv[c] = ar * xv[c] + br * yv[c];
    st_site_k<P>(y, g, k, yv);
  }
}

// x += a p, then (UPDATE_P) p = r + b p: the vector updates of a CG step
// whose hop epilogue updated r only (PX = NoX there); x in its own policy.
template <class P, class PX, bool UPDATE_P>
__global__ void __launch_bounds__(BLAS_THREADS)
k_cg_xp(Fld<PX> x, Fld<P> p, CFld<P> r, double a, double b, Geom g) {
  typedef typename P::R R;
  typedef typename PX::R RX;
  const R br = (R)b;
  DWFK_SITE_LOOP(k, g) {
    cx<R> pv[12];
    cx<RX> xv[12];
    ld_site<P>(pv, CFld<P>{p.d, p.sc}, g, k);
    ld_site<PX>(xv, CFld<PX>{x.d, x.sc}, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c)
      xv[c] = cx<RX>{(RX)(a * pv[c].re) + xv[c].re, (RX)(a * pv[c].im) + xv[c].im};
    st_site_k<PX>(x, g, k, xv);
    if (UPDATE_P) {
      cx<R> rv[12];
      ld_site<P>(rv, r, g, k);
#pragma unroll
      for (int c = 0; c < 12; ++c) pv[c] = rv[c] + br * pv[c];
      st_site_k<P>(p, g, k, pv);
    }
  }
}

// pn = r + b p (the next CG direction into another buffer)
template <class P>
__global__ void __launch_bounds__(BLAS_THREADS)
k_pnext(Fld<P> pn, CFld<P> r, CFld<P> p, const double* __restrict__ scal, Geom g) {
  typedef typename P::R R;
  // beta = |r|^2 / the previous |r|^2 from the device scalars (k_reduce_cg_step)
  const double rr = scal[0], ro = scal[2];
  const R br = (R)(ro != 0.0 ? rr / ro : 0.0);
  DWFK_SITE_LOOP(k, g) {
    cx<R> rv[12], pv[12];
    ld_site<P>(rv, r, g, k);
    ld_site<P>(pv, p, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) pv[c] = rv[c] + br * pv[c];
    st_site_k<P>(pn, g, k, pv);
  }
}

// x += a_0 p_0, then += a_1 p_1, ... (the order and the rounding of one
// update per CG step, k_cg_xp): the x updates of cnt steps in one pass
constexpr int XFLUSH_MAX = 16;
template <class P>
struct XFlush {
  const typename P::C* p[XFLUSH_MAX];
  const float* sc[XFLUSH_MAX];
  const double* alpha;  // device alphas of the steps (k_reduce_cg_step)
  int ai[XFLUSH_MAX];   // index into alpha of each pending step
  int cnt;
};
template <class P, class PX>
__global__ void __launch_bounds__(BLAS_THREADS)
k_x_flush(Fld<PX> x, XFlush<P> f, Geom g) {
  typedef typename P::R R;
  typedef typename PX::R RX;
  DWFK_SITE_LOOP(k, g) {
    cx<RX> xv[12];
    ld_site<PX>(xv, CFld<PX>{x.d, x.sc}, g, k);
    for (int j = 0; j < f.cnt; ++j) {
      cx<R> pv[12];
      ld_site<P>(pv, CFld<P>{f.p[j], f.sc[j]}, g, k);
      const double a = f.alpha[f.ai[j]];
#pragma unroll
      for (int c = 0; c < 12; ++c)
        xv[c] = cx<RX>{(RX)(a * pv[c].re) + xv[c].re, (RX)(a * pv[c].im) + xv[c].im};
    }
    st_site_k<PX>(x, g, k, xv);
  }
}

// z = a x + y, partial |z|^2 optional
template <class P, bool NORM>
__global__ void __launch_bounds__(BLAS_THREADS)
k_axpy_to(Fld<P> z, CFld<P> x, CFld<P> y, double a, Geom g, double* __restrict__ partial) {
  typedef typename P::R R;
  __shared__ double sh[16];
  double acc = 0.0;
  const R ar = (R)a;
  DWFK_SITE_LOOP(k, g) {
    cx<R> xv[12], yv[12];
    ld_site<P>(xv, x, g, k);
    ld_site<P>(yv, y, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) yv[c] = ar * xv[c] + yv[c];
    st_site_k<P>(z, g, k, yv);
    if (NORM) acc += norm12(yv);
  }
  if (NORM) {
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

template <class P>
__global__ void __launch_bounds__(BLAS_THREADS) k_scale(Fld<P> x, double a, Geom g) {
  typedef typename P::R R;
  const R ar = (R)a;
  DWFK_SITE_LOOP(k, g) {
    cx<R> v[12];
    ld_site<P>(v, CFld<P>{x.d, x.sc}, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) v[c] = ar * v[c];
    st_site_k<P>(x, g, k, v);
  }
}

// out = in in another storage precision
template <class PO, class PI>
__global__ void __launch_bounds__(BLAS_THREADS) k_convert(Fld<PO> out, CFld<PI> in, Geom g) {
  typedef typename PO::R RO;
  DWFK_SITE_LOOP(k, g) {
    cx<typename PI::R> v[12];
    cx<RO> w[12];
    ld_site<PI>(v, in, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) w[c] = cx<RO>{(RO)v[c].re, (RO)v[c].im};
    st_site_k<PO>(out, g, k, w);
  }
}

// x (FP64) += xs (any precision); xs = 0 afterwards
template <class PS>
__global__ void __launch_bounds__(BLAS_THREADS) k_accumulate(Fld<P64> x, Fld<PS> xs, Geom g) {
  DWFK_SITE_LOOP(k, g) {
    cx<typename PS::R> v[12], z[12];
    cx<double> w[12];
    ld_site<PS>(v, CFld<PS>{xs.d, xs.sc}, g, k);
    ld_site<P64>(w, CFld<P64>{x.d, x.sc}, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      w[c] = cx<double>{w[c].re + (double)v[c].re, w[c].im + (double)v[c].im};
      z[c] = cx<typename PS::R>{0, 0};
    }
    st_site_k<P64>(x, g, k, w);
    st_site_k<PS>(xs, g, k, z);
  }
}

// reference streaming copy (16-byte words)
__global__ void __launch_bounds__(256)
k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n16;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

// Bandwidth probes.  Each thread issues BW_UNROLL independent 16-byte loads
// per stream before it stores anything (grid-stride, a plain tail loop for
// the remainder); loads and stores are non-temporal.  MODE 0: d = a (copy);
// MODE 1: read a only (the xor of a thread's words goes to sink[thread], a
// grid-sized buffer, so the loads are kept); MODE 2: d = a ^ b ^ c (three
// reads per write, the read:write mix of the hop kernels).
constexpr int BW_UNROLL = 4;
template <int MODE>
__global__ void __launch_bounds__(256)
k_bw_probe(u32x4_t* __restrict__ d, const u32x4_t* __restrict__ a,
           const u32x4_t* __restrict__ b, const u32x4_t* __restrict__ c, size_t n16,
           u32x4_t* __restrict__ sink) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  u32x4_t acc = {0u, 0u, 0u, 0u};
  size_t k = tid;
  for (; k + (BW_UNROLL - 1) * stride < n16; k += BW_UNROLL * stride) {
    u32x4_t v[BW_UNROLL];
#pragma unroll
    for (int j = 0; j < BW_UNROLL; ++j) v[j] = __builtin_nontemporal_load(a + k + j * stride);
    if constexpr (MODE == 2) {
#pragma unroll
      for (int j = 0; j < BW_UNROLL; ++j)
        v[j] ^= __builtin_nontemporal_load(b + k + j * stride) ^
                __builtin_nontemporal_load(c + k + j * stride);
    }
    if constexpr (MODE == 1) {
#pragma unroll
      for (int j = 0; j < BW_UNROLL; ++j) acc ^= v[j];
    } else {
#pragma unroll
      for (int j = 0; j < BW_UNROLL; ++j) __builtin_nontemporal_store(v[j], d + k + j * stride);
    }
  }
  for (; k < n16; k += stride) {
    u32x4_t v = __builtin_nontemporal_load(a + k);
    if constexpr (MODE == 2) v ^= b[k] ^ c[k];
    if constexpr (MODE == 1)
      acc ^= v;
    else
      d[k] = v;
  }
  if constexpr (MODE == 1) sink[tid] = acc;
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

// End of one step of the sloppy CG: |r|^2 of the partials into scal[0],
// its previous value into scal[2] (for beta), and the step's alpha =
// |r|^2 / <p, A p> (the value cg_alpha gave the epilogue) into *alpha_out,
// all on the device so that the host need not wait for the step.
__global__ void __launch_bounds__(1024)
k_reduce_cg_step(const double* __restrict__ partial, int n, double* __restrict__ scal,
                 double* __restrict__ alpha_out) {
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
  if (threadIdx.x == 0) {
    const double rr = scal[0];
    *alpha_out = cg_alpha(scal);
    scal[2] = rr;
    scal[0] = t;
  }
}

// CPS half vector [s][cb site][24 reals] (FP64, host order) <-> device field
template <class P>
__global__ void __launch_bounds__(256)
k_cps_to_field(Fld<P> f, const double* __restrict__ cps, Geom g) {
  typedef typename P::R R;
  DWFK_SITE_LOOP(k, g) {
    const size_t s = k / g.vh;
    const int ic = (int)(k - s * g.vh);
    const double* src = cps + k * 24;
    cx<R> v[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) v[c] = cx<R>{(R)src[2 * c], (R)src[2 * c + 1]};
    st_site_k<P>(f, g, s * g.vh + site_of_cps(g, ic), v);
  }
}

template <class P>
__global__ void __launch_bounds__(256)
k_field_to_cps(double* __restrict__ cps, CFld<P> f, Geom g) {
  typedef typename P::R R;
  DWFK_SITE_LOOP(k, g) {
    const size_t s = k / g.vh;
    const int ic = (int)(k - s * g.vh);
    cx<R> v[12];
    ld_site<P>(v, f, g, s * g.vh + site_of_cps(g, ic));
    double* dst = cps + k * 24;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      dst[2 * c] = (double)v[c].re;
      dst[2 * c + 1] = (double)v[c].im;
    }
  }
}

// ---------------------------------------------------------------------------
// Host formats of whole fermion fields (FdwfBase::Fconvert, convert.C:639-831;
// Ffour2five / Ffive2four, f_dwf_base.C:961-1168).  `canon` is one 4D slice
// in CANONICAL order: site x + lx (y + ly (z + lz t)), 24 reals.  The 5D site
// (x, s) lives in the half `first` or `second` of a full vector:
//   DWF    (convert.C:728-760): first = 5D-odd half ((x+y+z+t+s) odd);
//   Mobius (convert.C:797-831): first = 4D-even half, every slice.
// up / lo select the upper (reals 0..11) and lower (12..23) spin pair.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool site_in_first(int p4, int s, bool dwf) {
  return dwf ? ((p4 + s) & 1) == 1 : p4 == 0;
}
__device__ __forceinline__ void canon_coords(const Geom& g, size_t k, int* c) {
  c[0] = (int)(k % g.lx);
  size_t r = k / g.lx;
  c[1] = (int)(r % g.ly);
  r /= g.ly;
  c[2] = (int)(r % g.lz);
  c[3] = (int)(r / g.lz);
}

// slice s of (first, second) = the spin pairs of canon selected by up / lo,
// zero elsewhere; canon == nullptr writes zeros
template <class P>
__global__ void __launch_bounds__(256)
k_canon_to_halves(Fld<P> first, Fld<P> second, const double* __restrict__ canon, Geom g, int s,
                  int dwf, int up, int lo) {
  typedef typename P::R R;
  const size_t v4 = (size_t)2 * g.vh;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < v4;
       k += (size_t)gridDim.x * blockDim.x) {
    int c[4];
    canon_coords(g, k, c);
    const int p4 = (c[0] + c[1] + c[2] + c[3]) & 1;
    const int i = site_index(g, c[0] >> 1, c[1], c[2], c[3]);
    cx<R> v[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      const bool take = canon && (q < 6 ? up : lo);
      v[q] = take ? cx<R>{(R)canon[k * 24 + 2 * q], (R)canon[k * 24 + 2 * q + 1]} : cx<R>{0, 0};
    }
    st_site_k<P>(site_in_first(p4, s, dwf) ? first : second, g, (size_t)s * g.vh + i, v);
  }
}

// the spin pairs selected by up / lo of canon = slice s of (first, second);
// the other reals of canon are left as they are
template <class P>
__global__ void __launch_bounds__(256)
k_halves_to_canon(double* __restrict__ canon, CFld<P> first, CFld<P> second, Geom g, int s,
                  int dwf, int up, int lo) {
  typedef typename P::R R;
  const size_t v4 = (size_t)2 * g.vh;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < v4;
       k += (size_t)gridDim.x * blockDim.x) {
    int c[4];
    canon_coords(g, k, c);
    const int p4 = (c[0] + c[1] + c[2] + c[3]) & 1;
    const int i = site_index(g, c[0] >> 1, c[1], c[2], c[3]);
    cx<R> v[12];
    ld_site<P>(v, site_in_first(p4, s, dwf) ? first : second, g, (size_t)s * g.vh + i);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      if (q < 6 ? up : lo) {
        canon[k * 24 + 2 * q] = (double)v[q].re;
        canon[k * 24 + 2 * q + 1] = (double)v[q].im;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Gauge observables of the FP64 device links (ReadLatticeParallel::
// CheckPlaqLinktrace, ReadLatticePar.C:396-450; Lattice::SumReTrPlaqNode,
// lattice_base.C:2878-2905): per block, the sum over its sites of
// ReTr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag for mu < nu, and of
// ReTr U_mu(x).  The anti-periodic sign carried by the device links cancels
// in every plaquette (a boundary link of direction mu enters it twice or
// not at all); the link trace divides it out.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_link64(cx<double>* m, const cx<double>* __restrict__ U,
                                            const Geom& g, int mu, const int* c) {
  const int par = (c[0] + c[1] + c[2] + c[3]) & 1;
  const int i = site_index(g, c[0] >> 1, c[1], c[2], c[3]);
  const cx<double>* src = U + (size_t)((mu * 2 + par) * 9) * g.vh + i;
#pragma unroll
  for (int e = 0; e < 9; ++e) m[e] = src[(size_t)e * g.vh];
}
__device__ __forceinline__ void mat_mul3(cx<double>* o, const cx<double>* a, const cx<double>* b) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      o[r * 3 + q] = cmul(a[r * 3 + 0], b[0 * 3 + q]) + cmul(a[r * 3 + 1], b[1 * 3 + q]) +
                     cmul(a[r * 3 + 2], b[2 * 3 + q]);
}

__global__ void __launch_bounds__(256)
k_plaquette(const cx<double>* __restrict__ U, Geom g, double* __restrict__ partial) {
  __shared__ double sh[16];
  const size_t v4 = (size_t)2 * g.vh;
  const int L[4] = {g.lx, g.ly, g.lz, g.lt};
  double plaq = 0.0, tr = 0.0;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < v4;
       k += (size_t)gridDim.x * blockDim.x) {
    int c[4];
    canon_coords(g, k, c);
    for (int mu = 0; mu < 4; ++mu) {
      cx<double> umu[9];
      load_link64(umu, U, g, mu, c);
      const bool neg = ((g.bcmask >> mu) & 1) && c[mu] == L[mu] - 1;
      tr += (neg ? -1.0 : 1.0) * (umu[0].re + umu[4].re + umu[8].re);
      for (int nu = mu + 1; nu < 4; ++nu) {
        int cm[4] = {c[0], c[1], c[2], c[3]}, cn[4] = {c[0], c[1], c[2], c[3]};
        cm[mu] = cm[mu] + 1 == L[mu] ? 0 : cm[mu] + 1;
        cn[nu] = cn[nu] + 1 == L[nu] ? 0 : cn[nu] + 1;
        cx<double> unu[9], a[9], b[9], t[9];
        load_link64(t, U, g, nu, cm);  // U_nu(x + mu)
        mat_mul3(a, umu, t);
        load_link64(unu, U, g, nu, c);
        load_link64(t, U, g, mu, cn);  // U_mu(x + nu)
        mat_mul3(b, unu, t);
        // ReTr(a b^dag) = sum_ij Re(a_ij conj(b_ij))
#pragma unroll
        for (int e = 0; e < 9; ++e) plaq += a[e].re * b[e].re + a[e].im * b[e].im;
      }
    }
  }
  const double tp = block_sum(plaq, sh);
  __syncthreads();
  const double tt = block_sum(tr, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = tp;
    partial[gridDim.x + blockIdx.x] = tt;
  }
}

// compressed links of policy P from the FP64 device links (rows 0 and 1)
template <class P>
__global__ void __launch_bounds__(256)
k_compress_links(typename P::S* __restrict__ dst, const cx<double>* __restrict__ src, Geom g) {
  const size_t n = (size_t)8 * g.vh;  // (mu, parity, site)
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const size_t mp = k / g.vh, i = k - mp * g.vh;
    const float inv = 1.0f;  // |link element| <= 1: bf16 as is, i16 normalised
#pragma unroll
    for (int e = 0; e < 6; ++e) {
      const cx<double> u = src[(mp * 9 + e) * g.vh + i];
      dst[(mp * 6 + e) * g.vh + i] = Pack<P>::pack(cx<typename P::R>{(typename P::R)u.re,
                                                                      (typename P::R)u.im}, inv);
    }
  }
}


// ---------------------------------------------------------------------------
// Multi-shift CG (DiracOp::MInvCG, d_op_base/noarch/minvcg.C:47-355): the
// updates of every shift in one pass per iteration.  For each shift s with
// bit s of psi_mask: psi_s += fpsi[s] p_s (the solution update that ends an
// iteration, :266-285); then with bit s of p_mask: p_s = zr[s] r + ap[s] p_s
// (the next iteration's direction, :210-222).  SINGLE keeps one psi that
// receives every shift's term.
// ---------------------------------------------------------------------------
constexpr int MAXSHIFT = 32;

template <class P>
struct ShiftArgs {
  typename P::C* psi[MAXSHIFT];
  typename P::C* p[MAXSHIFT];
  double fpsi[MAXSHIFT], zr[MAXSHIFT], ap[MAXSHIFT];
  unsigned psi_mask, p_mask;
  int nshift, single;
};

template <class P>
__global__ void __launch_bounds__(BLAS_THREADS)
k_minv_update(ShiftArgs<P> a, CFld<P> r, Geom g) {
  typedef typename P::R R;
  DWFK_SITE_LOOP(k, g) {
    cx<R> rv[12], acc[12];
    ld_site<P>(rv, r, g, k);
    if (a.single) ld_site<P>(acc, CFld<P>{a.psi[0], nullptr}, g, k);
    for (int s = 0; s < a.nshift; ++s) {
      const bool up_psi = (a.psi_mask >> s) & 1u, up_p = (a.p_mask >> s) & 1u;
      if (!up_psi && !up_p) continue;
      cx<R> pv[12];
      ld_site<P>(pv, CFld<P>{a.p[s], nullptr}, g, k);
      if (up_psi) {
        const R f = (R)a.fpsi[s];
        if (a.single) {
#pragma unroll
          for (int c = 0; c < 12; ++c) acc[c] = f * pv[c] + acc[c];
        } else {
          cx<R> ps[12];
          ld_site<P>(ps, CFld<P>{a.psi[s], nullptr}, g, k);
#pragma unroll
          for (int c = 0; c < 12; ++c) ps[c] = f * pv[c] + ps[c];
          st_site_k<P>(Fld<P>{a.psi[s], nullptr}, g, k, ps);
        }
      }
      if (up_p) {
        const R zr = (R)a.zr[s], ap = (R)a.ap[s];
#pragma unroll
        for (int c = 0; c < 12; ++c) pv[c] = zr * rv[c] + ap * pv[c];
        st_site_k<P>(Fld<P>{a.p[s], nullptr}, g, k, pv);
      }
    }
    if (a.single) st_site_k<P>(Fld<P>{a.psi[0], nullptr}, g, k, acc);
  }
}

// ---------------------------------------------------------------------------
// Many right-hand sides (DWF): one launch applies the operator to every
// active RHS.  A work-group holds ns sites x Ls slices x nr RHS (threadIdx.z);
// the links staged in LDS serve all of them.  Vectors of RHS j start at
// j * vs chunks (scales at j * vh * Ls); RHS j has its own mass and CG scalars
// scal[j * MR_NSCAL + ...]; active[j] == 0 makes every kernel skip it, which
// is how a converged RHS retires without changing any launch shape.
// ---------------------------------------------------------------------------
constexpr int MR_NSCAL = 8;
enum { MR_RR = 0, MR_D = 1, MR_RRNEW = 2, MR_BETA = 3, MR_STOP = 4, MR_RRUPD = 5, MR_DELTA2 = 6,
       MR_MAXIT = 7 };

template <class P, class PX>
struct MrArgs {
  Fld<P> out;
  CFld<P> in;
  CFld<P> x0;
  const typename P::S* U;
  double* partial;  // [nrhs][gridDim.x]
  Geom g;
  int cb_in;
  double two_a5, alpha;
  const double* mass;  // [nrhs]
  const int* active;   // [nrhs]
  size_t vs;           // chunks per vector
  int nrhs;
  // CG epilogue (x in PX)
  Fld<P> r;
  CFld<P> p;
  Fld<PX> x;
  double* scal;  // [nrhs][MR_NSCAL]
};

template <class P>
__device__ __forceinline__ Fld<P> fld_at(Fld<P> f, size_t vs, size_t vss, int j) {
  return Fld<P>{f.d + j * vs, f.sc ? f.sc + j * vss : nullptr};
}
template <class P>
__device__ __forceinline__ CFld<P> fld_at(CFld<P> f, size_t vs, size_t vss, int j) {
  return CFld<P>{f.d + j * vs, f.sc ? f.sc + j * vss : nullptr};
}

// per-RHS (threadIdx.z) sum over a work-group, fixed order; valid in the
// thread with threadIdx.x == threadIdx.y == 0 of each z
__device__ __forceinline__ double zslice_sum(double v, double* red) {
  const int nxy = blockDim.x * blockDim.y;
  const int t = threadIdx.x + blockDim.x * threadIdx.y;
  double* my = red + threadIdx.z * nxy;
  my[t] = v;
  __syncthreads();
  int n = nxy;
  while (n > 1) {
    const int h = (n + 1) >> 1;
    if (t < n - h) my[t] += my[t + h];
    __syncthreads();
    n = h;
  }
  return my[0];
}

template <class P, int DAG, int EPI, bool NORM, bool CG, class PX>
__global__ void __launch_bounds__(512) k_dwf_mrhs(MrArgs<P, PX> a) {
  typedef typename P::R R;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const Geom& g = a.g;
  const int ns = blockDim.x, il = threadIdx.x, s = threadIdx.y, q = threadIdx.z;
  const int nr = blockDim.z;
  const int j = blockIdx.y * nr + q;
  // the whole group leaves when none of its RHS is active (uniform branch)
  bool any = false;
  for (int k = 0; k < nr; ++k) {
    const int jj = blockIdx.y * nr + k;
    any = any || (jj < a.nrhs && a.active[jj]);
  }
  if (!any) return;
  int* lnb = reinterpret_cast<int*>(smem_raw);
  cx<R>* lu = reinterpret_cast<cx<R>*>(smem_raw + 2 * 8 * ns * 4);
  double* red = reinterpret_cast<double*>(smem_raw + 2 * 8 * ns * 4 + 2 * 8 * ns * 9 * sizeof(cx<R>));
  const int blk = xcd_block(g.nxcd, g.rev);
  const int i = blk * ns + il;
  constexpr bool kPQ = std::is_same<P, PI16>::value;  // integer hop
  if constexpr (kPQ)
    stage_links_pq(reinterpret_cast<LinkPQ*>(lu), lnb, a.U, g, blk, ns, 2, (1 + a.cb_in) & 1);
  else
    stage_links<P>(lu, lnb, a.U, g, blk, ns, 2, (1 + a.cb_in) & 1);
  __syncthreads();
  const bool act = j < a.nrhs && a.active[j] && i < g.vh;
  const double a_cg = (CG && j < a.nrhs) ? cg_alpha(a.scal + (size_t)j * MR_NSCAL) : 0.0;
  const size_t vss = (size_t)g.vh * g.ls;
  double nrm = 0.0;
  if (act) {
    const CFld<P> in = fld_at(a.in, a.vs, vss, j);
    const size_t slice = (size_t)nch<P>() * g.vh;
    cx<R> acc[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) acc[c] = cx<R>{0, 0};
    const PlainLoader<P> ld{in.d + s * slice, in.sc + (size_t)s * g.vh, g.vh};
    if constexpr (kPQ)
      hop4_pq<DAG>(acc, ld, reinterpret_cast<const LinkPQ*>(lu) + ((s & 1) * 8 * ns + il) * 9,
                   ns * 9, lnb + (s & 1) * 8 * ns + il, ns);
    else
      hop4<DAG>(acc, ld, lu + ((s & 1) * 8 * ns + il) * 9, ns * 9, lnb + (s & 1) * 8 * ns + il,
                ns);
    const int sm = s > 0 ? s - 1 : g.ls - 1, spl = s < g.ls - 1 ? s + 1 : 0;
    const double nm = -2.0 * a.mass[j] * (0.5 * a.two_a5);
    const R cm = (R)(s > 0 ? a.two_a5 : nm);
    const R cp = (R)(s < g.ls - 1 ? a.two_a5 : nm);
    const int lo_m = DAG ? 6 : 0, lo_p = DAG ? 0 : 6;
    {
      cx<R> vm[12], vp[12];
      ld_spinor<P>(in.d + sm * slice, in.sc + (size_t)sm * g.vh, g.vh, i, vm);
      ld_spinor<P>(in.d + spl * slice, in.sc + (size_t)spl * g.vh, g.vh, i, vp);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        acc[lo_m + c] = acc[lo_m + c] + cm * vm[lo_m + c];
        acc[lo_p + c] = acc[lo_p + c] + cp * vp[lo_p + c];
      }
    }
    if (EPI == 1) {
      const CFld<P> x0 = fld_at(a.x0, a.vs, vss, j);
      cx<R> xv[12];
      ld_spinor<P>(x0.d + s * slice, x0.sc + (size_t)s * g.vh, g.vh, i, xv);
      const R al = (R)a.alpha;
#pragma unroll
      for (int c = 0; c < 12; ++c) acc[c] = xv[c] + al * acc[c];
      if (NORM) nrm = norm12(acc);
      if (CG) {
        const CgEpi<P, PX> cg{fld_at(a.r, a.vs, vss, j), fld_at(a.p, a.vs, vss, j),
                              fld_at(a.x, a.vs * nch<PX>() / nch<P>(), vss, j),
                              a.scal + (size_t)j * MR_NSCAL};
        nrm = cg_epilogue<P, PX>(cg, g, s, i, acc, a_cg);
      }
    }
    if (!CG) {
      const Fld<P> out = fld_at(a.out, a.vs, vss, j);
      st_site<P>(out.d + s * slice, out.sc + (size_t)s * g.vh, g.vh, i, acc);
    }
  }
  if (NORM || CG) {
    const double t = zslice_sum(nrm, red);
    if (il == 0 && s == 0 && j < a.nrhs) a.partial[(size_t)j * gridDim.x + blockIdx.x] = t;
  }
}

// scal[j][slot] = sum of partial[j][0..n) for every RHS (one work-group each)
__global__ void __launch_bounds__(1024)
k_reduce_mrhs(const double* __restrict__ partial, int n, double* __restrict__ scal, int slot,
              const int* __restrict__ active) {
  __shared__ double sh[16];
  const int j = blockIdx.x;
  if (!active[j]) return;
  const double* pj = partial + (size_t)j * n;
  double a0 = 0.0, a1 = 0.0;
  int k = threadIdx.x;
  for (; k + 1024 < n; k += 2048) {
    a0 += pj[k];
    a1 += pj[k + 1024];
  }
  for (; k < n; k += 1024) a0 += pj[k];
  const double t = block_sum(a0 + a1, sh);
  if (threadIdx.x == 0) scal[(size_t)j * MR_NSCAL + slot] = t;
}

// end of a CG step for every active RHS: beta = rr_new / rr, rr = rr_new;
// an RHS whose residual meets its stop condition, or has dropped by delta
// since its last reliable update, pauses (active = 0) and raises flag = 1 for
// the host; <p, A p> == 0 pauses it with flag = 2, reaching the iteration
// cap with flag = 3.
__global__ void __launch_bounds__(256)
k_cg_scalars_mrhs(double* __restrict__ scal, int* __restrict__ active, int* __restrict__ flag,
                  int* __restrict__ iters, int nrhs) {
  for (int j = threadIdx.x; j < nrhs; j += blockDim.x) {
    if (!active[j]) continue;
    double* sc = scal + (size_t)j * MR_NSCAL;
    iters[j] += 1;
    if (sc[MR_D] == 0.0) {
      active[j] = 0;
      flag[j] = 2;
      continue;
    }
    const double rrn = sc[MR_RRNEW];
    sc[MR_BETA] = rrn / sc[MR_RR];
    sc[MR_RR] = rrn;
    if (rrn <= sc[MR_STOP] || rrn < sc[MR_DELTA2] * sc[MR_RRUPD]) {
      active[j] = 0;
      flag[j] = 1;
    } else if (iters[j] >= (int)sc[MR_MAXIT]) {
      active[j] = 0;
      flag[j] = 3;
    }
  }
}

// p_j = r_j + beta_j p_j for every active RHS (blockIdx.y = RHS)
template <class P>
__global__ void __launch_bounds__(BLAS_THREADS)
k_xpby_mrhs(Fld<P> p, CFld<P> r, const double* __restrict__ scal, const int* __restrict__ active,
            size_t vs, Geom g) {
  typedef typename P::R R;
  const int j = blockIdx.y;
  if (!active[j]) return;
  const size_t vss = (size_t)g.vh * g.ls;
  const Fld<P> pj = fld_at(p, vs, vss, j);
  const CFld<P> rj = fld_at(r, vs, vss, j);
  const R b = (R)scal[(size_t)j * MR_NSCAL + MR_BETA];
  DWFK_SITE_LOOP(k, g) {
    cx<R> pv[12], rv[12];
    ld_site<P>(pv, CFld<P>{pj.d, pj.sc}, g, k);
    ld_site<P>(rv, rj, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) pv[c] = rv[c] + b * pv[c];
    st_site_k<P>(pj, g, k, pv);
  }
}

// ---------------------------------------------------------------------------
// eigCG (DiracOp::InvEigCg, d_op_base/comsrc/inv_eig_cg.C:129-631) and the
// init-CG deflation.  These kernels treat FP64 vectors as flat arrays of n
// reals (the device layout permutes sites and components identically in every
// vector, so sums and linear combinations do not see it).  A set of K vectors
// is a base pointer with `stride` reals between vectors.
// ---------------------------------------------------------------------------
constexpr int EIG_MAXRANK = 32;  // 2 * nev
constexpr int EIG_DOTK = 16;     // vectors per batched-dot launch
constexpr int EIG_BLOCKS = 512;

// V[j] = sum_k V[k] qz[k + m j] for j < rank, in place (each thread reads
// every V[k] at its own index before it writes any V[j] there)
__global__ void __launch_bounds__(256)
k_eig_vmul(double* __restrict__ V, size_t stride, size_t n, int m, int rank,
           const double* __restrict__ qz) {
  extern __shared__ double sqz[];  // [m * rank]
  for (int t = threadIdx.x; t < m * rank; t += blockDim.x) sqz[t] = qz[t];
  __syncthreads();
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    double acc[EIG_MAXRANK];
#pragma unroll
    for (int j = 0; j < EIG_MAXRANK; ++j) acc[j] = 0.0;
    for (int k = 0; k < m; ++k) {
      const double v = V[k * stride + e];
#pragma unroll
      for (int j = 0; j < EIG_MAXRANK; ++j)
        if (j < rank) acc[j] = fma(v, sqz[k + m * j], acc[j]);
    }
#pragma unroll
    for (int j = 0; j < EIG_MAXRANK; ++j)
      if (j < rank) V[j * stride + e] = acc[j];
  }
}

// per block: sum over its complex elements of conj(W_k) x, k < K <= EIG_DOTK;
// partial[(2k + c) * gridDim.x + block], c = 0 real, 1 imaginary
template <typename TW>
__global__ void __launch_bounds__(256)
k_eig_mdot(const TW* __restrict__ W, size_t stride, int K, const double* __restrict__ x,
           size_t ncx, double* __restrict__ partial) {
  __shared__ double sh[16];
  double re[EIG_DOTK], im[EIG_DOTK];
#pragma unroll
  for (int k = 0; k < EIG_DOTK; ++k) re[k] = im[k] = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < ncx;
       e += (size_t)gridDim.x * blockDim.x) {
    const double xr = x[2 * e], xi = x[2 * e + 1];
#pragma unroll
    for (int k = 0; k < EIG_DOTK; ++k) {
      if (k < K) {
        const double wr = (double)W[k * stride + 2 * e], wi = (double)W[k * stride + 2 * e + 1];
        re[k] = fma(wr, xr, fma(wi, xi, re[k]));
        im[k] = fma(wr, xi, fma(-wi, xr, im[k]));
      }
    }
  }
  for (int k = 0; k < K; ++k) {
    const double a = block_sum(re[k], sh);
    __syncthreads();
    const double b = block_sum(im[k], sh);
    __syncthreads();
    if (threadIdx.x == 0) {
      partial[(2 * k) * gridDim.x + blockIdx.x] = a;
      partial[(2 * k + 1) * gridDim.x + blockIdx.x] = b;
    }
  }
}

// out[v] = sum of partial[v * nb .. v * nb + nb), one block per value
__global__ void __launch_bounds__(256)
k_eig_reduce_rows(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  __shared__ double sh[16];
  double a = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) a += partial[(size_t)blockIdx.x * nb + b];
  const double t = block_sum(a, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// out += sum_k c_k W_k with complex c_k (coef[2k], coef[2k+1]), k < K
template <typename TW>
__global__ void __launch_bounds__(256)
k_eig_mcombo(double* __restrict__ out, const TW* __restrict__ W, size_t stride, int K,
             const double* __restrict__ coef, size_t ncx) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < ncx;
       e += (size_t)gridDim.x * blockDim.x) {
    double orr = out[2 * e], oi = out[2 * e + 1];
    for (int k = 0; k < K; ++k) {
      const double cr = coef[2 * k], ci = coef[2 * k + 1];
      const double wr = (double)W[k * stride + 2 * e], wi = (double)W[k * stride + 2 * e + 1];
      orr = fma(cr, wr, fma(-ci, wi, orr));
      oi = fma(cr, wi, fma(ci, wr, oi));
    }
    out[2 * e] = orr;
    out[2 * e + 1] = oi;
  }
}

// u = v = (double)(float)v: the deflation space is kept at single precision
// (eig_QPropW.C:315)
__global__ void __launch_bounds__(256)
k_eig_round_f32(double* __restrict__ u, double* __restrict__ v, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const double r = (double)(float)v[e];
    u[e] = r;
    v[e] = r;
  }
}

// ---------------------------------------------------------------------------
// Implicitly restarted Lanczos (DiracOp::ImpResLanczos, d_op_base/comsrc/
// lanczos.C:193-840): the Lanczos vectors are single precision, as in CPS.
// ---------------------------------------------------------------------------
constexpr int LZ_MAXM = 256;   // nk + np
constexpr int LZ_TILE = 64;    // elements per work-group of k_lz_vq

// V[k] = sum_{j<m} V[j] Q[j * m + k] for k < ncol, in place: a work-group
// stages the m values of its 64 elements in LDS before it writes any
__global__ void __launch_bounds__(LZ_TILE)
k_lz_vq(float* __restrict__ V, size_t stride, size_t n, int m, int ncol,
        const double* __restrict__ Q) {
  extern __shared__ float sv[];  // [m][LZ_TILE]
  const int t = threadIdx.x;
  for (size_t base = (size_t)blockIdx.x * LZ_TILE; base < n; base += (size_t)gridDim.x * LZ_TILE) {
    const size_t e = base + t;
    const bool act = e < n;
    for (int j = 0; j < m; ++j) sv[j * LZ_TILE + t] = act ? V[j * stride + e] : 0.0f;
    __syncthreads();
    if (act)
      for (int k = 0; k < ncol; ++k) {
        double acc = 0.0;
        for (int j = 0; j < m; ++j) acc = fma((double)sv[j * LZ_TILE + t], Q[j * m + k], acc);
        V[k * stride + e] = (float)acc;
      }
    __syncthreads();
  }
}

// dst = (float)(a * src)
__global__ void __launch_bounds__(256)
k_lz_to_f32(float* __restrict__ dst, const double* __restrict__ src, double a, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    dst[e] = (float)(a * src[e]);
}
// dst = src
__global__ void __launch_bounds__(256)
k_lz_to_f64(double* __restrict__ dst, const float* __restrict__ src, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    dst[e] = (double)src[e];
}
// r -= c v with the complex c read from the device (c[0], c[1]): one step of
// the Gram-Schmidt of a new Lanczos residual, no host round trip
__global__ void __launch_bounds__(256)
k_lz_gs_axpy(double* __restrict__ r, const float* __restrict__ v, const double* __restrict__ c,
             size_t ncx) {
  const double cr = c[0], ci = c[1];
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < ncx;
       e += (size_t)gridDim.x * blockDim.x) {
    const double vr = v[2 * e], vi = v[2 * e + 1];
    r[2 * e] = r[2 * e] - (cr * vr - ci * vi);
    r[2 * e + 1] = r[2 * e + 1] - (cr * vi + ci * vr);
  }
}

}  // namespace dwfk
