// ptc_kernels.hpp -- gfx950 kernels of the gauge-equivariant parallel-
// transport network used as a preconditioner of the CG on MatPcDagMatPc.
//
// Network (no CPS counterpart): three layers Ls -> H -> H -> Ls channels,
//   out_o(x) = sum_{d=0..8} sum_i W[d][o][i] (T_d phi_i)(x)
// with T_0 = 1, T_{+mu} phi(x) = U_mu(x) phi(x+mu), T_{-mu} phi(x) =
// U_mu(x-mu)^dag phi(x-mu) (d = 1..4: +x..+t, 5..8: -x..-t) and W[d][o][i] a
// complex 4x4 spin matrix.  Colour is only ever moved by links, so the map is
// gauge covariant.
//
// Activations: the full 4D lattice, site S = parity*vh + i (i the tiled
// half-lattice index of the operator), per site C channels x 4 spins x 3
// colours of bf16 complex numbers (one uint32: re low, im high).
//
// Each layer is a dense contraction, run on the MFMA units:
//   C[M' x N] = sum_d A_d[M' x K'] B_d[K' x N]
// with M' = 8*Cout (re/im x channel x spin), K' = 8*Cin, N = 3 colours x 32
// sites of a tile, A_d = [[Re W_d, -Im W_d], [Im W_d, Re W_d]] in bf16 and
// B_d the transported neighbour fields (bf16, FP32 accumulation).  The
// weight gradient is the contraction the other way round,
//   P_d[M' x K'] = sum_{sites, colours} X[M' x N] Y_d[K' x N]^T
// with X the output gradient and Y_d the transported input.
// MFMA: v_mfma_f32_16x16x32_bf16 (lane l holds A[l%16][8(l/16)..+7],
// B[8(l/16)..+7][l%16] and C[4(l/16)..+3][l%16]); a K remainder of 16 takes
// one v_mfma_f32_16x16x16_bf16 (4 elements per lane).  The four waves of a
// work-group split the output blocks as a 2 x 2 grid: each wave keeps the A
// fragments of its block rows and the B fragments of its block columns in
// registers for one K step and issues every product between them, so each
// fragment is read from LDS once per step, not once per output block.
#pragma once
#include "dwf_kernels.hpp"

namespace ptck {

using dwfk::cx;
using dwfk::Geom;

constexpr int NDIR = 9;
constexpr int NT = 32;          // sites per tile
constexpr int NCOL = 3 * NT;    // GEMM columns per tile
constexpr int NTHR = 256;       // threads (4 waves) per work-group
constexpr int PADK = 8;         // LDS row padding of [col][K'] tiles (bf16 elements)
// row stride of the [K'][col] / [M'][col] tiles of the weight gradient: 100
// bf16 = 50 dwords, so the consecutive rows a wavefront writes while building
// a tile land in different LDS banks (a stride of 104 repeats every 16 rows)
constexpr int ROWC = NCOL + 4;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) { return (uint16_t)dwfk::bf16_rne(x); }
__device__ __forceinline__ float bf2f(uint32_t b16) { return __uint_as_float(b16 << 16); }
__device__ __forceinline__ uint32_t pack_c(cx<float> v) { return dwfk::bf16_pair(v.re, v.im); }
__device__ __forceinline__ cx<float> unpack_c(uint32_t w) {
  return {__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u)};
}

// 4D parity held by slice s of the CG half vector: DWF (s+1)%2, Mobius 1
__host__ __device__ __forceinline__ int slice_parity(int action, int s) {
  return action == 0 ? (s + 1) & 1 : 1;
}

// CG half vector (policy P, times `scale`) -> activations with C = ls
// channels; the other parity of every channel is cleared
template <class P>
__global__ void __launch_bounds__(256)
k_ptc_embed(uint32_t* __restrict__ act, dwfk::CFld<P> in, Geom g, int action, float scale) {
  const int C = g.ls;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < (size_t)g.vh * g.ls;
       k += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(k / g.vh), i = (int)(k - (size_t)s * g.vh);
    const int p = slice_parity(action, s);
    cx<typename P::R> v[12];
    dwfk::ld_spinor<P>(in.d + (size_t)s * dwfk::nch<P>() * g.vh, in.sc + (size_t)s * g.vh, g.vh,
                       i, v);
    uint32_t* dst = act + ((size_t)(p * g.vh + i) * C + s) * 12;
    uint32_t* zer = act + ((size_t)((1 - p) * g.vh + i) * C + s) * 12;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      dst[c] = pack_c(cx<float>{(float)v[c].re * scale, (float)v[c].im * scale});
      zer[c] = 0u;
    }
  }
}

// activations (C = ls) -> CG half vector (policy P, times `scale`)
template <class P>
__global__ void __launch_bounds__(256)
k_ptc_extract(dwfk::Fld<P> out, const uint32_t* __restrict__ act, Geom g, int action,
              float scale) {
  typedef typename P::R R;
  const int C = g.ls;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < (size_t)g.vh * g.ls;
       k += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(k / g.vh), i = (int)(k - (size_t)s * g.vh);
    const int p = slice_parity(action, s);
    const uint32_t* src = act + ((size_t)(p * g.vh + i) * C + s) * 12;
    cx<R> v[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      const cx<float> w = unpack_c(src[c]);
      v[c] = cx<R>{(R)(w.re * scale), (R)(w.im * scale)};
    }
    dwfk::st_site<P>(out.d + (size_t)s * dwfk::nch<P>() * g.vh, out.sc + (size_t)s * g.vh, g.vh,
                     i, v);
  }
}

// Training loss of one pair: diff = extract(a3) - tscale * target (FP32 CG
// vector), partial sums of |diff|^2, and the output gradient diff * gscale
// embedded as activations (gscale = 1/|tscale * target|^2: the derivative of
// the normalised loss with respect to conj(output)).
__global__ void __launch_bounds__(256)
k_ptc_loss(uint32_t* __restrict__ gact, const uint32_t* __restrict__ a3,
           dwfk::CFld<dwfk::P32> target, float tscale, Geom g, int action, float gscale,
           double* __restrict__ partial) {
  __shared__ double sh[16];
  const int C = g.ls;
  double acc = 0.0;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < (size_t)g.vh * g.ls;
       k += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(k / g.vh), i = (int)(k - (size_t)s * g.vh);
    const int p = slice_parity(action, s);
    const size_t base = ((size_t)(p * g.vh + i) * C + s) * 12;
    cx<float> t[12];
    dwfk::ld_spinor<dwfk::P32>(target.d + (size_t)s * 6 * g.vh, nullptr, g.vh, i, t);
    uint32_t* zer = gact + ((size_t)((1 - p) * g.vh + i) * C + s) * 12;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      const cx<float> o = unpack_c(a3[base + c]);
      const cx<float> d{o.re - tscale * t[c].re, o.im - tscale * t[c].im};
      acc += (double)d.re * d.re + (double)d.im * d.im;
      gact[base + c] = pack_c(cx<float>{d.re * gscale, d.im * gscale});
      zer[c] = 0u;
    }
  }
  const double tot = dwfk::block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// transported colour vector: d = 0 identity, 1..4 U phi, 5..8 U^dag phi
__device__ __forceinline__ void transport3(const cx<float>* u, int d, const cx<float>* f,
                                           cx<float>* v) {
  if (d == 0) {
    v[0] = f[0];
    v[1] = f[1];
    v[2] = f[2];
  } else if (d <= 4) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
      v[a] = dwfk::row_dot(u[a * 3 + 0], u[a * 3 + 1], u[a * 3 + 2], f[0], f[1], f[2]);
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a)
      v[a] = dwfk::row_dotc(u[0 * 3 + a], u[1 * 3 + a], u[2 * 3 + a], f[0], f[1], f[2]);
  }
}

// site of the activation neighbour of tile site n along d (S index)
__device__ __forceinline__ size_t nbr_site(const int* lnb, int d, int n, int par, int i0,
                                           const Geom& g) {
  if (d == 0) return (size_t)par * g.vh + i0 + n;
  return (size_t)(1 - par) * g.vh + lnb[(d - 1) * NT + n];
}

// Fill the transported field of direction d for the tile into LDS, as bf16,
// either [col][K'] (KMAJOR = false, the B operand of the forward GEMM) or
// [K'][col] (KMAJOR = true, the Y operand of the weight gradient).
// k' = part*(CIN*4) + i*4 + spin, col = n*3 + colour.
// Staging loops issue their loads CHUNK items at a time before using any
// (a few round trips to memory per tile instead of one per item, without
// holding a whole tile in registers next to the MFMA accumulators).
constexpr int CHUNK = 4;
// the largest chunk of at most `cap` items that divides `per` (items per thread)
constexpr int chunk_for(int per, int cap) {
  int c = cap;
  while (c > 1 && per % c != 0) --c;
  return c;
}

template <int CIN, bool KMAJOR>
__device__ __forceinline__ void build_transported(uint16_t* sB, const uint32_t* __restrict__ in,
                                                  const cx<float>* lu, const int* lnb, int d,
                                                  int par, int i0, const Geom& g) {
  constexpr int KP = 8 * CIN;
  constexpr int NITEM = NT * CIN * 4, PERT = (NITEM + NTHR - 1) / NTHR;
  constexpr int CH = chunk_for(PERT, CHUNK);
  // no item guards when the items split evenly into whole chunks
  constexpr bool EXACT = NITEM % NTHR == 0 && PERT % CH == 0;
#pragma unroll 1
  for (int j0 = 0; j0 < PERT; j0 += CH) {
    uint32_t raw[CH][3];
#pragma unroll
    for (int jj = 0; jj < CH; ++jj) {
      const int w = threadIdx.x + (j0 + jj) * NTHR;
      if (EXACT || (j0 + jj < PERT && w < NITEM)) {
        const int n = w / (CIN * 4), r = w - n * (CIN * 4);
        const int i = r >> 2, sp = r & 3;
        const size_t y = nbr_site(lnb, d, n, par, i0, g);
        const uint32_t* src = in + (y * CIN + i) * 12 + sp * 3;
        raw[jj][0] = src[0];
        raw[jj][1] = src[1];
        raw[jj][2] = src[2];
      }
    }
#pragma unroll
    for (int jj = 0; jj < CH; ++jj) {
      const int w = threadIdx.x + (j0 + jj) * NTHR;
      if (!EXACT && (j0 + jj >= PERT || w >= NITEM)) continue;
      const int n = w / (CIN * 4), r = w - n * (CIN * 4);
      const int i = r >> 2, sp = r & 3;
      cx<float> f[3], v[3];
      f[0] = unpack_c(raw[jj][0]);
      f[1] = unpack_c(raw[jj][1]);
      f[2] = unpack_c(raw[jj][2]);
      transport3(d == 0 ? nullptr : lu + ((d - 1) * NT + n) * 9, d, f, v);
      const int kre = i * 4 + sp, kim = CIN * 4 + i * 4 + sp;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int col = n * 3 + c;
        if (KMAJOR) {
          sB[kre * ROWC + col] = f2bf(v[c].re);
          sB[kim * ROWC + col] = f2bf(v[c].im);
        } else {
          sB[col * (KP + PADK) + kre] = f2bf(v[c].re);
          sB[col * (KP + PADK) + kim] = f2bf(v[c].im);
        }
      }
    }
  }
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) {
  const uint2 w = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(w.x & 0xffff);
  r[1] = (short)(w.x >> 16);
  r[2] = (short)(w.y & 0xffff);
  r[3] = (short)(w.y >> 16);
  return r;
}
// 8 consecutive bf16 (16-byte aligned: one LDS load)
__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}
// 8 consecutive bf16 at an 8-byte boundary (two LDS loads)
__device__ __forceinline__ bf16x8 ld8x2(const uint16_t* p) {
  const uint2 a = *reinterpret_cast<const uint2*>(p), b = *reinterpret_cast<const uint2*>(p + 4);
  return __builtin_bit_cast(bf16x8, uint4{a.x, a.y, b.x, b.y});
}

// LDS bytes of the forward kernel
template <int CIN, int COUT>
constexpr size_t fwd_lds() {
  constexpr size_t a = (size_t)(8 * COUT) * (8 * CIN + PADK) * 2;
  constexpr size_t b = (size_t)NCOL * (8 * CIN + PADK) * 2;
  constexpr size_t c = (size_t)(8 * COUT) * NCOL * 4;
  constexpr size_t links = (size_t)8 * NT * 9 * 8 + 8 * NT * 4;
  return (a + b > c ? a + b : c) + links;
}

// One layer on one tile of NT sites per work-group:
// out[S][o][spin][colour] = sum_d W_d T_d in.  A: [9][M'][K'] bf16.
// grid.x = 2 * vh / NT tiles (first all even-parity tiles, then odd).
template <int CIN, int COUT>
__global__ void __launch_bounds__(NTHR)
k_ptc_fwd(uint32_t* __restrict__ out, const uint32_t* __restrict__ in,
          const uint16_t* __restrict__ A, const typename dwfk::P32::S* __restrict__ U, Geom g) {
  constexpr int MP = 8 * COUT, KP = 8 * CIN;
  constexpr int MB = MP / 16, NB = NCOL / 16;
  constexpr int K32 = KP / 32, K16 = (KP % 32) / 16;  // K steps of 32, a tail of 16
  static_assert(KP % 16 == 0 && NB % 2 == 0 && ((KP + PADK) * 2) % 16 == 0, "tile shapes");
  // this wave's output blocks: rows [mb0, mb0 + RM) x columns [nb0, nb0 + RN)
  constexpr int RM = (MB + 1) / 2, RN = NB / 2;
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t* sA = reinterpret_cast<uint16_t*>(smem);
  uint16_t* sB = sA + MP * (KP + PADK);
  float* sC = reinterpret_cast<float*>(smem);  // after the last GEMM
  constexpr size_t big = fwd_lds<CIN, COUT>() - ((size_t)8 * NT * 9 * 8 + 8 * NT * 4);
  cx<float>* lu = reinterpret_cast<cx<float>*>(smem + big);
  int* lnb = reinterpret_cast<int*>(smem + big + (size_t)8 * NT * 9 * 8);

  const int tiles_per_par = g.vh / NT;
  const int par = blockIdx.x / tiles_per_par;
  const int i0 = (blockIdx.x - par * tiles_per_par) * NT;
  dwfk::stage_links<dwfk::P32>(lu, lnb, U, g, i0 / NT, NT, 1, par);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mb0 = (wave & 1) * RM, nb0 = (wave >> 1) * RN;
  f32x4 acc[RM][RN];
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int d = 0; d < NDIR; ++d) {
    __syncthreads();  // links staged / previous direction's operands consumed
    {
      // the direction's weights: every load issued before the stores
      constexpr int NA = MP * KP / 8, PA = (NA + NTHR - 1) / NTHR;
      constexpr int CH = chunk_for(PA, CHUNK);
      constexpr bool EXACT = NA % NTHR == 0 && PA % CH == 0;
      const uint4* src = reinterpret_cast<const uint4*>(A + (size_t)d * MP * KP);
#pragma unroll 1
      for (int j0 = 0; j0 < PA; j0 += CH) {
        uint4 wv[CH];
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) {
          const int w = threadIdx.x + (j0 + jj) * NTHR;
          if (EXACT || (j0 + jj < PA && w < NA)) wv[jj] = src[w];
        }
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) {
          const int w = threadIdx.x + (j0 + jj) * NTHR;
          if (EXACT || (j0 + jj < PA && w < NA)) {
            const int m = w / (KP / 8), k8 = w - m * (KP / 8);
            *reinterpret_cast<uint4*>(sA + m * (KP + PADK) + k8 * 8) = wv[jj];
          }
        }
      }
    }
    build_transported<CIN, false>(sB, in, lu, lnb, d, par, i0, g);
    __syncthreads();
    const uint16_t* pa = sA + (mb0 * 16 + (lane & 15)) * (KP + PADK);
    const uint16_t* pb = sB + (nb0 * 16 + (lane & 15)) * (KP + PADK);
#pragma unroll
    for (int kb = 0; kb < K32; ++kb) {
      const int k0 = kb * 32 + 8 * (lane >> 4);
      bf16x8 bfr[RN];
#pragma unroll
      for (int j = 0; j < RN; ++j) bfr[j] = ld8(pb + j * 16 * (KP + PADK) + k0);
#pragma unroll
      for (int i = 0; i < RM; ++i) {
        if (mb0 + i < MB) {
          const bf16x8 af = ld8(pa + i * 16 * (KP + PADK) + k0);
#pragma unroll
          for (int j = 0; j < RN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr[j], acc[i][j], 0, 0, 0);
        }
      }
    }
    if constexpr (K16 > 0) {
      const int k0 = K32 * 32 + 4 * (lane >> 4);
      bf16x4 bfr[RN];
#pragma unroll
      for (int j = 0; j < RN; ++j) bfr[j] = ld4(pb + j * 16 * (KP + PADK) + k0);
#pragma unroll
      for (int i = 0; i < RM; ++i) {
        if (mb0 + i < MB) {
          const bf16x4 af = ld4(pa + i * 16 * (KP + PADK) + k0);
#pragma unroll
          for (int j = 0; j < RN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bfr[j], acc[i][j], 0, 0, 0);
        }
      }
    }
  }
  __syncthreads();
  // C (lane: rows 4(l/16)..+3 of column l%16) -> LDS [M'][NCOL]
#pragma unroll
  for (int i = 0; i < RM; ++i) {
    if (mb0 + i < MB) {
#pragma unroll
      for (int j = 0; j < RN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          sC[((mb0 + i) * 16 + 4 * (lane >> 4) + r) * NCOL + (nb0 + j) * 16 + (lane & 15)] =
              acc[i][j][r];
    }
  }
  __syncthreads();
  // out[S][o][spin][colour]: re row o*4+spin, im row COUT*4 + o*4 + spin
  for (int w = threadIdx.x; w < NT * COUT * 12; w += NTHR) {
    const int n = w / (COUT * 12), r = w - n * (COUT * 12);
    const int os = r / 3, c = r - os * 3;
    const float re = sC[os * NCOL + n * 3 + c], im = sC[(COUT * 4 + os) * NCOL + n * 3 + c];
    out[((size_t)par * g.vh + i0 + n) * (COUT * 12) + r] = pack_c(cx<float>{re, im});
  }
}

template <int CIN, int COUT>
constexpr size_t gw_lds() {
  constexpr size_t x = (size_t)(8 * COUT) * ROWC * 2;
  constexpr size_t y = (size_t)(8 * CIN) * ROWC * 2;
  return x + y + (size_t)8 * NT * 9 * 8 + 8 * NT * 4;
}

// Weight gradient partials: for direction d = blockIdx.y and the tiles
// t = blockIdx.x, blockIdx.x + gridDim.x, ...:
// part[d][blockIdx.x][M'][K'] = sum X[M'][cols] Y_d[K'][cols]^T
template <int CIN, int COUT>
__global__ void __launch_bounds__(NTHR)
k_ptc_gw(float* __restrict__ part, const uint32_t* __restrict__ gout,
         const uint32_t* __restrict__ in, const typename dwfk::P32::S* __restrict__ U, Geom g) {
  constexpr int MP = 8 * COUT, KP = 8 * CIN;
  constexpr int MB = MP / 16, KB2 = KP / 16, CB = NCOL / 32;  // contraction over columns
  static_assert(NCOL % 32 == 0, "column steps of 32");
  // this wave's output blocks: rows [mb0, mb0 + RM) x columns [kb0, kb0 + RK)
  constexpr int RM = (MB + 1) / 2, RK = (KB2 + 1) / 2;
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t* sX = reinterpret_cast<uint16_t*>(smem);
  uint16_t* sY = sX + MP * ROWC;
  cx<float>* lu = reinterpret_cast<cx<float>*>(smem + (size_t)(MP + KP) * ROWC * 2);
  int* lnb = reinterpret_cast<int*>(reinterpret_cast<unsigned char*>(lu) + (size_t)8 * NT * 9 * 8);
  const int d = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_per_par = g.vh / NT, ntiles = 2 * tiles_per_par;
  const int mb0 = (wave & 1) * RM, kb0 = (wave >> 1) * RK;
  f32x4 acc[RM][RK];
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RK; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int par = t / tiles_per_par;
    const int i0 = (t - par * tiles_per_par) * NT;
    __syncthreads();  // previous tile's operands consumed
    dwfk::stage_links<dwfk::P32>(lu, lnb, U, g, i0 / NT, NT, 1, par);
    // X: the output gradient of the tile, [M'][col]; the tile is contiguous
    // in gout, and every load is issued before the stores
    {
      constexpr int NX = NT * COUT * 12, PX = (NX + NTHR - 1) / NTHR;
      constexpr int CH = chunk_for(PX, 2 * CHUNK);
      constexpr bool EXACT = NX % NTHR == 0 && PX % CH == 0;
      const uint32_t* src = gout + ((size_t)par * g.vh + i0) * (COUT * 12);
#pragma unroll 1
      for (int j0 = 0; j0 < PX; j0 += CH) {
        uint32_t xv[CH];
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) {
          const int w = threadIdx.x + (j0 + jj) * NTHR;
          if (EXACT || (j0 + jj < PX && w < NX)) xv[jj] = src[w];
        }
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) {
          const int w = threadIdx.x + (j0 + jj) * NTHR;
          if (EXACT || (j0 + jj < PX && w < NX)) {
            const int n = w / (COUT * 12), r = w - n * (COUT * 12);
            const int os = r / 3, c = r - os * 3;
            const cx<float> v = unpack_c(xv[jj]);
            sX[os * ROWC + n * 3 + c] = f2bf(v.re);
            sX[(COUT * 4 + os) * ROWC + n * 3 + c] = f2bf(v.im);
          }
        }
      }
    }
    __syncthreads();  // links staged
    build_transported<CIN, true>(sY, in, lu, lnb, d, par, i0, g);
    __syncthreads();
    const uint16_t* px = sX + (mb0 * 16 + (lane & 15)) * ROWC;
    const uint16_t* py = sY + (kb0 * 16 + (lane & 15)) * ROWC;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      // rows of ROWC = 100 bf16 are 8-byte aligned: two loads per fragment
      const int c0 = cb * 32 + 8 * (lane >> 4);
      bf16x8 yf[RK];
#pragma unroll
      for (int j = 0; j < RK; ++j)
        if (kb0 + j < KB2) yf[j] = ld8x2(py + j * 16 * ROWC + c0);
#pragma unroll
      for (int i = 0; i < RM; ++i) {
        if (mb0 + i < MB) {
          const bf16x8 xf = ld8x2(px + i * 16 * ROWC + c0);
#pragma unroll
          for (int j = 0; j < RK; ++j)
            if (kb0 + j < KB2)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xf, yf[j], acc[i][j], 0, 0, 0);
        }
      }
    }
  }
  float* dst = part + ((size_t)d * gridDim.x + blockIdx.x) * MP * KP;
#pragma unroll
  for (int i = 0; i < RM; ++i) {
    if (mb0 + i < MB) {
#pragma unroll
      for (int j = 0; j < RK; ++j)
        if (kb0 + j < KB2)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            dst[((mb0 + i) * 16 + 4 * (lane >> 4) + r) * KP + (kb0 + j) * 16 + (lane & 15)] =
                acc[i][j][r];
    }
  }
}

// G[d][o][i][a][b] (complex, FP32) = 2 * (P_re,re + P_im,im, P_im,re - P_re,im)
// summed over the nwg partials in a fixed order
__global__ void __launch_bounds__(256)
k_ptc_gw_reduce(cx<float>* __restrict__ G, const float* __restrict__ part, int nwg, int cin,
                int cout, int accumulate) {
  const int MP = 8 * cout, KP = 8 * cin;
  const int n = NDIR * cout * 4 * cin * 4;
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += gridDim.x * blockDim.x) {
    int r = w;
    const int b = r % 4;
    r /= 4;
    const int i = r % cin;
    r /= cin;
    const int a = r % 4;
    r /= 4;
    const int o = r % cout;
    const int d = r / cout;
    const int mre = o * 4 + a, mim = cout * 4 + o * 4 + a;
    const int kre = i * 4 + b, kim = cin * 4 + i * 4 + b;
    double re = 0.0, im = 0.0;
    for (int q = 0; q < nwg; ++q) {
      const float* P = part + ((size_t)d * nwg + q) * MP * KP;
      re += (double)P[mre * KP + kre] + (double)P[mim * KP + kim];
      im += (double)P[mim * KP + kre] - (double)P[mre * KP + kim];
    }
    // G index [d][o][i][a][b]
    cx<float>& gv = G[(((size_t)(d * cout + o) * cin + i) * 4 + a) * 4 + b];
    const cx<float> old = accumulate ? gv : cx<float>{0.f, 0.f};
    gv = cx<float>{old.re + (float)(2.0 * re), old.im + (float)(2.0 * im)};
  }
}

// Re <x, y> partial sums over a half vector
template <class P>
__global__ void __launch_bounds__(dwfk::BLAS_THREADS)
k_dot(dwfk::CFld<P> x, dwfk::CFld<P> y, Geom g, double* __restrict__ partial) {
  __shared__ double sh[16];
  double acc = 0.0;
  DWFK_SITE_LOOP(k, g) {
    cx<typename P::R> a[12], b[12];
    dwfk::ld_site<P>(a, x, g, k);
    dwfk::ld_site<P>(b, y, g, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) acc += (double)a[c].re * b[c].re + (double)a[c].im * b[c].im;
  }
  const double t = dwfk::block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

}  // namespace ptck
