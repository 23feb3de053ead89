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
// B[8(l/16)..+7][l%16] and C[4(l/16)..+3][l%16]).  The four waves of a
// work-group split the output blocks as a 2 x 2 grid (4 x 1 for the forward
// layers with 16 output channels): each wave keeps the A fragments of its
// block rows and the B fragments of its block columns in registers for one
// K step and issues every product between them, so each fragment is read
// once per step, not once per output block.  The forward
// layer reads its weight fragments straight from memory (pre-arranged in
// fragment order) and only the transported field goes through LDS.  The
// weight gradient keeps both operands as [col][row] tiles (written with 8-byte
// stores, like the forward's B tile) and reads its fragments with gfx950's
// transposing LDS loads (ds_read_b64_tr_b16).
// Transport: the links are staged as packed bf16 pairs so that each complex
// multiply-add of U phi is two v_dot2_f32_bf16 on the activation words as
// stored (FP32 accumulation).
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

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint32_t b16) { return __uint_as_float(b16 << 16); }
__device__ __forceinline__ uint32_t pack_c(cx<float> v) { return dwfk::bf16_pair(v.re, v.im); }
__device__ __forceinline__ cx<float> unpack_c(uint32_t w) {
  return {__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u)};
}

// 4D parity held by slice s of the CG half vector: DWF (s+1)%2, Mobius 1
__host__ __device__ __forceinline__ int slice_parity(int action, int s) {
  return action == 0 ? (s + 1) & 1 : 1;
}

// n unit Gaussians into an FP32 field (its reals in storage order): the
// counter-based construction of the host generator (SplitMix64 finaliser on
// a keyed counter, 53-bit uniforms, Box-Muller), evaluated with the device's
// FP64 log / sqrt / sincos, so that training pairs need no host vectors.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__global__ void __launch_bounds__(256) k_gauss_f32(float* __restrict__ out, size_t n,
                                                   uint64_t seed) {
  const uint64_t key = mix64(seed ^ 0xD1B54A32D192ED03ULL);
  const size_t npair = (n + 1) / 2;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < npair;
       k += (size_t)gridDim.x * blockDim.x) {
    const double u1 = ((mix64(key + 2 * k) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = ((mix64(key + 2 * k + 1) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(6.283185307179586476925286766559 * u2, &sn, &cs);
    out[2 * k] = (float)(r * cs);
    if (2 * k + 1 < n) out[2 * k + 1] = (float)(r * sn);
  }
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

// The tile's links as packed bf16 words for v_dot2_f32_bf16.  For the
// transport matrix M of network direction d (U_mu(x) for d = 1..4, U_mu(x-mu)^dag
// for d = 5..8), entry (a, b) becomes two words, R = (Re M_ab, -Im M_ab) and
// I = (Im M_ab, Re M_ab) (low half first), so that for a colour vector f held as
// activations (f_b = (Re f_b, Im f_b)):
//   Re (M f)_a = sum_b dot2(R_ab, f_b),   Im (M f)_a = sum_b dot2(I_ab, f_b).
// Each bf16 x bf16 product is exact in FP32.  The thread that staged link w
// (stage_links: thread w stages link w of the tile, w = direction * NT +
// site) rewrites its 9 complex entries in place, 18 words in the same 72
// bytes, so no barrier is needed in between (only_dir >= 0: only the NT
// links of that direction, staged compactly from entry 0, so entry
// threadIdx.x - only_dir * NT).
__device__ __forceinline__ void pack_links_dot2(cx<float>* lu, int only_dir = -1) {
  const int w = (int)threadIdx.x - (only_dir >= 0 ? only_dir * NT : 0);
  if (w < 0 || w >= (only_dir >= 0 ? NT : 8 * NT)) return;
  cx<float> u[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) u[e] = lu[w * 9 + e];
  // staged direction 4..7: backward links
  const bool dag = only_dir >= 0 ? only_dir >= 4 : w >= 4 * NT;
  uint32_t* q = reinterpret_cast<uint32_t*>(lu + w * 9);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const cx<float> m = dag ? cx<float>{u[b * 3 + a].re, -u[b * 3 + a].im} : u[a * 3 + b];
      q[(a * 3 + b) * 2] = dwfk::bf16_pair(m.re, -m.im);
      q[(a * 3 + b) * 2 + 1] = dwfk::bf16_pair(m.im, m.re);
    }
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float dot2(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, a),
                                         __builtin_bit_cast(bf16x2, b), c, false);
}

// Transport of the 4 spins of one channel: f[sp * 3 + b] (activation words),
// lw the 18 packed link words of (d, site) (unused for d = 0).  Out: for
// colour a, re[a][h] / im[a][h] = bf16 pairs of spins (2h, 2h + 1).
__device__ __forceinline__ void transport4(const uint32_t* lw, int d, const uint32_t* f,
                                           uint32_t (&re)[3][2], uint32_t (&im)[3][2]) {
  if (d == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t x = f[(2 * h) * 3 + a], y = f[(2 * h + 1) * 3 + a];
        re[a][h] = (x & 0xffffu) | (y << 16);
        im[a][h] = (x >> 16) | (y & 0xffff0000u);
      }
    return;
  }
  uint32_t L[18];
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const uint2 v = *reinterpret_cast<const uint2*>(lw + 2 * e);
    L[2 * e] = v.x;
    L[2 * e + 1] = v.y;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float vr[2], vi[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const uint32_t* fs = f + (2 * h + k) * 3;
        vr[k] = dot2(L[(a * 3 + 2) * 2], fs[2],
                     dot2(L[(a * 3 + 1) * 2], fs[1], dot2(L[(a * 3) * 2], fs[0], 0.f)));
        vi[k] = dot2(L[(a * 3 + 2) * 2 + 1], fs[2],
                     dot2(L[(a * 3 + 1) * 2 + 1], fs[1], dot2(L[(a * 3) * 2 + 1], fs[0], 0.f)));
      }
      re[a][h] = dwfk::bf16_pair(vr[0], vr[1]);
      im[a][h] = dwfk::bf16_pair(vi[0], vi[1]);
    }
}

// site of the activation neighbour of tile site n along d (S index)
// (lslot: the slot of the staged neighbour list holding direction d, d - 1
// when all eight directions are staged)
__device__ __forceinline__ size_t nbr_site(const int* lnb, int d, int n, int par, int i0,
                                           const Geom& g, int lslot) {
  if (d == 0) return (size_t)par * g.vh + i0 + n;
  return (size_t)(1 - par) * g.vh + lnb[lslot * NT + n];
}

// (site, channel) items of a tile spread over the work-group: PER per thread
template <int CIN>
struct Items {
  static constexpr int N = NT * CIN, PER = (N + NTHR - 1) / NTHR;
  static constexpr bool EXACT = N % NTHR == 0;
};

// Issue the loads of this thread's items of direction d: raw[j] holds the 12
// activation words (4 spins x 3 colours) of channel i at the neighbour of
// tile site n.  The neighbour list of the tile must be staged unless d = 0.
template <int CIN>
__device__ __forceinline__ void gather_items(u32x4 (&raw)[Items<CIN>::PER][3],
                                             const uint32_t* __restrict__ in, const int* lnb,
                                             int d, int par, int i0, const Geom& g,
                                             int lslot = -1) {
  if (lslot < 0) lslot = d - 1;
#pragma unroll
  for (int j = 0; j < Items<CIN>::PER; ++j) {
    const int w = threadIdx.x + j * NTHR;
    if (Items<CIN>::EXACT || w < Items<CIN>::N) {
      const int n = w / CIN, i = w - n * CIN;
      const size_t y = nbr_site(lnb, d, n, par, i0, g, lslot);
      const u32x4* src = reinterpret_cast<const u32x4*>(in + (y * CIN + i) * 12);
      raw[j][0] = src[0];
      raw[j][1] = src[1];
      raw[j][2] = src[2];
    }
  }
}

// Transport the gathered items of direction d and store them into LDS as
// bf16 in [col][K'] (row stride RS; the B operand of the forward GEMM
// and, read through the transposing loads, the Y operand of the weight
// gradient), with k' = part*(CIN*4) + i*4 + spin and col = n*3 + colour.  The
// link of an item is read from LDS once for its 4 spins, and the 4 spins of a
// colour's real (imaginary) parts are one 8-byte store.
template <int CIN, int RS = 8 * CIN + PADK>
__device__ __forceinline__ void store_transported(uint16_t* sB,
                                                  const u32x4 (&raw)[Items<CIN>::PER][3],
                                                  const uint32_t* lw, int d, int lslot = -1) {
  if (lslot < 0) lslot = d > 0 ? d - 1 : 0;
#pragma unroll
  for (int j = 0; j < Items<CIN>::PER; ++j) {
    const int w = threadIdx.x + j * NTHR;
    if (!Items<CIN>::EXACT && w >= Items<CIN>::N) continue;
    const int n = w / CIN, i = w - n * CIN;
    const uint32_t f[12] = {raw[j][0][0], raw[j][0][1], raw[j][0][2], raw[j][0][3],
                            raw[j][1][0], raw[j][1][1], raw[j][1][2], raw[j][1][3],
                            raw[j][2][0], raw[j][2][1], raw[j][2][2], raw[j][2][3]};
    uint32_t re[3][2], im[3][2];
    transport4(lw + (lslot * NT + n) * 18, d, f, re, im);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int col = n * 3 + a;
      *reinterpret_cast<uint2*>(sB + col * RS + i * 4) = uint2{re[a][0], re[a][1]};
      *reinterpret_cast<uint2*>(sB + col * RS + CIN * 4 + i * 4) = uint2{im[a][0], im[a][1]};
    }
  }
}

// 8 consecutive bf16 (16-byte aligned: one LDS load)
__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}
// Output rows of the forward layer kernel: M' = 8*COUT rows in MB = M'/16
// blocks, the real parts (rows o*4 + spin) in blocks [0, MB/2), the imaginary
// parts in [MB/2, MB).  When MB is a multiple of 4 ("paired"), the wave pair
// h = 0, 1 takes real blocks [h*MB/4, (h+1)*MB/4) and the imaginary blocks of
// the same channels, so a lane holds both halves of its output words and
// stores them from registers; otherwise the C tile goes through LDS.
template <int COUT>
constexpr bool fwd_paired() { return (8 * COUT / 16) % 4 == 0; }

// LDS bytes of the forward kernel: the transported tile [NCOL][K'] (+ the C
// tile [M'][NCOL] in FP32 when not paired, in the same bytes) and the links
__host__ __device__ constexpr size_t fwd_lds_links() { return (size_t)8 * NT * 9 * 8 + 8 * NT * 4; }
template <int CIN, int COUT>
constexpr size_t fwd_lds() {
  constexpr size_t b = (size_t)NCOL * (8 * CIN + PADK) * 2;
  constexpr size_t c = fwd_paired<COUT>() ? 0 : (size_t)(8 * COUT) * NCOL * 4;
  return (b > c ? b : c) + fwd_lds_links();
}

// Row block of local index i of wave pair h in the forward kernel (MB when
// that wave has none): paired shapes take real blocks [h*MB/4, (h+1)*MB/4)
// and the imaginary blocks of the same channels, others a contiguous range.
// With MB = 8 (16 output channels) the four waves are a 4 x 1 grid instead:
// wave w owns real block w and imaginary block 4 + w over all column blocks
// (h is then the wave index).  Each wave then holds a quarter of the weight
// fragments instead of half (32 VGPRs at K' = 128) and reads the whole B
// tile from LDS instead of half of it; the kernel fits three work-groups per
// CU and runs 11 % faster than on the 2 x 2 grid.
template <int MB>
constexpr bool fwd_g4() { return MB == 8; }
template <int RM, int MB>
__device__ __forceinline__ int fwd_block(int i, int h, bool paired) {
  if (fwd_g4<MB>()) return i == 0 ? h : MB / 2 + h;
  if (paired) return i < MB / 4 ? h * (MB / 4) + i : MB / 2 + h * (MB / 4) + (i - MB / 4);
  return h * RM + i < MB ? h * RM + i : MB;
}

// The weight fragments of direction d for this wave's row blocks
template <int RM, int K32, int MB>
__device__ __forceinline__ void load_frags(u32x4 (&af)[RM][K32], const u32x4* __restrict__ A,
                                           int d, int h, int lane, bool paired) {
#pragma unroll
  for (int i = 0; i < RM; ++i) {
    const int blk = fwd_block<RM, MB>(i, h, paired);
    if (blk < MB)
#pragma unroll
      for (int kb = 0; kb < K32; ++kb) af[i][kb] = A[((d * MB + blk) * K32 + kb) * 64 + lane];
  }
}

// One layer on one tile of NT sites per work-group:
// out[S][o][spin][colour] = sum_d W_d T_d in.
// A: the real forms [9][M'][K'] in MFMA fragment order, one 16-byte word per
// (d, row block, K step of 32, lane): A[d][16*blk + l%16][32*kb + 8*(l/16) ..
// + 7] at ((d*MB + blk)*K32 + kb)*64 + l, so that each wave loads its
// fragments straight from memory (no LDS stage; every work-group reads the
// same weights, 32 KB per direction at K' = M' = 128, which stay in L2).
// grid.x = 2 * vh / NT tiles (first all even-parity tiles, then odd).
// On the 2 x 2 grid a direction's weight fragments in registers (64 VGPRs
// at K' = M' = 128) hold the kernel to two waves per SIMD, as does the LDS
// of the unpaired shapes; the 4 x 1 grid of the 16-channel layers fits
// three.  Pipelining the transport of the next direction into a second B
// tile, or loading the fragments one K step ahead, timed the same or slower.
template <int CIN, int COUT>
__global__ void __launch_bounds__(NTHR, fwd_g4<8 * COUT / 16>() ? 3 : 2)
k_ptc_fwd(uint32_t* __restrict__ out, const uint32_t* __restrict__ in,
          const u32x4* __restrict__ A, const typename dwfk::P32::S* __restrict__ U, Geom g) {
  constexpr int MP = 8 * COUT, KP = 8 * CIN;
  constexpr int MB = MP / 16, NB = NCOL / 16;
  constexpr int K32 = KP / 32;
  constexpr bool PAIRED = fwd_paired<COUT>();
  static_assert(KP % 32 == 0 && NB % 2 == 0 && ((KP + PADK) * 2) % 16 == 0, "tile shapes");
  // this wave's output blocks: RM row blocks x columns [nb0, nb0 + RN)
  constexpr bool G4 = fwd_g4<MB>();
  constexpr int RM = G4 ? 2 : PAIRED ? MB / 2 : (MB + 1) / 2, RH = G4 ? 1 : MB / 4;
  constexpr int RN = G4 ? NB : NB / 2;
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t* sB = reinterpret_cast<uint16_t*>(smem);
  float* sC = reinterpret_cast<float*>(smem);  // after the last GEMM (not paired)
  constexpr size_t big = fwd_lds<CIN, COUT>() - fwd_lds_links();
  cx<float>* lu = reinterpret_cast<cx<float>*>(smem + big);
  int* lnb = reinterpret_cast<int*>(smem + big + (size_t)8 * NT * 9 * 8);

  // tile: each XCD sweeps a contiguous range, so neighbouring tiles share
  // its L2
  const int tiles_per_par = g.vh / NT, tile = dwfk::xcd_block(g.nxcd);
  const int par = tile / tiles_per_par;
  const int i0 = (tile - par * tiles_per_par) * NT;
  dwfk::stage_links<dwfk::P32>(lu, lnb, U, g, i0 / NT, NT, 1, par);
  pack_links_dot2(lu);
  const uint32_t* lw = reinterpret_cast<const uint32_t*>(lu);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = G4 ? wave : wave & 1;
  const int nb0 = G4 ? 0 : (wave >> 1) * RN;
  auto blk_of = [&](int i) { return fwd_block<RM, MB>(i, h, PAIRED); };
  f32x4 acc[RM][RN];
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // Per direction: the weight fragments are loaded first (their latency is
  // covered by the transport), the next direction's neighbour activations
  // after the B tile is built (covered by the products).
  u32x4 raw[Items<CIN>::PER][3];
  gather_items<CIN>(raw, in, lnb, 0, par, i0, g);  // d = 0: own sites
  for (int d = 0; d < NDIR; ++d) {
    u32x4 af[RM][K32];
    load_frags<RM, K32, MB>(af, A, d, h, lane, PAIRED);
    __syncthreads();  // links staged / previous direction's B tile consumed
    store_transported<CIN>(sB, raw, lw, d);
    __syncthreads();
    if (d + 1 < NDIR) gather_items<CIN>(raw, in, lnb, d + 1, par, i0, g);
    __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the products
    const uint16_t* pb = sB + (nb0 * 16 + (lane & 15)) * (KP + PADK);
#pragma unroll
    for (int kb = 0; kb < K32; ++kb) {
      const int k0 = kb * 32 + 8 * (lane >> 4);
      bf16x8 bfr[RN];
#pragma unroll
      for (int j = 0; j < RN; ++j) bfr[j] = ld8(pb + j * 16 * (KP + PADK) + k0);
#pragma unroll
      for (int i = 0; i < RM; ++i) {
        if (blk_of(i) < MB) {
          const bf16x8 a8 = __builtin_bit_cast(bf16x8, af[i][kb]);
#pragma unroll
          for (int j = 0; j < RN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a8, bfr[j], acc[i][j], 0, 0, 0);
        }
      }
    }
  }
  // C: lane l holds rows 4(l/16)..+3 of column l%16 of each block
  if constexpr (PAIRED) {
    // out[S][o][spin][colour] = (re row o*4+spin, im row COUT*4 + o*4+spin)
#pragma unroll
    for (int i = 0; i < RH; ++i) {
      const int o = (h * RH + i) * 4 + (lane >> 4);
#pragma unroll
      for (int j = 0; j < RN; ++j) {
        const int col = (nb0 + j) * 16 + (lane & 15), n = col / 3, c = col - n * 3;
        uint32_t* dst = out + (((size_t)par * g.vh + i0 + n) * COUT + o) * 12 + c;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          dst[r * 3] = pack_c(cx<float>{acc[i][j][r], acc[i + RH][j][r]});
      }
    }
  } else {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RM; ++i) {
      const int blk = blk_of(i);
      if (blk < MB) {
#pragma unroll
        for (int j = 0; j < RN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            sC[(blk * 16 + 4 * (lane >> 4) + r) * NCOL + (nb0 + j) * 16 + (lane & 15)] =
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
}

// Weight gradient LDS: the [col][row] tiles of X^T and Y^T (row padding
// PADW = 4 bf16 keeps three work-groups within a CU's 160 KB; the
// transposing loads showed no bank-conflict cost against a padding of 8 or
// 16) and the links and neighbours of the one direction it transports
constexpr int PADW = 4;
// work-groups per CU the weight-gradient kernel is built for: three (14 %
// faster than two with the next tile's output gradient prefetched), except
// the 12-channel shapes, whose registers do not fit three without spills
template <int CIN, int COUT>
constexpr int gw_occ() { return CIN != 12 && COUT != 12 ? 3 : 2; }
template <int CIN, int COUT>
constexpr size_t gw_lds() {
  constexpr size_t x = (size_t)NCOL * (8 * COUT + PADW) * 2;
  constexpr size_t y = (size_t)NCOL * (8 * CIN + PADW) * 2;
  return x + y + (size_t)NT * 9 * 8 + NT * 4;
}

// One MFMA operand fragment (8 bf16: rows r0 + l%16, columns c0 + 8(l/16) ..
// + 7 of the contraction) from a [col][row] LDS tile with row stride RS, by
// two transposing reads (ds_read_b64_tr_b16).  Within each 16-lane group,
// lane j receives in slot e element j%4 of the 8-byte chunk addressed by lane
// 4e + j/4, so lane i addresses column c0 + 8(l/16) + 4h + i/4 at rows
// r0 + 4(i%4) .. + 3 for the half h of the 8 columns.
typedef short v4i16 __attribute__((ext_vector_type(4)));
template <int RS>
__device__ __forceinline__ bf16x8 ld_frag_tr(const uint16_t* t, int r0, int c0, int lane) {
  const int i = lane & 15, c = c0 + 8 * (lane >> 4) + (i >> 2), r = r0 + 4 * (i & 3);
  typedef __attribute__((address_space(3))) v4i16 lds_v4i16;
  const v4i16 a = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4i16*)(t + c * RS + r));
  const v4i16 b = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4i16*)(t + (c + 4) * RS + r));
  typedef short v8i16 __attribute__((ext_vector_type(8)));
  return __builtin_bit_cast(bf16x8, v8i16{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]});
}

// The output gradient of one tile (contiguous in gout: NT sites x COUT
// channels x 12 words), one (site, channel) item of 12 words (three 16-byte
// loads) at a time: PER per thread
template <int COUT>
struct XTile {
  static constexpr int N = NT * COUT, PER = (N + NTHR - 1) / NTHR;
  static constexpr bool EXACT = N % NTHR == 0;
};
template <int COUT>
__device__ __forceinline__ void load_x(u32x4 (&xv)[XTile<COUT>::PER][3],
                                       const uint32_t* __restrict__ gout, int par, int i0,
                                       const Geom& g) {
  const u32x4* src = reinterpret_cast<const u32x4*>(gout + ((size_t)par * g.vh + i0) * (COUT * 12));
#pragma unroll
  for (int j = 0; j < XTile<COUT>::PER; ++j) {
    const int w = threadIdx.x + j * NTHR;
    if (XTile<COUT>::EXACT || w < XTile<COUT>::N) {
      xv[j][0] = src[3 * w];
      xv[j][1] = src[3 * w + 1];
      xv[j][2] = src[3 * w + 2];
    }
  }
}

// Weight gradient partials part[d][w][M'][K'] = sum X[M'][cols] Y_d[K'][cols]^T
// over the tiles of work-group w of direction d.  The nine work-groups of
// one w run on one XCD at the same time (work-group L goes to XCD L % nxcd),
// so the output-gradient tile and the neighbour fields they share are read
// into that XCD's L2 once; each XCD sweeps its own contiguous range of
// tiles, nwg / nxcd of them at a time.  A grid of 9 nwg work-groups; with
// nwg not a multiple of nxcd, one range over all tiles.
template <int CIN, int COUT>
__global__ void __launch_bounds__(NTHR, (gw_occ<CIN, COUT>()))
k_ptc_gw(float* __restrict__ part, int nwg, const uint32_t* __restrict__ gout,
         const uint32_t* __restrict__ in, const typename dwfk::P32::S* __restrict__ U, Geom g) {
  constexpr int MP = 8 * COUT, KP = 8 * CIN;
  constexpr int MB = MP / 16, KB2 = KP / 16, CB = NCOL / 32;  // contraction over columns
  constexpr int XS = MP + PADW, YS = KP + PADW;  // row strides of the [col][row] tiles
  static_assert(NCOL % 32 == 0 && (XS * 2) % 8 == 0 && (YS * 2) % 8 == 0, "tile shapes");
  // this wave's output blocks: rows [mb0, mb0 + RM) x columns [kb0, kb0 + RK)
  constexpr int RM = (MB + 1) / 2, RK = (KB2 + 1) / 2;
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t* sX = reinterpret_cast<uint16_t*>(smem);  // X^T: [col][M']
  uint16_t* sY = sX + NCOL * XS;                     // Y^T: [col][K']
  cx<float>* lu = reinterpret_cast<cx<float>*>(smem + (size_t)NCOL * (XS + YS) * 2);
  int* lnb = reinterpret_cast<int*>(reinterpret_cast<unsigned char*>(lu) + (size_t)NT * 9 * 8);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_per_par = g.vh / NT, ntiles = 2 * tiles_per_par;
  const int nx = nwg % g.nxcd == 0 ? g.nxcd : 1, m = nwg / nx;
  const int x = blockIdx.x % nx, slot = blockIdx.x / nx;
  const int d = slot % NDIR, k = slot / NDIR;  // k < m
  const int t_end = (int)((long)(x + 1) * ntiles / nx);
  const int t_first = (int)((long)x * ntiles / nx) + k;
  const int mb0 = (wave & 1) * RM, kb0 = (wave >> 1) * RK;
  f32x4 acc[RM][RK];
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RK; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // X of the next tile is loaded into registers before this tile's products
  // (PFX), or at the start of the tile when registers are short
  constexpr bool PFX = gw_occ<CIN, COUT>() == 2;
  u32x4 xv[XTile<COUT>::PER][3];
  if (PFX && t_first < t_end)
    load_x<COUT>(xv, gout, t_first / tiles_per_par,
                 (t_first - t_first / tiles_per_par * tiles_per_par) * NT, g);
  for (int t = t_first; t < t_end; t += m) {
    const int par = t / tiles_per_par;
    const int i0 = (t - par * tiles_per_par) * NT;
    if (!PFX) load_x<COUT>(xv, gout, par, i0, g);
    __syncthreads();  // previous tile's operands consumed
    // only direction d's links and neighbours are read (none for d = 0),
    // staged compactly in slot 0
    if (d > 0) {
      dwfk::stage_links<dwfk::P32>(lu, lnb, U, g, i0 / NT, NT, 1, par, d - 1, (d - 1) * NT);
      pack_links_dot2(lu, d - 1);
    }
    // X^T of the tile: item (site n, channel o) holds words spin*3 + colour;
    // the 4 spins of a colour are 4 consecutive rows o*4 + spin (real parts)
    // and COUT*4 + o*4 + spin (imaginary parts): two 8-byte stores per colour
#pragma unroll
    for (int j = 0; j < XTile<COUT>::PER; ++j) {
      const int w = threadIdx.x + j * NTHR;
      if (!XTile<COUT>::EXACT && w >= XTile<COUT>::N) continue;
      const int n = w / COUT, o = w - n * COUT;
      const uint32_t f[12] = {xv[j][0][0], xv[j][0][1], xv[j][0][2], xv[j][0][3],
                              xv[j][1][0], xv[j][1][1], xv[j][1][2], xv[j][1][3],
                              xv[j][2][0], xv[j][2][1], xv[j][2][2], xv[j][2][3]};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        uint16_t* row = sX + (n * 3 + c) * XS;
        *reinterpret_cast<uint2*>(row + o * 4) =
            uint2{(f[c] & 0xffffu) | (f[3 + c] << 16), (f[6 + c] & 0xffffu) | (f[9 + c] << 16)};
        *reinterpret_cast<uint2*>(row + COUT * 4 + o * 4) =
            uint2{(f[c] >> 16) | (f[3 + c] & 0xffff0000u), (f[6 + c] >> 16) | (f[9 + c] & 0xffff0000u)};
      }
    }
    __syncthreads();  // links staged
    {
      u32x4 raw[Items<CIN>::PER][3];
      gather_items<CIN>(raw, in, lnb, d, par, i0, g, 0);
      store_transported<CIN, YS>(sY, raw, reinterpret_cast<const uint32_t*>(lu), d, 0);
    }
    __syncthreads();
    {
      const int tn = t + m;
      if (PFX && tn < t_end)
        load_x<COUT>(xv, gout, tn / tiles_per_par, (tn - tn / tiles_per_par * tiles_per_par) * NT, g);
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the products
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      bf16x8 yf[RK];
#pragma unroll
      for (int j = 0; j < RK; ++j)
        if (kb0 + j < KB2) yf[j] = ld_frag_tr<YS>(sY, (kb0 + j) * 16, cb * 32, lane);
#pragma unroll
      for (int i = 0; i < RM; ++i) {
        if (mb0 + i < MB) {
          const bf16x8 xf = ld_frag_tr<XS>(sX, (mb0 + i) * 16, cb * 32, lane);
#pragma unroll
          for (int j = 0; j < RK; ++j)
            if (kb0 + j < KB2)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xf, yf[j], acc[i][j], 0, 0, 0);
        }
      }
    }
  }
  float* dst = part + ((size_t)d * nwg + x * m + k) * MP * KP;
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
