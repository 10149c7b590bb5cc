// denoise.hip -- the learned denoiser: a 3-level U-Net (32/64/128 channels,
// 3x3 convolutions, ReLU, 2x2 average pooling, nearest upsampling, skip
// concatenation) that maps a 1-sample render plus its first-hit features to
// the converged image.  It stands where the reference runs its median
// filter (src/gpu/median.glsl:51-75), with no reference counterpart.
//
// Every convolution -- forward, backward to the input and backward to the
// weights -- is an implicit GEMM on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16): BF16 operands, FP32 accumulation.  Activations
// are NHWC BF16, gradients FP32, master weights and Adam moments FP32.
//
// MFMA 16x16x32 fragment layout (lane l, register r): A[l%16][8*(l/16)+r],
// B[8*(l/16)+r][l%16] (8 bf16 each), D[4*(l/16)+r][l%16] (4 f32).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/convector.h"
#include "device_math.hpp"
#include "host.hpp"

using cvh::DevBuf;
using cvh::fail;
using cvh::hip_fail;

namespace cvd {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
  // round to nearest even (NaN kept quiet)
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));
// Two floats to bf16 (round to nearest even) in one v_cvt_pk_bf16_f32:
// a in the low half.
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{a, b}, bf2));
}


// ---------------------------------------------------------------------------
// Implicit-GEMM 3x3 convolution (stride 1, zero padding 1), NHWC.
//
// out[p][j] = sum_k A[p][k] * B[k][j], k = tap * kc + c over 9 taps and kc
// input channels.  Forward: A = input at p + offset(tap), B[k][j] = W[j][k]
// (weights stored [cout][kpad]).  Backward to the input (flip): A = dZ at
// p - offset(tap), B[k][j] = Wt[j][k] (stored [cin][kpad], k = tap * cout +
// co).  k beyond 9 * kc reads zeros.  Requirements (checked by the host):
// kc a multiple of 8, channel strides multiples of 8, kpad = roundup(9 *
// kc, 32).
// ---------------------------------------------------------------------------
struct ConvArgs {
  const uint16_t* x;  // A source (bf16), pixel stride xcs
  uint32_t xcs, kc;
  const uint16_t* w;  // B source, [NT*16 rows][kpad]
  uint32_t kpad;
  const float* bias;  // or null
  uint32_t nout;      // valid outputs (<= NT*16)
  uint16_t* y;        // bf16 output (pixel stride ycs) or null
  uint32_t ycs;
  float* yf;          // f32 output (pixel stride yfcs) or null
  uint32_t yfcs;
  uint32_t accumulate;  // yf += instead of =
  uint32_t relu;
  uint32_t flip;        // backward to the input: offsets mirrored
  uint32_t npix, h, w_;
  // non-null: the ReLU gradient of the layer whose output `mask` (bf16,
  // pixel stride mcs) is: outputs where mask <= 0 become 0 (written to y)
  const uint16_t* mask;
  uint32_t mcs;
  // non-null (k_conv_halo<..., UP>): input channels [0, 32 * up_chunks) are
  // the nearest 2x upsampling of xup (bf16, pixel stride xupcs, h / 2 x
  // w / 2 per image), read from it directly; the rest come from x
  const uint16_t* xup;
  uint32_t xupcs, up_chunks;
  // non-null (k_conv_halo<..., POOL>): the 2 x 2 average of the bf16
  // outputs y is written here too (pixel stride poolcs, h / 2 x w / 2 per
  // image), summed in the order of k_pool
  uint16_t* pool;
  uint32_t poolcs;
  // non-null (k_conv_halo<..., UG>, an input gradient): outputs [0, ug_ch)
  // are the gradient of a nearest 2x upsampling; their 2 x 2 sums (in the
  // order of k_upsample_grad), masked by the ReLU output ug_mask of the
  // layer below (pixel stride ug_mcs, h / 2 x w / 2), are written as that
  // layer's bf16 dZ (pixel stride ug_dzcs); the other outputs go to yf
  uint16_t* ug_dz;
  uint32_t ug_dzcs;
  const uint16_t* ug_mask;
  uint32_t ug_mcs, ug_ch;
};

// The convolution with the operands staged in LDS: a block of 4 waves
// computes 128 pixels x NT*16 outputs, each wave 32 pixels (two 16-row A
// fragments, so every B fragment feeds two MFMAs).  Per 32-wide K chunk the
// block gathers the input tile (128 pixels x 32 k, tap offsets and zero
// padding applied) and the weight tile (NT*16 rows x 32 k) into LDS with
// 16-byte loads, double buffered: the next chunk's global loads are in
// flight while the current chunk's MFMAs run.  Rows are padded to 40 bf16
// (80 B) so that the 16 rows of a fragment read fall in distinct banks.
constexpr uint32_t kConvPix = 128;
constexpr uint32_t kRow = 40;

template <int NT>
__global__ __launch_bounds__(256) void k_conv_lds(ConvArgs a) {
  constexpr uint32_t NB = (NT * 64 + 255) / 256;  // weight pieces per thread
  __shared__ __attribute__((aligned(16))) uint16_t As[2][kConvPix * kRow];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][NT * 16 * kRow];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t pblock = blockIdx.x * kConvPix;
  const uint32_t hw = a.h * a.w_;
  const uint32_t kvalid = 9 * a.kc, nchunks = a.kpad / 32;
  // the two input pieces of this thread: piece q = pixel q / 4, k offset (q % 4) * 8
  uint32_t abase[2];
  int ay[2], ax[2];
  bool aok[2];
  for (int i = 0; i < 2; i++) {
    const uint32_t q = tid + 256 * i;
    const uint32_t p = pblock + q / 4;
    aok[i] = p < a.npix;
    const uint32_t pp = aok[i] ? p : 0u;
    const uint32_t b = pp / hw, rem = pp % hw;
    ay[i] = (int)(rem / a.w_);
    ax[i] = (int)(rem % a.w_);
    abase[i] = b * hw;
  }
  static_assert(NB <= 3, "at most 3 weight pieces per thread");
  uint4 ra0, ra1, rb0, rb1, rb2;
  // pieces of chunk c into registers (weight pieces clamped, so that every
  // thread loads; the unused ones are not stored)
  auto a_piece = [&](int i, uint32_t c) -> uint4 {
    const uint32_t q = tid + 256 * i;
    const uint32_t k = c * 32 + (q % 4) * 8;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (aok[i] && k < kvalid) {
      const uint32_t tap = k / a.kc, ch = k % a.kc;
      int dy = (int)(tap / 3) - 1, dx = (int)(tap % 3) - 1;
      if (a.flip) {
        dy = -dy;
        dx = -dx;
      }
      const int sy = ay[i] + dy, sx = ax[i] + dx;
      if (sy >= 0 && sy < (int)a.h && sx >= 0 && sx < (int)a.w_)
        v = *reinterpret_cast<const uint4*>(
            a.x + ((size_t)abase[i] + (uint32_t)sy * a.w_ + (uint32_t)sx) * a.xcs + ch);
    }
    return v;
  };
  auto b_piece = [&](uint32_t j, uint32_t c) -> uint4 {
    const uint32_t q = min(tid + 256 * j, (uint32_t)(NT * 64 - 1));
    return *reinterpret_cast<const uint4*>(a.w + (size_t)(q / 4) * a.kpad + c * 32 + (q % 4) * 8);
  };
  auto load_chunk = [&](uint32_t c) {
    ra0 = a_piece(0, c);
    ra1 = a_piece(1, c);
    rb0 = b_piece(0, c);
    if (NB > 1) rb1 = b_piece(1, c);
    if (NB > 2) rb2 = b_piece(2, c);
  };
  auto put_b = [&](uint32_t buf, uint32_t j, uint4 v) {
    const uint32_t q = tid + 256 * j;
    if (q < NT * 64) *reinterpret_cast<uint4*>(&Bs[buf][(q / 4) * kRow + (q % 4) * 8]) = v;
  };
  auto store_chunk = [&](uint32_t buf) {
    *reinterpret_cast<uint4*>(&As[buf][(tid / 4) * kRow + (tid % 4) * 8]) = ra0;
    *reinterpret_cast<uint4*>(&As[buf][((tid + 256) / 4) * kRow + (tid % 4) * 8]) = ra1;
    put_b(buf, 0, rb0);
    if (NB > 1) put_b(buf, 1, rb1);
    if (NB > 2) put_b(buf, 2, rb2);
  };
  f4 acc[2][NT];
#pragma unroll
  for (int f = 0; f < 2; f++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[f][t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  load_chunk(0);
  store_chunk(0);
  __syncthreads();
  const uint32_t arow = wave * 32 + (lane & 15), koff = 8 * (lane >> 4);
  for (uint32_t c = 0; c < nchunks; c++) {
    const uint32_t buf = c & 1;
    if (c + 1 < nchunks) load_chunk(c + 1);
    const bf8 A0 = *reinterpret_cast<const bf8*>(&As[buf][arow * kRow + koff]);
    const bf8 A1 = *reinterpret_cast<const bf8*>(&As[buf][(arow + 16) * kRow + koff]);
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const bf8 B = *reinterpret_cast<const bf8*>(&Bs[buf][(t * 16 + (lane & 15)) * kRow + koff]);
      acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B, acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B, acc[1][t], 0, 0, 0);
    }
    if (c + 1 < nchunks) store_chunk(buf ^ 1u);
    __syncthreads();
  }
#pragma unroll
  for (int f = 0; f < 2; f++) {
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const uint32_t j = t * 16 + (lane & 15);
      const bool jok = j < a.nout;
      const float bj = (a.bias && jok) ? a.bias[j] : 0.0f;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const uint32_t p = pblock + wave * 32 + f * 16 + 4 * (lane >> 4) + r;
        if (!jok || p >= a.npix) continue;
        float v = acc[f][t][r] + bj;
        if (a.relu) v = fmaxf(v, 0.0f);
        if (a.mask && !(from_bf16(a.mask[(size_t)p * a.mcs + j]) > 0.0f)) v = 0.0f;
        if (a.y) a.y[(size_t)p * a.ycs + j] = to_bf16(v);
        if (a.yf) {
          float* o = a.yf + (size_t)p * a.yfcs + j;
          *o = a.accumulate ? *o + v : v;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// The same convolution from LDS halo tiles, for images whose rows are a
// multiple of 32 pixels wide.  A block of 4 waves computes an 8 x 32 pixel
// tile for NS * 16 outputs (blockIdx.y picks the slice of NS output tiles):
// each wave two rows, 4 fragments of 16 pixels.  The input channels run in
// chunks of 32 (16 when KC == 16).  A chunk's tile with its 1-pixel halo
// (zeros outside the image) and the chunk's weights for all 9 taps sit in
// LDS while the next chunk's global loads are in flight into registers, so
// the MFMAs wait only on LDS.  Per tap a wave reads 4 input and NS weight
// fragments and runs 4 * NS MFMAs.  With KC == 16 a 32-wide K step covers
// two taps (lanes 0-31 the even tap, 32-63 the odd one, zeros for the
// tenth).  The weights are the MFMA's A operand and the pixels its B
// operand, so each lane ends with 4 consecutive output channels of one
// pixel and stores them with one 8-byte (bf16) or 16-byte (f32) access.
// FLIP mirrors the taps (backward to the input, weights Wt).
// When the input is one chunk (KC <= 32) nothing would be in flight during
// the MFMAs, so a block is persistent (unless the epilogue reads a mask):
// it loads the weights and bias once, then walks tiles blockIdx.x,
// blockIdx.x + gridDim.x, ..., loading the next tile's halo into registers
// while it computes the current one.
// ---------------------------------------------------------------------------
constexpr uint32_t kHaloW = 32, kHaloH = 8;

template <int KC, int NS, bool FLIP, bool MASK = false, bool UP = false, bool POOL = false,
          bool UG = false>
__global__ __launch_bounds__(256, 2) void k_conv_halo(ConvArgs a) {
  constexpr int MF = 4;
  constexpr int HH = (int)kHaloH + 2, HW = (int)kHaloW + 2;
  constexpr int CC = KC == 16 ? 16 : 32;  // channels per chunk
  constexpr int NCH = KC / CC;
  constexpr int RS = CC + 8;  // LDS pixel stride (bf16): 16 bytes of padding
  constexpr int PP = CC / 8;  // 16-byte pieces per pixel and chunk
  constexpr int XP = HH * HW * PP;  // input pieces per chunk
  constexpr int KS = KC == 16 ? 5 : 9;  // 32-wide K steps per chunk
  constexpr int WR = NS * 16;           // weight rows of the slice
  constexpr int WS = 40;                // LDS weight row stride (bf16)
  constexpr int WP = KS * WR * 4;       // weight pieces per chunk
  constexpr int XPT = (XP + 255) / 256, WPT = (WP + 255) / 256;
  static_assert(KC % CC == 0, "channels per chunk");
  __shared__ __attribute__((aligned(16))) uint16_t xs[HH * HW * RS];
  __shared__ __attribute__((aligned(16))) uint16_t ws[KS * WR * WS];
  // (a masked epilogue reads the mask from global memory: behind the
  // prefetch, whose loads complete first, it would wait for all of them)
  constexpr bool PERSIST = NCH == 1 && !MASK;
  // persistent blocks: the slice's bias, read in the epilogue from LDS so
  // that it does not wait behind the next tile's global loads
  __shared__ float bias_s[PERSIST ? WR : 1];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t ntx = a.w_ / kHaloW, nty = (a.h + kHaloH - 1) / kHaloH;
  const uint32_t ntiles = (a.npix / (a.h * a.w_)) * ntx * nty;
  uint32_t tile = blockIdx.x;
  int x0 = 0, y0 = 0;
  size_t img_base = 0;
  const uint32_t row0 = blockIdx.y * WR;  // first output of the slice
  // this thread's input pieces: element offset of chunk 0, or ~0u outside
  // the image; its weight pieces: element offset of chunk 0
  uint32_t xoff[XPT], woff[WPT];
  // tile t's corner and its input pieces' offsets
  auto tile_offsets = [&](uint32_t t, int& tx0, int& ty0, size_t& tbase) {
    tx0 = (int)((t % ntx) * kHaloW);
    ty0 = (int)(((t / ntx) % nty) * kHaloH);
    tbase = (size_t)(t / (ntx * nty)) * a.h * a.w_;
#pragma unroll
    for (int i = 0; i < XPT; i++) {
      const int q = (int)tid + 256 * i;
      const int pix = q / PP, part = q % PP;
      const int sy = ty0 + pix / HW - 1, sx = tx0 + pix % HW - 1;
      xoff[i] = (q < XP && sy >= 0 && sy < (int)a.h && sx >= 0 && sx < (int)a.w_)
                    ? (uint32_t)((tbase + (size_t)sy * a.w_ + (uint32_t)sx) * a.xcs) + part * 8
                    : ~0u;
    }
  };
  tile_offsets(tile, x0, y0, img_base);
  // UP: offset in xup of input piece i of the current tile (inside the
  // image), recomputed at each load rather than held in registers
  auto up_offset = [&](int i) -> uint32_t {
    const int q = (int)tid + 256 * i;
    const int pix = q / PP, part = q % PP;
    const uint32_t sy = (uint32_t)(y0 + pix / HW - 1), sx = (uint32_t)(x0 + pix % HW - 1);
    return (uint32_t)((img_base / 4 + (size_t)(sy >> 1) * (a.w_ >> 1) + (sx >> 1)) * a.xupcs) +
           part * 8;
  };
#pragma unroll
  for (int i = 0; i < WPT; i++) {
    const int q = min((int)tid + 256 * i, WP - 1);
    const int st = q / (WR * 4), r = (q / 4) % WR, g = q % 4;
    // K step st of the chunk: tap st (KC >= 32) or taps 2 st, 2 st + 1 (KC == 16)
    const int k = KC == 16 ? st * 32 + 8 * g : st * KC + 8 * g;
    woff[i] = (row0 + (uint32_t)r) * a.kpad + (uint32_t)k;
  }
  u32x4 xr[XPT], wr[WPT];
  // Input pieces outside the image are loaded from the tensor's first
  // pixel (every load is unconditional, so that the compiler does not wait
  // for each one to merge it with a zero inside the load sequence) and
  // zeroed when stored to LDS.
  // (UP: the chunk's source, the lower level or the skip tensor, is chosen
  // by selects rather than a branch, for the same reason.)
#define CV_HALO_LOAD(c)                                                                   \
  {                                                                                       \
    const bool from_up = UP && (uint32_t)(c) < a.up_chunks;                               \
    const uint16_t* src = from_up ? a.xup : a.x;                                          \
    _Pragma("unroll") for (int i = 0; i < XPT; i++) xr[i] =                               \
        *reinterpret_cast<const u32x4*>(                                                  \
            src + (xoff[i] == ~0u ? 0u : from_up ? up_offset(i) : xoff[i]) + (c) * CC);   \
    _Pragma("unroll") for (int i = 0; i < WPT; i++) wr[i] =                               \
        *reinterpret_cast<const u32x4*>(a.w + woff[i] + (c) * 32);                        \
  }
#define CV_HALO_LOAD_X()                                                                  \
  {                                                                                       \
    _Pragma("unroll") for (int i = 0; i < XPT; i++) xr[i] =                               \
        *reinterpret_cast<const u32x4*>(a.x + (xoff[i] != ~0u ? xoff[i] : 0u));            \
  }
  // The scheduling barrier keeps the zeroing selects, which wait for the
  // loads, from being hoisted into the MFMAs that the loads overlap.
#define CV_HALO_STORE_X()                                                                 \
  {                                                                                       \
    __builtin_amdgcn_sched_barrier(0);                                                    \
    _Pragma("unroll") for (int i = 0; i < XPT; i++) {                                     \
      const int q = (int)tid + 256 * i;                                                   \
      if (q < XP)                                                                         \
        *reinterpret_cast<u32x4*>(&xs[(q / PP) * RS + (q % PP) * 8]) =                    \
            xoff[i] != ~0u ? xr[i] : u32x4{0u, 0u, 0u, 0u};                               \
    }                                                                                     \
  }
#define CV_HALO_STORE()                                                                   \
  {                                                                                       \
    CV_HALO_STORE_X();                                                                    \
    _Pragma("unroll") for (int i = 0; i < WPT; i++) {                                     \
      const int q = (int)tid + 256 * i;                                                   \
      if (q < WP) *reinterpret_cast<u32x4*>(&ws[(q / 4) * WS + (q % 4) * 8]) = wr[i];     \
    }                                                                                     \
  }
  const uint32_t col = lane & 15, kq = lane >> 4;
  // the wave's fragments: m -> row 2 * wave + m / 2, pixels 16 * (m % 2) + col
  const int wrow = 2 * (int)wave;
  const bf8 zero8 = {};
  CV_HALO_LOAD(0);
  CV_HALO_STORE();
  if constexpr (PERSIST) {
    if (tid < (uint32_t)WR)
      bias_s[tid] = (a.bias && row0 + tid < a.nout) ? a.bias[row0 + tid] : 0.0f;
  }
  __syncthreads();
  // Persistent blocks: a tile's halo is loaded two tiles ahead.  The loads
  // of tile t + 1 are issued before tile t - 1's epilogue and overlap it
  // and tile t's MFMAs; after those MFMAs they are stored to LDS and the
  // loads of tile t + 2 are issued, before tile t's epilogue (whose stores
  // they then do not wait behind).
  uint32_t next = tile + gridDim.x;
  bool more = PERSIST && next < ntiles;
  int nx0 = 0, ny0 = 0;
  size_t nbase = 0;
  if constexpr (PERSIST) {
    if (more) {
      tile_offsets(next, nx0, ny0, nbase);
      CV_HALO_LOAD_X();
    }
  }
  // lane: output channels row0 + t * 16 + 4 * kq + r (r = 0..3) of pixel
  // col of each fragment
  const bool vec_y = a.y && ((uintptr_t)a.y % 8 == 0) && (a.ycs % 4 == 0);
  const bool vec_f = a.yf && ((uintptr_t)a.yf % 16 == 0) && (a.yfcs % 4 == 0);
#pragma unroll 1
  for (;;) {
    f4 acc[MF][NS];
#pragma unroll
    for (int m = 0; m < MF; m++)
#pragma unroll
      for (int t = 0; t < NS; t++) acc[m][t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    // the MFMAs of the chunk in LDS
    auto chunk_mfma = [&]() {
#pragma unroll
      for (int st = 0; st < KS; st++) {
        // the lane's tap: with KC == 16 the upper 32 lanes take the odd one
        // (the tenth reads the centre and is zeroed)
        const int tp = KC == 16 ? 2 * st + (int)(kq >> 1) : st;
        const int tq = tp > 8 ? 4 : tp;
        int dy = tq / 3 - 1, dx = tq % 3 - 1;
        if (FLIP) {
          dy = -dy;
          dx = -dx;
        }
        const int coff = KC == 16 ? 8 * (int)(kq & 1) : 8 * (int)kq;
        bf8 Wf[NS];
#pragma unroll
        for (int t = 0; t < NS; t++)
          Wf[t] = *reinterpret_cast<const bf8*>(
              &ws[(st * WR + t * 16 + (int)col) * WS + 8 * kq]);
#pragma unroll
        for (int m = 0; m < MF; m++) {
          const int hy = wrow + m / 2 + 1 + dy, hx = 16 * (m % 2) + (int)col + 1 + dx;
          bf8 X = *reinterpret_cast<const bf8*>(&xs[(hy * HW + hx) * RS + coff]);
          if (KC == 16 && tp > 8) X = zero8;
#pragma unroll
          for (int t = 0; t < NS; t++)
            acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Wf[t], X, acc[m][t], 0, 0, 0);
        }
      }
    };
    // Every chunk but the last loads the next one during its MFMAs.  The
    // last is peeled off so that every path through the loop stores the
    // loaded registers: with the store under a condition, the compiler
    // takes the loads as possibly pending at the loop's head and waits
    // for all of them there, before the MFMAs they were meant to overlap.
#pragma unroll 1
    for (int c = 0; c + 1 < NCH; c++) {
      CV_HALO_LOAD(c + 1);
      // (the loads are issued before the MFMAs, not scheduled among them)
      __builtin_amdgcn_sched_barrier(0);
      chunk_mfma();
      __syncthreads();
      CV_HALO_STORE();
      __syncthreads();
    }
    chunk_mfma();
    uint32_t next2 = 0;
    bool more2 = false;
    int n2x0 = 0, n2y0 = 0;
    size_t n2base = 0;
    if constexpr (PERSIST) {
      __syncthreads();
      if (more) CV_HALO_STORE_X();
      __syncthreads();
      next2 = next + gridDim.x;
      more2 = more && next2 < ntiles;
      if (more2) {
        tile_offsets(next2, n2x0, n2y0, n2base);
        CV_HALO_LOAD_X();
      }
    }
    if constexpr (POOL) {
      // Fragments m (the upper row of the wave's pair) and m + 2 (the lower
      // row) together: both are stored, and the even lane of each column
      // pair, which holds pixel (2Y, 2X), takes its neighbour (2Y, 2X + 1)
      // by a DPP lane swap (quad_perm [1, 0, 3, 2]) and writes the 2 x 2
      // average of the stored bf16 values, summed in the order of k_pool.
      // (The host checks: y aligned for 8-byte stores, whole groups of 4.)
#pragma unroll
      for (int m = 0; m < 2; m++) {
        const int py = y0 + wrow;
        if (py >= (int)a.h) continue;
        const uint32_t px = (uint32_t)x0 + 16 * m + col;
        const size_t p = img_base + (size_t)py * a.w_ + px;
        const size_t q = img_base / 4 + (size_t)(py >> 1) * (a.w_ >> 1) + (px >> 1);
#pragma unroll
        for (int t = 0; t < NS; t++) {
          const uint32_t j0 = row0 + t * 16 + 4 * kq;
          if (j0 >= a.nout) continue;
          float u[4], l[4];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const float b = PERSIST ? bias_s[t * 16 + 4 * kq + r]
                                    : (a.bias ? a.bias[j0 + r] : 0.0f);
            u[r] = acc[m][t][r] + b;
            l[r] = acc[m + 2][t][r] + b;
            if (a.relu) {
              u[r] = fmaxf(u[r], 0.0f);
              l[r] = fmaxf(l[r], 0.0f);
            }
          }
          const uint2 pu = make_uint2(pack_bf16x2(u[0], u[1]), pack_bf16x2(u[2], u[3]));
          const uint2 pl = make_uint2(pack_bf16x2(l[0], l[1]), pack_bf16x2(l[2], l[3]));
          *reinterpret_cast<uint2*>(a.y + p * a.ycs + j0) = pu;
          *reinterpret_cast<uint2*>(a.y + (p + a.w_) * a.ycs + j0) = pl;
          const uint32_t wu[4] = {pu.x << 16, pu.x & 0xffff0000u, pu.y << 16, pu.y & 0xffff0000u};
          const uint32_t wl[4] = {pl.x << 16, pl.x & 0xffff0000u, pl.y << 16, pl.y & 0xffff0000u};
          float s4[4];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const float a0 = __uint_as_float(wu[r]), c0 = __uint_as_float(wl[r]);
            const float b0 = __int_as_float(
                __builtin_amdgcn_mov_dpp((int)wu[r], 0xB1, 0xF, 0xF, false));
            const float d0 = __int_as_float(
                __builtin_amdgcn_mov_dpp((int)wl[r], 0xB1, 0xF, 0xF, false));
            s4[r] = 0.25f * (((a0 + b0) + c0) + d0);
          }
          if ((col & 1) == 0)
            *reinterpret_cast<uint2*>(a.pool + q * a.poolcs + j0) =
                make_uint2(pack_bf16x2(s4[0], s4[1]), pack_bf16x2(s4[2], s4[3]));
        }
      }
    } else if (UG && row0 < a.ug_ch) {
      // the upsampling's gradient: the 2 x 2 sum of the f32 values (the
      // even lane of each column pair reads its neighbour by DPP, as in
      // the pooling above), masked by the layer below's ReLU output
#pragma unroll
      for (int m = 0; m < 2; m++) {
        const int py = y0 + wrow;
        if (py >= (int)a.h) continue;
        const uint32_t px = (uint32_t)x0 + 16 * m + col;
        const size_t q = img_base / 4 + (size_t)(py >> 1) * (a.w_ >> 1) + (px >> 1);
#pragma unroll
        for (int t = 0; t < NS; t++) {
          const uint32_t j0 = row0 + t * 16 + 4 * kq;
          float sv[4];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const float u = acc[m][t][r], l = acc[m + 2][t][r];
            const float un = __int_as_float(
                __builtin_amdgcn_mov_dpp(__float_as_int(u), 0xB1, 0xF, 0xF, false));
            const float ln = __int_as_float(
                __builtin_amdgcn_mov_dpp(__float_as_int(l), 0xB1, 0xF, 0xF, false));
            sv[r] = ((u + un) + l) + ln;
          }
          if ((col & 1) == 0) {
            const uint2 mk = *reinterpret_cast<const uint2*>(a.ug_mask + q * a.ug_mcs + j0);
            const uint32_t mw[4] = {mk.x << 16, mk.x & 0xffff0000u, mk.y << 16,
                                    mk.y & 0xffff0000u};
#pragma unroll
            for (int r = 0; r < 4; r++)
              if (!(__uint_as_float(mw[r]) > 0.0f)) sv[r] = 0.0f;
            *reinterpret_cast<uint2*>(a.ug_dz + q * a.ug_dzcs + j0) =
                make_uint2(pack_bf16x2(sv[0], sv[1]), pack_bf16x2(sv[2], sv[3]));
          }
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < MF; m++) {
        const int py = y0 + wrow + m / 2;
        if (py >= (int)a.h) continue;
        const size_t p = img_base + (size_t)py * a.w_ + (uint32_t)x0 + 16 * (m % 2) + col;
#pragma unroll
        for (int t = 0; t < NS; t++) {
          const uint32_t j0 = row0 + t * 16 + 4 * kq;
          if (j0 >= a.nout) continue;
          float v[4];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const uint32_t j = j0 + r;
            if constexpr (PERSIST)
              v[r] = acc[m][t][r] + bias_s[j - row0];
            else
              v[r] = acc[m][t][r] + ((a.bias && j < a.nout) ? a.bias[j] : 0.0f);
            if (a.relu) v[r] = fmaxf(v[r], 0.0f);
          }
          const bool whole = j0 + 4 <= a.nout;
          if constexpr (MASK) {
            // the ReLU gradient of the layer below: 4 bf16 of its output at once
            // (the host checks the mask's alignment and stride)
            if (whole) {
              const uint2 mk = *reinterpret_cast<const uint2*>(a.mask + p * a.mcs + j0);
              const uint32_t mw[4] = {mk.x << 16, mk.x & 0xffff0000u, mk.y << 16,
                                      mk.y & 0xffff0000u};
#pragma unroll
              for (int r = 0; r < 4; r++)
                if (!(__uint_as_float(mw[r]) > 0.0f)) v[r] = 0.0f;
            } else {
              for (int r = 0; r < 4; r++)
                if (j0 + r < a.nout && !(from_bf16(a.mask[p * a.mcs + j0 + r]) > 0.0f))
                  v[r] = 0.0f;
            }
          }
          if (a.y) {
            if (whole && vec_y) {
              uint2 pk;
              pk.x = pack_bf16x2(v[0], v[1]);
              pk.y = pack_bf16x2(v[2], v[3]);
              *reinterpret_cast<uint2*>(a.y + p * a.ycs + j0) = pk;
            } else {
              for (int r = 0; r < 4; r++)
                if (j0 + r < a.nout) a.y[p * a.ycs + j0 + r] = to_bf16(v[r]);
            }
          }
          if (a.yf) {
            float* o = a.yf + p * a.yfcs + j0;
            if (whole && vec_f) {
              float4 w4 = make_float4(v[0], v[1], v[2], v[3]);
              if (a.accumulate) {
                const float4 old = *reinterpret_cast<const float4*>(o);
                w4.x += old.x;
                w4.y += old.y;
                w4.z += old.z;
                w4.w += old.w;
              }
              *reinterpret_cast<float4*>(o) = w4;
            } else {
              for (int r = 0; r < 4; r++)
                if (j0 + r < a.nout) o[r] = a.accumulate ? o[r] + v[r] : v[r];
            }
          }
        }
      }
    }
    if (!more) break;
    tile = next;
    x0 = nx0;
    y0 = ny0;
    img_base = nbase;
    next = next2;
    nx0 = n2x0;
    ny0 = n2y0;
    nbase = n2base;
    more = more2;
  }
#undef CV_HALO_LOAD
#undef CV_HALO_LOAD_X
#undef CV_HALO_STORE
#undef CV_HALO_STORE_X
}

// ---------------------------------------------------------------------------
// Backward to the weights: dW[co][tap * cin + c] += sum_p dZ[p][co] *
// X[p + offset(tap)][c], and db[co] += sum_p dZ[p][co]: a GEMM whose K runs
// over the pixels, split over blocks (split-K).
// ---------------------------------------------------------------------------
struct WGradArgs {
  const uint16_t* dz;  // [p][zcs] bf16
  uint32_t zcs, cout;
  const uint16_t* x;   // [p][xcs] bf16
  uint32_t xcs, cin;
  float* dw;           // [cout][9 * cin]
  float* db;           // [cout] or null
  uint32_t npix, h, w_;
  uint32_t pix_per_block;
  float* part;         // k_wgrad_t: per pixel split [z][cout * 9 * cin + cout]
  uint32_t tiles_per_split;  // k_wgrad_halo: 8 x 32 pixel tiles per split
  // non-null (k_wgrad_halo<..., UP>): input channels [0, up_ch) are the
  // nearest 2x upsampling of xup (pixel stride xupcs, h / 2 x w / 2), read
  // from it directly (a multiple of the block's input channels)
  const uint16_t* xup;
  uint32_t xupcs, up_ch;
};

// 8 pixels x 8 bf16 channels (in[i] = pixel i) to 8 channels x 8 pixels
// (out[c] = channel c of pixels 0..7): halfword moves only.
__device__ __forceinline__ void transpose8x8(const uint4 (&in)[8], uint4 (&out)[8]) {
#pragma unroll
  for (int c = 0; c < 8; c++) {
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t lo = (&in[2 * k].x)[c / 2], hi = (&in[2 * k + 1].x)[c / 2];
      o[k] = (c & 1) ? (lo >> 16) | (hi & 0xffff0000u) : (lo & 0xffffu) | (hi << 16);
    }
    out[c] = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// Weight gradient, transposed in registers.  GEMM M = cout (NM 16-row
// tiles), N = 128 columns (tap, c) per block, K = pixels in chunks of 64.
// Waves 0-1 gather the shifted input (64 pixels x 128 columns), waves 2-3
// dZ (64 pixels x cout): each thread loads an 8-pixel x 8-channel tile with
// eight 16-byte loads, transposes it in registers and stores eight 16-byte
// rows (pixels contiguous), so every MFMA fragment is one 16-byte LDS read.
// Wave w owns column tiles 2w and 2w + 1: 4 * NM MFMAs per chunk.  The next
// chunk's loads are in flight while the current chunk's MFMAs run (LDS
// double buffered, one barrier per chunk).  Blocks split the pixels; each
// block stores its dW tile (and, in column tile 0, its db sums) into its
// split's partial slice, which k_wgrad_reduce adds up in split order: no
// atomics, so the gradient is the same on every run.
__device__ const uint4 kZero16 = {0u, 0u, 0u, 0u};
constexpr uint32_t kWgPix = 64;   // pixels per chunk
constexpr uint32_t kWgCols = 128; // columns per block
constexpr uint32_t kWgRow = 72;   // LDS row: 64 pixels + 8 pad (distinct banks per fragment)
template <int NM>
__global__ __launch_bounds__(256) void k_wgrad_t(WGradArgs a) {
  constexpr uint32_t CP = NM * 16;
  __shared__ __attribute__((aligned(16))) uint16_t At[2][CP * kWgRow];
  __shared__ __attribute__((aligned(16))) uint16_t Bt[2][kWgCols * kWgRow];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t ncols = 9 * a.cin;
  const uint32_t pbeg = blockIdx.z * a.pix_per_block;
  const uint32_t pend = min(a.npix, pbeg + a.pix_per_block);
  const uint32_t hw = a.h * a.w_;
  // tile of this thread: pixel group pg (8 pixels), channel group cg
  const bool is_b = tid < 128;
  const uint32_t t = is_b ? tid : tid - 128;
  const uint32_t pg = t % 8, cg = t / 8;  // cg < 16 (B) or < 2 * NM (A)
  const bool a_on = !is_b && cg < 2 * NM;
  const uint32_t bj = blockIdx.x * kWgCols + cg * 8;
  const bool b_on = is_b && bj < ncols;
  const uint32_t btap = b_on ? bj / a.cin : 0, bc = b_on ? bj % a.cin : 0;
  const int bdy = (int)(btap / 3) - 1, bdx = (int)(btap % 3) - 1;
  const bool do_bias = blockIdx.x == 0 && a_on;
  float* part = a.part + (size_t)blockIdx.z * (a.cout * ncols + a.cout);
  float bs[8];
#pragma unroll
  for (int e = 0; e < 8; e++) bs[e] = 0.0f;
  uint4 r[8];
  // Element offsets are 32-bit (the host checks npix * channels < 2^31).
  // Every load is issued: a padding pixel reads 16 zero bytes (kZero16), so
  // no branch joins the loaded registers and the loads stay in flight over
  // the MFMAs.  When the 8 pixels of a group lie in one row (always when
  // w % 8 == 0) the shifted row is checked once and each column once.
  auto pick = [&](bool ok, const uint16_t* q) -> uint4 {
    return *(ok ? reinterpret_cast<const uint4*>(q) : &kZero16);
  };
  auto load_chunk = [&](uint32_t pk) {
    const uint32_t p0 = pk + pg * 8;
    if (a_on) {
      const uint16_t* src = a.dz + cg * 8;
#pragma unroll
      for (int i = 0; i < 8; i++) r[i] = pick(p0 + i < pend, src + (p0 + i) * a.zcs);
    } else if (is_b) {
      const uint32_t pc = min(p0, pend - 1);
      const uint32_t b = pc / hw, rem = pc % hw;
      const uint32_t y = rem / a.w_, x = rem % a.w_;
      const int sy = (int)y + bdy;
      if (p0 + 7 < pend && x + 7 < a.w_) {
        const bool rowok = b_on && sy >= 0 && sy < (int)a.h;
        const int sx0 = (int)x + bdx;
        // wraps for sx0 = -1; pixel 0 is then not loaded and pixels 1..7
        // get their true offsets
        const uint32_t off0 = (b * hw + (uint32_t)sy * a.w_ + (uint32_t)sx0) * a.xcs + bc;
#pragma unroll
        for (int i = 0; i < 8; i++)
          r[i] = pick(rowok && (uint32_t)(sx0 + i) < a.w_, a.x + (off0 + (uint32_t)i * a.xcs));
      } else {
        uint32_t bb = b, yy = y, xx = x;
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const int syy = (int)yy + bdy, sxx = (int)xx + bdx;
          const bool ok = b_on && p0 + i < pend && syy >= 0 && syy < (int)a.h && sxx >= 0 &&
                          sxx < (int)a.w_;
          r[i] = pick(ok, a.x + ((bb * hw + (uint32_t)syy * a.w_ + (uint32_t)sxx) * a.xcs + bc));
          if (++xx == a.w_) {
            xx = 0;
            if (++yy == a.h) {
              yy = 0;
              bb++;
            }
          }
        }
      }
    }
  };
  auto store_chunk = [&](uint32_t buf) {
    if (!a_on && !is_b) return;
    uint4 o[8];
    transpose8x8(r, o);
    uint16_t* dst = is_b ? &Bt[buf][(cg * 8) * kWgRow + pg * 8] : &At[buf][(cg * 8) * kWgRow + pg * 8];
#pragma unroll
    for (int c = 0; c < 8; c++) *reinterpret_cast<uint4*>(dst + c * kWgRow) = o[c];
    if (do_bias) {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const uint32_t* w = &r[i].x;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          bs[2 * k] += from_bf16((uint16_t)(w[k] & 0xffffu));
          bs[2 * k + 1] += from_bf16((uint16_t)(w[k] >> 16));
        }
      }
    }
  };
  f4 acc[NM][2];
#pragma unroll
  for (int m = 0; m < NM; m++) acc[m][0] = acc[m][1] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  if (pbeg < pend) {
    load_chunk(pbeg);
    store_chunk(0);
  }
  __syncthreads();
  const uint32_t row = lane & 15, koff = 8 * (lane >> 4);
  uint32_t buf = 0;
  for (uint32_t pk = pbeg; pk < pend; pk += kWgPix) {
    const bool more = pk + kWgPix < pend;
    if (more) load_chunk(pk + kWgPix);
#pragma unroll
    for (uint32_t ks = 0; ks < kWgPix / 32; ks++) {
      bf8 B[2];
#pragma unroll
      for (int n = 0; n < 2; n++)
        B[n] = *reinterpret_cast<const bf8*>(&Bt[buf][((wave * 2 + n) * 16 + row) * kWgRow + ks * 32 + koff]);
#pragma unroll
      for (int m = 0; m < NM; m++) {
        const bf8 A = *reinterpret_cast<const bf8*>(&At[buf][(m * 16 + row) * kWgRow + ks * 32 + koff]);
#pragma unroll
        for (int n = 0; n < 2; n++)
          acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[n], acc[m][n], 0, 0, 0);
      }
    }
    if (more) store_chunk(buf ^ 1u);
    __syncthreads();
    buf ^= 1u;
  }
#pragma unroll
  for (int n = 0; n < 2; n++) {
    const uint32_t j = blockIdx.x * kWgCols + (wave * 2 + n) * 16 + row;
    if (j >= ncols) continue;
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t o = m * 16 + 4 * (lane >> 4) + q;
        if (o < a.cout) part[o * ncols + j] = acc[m][n][q];
      }
  }
  // db: the 8 pixel groups of a channel group are lanes 8 cg .. 8 cg + 7
  if (!is_b) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      bs[e] += __shfl_xor(bs[e], 1);
      bs[e] += __shfl_xor(bs[e], 2);
      bs[e] += __shfl_xor(bs[e], 4);
    }
    if (do_bias && pg == 0) {
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (cg * 8 + e < a.cout) part[a.cout * ncols + cg * 8 + e] = bs[e];
    }
  }
}

// ---------------------------------------------------------------------------
// Backward to the weights from transposed LDS tiles, for images whose rows
// are a multiple of 32 pixels wide: dW[co][tap][ci] = sum_p dZ[p][co] *
// X[p + offset(tap)][ci] as MFMAs with K = 32 pixels of one row.  A block
// takes NS output tiles and CT input tiles of 16 channels (one wave per
// pair) and a split of the 8 x 32 pixel tiles.  Per tile it stages dZ
// (8 rows x 32 pixels) and X with its halo (10 rows x 34 pixels, zeros
// outside the image) channel-major into LDS -- pixel pairs transposed into
// 32-bit words -- while the next tile's loads are in flight.  A wave keeps
// its 8 dZ fragments (one per row) in registers and reads each halo row of
// its input channels once: three aligned 8-pixel windows per lane, from
// which funnel shifts make the windows one pixel left and right; every
// such row feeds the 3 output rows and 3 column offsets it belongs to, 72
// MFMAs per tile.  The bias gradient is an MFMA of dZ against ones.  Each
// split stores its partial tile, which k_wgrad_reduce adds in split order.
// ---------------------------------------------------------------------------
constexpr uint32_t kWhCols = 48;                    // X^T row: pixels x0 - 8 .. x0 + 39
constexpr uint32_t kWhCi = 10 * kWhCols + 8;        // X^T stride per input channel (bf16)
constexpr uint32_t kWhCo = 8 * 32 + 8;              // dZ^T stride per output channel

template <int NS, int CT, bool UP = false>
__global__ __launch_bounds__(64 * NS * CT) void k_wgrad_halo(WGradArgs a) {
  constexpr int BS = 64 * NS * CT;
  constexpr int CC = 16 * CT, CO = 16 * NS;
  constexpr int XG = CC / 8, ZG = CO / 8;  // 8-channel groups
  constexpr int XP = 10 * 18 * XG;         // X pixel-pair pieces per tile
  constexpr int ZP = 8 * 16 * ZG;          // dZ pixel-pair pieces per tile
  constexpr int XPT = (XP + BS - 1) / BS, ZPT = (ZP + BS - 1) / BS;
  // Each 8-channel group's rows start 32 / groups dwords further on (a
  // multiple of 16 bytes, as the 16-byte reads need): the transposing
  // stores write one channel row per group at a time, and with the row
  // strides alone (multiples of 32 dwords apart for 8 channels) every
  // group would hit the same LDS banks.
  constexpr int XSH = 64 / XG, ZSH = 64 / ZG;  // bf16 elements
  __shared__ __attribute__((aligned(16))) uint16_t xs[CC * kWhCi + (XG - 1) * XSH];
  __shared__ __attribute__((aligned(16))) uint16_t zs[CO * kWhCo + (ZG - 1) * ZSH];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t ct = wave / CT, it = wave % CT;
  const uint32_t ln = lane & 15, g = lane >> 4;
  const uint32_t ci0 = blockIdx.x * CC, co0 = blockIdx.y * CO;
  const uint32_t ncols = 9 * a.cin;
  const uint32_t ntx = a.w_ / 32, nty = (a.h + 7) / 8;
  const uint32_t ntiles = (a.npix / (a.h * a.w_)) * ntx * nty;
  const uint32_t tb = blockIdx.z * a.tiles_per_split;
  const uint32_t te = min(ntiles, tb + a.tiles_per_split);
  u32x4 xr[XPT][2], zr[ZPT][2];
  // the tile's pieces into registers: a piece is 8 channels of two
  // neighbouring pixels (X: columns 6 + 2 k of the halo row, i.e. pixels
  // x0 - 2 + 2 k; dZ: pixels x0 + 2 k)
  // UP: this block's input channels are upsampled ones (block-uniform)
  const bool from_up = UP && ci0 < a.up_ch;
  auto load_tile = [&](uint32_t t) {
    const uint32_t tx = t % ntx, ty = (t / ntx) % nty, img = t / (ntx * nty);
    const int x0 = (int)(tx * 32), y0 = (int)(ty * 8);
    const size_t base = (size_t)img * a.h * a.w_;
#pragma unroll
    for (int i = 0; i < XPT; i++) {
      const int q = (int)tid + BS * i;
      const int g8 = q % XG, pr = (q / XG) % 18, row = q / XG / 18;
      const int sy = y0 + row - 1;
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int px = x0 - 2 + 2 * pr + e;
        const bool in = q < XP && sy >= 0 && sy < (int)a.h && px >= 0 && px < (int)a.w_;
        if (from_up)
          xr[i][e] = in ? *reinterpret_cast<const u32x4*>(
                              a.xup +
                              (base / 4 + (size_t)(sy >> 1) * (a.w_ >> 1) + ((uint32_t)px >> 1)) *
                                  a.xupcs +
                              ci0 + 8 * g8)
                        : u32x4{0u, 0u, 0u, 0u};
        else
          xr[i][e] = in ? *reinterpret_cast<const u32x4*>(
                              a.x + (base + (size_t)sy * a.w_ + (uint32_t)px) * a.xcs + ci0 +
                              8 * g8)
                        : u32x4{0u, 0u, 0u, 0u};
      }
    }
#pragma unroll
    for (int i = 0; i < ZPT; i++) {
      const int q = (int)tid + BS * i;
      const int g8 = q % ZG, pr = (q / ZG) % 16, row = q / ZG / 16;
      const int sy = y0 + row;
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int px = x0 + 2 * pr + e;
        zr[i][e] = (q < ZP && sy < (int)a.h)
                       ? *reinterpret_cast<const u32x4*>(
                             a.dz + (base + (size_t)sy * a.w_ + (uint32_t)px) * a.zcs + co0 + 8 * g8)
                       : u32x4{0u, 0u, 0u, 0u};
      }
    }
  };
  // transposed into LDS: word (channel c, pixel pair) = c of the left pixel
  // | c of the right pixel << 16
  auto store_tile = [&]() {
#pragma unroll
    for (int i = 0; i < XPT; i++) {
      const int q = (int)tid + BS * i;
      if (q >= XP) continue;
      const int g8 = q % XG, pr = (q / XG) % 18, row = q / XG / 18;
      uint32_t* d =
          reinterpret_cast<uint32_t*>(xs + (8 * g8) * kWhCi + g8 * XSH + row * kWhCols + 6 + 2 * pr);
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const uint32_t lo = xr[i][0][e / 2], hi = xr[i][1][e / 2];
        d[e * (kWhCi / 2)] = (e & 1) ? (lo >> 16) | (hi & 0xffff0000u) : (lo & 0xffffu) | (hi << 16);
      }
    }
#pragma unroll
    for (int i = 0; i < ZPT; i++) {
      const int q = (int)tid + BS * i;
      if (q >= ZP) continue;
      const int g8 = q % ZG, pr = (q / ZG) % 16, row = q / ZG / 16;
      uint32_t* d = reinterpret_cast<uint32_t*>(zs + (8 * g8) * kWhCo + g8 * ZSH + row * 32 + 2 * pr);
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const uint32_t lo = zr[i][0][e / 2], hi = zr[i][1][e / 2];
        d[e * (kWhCo / 2)] = (e & 1) ? (lo >> 16) | (hi & 0xffff0000u) : (lo & 0xffffu) | (hi << 16);
      }
    }
  };
  f4 acc[9], accb = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < 9; k++) acc[k] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  bf8 ones;
#pragma unroll
  for (int k = 0; k < 8; k++) ones[k] = (__bf16)1.0f;
  if (tb < te) load_tile(tb);
#pragma unroll 1
  for (uint32_t t = tb; t < te; t++) {
    store_tile();
    __syncthreads();
    if (t + 1 < te) load_tile(t + 1);
    bf8 A[8];
#pragma unroll
    for (int r = 0; r < 8; r++)
      A[r] = *reinterpret_cast<const bf8*>(
          &zs[(ct * 16 + ln) * kWhCo + ((ct * 16 + ln) / 8) * ZSH + r * 32 + 8 * g]);
    if (it == 0) {
#pragma unroll
      for (int r = 0; r < 8; r++) accb = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[r], ones, accb, 0, 0, 0);
    }
    const uint16_t* xrow0 = &xs[(it * 16 + ln) * kWhCi + ((it * 16 + ln) / 8) * XSH + 8 * g];
#pragma unroll
    for (int hr = 0; hr < 10; hr++) {
      // windows of pixels x0 + 8 g - 8, x0 + 8 g, x0 + 8 g + 8 (8 each)
      const u32x4 w0 = *reinterpret_cast<const u32x4*>(xrow0 + hr * kWhCols);
      const u32x4 w1 = *reinterpret_cast<const u32x4*>(xrow0 + hr * kWhCols + 8);
      const u32x4 w2 = *reinterpret_cast<const u32x4*>(xrow0 + hr * kWhCols + 16);
      // one pixel left (dx = -1) and right (dx = +1) of the centre window
      const uint32_t s0 = __builtin_amdgcn_alignbit(w1[0], w0[3], 16);
      const uint32_t s1 = __builtin_amdgcn_alignbit(w1[1], w1[0], 16);
      const uint32_t s2 = __builtin_amdgcn_alignbit(w1[2], w1[1], 16);
      const uint32_t s3 = __builtin_amdgcn_alignbit(w1[3], w1[2], 16);
      const uint32_t s4 = __builtin_amdgcn_alignbit(w2[0], w1[3], 16);
      const u32x4 fl = {s0, s1, s2, s3}, fr = {s1, s2, s3, s4};
      const bf8 F[3] = {__builtin_bit_cast(bf8, fl), __builtin_bit_cast(bf8, w1),
                        __builtin_bit_cast(bf8, fr)};
#pragma unroll
      for (int dyi = 0; dyi < 3; dyi++) {
        // halo row hr is row hr - dyi of the tile for dy = dyi - 1
        const int r = hr - dyi;
        if (r < 0 || r >= 8) continue;
#pragma unroll
        for (int dxi = 0; dxi < 3; dxi++)
          acc[dyi * 3 + dxi] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[r], F[dxi], acc[dyi * 3 + dxi], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // the split's partial: lane holds output channels 4 g + r of input channel ln
  float* part = a.part + (size_t)blockIdx.z * (a.cout * ncols + a.cout);
  const uint32_t ci = ci0 + it * 16 + ln;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint32_t co = co0 + ct * 16 + 4 * g + r;
    if (co >= a.cout) continue;
#pragma unroll
    for (int tap = 0; tap < 9; tap++) part[(size_t)co * ncols + tap * a.cin + ci] = acc[tap][r];
    if (it == 0 && blockIdx.x == 0 && ln == 0) part[(size_t)a.cout * ncols + co] = accb[r];
  }
}


// dW += sum over the pixel splits of the partial slices, and db likewise
// from the slices' tails.  A block takes 64 consecutive entries; its 16
// waves take every 16th split (four independent sums each, coalesced
// 256-byte reads), then wave 0 adds the 16 wave sums: a fixed order.
__global__ __launch_bounds__(1024) void k_wgrad_reduce(const float* part, uint32_t nsplit,
                                                       uint32_t nw, uint32_t nb, float* dw,
                                                       float* db) {
  __shared__ float red[16][64];
  const uint32_t lx = threadIdx.x & 63, zg = threadIdx.x >> 6;
  const uint32_t n = nw + nb, e = blockIdx.x * 64 + lx;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
  if (e < n) {
    uint32_t z = zg;
    for (; z + 48 < nsplit; z += 64) {
      s0 += part[(size_t)z * n + e];
      s1 += part[(size_t)(z + 16) * n + e];
      s2 += part[(size_t)(z + 32) * n + e];
      s3 += part[(size_t)(z + 48) * n + e];
    }
    for (; z < nsplit; z += 16) s0 += part[(size_t)z * n + e];
  }
  red[zg][lx] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (zg == 0 && e < n) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; k++) s += red[k][lx];
    if (e < nw)
      dw[e] += s;
    else
      db[e - nw] += s;
  }
}

// ---------------------------------------------------------------------------
// Elementwise and resampling kernels
// ---------------------------------------------------------------------------

// The layout kernels below move 8 BF16 (16 B) or 4 FP32 channels per
// thread with 32-bit index arithmetic: every channel count and channel
// stride they are given is a multiple of 8, every base pointer 16-B aligned.
__device__ __forceinline__ void unpack8(uint4 v, float* f) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    f[2 * k] = from_bf16((uint16_t)(w[k] & 0xffffu));
    f[2 * k + 1] = from_bf16((uint16_t)(w[k] >> 16));
  }
}
__device__ __forceinline__ uint4 pack8(const float* f) {
  uint32_t w[4];
#pragma unroll
  for (int k = 0; k < 4; k++) w[k] = (uint32_t)to_bf16(f[2 * k]) | ((uint32_t)to_bf16(f[2 * k + 1]) << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// dZ = dY * (Y > 0) for ReLU layers (or dY for linear ones), as BF16 with
// the padded channels [c, zcs) zeroed (zcs a multiple of 8, dz 16-B
// aligned): one thread per pixel, 8 channels per store.
__global__ void k_relu_grad(const float* dy, uint32_t dycs, const uint16_t* y, uint32_t ycs,
                            uint32_t c, uint16_t* dz, uint32_t zcs, uint32_t npix, uint32_t relu) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += gridDim.x * blockDim.x) {
    for (uint32_t g = 0; g < zcs; g += 8) {
      float v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint32_t ch = g + k;
        float x = 0.0f;
        if (ch < c) {
          x = dy[(size_t)p * dycs + ch];
          if (relu && !(from_bf16(y[(size_t)p * ycs + ch]) > 0.0f)) x = 0.0f;
        }
        v[k] = x;
      }
      *reinterpret_cast<uint4*>(dz + (size_t)p * zcs + g) = pack8(v);
    }
  }
}
__global__ void k_relu_grad8(const float* dy, uint32_t dycs, const uint16_t* y, uint32_t ycs,
                             uint32_t c, uint16_t* dz, uint32_t npix, uint32_t relu) {
  const uint32_t g = c / 8, n = npix * g;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t p = i / g, ch = (i % g) * 8;
    const float4 a = *reinterpret_cast<const float4*>(dy + (size_t)p * dycs + ch);
    const float4 b = *reinterpret_cast<const float4*>(dy + (size_t)p * dycs + ch + 4);
    float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    if (relu) {
      float yv[8];
      unpack8(*reinterpret_cast<const uint4*>(y + (size_t)p * ycs + ch), yv);
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (!(yv[k] > 0.0f)) v[k] = 0.0f;
    }
    *reinterpret_cast<uint4*>(dz + (size_t)p * c + ch) = pack8(v);
  }
}

// 2x2 average pooling, BF16 in and out.
__global__ void k_pool(const uint16_t* x, uint32_t xcs, uint16_t* y, uint32_t ycs, uint32_t c,
                       uint32_t nb, uint32_t h2, uint32_t w2) {
  const uint32_t g = c / 8, n = nb * h2 * w2 * g;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t ch = (i % g) * 8, q = i / g;
    const uint32_t X = q % w2, t = q / w2, Y = t % h2, B = t / h2;
    const uint32_t base = (B * (2 * h2) + 2 * Y) * (2 * w2) + 2 * X;
    float a[8], b[8], c2[8], d[8], r[8];
    unpack8(*reinterpret_cast<const uint4*>(x + (size_t)base * xcs + ch), a);
    unpack8(*reinterpret_cast<const uint4*>(x + (size_t)(base + 1) * xcs + ch), b);
    unpack8(*reinterpret_cast<const uint4*>(x + (size_t)(base + 2 * w2) * xcs + ch), c2);
    unpack8(*reinterpret_cast<const uint4*>(x + (size_t)(base + 2 * w2 + 1) * xcs + ch), d);
#pragma unroll
    for (int k = 0; k < 8; k++) r[k] = 0.25f * (((a[k] + b[k]) + c2[k]) + d[k]);
    *reinterpret_cast<uint4*>(y + (size_t)q * ycs + ch) = pack8(r);
  }
}

// Backward of the pooling: dx += dy / 4 on each of the 4 inputs (FP32).
__global__ void k_pool_grad(const float* dy, uint32_t dycs, float* dx, uint32_t dxcs, uint32_t c,
                            uint32_t nb, uint32_t h2, uint32_t w2) {
  const uint32_t g = c / 4, n = nb * h2 * w2 * g;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t ch = (i % g) * 4, q = i / g;
    const uint32_t X = q % w2, t = q / w2, Y = t % h2, B = t / h2;
    const uint32_t base = (B * (2 * h2) + 2 * Y) * (2 * w2) + 2 * X;
    const float4 v = *reinterpret_cast<const float4*>(dy + (size_t)q * dycs + ch);
    const float4 gv = make_float4(0.25f * v.x, 0.25f * v.y, 0.25f * v.z, 0.25f * v.w);
    const uint32_t px[4] = {base, base + 1, base + 2 * w2, base + 2 * w2 + 1};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      float4* o = reinterpret_cast<float4*>(dx + (size_t)px[k] * dxcs + ch);
      float4 u = *o;
      u.x += gv.x;
      u.y += gv.y;
      u.z += gv.z;
      u.w += gv.w;
      *o = u;
    }
  }
}

// k_pool_grad followed by k_relu_grad8 in one pass: for each pixel of the
// full-resolution gradient g (c channels, stride gcs) the pooled gradient's
// share is added as k_pool_grad adds it (g + 0.25 * gp of the parent pixel,
// stride gpcs, h2 x w2 per image), masked by the ReLU output y (stride ycs)
// and written as bf16 dZ (stride c), 8 channels per thread.
__global__ void k_pool_relu_grad(const float* g, uint32_t gcs, const float* gp, uint32_t gpcs,
                                 const uint16_t* y, uint32_t ycs, uint16_t* dz, uint32_t c,
                                 uint32_t nb, uint32_t h2, uint32_t w2) {
  const uint32_t gr = c / 8, n = nb * (2 * h2) * (2 * w2) * gr;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t ch = (i % gr) * 8, p = i / gr;
    const uint32_t X = p % (2 * w2), t = p / (2 * w2), Y = t % (2 * h2), B = t / (2 * h2);
    const uint32_t q = (B * h2 + Y / 2) * w2 + X / 2;
    const float4 a0 = *reinterpret_cast<const float4*>(g + (size_t)p * gcs + ch);
    const float4 a1 = *reinterpret_cast<const float4*>(g + (size_t)p * gcs + ch + 4);
    const float4 b0 = *reinterpret_cast<const float4*>(gp + (size_t)q * gpcs + ch);
    const float4 b1 = *reinterpret_cast<const float4*>(gp + (size_t)q * gpcs + ch + 4);
    float v[8] = {a0.x + 0.25f * b0.x, a0.y + 0.25f * b0.y, a0.z + 0.25f * b0.z,
                  a0.w + 0.25f * b0.w, a1.x + 0.25f * b1.x, a1.y + 0.25f * b1.y,
                  a1.z + 0.25f * b1.z, a1.w + 0.25f * b1.w};
    float yv[8];
    unpack8(*reinterpret_cast<const uint4*>(y + (size_t)p * ycs + ch), yv);
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (!(yv[k] > 0.0f)) v[k] = 0.0f;
    *reinterpret_cast<uint4*>(dz + (size_t)p * c + ch) = pack8(v);
  }
}

// Nearest 2x upsampling, BF16.
__global__ void k_upsample(const uint16_t* x, uint32_t xcs, uint16_t* y, uint32_t ycs, uint32_t c,
                           uint32_t nb, uint32_t h, uint32_t w) {
  const uint32_t g = c / 8, n = nb * h * w * g;  // output vectors (h, w: output size)
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t ch = (i % g) * 8, q = i / g;
    const uint32_t X = q % w, t = q / w, Y = t % h, B = t / h;
    const uint32_t src = (B * (h / 2) + Y / 2) * (w / 2) + X / 2;
    *reinterpret_cast<uint4*>(y + (size_t)q * ycs + ch) =
        *reinterpret_cast<const uint4*>(x + (size_t)src * xcs + ch);
  }
}

// Backward of the upsampling: dx = sum of the 4 output gradients (FP32).
__global__ void k_upsample_grad(const float* dy, uint32_t dycs, float* dx, uint32_t dxcs,
                                uint32_t c, uint32_t nb, uint32_t h2, uint32_t w2) {
  const uint32_t g = c / 4, n = nb * h2 * w2 * g;  // input vectors (h2, w2: input size)
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t ch = (i % g) * 4, q = i / g;
    const uint32_t X = q % w2, t = q / w2, Y = t % h2, B = t / h2;
    const uint32_t base = (B * (2 * h2) + 2 * Y) * (2 * w2) + 2 * X;
    const float4 a = *reinterpret_cast<const float4*>(dy + (size_t)base * dycs + ch);
    const float4 b = *reinterpret_cast<const float4*>(dy + (size_t)(base + 1) * dycs + ch);
    const float4 e = *reinterpret_cast<const float4*>(dy + (size_t)(base + 2 * w2) * dycs + ch);
    const float4 f = *reinterpret_cast<const float4*>(dy + (size_t)(base + 2 * w2 + 1) * dycs + ch);
    *reinterpret_cast<float4*>(dx + (size_t)q * dxcs + ch) =
        make_float4(((a.x + b.x) + e.x) + f.x, ((a.y + b.y) + e.y) + f.y,
                    ((a.z + b.z) + e.z) + f.z, ((a.w + b.w) + e.w) + f.w);
  }
}

// Input features (row-major f32, 10 channels) -> BF16 [p][16], zero padded;
// one pixel per thread, two 16-B stores.
__global__ void k_to_input(const float* in, uint16_t* x, uint32_t npix) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    float f[16];
#pragma unroll
    for (int k = 0; k < 10; k++) f[k] = in[(size_t)p * 10 + k];
#pragma unroll
    for (int k = 10; k < 16; k++) f[k] = 0.0f;
    uint4* o = reinterpret_cast<uint4*>(x + (size_t)p * 16);
    o[0] = pack8(f);
    o[1] = pack8(f + 8);
  }
}

// L1 loss in log space: out (predicted log1p radiance, [p][3]) against
// log1p(target); writes dL/dout / npix*3 into dout ([p][3]) and block b's
// share of the loss to part[b] (k_sum_parts adds them in a fixed order: no
// atomics, which on one address serialise, and the same sum every run).
__global__ void k_l1_loss(const float* out, const float* target, float* dout, uint32_t npix,
                          float* part_out) {
  __shared__ float part[256];
  float s = 0.0f;
  const float scale = 1.0f / (3.0f * (float)npix);
  const size_t n = (size_t)npix * 3;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float d = out[i] - log1pf(fmaxf(target[i], 0.0f));
    s += fabsf(d);
    dout[i] = d > 0.0f ? scale : (d < 0.0f ? -scale : 0.0f);
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part_out[blockIdx.x] = part[0] * scale;
}

// *sum = part[0] + ... + part[n - 1] (n <= 1024) by one 256-thread block in
// a fixed order.
__global__ __launch_bounds__(256) void k_sum_parts(const float* part, uint32_t n, float* sum) {
  __shared__ float red[256];
  float s = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *sum = red[0];
}

// Adam (Kingma & Ba) on the flat FP32 parameters.
__global__ void k_adam(float* w, const float* g, float* m, float* v, uint32_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// FP32 master weights [cout][9][cin] -> BF16 W [cout_pad][kpad] (k = tap *
// cin + c) and Wt [cin][kpad_t] (k = tap * cout_pad + co), zero padded.
// One layer of k_pack: its weights in the parameter vector and its two
// packed copies.
struct PackLayer {
  const float* w;
  uint16_t* W;
  uint16_t* Wt;
  uint32_t cout, cin, cout_pad, kpad, kpad_t;
};
struct PackArgs {
  PackLayer l[11];
};
// blockIdx.y: the layer
__global__ void k_pack(PackArgs pa) {
  const PackLayer& P = pa.l[blockIdx.y];
  const float* w = P.w;
  uint16_t *W = P.W, *Wt = P.Wt;
  const uint32_t cout = P.cout, cin = P.cin, cout_pad = P.cout_pad, kpad = P.kpad,
                 kpad_t = P.kpad_t;
  const uint32_t n1 = cout_pad * kpad, n2 = cin * kpad_t;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n1 + n2;
       i += gridDim.x * blockDim.x) {
    if (i < n1) {
      const uint32_t co = i / kpad, k = i % kpad;
      float v = 0.0f;
      if (co < cout && k < 9 * cin) v = w[((size_t)co * 9 + k / cin) * cin + k % cin];
      W[i] = to_bf16(v);
    } else {
      const uint32_t e = i - n1;
      const uint32_t c = e / kpad_t, k = e % kpad_t;
      float v = 0.0f;
      if (k < 9 * cout_pad) {
        const uint32_t tap = k / cout_pad, co = k % cout_pad;
        if (co < cout) v = w[((size_t)co * 9 + tap) * cin + c];
      }
      Wt[e] = to_bf16(v);
    }
  }
}

__global__ void k_fill(float* p, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Renderer outputs -> network input: log1p of the mean radiance (16x4-block
// HDR layout / spp), albedo, octahedral normal, log1p depth, spp -- the
// features summed over spp frames, 8 f32 per pixel (albedo rgb, normal xyz,
// depth, 0) -- as 10 f32 per pixel into `out` (rows y from 0) and/or as the
// network's 16-channel bf16 input into `x` (either may be null).
__global__ void k_features(const float* hdr, const float* feat, uint32_t width, uint32_t height,
                           uint32_t spp, float* out, uint16_t* x) {
  const uint32_t npix = width * height;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    const uint32_t px = p % width, py = p / width;
    const float* h = hdr + cvk::hdr_offset(px, py, width);
    const float inv = 1.0f / (float)spp;
    const float* f = feat + (size_t)p * 8;
    float o[10];
    o[0] = log1pf(fmaxf(h[0] * inv, 0.0f));
    o[1] = log1pf(fmaxf(h[8] * inv, 0.0f));
    o[2] = log1pf(fmaxf(h[16] * inv, 0.0f));
    o[3] = f[0] * inv;
    o[4] = f[1] * inv;
    o[5] = f[2] * inv;
    float nx = f[3] * inv, ny = f[4] * inv, nz = f[5] * inv;
    const float l1 = fabsf(nx) + fabsf(ny) + fabsf(nz);
    float ox = 0.0f, oy = 0.0f;
    if (l1 > 0.0f) {
      ox = nx / l1;
      oy = ny / l1;
      if (nz < 0.0f) {
        const float tx = (1.0f - fabsf(oy)) * (ox >= 0.0f ? 1.0f : -1.0f);
        const float ty = (1.0f - fabsf(ox)) * (oy >= 0.0f ? 1.0f : -1.0f);
        ox = tx;
        oy = ty;
      }
    }
    o[6] = ox;
    o[7] = oy;
    o[8] = log1pf(fmaxf(f[6] * inv, 0.0f));
    o[9] = (float)spp;
    if (out) {
#pragma unroll
      for (int k = 0; k < 10; k++) out[(size_t)p * 10 + k] = o[k];
    }
    if (x) {
      // the network's bf16 input directly (what k_to_input makes of `out`)
      float f[16];
#pragma unroll
      for (int k = 0; k < 16; k++) f[k] = k < 10 ? o[k] : 0.0f;
      uint4* xo = reinterpret_cast<uint4*>(x + (size_t)p * 16);
      xo[0] = pack8(f);
      xo[1] = pack8(f + 8);
    }
  }
}

// Network output (log1p radiance, [p][3]) -> 16x4-block HDR holding
// spp * radiance, so that it converts like an accumulation of spp frames.
__global__ void k_output(const float* out, uint32_t width, uint32_t height, uint32_t spp,
                         float* hdr) {
  const uint32_t npix = width * height;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    const uint32_t px = p % width, py = p / width;
    float* h = hdr + cvk::hdr_offset(px, py, width);
    for (int c = 0; c < 3; c++) h[8 * c] = expm1f(fmaxf(out[(size_t)p * 3 + c], 0.0f)) * (float)spp;
  }
}

}  // namespace cvd

// ---------------------------------------------------------------------------
// Host side: the network, its buffers and the training step.
// ---------------------------------------------------------------------------

namespace {

inline uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }
inline uint32_t grid1(size_t n) { return cvh::grid_for(n, 256, 65536); }

struct Conv {
  uint32_t cin, cout, cout_pad, kpad, kpad_t;
  bool relu;
  size_t w_off, b_off;  // offsets into the flat parameter vector
  DevBuf<uint16_t> W, Wt;
};

// A BF16 NHWC tensor view and its FP32 gradient view.
struct Act {
  uint16_t* p = nullptr;
  uint32_t cs = 0;  // pixel stride in elements
  float* g = nullptr;
  uint32_t gcs = 0;
};

// The 11 convolutions: enc0a enc0b enc1a enc1b bot2a bot2b dec1a dec1b
// dec0a dec0b out.
const uint32_t kLayers[11][3] = {
    {16, 32, 1},  {32, 32, 1},   {32, 64, 1},  {64, 64, 1},  {64, 128, 1}, {128, 128, 1},
    {192, 64, 1}, {64, 64, 1},   {96, 32, 1},  {32, 32, 1},  {32, 3, 0}};
// blocks of the loss kernel (its partials follow the loss word)
constexpr uint32_t kLossParts = 1024;

}  // namespace

struct cv_denoiser {
  hipStream_t stream = nullptr;
  Conv conv[11];
  size_t nparams = 0;
  DevBuf<float> params, grads, m, v;
  uint64_t step = 0;
  // activations for the current size
  uint32_t nb = 0, h = 0, w = 0;
  size_t cap_pixels = 0;
  // dZ of the layer in backward and of the layer below it (two buffers, so
  // that one layer's dX convolution can write the next layer's dZ)
  DevBuf<uint16_t> x0, e0a, cat0, p0, e1a, cat1, p1, b2a, b2b, d1a, d1b, d0a, d0b, dz, dz2;
  // f32 gradients where they are summed or pooled / upsampled (the other
  // layers hand their dZ on directly, see backward)
  DevBuf<float> out, dout, g_cat0, g_p0, g_cat1, g_p1, g_b2b, g_d1b, loss;
  DevBuf<float> in_f32, tgt_f32;
  DevBuf<float> wpart;  // weight-gradient partials of the pixel splits
  // convolutions from LDS halo tiles where the shape allows (else k_conv_lds)
  bool halo_convs = true;
  std::mutex mu;
};

namespace {

// The halo-tile kernel for the network's shapes (rows a multiple of 32
// pixels, the input channels a multiple of 16 in a dense channel range):
// NT output tiles in slices of NS per block (grid.y = NT / NS), at most 4,
// so that a chunk's weights and tile fit in LDS beside other blocks.
// Workgroups of a halo-tile kernel resident on the device at once (CUs x
// blocks per CU), taken once per instantiation.
template <int KC, int NS, bool FLIP, bool MASK, bool UP, bool POOL, bool UG>
uint32_t halo_resident_blocks() {
  static const uint32_t n = [] {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, cvd::k_conv_halo<KC, NS, FLIP, MASK, UP, POOL, UG>, 256, 0) != hipSuccess ||
        cus <= 0 || per <= 0)
      return 512u;
    return (uint32_t)(cus * per);
  }();
  return n;
}

template <int KC, int NS, bool FLIP, bool MASK = false, bool UP = false, bool POOL = false,
          bool UG = false>
bool try_halo(cv_denoiser* d, const cvd::ConvArgs& a, uint32_t nt, uint32_t ns) {
  if (a.kc != (uint32_t)KC || ns != (uint32_t)NS || (a.flip != 0) != FLIP ||
      (a.mask != nullptr) != MASK || (a.xup != nullptr) != UP || (a.pool != nullptr) != POOL ||
      (a.ug_dz != nullptr) != UG)
    return false;
  // the upsampling's gradient covers whole slices of NS * 16 outputs
  if (UG && a.ug_ch % (NS * 16)) return false;
  const uint32_t imgs = a.npix / (a.h * a.w_);
  const uint32_t tiles =
      imgs * (a.w_ / cvd::kHaloW) * ((a.h + cvd::kHaloH - 1) / cvd::kHaloH);
  uint32_t blocks = tiles;
  if (KC <= 32 && !MASK) {
    // persistent blocks (one input chunk): as many as are resident with
    // every slice's, each taking an equal share of the tiles
    const uint32_t cap =
        std::max(1u, halo_resident_blocks<KC, NS, FLIP, MASK, UP, POOL, UG>() / (nt / NS));
    const uint32_t per = (tiles + cap - 1) / cap;
    blocks = (tiles + per - 1) / per;
  }
  hipLaunchKernelGGL((cvd::k_conv_halo<KC, NS, FLIP, MASK, UP, POOL, UG>), dim3(blocks, nt / NS),
                     dim3(256), 0, d->stream, a);
  return true;
}

bool launch_conv_halo(cv_denoiser* d, const cvd::ConvArgs& a, uint32_t nt) {
  // whole 32-pixel rows, whole images, and every pixel's KC channels within
  // its stride (the tile loads read channels [0, KC) of each pixel)
  if (a.w_ % cvd::kHaloW || a.npix % (a.h * a.w_) || a.xcs % 8 || a.xcs < a.kc) return false;
  // a fused ReLU mask is read 4 channels at a time
  if (a.mask && (a.mcs % 4 || (uintptr_t)a.mask % 8)) return false;
  // a fused pooling: even sizes, whole groups of 4 outputs, 8-byte stores
  if (a.pool && (a.h % 2 || a.w_ % 2 || a.nout % 4 || a.poolcs % 4 || (uintptr_t)a.pool % 8 ||
                 !a.y || a.ycs % 4 || (uintptr_t)a.y % 8 || a.yf || a.mask))
    return false;
  // a fused upsampling gradient: even sizes, 8-byte mask reads and dZ stores
  if (a.ug_dz && (a.h % 2 || a.w_ % 2 || !a.ug_mask || a.ug_mcs % 4 || a.ug_dzcs % 4 ||
                  (uintptr_t)a.ug_mask % 8 || (uintptr_t)a.ug_dz % 8 || a.y || a.mask ||
                  a.ug_ch > a.nout))
    return false;
  // a fused upsampling: even sizes, whole 16-byte pieces of xup
  if (a.xup && (a.h % 2 || a.w_ % 2 || a.xupcs % 8 || a.xupcs < 32 * a.up_chunks ||
                32 * a.up_chunks > a.kc))
    return false;
  const uint32_t ns = nt % 4 == 0 ? 4 : nt % 2 == 0 ? 2 : nt % 3 == 0 ? 3 : 1;
  return try_halo<16, 2, false>(d, a, nt, ns) || try_halo<32, 1, false>(d, a, nt, ns) ||
         try_halo<32, 2, false>(d, a, nt, ns) || try_halo<32, 4, false>(d, a, nt, ns) ||
         try_halo<64, 4, false>(d, a, nt, ns) || try_halo<96, 2, false>(d, a, nt, ns) ||
         try_halo<128, 4, false>(d, a, nt, ns) || try_halo<192, 4, false>(d, a, nt, ns) ||
         try_halo<16, 2, true>(d, a, nt, ns) || try_halo<32, 2, true>(d, a, nt, ns) ||
         try_halo<32, 3, true>(d, a, nt, ns) || try_halo<64, 2, true>(d, a, nt, ns) ||
         try_halo<64, 4, true>(d, a, nt, ns) || try_halo<128, 4, true>(d, a, nt, ns) ||
         try_halo<16, 2, true, true>(d, a, nt, ns) || try_halo<32, 2, true, true>(d, a, nt, ns) ||
         try_halo<64, 4, true, true>(d, a, nt, ns) || try_halo<128, 4, true, true>(d, a, nt, ns) ||
         try_halo<96, 2, false, false, true>(d, a, nt, ns) ||
         try_halo<192, 4, false, false, true>(d, a, nt, ns) ||
         try_halo<32, 2, false, false, false, true>(d, a, nt, ns) ||
         try_halo<64, 4, false, false, false, true>(d, a, nt, ns) ||
         try_halo<32, 2, true, false, false, false, true>(d, a, nt, ns) ||
         try_halo<64, 4, true, false, false, false, true>(d, a, nt, ns);
}

cv_status launch_conv(cv_denoiser* d, const cvd::ConvArgs& a, uint32_t nt) {
  if (d->halo_convs && launch_conv_halo(d, a, nt)) {
    HIPTRY(hipGetLastError(), "conv");
    return CV_OK;
  }
  // the general kernel reads its whole input from x and writes y only
  if (a.xup || a.pool || a.ug_dz)
    return fail(CV_ERR_INVALID_ARGUMENT, "fused upsampling / pooling needs the halo-tile kernel");
  const uint32_t blocks = (a.npix + cvd::kConvPix - 1) / cvd::kConvPix;
  hipStream_t s = d->stream;
  switch (nt) {
    case 1: hipLaunchKernelGGL(cvd::k_conv_lds<1>, dim3(blocks), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(cvd::k_conv_lds<2>, dim3(blocks), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(cvd::k_conv_lds<4>, dim3(blocks), dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL(cvd::k_conv_lds<6>, dim3(blocks), dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL(cvd::k_conv_lds<8>, dim3(blocks), dim3(256), 0, s, a); break;
    case 12: hipLaunchKernelGGL(cvd::k_conv_lds<12>, dim3(blocks), dim3(256), 0, s, a); break;
    default: return fail(CV_ERR_INVALID_ARGUMENT, "unsupported conv width");
  }
  HIPTRY(hipGetLastError(), "conv");
  return CV_OK;
}

// y = relu(conv(x) + b) at level size (h, w).
cv_status conv_fwd(cv_denoiser* d, const Conv& L, const uint16_t* x, uint32_t xcs, uint16_t* y,
                   uint32_t ycs, float* yf, uint32_t yfcs, uint32_t h, uint32_t w,
                   const uint16_t* xup = nullptr, uint32_t xupcs = 0, uint32_t up_chunks = 0,
                   uint16_t* pool = nullptr, uint32_t poolcs = 0) {
  cvd::ConvArgs a{};
  a.x = x;
  a.xcs = xcs;
  a.pool = pool;
  a.poolcs = poolcs;
  a.xup = xup;
  a.xupcs = xupcs;
  a.up_chunks = up_chunks;
  a.kc = L.cin;
  a.w = L.W.p;
  a.kpad = L.kpad;
  a.bias = d->params.p + L.b_off;
  a.nout = L.cout;
  a.y = y;
  a.ycs = ycs;
  a.yf = yf;
  a.yfcs = yfcs;
  a.relu = L.relu ? 1u : 0u;
  a.npix = d->nb * h * w;
  a.h = h;
  a.w_ = w;
  return launch_conv(d, a, L.cout_pad / 16);
}

// The gradient of a nearest 2x upsampling fused into an input-gradient
// convolution: its first `ch` outputs are summed 2 x 2, masked by the ReLU
// output y (pixel stride ycs) of the layer below the upsampling and written
// as that layer's bf16 dZ (pixel stride dzcs).
struct UpGrad {
  uint16_t* dz;
  uint32_t dzcs;
  const uint16_t* y;
  uint32_t ycs, ch;
};

// A layer input whose first `ch` channels are the nearest 2x upsampling of
// x (pixel stride cs) and not written out: the weight gradient reads them
// from x.
struct UpIn {
  const uint16_t* x;
  uint32_t cs, ch;
};

cv_status conv_bwd_input(cv_denoiser* d, const Conv& L, const uint16_t* dz, float* gx,
                         uint32_t gxcs, bool accumulate, const uint16_t* x, uint32_t xcs,
                         uint16_t* dz_next, uint32_t npix, uint32_t h, uint32_t w,
                         const UpGrad* ug);

// Backward of one convolution: dZ (dY at y's shape, f32, through the ReLU
// mask; or, with dz_ready, already in dz), dW and db, and dX (+= into gx
// when accumulate).  With dz_next the layer below -- whose ReLU output is
// this layer's input x, with L.cin channels -- gets its dZ from the dX
// convolution directly: masked by x and written as bf16 to dz_next, in
// place of dX.
cv_status conv_bwd(cv_denoiser* d, const Conv& L, const uint16_t* x, uint32_t xcs,
                   const uint16_t* y, uint32_t ycs, const float* gy, uint32_t gycs, uint16_t* dz,
                   bool dz_ready, float* gx, uint32_t gxcs, bool accumulate, uint16_t* dz_next,
                   uint32_t h, uint32_t w, const UpGrad* ug = nullptr,
                   const UpIn* up = nullptr) {
  const uint32_t npix = d->nb * h * w;
  hipStream_t s = d->stream;
  if (!dz_ready) {
    if (L.cout == L.cout_pad && L.cout % 8 == 0 && gycs % 8 == 0 && (!y || ycs % 8 == 0))
      hipLaunchKernelGGL(cvd::k_relu_grad8, dim3(grid1((size_t)npix * L.cout / 8)), dim3(256), 0,
                         s, gy, gycs, y, ycs, L.cout, dz, npix, L.relu ? 1u : 0u);
    else
      hipLaunchKernelGGL(cvd::k_relu_grad, dim3(grid1(npix)), dim3(256), 0, s,
                         gy, gycs, y, ycs, L.cout, dz, L.cout_pad, npix, L.relu ? 1u : 0u);
    HIPTRY(hipGetLastError(), "relu grad");
  }
  if ((uint64_t)npix * std::max(xcs, L.cout_pad) >= (1ull << 31))
    return fail(CV_ERR_INVALID_ARGUMENT, "training batch too large for 32-bit offsets");
  cvd::WGradArgs g{};
  g.dz = dz;
  g.zcs = L.cout_pad;
  g.cout = L.cout;
  g.x = x;
  g.xcs = xcs;
  g.cin = L.cin;
  g.dw = d->grads.p + L.w_off;
  g.db = d->grads.p + L.b_off;
  g.npix = npix;
  g.h = h;
  g.w_ = w;
  if (up) {
    g.xup = up->x;
    g.xupcs = up->cs;
    g.up_ch = up->ch;
  }
  const uint32_t nw = L.cout * 9 * L.cin;
  if (w % 32 == 0 && npix % (h * w) == 0 && L.cin % 16 == 0 && L.cout_pad % 16 == 0 &&
      xcs % 8 == 0 && (uintptr_t)x % 16 == 0) {
    // transposed LDS tiles: NS output and CT input tiles of 16 channels per
    // block, about 512 blocks over pixel splits of at least 2 tiles (the
    // splits' partial tiles are written and added up again: with 2048
    // blocks that traffic made the reduction 0.14 ms of a 2.0 ms training
    // step; 512 blocks measured fastest, against 2048, 1024, 256 and 128)
    const uint32_t NS = L.cout_pad % 32 == 0 ? 2u : 1u, CT = L.cin % 32 == 0 ? 2u : 1u;
    const uint32_t bps = (L.cin / (16 * CT)) * (L.cout_pad / (16 * NS));
    const uint32_t tiles = (npix / (h * w)) * (w / 32) * ((h + 7) / 8);
    uint32_t splits = std::max(1u, std::min(tiles / 2, std::max(1u, 512u / bps)));
    g.tiles_per_split = (tiles + splits - 1) / splits;
    splits = (tiles + g.tiles_per_split - 1) / g.tiles_per_split;
    HIPTRY(d->wpart.reserve((size_t)splits * (nw + L.cout)), "alloc");
    g.part = d->wpart.p;
    const dim3 grid(L.cin / (16 * CT), L.cout_pad / (16 * NS), splits), block(64 * NS * CT);
    if (up) {
      // whole blocks of 32 input channels upsampled or not, 16-byte pieces
      if (NS != 2 || CT != 2 || up->ch % 32 || up->cs % 8 || (uintptr_t)up->x % 16 || h % 2 ||
          w % 2)
        return fail(CV_ERR_INVALID_ARGUMENT, "upsampled input needs the halo weight gradient");
      hipLaunchKernelGGL((cvd::k_wgrad_halo<2, 2, true>), grid, block, 0, s, g);
    } else if (NS == 2 && CT == 2)
      hipLaunchKernelGGL((cvd::k_wgrad_halo<2, 2>), grid, block, 0, s, g);
    else if (NS == 2)
      hipLaunchKernelGGL((cvd::k_wgrad_halo<2, 1>), grid, block, 0, s, g);
    else if (CT == 2)
      hipLaunchKernelGGL((cvd::k_wgrad_halo<1, 2>), grid, block, 0, s, g);
    else
      hipLaunchKernelGGL((cvd::k_wgrad_halo<1, 1>), grid, block, 0, s, g);
    HIPTRY(hipGetLastError(), "wgrad");
    hipLaunchKernelGGL(cvd::k_wgrad_reduce, dim3((nw + L.cout + 63) / 64), dim3(1024), 0, s,
                       d->wpart.p, splits, nw, L.cout, g.dw, g.db);
    HIPTRY(hipGetLastError(), "wgrad reduce");
    return conv_bwd_input(d, L, dz, gx, gxcs, accumulate, x, xcs, dz_next, npix, h, w, ug);
  }
  if (up) return fail(CV_ERR_INVALID_ARGUMENT, "upsampled input needs the halo weight gradient");
  // about 2048 blocks: pixel splits of a multiple of 64 pixels
  const uint32_t col_tiles = (9 * L.cin + cvd::kWgCols - 1) / cvd::kWgCols;
  uint32_t splits = std::max(1u, std::min(npix / 2048u, 2048u / col_tiles));
  g.pix_per_block = (npix / splits + cvd::kWgPix - 1) / cvd::kWgPix * cvd::kWgPix;
  splits = (npix + g.pix_per_block - 1) / g.pix_per_block;
  HIPTRY(d->wpart.reserve((size_t)splits * (nw + L.cout)), "alloc");
  g.part = d->wpart.p;
  dim3 grid(col_tiles, 1, splits);
  switch (L.cout_pad / 16) {
    case 1: hipLaunchKernelGGL(cvd::k_wgrad_t<1>, grid, dim3(256), 0, s, g); break;
    case 2: hipLaunchKernelGGL(cvd::k_wgrad_t<2>, grid, dim3(256), 0, s, g); break;
    case 4: hipLaunchKernelGGL(cvd::k_wgrad_t<4>, grid, dim3(256), 0, s, g); break;
    case 8: hipLaunchKernelGGL(cvd::k_wgrad_t<8>, grid, dim3(256), 0, s, g); break;
    default: return fail(CV_ERR_INVALID_ARGUMENT, "unsupported conv width");
  }
  HIPTRY(hipGetLastError(), "wgrad");
  hipLaunchKernelGGL(cvd::k_wgrad_reduce, dim3((nw + L.cout + 63) / 64), dim3(1024), 0, s,
                     d->wpart.p, splits, nw, L.cout, g.dw, g.db);
  HIPTRY(hipGetLastError(), "wgrad reduce");
  return conv_bwd_input(d, L, dz, gx, gxcs, accumulate, x, xcs, dz_next, npix, h, w, ug);
}

// dX of one convolution (+= into gx when accumulate): the convolution of
// dZ with the flipped weights.
cv_status conv_bwd_input(cv_denoiser* d, const Conv& L, const uint16_t* dz, float* gx,
                         uint32_t gxcs, bool accumulate, const uint16_t* x, uint32_t xcs,
                         uint16_t* dz_next, uint32_t npix, uint32_t h, uint32_t w,
                         const UpGrad* ug) {
  if (gx || dz_next) {
    cvd::ConvArgs a{};
    a.x = dz;
    a.xcs = L.cout_pad;
    a.kc = L.cout_pad;
    a.w = L.Wt.p;
    a.kpad = L.kpad_t;
    a.bias = nullptr;
    a.nout = L.cin;
    if (dz_next) {
      if (accumulate || L.cin % 16)
        return fail(CV_ERR_INVALID_ARGUMENT, "fused ReLU gradient needs a plain dX");
      a.y = dz_next;
      a.ycs = L.cin;
      a.mask = x;
      a.mcs = xcs;
    } else {
      a.yf = gx;
      a.yfcs = gxcs;
    }
    if (ug) {
      if (dz_next) return fail(CV_ERR_INVALID_ARGUMENT, "fused upsampling gradient needs a plain dX");
      a.ug_dz = ug->dz;
      a.ug_dzcs = ug->dzcs;
      a.ug_mask = ug->y;
      a.ug_mcs = ug->ycs;
      a.ug_ch = ug->ch;
    }
    a.accumulate = accumulate ? 1u : 0u;
    a.flip = 1;
    a.npix = npix;
    a.h = h;
    a.w_ = w;
    cv_status cs = launch_conv(d, a, round_up(L.cin, 16) / 16);
    if (cs != CV_OK) return cs;
  }
  return CV_OK;
}

// Activation buffers are indexed by pixel only (NHWC), so one allocation
// serves every shape with at most as many pixels; it grows, never shrinks.
cv_status reserve_acts(cv_denoiser* d, uint32_t nb, uint32_t h, uint32_t w) {
  const size_t need = (size_t)nb * h * w;
  if (need <= d->cap_pixels) {
    d->nb = nb;
    d->h = h;
    d->w = w;
    return CV_OK;
  }
  const size_t P0 = need, P1 = P0 / 4, P2 = P0 / 16;
  HIPTRY(d->x0.alloc(P0 * 16), "alloc");
  HIPTRY(d->e0a.alloc(P0 * 32), "alloc");
  HIPTRY(d->cat0.alloc(P0 * 96), "alloc");
  HIPTRY(d->p0.alloc(P1 * 32), "alloc");
  HIPTRY(d->e1a.alloc(P1 * 64), "alloc");
  HIPTRY(d->cat1.alloc(P1 * 192), "alloc");
  HIPTRY(d->p1.alloc(P2 * 64), "alloc");
  HIPTRY(d->b2a.alloc(P2 * 128), "alloc");
  HIPTRY(d->b2b.alloc(P2 * 128), "alloc");
  HIPTRY(d->d1a.alloc(P1 * 64), "alloc");
  HIPTRY(d->d1b.alloc(P1 * 64), "alloc");
  HIPTRY(d->d0a.alloc(P0 * 32), "alloc");
  HIPTRY(d->d0b.alloc(P0 * 32), "alloc");
  HIPTRY(d->dz.alloc(P0 * 128), "alloc");
  HIPTRY(d->dz2.alloc(P0 * 128), "alloc");
  HIPTRY(d->out.alloc(P0 * 3), "alloc");
  HIPTRY(d->dout.alloc(P0 * 3), "alloc");
  HIPTRY(d->g_cat0.alloc(P0 * 96), "alloc");
  HIPTRY(d->g_p0.alloc(P1 * 32), "alloc");
  HIPTRY(d->g_cat1.alloc(P1 * 192), "alloc");
  HIPTRY(d->g_p1.alloc(P2 * 64), "alloc");
  HIPTRY(d->g_b2b.alloc(P2 * 128), "alloc");
  HIPTRY(d->g_d1b.alloc(P1 * 64), "alloc");
  HIPTRY(d->loss.alloc(1 + kLossParts), "alloc");
  d->cap_pixels = need;
  d->nb = nb;
  d->h = h;
  d->w = w;
  return CV_OK;
}

cv_status pack_weights(cv_denoiser* d) {
  cvd::PackArgs pa{};
  uint32_t most = 0;
  for (int i = 0; i < 11; i++) {
    const Conv& L = d->conv[i];
    pa.l[i] = {d->params.p + L.w_off, L.W.p, L.Wt.p, L.cout, L.cin, L.cout_pad, L.kpad, L.kpad_t};
    most = std::max(most, L.cout_pad * L.kpad + L.cin * L.kpad_t);
  }
  // every layer in one launch: a row of blocks per layer
  hipLaunchKernelGGL(cvd::k_pack, dim3(std::min<uint32_t>((most + 1023) / 1024, 128u), 11),
                     dim3(256), 0, d->stream, pa);
  HIPTRY(hipGetLastError(), "pack");
  return CV_OK;
}

// Forward pass from d->x0 to `out` (f32 log1p radiance, [p][3]), d->out
// when null.  With `materialize` the upsampled halves of cat0 / cat1 are
// written out by k_upsample; otherwise the decoder convolutions read them
// from the lower level directly (the same values; a training step's weight
// gradients then read them there too).  With materialize_pool k_pool
// writes the pooled encoder outputs; otherwise the encoder convolutions'
// epilogues do.
cv_status forward(cv_denoiser* d, bool materialize, bool materialize_pool, float* out = nullptr) {
  const uint32_t H = d->h, W = d->w, H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4;
  auto& L = d->conv;
  hipStream_t s = d->stream;
  cv_status cs;
#define CVD_TRY(x) \
  if ((cs = (x)) != CV_OK) return cs
  // the pooled encoder outputs come from the convolutions' epilogues (the
  // same bits as k_pool, which the materialized path runs)
  const bool fuse_pool = !materialize_pool && d->halo_convs && W % (2 * cvd::kHaloW) == 0;
  CVD_TRY(conv_fwd(d, L[0], d->x0.p, 16, d->e0a.p, 32, nullptr, 0, H, W));
  if (fuse_pool) {
    CVD_TRY(conv_fwd(d, L[1], d->e0a.p, 32, d->cat0.p + 64, 96, nullptr, 0, H, W, nullptr, 0, 0,
                     d->p0.p, 32));
  } else {
    CVD_TRY(conv_fwd(d, L[1], d->e0a.p, 32, d->cat0.p + 64, 96, nullptr, 0, H, W));
    hipLaunchKernelGGL(cvd::k_pool, dim3(grid1((size_t)d->nb * H1 * W1 * 4)), dim3(256), 0, s,
                       d->cat0.p + 64, 96u, d->p0.p, 32u, 32u, d->nb, H1, W1);
  }
  CVD_TRY(conv_fwd(d, L[2], d->p0.p, 32, d->e1a.p, 64, nullptr, 0, H1, W1));
  if (fuse_pool) {
    CVD_TRY(conv_fwd(d, L[3], d->e1a.p, 64, d->cat1.p + 128, 192, nullptr, 0, H1, W1, nullptr, 0,
                     0, d->p1.p, 64));
  } else {
    CVD_TRY(conv_fwd(d, L[3], d->e1a.p, 64, d->cat1.p + 128, 192, nullptr, 0, H1, W1));
    hipLaunchKernelGGL(cvd::k_pool, dim3(grid1((size_t)d->nb * H2 * W2 * 8)), dim3(256), 0, s,
                       d->cat1.p + 128, 192u, d->p1.p, 64u, 64u, d->nb, H2, W2);
  }
  CVD_TRY(conv_fwd(d, L[4], d->p1.p, 64, d->b2a.p, 128, nullptr, 0, H2, W2));
  CVD_TRY(conv_fwd(d, L[5], d->b2a.p, 128, d->b2b.p, 128, nullptr, 0, H2, W2));
  const bool fuse_up = !materialize && d->halo_convs && W1 % cvd::kHaloW == 0;
  if (fuse_up) {
    CVD_TRY(conv_fwd(d, L[6], d->cat1.p, 192, d->d1a.p, 64, nullptr, 0, H1, W1, d->b2b.p, 128, 4));
  } else {
    hipLaunchKernelGGL(cvd::k_upsample, dim3(grid1((size_t)d->nb * H1 * W1 * 16)), dim3(256), 0,
                       s, d->b2b.p, 128u, d->cat1.p, 192u, 128u, d->nb, H1, W1);
    CVD_TRY(conv_fwd(d, L[6], d->cat1.p, 192, d->d1a.p, 64, nullptr, 0, H1, W1));
  }
  CVD_TRY(conv_fwd(d, L[7], d->d1a.p, 64, d->d1b.p, 64, nullptr, 0, H1, W1));
  if (fuse_up) {
    CVD_TRY(conv_fwd(d, L[8], d->cat0.p, 96, d->d0a.p, 32, nullptr, 0, H, W, d->d1b.p, 64, 2));
  } else {
    hipLaunchKernelGGL(cvd::k_upsample, dim3(grid1((size_t)d->nb * H * W * 8)), dim3(256), 0, s,
                       d->d1b.p, 64u, d->cat0.p, 96u, 64u, d->nb, H, W);
    CVD_TRY(conv_fwd(d, L[8], d->cat0.p, 96, d->d0a.p, 32, nullptr, 0, H, W));
  }
  CVD_TRY(conv_fwd(d, L[9], d->d0a.p, 32, d->d0b.p, 32, nullptr, 0, H, W));
  CVD_TRY(conv_fwd(d, L[10], d->d0b.p, 32, nullptr, 0, out ? out : d->out.p, 3, H, W));
  HIPTRY(hipGetLastError(), "forward");
  return CV_OK;
}

// Backward pass from d->dout to d->grads (zeroed here first).
// With `materialize` the upsampling gradients run as kernels of their own
// (k_upsample_grad, then the ReLU gradient of the layer below); otherwise
// the input-gradient convolutions of the two decoder layers above an
// upsampling write the layer below's dZ directly (the same bits).
cv_status backward(cv_denoiser* d, bool materialize) {
  const uint32_t H = d->h, W = d->w, H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4;
  auto& L = d->conv;
  hipStream_t s = d->stream;
  const uint32_t nb = d->nb;
  cv_status cs;
  hipLaunchKernelGGL(cvd::k_fill, dim3(grid1(d->nparams)), dim3(256), 0, s, d->grads.p,
                     d->nparams, 0.0f);
  // Where a layer's input is the ReLU output of the layer below (same
  // buffer, same channels), the dX convolution writes that layer's dZ
  // directly (dz_next); dZ alternates between the two buffers.
  uint16_t* const dzA = d->dz.p;
  uint16_t* const dzB = d->dz2.p;
  // out <- d0b
  CVD_TRY(conv_bwd(d, L[10], d->d0b.p, 32, nullptr, 0, d->dout.p, 3, dzA, false, nullptr, 0,
                   false, dzB, H, W));
  CVD_TRY(conv_bwd(d, L[9], d->d0a.p, 32, d->d0b.p, 32, nullptr, 0, dzB, true, nullptr, 0, false,
                   dzA, H, W));
  const bool fuse = !materialize && d->halo_convs;
  // cat0[0:64] = up(d1b): with fuse, L8's dX writes L7's dZ into dzB
  const UpGrad ug0{dzB, 64, d->d1b.p, 64, 64};
  // with fuse the forward did not write cat0[0:64]: the weight gradient
  // reads the upsampled channels from d1b
  const UpIn up0{d->d1b.p, 64, 64};
  CVD_TRY(conv_bwd(d, L[8], d->cat0.p, 96, d->d0a.p, 32, nullptr, 0, dzA, true, d->g_cat0.p, 96,
                   false, nullptr, H, W, fuse ? &ug0 : nullptr, fuse ? &up0 : nullptr));
  if (!fuse)
    hipLaunchKernelGGL(cvd::k_upsample_grad, dim3(grid1((size_t)nb * H1 * W1 * 16)), dim3(256), 0,
                       s, d->g_cat0.p, 96u, d->g_d1b.p, 64u, 64u, nb, H1, W1);
  CVD_TRY(conv_bwd(d, L[7], d->d1a.p, 64, d->d1b.p, 64, d->g_d1b.p, 64, dzB, fuse, nullptr, 0,
                   false, dzA, H1, W1));
  // cat1[0:128] = up(b2b): with fuse, L6's dX writes L5's dZ into dzB
  const UpGrad ug1{dzB, 128, d->b2b.p, 128, 128};
  const UpIn up1{d->b2b.p, 128, 128};
  CVD_TRY(conv_bwd(d, L[6], d->cat1.p, 192, d->d1a.p, 64, nullptr, 0, dzA, true, d->g_cat1.p, 192,
                   false, nullptr, H1, W1, fuse ? &ug1 : nullptr, fuse ? &up1 : nullptr));
  if (!fuse)
    hipLaunchKernelGGL(cvd::k_upsample_grad, dim3(grid1((size_t)nb * H2 * W2 * 32)), dim3(256), 0,
                       s, d->g_cat1.p, 192u, d->g_b2b.p, 128u, 128u, nb, H2, W2);
  CVD_TRY(conv_bwd(d, L[5], d->b2a.p, 128, d->b2b.p, 128, d->g_b2b.p, 128, dzB, fuse, nullptr, 0,
                   false, dzA, H2, W2));
  CVD_TRY(conv_bwd(d, L[4], d->p1.p, 64, d->b2a.p, 128, nullptr, 0, dzA, true, d->g_p1.p, 64,
                   false, nullptr, H2, W2));
  // p1 = pool(cat1[128:192]): its gradient joins the skip gradient in g_cat1
  // (with fuse, in the same pass as L3's ReLU gradient, into dzB)
  if (fuse)
    hipLaunchKernelGGL(cvd::k_pool_relu_grad, dim3(grid1((size_t)nb * H1 * W1 * 8)), dim3(256), 0,
                       s, d->g_cat1.p + 128, 192u, d->g_p1.p, 64u, d->cat1.p + 128, 192u, dzB,
                       64u, nb, H2, W2);
  else
    hipLaunchKernelGGL(cvd::k_pool_grad, dim3(grid1((size_t)nb * H2 * W2 * 16)), dim3(256), 0, s,
                       d->g_p1.p, 64u, d->g_cat1.p + 128, 192u, 64u, nb, H2, W2);
  CVD_TRY(conv_bwd(d, L[3], d->e1a.p, 64, d->cat1.p + 128, 192, d->g_cat1.p + 128, 192, dzB,
                   fuse, nullptr, 0, false, dzA, H1, W1));
  CVD_TRY(conv_bwd(d, L[2], d->p0.p, 32, d->e1a.p, 64, nullptr, 0, dzA, true, d->g_p0.p, 32,
                   false, nullptr, H1, W1));
  if (fuse)
    hipLaunchKernelGGL(cvd::k_pool_relu_grad, dim3(grid1((size_t)nb * H * W * 4)), dim3(256), 0, s,
                       d->g_cat0.p + 64, 96u, d->g_p0.p, 32u, d->cat0.p + 64, 96u, dzB, 32u, nb,
                       H1, W1);
  else
    hipLaunchKernelGGL(cvd::k_pool_grad, dim3(grid1((size_t)nb * H1 * W1 * 8)), dim3(256), 0, s,
                       d->g_p0.p, 32u, d->g_cat0.p + 64, 96u, 32u, nb, H1, W1);
  CVD_TRY(conv_bwd(d, L[1], d->e0a.p, 32, d->cat0.p + 64, 96, d->g_cat0.p + 64, 96, dzB, fuse,
                   nullptr, 0, false, dzA, H, W));
  CVD_TRY(conv_bwd(d, L[0], d->x0.p, 16, d->e0a.p, 32, nullptr, 0, dzA, true, nullptr, 0, false,
                   nullptr, H, W));
  HIPTRY(hipGetLastError(), "backward");
  return CV_OK;
#undef CVD_TRY
}

cv_status check_shape(uint32_t nb, uint32_t h, uint32_t w) {
  if (nb == 0 || h == 0 || w == 0 || (h & 3) || (w & 63))
    return fail(CV_ERR_INVALID_ARGUMENT, "denoiser sizes: height a multiple of 4, width of 64");
  if ((uint64_t)nb * h * w >= (1ull << 28)) return fail(CV_ERR_INVALID_ARGUMENT, "batch too large");
  return CV_OK;
}

// Uploads a host or device f32 array into buf.
cv_status stage_in(cv_denoiser* d, DevBuf<float>& buf, const float* src, size_t n, bool dev,
                   const float** out) {
  if (dev) {
    *out = src;
    return CV_OK;
  }
  if (buf.n < n) HIPTRY(buf.alloc(n), "alloc");
  HIPTRY(hipMemcpyAsync(buf.p, src, n * 4, hipMemcpyHostToDevice, d->stream), "upload");
  *out = buf.p;
  return CV_OK;
}

}  // namespace

namespace cvh {

cv_status denoise_image(cv_denoiser* d, const float* hdr, const float* features, uint32_t width,
                        uint32_t height, uint32_t spp, float* dst, bool dst_device) {
  if (!d || !hdr || !features || !dst || spp == 0)
    return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  cv_status cs = check_shape(1, height, width);
  if (cs != CV_OK) return cs;
  std::lock_guard<std::mutex> g(d->mu);
  if ((cs = reserve_acts(d, 1, height, width)) != CV_OK) return cs;
  const size_t P = (size_t)width * height;
  if (!dst_device) HIPTRY(d->tgt_f32.reserve(P * 3), "alloc");
  hipStream_t s = d->stream;
  // the features straight into the network's bf16 input
  hipLaunchKernelGGL(cvd::k_features, dim3(grid1(P)), dim3(256), 0, s, hdr, features, width,
                     height, spp, nullptr, d->x0.p);
  if ((cs = forward(d, false, false)) != CV_OK) return cs;
  // a device destination receives the HDR image directly
  hipLaunchKernelGGL(cvd::k_output, dim3(grid1(P)), dim3(256), 0, s, d->out.p, width, height, spp,
                     dst_device ? dst : d->tgt_f32.p);
  HIPTRY(hipGetLastError(), "denoise");
  if (!dst_device)
    HIPTRY(hipMemcpyAsync(dst, d->tgt_f32.p, P * 12, hipMemcpyDeviceToHost, s), "download");
  HIPTRY(hipStreamSynchronize(s), "sync");
  return CV_OK;
}

}  // namespace cvh

extern "C" {

size_t cv_denoiser_num_params(void) {
  size_t n = 0;
  for (auto& l : kLayers) n += (size_t)l[1] * 9 * l[0] + l[1];
  return n;
}

cv_status cv_denoiser_create(uint64_t seed, cv_denoiser** out) {
  if (!out) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  cv_status cs = cvh::ensure_device();
  if (cs != CV_OK) return cs;
  cv_denoiser* d = new cv_denoiser();
  auto bail = [&](cv_status st) {
    cv_denoiser_destroy(d);
    return st;
  };
  hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreate"));
  size_t off = 0;
  for (int i = 0; i < 11; i++) {
    Conv& L = d->conv[i];
    L.cin = kLayers[i][0];
    L.cout = kLayers[i][1];
    L.relu = kLayers[i][2] != 0;
    L.cout_pad = round_up(L.cout, 16);
    L.kpad = round_up(9 * L.cin, 32);
    L.kpad_t = round_up(9 * L.cout_pad, 32);
    L.w_off = off;
    off += (size_t)L.cout * 9 * L.cin;
    L.b_off = off;
    off += L.cout;
    if ((e = L.W.alloc((size_t)L.cout_pad * L.kpad)) || (e = L.Wt.alloc((size_t)L.cin * L.kpad_t)))
      return bail(hip_fail(e, "alloc weights"));
  }
  d->nparams = off;
  if ((e = d->params.alloc(off)) || (e = d->grads.alloc(off)) || (e = d->m.alloc(off)) ||
      (e = d->v.alloc(off)))
    return bail(hip_fail(e, "alloc parameters"));
  // He-normal initialisation from a splitmix64 stream (Box-Muller), zero bias
  std::vector<float> h(off, 0.0f);
  uint64_t z = seed;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  };
  for (const Conv& L : d->conv) {
    const float sd = sqrtf(2.0f / (9.0f * (float)(L.cin == 16 ? 10 : L.cin)));
    for (size_t k = 0; k < (size_t)L.cout * 9 * L.cin; k++) {
      double u1 = ((next() >> 11) + 1) * (1.0 / 9007199254740993.0);
      double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
      h[L.w_off + k] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
  }
  if ((e = hipMemcpyAsync(d->params.p, h.data(), off * 4, hipMemcpyHostToDevice, d->stream)))
    return bail(hip_fail(e, "upload"));
  hipLaunchKernelGGL(cvd::k_fill, dim3(grid1(off)), dim3(256), 0, d->stream, d->m.p, off, 0.0f);
  hipLaunchKernelGGL(cvd::k_fill, dim3(grid1(off)), dim3(256), 0, d->stream, d->v.p, off, 0.0f);
  if ((cs = pack_weights(d)) != CV_OK) return bail(cs);
  if ((e = hipStreamSynchronize(d->stream))) return bail(hip_fail(e, "init"));
  *out = d;
  return CV_OK;
}

void cv_denoiser_destroy(cv_denoiser* d) {
  if (!d) return;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
}

cv_status cv_denoiser_get_params(cv_denoiser* d, float* params, float* grads) {
  if (!d) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(d->mu);
  if (params)
    HIPTRY(hipMemcpyAsync(params, d->params.p, d->nparams * 4, hipMemcpyDeviceToHost, d->stream),
           "download");
  if (grads)
    HIPTRY(hipMemcpyAsync(grads, d->grads.p, d->nparams * 4, hipMemcpyDeviceToHost, d->stream),
           "download");
  HIPTRY(hipStreamSynchronize(d->stream), "sync");
  return CV_OK;
}

cv_status cv_denoiser_set_params(cv_denoiser* d, const float* params) {
  if (!d || !params) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(d->mu);
  HIPTRY(hipMemcpyAsync(d->params.p, params, d->nparams * 4, hipMemcpyHostToDevice, d->stream),
         "upload");
  hipLaunchKernelGGL(cvd::k_fill, dim3(grid1(d->nparams)), dim3(256), 0, d->stream, d->m.p,
                     d->nparams, 0.0f);
  hipLaunchKernelGGL(cvd::k_fill, dim3(grid1(d->nparams)), dim3(256), 0, d->stream, d->v.p,
                     d->nparams, 0.0f);
  d->step = 0;
  cv_status cs = pack_weights(d);
  if (cs != CV_OK) return cs;
  HIPTRY(hipStreamSynchronize(d->stream), "sync");
  return CV_OK;
}

cv_status cv_denoiser_forward(cv_denoiser* d, const float* input, float* output, uint32_t nb,
                              uint32_t h, uint32_t w, uint32_t flags) {
  if (!d || !input || !output) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  cv_status cs = check_shape(nb, h, w);
  if (cs != CV_OK) return cs;
  std::lock_guard<std::mutex> g(d->mu);
  const bool dev = flags & CV_PTR_DEVICE;
  if ((cs = reserve_acts(d, nb, h, w)) != CV_OK) return cs;
  const size_t P = (size_t)nb * h * w;
  const float* in = nullptr;
  if ((cs = stage_in(d, d->in_f32, input, P * 10, dev, &in)) != CV_OK) return cs;
  hipLaunchKernelGGL(cvd::k_to_input, dim3(grid1(P)), dim3(256), 0, d->stream, in, d->x0.p,
                     (uint32_t)P);
  const bool mat = flags & CV_DENOISE_MATERIALIZE;
  // device output: the last layer writes it directly
  if ((cs = forward(d, mat, mat, dev ? output : nullptr)) != CV_OK) return cs;
  if (!dev)
    HIPTRY(hipMemcpyAsync(output, d->out.p, P * 12, hipMemcpyDeviceToHost, d->stream),
           "download");
  HIPTRY(hipStreamSynchronize(d->stream), "sync");
  return CV_OK;
}

cv_status cv_denoiser_train_step(cv_denoiser* d, const float* input, const float* target,
                                 uint32_t nb, uint32_t h, uint32_t w, float lr, uint32_t flags,
                                 float* loss) {
  if (!d || !input || !target) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  cv_status cs = check_shape(nb, h, w);
  if (cs != CV_OK) return cs;
  std::lock_guard<std::mutex> g(d->mu);
  const bool dev = flags & CV_PTR_DEVICE;
  if ((cs = reserve_acts(d, nb, h, w)) != CV_OK) return cs;
  const size_t P = (size_t)nb * h * w;
  const float *in = nullptr, *tg = nullptr;
  if ((cs = stage_in(d, d->in_f32, input, P * 10, dev, &in)) != CV_OK) return cs;
  if ((cs = stage_in(d, d->tgt_f32, target, P * 3, dev, &tg)) != CV_OK) return cs;
  hipStream_t s = d->stream;
  hipLaunchKernelGGL(cvd::k_to_input, dim3(grid1(P)), dim3(256), 0, s, in, d->x0.p,
                     (uint32_t)P);
  const bool mat = flags & CV_DENOISE_MATERIALIZE;
  if ((cs = forward(d, mat, mat)) != CV_OK) return cs;
  // the loss: block partials after the sum's word, then their sum
  const uint32_t lblocks = (uint32_t)std::min<size_t>(grid1(P * 3), kLossParts);
  hipLaunchKernelGGL(cvd::k_l1_loss, dim3(lblocks), dim3(256), 0, s, d->out.p, tg, d->dout.p,
                     (uint32_t)P, d->loss.p + 1);
  hipLaunchKernelGGL(cvd::k_sum_parts, dim3(1), dim3(256), 0, s, d->loss.p + 1, lblocks,
                     d->loss.p);
  if ((cs = backward(d, mat)) != CV_OK) return cs;
  if (lr > 0.0f) {
    d->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)d->step), c2 = 1.0f - powf(b2, (float)d->step);
    hipLaunchKernelGGL(cvd::k_adam, dim3(grid1(d->nparams)), dim3(256), 0, s, d->params.p,
                       d->grads.p, d->m.p, d->v.p, (uint32_t)d->nparams, lr, b1, b2, 1e-8f, c1, c2);
    HIPTRY(hipGetLastError(), "adam");
    if ((cs = pack_weights(d)) != CV_OK) return cs;
  }
  float l = 0.0f;
  HIPTRY(hipMemcpyAsync(&l, d->loss.p, 4, hipMemcpyDeviceToHost, s), "download");
  HIPTRY(hipStreamSynchronize(s), "sync");
  if (loss) *loss = l;
  return CV_OK;
}

cv_status cv_denoiser_prepare_input(const float* hdr, const float* features, uint32_t width,
                                    uint32_t height, uint32_t spp, float* input, uint32_t flags) {
  if (!hdr || !features || !input || spp == 0) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  if ((width & 15) || (height & 3)) return fail(CV_ERR_NOT_MULTIPLE_OF_16, "block layout sizes");
  cv_status cs = cvh::ensure_device();
  if (cs != CV_OK) return cs;
  const bool dev = flags & CV_PTR_DEVICE;
  const size_t P = (size_t)width * height;
  DevBuf<float> h, f, o;
  const float *hp = hdr, *fp = features;
  float* op = input;
  if (!dev) {
    HIPTRY(h.alloc(P * 3), "alloc");
    HIPTRY(f.alloc(P * 8), "alloc");
    HIPTRY(o.alloc(P * 10), "alloc");
    HIPTRY(hipMemcpy(h.p, hdr, P * 12, hipMemcpyHostToDevice), "upload");
    HIPTRY(hipMemcpy(f.p, features, P * 32, hipMemcpyHostToDevice), "upload");
    hp = h.p;
    fp = f.p;
    op = o.p;
  }
  hipLaunchKernelGGL(cvd::k_features, dim3(grid1(P)), dim3(256), 0, 0, hp, fp, width, height, spp,
                     op, nullptr);
  HIPTRY(hipGetLastError(), "features");
  if (!dev) HIPTRY(hipMemcpy(input, op, P * 40, hipMemcpyDeviceToHost), "download");
  HIPTRY(hipDeviceSynchronize(), "sync");
  return CV_OK;
}

cv_status cv_denoiser_output_to_hdr(const float* output, uint32_t width, uint32_t height,
                                    uint32_t spp, float* hdr, uint32_t flags) {
  if (!output || !hdr || spp == 0) return fail(CV_ERR_INVALID_ARGUMENT, "null argument");
  if ((width & 15) || (height & 3)) return fail(CV_ERR_NOT_MULTIPLE_OF_16, "block layout sizes");
  cv_status cs = cvh::ensure_device();
  if (cs != CV_OK) return cs;
  const bool dev = flags & CV_PTR_DEVICE;
  const size_t P = (size_t)width * height;
  DevBuf<float> o, h;
  const float* op = output;
  float* hp = hdr;
  if (!dev) {
    HIPTRY(o.alloc(P * 3), "alloc");
    HIPTRY(h.alloc(P * 3), "alloc");
    HIPTRY(hipMemcpy(o.p, output, P * 12, hipMemcpyHostToDevice), "upload");
    op = o.p;
    hp = h.p;
  }
  hipLaunchKernelGGL(cvd::k_output, dim3(grid1(P)), dim3(256), 0, 0, op, width, height, spp, hp);
  HIPTRY(hipGetLastError(), "output");
  if (!dev) HIPTRY(hipMemcpy(hdr, hp, P * 12, hipMemcpyDeviceToHost), "download");
  HIPTRY(hipDeviceSynchronize(), "sync");
  return CV_OK;
}

}  // extern "C"
