// k_cnn.hip -- the proposal network: forward, backward and Adam on gfx950.
//
// Every matrix stage is one instantiation of k_mm, a wave-tiled bfloat16
// MFMA product C[M][N] = sum_k A[m][k] B[k][n] (v_mfma_f32_16x16x32_bf16,
// float32 accumulation).  The operands are not materialised: each stage's
// A and B functors gather them, four consecutive k at a time (implicit
// im2col for the convolutions, the transposed convolution for the data
// gradients), and its C functor applies the epilogue (bias + ReLU +
// bfloat16 store, ReLU-masked gradient store, or a float32 split-K partial).
// Where every wave of a block shares its B columns (the convolutions'
// weights), the block stages B in LDS once.
//
// Register layout of v_mfma_f32_16x16x32_bf16 for lane l (0..63), with
// r = l % 16 and g = l / 16 (checked against a CPU product on gfx950):
//   A operand: A[r][8g..8g+7]    B operand: B[8g..8g+7][r]
//   result:    C[4g..4g+3][r]
//
// Activations are NHWC (channels contiguous), bfloat16.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "hd/proposal.h"
#include "hd_device.hpp"
#include "hd_internal.hpp"

namespace hd {
namespace cnn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TILE = 64, CH = 32, HID = 64, FEAT = 16 * CH;
constexpr int CIN1 = 2;         // input channels: coarse plane, 4x wider time scale
constexpr int KL1 = 9 * CIN1;   // layer-1 reduction (tap*2 + channel)
constexpr int K1 = 32;          // ... padded to the MFMA step
constexpr int KC = 9 * CH;      // 288

// Flat float32 parameter offsets.
constexpr size_t O_W1 = 0, O_B1 = O_W1 + CH * KL1;
constexpr size_t O_W2 = O_B1 + CH, O_B2 = O_W2 + CH * KC;
constexpr size_t O_W3 = O_B2 + CH, O_B3 = O_W3 + CH * KC;
constexpr size_t O_W4 = O_B3 + CH, O_B4 = O_W4 + CH * KC;
constexpr size_t O_F1W = O_B4 + CH, O_F1B = O_F1W + (size_t)HID * FEAT;
constexpr size_t O_F2W = O_F1B + HID, O_F2B = O_F2W + HID;
constexpr size_t NPARAM = O_F2B + 1;
static_assert(NPARAM == 61249, "parameter count");

// bfloat16 weight copies (element offsets in one uint16 array).
constexpr size_t Q_W1 = 0;                          // [32][32]
constexpr size_t Q_W2 = Q_W1 + CH * K1;             // [32][288] x3
constexpr size_t Q_W3 = Q_W2 + CH * KC;
constexpr size_t Q_W4 = Q_W3 + CH * KC;
constexpr size_t Q_T2 = Q_W4 + CH * KC;             // [9][32 in][32 out] x3
constexpr size_t Q_T3 = Q_T2 + CH * KC;
constexpr size_t Q_T4 = Q_T3 + CH * KC;
constexpr size_t Q_F1 = Q_T4 + CH * KC;             // [64][512]
constexpr size_t Q_F1T = Q_F1 + (size_t)HID * FEAT; // [512][64]
constexpr size_t Q_F2 = Q_F1T + (size_t)HID * FEAT; // [64]
constexpr size_t NQ = Q_F2 + HID;

__device__ __forceinline__ bf16x4 zero4() {
  bf16x4 v;
  v[0] = v[1] = v[2] = v[3] = 0;
  return v;
}
__device__ __forceinline__ bf16x4 load4(const uint16_t* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}
constexpr uint16_t BF_ONE = 0x3F80;

// ---------------------------------------------------------------------------
// The generic product, on gfx950's K = 32 step (v_mfma_f32_16x16x32_bf16):
// lane l = 16 g + r holds A[r][8g .. 8g+7] and B[8g .. 8g+7][r] of the step,
// and gets C[4g .. 4g+3][r] (the result layout of the K = 16 form).  The
// 8-element operands are two 4-element functor gathers.
//
// A block is four waves; wave w owns output tile blockIdx.x*4 + w of
// (16*TM) x (16*TN) elements, tiles ordered n fastest.  blockIdx.y selects a
// K chunk [y*kchunk, min((y+1)*kchunk, K)) (split-K).  K and kchunk are
// multiples of 16 (the second half of a step past the chunk's end is zero,
// and its gathers are not made); the functors return zeros past their data,
// and the C functor drops rows >= M and columns >= N.
//
// STAGE_B: every wave of the block has the same columns (N <= 16*TN), so the
// block gathers its B chunk once into LDS as [column][k] rows of kchunk + 8
// elements (16-byte aligned rows; 16-byte reads of 16 rows hit distinct
// banks) and the waves read their B operands there: the convolutions'
// weights, shared by the block's 4 * TM row tiles.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 cat8(bf16x4 lo, bf16x4 hi) {
  const s16x8 v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8, v);
}

template <int TM, int TN, bool STAGE_B, class P>
__global__ __launch_bounds__(256) void k_mm(P p) {
  extern __shared__ uint16_t bsh[];  // STAGE_B: [16*TN][kchunk + 8]
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const long tiles_n = (p.N + 16 * TN - 1) / (16 * TN);
  const long kb = (long)blockIdx.y * p.kchunk;
  const long ke = kb + p.kchunk < p.K ? kb + p.kchunk : p.K;
  const int ldb = (int)p.kchunk + 8;
  if (STAGE_B) {  // the block's columns are 0 .. 16*TN-1 (tiles_n == 1)
    const int items = 16 * TN * (int)(p.kchunk / 4);
    for (int it = threadIdx.x; it < items; it += 256) {
      const int n = it % (16 * TN), kq = it / (16 * TN);
      const long k = kb + 4 * kq;
      *reinterpret_cast<bf16x4*>(bsh + n * ldb + 4 * kq) = k < ke ? p.B(k, n) : zero4();
    }
    __syncthreads();
  }
  const long tile = (long)blockIdx.x * 4 + wave;
  const long tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const long m0 = tm * 16 * TM, n0 = tn * 16 * TN;
  if (m0 >= p.M) return;  // wave-uniform, after the block's only barrier
  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long k0 = kb; k0 < ke; k0 += 32) {
    const long k = k0 + 8 * g;
    const bool in = k < ke;  // ke - k0 is a multiple of 16: all 8 or none
    bf16x8 a[TM], b[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const long m = m0 + 16 * i + r;
      a[i] = in ? cat8(p.A(m, k), p.A(m, k + 4)) : cat8(zero4(), zero4());
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      if (STAGE_B) {
        b[j] = in ? __builtin_bit_cast(bf16x8, *reinterpret_cast<const s16x8*>(
                                                   bsh + (16 * j + r) * ldb + (k - kb)))
                  : cat8(zero4(), zero4());
      } else {
        const long n = n0 + 16 * j + r;
        b[j] = in ? cat8(p.B(k, n), p.B(k + 4, n)) : cat8(zero4(), zero4());
      }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        p.C(m0 + 16 * i + 4 * g + e, n0 + 16 * j + r, acc[i][j][e], (int)blockIdx.y);
}

template <int TM, int TN, class P>
hipError_t run_mm(const P& p, hipStream_t s) {
  const long tiles = ((p.M + 16 * TM - 1) / (16 * TM)) * ((p.N + 16 * TN - 1) / (16 * TN));
  const long splits = (p.K + p.kchunk - 1) / p.kchunk;
  if (tiles <= 0 || splits <= 0) return hipSuccess;
  if (p.K % 16 || p.kchunk % 16) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)splits);
  const size_t lds = (size_t)16 * TN * (p.kchunk + 8) * sizeof(uint16_t);
  // (one step of K = 32 reads each B element once: staging it would only add a barrier)
  if (p.N <= 16 * TN && p.kchunk >= 64 && lds <= 64 * 1024)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mm<TM, TN, true, P>), grid, dim3(256), lds, s, p);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mm<TM, TN, false, P>), grid, dim3(256), 0, s, p);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Forward stages.

// Layer 1: x [B][64][64][2] -> y1 [B][32][32][32].  k = tap*2 + channel (18 of 32 used).
struct Conv1Fwd {
  long M, N, K, kchunk;
  const uint16_t* x;
  const uint16_t* w;    // [32][16]
  const float* bias;
  uint16_t* y;
  __device__ bf16x4 A(long m, long k) const {
    if (m >= M) return zero4();
    const long b = m >> 10;
    const int oy = (int)((m >> 5) & 31), ox = (int)(m & 31);
    bf16x4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int kk = (int)k + e;
      const int tap = kk >> 1, ci = kk & 1;
      uint16_t h = 0;
      if (tap < 9) {
        const int iy = 2 * oy - 1 + tap / 3, ix = 2 * ox - 1 + tap % 3;
        if (iy >= 0 && iy < TILE && ix >= 0 && ix < TILE)
          h = x[((b * TILE + iy) * TILE + ix) * CIN1 + ci];
      }
      v[e] = (short)h;
    }
    return v;
  }
  __device__ bf16x4 B(long k, long n) const { return load4(w + n * K1 + k); }
  __device__ void C(long m, long n, float v, int) const {
    if (m >= M) return;
    const float t = v + bias[n];
    y[m * CH + n] = f2bf(t > 0.f ? t : 0.f);
  }
};

// Layers 2-4: x [B][H][H][32] -> y [B][H/2][H/2][32].  k = tap*32 + ci.
struct ConvFwd {
  long M, N, K, kchunk;
  int H, Ho;
  const uint16_t* x;
  const uint16_t* w;    // [32][288]
  const float* bias;
  uint16_t* y;
  __device__ bf16x4 A(long m, long k) const {
    if (m >= M) return zero4();
    const long b = m / (Ho * Ho);
    const int rem = (int)(m - b * Ho * Ho);
    const int oy = rem / Ho, ox = rem % Ho;
    const int tap = (int)(k >> 5), ci = (int)(k & 31);
    const int iy = 2 * oy - 1 + tap / 3, ix = 2 * ox - 1 + tap % 3;
    if (iy < 0 || iy >= H || ix < 0 || ix >= H) return zero4();
    return load4(x + ((b * H + iy) * H + ix) * CH + ci);
  }
  __device__ bf16x4 B(long k, long n) const { return load4(w + n * KC + k); }
  __device__ void C(long m, long n, float v, int) const {
    if (m >= M) return;
    const float t = v + bias[n];
    y[m * CH + n] = f2bf(t > 0.f ? t : 0.f);
  }
};

// fc1: f [B][512] -> h [B][64].
struct Fc1Fwd {
  long M, N, K, kchunk;
  const uint16_t* f;
  const uint16_t* w;    // [64][512]
  const float* bias;
  uint16_t* h;
  __device__ bf16x4 A(long m, long k) const {
    if (m >= M) return zero4();
    return load4(f + m * FEAT + k);
  }
  __device__ bf16x4 B(long k, long n) const { return load4(w + n * FEAT + k); }
  __device__ void C(long m, long n, float v, int) const {
    if (m >= M) return;
    const float t = v + bias[n];
    h[m * HID + n] = f2bf(t > 0.f ? t : 0.f);
  }
};

// fc2: one thread per tile, float32 accumulation in index order.
__global__ void k_fc2(const uint16_t* __restrict__ h, const uint16_t* __restrict__ w2,
                      const float* __restrict__ b2, int n, float* __restrict__ z) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  float s = 0.f;
  for (int j = 0; j < HID; ++j) {
    const float p = bf2f(h[(size_t)b * HID + j]) * bf2f(w2[j]);
    s = s + p;
  }
  z[b] = s + b2[0];
}

// ---------------------------------------------------------------------------
// Backward stages.

// Loss head, one block: dz = w (sigmoid(z) - y) / sum(w) for every tile,
// the loss, and the fc2 gradients (float32, fixed order).
constexpr int HEAD_BS = 256;
__global__ __launch_bounds__(HEAD_BS) void k_head(const float* __restrict__ z,
                                                  const float* __restrict__ lab,
                                                  const float* __restrict__ wt, int n,
                                                  const uint16_t* __restrict__ h,
                                                  float* __restrict__ dz,
                                                  float* __restrict__ loss_out,
                                                  float* __restrict__ g_f2w,
                                                  float* __restrict__ g_f2b) {
  __shared__ float red[HEAD_BS];
  __shared__ float red2[HEAD_BS];
  float sw = 0.f, sl = 0.f;
  for (int b = threadIdx.x; b < n; b += HEAD_BS) {
    const float zz = z[b];
    sw += wt[b];
    const float sp = fmaxf(zz, 0.f) + log1pf(expf(-fabsf(zz)));
    sl += wt[b] * (sp - lab[b] * zz);
  }
  red[threadIdx.x] = sw;
  red2[threadIdx.x] = sl;
  __syncthreads();
  for (int o = HEAD_BS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[threadIdx.x] += red[threadIdx.x + o];
      red2[threadIdx.x] += red2[threadIdx.x + o];
    }
    __syncthreads();
  }
  const float W = red[0];
  const float inv = W > 0.f ? 1.f / W : 0.f;
  if (threadIdx.x == 0) loss_out[0] = red2[0] * inv;
  for (int b = threadIdx.x; b < n; b += HEAD_BS) {
    const float sg = 1.f / (1.f + expf(-z[b]));
    dz[b] = wt[b] * (sg - lab[b]) * inv;
  }
  __syncthreads();
  if (threadIdx.x < HID) {
    const int j = threadIdx.x;
    float g = 0.f;
    for (int b = 0; b < n; ++b) g += dz[b] * bf2f(h[(size_t)b * HID + j]);
    g_f2w[j] = g;
  } else if (threadIdx.x == HID) {
    float g = 0.f;
    for (int b = 0; b < n; ++b) g += dz[b];
    g_f2b[0] = g;
  }
}

// dh = dz * w2 where h > 0 (bfloat16).
__global__ void k_dh(const float* __restrict__ dz, const uint16_t* __restrict__ w2,
                     const uint16_t* __restrict__ h, int n, uint16_t* __restrict__ dh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * HID) return;
  const int b = i / HID, j = i % HID;
  const float v = dz[b] * bf2f(w2[j]);
  dh[i] = h[i] != 0 && !(h[i] & 0x8000) ? f2bf(v) : (uint16_t)0;
}

// Weight gradients of fc1 with the bias as column 512: M = 64, N = 513,
// K = tiles; float32 split-K partials [split][64][NP].
struct Fc1Wgrad {
  long M, N, K, kchunk;
  int n;                 // tiles
  long NP;               // partial row length
  const uint16_t* dh;    // [n][64]
  const uint16_t* f;     // [n][512]
  float* part;
  __device__ bf16x4 A(long m, long k) const {
    bf16x4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = (short)((k + e < n) ? dh[(k + e) * HID + m] : 0);
    return v;
  }
  __device__ bf16x4 B(long k, long col) const {
    bf16x4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      uint16_t x = 0;
      if (k + e < n) x = col < FEAT ? f[(k + e) * FEAT + col] : (col == FEAT ? BF_ONE : 0);
      v[e] = (short)x;
    }
    return v;
  }
  __device__ void C(long m, long col, float v, int split) const {
    if (col < N) part[((size_t)split * M + m) * NP + col] = v;
  }
};

// Data gradient of fc1, masked by layer 4's ReLU: M = tiles, N = 512, K = 64.
struct Fc1Dgrad {
  long M, N, K, kchunk;
  const uint16_t* dh;    // [n][64]
  const uint16_t* wt;    // [512][64]
  const uint16_t* f;     // y4 [n][512]
  uint16_t* d4;
  __device__ bf16x4 A(long m, long k) const {
    if (m >= M) return zero4();
    return load4(dh + m * HID + k);
  }
  __device__ bf16x4 B(long k, long n) const { return load4(wt + n * HID + k); }
  __device__ void C(long m, long n, float v, int) const {
    if (m >= M) return;
    const uint16_t y = f[m * FEAT + n];
    d4[m * FEAT + n] = (y != 0 && !(y & 0x8000)) ? f2bf(v) : (uint16_t)0;
  }
};

// Data gradient of conv layer l (input side H): M = B*H*H input pixels,
// N = 32 input channels, K = 288 (tap*32 + output channel), masked by the
// ReLU of the layer's input activation.
struct ConvDgrad {
  long M, N, K, kchunk;
  int H, Ho;
  const uint16_t* d;     // gradient at the layer output [B][Ho][Ho][32]
  const uint16_t* wt;    // [9][32 in][32 out]
  const uint16_t* xin;   // layer input activation [B][H][H][32]
  uint16_t* dx;
  __device__ bf16x4 A(long m, long k) const {
    if (m >= M) return zero4();
    const long b = m / (H * H);
    const int rem = (int)(m - b * H * H);
    const int iy = rem / H, ix = rem % H;
    const int tap = (int)(k >> 5), co = (int)(k & 31);
    const int ty = iy + 1 - tap / 3, tx = ix + 1 - tap % 3;
    if (ty < 0 || tx < 0 || (ty & 1) || (tx & 1)) return zero4();
    const int oy = ty >> 1, ox = tx >> 1;
    if (oy >= Ho || ox >= Ho) return zero4();
    return load4(d + ((b * Ho + oy) * Ho + ox) * CH + co);
  }
  __device__ bf16x4 B(long k, long n) const {
    const int tap = (int)(k >> 5), co = (int)(k & 31);
    return load4(wt + ((size_t)tap * CH + n) * CH + co);
  }
  __device__ void C(long m, long n, float v, int) const {
    if (m >= M) return;
    const uint16_t y = xin[m * CH + n];
    dx[m * CH + n] = (y != 0 && !(y & 0x8000)) ? f2bf(v) : (uint16_t)0;
  }
};

// Weight gradients of conv layer l with the bias as column KL: M = 32 output
// channels, N = KL + 1, K = pixels of the layer output; float32 partials.
// CIN == 2 reads the input tiles, CIN == 32 an NHWC activation.
template <int CIN>
struct ConvWgrad {
  long M, N, K, kchunk;
  int H, Ho;
  long NP;
  const uint16_t* d;     // [B][Ho][Ho][32]
  const uint16_t* xin;   // CIN == 2: [B][64][64][2]; else [B][H][H][32]
  float* part;
  static constexpr int KL = 9 * CIN;
  __device__ bf16x4 A(long m, long k) const {
    bf16x4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = (short)d[(k + e) * CH + m];
    return v;
  }
  __device__ bf16x4 B(long k, long col) const {
    bf16x4 v;
    const int tap = (int)(col / CIN), ci = (int)(col % CIN);
    const int ky = tap / 3, kx = tap % 3;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      uint16_t x = 0;
      if (col < KL) {
        const long p = k + e;
        const long b = p / (Ho * Ho);
        const int rem = (int)(p - b * Ho * Ho);
        const int oy = rem / Ho, ox = rem % Ho;
        const int iy = 2 * oy - 1 + ky, ix = 2 * ox - 1 + kx;
        if (iy >= 0 && iy < H && ix >= 0 && ix < H)
          x = xin[((b * H + iy) * H + ix) * CIN + ci];
      } else if (col == KL) {
        x = BF_ONE;
      }
      v[e] = (short)x;
    }
    return v;
  }
  __device__ void C(long m, long col, float v, int split) const {
    if (col < N) part[((size_t)split * M + m) * NP + col] = v;
  }
};

// Sums split-K partials [splits][M][NP] into the weight gradient [M][KL]
// and the bias gradient [M] (column KL), in split order.
__global__ void k_reduce_parts(const float* __restrict__ part, int splits, int M, int NP, int KL,
                               float* __restrict__ gw, float* __restrict__ gb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * (KL + 1)) return;
  const int m = i / (KL + 1), col = i % (KL + 1);
  float s = 0.f;
  for (int sp = 0; sp < splits; ++sp) s += part[((size_t)sp * M + m) * NP + col];
  if (col < KL) gw[(size_t)m * KL + col] = s;
  else gb[m] = s;
}

// bfloat16 copies of the master weights in the layouts the products read.
__global__ void k_quantize(const float* __restrict__ p, uint16_t* __restrict__ q) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)CH * K1) {
    const int co = (int)(i / K1), t = (int)(i % K1);
    q[Q_W1 + i] = t < KL1 ? f2bf(p[O_W1 + co * KL1 + t]) : (uint16_t)0;
  }
  if (i < (size_t)CH * KC) {
    const size_t ow[3] = {O_W2, O_W3, O_W4};
    const size_t qw[3] = {Q_W2, Q_W3, Q_W4};
    const size_t qt[3] = {Q_T2, Q_T3, Q_T4};
    const int co = (int)(i / KC), rem = (int)(i % KC);
    const int tap = rem / CH, ci = rem % CH;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const uint16_t h = f2bf(p[ow[l] + i]);
      q[qw[l] + i] = h;
      q[qt[l] + ((size_t)tap * CH + ci) * CH + co] = h;
    }
  }
  if (i < (size_t)HID * FEAT) {
    const int n = (int)(i / FEAT), k = (int)(i % FEAT);
    const uint16_t h = f2bf(p[O_F1W + i]);
    q[Q_F1 + i] = h;
    q[Q_F1T + (size_t)k * HID + n] = h;
  }
  if (i < (size_t)HID) q[Q_F2 + i] = f2bf(p[O_F2W + i]);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float wd, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float upd = (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = p[i] - lr * (upd + wd * p[i]);
}

}  // namespace cnn
}  // namespace hd

// ---------------------------------------------------------------------------
// Host object.

using namespace hd::cnn;

struct hd_cnn_t {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t max_batch = 0, n = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* q = nullptr;
  uint16_t *y1 = nullptr, *y2 = nullptr, *y3 = nullptr, *y4 = nullptr, *h = nullptr;
  uint16_t *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *d4 = nullptr, *dh = nullptr;
  float *z = nullptr, *dz = nullptr, *part = nullptr, *loss = nullptr;
  size_t part_words = 0;
  const uint16_t* x = nullptr;   // tiles of the last forward
  long step = 0;
};

namespace {

constexpr long WG_CHUNK = 4096;  // pixels per split of the weight-gradient products
constexpr long F1_CHUNK = 256;   // tiles per split of the fc1 weight gradient

struct Guard {
  int prev = -1;
  explicit Guard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~Guard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

#define CK(...)                                                  \
  do {                                                           \
    if ((__VA_ARGS__) != hipSuccess) return HD_INTERNAL_GPU_ERROR; \
  } while (0)

size_t part_words_for(size_t B) {
  size_t w = 0;
  const long pix[4] = {1024, 256, 64, 16};
  for (int l = 0; l < 4; ++l) {
    const long P = (long)B * pix[l];
    const long splits = (P + WG_CHUNK - 1) / WG_CHUNK;
    const long np = l == 0 ? K1 : 304;
    w = std::max<size_t>(w, (size_t)(splits * CH * np));
  }
  const long s1 = ((long)B + F1_CHUNK - 1) / F1_CHUNK;
  w = std::max<size_t>(w, (size_t)(s1 * HID * 528));
  return w;
}

template <typename T>
hipError_t alloc(T** p, size_t count) {
  return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

}  // namespace

extern "C" hd_size hd_cnn_param_count(void) { return NPARAM; }

extern "C" void hd_cnn_destroy(hd_cnn net) {
  if (!net) return;
  Guard gd(net->device);
  if (net->stream) (void)hipStreamSynchronize(net->stream);
  void* bufs[] = {net->p, net->g, net->m, net->v, net->q, net->y1, net->y2, net->y3, net->y4,
                  net->h, net->d1, net->d2, net->d3, net->d4, net->dh, net->z, net->dz,
                  net->part, net->loss};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
}

extern "C" hd_error hd_cnn_create(hd_cnn* out, int gpu_id, hd_size max_batch) {
  if (!out) return HD_INVALID_POINTER;
  *out = nullptr;
  if (max_batch == 0) return HD_INVALID_STRIDE;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || gpu_id < 0 || gpu_id >= ndev)
    return HD_INVALID_DEVICE_INDEX;
  Guard gd(gpu_id);
  hd_cnn net = new hd_cnn_t;
  net->device = gpu_id;
  net->max_batch = max_batch;
  const size_t B = max_batch;
  bool ok = hipStreamCreate(&net->stream) == hipSuccess;
  ok = ok && alloc(&net->p, NPARAM) == hipSuccess && alloc(&net->g, NPARAM) == hipSuccess &&
       alloc(&net->m, NPARAM) == hipSuccess && alloc(&net->v, NPARAM) == hipSuccess &&
       alloc(&net->q, NQ) == hipSuccess;
  ok = ok && alloc(&net->y1, B * 32 * 32 * CH) == hipSuccess &&
       alloc(&net->y2, B * 16 * 16 * CH) == hipSuccess &&
       alloc(&net->y3, B * 8 * 8 * CH) == hipSuccess && alloc(&net->y4, B * FEAT) == hipSuccess &&
       alloc(&net->h, B * HID) == hipSuccess;
  ok = ok && alloc(&net->d1, B * 32 * 32 * CH) == hipSuccess &&
       alloc(&net->d2, B * 16 * 16 * CH) == hipSuccess &&
       alloc(&net->d3, B * 8 * 8 * CH) == hipSuccess && alloc(&net->d4, B * FEAT) == hipSuccess &&
       alloc(&net->dh, B * HID) == hipSuccess;
  net->part_words = part_words_for(B);
  ok = ok && alloc(&net->z, B) == hipSuccess && alloc(&net->dz, B) == hipSuccess &&
       alloc(&net->part, net->part_words) == hipSuccess && alloc(&net->loss, 4) == hipSuccess;
  if (!ok) {
    hd_cnn_destroy(net);
    return HD_MEM_ALLOC_FAILED;
  }
  *out = net;
  return hd_cnn_init_params(net, 1);
}

namespace {
hd_error upload_params(hd_cnn net, const float* host, bool zero_moments) {
  Guard gd(net->device);
  hipStream_t s = net->stream;
  CK(hipMemcpyAsync(net->p, host, NPARAM * sizeof(float), hipMemcpyHostToDevice, s));
  if (zero_moments) {
    CK(hipMemsetAsync(net->m, 0, NPARAM * sizeof(float), s));
    CK(hipMemsetAsync(net->v, 0, NPARAM * sizeof(float), s));
    net->step = 0;
  }
  const size_t nq = (size_t)HID * FEAT;
  hipLaunchKernelGGL(k_quantize, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, net->p,
                     net->q);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(s));
  return HD_NO_ERROR;
}
}  // namespace

extern "C" hd_error hd_cnn_init_params(hd_cnn net, unsigned long long seed) {
  if (!net) return HD_INVALID_POINTER;
  std::vector<float> p(NPARAM, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](size_t off, size_t n, double fan_in) {
    std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / fan_in));
    for (size_t i = 0; i < n; ++i) p[off + i] = (float)nd(rng);
  };
  fill(O_W1, CH * KL1, KL1);
  fill(O_W2, CH * KC, KC);
  fill(O_W3, CH * KC, KC);
  fill(O_W4, CH * KC, KC);
  fill(O_F1W, (size_t)HID * FEAT, FEAT);
  fill(O_F2W, HID, HID);
  return upload_params(net, p.data(), true);
}

extern "C" hd_error hd_cnn_set_params(hd_cnn net, const hd_float* host) {
  if (!net || !host) return HD_INVALID_POINTER;
  return upload_params(net, host, true);
}

extern "C" hd_error hd_cnn_get_params(hd_cnn net, hd_float* host) {
  if (!net || !host) return HD_INVALID_POINTER;
  Guard gd(net->device);
  CK(hipMemcpyAsync(host, net->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  CK(hipStreamSynchronize(net->stream));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_cnn_get_grads(hd_cnn net, hd_float* host) {
  if (!net || !host) return HD_INVALID_POINTER;
  Guard gd(net->device);
  CK(hipMemcpyAsync(host, net->g, NPARAM * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  CK(hipStreamSynchronize(net->stream));
  return HD_NO_ERROR;
}

namespace hd {
// Forward over n <= max_batch tiles on stream s (no host synchronisation).
hipError_t cnn_forward_async(hd_cnn net, const uint16_t* x, size_t n, float* logits,
                             hipStream_t s) {
  const uint16_t* q = net->q;
  hipError_t e;
  {
    Conv1Fwd p{(long)n * 1024, CH, K1, K1, x, q + Q_W1, net->p + O_B1, net->y1};
    if ((e = run_mm<2, 2>(p, s)) != hipSuccess) return e;
  }
  {
    ConvFwd p{(long)n * 256, CH, KC, KC, 32, 16, net->y1, q + Q_W2, net->p + O_B2, net->y2};
    if ((e = run_mm<2, 2>(p, s)) != hipSuccess) return e;
  }
  {
    ConvFwd p{(long)n * 64, CH, KC, KC, 16, 8, net->y2, q + Q_W3, net->p + O_B3, net->y3};
    if ((e = run_mm<1, 2>(p, s)) != hipSuccess) return e;
  }
  {
    ConvFwd p{(long)n * 16, CH, KC, KC, 8, 4, net->y3, q + Q_W4, net->p + O_B4, net->y4};
    if ((e = run_mm<1, 2>(p, s)) != hipSuccess) return e;
  }
  {
    Fc1Fwd p{(long)n, HID, FEAT, FEAT, net->y4, q + Q_F1, net->p + O_F1B, net->h};
    if ((e = run_mm<1, 2>(p, s)) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_fc2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, net->h,
                     q + Q_F2, net->p + O_F2B, (int)n, logits ? logits : net->z);
  return hipGetLastError();
}
size_t cnn_max_batch(hd_cnn net) { return net->max_batch; }
}  // namespace hd

extern "C" hd_error hd_cnn_forward(hd_cnn net, const unsigned short* d_tiles, hd_size n,
                                   hd_float* d_logits) {
  if (!net || !d_tiles) return HD_INVALID_POINTER;
  if (n == 0 || n > net->max_batch) return HD_INVALID_STRIDE;
  Guard gd(net->device);
  net->n = n;
  net->x = d_tiles;
  CK(hd::cnn_forward_async(net, d_tiles, n, net->z, net->stream));
  if (d_logits)
    CK(hipMemcpyAsync(d_logits, net->z, n * sizeof(float), hipMemcpyDeviceToDevice, net->stream));
  CK(hipStreamSynchronize(net->stream));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_cnn_get_logits(hd_cnn net, hd_float* host, hd_size n) {
  if (!net || !host) return HD_INVALID_POINTER;
  if (n > net->n) return HD_INVALID_STRIDE;
  Guard gd(net->device);
  CK(hipMemcpyAsync(host, net->z, n * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  CK(hipStreamSynchronize(net->stream));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_cnn_backward(hd_cnn net, const hd_float* d_labels,
                                    const hd_float* d_weights, hd_float* loss) {
  if (!net || !d_labels || !d_weights) return HD_INVALID_POINTER;
  if (net->n == 0 || !net->x) return HD_INVALID_STRIDE;
  Guard gd(net->device);
  hipStream_t s = net->stream;
  const long n = (long)net->n;
  const uint16_t* q = net->q;
  float* g = net->g;
  hipLaunchKernelGGL(k_head, dim3(1), dim3(HEAD_BS), 0, s, net->z, d_labels, d_weights, (int)n,
                     net->h, net->dz, net->loss, g + O_F2W, g + O_F2B);
  CK(hipGetLastError());
  hipLaunchKernelGGL(k_dh, dim3((unsigned)((n * HID + 255) / 256)), dim3(256), 0, s, net->dz,
                     q + Q_F2, net->h, (int)n, net->dh);
  CK(hipGetLastError());
  auto reduce = [&](long splits, int M, int NP, int KL, float* gw, float* gb) {
    const int tot = M * (KL + 1);
    hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       net->part, (int)splits, M, NP, KL, gw, gb);
    return hipGetLastError();
  };
  const long kpad = (n + 15) / 16 * 16;
  {  // fc1 weights (+ bias column)
    Fc1Wgrad p{HID, FEAT + 1, kpad, F1_CHUNK, (int)n, 528, net->dh, net->y4, net->part};
    CK(run_mm<1, 1>(p, s));
    CK(reduce((kpad + F1_CHUNK - 1) / F1_CHUNK, HID, 528, FEAT, g + O_F1W, g + O_F1B));
  }
  {  // gradient at y4
    Fc1Dgrad p{n, FEAT, HID, HID, net->dh, q + Q_F1T, net->y4, net->d4};
    CK(run_mm<1, 2>(p, s));
  }
  // conv layers 4, 3, 2: weights then the data gradient of the layer input
  struct L {
    int H, Ho;
    const uint16_t* xin;
    uint16_t* d;
    uint16_t* dprev;
    size_t ow, ob, qt;
  };
  const L layers[3] = {{8, 4, net->y3, net->d4, net->d3, O_W4, O_B4, Q_T4},
                       {16, 8, net->y2, net->d3, net->d2, O_W3, O_B3, Q_T3},
                       {32, 16, net->y1, net->d2, net->d1, O_W2, O_B2, Q_T2}};
  for (const L& ly : layers) {
    const long P = n * ly.Ho * ly.Ho;
    ConvWgrad<CH> pw{CH, KC + 1, P, WG_CHUNK, ly.H, ly.Ho, 304, ly.d, ly.xin, net->part};
    CK(run_mm<1, 1>(pw, s));
    CK(reduce((P + WG_CHUNK - 1) / WG_CHUNK, CH, 304, KC, g + ly.ow, g + ly.ob));
    ConvDgrad pd{n * ly.H * ly.H, CH, KC, KC, ly.H, ly.Ho, ly.d, q + ly.qt, ly.xin, ly.dprev};
    CK(run_mm<2, 2>(pd, s));
  }
  {  // conv1 weights (input tiles)
    const long P = n * 1024;
    ConvWgrad<CIN1> pw{CH, KL1 + 1, P, WG_CHUNK, TILE, 32, K1, net->d1, net->x, net->part};
    CK(run_mm<1, 1>(pw, s));
    CK(reduce((P + WG_CHUNK - 1) / WG_CHUNK, CH, K1, KL1, g + O_W1, g + O_B1));
  }
  if (loss) CK(hipMemcpyAsync(loss, net->loss, sizeof(float), hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_cnn_adam(hd_cnn net, hd_float lr, hd_float beta1, hd_float beta2,
                                hd_float eps, hd_float weight_decay) {
  if (!net) return HD_INVALID_POINTER;
  Guard gd(net->device);
  hipStream_t s = net->stream;
  net->step += 1;
  const float c1 = (float)(1.0 - std::pow((double)beta1, (double)net->step));
  const float c2 = (float)(1.0 - std::pow((double)beta2, (double)net->step));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((NPARAM + 255) / 256)), dim3(256), 0, s, net->p,
                     net->g, net->m, net->v, NPARAM, lr, beta1, beta2, eps, weight_decay, c1, c2);
  CK(hipGetLastError());
  const size_t nq = (size_t)HID * FEAT;
  hipLaunchKernelGGL(k_quantize, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, net->p,
                     net->q);
  CK(hipGetLastError());
  return HD_NO_ERROR;
}
