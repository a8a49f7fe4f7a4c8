// Learned nonlinear column step: a 4-layer 1-D conv net over t that stands in
// for one dZ of the KT update (ASS:833-839) on smooth columns.
//
//   A = max_t |u|, x = u / A, X0[t] = (x, x^2, 0...)          32 channels, bf16
//   X_j = relu(conv9(X_{j-1}, W_j) + b_j), j = 1..3           32 channels, bf16
//   y   = conv9(X_3, W_4) + b_4 (channel 0), u_new = u + c A^2 y,  c = dZ N / dT
//
// conv9 is the zero-padded cross-correlation out[t] = sum_k W_k in[t + k - 4].
// Every layer is a 32 x 32 x 9 convolution (layer 1 reads channels 0-1, layer
// 4 writes channel 0; the other weights are zero), run on the matrix cores as
// D[co][t] += W_k[co][ci] X[t + k - 4][ci] with v_mfma_f32_16x16x32_bf16:
// lane l holds A[l%16][8(l/16) + j], B[8(l/16) + j][l%16] and
// D[4(l/16) + i][l%16].  Activations are [t][32 channels] rows of bfloat16;
// accumulation is float32, weight gradients included.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ash {
namespace net {

constexpr int C = 32;       // channels
constexpr int KW = 9;       // kernel width
constexpr int HW = KW / 2;  // half width
constexpr int NLAY = 4;
constexpr int W_SIZE = C * C * KW;
constexpr int LAYER_SIZE = W_SIZE + C;           // W[co][ci][k] then b[co], float32
constexpr int PARAM_COUNT = NLAY * LAYER_SIZE;
constexpr int RS = 40;  // LDS row stride in bf16 (80 B: 16-byte aligned, spreads the banks)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16;

__device__ __forceinline__ u16 f2bf(float x) { return __builtin_bit_cast(u16, (__bf16)x); }
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d) {
  return make_uint2((uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16), (uint32_t)f2bf(c) | ((uint32_t)f2bf(d) << 16));
}

// A operands of one layer for this lane: a[mt][k] = W_k[16 mt + l%16][8 (l/16) .. +8).
// Wl is [KW][C][C] bf16 (LDS or global), 16-byte aligned.
template <int NMT>
__device__ __forceinline__ void load_a(const u16* Wl, int lane, bf16x8 (&a)[2][KW]) {
#pragma unroll
  for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
    for (int k = 0; k < KW; ++k)
      a[mt][k] = *reinterpret_cast<const bf16x8*>(Wl + (k * C + 16 * mt + (lane & 15)) * C + 8 * (lane >> 4));
}

// 32 (or 16) output channels x 16 positions p0..p0+15 from LDS rows p0-4 .. p0+19.
template <int NMT>
__device__ __forceinline__ void conv_tile(const u16* in, int p0, int lane, const bf16x8 (&a)[2][KW], f32x4 (&acc)[2]) {
#pragma unroll
  for (int mt = 0; mt < NMT; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < KW; ++k) {
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(in + (p0 + (lane & 15) + k - HW) * RS + 8 * (lane >> 4));
#pragma unroll
    for (int mt = 0; mt < NMT; ++mt) acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][k], b, acc[mt], 0, 0, 0);
  }
}

// ---------------------------------------------------------------------------
// Fused inference over real columns R[c][t] (in place), persistent blocks.
// Each wave works on its own columns and never waits for another wave (one
// block barrier, after the weights are staged).  Per column: A = max|u| and the
// largest step D = max|u[t+1]-u[t]|.  A column with D / dT > dpdt_max is
// flagged (flag[c] = 1) and left to the KT kernel.  A column with c D <
// NEGLIGIBLE is left unchanged: the Burgers increment of one step is bounded by
// about c A D (c = dz N / dT), so there it is below 1e-5 of the column's own
// amplitude.  Otherwise the four layers run on segments of SEG outputs with a
// 16-sample halo on each side (the receptive field of four 9-tap layers); a
// segment whose window stays below QUIET * A is left unchanged.
//
// The layers of a segment run in place in the wave's LDS buffer: the tile of
// positions p0 .. p0+15 of layer j reads the rows of positions p0-4 .. p0+19
// of layer j-1 and writes its outputs four rows lower, over rows that no later
// tile of the layer reads (the wave walks the tiles in increasing p0).  With
// the features of positions t0-16 .. t0+SEG+15 in rows 0 .. SEG+31, layer j
// position q sits in row q - 4j (position index relative to t0-16).
// ---------------------------------------------------------------------------
constexpr int INF_WAVES = 8;
constexpr int INF_THREADS = 64 * INF_WAVES;
constexpr int SEG = 128;                  // outputs per segment
constexpr int WROWS = SEG + 40;           // rows of a wave's buffer: the window, and the last tiles' overrun
constexpr float QUIET = 1e-4f;
constexpr float NEGLIGIBLE = 1e-5f;

// First layer with its two input channels (x, x^2): the (tap, channel) pairs
// K = 2 k + ci (18 of them, zero-padded to 32) form the reduction axis, so a
// 16-position tile is one matrix-core product per 16 output channels instead
// of nine.  W0p[co][K] is the packed weight, lane l's B operand holds
// K = 8 (l/16) .. +8 of position p0 + l%16: the (x, x^2) pairs of taps
// 4 (l/16) .. 4 (l/16) + 3 (zero past the last tap).  a0[mt] holds the lane's
// A fragment W0p[16 mt + l%16][8 (l/16) .. +8).
__device__ __forceinline__ void conv_first_tile(const u16* in, const bf16x8 (&a0)[2], int p0, int lane, f32x4 (&acc)[2]) {
  const int n = lane & 15, b = lane >> 4;
  uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = 4 * b + i;
    if (k < KW) w[i] = *reinterpret_cast<const uint32_t*>(in + (p0 + n + k - HW) * RS);
  }
  const bf16x8 bv = __builtin_bit_cast(bf16x8, make_uint4(w[0], w[1], w[2], w[3]));
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0[mt], bv, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
}

template <int NT_>
__global__ __launch_bounds__(INF_THREADS) void k_net_infer(float* __restrict__ R, int ncols, const u16* __restrict__ Wm,
                                                           const float* __restrict__ params, float c, float dpdt_max,
                                                           float inv_dt, int32_t* __restrict__ flag) {
  static_assert(NT_ % SEG == 0 && NT_ % 64 == 0, "column length must be a multiple of the segment");
  // hidden layers 1 and 2 (network layers 2 and 3) as [layer][k][co][RS]: rows
  // padded to 80 bytes, so the sixteen co rows a fragment load touches fall in
  // distinct banks
  __shared__ __align__(16) u16 sW[2 * KW * C * RS];
  __shared__ __align__(16) u16 sW0[C * C];
  __shared__ __align__(16) u16 sW3[16 * C];
  __shared__ __align__(16) u16 buf[INF_WAVES][WROWS * RS];
  __shared__ float sbias[NLAY][C];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  {
    // Wm is [layer][k][co][ci]
    for (int i = tid; i < 2 * KW * C * (C / 8); i += INF_THREADS) {
      const int row = i / (C / 8), q = i % (C / 8);  // row = (layer, k, co)
      reinterpret_cast<uint4*>(sW + row * RS)[q] = reinterpret_cast<const uint4*>(Wm + W_SIZE + row * C)[q];
    }
    for (int i = tid; i < C * C; i += INF_THREADS) {
      const int co = i / C, kk = i % C, k = kk >> 1, ci = kk & 1;
      sW0[i] = kk < 2 * KW ? Wm[(k * C + co) * C + ci] : (u16)0;
    }
    // output layer (its only output channel is 0) with the taps as rows: sW3[k][ci]
    for (int i = tid; i < 16 * C; i += INF_THREADS) {
      const int k = i / C, ci = i % C;
      sW3[i] = k < KW ? Wm[3 * W_SIZE + (k * C + 0) * C + ci] : (u16)0;
    }
    for (int i = tid; i < NLAY * C; i += INF_THREADS) sbias[i / C][i % C] = params[(i / C) * LAYER_SIZE + W_SIZE + i % C];
  }
  __syncthreads();
  u16* wb = buf[wave];
  const int n = lane & 15, g4 = lane >> 4;
  const float b3 = sbias[3][0];
  // ReLU(acc + bias) as bf16 into rows p0 + n - shift (positions p0 + n), zero
  // where the position lies outside the column (the zero padding of the next layer)
  auto emit = [&](int j, int p0, int t0, int shift, const f32x4 (&acc)[2]) {
    const int p = p0 + n, t = t0 - 16 + p;
    const bool inside = t >= 0 && t < NT_;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const int co = 16 * mt + 4 * g4;
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = inside ? fmaxf(acc[mt][i] + sbias[j - 1][co + i], 0.f) : 0.f;
      *reinterpret_cast<uint2*>(wb + (p - shift) * RS + co) = pack4(v[0], v[1], v[2], v[3]);
    }
  };
  const int gwave = blockIdx.x * INF_WAVES + wave, nwaves = gridDim.x * INF_WAVES;
  for (int col = gwave; col < ncols; col += nwaves) {
    float* g = R + (long)col * NT_;
    // column statistics: lane l reads samples E l .. E l + E - 1
    constexpr int E = NT_ / 64;
    float am = 0.f, dm = 0.f, first = 0.f, last = 0.f;
#pragma unroll
    for (int q = 0; q < E / 4; ++q) {
      const float4 v = reinterpret_cast<const float4*>(g + E * lane)[q];
      const float w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        am = fmaxf(am, fabsf(w[i]));
        if (q + i > 0) dm = fmaxf(dm, fabsf(w[i] - last));
        if (q + i == 0) first = w[i];
        last = w[i];
      }
    }
    {
      const float nxt = __shfl(first, lane + 1, 64);  // the next lane's first sample
      if (lane < 63) dm = fmaxf(dm, fabsf(nxt - last));
    }
    for (int o = 32; o > 0; o >>= 1) {
      am = fmaxf(am, __shfl_xor(am, o, 64));
      dm = fmaxf(dm, __shfl_xor(dm, o, 64));
    }
    const bool fallback = dm * inv_dt > dpdt_max;  // wave-uniform
    if (lane == 0) flag[col] = fallback ? 1 : 0;
    if (fallback || am == 0.f || c * dm < NEGLIGIBLE) continue;
    const float A = am, gain = c * A * A;
    // The window of a segment (samples t0-16 .. t0+SEG+15) is read into registers
    // while the previous segment is computed, before that segment writes its
    // outputs: every layer input is an original sample.  The features go to the
    // rows as bf16 (x, x^2) and the sample itself, in float32, to bytes 64-67 of
    // its row (the row padding, which no layer writes), where the output update
    // finds it.
    constexpr int FQ = (SEG + 32 + 63) / 64;
    float fnext[FQ];
    auto fetch = [&](int ts) {
#pragma unroll
      for (int q = 0; q < FQ; ++q) {
        const int t = ts - 16 + lane + 64 * q;
        fnext[q] = g[min(max(t, 0), NT_ - 1)];  // clamped address; outside the column masked below
      }
    };
    fetch(0);
    for (int t0 = 0; t0 < NT_; t0 += SEG) {
      float fcur[FQ];
#pragma unroll
      for (int q = 0; q < FQ; ++q) fcur[q] = fnext[q];
      if (t0 + SEG < NT_) fetch(t0 + SEG);
      float wm = 0.f;
#pragma unroll
      for (int q = 0; q < FQ; ++q) {
        const int p = lane + 64 * q, t = t0 - 16 + p;
        if (p < SEG + 32) {
          const float v = (t >= 0 && t < NT_) ? fcur[q] : 0.f;
          const float x = v / A;
          wm = fmaxf(wm, fabsf(x));
          *reinterpret_cast<uint32_t*>(wb + p * RS) = (uint32_t)f2bf(x) | ((uint32_t)f2bf(x * x) << 16);
          *reinterpret_cast<float*>(wb + p * RS + C) = v;
        }
      }
      for (int o = 32; o > 0; o >>= 1) wm = fmaxf(wm, __shfl_xor(wm, o, 64));
      if (wm < QUIET) continue;  // wave-uniform; the segment stays as it is
      wave_sync();
      // layer 1 (two input channels, packed taps): positions 4 .. SEG+27 -> rows 0 ..
      {
        bf16x8 a0[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) a0[mt] = *reinterpret_cast<const bf16x8*>(sW0 + (16 * mt + n) * C + 8 * g4);
        for (int p0 = 4; p0 < SEG + 28; p0 += 16) {
          f32x4 acc[2];
          conv_first_tile(wb, a0, p0, lane, acc);
          wave_sync();  // every lane has read the rows before any is overwritten
          emit(1, p0, t0, 4, acc);
          wave_sync();
        }
      }
      // layers 2 and 3: positions 4j .. SEG+31-4j from rows shifted by 4(j-1)
#pragma unroll 1
      for (int j = 2; j <= 3; ++j) {
        bf16x8 a[2][KW];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int k = 0; k < KW; ++k)
            a[mt][k] = *reinterpret_cast<const bf16x8*>(sW + (((j - 2) * KW + k) * C + 16 * mt + n) * RS + 8 * g4);
        const int sh_in = 4 * (j - 1);
        for (int p0 = 4 * j; p0 < SEG + 32 - 4 * j; p0 += 16) {
          // rows of positions p0-4 .. p0+19 of layer j-1
          const u16* rowp = wb + (p0 - HW + n - sh_in) * RS + 8 * g4;
          bf16x8 cur[KW];
#pragma unroll
          for (int k = 0; k < KW; ++k) cur[k] = *reinterpret_cast<const bf16x8*>(rowp + k * RS);
          f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
          for (int k = 0; k < KW; ++k) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][k], cur[k], acc[mt], 0, 0, 0);
          }
          // scheduling groups: the nine LDS reads, then the eighteen products
          __builtin_amdgcn_sched_group_barrier(0x100, KW, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, 2 * KW, 0);
          wave_sync();  // every lane has read the rows before any is overwritten
          emit(j, p0, t0, 4 * j, acc);
          wave_sync();
        }
      }
      // output layer: positions 16 .. SEG+15 are t0 .. t0+SEG-1, from the layer-3
      // rows of positions p0-4 .. p0+19 (row = position - 12), nine taps on the
      // matrix cores with the single output channel in matrix row 0 (the other
      // rows of A are zero)
      {
        bf16x8 a3[KW];
        const bf16x8 zero8 = __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u));
#pragma unroll
        for (int k = 0; k < KW; ++k) a3[k] = n == 0 ? *reinterpret_cast<const bf16x8*>(sW3 + k * C + 8 * g4) : zero8;
        for (int p0 = 16; p0 < SEG + 16; p0 += 16) {
          const u16* rowp = wb + (p0 - 16 + n) * RS + 8 * g4;
          bf16x8 cur[KW];
#pragma unroll
          for (int k = 0; k < KW; ++k) cur[k] = *reinterpret_cast<const bf16x8*>(rowp + k * RS);
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int k = 0; k < KW; ++k) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3[k], cur[k], acc, 0, 0, 0);
          if (g4 == 0) {
            const int t = t0 + p0 - 16 + n;
            g[t] = *reinterpret_cast<const float*>(wb + (p0 + n) * RS + C) + gain * (acc[0] + b3);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Training path: per-layer kernels over a batch of B columns, activations in
// global memory as [B][nT][32] bf16.
// ---------------------------------------------------------------------------
constexpr int TT = 128;  // outputs per block of the per-layer kernels

// Features and per-column amplitude: one block (256 threads) per column.
__global__ __launch_bounds__(256) void k_net_feat(const float* __restrict__ u, int nT, u16* __restrict__ X0,
                                                  float* __restrict__ amp) {
  __shared__ float sred[4];
  const int b = blockIdx.x, tid = threadIdx.x;
  const float* g = u + (long)b * nT;
  float am = 0.f;
  for (int t = tid; t < nT; t += 256) am = fmaxf(am, fabsf(g[t]));
  for (int o = 32; o > 0; o >>= 1) am = fmaxf(am, __shfl_xor(am, o, 64));
  if ((tid & 63) == 0) sred[tid >> 6] = am;
  __syncthreads();
  const float A = fmaxf(fmaxf(sred[0], sred[1]), fmaxf(sred[2], sred[3]));
  if (tid == 0) amp[b] = A;
  const float As = A > 0.f ? A : 1.f;
  for (int t = tid; t < nT; t += 256) {
    const float x = g[t] / As;
    uint4* row = reinterpret_cast<uint4*>(X0 + ((long)b * nT + t) * C);
    row[0] = make_uint4((uint32_t)f2bf(x) | ((uint32_t)f2bf(x * x) << 16), 0u, 0u, 0u);
    row[1] = row[2] = row[3] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// MODE 0: out = bf16(relu(conv + bias)); MODE 1: y = conv[ch 0] + bias[0] (float);
// MODE 2: out = bf16(conv * (mask > 0)) (input gradient through a relu).
// grid (nT / TT, B), 256 threads.
template <int MODE>
__global__ __launch_bounds__(256) void k_net_conv(const u16* __restrict__ in, int nT, const u16* __restrict__ W,
                                                  const float* __restrict__ bias, const u16* __restrict__ mask,
                                                  u16* __restrict__ out, float* __restrict__ y) {
  __shared__ __align__(16) u16 win[(TT + 2 * HW) * RS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.y, t0 = blockIdx.x * TT;
  const long cb = (long)b * nT;
  // window rows p = 0 .. TT+7 hold t = t0 - 4 + p (zero outside the column)
  for (int i = tid; i < (TT + 2 * HW) * 4; i += 256) {
    const int p = i >> 2, q = i & 3, t = t0 - HW + p;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (t >= 0 && t < nT) v = reinterpret_cast<const uint4*>(in + (cb + t) * C)[q];
    reinterpret_cast<uint4*>(win + p * RS)[q] = v;
  }
  __syncthreads();
  constexpr int NMT = MODE == 1 ? 1 : 2;
  bf16x8 a[2][KW];
  load_a<NMT>(W, lane, a);
  for (int tt = wave; tt < TT / 16; tt += 4) {
    f32x4 acc[2];
    conv_tile<NMT>(win, HW + 16 * tt, lane, a, acc);
    const int t = t0 + 16 * tt + (lane & 15);
    if (t >= nT) continue;
    if constexpr (MODE == 1) {
      if ((lane >> 4) == 0) y[cb + t] = acc[0][0] + bias[0];
    } else {
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        const int co = 16 * mt + 4 * (lane >> 4);
        float v[4];
        if constexpr (MODE == 0) {
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] = fmaxf(acc[mt][i] + bias[co + i], 0.f);
        } else {
          const uint2 mk = *reinterpret_cast<const uint2*>(mask + (cb + t) * C + co);
          const float m0 = bf2f((u16)(mk.x & 0xffff)), m1 = bf2f((u16)(mk.x >> 16));
          const float m2 = bf2f((u16)(mk.y & 0xffff)), m3 = bf2f((u16)(mk.y >> 16));
          v[0] = m0 > 0.f ? acc[mt][0] : 0.f;
          v[1] = m1 > 0.f ? acc[mt][1] : 0.f;
          v[2] = m2 > 0.f ? acc[mt][2] : 0.f;
          v[3] = m3 > 0.f ? acc[mt][3] : 0.f;
        }
        *reinterpret_cast<uint2*>(out + (cb + t) * C + co) = pack4(v[0], v[1], v[2], v[3]);
      }
    }
  }
}

// Loss (mean square over all samples) and its gradient, in channel 0 of dY:
// y* = (u_next - u) / (c A^2).  grid (nT / 256, B), 256 threads.
__global__ __launch_bounds__(256) void k_net_loss(const float* __restrict__ y, const float* __restrict__ u,
                                                  const float* __restrict__ un, const float* __restrict__ amp,
                                                  int nT, float c, float inv_count, u16* __restrict__ dY,
                                                  double* __restrict__ loss) {
  __shared__ double sred[4];
  const int b = blockIdx.y, t = blockIdx.x * 256 + threadIdx.x;
  double l2 = 0.0;
  if (t < nT) {
    const long i = (long)b * nT + t;
    const float A = amp[b];
    const float s = A > 0.f ? c * A * A : 1.f;
    const float ys = A > 0.f ? (un[i] - u[i]) / s : 0.f;
    const float r = y[i] - ys;
    l2 = (double)r * (double)r;
    uint4* row = reinterpret_cast<uint4*>(dY + i * C);
    row[0] = make_uint4((uint32_t)f2bf(2.f * r * inv_count), 0u, 0u, 0u);
    row[1] = row[2] = row[3] = make_uint4(0u, 0u, 0u, 0u);
  }
  for (int o = 32; o > 0; o >>= 1) l2 += __shfl_xor(l2, o, 64);
  if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = l2;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (sred[0] + sred[1] + sred[2] + sred[3]) * (double)inv_count);
}

// Weight and bias gradients of one layer on the matrix cores: for every tap k,
// gW[co][ci][k] += sum_t D[t][co] X[t+k-4][ci] is a 32 x 32 product with the
// sum over t as the reduction dimension.  The block's tiles are staged in LDS
// transposed ([channel][t], bf16): A operands (8 consecutive t of one co) are
// single 16-byte reads; B operands start at an odd offset for odd k and are
// assembled from 2-byte reads.  grid (nT / TT, B), 256 threads.
constexpr int WS_D = TT + 8;   // row stride of dT (272 B, a multiple of 16 B)
constexpr int WS_X = TT + 16;  // row stride of xT (288 B)
__global__ __launch_bounds__(256) void k_net_wgrad(const u16* __restrict__ Dg, const u16* __restrict__ X, int nT,
                                                   float* __restrict__ gW, float* __restrict__ gb) {
  __shared__ __align__(16) u16 dT[C * WS_D];  // dT[co][p] = D[t0 + p][co]
  __shared__ __align__(16) u16 xT[C * WS_X];  // xT[ci][p] = X[t0 - 4 + p][ci]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.y, t0 = blockIdx.x * TT;
  const long cb = (long)b * nT;
  for (int i = tid; i < TT * C; i += 256) {
    const int pp = i / C, ch = i % C, t = t0 + pp;
    dT[ch * WS_D + pp] = t < nT ? Dg[(cb + t) * C + ch] : (u16)0;
  }
  for (int i = tid; i < (TT + 2 * HW) * C; i += 256) {
    const int pp = i / C, ch = i % C, t = t0 - HW + pp;
    xT[ch * WS_X + pp] = (t >= 0 && t < nT) ? X[(cb + t) * C + ch] : (u16)0;
  }
  __syncthreads();
  const int g = lane >> 4, r = lane & 15;
  // 9 taps x 2 co tiles x 2 ci tiles of 16 x 16
  for (int u = wave; u < KW * 4; u += 4) {
    const int k = u >> 2, mt = (u >> 1) & 1, nt = u & 1;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < TT / 32; ++kb) {
      const int tb = 32 * kb + 8 * g;
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(dT + (16 * mt + r) * WS_D + tb);
      const u16* xr = xT + (16 * nt + r) * WS_X + tb + k;
      bf16x8 bv;
#pragma unroll
      for (int j = 0; j < 8; ++j) bv[j] = __builtin_bit_cast(__bf16, xr[j]);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    // D[co = 16 mt + 4 g + i][ci = 16 nt + r]
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(&gW[((16 * mt + 4 * g + i) * C + 16 * nt + r) * KW + k], acc[i]);
  }
  if (tid < C) {
    float s = 0.f;
    for (int pp = 0; pp < TT; ++pp) s += bf2f(dT[tid * WS_D + pp]);
    atomicAdd(&gb[tid], s);
  }
}

// Adam on the trainable entries (mask 1); masked entries stay at zero.
__global__ void k_net_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, const float* __restrict__ mask, int n, float lr, float b1, float b2,
                           float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mask[i] == 0.f) { p[i] = 0.f; return; }
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// bf16 weights for the matrix cores: Wm[j][k][co][ci] = W_j[co][ci][k] and the
// flipped transpose Wt[j][k][ci][co] = W_j[co][ci][KW-1-k] (input gradients).
__global__ void k_net_pack(const float* __restrict__ p, u16* __restrict__ Wm, u16* __restrict__ Wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NLAY * W_SIZE) return;
  const int j = i / W_SIZE, r = i % W_SIZE;
  const int co = r / (C * KW), ci = (r / KW) % C, k = r % KW;
  const u16 w = f2bf(p[j * LAYER_SIZE + r]);
  Wm[j * W_SIZE + (k * C + co) * C + ci] = w;
  Wt[j * W_SIZE + ((KW - 1 - k) * C + ci) * C + co] = w;
}

}  // namespace net
}  // namespace ash
