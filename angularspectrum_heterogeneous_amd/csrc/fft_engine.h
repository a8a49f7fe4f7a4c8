// Stockham FFT building blocks for gfx950 (wave64), shared by every sweep.
//
// A transform of length L is a list of radix-R passes (R in {2,3,4,8,16}).
// Each pass is executed by butterflies: butterfly j of a line reads elements
// j + r*L/R (r = 0..R-1), multiplies by the Stockham twiddle W_{Ns R}^{(j%Ns) r},
// runs an R-point DFT in registers and writes element (j/Ns)*Ns*R + j%Ns + r*Ns
// (Ns = product of the radices of the earlier passes).  The last pass therefore
// leaves butterfly j holding outputs j + r*L/R in natural order, which is the
// same pattern the first pass reads: a forward transform followed by an
// inverse transform whose first radix equals the forward's last radix needs no
// data exchange in between (used by the fused y-pass).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ash {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// Small DFTs in registers.  Forward kernel exp(-2 pi i n k / R); INV flips sign.
// ---------------------------------------------------------------------------
template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft3(float2& a, float2& b, float2& c) {
  const float c1 = -0.5f;
  const float s1 = INV ? 0.86602540378443864676f : -0.86602540378443864676f;
  float2 t1 = cadd(b, c);
  float2 t2 = csub(b, c);
  float2 y0 = cadd(a, t1);
  float2 m1 = make_float2(a.x + c1 * t1.x, a.y + c1 * t1.y);
  // (b - c) * i*s1
  float2 m2 = make_float2(-s1 * t2.y, s1 * t2.x);
  a = y0;
  b = cadd(m1, m2);
  c = csub(m1, m2);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a, float2& b, float2& c, float2& d) {
  float2 t0 = cadd(a, c), t1 = csub(a, c);
  float2 t2 = cadd(b, d), t3 = rot90<INV>(csub(b, d));
  a = cadd(t0, t2);
  c = csub(t0, t2);
  b = cadd(t1, t3);
  d = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
  const float h = 0.70710678118654752440f;
  // radix-2 first stage (decimation in frequency)
  float2 a0 = cadd(v[0], v[4]), a4 = csub(v[0], v[4]);
  float2 a1 = cadd(v[1], v[5]), a5 = csub(v[1], v[5]);
  float2 a2 = cadd(v[2], v[6]), a6 = csub(v[2], v[6]);
  float2 a3 = cadd(v[3], v[7]), a7 = csub(v[3], v[7]);
  // twiddles W8^1, W8^2, W8^3 on the odd half
  a5 = INV ? make_float2(h * (a5.x - a5.y), h * (a5.x + a5.y)) : make_float2(h * (a5.x + a5.y), h * (a5.y - a5.x));
  a6 = rot90<INV>(a6);
  a7 = INV ? make_float2(-h * (a7.x + a7.y), h * (a7.x - a7.y)) : make_float2(h * (a7.y - a7.x), -h * (a7.x + a7.y));
  dft4<INV>(a0, a1, a2, a3);
  dft4<INV>(a4, a5, a6, a7);
  v[0] = a0; v[2] = a1; v[4] = a2; v[6] = a3;
  v[1] = a4; v[3] = a5; v[5] = a6; v[7] = a7;
}

template <bool INV>
__device__ __forceinline__ void dft16(float2* v) {
  // 4 x 4 decomposition: n = 4 n1 + n2, k = k1 + 4 k2
  const float c1 = 0.92387953251128675613f;  // cos(pi/8)
  const float s1 = 0.38268343236508977173f;  // sin(pi/8)
  const float h = 0.70710678118654752440f;
  float2 w[16];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 a = v[n2], b = v[n2 + 4], c = v[n2 + 8], d = v[n2 + 12];
    dft4<INV>(a, b, c, d);
    w[n2 * 4 + 0] = a; w[n2 * 4 + 1] = b; w[n2 * 4 + 2] = c; w[n2 * 4 + 3] = d;
  }
  // twiddle W16^(n2*k1)
  const float sg = INV ? 1.f : -1.f;
  const float2 t1 = make_float2(c1, sg * s1), t2 = make_float2(h, sg * h), t3 = make_float2(s1, sg * c1);
  const float2 t6 = make_float2(-h, sg * h), t9 = make_float2(-c1, -sg * s1);
  // n2=1: k1=1,2,3 -> W^1, W^2, W^3 ; n2=2: W^2, W^4, W^6 ; n2=3: W^3, W^6, W^9
  w[5] = cmul(w[5], t1); w[6] = cmul(w[6], t2); w[7] = cmul(w[7], t3);
  w[9] = cmul(w[9], t2); w[10] = rot90<INV>(w[10]); w[11] = cmul(w[11], t6);
  w[13] = cmul(w[13], t3); w[14] = cmul(w[14], t6); w[15] = cmul(w[15], t9);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 a = w[k1], b = w[4 + k1], c = w[8 + k1], d = w[12 + k1];
    dft4<INV>(a, b, c, d);
    v[k1] = a; v[k1 + 4] = b; v[k1 + 8] = c; v[k1 + 12] = d;
  }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) dft2<INV>(v[0], v[1]);
  else if constexpr (R == 3) dft3<INV>(v[0], v[1], v[2]);
  else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 8) dft8<INV>(v);
  else if constexpr (R == 16) dft16<INV>(v);
  else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------------------
// Pass lists.  Forward order R0, R1, [R2]; the inverse runs them reversed.
// ---------------------------------------------------------------------------
template <int L> struct Radix;
template <> struct Radix<64>   { static constexpr int NP = 2; static constexpr int R0 = 8,  R1 = 8,  R2 = 1; };
template <> struct Radix<128>  { static constexpr int NP = 2; static constexpr int R0 = 16, R1 = 8,  R2 = 1; };
template <> struct Radix<192>  { static constexpr int NP = 3; static constexpr int R0 = 16, R1 = 4,  R2 = 3; };
template <> struct Radix<256>  { static constexpr int NP = 2; static constexpr int R0 = 16, R1 = 16, R2 = 1; };
template <> struct Radix<384>  { static constexpr int NP = 3; static constexpr int R0 = 16, R1 = 8,  R2 = 3; };
template <> struct Radix<512>  { static constexpr int NP = 3; static constexpr int R0 = 8,  R1 = 8,  R2 = 8; };
template <> struct Radix<1024> { static constexpr int NP = 3; static constexpr int R0 = 16, R1 = 8,  R2 = 8; };

template <int L, int P, bool INV>
struct PassR {
  using RX = Radix<L>;
  static constexpr int NP = RX::NP;
  // forward pass p uses R_p; inverse pass p uses R_{NP-1-p}
  static constexpr int idx = INV ? (NP - 1 - P) : P;
  static constexpr int R = idx == 0 ? RX::R0 : (idx == 1 ? RX::R1 : RX::R2);
  static constexpr int r_at(int i) { return i == 0 ? RX::R0 : (i == 1 ? RX::R1 : RX::R2); }
  static constexpr int ns() {
    int n = 1;
    for (int q = 0; q < P; ++q) n *= r_at(INV ? (NP - 1 - q) : q);
    return n;
  }
  static constexpr int Ns = ns();
  static constexpr int NB = L / R;
};

// Twiddle table tw[i] = exp(-2 pi i i / L) (float2) held in LDS.
template <int L, int R, int Ns, bool INV>
__device__ __forceinline__ void twiddle(float2* v, int j, const float2* tw) {
  if constexpr (Ns > 1) {
    const int k = j % Ns;
    constexpr int step = L / (Ns * R);
#pragma unroll
    for (int r = 1; r < R; ++r) {
      float2 w = tw[k * r * step];
      if (INV) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
}

template <int L, int P, bool INV>
__device__ __forceinline__ int out_pos(int j, int r) {
  using PR = PassR<L, P, INV>;
  constexpr int Ns = PR::Ns, R = PR::R;
  return (j / Ns) * Ns * R + (j % Ns) + r * Ns;
}

}  // namespace ash
