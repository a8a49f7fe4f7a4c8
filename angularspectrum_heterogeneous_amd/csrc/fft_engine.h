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

// Irrational butterfly constants as double-float pairs (hi + lo equals the
// constant to ~1e-16).  Multiplying as fma(x, hi, x*lo) rounds once, without
// bias.  With the constants rounded to float instead, sqrt(1/2), cos(pi/8) and
// sin(pi/8) all round down, every radix-8/16 butterfly then loses energy by a
// fixed factor (~3e-8 per path), and over the ~10^3 transforms of a march that
// bias accumulates into a coherent gain error on the beam.
constexpr float H_HI = 0.7071067690849304f, H_LO = 1.2101617485882343e-08f;   // sqrt(1/2)
constexpr float C1_HI = 0.9238795042037964f, C1_LO = 2.830748968563057e-08f;  // cos(pi/8)
constexpr float S1_HI = 0.3826834261417389f, S1_LO = 6.2233507236442165e-09f; // sin(pi/8)
constexpr float S3_HI = 0.8660253882408142f, S3_LO = 1.5543625053737742e-08f; // sqrt(3)/2

__device__ __forceinline__ float mc(float x, float hi, float lo) { return fmaf(x, hi, x * lo); }

// w * (sp P + i sq Q), P and Q positive constants given as hi/lo pairs, sp, sq = +-1
__device__ __forceinline__ float2 cmul_c(float2 w, float ph, float pl, float sp, float qh, float ql, float sq) {
  const float xp = sp * mc(w.x, ph, pl), yq = sq * mc(w.y, qh, ql);
  const float xq = sq * mc(w.x, qh, ql), yp = sp * mc(w.y, ph, pl);
  return make_float2(xp - yq, xq + yp);
}

template <bool INV>
__device__ __forceinline__ void dft3(float2& a, float2& b, float2& c) {
  const float sg = INV ? 1.f : -1.f;  // sin(2 pi/3) sign
  float2 t1 = cadd(b, c);
  float2 t2 = csub(b, c);
  float2 y0 = cadd(a, t1);
  float2 m1 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
  // (b - c) * i * sg * sqrt(3)/2
  float2 m2 = make_float2(-sg * mc(t2.y, S3_HI, S3_LO), sg * mc(t2.x, S3_HI, S3_LO));
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
  // radix-2 first stage (decimation in frequency)
  float2 a0 = cadd(v[0], v[4]), a4 = csub(v[0], v[4]);
  float2 a1 = cadd(v[1], v[5]), a5 = csub(v[1], v[5]);
  float2 a2 = cadd(v[2], v[6]), a6 = csub(v[2], v[6]);
  float2 a3 = cadd(v[3], v[7]), a7 = csub(v[3], v[7]);
  // twiddles W8^1, W8^2, W8^3 on the odd half
  if (INV) {
    a5 = make_float2(mc(a5.x - a5.y, H_HI, H_LO), mc(a5.x + a5.y, H_HI, H_LO));
    a7 = make_float2(-mc(a7.x + a7.y, H_HI, H_LO), mc(a7.x - a7.y, H_HI, H_LO));
  } else {
    a5 = make_float2(mc(a5.x + a5.y, H_HI, H_LO), mc(a5.y - a5.x, H_HI, H_LO));
    a7 = make_float2(mc(a7.y - a7.x, H_HI, H_LO), -mc(a7.x + a7.y, H_HI, H_LO));
  }
  a6 = rot90<INV>(a6);
  dft4<INV>(a0, a1, a2, a3);
  dft4<INV>(a4, a5, a6, a7);
  v[0] = a0; v[2] = a1; v[4] = a2; v[6] = a3;
  v[1] = a4; v[3] = a5; v[5] = a6; v[7] = a7;
}

template <bool INV>
__device__ __forceinline__ void dft16(float2* v) {
  // 4 x 4 decomposition: n = 4 n1 + n2, k = k1 + 4 k2
  float2 w[16];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 a = v[n2], b = v[n2 + 4], c = v[n2 + 8], d = v[n2 + 12];
    dft4<INV>(a, b, c, d);
    w[n2 * 4 + 0] = a; w[n2 * 4 + 1] = b; w[n2 * 4 + 2] = c; w[n2 * 4 + 3] = d;
  }
  // twiddle W16^(n2*k1); sg is the sign of the imaginary parts (forward: -)
  const float sg = INV ? 1.f : -1.f;
  // W^1 = (c1, sg s1), W^2 = (h, sg h), W^3 = (s1, sg c1), W^6 = (-h, sg h), W^9 = (-c1, -sg s1)
  w[5] = cmul_c(w[5], C1_HI, C1_LO, 1.f, S1_HI, S1_LO, sg);
  w[6] = cmul_c(w[6], H_HI, H_LO, 1.f, H_HI, H_LO, sg);
  w[7] = cmul_c(w[7], S1_HI, S1_LO, 1.f, C1_HI, C1_LO, sg);
  w[9] = cmul_c(w[9], H_HI, H_LO, 1.f, H_HI, H_LO, sg);
  w[10] = rot90<INV>(w[10]);
  w[11] = cmul_c(w[11], H_HI, H_LO, -1.f, H_HI, H_LO, sg);
  w[13] = cmul_c(w[13], S1_HI, S1_LO, 1.f, C1_HI, C1_LO, sg);
  w[14] = cmul_c(w[14], H_HI, H_LO, -1.f, H_HI, H_LO, sg);
  w[15] = cmul_c(w[15], C1_HI, C1_LO, -1.f, S1_HI, S1_LO, -sg);
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
// Two radix orders per line length (V = 0, 1); a plan's autotuner picks one.
// ---------------------------------------------------------------------------
template <int L, int V = 0> struct Radix;
template <> struct Radix<64, 0>   { static constexpr int NP = 2; static constexpr int R0 = 8,  R1 = 8,  R2 = 1; };
template <> struct Radix<64, 1>   { static constexpr int NP = 2; static constexpr int R0 = 16, R1 = 4,  R2 = 1; };
template <> struct Radix<128, 0>  { static constexpr int NP = 2; static constexpr int R0 = 16, R1 = 8,  R2 = 1; };
template <> struct Radix<128, 1>  { static constexpr int NP = 3; static constexpr int R0 = 8,  R1 = 4,  R2 = 4; };
template <> struct Radix<192, 0>  { static constexpr int NP = 3; static constexpr int R0 = 16, R1 = 4,  R2 = 3; };
template <> struct Radix<192, 1>  { static constexpr int NP = 3; static constexpr int R0 = 8,  R1 = 8,  R2 = 3; };
template <> struct Radix<256, 0>  { static constexpr int NP = 2; static constexpr int R0 = 16, R1 = 16, R2 = 1; };
template <> struct Radix<256, 1>  { static constexpr int NP = 3; static constexpr int R0 = 8,  R1 = 8,  R2 = 4; };
template <> struct Radix<384, 0>  { static constexpr int NP = 3; static constexpr int R0 = 16, R1 = 8,  R2 = 3; };
template <> struct Radix<384, 1>  { static constexpr int NP = 3; static constexpr int R0 = 8,  R1 = 16, R2 = 3; };
template <> struct Radix<512, 0>  { static constexpr int NP = 3; static constexpr int R0 = 8,  R1 = 8,  R2 = 8; };
template <> struct Radix<1024, 0> { static constexpr int NP = 3; static constexpr int R0 = 16, R1 = 8,  R2 = 8; };

template <int L, int P, bool INV, int V = 0>
struct PassR {
  using RX = Radix<L, V>;
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

template <int L, int P, bool INV, int V = 0>
__device__ __forceinline__ int out_pos(int j, int r) {
  using PR = PassR<L, P, INV, V>;
  constexpr int Ns = PR::Ns, R = PR::R;
  return (j / Ns) * Ns * R + (j % Ns) + r * Ns;
}

}  // namespace ash
