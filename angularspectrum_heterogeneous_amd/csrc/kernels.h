// gfx950 kernels of one z-step of the nonlinear angular-spectrum march.
//
// State between kernels: the rfft_t half spectrum S[x][y][f] (complex64),
// f fastest, row length MP (>= M+1, multiple of 16), M = nT/2.  Bin M is the
// Nyquist bin, bins > M are padding: only the y pass touches them (it transforms
// them like any other line), and no valid bin depends on them.
//
// Line sweeps (x and y) work on tiles of one line family: all positions along
// the transform axis for LF=16 consecutive f bins (128-byte segments).  The
// t-column kernel gives each column to LPC lanes of one wave (column.h).
#pragma once
#include "fft_engine.h"

namespace ash {

constexpr int LF = 16;    // f bins per line-sweep tile
constexpr int LNT = 256;  // threads per line-sweep block

// ---------------------------------------------------------------------------
// Generic multi-line Stockham transform.  Layout policy:
//   LINE_FAST=true : LDS index pos*NL + line, butterfly b -> (line=b%NL, j=b/NL)
//   LINE_FAST=false: LDS index line*LS + pos, butterfly b -> (line=b/NB, j=b%NB)
// load(line,pos) feeds pass 0, store(line,pos,v) receives the last pass.
// ---------------------------------------------------------------------------
template <int L, int NL, int LS, bool LINE_FAST>
struct Lay {
  __device__ static __forceinline__ int idx(int line, int pos) {
    return LINE_FAST ? pos * NL + line : line * LS + pos;
  }
  template <int NB>
  __device__ static __forceinline__ void map(int b, int& line, int& j) {
    if (LINE_FAST) { line = b % NL; j = b / NL; }
    else { line = b / NB; j = b % NB; }
  }
};

// One Stockham pass.  IN: 0 = load functor, 1 = LDS; OUT: 0 = store functor, 1 = LDS.
template <int L, int P, bool INV, int NT, int NL, class LY, int IN, int OUT, int V, class LoadF, class StoreF>
__device__ __forceinline__ void fft_pass_io(float2* buf, const float2* tw, int tid, LoadF& load, StoreF& store) {
  using PR = PassR<L, P, INV, V>;
  constexpr int R = PR::R, NB = PR::NB, Ns = PR::Ns;
  constexpr int TOT = NB * NL;
  constexpr int PER = (TOT + NT - 1) / NT;
  float2 v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = tid + q * NT;
    if (TOT % NT == 0 || b < TOT) {
      int line, j;
      LY::template map<NB>(b, line, j);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (IN == 0) v[q][r] = load(line, j + r * NB);
        else v[q][r] = buf[LY::idx(line, j + r * NB)];
      }
      twiddle<L, R, Ns, INV>(v[q], j, tw);
      dft<R, INV>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = tid + q * NT;
    if (TOT % NT == 0 || b < TOT) {
      int line, j;
      LY::template map<NB>(b, line, j);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int pos = out_pos<L, P, INV, V>(j, r);
        if constexpr (OUT == 0) store(line, pos, v[q][r]);
        else buf[LY::idx(line, pos)] = v[q][r];
      }
    }
  }
  if constexpr (OUT == 1) __syncthreads();
}

template <int L, int P, bool INV, int NT, int NL, class LY, int V, class LoadF, class StoreF>
__device__ __forceinline__ void fft_pass(float2* buf, const float2* tw, int tid, LoadF& load, StoreF& store) {
  constexpr int NP = PassR<L, P, INV, V>::NP;
  fft_pass_io<L, P, INV, NT, NL, LY, P == 0 ? 0 : 1, P == NP - 1 ? 0 : 1, V>(buf, tw, tid, load, store);
}

// V selects the radix order (Radix<L, V>).
template <int L, bool INV, int NT, int NL, class LY, int V = 0, class LoadF, class StoreF>
__device__ __forceinline__ void fft_run(float2* buf, const float2* tw, int tid, LoadF load, StoreF store) {
  constexpr int NP = Radix<L, V>::NP;
  fft_pass<L, 0, INV, NT, NL, LY, V>(buf, tw, tid, load, store);
  if constexpr (NP > 1) fft_pass<L, 1, INV, NT, NL, LY, V>(buf, tw, tid, load, store);
  if constexpr (NP > 2) fft_pass<L, 2, INV, NT, NL, LY, V>(buf, tw, tid, load, store);
}

template <int L>
__device__ __forceinline__ void load_tw(float2* tw, const float2* __restrict__ g, int tid, int nt) {
  for (int i = tid; i < L; i += nt) tw[i] = g[i];
}

// Line-sweep element access at a 32-bit byte offset from the tile's base (a
// block-uniform pointer): the address is one 32-bit VGPR plus the base in
// SGPRs, instead of a 64-bit per-lane pointer.  The spectrum of the largest
// grid (384 x 384 x 2048, 1.2 GB) stays below 4 GiB.
__device__ __forceinline__ float2 gld(const float2* base, uint32_t off) {
  return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ void gst(float2* base, uint32_t off, float2 v) {
  *reinterpret_cast<float2*>(reinterpret_cast<char*>(base) + off) = v;
}

// ---------------------------------------------------------------------------
// Per-plan constant tables (device memory, built on the host in FP64).
// ---------------------------------------------------------------------------
struct Tables {
  const double* kx;      // [nX] natural-order kx (ASS:403-408 centred grid)
  const double* ky;      // [nY]
  const double* kt;      // [MP] ASM t-wavenumber, 0 where the bin is not kept
  const double* ka;      // [MP] attenuation |k| = 2 pi f_m / c0
  const double* alpha;   // [MP] alpha(f_m)   (Np/m)
  const double* astar;   // [MP] alphaStar(f_m)
  const float* tapx_h;   // [nX] sqrt(taper_x)   (split-step abl_half factor)
  const float* tapy_h;   // [nY]
  const float* tapt_h;   // [nT]
  const float* tapx_f;   // [nX] taper_x        (sequential abl factor)
  const float* tapy_f;
  const float* tapt_f;
  const float* lgx;      // [nX] log2(taper_x) for the frequency-weighted mask
  const float* lgy;      // [nY]
  const float* fw_exp;   // [MP] exponent clip(f0/max(f_m,f_1),0.1,10)
  const float* sa_wx;    // [nX] super-absorbing weights
  const float* sa_wy;    // [nY]
  const float* sa_cw;    // [MP] c0/omega_m (0 where the coefficient is zeroed)
  const float2* twx;     // [nX] exp(-2 pi i k / nX)
  const float2* twy;     // [nY]
  const float2* twc;     // column-transform pass twiddles: forward passes, then inverse passes
  const float2* twgx;    // [nX] exp(-2 pi i k / nX), general-size x kernels
  const float2* twgy;    // [nY]
  const float2* twgt;    // [nT] exp(-2 pi i k / nT): column packing, general-length column kernel
};

struct Geom {
  int nX, nY, nT;
  int M;    // last rfft bin, nT / 2
  int MP;   // spectrum row length (>= M + 1, multiple of LF)
  int nfb;  // LF-bin tiles per row
  int nyq;  // the Nyquist bin (M) for even nT, -1 for odd nT
};

// Operator descriptor for one k-space visit.
struct OpDesc {
  int use_prop;       // angular-spectrum propagator with step dz_prop
  int use_att;        // attenuation/dispersion with step dz_att
  int obliquity;      // k/k_z path-length factor
  int clip_negative;  // water branch: zero the filter where Re < 0
  double dz_prop;
  double dz_att;
  double k_start, k_max;  // transverse cosine filter band
  float norm;             // 1/(nX nY) of the inverse 2-D FFT
  double k_start2, k_max2;  // squared band edges
  float k_startf, inv_band, dzp;
  int precise;              // evaluate in FP64 and round once (accuracy reference)
};

// sin and cos for the bounded phases of the visit operator (|x| < ~1e4):
// Cody-Waite reduction by pi/2 with a three-part constant, then the minimax
// polynomials of the Cephes library (S. Moshier) on [-pi/4, pi/4].  Branch
// free; max abs error ~9e-8 and no magnitude bias (sin^2 + cos^2 - 1 averages
// ~1e-9), which matters because the same operator is applied every step.
__device__ __forceinline__ void bounded_sincos(float x, float* sn, float* cs) {
  const float k = rintf(x * 0.636619772367581343f);
  float r = fmaf(-k, 1.5703125f, x);
  r = fmaf(-k, 4.837512969970703125e-4f, r);
  r = fmaf(-k, 7.54978995489188216e-8f, r);
  const float z = r * r;
  const float ps = fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f);
  const float s = fmaf(z * r, ps, r);
  const float pc = fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f);
  const float c = fmaf(z * z, pc, fmaf(-0.5f, z, 1.0f));
  const int q = ((int)k) & 3;
  const float s1 = (q & 1) ? c : s;
  const float c1 = (q & 1) ? s : c;
  *sn = (q & 2) ? -s1 : s1;
  *cs = ((q + 1) & 2) ? -c1 : c1;
}

// kx^2 + ky^2 and k^2 - kk with every product and sum rounded on its own, as
// NumPy evaluates them (ASS:492, ASS:506, ASS:608, ASS:620).  A fused
// multiply-add rounds once and can move the result by an ulp; on grids where a
// mode sits exactly at the evanescent cutoff k_z = 0 that ulp flips the mode
// between the attenuation's obliquity factor exp(-alpha dz 1e6) and 1.
__device__ __forceinline__ double kperp2(double kx, double ky) {
#pragma clang fp contract(off)
  return kx * kx + ky * ky;
}
__device__ __forceinline__ double kz2_of(double k, double kk) {
#pragma clang fp contract(off)
  return k * k - kk;
}

// Per-f constants of one visit, staged in LDS by the y-pass.
struct LineConst {
  double kt2, ka2;   // squared ASM and attenuation wavenumbers (FP64)
  double ktd;        // ASM wavenumber (FP64)
  double ald, asd;   // alpha dz_att, alphaStar dz_att (FP64)
  float kt, ka;      // the same, single precision
  float al, as;      // alpha(f) dz_att, alphaStar(f) dz_att
  int kept;          // propagator keeps this bin
};

// Propagator x attenuation at natural bin (kx, ky, m); kk = kx^2 + ky^2 in FP64.
// Only the differences kt^2 - kk and ka^2 - kk need FP64 (they cancel near the
// evanescent cutoff); everything else is single precision.  The propagator
// phase dz (kt - kz) = dz kk / (kt + kz) avoids the cancellation of kt - kz,
// and the attenuation phase is folded into one sincos with it.
__device__ __forceinline__ float2 visit_operator_fp64(const OpDesc& op, double kk, const LineConst& lc, int m) {
  double ph = 0.0, lg = 0.0, amp = op.norm;
  if (op.use_prop) {
    if (!lc.kept || kk > op.k_max2) return make_float2(0.f, 0.f);
    if (kk > op.k_start2) amp *= 0.5 * (1.0 + cospi((sqrt(kk) - op.k_start) / (op.k_max - op.k_start)));
    const double d2 = lc.kt2 - kk;
    if (d2 > 0.0) ph = op.dz_prop * (kk / (lc.ktd + sqrt(d2)));
    else { ph = op.dz_prop * lc.ktd; lg = -op.dz_prop * sqrt(-d2); }
  }
  if (op.use_att && m > 0) {
    const double ka = sqrt(lc.ka2);
    double re = 0.0, im = 0.0;
    bool apply = true;
    if (op.obliquity) {
      const double kz2 = lc.ka2 - kk;
      if (kz2 > 0.0) {
        const double obl = ka / fmax(sqrt(kz2), ka * 1e-6);
        re = -lc.ald * obl; im = -lc.asd * obl;
      } else apply = false;
    } else { re = -lc.ald; im = -lc.asd; }
    if (apply) {
      if (op.clip_negative && cos(im) < 0.0) return make_float2(0.f, 0.f);
      lg += re; ph += im;
    }
  }
  amp *= exp(lg);
  double s, c;
  sincos(ph, &s, &c);
  return make_float2((float)(amp * c), (float)(amp * s));
}

__device__ __forceinline__ float2 visit_operator(const OpDesc& op, double kk, const LineConst& lc, int m) {
  float ph = 0.f, lg = 0.f, amp = op.norm;
  const float kkf = (float)kk;
  if (op.use_prop) {
    if (!lc.kept || kk > op.k_max2) return make_float2(0.f, 0.f);
    if (kk > op.k_start2) {
      // pos in [0, 1] (kk <= k_max2 here): the bounded sincos, not the general cospi
      const float pos = (__builtin_amdgcn_sqrtf(kkf) - op.k_startf) * op.inv_band;
      float sn, cs;
      bounded_sincos(3.14159265358979f * pos, &sn, &cs);
      amp *= 0.5f * (1.f + cs);
    }
    const float d2 = (float)(lc.kt2 - kk);
    if (d2 > 0.f) {
      ph = op.dzp * kkf * __builtin_amdgcn_rcpf(lc.kt + __builtin_amdgcn_sqrtf(d2));
    } else {
      ph = op.dzp * lc.kt;
      lg = -op.dzp * __builtin_amdgcn_sqrtf(-d2);
    }
  }
  if (op.use_att && m > 0) {
    float re, im;
    bool apply = true;
    if (op.obliquity) {
      const float kz2 = (float)(lc.ka2 - kk);
      if (kz2 > 0.f) {
        const float obl = lc.ka * __builtin_amdgcn_rsqf(fmaxf(kz2, lc.ka * lc.ka * 1e-12f));
        re = -lc.al * obl;
        im = -lc.as * obl;
      } else {
        apply = false;
      }
    } else {
      re = -lc.al;
      im = -lc.as;
    }
    if (apply) {
      if (op.clip_negative) {
        float sn, cs;
        bounded_sincos(im, &sn, &cs);
        if (cs < 0.f) return make_float2(0.f, 0.f);
      }
      lg += re;
      ph += im;
    }
  }
  // lg <= 0: the hardware base-2 exponential (v_exp_f32), ~1 ulp
  if (lg != 0.f) amp *= __builtin_amdgcn_exp2f(lg * 1.44269504088896341f);
  float s, c;
  bounded_sincos(ph, &s, &c);
  return make_float2(amp * c, amp * s);
}

// ---------------------------------------------------------------------------
// x-sweep, forward FFT along x.  Tile = (y, f-block).
// ---------------------------------------------------------------------------
template <int L, int V>
__global__ __launch_bounds__(LNT) void k_xfft_fwd(float2* __restrict__ spec, Tables T, Geom g) {
  __shared__ float2 tw[L];
  __shared__ float2 buf[L * LF];
  const int tid = threadIdx.x;
  const int y = blockIdx.x / g.nfb, fb = blockIdx.x % g.nfb;
  const int f0 = fb * LF;
  const uint32_t sxb = (uint32_t)g.nY * g.MP * 8u;  // x stride in bytes
  float2* base = spec + (long)y * g.MP + f0;
  const int nvalid = min(LF, g.M + 1 - f0);
  load_tw<L>(tw, T.twx, tid, LNT);
  __syncthreads();
  using LY = Lay<L, LF, 0, true>;
  auto ld = [&](int line, int pos) -> float2 {
    return line < nvalid ? gld(base, pos * sxb + line * 8u) : make_float2(0.f, 0.f);
  };
  auto st = [&](int line, int pos, float2 v) {
    if (line < nvalid) gst(base, pos * sxb + line * 8u, v);
  };
  fft_run<L, false, LNT, LF, LY, V>(buf, tw, tid, ld, st);
}

// f bins per y-pass tile: 8 for the 384-point line, whose 16-bin tile (78 KB of
// LDS with the operator table) would hold the CU to two blocks
template <int L> constexpr int ypass_lines() { return L == 384 ? 8 : LF; }
// threads of a y-pass block: 512 for the 256-point 8*8*4 order, whose every pass
// has 512 or more butterflies per tile (twice the waves per CU in the same LDS),
// else LNT
template <int L, int V> constexpr int ypass_threads() { return (L == 256 && V == 1) ? 512 : LNT; }

// ---------------------------------------------------------------------------
// y-pass: forward FFT along y, multiply by the visit operator, inverse FFT.
// Tile = (kx, f-block).  The 1/(nX nY) normalisation is in the operator.
// ---------------------------------------------------------------------------
// One block handles a pair of x rows.  The centred wavenumber grid
// (ASS:403-408) has half-integer offsets, so natural bins b and n-1-b carry
// wavenumbers of equal magnitude: rows x and nX-1-x see the same kx^2 and
// positions y and nY-1-y the same ky^2.  The operator depends on (kx, ky)
// only through kk = kx^2 + ky^2, so one block evaluates it once per
// (|ky| index, f bin) into an LDS table and applies it to both rows: a
// quarter of one evaluation per element.
template <int L, bool PRECISE, int V>
__global__ __launch_bounds__((ypass_threads<L, V>())) void k_ypass(float2* __restrict__ spec, Tables T, Geom g, OpDesc op) {
  constexpr int H = L / 2;
  constexpr int NTY = ypass_threads<L, V>();
  constexpr int LFY = ypass_lines<L>();
  __shared__ float2 tw[L];
  __shared__ float2 buf[L * LFY];
  __shared__ float2 opt[H * LFY];  // [i * LFY + line]: operator at natural y bins i and L-1-i
  __shared__ LineConst slc[LFY];
  const int tid = threadIdx.x;
  const int nfb = g.MP / LFY;
  const int ip = blockIdx.x / nfb, fb = blockIdx.x % nfb;
  const int f0 = fb * LFY;
  const uint32_t syb = (uint32_t)g.MP * 8u;  // y stride in bytes
  // lines past the Nyquist bin (padding, f > M) are transformed like the others
  // (no guard branches around the loads and stores): each line is an independent
  // transform, so padding values never reach a valid bin
  load_tw<L>(tw, T.twy, tid, NTY);
  if (tid < LFY) {
    const int m = min(f0 + tid, g.MP - 1);
    const double kt = T.kt[m], ka = T.ka[m];
    LineConst lc;
    lc.kt2 = kt * kt; lc.ka2 = ka * ka;
    lc.ktd = kt;
    lc.kt = (float)kt; lc.ka = (float)ka;
    lc.ald = T.alpha[m] * op.dz_att;
    lc.asd = T.astar[m] * op.dz_att;
    lc.al = (float)lc.ald;
    lc.as = (float)lc.asd;
    lc.kept = kt > 0.0;
    slc[tid] = lc;
  }
  __syncthreads();
  static_assert(NTY % LFY == 0, "the line of a thread must be fixed");
  const int myline = tid % LFY;
  const LineConst lc = slc[myline];
  using LY = Lay<L, LFY, 0, true>;
  // The first pass's elements of a row are read into registers ahead of use:
  // row 0's while the operator table is computed, row 1's while row 0 is
  // transformed back (when every thread owns whole butterflies of that pass;
  // otherwise the pass reads memory itself).
  using P0 = PassR<L, 0, false, V>;
  constexpr int R0 = P0::R, NB0 = P0::NB, TOT0 = NB0 * LFY;
  constexpr bool PREF = TOT0 % NTY == 0;
  constexpr int NPRE = PREF ? (TOT0 / NTY) * R0 : 1;
  float2 pre[NPRE];
  auto prefetch = [&](int x) {
    if constexpr (PREF) {
      const float2* bx = spec + (long)x * g.nY * g.MP + f0;
#pragma unroll
      for (int q = 0; q < TOT0 / NTY; ++q) {
        int line, j;
        LY::template map<NB0>(tid + q * NTY, line, j);
#pragma unroll
        for (int r = 0; r < R0; ++r) pre[q * R0 + r] = gld(bx, (j + r * NB0) * syb + line * 8u);
      }
    }
  };
  prefetch(ip);
  {
    const double kx = T.kx[ip];
    // entry e = tid + NTY k is on line e % LF = myline
#pragma unroll 2
    for (int e = tid; e < H * LFY; e += NTY) {
      const double ky = T.ky[e / LFY];
      const double kk = kperp2(kx, ky);
      if constexpr (PRECISE) opt[e] = visit_operator_fp64(op, kk, lc, f0 + myline);
      else opt[e] = visit_operator(op, kk, lc, f0 + myline);
    }
  }
  // the table is read after the barriers of the forward transform
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
    // an opaque copy of the thread index per row: without it the compiler
    // keeps every thread-dependent LDS address and twiddle of both transforms
    // live across the rows (~70 more VGPRs, one block fewer per CU)
    int t = tid;
    asm volatile("" : "+v"(t));
    const int x = h == 0 ? ip : g.nX - 1 - ip;
    float2* base = spec + (long)x * g.nY * g.MP + f0;
    int k = 0;  // the first pass asks for its elements in prefetch order
    auto ld = [&](int line, int pos) -> float2 {
      if constexpr (PREF) return pre[k++];
      else return gld(base, pos * syb + line * 8u);
    };
    auto st_lds = [&](int line, int pos, float2 v) { buf[LY::idx(line, pos)] = v; };
    // pass 0 stores to LDS only after its barrier, which follows every read of
    // the previous row's last pass
    fft_run<L, false, NTY, LFY, LY, V>(buf, tw, t, ld, st_lds);
    if (h == 0) prefetch(g.nX - 1 - ip);
    auto ld_op = [&](int line, int pos) -> float2 {
      const int i = pos < H ? pos : L - 1 - pos;
      return cmul(buf[LY::idx(line, pos)], opt[i * LFY + line]);
    };
    auto st = [&](int line, int pos, float2 v) {
      gst(base, pos * syb + line * 8u, v);
    };
    __syncthreads();
    fft_run<L, true, NTY, LFY, LY, V>(buf, tw, t, ld_op, st);
  }
}

// ---------------------------------------------------------------------------
// x-sweep, inverse FFT along x + epilogue.
//   EPI 0: none
//   EPI 1: multiply by tapx[x] tapy[y] (abl factor in x,y)
//   EPI 2: project DC/Nyquist to real, frequency-weighted mask, super-absorbing x pass
// ---------------------------------------------------------------------------
struct EpiDesc {
  int half;           // EPI 1: use the sqrt tapers (split-step) instead of the full ones
  int fw;             // EPI 2: frequency-weighted boundary on
  int sa;             // EPI 2: super-absorbing x pass on
  float sa_strength;
  float inv2dx, invdx;  // 1/(2 dX), 1/dX
  float scale;          // EPI 2: factor on the transformed values (1/nX when no y-pass normalised them)
};

template <int L, int EPI, int V>
__global__ __launch_bounds__(LNT) void k_xfft_inv(float2* __restrict__ spec, Tables T, Geom g, EpiDesc e) {
  __shared__ float2 tw[L];
  __shared__ float2 buf[L * LF];
  const int tid = threadIdx.x;
  const int y = blockIdx.x / g.nfb, fb = blockIdx.x % g.nfb;
  const int f0 = fb * LF;
  const uint32_t sxb = (uint32_t)g.nY * g.MP * 8u;  // x stride in bytes
  float2* base = spec + (long)y * g.MP + f0;
  const int nvalid = min(LF, g.M + 1 - f0);
  load_tw<L>(tw, T.twx, tid, LNT);
  __syncthreads();
  using LY = Lay<L, LF, 0, true>;
  auto ld = [&](int line, int pos) -> float2 {
    return line < nvalid ? gld(base, pos * sxb + line * 8u) : make_float2(0.f, 0.f);
  };
  if constexpr (EPI == 0 || EPI == 1) {
    float ty = 1.f;
    const float* tx = nullptr;
    if constexpr (EPI == 1) {
      ty = e.half ? T.tapy_h[y] : T.tapy_f[y];
      tx = e.half ? T.tapx_h : T.tapx_f;
    }
    // EPI 1: the taper factors of the positions this thread's last pass stores,
    // read into registers ahead of the transform (when every thread owns whole
    // butterflies of that pass), so the stores wait on no load
    using PL = PassR<L, Radix<L, V>::NP - 1, true, V>;
    constexpr int RL = PL::R, NBL = PL::NB, TOTL = NBL * LF;
    constexpr bool TPRE = EPI == 1 && TOTL % LNT == 0;
    float tf[TPRE ? (TOTL / LNT) * RL : 1];
    if constexpr (TPRE) {
#pragma unroll
      for (int q = 0; q < TOTL / LNT; ++q) {
        int line, j;
        LY::template map<NBL>(tid + q * LNT, line, j);
#pragma unroll
        for (int r = 0; r < RL; ++r) tf[q * RL + r] = tx[out_pos<L, Radix<L, V>::NP - 1, true, V>(j, r)] * ty;
      }
    }
    int k = 0;  // the last pass stores in the order the factors were read
    auto st = [&](int line, int pos, float2 v) {
      if constexpr (TPRE) v = cscale(v, tf[k++]);
      else if constexpr (EPI == 1) v = cscale(v, tx[pos] * ty);
      if (line < nvalid) gst(base, pos * sxb + line * 8u, v);
    };
    fft_run<L, true, LNT, LF, LY, V>(buf, tw, tid, ld, st);
  } else {
    // the inverse transform's last pass stores into LDS with the scale, the
    // DC/Nyquist projection and the frequency-weighted mask already applied,
    // so the SA stencil below reads masked neighbours without recomputing them
    // x tables in LDS while that keeps the blocks per CU (L <= 256), else read
    // through the cache
    constexpr bool TLDS = L <= 256;
    __shared__ float slgx[TLDS ? L : 1], swx[TLDS ? L : 1];
    __shared__ float sfwe[LF];
    if constexpr (TLDS) {
      for (int i = tid; i < L; i += LNT) {
        slgx[i] = T.lgx[i];
        swx[i] = T.sa_wx[i];
      }
    }
    const float* lgx = TLDS ? slgx : T.lgx;
    const float* wx = TLDS ? swx : T.sa_wx;
    if (tid < LF) sfwe[tid] = T.fw_exp[min(f0 + tid, g.MP - 1)];
    const float lgy = T.lgy[y];
    // (the barrier of the transform's first pass orders these stores before the reads)
    auto st_lds = [&](int line, int pos, float2 v) {
      const int m = f0 + line;
      v = cscale(v, e.scale);
      if (m == 0 || m == g.nyq) v.y = 0.f;  // irfft/rfft round trip of the reference
      if (e.fw) v = cscale(v, exp2f(sfwe[line] * (lgx[pos] + lgy)));
      buf[LY::idx(line, pos)] = v;
    };
    fft_run<L, true, LNT, LF, LY, V>(buf, tw, tid, ld, st_lds);
    __syncthreads();
    // element e -> (line, pos), line fastest: coalesced 128-byte rows; the
    // line of a thread is fixed (LNT is a multiple of LF)
    constexpr int PER = (L * LF) / LNT;
    const int line = tid % LF;
    const float cw = T.sa_cw[min(f0 + line, g.MP - 1)];
    float2 res[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int pos = (tid + q * LNT) / LF;
      float2 v = buf[LY::idx(line, pos)];
      if (e.sa) {
        const float w = wx[pos];
        if (w != 0.f) {
          float2 d;
          if (pos == 0) d = cscale(csub(buf[LY::idx(line, 1)], v), e.invdx);
          else if (pos == L - 1) d = cscale(csub(v, buf[LY::idx(line, L - 2)]), e.invdx);
          else d = cscale(csub(buf[LY::idx(line, pos + 1)], buf[LY::idx(line, pos - 1)]), e.inv2dx);
          // coeff * dF with coeff = -i c0/omega
          const float2 cd = make_float2(cw * d.y, -cw * d.x);
          const float2 pin = pos < L / 2 ? cscale(cadd(v, cd), 0.5f) : cscale(csub(v, cd), 0.5f);
          v = csub(v, cscale(pin, e.sa_strength * w));
        }
      }
      res[q] = v;
    }
    if (line < nvalid) {
#pragma unroll
      for (int q = 0; q < PER; ++q) gst(base, ((tid + q * LNT) / LF) * sxb + line * 8u, res[q]);
    }
  }
}

// Rows k_ysa keeps in LDS: all of them with a screen, else the two boundary bands
// (rows 0..band and L-1-band..L-1).
__host__ __device__ inline int ysa_rows(int L, int band, bool screen) {
  return (!screen && 2 * (band + 1) < L) ? 2 * (band + 1) : L;
}
// dynamic LDS of k_ysa: the handled entries, then three floats per handled row
__host__ __device__ inline int ysa_lds_bytes(int L, int band, bool screen) {
  const int n = ysa_rows(L, band, screen);
  return n * LF * 8 + n * 3 * 4;
}

// ---------------------------------------------------------------------------
// y-sweep without FFT: super-absorbing y pass, DC/Nyquist projection and an
// optional phase/amplitude screen.  Tile = (x, f-block).
// ---------------------------------------------------------------------------
struct ScreenDesc {
  const float* phase;  // [nX*nY] or null
  const float* amp;    // [nX*nY] or null
  float f0b;           // max(f0_bin, 1)
  int sa;
  float sa_strength;
  float inv2dy, invdy;
  int band;            // rows 0..band-1 (and mirrored) carry nonzero super-absorbing weight
};

template <int L>
__global__ __launch_bounds__(LNT) void k_ysa(float2* __restrict__ spec, Tables T, Geom g, ScreenDesc s) {
  // rows handled x LF entries, sized at launch (ysa_lds_bytes): the two boundary
  // bands need a fraction of a full line, and the smaller buffer lets more blocks
  // share a CU.  After the entries: per handled row its super-absorbing weight and
  // the screen's phase and amplitude.
  extern __shared__ float2 buf[];
  const int tid = threadIdx.x;
  const int x = blockIdx.x / g.nfb, fb = blockIdx.x % g.nfb;
  const int f0 = fb * LF;
  float2* base = spec + (long)x * g.nY * g.MP + f0;
  const long sy = g.MP;
  const int nvalid = min(LF, g.M + 1 - f0);
  // rows handled: all of them with a screen, else only the two boundary bands
  // (rows 0..band and L-1-band..L-1, which hold every row with a nonzero
  // super-absorbing weight and its neighbours)
  const int nrows = ysa_rows(L, s.band, s.phase != nullptr);
  const bool band = nrows < L;
  auto row_of = [&](int i) { return (band && i > s.band) ? L - nrows + i : i; };
  const int nel = nrows * LF;
  float* s_w = reinterpret_cast<float*>(buf + nel);
  float* s_ph = s_w + nrows;
  float* s_am = s_ph + nrows;
  constexpr int PER = (L * LF) / LNT;
  static_assert(LNT % LF == 0, "the line of a thread must be fixed");
  const int line = tid % LF, m = f0 + line;
  const int nq = (nel + LNT - 1) / LNT;  // same for the whole block
  {
    // every load is issued before the first LDS store waits for one; rows past
    // nrows are clamped to a handled row (the line always lies inside the
    // spectrum row: f0 + LF <= MP)
    float2 v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q)
      if (q < nq) v[q] = base[row_of(min((tid + q * LNT) / LF, nrows - 1)) * sy + line];
    for (int i = tid; i < nrows; i += LNT) {
      const int pos = row_of(i);
      s_w[i] = s.sa ? T.sa_wy[pos] : 0.f;
      if (s.phase) {
        const long idx = (long)x * g.nY + pos;
        s_ph[i] = s.phase[idx];
        s_am[i] = s.amp ? fminf(fmaxf(s.amp[idx], 1e-6f), 1.f) : 1.f;
      }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int el = tid + q * LNT;
      if (q < nq && el < nel) buf[el] = line < nvalid ? v[q] : make_float2(0.f, 0.f);
    }
  }
  const float cw = s.sa ? T.sa_cw[min(m, g.MP - 1)] : 0.f;
  const float fs = (float)m / s.f0b;
  __syncthreads();
  float2 res[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int el = tid + q * LNT;
    if (el >= nel) break;
    const int i = el / LF;
    const int pos = row_of(i);
    float2 u = buf[i * LF + line];
    const float w = s_w[i];
    if (w != 0.f) {
      float2 d;
      if (pos == 0) d = cscale(csub(buf[(i + 1) * LF + line], u), s.invdy);
      else if (pos == L - 1) d = cscale(csub(u, buf[(i - 1) * LF + line]), s.invdy);
      else d = cscale(csub(buf[(i + 1) * LF + line], buf[(i - 1) * LF + line]), s.inv2dy);
      const float2 cd = make_float2(cw * d.y, -cw * d.x);
      const float2 pin = pos < L / 2 ? cscale(cadd(u, cd), 0.5f) : cscale(csub(u, cd), 0.5f);
      u = csub(u, cscale(pin, s.sa_strength * w));
    }
    if (m == 0 || m == g.nyq) u.y = 0.f;
    if (s.phase) {
      const float th = s_ph[i] * fs;
      float sn, cs;
      sincosf(th, &sn, &cs);
      u = cmul(u, make_float2(cs, -sn));
      if (s.amp) u = cscale(u, powf(s_am[i], fs));
      if (m == 0 || m == g.nyq) u.y = 0.f;
    }
    res[q] = u;
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int el = tid + q * LNT;
    if (el >= nel) break;
    if (line < nvalid) base[row_of(el / LF) * sy + line] = res[q];
  }
}

// ---------------------------------------------------------------------------
// t-column kernel: irfft -> real-space column work -> rfft.
// ---------------------------------------------------------------------------
enum ColMode : int {
  COL_LOAD = 0,    // real field in -> spectrum (+ I_before, max|p| of the input)
  COL_NL = 1,      // taper * nonlinear step
  COL_TAPER = 2,   // taper only
  COL_FINAL = 3,   // sources + reductions (+ spectrum rewrite if a source was added)
  COL_STORE = 4,   // spectrum -> real field
  COL_RAWNL = 5,   // real field in -> nonlinear step -> real field out (no FFT)
  COL_FINAL_SRC = 6,  // COL_FINAL on a step that injects source slices
  COL_TOREAL = 7,  // spectrum -> taper -> real field (the learned step's input)
  COL_MAXABS = 8,  // spectrum -> taper -> per-block max|p| into ColDesc::partial (read only)
  COL_TOF = 9,     // spectrum -> time of flight of the column into ColDesc::tof (read only)
};

// Time-of-flight estimators (ColDesc::tof_mode)
enum TofMode : int { TOF_OFF = 0, TOF_ENVELOPE = 1, TOF_MATCHED = 2 };

// CFL sub-cycling of the KT update (ColDesc::subcycle)
enum Subcycle : int { SUBCYCLE_OFF = 0, SUBCYCLE_COLUMN = 1, SUBCYCLE_GLOBAL = 2 };

enum NLKind : int { NL_NONE = 0, NL_KT = 1, NL_RUSANOV = 2, NL_TVD = 3, NL_KT_MINMOD = 4 };

struct ColDesc {
  int mode;
  int nl;            // NLKind
  int taper;         // 0 none, 1 sqrt taper (split), 2 full taper (sequential)
  int subcycle;      // Subcycle
  float cfl_target;
  const unsigned* gmax;  // SUBCYCLE_GLOBAL: float bits of the field's max|p|
  unsigned* subcap;      // when set, += 1 per column (or field) whose sub-steps hit MAX_SUBSTEPS
  unsigned* partial;     // COL_MAXABS: per-block max|p| bits, [gridDim.x]
  // time of flight (COL_TOF): retarded arrival time per column into tof[row][c]
  float* tof;
  int tof_mode;          // TofMode
  float tof_ratio;       // envelope threshold fraction
  double tof_t0, tof_dt; // time of sample 0 and the sample spacing of the time axis
  double tof_tref;       // matched filter: time at which the reference trace peaks
  const float2* tofref;  // matched filter: rfft of the reference trace [MP]
  double N, dT, dz;  // Burgers coefficient, time step, z step of the nonlinear update
  float beta_tvd;
  // real-field io
  const float* rin;  // COL_LOAD / COL_RAWNL input [nX*nY*nT]
  float* rout;       // COL_STORE / COL_RAWNL / COL_TOREAL / COL_FINAL(optional) output
  const int32_t* cflag;  // when set, only columns with cflag[c] != 0 are processed
  const double* nscale;  // when set, N is scaled by nscale[0] / (nscale[1] + 1e-30) (beam obliquity)
  // reductions (COL_FINAL / COL_LOAD)
  const int* row;    // device step counter; outputs go to row `*row`
  int max_rows;
  float* pnp; float* ppp; float* pI; float* pIloss; float* pax;
  unsigned* stats;   // [max_rows+1] max|p| bits; COL_LOAD writes slot 0, COL_FINAL slot row+1
  double* ibefore;   // [nX*nY] previous-step intensity (FP64)
  int cax;           // index of the on-axis column (nX/2, nY/2)
  // sources (COL_FINAL)
  const float* src[4];
  int n_src;
  const unsigned char* smask;
  const double* sdelay;
  const double* taxis;
  double s_f0, s_p0, s_ncyc, s_dur;
};

// Tone burst of the bowl and sparse-array sources in FP64, rounded to float32 once:
// p0 exp(-(1.05 tm w0 / (ncycles pi))^(2 dur)) sin(w0 tm), tm = t - delay
// (ASS:1061-1064, ASS:1196-1200, validate_transcranial.py:307-309).
__device__ __forceinline__ float tone_burst(double tm, double w0, double ncyc, double cexp, double p0) {
  const double env = exp(-pow(1.05 * tm * w0 / (ncyc * 3.14159265358979323846), cexp));
  return (float)(env * sin(w0 * tm) * p0);
}

// Dense field of a sparse source: out[c][t] = mask[c] ? tone_burst(taxis[t] - delay[c]) : 0.
__global__ __launch_bounds__(256) void k_source_field(const unsigned char* __restrict__ mask,
                                                      const double* __restrict__ delay,
                                                      const double* __restrict__ taxis, long ncols, int nT,
                                                      double w0, double ncyc, double cexp, double p0,
                                                      float* __restrict__ out) {
  const long n = ncols * (long)nT;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long c = e / nT;
    const int t = (int)(e - c * nT);
    out[e] = mask[c] ? tone_burst(taxis[t] - delay[c], w0, ncyc, cexp, p0) : 0.f;
  }
}

// Spectrum bin fed to the inverse transform of a column pass: the field itself
// (sel 0), its Hilbert transform (sel 1: -i X on bins 1..ceil(nT/2)-1, the DC and
// Nyquist bins dropped) or its circular cross-correlation with the reference
// trace (sel 2: X conj(R)).
__device__ __forceinline__ float2 tof_bin(int sel, int k, float2 x, int nyq, const float2* ref) {
  if (sel == 1) return (k == 0 || k == nyq) ? make_float2(0.f, 0.f) : make_float2(x.y, -x.x);
  if (sel == 2) return cmul(x, cconj(ref[k]));
  return x;
}

__device__ __forceinline__ float sgnf(float a) { return a > 0.f ? 1.f : (a < 0.f ? -1.f : 0.f); }

// Half the limited slope of a cell, 0.5 sigma, from its differences dm = u_i -
// u_{i-1} and dp = u_{i+1} - u_i (ASS:727-755).  minmod(a, b) (a b > 0 ? the one
// of smaller magnitude : 0) is the median of a, b and 0, one v_med3_f32, and
// minmod(a, b, c) = minmod(a, minmod(b, c)).  With powers of two scaling exactly:
//   MC:     0.5 minmod(2 dm, (dm + dp)/2, 2 dp) = minmod(dm, (dm + dp)/4, dp)
//   minmod: 0.5 minmod(dm, dp)
// (the reference tests a b > 0 on the products, which can underflow to zero for
// differences below ~1e-19 and return 0 where the median returns the difference).
__device__ __forceinline__ float med3f(float a, float b, float c) { return __builtin_amdgcn_fmed3f(a, b, c); }
template <bool MC>
__device__ __forceinline__ float half_slope(float dm, float dp) {
  if constexpr (MC) return med3f(dm, med3f(0.25f * (dm + dp), dp, 0.f), 0.f);
  return 0.5f * med3f(dm, dp, 0.f);
}

// KT central-upwind flux at an interface (ASS:771-793),
//   F = (a+ f- - a- f+ - a+ a- (u+ - u-)) / (a+ - a-),  f(u) = -u^2/2,
//   a+ = max(-u-, -u+, 0), a- = min(-u-, -u+, 0),
// returned as G = -2 F.  For Burgers the quotient has closed forms: u-^2 with
// both speeds >= 0 (u-, u+ <= 0), u+^2 with both <= 0, -3 u- u+ for u- > 0 > u+
// and u-^2 + u+^2 + 3 u- u+ for u- < 0 < u+.  With P = min(u-, 0), Q = max(u+, 0)
// and X = u- u+ for u- < 0 < u+, -u- u+ for u- > 0 > u+, 0 for states of one
// sign, all four are P^2 + Q^2 + 3 X, without a branch or a select.  X is
// copysign(min(u- u+, 0), u-): the product is negative exactly when the signs
// differ, and its magnitude takes the sign of u-.  (X equals P Q - R with
// R = max(u-, 0) min(u+, 0), bit for bit: both round the one product u- u+.)
// The Lax-Friedrichs fallback of ASS:792-796 acts only where both states lie
// below 1e-14 in magnitude; there both fluxes are below 1e-28 and change no
// single-precision sample through dz N/dT dF, so the closed form stands for it.
__device__ __forceinline__ float kt_iface_g(float um, float up) {
  const float P = fminf(um, 0.f), Q = fmaxf(up, 0.f);
  const float X = copysignf(fminf(um * up, 0.f), um);
  return fmaf(3.f, X, fmaf(Q, Q, P * P));
}

// Semi-discrete KT rhs for S samples t0..t0+S-1 from u[-2..S+1] (index offset 2),
// for columns whose ends fall anywhere in a lane (the general-length kernel).
template <int S, bool MC>
__device__ __forceinline__ void kt_rhs_local(const float* u, int t0, int nT, float coef, float* rhs) {
  float hs[S + 2];
#pragma unroll
  for (int i = 0; i < S + 2; ++i) {
    const int t = t0 - 1 + i;
    hs[i] = (t >= 1 && t <= nT - 2) ? half_slope<MC>(u[i + 1] - u[i], u[i + 2] - u[i + 1]) : 0.f;
  }
  float G[S + 1];
#pragma unroll
  for (int i = 0; i < S + 1; ++i) {
    // interface between t = t0-1+i and t0+i ; valid for 0 <= t0-1+i <= nT-2
    G[i] = kt_iface_g(u[i + 1] + hs[i], u[i + 2] - hs[i + 1]);
  }
  const float hc = 0.5f * coef;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const int t = t0 + i;
    rhs[i] = (t >= 1 && t <= nT - 2) ? hc * (G[i + 1] - G[i]) : 0.f;
  }
}

// Flux differences of the semi-discrete KT update for the S contiguous samples of
// one lane of an FFT column (t0 = S ln, nT = S LPC), from h = u[t0-2 .. t0+S+1].
// Output dG[k] = G(k, k+1) - G(k-1, k) with the interface flux F = -G/2, so the
// rhs of ASS:757-799 is N/(2 dT) dG.  The arithmetic of kt_rhs_local with the
// differences shared between neighbouring slopes.  The slope and the rhs at
// t = 0 and t = nT-1 are zero (ASS:765-766, 799); only lanes `first` (ln = 0)
// and `last` (ln = LPC-1) hold those samples.  The caller gives those lanes the
// end sample again as the first sample past the column's end (h[1] = u(0) on
// `first`, h[S+2] = u(nT-1) on `last`): the difference across the end is then
// zero, and a limited slope with one zero difference is zero, so the end slopes
// vanish without a select.  The rhs at the ends is masked here.
template <int S, bool MC>
__device__ __forceinline__ void kt_dflux_lane(const float* h, bool first, bool last, float* dG) {
  float d[S + 3];
#pragma unroll
  for (int j = 0; j < S + 3; ++j) d[j] = h[j + 1] - h[j];
  // half slopes of samples k = -1 .. S (index k + 1)
  float hs[S + 2];
#pragma unroll
  for (int k = 0; k < S + 2; ++k) hs[k] = half_slope<MC>(d[k], d[k + 1]);
  // G(k, k+1), k = -1 .. S-1 (index k + 1), between the states
  // u_k + sigma_k/2 and u_{k+1} - sigma_{k+1}/2
  float G[S + 1];
#pragma unroll
  for (int k = 0; k < S + 1; ++k) G[k] = kt_iface_g(h[k + 1] + hs[k], h[k + 2] - hs[k + 1]);
#pragma unroll
  for (int k = 0; k < S; ++k) dG[k] = G[k + 1] - G[k];
  if (first) dG[0] = 0.f;
  if (last) dG[S - 1] = 0.f;
}

}  // namespace ash

#include "column.h"

namespace ash {

__global__ void k_set_row(int* row, int v) { *row = v; }
__global__ void k_set_u32(unsigned* a, unsigned v) { *a = v; }
__global__ void k_copy_u32x2(const unsigned* a, const unsigned* b, unsigned* out) { out[0] = *a; out[1] = *b; }

// Partial maxima of |u| over a real field of ncols columns of nT samples: block b
// covers columns [64 b, 64 b + 64) and writes partial[b] (float bits).  The
// maximum is taken on the bit patterns of |u|, which order like the values and
// put a NaN above +inf, so a NaN sample shows in the result.
constexpr int ABSMAX_COLS = 64;
__global__ __launch_bounds__(256) void k_absmax_cols(const float* __restrict__ u, long ncols, int nT,
                                                     unsigned* __restrict__ partial) {
  __shared__ unsigned s_red[4];
  const long c0 = (long)blockIdx.x * ABSMAX_COLS;
  const long c1 = c0 + ABSMAX_COLS < ncols ? c0 + ABSMAX_COLS : ncols;
  const long e0 = c0 * nT, e1 = c1 * nT;
  unsigned m = 0u;
  for (long e = e0 + threadIdx.x; e < e1; e += 256) m = max(m, __float_as_uint(fabsf(u[e])));
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = max(max(s_red[0], s_red[1]), max(s_red[2], s_red[3]));
}

// out[0] = max of n partial maxima (float bits; all finite or +inf values are
// non-negative floats, whose bit patterns order like the values).  One block.
__global__ __launch_bounds__(1024) void k_reduce_max(const unsigned* __restrict__ partial, int n,
                                                     unsigned* __restrict__ out) {
  __shared__ unsigned s_red[16];
  unsigned m = 0u;
  for (int i = threadIdx.x; i < n; i += 1024) m = max(m, partial[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned r = s_red[0];
    for (int w = 1; w < 16; ++w) r = max(r, s_red[w]);
    out[0] = r;
  }
}
__global__ void k_zero2(double* a) { a[0] = 0.0; a[1] = 0.0; }

// y transform of one (kx row, f block) tile and the power-weighted obliquity
// sums of its plane-wave content (ASS:630-643): acc[0] += sum P k/k_z and
// acc[1] += sum P over the propagating modes, P = |F|^2, k = 2 pi f / c0,
// k_z = max(sqrt(k^2 - kx^2 - ky^2), 1e-6 k); f = 0 and evanescent modes
// weigh 0 (ASS:598-627).  The state is only read.
template <int L, int V>
__global__ __launch_bounds__(LNT) void k_ypower(const float2* __restrict__ spec, Tables T, Geom g,
                                                double* __restrict__ acc) {
  __shared__ float2 tw[L];
  __shared__ float2 buf[L * LF];
  __shared__ double sred[2][LNT / 64];
  const int tid = threadIdx.x;
  const int x = blockIdx.x / g.nfb, fb = blockIdx.x % g.nfb;
  const int f0 = fb * LF;
  const float2* base = spec + (long)x * g.nY * g.MP + f0;
  const int nvalid = min(LF, g.M + 1 - f0);
  load_tw<L>(tw, T.twy, tid, LNT);
  __syncthreads();
  using LY = Lay<L, LF, 0, true>;
  auto ld = [&](int line, int pos) -> float2 {
    return line < nvalid ? base[(long)pos * g.MP + line] : make_float2(0.f, 0.f);
  };
  auto st = [&](int line, int pos, float2 v) { buf[LY::idx(line, pos)] = v; };
  fft_run<L, false, LNT, LF, LY, V>(buf, tw, tid, ld, st);
  __syncthreads();
  static_assert(LNT % LF == 0, "the line of a thread must be fixed");
  const int line = tid % LF, m = f0 + line;
  const double km = (line < nvalid && m > 0) ? T.ka[m] : 0.0;
  const double kx = T.kx[x];
  double num = 0.0, den = 0.0;
  if (km > 0.0) {
    for (int pos = tid / LF; pos < L; pos += LNT / LF) {
      const double ky = T.ky[pos];
      const double kz2 = kz2_of(km, kperp2(kx, ky));
      if (kz2 > 0.0) {
        const float2 v = buf[LY::idx(line, pos)];
        const double P = (double)v.x * v.x + (double)v.y * v.y;
        num += P * km / fmax(sqrt(kz2), km * 1e-6);
        den += P;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    num += __shfl_xor(num, o, 64);
    den += __shfl_xor(den, o, 64);
  }
  if ((tid & 63) == 0) { sred[0][tid >> 6] = num; sred[1][tid >> 6] = den; }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < LNT / 64; ++w) { a += sred[0][w]; b += sred[1][w]; }
    atomicAdd(acc, a);
    atomicAdd(acc + 1, b);
  }
}
__global__ void k_advance_row(int* row) { *row += 1; }

}  // namespace ash
