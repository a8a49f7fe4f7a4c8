// t-column kernel: irfft_t -> real-space column work -> rfft_t, wave-synchronous.
//
// LPC lanes own one column (LPC = min(64, M/8)); lane ln holds the complex
// samples at positions ln + LPC*e (e < E = M/LPC) between FFT passes, so the
// Stockham passes exchange data through a per-column LDS region only between
// passes.  The LDS index n + n/8 keeps the strided transposes free of bank
// conflicts.  For the real-space work a lane owns S = 2E contiguous samples;
// the stencil halo comes from the neighbour lanes by shuffles.  No block-level
// barrier is used inside the column loop; each wave walks its own columns.
#pragma once
#include "kernels.h"

namespace ash {

// Most KT sub-steps taken in one dZ by CFL sub-cycling.  A column (per-column
// mode) or field (global mode) that would need more takes this many and is
// counted in ColDesc::subcap, so the host can report it.
constexpr int MAX_SUBSTEPS = 1024;

// Sub-steps for a CFL number cfl at target `target` (_kt_flux_adaptive,
// ASS:849-856): 1 at or below the target or for a field below 1e-30, else
// ceil(cfl / target), bounded by MAX_SUBSTEPS.  *capped is set when the bound
// (or a non-finite CFL number) took over.
__device__ __forceinline__ int substeps_for(double cfl, double target, float umax, bool* capped) {
  *capped = umax != umax;  // NaN amplitude: one step, reported
  if (!(umax >= 1e-30f)) return 1;
  const double q = cfl / target;
  if (q <= 1.0) return 1;
  if (!(q <= (double)MAX_SUBSTEPS)) {  // beyond the bound, or NaN
    *capped = true;
    return q > (double)MAX_SUBSTEPS ? MAX_SUBSTEPS : 1;
  }
  return (int)ceil(q);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#ifndef ASH_COL_PAD_SHIFT
#define ASH_COL_PAD_SHIFT 3
#endif
__device__ __forceinline__ int pidx(int n) { return n + (n >> ASH_COL_PAD_SHIFT); }

template <int M> struct ColRadix;
template <> struct ColRadix<128>  { static constexpr int NP = 3, R0 = 8,  R1 = 4, R2 = 4; };
template <> struct ColRadix<256>  { static constexpr int NP = 3, R0 = 8,  R1 = 8, R2 = 4; };
template <> struct ColRadix<512>  { static constexpr int NP = 3, R0 = 8,  R1 = 8, R2 = 8; };
template <> struct ColRadix<1024> { static constexpr int NP = 3, R0 = 16, R1 = 8, R2 = 8; };

template <int M>
struct ColCfg {
  static constexpr int LPC = (M / 8) < 64 ? (M / 8) : 64;  // lanes per column
  static constexpr int E = M / LPC;                          // complex samples per lane
  static constexpr int CPW = 64 / LPC;                       // columns per wave
  static constexpr int S = 2 * E;                            // real samples per lane
  static constexpr int LDSN = M + (M >> ASH_COL_PAD_SHIFT) + 4;  // padded per-column LDS (float2)
  static constexpr int WPB = 4;                              // waves per block
};

// Pass metadata for the column transform; direction INV runs the radices reversed.
template <int M, bool INV>
struct CDir {
  using RX = ColRadix<M>;
  static constexpr int NP = RX::NP;
  static constexpr int rad(int p) {
    const int i = INV ? NP - 1 - p : p;
    return i == 0 ? RX::R0 : (i == 1 ? RX::R1 : RX::R2);
  }
  static constexpr int ns(int p) {
    int n = 1;
    for (int q = 0; q < p; ++q) n *= rad(q);
    return n;
  }
  // twiddle table of pass p (p >= 1): [r * Ns + k] = W_{Ns R}^{k r}
  static constexpr int toff(int p) {
    int o = 0;
    for (int q = 1; q < p; ++q) o += ns(q) * rad(q);
    return o;
  }
  static constexpr int total() { return toff(NP); }
};

template <int M>
constexpr int col_tw_total() { return CDir<M, false>::total() + CDir<M, true>::total(); }
// A radix list that reads the same reversed gives the inverse transform the
// forward table (the direction only conjugates at use): one copy in LDS.
template <int M>
constexpr bool col_tw_shared() {
  return CDir<M, false>::rad(0) == CDir<M, true>::rad(0) && CDir<M, false>::rad(1) == CDir<M, true>::rad(1) &&
         CDir<M, false>::rad(2) == CDir<M, true>::rad(2);
}

// One Stockham pass of a column transform; data in registers v[e] at position ln + LPC e.
template <int M, int P, bool INV>
__device__ __forceinline__ void cpass(float2* v, float2* cb, const float2* tw, int ln) {
  using C = ColCfg<M>;
  using D = CDir<M, INV>;
  constexpr int R = D::rad(P), Ns = D::ns(P), Q = C::E / R, LPC = C::LPC;
  constexpr bool LAST = (P == D::NP - 1);
  float2 w[Q][R];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = ln + q * LPC;
#pragma unroll
    for (int r = 0; r < R; ++r) w[q][r] = v[q + r * Q];
    if constexpr (Ns > 1) {
      const int k = j % Ns;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 t = tw[D::toff(P) + r * Ns + k];
        if (INV) t.y = -t.y;
        w[q][r] = cmul(w[q][r], t);
      }
    }
    dft<R, INV>(w[q]);
  }
  if constexpr (LAST) {
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * Q] = w[q][r];
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = ln + q * LPC;
#pragma unroll
      for (int r = 0; r < R; ++r) cb[pidx((j / Ns) * Ns * R + (j % Ns) + r * Ns)] = w[q][r];
    }
    wave_sync();
#pragma unroll
    for (int e = 0; e < C::E; ++e) v[e] = cb[pidx(ln + LPC * e)];
    wave_sync();
  }
}

template <int M, bool INV>
__device__ __forceinline__ void cfft(float2* v, float2* cb, const float2* twdir, int ln) {
  cpass<M, 0, INV>(v, cb, twdir, ln);
  cpass<M, 1, INV>(v, cb, twdir, ln);
  cpass<M, 2, INV>(v, cb, twdir, ln);
}

// shuffle within the wave from lane src (caller masks column edges)
__device__ __forceinline__ float shf(float x, int src) { return __shfl(x, src, 64); }

// Cross-lane moves by DPP (no LDS traffic): x from the lane that CTRL names, in
// the rows RMASK selects; lanes without a source (or outside RMASK) take id.
template <int CTRL, int RMASK>
__device__ __forceinline__ float dpp_f(float x, float id) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(id), __float_as_int(x), CTRL, RMASK, 0xF, false));
}
template <int CTRL, int RMASK>
__device__ __forceinline__ double dpp_d0(double x) {  // identity 0
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, RMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, RMASK, 0xF, false);
  return __hiloint2double(hi, lo);
}
// KT halo of a lane, h[0..1] and h[S+2..S+3] around its samples x[0..S): the
// previous lane's last two samples and the next lane's first two; past the
// column's ends the end sample again, then zero (the end slopes vanish).  With
// one column per wave the neighbours come by DPP wave shifts and the wave's end
// lanes keep the fallback; narrower columns shuffle and mask.
template <int S, int LPC>
__device__ __forceinline__ void kt_halo(float* h, const float* x, int lsrc, int rsrc, bool has_l, bool has_r) {
  if constexpr (LPC == 64) {
    h[0] = dpp_f<0x138, 0xF>(x[S - 2], 0.f);       // wave_shr:1, lane i from lane i-1
    h[1] = dpp_f<0x138, 0xF>(x[S - 1], x[0]);
    h[S + 2] = dpp_f<0x130, 0xF>(x[0], x[S - 1]);  // wave_shl:1, lane i from lane i+1
    h[S + 3] = dpp_f<0x130, 0xF>(x[1], 0.f);
  } else {
    const float a = shf(x[S - 2], lsrc), b = shf(x[S - 1], lsrc);
    const float cc = shf(x[0], rsrc), dd = shf(x[1], rsrc);
    h[0] = has_l ? a : 0.f; h[1] = has_l ? b : x[0];
    h[S + 2] = has_r ? cc : x[S - 1]; h[S + 3] = has_r ? dd : 0.f;
  }
#pragma unroll
  for (int i = 0; i < S; ++i) h[i + 2] = x[i];
}

template <int CTRL, int RMASK>
__device__ __forceinline__ void col_red_step(float& mn, float& mx, double& e2) {
  mn = fminf(mn, dpp_f<CTRL, RMASK>(mn, INFINITY));
  mx = fmaxf(mx, dpp_f<CTRL, RMASK>(mx, -INFINITY));
  e2 += dpp_d0<CTRL, RMASK>(e2);
}
// Minimum, maximum and sum over the LPC lanes of each column (aligned groups of
// 16, 32 or 64 lanes): an inclusive scan along each 16-lane row by shifts of 1,
// 2, 4 and 8 lanes, then lane 15 of each row broadcast into the next row and
// lane 31 into rows 2 and 3.  The column's result lands in its last lane
// (ln = LPC - 1); the other lanes hold partial results.
template <int LPC>
__device__ __forceinline__ void col_reduce_last(float& mn, float& mx, double& e2) {
  static_assert(LPC == 16 || LPC == 32 || LPC == 64, "columns of 16, 32 or 64 lanes");
  col_red_step<0x111, 0xF>(mn, mx, e2);  // row_shr:1
  col_red_step<0x112, 0xF>(mn, mx, e2);  // row_shr:2
  col_red_step<0x114, 0xF>(mn, mx, e2);  // row_shr:4
  col_red_step<0x118, 0xF>(mn, mx, e2);  // row_shr:8
  if constexpr (LPC >= 32) col_red_step<0x142, 0xA>(mn, mx, e2);  // row_bcast:15 into rows 1 and 3
  if constexpr (LPC == 64) col_red_step<0x143, 0xC>(mn, mx, e2);  // row_bcast:31 into rows 2 and 3
}

// Lane view of one column: this lane's S real samples u[i] sit at t = t(i).
// STRIDED: the FFT's ownership (pairs 2(ln + LPC e), +1); else contiguous
// samples t0 .. t0 + S - 1.  BOUNDED: the column has nT <= LPC S samples and
// the lane's samples at t >= nT are padding (zero, never read or written).
template <int S, int LPC, bool STRIDED, bool BOUNDED>
struct LaneCol {
  int ln, lane, t0, nT;
  __device__ __forceinline__ int t(int i) const {
    if constexpr (STRIDED) return 2 * (ln + LPC * (i >> 1)) + (i & 1);
    else return t0 + i;
  }
  __device__ __forceinline__ bool in(int i) const { return !BOUNDED || t(i) < nT; }
  // store u[0..S) into the real column starting at dst
  __device__ __forceinline__ void store(float* dst, const float* u) const {
    static_assert(!(STRIDED && BOUNDED), "bounded columns use contiguous samples");
    if constexpr (STRIDED) {
      float2* d2 = reinterpret_cast<float2*>(dst);
#pragma unroll
      for (int e = 0; e < S / 2; ++e) d2[ln + LPC * e] = make_float2(u[2 * e], u[2 * e + 1]);
    } else if constexpr (!BOUNDED) {
      float4* d4 = reinterpret_cast<float4*>(dst + t0);
#pragma unroll
      for (int q = 0; q < S / 4; ++q) d4[q] = make_float4(u[4 * q], u[4 * q + 1], u[4 * q + 2], u[4 * q + 3]);
    } else {
#pragma unroll
      for (int i = 0; i < S; ++i)
        if (t0 + i < nT) dst[t0 + i] = u[i];
    }
  }
};

// Real-space part of a column pass (after the inverse transform, before the
// forward one): taper, nonlinear update with its sub-cycling, source injection,
// reductions and real outputs, on the S samples of this lane.  Shared by the
// FFT column kernel and the general-length column kernel.
template <int S, int LPC, bool STRIDED, bool BOUNDED, int MODE, int NLK>
__device__ __forceinline__ void column_realspace(float (&u)[S], const LaneCol<S, LPC, STRIDED, BOUNDED>& L,
                                                 const ColDesc& d, const Tables& T, int c, bool live, int ncols,
                                                 int row, double Nn, int nsub_global, float& wmax) {
  constexpr bool FIN = (MODE == COL_FINAL || MODE == COL_FINAL_SRC);
  const float* tt = d.taper == 1 ? T.tapt_h : T.tapt_f;
  if constexpr (MODE == COL_NL || MODE == COL_TAPER || MODE == COL_RAWNL) {
    if (d.taper) {
#pragma unroll
      for (int i = 0; i < S; ++i) u[i] *= (L.in(i) ? tt[L.t(i)] : 0.f);
    }
    if constexpr (NLK != NL_NONE && MODE != COL_TAPER) {
      int nsub = 1;
      if (d.subcycle == SUBCYCLE_COLUMN) {
        float mx = 0.f;
#pragma unroll
        for (int i = 0; i < S; ++i) mx = fmaxf(mx, fabsf(u[i]));
        for (int o = LPC / 2; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        bool capped;
        nsub = substeps_for(Nn * (double)mx * d.dz / d.dT, (double)d.cfl_target, mx, &capped);
        if (capped && live && L.ln == 0 && d.subcap) atomicAdd(d.subcap, 1u);
      } else if (d.subcycle == SUBCYCLE_GLOBAL) {
        nsub = nsub_global;
      }
      int nsub_w = nsub;
      for (int o = 32; o >= LPC; o >>= 1) nsub_w = max(nsub_w, __shfl_xor(nsub_w, o, 64));
      const float dz = (float)(d.dz / nsub);
      const float coef = (float)(Nn / d.dT);
      // dz N / (2 dT) of one sub-step: u + dz rhs = u + hcdz dG (no division
      // for the usual single sub-step)
      const double hc1 = 0.5 * Nn / d.dT;
      double hcd = hc1 * d.dz;
      if (nsub != 1) hcd = hc1 * (d.dz / nsub);
      const float hcdz = (float)hcd, hh = 0.5f * hcdz;
      const int lsrc = L.lane - 1, rsrc = L.lane + 1;
      const bool has_l = L.ln > 0, has_r = L.ln < LPC - 1;
      for (int sub = 0; sub < nsub_w; ++sub) {
        // (with one column per wave nsub is the wave's count: every sub-step is taken)
        const bool act = LPC == 64 || sub < nsub;
        if constexpr ((NLK == NL_KT || NLK == NL_KT_MINMOD) && !BOUNDED) {
          // SSP-RK2 as s1 = u + hcdz dG(u), u' = (u + s1 + hcdz dG(s1)) / 2, kept
          // as w = (u + s1) / 2 = u + hh dG(u) and u' = w + hh dG(s1), hh = hcdz / 2
          float h[S + 4], dG[S], s1[S];
          const bool first = L.ln == 0, last = L.ln == LPC - 1;
          kt_halo<S, LPC>(h, u, lsrc, rsrc, has_l, has_r);
          kt_dflux_lane<S, NLK == NL_KT>(h, first, last, dG);
#pragma unroll
          for (int i = 0; i < S; ++i) {
            s1[i] = act ? fmaf(hcdz, dG[i], u[i]) : u[i];
            u[i] = act ? fmaf(hh, dG[i], u[i]) : u[i];  // w
          }
          kt_halo<S, LPC>(h, s1, lsrc, rsrc, has_l, has_r);
          kt_dflux_lane<S, NLK == NL_KT>(h, first, last, dG);
          if (act) {
#pragma unroll
            for (int i = 0; i < S; ++i) u[i] = fmaf(hh, dG[i], u[i]);
          }
        } else if constexpr (NLK == NL_KT || NLK == NL_KT_MINMOD) {
          float h[S + 4], r0[S], s1[S];
          // halo: 2 samples from each neighbour lane (zero outside the column)
          {
            const float a = shf(u[S - 2], lsrc), b = shf(u[S - 1], lsrc);
            const float cc = shf(u[0], rsrc), dd = shf(u[1], rsrc);
            h[0] = has_l ? a : 0.f; h[1] = has_l ? b : 0.f;
            h[S + 2] = has_r ? cc : 0.f; h[S + 3] = has_r ? dd : 0.f;
#pragma unroll
            for (int i = 0; i < S; ++i) h[i + 2] = u[i];
          }
          kt_rhs_local<S, NLK == NL_KT>(h, L.t0, L.nT, coef, r0);
#pragma unroll
          for (int i = 0; i < S; ++i) s1[i] = act ? u[i] + dz * r0[i] : u[i];
          {
            const float a = shf(s1[S - 2], lsrc), b = shf(s1[S - 1], lsrc);
            const float cc = shf(s1[0], rsrc), dd = shf(s1[1], rsrc);
            h[0] = has_l ? a : 0.f; h[1] = has_l ? b : 0.f;
            h[S + 2] = has_r ? cc : 0.f; h[S + 3] = has_r ? dd : 0.f;
#pragma unroll
            for (int i = 0; i < S; ++i) h[i + 2] = s1[i];
          }
          kt_rhs_local<S, NLK == NL_KT>(h, L.t0, L.nT, coef, r0);
          if (act) {
#pragma unroll
            for (int i = 0; i < S; ++i) u[i] = 0.5f * u[i] + 0.5f * (s1[i] + dz * r0[i]);
          }
        } else {
          float h[S + 2];
          {
            const float a = shf(u[S - 1], lsrc), b = shf(u[0], rsrc);
            h[0] = has_l ? a : 0.f;
            h[S + 1] = has_r ? b : 0.f;
#pragma unroll
            for (int i = 0; i < S; ++i) h[i + 1] = u[i];
          }
          const float cf = (float)(Nn * (double)dz / d.dT);
          float fl[S + 1];
#pragma unroll
          for (int i = 0; i < S + 1; ++i) {
            const float uL = h[i], uR = h[i + 1];
            const float a = fmaxf(fabsf(uL), fabsf(uR));
            if constexpr (NLK == NL_RUSANOV) fl[i] = -(uL * uL + uR * uR) / 2.f - a * (uR - uL);
            else fl[i] = 0.5f * (-0.5f * (uL * uL + uR * uR)) - 0.5f * a * (uR - uL);
          }
#pragma unroll
          for (int i = 0; i < S; ++i) {
            const int t = L.t0 + i;
            if (act && t >= 1 && t <= L.nT - 2) {
              const float fd = fl[i + 1] - fl[i];
              if constexpr (NLK == NL_RUSANOV) {
                u[i] += -cf * fd;
              } else {
                const float dm = h[i + 1] - h[i], dp = h[i + 2] - h[i + 1];
                const float eps = 1e-30f, b = d.beta_tvd;
                const float r = dp / (dm + eps * sgnf(dm + eps));
                const float ri = dm / (dp + eps * sgnf(dp + eps));
                const float pp = fmaxf(0.f, fminf(fminf(b * r, 1.f), fminf(r, b)));
                const float pm = fmaxf(0.f, fminf(fminf(b * ri, 1.f), fminf(ri, b)));
                u[i] += -cf * (0.5f * (pp + pm) * fd);
              }
            }
          }
        }
      }
    }
    if (MODE == COL_RAWNL && live) L.store(d.rout + (long)c * L.nT, u);
  } else if constexpr (FIN) {
    if constexpr (MODE == COL_FINAL_SRC) {
      // loads without a test each (clamped into the column, padding masked after),
      // so they issue back to back
      if (live) {
        for (int k = 0; k < d.n_src; ++k) {
          const float* sp = d.src[k] + (long)c * L.nT;
          float sv[S];
#pragma unroll
          for (int i = 0; i < S; ++i) sv[i] = sp[min(L.t(i), L.nT - 1)];
#pragma unroll
          for (int i = 0; i < S; ++i) u[i] += L.in(i) ? sv[i] : 0.f;
        }
      }
      if (d.smask && live && d.smask[c]) {
        const double dl = d.sdelay[c];
        const double w0 = 2.0 * 3.14159265358979323846 * d.s_f0;
        const double cexp = 2.0 * d.s_dur;
        float tv[S];
#pragma unroll
        for (int i = 0; i < S; ++i) tv[i] = d.taxis[min(L.t(i), L.nT - 1)];
#pragma unroll
        for (int i = 0; i < S; ++i) u[i] += L.in(i) ? tone_burst(tv[i] - dl, w0, d.s_ncyc, cexp, d.s_p0) : 0.f;
      }
    }
    // (padding samples of a bounded lane are zero: they add nothing to e2 but are
    // kept out of the extrema)
    float mn = BOUNDED ? INFINITY : u[0], mx = BOUNDED ? -INFINITY : u[0];
    double e2 = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      mn = fminf(mn, L.in(i) ? u[i] : INFINITY);
      mx = fmaxf(mx, L.in(i) ? u[i] : -INFINITY);
      e2 += (double)u[i] * (double)u[i];
    }
    // (the column's totals in its last lane; wmax below takes the block's maximum
    // over every lane, partial results included)
    col_reduce_last<LPC>(mn, mx, e2);
    if (live && row < d.max_rows) {
      if (L.ln == LPC - 1) {
        const long o = (long)row * ncols + c;
        const double ib = d.ibefore[c];
        d.pnp[o] = mn;
        d.ppp[o] = mx;
        d.pI[o] = (float)e2;
        d.pIloss[o] = (float)fmax(0.0, ib - e2);
        d.ibefore[c] = e2;
      }
      if (c == d.cax) L.store(d.pax + (long)row * L.nT, u);
    }
    if (live) wmax = fmaxf(wmax, fmaxf(mx, -mn));
    if (d.rout && live) L.store(d.rout + (long)c * L.nT, u);
  } else if constexpr (MODE == COL_LOAD) {
    float mx = 0.f;
    double e2 = 0.0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      mx = fmaxf(mx, fabsf(u[i]));
      e2 += (double)u[i] * (double)u[i];
    }
    for (int o = LPC / 2; o > 0; o >>= 1) {
      mx = fmaxf(mx, __shfl_xor(mx, o, 64));
      e2 += __shfl_xor(e2, o, 64);
    }
    if (live && L.ln == 0 && d.ibefore) d.ibefore[c] = e2;
    if (live) wmax = fmaxf(wmax, mx);
  } else if constexpr (MODE == COL_STORE) {
    if (live) L.store(d.rout + (long)c * L.nT, u);
  } else if constexpr (MODE == COL_TOREAL) {
    if (d.taper) {
#pragma unroll
      for (int i = 0; i < S; ++i) u[i] *= (L.in(i) ? tt[L.t(i)] : 0.f);
    }
    if (live) L.store(d.rout + (long)c * L.nT, u);
  } else if constexpr (MODE == COL_MAXABS) {
    // the field the nonlinear pass will see (same transform, same taper), read only
    if (d.taper) {
#pragma unroll
      for (int i = 0; i < S; ++i) u[i] *= (L.in(i) ? tt[L.t(i)] : 0.f);
    }
#pragma unroll
    for (int i = 0; i < S; ++i) wmax = fmaxf(wmax, live ? fabsf(u[i]) : 0.f);
  }
}

// True on every lane of the column when any of its rfft bins 0..M is nonzero.
template <int LPC>
__device__ __forceinline__ bool spectrum_nonzero(const float2* sc, int M, int ln) {
  bool nz = false;
  for (int k = ln; k <= M; k += LPC) {
    const float2 x = sc[k];
    nz = nz || x.x != 0.f || x.y != 0.f;
  }
  int f = nz ? 1 : 0;
  for (int o = LPC / 2; o > 0; o >>= 1) f |= __shfl_xor(f, o, 64);
  return f != 0;
}

// Time of flight of one column (retarded time; the host adds z / c0), written to
// d.tof[row][c] by the column's first lane.
//   TOF_ENVELOPE: u = p, q = its Hilbert transform; the first sample where the
//     envelope sqrt(u^2 + q^2) reaches tof_ratio of its maximum, interpolated
//     linearly from the sample before (the first sample itself when it is 0 or
//     when nothing reaches the threshold).
//   TOF_MATCHED: u = circular cross-correlation with the reference; its first
//     maximum k refined by the vertex of the parabola through k-1, k, k+1
//     (circular), the lag in (-nT/2, nT/2], plus the reference's peak time.
//   A column without signal (nonzero = false) gets NaN.
// The values pass through the column's LDS region `lds` (nT floats).
template <int S, int LPC, bool STRIDED, bool BOUNDED>
__device__ __forceinline__ void column_tof(const float* u, const float* q, const LaneCol<S, LPC, STRIDED, BOUNDED>& L,
                                           float* lds, const ColDesc& d, int c, bool live, int ncols, int row,
                                           bool nonzero) {
  const bool env = d.tof_mode == TOF_ENVELOPE;
  const int nT = L.nT;
  float val[S];
  float best = -INFINITY;
  int bidx = 0x7fffffff;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const float v = env ? sqrtf(u[i] * u[i] + q[i] * q[i]) : u[i];
    val[i] = L.in(i) ? v : -INFINITY;
    const int t = L.t(i);
    if (val[i] > best || (val[i] == best && t < bidx)) { best = val[i]; bidx = t; }
  }
  for (int o = LPC / 2; o > 0; o >>= 1) {
    const float ob = __shfl_xor(best, o, 64);
    const int oi = __shfl_xor(bidx, o, 64);
    if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
  }
#pragma unroll
  for (int i = 0; i < S; ++i)
    if (L.in(i)) lds[L.t(i)] = val[i];
  int k = bidx;
  float thr = 0.f;
  if (env) {
    thr = d.tof_ratio * best;
    k = 0x7fffffff;
#pragma unroll
    for (int i = 0; i < S; ++i)
      if (L.in(i) && val[i] >= thr) k = min(k, L.t(i));
    for (int o = LPC / 2; o > 0; o >>= 1) k = min(k, __shfl_xor(k, o, 64));
    if (k == 0x7fffffff) k = 0;
  }
  wave_sync();
  if (live && L.ln == 0 && row < d.max_rows) {
    double res;
    if (env) {
      if (k == 0) {
        res = d.tof_t0;
      } else {
        const float e0 = lds[k - 1], e1 = lds[k];
        const float frac = e1 > e0 ? (thr - e0) / (e1 - e0) : 1.f;
        res = d.tof_t0 + ((double)(k - 1) + (double)frac) * d.tof_dt;
      }
      if (!(best > 0.f)) res = NAN;
    } else {
      const float cm = lds[k == 0 ? nT - 1 : k - 1], cp = lds[k == nT - 1 ? 0 : k + 1];
      const float den = cm - 2.f * best + cp;
      const float delta = den < 0.f ? 0.5f * (cm - cp) / den : 0.f;
      double lag = (double)k + (double)delta;
      if (lag > 0.5 * nT) lag -= nT;
      res = d.tof_tref + lag * d.tof_dt;
      if (!nonzero) res = NAN;
    }
    d.tof[(long)row * ncols + c] = (float)res;
  }
  wave_sync();
}

template <int M, int MODE, int NLK>
__global__ __launch_bounds__(256) void k_column(float2* __restrict__ spec, Tables T, Geom g, ColDesc d) {
  using C = ColCfg<M>;
  constexpr int E = C::E, LPC = C::LPC, CPW = C::CPW, S = C::S, NT_ = 2 * M;
  constexpr int TWF = CDir<M, false>::total();
  constexpr int TWT = col_tw_shared<M>() ? TWF : col_tw_total<M>();
  __shared__ float2 s_tw[TWT];
  __shared__ float2 s_twn[M / 2 + 1];  // W_N^k / 2, k <= M/2 (the rest by symmetry)
  __shared__ float2 s_buf[C::WPB * CPW * C::LDSN];
  __shared__ float s_red[C::WPB];
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int cw = lane / LPC, ln = lane % LPC;
  constexpr bool spectral_in = !(MODE == COL_LOAD || MODE == COL_RAWNL);
  constexpr bool rewrite = (MODE == COL_NL || MODE == COL_TAPER || MODE == COL_LOAD || MODE == COL_FINAL_SRC);
  if constexpr (spectral_in || rewrite) {
    // (COL_RAWNL runs no transform and may be launched without tables)
    for (int i = tid; i < TWT; i += 256) s_tw[i] = T.twc[i];
    for (int i = tid; i <= M / 2; i += 256) s_twn[i] = cscale(T.twgt[i], 0.5f);  // exp(-2 pi i k / nT) / 2
  }
  __syncthreads();
  const float2* twf = s_tw;
  const float2* twi = col_tw_shared<M>() ? s_tw : s_tw + TWF;
  // W_N^k / 2 at bin k = ln + LPC e: bins of the lower half (e < E/2) from the
  // table, the upper half from W_N^k = -conj(W_N^(M-k)), M-k <= M/2 (e is a
  // compile-time index, so no select).  M-k is addressed from its smallest
  // value up, so every read is one base register plus a constant offset.
  const float2* twlo = s_twn + ln;
  const float2* twhi = s_twn + (M - ln - LPC * (E - 1));
  auto twn = [&](int e) -> float2 {
    if (e < E / 2) return twlo[LPC * e];
    const float2 w = twhi[LPC * (E - 1 - e)];
    return make_float2(-w.x, w.y);
  };
  float2* cb = s_buf + (wave * CPW + cw) * C::LDSN;
  const int ncols = g.nX * g.nY;
  const int per_iter = C::WPB * CPW;
  const int t0 = ln * S;
  const float invN = 1.f / (float)NT_;
  // pointwise modes keep the FFT's strided ownership (lane ln holds samples
  // 2(ln + LPC e) and 2(ln + LPC e) + 1); only the t-stencil modes transpose to
  // contiguous samples
  constexpr bool STRIDED = !(MODE == COL_NL || MODE == COL_RAWNL);
  const LaneCol<S, LPC, STRIDED, false> lc{ln, lane, t0, NT_};
  constexpr bool FIN = (MODE == COL_FINAL || MODE == COL_FINAL_SRC);
  const int row = (FIN || MODE == COL_TOF) ? *d.row : 0;
  // Burgers coefficient, scaled by the beam obliquity when requested (ASS:934)
  const double Nn = d.nscale ? d.N * (d.nscale[0] / (d.nscale[1] + 1e-30)) : d.N;
  float wmax = 0.f;  // running max|p| of this lane's columns (stats, COL_MAXABS)
  // global-max sub-cycling (_kt_flux_adaptive, ASS:842-859): one count for the field
  int nsub_global = 1;
  if constexpr (NLK != NL_NONE) {
    if (d.subcycle == SUBCYCLE_GLOBAL) {
      const float gm = __uint_as_float(*d.gmax);
      bool capped;
      nsub_global = substeps_for(Nn * (double)gm * d.dz / d.dT, (double)d.cfl_target, gm, &capped);
      if (capped && blockIdx.x == 0 && tid == 0 && d.subcap) atomicAdd(d.subcap, 1u);
    }
  }

  const float2* sc_cur = nullptr;  // spectrum row of the current column (null: not live)
  // irfft of the column's spectrum (bins modified by tof_bin(sel)) into this lane's
  // samples u[0..S), in the ownership of this mode
  auto inv_spec = [&](float (&uo)[S], int sel) {
    // all loads of the column first, under one branch (a test per load would
    // wait for each load before issuing the next)
    float2 v[E], xn;
    if (sc_cur) {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = sc_cur[ln + LPC * e];
      xn = sc_cur[M];
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = make_float2(0.f, 0.f);
      xn = make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = ln + LPC * e;
      float2 x = v[e];
      if (k == 0) x.y = 0.f;
      x = tof_bin(sel, k, x, M, d.tofref);
      v[e] = x;
      cb[pidx(k)] = x;
    }
    if (ln == 0) {
      xn.y = 0.f;
      xn = tof_bin(sel, M, xn, M, d.tofref);
      cb[pidx(M)] = xn;
    }
    wave_sync();
    // Z[k] = (X_k + conj X_{M-k}) + i W_N^{-k} (X_k - conj X_{M-k}), scaled by 1/N
    // (the table holds W_N^k / 2: t is half the product, and 2 t is exact)
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = ln + LPC * e;
      const float2 a = v[e], b = cconj(cb[pidx(M - k)]);
      const float2 sm = cadd(a, b), df = csub(a, b);
      const float2 t = cmul(cconj(twn(e)), df);
      v[e] = make_float2(fmaf(-2.f, t.y, sm.x) * invN, fmaf(2.f, t.x, sm.y) * invN);
    }
    wave_sync();
    cfft<M, true>(v, cb, twi, ln);
    if constexpr (STRIDED) {
#pragma unroll
      for (int e = 0; e < E; ++e) { uo[2 * e] = v[e].x; uo[2 * e + 1] = v[e].y; }
    } else {
      // transpose to contiguous ownership: lane ln gets z[E ln .. E ln + E)
#pragma unroll
      for (int e = 0; e < E; ++e) cb[pidx(ln + LPC * e)] = v[e];
      wave_sync();
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const float2 z = cb[pidx(E * ln + q)];
        uo[2 * q] = z.x;
        uo[2 * q + 1] = z.y;
      }
      wave_sync();
    }
  };
  for (int cbase = blockIdx.x * per_iter; cbase < ncols; cbase += gridDim.x * per_iter) {
    const int c = cbase + wave * CPW + cw;
    const bool live = c < ncols && (!d.cflag || d.cflag[c] != 0);
    if (CPW == 1 && !live) continue;  // one column per wave: skip it whole
    float2* sc = spec + (long)c * g.MP;
    sc_cur = live ? sc : nullptr;
    float u[S];
    float uq[MODE == COL_TOF ? S : 1];

    // ---------------- stage 1: real samples t0 .. t0+S-1 of the column ----------------
    if constexpr (!spectral_in && STRIDED) {
      const float2* src = reinterpret_cast<const float2*>(d.rin + (long)c * NT_);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const float2 x = live ? src[ln + LPC * e] : make_float2(0.f, 0.f);
        u[2 * e] = x.x; u[2 * e + 1] = x.y;
      }
    } else if constexpr (!spectral_in) {
      const float4* src = reinterpret_cast<const float4*>(d.rin + (long)c * NT_ + t0);
#pragma unroll
      for (int q = 0; q < S / 4; ++q) {
        const float4 x = live ? src[q] : make_float4(0.f, 0.f, 0.f, 0.f);
        u[4 * q] = x.x; u[4 * q + 1] = x.y; u[4 * q + 2] = x.z; u[4 * q + 3] = x.w;
      }
    } else if constexpr (MODE == COL_TOF) {
      // the field (envelope: and its Hilbert transform) or its correlation with the reference
      if (d.tof_mode == TOF_ENVELOPE) {
        inv_spec(u, 0);
        inv_spec(uq, 1);
      } else {
        inv_spec(u, 2);
      }
    } else {
      inv_spec(u, 0);
    }

    // ---------------- stage 2: real-space work ----------------
    if constexpr (MODE == COL_TOF) {
      const bool nz = spectrum_nonzero<LPC>(sc, live ? M : -1, ln);
      column_tof<S, LPC, STRIDED, false>(u, uq, lc, reinterpret_cast<float*>(cb), d, c, live, ncols, row, nz);
    } else {
      column_realspace<S, LPC, STRIDED, false, MODE, NLK>(u, lc, d, T, c, live, ncols, row, Nn, nsub_global, wmax);
    }

    // ---------------- stage 3: rfft back into the spectrum ----------------
    if constexpr (rewrite) {
      float2 v[E];
      if constexpr (STRIDED) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = make_float2(u[2 * e], u[2 * e + 1]);
      } else {
#pragma unroll
        for (int q = 0; q < E; ++q) cb[pidx(E * ln + q)] = make_float2(u[2 * q], u[2 * q + 1]);
        wave_sync();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = cb[pidx(ln + LPC * e)];
        wave_sync();
      }
      cfft<M, false>(v, cb, twf, ln);
#pragma unroll
      for (int e = 0; e < E; ++e) cb[pidx(ln + LPC * e)] = v[e];
      wave_sync();
      const float2 z0 = cb[0];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int k = ln + LPC * e;
        const float2 zk = v[e];
        const float2 zm = cconj(cb[pidx(k == 0 ? 0 : M - k)]);
        // X_k = (Z_k + conj Z_{M-k}) / 2 - i W_N^k (Z_k - conj Z_{M-k}) / 2, the
        // halves folded into the table and the fmas
        const float2 df = csub(zk, zm);
        const float2 t = cmul(twn(e), df);
        if (live) sc[k] = make_float2(fmaf(0.5f, zk.x, fmaf(0.5f, zm.x, t.y)), fmaf(0.5f, zk.y, fmaf(0.5f, zm.y, -t.x)));
      }
      if (ln == 0 && live) sc[M] = make_float2(z0.x - z0.y, 0.f);
      wave_sync();
    }
  }

  // max|p| of the block -> stats slot (one atomic per block), or for COL_MAXABS
  // -> this block's slot of the partial maxima (no atomics: a reduction kernel follows)
  if constexpr (FIN || MODE == COL_LOAD || MODE == COL_MAXABS) {
    if (MODE != COL_MAXABS && !d.stats) return;
    for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o, 64));
    if (lane == 0) s_red[wave] = wmax;
    __syncthreads();
    if (tid == 0) {
      float m = s_red[0];
      for (int k = 1; k < C::WPB; ++k) m = fmaxf(m, s_red[k]);
      if constexpr (MODE == COL_MAXABS) {
        d.partial[blockIdx.x] = __float_as_uint(m);
      } else {
        unsigned* slot = FIN ? d.stats + row + 1 : d.stats;
        if (!FIN || row < d.max_rows) atomicMax(slot, __float_as_uint(m));
      }
    }
  }
}

}  // namespace ash
