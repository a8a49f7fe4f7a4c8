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

// Most KT sub-steps one column takes in one dZ with per-column CFL sub-cycling
constexpr int MAX_SUBSTEPS = 1024;

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

template <int M, int MODE, int NLK>
__global__ __launch_bounds__(256) void k_column(float2* __restrict__ spec, Tables T, Geom g, ColDesc d) {
  using C = ColCfg<M>;
  constexpr int E = C::E, LPC = C::LPC, CPW = C::CPW, S = C::S, NT_ = 2 * M;
  constexpr int TWF = CDir<M, false>::total();
  constexpr int TWT = col_tw_total<M>();
  __shared__ float2 s_tw[TWT];
  __shared__ float2 s_twn[M + 1];
  __shared__ float2 s_buf[C::WPB * CPW * C::LDSN];
  __shared__ float s_red[C::WPB];
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int cw = lane / LPC, ln = lane % LPC;
  for (int i = tid; i < TWT; i += 256) s_tw[i] = T.twc[i];
  for (int i = tid; i <= M; i += 256) s_twn[i] = T.twn[i];
  __syncthreads();
  const float2* twf = s_tw;
  const float2* twi = s_tw + TWF;
  float2* cb = s_buf + (wave * CPW + cw) * C::LDSN;
  const int ncols = g.nX * g.nY;
  const int per_iter = C::WPB * CPW;
  const int t0 = ln * S;
  const float invN = 1.f / (float)NT_;
  constexpr bool spectral_in = !(MODE == COL_LOAD || MODE == COL_RAWNL);
  // pointwise modes keep the FFT's strided ownership (lane ln holds samples
  // 2(ln + LPC e) and 2(ln + LPC e) + 1); only the t-stencil modes transpose to
  // contiguous samples
  constexpr bool STRIDED = !(MODE == COL_NL || MODE == COL_RAWNL);
  auto tsamp = [&](int i) -> int {
    if constexpr (STRIDED) return 2 * (ln + LPC * (i >> 1)) + (i & 1);
    else return t0 + i;
  };
  // store this lane's samples u[0..S) of a real column starting at dst
  auto store_real = [&](float* dst, const float* u) {
    if constexpr (STRIDED) {
      float2* d2 = reinterpret_cast<float2*>(dst);
#pragma unroll
      for (int e = 0; e < E; ++e) d2[ln + LPC * e] = make_float2(u[2 * e], u[2 * e + 1]);
    } else {
      float4* d4 = reinterpret_cast<float4*>(dst + t0);
#pragma unroll
      for (int q = 0; q < S / 4; ++q) d4[q] = make_float4(u[4 * q], u[4 * q + 1], u[4 * q + 2], u[4 * q + 3]);
    }
  };
  constexpr bool FIN = (MODE == COL_FINAL || MODE == COL_FINAL_SRC);
  constexpr bool rewrite = (MODE == COL_NL || MODE == COL_TAPER || MODE == COL_LOAD || MODE == COL_FINAL_SRC);
  const int row = FIN ? *d.row : 0;
  const float* tt = d.taper == 1 ? T.tapt_h : T.tapt_f;
  // Burgers coefficient, scaled by the beam obliquity when requested (ASS:934)
  const double Nn = d.nscale ? d.N * (d.nscale[0] / (d.nscale[1] + 1e-30)) : d.N;
  float wmax = 0.f;  // running max|p| of this lane's columns (stats)

  for (int cbase = blockIdx.x * per_iter; cbase < ncols; cbase += gridDim.x * per_iter) {
    const int c = cbase + wave * CPW + cw;
    const bool live = c < ncols && (!d.cflag || d.cflag[c] != 0);
    if (CPW == 1 && !live) continue;  // one column per wave: skip it whole
    float2* sc = spec + (long)c * g.MP;
    float u[S];

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
    } else {
      float2 v[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int k = ln + LPC * e;
        float2 x = live ? sc[k] : make_float2(0.f, 0.f);
        if (k == 0) x.y = 0.f;
        v[e] = x;
        cb[pidx(k)] = x;
      }
      if (ln == 0) {
        float2 xn = live ? sc[M] : make_float2(0.f, 0.f);
        xn.y = 0.f;
        cb[pidx(M)] = xn;
      }
      wave_sync();
      // Z[k] = (X_k + conj X_{M-k}) + i W_N^{-k} (X_k - conj X_{M-k}), scaled by 1/N
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int k = ln + LPC * e;
        const float2 a = v[e], b = cconj(cb[pidx(M - k)]);
        const float2 sm = cadd(a, b), df = csub(a, b);
        const float2 t = cmul(cconj(s_twn[k]), df);
        v[e] = make_float2((sm.x - t.y) * invN, (sm.y + t.x) * invN);
      }
      wave_sync();
      cfft<M, true>(v, cb, twi, ln);
      if constexpr (STRIDED) {
#pragma unroll
        for (int e = 0; e < E; ++e) { u[2 * e] = v[e].x; u[2 * e + 1] = v[e].y; }
      } else {
        // transpose to contiguous ownership: lane ln gets z[E ln .. E ln + E)
#pragma unroll
        for (int e = 0; e < E; ++e) cb[pidx(ln + LPC * e)] = v[e];
        wave_sync();
#pragma unroll
        for (int q = 0; q < E; ++q) {
          const float2 z = cb[pidx(E * ln + q)];
          u[2 * q] = z.x;
          u[2 * q + 1] = z.y;
        }
        wave_sync();
      }
    }

    // ---------------- stage 2: real-space work ----------------
    if constexpr (MODE == COL_NL || MODE == COL_TAPER || MODE == COL_RAWNL) {
      if (d.taper) {
#pragma unroll
        for (int i = 0; i < S; ++i) u[i] *= tt[tsamp(i)];
      }
      if constexpr (NLK != NL_NONE && MODE != COL_TAPER) {
        int nsub = 1;
        if (d.subcycle) {
          float mx = 0.f;
#pragma unroll
          for (int i = 0; i < S; ++i) mx = fmaxf(mx, fabsf(u[i]));
          for (int o = LPC / 2; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
          const double cfl = Nn * (double)mx * d.dz / d.dT;
          // ceil(cfl / target) sub-steps above the target, bounded so that a blown-up
          // column (inf) cannot keep the wave looping; NaN takes one step
          const double q = cfl / (double)d.cfl_target;
          if (q > 1.0 && mx >= 1e-30f) nsub = q < (double)MAX_SUBSTEPS ? (int)ceil(q) : MAX_SUBSTEPS;
        }
        int nsub_w = nsub;
        for (int o = 32; o >= LPC; o >>= 1) nsub_w = max(nsub_w, __shfl_xor(nsub_w, o, 64));
        const float dz = (float)(d.dz / nsub);
        const float coef = (float)(Nn / d.dT);
        const int lsrc = lane - 1, rsrc = lane + 1;
        const bool has_l = ln > 0, has_r = ln < LPC - 1;
        for (int sub = 0; sub < nsub_w; ++sub) {
          const bool act = sub < nsub;
          if constexpr (NLK == NL_KT || NLK == NL_KT_MINMOD) {
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
            kt_rhs_local<S, NLK == NL_KT>(h, t0, NT_, coef, r0);
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
            kt_rhs_local<S, NLK == NL_KT>(h, t0, NT_, coef, r0);
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
              const int t = t0 + i;
              if (act && t >= 1 && t <= NT_ - 2) {
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
      if (MODE == COL_RAWNL && live) store_real(d.rout + (long)c * NT_, u);
    } else if constexpr (FIN) {
      if constexpr (MODE == COL_FINAL_SRC) {
        for (int k = 0; k < d.n_src; ++k) {
          const float* sp = d.src[k] + (long)c * NT_;
#pragma unroll
          for (int i = 0; i < S; ++i) u[i] += live ? sp[tsamp(i)] : 0.f;
        }
        if (d.smask && live && d.smask[c]) {
          const double dl = d.sdelay[c];
          const double w0 = 2.0 * 3.14159265358979323846 * d.s_f0;
          const double cexp = 2.0 * d.s_dur;
#pragma unroll
          for (int i = 0; i < S; ++i) u[i] += tone_burst(d.taxis[tsamp(i)] - dl, w0, d.s_ncyc, cexp, d.s_p0);
        }
      }
      float mn = u[0], mx = u[0];
      double e2 = 0.0;
#pragma unroll
      for (int i = 0; i < S; ++i) {
        mn = fminf(mn, u[i]);
        mx = fmaxf(mx, u[i]);
        e2 += (double)u[i] * (double)u[i];
      }
      for (int o = LPC / 2; o > 0; o >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, o, 64));
        mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        e2 += __shfl_xor(e2, o, 64);
      }
      if (live && row < d.max_rows) {
        if (ln == 0) {
          const long o = (long)row * ncols + c;
          const double ib = d.ibefore[c];
          d.pnp[o] = mn;
          d.ppp[o] = mx;
          d.pI[o] = (float)e2;
          d.pIloss[o] = (float)fmax(0.0, ib - e2);
          d.ibefore[c] = e2;
        }
        if (c == d.cax) store_real(d.pax + (long)row * NT_, u);
      }
      if (live) wmax = fmaxf(wmax, fmaxf(mx, -mn));
      if (d.rout && live) store_real(d.rout + (long)c * NT_, u);
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
      if (live && ln == 0 && d.ibefore) d.ibefore[c] = e2;
      if (live) wmax = fmaxf(wmax, mx);
    } else if constexpr (MODE == COL_STORE) {
      if (live) store_real(d.rout + (long)c * NT_, u);
    } else if constexpr (MODE == COL_TOREAL) {
      if (d.taper) {
#pragma unroll
        for (int i = 0; i < S; ++i) u[i] *= tt[tsamp(i)];
      }
      if (live) store_real(d.rout + (long)c * NT_, u);
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
        const float2 sm = cadd(zk, zm), df = csub(zk, zm);
        const float2 t = cmul(s_twn[k], df);
        if (live) sc[k] = make_float2(0.5f * (sm.x + t.y), 0.5f * (sm.y - t.x));
      }
      if (ln == 0 && live) sc[M] = make_float2(z0.x - z0.y, 0.f);
      wave_sync();
    }
  }

  // max|p| of the block -> stats slot (one atomic per block)
  if constexpr (FIN || MODE == COL_LOAD) {
    if (!d.stats) return;
    for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o, 64));
    if (lane == 0) s_red[wave] = wmax;
    __syncthreads();
    if (tid == 0) {
      float m = s_red[0];
      for (int k = 1; k < C::WPB; ++k) m = fmaxf(m, s_red[k]);
      unsigned* slot = FIN ? d.stats + row + 1 : d.stats;
      if (!FIN || row < d.max_rows) atomicMax(slot, __float_as_uint(m));
    }
  }
}

}  // namespace ash
