// General-size kernels: the same sweeps as kernels.h / column.h for grid
// sizes outside the FFT kernels' set (any line length 2..1024, any nT 4..2048),
// e.g. the odd and prime grids of the reference's validation scripts.
//
// The transforms are direct DFTs through LDS, O(n^2) per line, with a
// full-length twiddle table exp(-2 pi i k / n) (FP64-accurate, rounded once)
// indexed by (j k) mod n, so every twiddle is exact to one rounding.  Each
// sweep reads and writes the spectrum once, like its FFT counterpart; these
// kernels trade arithmetic for generality and are not the fast path.
#pragma once
#include "column.h"

namespace ash {

constexpr int GNT = 256;  // threads per general-size line block

// f bins per general-size line tile: the tile (two n x lines complex buffers)
// stays within 32 KB of LDS
__host__ __device__ constexpr int gen_lines(int n) { return n <= 128 ? 16 : n <= 256 ? 8 : n <= 512 ? 4 : 2; }
constexpr int GEN_MAXL = 1024;
constexpr int GEN_TILE = 2048;  // n * lines <= 2048 for every n <= 1024

// out[k][l] = sum_j in[j][l] W^(j k), W = exp(-+2 pi i / n); [pos][line] layout, line fastest.
template <bool INV>
__device__ __forceinline__ void tile_dft(const float2* in, float2* out, const float2* tw, int n, int lines,
                                         int tid) {
  const int tot = n * lines;
  for (int e = tid; e < tot; e += GNT) {
    const int l = e % lines, k = e / lines;
    float ax = 0.f, ay = 0.f;
    int idx = 0;
    for (int j = 0; j < n; ++j) {
      const float2 x = in[j * lines + l];
      float2 w = tw[idx];
      if (INV) w.y = -w.y;
      ax = fmaf(x.x, w.x, fmaf(-x.y, w.y, ax));
      ay = fmaf(x.x, w.y, fmaf(x.y, w.x, ay));
      idx += k;
      if (idx >= n) idx -= n;
    }
    out[e] = make_float2(ax, ay);
  }
}

// x-sweep, forward DFT along x.  Tile = (y, block of `lines` f bins).
__global__ __launch_bounds__(GNT) void k_gx_fwd(float2* __restrict__ spec, Tables T, Geom g) {
  __shared__ float2 tw[GEN_MAXL];
  __shared__ float2 a[GEN_TILE], b[GEN_TILE];
  const int n = g.nX, lines = gen_lines(n), nfb = g.MP / lines;
  const int tid = threadIdx.x;
  const int y = blockIdx.x / nfb, f0 = (blockIdx.x % nfb) * lines;
  const long sx = (long)g.nY * g.MP;
  float2* base = spec + (long)y * g.MP + f0;
  const int nvalid = min(lines, g.M + 1 - f0);
  for (int i = tid; i < n; i += GNT) tw[i] = T.twgx[i];
  for (int e = tid; e < n * lines; e += GNT) {
    const int l = e % lines, j = e / lines;
    a[e] = l < nvalid ? base[j * sx + l] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  tile_dft<false>(a, b, tw, n, lines, tid);
  __syncthreads();
  for (int e = tid; e < n * lines; e += GNT) {
    const int l = e % lines, k = e / lines;
    if (l < nvalid) base[k * sx + l] = b[e];
  }
}

// x-sweep, inverse DFT along x + the epilogues of k_xfft_inv (EPI 0, 1, 2).
template <int EPI>
__global__ __launch_bounds__(GNT) void k_gx_inv(float2* __restrict__ spec, Tables T, Geom g, EpiDesc ep) {
  __shared__ float2 tw[GEN_MAXL];
  __shared__ float2 a[GEN_TILE], b[GEN_TILE];
  const int n = g.nX, lines = gen_lines(n), nfb = g.MP / lines;
  const int tid = threadIdx.x;
  const int y = blockIdx.x / nfb, f0 = (blockIdx.x % nfb) * lines;
  const long sx = (long)g.nY * g.MP;
  float2* base = spec + (long)y * g.MP + f0;
  const int nvalid = min(lines, g.M + 1 - f0);
  for (int i = tid; i < n; i += GNT) tw[i] = T.twgx[i];
  for (int e = tid; e < n * lines; e += GNT) {
    const int l = e % lines, j = e / lines;
    a[e] = l < nvalid ? base[j * sx + l] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  tile_dft<true>(a, b, tw, n, lines, tid);
  __syncthreads();
  if constexpr (EPI == 0 || EPI == 1) {
    float ty = 1.f;
    const float* tx = nullptr;
    if constexpr (EPI == 1) {
      ty = ep.half ? T.tapy_h[y] : T.tapy_f[y];
      tx = ep.half ? T.tapx_h : T.tapx_f;
    }
    for (int e = tid; e < n * lines; e += GNT) {
      const int l = e % lines, k = e / lines;
      float2 v = b[e];
      if constexpr (EPI == 1) v = cscale(v, tx[k] * ty);
      if (l < nvalid) base[k * sx + l] = v;
    }
  } else {
    // scale, DC/Nyquist projection and frequency-weighted mask, then the SA x stencil
    // on the masked values (as k_xfft_inv<., 2>)
    const float lgy = T.lgy[y];
    for (int e = tid; e < n * lines; e += GNT) {
      const int l = e % lines, k = e / lines, m = f0 + l;
      float2 v = cscale(b[e], ep.scale);
      if (m == 0 || m == g.nyq) v.y = 0.f;
      if (ep.fw) v = cscale(v, exp2f(T.fw_exp[min(m, g.MP - 1)] * (T.lgx[k] + lgy)));
      b[e] = v;
    }
    __syncthreads();
    for (int e = tid; e < n * lines; e += GNT) {
      const int l = e % lines, k = e / lines, m = f0 + l;
      float2 v = b[e];
      if (ep.sa) {
        const float w = T.sa_wx[k];
        if (w != 0.f) {
          float2 d;
          if (k == 0) d = cscale(csub(b[lines + l], v), ep.invdx);
          else if (k == n - 1) d = cscale(csub(v, b[(n - 2) * lines + l]), ep.invdx);
          else d = cscale(csub(b[(k + 1) * lines + l], b[(k - 1) * lines + l]), ep.inv2dx);
          const float cw = T.sa_cw[min(m, g.MP - 1)];
          const float2 cd = make_float2(cw * d.y, -cw * d.x);
          const float2 pin = k < n / 2 ? cscale(cadd(v, cd), 0.5f) : cscale(csub(v, cd), 0.5f);
          v = csub(v, cscale(pin, ep.sa_strength * w));
        }
      }
      if (l < nvalid) base[k * sx + l] = v;
    }
  }
}

// y-pass: forward DFT along y, the visit operator, inverse DFT.  Tile = (x, f block).
__global__ __launch_bounds__(GNT) void k_gy_pass(float2* __restrict__ spec, Tables T, Geom g, OpDesc op) {
  __shared__ float2 tw[GEN_MAXL];
  __shared__ float2 a[GEN_TILE], b[GEN_TILE];
  __shared__ LineConst slc[16];
  const int n = g.nY, lines = gen_lines(n), nfb = g.MP / lines;
  const int tid = threadIdx.x;
  const int x = blockIdx.x / nfb, f0 = (blockIdx.x % nfb) * lines;
  float2* base = spec + (long)x * g.nY * g.MP + f0;
  const int nvalid = min(lines, g.M + 1 - f0);
  for (int i = tid; i < n; i += GNT) tw[i] = T.twgy[i];
  if (tid < lines) {
    const int m = min(f0 + tid, g.MP - 1);
    const double kt = T.kt[m], ka = T.ka[m];
    LineConst lc;
    lc.kt2 = kt * kt;
    lc.ka2 = ka * ka;
    lc.ktd = kt;
    lc.kt = (float)kt;
    lc.ka = (float)ka;
    lc.ald = T.alpha[m] * op.dz_att;
    lc.asd = T.astar[m] * op.dz_att;
    lc.al = (float)lc.ald;
    lc.as = (float)lc.asd;
    lc.kept = kt > 0.0;
    slc[tid] = lc;
  }
  for (int e = tid; e < n * lines; e += GNT) {
    const int l = e % lines, j = e / lines;
    a[e] = l < nvalid ? base[(long)j * g.MP + l] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  tile_dft<false>(a, b, tw, n, lines, tid);
  __syncthreads();
  const double kx = T.kx[x];
  for (int e = tid; e < n * lines; e += GNT) {
    const int l = e % lines, k = e / lines;
    const double ky = T.ky[k];
    const double kk = kperp2(kx, ky);
    const float2 o = op.precise ? visit_operator_fp64(op, kk, slc[l], f0 + l) : visit_operator(op, kk, slc[l], f0 + l);
    b[e] = cmul(b[e], o);
  }
  __syncthreads();
  tile_dft<true>(b, a, tw, n, lines, tid);
  __syncthreads();
  for (int e = tid; e < n * lines; e += GNT) {
    const int l = e % lines, k = e / lines;
    if (l < nvalid) base[(long)k * g.MP + l] = a[e];
  }
}

// Power-weighted obliquity sums of a (kx row, f block) tile (as k_ypower): forward
// DFT along y, then acc[0] += sum P k/k_z, acc[1] += sum P over propagating modes.
__global__ __launch_bounds__(GNT) void k_gy_power(const float2* __restrict__ spec, Tables T, Geom g,
                                                  double* __restrict__ acc) {
  __shared__ float2 tw[GEN_MAXL];
  __shared__ float2 a[GEN_TILE], b[GEN_TILE];
  __shared__ double sred[2][GNT / 64];
  const int n = g.nY, lines = gen_lines(n), nfb = g.MP / lines;
  const int tid = threadIdx.x;
  const int x = blockIdx.x / nfb, f0 = (blockIdx.x % nfb) * lines;
  const float2* base = spec + (long)x * g.nY * g.MP + f0;
  const int nvalid = min(lines, g.M + 1 - f0);
  for (int i = tid; i < n; i += GNT) tw[i] = T.twgy[i];
  for (int e = tid; e < n * lines; e += GNT) {
    const int l = e % lines, j = e / lines;
    a[e] = l < nvalid ? base[(long)j * g.MP + l] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  tile_dft<false>(a, b, tw, n, lines, tid);
  __syncthreads();
  const double kx = T.kx[x];
  double num = 0.0, den = 0.0;
  for (int e = tid; e < n * lines; e += GNT) {
    const int l = e % lines, k = e / lines, m = f0 + l;
    if (l >= nvalid || m == 0) continue;
    const double km = T.ka[m];
    const double ky = T.ky[k];
    const double kz2 = kz2_of(km, kperp2(kx, ky));
    if (kz2 > 0.0) {
      const float2 v = b[e];
      const double P = (double)v.x * v.x + (double)v.y * v.y;
      num += P * km / fmax(sqrt(kz2), km * 1e-6);
      den += P;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    num += __shfl_xor(num, o, 64);
    den += __shfl_xor(den, o, 64);
  }
  if ((tid & 63) == 0) { sred[0][tid >> 6] = num; sred[1][tid >> 6] = den; }
  __syncthreads();
  if (tid == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int w = 0; w < GNT / 64; ++w) { s0 += sred[0][w]; s1 += sred[1][w]; }
    atomicAdd(acc, s0);
    atomicAdd(acc + 1, s1);
  }
}

// y-sweep without a transform (as k_ysa) for any nY: super-absorbing y pass,
// DC/Nyquist projection and an optional phase/amplitude screen.  Tile = (x, f block
// of gen_lines(nY) bins); the rows handled (all with a screen, else the two boundary
// bands) sit in LDS.
__global__ __launch_bounds__(GNT) void k_gy_sa(float2* __restrict__ spec, Tables T, Geom g, ScreenDesc s) {
  __shared__ float2 buf[GEN_TILE];
  const int L = g.nY, lines = gen_lines(L), nfb = g.MP / lines;
  const int tid = threadIdx.x;
  const int x = blockIdx.x / nfb, f0 = (blockIdx.x % nfb) * lines;
  float2* base = spec + (long)x * g.nY * g.MP + f0;
  const int nvalid = min(lines, g.M + 1 - f0);
  const int nrows = ysa_rows(L, s.band, s.phase != nullptr);
  const bool band = nrows < L;
  auto row_of = [&](int i) { return (band && i > s.band) ? L - nrows + i : i; };
  const int nel = nrows * lines;
  for (int el = tid; el < nel; el += GNT) {
    const int l = el % lines, i = el / lines;
    buf[el] = l < nvalid ? base[(long)row_of(i) * g.MP + l] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  for (int el = tid; el < nel; el += GNT) {
    const int l = el % lines, i = el / lines;
    const int pos = row_of(i), m = f0 + l;
    float2 u = buf[el];
    if (s.sa) {
      const float w = T.sa_wy[pos];
      if (w != 0.f) {
        float2 d;
        if (pos == 0) d = cscale(csub(buf[(i + 1) * lines + l], u), s.invdy);
        else if (pos == L - 1) d = cscale(csub(u, buf[(i - 1) * lines + l]), s.invdy);
        else d = cscale(csub(buf[(i + 1) * lines + l], buf[(i - 1) * lines + l]), s.inv2dy);
        const float cw = T.sa_cw[min(m, g.MP - 1)];
        const float2 cd = make_float2(cw * d.y, -cw * d.x);
        const float2 pin = pos < L / 2 ? cscale(cadd(u, cd), 0.5f) : cscale(csub(u, cd), 0.5f);
        u = csub(u, cscale(pin, s.sa_strength * w));
      }
    }
    if (m == 0 || m == g.nyq) u.y = 0.f;
    if (s.phase) {
      const long idx = (long)x * g.nY + pos;
      const float fs = (float)m / s.f0b;
      const float th = s.phase[idx] * fs;
      float sn, cs;
      sincosf(th, &sn, &cs);
      u = cmul(u, make_float2(cs, -sn));
      if (s.amp) {
        const float amp = fminf(fmaxf(s.amp[idx], 1e-6f), 1.f);
        u = cscale(u, powf(amp, fs));
      }
      if (m == 0 || m == g.nyq) u.y = 0.f;
    }
    if (l < nvalid) base[(long)pos * g.MP + l] = u;
  }
}

// ---------------------------------------------------------------------------
// General-length t-column kernel: one wave per column, lane ln holds the S
// contiguous samples t = S ln .. S ln + S - 1 (padding beyond nT is zero).
// irfft and rfft are direct real DFTs through the wave's LDS region; the
// real-space stage is column_realspace, as in k_column.  Semantics of the
// transforms (ASS:664, :685): the inverse ignores the imaginary parts of the DC
// bin and, for even nT, of the Nyquist bin; the forward writes them as zero.
// ---------------------------------------------------------------------------
constexpr int GC_WPB = 4;

template <int S, int MODE, int NLK>
__global__ __launch_bounds__(256) void k_gcolumn(float2* __restrict__ spec, Tables T, Geom g, ColDesc d) {
  constexpr int NTMAX = 64 * S;
  __shared__ float2 s_tw[NTMAX];
  __shared__ float2 s_buf[GC_WPB][NTMAX / 2 + 8];
  __shared__ float s_red[GC_WPB];
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64, ln = lane;
  const int nT = g.nT, M = g.M;  // bins 0..M, M = nT / 2
  constexpr bool spectral_in = !(MODE == COL_LOAD || MODE == COL_RAWNL);
  constexpr bool rewrite = (MODE == COL_NL || MODE == COL_TAPER || MODE == COL_LOAD || MODE == COL_FINAL_SRC);
  if constexpr (spectral_in || rewrite) {
    for (int i = tid; i < nT; i += 256) s_tw[i] = T.twgt[i];
  }
  __syncthreads();
  float2* cb = s_buf[wave];
  float* rb = reinterpret_cast<float*>(cb);
  const int ncols = g.nX * g.nY;
  const int t0 = ln * S;
  const float invN = 1.f / (float)nT;
  const LaneCol<S, 64, false, true> lc{ln, lane, t0, nT};
  constexpr bool FIN = (MODE == COL_FINAL || MODE == COL_FINAL_SRC);
  const int row = (FIN || MODE == COL_TOF) ? *d.row : 0;
  const double Nn = d.nscale ? d.N * (d.nscale[0] / (d.nscale[1] + 1e-30)) : d.N;
  float wmax = 0.f;
  int nsub_global = 1;
  if constexpr (NLK != NL_NONE) {
    if (d.subcycle == SUBCYCLE_GLOBAL) {
      const float gm = __uint_as_float(*d.gmax);
      bool capped;
      nsub_global = substeps_for(Nn * (double)gm * d.dz / d.dT, (double)d.cfl_target, gm, &capped);
      if (capped && blockIdx.x == 0 && tid == 0 && d.subcap) atomicAdd(d.subcap, 1u);
    }
  }
  const bool even = (nT & 1) == 0;
  const int mh = (nT - 1) / 2;  // last bin with a conjugate partner

  const float2* sc_cur = nullptr;  // spectrum row of the current column
  // inverse real DFT of the current column's bins (modified by tof_bin(sel)) into
  // this lane's samples: u[t] = (X_0 + 2 sum_{m=1}^{mh} Re(X_m e^{2 pi i m t / nT})
  //                             + [even nT] X_{nT/2} (-1)^t) / nT
  auto inv_spec = [&](float (&uo)[S], int sel) {
    for (int m = ln; m <= M; m += 64) {
      float2 x = sc_cur[m];
      if (m == 0 || (even && m == M)) x.y = 0.f;
      cb[m] = tof_bin(sel, m, x, even ? M : -1, d.tofref);
    }
    wave_sync();
    const float x0 = cb[0].x;
    const float xn = even ? cb[M].x : 0.f;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int t = t0 + i;
      float acc = 0.f;
      if (t < nT) {
        int idx = 0;
        for (int m = 1; m <= mh; ++m) {
          idx += t;
          if (idx >= nT) idx -= nT;
          const float2 x = cb[m], w = s_tw[idx];  // e^{+i} = conj(w)
          acc = fmaf(x.x, w.x, fmaf(x.y, w.y, acc));
        }
        acc = fmaf(2.f, acc, x0 + ((t & 1) ? -xn : xn));
      }
      uo[i] = t < nT ? acc * invN : 0.f;
    }
    wave_sync();
  };

  for (int cbase = blockIdx.x * GC_WPB; cbase < ncols; cbase += gridDim.x * GC_WPB) {
    const int c = cbase + wave;
    const bool live = c < ncols && (!d.cflag || d.cflag[c] != 0);
    if (!live) continue;
    float2* sc = spec + (long)c * g.MP;
    float u[S];
    sc_cur = sc;
    float uq[MODE == COL_TOF ? S : 1];
    if constexpr (!spectral_in) {
      const float* src = d.rin + (long)c * nT;
#pragma unroll
      for (int i = 0; i < S; ++i) u[i] = (t0 + i < nT) ? src[t0 + i] : 0.f;
    } else if constexpr (MODE == COL_TOF) {
      if (d.tof_mode == TOF_ENVELOPE) {
        inv_spec(u, 0);
        inv_spec(uq, 1);
      } else {
        inv_spec(u, 2);
      }
    } else {
      inv_spec(u, 0);
    }

    if constexpr (MODE == COL_TOF) {
      const bool nz = spectrum_nonzero<64>(sc, M, ln);
      column_tof<S, 64, false, true>(u, uq, lc, rb, d, c, live, ncols, row, nz);
    } else {
      column_realspace<S, 64, false, true, MODE, NLK>(u, lc, d, T, c, live, ncols, row, Nn, nsub_global, wmax);
    }

    if constexpr (rewrite) {
#pragma unroll
      for (int i = 0; i < S; ++i)
        if (t0 + i < nT) rb[t0 + i] = u[i];
      wave_sync();
      // X_m = sum_t u[t] e^{-2 pi i m t / nT}, m = ln, ln + 64, ...
      float2 res[(NTMAX / 2 + 64) / 64];
#pragma unroll
      for (int q = 0; q < (NTMAX / 2 + 64) / 64; ++q) {
        const int m = ln + 64 * q;
        float ax = 0.f, ay = 0.f;
        if (m <= M) {
          int idx = 0;
          for (int t = 0; t < nT; ++t) {
            const float v = rb[t];
            const float2 w = s_tw[idx];
            ax = fmaf(v, w.x, ax);
            ay = fmaf(v, w.y, ay);
            idx += m;
            if (idx >= nT) idx -= nT;
          }
          if (m == 0 || (even && m == M)) ay = 0.f;
        }
        res[q] = make_float2(ax, ay);
      }
      wave_sync();
#pragma unroll
      for (int q = 0; q < (NTMAX / 2 + 64) / 64; ++q) {
        const int m = ln + 64 * q;
        if (m <= M) sc[m] = res[q];
      }
    }
  }

  if constexpr (FIN || MODE == COL_LOAD || MODE == COL_MAXABS) {
    if (MODE != COL_MAXABS && !d.stats) return;
    for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o, 64));
    if (lane == 0) s_red[wave] = wmax;
    __syncthreads();
    if (tid == 0) {
      float m = s_red[0];
      for (int k = 1; k < GC_WPB; ++k) m = fmaxf(m, s_red[k]);
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
