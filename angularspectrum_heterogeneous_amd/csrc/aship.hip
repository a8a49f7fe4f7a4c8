// Host side of the aship C ABI: plans, FP64 operator tables, launch sequences
// and the per-step hipGraph.  Kernels live in kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/aship.h"
#include "kernels.h"
#include "generic.h"
#include "learned.h"

using namespace ash;

// Scratch of ash_nonlinear_step_noalloc, the entry point that allocates nothing:
// [0, NOALLOC_PARTIALS) per-block max|p| partials (one per 64 columns), then the
// global maximum and the sub-cycling cap counter.  In host code `scratch` names
// a host-side stand-in; kernels get its device address from hipGetSymbolAddress
// (one address per device, looked up on the first call on that device).
constexpr int NOALLOC_PARTIALS = 16384;  // columns: at most 64 * 16384 = 1,048,576
constexpr int NOALLOC_GMAX = NOALLOC_PARTIALS;
constexpr int NOALLOC_SUBCAP = NOALLOC_PARTIALS + 1;
static __device__ __align__(256) unsigned scratch[NOALLOC_PARTIALS + 64];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

void set_error_nothrow(const char* msg) noexcept {
  try {
    g_err = msg;
  } catch (...) {
    g_err.clear();
  }
}

// Every int-returning entry point is a function-try-block ending in this handler:
// no C++ exception (std::bad_alloc from a table or a graph, a throwing library call)
// may unwind through the C ABI into a foreign caller.
int caught_exception() noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    set_error_nothrow("host memory allocation failed");
  } catch (const std::exception& e) {
    set_error_nothrow(e.what());
  } catch (...) {
    set_error_nothrow("unknown C++ exception");
  }
  return ASH_E_HOST;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fail(ASH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Sizes of the FFT kernels; every other size runs the general-size kernels (generic.h).
bool line_supported(int n) { return n == 64 || n == 128 || n == 192 || n == 256 || n == 384; }
bool col_supported(int m) { return m == 128 || m == 256 || m == 512 || m == 1024; }
constexpr int GEN_MAX_LINE = GEN_MAXL;  // general-size line kernels: 2 .. 1024
constexpr int GEN_MAX_NT = 2048;         // general-length column kernel: 4 .. 2048

// Python's round(): half to even
long py_round(double v) {
  double r = std::nearbyint(v);  // default rounding mode is to-nearest-even
  return (long)r;
}

// Tapers of ASS:82-110 in FP64.
std::vector<double> taper(int N, int n, int profile) {
  std::vector<double> v(N, 1.0);
  for (int i = 0; i < N; ++i) {
    double s = -1.0;
    if (i < n) s = (double)(n - i - 1) / n;
    if (i >= N - n) s = (double)(i - (N - n - 1)) / n;
    if (s < 0.0) continue;
    if (profile == ASH_PROFILE_WENDLAND) {
      double w = 1.0 - 10 * s * s * s + 15 * s * s * s * s - 6 * s * s * s * s * s;
      v[i] = std::fmin(std::fmax(w, 0.0), 1.0);
    } else {
      v[i] = 1.0 - s * s;
    }
  }
  return v;
}

// Sum in NumPy's order (its pairwise summation of a contiguous float64 array):
// below 8 terms a plain loop; up to 128 eight interleaved partial sums combined as
// ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then the remainder; above, the halves
// (split at a multiple of 8) recursively.  The reference builds its wavenumber
// grids with np.mean (ASS:404-408); a sum in another order can move a grid value
// by an ulp, and on grids where some mode sits exactly at the evanescent cutoff
// (k_z = 0, where the obliquity factor of the attenuation jumps from 0 to 1) that
// ulp decides the mode.
double numpy_sum(const double* a, long n) {
  if (n < 8) {
    double r = 0.0;
    for (long i = 0; i < n; ++i) r += a[i];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    long i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  long n2 = n / 2;
  n2 -= n2 % 8;
  return numpy_sum(a, n2) + numpy_sum(a + n2, n - n2);
}

// (n-1)-normalised mean-subtracted grid (ASS:403-408), natural FFT order.
std::vector<double> centred_natural(int n, double d, double extra_div) {
  std::vector<double> k(n);
  for (int j = 0; j < n; ++j) k[j] = (double)j / (n - 1) / d * 2 * M_PI / extra_div;
  const double mean = numpy_sum(k.data(), n) / n;
  for (int j = 0; j < n; ++j) k[j] -= mean;
  std::vector<double> nat(n);
  for (int i = 0; i < n; ++i) nat[i] = k[(i + n / 2) % n];
  return nat;
}

}  // namespace

struct ash_plan {
  ash_desc d;
  Geom g;
  double N = 0.0;     // beta / (2 c0^3 rho0)
  double dz = 0.0;
  double lambda_char = 0.0;
  int kt_nonzero = 0;
  float2* spec = nullptr;
  double* ibefore = nullptr;
  int* d_row = nullptr;
  double* obl_acc = nullptr;  // [2] power-weighted obliquity sums of the current step
  unsigned* red = nullptr;    // [RED_WORDS] max|p| partials, global max, sub-cycling cap counter
  Tables T{};
  std::vector<void*> allocs;
  int64_t ws_bytes = 0;
  bool loaded = false;
  int host_row = -1;
  bool gen_x = false, gen_y = false, gen_t = false;  // axis runs the general-size kernels
  int sa_band = 0;      // last row index (from each edge) with a nonzero super-absorbing y weight
  int precise_ops = 0;  // evaluate the k-space operators in FP64 (set from ASH_PRECISE_OPS)
  // graph of a plain step
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipStream_t cap_stream = nullptr;
  ash_step_io gkey{};
  int epoch = 0;  // bumped when the step sequence changes (captured graphs are stale)
  // launch configuration (ash_plan_autotune)
  int rvx = 0, rvy = 0;  // radix order of the x- and y-line transforms
  int col_grid = 2048;   // blocks of the persistent column kernels
  int fin_grid = 2048;   // blocks of the final column pass: as many as are resident at once
  // per-kernel timing (ash_profile_step): events around each launch
  bool prof_on = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_ev;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_next = 0;
  // learned nonlinear step (ash_plan_set_learned)
  ash_net* net = nullptr;
  double dpdt_max = 0.0;     // KT where max|dp/dt| of a column exceeds this (Pa/s)
  float* rscratch = nullptr;  // real field between the column passes [nX*nY*nT]
  int32_t* nflag = nullptr;   // 1 where a column took the KT update [nX*nY]
  // time of flight (ash_plan_set_tof)
  int tof_mode = 0;
  double tof_ratio = 1.0, tof_t0 = 0.0, tof_dt = 0.0, tof_tref = 0.0;
  float2* tofref = nullptr;   // rfft of the matched-filter reference trace [MP]
};

struct ash_net {
  int nT = 0, max_cols = 0, n_cu = 0;
  float* params = nullptr;  // float32 master parameters
  float* m = nullptr;       // Adam moments
  float* v = nullptr;
  float* mask = nullptr;    // 1 on trainable entries
  uint16_t* Wm = nullptr;   // bf16 [layer][k][co][ci]
  uint16_t* Wt = nullptr;   // bf16 [layer][k][ci][co], flipped in k
  uint16_t* X[4] = {};      // training activations [max_cols][nT][32]
  uint16_t* Dg[2] = {};     // back-propagated gradients, same shape
  float* y = nullptr;       // [max_cols][nT]
  float* amp = nullptr;     // [max_cols]
  std::vector<void*> allocs;
};

struct ash_batch {
  std::vector<ash_plan*> plans;    // set of the captured graph
  std::vector<ash_step_io> keys;
  std::vector<int> epochs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipStream_t cap = nullptr;
  // concurrent branches of the captured step: instance i runs on branch[lane[i]]
  int nbranch = 4;
  std::vector<hipStream_t> branch;
  std::vector<hipEvent_t> fork_join;  // [0] fork, [1 + k] join of branch k
};

namespace {

// Layout of ash_plan::red (and of the no-allocation scratch, same kernels):
// per-block max|p| partials, then the global maximum, then the count of columns
// (or fields) whose sub-cycling hit MAX_SUBSTEPS.
constexpr int RED_PARTIALS = 65536;  // >= any column grid (<= 65535) and nX nY / 64 (<= 2304)
constexpr int RED_GMAX = RED_PARTIALS;
constexpr int RED_SUBCAP = RED_PARTIALS + 1;
constexpr int RED_WORDS = RED_PARTIALS + 64;

template <class T>
int upload(ash_plan* p, const std::vector<T>& h, const T** out) {
  void* d = nullptr;
  p->allocs.reserve(p->allocs.size() + 1);  // the push_back below cannot throw
  HIP_TRY(hipMalloc(&d, h.size() * sizeof(T) + 16));
  p->allocs.push_back(d);
  HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  p->ws_bytes += (int64_t)(h.size() * sizeof(T));
  *out = (const T*)d;
  return ASH_OK;
}

std::vector<float2> twiddles(int L) {
  std::vector<float2> t(L);
  for (int k = 0; k < L; ++k) {
    double a = -2.0 * M_PI * (double)k / (double)L;
    t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    // exact values at the multiples of a quarter turn
    if ((4L * k) % L == 0) {
      const int q = (int)((4L * k) / L);
      const float c[4] = {1.f, 0.f, -1.f, 0.f}, sn[4] = {0.f, -1.f, 0.f, 1.f};
      t[k] = make_float2(c[q], sn[q]);
    }
  }
  return t;
}

int build_tables(ash_plan* p) {
  const ash_desc& d = p->d;
  const int nX = d.nX, nY = d.nY, nT = d.nT, M = p->g.M, MP = p->g.MP;
  int rc;
  std::vector<double> kx = centred_natural(nX, d.dX, 1.0), ky = centred_natural(nY, d.dY, 1.0);
  if ((rc = upload(p, kx, &p->T.kx))) return rc;
  if ((rc = upload(p, ky, &p->T.ky))) return rc;
  // ASM t-wavenumber of the kept (positive) bins: fftshifted index (m + nT/2) % nT
  std::vector<double> ktc(nT);
  {
    for (int j = 0; j < nT; ++j) ktc[j] = (double)j / (nT - 1) / d.dT * 2 * M_PI / d.c0;
    const double mean = numpy_sum(ktc.data(), nT) / nT;
    std::vector<double> nzabs;
    for (int j = 0; j < nT; ++j) {
      ktc[j] -= mean;
      if (ktc[j] != 0.0) nzabs.push_back(std::fabs(ktc[j]));
    }
    const int cnt = (int)nzabs.size();
    p->kt_nonzero = cnt;
    // lambda_char = c0 / mean|kt| over the nonzero bins (ASS:415)
    p->lambda_char = cnt ? d.c0 / (numpy_sum(nzabs.data(), cnt) / cnt) : d.c0 / (1.0 / d.dT);
  }
  std::vector<double> kt(MP, 0.0), ka(MP, 0.0), al(MP, 0.0), as(MP, 0.0);
  const int kept_hi = (nT + 1) / 2 - 1;  // natural bins 1..ceil(nT/2)-1 are kept
  for (int m = 1; m <= kept_hi && m < MP; ++m) kt[m] = ktc[(m + nT / 2) % nT];
  // attenuation (ASS:523-592)
  const double cl = 20.0 * std::log10(M_E);
  std::vector<double> f(nT);
  for (int m = 0; m < nT; ++m) f[m] = (double)m / (nT * d.dT);
  if (d.alpha0 < 0) {
    const double conv = 2.17e-3 / 1e12 * 1e2 / cl;
    for (int m = 0; m <= M; ++m) {
      al[m] = conv * f[m] * f[m];
      const double a0 = (conv / std::pow(2 * M_PI, 2)) * std::tan(M_PI) * ((2 * M_PI * f[m]) - (2 * M_PI * d.f0));
      as[m] = m == 0 ? 0.0 : 2 * M_PI * a0 * f[m];
    }
  } else {
    const double pw = d.attenPow;
    const double conv = d.alpha0 / std::pow(1e6, pw) * 1e2 / cl;
    const bool logb = std::fmod(pw, 2.0) == 1.0;
    for (int m = 0; m <= M; ++m) {
      al[m] = conv * std::pow(f[m], pw);
      double a0;
      if (logb) {
        const double fs = m == 0 ? (nT > 1 ? f[1] / 2 : 0.0) : f[m];
        a0 = (-2 * conv / std::pow(2 * M_PI, pw) / M_PI) * (std::log(2 * M_PI * fs) - std::log(2 * M_PI * d.f0));
      } else {
        a0 = (conv / std::pow(2 * M_PI, pw)) * std::tan(M_PI * pw / 2) *
             (std::pow(2 * M_PI * f[m], pw - 1) - std::pow(2 * M_PI * d.f0, pw - 1));
      }
      as[m] = m == 0 ? 0.0 : 2 * M_PI * a0 * f[m];
    }
  }
  for (int m = 0; m <= M; ++m) ka[m] = 2 * M_PI * f[m] / d.c0;
  if ((rc = upload(p, kt, &p->T.kt))) return rc;
  if ((rc = upload(p, ka, &p->T.ka))) return rc;
  if ((rc = upload(p, al, &p->T.alpha))) return rc;
  if ((rc = upload(p, as, &p->T.astar))) return rc;
  // boundary tapers (ASS:116-131)
  const long bx = std::max(py_round(nX * d.boundaryFactor), 1L);
  const long by = std::max(py_round(nY * d.boundaryFactor), 1L);
  const long bt = std::max(py_round(nT * d.boundaryFactor), 1L);
  std::vector<double> vx(nX, 1.0), vy(nY, 1.0), vt(nT, 1.0);
  if (d.useBoundaryLayer) {
    vx = taper(nX, (int)bx, d.boundaryProfile);
    vy = taper(nY, (int)by, d.boundaryProfile);
    vt = taper(nT, (int)bt, d.boundaryProfile);
  }
  auto tof = [](const std::vector<double>& v, bool sq) {
    std::vector<float> o(v.size());
    for (size_t i = 0; i < v.size(); ++i) o[i] = (float)(sq ? std::sqrt(v[i]) : v[i]);
    return o;
  };
  if ((rc = upload(p, tof(vx, true), &p->T.tapx_h))) return rc;
  if ((rc = upload(p, tof(vy, true), &p->T.tapy_h))) return rc;
  if ((rc = upload(p, tof(vt, true), &p->T.tapt_h))) return rc;
  if ((rc = upload(p, tof(vx, false), &p->T.tapx_f))) return rc;
  if ((rc = upload(p, tof(vy, false), &p->T.tapy_f))) return rc;
  if ((rc = upload(p, tof(vt, false), &p->T.tapt_f))) return rc;
  // frequency-weighted mask (ASS:144-179): always from the profile tapers
  {
    std::vector<double> wx = taper(nX, (int)bx, d.boundaryProfile), wy = taper(nY, (int)by, d.boundaryProfile);
    std::vector<float> lx(nX), ly(nY);
    for (int i = 0; i < nX; ++i) lx[i] = (float)std::log2(wx[i]);
    for (int i = 0; i < nY; ++i) ly[i] = (float)std::log2(wy[i]);
    std::vector<float> fe(MP, 0.f);
    const double f1 = M >= 1 ? 1.0 / (nT * d.dT) : 1.0;
    for (int m = 0; m <= M; ++m) {
      const double fm = std::fmax((double)m / (nT * d.dT), f1);
      fe[m] = (float)std::fmin(std::fmax(d.f0 / fm, 0.1), 10.0);
    }
    if ((rc = upload(p, lx, &p->T.lgx))) return rc;
    if ((rc = upload(p, ly, &p->T.lgy))) return rc;
    if ((rc = upload(p, fe, &p->T.fw_exp))) return rc;
  }
  // super-absorbing weights (ASS:275-284) and c0/omega (ASS:219-228)
  {
    auto ramp = [](int N, int n) {
      std::vector<float> w(N, 0.f);
      for (int i = 0; i < N; ++i) {
        if (i < n) w[i] = (float)((double)(n - i - 1) / n);
        if (i >= N - n) w[i] = (float)((double)(i - (N - n - 1)) / n);
      }
      return w;
    };
    p->sa_band = (int)std::min<long>(by, nY);
    if ((rc = upload(p, ramp(nX, (int)bx), &p->T.sa_wx))) return rc;
    if ((rc = upload(p, ramp(nY, (int)by), &p->T.sa_wy))) return rc;
    std::vector<float> cw(MP, 0.f);
    const double w1 = 2 * M_PI * (1.0 / (nT * d.dT));
    for (int m = 1; m <= M; ++m) {
      const double w = 2 * M_PI * ((double)m / (nT * d.dT));
      if (std::fabs(w) > 0.5 * w1) cw[m] = (float)(d.c0 / w);
    }
    if ((rc = upload(p, cw, &p->T.sa_cw))) return rc;
  }
  if ((rc = upload(p, twiddles(nX), &p->T.twx))) return rc;
  if ((rc = upload(p, twiddles(nY), &p->T.twy))) return rc;
  p->T.twgx = p->T.twx;  // the same full-length tables serve the general-size kernels
  p->T.twgy = p->T.twy;
  if ((rc = upload(p, twiddles(nT), &p->T.twgt))) return rc;
  if (!p->gen_t) {
    // pass-major twiddles of the column transform, [r * Ns + k] = W_{Ns R}^{k r}
    auto radices = [&](bool inv) {
      std::vector<int> r;
      switch (M) {
        case 128: r = {8, 4, 4}; break;
        case 256: r = {8, 8, 4}; break;
        case 512: r = {8, 8, 8}; break;
        default: r = {16, 8, 8}; break;
      }
      if (inv) std::reverse(r.begin(), r.end());
      return r;
    };
    std::vector<float2> tc;
    for (int inv = 0; inv < 2; ++inv) {
      std::vector<int> r = radices(inv);
      int ns = r[0];
      for (size_t pz = 1; pz < r.size(); ++pz) {
        const int R = r[pz];
        for (int rr = 0; rr < R; ++rr)
          for (int k = 0; k < ns; ++k) {
            const double a = -2.0 * M_PI * (double)k * rr / ((double)ns * R);
            tc.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
          }
        ns *= R;
      }
    }
    if ((rc = upload(p, tc, &p->T.twc))) return rc;
  }
  return ASH_OK;
}

// transverse cosine filter band of ASS:411-426 for step dz
void filter_band(const ash_plan* p, double dz, double* k_start, double* k_max) {
  const ash_desc& d = p->d;
  const double kmax = std::min(M_PI / d.dX, M_PI / d.dY);
  const double norm_step = dz / p->lambda_char;
  if (d.useAdaptiveFiltering) {
    const double cutoff = d.freqFilterThreshold * (1 + d.adaptiveFilterStrength * norm_step * 5);
    *k_max = std::min(kmax * (1 - cutoff), kmax * 0.95);
    *k_start = *k_max * 0.8;
  } else {
    *k_max = kmax * 0.95;
    *k_start = *k_max * 0.9;
  }
}

OpDesc make_op(const ash_plan* p, bool prop, double dz_prop, bool att, double dz_att, double dz_filter) {
  OpDesc o{};
  o.use_prop = prop;
  o.use_att = att;
  o.obliquity = p->d.useObliquityCorrection;
  o.clip_negative = p->d.alpha0 < 0;
  o.dz_prop = dz_prop;
  o.dz_att = dz_att;
  filter_band(p, dz_filter, &o.k_start, &o.k_max);
  o.norm = (float)(1.0 / ((double)p->d.nX * p->d.nY));
  o.k_start2 = o.k_start * o.k_start;
  o.k_max2 = o.k_max * o.k_max;
  o.k_startf = (float)o.k_start;
  o.inv_band = (float)(1.0 / (o.k_max - o.k_start));
  o.dzp = (float)dz_prop;
  o.precise = p->precise_ops;
  return o;
}

// ---- per-kernel timing -------------------------------------------------------
hipEvent_t prof_event(ash_plan* p) {
  if (p->ev_next == p->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    p->ev_pool.push_back(e);
  }
  return p->ev_pool[p->ev_next++];
}
struct ProfScope {
  ash_plan* p;
  hipStream_t s;
  int kind;
  hipEvent_t a = nullptr;
  ProfScope(ash_plan* p_, int kind_, hipStream_t s_) : p(p_), s(s_), kind(kind_) {
    if (p->prof_on) {
      a = prof_event(p);
      if (a) (void)hipEventRecord(a, s);
    }
  }
  ~ProfScope() {
    if (p->prof_on && a) {
      hipEvent_t b = prof_event(p);
      if (b) {
        (void)hipEventRecord(b, s);
        p->prof_ev.push_back({kind, {a, b}});
      }
    }
  }
};

// ---- launch helpers --------------------------------------------------------
// rvx / rvy: radix order of the x-line and y-line transforms (Radix<L, V>).
template <int L>
void launch_xfwd_L(ash_plan* p, hipStream_t s) {
  const dim3 grid(p->g.nY * p->g.nfb);
  if (p->rvx) hipLaunchKernelGGL((k_xfft_fwd<L, 1>), grid, dim3(LNT), 0, s, p->spec, p->T, p->g);
  else hipLaunchKernelGGL((k_xfft_fwd<L, 0>), grid, dim3(LNT), 0, s, p->spec, p->T, p->g);
}
template <int L>
void launch_ypass_L(ash_plan* p, const OpDesc& op, hipStream_t s) {
  const dim3 grid((p->g.nX / 2) * (p->g.MP / ypass_lines<L>()));  // one block per (x row pair, f group)
  // the FP64-operator reference runs with the default radix order
  if (op.precise) hipLaunchKernelGGL((k_ypass<L, true, 0>), grid, dim3(ypass_threads<L, 0>()), 0, s, p->spec, p->T, p->g, op);
  else if (p->rvy) hipLaunchKernelGGL((k_ypass<L, false, 1>), grid, dim3(ypass_threads<L, 1>()), 0, s, p->spec, p->T, p->g, op);
  else hipLaunchKernelGGL((k_ypass<L, false, 0>), grid, dim3(ypass_threads<L, 0>()), 0, s, p->spec, p->T, p->g, op);
}
template <int L, int EPI>
void launch_xinv_L(ash_plan* p, const EpiDesc& e, hipStream_t s) {
  const dim3 grid(p->g.nY * p->g.nfb);
  if (p->rvx) hipLaunchKernelGGL((k_xfft_inv<L, EPI, 1>), grid, dim3(LNT), 0, s, p->spec, p->T, p->g, e);
  else hipLaunchKernelGGL((k_xfft_inv<L, EPI, 0>), grid, dim3(LNT), 0, s, p->spec, p->T, p->g, e);
}
template <int L>
void launch_ysa_L(ash_plan* p, const ScreenDesc& sd, hipStream_t s) {
  const size_t lds = (size_t)ysa_lds_bytes(L, sd.band, sd.phase != nullptr);
  hipLaunchKernelGGL(k_ysa<L>, dim3(p->g.nX * p->g.nfb), dim3(LNT), lds, s, p->spec, p->T, p->g, sd);
}
template <int M, int MODE, int NLK>
void launch_col_MN(ash_plan* p, const ColDesc& cd, hipStream_t s) {
  const int ncols = p->g.nX * p->g.nY;
  const int per_block = ColCfg<M>::WPB * ColCfg<M>::CPW;
  int blocks = (ncols + per_block - 1) / per_block;
  // persistent: each wave walks several columns.  The final pass runs one
  // resident round of blocks (measured fastest for it: 74 against 83 us at
  // 256x256x1024 with 4096 blocks); the other passes take the tuned grid.
  const bool fin = MODE == COL_FINAL || MODE == COL_FINAL_SRC;
  blocks = std::min(blocks, fin ? p->fin_grid : p->col_grid);
  hipLaunchKernelGGL((k_column<M, MODE, NLK>), dim3(blocks), dim3(256), 0, s, p->spec, p->T, p->g, cd);
}
template <int M, int MODE>
void launch_col_MM(ash_plan* p, const ColDesc& cd, hipStream_t s) {
  switch (cd.nl) {
    case NL_KT: launch_col_MN<M, MODE, NL_KT>(p, cd, s); break;
    case NL_KT_MINMOD: launch_col_MN<M, MODE, NL_KT_MINMOD>(p, cd, s); break;
    case NL_RUSANOV: launch_col_MN<M, MODE, NL_RUSANOV>(p, cd, s); break;
    case NL_TVD: launch_col_MN<M, MODE, NL_TVD>(p, cd, s); break;
    default: launch_col_MN<M, MODE, NL_NONE>(p, cd, s); break;
  }
}
template <int M>
void launch_col_M(ash_plan* p, const ColDesc& cd, hipStream_t s) {
  switch (cd.mode) {
    case COL_LOAD: launch_col_MN<M, COL_LOAD, NL_NONE>(p, cd, s); break;
    case COL_NL: launch_col_MM<M, COL_NL>(p, cd, s); break;
    case COL_TAPER: launch_col_MN<M, COL_TAPER, NL_NONE>(p, cd, s); break;
    case COL_FINAL:
      if (cd.n_src > 0 || cd.smask) launch_col_MN<M, COL_FINAL_SRC, NL_NONE>(p, cd, s);
      else launch_col_MN<M, COL_FINAL, NL_NONE>(p, cd, s);
      break;
    case COL_STORE: launch_col_MN<M, COL_STORE, NL_NONE>(p, cd, s); break;
    case COL_RAWNL: launch_col_MM<M, COL_RAWNL>(p, cd, s); break;
    case COL_TOREAL: launch_col_MN<M, COL_TOREAL, NL_NONE>(p, cd, s); break;
    case COL_MAXABS: launch_col_MN<M, COL_MAXABS, NL_NONE>(p, cd, s); break;
    case COL_TOF: launch_col_MN<M, COL_TOF, NL_NONE>(p, cd, s); break;
    default: break;
  }
}

#define DISPATCH_L(n, CALL)                    \
  switch (n) {                                 \
    case 64: { constexpr int L = 64; CALL; } break;   \
    case 128: { constexpr int L = 128; CALL; } break; \
    case 192: { constexpr int L = 192; CALL; } break; \
    case 256: { constexpr int L = 256; CALL; } break; \
    case 384: { constexpr int L = 384; CALL; } break; \
    default: break;                            \
  }

dim3 gen_grid(const ash_plan* p, int n_line, int n_other) { return dim3(n_other * (p->g.MP / gen_lines(n_line))); }

void xfwd(ash_plan* p, hipStream_t s) {
  ProfScope ps(p, ASH_K_XFWD, s);
  if (p->gen_x) {
    hipLaunchKernelGGL(k_gx_fwd, gen_grid(p, p->g.nX, p->g.nY), dim3(GNT), 0, s, p->spec, p->T, p->g);
    return;
  }
  DISPATCH_L(p->g.nX, launch_xfwd_L<L>(p, s));
}
void ypass(ash_plan* p, const OpDesc& op, hipStream_t s) {
  ProfScope ps(p, ASH_K_YPASS, s);
  if (p->gen_y) {
    hipLaunchKernelGGL(k_gy_pass, gen_grid(p, p->g.nY, p->g.nX), dim3(GNT), 0, s, p->spec, p->T, p->g, op);
    return;
  }
  DISPATCH_L(p->g.nY, launch_ypass_L<L>(p, op, s));
}
void xinv(ash_plan* p, int epi, const EpiDesc& e, hipStream_t s) {
  ProfScope ps(p, ASH_K_XINV, s);
  if (p->gen_x) {
    const dim3 grid = gen_grid(p, p->g.nX, p->g.nY);
    if (epi == 0) hipLaunchKernelGGL(k_gx_inv<0>, grid, dim3(GNT), 0, s, p->spec, p->T, p->g, e);
    else if (epi == 1) hipLaunchKernelGGL(k_gx_inv<1>, grid, dim3(GNT), 0, s, p->spec, p->T, p->g, e);
    else hipLaunchKernelGGL(k_gx_inv<2>, grid, dim3(GNT), 0, s, p->spec, p->T, p->g, e);
    return;
  }
  if (epi == 0) { DISPATCH_L(p->g.nX, (launch_xinv_L<L, 0>(p, e, s))); }
  else if (epi == 1) { DISPATCH_L(p->g.nX, (launch_xinv_L<L, 1>(p, e, s))); }
  else { DISPATCH_L(p->g.nX, (launch_xinv_L<L, 2>(p, e, s))); }
}
template <int L>
void launch_ypower_L(ash_plan* p, hipStream_t s) {
  const dim3 grid(p->g.nX * p->g.nfb);
  if (p->rvy) hipLaunchKernelGGL((k_ypower<L, 1>), grid, dim3(LNT), 0, s, p->spec, p->T, p->g, p->obl_acc);
  else hipLaunchKernelGGL((k_ypower<L, 0>), grid, dim3(LNT), 0, s, p->spec, p->T, p->g, p->obl_acc);
}
void ypower(ash_plan* p, hipStream_t s) {
  ProfScope ps(p, ASH_K_YPASS, s);
  if (p->gen_y) {
    hipLaunchKernelGGL(k_gy_power, gen_grid(p, p->g.nY, p->g.nX), dim3(GNT), 0, s, p->spec, p->T, p->g, p->obl_acc);
    return;
  }
  DISPATCH_L(p->g.nY, launch_ypower_L<L>(p, s));
}
void ysa(ash_plan* p, const ScreenDesc& sd, hipStream_t s) {
  ProfScope ps(p, ASH_K_YSA, s);
  if (p->gen_y) {
    hipLaunchKernelGGL(k_gy_sa, gen_grid(p, p->g.nY, p->g.nX), dim3(GNT), 0, s, p->spec, p->T, p->g, sd);
    return;
  }
  DISPATCH_L(p->g.nY, launch_ysa_L<L>(p, sd, s));
}

// General-length column kernel: S real samples per lane, 64 S >= nT.
int gen_col_samples(int nT) { return nT <= 128 ? 2 : nT <= 256 ? 4 : nT <= 512 ? 8 : nT <= 1024 ? 16 : 32; }
template <int S, int MODE, int NLK>
void launch_gcol_SMN(ash_plan* p, const ColDesc& cd, hipStream_t s) {
  const int ncols = p->g.nX * p->g.nY;
  const int blocks = std::min((ncols + GC_WPB - 1) / GC_WPB, p->col_grid);
  hipLaunchKernelGGL((k_gcolumn<S, MODE, NLK>), dim3(blocks), dim3(256), 0, s, p->spec, p->T, p->g, cd);
}
template <int S, int MODE>
void launch_gcol_SM(ash_plan* p, const ColDesc& cd, hipStream_t s) {
  switch (cd.nl) {
    case NL_KT: launch_gcol_SMN<S, MODE, NL_KT>(p, cd, s); break;
    case NL_KT_MINMOD: launch_gcol_SMN<S, MODE, NL_KT_MINMOD>(p, cd, s); break;
    case NL_RUSANOV: launch_gcol_SMN<S, MODE, NL_RUSANOV>(p, cd, s); break;
    case NL_TVD: launch_gcol_SMN<S, MODE, NL_TVD>(p, cd, s); break;
    default: launch_gcol_SMN<S, MODE, NL_NONE>(p, cd, s); break;
  }
}
template <int S>
void launch_gcol_S(ash_plan* p, const ColDesc& cd, hipStream_t s) {
  switch (cd.mode) {
    case COL_LOAD: launch_gcol_SMN<S, COL_LOAD, NL_NONE>(p, cd, s); break;
    case COL_NL: launch_gcol_SM<S, COL_NL>(p, cd, s); break;
    case COL_TAPER: launch_gcol_SMN<S, COL_TAPER, NL_NONE>(p, cd, s); break;
    case COL_FINAL:
      if (cd.n_src > 0 || cd.smask) launch_gcol_SMN<S, COL_FINAL_SRC, NL_NONE>(p, cd, s);
      else launch_gcol_SMN<S, COL_FINAL, NL_NONE>(p, cd, s);
      break;
    case COL_STORE: launch_gcol_SMN<S, COL_STORE, NL_NONE>(p, cd, s); break;
    case COL_RAWNL: launch_gcol_SM<S, COL_RAWNL>(p, cd, s); break;
    case COL_TOREAL: launch_gcol_SMN<S, COL_TOREAL, NL_NONE>(p, cd, s); break;
    case COL_MAXABS: launch_gcol_SMN<S, COL_MAXABS, NL_NONE>(p, cd, s); break;
    case COL_TOF: launch_gcol_SMN<S, COL_TOF, NL_NONE>(p, cd, s); break;
    default: break;
  }
}

// Blocks of the final column pass: one resident round (CUs x blocks per CU).
template <int M>
int final_grid_M(ash_plan* p) {
  int dev = 0, ncu = 0, per_cu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_column<M, COL_FINAL, NL_NONE>, 256, 0));
  p->fin_grid = std::max(1, ncu * per_cu);
  return ASH_OK;
}
int set_final_grid(ash_plan* p) {
  switch (p->g.M) {
    case 128: return final_grid_M<128>(p);
    case 256: return final_grid_M<256>(p);
    case 512: return final_grid_M<512>(p);
    case 1024: return final_grid_M<1024>(p);
    default: return ASH_OK;
  }
}

void column(ash_plan* p, const ColDesc& cd, hipStream_t s) {
  const int kind = cd.mode == COL_NL ? ASH_K_COL_NL : cd.mode == COL_TAPER ? ASH_K_COL_TAPER
                 : cd.mode == COL_FINAL ? ASH_K_COL_FINAL : ASH_K_COL_OTHER;
  ProfScope ps(p, kind, s);
  if (p->gen_t) {
    switch (gen_col_samples(p->g.nT)) {
      case 2: launch_gcol_S<2>(p, cd, s); break;
      case 4: launch_gcol_S<4>(p, cd, s); break;
      case 8: launch_gcol_S<8>(p, cd, s); break;
      case 16: launch_gcol_S<16>(p, cd, s); break;
      default: launch_gcol_S<32>(p, cd, s); break;
    }
    return;
  }
  switch (p->g.M) {
    case 128: launch_col_M<128>(p, cd, s); break;
    case 256: launch_col_M<256>(p, cd, s); break;
    case 512: launch_col_M<512>(p, cd, s); break;
    case 1024: launch_col_M<1024>(p, cd, s); break;
    default: break;
  }
}

ColDesc col_base(const ash_plan* p) {
  ColDesc c{};
  c.N = p->N;
  c.dT = p->d.dT;
  c.dz = p->dz;
  c.cfl_target = (float)p->d.cflTarget;
  c.subcycle = 0;
  c.gmax = p->red + RED_GMAX;
  c.subcap = p->red + RED_SUBCAP;
  c.cax = (p->d.nX / 2) * p->d.nY + p->d.nY / 2;
  return c;
}

int nl_kind(const ash_plan* p) {
  if (!p->d.useSplitStep) return NL_RUSANOV;
  if (p->d.fluxScheme == 1) return NL_KT;
  return p->d.useTVD ? NL_TVD : NL_RUSANOV;
}

bool is_kt(int nl) { return nl == NL_KT || nl == NL_KT_MINMOD; }

int col_blocks(const ash_plan* p, int M) {
  if (p->gen_t) return std::min((p->g.nX * p->g.nY + GC_WPB - 1) / GC_WPB, p->col_grid);
  const int per_block = M == 128 ? ColCfg<128>::WPB * ColCfg<128>::CPW
                      : M == 256 ? ColCfg<256>::WPB * ColCfg<256>::CPW
                      : M == 512 ? ColCfg<512>::WPB * ColCfg<512>::CPW
                                 : ColCfg<1024>::WPB * ColCfg<1024>::CPW;
  const int ncols = p->g.nX * p->g.nY;
  return std::min((ncols + per_block - 1) / per_block, p->col_grid);
}

// Global max|p| of the field the nonlinear column pass `cd` will see (same
// transform and taper), into red[RED_GMAX]: a read-only column pass writing one
// partial per block, then a one-block reduction.
void global_max_of_nl_input(ash_plan* p, const ColDesc& cd, hipStream_t s) {
  ColDesc probe = cd;
  probe.mode = COL_MAXABS;
  probe.nl = NL_NONE;
  probe.partial = p->red;
  column(p, probe, s);
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, s, p->red, col_blocks(p, p->g.M), p->red + RED_GMAX);
}

double beta_tvd(const ash_plan* p, double dz) {
  const double norm_step = dz / (p->d.c0 / p->d.f0);
  return std::max(1.0, 2.0 - p->d.adaptiveFilterStrength * norm_step * 10);
}

EpiDesc epi_base(const ash_plan* p) {
  EpiDesc e{};
  e.sa_strength = (float)p->d.superAbsorbingStrength;
  e.inv2dx = (float)(1.0 / (2 * p->d.dX));
  e.invdx = (float)(1.0 / p->d.dX);
  e.scale = 1.f;
  return e;
}

int net_apply(ash_net* n, float* u, int ncols, double c, double dpdt_max, double dT, int32_t* flag,
              hipStream_t s) {
  const int grid = std::min(ncols, std::max(n->n_cu, 1));  // one ~154 KB LDS block (8 waves) per CU
  const float inv = (float)(1.0 / dT);
  switch (n->nT) {
    case 256: hipLaunchKernelGGL(net::k_net_infer<256>, dim3(grid), dim3(net::INF_THREADS), 0, s, u, ncols, n->Wm, n->params, (float)c, (float)dpdt_max, inv, flag); break;
    case 512: hipLaunchKernelGGL(net::k_net_infer<512>, dim3(grid), dim3(net::INF_THREADS), 0, s, u, ncols, n->Wm, n->params, (float)c, (float)dpdt_max, inv, flag); break;
    case 1024: hipLaunchKernelGGL(net::k_net_infer<1024>, dim3(grid), dim3(net::INF_THREADS), 0, s, u, ncols, n->Wm, n->params, (float)c, (float)dpdt_max, inv, flag); break;
    case 2048: hipLaunchKernelGGL(net::k_net_infer<2048>, dim3(grid), dim3(net::INF_THREADS), 0, s, u, ncols, n->Wm, n->params, (float)c, (float)dpdt_max, inv, flag); break;
    default: return fail(ASH_E_UNSUPPORTED, "network column length must be 256, 512, 1024 or 2048");
  }
  return ASH_OK;
}

// Learned nonlinear update in place of the KT column pass: irfft * taper to a
// real field, network on the smooth columns, KT on the flagged ones, rfft.
void learned_nl(ash_plan* p, ColDesc cd, double dz, hipStream_t s) {
  cd.mode = COL_TOREAL;
  cd.taper = 1;
  cd.rout = p->rscratch;
  column(p, cd, s);
  {
    ProfScope ps(p, ASH_K_NET, s);
    net_apply(p->net, p->rscratch, p->g.nX * p->g.nY, dz * p->N / p->d.dT, p->dpdt_max, p->d.dT, p->nflag, s);
  }
  cd.mode = COL_RAWNL;
  cd.taper = 0;
  cd.rin = p->rscratch;
  cd.rout = p->rscratch;
  cd.cflag = p->nflag;
  column(p, cd, s);
  cd.mode = COL_LOAD;
  cd.rin = p->rscratch;
  cd.rout = nullptr;
  cd.cflag = nullptr;
  cd.ibefore = nullptr;
  cd.stats = nullptr;
  column(p, cd, s);
}

// Beam-averaged nonlinear obliquity (ASS:923-965): the absorbing-layer taper in
// t is applied here (the reference scales N on the field after abl), then the
// power-weighted <k/k_z> of the field's 3-D spectrum is summed into obl_acc by
// an x transform, a read-only y transform and the inverse x transform; the
// nonlinear column pass that follows reads N <k/k_z> from obl_acc.
void nl_obliquity(ash_plan* p, ColDesc& cd, hipStream_t s) {
  ColDesc c = cd;
  c.mode = COL_TAPER;
  c.nl = NL_NONE;
  column(p, c, s);
  hipLaunchKernelGGL(k_zero2, dim3(1), dim3(1), 0, s, p->obl_acc);
  xfwd(p, s);
  ypower(p, s);
  EpiDesc e = epi_base(p);
  e.scale = 1.f / (float)p->g.nX;
  xinv(p, 2, e, s);  // fw = sa = 0: scale and DC/Nyquist projection only
  cd.taper = 0;
  cd.nscale = p->obl_acc;
}

// The composite march step (ASS:874-913) on the spectral state, leaving the
// state after the last attenuation visit with epilogue `epi_last` fused.
void enqueue_composite(ash_plan* p, double dz, int epi_last, const EpiDesc& elast, hipStream_t s) {
  EpiDesc e = epi_base(p);
  ColDesc cd = col_base(p);
  cd.dz = dz;
  cd.beta_tvd = (float)beta_tvd(p, dz);
  cd.nl = nl_kind(p);
  cd.subcycle = is_kt(cd.nl) ? p->d.subcycle : 0;  // only the KT update sub-cycles (ASS:842-859)
  if (p->d.useSplitStep) {
    // attenuation half + propagator half, merged (both diagonal in k-space)
    xfwd(p, s);
    ypass(p, make_op(p, true, dz / 2, true, dz * 0.5, dz), s);
    e.half = 1;
    xinv(p, 1, e, s);
    cd.taper = 1;
    if (p->d.useNonlinearityObliquity) nl_obliquity(p, cd, s);
    if (p->net) {
      learned_nl(p, cd, dz, s);
    } else {
      if (cd.subcycle == SUBCYCLE_GLOBAL) global_max_of_nl_input(p, cd, s);
      cd.mode = COL_NL;
      column(p, cd, s);
    }
    xfwd(p, s);
    ypass(p, make_op(p, true, dz / 2, false, 0.0, dz), s);
    xinv(p, 1, e, s);
    cd.mode = COL_TAPER;
    cd.nl = NL_NONE;
    cd.taper = 1;
    cd.nscale = nullptr;
    column(p, cd, s);
    xfwd(p, s);
    ypass(p, make_op(p, false, 0.0, true, dz * 0.5, dz), s);
    xinv(p, epi_last, elast, s);
  } else {
    xfwd(p, s);
    ypass(p, make_op(p, true, dz, false, 0.0, dz), s);
    e.half = 0;
    xinv(p, 1, e, s);
    cd.taper = 2;
    if (p->d.useNonlinearityObliquity) nl_obliquity(p, cd, s);
    cd.mode = COL_NL;
    column(p, cd, s);
    xfwd(p, s);
    ypass(p, make_op(p, false, 0.0, true, dz * 1.0, dz), s);
    xinv(p, epi_last, elast, s);
  }
}

void enqueue_step(ash_plan* p, const ash_step_io* io, hipStream_t s) {
  EpiDesc el = epi_base(p);
  el.fw = p->d.useFreqWeightedBoundary;
  el.sa = p->d.useSuperAbsorbing;
  enqueue_composite(p, io->dz, 2, el, s);
  ScreenDesc sd{};
  sd.sa_strength = (float)p->d.superAbsorbingStrength;
  sd.inv2dy = (float)(1.0 / (2 * p->d.dY));
  sd.invdy = (float)(1.0 / p->d.dY);
  const double f0_bin = p->d.f0 / (1.0 / (p->d.nT * p->d.dT));
  sd.f0b = (float)std::fmax(f0_bin, 1.0);
  sd.band = p->sa_band;
  if (p->d.useSuperAbsorbing || io->n_screens > 0) {
    sd.sa = p->d.useSuperAbsorbing;
    if (io->n_screens > 0) { sd.phase = io->screen_phase_dev[0]; sd.amp = io->screen_amp_dev[0]; }
    ysa(p, sd, s);
    for (int k = 1; k < io->n_screens; ++k) {
      ScreenDesc s2 = sd;
      s2.sa = 0;
      s2.phase = io->screen_phase_dev[k];
      s2.amp = io->screen_amp_dev[k];
      ysa(p, s2, s);
    }
  }
  ColDesc cd = col_base(p);
  cd.mode = COL_FINAL;
  cd.row = p->d_row;
  cd.max_rows = io->max_rows;
  cd.pnp = io->pnp_dev; cd.ppp = io->ppp_dev; cd.pI = io->pI_dev; cd.pIloss = io->pIloss_dev; cd.pax = io->pax_dev;
  cd.stats = io->stats_dev;
  cd.ibefore = p->ibefore;
  cd.rout = io->field_out_dev;
  cd.n_src = io->n_src_dense;
  for (int k = 0; k < io->n_src_dense && k < 4; ++k) cd.src[k] = io->src_dense_dev[k];
  cd.smask = io->src_mask_dev;
  cd.sdelay = io->src_delay_dev;
  cd.taxis = io->taxis_dev;
  cd.s_f0 = io->src_f0; cd.s_p0 = io->src_p0; cd.s_ncyc = io->src_ncycles; cd.s_dur = io->src_dur;
  column(p, cd, s);
  if (p->tof_mode != 0 && io->tof_dev) {
    // time of flight of the field after the step (ASS:1538-1545), read only
    ColDesc ct = col_base(p);
    ct.mode = COL_TOF;
    ct.row = p->d_row;
    ct.max_rows = io->max_rows;
    ct.tof = io->tof_dev;
    ct.tof_mode = p->tof_mode;
    ct.tof_ratio = (float)p->tof_ratio;
    ct.tof_t0 = p->tof_t0;
    ct.tof_dt = p->tof_dt;
    ct.tof_tref = p->tof_tref;
    ct.tofref = p->tofref;
    column(p, ct, s);
  }
  hipLaunchKernelGGL(k_advance_row, dim3(1), dim3(1), 0, s, p->d_row);
}

int load_real(ash_plan* p, const float* in, uint32_t* stats, double* ib, hipStream_t s) {
  ColDesc cd = col_base(p);
  cd.mode = COL_LOAD;
  cd.rin = in;
  cd.stats = stats;
  cd.ibefore = ib;
  column(p, cd, s);
  return ASH_OK;
}

int store_real(ash_plan* p, float* out, hipStream_t s) {
  ColDesc cd = col_base(p);
  cd.mode = COL_STORE;
  cd.rout = out;
  column(p, cd, s);
  return ASH_OK;
}

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ASH_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return ASH_OK;
}

}  // namespace

extern "C" {

int ash_abi_version(void) { return ASH_ABI_VERSION; }

const char* ash_last_error(void) { return g_err.c_str(); }

int ash_centred_wavenumbers(int32_t n, double d, double* out) try {
  if (!out) return fail(ASH_E_ARG, "null argument");
  if (n < 2 || !(d > 0)) return fail(ASH_E_ARG, "n must be >= 2 and d positive");
  const std::vector<double> k = centred_natural(n, d, 1.0);
  std::copy(k.begin(), k.end(), out);
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

int ash_plan_create(const ash_desc* desc, ash_plan** out) try {
  if (!desc || !out) return fail(ASH_E_ARG, "null argument");
  *out = nullptr;
  const ash_desc& d = *desc;
  if (d.nX <= 0 || d.nY <= 0 || d.nT <= 0) return fail(ASH_E_ARG, "grid sizes must be positive");
  if (!(d.dX > 0 && d.dY > 0 && d.dT > 0 && d.c0 > 0 && d.rho0 > 0))
    return fail(ASH_E_ARG, "dX, dY, dT, c0 and rho0 must be positive");
  if (d.nX < 2 || d.nY < 2 || d.nX > GEN_MAX_LINE || d.nY > GEN_MAX_LINE)
    return fail(ASH_E_UNSUPPORTED, "nX and nY must be between 2 and 1024");
  if (d.nT < 4 || d.nT > GEN_MAX_NT) return fail(ASH_E_UNSUPPORTED, "nT must be between 4 and 2048");
  if (d.subcycle < ASH_SUBCYCLE_OFF || d.subcycle > ASH_SUBCYCLE_GLOBAL)
    return fail(ASH_E_ARG, "subcycle must be ASH_SUBCYCLE_OFF, _COLUMN or _GLOBAL");
  if (d.subcycle != ASH_SUBCYCLE_OFF && !(d.useSplitStep && d.fluxScheme == 1))
    return fail(ASH_E_ARG, "CFL sub-cycling applies to the split-step KT update only (fluxScheme kt)");
  // destroys the partial plan on every early return and on an exception
  std::unique_ptr<ash_plan, void (*)(ash_plan*)> guard(new ash_plan(), ash_plan_destroy);
  ash_plan* p = guard.get();
  p->d = d;
  if (p->d.cflTarget <= 0) p->d.cflTarget = 0.5;
  p->g.nX = d.nX; p->g.nY = d.nY; p->g.nT = d.nT;
  p->g.M = d.nT / 2;
  p->g.MP = ((p->g.M + 1 + LF - 1) / LF) * LF;
  p->g.nfb = p->g.MP / LF;
  p->g.nyq = d.nT % 2 == 0 ? p->g.M : -1;
  p->gen_x = !line_supported(d.nX);
  // the FFT y pass works on row pairs x, nX-1-x, whose |kx| agree only for even nX
  p->gen_y = !line_supported(d.nY) || d.nX % 2 != 0;
  p->gen_t = !(d.nT % 2 == 0 && col_supported(d.nT / 2));
  p->N = d.beta / (2 * d.c0 * d.c0 * d.c0 * d.rho0);
  if (const char* ev = std::getenv("ASH_PRECISE_OPS")) p->precise_ops = std::atoi(ev);
  p->dz = 0.0;
  int rc = build_tables(p);
  if (!rc && !p->gen_t) rc = set_final_grid(p);
  if (rc) return rc;
  const size_t nspec = (size_t)d.nX * d.nY * p->g.MP;
  // the line sweeps address the spectrum with 32-bit byte offsets
  if (nspec * sizeof(float2) >= (size_t(1) << 32))
    return fail(ASH_E_UNSUPPORTED, "spectrum larger than 4 GiB");
  // each buffer is recorded as soon as it exists, so a later failure frees it
  auto dalloc = [p](void** ptr, size_t bytes) -> hipError_t {
    p->allocs.reserve(p->allocs.size() + 1);
    hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipSuccess) p->allocs.push_back(*ptr);
    return e;
  };
  hipError_t e;
  if ((e = dalloc((void**)&p->spec, nspec * sizeof(float2))) != hipSuccess ||
      (e = dalloc((void**)&p->ibefore, (size_t)d.nX * d.nY * sizeof(double))) != hipSuccess ||
      (e = dalloc((void**)&p->d_row, 64)) != hipSuccess ||
      (e = dalloc((void**)&p->obl_acc, 64)) != hipSuccess ||
      (e = dalloc((void**)&p->red, RED_WORDS * sizeof(unsigned))) != hipSuccess)
    return fail(ASH_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  // padding bins are read (masked) but never written: clear once
  if ((e = hipMemset(p->spec, 0, nspec * sizeof(float2))) != hipSuccess ||
      (e = hipMemset(p->d_row, 0, 64)) != hipSuccess ||
      (e = hipMemset(p->red, 0, RED_WORDS * sizeof(unsigned))) != hipSuccess ||
      (e = hipDeviceSynchronize()) != hipSuccess)
    return fail(ASH_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
  p->ws_bytes += (int64_t)(nspec * sizeof(float2) + (size_t)d.nX * d.nY * sizeof(double) + RED_WORDS * sizeof(unsigned));
  *out = guard.release();
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

void ash_plan_destroy(ash_plan* p) {
  if (!p) return;
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  for (void* a : p->allocs) (void)hipFree(a);
  for (hipEvent_t e : p->ev_pool) (void)hipEventDestroy(e);
  delete p;
}

int ash_plan_set_dz(ash_plan* p, double dz) try {
  if (!p) return fail(ASH_E_ARG, "null plan");
  if (!(dz > 0)) return fail(ASH_E_ARG, "dz must be positive");
  p->dz = dz;
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

int ash_load_field(ash_plan* p, const float* field, uint32_t* stats, void* stream) try {
  if (!p || !field) return fail(ASH_E_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  load_real(p, field, stats, p->ibefore, s);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, p->red + RED_SUBCAP, 0u);  // a new march
  p->loaded = true;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_store_field(ash_plan* p, float* field, void* stream) try {
  if (!p || !field) return fail(ASH_E_ARG, "null argument");
  if (!p->loaded) return fail(ASH_E_STATE, "no field loaded");
  store_real(p, field, (hipStream_t)stream);
  return check_launch();
} catch (...) {
  return caught_exception();
}

namespace {
int check_io(const ash_plan* p, const ash_step_io* io) {
  if (!p || !io) return fail(ASH_E_ARG, "null argument");
  if (!p->loaded) return fail(ASH_E_STATE, "ash_step before ash_load_field");
  if (!(io->dz > 0)) return fail(ASH_E_ARG, "dz must be positive");
  if (io->row < 0 || io->row >= io->max_rows) return fail(ASH_E_ARG, "row out of range");
  if (!io->pnp_dev || !io->ppp_dev || !io->pI_dev || !io->pIloss_dev || !io->pax_dev || !io->stats_dev)
    return fail(ASH_E_ARG, "output buffers must be non-null");
  if (io->n_screens < 0 || io->n_screens > 4 || io->n_src_dense < 0 || io->n_src_dense > 4)
    return fail(ASH_E_ARG, "at most 4 screens and 4 dense sources per step");
  if (io->src_mask_dev && (!io->src_delay_dev || !io->taxis_dev))
    return fail(ASH_E_ARG, "sparse source needs delay and taxis arrays");
  return ASH_OK;
}

// A plain step (no screen, no source, the plan's dz) can replay a captured graph.
bool plain_step(const ash_plan* p, const ash_step_io* io) {
  return io->n_screens == 0 && io->n_src_dense == 0 && !io->src_mask_dev && io->dz == p->dz;
}

// The io fields a captured step depends on.
bool same_io(const ash_step_io& a, const ash_step_io& b) {
  return a.pnp_dev == b.pnp_dev && a.ppp_dev == b.ppp_dev && a.pI_dev == b.pI_dev && a.pIloss_dev == b.pIloss_dev &&
         a.pax_dev == b.pax_dev && a.stats_dev == b.stats_dev && a.max_rows == b.max_rows &&
         a.field_out_dev == b.field_out_dev && a.dz == b.dz && a.tof_dev == b.tof_dev;
}
}  // namespace

int ash_step(ash_plan* p, const ash_step_io* io, void* stream) try {
  int rc = check_io(p, io);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (io->row != p->host_row) {
    hipLaunchKernelGGL(k_set_row, dim3(1), dim3(1), 0, s, p->d_row, io->row);
  }
  if (plain_step(p, io)) {
    const bool same = p->gexec && same_io(p->gkey, *io);
    if (!same) {
      if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
      if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
      if (!p->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
      HIP_TRY(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
      enqueue_step(p, io, p->cap_stream);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
      if (e != hipSuccess) return fail(ASH_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      p->graph = g;
      HIP_TRY(hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0));
      p->gkey = *io;
    }
    HIP_TRY(hipGraphLaunch(p->gexec, s));
  } else {
    enqueue_step(p, io, s);
  }
  p->host_row = io->row + 1;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_plan_autotune(ash_plan* p, int32_t* choices, void* stream) try {
  if (!p) return fail(ASH_E_ARG, "null plan");
  hipStream_t s = (hipStream_t)stream;
  const size_t nbytes = (size_t)p->g.nX * p->g.nY * p->g.MP * sizeof(float2);
  // the saved spectrum and the timing events are released on every exit path
  std::unique_ptr<void, void (*)(void*)> save(nullptr, [](void* q) { if (q) (void)hipFree(q); });
  auto ev_del = [](hipEvent_t e) { if (e) (void)hipEventDestroy(e); };
  std::unique_ptr<std::remove_pointer<hipEvent_t>::type, decltype(ev_del)> e0(nullptr, ev_del), e1(nullptr, ev_del);
  if (p->loaded) {
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, nbytes));
    save.reset(q);
    HIP_TRY(hipMemcpyAsync(q, p->spec, nbytes, hipMemcpyDeviceToDevice, s));
  } else {
    HIP_TRY(hipMemsetAsync(p->spec, 0, nbytes, s));  // finite data for the trial sweeps
  }
  {
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    e0.reset(a);
    HIP_TRY(hipEventCreate(&b));
    e1.reset(b);
  }
  int rc = ASH_OK;
  auto timed = [&](auto&& launch) -> float {
    launch();  // warm-up
    (void)hipEventRecord(e0.get(), s);
    for (int r = 0; r < 5; ++r) launch();
    (void)hipEventRecord(e1.get(), s);
    (void)hipEventSynchronize(e1.get());
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0.get(), e1.get()) != hipSuccess) rc = fail(ASH_E_HIP, "autotune timing failed");
    return ms / 5.f;
  };
  const double dz = p->dz > 0 ? p->dz : p->d.dX;
  // x-line transforms: forward + inverse per radix order (the general-size
  // kernels have no alternatives)
  if (!p->gen_x) {
    float best = 1e30f;
    int pick = 0;
    for (int v = 0; v < 2; ++v) {
      p->rvx = v;
      EpiDesc e = epi_base(p);
      e.half = 1;
      const float t = timed([&] { xfwd(p, s); xinv(p, 1, e, s); });
      if (t < best) { best = t; pick = v; }
    }
    p->rvx = pick;
  }
  // y pass per radix order
  if (!p->gen_y) {
    float best = 1e30f;
    int pick = 0;
    const OpDesc op = make_op(p, true, dz / 2, true, dz * 0.5, dz);
    for (int v = 0; v < 2; ++v) {
      p->rvy = v;
      const float t = timed([&] { ypass(p, op, s); });
      if (t < best) { best = t; pick = v; }
    }
    p->rvy = pick;
  }
  // persistent grid of the column kernels (timed on the nonlinear pass)
  {
    ColDesc cd = col_base(p);
    cd.dz = dz;
    cd.beta_tvd = (float)beta_tvd(p, dz);
    cd.nl = nl_kind(p);
    cd.subcycle = 0;
    cd.taper = 1;
    cd.mode = COL_NL;
    float best = 1e30f;
    int pick = 2048;
    for (int g : {512, 1024, 2048, 4096}) {
      p->col_grid = g;
      const float t = timed([&] { column(p, cd, s); });
      if (t < best) { best = t; pick = g; }
    }
    p->col_grid = pick;
  }
  if (save) {
    // restore the loaded field (also after a failed timing)
    HIP_TRY(hipMemcpyAsync(p->spec, save.get(), nbytes, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  ++p->epoch;
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
  if (choices) { choices[0] = p->rvx; choices[1] = p->rvy; choices[2] = p->col_grid; }
  if (rc) return rc;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_plan_set_launch(ash_plan* p, const int32_t* choices) try {
  if (!p || !choices) return fail(ASH_E_ARG, "null argument");
  if (choices[0] < 0 || choices[0] > 1 || choices[1] < 0 || choices[1] > 1)
    return fail(ASH_E_ARG, "radix orders are 0 or 1");
  if (choices[2] < 1 || choices[2] > 65535) return fail(ASH_E_ARG, "column grid out of range");
  p->rvx = choices[0];
  p->rvy = choices[1];
  p->col_grid = choices[2];
  ++p->epoch;
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

int ash_batch_create(ash_batch** out) try {
  if (!out) return fail(ASH_E_ARG, "null argument");
  *out = new ash_batch();
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

void ash_batch_destroy(ash_batch* b) {
  if (!b) return;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  for (hipStream_t st : b->branch) (void)hipStreamDestroy(st);
  for (hipEvent_t e : b->fork_join) (void)hipEventDestroy(e);
  delete b;
}

int ash_batch_set_branches(ash_batch* b, int32_t n) try {
  if (!b) return fail(ASH_E_ARG, "null batch");
  if (n < 1 || n > ASH_BATCH_MAX_BRANCHES) return fail(ASH_E_ARG, "branches must be 1..4");
  if (n != b->nbranch) {
    b->nbranch = n;
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  }
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

int ash_batch_step(ash_batch* b, ash_plan* const* plans, const ash_step_io* ios, int32_t n, void* stream) try {
  if (!b || !plans || !ios || n < 1) return fail(ASH_E_ARG, "bad argument");
  bool all_plain = true;
  for (int i = 0; i < n; ++i) {
    int rc = check_io(plans[i], &ios[i]);
    if (rc) return rc;
    for (int j = 0; j < i; ++j)
      if (plans[j] == plans[i]) return fail(ASH_E_ARG, "a plan appears twice in the batch");
    all_plain = all_plain && plain_step(plans[i], &ios[i]);
  }
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n; ++i)
    if (ios[i].row != plans[i]->host_row)
      hipLaunchKernelGGL(k_set_row, dim3(1), dim3(1), 0, s, plans[i]->d_row, ios[i].row);
  if (all_plain) {
    bool same = b->gexec && (int)b->plans.size() == n;
    for (int i = 0; same && i < n; ++i)
      same = b->plans[i] == plans[i] && b->epochs[i] == plans[i]->epoch && same_io(b->keys[i], ios[i]);
    if (!same) {
      if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
      if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
      if (!b->cap) HIP_TRY(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
      // instances go to nbranch concurrent branches of the graph (fork/join on
      // events), balanced by their voxel counts (largest first)
      const int nb = std::min(b->nbranch, n);
      while ((int)b->branch.size() < nb) {
        hipStream_t st = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
        b->branch.push_back(st);
      }
      while ((int)b->fork_join.size() < nb + 1) {
        hipEvent_t ev = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->fork_join.push_back(ev);
      }
      std::vector<int> order(n), lane(n, 0);
      std::vector<double> load(nb, 0.0);
      for (int i = 0; i < n; ++i) order[i] = i;
      auto work = [&](int i) { return (double)plans[i]->g.nX * plans[i]->g.nY * plans[i]->g.nT; };
      std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return work(a) > work(c); });
      for (int i : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lane[i] = k;
        load[k] += work(i);
      }
      HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
      hipError_t ce = hipSuccess;  // first error inside the capture (the capture is ended either way)
      auto keep = [&ce](hipError_t e) { if (ce == hipSuccess) ce = e; };
      if (nb > 1) {
        keep(hipEventRecord(b->fork_join[0], b->cap));
        for (int k = 0; k < nb; ++k) keep(hipStreamWaitEvent(b->branch[k], b->fork_join[0], 0));
        for (int i = 0; i < n; ++i) enqueue_step(plans[i], &ios[i], b->branch[lane[i]]);
        for (int k = 0; k < nb; ++k) {
          keep(hipEventRecord(b->fork_join[1 + k], b->branch[k]));
          keep(hipStreamWaitEvent(b->cap, b->fork_join[1 + k], 0));
        }
      } else {
        for (int i = 0; i < n; ++i) enqueue_step(plans[i], &ios[i], b->cap);
      }
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->cap, &g);
      if (ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(ASH_E_HIP, std::string("batch capture: ") + hipGetErrorString(ce));
      }
      if (e != hipSuccess) return fail(ASH_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      b->graph = g;
      HIP_TRY(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      b->plans.assign(plans, plans + n);
      b->keys.assign(ios, ios + n);
      b->epochs.resize(n);
      for (int i = 0; i < n; ++i) b->epochs[i] = plans[i]->epoch;
    }
    HIP_TRY(hipGraphLaunch(b->gexec, s));
  } else {
    for (int i = 0; i < n; ++i) enqueue_step(plans[i], &ios[i], s);
  }
  for (int i = 0; i < n; ++i) plans[i]->host_row = ios[i].row + 1;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_march_step(ash_plan* p, const float* in, float* out, void* stream) try {
  if (!p || !in || !out) return fail(ASH_E_ARG, "null argument");
  if (!(p->dz > 0)) return fail(ASH_E_STATE, "dz not set");
  hipStream_t s = (hipStream_t)stream;
  load_real(p, in, nullptr, nullptr, s);
  EpiDesc e = epi_base(p);
  enqueue_composite(p, p->dz, 0, e, s);
  store_real(p, out, s);
  p->loaded = true;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_attenuation_step(ash_plan* p, const float* in, float* out, double scale, void* stream) try {
  if (!p || !in || !out) return fail(ASH_E_ARG, "null argument");
  if (!(p->dz > 0)) return fail(ASH_E_STATE, "dz not set");
  hipStream_t s = (hipStream_t)stream;
  load_real(p, in, nullptr, nullptr, s);
  xfwd(p, s);
  ypass(p, make_op(p, false, 0.0, true, p->dz * scale, p->dz), s);
  xinv(p, 0, epi_base(p), s);
  store_real(p, out, s);
  p->loaded = true;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_angular_spectrum_step(ash_plan* p, const float* in, float* out, int32_t half, void* stream) try {
  if (!p || !in || !out) return fail(ASH_E_ARG, "null argument");
  if (!(p->dz > 0)) return fail(ASH_E_STATE, "dz not set");
  hipStream_t s = (hipStream_t)stream;
  load_real(p, in, nullptr, nullptr, s);
  xfwd(p, s);
  ypass(p, make_op(p, true, half ? p->dz / 2 : p->dz, false, 0.0, p->dz), s);
  EpiDesc e = epi_base(p);
  e.half = half ? 1 : 0;
  xinv(p, 1, e, s);
  ColDesc cd = col_base(p);
  cd.mode = COL_TAPER;
  cd.taper = half ? 1 : 2;
  column(p, cd, s);
  store_real(p, out, s);
  p->loaded = true;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_nonlinear_step(ash_plan* p, const float* in, float* out, int32_t scheme, double dz, double beta,
                       int32_t subcycle, void* stream) try {
  if (!p || !in || !out) return fail(ASH_E_ARG, "null argument");
  if (scheme < ASH_NL_NONE || scheme > ASH_NL_KT_MINMOD) return fail(ASH_E_ARG, "unknown scheme");
  if (subcycle < ASH_SUBCYCLE_OFF || subcycle > ASH_SUBCYCLE_GLOBAL) return fail(ASH_E_ARG, "unknown subcycle mode");
  if (subcycle != ASH_SUBCYCLE_OFF && !is_kt(scheme))
    return fail(ASH_E_ARG, "CFL sub-cycling applies to the KT schemes only");
  hipStream_t s = (hipStream_t)stream;
  const long ncols = (long)p->g.nX * p->g.nY;
  ColDesc cd = col_base(p);
  cd.mode = COL_RAWNL;
  cd.nl = scheme;
  cd.dz = dz;
  cd.beta_tvd = (float)beta;
  cd.subcycle = subcycle;
  cd.rin = in;
  cd.rout = out;
  if (subcycle == ASH_SUBCYCLE_GLOBAL) {
    const int nb = (int)((ncols + ABSMAX_COLS - 1) / ABSMAX_COLS);
    hipLaunchKernelGGL(k_absmax_cols, dim3(nb), dim3(256), 0, s, in, ncols, p->g.nT, p->red);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, s, p->red, nb, p->red + RED_GMAX);
  }
  column(p, cd, s);
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_max_abs(ash_plan* p, const float* in, uint32_t* out_bits, void* stream) try {
  if (!p || !in || !out_bits) return fail(ASH_E_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const long ncols = (long)p->g.nX * p->g.nY;
  const int nb = (int)((ncols + ABSMAX_COLS - 1) / ABSMAX_COLS);
  hipLaunchKernelGGL(k_absmax_cols, dim3(nb), dim3(256), 0, s, in, ncols, p->g.nT, p->red);
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, s, p->red, nb, out_bits);
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_plan_subcycle_capped(ash_plan* p, int64_t* out) try {
  if (!p || !out) return fail(ASH_E_ARG, "null argument");
  unsigned v = 0;
  HIP_TRY(hipMemcpy(&v, p->red + RED_SUBCAP, sizeof(v), hipMemcpyDeviceToHost));
  *out = (int64_t)v;
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

}  // extern "C"

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_on_current_device(unsigned** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return fail(ASH_E_UNSUPPORTED, "device index beyond the no-allocation scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = static_cast<unsigned*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return ASH_OK;
}

// k_column<M, COL_RAWNL, NLK> on ncols real columns without a plan (no tables:
// the pass runs no transform).
template <int M>
void launch_rawnl_M(const ColDesc& cd, int ncols, hipStream_t s) {
  Geom g{};
  g.nX = ncols;
  g.nY = 1;
  g.nT = 2 * M;
  g.M = M;
  g.MP = 0;
  const Tables T{};
  const int per_block = ColCfg<M>::WPB * ColCfg<M>::CPW;
  const int blocks = std::min((ncols + per_block - 1) / per_block, 2048);
  switch (cd.nl) {
    case NL_KT: hipLaunchKernelGGL((k_column<M, COL_RAWNL, NL_KT>), dim3(blocks), dim3(256), 0, s, nullptr, T, g, cd); break;
    case NL_KT_MINMOD: hipLaunchKernelGGL((k_column<M, COL_RAWNL, NL_KT_MINMOD>), dim3(blocks), dim3(256), 0, s, nullptr, T, g, cd); break;
    case NL_RUSANOV: hipLaunchKernelGGL((k_column<M, COL_RAWNL, NL_RUSANOV>), dim3(blocks), dim3(256), 0, s, nullptr, T, g, cd); break;
    case NL_TVD: hipLaunchKernelGGL((k_column<M, COL_RAWNL, NL_TVD>), dim3(blocks), dim3(256), 0, s, nullptr, T, g, cd); break;
    default: hipLaunchKernelGGL((k_column<M, COL_RAWNL, NL_NONE>), dim3(blocks), dim3(256), 0, s, nullptr, T, g, cd); break;
  }
}

}  // namespace

extern "C" {

int ash_nonlinear_step_noalloc(int64_t ncols, int32_t nT, double N, double dT, const float* in, float* out,
                               int32_t scheme, double dz, double beta_tvd, int32_t subcycle, double cfl_target,
                               uint32_t* report_dev, void* stream) try {
  if (!in || !out) return fail(ASH_E_ARG, "null argument");
  if (ncols < 1) return fail(ASH_E_ARG, "ncols must be positive");
  if (ncols > (int64_t)NOALLOC_PARTIALS * ABSMAX_COLS)
    return fail(ASH_E_UNSUPPORTED, "ncols exceeds the no-allocation scratch (at most 1048576 columns)");
  if (!(nT == 256 || nT == 512 || nT == 1024 || nT == 2048))
    return fail(ASH_E_UNSUPPORTED, "nT must be one of 256, 512, 1024, 2048");
  if (!(N >= 0 && dT > 0 && dz > 0)) return fail(ASH_E_ARG, "N must be >= 0, dT and dz positive");
  if (scheme < ASH_NL_NONE || scheme > ASH_NL_KT_MINMOD) return fail(ASH_E_ARG, "unknown scheme");
  if (subcycle < ASH_SUBCYCLE_OFF || subcycle > ASH_SUBCYCLE_GLOBAL) return fail(ASH_E_ARG, "unknown subcycle mode");
  if (subcycle != ASH_SUBCYCLE_OFF && !is_kt(scheme))
    return fail(ASH_E_ARG, "CFL sub-cycling applies to the KT schemes only");
  if (subcycle != ASH_SUBCYCLE_OFF && !(cfl_target > 0)) return fail(ASH_E_ARG, "cfl_target must be positive");
  unsigned* sc = nullptr;
  int rc = scratch_on_current_device(&sc);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  ColDesc cd{};
  cd.mode = COL_RAWNL;
  cd.nl = scheme;
  cd.N = N;
  cd.dT = dT;
  cd.dz = dz;
  cd.beta_tvd = (float)beta_tvd;
  cd.subcycle = subcycle;
  cd.cfl_target = (float)cfl_target;
  cd.rin = in;
  cd.rout = out;
  cd.gmax = sc + NOALLOC_GMAX;
  cd.subcap = sc + NOALLOC_SUBCAP;
  // the scratch keeps the previous call's contents: the counter and (outside the
  // global mode) the reported maximum start at zero
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, sc + NOALLOC_SUBCAP, 0u);
  if (subcycle == ASH_SUBCYCLE_GLOBAL) {
    const int nb = (int)((ncols + ABSMAX_COLS - 1) / ABSMAX_COLS);  // <= NOALLOC_PARTIALS (checked above)
    hipLaunchKernelGGL(k_absmax_cols, dim3(nb), dim3(256), 0, s, in, (long)ncols, (int)nT, sc);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, s, sc, nb, sc + NOALLOC_GMAX);
  } else {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, sc + NOALLOC_GMAX, 0u);
  }
  switch (nT / 2) {
    case 128: launch_rawnl_M<128>(cd, (int)ncols, s); break;
    case 256: launch_rawnl_M<256>(cd, (int)ncols, s); break;
    case 512: launch_rawnl_M<512>(cd, (int)ncols, s); break;
    default: launch_rawnl_M<1024>(cd, (int)ncols, s); break;
  }
  if (report_dev) hipLaunchKernelGGL(k_copy_u32x2, dim3(1), dim3(1), 0, s, sc + NOALLOC_GMAX, sc + NOALLOC_SUBCAP, report_dev);
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_freq_weighted_boundary_step(ash_plan* p, const float* in, float* out, void* stream) try {
  if (!p || !in || !out) return fail(ASH_E_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  load_real(p, in, nullptr, nullptr, s);
  xfwd(p, s);
  EpiDesc e = epi_base(p);
  e.fw = 1;
  e.scale = (float)(1.0 / p->d.nX);
  xinv(p, 2, e, s);
  store_real(p, out, s);
  p->loaded = true;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_super_absorbing_step(ash_plan* p, const float* in, float* out, void* stream) try {
  if (!p || !in || !out) return fail(ASH_E_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  load_real(p, in, nullptr, nullptr, s);
  xfwd(p, s);
  EpiDesc e = epi_base(p);
  e.sa = 1;
  e.scale = (float)(1.0 / p->d.nX);
  xinv(p, 2, e, s);
  ScreenDesc sd{};
  sd.sa = 1;
  sd.band = p->sa_band;
  sd.sa_strength = (float)p->d.superAbsorbingStrength;
  sd.inv2dy = (float)(1.0 / (2 * p->d.dY));
  sd.invdy = (float)(1.0 / p->d.dY);
  sd.f0b = 1.f;
  ysa(p, sd, s);
  store_real(p, out, s);
  p->loaded = true;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_apply_phase_screen(ash_plan* p, const float* in, float* out, const float* phase, const float* amp,
                           void* stream) try {
  if (!p || !in || !out || !phase) return fail(ASH_E_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  load_real(p, in, nullptr, nullptr, s);
  ScreenDesc sd{};
  sd.sa = 0;
  sd.phase = phase;
  sd.amp = amp;
  const double f0_bin = p->d.f0 / (1.0 / (p->d.nT * p->d.dT));
  sd.f0b = (float)std::fmax(f0_bin, 1.0);
  ysa(p, sd, s);
  store_real(p, out, s);
  p->loaded = true;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_source_field(const uint8_t* mask, const double* delay, const double* taxis, int64_t ncols, int32_t nT,
                     double f0, double p0, double ncycles, double dur, float* out, void* stream) try {
  if (!mask || !delay || !taxis || !out) return fail(ASH_E_ARG, "null argument");
  if (ncols <= 0 || nT <= 0) return fail(ASH_E_ARG, "ncols and nT must be positive");
  if (!(ncycles != 0.0)) return fail(ASH_E_ARG, "ncycles must be nonzero");
  const long n = (long)ncols * nT;
  const long blocks = std::min<long>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_source_field, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, mask, delay, taxis,
                     (long)ncols, (int)nT, 2.0 * M_PI * f0, ncycles, 2.0 * dur, p0, out);
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_profile_step(ash_plan* p, const ash_step_io* io, int32_t n_iter, float* ms_out, int32_t* count_out,
                     void* stream) try {
  if (!p || !io || !ms_out || !count_out || n_iter <= 0) return fail(ASH_E_ARG, "bad argument");
  if (!p->loaded) return fail(ASH_E_STATE, "ash_profile_step before ash_load_field");
  if (io->row < 0 || io->row + n_iter > io->max_rows) return fail(ASH_E_ARG, "rows out of range");
  hipStream_t s = (hipStream_t)stream;
  for (int k = 0; k < ASH_K_COUNT; ++k) { ms_out[k] = 0.f; count_out[k] = 0; }
  p->prof_ev.clear();
  p->ev_next = 0;
  p->prof_on = true;
  hipLaunchKernelGGL(k_set_row, dim3(1), dim3(1), 0, s, p->d_row, io->row);
  for (int it = 0; it < n_iter; ++it) enqueue_step(p, io, s);
  p->prof_on = false;
  p->host_row = io->row + n_iter;
  int rc = check_launch();
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  for (auto& e : p->prof_ev) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e.second.first, e.second.second));
    ms_out[e.first] += ms;
    count_out[e.first] += 1;
  }
  for (int k = 0; k < ASH_K_COUNT; ++k) {
    if (count_out[k]) ms_out[k] /= (float)count_out[k];
    count_out[k] /= n_iter;  // launches per step
  }
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

int64_t ash_plan_workspace_bytes(const ash_plan* p) { return p ? p->ws_bytes : 0; }

// ---------------------------------------------------------------------------
// Learned nonlinear step
// ---------------------------------------------------------------------------
int64_t ash_net_param_count(void) { return net::PARAM_COUNT; }

int ash_net_create(int32_t nT, int32_t max_cols, ash_net** out) try {
  if (!out) return fail(ASH_E_ARG, "null argument");
  *out = nullptr;
  if (!(nT == 256 || nT == 512 || nT == 1024 || nT == 2048))
    return fail(ASH_E_UNSUPPORTED, "network column length must be 256, 512, 1024 or 2048");
  if (max_cols < 1) return fail(ASH_E_ARG, "max_cols must be positive");
  // destroys the partial network on every early return and on an exception
  std::unique_ptr<ash_net, void (*)(ash_net*)> guard(new ash_net(), ash_net_destroy);
  ash_net* n = guard.get();
  n->nT = nT;
  n->max_cols = max_cols;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&n->n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  auto alloc = [&](void** ptr, size_t bytes) -> int {
    n->allocs.reserve(n->allocs.size() + 1);  // the push_back below cannot throw
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) return fail(ASH_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    n->allocs.push_back(*ptr);
    e = hipMemset(*ptr, 0, bytes);
    if (e != hipSuccess) return fail(ASH_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    return ASH_OK;
  };
  const size_t np = net::PARAM_COUNT, act = (size_t)max_cols * nT * net::C;
  int rc = alloc((void**)&n->params, np * 4);
  if (!rc) rc = alloc((void**)&n->m, np * 4);
  if (!rc) rc = alloc((void**)&n->v, np * 4);
  if (!rc) rc = alloc((void**)&n->mask, np * 4);
  if (!rc) rc = alloc((void**)&n->Wm, (size_t)net::NLAY * net::W_SIZE * 2);
  if (!rc) rc = alloc((void**)&n->Wt, (size_t)net::NLAY * net::W_SIZE * 2);
  for (int j = 0; j < 4 && !rc; ++j) rc = alloc((void**)&n->X[j], act * 2);
  for (int j = 0; j < 2 && !rc; ++j) rc = alloc((void**)&n->Dg[j], act * 2);
  if (!rc) rc = alloc((void**)&n->y, (size_t)max_cols * nT * 4);
  if (!rc) rc = alloc((void**)&n->amp, (size_t)max_cols * 4);
  if (!rc) {
    // trainable entries: W1 input channels 0-1, W4 and b4 output channel 0, all of layers 2-3
    std::vector<float> mk(np, 1.f);
    for (int co = 0; co < net::C; ++co)
      for (int ci = 2; ci < net::C; ++ci)
        for (int k = 0; k < net::KW; ++k) mk[(co * net::C + ci) * net::KW + k] = 0.f;
    const size_t o4 = 3 * (size_t)net::LAYER_SIZE;
    for (int co = 1; co < net::C; ++co) {
      for (int i = 0; i < net::C * net::KW; ++i) mk[o4 + co * net::C * net::KW + i] = 0.f;
      mk[o4 + net::W_SIZE + co] = 0.f;
    }
    hipError_t e = hipMemcpy(n->mask, mk.data(), np * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = fail(ASH_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  if (rc) return rc;
  *out = guard.release();
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

void ash_net_destroy(ash_net* n) {
  if (!n) return;
  for (void* a : n->allocs) (void)hipFree(a);
  delete n;
}

int ash_net_set_params(ash_net* n, const float* params, void* stream) try {
  if (!n || !params) return fail(ASH_E_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpyAsync(n->params, params, net::PARAM_COUNT * 4, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(net::k_net_pack, dim3((net::NLAY * net::W_SIZE + 255) / 256), dim3(256), 0, s, n->params, n->Wm,
                     n->Wt);
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_net_get_params(ash_net* n, float* params, void* stream) try {
  if (!n || !params) return fail(ASH_E_ARG, "null argument");
  HIP_TRY(hipMemcpyAsync(params, n->params, net::PARAM_COUNT * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

namespace {
int net_forward(ash_net* n, const float* u, int ncols, float* y, float* amp, hipStream_t s) {
  const int nT = n->nT;
  hipLaunchKernelGGL(net::k_net_feat, dim3(ncols), dim3(256), 0, s, u, nT, n->X[0], amp);
  const dim3 grid(nT / net::TT, ncols);
  for (int j = 0; j < 3; ++j)
    hipLaunchKernelGGL(net::k_net_conv<0>, grid, dim3(256), 0, s, n->X[j], nT, n->Wm + j * net::W_SIZE,
                       n->params + j * net::LAYER_SIZE + net::W_SIZE, (const uint16_t*)nullptr, n->X[j + 1],
                       (float*)nullptr);
  hipLaunchKernelGGL(net::k_net_conv<1>, grid, dim3(256), 0, s, n->X[3], nT, n->Wm + 3 * net::W_SIZE,
                     n->params + 3 * net::LAYER_SIZE + net::W_SIZE, (const uint16_t*)nullptr, (uint16_t*)nullptr, y);
  return check_launch();
}
}  // namespace

int ash_net_forward(ash_net* n, const float* u, int32_t ncols, float* y, float* amp, void* stream) try {
  if (!n || !u || !y || !amp) return fail(ASH_E_ARG, "null argument");
  if (ncols < 1 || ncols > n->max_cols) return fail(ASH_E_ARG, "ncols out of range");
  return net_forward(n, u, ncols, y, amp, (hipStream_t)stream);
} catch (...) {
  return caught_exception();
}

int ash_net_loss_grad(ash_net* n, const float* u, const float* un, int32_t ncols, double c, double* loss,
                      float* grads, void* stream) try {
  if (!n || !u || !un || !loss || !grads) return fail(ASH_E_ARG, "null argument");
  if (ncols < 1 || ncols > n->max_cols) return fail(ASH_E_ARG, "ncols out of range");
  if (!(c > 0)) return fail(ASH_E_ARG, "c must be positive");
  hipStream_t s = (hipStream_t)stream;
  const int nT = n->nT;
  int rc = net_forward(n, u, ncols, n->y, n->amp, s);
  if (rc) return rc;
  hipLaunchKernelGGL(net::k_net_loss, dim3((nT + 255) / 256, ncols), dim3(256), 0, s, n->y, u, un, n->amp, nT,
                     (float)c, (float)(1.0 / ((double)ncols * nT)), n->Dg[0], loss);
  HIP_TRY(hipMemsetAsync(grads, 0, net::PARAM_COUNT * 4, s));
  const dim3 grid(nT / net::TT, ncols);
  int cur = 0;
  for (int j = 3; j >= 0; --j) {
    float* gW = grads + j * net::LAYER_SIZE;
    hipLaunchKernelGGL(net::k_net_wgrad, grid, dim3(256), 0, s, n->Dg[cur], n->X[j], nT, gW, gW + net::W_SIZE);
    if (j > 0) {
      hipLaunchKernelGGL(net::k_net_conv<2>, grid, dim3(256), 0, s, n->Dg[cur], nT, n->Wt + j * net::W_SIZE,
                         (const float*)nullptr, n->X[j], n->Dg[1 - cur], (float*)nullptr);
      cur = 1 - cur;
    }
  }
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_net_adam(ash_net* n, const float* grads, float lr, float b1, float b2, float eps, int32_t step,
                 void* stream) try {
  if (!n || !grads) return fail(ASH_E_ARG, "null argument");
  if (step < 1) return fail(ASH_E_ARG, "step counts from 1");
  hipStream_t s = (hipStream_t)stream;
  const float bc1 = (float)(1.0 - std::pow((double)b1, step)), bc2 = (float)(1.0 - std::pow((double)b2, step));
  hipLaunchKernelGGL(net::k_net_adam, dim3((net::PARAM_COUNT + 255) / 256), dim3(256), 0, s, n->params, grads, n->m,
                     n->v, n->mask, net::PARAM_COUNT, lr, b1, b2, eps, bc1, bc2);
  hipLaunchKernelGGL(net::k_net_pack, dim3((net::NLAY * net::W_SIZE + 255) / 256), dim3(256), 0, s, n->params, n->Wm,
                     n->Wt);
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_net_apply(ash_net* n, float* u, int32_t ncols, double c, double dpdt_max, double dT, int32_t* flag,
                  void* stream) try {
  if (!n || !u || !flag) return fail(ASH_E_ARG, "null argument");
  if (ncols < 1) return fail(ASH_E_ARG, "ncols must be positive");
  if (!(dT > 0)) return fail(ASH_E_ARG, "dT must be positive");
  int rc = net_apply(n, u, ncols, c, dpdt_max, dT, flag, (hipStream_t)stream);
  if (rc) return rc;
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_plan_set_tof(ash_plan* p, int32_t mode, double env_ratio, double t0, double dt, const float* ref_dev,
                     double t_ref_peak, void* stream) try {
  if (!p) return fail(ASH_E_ARG, "null plan");
  if (mode < ASH_TOF_OFF || mode > ASH_TOF_MATCHED) return fail(ASH_E_ARG, "unknown TOF mode");
  if (mode != ASH_TOF_OFF && !(dt > 0)) return fail(ASH_E_ARG, "the time axis spacing must be positive");
  if (mode == ASH_TOF_MATCHED && !ref_dev) return fail(ASH_E_ARG, "the matched filter needs a reference trace");
  hipStream_t s = (hipStream_t)stream;
  if (mode == ASH_TOF_MATCHED) {
    if (!p->tofref) {
      void* q = nullptr;
      p->allocs.reserve(p->allocs.size() + 1);  // the push_back below cannot throw
      HIP_TRY(hipMalloc(&q, (size_t)p->g.MP * sizeof(float2)));
      p->allocs.push_back(q);
      p->tofref = static_cast<float2*>(q);
      p->ws_bytes += (int64_t)p->g.MP * sizeof(float2);
    }
    // rfft of the reference trace: the load pass of the column kernels on one column
    const Geom gsave = p->g;
    float2* ssave = p->spec;
    p->g.nX = 1;
    p->g.nY = 1;
    p->spec = p->tofref;
    ColDesc cd = col_base(p);
    cd.mode = COL_LOAD;
    cd.rin = ref_dev;
    column(p, cd, s);
    p->g = gsave;
    p->spec = ssave;
  }
  p->tof_mode = mode;
  p->tof_ratio = env_ratio;
  p->tof_t0 = t0;
  p->tof_dt = dt;
  p->tof_tref = t_ref_peak;
  ++p->epoch;  // the captured step graphs hold the previous TOF pass
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
  return check_launch();
} catch (...) {
  return caught_exception();
}

int ash_plan_set_learned(ash_plan* p, ash_net* n, double dpdt_max) try {
  if (!p) return fail(ASH_E_ARG, "null plan");
  if (n) {
    if (!p->d.useSplitStep || p->d.fluxScheme != 1)
      return fail(ASH_E_UNSUPPORTED, "the learned step replaces the split-step KT update only");
    if (p->d.useNonlinearityObliquity)
      return fail(ASH_E_UNSUPPORTED, "the learned step does not take the beam obliquity scale");
    if (p->d.subcycle != ASH_SUBCYCLE_OFF)
      return fail(ASH_E_UNSUPPORTED, "the learned step replaces the un-sub-cycled KT update only");
    if (n->nT != p->d.nT) return fail(ASH_E_ARG, "network column length differs from the plan's nT");
    const size_t ncol = (size_t)p->d.nX * p->d.nY;
    // each buffer is allocated while missing, so a call after a failed one completes the set
    p->allocs.reserve(p->allocs.size() + 2);  // the push_backs below cannot throw
    if (!p->rscratch) {
      HIP_TRY(hipMalloc(&p->rscratch, ncol * p->d.nT * sizeof(float)));
      p->allocs.push_back(p->rscratch);
      p->ws_bytes += (int64_t)(ncol * p->d.nT * sizeof(float));
    }
    if (!p->nflag) {
      HIP_TRY(hipMalloc(&p->nflag, ncol * sizeof(int32_t)));
      p->allocs.push_back(p->nflag);
      p->ws_bytes += (int64_t)(ncol * sizeof(int32_t));
      HIP_TRY(hipMemset(p->nflag, 0, ncol * sizeof(int32_t)));
    }
  }
  p->net = n;
  p->dpdt_max = dpdt_max;
  // the captured step graphs hold the other update
  ++p->epoch;
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
  return ASH_OK;
} catch (...) {
  return caught_exception();
}

}  // extern "C"
