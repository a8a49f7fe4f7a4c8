/*
 * aship: C ABI of the MI355X (gfx950) nonlinear angular-spectrum marching library.
 *
 * The library replaces the per-z marching step of the reference solver
 * pinton-lab/angularspectrum_heterogeneous, module angular_spectrum_solver.py
 * (cited as ASS:<line>).  The reference has no FFI; its boundary is the Python
 * API (SolverParams, angular_spectrum_solve) and the jitted step functions.
 * Each entry point below names the reference function whose behaviour it
 * reproduces.
 *
 * Conventions
 *   - Every pointer argument named *_dev is device memory owned by the caller
 *     (PyTorch tensors in the Python host code).  Real fields are float32,
 *     C-contiguous (nX, nY, nT) with t fastest, exactly the reference layout.
 *   - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 *     All work is enqueued asynchronously on it.
 *   - Return value 0 on success, a negative ASH_E_* code on failure; the text
 *     of the last error of the calling thread is returned by ash_last_error().
 *     No exception or abort crosses the ABI.
 *   - A plan is not thread-safe: calls on one plan must be serialised by the
 *     caller.  Distinct plans are independent.
 *   - The plan owns only its workspace (the spectral state of the field, FP64
 *     intensity carry, operator tables and the captured step graph).
 */
#ifndef ASHIP_H
#define ASHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASH_ABI_VERSION 3

#define ASH_OK 0
#define ASH_E_ARG (-1)         /* invalid argument */
#define ASH_E_UNSUPPORTED (-2) /* grid size or option not implemented */
#define ASH_E_HIP (-3)         /* HIP runtime error */
#define ASH_E_STATE (-4)       /* call out of order (e.g. step before load) */
#define ASH_E_HOST (-5)        /* host-side failure: out of host memory, internal exception */

/* boundaryProfile */
#define ASH_PROFILE_QUADRATIC 0
#define ASH_PROFILE_WENDLAND 1
/* nonlinear scheme (dispatch of ASS:1454-1492) */
#define ASH_NL_NONE 0
#define ASH_NL_KT 1          /* _kt_flux, ASS:833-839 */
#define ASH_NL_RUSANOV 2     /* _rusanov_flux_standard, ASS:688-696 */
#define ASH_NL_TVD 3         /* _rusanov_flux_tvd, ASS:699-724 */
#define ASH_NL_KT_MINMOD 4   /* _kt_flux_minmod, ASS:862-868 */
/* CFL sub-cycling of the KT update (0 = none, as the reference march).  A step
 * whose CFL number N max|u| dz / dT exceeds cflTarget is split into
 * ceil(CFL / cflTarget) SSP-RK2 sub-steps of dz / n.
 *   GLOBAL: one count for the whole field from its max|u|, as _kt_flux_adaptive
 *           (ASS:842-859), with the maximum reduced on the device (no host sync).
 *   COLUMN: each (x, y) column from its own max|u| (extension; equal to GLOBAL
 *           wherever the field's CFL number is at or below the target).
 * Sub-cycling applies to the KT schemes only.  The kernels take at most 1024
 * sub-steps per dz; a column (or field) that would need more is counted, see
 * ash_plan_subcycle_capped. */
#define ASH_SUBCYCLE_OFF 0
#define ASH_SUBCYCLE_COLUMN 1
#define ASH_SUBCYCLE_GLOBAL 2
#define ASH_MAX_SUBSTEPS 1024

/* Solver description: the fields of SolverParams (ASS:35-76) that shape the
 * march, plus the grid.  Flags are 0/1. */
typedef struct ash_desc {
  int32_t nX, nY, nT;
  double dX, dY, dT;
  double c0, rho0, beta;
  double alpha0;          /* dB/MHz^pow/cm; negative selects water (ASS:1316) */
  double attenPow;
  double f0;
  double boundaryFactor;
  int32_t useSplitStep;
  int32_t useAdaptiveFiltering;
  int32_t useTVD;
  double freqFilterThreshold;
  double adaptiveFilterStrength;
  int32_t useBoundaryLayer;
  int32_t boundaryProfile;         /* ASH_PROFILE_* */
  int32_t useFreqWeightedBoundary;
  int32_t useSuperAbsorbing;
  double superAbsorbingStrength;
  int32_t fluxScheme;              /* 0 rusanov, 1 kt (ASS:64) */
  int32_t useObliquityCorrection;
  int32_t subcycle;                /* ASH_SUBCYCLE_* (extension, default off; split-step KT only) */
  double cflTarget;                /* target CFL of sub-cycling (0.5 in ASS:842) */
  int32_t useNonlinearityObliquity; /* N scaled by the beam's power-weighted <k/k_z> (ASS:923-965) */
} ash_desc;

typedef struct ash_plan ash_plan;

/* Per-step inputs and outputs of ash_step.  Output arrays are device buffers
 * laid out (max_rows, nX, nY) (and (max_rows, nT) for pax): row r holds step
 * r.  The reference stores the same values at [:, :, r] (ASS:1530-1535). */
typedef struct ash_step_io {
  int32_t row;                 /* step index cc */
  int32_t max_rows;
  float* pnp_dev;
  float* ppp_dev;
  float* pI_dev;
  float* pIloss_dev;
  float* pax_dev;
  uint32_t* stats_dev;         /* [max_rows+1] float bits of max|p|; slot r+1 after step r */
  float* field_out_dev;        /* optional: real field after the step (NULL = not written) */
  double dz;                   /* this step's dZ (the last step may be shorter) */
  /* phase/amplitude screens crossed in this step (ASS:1505-1518), in order */
  int32_t n_screens;
  const float* screen_phase_dev[4];
  const float* screen_amp_dev[4];  /* entries may be NULL */
  /* dense source slices crossed in this step (ASS:1521-1527), in order */
  int32_t n_src_dense;
  const float* src_dense_dev[4];
  /* sparse bowl slice (mask + delay form of make_bowl_source_planes slices) */
  const uint8_t* src_mask_dev;     /* [nX*nY] or NULL */
  const double* src_delay_dev;     /* [nX*nY] */
  const double* taxis_dev;         /* [nT] */
  double src_f0, src_p0, src_ncycles, src_dur;
  /* time of flight after the step (ash_plan_set_tof): retarded arrival time of each
   * column, (max_rows, nX, nY); NULL = not computed */
  float* tof_dev;
} ash_step_io;

/* ABI version (ASH_ABI_VERSION). */
int ash_abi_version(void);

/* Text of the last error raised on the calling thread. */
const char* ash_last_error(void);

/* The (n-1)-normalised, mean-subtracted wavenumber grid of ASS:403-408
 * (2 pi j / ((n-1) d) minus its NumPy mean) in natural FFT order, as a plan's
 * tables hold it; out is host memory of n doubles.  Host only, no device use. */
int ash_centred_wavenumbers(int32_t n, double d, double* out);

/* Creates a plan; replaces the operator precompute of ASS:1301-1371
 * (_build_operators / _to_jax).  Operators are evaluated on the fly, so the
 * plan holds only 1-D tables; the field state needs nX*nY*MP*8 bytes. */
int ash_plan_create(const ash_desc* desc, ash_plan** out);

/* Releases the plan and its device workspace. */
void ash_plan_destroy(ash_plan* plan);

/* Sets the z step used by the following calls (ASS:1292-1296 and the partial
 * last step of ASS:1447-1450, which rebuilds the operators at dZ_step). */
int ash_plan_set_dz(ash_plan* plan, double dz);

/* Uploads the initial field into the plan's spectral state (ASS:1408) and
 * records, for the first step, the FP64 intensity I_before = sum_t p^2
 * (ASS:1445) and max|p| into stats_dev[0] (ASS:1418-1420).  stats_dev may be
 * NULL. */
int ash_load_field(ash_plan* plan, const float* field_dev, uint32_t* stats_dev, void* stream);

/* Writes the current real field (the np.array(field) of ASS:1567). */
int ash_store_field(ash_plan* plan, float* field_dev, void* stream);

/* One full z-step of the loop body ASS:1443-1535: the composite march step,
 * frequency-weighted and super-absorbing boundaries, screens, sources and the
 * pnp/ppp/pI/pIloss/pax reductions.  Plain steps (no screen, no source,
 * io->dz equal to the plan dZ) replay one captured hipGraph. */
int ash_step(ash_plan* plan, const ash_step_io* io, void* stream);

/* ---- Operator-level entry points on real fields (parity tests) ---------- */

/* march_step_split_kt / _split_tvd / _split_standard (ASS:883-913) or
 * march_step_sequential (ASS:874-880), selected by the plan flags, at the
 * plan dZ.  in_dev and out_dev may alias. */
int ash_march_step(ash_plan* plan, const float* in_dev, float* out_dev, void* stream);

/* _attenuation_step (ASS:667-685) with the filter of precalculate_ad_pow2 /
 * precalculate_ad (ASS:523-592) at scale `scale` (0.5 = afilt3d_half). */
int ash_attenuation_step(ash_plan* plan, const float* in_dev, float* out_dev, double scale, void* stream);

/* _angular_spectrum_step (ASS:649-655) with HH (half=0) or HH_half (half=1)
 * of precalculate_mas (ASS:396-462) and abl / abl_half (ASS:116-131). */
int ash_angular_spectrum_step(ash_plan* plan, const float* in_dev, float* out_dev, int32_t half, void* stream);

/* _kt_flux, _kt_flux_minmod, _rusanov_flux_standard, _rusanov_flux_tvd
 * (ASS:688-868) on every (x, y) column, with step dz.  subcycle =
 * ASH_SUBCYCLE_GLOBAL is _kt_flux_adaptive (ASS:842-859) at the plan's
 * cflTarget; ASH_SUBCYCLE_COLUMN sub-cycles per column.  Sub-cycling with a
 * non-KT scheme is ASH_E_ARG.  in_dev and out_dev may alias. */
int ash_nonlinear_step(ash_plan* plan, const float* in_dev, float* out_dev, int32_t scheme, double dz,
                       double beta_tvd, int32_t subcycle, void* stream);

/* Global maximum of |field| (the float(jnp.max(jnp.abs(field))) of ASS:849),
 * written as float bits into *out_bits_dev.  The plan's field state is not
 * touched. */
int ash_max_abs(ash_plan* plan, const float* in_dev, uint32_t* out_bits_dev, void* stream);

/* Columns (per-column mode) or fields (global mode) whose sub-cycling needed
 * more than ASH_MAX_SUBSTEPS sub-steps since the last ash_load_field (or plan
 * creation), counted on the device; they took ASH_MAX_SUBSTEPS (one step for a
 * NaN amplitude).  The reference has no bound (its loop runs on the host).
 * Synchronises with the device. */
int ash_plan_subcycle_capped(ash_plan* plan, int64_t* out);

/* ash_nonlinear_step without a plan and without any allocation.  No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code under hipGraph stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The global-max reduction and the sub-step counter use a
 * fixed device array in the library (64 KiB per GPU, held until the process
 * ends), so ncols is at most 1048576 (larger: ASH_E_UNSUPPORTED); nT is 256,
 * 512, 1024 or 2048.  N = beta / (2 c0^3 rho0).  report_dev (optional, 2
 * words, device memory) receives the float bits of the field's max|u| (global
 * mode; 0 otherwise) and the number of columns (or fields) whose sub-steps hit
 * ASH_MAX_SUBSTEPS.  Calls that overlap on one GPU (concurrent streams or
 * threads) race on that array; calls in stream order do not. */
int ash_nonlinear_step_noalloc(int64_t ncols, int32_t nT, double N, double dT, const float* in_dev,
                               float* out_dev, int32_t scheme, double dz, double beta_tvd, int32_t subcycle,
                               double cfl_target, uint32_t* report_dev, void* stream);

/* _freq_weighted_boundary_step (ASS:658-664) with precalculate_abl_freq. */
int ash_freq_weighted_boundary_step(ash_plan* plan, const float* in_dev, float* out_dev, void* stream);

/* _super_absorbing_step (ASS:185-272) with _make_boundary_weights. */
int ash_super_absorbing_step(ash_plan* plan, const float* in_dev, float* out_dev, void* stream);

/* _apply_phase_screen (ASS:352-390); amp_dev may be NULL. */
int ash_apply_phase_screen(ash_plan* plan, const float* in_dev, float* out_dev, const float* phase_dev,
                           const float* amp_dev, void* stream);

/* Dense initial field of a sparse source, synthesised on the device:
 * out[c][t] = mask[c] ? p0 exp(-(1.05 (taxis[t]-delay[c]) w0 / (ncycles pi))^(2 dur))
 *                       sin(w0 (taxis[t]-delay[c])) : 0,  w0 = 2 pi f0,
 * in FP64 rounded once to float32 (c = x*nY + y, ncols = nX*nY, out is
 * [ncols][nT]).  Replaces the NumPy field synthesis of make_bowl_source
 * (ASS:1059-1064), make_bowl_source_planes (ASS:1193-1202) and the sparse-array
 * builder (validate_transcranial.py:307-311); all pointers are device memory. */
int ash_source_field(const uint8_t* mask_dev, const double* delay_dev, const double* taxis_dev, int64_t ncols,
                     int32_t nT, double f0, double p0, double ncycles, double dur, float* out_dev, void* stream);

/* Kernel kinds reported by ash_profile_step */
#define ASH_K_XFWD 0        /* forward FFT along x */
#define ASH_K_YPASS 1       /* FFT along y, operator multiply, inverse FFT */
#define ASH_K_XINV 2        /* inverse FFT along x + epilogue */
#define ASH_K_YSA 3         /* super-absorbing y pass / screen */
#define ASH_K_COL_NL 4      /* t column: irfft, taper, nonlinear step, rfft */
#define ASH_K_COL_TAPER 5   /* t column: irfft, taper, rfft */
#define ASH_K_COL_FINAL 6   /* t column: irfft, sources, reductions */
#define ASH_K_COL_OTHER 7
#define ASH_K_NET 8         /* learned nonlinear step (network inference) */
#define ASH_K_COUNT 9

/* Launch-configuration autotuner, no counterpart in the reference: times the
 * two radix orders of the x-line and of the y-line transforms and four
 * persistent grid sizes of the column kernels on this plan, and keeps the
 * fastest of each.  A loaded field is saved and restored; before
 * ash_load_field the spectrum is cleared first.  The radix order changes the
 * float32 rounding of the transforms (not their result).  choices (optional,
 * 3 entries): x radix order, y radix order, column grid. */
int ash_plan_autotune(ash_plan* plan, int32_t* choices, void* stream);

/* Pins the launch configuration (same 3 entries as ash_plan_autotune reports),
 * e.g. to repeat a tuned run bit for bit.  Defaults: 0, 0, 2048. */
int ash_plan_set_launch(ash_plan* plan, const int32_t* choices);

/* Batched march: one z-step of n independent plans (plan i with ios[i]), for
 * solves of many transmit configurations.  No counterpart in the reference,
 * whose solver runs one configuration per call (ASS:1217).  When every step
 * is plain (as for ash_step), the steps of all plans are captured into one
 * hipGraph and one graph launch advances the whole batch; the graph is
 * replayed while the plan set and the io buffers stay the same.  Plans may
 * have different grid sizes. */
typedef struct ash_batch ash_batch;
int ash_batch_create(ash_batch** out);
void ash_batch_destroy(ash_batch* batch);
int ash_batch_step(ash_batch* batch, ash_plan* const* plans, const ash_step_io* ios, int32_t n, void* stream);

/* Concurrent branches of the batch's captured step (1 = one serial chain; default
 * and maximum ASH_BATCH_MAX_BRANCHES, the hardware queues of a process): the
 * instances are spread over the branches by voxel count, largest first, and the
 * branches fork from and join into the launch stream, so small grids of
 * different instances overlap on the GPU.  The results do not depend on it. */
#define ASH_BATCH_MAX_BRANCHES 4
int ash_batch_set_branches(ash_batch* batch, int32_t n);

/* Timing aid with no counterpart in the reference (whose only timing is the
 * wall clock of ASS:1411, :1552): runs n_iter steps like ash_step, without the
 * graph, with HIP events recorded on `stream` around every kernel launch.
 * ms_out[k] receives the average duration (ms) of one launch of kind k and
 * count_out[k] the launches of kind k per step.  Rows io->row ..
 * io->row + n_iter - 1 of the outputs are written. */
int ash_profile_step(ash_plan* plan, const ash_step_io* io, int32_t n_iter, float* ms_out,
                     int32_t* count_out, void* stream);

/* Bytes of device workspace held by the plan. */
int64_t ash_plan_workspace_bytes(const ash_plan* plan);

/* ---------------------------------------------------------------------------
 * Learned nonlinear step.  No counterpart in the reference: a 4-layer 1-D conv
 * net over t (kernel 9, 32 channels, bf16 activations and weights on the
 * matrix cores, float32 accumulation) that stands in for the KT update over
 * one dZ (_kt_flux, ASS:833-839) on smooth columns.  With A = max|u| of a
 * column, the net maps x = u/A (features x, x^2) to y and the step is
 * u += c A^2 y with c = dZ N / dT.  Columns whose largest slope
 * max|u[t+1]-u[t]| / dT exceeds a threshold (Pa/s) take the KT update.
 * Parameters: float32, ash_net_param_count() entries, four layers of
 * W[co][ci][k] (32 x 32 x 9) followed by b[co] (32); layer 1 reads input
 * channels 0-1 only and layer 4 writes output channel 0 only (the other
 * entries are held at zero).
 * ------------------------------------------------------------------------- */
typedef struct ash_net ash_net;

int64_t ash_net_param_count(void);

/* nT: column length (256, 512, 1024 or 2048); max_cols: largest training
 * batch (columns) for ash_net_forward / ash_net_loss_grad. */
int ash_net_create(int32_t nT, int32_t max_cols, ash_net** out);
void ash_net_destroy(ash_net* net);

/* Copy float32 parameters (device) in or out. */
int ash_net_set_params(ash_net* net, const float* params_dev, void* stream);
int ash_net_get_params(ash_net* net, float* params_dev, void* stream);

/* Training-path forward on ncols columns u[ncols][nT]: y[ncols][nT] is the
 * normalised output, amp[ncols] the column amplitudes A. */
int ash_net_forward(ash_net* net, const float* u_dev, int32_t ncols, float* y_dev, float* amp_dev, void* stream);

/* Mean-square loss of y against y* = (u_next - u) / (c A^2) over all samples,
 * added to *loss_dev (double), and its gradient in the parameters written to
 * grads_dev (float32, ash_net_param_count() entries). */
int ash_net_loss_grad(ash_net* net, const float* u_dev, const float* u_next_dev, int32_t ncols, double c,
                      double* loss_dev, float* grads_dev, void* stream);

/* One Adam update (step counts from 1) of the parameters with grads_dev. */
int ash_net_adam(ash_net* net, const float* grads_dev, float lr, float beta1, float beta2, float eps,
                 int32_t step, void* stream);

/* Fused inference in place on real columns u[ncols][nT]: u += c A^2 y for the
 * columns whose max|u[t+1]-u[t]| / dT is at most dpdt_max; flag[c] = 1 marks
 * the others (left unchanged).  A column with c max|u[t+1]-u[t]| < 1e-5 (a
 * Burgers increment below 1e-5 of its amplitude) is left unchanged with
 * flag 0.  Within a column, a 512-sample chunk (the whole column when
 * shorter) whose window (16 samples of halo each side) stays below 1e-4 A is
 * left unchanged. */
int ash_net_apply(ash_net* net, float* u_dev, int32_t ncols, double c, double dpdt_max, double dT,
                  int32_t* flag_dev, void* stream);

/* Time of flight per (x, y) column after every step (ash_step_io.tof_dev),
 * replacing the reference's calls of its tof_extraction module (imported at
 * ASS:26-29, called at ASS:1538-1545; the module is not shipped and its
 * estimators are documented only at ASS:1244-1264).  The value written is the
 * retarded arrival time; the caller adds z / c0 for the arrival time at depth z.
 *   ASH_TOF_ENVELOPE: the first time the Hilbert envelope of the column reaches
 *     env_ratio of its maximum, interpolated linearly from the sample before
 *     (env_ratio = 1: the envelope-peak sample);
 *   ASH_TOF_MATCHED: t_ref_peak plus the lag of the first maximum of the
 *     circular cross-correlation with ref_dev (nT samples, device memory),
 *     refined by parabolic interpolation over the peak and its two neighbours.
 * Sample k is at time t0 + k dt.  A column that is zero everywhere gets NaN.
 * ASH_TOF_OFF disables the pass. */
#define ASH_TOF_OFF 0
#define ASH_TOF_ENVELOPE 1
#define ASH_TOF_MATCHED 2
int ash_plan_set_tof(ash_plan* plan, int32_t mode, double env_ratio, double t0, double dt, const float* ref_dev,
                     double t_ref_peak, void* stream);

/* Use the network for the nonlinear update of the plan's steps (null: the KT
 * update).  Split-step KT plans only; nT must match.  The plan keeps a real
 * scratch field (nX nY nT float32) while a network is set. */
int ash_plan_set_learned(ash_plan* plan, ash_net* net, double dpdt_max);

#ifdef __cplusplus
}
#endif
#endif
