/*
 * kwave_hip.h -- C ABI of the MI355X-native k-space first-order PSTD time loop.
 *
 * This library replaces the time loop of KWave.jl's
 *   kspace_first_order(kgrid, medium, source, sensor; kwargs...) -> SimulationOutput
 * (KWave.jl/src/solver/first_order.jl:92-281 (1D), 287-499 (2D), 505-728 (3D))
 * and its per-step operators (KWave.jl/src/solver/first_order_steps.jl:1-825).
 * The host wrapper (Julia `ccall`, Python `ctypes`) keeps the reference's
 * objects and validation; everything from table generation to the last
 * recorded sample runs behind this boundary on one GPU.
 *
 * Conventions
 *   - Arrays are column-major with x fastest (Julia layout); linear indices
 *     are 0-based and, for masks, in `findall` (column-major) order.
 *   - Host pointers are borrowed for the duration of the call only.
 *   - Every function returns KW_OK (0) or a negative status; the message of
 *     the last failure on the calling thread is kw_last_error().
 *   - One kw_sim is not reentrant; distinct sims may be driven from
 *     different host threads.  Each sim owns one HIP stream.
 */
#ifndef KWAVE_HIP_H
#define KWAVE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KW_ABI_VERSION 1

enum kw_status {
  KW_OK = 0,
  KW_ERR_INVALID = -1,     /* bad descriptor / argument (reference: error(...)) */
  KW_ERR_UNSUPPORTED = -2, /* valid input this build cannot run (e.g. FFT length) */
  KW_ERR_HIP = -3,         /* HIP runtime failure */
  KW_ERR_OOM = -4,         /* device allocation failed */
  KW_ERR_STATE = -5        /* call out of order (e.g. sensor data before any step) */
};

/* SourceMode, KWave.jl/src/types.jl:14 (same order). */
enum kw_source_mode { KW_DIRICHLET = 0, KW_ADDITIVE = 1, KW_ADDITIVE_NO_CORRECTION = 2 };

/* alpha_mode symbols, KWave.jl/src/medium.jl:66; any other symbol acts as
 * :stokes (first_order_steps.jl:252). */
enum kw_alpha_mode { KW_NO_ABSORPTION = 0, KW_NO_DISPERSION = 1, KW_STOKES = 2 };

/* Record fields, KWave.jl/src/sensor.jl:41-45 (bit per symbol). */
enum kw_record {
  KW_REC_P = 1u << 0,
  KW_REC_P_MAX = 1u << 1,
  KW_REC_P_MIN = 1u << 2,
  KW_REC_P_RMS = 1u << 3,
  KW_REC_P_FINAL = 1u << 4,
  KW_REC_UX = 1u << 5,
  KW_REC_UY = 1u << 6,
  KW_REC_UZ = 1u << 7,
  KW_REC_U_MAX = 1u << 8,
  KW_REC_U_RMS = 1u << 9,
  KW_REC_U_FINAL = 1u << 10,
  KW_REC_I_AVG = 1u << 11,
  KW_REC_I_MAX = 1u << 12
};

/* Fields read by kw_sim_get_field and written by kw_sim_set_field.  KW_FIELD_RHO_PREV is the
 * previous total density of the homogeneous :stokes branch (rho_total_prev,
 * first_order_steps.jl:255-259, 356-358), present only when that branch runs. */
enum kw_field { KW_FIELD_P = 0, KW_FIELD_UX, KW_FIELD_UY, KW_FIELD_UZ, KW_FIELD_RHOX, KW_FIELD_RHOY, KW_FIELD_RHOZ,
                KW_FIELD_RHO_PREV };

/* Storage of the static medium maps (c0, rho0, B/A). */
enum kw_map_storage { KW_MAPS_NATIVE = 0, KW_MAPS_FP16 = 2 };

typedef struct kw_ctx kw_ctx;
typedef struct kw_sim kw_sim;

/*
 * Simulation descriptor: everything kspace_first_order derives from
 * (kgrid, medium, source, sensor, kwargs) before its time loop
 * (first_order.jl:311-440 and 529-669).  Scalars that the reference keeps
 * as Float64 stay double here; array inputs are Float64, column-major.
 */
typedef struct kw_sim_desc {
  int32_t ndim;          /* 1, 2 or 3 (KWaveGrid1D/2D/3D, grid.jl:21-108) */
  int64_t N[3];          /* Nx, Ny, Nz (unused dims = 1) */
  double d[3];           /* dx, dy, dz [m] */
  double dt;             /* kgrid.dt[] */
  int64_t Nt;            /* kgrid.Nt[] (time reversal: columns of the TR data) */
  int32_t precision;     /* 4 = Float32 kernels, 8 = Float64 kernels (data_cast) */
  int32_t map_storage;   /* kw_map_storage; KW_MAPS_FP16 stores c0 - c0_offset and
                            rho0 - rho0_offset as FP16 (arithmetic stays in `precision`) */
  int32_t pml_size[3];   /* per axis (first_order.jl:322, 540) */
  double pml_alpha[3];
  double c_ref;          /* mean(c0) as the host computed it (first_order.jl:361, 581) */
  double c0_offset;      /* KW_MAPS_FP16 only */
  double rho0_offset;    /* KW_MAPS_FP16 only */

  const double* c0;      /* sound speed map or NULL (then c0_scalar) */
  double c0_scalar;
  const double* rho0;    /* density map or NULL (then rho0_scalar) */
  double rho0_scalar;
  int32_t nonlinear;     /* medium.BonA !== nothing */
  const double* bona;    /* B/A map or NULL (then bona_scalar) */
  double bona_scalar;
  int32_t absorbing;     /* !is_lossless(medium) (medium.jl:90) */
  int32_t alpha_mode;    /* kw_alpha_mode */
  double alpha_coeff;    /* dB/(MHz^y cm), scalar (first_order.jl:588-592) */
  double alpha_power;    /* y */

  const double* p0;      /* initial pressure (full grid) or NULL */
  int32_t smooth_p0;     /* apply smooth(p0; restore_max=true) (filters.jl:20-60) */

  /* time-varying pressure source (source.jl:24-34; first_order_steps.jl:561-715) */
  int64_t n_psrc;
  const int64_t* psrc_index;   /* findall(p_mask), 0-based */
  const double* psrc_signal;   /* n_psrc x psrc_ncols, column-major */
  int64_t psrc_ncols;
  int32_t p_mode;              /* kw_source_mode */

  /* time-varying velocity source (first_order_steps.jl:541-559, 586-617, 646-690) */
  int64_t n_usrc;
  const int64_t* usrc_index;   /* findall(u_mask), 0-based */
  const double* usrc_signal[3];/* ux, uy, uz: n_usrc x usrc_ncols[d] each, or NULL */
  int64_t usrc_ncols[3];
  int32_t u_mode;

  /* sensor (sensor.jl:31-38, 94-136) */
  int64_t n_sensor;
  const int64_t* sensor_index; /* findall(mask), 0-based; NULL with n_sensor>0 = Cartesian (never filled) */
  uint32_t record;             /* kw_record bits */

  /* time reversal (first_order.jl:426-433, 450-460): n_sensor x Nt data injected
   * (Dirichlet) at sensor_index, reversed in time; NULL if not TR */
  const double* tr_data;

  /* sensor post-filters on the recorded p (first_order_steps.jl:836-910) */
  int32_t directivity;              /* directivity_angle and directivity_size both set */
  const double* directivity_angle;  /* per-sensor angles [rad] or NULL (then the scalar) */
  double directivity_angle_scalar;
  double directivity_size;          /* [m] */
  int32_t freq_response;            /* frequency_response set */
  double freq_center;               /* [Hz] */
  double freq_bandwidth;            /* [Hz] */

  /* kspace_first_order_as (axisymmetric.jl:26-227): 2-D only, x axial, y
   * radial with r = |y_vec| (0 replaced by dy/2); the radial density update
   * adds ur/r.  tr_data is ignored (that solver has no time reversal). */
  int32_t axisymmetric;
} kw_sim_desc;

/* Library identity and errors. */
int32_t kw_abi_version(void);
const char* kw_last_error(void);
/* sizeof(kw_sim_desc) and sizeof(kw_elastic_desc) as this library was compiled: a binding
 * checks them against its own structure layout when it loads the library (a library and a
 * binding built from different versions of this header would otherwise misread pointers). */
int32_t kw_desc_sizes(int64_t* sim_desc_bytes, int64_t* elastic_desc_bytes);

/* Device context: owns nothing but the device selection. */
int32_t kw_device_count(int32_t* count);
int32_t kw_ctx_create(int32_t device, kw_ctx** out);
int32_t kw_ctx_destroy(kw_ctx* ctx);

/* How an axis of length n runs in kw_sim_create (the acoustic solver): 1 on a compiled
 * line-FFT plan, 2 as chirp-z (Bluestein) lines on a power-of-two plan (any length up to
 * 2048: 2 to 4x the line-FFT work, the same memory traffic), 0 not at all (n < 1, or longer
 * than 2048 without a plan).  Batches, the adjoint, the elastic solver and the operator entry
 * points take compiled lengths (1) only.  FFTW plans any length (fft_utils.jl:44-62). */
int32_t kw_fft_length_supported(int64_t n);

/*
 * 1 if a grid of ndim axes N[0..ndim-1] runs in `precision` (4 or 8), else 0:
 * every axis length runs (kw_fft_length_supported != 0) and no strided line spans more than
 * 4 GiB (the line passes use 32-bit per-lane byte offsets from a 64-bit
 * workgroup base; e.g. 1024^3 in Float32 and 1024^2 x 512 in Float64 exceed
 * it).  kw_sim_create / kw_elastic_create reject such grids with
 * KW_ERR_UNSUPPORTED before allocating.  Host-only; no device is touched.
 * No reference counterpart (FFTW plans any size that fits in memory).
 */
int32_t kw_grid_supported(int32_t ndim, const int64_t* N, int32_t precision);

/*
 * Build a simulation: host-side FP64 tables (k vectors, staggered shifts,
 * kappa, absorption, PML, twiddles), device buffers, p0 initialisation
 * (initialize_p0_*, first_order_steps.jl:58-191).  Replaces
 * first_order.jl:146-224 / 346-440 / 564-669.
 */
int32_t kw_sim_create(kw_ctx* ctx, const kw_sim_desc* desc, kw_sim** out);

/* Advance n_steps time steps (the loop body of first_order.jl:232-266,
 * 448-484, 677-713); returns when they are enqueued (use kw_sim_synchronize). */
int32_t kw_sim_run(kw_sim* sim, int64_t n_steps);
int32_t kw_sim_synchronize(kw_sim* sim);
int32_t kw_sim_steps_done(const kw_sim* sim, int64_t* steps);

/* Copy a full field (column-major, Float64) into caller memory. */
int32_t kw_sim_get_field(kw_sim* sim, int32_t which, double* out);
/* Overwrite a full field of the current state (Float64 in, cast to the kernel precision):
 * with kw_sim_run(sim, 1) this advances one time_step_{1,2,3}d! (first_order_steps.jl:296-535)
 * from a given state -- the operator-level form of the time step.  The step counter is not
 * changed (sources and recording continue from steps_done + 1). */
int32_t kw_sim_set_field(kw_sim* sim, int32_t which, const double* in);

/*
 * Finalised sensor data (finalize_sensor_data!, first_order_steps.jl:818-825)
 * as Float64 in the reference's shapes: time series n_sensor x Nt
 * (column-major), reductions n_sensor.  In time-reversal mode KW_REC_P_FINAL
 * returns the full field (first_order.jl:721-723).
 */
int32_t kw_sim_sensor_shape(kw_sim* sim, uint32_t field, int64_t* rows, int64_t* cols);
int32_t kw_sim_get_sensor(kw_sim* sim, uint32_t field, double* out);

/*
 * Timing support for benchmarks: run n_steps and return the per-kernel-class
 * average device time in milliseconds (HIP events on the sim's stream).
 * names/ms have room for max_classes entries; *n_classes receives the count.
 */
int32_t kw_sim_profile(kw_sim* sim, int64_t n_steps, int32_t max_classes, int32_t* n_classes,
                       const char** names, double* ms, int64_t* launches);

/*
 * Launch-shape autotuner (no counterpart in the reference, whose FFTW plans
 * are tuned by FFTW.MEASURE in create_fft_plans, fft_utils.jl:44-62).  Runs
 * timed trials (steps timed steps each; time reversal ignored) with 1, 1/2,
 * 1/4 and 1/8 of the default lines per workgroup and with twice the default
 * (up to the pass's 512-thread limit, 256 for the complex-to-real pass, and
 * the workgroup LDS) for every pass and keeps
 * the fastest per (precision, pass, line length) in a process-wide table
 * that later kw_sim_create / kw_batch_create calls read.  With n_sims == 1
 * the trials run descs[0] as a single simulation and, for passes whose line
 * length is 256 or 512, also time the alternative plan families 0, 1, 2
 * (radix schedule and threads per line), which single simulations then use;
 * with n_sims > 1 they run the whole array as one batch (kw_batch_create),
 * so the table fits that batch (batches keep the default plan families).
 * kw_tune_export writes the table as text lines "precision pass n lines"
 * with a fifth column "plan" where a plan family other than the pass's
 * default was chosen (*len = its length; buf gets at most cap-1 bytes plus
 * a NUL); kw_tune_import merges such text (a line without the fifth column
 * resets that entry to the default family; KW_ERR_INVALID for a plan family
 * the line length does not have); kw_tune_clear empties the table.
 */
int32_t kw_autotune(kw_ctx* ctx, const kw_sim_desc* descs, int32_t n_sims, int32_t steps);
int32_t kw_tune_export(char* buf, int64_t cap, int64_t* len);
int32_t kw_tune_import(const char* text);
int32_t kw_tune_clear(void);

/*
 * Misfit gradient w.r.t. the sound-speed map (no counterpart in the
 * reference, which has no adjoint; it stands in for differentiating
 * kspace_first_order (first_order.jl:505-728) through an AD tool).
 * Runs the whole simulation from step 0: the forward with state checkpoints
 * every `checkpoint_every` steps, then the discrete adjoint backwards in time,
 * recomputing each segment from its checkpoint.  Loss L = 1/2 sum (p - p_obs)^2
 * over the recorded sensor p (p_obs: n_sensor x Nt, column-major).  Outputs the
 * loss and dL/dc0 over the whole grid (column-major, Float64), including the
 * 1/N term through c_ref = mean(c0) (first_order.jl:361) and the 1/c0^2 of
 * pressure-source injection (first_order_steps.jl:630-631).
 * Requirements: c0 given as a map, linear medium (no B/A), additive sources,
 * no time reversal, maps stored in the kernel precision, p recorded at a
 * binary sensor mask.  After the call the simulation has run Nt steps and
 * its sensor outputs are those of the forward run.
 */
int32_t kw_sim_grad_c0(kw_sim* sim, const double* p_obs, int64_t checkpoint_every, double* grad, double* loss);

/*
 * Stacked firings (the "one simulation per element, batched" form of transmit sequences and of the
 * adjoint gradient over firings): n_firings 2-D descriptors that share the grid, time step, medium,
 * PML and record fields and differ in their time-varying sources and binary sensor masks.  They run
 * as one simulation whose passes cover every firing at once (each firing is an independent 2-D
 * problem; nothing couples them).  The resulting kw_sim is driven like any other:
 *   kw_sim_get_sensor / kw_sim_sensor_shape: rows of every firing's sensor points, firing after
 *     firing (each in its own findall order);
 *   kw_sim_get_field: Nx x Ny x n_firings;
 *   kw_sim_grad_c0: p_obs holds those rows (sum of n_sensor) x Nt; grad is Nx x Ny, the sum of the
 *     firings' gradients; loss is summed.
 * Source signals of different lengths are merged as the reference samples them (column
 * min(t, ncols)).  p0, time reversal, post-filters and Cartesian masks are rejected.
 */
int32_t kw_sim_create_firings(kw_ctx* ctx, const kw_sim_desc* descs, int32_t n_firings, kw_sim** out);

int32_t kw_sim_destroy(kw_sim* sim);

/*
 * Batches of independent simulations (no reference counterpart: the
 * reference runs one simulation per call).  Every pass is one launch per
 * line length over all simulations, the step index lives on the device and,
 * with KW_BATCH_GRAPH in use_graph (a flags word; 0 and 1 keep their old
 * meaning), one step is captured in a hipGraph and replayed.  With
 * KW_BATCH_PERSISTENT each batched pass runs as a persistent grid (as many
 * workgroups as fit the device at once, each looping over the pass's
 * workgroup tiles) instead of one workgroup per tile; KW_BATCH_GRID8(n)
 * (n >= 1, with KW_BATCH_PERSISTENT) caps that grid at 8 n workgroups, e.g.
 * to leave room on the device for other work.  Other bits: KW_ERR_INVALID.  All simulations share precision, dimensionality and Nt;
 * no time reversal.  Sensor data and fields are read per simulation like
 * kw_sim_get_*.
 */
#define KW_BATCH_GRAPH 1
#define KW_BATCH_PERSISTENT 2
#define KW_BATCH_GRID8(n) ((int32_t)(n) << 8)
typedef struct kw_batch kw_batch;
int32_t kw_batch_create(kw_ctx* ctx, const kw_sim_desc* descs, int32_t n_sims, int32_t use_graph, kw_batch** out);
int32_t kw_batch_run(kw_batch* batch, int64_t n_steps);
int32_t kw_batch_synchronize(kw_batch* batch);
int32_t kw_batch_sensor_shape(kw_batch* batch, int32_t sim, uint32_t field, int64_t* rows, int64_t* cols);
int32_t kw_batch_get_sensor(kw_batch* batch, int32_t sim, uint32_t field, double* out);
int32_t kw_batch_get_field(kw_batch* batch, int32_t sim, int32_t which, double* out);
int32_t kw_batch_destroy(kw_batch* batch);

/*
 * Operator-level entry points (parity tests of single operators).
 * kw_test_axis_table: the per-axis tables the solvers build on the host, FP64:
 *   KW_TAB_K       k vector (grid.jl:120-130), n values
 *   KW_TAB_PML     get_pml (pml.jl:25-55), n values; KW_TAB_PML_SG the staggered profile
 *   KW_TAB_DPLUS   i k exp(+i k d/2) (grid.jl:149-153), 2n values (re, im); KW_TAB_DMINUS with exp(-i k d/2);
 *                  real part only at the even-n Nyquist bin (what the complex-to-real inverse keeps)
 * kw_test_eos_table: the EOS multiplier of _compute_pressure! (first_order_steps.jl:245-289) on the
 *   rfft grid (Nx/2+1, Ny, Nz), column-major, before the inverse transform's 1/N: which 0 =
 *   c^2 kappa (1 + nabla1) (c^2 = 1 for a c0 map), which 1 = -c^2 kappa nabla2 (homogeneous :stokes).
 * kw_test_spectral_gradient: spectral_gradient! (fft_utils.jl:126-182) of a real field along one axis
 *   on the GPU: sign +1 / -1 = the staggered multipliers above, 0 = unstaggered i k (x axis, or an
 *   odd-length axis: the N-D transform is line-separable there).
 */
enum kw_axis_table { KW_TAB_K = 0, KW_TAB_PML, KW_TAB_PML_SG, KW_TAB_DPLUS, KW_TAB_DMINUS };
int32_t kw_test_axis_table(int64_t n, double d, int32_t pml_size, double pml_alpha, double dt, int32_t which,
                           double* out);
int32_t kw_test_eos_table(const kw_sim_desc* desc, int32_t which, double* out);
int32_t kw_test_spectral_gradient(kw_ctx* ctx, int32_t precision, int32_t ndim, const int64_t* N, const double* d,
                                  int32_t axis, int32_t sign, const double* f, double* out);

/*
 * kw_spectral_gradient_noalloc: kw_test_spectral_gradient on device arrays, allocating nothing.
 * No counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  f and out are device
 * arrays of prod(N) elements of the precision's type (float or double), column-major; the call is
 * enqueued on `stream` (a hipStream_t, or NULL for the null stream) and returns without waiting.
 * The line FFT's twiddles and the derivative multiplier are generated on the device into a
 * per-device static scratch array of KW_NOALLOC_SCRATCH_BYTES (2 x 4096 complex FP64), whose address
 * kw_ctx_create resolves; every supported axis length fits.  All calls on one GPU share that array:
 * calls that overlap on one GPU (different streams or threads) race and must be serialised by the
 * caller.  Errors as kw_test_spectral_gradient (argument checks run before anything is enqueued).
 */
#define KW_NOALLOC_SCRATCH_BYTES 131072
int32_t kw_spectral_gradient_noalloc(kw_ctx* ctx, int32_t precision, int32_t ndim, const int64_t* N,
                                     const double* d, int32_t axis, int32_t sign, const void* f, void* out,
                                     void* stream);

/*
 * Elastic stress-velocity PSTD: replaces the time loops of
 * pstd_elastic_2d (KWave.jl/src/solver/elastic.jl:133-368) and
 * pstd_elastic_3d (elastic.jl:386-588).  Velocities v_a and stresses s_ab
 * on the staggered grid, no k-space correction; pressure p = -tr(s)/ndim
 * is what the sensor records.  Axis lengths as for kw_sim_desc
 * (kw_fft_length_supported); Cartesian sensor masks are preallocated and
 * never filled, as in the reference.
 */
typedef struct kw_elastic kw_elastic;
enum kw_elastic_field {
  KW_EL_VX = 0, KW_EL_VY, KW_EL_VZ, KW_EL_SXX, KW_EL_SYY, KW_EL_SZZ, KW_EL_SXY, KW_EL_SXZ, KW_EL_SYZ
};
typedef struct kw_elastic_desc {
  int32_t ndim;          /* 2 or 3 */
  int64_t N[3];
  double d[3];
  double dt;
  int64_t Nt;
  int32_t precision;     /* 4 or 8 */
  int32_t pml_size[3];   /* defaults 20 (2-D) / 10 (3-D) are the caller's (elastic.jl:138, 391) */
  double pml_alpha[3];
  const double* cp;      /* sound_speed_compression map or NULL (then cp_scalar) */
  double cp_scalar;
  const double* cs;      /* sound_speed_shear map or NULL (then cs_scalar) */
  double cs_scalar;
  const double* rho0;    /* density map or NULL (then rho0_scalar) */
  double rho0_scalar;
  double chi_p;          /* Kelvin-Voigt factors (elastic.jl:193-210); 0 where the reference */
  double chi_s;          /*   applies none (3-D, or any heterogeneous Lame parameter) */
  const double* p0;      /* normal stresses start at -p0 (elastic.jl:217-222, 458-464) or NULL */
  /* velocity source at u_mask (elastic.jl:248-275): ux, uy, uz are n_usrc x ncols */
  int64_t n_usrc;
  const int64_t* usrc_index;      /* findall(u_mask), 0-based */
  const double* usrc_signal[3];
  int64_t usrc_ncols[3];
  int32_t u_mode;                 /* kw_source_mode (Dirichlet or additive) */
  /* stress source at s_mask (elastic.jl:312-345): sxx syy szz sxy sxz syz */
  int64_t n_ssrc;
  const int64_t* ssrc_index;      /* findall(s_mask), 0-based */
  const double* ssrc_signal[6];
  int64_t ssrc_ncols[6];
  int32_t s_mode;
  /* sensor (record_sensor_data!, first_order_steps.jl) */
  int64_t n_sensor;
  const int64_t* sensor_index;    /* findall(mask), 0-based; NULL with n_sensor>0 = Cartesian */
  uint32_t record;                /* kw_record bits */
} kw_elastic_desc;
int32_t kw_elastic_create(kw_ctx* ctx, const kw_elastic_desc* desc, kw_elastic** out);
int32_t kw_elastic_run(kw_elastic* sim, int64_t n_steps);
int32_t kw_elastic_synchronize(kw_elastic* sim);
int32_t kw_elastic_steps_done(const kw_elastic* sim, int64_t* steps);
/* which: kw_elastic_field; the full grid, column-major */
int32_t kw_elastic_get_field(kw_elastic* sim, int32_t which, double* out);
/* p(t) of the pressure -tr(s)/ndim over the whole grid (progress callbacks) */
int32_t kw_elastic_get_pressure(kw_elastic* sim, double* out);
int32_t kw_elastic_sensor_shape(kw_elastic* sim, uint32_t field, int64_t* rows, int64_t* cols);
int32_t kw_elastic_get_sensor(kw_elastic* sim, uint32_t field, double* out);
/* Timing support as kw_sim_profile: per-pass-class mean device time (ms) over n_steps. */
int32_t kw_elastic_profile(kw_elastic* sim, int64_t n_steps, int32_t max_classes, int32_t* n_classes,
                           const char** names, double* ms, int64_t* launches);
int32_t kw_elastic_destroy(kw_elastic* sim);

#ifdef __cplusplus
}
#endif

#endif /* KWAVE_HIP_H */
