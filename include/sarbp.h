/*
 * sarbp.h -- C ABI of the MI355X (gfx950) time-domain back-projection imager.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns a
 * sarbp_status and never exits or throws.  Device pointers are caller-owned
 * HIP device allocations; `stream` is a hipStream_t passed as void* (NULL is
 * the null stream).  Work is enqueued asynchronously on that stream unless the
 * function says otherwise.
 *
 * Reference interface replaced (reference repository wenj-yan/
 * SAR_CambrianAccelerationCard, code/SAR_BP/CUDA):
 *   void launchBackProjection(cufftComplex* Srnm, cufftComplex* Image,
 *                             const SARParams params, int Na, int Nr);
 *     -- BPA_CUDA_kernels.cu:91-99, declared at BPA_CUDA.cpp:8-9, called from
 *        SAR_Processor::applyBPAlgorithm (BPA_CUDA.cpp:83-90).
 *   struct SARParams -- BPA_CUDA.h:28-50.
 *   CHECK_CUDA (print + exit) -- BPA_CUDA.h:11-17, replaced by status codes.
 * The arithmetic follows the BP loops of code/SAR_BP/matlab/BPA_SAR.m:133-153
 * and :178-197 (pixel value = sum over pulses of the range-compressed sample at
 * the pixel's slant range times exp(+j 4 pi fc R / c)), not the CUDA kernel,
 * whose range does not depend on the pulse (BPA_CUDA_kernels.cu:56-64).
 */
#ifndef SARBP_H
#define SARBP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SARBP_ABI_VERSION 4

typedef enum sarbp_status {
  SARBP_OK = 0,
  SARBP_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size or enum value    */
  SARBP_ERR_OUT_OF_MEMORY = 2,    /* a device or host allocation failed      */
  SARBP_ERR_OVER_BUDGET = 3,      /* job can never fit the server HBM budget */
  SARBP_ERR_DEVICE = 4,           /* a HIP runtime call failed               */
  SARBP_ERR_UNSUPPORTED = 5,      /* valid request this build cannot run     */
  SARBP_ERR_NOT_READY = 6,        /* job still queued or running             */
  SARBP_ERR_NOT_FOUND = 7,        /* unknown job id                          */
  SARBP_ERR_CANCELLED = 8,        /* job was cancelled                       */
  SARBP_ERR_TIMEOUT = 9,          /* wait or watchdog expired                */
  SARBP_ERR_SHUTDOWN = 10         /* server is shutting down                 */
} sarbp_status;

/* Interleaved complex64: the layout of cufftComplex / std::complex<float>. */
typedef struct sarbp_cf32 {
  float re, im;
} sarbp_cf32;

typedef enum sarbp_interp {
  SARBP_INTERP_NEAREST = 0, /* MATLAB round() of the bin (BPA_SAR.m:141,184) */
  SARBP_INTERP_LINEAR = 1,  /* two-tap linear                                */
  SARBP_INTERP_SINC8 = 2    /* 8-tap Kaiser-windowed sinc                    */
} sarbp_interp;

typedef enum sarbp_precision {
  /* FP64 per-(tile, pulse) reference range and phase, FP32 per-pixel
   * differential with exact modular phase reduction.  Default.  The FP32
   * differential bounds a tile's radius (48 m at 5 km, more at longer
   * range); a job whose grid no tile shape fits runs SARBP_PREC_DF. */
  SARBP_PREC_REF = 0,
  SARBP_PREC_F64 = 1, /* FP64 slant range and phase for every pixel-pulse    */
  SARBP_PREC_F32 = 2, /* FP32 positions, range and phase (inaccurate)         */
  /* FP64 per-(tile, pulse) reference plus a compensated double-float
   * (FP32 pair) per-pixel differential range and phase: no tile-radius
   * limit, about half the REF rate. */
  SARBP_PREC_DF = 3
} sarbp_precision;

typedef enum sarbp_dtype {
  SARBP_DTYPE_C64 = 0,   /* profiles as complex64 (8 B / sample)             */
  SARBP_DTYPE_BF16X2 = 1 /* profiles as bf16 re, bf16 im (4 B / sample; re in
                            the low half-word)                               */
} sarbp_dtype;

/* Image grid: pixel (ix, iy) sits at origin + ix*ux + iy*uy (metres, FP64).
 * The image is stored row-major [ny][nx], ix fastest.  In the reference the
 * rows are azimuth and the columns range (BPA_CUDA_kernels.cu:53-54). */
typedef struct sarbp_grid {
  double origin[3];
  double ux[3];
  double uy[3];
  int32_t nx, ny;
} sarbp_grid;

/* Radar / processing parameters (the SARParams of BPA_CUDA.h:28-50, reduced to
 * what BP reads, in FP64). */
typedef struct sarbp_params {
  double c;           /* propagation speed, m/s                               */
  double fc;          /* carrier frequency, Hz (phase = 4 pi fc R / c)        */
  double r0;          /* slant range of range sample 0, m                     */
  double dr;          /* range sample spacing, m                              */
  double kaiser_beta; /* Kaiser window shape for SARBP_INTERP_SINC8           */
  int32_t n_pulses;   /* profiles rows, positions rows                        */
  int32_t n_samples;  /* samples per profile                                  */
  int32_t interp;     /* sarbp_interp                                         */
  int32_t precision;  /* sarbp_precision                                      */
  int32_t dtype;      /* sarbp_dtype                                          */
  int32_t accumulate; /* 0: image overwritten; 1: sum added into the image    */
} sarbp_params;

typedef struct sarbp_context sarbp_context;

/* Human-readable text for a status code (static storage, never NULL). */
const char* sarbp_strerror(sarbp_status status);

/* ABI version of the loaded library (SARBP_ABI_VERSION it was built with). */
int32_t sarbp_abi_version(void);

/* Creates a context bound to HIP device `device`.  Kernel tables are uploaded
 * lazily on that device.  A context may be used from several host threads; calls
 * are serialised internally. */
sarbp_status sarbp_create(sarbp_context** ctx, int32_t device);
void sarbp_destroy(sarbp_context* ctx);

/*
 * Back-projects one image.  Replaces launchBackProjection (BPA_CUDA_kernels.cu:91).
 *   d_profiles : [n_pulses][n_samples] range-compressed pulses (dtype), device
 *   d_pos      : [n_pulses][3] FP64 antenna phase-centre positions, device
 *   grid       : host struct
 *   params     : host struct
 *   d_image    : [ny][nx] complex64, device; overwritten or accumulated into
 * Range sample k of pulse m is at slant range r0 + k*dr; samples outside
 * [0, n_samples) read as zero.  Asynchronous on `stream`.
 */
sarbp_status sarbp_backproject(sarbp_context* ctx, const void* d_profiles,
                               const double* d_pos, const sarbp_grid* grid,
                               const sarbp_params* params, sarbp_cf32* d_image,
                               void* stream);

/* Same, then waits for the stream (the synchronous behaviour of
 * launchBackProjection, BPA_CUDA_kernels.cu:97-98). */
sarbp_status sarbp_backproject_sync(sarbp_context* ctx, const void* d_profiles,
                                    const double* d_pos, const sarbp_grid* grid,
                                    const sarbp_params* params,
                                    sarbp_cf32* d_image, void* stream);

/* ---- reference-compatible entry point ------------------------------------------ */

/* The reference's SARParams (code/SAR_BP/CUDA/BPA_CUDA.h:28-50), field for
 * field, with its defaults documented there; lambda = c / fc. */
typedef struct sarbp_sar_params {
  float c, fc, lambda, Kr, Tr, Br, Fr, vr, H, La, Rg0, Az0;
  int32_t Nfast, Mslow, Ntarget;
} sarbp_sar_params;

/*
 * Drop-in for launchBackProjection(Srnm, Image, params, Na, Nr)
 * (BPA_CUDA_kernels.cu:91-99): same arguments (device Srnm [Mslow][Nfast] and
 * Image [Na][Nr], params by value), synchronous like the original
 * (cudaDeviceSynchronize at :97-98), but a status code instead of exit().
 * Geometry as the reference's host driver builds it (BPA_CUDA.cpp:35-48): pulse
 * m at (vr ta_m, 0, H), ta = linspace(-La/(2 vr), La/(2 vr), Mslow); range
 * sample k at c k / (2 Fr).  Image row i / column j is the ground point
 * (Az0 + i - Na/2, Rg0 + j - Nr/2, 0), 1 m pixels as the reference kernel's
 * x = j - Nr/2, y = i - Na/2 (BPA_CUDA_kernels.cu:56-57).  The sum follows
 * BPA_SAR.m:133-153 (nearest bin, exp(+j 4 pi R / lambda)); the reference
 * kernel's pulse-invariant range is not reproduced.  Uses a per-device
 * default context; the current HIP device.
 */
sarbp_status sarbp_launch_backprojection(sarbp_cf32* Srnm, sarbp_cf32* Image,
                                         sarbp_sar_params params, int32_t Na,
                                         int32_t Nr);

/*
 * Drop-in for launchGenerateEcho(Srnm, Ptarget, ta, tr, params, Mslow, Nfast)
 * (BPA_CUDA_kernels.cu:79-89, kernel generateEchoKernel :3-45): device Srnm
 * [Mslow][Nfast] (written), device Ptarget [Ntarget][3] (azimuth, ground
 * range, sigma), device ta [Mslow] and tr [Nfast] (slow and fast time, s);
 * params by value (c, lambda, Kr, Tr, vr, H, La and Ntarget are used).
 * Sample (i, j) = sum over targets k with 0 < t < Tr and |Az| < La/2 of
 * sigma_k exp(j (pi Kr t^2 - 4 pi R / lambda)), Az = ta[i] vr - x_k,
 * R = sqrt(Az^2 + y_k^2 + H^2), t = tr[j] - 2 R / c.  Range, gate and phase
 * are evaluated in FP64 (the reference's FP32 phase, ~1e6 rad, carries
 * ~0.1 rad of error).  Synchronous like the original (null stream, then
 * hipDeviceSynchronize), a status code instead of exit(); the current HIP
 * device.
 */
sarbp_status sarbp_launch_generate_echo(sarbp_cf32* Srnm, const float* Ptarget,
                                        const float* ta, const float* tr,
                                        sarbp_sar_params params, int32_t Mslow,
                                        int32_t Nfast);

/* ---- plans ------------------------------------------------------------------ */

/*
 * A plan fixes the launch configuration for one geometry.  sarbp_backproject
 * plans from the grid alone (range window sized by the triangle inequality,
 * second-order range expansion); a plan made from the host copy of the
 * antenna positions bounds the line-of-sight directions and the minimum slant
 * range over the whole scene, which gives shorter LDS windows and, where the
 * bound allows it, the cheaper first-order range expansion.  The plan stays
 * valid for any positions inside the bounds computed at creation (the device
 * positions passed to sarbp_plan_execute must be the ones planned for).
 */
typedef struct sarbp_plan sarbp_plan;

typedef struct sarbp_plan_info {
  int32_t tiles_x, tiles_y;     /* workgroups of the plain launch            */
  int32_t tile_nx, tile_ny;     /* pixels per tile                           */
  int32_t window;               /* LDS range window per pulse, samples       */
  int32_t pulses_per_stage;     /* pulses staged per barrier interval        */
  int32_t rows_per_lane, cols_per_lane;
  int32_t range_order;          /* 1 if the first-order expansion holds over
                                   the whole scene, else 0 (chosen per stage) */
  int32_t lds_bytes;            /* dynamic LDS per workgroup                 */
  double max_phase_err_rad;     /* bound on the expansion truncation error   */
  int32_t sinc8_halves;         /* SINC8 kernel: 2 = half-bin degree-5 tap
                                   polynomials, 1 = whole-bin degree 7; 0 for
                                   other interpolators (chosen per launch)   */
  int32_t precision;            /* arithmetic the kernels run: the requested
                                   sarbp_precision, or SARBP_PREC_DF for a
                                   SARBP_PREC_REF grid too coarse for its
                                   tile-radius limit                        */
} sarbp_plan_info;

sarbp_status sarbp_plan_create(sarbp_context* ctx, const double* h_pos,
                               const sarbp_grid* grid,
                               const sarbp_params* params, sarbp_plan** plan);
sarbp_status sarbp_plan_execute(sarbp_plan* plan, const void* d_profiles,
                                const double* d_pos, sarbp_cf32* d_image,
                                void* stream);
sarbp_status sarbp_plan_get_info(const sarbp_plan* plan, sarbp_plan_info* info);
void sarbp_plan_destroy(sarbp_plan* plan);

/* Detail text of the last error raised on the calling thread ("" if none). */
const char* sarbp_last_error_detail(void);

/* ---- launch tuning -------------------------------------------------------- */

/* Kernel configuration knobs chosen by the autotuner.  A workgroup's tile is
 * (64 cols_per_lane) x (4 rows_per_lane) pixels: each lane accumulates
 * cols_per_lane columns (64 apart) of rows_per_lane rows.  pulses_per_stage
 * is the number of pulses staged in LDS per barrier interval, an upper bound:
 * the planner reduces it until the stage fits 64 KiB of LDS, then by one
 * pulse at a time while one pulse fewer lets one more workgroup share a CU
 * (a job's window sets the LDS per pulse), down to 4 pulses for a plain
 * launch (sarbp_plan_get_info reports the depth planned) and to 1 for a job
 * of a batch or a server, whose persistent launch gains more from occupancy.
 * The depth changes no bit of an image.  Shapes: rows 2, 4 or 8 with
 * cols 1 or 2; rows 2 or 4 with cols 4 (SARBP_PREC_REF); cols 1 otherwise. */
typedef struct sarbp_tuning {
  int32_t rows_per_lane;
  int32_t cols_per_lane;
  int32_t pulses_per_stage;
} sarbp_tuning;

/* Size classes of the tuning cache.  A job's class is set by the range
 * window w (samples per pulse staged in LDS) that a tile of 64 x 32 pixels
 * needs, as the planner bounds it -- from the host positions when the call
 * has them, else from the grid alone: class 0 for w < 80, 1 for w < 120, 2
 * for w < 176, 3 for w < 272, 4 above.  w grows with the pixel spacing over
 * dr, with the aperture and at close range (X band at 5 km: 0.25 / 0.35 /
 * 0.5 / 0.7 / 1 m pixels fall in classes 0-4; 0.25 m pixels under an
 * 8192-pulse aperture in class 1).  It sets a stage's LDS footprint per pulse
 * and the sinc8 precompute per pixel, so each class has its own best launch
 * configuration. */
#define SARBP_SIZE_CLASSES 5

/* The size class of a job (host only); h_pos: optional host copy of the
 * [n_pulses][3] positions (planned bound).  -1 for invalid arguments. */
int32_t sarbp_size_class(const sarbp_grid* grid, const sarbp_params* params,
                         const double* h_pos);

/* Sets the configuration used by later calls with this (interp, precision,
 * dtype) triple on this context, for every size class (replacing per-class
 * entries); all-zero restores the built-in per-class defaults.
 * With SARBP_PREC_REF the tile shape is an upper bound: a call whose grid
 * would put tile pixels more than 48 m from the tile centre (FP32
 * differential range) runs with the largest compiled shape within it that
 * fits (plan info reports the shape used), in a batch too.  Only a grid
 * too coarse for every shape is refused
 * (SARBP_ERR_UNSUPPORTED). */
sarbp_status sarbp_set_tuning(sarbp_context* ctx, int32_t interp,
                              int32_t precision, int32_t dtype,
                              const sarbp_tuning* tuning);
/* The configuration in effect for size class 0. */
sarbp_status sarbp_get_tuning(sarbp_context* ctx, int32_t interp,
                              int32_t precision, int32_t dtype,
                              sarbp_tuning* tuning);
/* Per size class: set (all-zero removes the class's entry) and get the
 * configuration in effect (the class's entry, else the all-class one, else
 * the built-in default) for calls that give host positions (plans, batch
 * jobs and server requests with h_pos).  Calls without them (the unplanned
 * sarbp_backproject, plans and jobs with h_pos NULL) plan every window from
 * the grid alone -- a disc of the tile's radius, other classes and other
 * best tunings -- and keep their own per-class entries: the built-in
 * position-free defaults, or what sarbp_autotune without h_pos stored; an
 * all-class sarbp_set_tuning applies to both. */
sarbp_status sarbp_set_class_tuning(sarbp_context* ctx, int32_t interp,
                                    int32_t precision, int32_t dtype,
                                    int32_t size_class, const sarbp_tuning* tuning);
sarbp_status sarbp_get_class_tuning(sarbp_context* ctx, int32_t interp,
                                    int32_t precision, int32_t dtype,
                                    int32_t size_class, sarbp_tuning* tuning);

/* Result of sarbp_autotune. */
typedef struct sarbp_autotune_result {
  sarbp_tuning best;      /* stored as this context's tuning               */
  double best_ms;         /* per launch on the timing crop                 */
  double default_ms;      /* the built-in default on the same crop         */
  int32_t n_candidates;   /* configurations timed                          */
  int32_t crop_nx, crop_ny;
} sarbp_autotune_result;

/*
 * Project autotuner.  Times every (tile shape, pulses_per_stage)
 * configuration whose LDS footprint fits, on a centred crop of the given job
 * (all pulses, the plan the optional host positions allow) that keeps up to
 * max(max_crop_px, 4096) columns and holds about max_crop_px^2 pixels -- enough
 * tiles to fill the device -- with HIP events on `stream`, and stores the
 * fastest as this context's tuning for the job's (interp, precision, dtype).
 * d_scratch receives the crop images and must hold scratch_pixels >=
 * crop_nx * crop_ny complex64 values.  The built-in default of the job's
 * size class stays unless a candidate is more than 2% faster.  The result is
 * stored for the job's size class (sarbp_size_class) only, for calls with
 * host positions if h_pos is given, else for calls without (see
 * sarbp_set_class_tuning).  Synchronous.
 */
sarbp_status sarbp_autotune(sarbp_context* ctx, const void* d_profiles,
                            const double* d_pos, const double* h_pos,
                            const sarbp_grid* grid, const sarbp_params* params,
                            sarbp_cf32* d_scratch, size_t scratch_pixels,
                            int32_t max_crop_px, void* stream,
                            sarbp_autotune_result* result);

/* ---- batched launch --------------------------------------------------------- */

typedef struct sarbp_job {
  const void* d_profiles;
  const double* d_pos;
  const double* h_pos; /* optional host copy of the positions (tighter plan)  */
  sarbp_grid grid;
  sarbp_params params;
  sarbp_cf32* d_image;
} sarbp_job;

/*
 * Back-projects n_jobs independent images of arbitrary, uneven sizes in one
 * persistent launch per tile shape: a device work list of (job, tile) items
 * is drained by a fixed grid of workgroups.  Each job uses the tile shape a
 * plain launch of it would (sarbp_set_tuning); jobs whose grids need a
 * smaller shape form their own launch, queued on the same stream.  All jobs
 * must share interp, precision and dtype.  Each job's pixel sums are formed
 * in a fixed pulse order, so a job's image is bitwise identical to its plain
 * launch whatever else is in the batch.  The work list lives in
 * context-owned device memory.
 */
sarbp_status sarbp_backproject_batch(sarbp_context* ctx, const sarbp_job* jobs,
                                     int32_t n_jobs, void* stream);

/*
 * sarbp_backproject_batch without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same jobs, kernels and results -- each image bitwise that of
 * sarbp_backproject_batch -- but the work list lives in a fixed device array
 * of the library (one per GPU, kept until the process ends; its address is
 * looked up by sarbp_create), written by one host-to-device copy on `stream`
 * from a static host staging area, and the persistent kernels are launched
 * directly instead of from captured graphs.  Calls no hipMalloc, hipFree or
 * other allocating routine.  Limits: at most 256 jobs per call (more:
 * SARBP_ERR_INVALID_ARGUMENT) and 8 tile shapes (SARBP_ERR_UNSUPPORTED).
 * Every call on one GPU shares that array: calls whose work overlaps on the
 * device (other streams, or other host threads) race and must be serialised
 * by the caller; calls on one stream are ordered.
 */
sarbp_status sarbp_backproject_batch_noalloc(sarbp_context* ctx, const sarbp_job* jobs,
                                             int32_t n_jobs, void* stream);

/* ---- serving --------------------------------------------------------------------
 *
 * A server owns one device.  Requests are queued on submit; an admission
 * thread moves them, in arrival order, into a device arena of
 * hbm_budget_bytes (profiles, positions and image resident in HBM) as memory
 * and job slots free up; a scheduler thread runs epochs: one batched
 * persistent launch (replayed from a captured hipGraph) in which every
 * resident job back-projects its next chunk of pulses_per_epoch pulses.  Jobs
 * admitted while others are mid-flight join the next epoch (continuous
 * batching); a completion thread copies each finished image to the caller's
 * host buffer and frees its arena space.  A job's pixel sums are formed
 * chunk by chunk in a fixed order, so its image is bitwise the same whatever
 * else is in flight.  No counterpart in the reference, which images one
 * simulated scene per process (code/SAR_BP/CUDA/main.cpp:3-19).
 */
typedef struct sarbp_server sarbp_server;

typedef struct sarbp_server_config {
  int32_t device;
  int32_t interp, precision, dtype; /* shared by every job of the server      */
  size_t hbm_budget_bytes;          /* device arena for resident jobs         */
  int32_t max_active_jobs;          /* resident jobs at once (0: no limit)    */
  int32_t pulses_per_epoch;         /* pulse chunk per job per epoch (0: all) */
  int32_t rows_per_lane;            /* launch tuning for every size class (0:
                                       the built-in per-class defaults, or
                                       sarbp_server_set_class_tuning)         */
  int32_t cols_per_lane;
  int32_t pulses_per_stage;
  double epoch_watchdog_s;          /* epochs slower than this fail their jobs
                                       with SARBP_ERR_TIMEOUT (0: off)        */
} sarbp_server_config;

/* Host buffers must stay valid until the job is done, failed or cancelled. */
typedef struct sarbp_request {
  const void* h_profiles; /* [n_pulses][n_samples], the server's dtype     */
  const double* h_pos;    /* [n_pulses][3]                                  */
  sarbp_grid grid;
  sarbp_params params;    /* interp / precision / dtype = the server's      */
  sarbp_cf32* h_image;    /* [ny][nx] output                                */
} sarbp_request;

typedef enum sarbp_job_state {
  SARBP_JOB_QUEUED = 0,
  SARBP_JOB_RESIDENT = 1, /* data in HBM, no epoch started yet */
  SARBP_JOB_RUNNING = 2,
  SARBP_JOB_DONE = 3,
  SARBP_JOB_FAILED = 4,
  SARBP_JOB_CANCELLED = 5
} sarbp_job_state;

typedef struct sarbp_job_info {
  int32_t state;        /* sarbp_job_state                                   */
  int32_t status;       /* sarbp_status of a failed job, else SARBP_OK       */
  int32_t epochs;       /* epochs the job took part in                       */
  int32_t reserved;
  double submit_s, admit_s, first_epoch_s, done_s; /* server clock, seconds  */
  double work;          /* pixel-pulses                                      */
} sarbp_job_info;

typedef struct sarbp_server_stats {
  int64_t submitted, completed, failed, cancelled, epochs;
  int32_t queued, active;
  double work_done;     /* pixel-pulses of completed jobs                    */
  double gpu_busy_s;    /* summed device time of the epochs                  */
  double uptime_s;
  size_t hbm_in_use, hbm_peak;
} sarbp_server_stats;

sarbp_status sarbp_server_create(const sarbp_server_config* cfg, sarbp_server** server);
/* Launch configuration of later submissions of one size class (for example
 * a context's autotuned sarbp_get_class_tuning); all-zero restores the
 * built-in default of the class. */
sarbp_status sarbp_server_set_class_tuning(sarbp_server* server, int32_t size_class,
                                           const sarbp_tuning* tuning);
/* Rejects requests that can never fit the arena with SARBP_ERR_OVER_BUDGET. */
sarbp_status sarbp_server_submit(sarbp_server* server, const sarbp_request* req,
                                 uint64_t* job_id);
/* A request that carries raw echoes instead of range-compressed profiles.
 * The server copies the raw echoes (n_raw samples per pulse instead of n_out)
 * and range-compresses them in HBM (sarbp_range_compress) before imaging.
 * params.n_samples, r0 and dr are set from `range` (sarbp_range_info);
 * range.dtype is set to the server's dtype.  Declared further down. */
typedef struct sarbp_raw_request sarbp_raw_request;
sarbp_status sarbp_server_submit_raw(sarbp_server* server, const sarbp_raw_request* req,
                                     uint64_t* job_id);
sarbp_status sarbp_server_poll(sarbp_server* server, uint64_t job_id,
                               sarbp_job_info* info);
/* Waits until the job is done, failed or cancelled (SARBP_ERR_TIMEOUT after
 * timeout_s; timeout_s < 0 waits forever).  A failed job returns its failure
 * status, and sarbp_last_error_detail then names the call that failed it. */
sarbp_status sarbp_server_wait(sarbp_server* server, uint64_t job_id,
                               double timeout_s, sarbp_job_info* info);
sarbp_status sarbp_server_cancel(sarbp_server* server, uint64_t job_id);
/* Forgets a finished job's record (SARBP_ERR_NOT_READY while it runs). */
sarbp_status sarbp_server_release(sarbp_server* server, uint64_t job_id);
sarbp_status sarbp_server_get_stats(sarbp_server* server, sarbp_server_stats* stats);
/* Waits until no job is queued or active (SARBP_ERR_TIMEOUT after timeout_s). */
sarbp_status sarbp_server_drain(sarbp_server* server, double timeout_s);
/* Stops accepting work, cancels queued jobs, waits for in-flight epochs. */
void sarbp_server_destroy(sarbp_server* server);

/* ---- range compression ---------------------------------------------------------
 *
 * The step before BP (code/SAR_BP/matlab/BPA_SAR.m:93-100 matched filter,
 * :106-111 upsampling by zero insertion): raw LFM echoes -> range-compressed,
 * L-times upsampled, cropped profiles that sarbp_backproject consumes.  The
 * reference does this in MATLAB only (its CUDA driver feeds raw echoes to the
 * BP kernel, BPA_CUDA.cpp:70-84); there is no compiled counterpart.
 *
 * Per pulse, with nf = smallest power of two >= n_raw + ceil(Tp fs):
 *   h[n]   = exp(j pi Kr (n/fs - Tp/2)^2) for 0 < n/fs < Tp (and n < n_raw),
 *            else 0; Kr = B / Tp (chirp centred at baseband)
 *   S      = FFT_nf(raw zero-padded to nf);  Y = S . conj(FFT_nf(h))
 *   Z      = Y with bins [0, (nf+1)/2) kept at the front and the rest moved to
 *            the end of an L*nf spectrum (zeros between)
 *   out[i] = IFFT_{L nf}(Z)[crop_offset + i] / nf,   i in [0, n_out)
 * The 1/nf scale keeps the compressed amplitude (a unit chirp compresses to
 * Tp*fs at its delay) and makes out[L q - crop_offset] the unupsampled
 * compressed sample q.  Output sample i lies at slant range r0 + i*dr with
 * r0 = c/2 (t0 + crop_offset / (L fs)), dr = c / (2 L fs).
 */
typedef struct sarbp_range_params {
  double c;             /* propagation speed, m/s                          */
  double bandwidth;     /* chirp bandwidth B, Hz                           */
  double pulse_width;   /* chirp duration Tp, s                            */
  double fs;            /* raw complex sampling rate, Hz                   */
  double t0;            /* fast time of raw sample 0, s (range axis only)  */
  int32_t n_raw;        /* raw samples per pulse                           */
  int32_t upsample;     /* L, 1..16                                        */
  int32_t crop_offset;  /* first upsampled sample kept, >= 0               */
  int32_t n_out;        /* samples kept per pulse, crop_offset + n_out <= L nf */
  int32_t dtype;        /* sarbp_dtype of the output profiles              */
  int32_t reserved;     /* 0                                               */
} sarbp_range_params;

/* FFT size nf and output range axis (r0, dr) of a parameter set; validates it
 * (nf <= 8192 on the device path).  Host-only; NULL outputs are skipped. */
sarbp_status sarbp_range_info(const sarbp_range_params* rp, int32_t* nf,
                              double* r0, double* dr);

/* Range-compresses d_raw [n_pulses][n_raw] complex64 into d_profiles
 * [n_pulses][n_out] (complex64, or bf16x2 words rounded to nearest even).
 * One workgroup per pulse, all FFTs in LDS, FP32 arithmetic with FP64-derived
 * tables.  The tables are built once per parameter set and kept by the
 * context. */
sarbp_status sarbp_range_compress(sarbp_context* ctx, const sarbp_cf32* d_raw,
                                  int32_t n_pulses, const sarbp_range_params* rp,
                                  void* d_profiles, void* stream);

/* Served job from raw echoes (sarbp_server_submit_raw).  Host buffers must
 * stay valid until the job is done, failed or cancelled. */
struct sarbp_raw_request {
  const sarbp_cf32* h_raw;   /* [n_pulses][range.n_raw] complex64 raw echoes */
  sarbp_range_params range;  /* their range compression                     */
  const double* h_pos;       /* [n_pulses][3]                                */
  sarbp_grid grid;
  sarbp_params params;       /* as sarbp_request; n_samples, r0, dr ignored  */
  sarbp_cf32* h_image;       /* [ny][nx] output                              */
};

/* ---- omega-K image formation ---------------------------------------------------
 *
 * The reference's second imaging path (code/SAR_WK/wk_v5.2.3, process_wk at
 * wk.cc:43-343; MATLAB prototype code/SAR_WK/matlab/mian_wka_sinc.m).  Rows are
 * azimuth lines (Na), columns range samples (Nr), row-major complex64.
 *   fr[n] = fftshift(-Nr/2 : Nr/2-1)[n] * fs / Nr           (wk.cc:197-208)
 *   fa[m] = f_nc + fftshift(-Na/2 : Na/2-1)[m] * prf / Na   (wk.cc:211-221)
 *   s     = sqrt(max(0, (f0 + fr)^2 - c^2 fa^2 / (4 vr^2)))
 * Stolt interpolation (stolt.cc:324-727): output (m, n) is the sum over the taps
 * idx of in(m, idx) sinc(cen - idx), sinc(x) = sin(pi x)/(pi x), sinc(0) = 1,
 * at cen = n - (s - f0 - fr[n]) Nr / fs.
 */
typedef enum sarbp_stolt_window {
  /* idx = floor(cen) + p, p = -P/2 .. P/2-1, each clamped to [0, Nr-1] with the
   * distance taken to the clamped index: the MLU code (stolt.cc:522-583). */
  SARBP_STOLT_FLOOR = 0,
  /* every idx in [0, Nr-1] with |cen - idx| <= P/2: the MATLAB prototype
   * (mian_wka_sinc.m:98-111). */
  SARBP_STOLT_CENTERED = 1
} sarbp_stolt_window;

/* Stolt interpolation along range frequency of d_in [na][nr] into d_out (may be
 * d_in).  Replaces StoltInterp_sinc (code/SAR_WK/wk_v5.2.3/task/stolt.cc:324,
 * declared task/task.h:37-44): h_fr_axis [nr] and h_fa_axis [na] are the
 * range- and azimuth-frequency axes (host, Hz), fs the range sampling rate
 * (its argument `fr`), taps = P (even, 2..16), window a sarbp_stolt_window;
 * nr <= 8192.  The axes are staged before the call returns (the host arrays
 * may then be reused); a call whose axes equal the previous call's on this
 * context only enqueues the kernel, any other first waits for the previous
 * call's kernel to finish.
 * One workgroup per azimuth line: the line is staged in LDS, the mapping is
 * evaluated in FP64, the taps in FP32. */
sarbp_status sarbp_stolt_sinc(sarbp_context* ctx, const sarbp_cf32* d_in, sarbp_cf32* d_out,
                              const double* h_fr_axis, const double* h_fa_axis, double f0, double c,
                              double vr, int32_t na, int32_t nr, double fs, int32_t taps,
                              int32_t window, void* stream);

/* Range reference of sarbp_omega_k's reference function. */
typedef enum sarbp_wk_range_ref {
  /* The reference's pipeline (compression.cc:17-58): the range spectrum keeps
   * the DFT's time origin at padded sample 0 and the reference function uses
   * r0 as given.  It focuses a target only when the two-way time of padded
   * sample 0 is a whole number of padded range windows (Nr / fs) after
   * 2 r0 / c and the target lies within Nr/2 samples after padded sample 0;
   * a target in the far half of the padded window defocuses. */
  SARBP_WK_RANGE_REF_R0 = 0,
  /* No counterpart in the reference: r0 is the slant range of raw echo sample
   * 0 (padded sample 0 at fast time ts = 2 r0/c - pad/fs, pad the leading
   * zero samples).  The range spectrum is moved to absolute fast time (times
   * exp(-j 2 pi fr ts)), the reference function uses the range of padded
   * sample floor(Nr/2), and after the Stolt step the origin goes back to
   * padded sample 0 (times exp(-j 2 pi fr floor(Nr/2) / fs)): every target in
   * the padded window focuses, at its own padded range sample, for any r0. */
  SARBP_WK_RANGE_REF_CENTRE = 1
} sarbp_wk_range_ref;

/* The reference's other Stolt step, by grid sampling: performStoltInterpolation
 * (code/SAR_WK/wk_v5.2.3/task/stolt.cc:38-318, declared task/task.h:27-34;
 * its call in process_wk, wk.cc:272, is commented out), on device
 * d_in [na][nr] into d_out (may be d_in).  Each line is rotated by nr/2
 * (column w holds input (w + nr - nr/2) mod nr) and output n is the bilinear,
 * zero-padded, align-corners grid sample of it at
 * x = (sqrt((f0 + fr_axis[n])^2 - c^2 fa_axis[m]^2 / (4 vr^2)) - f0) / (fr nr / 2),
 * on its own row: linear interpolation along the line at
 * ((x + 1) / 2) (nr - 1), zeros outside.  The mapping is FP64 rounded to
 * FP32 and the grid arithmetic FP32, as the reference's.  fr is the
 * frequency step the reference is called with (Fr / Nr); the output is not
 * rotated back.  Axes as for sarbp_stolt_sinc; nr <= 8192. */
sarbp_status sarbp_stolt_gridsample(sarbp_context* ctx, const sarbp_cf32* d_in,
                                    sarbp_cf32* d_out, const double* h_fr_axis,
                                    const double* h_fa_axis, double f0, double c,
                                    double vr, int32_t na, int32_t nr, double fr,
                                    void* stream);

typedef struct sarbp_wk_params {
  double c;       /* propagation speed, m/s                                   */
  double f0;      /* carrier frequency, Hz                                    */
  double fs;      /* range sampling rate Fr, Hz                               */
  double prf;     /* azimuth sampling rate, Hz                                */
  double kr;      /* signed chirp FM rate, Hz/s (negative for a down-chirp:
                     the paras.txt convention, read.cc:54-82)                 */
  double vr;      /* effective radar velocity, m/s                            */
  double r0;      /* reference slant range, m                                 */
  double f_nc;    /* Doppler centroid, Hz                                     */
  int32_t taps;   /* Stolt taps P (6 in the reference, wk.cc:273)             */
  int32_t window; /* sarbp_stolt_window                                       */
  int32_t range_ref; /* sarbp_wk_range_ref (SARBP_WK_RANGE_REF_R0: reference) */
  int32_t reserved;  /* must be 0                                             */
} sarbp_wk_params;

/* Padded size of an na_in x nr_in echo: round(na_in/6) zero lines before and
 * after, round(nr_in/3) zero samples before and after each line
 * (code/SAR_WK/wk_v5.2.3/task/preprocess.cc:21-24).  Host-only. */
sarbp_status sarbp_wk_padded_size(int32_t na_in, int32_t nr_in, int32_t* na, int32_t* nr);

/* The steps of process_wk one by one, for callers that drive them as the
 * reference's task API does (task/task.h:46-73); sarbp_omega_k runs them all
 * in one pass.  Device buffers, asynchronous on `stream`.
 *
 * preprocess_wk (preprocess.cc:17-166): d_echo [na_in][nr_in] -> d_out
 * [na][nr] (the padded size): zeros around the echo, line m times
 * exp(-j 2 pi f_nc ta_m), ta_m = (m - na/2) / prf. */
sarbp_status sarbp_wk_preprocess(sarbp_context* ctx, const sarbp_cf32* d_echo,
                                 int32_t na_in, int32_t nr_in, double prf,
                                 double f_nc, sarbp_cf32* d_out, void* stream);

/* perform2DFFT / perform2DIFFT (fft.cc:18-359): the 2-D DFT of d_data
 * [na][nr] in place, natural order; inverse != 0: the inverse transform
 * scaled by 1/(na nr).  Plans are kept by the context; calls on one context
 * are ordered on the device like sarbp_omega_k's. */
sarbp_status sarbp_wk_fft2(sarbp_context* ctx, sarbp_cf32* d_data, int32_t na,
                           int32_t nr, int32_t inverse, void* stream);

/* generateReferenceFunction_CPU (compression.cc:17-58), on the device: d_theta
 * [na][nr] = exp(j theta), theta = 4 pi r0/c sqrt(max(0, (f0 + fr)^2 -
 * c^2 fa^2 / (4 vr^2))) - pi fr^2 / kr, with kr as that function takes it
 * (process_wk passes -Kr of paras.txt, wk.cc:229).  Axes as for
 * sarbp_stolt_sinc (host FP64, staged before the call returns). */
sarbp_status sarbp_wk_reference_function(sarbp_context* ctx, sarbp_cf32* d_theta,
                                         const double* h_fr_axis,
                                         const double* h_fa_axis, double f0,
                                         double c, double vr, double kr,
                                         double r0, int32_t na, int32_t nr,
                                         void* stream);

/* omega-K image of raw echoes d_echo [na_in][nr_in] into d_image [na][nr]
 * (the padded size; the buffer is also the transform workspace).  The steps
 * of process_wk (wk.cc:87-285): zero padding and azimuth down-conversion by
 * exp(-j 2 pi f_nc ta), ta[m] = (m - na/2) / prf (preprocess.cc:17-166); 2-D
 * FFT (fft.cc:18-198); reference function exp(j (4 pi r0/c s + pi fr^2 / kr))
 * (compression.cc:17-58; wp->range_ref selects the range origin, above);
 * Stolt interpolation (above); 2-D inverse FFT scaled
 * by 1/(na nr) (fft.cc:200-359).  The range DFT, reference function, Stolt
 * step and range inverse DFT run as one kernel per azimuth line (chirp-z
 * transforms in LDS; lines over 4096 samples need 128 KiB of LDS per
 * workgroup, else SARBP_ERR_UNSUPPORTED) between azimuth transforms (LDS
 * FFTs when na is a power of two from 1024 to 8192, LDS chirp-z transforms for
 * other na up to 4096, hipFFT beyond).  Plans,
 * twiddle and chirp tables
 * and a workspace of na x nr complex64 words are kept by the
 * context (the first call of a size builds them).  Calls on one context share
 * them and are ordered on the device: each call's stream first waits for the
 * previous call's work, whatever streams they use.  nr <= 8192. */
sarbp_status sarbp_omega_k(sarbp_context* ctx, const sarbp_cf32* d_echo, int32_t na_in, int32_t nr_in,
                           const sarbp_wk_params* wp, sarbp_cf32* d_image, void* stream);

/* ---- diagnostics -------------------------------------------------------------- */

/* The 96 coefficients of the SINC8 interpolator's half-bin kernel for Kaiser
 * shape `beta`: out[(h * 6 + k) * 8 + t] multiplies e^k in the weight of tap t
 * (sample floor(u) - 3 + t) at fractional bin f = u - floor(u) in half
 * h = floor(2 f) of the bin, e = 2 f - h - 1/2.  The halves mirror each
 * other exactly: out[((1 - h) * 6 + k) * 8 + 7 - t] = (-1)^k
 * out[(h * 6 + k) * 8 + t].  Host-only, no device needed. */
sarbp_status sarbp_sinc8_coefficients(double beta, float* out);
/* The 64 coefficients of its whole-bin kernel: out[k * 8 + t] multiplies e^k,
 * e = f - 1/2, in the weight of tap t; out[k * 8 + 7 - t] = (-1)^k
 * out[k * 8 + t].  Host-only. */
sarbp_status sarbp_sinc8_coefficients_whole(double beta, float* out);

/* Vector-ALU issue-rate probe (roofline calibration; no counterpart in the
 * reference).  Launches n_blocks workgroups of 256 threads; each thread runs
 * `iters` rounds of 128 VALU instructions over 8 independent FP32 chains and
 * writes one float to d_out[n_blocks * 256] (device).  kind 0: 128 v_fma_f32
 * (VOP3) per round; 1: 128 v_sin_f32; 2: 120 v_fma_f32 + 8 v_sin_f32; 3: 112
 * v_fma_f32 + 16 v_sin_f32; 4: 128 v_fmac_f32 (VOP2); 5: 128 v_mul_f32 (VOP2);
 * 6: 128 v_add_f32 (VOP2); 7: 128 v_fract_f32 (VOP1); 8: 128 v_med3_f32
 * (VOP3); 9: 64 v_fma_f32 + 64 v_fmac_f32 alternating; 10-16: 64 v_fma_f32
 * interleaved with 64 of v_add_u32, v_mad_u32_u24, v_cvt_f32_i32, v_fract_f32,
 * v_med3_f32, v_mul_f32, v_mov_b32, v_add_f32, v_lshrrev_b32, v_and_b32,
 * v_or_b32, v_min_u32, v_cndmask_b32, v_lshl_add_u32, v_add3_u32,
 * v_cvt_i32_f32, v_max_f32, v_floor_f32, v_sub_u32, v_bfe_u32, v_ldexp_f32,
 * v_ashrrev_i32 respectively.
 * Asynchronous on `stream`; the caller times it. */
sarbp_status sarbp_valu_probe(int32_t kind, int32_t n_blocks, int32_t iters,
                              float* d_out, void* stream);

/* The BP kernels' phase rotation evaluated on n phases (revolutions, FP32,
 * device memory): d_out[2 i] = cos(2 pi ph[i]), d_out[2 i + 1] = sin(2 pi
 * ph[i]) as the pixel loop computes them (hardware v_cos_f32 / v_sin_f32 on
 * revolutions, the rotation of 1 + 0j).  For accuracy tests against FP64.
 * No counterpart in the reference, whose kernel calls sincosf
 * (code/SAR_BP/CUDA/BPA_CUDA_kernels.cu:67).  Asynchronous on `stream`. */
sarbp_status sarbp_rotation_probe(const float* d_phase, float* d_out, int64_t n,
                                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SARBP_H */
