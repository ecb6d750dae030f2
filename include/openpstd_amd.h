/*
 * openpstd_amd — MI355X-native PSTD update loop behind openPSTD's KernelInterface.
 *
 * This header is the drop-in boundary: a plain C ABI (no C++ or torch types).
 *
 *  1. Configuration types mirroring openPSTD's PSTDSettings / DomainConf /
 *     PSTDConfiguration (kernel/KernelInterface.h:44-241).
 *  2. The device solver context (pstd_ctx_*): what openPSTD's GPU solver slot,
 *     GPUSingleThreadSolver::compute_propagation (kernel/Solver.cpp:153-156) and
 *     GPUMultiThreadSolver::compute_propagation (kernel/Solver.cpp:157-160),
 *     binds to advance a scene on the GPU.  The scene is handed over as a flat
 *     plan (domains, line segments, receivers) built from openPSTD's Scene.
 *  3. The kernel facade (pstd_kernel_*): the C form of
 *     KernelInterface::{initialize_kernel, run, get_metadata}
 *     (kernel/KernelInterface.h:332-351) as implemented by
 *     PSTDKernel(bool GPU, bool MCPU) (kernel/PSTDKernel.h:52-150), for callers
 *     that are not C++ (the Python package binds this through ctypes).
 *
 * All functions return PSTD_OK (0) on success and a negative status on
 * failure; the message is available from pstd_ctx_last_error /
 * pstd_kernel_last_error.  No exception crosses this boundary.
 */
#ifndef OPENPSTD_AMD_H
#define OPENPSTD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define PSTD_OK 0
#define PSTD_ERR_INVALID (-1)        /* invalid argument or unsupported scene */
#define PSTD_ERR_NOT_CONFIGURED (-2) /* PSTDKernelNotConfiguredException (kernel/KernelInterface.h:353-356) */
#define PSTD_ERR_DEVICE (-3)         /* HIP runtime error */
#define PSTD_ERR_UNSUPPORTED (-4)    /* solver flavour not provided by this library */
#define PSTD_ERR_ALLOC (-5)          /* device or host allocation failed */

/* -------------------------------------------------------- configuration */
/* Mirrors PSTDSettings (kernel/KernelInterface.h:44-158); same fields, same
 * units (SI), same serialisation order (kernel/KernelInterface.h:80-97). */
typedef struct pstd_settings {
    float calctime;         /* simulation length [s]; steps = (int)(calctime/dt) */
    float c1;               /* speed of sound [m/s] */
    float ampmax;           /* PML attenuation amplitude */
    float freqmax;          /* maximum frequency [Hz] (not used by the loop) */
    float rho;              /* density of air [kg/m^3] */
    float patcherror;       /* patch error [dB]; window = round(round((pe*0.7-17)/2)*2) */
    float tfactRK;          /* CFL factor; dt = tfactRK*dx/c1 */
    float grid_spacing;     /* dx [m] */
    float band_width;       /* Gaussian source bandwidth beta */
    float wave_length;      /* only used by spectral interpolation (disabled in openPSTD) */
    float rk_coefficients[6];
    int32_t spectral_interpolation;
    int32_t pml_cells;
    int32_t save_nth;
} pstd_settings;

/* DomainConfEdge (kernel/KernelInterface.h:163-176) */
typedef struct pstd_edge {
    float absorption; /* in [0, 1] */
    int32_t lr;       /* locally reacting flag */
} pstd_edge;

/* DomainConf (kernel/KernelInterface.h:181-206); world metres. */
typedef struct pstd_domain_conf {
    float top_left[2];
    float size[2];
    pstd_edge T, L, B, R;
} pstd_domain_conf;

/* Receiver sampling rule.  openPSTD's nearest-cell receiver reads
 * p0(rel.x, rel.y) with p0 stored (row = y, col = x)
 * (kernel/core/Receiver.cpp:80-83), i.e. the transposed cell.
 * PSTD_RECEIVER_PHYSICAL (default) samples p0[rel.y][rel.x];
 * PSTD_RECEIVER_REFERENCE reads what openPSTD's p0(rel.x, rel.y) reads from
 * its column-major ny x nx array in an NDEBUG build: storage element
 * rel.x + rel.y * ny, i.e. the cell (row (rel.x + rel.y*ny) % ny, column
 * (rel.x + rel.y*ny) / ny) -- the transposed cell when rel.x < ny -- and
 * rejects receivers whose element lies past the array (a build with
 * assertions aborts on every rel.x >= ny). */
#define PSTD_RECEIVER_PHYSICAL 0
#define PSTD_RECEIVER_REFERENCE 1
/* Spectral interpolation at the receiver's exact location
 * (Receiver::compute_with_si, kernel/core/Receiver.cpp:85-121, which openPSTD
 * keeps switched off, :48): the pressure of the receiver's domain shifted
 * along X by the fractional part of the receiver's x (every row of the domain
 * and the window rows of its Y neighbours, each line with its own neighbours
 * and stencil as Domain::calc resolves them), then the receiver's column
 * shifted along Y by the fractional part of y, read at the receiver's cell.
 * Fractional parts are location - (int)location (openPSTD's constructor
 * stores the integer cell there, :41); wavenumbers are those of each line's
 * FFT length.  PSTDSettings::spectral_interpolation alone does not select it,
 * as in openPSTD. */
#define PSTD_RECEIVER_SPECTRAL 2

/* PSTDConfiguration (kernel/KernelInterface.h:214-241). */
typedef struct pstd_config {
    pstd_settings settings;
    int32_t n_speakers;
    const float *speakers;   /* n_speakers x {x, y, z} world metres */
    int32_t n_receivers;
    const float *receivers;  /* n_receivers x {x, y, z} world metres */
    int32_t n_domains;
    const pstd_domain_conf *domains;
    int32_t receiver_mode;   /* PSTD_RECEIVER_* */
} pstd_config;

/* ------------------------------------------------------- device solver */
/* Field identifiers for pstd_ctx_set_field / pstd_ctx_get_field; they name
 * FieldValues members (kernel/core/Domain.h:57-63). */
#define PSTD_FIELD_P0 0
#define PSTD_FIELD_PX 1
#define PSTD_FIELD_PY 2
#define PSTD_FIELD_VX 3
#define PSTD_FIELD_VY 4

/* One rectangular domain of the scene (kernel/core/Domain.h:103-153).
 * Arrays are row-major (row = y, column = x), the frame order of
 * Solver::get_pressure_vector (kernel/Solver.cpp:208-217):
 * p0/px/py ny x nx, vx ny x (nx+1), vy (ny+1) x nx. */
typedef struct pstd_plan_domain {
    int32_t nx, ny;
    float rho;               /* Domain::rho */
    int32_t is_pml;          /* receives PML attenuation after every step */
    int32_t is_rigid;        /* Domain::is_rigid(): never updated */
    const float *pml_px;     /* ny x nx PML factors, or NULL when !is_pml */
    const float *pml_py;     /* ny x nx */
    const float *pml_vx;     /* ny x (nx+1) */
    const float *pml_vy;     /* (ny+1) x nx */
} pstd_plan_domain;

/* A run of lines of one domain that share neighbours and stencil: one
 * spatderp3 call of Domain::calc (kernel/core/Domain.cpp:149-370).
 * direction 0 = X (lines are rows), 1 = Y (lines are columns);
 * field 0 = pressure (derivative of p0 -> updates vx/vy),
 *       1 = velocity (derivative of vx/vy -> updates px/py). */
typedef struct pstd_plan_segment {
    int32_t domain;
    int32_t direction;
    int32_t field;
    int32_t line_begin, line_end;   /* local line range [begin, end) */
    int32_t side1, side2;           /* neighbour domains (left/bottom, right/top) or -1 = zeros */
    int32_t side1_offset, side2_offset; /* neighbour-local line = line + offset */
    int32_t wlen;                   /* window half length w (Domain::calc halves it for short domains) */
    int32_t table_len;              /* length of the wavenumber table derfact is read from */
    float c_side1, c_main1, c_side2, c_main2; /* get_rho_array stencil coefficients */
    int32_t interp_only;            /* 1: not part of the time step; the X pressure lines of a domain
                                       the RK loop does not update along X (rigid, or a PML strip
                                       that skips X), present for spectral receivers beside it */
} pstd_plan_segment;

typedef struct pstd_plan_receiver {
    int32_t domain;
    int32_t index;                  /* row-major index into the domain's p0 */
    int32_t spectral;               /* 1: spectral interpolation (PSTD_RECEIVER_SPECTRAL) */
    float offset_x, offset_y;       /* fractional offsets from the cell, in cells, in [0, 1) */
} pstd_plan_receiver;

typedef struct pstd_plan_desc {
    int32_t n_domains;
    const pstd_plan_domain *domains;
    int32_t n_segments;
    const pstd_plan_segment *segments;
    int32_t n_receivers;
    const pstd_plan_receiver *receivers;
    float dx;                       /* grid spacing */
    float dt;                       /* PSTDSettings::GetTimeStep() */
    float c1;                       /* speed of sound */
    float patch_error;              /* window shape */
    float rk_alpha[6];              /* low-storage RK coefficients */
} pstd_plan_desc;

typedef struct pstd_ctx pstd_ctx;

/* Creates a solver context on HIP device `device`. */
int pstd_ctx_create(int device, pstd_ctx **out);
/* Uploads the scene plan: allocates device state (zeroed), builds the device
 * plan cache (derivative filters, windows, twiddles, line tiles). */
int pstd_ctx_upload_plan(pstd_ctx *ctx, const pstd_plan_desc *plan);
/* Derivative form (pstd_ctx_set_form, before the plan is uploaded).
 * PSTD_FORM_FFT: every line derivative as openPSTD computes it (windowed line,
 *   FFT, spectral factor, inverse FFT) in FP32.
 * PSTD_FORM_GEMM_F16: segments whose lines and outputs fill blocks of 128
 *   lines x 256 outputs at least to 3/4 apply the segment's dense derivative
 *   matrix (pstd_derivative_matrix) on MFMA with FP16 hi + lo split operands
 *   (22-bit significands; the inputs scaled per scene by a power of two, from
 *   the largest value the previous RK stage wrote) and FP32 accumulation; the
 *   last output of a pressure line (n + 1 outputs) is an FP32 dot product;
 *   the other segments (thin PML strips) keep the FFT form.
 * PSTD_FORM_GEMM_BF16: the same with BF16 hi + lo operands (16-bit
 *   significands, no range limit). */
#define PSTD_FORM_FFT 0
#define PSTD_FORM_GEMM_F16 1
#define PSTD_FORM_GEMM_BF16 2
/* PSTD_FORM_AUTO (no counterpart in openPSTD): the plan is uploaded and
 * autotuned in the FFT form and in the FP16 GEMM form, and the form with the
 * shorter RK stage is kept (per device and plan signature, so a later upload
 * of a plan of the same geometry uploads once); pstd_ctx_get_form reports it. */
#define PSTD_FORM_AUTO 3
int pstd_ctx_set_form(pstd_ctx *ctx, int form);
/* The form the uploaded plan runs in (PSTD_FORM_FFT / _GEMM_F16 / _GEMM_BF16). */
int pstd_ctx_get_form(const pstd_ctx *ctx);
/* Uploads several independent scenes as one plan: every RK stage of all of
 * them is one launch.  Domains and receivers are numbered across the scenes in
 * order (scene 0's domains first); segments name domains of their own scene.
 * The scenes must share dx, dt, c1, patch error and RK coefficients.  No
 * openPSTD counterpart: openPSTD's Solver advances one Scene
 * (kernel/Solver.cpp:62-93); a batch advances each scene exactly as that loop
 * does. */
int pstd_ctx_upload_plans(pstd_ctx *ctx, const pstd_plan_desc *plans, int n_plans);
/* Host <-> device copies of one field of one domain (FieldValues::current). */
int pstd_ctx_set_field(pstd_ctx *ctx, int domain, int field, const float *host, size_t count);
int pstd_ctx_get_field(pstd_ctx *ctx, int domain, int field, float *host, size_t count);
/* Advances `steps` time steps (compute_timestep, kernel/Solver.cpp:71-93):
 * 6 RK stages, PML attenuation, receiver sampling.  When `samples` is not
 * NULL it receives steps x n_receivers receiver values (one per step, in
 * receiver order).  Blocks until the samples are on the host. */
int pstd_ctx_advance(pstd_ctx *ctx, int steps, float *samples);
/* Like pstd_ctx_advance but does not synchronise and keeps the samples on the
 * device (bench and pipelined callers); pstd_ctx_synchronize waits. */
int pstd_ctx_advance_async(pstd_ctx *ctx, int steps);
/* Allocation-free variant of pstd_ctx_advance_async, with no counterpart in
 * the reference: for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid it).  It calls no hipMalloc, hipFree or other
 * allocating routine: the launches are issued one by one (no graph is
 * instantiated), and the GEMM form's scratch -- the per-scene maxima that
 * choose the split scaling -- lives in a file-scope device array of
 * 3 x 2 x 1024 words per device instead of the context's buffer; more than
 * 1024 scenes in a GEMM-form context is PSTD_ERR_INVALID.  Every call starts
 * that array from zero and from the maxima of the current state, so results
 * equal pstd_ctx_advance_async's from the same state (bitwise when that call
 * also starts from host-written state).  All calls on one device share the
 * array: calls that overlap on one device race. */
int pstd_ctx_advance_static(pstd_ctx *ctx, int steps);
int pstd_ctx_synchronize(pstd_ctx *ctx);
/* Current p0 of one domain, row-major (Solver::get_pressure_vector). */
int pstd_ctx_read_frame(pstd_ctx *ctx, int domain, float *out, size_t count);
/* Advances `steps` time steps like pstd_ctx_advance, with one HIP event pair
 * on the solver stream around the whole run, and waits for it; returns the
 * GPU time of the run divided by its 6*steps RK-stage launches (ms) and that
 * launch count.  For measurement; the state advances exactly as with
 * pstd_ctx_advance (samples stay on the device). */
int pstd_ctx_profile_stages(pstd_ctx *ctx, int steps, double *mean_stage_ms, int *launches);
/* Advances `steps` time steps like pstd_ctx_advance_async, issuing every
 * launch eagerly with a HIP event pair on the solver stream around each
 * launch of a kernel class, and waits.  ms[c] / launches[c] receive the summed
 * GPU time and launch count of class c: 0 the FFT-form kernel of lines up to
 * 2048 points (with the passthrough and receiver tiles), 1 the 4096-point
 * kernel, 2 the 8192-point kernel, 3 the GEMM-form kernel; ms[4] the whole
 * run, launches[4] its 6*steps RK stages.  For measurement (no counterpart in
 * the reference). */
int pstd_ctx_profile_kernels(pstd_ctx *ctx, int steps, double ms[5], int launches[5]);
/* Replays whole blocks of `steps` time steps as one captured hipGraph (0 = the
 * stage launches are issued one by one; default 20).  The graph is captured on
 * first use, when p0 comes from the stage outputs (the first step after an
 * upload or a host write runs eagerly), and dropped on re-upload or a change of
 * launch shape.  Receiver samples of a replay are copied into place on the
 * device.  Results are identical either way. */
int pstd_ctx_set_graph_steps(pstd_ctx *ctx, int steps);
/* Captures and instantiates the step graph now instead of at its first use
 * (no time step runs), so that a timed loop does not include the capture.
 * PSTD_ERR_INVALID before the first time step after an upload or a host
 * write; a no-op when graphs are off or already captured.  No counterpart in
 * the reference. */
int pstd_ctx_capture_graph(pstd_ctx *ctx);
/* on = 1: in a stage with GEMM-form tiles, the FFT-form tiles (thin
 * segments, PML strips, passthrough and receiver tiles) run on a second
 * stream of the context, concurrently with the GEMM kernel, forked after and
 * joined before the neighbouring stages.  Default 0: issued in order on one
 * stream (the GEMM kernel occupies every CU's LDS, so concurrency gains
 * nothing measurable).  Results are identical either way.  No counterpart in
 * the reference. */
int pstd_ctx_set_stage_overlap(pstd_ctx *ctx, int on);
/* Launch shape of the stage kernel (no counterpart in the reference, whose
 * FFTW plans are made with FFTW_ESTIMATE on every call,
 * kernel/core/kernel_functions.cpp:175-186).  pstd_ctx_upload_plan(s) times
 * the candidates on the zeroed state and keeps the fastest (the autotuner);
 * these calls override or report its choice.  Results are identical for
 * every choice.
 *  - tile order: 1 gives every XCD a contiguous run of tiles (neighbouring
 *    lines share its L2), 0 the plain round-robin order;
 *  - kernel flavour: 0 loads every descriptor-dependent operand (filter,
 *    twiddles, t_n) before the transforms (most registers), 1 loads them
 *    where they are used (fewest registers, most resident waves), 2 loads
 *    filter and twiddles up front and t_n where used;
 *  - tile threads: block size of the kernel for lines of FFT length <= 2048,
 *    128, 256 or 512 (512 puts twice the lines of 256 in a tile: for columns,
 *    twice the contiguous bytes of each row per wave access; 128 spreads a
 *    small scene over more CUs and needs every such line to have F <= 1024,
 *    else PSTD_ERR_INVALID).
 * The choice is kept per (device, plan signature: form, scenes, every tile's
 * line class, line count and direction): a later upload of a plan with the
 * same signature on the same device takes it without trials.
 * pstd_ctx_get_tuning reports the choice and its timed GPU ms per stage,
 * pstd_ctx_get_tile_threads the block size, pstd_ctx_get_autotune the time
 * steps the trials ran on this context (launches of the stage kernels that are
 * not part of the caller's steps) and whether the choice came from the cache. */
int pstd_ctx_set_tile_order(pstd_ctx *ctx, int xcd_contiguous);
int pstd_ctx_set_kernel_flavour(pstd_ctx *ctx, int flavour);
int pstd_ctx_set_tile_threads(pstd_ctx *ctx, int threads);
int pstd_ctx_get_tile_threads(const pstd_ctx *ctx);
int pstd_ctx_get_tuning(const pstd_ctx *ctx, int32_t *flavour, int32_t *xcd_contiguous, double *stage_ms);
int pstd_ctx_get_autotune(const pstd_ctx *ctx, int32_t *trial_steps, int32_t *from_cache);
/* Work of the GEMM-form segments in one RK stage: algorithmic MFMA flops
 * (3 split products x 2 x K x outputs x lines) and line outputs. */
int pstd_ctx_gemm_work(const pstd_ctx *ctx, double *flops_per_stage, long long *outputs_per_stage);
/* Cells updated per RK stage: sum of nx*ny over the non-rigid domains. */
long long pstd_ctx_cells_per_stage(const pstd_ctx *ctx);
/* Number of time steps advanced since upload. */
int pstd_ctx_steps_done(const pstd_ctx *ctx);
const char *pstd_ctx_last_error(const pstd_ctx *ctx);
void pstd_ctx_destroy(pstd_ctx *ctx);

/* ------------------------------------------------------- kernel facade */
/* KernelCallback (kernel/KernelInterface.h:243-304) as C function pointers.
 * write_frame: WriteFrame(frame, domain, data) — fresh row-major p0 per call.
 * write_sample: WriteSample(startSample, receiver, data).
 * log: level 0 Fatal, 1 Error, 2 Warning, 3 Info, 4 Debug. */
typedef struct pstd_callbacks {
    void *user;
    void (*write_frame)(void *user, int frame, int domain, const float *data, size_t n);
    void (*write_sample)(void *user, int start_sample, int receiver, const float *data, size_t n);
    void (*log)(void *user, int level, const char *message);
} pstd_callbacks;

typedef struct pstd_kernel pstd_kernel;

/* PSTDKernel(bool GPU, bool MCPU) (kernel/PSTDKernel.cpp:44-47).  This
 * library provides the GPU solver slot: GPU=0 is PSTD_ERR_UNSUPPORTED at run. */
int pstd_kernel_create(int gpu, int mcpu, pstd_kernel **out);
/* initialize_kernel (kernel/PSTDKernel.cpp:49-59). */
int pstd_kernel_initialize(pstd_kernel *k, const pstd_config *conf, const pstd_callbacks *log);
/* run (kernel/PSTDKernel.cpp:121-148): blocking; callbacks on the caller's thread. */
int pstd_kernel_run(pstd_kernel *k, const pstd_callbacks *cb);
/* get_metadata (kernel/PSTDKernel.cpp:154-168): sizes of every domain
 * (PML included) as x,y,z triples, and the frame count. */
int pstd_kernel_get_metadata(pstd_kernel *k, int32_t *n_domains, int32_t *sizes_xyz, int32_t max_domains,
                             int32_t *framecount);
/* Scene inspection (the test-only Scene access of PSTDKernel::get_scene,
 * kernel/PSTDKernel.cpp:150-152).  info receives 16 ints:
 * id, tl.x, tl.y, size.x, size.y, is_pml, is_secondary_pml, is_rigid, local,
 * n_left, n_right, n_top, n_bottom, update_x, update_y, pml_for_count;
 * finfo receives alpha, impedance, rho. */
int pstd_kernel_domain_info(pstd_kernel *k, int index, int32_t *info, float *finfo);
/* Current field of domain `index` (after initialisation, or after run). */
int pstd_kernel_read_field(pstd_kernel *k, int index, int field, float *out, size_t count);
/* PML factor array (0 px, 1 py, 2 vx, 3 vy) of domain `index`. */
int pstd_kernel_read_pml(pstd_kernel *k, int index, int which, float *out, size_t count);
/* Builds the device context of the configured scene (plan uploaded, initial
 * fields set, frame 0 next) without running it, and returns it in *ctx (owned
 * by the kernel; valid until the next prepare/run/destroy).  For callers that
 * drive the time loop themselves (benchmarks, stepwise inspection). */
int pstd_kernel_prepare(pstd_kernel *k, pstd_ctx **ctx);
/* Derivative form of the device contexts this kernel creates (PSTD_FORM_*,
 * default PSTD_FORM_FFT); takes effect at the next run / prepare. */
int pstd_kernel_set_form(pstd_kernel *k, int form);
/* Builds one device context that advances the scenes of `n` configured
 * kernels together (pstd_ctx_upload_plans), initial fields set, frame 0 next.
 * The caller owns *ctx (pstd_ctx_destroy).  Domain d of scene i is domain
 * (sum of the earlier scenes' domain counts) + d; receivers likewise.  The
 * derivative form is that of kernels[0]. */
int pstd_batch_prepare(pstd_kernel *const *kernels, int n, int device, pstd_ctx **ctx);
/* Dense derivative matrix of one segment's lines (n_out x K row-major, in
 * double; n_out = n_main + 1 for pressure, n_main - 1 for velocity;
 * K = n_main + 2*wlen), the operator the GEMM form multiplies: row i applied to
 * [side1 tail (wlen) | main (n_main) | side2 head (wlen)] gives output i of
 * spatderp3 (kernel/core/kernel_functions.cpp:123-318) for that line.
 * coef = {c_side1, c_main1, c_side2, c_main2} (pstd_plan_segment).  Returns
 * the element count (out may be NULL to query it) or a negative status. */
long long pstd_derivative_matrix(int F, int table_len, int field, double dx, int wlen, int n_main, double patch_error,
                                 const double coef[4], double *out, size_t cap);
/* ---------------------------------------------------------- .pstd files */
/* openPSTD's scene/results file (shared/PSTDFile.{h,cpp}): an unqlite 1.1.6
 * key/value database (read whole into memory, written back whole by
 * pstd_file_commit) holding the scene as a Boost text archive of
 * PSTDConfiguration (key {1}), the results scene ({100}), frame counts
 * ({102, domain}), frames ({103, domain, frame}: float32 p0, row-major) and
 * receiver series ({104, receiver}: float32, appended per step), version
 * ({10000} = 4). */
typedef struct pstd_file pstd_file;
/* PSTDFile::Open (shared/PSTDFile.cpp:165-178): reads the file, checks the version. */
int pstd_file_open(const char *path, pstd_file **out);
/* PSTDFile::New (:180-199): version, the default scene, an empty results scene; written at once. */
int pstd_file_new(const char *path, pstd_file **out);
/* PSTDFile::GetSceneConf / GetResultsSceneConf (:206-209, 476-479): the arrays
 * `out` points at belong to the file until the next call or pstd_file_close;
 * receiver_mode is PSTD_RECEIVER_PHYSICAL (the file has none); an archive
 * with other than 6 RK coefficients (an empty scene has none) fills the
 * first ones and zeros. */
int pstd_file_get_scene(pstd_file *f, int results, pstd_config *out);
/* PSTDFile::SetSceneConf (:211-214, 433-452). */
int pstd_file_set_scene(pstd_file *f, const pstd_config *conf);
/* PSTDFile::GetResultsFrameCount / GetResultsFrame / GetReceiverData
 * (:221-231, 487-492): counts, or element counts (out may be NULL to query). */
int pstd_file_frame_count(pstd_file *f, int domain);
long long pstd_file_get_frame(pstd_file *f, int frame, int domain, float *out, size_t cap);
long long pstd_file_get_receiver(pstd_file *f, int receiver, float *out, size_t cap);
/* Raw records (PSTDFile::GetRawValue / SetRawValue / AppendRawValue /
 * DeleteValue, shared/PSTDFile.cpp:321-342, 368-402, 456-466); keys are
 * native 32-bit integers {prefix, ...} (PSTDFile::CreateKey, :344-366).
 * get_raw returns the value's size (out may be NULL); set_raw with append
 * extends the value.  record_count takes a list of the keys in byte order
 * (PSTDFile::OutputDebugInfo, :261-300); record_key reads key i of that
 * list. */
long long pstd_file_get_raw(pstd_file *f, const void *key, int nkey, void *out, size_t cap);
int pstd_file_set_raw(pstd_file *f, const void *key, int nkey, const void *data, long long n, int append);
int pstd_file_delete_raw(pstd_file *f, const void *key, int nkey);
int pstd_file_record_count(pstd_file *f);
int pstd_file_record_key(pstd_file *f, int i, void *out, int cap);
/* PSTDFile::Commit (:494-501): writes the whole database to the file. */
int pstd_file_commit(pstd_file *f);
/* The CLI's `run` (RunCommand::execute, CLI/main.cpp:422-452): delete old
 * results, initialise new ones, run PSTDKernel(gpu, mcpu) on the file's scene
 * with frames and receiver samples written into the file (CLIOutput,
 * CLI/output.cpp:30-39), commit.  Log lines go to log->log. */
int pstd_file_run(pstd_file *f, int gpu, int mcpu, int receiver_mode, const pstd_callbacks *log);
const char *pstd_file_last_error(const pstd_file *f);
void pstd_file_close(pstd_file *f);

/* One row of the spectral receivers' shift operator: the K = n_main + 2w
 * weights of [side1 tail | main | side2 head] whose dot product with a
 * pressure line is spatderp3 (kernel/core/kernel_functions.cpp:123-318) of
 * that line with the shift factors of Receiver::get_fft_factors
 * (kernel/core/Receiver.cpp:57-73, offset in cells) in place of derfact,
 * read at output `row` (< n_main + 1).  Returns K (out may be NULL to query
 * it) or a negative status. */
long long pstd_shift_operator_row(int F, int table_len, double dx, int wlen, int n_main, double patch_error,
                                  const double coef[4], double offset, int row, double *out, size_t cap);
/* Plan the kernel hands to the device (segments etc.) — for tests. */
int pstd_kernel_segment_count(pstd_kernel *k);
int pstd_kernel_get_segment(pstd_kernel *k, int i, pstd_plan_segment *out);
const char *pstd_kernel_last_error(const pstd_kernel *k);
void pstd_kernel_destroy(pstd_kernel *k);

#ifdef __cplusplus
}
#endif
#endif /* OPENPSTD_AMD_H */
