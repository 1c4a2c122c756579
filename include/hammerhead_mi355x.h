/*
 * hammerhead_mi355x.h — C ABI of the MI355X (gfx950) WIDIM correlation backend.
 *
 * This library replaces the device-engine hooks of Hammerhead.jl's backend seam
 * (a private multiple-dispatch hook set resolved from a public Symbol,
 * docs/src/reference/backends.md:40-66, src/types.jl:61-111). A Julia shim that
 * registers `backend = :mi355x` binds every entry point below with `ccall`;
 * the Python host package binds them with ctypes. Plain pointers and sizes
 * only; no HIP, torch or Julia types cross the boundary.
 *
 * Conventions (identical to the reference):
 *   - images, vector grids and correlation planes are COLUMN-MAJOR;
 *   - window origins, grid indices and peak locations are 1-BASED;
 *   - u is the column (x) displacement, v the row (y) displacement; a particle
 *     at (row, col) in A is found at (row + v, col + u) in B
 *     (src/correlators.jl:4-7, src/types.jl:322-325).
 *   - arithmetic type T = float32 (the reference's Float32 pipeline,
 *     src/types.jl:314-317) with float64 islands where the reference keeps them
 *     (correlation-moment sums, uncertainty sums, substitution/replacement
 *     medians, the UOD residual division).
 *
 * Every call returns an int status (HH_OK = 0). On error, hh_last_error(ctx)
 * returns a message. Options the device engine does not implement are rejected
 * before any work with HH_EUNSUPPORTED, mirroring the reference's up-front
 * ArgumentError (src/ka_backend.jl:13-29); shape errors return HH_EDIM
 * (DimensionMismatch). A context owns one HIP stream; it is not thread-safe,
 * distinct contexts are.
 */
#ifndef HAMMERHEAD_MI355X_H
#define HAMMERHEAD_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
    HH_OK = 0,
    HH_EINVAL = 1,       /* ArgumentError: invalid parameter value            */
    HH_EDIM = 2,         /* DimensionMismatch                                 */
    HH_EUNSUPPORTED = 3, /* ArgumentError: option not implemented on device   */
    HH_EDEVICE = 4,      /* HIP runtime error                                 */
    HH_ENOMEM = 5,       /* device allocation failed                          */
    HH_ESTATE = 6        /* call-order error (e.g. hh_deform before a pair)   */
};

/* ---- enumerations mirroring PIVParameters symbols (src/types.jl:212-231) */
enum { HH_CROSS = 0, HH_PHASE = 1 };                   /* correlation_method */
enum { HH_APOD_NONE = 0, HH_APOD_GAUSS = 1 };           /* apodization        */
enum { HH_GAUSS3 = 0, HH_GAUSS9 = 1, HH_GAUSS2D = 2, HH_SUBPIX_NONE = 3 };
enum { HH_REGIONALMAX = 0, HH_EXCLUSION = 1 };          /* peak_finder        */
/* extra validators of PIVParameters.validation (src/quality.jl:383-428) */
enum { HH_V_PEAK_RATIO = 1, HH_V_MOMENT = 2, HH_V_VELMAG = 3, HH_V_UOD = 4 };
/* correlation arithmetic form */
enum { HH_FORM_FFT32 = 0, HH_FORM_MFMA16 = 1 };
/* window source of a correlation call */
enum { HH_SRC_RAW = 0, HH_SRC_WARP = 1 };

#define HH_MAX_VALIDATORS 4
#define HH_UQ_NSTATS 44 /* UQ_NSTATS, src/uncertainty.jl:35 */

/* One entry of PIVParameters.validation (src/quality.jl:383-428):
 *   HH_V_PEAK_RATIO  a = threshold
 *   HH_V_MOMENT      a = threshold
 *   HH_V_VELMAG      a = min, b = max
 *   HH_V_UOD         a = threshold, b = epsilon, neighborhood = layers       */
typedef struct hh_validator {
    int32_t kind;
    int32_t neighborhood;
    double a;
    double b;
} hh_validator;

/* PIVParameters (src/types.jl:212-291), field for field. */
typedef struct hh_pass {
    int32_t window_size[2];      /* (rows, cols)                           */
    int32_t search_area_size[2]; /* (rows, cols)                           */
    int32_t overlap[2];          /* (rows, cols)                           */
    int32_t correlation_method;  /* HH_CROSS | HH_PHASE                    */
    int32_t padding;             /* bool                                   */
    int32_t apodization;         /* HH_APOD_*                              */
    int32_t subpixel_method;     /* HH_GAUSS3 | HH_GAUSS9 | ...            */
    int32_t n_peaks;
    int32_t peak_finder;         /* HH_REGIONALMAX | HH_EXCLUSION          */
    int32_t uncertainty;         /* bool                                   */
    int32_t uod_enable;          /* bool                                   */
    double uod_threshold;
    int32_t uod_neighborhood;
    int32_t replace_outliers;    /* bool                                   */
    double min_peak_ratio;
    int32_t max_iterations;
    int32_t keep_correlation_planes; /* bool (rejected on device)          */
    double convergence_tol;
    int32_t n_validators;
    int32_t reserved;
    hh_validator validators[HH_MAX_VALIDATORS];
} hh_pass;

/* PIVResult{Float32} (src/types.jl:359-373) as caller-owned host arrays.
 * Grids are ny x nx column-major; x has nx entries, y has ny. A NULL array
 * is skipped. outliers/mask are one byte per cell (0/1).                     */
typedef struct hh_field {
    int32_t ny, nx;
    float *x, *y;
    float *u, *v, *peak_ratio, *correlation_moment;
    float *uncertainty_u, *uncertainty_v;
    uint8_t *outliers, *mask;
} hh_field;

typedef struct hh_ctx hh_ctx;

/* ---- lifecycle ------------------------------------------------------- */

/* Create a context on HIP device `device` (piv_workspace(; backend),
 * src/pipeline.jl:44-47). */
int hh_create(hh_ctx **out, int device);
/* Release all device memory (release_piv_engines!, src/pipeline.jl:121). */
int hh_destroy(hh_ctx *ctx);
/* Message of the last failing call on ctx ("" if none); valid until the
 * next call on ctx. ctx may be NULL (reports hh_create failures). */
const char *hh_last_error(const hh_ctx *ctx);
const char *hh_version(void);

/* Select the correlation arithmetic: HH_FORM_FFT32 (FP32 radix FFT) or
 * HH_FORM_MFMA16 (split-FP16 DFT products on MFMA). */
int hh_set_form(hh_ctx *ctx, int form);

/* ---- parameter checks (no device work) --------------------------------- */

/* _check_backend_params(b, passes) (src/types.jl:110, src/ka_backend.jl:13-29):
 * HH_OK, or HH_EINVAL / HH_EUNSUPPORTED with a message, before any work.
 * The device scope: square windows and search areas, FFT side (the search
 * area, doubled when padded) 16, 32, 64 or 128 (radix-2) or 24, 48 or 96
 * (direct DFT); an enlarged search area
 * (search_area_size > window_size) takes the general window load of
 * src/correlators.jl:230-262 and restrict_search_lags! (:347-355). */
int hh_check_passes(hh_ctx *ctx, const hh_pass *passes, int npasses);
/* Grid counts of pass_grid (src/pipeline.jl:529-555); *ny = #row starts. */
int hh_grid_size(const hh_pass *pass, int nr, int nc, int *ny, int *nx);

/* ---- image pair -------------------------------------------------------- */

/* Copy a host column-major float32 pair (nr x nc) to the device
 * (the engine staging of ext/HammerheadAMDGPUExt.jl:375-396). */
int hh_set_pair(hh_ctx *ctx, const float *A, const float *B, int nr, int nc);
/* Borrow a pair already resident in device memory (column-major float32,
 * nr x nc). The pointers must stay valid while ctx uses them. */
int hh_set_pair_device(hh_ctx *ctx, const float *dA, const float *dB, int nr, int nc);

/* ---- dewarping (stereo) ------------------------------------------------ */

/* ImageDewarper(cam, grid, image_size) (src/dewarp.jl:196-221) as its
 * precomputed fields: for every node of the ny x nx world grid the 1-based
 * source row and column in the camera frame (rows, cols: column-major
 * float64, as the reference stores them) and the out-of-view mask (nullable;
 * nonzero = out of view). Nodes that are masked, non-finite or outside
 * [1, nr] x [1, nc] are out of view. nr x nc: the camera's frame size (at
 * most 32767 per side); the map is kept on the device with each coordinate
 * resolved to 2^-16 px. */
int hh_set_dewarp_map(hh_ctx *ctx, const double *rows, const double *cols, const uint8_t *mask, int ny, int nx,
                      int nr, int nc);
/* dewarp!(out, dw, img) (src/dewarp.jl:233-248) on the device: the frame
 * (host column-major float32, nr x nc = the map's camera size, HH_EDIM
 * otherwise) is prefiltered (cubic B-spline, Line(OnGrid())) and sampled at
 * the map; out (ny x nx host) holds 0 at out-of-view nodes. */
int hh_dewarp(hh_ctx *ctx, const float *img, int nr, int nc, float *out);
/* Both frames of one camera's pair dewarped as hh_dewarp does, the result
 * kept on the device as the context's current pair (ny x nx), as
 * _run_piv_stereo! (src/stereo.jl:186-192) dewarps before each camera's
 * run_piv: a following hh_run_piv (or the hooks) correlates it without a
 * host round trip. */
int hh_dewarp_pair(hh_ctx *ctx, const float *A, const float *B, int nr, int nc);

/* ---- deformation ------------------------------------------------------- */

/* Cubic B-spline prefilter of both images on the device: the coefficients of
 * Interpolations' BSpline(Cubic(Line(OnGrid()))) (image_interpolant,
 * src/pipeline.jl:909-915), padded (nr+2) x (nc+2). */
int hh_prefilter(hh_ctx *ctx);
/* _deform_context(b, workspace, itpA, itpB, imgsize, T) (src/pipeline.jl:
 * 602-611; device form _ka_deform_context, src/ka_backend.jl:335-351): stage
 * the host-prefiltered interpolants' coefficients, parent(itp.itp.coefs) of
 * image_interpolant (src/pipeline.jl:909-915), (nr+2) x (nc+2) column-major
 * float32 each, logical index k at parent k + 1. Replaces the context's
 * coefficients (as hh_prefilter does) and fixes the image size for hh_deform;
 * a raw pair of another size is dropped. */
int hh_set_coefs(hh_ctx *ctx, const float *cA, const float *cB, int nr, int nc);
/* Symmetric deformation by a predictor grid (apply_predictor/deform_images,
 * src/pipeline.jl:575-600,936-965; _ka_deform!, src/ka_backend.jl:293-312).
 * py (gny knots, uniform), px (gnx knots, uniform), pu/pv gny x gnx host
 * column-major. Writes the device-resident warped pair (HH_SRC_WARP). */
int hh_deform(hh_ctx *ctx, const float *py, int gny, const float *px, int gnx,
              const float *pu, const float *pv);
/* Copy the current warped pair (or raw pair for src = HH_SRC_RAW) to host. */
int hh_get_images(hh_ctx *ctx, int src, float *A, float *B);
/* Copy the prefiltered coefficients ((nr+2) x (nc+2) each) to host. */
int hh_get_coefs(hh_ctx *ctx, float *cA, float *cB);

/* ---- per-window correlation ------------------------------------------- */

/* process_windows! (src/pipeline.jl:972-1027; device engine
 * ext/HammerheadAMDGPUExt.jl:398-493). jobs: int32 4 x njobs column-major
 * (gi, gj, rs, cs), 1-based, as pass_grid builds them. Writes RESIDUAL
 * displacements into u, v and peak_ratio, correlation_moment (ny x nx) and
 * the gauss3 alternatives into alt_u/alt_v (ny x nx x (n_peaks-1), entries
 * of absent peaks untouched). unc_u/unc_v (nullable) receive the fused
 * Wieneke uncertainty. Debug outputs (nullable): peak_locs int32
 * 2 x K x njobs (1-based plane row, col), peak_vals float K x njobs, found
 * int32 njobs, with K = max(n_peaks, 2). */
int hh_process_windows(hh_ctx *ctx, const hh_pass *pass, int src,
                       const int32_t *jobs, int njobs, int ny, int nx,
                       float *u, float *v, float *peak_ratio, float *correlation_moment,
                       float *alt_u, float *alt_v, float *unc_u, float *unc_v,
                       int32_t *peak_locs, float *peak_vals, int32_t *found);

/* correlation_plane! (src/correlators.jl:326-341) for each job: the
 * fftshifted, overlap-normalized plane, M x M x njobs float32 (M = 2w when
 * padded). The keep_correlation_planes inspection path. */
int hh_correlation_planes(hh_ctx *ctx, const hh_pass *pass, int src,
                          const int32_t *jobs, int njobs, float *planes);

/* uncertainty_sweep! (src/pipeline.jl:873-899; src/uncertainty.jl:42-179):
 * sigma_u/sigma_v per job; stats (nullable) float64 2 x 44 x njobs. */
int hh_uncertainty_sweep(hh_ctx *ctx, const hh_pass *pass, int src,
                         const int32_t *jobs, int njobs, int ny, int nx,
                         float *unc_u, float *unc_v, double *stats);

/* ---- validation and predictor ------------------------------------------ */

/* validate_and_replace! (src/pipeline.jl:617-645): UOD, peak-ratio and
 * extra validators, mask exclusion, peak substitution (alt_u/alt_v nullable:
 * ny x nx x (n_peaks-1), NaN where absent) and local-median replacement
 * when force_replace || replace_outliers. u, v edited in place; outliers
 * written. mask nullable (all false). */
int hh_validate_and_replace(hh_ctx *ctx, const hh_pass *pass, int ny, int nx,
                            int force_replace, float *u, float *v,
                            const float *peak_ratio, const float *correlation_moment,
                            const float *alt_u, const float *alt_v,
                            const uint8_t *mask, uint8_t *outliers);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). hh_validate_and_replace on DEVICE-resident grids: every
 * pointer is device memory (u, v in/out; peak_ratio, correlation_moment in;
 * alt_u/alt_v nullable, ny x nx x (n_peaks-1); mask nullable, one byte per
 * cell; outliers out), the work is enqueued on the context stream
 * (hh_stream) and the call returns without waiting. Its scratch (a counter
 * and the substituted/replaced u, v and flags) is a fixed device array of the
 * library, one per GPU, for grids of at most HH_NOALLOC_MAX_CELLS cells
 * (larger grids: HH_EINVAL); it allocates nothing. Calls that overlap on one
 * GPU (from several contexts or streams) share that array and race; order
 * them on one stream. */
#define HH_NOALLOC_MAX_CELLS (1 << 20)
int hh_validate_and_replace_noalloc(hh_ctx *ctx, const hh_pass *pass, int ny, int nx, int force_replace, float *u,
                                    float *v, const float *peak_ratio, const float *correlation_moment,
                                    const float *alt_u, const float *alt_v, const uint8_t *mask, uint8_t *outliers);
/* smooth_field (src/quality.jl:123-136). */
int hh_smooth_field(hh_ctx *ctx, int ny, int nx, const float *f, float *out);

/* ---- whole schedule ---------------------------------------------------- */

/* run_piv(imgA, imgB, passes; predictor_smoothing) (src/pipeline.jl:361-449)
 * on the current pair, entirely on the device. If result != NULL the final
 * PIVResult is copied to its host arrays (result->ny/nx must match the final
 * pass grid); with NULL the result stays device-resident (hh_fetch_result). */
int hh_run_piv(hh_ctx *ctx, const hh_pass *passes, int npasses,
               int predictor_smoothing, hh_field *result);
/* Copy the device-resident result of the last hh_run_piv or
 * hh_run_piv_ensemble to host. */
int hh_fetch_result(hh_ctx *ctx, hh_field *result);

/* The mask and mask_threshold keywords of run_piv (src/pipeline.jl:361-449,
 * checks at 393-396), run_piv_sequence (src/io.jl:229-247) and
 * run_piv_ensemble (src/ensemble.jl:47-84,151-154): an nr x nc column-major
 * pixel mask (nonzero = masked) applied to every following hh_run_piv /
 * hh_run_piv_batch / hh_run_piv_ensemble, whose pairs must all be nr x nc
 * (HH_EDIM at the run otherwise). Windows whose masked-pixel fraction reaches
 * mask_threshold produce no vector (pass_grid, src/pipeline.jl:542-549):
 * u, v, peak_ratio, correlation_moment and the uncertainties hold NaN there,
 * they are never outliers and result->mask flags them; windows with fewer
 * masked pixels are correlated from the mean of their valid pixels with the
 * masked ones as zero (load_windows!, src/correlators.jl:214-227). mask NULL
 * (or one with no pixel set) removes the mask; the threshold must be in
 * (0, 1] (HH_EINVAL). The mask stays set across hh_set_pair* and
 * hh_set_pairs; setting an identical mask again keeps the planned schedule.
 * The per-hook entry points (hh_process_windows, hh_correlation_planes,
 * hh_uncertainty_sweep, hh_ensemble_add) load every job's window with it as
 * load_windows! does (valid-pixel means, masked pixels zero); their jobs are
 * the caller's (pass_grid already left the masked windows out). */
int hh_set_mask(hh_ctx *ctx, const uint8_t *mask, int nr, int nc, double mask_threshold);

/* ---- batches of pairs of uneven sizes ----------------------------------- */

/* The frames of run_piv_sequence(pairs, params) (src/io.jl:229-247): npairs
 * host column-major pairs, pair q of nr[q] x nc[q], copied into one
 * device-resident image atlas. Replaces the context's current pair. With
 * the same pair sizes in the same order as the resident batch, only the
 * frames change: the planned schedule (and its captured graph) is kept. */
int hh_set_pairs(hh_ctx *ctx, const float *const *A, const float *const *B, const int32_t *nr,
                 const int32_t *nc, int npairs);
/* run_piv(A[q], B[q], passes; predictor_smoothing) for every pair of the
 * batch (the per-pair loop of run_piv_sequence, src/io.jl:229-247, with the
 * body of src/pipeline.jl:361-449): each sweep runs one launch per stage over
 * all pairs. Results stay on the device; results may be NULL or an array of
 * npairs fields, filled as hh_fetch_pair does. Each pair keeps its own
 * convergence (field_change, src/pipeline.jl:842-866). */
int hh_run_piv_batch(hh_ctx *ctx, const hh_pass *passes, int npasses, int predictor_smoothing,
                     hh_field *results);
/* The PIVResult of pair q of the last hh_run_piv_batch (or q = 0 after
 * hh_run_piv). */
int hh_fetch_pair(hh_ctx *ctx, int q, hh_field *result);
/* Grid size of pair q's final pass for the last planned schedule. */
int hh_pair_grid(hh_ctx *ctx, int q, int *ny, int *nx);

/* Schedules whose passes all have a fixed sweep count (max_iterations = 1
 * or convergence_tol = 0) are captured once into a hipGraph per context and
 * replayed; 0 disables capture (every call launches eagerly). Returns the
 * number of graph replays so far in *replays when non-NULL. */
int hh_set_graphs(hh_ctx *ctx, int enable);
int hh_graph_replays(hh_ctx *ctx, int64_t *replays);

/* Autotuner of the correlation kernels (level 1 by default): while a
 * schedule is planned, each pass's correlation kernel is timed over its own
 * job list for every candidate and the fastest is kept per context.
 * Level 1 tunes the launch geometry: resident MFMA waves per SIMD (which
 * sets the windows each persistent wave and workgroup walks, and for the
 * 16-px matrix-core kernel its register budget: 3 waves at 168 registers or
 * 4 at 128), workgroup size of the 32-px matrix-core kernel (4 or 8 waves),
 * window-to-XCD order;
 * results do not depend on the choice. Level 2 also chooses the kernel
 * family per FFT side among the form's own (the split-FP16 matrix-core DFT
 * in the MFMA form), the register FFT and the LDS radix-2 FFT / direct DFT
 * kernels, the uncertainty launch included in each family's time; results
 * then differ between choices at the arithmetic's round-off (each family
 * meets the same parity bars). 0 disables tuning. hh_tune_report writes the
 * choices and timings as a JSON object (NUL-terminated, truncated to len;
 * *needed = full size). No counterpart in the reference (its device batch
 * size is a fixed keyword, src/ka_backend.jl). */
int hh_set_autotune(hh_ctx *ctx, int enable);
int hh_tune_report(hh_ctx *ctx, char *buf, size_t len, size_t *needed);
/* Preload one autotuner choice (a key of hh_tune_report) so that plans use it
 * without timing candidates, e.g. choices saved by an earlier process. */
int hh_tune_set(hh_ctx *ctx, const char *key, int waves_per_simd, int linear_order);
/* Preload the kernel family of a key (0 the form's own, 1 register FFT,
 * 2 LDS FFT / direct DFT), as the "family" of hh_tune_report. */
int hh_tune_set_family(hh_ctx *ctx, const char *key, int family);

/* ---- ensemble (sum-of-correlation) PIV --------------------------------- */

/* _plane_accumulator / _uncertainty_accumulator of ensemble_pass
 * (src/ensemble.jl:165-190; device form ext/HammerheadAMDGPUExt.jl:550-560):
 * start FP64 per-window plane sums for jobs (int32 4 x njobs, 1-based) of
 * one pass; with uncertainty the Wieneke statistics are pooled too. The sums
 * live on the device until hh_ensemble_analyze. */
int hh_ensemble_begin(hh_ctx *ctx, const hh_pass *pass, const int32_t *jobs, int njobs, int uncertainty);
/* accumulate_planes! (src/ensemble.jl:317-343; device form
 * ext/HammerheadAMDGPUExt.jl:561-621): add the correlation planes of the
 * current pair (src = HH_SRC_RAW, or HH_SRC_WARP after hh_deform). */
int hh_ensemble_add(hh_ctx *ctx, int src);
/* ensemble_analyze! (src/ensemble.jl:284-313; device form
 * ext/HammerheadAMDGPUExt.jl:622-670): peaks of the summed planes. u, v
 * (ny x nx) hold the shared predictor at the nodes on entry and the total
 * displacement on exit; alt_u/alt_v (nullable, ny x nx x (n_peaks-1)) get
 * predictor-relative alternatives (NaN where absent); unc_u/unc_v receive
 * the pooled uncertainty when the ensemble was begun with it. */
int hh_ensemble_analyze(hh_ctx *ctx, int ny, int nx, float *u, float *v, float *peak_ratio,
                        float *correlation_moment, float *alt_u, float *alt_v, float *unc_u, float *unc_v);
/* The current FP64 plane sums, M x M x njobs (fftshifted, as
 * correlation_plane!): the keep_correlation_planes view of an ensemble. */
int hh_ensemble_planes(hh_ctx *ctx, double *planes);
/* run_piv_ensemble(pairs, passes; predictor_smoothing)
 * (src/ensemble.jl:47-84) on the device: A[q], B[q] are host column-major
 * float32 nr x nc frames of pair q, uploaded once and kept resident in HBM
 * for every pass. With A = B = NULL the pairs of the previous call (same
 * npairs, nr, nc) are used again without an upload. The context's pixel
 * mask (hh_set_mask), when set, applies as the mask keyword does. The final
 * result is copied to `result` when non-NULL, and stays available to
 * hh_fetch_result. */
int hh_run_piv_ensemble(hh_ctx *ctx, const float *const *A, const float *const *B, int npairs, int nr, int nc,
                        const hh_pass *passes, int npasses, int predictor_smoothing, hh_field *result);

/* ---- correlation energy and its adjoint -------------------------------- */
/* No counterpart in the reference, which has no gradient path (nearest:
 * the non-gradient correlate_deformable, src/correlators.jl:620-638). For a
 * node field pu, pv (gny x gnx, column-major) on uniform knots py, px, the
 * current pair is warped symmetrically as in deform_images
 * (src/pipeline.jl:936-965: bilinear Flat() node values, cubic B-spline
 * samples A(r - v/2, c - u/2), B(r + v/2, c + u/2); bilinear samples of the
 * pair instead after hh_set_energy_sampling(HH_SAMPLE_BILINEAR)), and over the windows of
 * `pass` (square, 8/16/32 px; grid and apodization as pass_grid)
 *     E = - sum_w sum_{x in w} p_x (A^_x - mean_w A^) (B^_x - mean_w B^).
 * The prefiltered pair is computed on first use (hh_prefilter). */
int hh_energy(hh_ctx *ctx, const hh_pass *pass, const float *py, int gny, const float *px, int gnx,
              const float *pu, const float *pv, double *energy);
/* Image sampling of the warped windows of hh_energy, hh_grad_energy and
 * hh_refine_nodes: HH_SAMPLE_CUBIC (default; the cubic B-spline interpolant
 * of the deformation path) or HH_SAMPLE_BILINEAR (bilinear sampling of the
 * pair itself, zero outside [1, n]). No counterpart in the reference. */
enum { HH_SAMPLE_CUBIC = 0, HH_SAMPLE_BILINEAR = 1 };
int hh_set_energy_sampling(hh_ctx *ctx, int sampling);
/* E and its gradient dE/dpu, dE/dpv (gny x gnx) by the hand-written adjoint,
 * which recomputes the warped windows instead of storing them. energy is
 * nullable. */
int hh_grad_energy(hh_ctx *ctx, const hh_pass *pass, const float *py, int gny, const float *px, int gnx,
                   const float *pu, const float *pv, double *energy, float *grad_u, float *grad_v);
/* Gradient-based refinement of a node field: niter descent steps with a
 * backtracking step length (the largest node moves at most 1 px per step);
 * pu, pv updated in place; energies (nullable, niter + 1) receives E before
 * each step and at the end. */
int hh_refine_nodes(hh_ctx *ctx, const hh_pass *pass, const float *py, int gny, const float *px, int gnx,
                    float *pu, float *pv, int niter, double *energies);

/* ---- stream and timing ------------------------------------------------- */

/* The context's HIP stream (hipStream_t), for callers that time or order
 * work around the library. */
void *hh_stream(hh_ctx *ctx);
int hh_synchronize(hh_ctx *ctx);
/* Per-kernel timing with HIP events on the context stream: enable, then read
 * total milliseconds and launch count of a kernel family by name
 * ("correlate", "deform", "prefilter", "validate", "uq"). */
int hh_set_timing(hh_ctx *ctx, int enable);
/* Timing of one kernel family and window only (family NULL: every family,
 * as hh_set_timing(ctx, 1)); the other launches carry no events. No
 * counterpart in the reference: for benchmarks that time one kernel inside
 * an otherwise uninstrumented schedule. */
int hh_set_timing_only(hh_ctx *ctx, const char *family, int window);
int hh_kernel_time(hh_ctx *ctx, const char *family, int window, double *ms_total,
                   int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* HAMMERHEAD_MI355X_H */
