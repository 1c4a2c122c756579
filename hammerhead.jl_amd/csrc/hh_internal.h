// hh_internal.h — shared definitions of the gfx950 WIDIM backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hammerhead_mi355x.h"

namespace hh {

// Error carrier: every launcher returns HH_OK or sets a message.
struct Status {
    int code = HH_OK;
    std::string msg;
};

#define HH_HIP(expr)                                                                    \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            throw ::hh::Error(HH_EDEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
        }                                                                               \
    } while (0)

struct Error {
    int code;
    std::string msg;
    Error(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// ---- per-pass correlation configuration ------------------------------------
constexpr int kMaxPeaks = 8;   // max(n_peaks, 2) supported on the device
constexpr int kMaxM = 128;     // largest FFT side

struct CorrConfig {
    int w;          // square window side
    int s;          // square search-area side (frame B), s >= w; s = w: equal areas
    int M;          // FFT side: s, or 2s when padded
    int padded;
    int method;     // HH_CROSS / HH_PHASE
    int subpixel;   // HH_GAUSS3 / HH_GAUSS9 / HH_SUBPIX_NONE
    int finder;     // HH_REGIONALMAX / HH_EXCLUSION
    int n_peaks;
    int K;          // max(n_peaks, 2)
    // scan box of the analysis in fftshifted plane coordinates (0-based,
    // inclusive): outside the overlap-gain support the plane is exactly 0
    int box_lo, box_hi;
    const float *apod;  // w*w device
    const float *gain;  // M*M device (fftshifted), nullptr if not padded
    const float *wfilt; // M*M device phase filter (unshifted), or nullptr
    int form;           // HH_FORM_FFT32 / HH_FORM_MFMA16
    // launch geometry chosen by the autotuner (0: the kernel's default);
    // results do not depend on it
    int waves_per_simd; // persistent MFMA kernel: resident waves per SIMD
    int linear_order;   // 1: windows in launch order instead of XCD bands
    // kernel family (the autotuner's algorithm choice): 0 the form's own
    // (matrix-core DFT where it covers the pass in the MFMA form, else the
    // register FFT, else the LDS kernels), 1 the register FFT (no matrix
    // cores), 2 the LDS radix-2 FFT / direct DFT kernels
    int family;
};

// Kernel family a correlation launch of cfg takes ("mfma32", "regL1",
// "regL2", "regL4", "lds"), with or without fused uncertainty.
const char *correlate_family(const CorrConfig &cfg, bool uq);

// Outputs of a correlation launch; any pointer may be null.
struct CorrOut {
    float *u, *v, *pr, *cm;          // ny*nx grids
    float *alt_u, *alt_v;            // ny*nx*(n_peaks-1)
    const float *u0, *v0;            // predictor node values added (nullable)
    int32_t *locs;                   // 2*K*njobs
    float *vals;                     // K*njobs
    int32_t *found;                  // njobs
    float *planes;                   // M*M*njobs
    float *unc_u, *unc_v;            // fused Wieneke uncertainty (nullable)
    // ensemble accumulation: when set, the launch only adds each window's
    // plane into its FP64 sum (layout of correlate_box_layout) and writes
    // nothing else
    double *acc;
    int ny, nx;
};

// ---- launchers (k_*.hip) ----------------------------------------------------
// A second stream on which a launch may run independent work concurrently
// with the caller's stream: forked after the work already queued on the
// caller's stream and joined back before anything queued after the launch
// (both streams at the default priority).
struct SideStream {
    hipStream_t stream;
    hipEvent_t fork, join;
    // the caller makes its stream wait for `join` itself, later
    bool defer_join;
};

// Returns true when out.unc_u/unc_v were filled by the same launch (fused
// uncertainty); otherwise the caller runs launch_uq. With `side`, the
// statistics of an MFMA-form sweep run on the side stream beside the
// correlation.
bool launch_correlate(const CorrConfig &cfg, const float *A, const float *B, int nr, int nc, const int32_t *jobs,
                      int njobs, const CorrOut &out, hipStream_t s, const SideStream *side = nullptr);

// Builds the constant tables a correlation launch of cfg may need (device
// allocations happen here, never inside a launch sequence).
void correlate_prepare(const CorrConfig &cfg, hipStream_t s);

// Plane-sum layout of the accumulate mode: true when the specialized kernels
// take the configuration and add only the analysis box [box_lo, box_hi]^2
// (box side nb = box_hi - box_lo + 1, column-major, nb^2 doubles per
// window; the plane is exactly zero outside it), false for the whole
// M x M plane.
bool correlate_box_layout(const CorrConfig &cfg);

// ensemble_analyze!: peaks, subpixel, ratio, moment and alternatives of
// summed planes (acc in the layout above, rounded to float), written like a
// correlation launch (out.u0/v0 the shared predictor).
void launch_plane_analyze(const CorrConfig &cfg, const double *acc, bool box, const int32_t *jobs, int njobs,
                          const CorrOut &out, hipStream_t s);

// Compute units of the current device (cached after the first call).
int device_cu_count();

// Split-FP16 MFMA DFT kernels (HH_FORM_MFMA16); false when not covered.
bool launch_correlate_mfma(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs,
                           int njobs, const CorrOut &out, hipStream_t s);
// The 64-px matrix-core kernel (circular length 96, cross correlation), its
// coverage and its device tables (built by mfma96_prepare outside capture).
bool mfma96_covers(const CorrConfig &cfg);
void mfma96_prepare(hipStream_t s);
bool launch_correlate_mfma96(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs,
                             int njobs, const CorrOut &out, hipStream_t s);

// Specialized kernels; return false when the configuration is not covered.
bool launch_correlate_fast(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs,
                           int njobs, const CorrOut &out, hipStream_t s, bool *uq_done);

// ---- vector grids of one or several pairs -----------------------------------
// The cells of several vector grids stored back to back in one set of
// buffers: segment s is one pair's ny x nx column-major grid at buffer cell
// `off`. A launch enumerates the cells of its segments (all of them, or the
// pairs still iterating) as 0..total-1, segment s from `start`. Kernels over
// grids map an enumerated cell to (segment, local row, local column) and
// its buffer cell off + local; neighbourhoods never cross a segment.
struct GridSeg {
    int start, off, ny, nx, pair;
    int pad[3];
};
struct Grids {
    const GridSeg *seg;  // device table of nseg segments, or nullptr: one grid
    int nseg;
    int ny, nx;          // the single grid when seg is null
    int total;           // enumerated cells
    int stride;          // buffer cells (the stride of the alternatives)
};
inline Grids one_grid(int ny, int nx) { return Grids{nullptr, 1, ny, nx, ny * nx, ny * nx}; }

// Segment of enumerated cell e: the last s with seg[s].start <= e.
__device__ __forceinline__ GridSeg grid_seg(const Grids &g, int e, int &s)
{
    if (!g.seg) {
        s = 0;
        return GridSeg{0, 0, g.ny, g.nx, 0, {0, 0, 0}};
    }
    int lo = 0, hi = g.nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g.seg[mid].start <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    s = lo;
    return g.seg[lo];
}

// Enumerated cell e: its segment, local row and column, and buffer cell.
struct Cell {
    GridSeg sg;
    int si, r, c, b;
};
__device__ __forceinline__ Cell cell_of(const Grids &g, int e)
{
    Cell k;
    k.sg = grid_seg(g, e, k.si);
    const int el = e - k.sg.start;
    k.r = el % k.sg.ny;
    k.c = el / k.sg.ny;
    k.b = k.sg.off + el;
    return k;
}

// Deformation of one pair of a batch (deform_images, src/pipeline.jl:936-965):
// its prefiltered coefficients, the predictor grid on uniform knots, and
// where its warped pixels go in the (column-major, leading dimension ld)
// warp atlas.
struct DeformPair {
    const float *cA, *cB;  // (nr+2) x (nc+2) each
    const float *pu, *pv;  // gny x gnx
    int nr, nc, gny, gnx;
    float y0, ysp, x0, xsp;
    long long out_off;     // atlas offset of pixel (1, 1)
    int blk0;              // first block of the pair in a batched launch
    int pad;
};
// Blocks of 256 threads per column of an nr-row image in k_deform (one
// block per 256-row strip of a column).
__host__ __device__ inline int deform_strips(int nr) { return (nr + 255) / 256; }
// The predictor node values of a pass (predictor_node_values on the pass's
// window centres, see launch_node_values_batch), optionally run by extra
// blocks of the deformation launch: the two read the same predictor and are
// independent, so one launch serves both.
struct NodeValuesJob {
    Grids g;
    const DeformPair *pairs;  // device table (segment s reads pairs[seg.pair])
    int r0, rstep, c0, cstep;
    float half;
    float *u, *v;
    int nblocks;              // 0: none
};
struct DeformTab {
    const DeformPair *pairs;  // device table, or nullptr: `one`
    int npairs;
    int nblocks;
    DeformPair one;
    NodeValuesJob nv;
};

// img: nr x nc with leading dimension ld.
// Both frames of a pair, two launches (one per axis) covering both frames;
// tmp holds 2 x (nr+2) x nc floats.
void launch_prefilter_pair(const float *imgA, const float *imgB, int ld, int nr, int nc, float *coefA, float *coefB,
                           float *tmp, hipStream_t s);
void launch_deform(const float *cA, const float *cB, int nr, int nc, const float *pu, const float *pv, int gny,
                   int gnx, float y0, float ysp, float x0, float xsp, float *wA, float *wB, hipStream_t s);
// All pairs of a batch in one launch (ld: leading dimension of the atlas).
void launch_deform_batch(const DeformTab &tab, int ld, float *wA, float *wB, hipStream_t s);
void launch_node_values(const float *pu, const float *pv, int gny, int gnx, float y0, float ysp, float x0, float xsp,
                        const float *y, int ny, const float *x, int nx, float *u, float *v, hipStream_t s);
// predictor_node_values of every segment of g (src/pipeline.jl:561-569):
// the window centres of a pass are y = r0 + i rstep + half, x = c0 + j cstep
// + half (0-based i, j); segment s reads the predictor of pairs[seg.pair].
void launch_node_values_batch(const Grids &g, const DeformPair *pairs, int r0, int rstep, int c0, int cstep,
                              float half, float *u, float *v, hipStream_t s);
// The same work as a job for extra blocks of a deformation launch
// (DeformTab::nv).
NodeValuesJob node_values_job(const Grids &g, const DeformPair *pairs, int r0, int rstep, int c0, int cstep,
                              float half, float *u, float *v);

struct ValidateArgs {
    Grids grids;
    const hh_pass *pass;     // host copy (scalars are passed by value)
    int force_replace;
    float *u, *v;            // in/out
    const float *pr, *cm;
    const float *alt_u, *alt_v;  // nullable, grids.stride*(n_peaks-1)
    uint8_t *outliers;       // out
    // scratch (device): grids.stride each
    float *su, *sv;
    uint8_t *sflag;
    int *count;  // grids.nseg device ints: valid cells per segment, saturating at 3
    // grid mask (nullable, grids.stride bytes): masked cells are excluded
    // from the UOD, never flagged or substituted, and NaN after a replacement
    const uint8_t *mask;
    // predictor (nullable): with no mask and a replacement stage, the
    // replacement launch also writes build_predictor's smoothing of the
    // replaced field here (launch_validate then returns true)
    float *smooth_u, *smooth_v;
};
bool launch_validate(const ValidateArgs &a, hipStream_t s);
// u, v, peak ratio and correlation moment of the masked cells set to NaN.
void launch_mask_nan(const Grids &g, const uint8_t *mask, float *u, float *v, float *pr, float *cm, hipStream_t s);
// build_predictor's fill (src/pipeline.jl:454-459): the masked cells of
// (u, v) replaced from the unmasked ones (replace_vectors!) into (ou, ov);
// count: grids.nseg device ints of scratch.
void launch_fill_masked(const Grids &g, const float *u, const float *v, const uint8_t *mask, int *count, float *ou,
                        float *ov, hipStream_t s);
// Windows with masked pixels (load_windows!, src/correlators.jl:214-262):
// for each of the nw entries of mjobs (4 ints, 0-based: image row, image
// column, mask row, mask column of the window's first pixel) an sz x sz
// tile of each frame at column t sz of the tile atlases tA, tB (leading
// dimension sz): frame A's w x w window at offset (sz - w) / 2, frame B's
// sz x sz area around it, valid pixels minus the mean of their region's
// valid pixels, masked pixels (and the rest of tile A) 0. sz = w gives
// window tiles for both frames (the uncertainty's load). mask: column-major,
// leading dimension mld.
void launch_mask_tiles(const float *A, const float *B, int ld, const uint8_t *mask, int mld, const int32_t *mjobs,
                       int nw, int w, int sz, float *tA, float *tB, hipStream_t s);
void launch_smooth(const Grids &g, const float *f, float *out, hipStream_t s);
// Two fields (u and v of a predictor) in one launch.
void launch_smooth2(const Grids &g, const float *f, const float *f2, float *out, float *out2, hipStream_t s);

// stats (nullable): 2 x 44 FP64 per job, overwritten, or added to when
// accumulate (pooled statistics of the ensemble).
void launch_uq(int w, const float *apod, const float *A, const float *B, int nr, const int32_t *jobs, int njobs,
               int ny, float *unc_u, float *unc_v, double *stats, bool accumulate, hipStream_t s);
// finalize_uncertainty of pooled statistics (2 x 44 per job).
void launch_uq_finalize(const double *stats, const int32_t *jobs, int njobs, int ny, float *unc_u, float *unc_v,
                        hipStream_t s);

// ---- correlation energy and its adjoint (k_adjoint.hip) ---------------------
struct EnergyArgs {
    const float *cA, *cB;  // prefiltered pair, (nr+2) x (nc+2) each (cubic sampling)
    const float *A, *B;    // the pair itself, nr x nc each (bilinear sampling)
    int bilinear;          // image sampling: 0 cubic B-spline, 1 bilinear
    int nr, nc;
    const float *pu, *pv;  // node field gny x gnx on uniform knots
    int gny, gnx;
    float y0, ysp, x0, xsp;
    int w, ny, nx, r0, c0, st_r, st_c;  // windows of the pass
    const float *apod;                   // w x w
};
// Per-window statistics (mA, mB, PA/N, PB/N) and energies, and E in *esum.
void launch_energy(const EnergyArgs &g, float4 *stats, float *ew, double *esum, hipStream_t s);
// dE/dpu, dE/dpv from the statistics of launch_energy; rows/cols: the
// pixel range of each node interval (ncy, ncx of them), part: 8 per cell.
void launch_energy_adjoint(const EnergyArgs &g, const float4 *stats, const int2 *rows, const int2 *cols, int ncy,
                           int ncx, float *part, float *gu, float *gv, hipStream_t s);

// dewarp! of one or two prefiltered frames (cA/cB: (nr+2) x (nc+2)) at the
// 16.16 fixed-point source map (ny x nx int2, INT32_MIN = out of view) into
// oA/oB (ny x nx); cB/oB nullable.
void launch_dewarp(const int2 *map, int ny, int nx, const float *cA, const float *cB, int nr, int nc, float *oA,
                   float *oB, hipStream_t s);
}  // namespace hh
