// hh_ctx.h — host-side state of a context (the engine pool of one
// PIVWorkspace, src/pipeline.jl:104-137) and the helpers shared by the C ABI
// (hh_api.cpp) and the schedule engine (hh_engine.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "hh_internal.h"

namespace hh {
namespace host {

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T *ensure(size_t m)
    {
        if (m > n) {
            release();
            if (m > 0) {
                hipError_t e = hipMalloc(&p, sizeof(T) * m);
                if (e != hipSuccess) {
                    p = nullptr;
                    throw Error(HH_ENOMEM, std::string("device allocation failed: ") + hipGetErrorString(e));
                }
            }
            n = m;
        }
        return p;
    }
};

// Per window-configuration device constants (apodization, overlap gain,
// phase filter), the engine pool of piv_correlation_engines
// (src/pipeline.jl:104-137).
struct ConfigBufs {
    DBuf<float> apod, gain, wfilt;
    int box_lo = 0, box_hi = 0;
};

struct TimedLaunch {
    std::string family;
    int w;
    hipEvent_t a, b;
};

struct PassGrid {
    int ny = 0, nx = 0;
    std::vector<float> x, y;
    std::vector<int32_t> jobs;
};

// ---- schedule engine state (hh_engine.cpp) ----------------------------------

// The pairs of one run: a single pair (hh_set_pair*) or a batch of pairs of
// uneven sizes (hh_set_pairs) packed into one column-major image atlas.
struct Layout {
    int npairs = 0;
    int ld = 0;                                   // atlas rows (leading dimension)
    long long ncols = 0;                          // atlas columns
    std::vector<int> nr, nc, roff, coff;          // per pair: size and atlas origin (0-based)
    std::vector<size_t> coef_off;                 // per pair: offset of its coefficient arrays
    size_t coef_total = 0;                        // floats per image over all pairs
    const float *A = nullptr, *B = nullptr;       // atlas of the first / second frames (device)
    uint64_t gen = 0;                             // changes whenever the layout changes
};

// Device tables of one pass over a set of pairs.
struct PassTables {
    DBuf<int32_t> jobs;       // 4 x njobs: (global cell + 1, 1, atlas row, atlas column)
    int njobs = 0;
    DBuf<GridSeg> segs;
    Grids grids{};
    DBuf<DeformPair> dprev;   // predictor: the previous pass's grids
    DBuf<DeformPair> dself;   // predictor: this pass's grids
    DeformPair hprev{}, hself{};  // host copies when the tables hold one pair
    int blocks = 0;           // deformation blocks
    DBuf<int> count;          // one valid-cell counter per segment (saturating at 3)
    // windows with masked pixels below the grid-mask threshold: the tile
    // builder's list (4 x npart, 0-based image row / column and mask row /
    // column of the window), the correlation jobs on the s x s tiles (cell +
    // 1, 1, 1 + m, t s + 1 + m; m the search margin) and the uncertainty's
    // jobs on w x w window tiles (cell + 1, 1, 1, t w + 1)
    DBuf<int32_t> mjobs, tjobs, ujobs;
    int npart = 0;
};

struct PassPlan {
    hh_pass p{};
    CorrConfig cfg{};       // plain sweeps
    CorrConfig cfg_uq{};    // the sweep with fused uncertainty
    int r0 = 1, rstep = 1, c0 = 1, cstep = 1;
    float half = 0.f;                        // (w - 1) / 2
    std::vector<int> ny, nx, off;            // per pair grid and cell offset
    int total = 0;                           // cells over all pairs
    bool fixed = true;                       // sweep count known in advance
    PassTables all;
    // grid mask of pass_grid (src/pipeline.jl:542-549): every pair has the
    // mask's size, so one ny x nx grid (host) serves all; the device copy
    // covers the buffer cells of every pair
    std::vector<uint8_t> gm;
    DBuf<uint8_t> gmask;
    bool anyg = false;
};

// Host copies of one pass's job lists (rebuilt into subsets when pairs stop
// iterating).
struct HostJobs {
    std::vector<int32_t> jobs;
    std::vector<int> off, cnt;  // per pair: index of its first job, job count
    // windows with masked pixels, 5 ints each: buffer cell, image row and
    // column in the atlas, row and column in the mask (0-based)
    std::vector<int32_t> part;
    std::vector<int> poff, pcnt;
};

// Autotuned launch geometry, keyed by kernel family and problem shape: the
// chosen candidate and the measured milliseconds of every candidate.
struct TuneEntry {
    int waves_per_simd = 0, linear_order = 0;
    int family = 0;  // CorrConfig::family of the choice
    std::vector<std::pair<std::string, float>> ms;
};
struct Tuner {
    bool enabled = true;
    // 1: launch geometry only (results do not depend on the choice);
    // 2: also the kernel family per FFT side (results may then differ at
    // the arithmetic's round-off between choices)
    int level = 1;
    std::map<std::string, TuneEntry> entries;
};

struct Engine {
    Layout lay;
    DBuf<float> atlas;       // owned batch atlas: frames A then B
    DBuf<float> stage;       // upload staging of a frame narrower than the atlas
    DBuf<float> coefs;       // coefficient arrays: all A, then all B
    DBuf<float> warp;        // warp atlas: A then B
    DBuf<float> ptmp;        // prefilter scratch
    DBuf<float> u, v, pr, cm, altu, altv, uu, uv, u0, v0, su, sv, mu, mv, prevu, prevv, pu, pv, ch;
    DBuf<float> zero;        // zeros for the first sweep's predictor (cleared when planned, never written)
    DBuf<uint8_t> out, sflag;
    // current plan
    std::vector<uint8_t> key;
    std::vector<std::unique_ptr<PassPlan>> plan;
    std::vector<HostJobs> hjobs;
    bool deforms = false;
    int smoothing = 1;
    size_t maxcells = 0, maxalt = 0;
    hipGraphExec_t graph = nullptr;
    bool graph_failed = false;
    bool use_graphs = true;
    int graph_replays = 0;
    bool have_result = false;
    // tables of an active subset of a sweep (dynamic convergence)
    PassTables sub;
    // pixel mask of run_piv(...; mask, mask_threshold), shared by every pair
    // (host and device copies, column-major mnr x mnc), its threshold, and a
    // generation that changes whenever either changes
    std::vector<uint8_t> hmask;
    int mnr = 0, mnc = 0;
    double mthr = 0.5;
    uint64_t mgen = 0;
    DBuf<uint8_t> dmask;
    DBuf<float> tiles;       // tile atlas of the masked windows: A then B
    size_t maxtile = 0;      // floats per frame
    DBuf<int> mcount;        // scratch counters of the predictor's fill
};

}  // namespace host
}  // namespace hh

struct hh_ctx {
    int device = 0;
    int ncu = 256;
    hipStream_t stream = nullptr;
    // the side stream of concurrent launches within a schedule (default
    // priority), with its fork and join events, and the join event of the
    // prefilter that runs beside the first pass
    hh::SideStream side{};
    hipEvent_t pre_join = nullptr;
    // device address of this GPU's no-allocation scratch (hh_create)
    unsigned char *scratch = nullptr;
    // dewarping (hh_set_dewarp_map): 16.16 fixed-point source map of the
    // dw_ny x dw_nx world grid into dw_nr x dw_nc camera frames; raw frames,
    // their coefficients and the prefilter scratch
    hh::host::DBuf<int2> dwmap;
    int dw_ny = 0, dw_nx = 0, dw_nr = 0, dw_nc = 0;
    hh::host::DBuf<float> dwraw, dwcoef, dwtmp, dwout;
    // image sampling of the correlation energy (hh_set_energy_sampling)
    bool energy_bilinear = false;
    std::string err;
    int form = HH_FORM_FFT32;
    int nr = 0, nc = 0;
    const float *A = nullptr, *B = nullptr;  // current pair (owned or borrowed)
    hh::host::DBuf<float> ownA, ownB;
    hh::host::DBuf<float> coefA, coefB, ptmp;
    bool coefs_valid = false;
    hh::host::DBuf<float> warpA, warpB;
    bool warp_valid = false;
    // predictor grid (device) + knots
    hh::host::DBuf<float> pu, pv;
    int gny = 0, gnx = 0;
    float y0 = 0, ysp = 1, x0 = 0, xsp = 1;
    std::map<std::tuple<int, int, int, int, int>, std::unique_ptr<hh::host::ConfigBufs>> configs;
    // job lists per (w, search side, overlap, nr, nc)
    std::map<std::tuple<int, int, int, int, int, int>, std::unique_ptr<hh::host::DBuf<int32_t>>> jobcache;
    // chain buffers of the ensemble driver
    hh::host::DBuf<float> gu, gv, gpr, gcm, galtu, galtv, guu, guv, gsu, gsv, gpx, gpy;
    hh::host::DBuf<uint8_t> gout, gsflag, gmask;
    hh::host::DBuf<int> gcount;
    hh::host::DBuf<int32_t> hookjobs, hookmask[2];
    hh::host::DBuf<float> hooktiles[2];  // masked-window tiles of the per-hook entry points (two loads)
    hh::host::DBuf<float> hk[12];
    hh::host::DBuf<double> hkd;
    hh::host::DBuf<int32_t> hki[2];
    // ensemble accumulation (run_piv_ensemble): FP64 plane sums per window
    // in the layout of correlate_box_layout, pooled uncertainty statistics
    struct {
        bool active = false;
        hh_pass pass{};
        hh::CorrConfig cfg{};
        bool box = false, unc = false;
        size_t per = 0;
        int njobs = 0, npairs = 0;
        const int32_t *jobs = nullptr;       // device
        std::vector<int32_t> hjobs;          // host copy for bounds checks
        // with a pixel mask: the first nclean jobs read the images, the
        // npart after them are windows with masked pixels, accumulated from
        // centred tiles (tjobs: their jobs on the tile atlas, mjobs: the
        // tile builder's list)
        int nclean = 0, npart = 0;
        const int32_t *tjobs = nullptr, *mjobs = nullptr, *ujobs = nullptr;  // ujobs: window tiles (statistics)
    } ens;
    hh::host::DBuf<int32_t> ensjobs, enspart;
    hh::host::DBuf<float> enstiles;
    std::vector<uint8_t> final_gm;  // final grid mask of hh_run_piv_ensemble (empty: none)
    hh::host::DBuf<double> ensacc, ensstats;
    hh::host::DBuf<float> arena;  // the pairs of hh_run_piv_ensemble, resident in HBM
    int arena_pairs = 0, arena_nr = 0, arena_nc = 0;
    hipStream_t copy_stream = nullptr;        // pair uploads overlapped with the first pass
    std::vector<hipEvent_t> upload_ev;        // ring of upload-complete events
    // correlation energy and its adjoint
    hh::host::DBuf<float4> estats;
    hh::host::DBuf<float> ew, epart, egrad, enode;
    hh::host::DBuf<double> esum;
    hh::host::DBuf<int2> ecells;
    // final result of hh_run_piv_ensemble (ensemble chain buffers)
    hh::host::PassGrid final_grid;
    bool have_result = false;
    // run_piv / batch engine
    hh::host::Engine eng;
    hh::host::Tuner tune;
    // timing
    bool timing = false;
    std::string timing_family;  // non-empty: time this family and window only
    int timing_w = 0;
    std::vector<hh::host::TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::pair<std::string, int>, std::pair<double, int64_t>> ktime;
};

namespace hh {
namespace host {

hipEvent_t get_event(hh_ctx *c);

// Run a launch, bracketed by events when timing is on.
template <typename F>
void timed(hh_ctx *c, const char *family, int w, F &&f)
{
    if (!c->timing || (!c->timing_family.empty() && (c->timing_family != family || c->timing_w != w))) {
        f();
        return;
    }
    TimedLaunch t{family, w, get_event(c), get_event(c)};
    HH_HIP(hipEventRecord(t.a, c->stream));
    f();
    HH_HIP(hipEventRecord(t.b, c->stream));
    c->pending.push_back(t);
}

void check_pass_valid(const hh_pass &p);
void check_pass_device(const hh_pass &p, bool for_schedule);
PassGrid make_grid(const hh_pass &p, int nr, int nc);
ConfigBufs &get_config(hh_ctx *c, const hh_pass &p);
CorrConfig make_corr(hh_ctx *c, const hh_pass &p);
void fill(hh_ctx *c, float *p, size_t n, float v);
float quantile95(std::vector<float> &v);

// ---- schedule engine (hh_engine.cpp) ----------------------------------------
// The layout of the context's current single pair (hh_set_pair*).
void engine_single_layout(hh_ctx *c);
// Upload a batch of host pairs into the owned atlas.
void engine_set_pairs(hh_ctx *c, const float *const *A, const float *const *B, const int32_t *nr,
                      const int32_t *nc, int npairs);
// run_piv of every pair of the layout (src/pipeline.jl:361-449).
void engine_run(hh_ctx *c, const hh_pass *passes, int np, int smoothing);
// The final PIVResult of pair q.
void engine_fetch(hh_ctx *c, int q, hh_field *r);
// The pixel mask of every following run (mask null: none).
void engine_set_mask(hh_ctx *c, const uint8_t *mask, int nr, int nc, double threshold);
void engine_release(hh_ctx *c);
void engine_set_graphs(hh_ctx *c, bool enable);
std::string engine_tune_report(hh_ctx *c);

}  // namespace host
}  // namespace hh
