// hh_api.cpp — C ABI and pass orchestration of the gfx950 WIDIM backend.
//
// Implements include/hammerhead_mi355x.h. The orchestration of hh_run_piv
// follows run_piv / piv_pass (src/pipeline.jl:361-449,655-832): per pass the
// predictor is built from the previous result (build_predictor,
// src/pipeline.jl:454-464), every sweep deforms the prefiltered pair by it,
// correlates all windows, adds the predictor back, validates and replaces,
// and iterating passes rebuild their own predictor between sweeps; the
// Wieneke uncertainty runs once on the final warp. Everything stays on the
// context's HIP stream; with convergence_tol = 0 there is no host
// synchronization inside the schedule.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "hh_ctx.h"

using namespace hh;

using namespace hh::host;

namespace hh {
namespace host {

hipEvent_t get_event(hh_ctx *c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HH_HIP(hipEventCreate(&e));
    return e;
}

}  // namespace host
}  // namespace hh

namespace {

thread_local std::string g_create_err;

// ---- scratch of hh_validate_and_replace_noalloc ---------------------------
// Device memory the validation chain needs beyond its inputs and outputs:
// the valid-cell counter (256 bytes), the substituted / replaced u and v
// (float each) and flags (one byte) of up to HH_NOALLOC_MAX_CELLS cells. One
// copy per GPU for the life of the process; in host code `scratch` only
// names the symbol, its device address comes from hipGetSymbolAddress.
constexpr size_t kNoAllocCells = HH_NOALLOC_MAX_CELLS;
constexpr size_t kScratchCount = 256;
constexpr size_t kScratchBytes = kScratchCount + kNoAllocCells * (2 * sizeof(float) + sizeof(uint8_t));
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::vector<unsigned char *> g_scratch_ptr;  // per device, resolved once

// The device address of this GPU's scratch (the calling thread's current
// device must be `device`); resolved on the first call per device.
unsigned char *scratch_on(int device, std::string &err)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("scratch symbol lookup failed: ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    }
    return g_scratch_ptr[device];
}

void drain_timing(hh_ctx *c)
{
    if (c->pending.empty()) return;
    HH_HIP(hipStreamSynchronize(c->stream));
    for (auto &t : c->pending) {
        float ms = 0.f;
        HH_HIP(hipEventElapsedTime(&ms, t.a, t.b));
        auto &acc = c->ktime[{t.family, t.w}];
        acc.first += ms;
        acc.second += 1;
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->pending.clear();
}

int fail(hh_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

template <typename F>
int guarded(hh_ctx *c, F &&f)
{
    if (!c) return HH_EINVAL;
    c->err.clear();
    try {
        HH_HIP(hipSetDevice(c->device));
        f();
        return HH_OK;
    } catch (const Error &e) {
        c->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        c->err = "host allocation failed";
        return HH_ENOMEM;
    }
}

}  // namespace

namespace hh {
namespace host {

// ---- parameter checks ---------------------------------------------------------

void check_pass_valid(const hh_pass &p)
{
    auto bad = [](const std::string &m) { throw Error(HH_EINVAL, m); };
    if (p.window_size[0] < 4 || p.window_size[1] < 4) bad("window_size must be at least 4 in each dimension");
    for (int d = 0; d < 2; ++d) {
        if (p.search_area_size[d] < p.window_size[d])
            bad("search_area_size must be at least window_size in each dimension");
        if ((p.search_area_size[d] - p.window_size[d]) % 2)
            bad("search_area_size - window_size must be even in each dimension");
        if (!(0 <= p.overlap[d] && p.overlap[d] < p.window_size[d]))
            bad("overlap must satisfy 0 <= overlap < window_size");
    }
    if (p.correlation_method != HH_CROSS && p.correlation_method != HH_PHASE)
        bad("correlation_method must be :cross or :phase");
    if (p.apodization != HH_APOD_NONE && p.apodization != HH_APOD_GAUSS) bad("apodization must be :none or :gauss");
    if (p.subpixel_method < HH_GAUSS3 || p.subpixel_method > HH_SUBPIX_NONE) bad("unknown subpixel_method");
    if (p.n_peaks < 1) bad("n_peaks must be at least 1");
    if (p.peak_finder != HH_REGIONALMAX && p.peak_finder != HH_EXCLUSION)
        bad("peak_finder must be :exclusion or :regionalmax");
    if (!(p.uod_threshold > 0)) bad("uod_threshold must be positive");
    if (p.uod_neighborhood < 1) bad("uod_neighborhood must be at least 1");
    if (!(p.min_peak_ratio >= 0)) bad("min_peak_ratio must be non-negative");
    if (p.max_iterations < 1) bad("max_iterations must be at least 1");
    if (!(p.convergence_tol >= 0)) bad("convergence_tol must be non-negative");
    if (p.n_validators < 0 || p.n_validators > HH_MAX_VALIDATORS) bad("at most 4 extra validators");
}

// The device engine's scope (the counterpart of _ka_scope_check,
// src/ka_backend.jl:13-29): rejected up front with HH_EUNSUPPORTED.
void check_pass_device(const hh_pass &p, bool for_schedule)
{
    check_pass_valid(p);
    auto uns = [](const std::string &m) { throw Error(HH_EUNSUPPORTED, "backend :mi355x " + m); };
    if (p.window_size[0] != p.window_size[1]) uns("supports square windows only");
    if (p.search_area_size[0] != p.search_area_size[1]) uns("supports square search areas only");
    const int w = p.window_size[0], sz = p.search_area_size[0];
    const int M = p.padding ? 2 * sz : sz;
    if (!(M == 16 || M == 24 || M == 32 || M == 48 || M == 64 || M == 96 || M == 128))
        uns("supports FFT sides 16, 24, 32, 48, 64, 96 and 128 (the search area, doubled when padded)");
    if (p.subpixel_method == HH_GAUSS2D) uns("supports subpixel_method :gauss3, :gauss9 or :none only");
    if (p.n_peaks > kMaxPeaks) uns("supports n_peaks up to 8");
    if (p.uod_neighborhood > 3) uns("supports uod_neighborhood 1, 2 or 3");
    for (int q = 0; q < p.n_validators; ++q) {
        const hh_validator &v = p.validators[q];
        if (v.kind < HH_V_PEAK_RATIO || v.kind > HH_V_UOD) throw Error(HH_EINVAL, "unknown validator kind");
        if (v.kind == HH_V_UOD && (v.neighborhood < 1 || v.neighborhood > 3))
            uns("supports UOD validators with neighborhood 1, 2 or 3");
        if (v.kind == HH_V_UOD && !(v.b > 0)) throw Error(HH_EINVAL, "epsilon must be positive");
    }
    if (p.uncertainty && w > 64) uns("supports uncertainty for windows up to 64 px");
    if (for_schedule && p.keep_correlation_planes)
        uns("does not support keep_correlation_planes in run_piv; use hh_correlation_planes");
}

PassGrid make_grid(const hh_pass &p, int nr, int nc)
{
    const int wr = p.window_size[0], wc = p.window_size[1];
    const int sr = p.search_area_size[0], sc = p.search_area_size[1];
    const int mr = (sr - wr) / 2, mc = (sc - wc) / 2;
    if (sr > nr || sc > nc) throw Error(HH_EINVAL, "search area exceeds image size");
    const int r0 = 1 + mr, rstep = wr - p.overlap[0], rstop = nr - wr - mr + 1;
    const int c0 = 1 + mc, cstep = wc - p.overlap[1], cstop = nc - wc - mc + 1;
    PassGrid g;
    g.ny = rstop >= r0 ? (rstop - r0) / rstep + 1 : 0;
    g.nx = cstop >= c0 ? (cstop - c0) / cstep + 1 : 0;
    g.x.resize(g.nx);
    g.y.resize(g.ny);
    for (int j = 0; j < g.nx; ++j) g.x[j] = (float)(c0 + j * cstep) + (float)((wc - 1) / 2.0);
    for (int i = 0; i < g.ny; ++i) g.y[i] = (float)(r0 + i * rstep) + (float)((wr - 1) / 2.0);
    g.jobs.resize((size_t)4 * g.ny * g.nx);
    size_t k = 0;
    for (int j = 0; j < g.nx; ++j)
        for (int i = 0; i < g.ny; ++i) {
            g.jobs[4 * k] = i + 1;
            g.jobs[4 * k + 1] = j + 1;
            g.jobs[4 * k + 2] = r0 + i * rstep;
            g.jobs[4 * k + 3] = c0 + j * cstep;
            ++k;
        }
    return g;
}

// apodization_window (src/correlators.jl:18-25), Float64 then float32.
std::vector<float> host_apod(int w, int apod)
{
    std::vector<float> a((size_t)w * w);
    const double c = (w + 1) / 2.0, s = w / 4.0;
    for (int j = 1; j <= w; ++j)
        for (int i = 1; i <= w; ++i)
            a[(i - 1) + (size_t)(j - 1) * w] =
                apod == HH_APOD_GAUSS
                    ? (float)std::exp(-((i - c) * (i - c) / (2 * s * s) + (j - c) * (j - c) / (2 * s * s)))
                    : 1.0f;
    return a;
}

ConfigBufs &get_config(hh_ctx *c, const hh_pass &p)
{
    const int w = p.window_size[0], sz = p.search_area_size[0];
    auto key = std::make_tuple(w, sz, (int)p.padding, (int)p.apodization, (int)p.correlation_method);
    auto it = c->configs.find(key);
    if (it != c->configs.end()) return *it->second;
    auto cb = std::make_unique<ConfigBufs>();
    const int M = p.padding ? 2 * sz : sz;
    std::vector<float> apod = host_apod(w, p.apodization);
    cb->apod.ensure(apod.size());
    HH_HIP(hipMemcpy(cb->apod.p, apod.data(), sizeof(float) * apod.size(), hipMemcpyHostToDevice));
    cb->box_lo = 0;
    cb->box_hi = M - 1;
    // the overlap gain belongs to padded equal areas (src/correlators.jl:73)
    if (p.padding && sz == w) {
        // overlap_gain! (src/correlators.jl:170-181): the autocorrelation of the
        // apodization (direct sums in Float64 of the float32 window), in the
        // fftshifted layout; wmax / w where w >= wmax / 2, else 0.
        std::vector<double> ac((size_t)M * M, 0.0);
        const int H = M / 2;
        for (int l2 = -(w - 1); l2 <= w - 1; ++l2)
            for (int l1 = -(w - 1); l1 <= w - 1; ++l1) {
                double s = 0.0;
                for (int j = std::max(0, -l2); j < std::min(w, w - l2); ++j)
                    for (int i = std::max(0, -l1); i < std::min(w, w - l1); ++i)
                        s += (double)apod[i + (size_t)j * w] * (double)apod[(i + l1) + (size_t)(j + l2) * w];
                ac[(l1 + H) + (size_t)(l2 + H) * M] = s;
            }
        double wmax = 0.0;
        for (double v : ac) wmax = std::max(wmax, v);
        std::vector<float> gain((size_t)M * M);
        int lo = M, hi = -1;
        for (int jp = 0; jp < M; ++jp)
            for (int ip = 0; ip < M; ++ip) {
                const double v = ac[ip + (size_t)jp * M];
                const bool on = v >= wmax / 2;
                gain[ip + (size_t)jp * M] = on ? (float)(wmax / v) : 0.0f;
                if (on) {
                    lo = std::min(lo, std::min(ip, jp));
                    hi = std::max(hi, std::max(ip, jp));
                }
            }
        cb->box_lo = std::max(0, lo - 1);
        cb->box_hi = std::min(M - 1, hi + 1);
        cb->gain.ensure(gain.size());
        HH_HIP(hipMemcpy(cb->gain.p, gain.data(), sizeof(float) * gain.size(), hipMemcpyHostToDevice));
    }
    if (p.correlation_method == HH_PHASE) {
        const double sigma = M / 8.0;
        std::vector<float> wf((size_t)M * M);
        for (int j = 0; j < M; ++j)
            for (int i = 0; i < M; ++i) {
                const double fi = i < (M + 1) / 2 ? i : i - M, fj = j < (M + 1) / 2 ? j : j - M;
                wf[i + (size_t)j * M] = (float)std::exp(-(fi * fi + fj * fj) / (2 * sigma * sigma));
            }
        cb->wfilt.ensure(wf.size());
        HH_HIP(hipMemcpy(cb->wfilt.p, wf.data(), sizeof(float) * wf.size(), hipMemcpyHostToDevice));
    }
    ConfigBufs &ref = *cb;
    c->configs.emplace(key, std::move(cb));
    return ref;
}

CorrConfig make_corr(hh_ctx *c, const hh_pass &p)
{
    ConfigBufs &cb = get_config(c, p);
    CorrConfig cfg{};
    cfg.w = p.window_size[0];
    cfg.s = p.search_area_size[0];
    cfg.M = p.padding ? 2 * cfg.s : cfg.s;
    cfg.padded = p.padding;
    cfg.method = p.correlation_method;
    cfg.subpixel = p.subpixel_method;
    cfg.finder = p.peak_finder;
    cfg.n_peaks = p.n_peaks;
    cfg.K = std::max(p.n_peaks, 2);
    cfg.box_lo = cb.box_lo;
    cfg.box_hi = cb.box_hi;
    cfg.apod = cb.apod.p;
    cfg.gain = p.padding && cfg.s == cfg.w ? cb.gain.p : nullptr;
    cfg.wfilt = p.correlation_method == HH_PHASE ? cb.wfilt.p : nullptr;
    cfg.form = c->form;
    return cfg;
}

const int32_t *device_jobs(hh_ctx *c, const hh_pass &p, const PassGrid &g)
{
    auto key = std::make_tuple(p.window_size[0], p.search_area_size[0], p.overlap[0], p.overlap[1], c->nr, c->nc);
    auto it = c->jobcache.find(key);
    if (it != c->jobcache.end()) return it->second->p;
    auto b = std::make_unique<DBuf<int32_t>>();
    b->ensure(std::max<size_t>(g.jobs.size(), 4));
    if (!g.jobs.empty())
        HH_HIP(hipMemcpy(b->p, g.jobs.data(), sizeof(int32_t) * g.jobs.size(), hipMemcpyHostToDevice));
    const int32_t *ptr = b->p;
    c->jobcache.emplace(key, std::move(b));
    return ptr;
}

void need_pair(hh_ctx *c)
{
    if (!c->A) throw Error(HH_ESTATE, "no image pair: call hh_set_pair first");
}

// The frame-A window and the frame-B search area of a job (1-based window
// origin rs, cs; the search area has the pass's margins around it) lie in
// the current pair.
void check_job_area(hh_ctx *c, const hh_pass &p, int rs, int cs)
{
    const int sz = p.search_area_size[0], m = (sz - p.window_size[0]) / 2;
    if (rs - m < 1 || cs - m < 1 || rs - m + sz - 1 > c->nr || cs - m + sz - 1 > c->nc)
        throw Error(HH_EDIM, m ? "search area outside the image" : "window outside the image");
}

const float *src_image(hh_ctx *c, int src, bool a)
{
    if (src == HH_SRC_WARP) {
        if (!c->warp_valid) throw Error(HH_ESTATE, "no warped pair: call hh_deform first");
        return a ? c->warpA.p : c->warpB.p;
    }
    need_pair(c);
    if (src != HH_SRC_RAW) throw Error(HH_EINVAL, "src must be HH_SRC_RAW or HH_SRC_WARP");
    return a ? c->A : c->B;
}

void do_prefilter(hh_ctx *c)
{
    need_pair(c);
    const size_t n = (size_t)(c->nr + 2) * (c->nc + 2);
    c->coefA.ensure(n);
    c->coefB.ensure(n);
    c->ptmp.ensure((size_t)2 * (c->nr + 2) * c->nc);
    timed(c, "prefilter", 0, [&] {
        launch_prefilter_pair(c->A, c->B, c->nr, c->nr, c->nc, c->coefA.p, c->coefB.p, c->ptmp.p, c->stream);
    });
    c->coefs_valid = true;
}

void do_deform(hh_ctx *c)
{
    const size_t n = (size_t)c->nr * c->nc;
    c->warpA.ensure(n);
    c->warpB.ensure(n);
    timed(c, "deform", 0, [&] {
        launch_deform(c->coefA.p, c->coefB.p, c->nr, c->nc, c->pu.p, c->pv.p, c->gny, c->gnx, c->y0, c->ysp, c->x0,
                      c->xsp, c->warpA.p, c->warpB.p, c->stream);
    });
    c->warp_valid = true;
}

__global__ void k_fill_f32(float *p, size_t n, float v)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = v;
}

__global__ void k_restore(float *u, float *v, const float *mu, const float *mv, const uint8_t *out, int n)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n && out[e]) {
        u[e] = mu[e];
        v[e] = mv[e];
    }
}

// max(|du|, |dv|) per unmasked cell for field_change (src/pipeline.jl:849-866);
// NaN/finite flips give +Inf, NaN in both sweeps gives NaN (dropped on host).
__global__ void k_change(const float *u, const float *v, const float *pu, const float *pv, int n, float *out)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float du = fabsf(u[e] - pu[e]), dv = fabsf(v[e] - pv[e]);
    if (isnan(du) || isnan(dv)) {
        const bool same = (isnan(u[e]) == isnan(pu[e])) && (isnan(v[e]) == isnan(pv[e]));
        out[e] = same ? NAN : INFINITY;
        return;
    }
    out[e] = du > dv ? du : dv;
}

void fill(hh_ctx *c, float *p, size_t n, float v)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p, n, v);
    HH_HIP(hipGetLastError());
}

// quantile!(buf, 0.95), type 7 (Statistics.jl semantics).
float quantile95(std::vector<float> &v)
{
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    const double p = 0.95;
    const double aleph = n * p + (1.0 - p);
    long j = (long)aleph;
    if (j > (long)n - 1) j = (long)n - 1;
    if (j < 1) j = 1;
    float gamma = (float)(aleph - (double)j);
    gamma = std::min(1.f, std::max(0.f, gamma));
    float a, b;
    if (n == 1) {
        a = b = v[0];
    } else {
        a = v[j - 1];
        b = v[j];
    }
    if (std::isfinite(a) && std::isfinite(b)) return a + gamma * (b - a);
    return (1 - gamma) * a + gamma * b;
}

// gmask (nullable, device): masked cells, filled from their unmasked
// neighbours (into the validation scratch) before the smoothing.
void set_predictor_from_grid(hh_ctx *c, const PassGrid &g, const float *u, const float *v, bool smoothing,
                             const uint8_t *gmask = nullptr)
{
    const size_t n = (size_t)g.ny * g.nx;
    c->pu.ensure(n);
    c->pv.ensure(n);
    if (gmask) {
        float *fu = c->gsu.ensure(n), *fv = c->gsv.ensure(n);
        launch_fill_masked(one_grid(g.ny, g.nx), u, v, gmask, c->gcount.ensure(1), fu, fv, c->stream);
        u = fu;
        v = fv;
    }
    if (smoothing) {
        launch_smooth(one_grid(g.ny, g.nx), u, c->pu.p, c->stream);
        launch_smooth(one_grid(g.ny, g.nx), v, c->pv.p, c->stream);
    } else {
        HH_HIP(hipMemcpyAsync(c->pu.p, u, sizeof(float) * n, hipMemcpyDeviceToDevice, c->stream));
        HH_HIP(hipMemcpyAsync(c->pv.p, v, sizeof(float) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    c->gny = g.ny;
    c->gnx = g.nx;
    c->y0 = g.y[0];
    c->x0 = g.x[0];
    c->ysp = g.ny > 1 ? g.y[1] - g.y[0] : 1.f;
    c->xsp = g.nx > 1 ? g.x[1] - g.x[0] : 1.f;
}

}  // namespace host
}  // namespace hh

namespace {

template <typename T>
T *upload(hh_ctx *c, DBuf<T> &b, const T *h, size_t n)
{
    T *p = b.ensure(std::max<size_t>(n, 1));
    if (h && n) HH_HIP(hipMemcpyAsync(p, h, sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
    return p;
}

template <typename T>
void download(hh_ctx *c, T *h, const T *d, size_t n)
{
    if (h && n) HH_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
}

// The windows a per-hook entry point correlates. Without a pixel mask
// (hh_set_mask) they are read from the images; with one, every job's window
// is first copied into a centred tile (k_mask_tiles: the mean of its valid
// pixels subtracted, masked pixels 0, as load_windows! loads a window with a
// mask, src/correlators.jl:199-227) and the launch reads the tiles through
// jobs (gi, gj, 1, t w + 1) in the caller's order, so every per-job output
// keeps its index.
struct HookSource {
    const float *A, *B;
    int ld;
    const int32_t *jobs;
};

// window_only: the uncertainty's load (window-sized second frame at the
// window origin, src/pipeline.jl:984-989) instead of the search area. The
// tiles go to buffer `slot` (0 or 1), so both loads of one call can coexist.
HookSource hook_source(hh_ctx *c, const hh_pass &p, const int32_t *hjobs, int njobs, const float *dA,
                       const float *dB, const int32_t *djobs, bool window_only = false, int slot = 0)
{
    const Engine &E = c->eng;
    const int w = p.window_size[0];
    if (E.hmask.empty() || njobs <= 0) return HookSource{dA, dB, c->nr, djobs};
    if (E.mnr != c->nr || E.mnc != c->nc)
        throw Error(HH_EDIM, "mask must have the same size as the images, got (" + std::to_string(E.mnr) + ", " +
                                 std::to_string(E.mnc) + ")");
    const int sz = window_only ? w : p.search_area_size[0], mg = (sz - w) / 2;
    if ((size_t)njobs * sz >= (size_t)1 << 31) throw Error(HH_EINVAL, "too many windows for the masked tiles");
    std::vector<int32_t> t((size_t)8 * njobs);
    for (int k = 0; k < njobs; ++k) {
        const int32_t *j = hjobs + 4 * (size_t)k;
        int32_t *tj = &t[4 * (size_t)k], *mj = &t[4 * ((size_t)njobs + k)];
        tj[0] = j[0];
        tj[1] = j[1];
        tj[2] = 1 + mg;
        tj[3] = k * sz + 1 + mg;
        mj[0] = mj[2] = j[2] - 1;
        mj[1] = mj[3] = j[3] - 1;
    }
    const int32_t *d = upload(c, c->hookmask[slot], t.data(), t.size());
    const size_t tp = (size_t)njobs * sz * sz;
    float *tiles = c->hooktiles[slot].ensure(2 * tp);
    launch_mask_tiles(dA, dB, c->nr, E.dmask.p, E.mnr, d + 4 * (size_t)njobs, njobs, w, sz, tiles, tiles + tp,
                      c->stream);
    HH_HIP(hipStreamSynchronize(c->stream));  // the upload reads `t`
    return HookSource{tiles, tiles + tp, sz, d};
}

// ---- ensemble (sum-of-correlation) PIV: src/ensemble.jl:133-343 -----------

__global__ void k_fill_f64(double *p, size_t n, double v)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = v;
}

void fill64(hh_ctx *c, double *p, size_t n, double v)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p, n, v);
    HH_HIP(hipGetLastError());
}

// _plane_accumulator / _uncertainty_accumulator (src/ensemble.jl:165-190):
// zeroed FP64 plane sums for the jobs (device) of one pass.
void ens_begin(hh_ctx *c, const hh_pass &p, const int32_t *djobs, int njobs, bool unc)
{
    check_pass_device(p, true);
    if (unc && p.window_size[0] > 64) throw Error(HH_EUNSUPPORTED, "uncertainty supports windows up to 64 px");
    auto &e = c->ens;
    e.active = false;
    e.pass = p;
    e.cfg = make_corr(c, p);
    e.box = correlate_box_layout(e.cfg);
    const int nb = e.cfg.box_hi - e.cfg.box_lo + 1;
    e.per = e.box ? (size_t)nb * nb : (size_t)e.cfg.M * e.cfg.M;
    e.njobs = njobs;
    e.jobs = djobs;
    e.nclean = njobs;
    e.npart = 0;
    e.tjobs = e.mjobs = e.ujobs = nullptr;
    e.unc = unc;
    e.npairs = 0;
    fill64(c, c->ensacc.ensure(std::max<size_t>(e.per * njobs, 1)), e.per * njobs, 0.0);
    if (unc) fill64(c, c->ensstats.ensure(std::max<size_t>((size_t)2 * HH_UQ_NSTATS * njobs, 1)),
                    (size_t)2 * HH_UQ_NSTATS * njobs, 0.0);
    e.active = true;
}

// accumulate_planes! (src/ensemble.jl:317-343) of one pair (device images;
// with src given, the clean windows are correlated through src, and their
// statistics taken through usrc, instead of A, B and the ensemble's jobs).
void ens_add(hh_ctx *c, const float *A, const float *B, const HookSource *src = nullptr,
             const HookSource *usrc = nullptr)
{
    auto &e = c->ens;
    if (!e.active) throw Error(HH_ESTATE, "no ensemble: call hh_ensemble_begin first");
    CorrOut out{};
    out.acc = c->ensacc.p;
    out.ny = out.nx = 1;
    const float *cA = src ? src->A : A, *cB = src ? src->B : B;
    const int cld = src ? src->ld : c->nr;
    const int32_t *cj = src ? src->jobs : e.jobs;
    timed(c, "ensemble_add", e.cfg.w,
          [&] { launch_correlate(e.cfg, cA, cB, cld, 0, cj, e.nclean, out, c->stream); });
    const int w = e.cfg.w, sz = e.cfg.s;
    const size_t tp = (size_t)e.npart * sz * sz;
    float *tA = e.npart ? c->enstiles.p : nullptr, *tB = tA ? tA + tp : nullptr;
    if (e.npart) {
        // windows with masked pixels: centred tiles, summed into their own
        // slots after the clean windows'
        timed(c, "mask_tiles", w, [&] {
            launch_mask_tiles(A, B, c->nr, c->eng.dmask.p, c->eng.mnr, e.mjobs, e.npart, w, sz, tA, tB, c->stream);
        });
        CorrOut o2 = out;
        o2.acc = c->ensacc.p + e.per * e.nclean;
        timed(c, "ensemble_add_masked", w,
              [&] { launch_correlate(e.cfg, tA, tB, sz, 0, e.tjobs, e.npart, o2, c->stream); });
    }
    if (e.unc) {
        const ConfigBufs &cb = get_config(c, e.pass);
        const float *uA = usrc ? usrc->A : cA, *uB = usrc ? usrc->B : cB;
        const int uld = usrc ? usrc->ld : cld;
        const int32_t *uj = usrc ? usrc->jobs : cj;
        timed(c, "uq", e.cfg.w, [&] {
            launch_uq(e.cfg.w, cb.apod.p, uA, uB, uld, uj, e.nclean, 1, nullptr, nullptr, c->ensstats.p, true,
                      c->stream);
            if (e.npart) {
                // the statistics read window tiles of both frames
                if (sz != w)
                    launch_mask_tiles(A, B, c->nr, c->eng.dmask.p, c->eng.mnr, e.mjobs, e.npart, w, w, tA, tB,
                                      c->stream);
                launch_uq(w, cb.apod.p, tA, tB, w, sz != w ? e.ujobs : e.tjobs, e.npart, 1, nullptr, nullptr,
                          c->ensstats.p + (size_t)2 * HH_UQ_NSTATS * e.nclean, true, c->stream);
            }
        });
    }
    ++e.npairs;
}

// ensemble_analyze! (src/ensemble.jl:284-313) into device grids; u, v hold
// the shared predictor on entry and the total displacement on exit.
void ens_analyze(hh_ctx *c, int ny, int nx, float *u, float *v, float *pr, float *cm, float *au, float *av, float *uu,
                 float *uv)
{
    auto &e = c->ens;
    if (!e.active) throw Error(HH_ESTATE, "no ensemble: call hh_ensemble_begin first");
    CorrOut out{};
    out.u = u;
    out.v = v;
    out.u0 = u;
    out.v0 = v;
    out.pr = pr;
    out.cm = cm;
    out.alt_u = au;
    out.alt_v = av;
    out.ny = ny;
    out.nx = nx;
    timed(c, "ensemble_analyze", e.cfg.w,
          [&] { launch_plane_analyze(e.cfg, c->ensacc.p, e.box, e.jobs, e.njobs, out, c->stream); });
    if (e.unc && (uu || uv)) launch_uq_finalize(c->ensstats.p, e.jobs, e.njobs, ny, uu, uv, c->stream);
}

// run_piv_ensemble(pairs, passes; mask, mask_threshold) (src/ensemble.jl:
// 47-84) with ensemble_pass (src/ensemble.jl:133-265), entirely on the
// device. The pairs are uploaded once and stay resident across the passes.
// The context's pixel mask (hh_set_mask), when set, removes the windows of
// the grid mask, loads partly masked windows from their valid pixels and
// validates and fills the masked cells as run_piv does.
void run_ensemble(hh_ctx *c, const float *const *hA, const float *const *hB, int npairs, int nr, int nc,
                  const hh_pass *passes, int np, int smoothing)
{
    if (npairs < 1) throw Error(HH_EINVAL, "pairs must not be empty");
    if (np < 1) throw Error(HH_EINVAL, "at least one pass is required");
    if (nr < 2 || nc < 2) throw Error(HH_EDIM, "images must be at least 2 x 2");
    for (int k = 0; k < np; ++k) check_pass_device(passes[k], true);
    const Engine &EM = c->eng;
    const bool masked = !EM.hmask.empty();
    if (masked && (EM.mnr != nr || EM.mnc != nc))
        throw Error(HH_EDIM, "mask must have the same size as the images, got (" + std::to_string(EM.mnr) + ", " +
                                 std::to_string(EM.mnc) + ") for images of (" + std::to_string(nr) + ", " +
                                 std::to_string(nc) + ")");
    const size_t npx = (size_t)nr * nc;
    float *arena = nullptr;
    bool upload = false;
    if (!hA && !hB) {
        // the pairs of the previous call, still resident
        if (c->arena_pairs != npairs || c->arena_nr != nr || c->arena_nc != nc)
            throw Error(HH_ESTATE, "no resident pairs of this count and size: pass the frames");
        arena = c->arena.p;
    } else {
        if (!hA || !hB) throw Error(HH_EINVAL, "null pair list");
        for (int q = 0; q < npairs; ++q)
            if (!hA[q] || !hB[q]) throw Error(HH_EINVAL, "null image in the pair list");
        c->arena_pairs = 0;
        arena = c->arena.ensure(2 * npx * npairs);
        // uploads go on a copy stream, pair by pair, and the first pass's
        // accumulation of pair q waits only for pair q: the transfer of the
        // next pairs overlaps the correlation of the earlier ones
        if (!c->copy_stream) HH_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        while (c->upload_ev.size() < 4) {
            hipEvent_t e;
            HH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            c->upload_ev.push_back(e);
        }
        upload = true;
    }
    if (nr != c->nr || nc != c->nc) c->jobcache.clear();
    c->nr = nr;
    c->nc = nc;
    c->A = arena;
    c->B = arena + npx;
    c->coefs_valid = c->warp_valid = c->have_result = false;
    c->eng.have_result = false;
    size_t maxng = 0, maxalt = 0;
    int maxny = 0, maxnx = 0;
    for (int k = 0; k < np; ++k) {
        int ny = 0, nx = 0;
        hh_grid_size(&passes[k], nr, nc, &ny, &nx);
        maxng = std::max(maxng, (size_t)ny * nx);
        maxalt = std::max(maxalt, (size_t)ny * nx * std::max(passes[k].n_peaks - 1, 0));
        maxny = std::max(maxny, ny);
        maxnx = std::max(maxnx, nx);
    }
    for (DBuf<float> *b : {&c->gu, &c->gv, &c->gpr, &c->gcm, &c->guu, &c->guv, &c->gsu, &c->gsv, &c->pu, &c->pv})
        b->ensure(std::max<size_t>(maxng, 1));
    c->galtu.ensure(std::max<size_t>(maxalt, 1));
    c->galtv.ensure(std::max<size_t>(maxalt, 1));
    c->gout.ensure(std::max<size_t>(maxng, 1));
    c->gsflag.ensure(std::max<size_t>(maxng, 1));
    c->gpx.ensure(std::max(maxnx, 1));
    c->gpy.ensure(std::max(maxny, 1));
    PassGrid prev;
    bool have_pred = false, prev_anyg = false;
    c->final_gm.clear();
    for (int k = 0; k < np; ++k) {
        const hh_pass &p = passes[k];
        const bool force_replace = k < np - 1;
        PassGrid g = make_grid(p, nr, nc);
        const int ny = g.ny, nx = g.nx;
        const size_t ng = (size_t)ny * nx;
        if (ng == 0) throw Error(HH_EINVAL, "the pass grid is empty");
        if (k > 0) {
            // c->gmask still holds the previous pass's grid mask
            set_predictor_from_grid(c, prev, c->gu.p, c->gv.p, smoothing, prev_anyg ? c->gmask.p : nullptr);
            have_pred = true;
        }
        const bool unc = p.uncertainty && !force_replace;
        const int32_t *djobs = nullptr;
        int nclean = (int)ng, npart = 0;
        bool anyg = false;
        std::vector<uint8_t> gm;
        if (masked) {
            // pass_grid's grid mask (src/pipeline.jl:542-549): the clean
            // windows first, then those with masked pixels; masked ones
            // have no job
            const int w = p.window_size[0], sz = p.search_area_size[0], mg = (sz - w) / 2;
            std::vector<int32_t> clean, part, tj, mj, uj;
            gm.assign(ng, 0);
            auto count = [&](int r0, int c0, int side) {  // masked pixels of a block (1-based origin)
                long n = 0;
                for (int cc = 0; cc < side; ++cc) {
                    const uint8_t *col = EM.hmask.data() + (size_t)(c0 - 1 + cc) * nr + (r0 - 1);
                    for (int i = 0; i < side; ++i) n += col[i];
                }
                return n;
            };
            for (size_t j = 0; j < ng; ++j) {
                const int32_t *jb = &g.jobs[4 * j];
                const int rs = jb[2], cs = jb[3];
                const long nm = count(rs, cs, w), ns = sz == w ? nm : count(rs - mg, cs - mg, sz);
                const bool gmj = (double)nm / (double)(w * w) >= EM.mthr || (double)ns / (double)(sz * sz) >= EM.mthr;
                gm[j] = gmj;
                anyg |= gmj;
                if (gmj) continue;
                if (nm == 0 && ns == 0) {
                    clean.insert(clean.end(), jb, jb + 4);
                } else {
                    const int t = (int)(tj.size() / 4);
                    part.insert(part.end(), jb, jb + 4);
                    tj.insert(tj.end(), {jb[0], jb[1], 1 + mg, t * sz + 1 + mg});
                    uj.insert(uj.end(), {jb[0], jb[1], 1, t * w + 1});
                    mj.insert(mj.end(), {rs - 1, cs - 1, rs - 1, cs - 1});
                }
            }
            nclean = (int)(clean.size() / 4);
            npart = (int)(part.size() / 4);
            if ((size_t)npart * sz >= (size_t)1 << 31) throw Error(HH_EINVAL, "too many windows with masked pixels");
            clean.insert(clean.end(), part.begin(), part.end());
            tj.insert(tj.end(), mj.begin(), mj.end());
            tj.insert(tj.end(), uj.begin(), uj.end());
            djobs = ::upload(c, c->ensjobs, clean.data(), clean.size());
            const int32_t *pj = ::upload(c, c->enspart, tj.data(), tj.size());
            uint8_t *dgm = c->gmask.ensure(ng);
            HH_HIP(hipMemcpyAsync(dgm, gm.data(), ng, hipMemcpyHostToDevice, c->stream));
            // the uploads read host vectors of this scope
            HH_HIP(hipStreamSynchronize(c->stream));
            if (npart) c->enstiles.ensure(2 * (size_t)npart * sz * sz);
            ens_begin(c, p, djobs, nclean + npart, unc);
            c->ens.nclean = nclean;
            c->ens.npart = npart;
            c->ens.tjobs = pj;
            c->ens.mjobs = pj + 4 * (size_t)npart;
            c->ens.ujobs = pj + 8 * (size_t)npart;
        } else {
            djobs = device_jobs(c, p, g);
            ens_begin(c, p, djobs, (int)ng, unc);
        }
        for (int q = 0; q < npairs; ++q) {
            c->A = arena + 2 * npx * q;
            c->B = arena + 2 * npx * q + npx;
            if (upload && k == 0) {
                HH_HIP(hipMemcpyAsync(arena + 2 * npx * q, hA[q], sizeof(float) * npx, hipMemcpyHostToDevice,
                                      c->copy_stream));
                HH_HIP(hipMemcpyAsync(arena + 2 * npx * q + npx, hB[q], sizeof(float) * npx, hipMemcpyHostToDevice,
                                      c->copy_stream));
                // a ring event: the stream's wait takes the state recorded
                // now, so re-recording it for a later pair is safe
                hipEvent_t ev = c->upload_ev[q % c->upload_ev.size()];
                HH_HIP(hipEventRecord(ev, c->copy_stream));
                HH_HIP(hipStreamWaitEvent(c->stream, ev, 0));
            }
            if (have_pred) {
                do_prefilter(c);
                do_deform(c);
                ens_add(c, c->warpA.p, c->warpB.p);
            } else {
                ens_add(c, c->A, c->B);
            }
        }
        if (upload && k == 0) {
            c->arena_pairs = npairs;
            c->arena_nr = nr;
            c->arena_nc = nc;
        }
        float *u = c->gu.p, *v = c->gv.p;
        if (have_pred) {
            float *gx = c->gpx.p, *gy = c->gpy.p;
            HH_HIP(hipMemcpyAsync(gx, g.x.data(), sizeof(float) * nx, hipMemcpyHostToDevice, c->stream));
            HH_HIP(hipMemcpyAsync(gy, g.y.data(), sizeof(float) * ny, hipMemcpyHostToDevice, c->stream));
            launch_node_values(c->pu.p, c->pv.p, c->gny, c->gnx, c->y0, c->ysp, c->x0, c->xsp, gy, ny, gx, nx, u, v,
                               c->stream);
            // the copies read host vectors that this scope owns
            HH_HIP(hipStreamSynchronize(c->stream));
        } else {
            fill(c, u, ng, 0.f);
            fill(c, v, ng, 0.f);
        }
        const int nalt = p.n_peaks - 1;
        float *au = nalt > 0 ? c->galtu.p : nullptr, *av = nalt > 0 ? c->galtv.p : nullptr;
        fill(c, c->guu.p, ng, NAN);
        fill(c, c->guv.p, ng, NAN);
        ens_analyze(c, ny, nx, u, v, c->gpr.p, c->gcm.p, au, av, c->guu.p, c->guv.p);
        if (anyg) launch_mask_nan(one_grid(ny, nx), c->gmask.p, u, v, c->gpr.p, c->gcm.p, c->stream);
        ValidateArgs va{};
        va.grids = one_grid(ny, nx);
        va.pass = &p;
        va.force_replace = force_replace;
        va.u = u;
        va.v = v;
        va.pr = c->gpr.p;
        va.cm = c->gcm.p;
        va.alt_u = au;
        va.alt_v = av;
        va.outliers = c->gout.p;
        va.su = c->gsu.p;
        va.sv = c->gsv.p;
        va.sflag = c->gsflag.p;
        va.count = c->gcount.ensure(1);
        va.mask = anyg ? c->gmask.p : nullptr;
        timed(c, "validate", p.window_size[0], [&] { launch_validate(va, c->stream); });
        prev = std::move(g);
        prev_anyg = anyg;
        if (k == np - 1 && anyg) c->final_gm = gm;
    }
    HH_HIP(hipGetLastError());
    c->final_grid = prev;
    c->have_result = true;
}

// ---- windowed correlation energy and its adjoint ------------------------------

// lin_axis of the device (interp.h), in the same float arithmetic.
int host_lin_axis(float q0, float qstep, int n, float q)
{
    if (n == 1) return 0;
    float t = (q - q0) / qstep;
    t = t < 0.f ? 0.f : t;
    t = t > (float)(n - 1) ? (float)(n - 1) : t;
    int i = (int)std::floor(t);
    return i > n - 2 ? n - 2 : i;
}

struct EnergySetup {
    EnergyArgs g{};
    int ncy = 0, ncx = 0;
    const int2 *rows = nullptr, *cols = nullptr;
};

// Node grid, windows and node-interval pixel ranges of one energy call; the
// node field itself is uploaded by energy_nodes.
// A predictor or node value that is not finite would give NaN sample
// positions; they are rejected before anything reaches the device.
void check_finite(const float *a, size_t n, const char *what)
{
    for (size_t k = 0; k < n; ++k)
        if (!std::isfinite(a[k])) throw Error(HH_EINVAL, std::string(what) + " holds a non-finite value");
}

// Uniform knots (the first two define the spacing): finite and increasing.
void check_knots(const float *py, int gny, const float *px, int gnx)
{
    check_finite(py, (size_t)gny, "row knots");
    check_finite(px, (size_t)gnx, "column knots");
    if ((gny > 1 && !(py[1] > py[0])) || (gnx > 1 && !(px[1] > px[0])))
        throw Error(HH_EINVAL, "node knots must increase");
}

EnergySetup energy_setup(hh_ctx *c, const hh_pass &p, const float *py, int gny, const float *px, int gnx)
{
    need_pair(c);
    check_pass_valid(p);
    if (p.search_area_size[0] != p.window_size[0] || p.search_area_size[1] != p.window_size[1] ||
        p.window_size[0] != p.window_size[1])
        throw Error(HH_EUNSUPPORTED, "the correlation-energy adjoint takes square windows without search margins");
    const int w = p.window_size[0];
    if (w != 8 && w != 16 && w != 32)
        throw Error(HH_EUNSUPPORTED, "the correlation-energy adjoint supports 8, 16 and 32 px windows");
    if (gny < 1 || gnx < 1 || !py || !px) throw Error(HH_EINVAL, "empty node grid");
    check_knots(py, gny, px, gnx);
    if (!c->energy_bilinear && !c->coefs_valid) do_prefilter(c);
    PassGrid grid = make_grid(p, c->nr, c->nc);
    if (grid.ny * grid.nx == 0) throw Error(HH_EINVAL, "the pass grid is empty");
    EnergySetup s;
    EnergyArgs &g = s.g;
    g.bilinear = c->energy_bilinear ? 1 : 0;
    g.cA = c->energy_bilinear ? nullptr : c->coefA.p;
    g.cB = c->energy_bilinear ? nullptr : c->coefB.p;
    g.A = c->A;
    g.B = c->B;
    g.nr = c->nr;
    g.nc = c->nc;
    g.gny = gny;
    g.gnx = gnx;
    g.y0 = py[0];
    g.ysp = gny > 1 ? py[1] - py[0] : 1.f;
    g.x0 = px[0];
    g.xsp = gnx > 1 ? px[1] - px[0] : 1.f;
    g.w = w;
    g.ny = grid.ny;
    g.nx = grid.nx;
    g.r0 = grid.jobs[2];
    g.c0 = grid.jobs[3];
    g.st_r = w - p.overlap[0];
    g.st_c = w - p.overlap[1];
    g.apod = get_config(c, p).apod.p;
    s.ncy = std::max(gny - 1, 1);
    s.ncx = std::max(gnx - 1, 1);
    std::vector<int2> cells(s.ncy + s.ncx, make_int2(INT32_MAX, INT32_MIN));
    for (int r = 1; r <= c->nr; ++r) {
        int2 &q = cells[host_lin_axis(g.y0, g.ysp, gny, (float)r)];
        q.x = std::min(q.x, r);
        q.y = std::max(q.y, r);
    }
    for (int cc = 1; cc <= c->nc; ++cc) {
        int2 &q = cells[s.ncy + host_lin_axis(g.x0, g.xsp, gnx, (float)cc)];
        q.x = std::min(q.x, cc);
        q.y = std::max(q.y, cc);
    }
    // one pixel of slack each side (the kernel re-checks membership)
    for (int k = 0; k < s.ncy + s.ncx; ++k) {
        const int n = k < s.ncy ? c->nr : c->nc;
        int2 &q = cells[k];
        if (q.x > q.y) {
            q = make_int2(1, 0);  // empty interval
            continue;
        }
        q.x = std::max(1, q.x - 1);
        q.y = std::min(n, q.y + 1);
    }
    int2 *dcells = c->ecells.ensure(cells.size());
    HH_HIP(hipMemcpyAsync(dcells, cells.data(), sizeof(int2) * cells.size(), hipMemcpyHostToDevice, c->stream));
    HH_HIP(hipStreamSynchronize(c->stream));
    s.rows = dcells;
    s.cols = dcells + s.ncy;
    const size_t nw = (size_t)grid.ny * grid.nx;
    c->estats.ensure(nw);
    c->ew.ensure(nw);
    c->esum.ensure(1);
    c->epart.ensure((size_t)8 * s.ncy * s.ncx);
    c->egrad.ensure((size_t)2 * gny * gnx);
    c->enode.ensure((size_t)2 * gny * gnx);
    return s;
}

void energy_nodes(hh_ctx *c, EnergySetup &s, const float *pu, const float *pv)
{
    if (!pu || !pv) throw Error(HH_EINVAL, "null node field");
    const size_t n = (size_t)s.g.gny * s.g.gnx;
    check_finite(pu, n, "node field u");
    check_finite(pv, n, "node field v");
    HH_HIP(hipMemcpyAsync(c->enode.p, pu, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HH_HIP(hipMemcpyAsync(c->enode.p + n, pv, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    s.g.pu = c->enode.p;
    s.g.pv = c->enode.p + n;
}

double energy_eval(hh_ctx *c, EnergySetup &s, const float *pu, const float *pv, float *gu, float *gv)
{
    energy_nodes(c, s, pu, pv);
    timed(c, "energy", s.g.w, [&] { launch_energy(s.g, c->estats.p, c->ew.p, c->esum.p, c->stream); });
    const size_t n = (size_t)s.g.gny * s.g.gnx;
    if (gu || gv) {
        timed(c, "energy_adjoint", s.g.w, [&] {
            launch_energy_adjoint(s.g, c->estats.p, s.rows, s.cols, s.ncy, s.ncx, c->epart.p, c->egrad.p,
                                  c->egrad.p + n, c->stream);
        });
        if (gu) HH_HIP(hipMemcpyAsync(gu, c->egrad.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
        if (gv) HH_HIP(hipMemcpyAsync(gv, c->egrad.p + n, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    }
    double E = 0.0;
    HH_HIP(hipMemcpyAsync(&E, c->esum.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HH_HIP(hipStreamSynchronize(c->stream));
    return E;
}

// Gradient descent on the node field with a backtracking step: the first
// trial moves the largest node by tau px (0.25, then twice the last accepted
// move, at most 1 px), halved until the energy decreases.
void refine_nodes(hh_ctx *c, EnergySetup &s, float *pu, float *pv, int niter, double *energies)
{
    const size_t n = (size_t)s.g.gny * s.g.gnx;
    std::vector<float> gu(n), gv(n), tu(n), tv(n);
    double E = energy_eval(c, s, pu, pv, gu.data(), gv.data());
    double tau = 0.25;
    int it = 0;
    for (; it < niter; ++it) {
        if (energies) energies[it] = E;
        double gmax = 0.0;
        for (size_t k = 0; k < n; ++k) gmax = std::max(gmax, (double)std::max(std::fabs(gu[k]), std::fabs(gv[k])));
        if (!(gmax > 0.0) || !std::isfinite(gmax) || !std::isfinite(E)) break;
        double alpha = tau / gmax, Et = E;
        bool ok = false;
        for (int bt = 0; bt < 8 && !ok; ++bt) {
            for (size_t k = 0; k < n; ++k) {
                tu[k] = (float)(pu[k] - alpha * gu[k]);
                tv[k] = (float)(pv[k] - alpha * gv[k]);
            }
            Et = energy_eval(c, s, tu.data(), tv.data(), nullptr, nullptr);
            if (Et < E)
                ok = true;
            else
                alpha /= 2;
        }
        if (!ok) break;
        std::copy(tu.begin(), tu.end(), pu);
        std::copy(tv.begin(), tv.end(), pv);
        E = energy_eval(c, s, pu, pv, gu.data(), gv.data());
        tau = std::min(1.0, 2.0 * alpha * gmax);
    }
    if (energies)
        for (int k = it; k <= niter; ++k) energies[k] = E;
}

void fetch(hh_ctx *c, hh_field *r)
{
    if (!c->have_result) throw Error(HH_ESTATE, "no result: call hh_run_piv first");
    const PassGrid &g = c->final_grid;  // the ensemble driver's grid
    if (r->ny != g.ny || r->nx != g.nx)
        throw Error(HH_EDIM, "result grid is " + std::to_string(g.ny) + "x" + std::to_string(g.nx) + ", got " +
                                 std::to_string(r->ny) + "x" + std::to_string(r->nx));
    const size_t ng = (size_t)g.ny * g.nx;
    auto d2h = [&](void *dst, const void *src, size_t bytes) {
        if (dst) HH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    };
    d2h(r->u, c->gu.p, sizeof(float) * ng);
    d2h(r->v, c->gv.p, sizeof(float) * ng);
    d2h(r->peak_ratio, c->gpr.p, sizeof(float) * ng);
    d2h(r->correlation_moment, c->gcm.p, sizeof(float) * ng);
    d2h(r->uncertainty_u, c->guu.p, sizeof(float) * ng);
    d2h(r->uncertainty_v, c->guv.p, sizeof(float) * ng);
    d2h(r->outliers, c->gout.p, ng);
    HH_HIP(hipStreamSynchronize(c->stream));
    if (r->x) std::memcpy(r->x, g.x.data(), sizeof(float) * g.nx);
    if (r->y) std::memcpy(r->y, g.y.data(), sizeof(float) * g.ny);
    if (r->mask) {
        if (c->final_gm.size() == ng)
            std::memcpy(r->mask, c->final_gm.data(), ng);
        else
            std::memset(r->mask, 0, ng);
    }
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char *hh_version(void) { return "hammerhead-mi355x 0.1 (gfx950)"; }

int hh_create(hh_ctx **out, int device)
{
    g_create_err.clear();
    if (!out) return HH_EINVAL;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        g_create_err = "no HIP device available";
        return HH_EDEVICE;
    }
    if (device < 0 || device >= n) {
        g_create_err = "device index out of range";
        return HH_EINVAL;
    }
    auto *c = new (std::nothrow) hh_ctx();
    if (!c) return HH_ENOMEM;
    c->device = device;
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side.fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side.join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->pre_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        g_create_err = std::string("stream creation failed: ") + hipGetErrorString(e);
        if (c->pre_join) (void)hipEventDestroy(c->pre_join);
        if (c->side.join) (void)hipEventDestroy(c->side.join);
        if (c->side.fork) (void)hipEventDestroy(c->side.fork);
        if (c->side.stream) (void)hipStreamDestroy(c->side.stream);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return HH_EDEVICE;
    }
    // this device's no-allocation scratch, resolved at set-up with the
    // device current
    c->scratch = scratch_on(device, g_create_err);
    if (!c->scratch) {
        (void)hipEventDestroy(c->pre_join);
        (void)hipEventDestroy(c->side.join);
        (void)hipEventDestroy(c->side.fork);
        (void)hipStreamDestroy(c->side.stream);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return HH_EDEVICE;
    }
    *out = c;
    return HH_OK;
}

int hh_destroy(hh_ctx *c)
{
    if (!c) return HH_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &t : c->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    engine_release(c);
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    if (c->copy_stream) {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    for (auto e : c->upload_ev) (void)hipEventDestroy(e);
    (void)hipStreamSynchronize(c->side.stream);
    (void)hipEventDestroy(c->pre_join);
    (void)hipEventDestroy(c->side.join);
    (void)hipEventDestroy(c->side.fork);
    (void)hipStreamDestroy(c->side.stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return HH_OK;
}

const char *hh_last_error(const hh_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int hh_set_form(hh_ctx *c, int form)
{
    return guarded(c, [&] {
        if (form != HH_FORM_FFT32 && form != HH_FORM_MFMA16) throw Error(HH_EINVAL, "unknown correlation form");
        c->form = form;
    });
}

int hh_check_passes(hh_ctx *c, const hh_pass *passes, int npasses)
{
    return guarded(c, [&] {
        if (npasses < 1) throw Error(HH_EINVAL, "at least one pass is required");
        for (int k = 0; k < npasses; ++k) check_pass_device(passes[k], true);
    });
}

int hh_grid_size(const hh_pass *p, int nr, int nc, int *ny, int *nx)
{
    if (!p || !ny || !nx) return HH_EINVAL;
    try {
        check_pass_valid(*p);
        PassGrid g = make_grid(*p, nr, nc);
        *ny = g.ny;
        *nx = g.nx;
        return HH_OK;
    } catch (const Error &e) {
        return e.code;
    }
}

int hh_set_pair(hh_ctx *c, const float *A, const float *B, int nr, int nc)
{
    return guarded(c, [&] {
        if (!A || !B) throw Error(HH_EINVAL, "null image");
        if (nr < 2 || nc < 2) throw Error(HH_EDIM, "images must be at least 2 x 2");
        const size_t n = (size_t)nr * nc;
        c->ownA.ensure(n);
        c->ownB.ensure(n);
        HH_HIP(hipMemcpyAsync(c->ownA.p, A, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        HH_HIP(hipMemcpyAsync(c->ownB.p, B, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        HH_HIP(hipStreamSynchronize(c->stream));
        c->A = c->ownA.p;
        c->B = c->ownB.p;
        if (nr != c->nr || nc != c->nc) c->jobcache.clear();
        c->nr = nr;
        c->nc = nc;
        c->coefs_valid = c->warp_valid = c->have_result = false;
        c->eng.have_result = false;
    });
}

int hh_set_pair_device(hh_ctx *c, const float *dA, const float *dB, int nr, int nc)
{
    return guarded(c, [&] {
        if (!dA || !dB) throw Error(HH_EINVAL, "null image");
        if (nr < 2 || nc < 2) throw Error(HH_EDIM, "images must be at least 2 x 2");
        c->A = dA;
        c->B = dB;
        if (nr != c->nr || nc != c->nc) c->jobcache.clear();
        c->nr = nr;
        c->nc = nc;
        c->coefs_valid = c->warp_valid = c->have_result = false;
        c->eng.have_result = false;
    });
}

int hh_set_dewarp_map(hh_ctx *c, const double *rows, const double *cols, const uint8_t *mask, int ny, int nx, int nr,
                      int nc)
{
    return guarded(c, [&] {
        if (!rows || !cols) throw Error(HH_EINVAL, "null coordinate map");
        if (ny < 1 || nx < 1) throw Error(HH_EDIM, "empty dewarp grid");
        if (nr < 2 || nc < 2 || nr > 32767 || nc > 32767)
            throw Error(HH_EDIM, "camera frames must be 2 to 32767 pixels per side");
        const size_t n = (size_t)ny * nx;
        if (n >= (size_t)1 << 31) throw Error(HH_EDIM, "dewarp grid too large");
        std::vector<int2> m(n);
        for (size_t e = 0; e < n; ++e) {
            const double r = rows[e], q = cols[e];
            // ImageDewarper's view test (src/dewarp.jl:205-215): finite and
            // within [1, nr] x [1, nc]; the mask marks the others too
            const bool in = !(mask && mask[e]) && std::isfinite(r) && std::isfinite(q) && r >= 1.0 &&
                            r <= (double)nr && q >= 1.0 && q <= (double)nc;
            m[e] = in ? make_int2((int)std::llround(r * 65536.0), (int)std::llround(q * 65536.0))
                      : make_int2(INT32_MIN, INT32_MIN);
        }
        c->dwmap.ensure(n);
        HH_HIP(hipMemcpyAsync(c->dwmap.p, m.data(), sizeof(int2) * n, hipMemcpyHostToDevice, c->stream));
        HH_HIP(hipStreamSynchronize(c->stream));
        c->dw_ny = ny;
        c->dw_nx = nx;
        c->dw_nr = nr;
        c->dw_nc = nc;
    });
}

namespace {

// Upload nframes host frames of the dewarper's camera size, prefilter them
// on the device and resample them onto the grid into oA (and oB).
void dewarp_frames(hh_ctx *c, const float *A, const float *B, int nr, int nc, float *oA, float *oB)
{
    if (c->dw_ny == 0) throw Error(HH_ESTATE, "no dewarp map: call hh_set_dewarp_map first");
    if (nr != c->dw_nr || nc != c->dw_nc) throw Error(HH_EDIM, "frame size does not match the dewarp map's camera");
    const size_t npx = (size_t)nr * nc, ncf = (size_t)(nr + 2) * (nc + 2);
    float *raw = c->dwraw.ensure(2 * npx);
    float *coef = c->dwcoef.ensure(2 * ncf);
    float *tmp = c->dwtmp.ensure((size_t)2 * (nr + 2) * nc);
    HH_HIP(hipMemcpyAsync(raw, A, sizeof(float) * npx, hipMemcpyHostToDevice, c->stream));
    if (B) HH_HIP(hipMemcpyAsync(raw + npx, B, sizeof(float) * npx, hipMemcpyHostToDevice, c->stream));
    timed(c, "prefilter", 0, [&] {
        launch_prefilter_pair(raw, B ? raw + npx : raw, nr, nr, nc, coef, coef + ncf, tmp, c->stream);
    });
    timed(c, "dewarp", 0, [&] {
        launch_dewarp(c->dwmap.p, c->dw_ny, c->dw_nx, coef, B ? coef + ncf : nullptr, nr, nc, oA, B ? oB : nullptr,
                      c->stream);
    });
}

}  // namespace

int hh_dewarp(hh_ctx *c, const float *img, int nr, int nc, float *out)
{
    return guarded(c, [&] {
        if (!img || !out) throw Error(HH_EINVAL, "null argument");
        const size_t n = (size_t)c->dw_ny * c->dw_nx;
        float *o = c->dwout.ensure(n > 0 ? n : 1);
        dewarp_frames(c, img, nullptr, nr, nc, o, nullptr);
        HH_HIP(hipMemcpyAsync(out, o, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_dewarp_pair(hh_ctx *c, const float *A, const float *B, int nr, int nc)
{
    return guarded(c, [&] {
        if (!A || !B) throw Error(HH_EINVAL, "null image");
        const int ny = c->dw_ny, nx = c->dw_nx;
        if (ny > 0 && (ny < 2 || nx < 2)) throw Error(HH_EDIM, "the dewarped pair must be at least 2 x 2");
        const size_t n = (size_t)ny * nx;
        c->ownA.ensure(n > 0 ? n : 1);
        c->ownB.ensure(n > 0 ? n : 1);
        dewarp_frames(c, A, B, nr, nc, c->ownA.p, c->ownB.p);
        HH_HIP(hipStreamSynchronize(c->stream));
        c->A = c->ownA.p;
        c->B = c->ownB.p;
        if (ny != c->nr || nx != c->nc) c->jobcache.clear();
        c->nr = ny;
        c->nc = nx;
        c->coefs_valid = c->warp_valid = c->have_result = false;
        c->eng.have_result = false;
    });
}

int hh_prefilter(hh_ctx *c)
{
    return guarded(c, [&] { do_prefilter(c); });
}

int hh_set_coefs(hh_ctx *c, const float *cA, const float *cB, int nr, int nc)
{
    return guarded(c, [&] {
        if (!cA || !cB) throw Error(HH_EINVAL, "null coefficient array");
        if (nr < 2 || nc < 2) throw Error(HH_EDIM, "images must be at least 2 x 2");
        const size_t n = (size_t)(nr + 2) * (nc + 2);
        c->coefA.ensure(n);
        c->coefB.ensure(n);
        HH_HIP(hipMemcpyAsync(c->coefA.p, cA, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        HH_HIP(hipMemcpyAsync(c->coefB.p, cB, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        HH_HIP(hipStreamSynchronize(c->stream));
        if (nr != c->nr || nc != c->nc) {
            // a raw pair of another size no longer describes these images
            c->A = c->B = nullptr;
            c->jobcache.clear();
        }
        c->nr = nr;
        c->nc = nc;
        c->coefs_valid = true;
        c->warp_valid = c->have_result = false;
        c->eng.have_result = false;
    });
}

int hh_deform(hh_ctx *c, const float *py, int gny, const float *px, int gnx, const float *pu, const float *pv)
{
    return guarded(c, [&] {
        if (!c->coefs_valid) throw Error(HH_ESTATE, "no interpolant: call hh_prefilter or hh_set_coefs first");
        if (gny < 1 || gnx < 1 || !py || !px || !pu || !pv) throw Error(HH_EINVAL, "empty predictor grid");
        const size_t n = (size_t)gny * gnx;
        check_knots(py, gny, px, gnx);
        check_finite(pu, n, "predictor u");
        check_finite(pv, n, "predictor v");
        upload(c, c->pu, pu, n);
        upload(c, c->pv, pv, n);
        c->gny = gny;
        c->gnx = gnx;
        c->y0 = py[0];
        c->x0 = px[0];
        c->ysp = gny > 1 ? py[1] - py[0] : 1.f;
        c->xsp = gnx > 1 ? px[1] - px[0] : 1.f;
        do_deform(c);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_get_images(hh_ctx *c, int src, float *A, float *B)
{
    return guarded(c, [&] {
        const size_t n = (size_t)c->nr * c->nc;
        download(c, A, src_image(c, src, true), n);
        download(c, B, src_image(c, src, false), n);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_get_coefs(hh_ctx *c, float *cA, float *cB)
{
    return guarded(c, [&] {
        if (!c->coefs_valid) throw Error(HH_ESTATE, "no interpolant: call hh_prefilter first");
        const size_t n = (size_t)(c->nr + 2) * (c->nc + 2);
        download(c, cA, c->coefA.p, n);
        download(c, cB, c->coefB.p, n);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_process_windows(hh_ctx *c, const hh_pass *pass, int src, const int32_t *jobs, int njobs, int ny, int nx,
                       float *u, float *v, float *peak_ratio, float *correlation_moment, float *alt_u, float *alt_v,
                       float *unc_u, float *unc_v, int32_t *peak_locs, float *peak_vals, int32_t *found)
{
    return guarded(c, [&] {
        if (!pass) throw Error(HH_EINVAL, "null pass");
        check_pass_device(*pass, false);
        if (njobs < 0 || (njobs > 0 && !jobs)) throw Error(HH_EINVAL, "invalid job list");
        if (!u || !v || !peak_ratio || !correlation_moment) throw Error(HH_EINVAL, "null output grid");
        const float *dA = src_image(c, src, true), *dB = src_image(c, src, false);
        const int w = pass->window_size[0];
        for (int k = 0; k < njobs; ++k) {
            const int gi = jobs[4 * k], gj = jobs[4 * k + 1], rs = jobs[4 * k + 2], cs = jobs[4 * k + 3];
            if (gi < 1 || gi > ny || gj < 1 || gj > nx) throw Error(HH_EDIM, "job grid index out of range");
            check_job_area(c, *pass, rs, cs);
        }
        if (njobs == 0) return;
        const size_t ng = (size_t)ny * nx;
        const int nalt = pass->n_peaks - 1;
        const int K = std::max(pass->n_peaks, 2);
        CorrConfig cfg = make_corr(c, *pass);
        const int32_t *dj = upload(c, c->hookjobs, jobs, (size_t)4 * njobs);
        const HookSource hs = hook_source(c, *pass, jobs, njobs, dA, dB, dj);
        CorrOut out{};
        out.u = upload(c, c->hk[0], u, ng);
        out.v = upload(c, c->hk[1], v, ng);
        out.pr = upload(c, c->hk[2], peak_ratio, ng);
        out.cm = upload(c, c->hk[3], correlation_moment, ng);
        if (alt_u && alt_v && nalt > 0) {
            out.alt_u = upload(c, c->hk[4], alt_u, ng * nalt);
            out.alt_v = upload(c, c->hk[5], alt_v, ng * nalt);
        }
        if (peak_locs) out.locs = upload<int32_t>(c, c->hki[0], nullptr, (size_t)2 * K * njobs);
        if (peak_vals) out.vals = upload<float>(c, c->hk[6], nullptr, (size_t)K * njobs);
        if (found) out.found = upload<int32_t>(c, c->hki[1], nullptr, (size_t)njobs);
        out.ny = ny;
        out.nx = nx;
        float *duu = nullptr, *duv = nullptr;
        if (unc_u || unc_v) {
            duu = upload(c, c->hk[7], unc_u, ng);
            duv = upload(c, c->hk[8], unc_v, ng);
            out.unc_u = duu;
            out.unc_v = duv;
        }
        bool fused = false;
        timed(c, out.unc_u ? "correlate_uq" : "correlate", w,
              [&] { fused = launch_correlate(cfg, hs.A, hs.B, hs.ld, 0, hs.jobs, njobs, out, c->stream); });
        if ((unc_u || unc_v) && !fused) {
            const ConfigBufs &cb = get_config(c, *pass);
            const HookSource us = pass->search_area_size[0] == w ? hs
                                                                 : hook_source(c, *pass, jobs, njobs, dA, dB, dj, true, 1);
            timed(c, "uq", w, [&] {
                launch_uq(w, cb.apod.p, us.A, us.B, us.ld, us.jobs, njobs, ny, unc_u ? duu : nullptr,
                          unc_v ? duv : nullptr, nullptr, false, c->stream);
            });
        }
        download(c, u, out.u, ng);
        download(c, v, out.v, ng);
        download(c, peak_ratio, out.pr, ng);
        download(c, correlation_moment, out.cm, ng);
        if (out.alt_u) {
            download(c, alt_u, out.alt_u, ng * nalt);
            download(c, alt_v, out.alt_v, ng * nalt);
        }
        if (unc_u) download(c, unc_u, duu, ng);
        if (unc_v) download(c, unc_v, duv, ng);
        if (peak_locs) download(c, peak_locs, out.locs, (size_t)2 * K * njobs);
        if (peak_vals) download(c, peak_vals, out.vals, (size_t)K * njobs);
        if (found) download(c, found, out.found, (size_t)njobs);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_correlation_planes(hh_ctx *c, const hh_pass *pass, int src, const int32_t *jobs, int njobs, float *planes)
{
    return guarded(c, [&] {
        if (!pass || !planes) throw Error(HH_EINVAL, "null argument");
        check_pass_device(*pass, false);
        if (njobs < 0 || (njobs > 0 && !jobs)) throw Error(HH_EINVAL, "invalid job list");
        const float *dA = src_image(c, src, true), *dB = src_image(c, src, false);
        const int M = pass->padding ? 2 * pass->search_area_size[0] : pass->search_area_size[0];
        int maxgi = 1, maxgj = 1;
        for (int k = 0; k < njobs; ++k) {
            const int rs = jobs[4 * k + 2], cs = jobs[4 * k + 3];
            if (jobs[4 * k] < 1 || jobs[4 * k + 1] < 1) throw Error(HH_EDIM, "job grid index out of range");
            check_job_area(c, *pass, rs, cs);
            maxgi = std::max(maxgi, jobs[4 * k]);
            maxgj = std::max(maxgj, jobs[4 * k + 1]);
        }
        if (njobs <= 0) return;
        CorrConfig cfg = make_corr(c, *pass);
        const int32_t *dj = upload(c, c->hookjobs, jobs, (size_t)4 * njobs);
        const HookSource hs = hook_source(c, *pass, jobs, njobs, dA, dB, dj);
        const size_t ng = (size_t)maxgi * maxgj;
        CorrOut out{};
        out.u = upload<float>(c, c->hk[0], nullptr, ng);
        out.v = upload<float>(c, c->hk[1], nullptr, ng);
        out.pr = upload<float>(c, c->hk[2], nullptr, ng);
        out.cm = upload<float>(c, c->hk[3], nullptr, ng);
        out.planes = upload<float>(c, c->hk[9], nullptr, (size_t)M * M * njobs);
        out.ny = maxgi;
        out.nx = maxgj;
        launch_correlate(cfg, hs.A, hs.B, hs.ld, 0, hs.jobs, njobs, out, c->stream);
        download(c, planes, out.planes, (size_t)M * M * njobs);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_uncertainty_sweep(hh_ctx *c, const hh_pass *pass, int src, const int32_t *jobs, int njobs, int ny, int nx,
                         float *unc_u, float *unc_v, double *stats)
{
    return guarded(c, [&] {
        if (!pass) throw Error(HH_EINVAL, "null pass");
        check_pass_device(*pass, false);
        if (pass->window_size[0] > 64) throw Error(HH_EUNSUPPORTED, "uncertainty supports windows up to 64 px");
        const float *dA = src_image(c, src, true), *dB = src_image(c, src, false);
        const int w = pass->window_size[0];
        for (int k = 0; k < njobs; ++k) {
            const int gi = jobs[4 * k], gj = jobs[4 * k + 1], rs = jobs[4 * k + 2], cs = jobs[4 * k + 3];
            if (gi < 1 || gi > ny || gj < 1 || gj > nx) throw Error(HH_EDIM, "job grid index out of range");
            if (rs < 1 || cs < 1 || rs + w - 1 > c->nr || cs + w - 1 > c->nc)
                throw Error(HH_EDIM, "window outside the image");
        }
        if (njobs <= 0) return;
        const size_t ng = (size_t)ny * nx;
        const int32_t *dj = upload(c, c->hookjobs, jobs, (size_t)4 * njobs);
        const HookSource hs = hook_source(c, *pass, jobs, njobs, dA, dB, dj, true);
        float *duu = upload(c, c->hk[7], unc_u, ng);
        float *duv = upload(c, c->hk[8], unc_v, ng);
        double *dst = stats ? c->hkd.ensure((size_t)2 * HH_UQ_NSTATS * njobs) : nullptr;
        const ConfigBufs &cb = get_config(c, *pass);
        timed(c, "uq", w, [&] {
            launch_uq(w, cb.apod.p, hs.A, hs.B, hs.ld, hs.jobs, njobs, ny, unc_u ? duu : nullptr,
                      unc_v ? duv : nullptr, dst, false, c->stream);
        });
        if (unc_u) download(c, unc_u, duu, ng);
        if (unc_v) download(c, unc_v, duv, ng);
        if (stats) download(c, stats, dst, (size_t)2 * HH_UQ_NSTATS * njobs);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_validate_and_replace(hh_ctx *c, const hh_pass *pass, int ny, int nx, int force_replace, float *u, float *v,
                            const float *peak_ratio, const float *correlation_moment, const float *alt_u,
                            const float *alt_v, const uint8_t *mask, uint8_t *outliers)
{
    return guarded(c, [&] {
        if (!pass || !u || !v || !peak_ratio || !correlation_moment || !outliers)
            throw Error(HH_EINVAL, "null argument");
        check_pass_valid(*pass);
        if (pass->uod_neighborhood > 3) throw Error(HH_EUNSUPPORTED, "uod_neighborhood up to 3 on the device");
        if (ny < 1 || nx < 1) throw Error(HH_EDIM, "empty grid");
        const size_t ng = (size_t)ny * nx;
        bool anym = false;
        if (mask)
            for (size_t k = 0; k < ng && !anym; ++k) anym = mask[k] != 0;
        const int nalt = pass->n_peaks - 1;
        ValidateArgs va{};
        va.grids = one_grid(ny, nx);
        va.pass = pass;
        va.force_replace = force_replace;
        va.u = upload(c, c->hk[0], u, ng);
        va.v = upload(c, c->hk[1], v, ng);
        va.pr = upload(c, c->hk[2], peak_ratio, ng);
        va.cm = upload(c, c->hk[3], correlation_moment, ng);
        if (alt_u && alt_v && nalt > 0) {
            va.alt_u = upload(c, c->hk[4], alt_u, ng * nalt);
            va.alt_v = upload(c, c->hk[5], alt_v, ng * nalt);
        }
        va.outliers = c->gout.ensure(ng);
        va.su = c->gsu.ensure(ng);
        va.sv = c->gsv.ensure(ng);
        va.sflag = c->gsflag.ensure(ng);
        va.count = c->gcount.ensure(1);
        if (anym) {
            uint8_t *dm = c->gmask.ensure(ng);
            std::vector<uint8_t> m(ng);
            for (size_t k = 0; k < ng; ++k) m[k] = mask[k] != 0;
            HH_HIP(hipMemcpyAsync(dm, m.data(), ng, hipMemcpyHostToDevice, c->stream));
            HH_HIP(hipStreamSynchronize(c->stream));
            va.mask = dm;
        }
        c->have_result = false;
        timed(c, "validate", pass->window_size[0], [&] { launch_validate(va, c->stream); });
        download(c, u, va.u, ng);
        download(c, v, va.v, ng);
        download(c, outliers, va.outliers, ng);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_validate_and_replace_noalloc(hh_ctx *c, const hh_pass *pass, int ny, int nx, int force_replace, float *u,
                                    float *v, const float *peak_ratio, const float *correlation_moment,
                                    const float *alt_u, const float *alt_v, const uint8_t *mask, uint8_t *outliers)
{
    return guarded(c, [&] {
        if (!pass || !u || !v || !peak_ratio || !correlation_moment || !outliers)
            throw Error(HH_EINVAL, "null argument");
        check_pass_valid(*pass);
        if (pass->uod_neighborhood > 3) throw Error(HH_EUNSUPPORTED, "uod_neighborhood up to 3 on the device");
        if (ny < 1 || nx < 1) throw Error(HH_EDIM, "empty grid");
        if ((size_t)ny * nx > kNoAllocCells)
            throw Error(HH_EINVAL, "grid of " + std::to_string((size_t)ny * nx) + " cells exceeds the " +
                                       std::to_string(kNoAllocCells) + "-cell scratch of the no-allocation entry point");
        if (!c->scratch) throw Error(HH_ESTATE, "no scratch on this device");
        ValidateArgs va{};
        va.grids = one_grid(ny, nx);
        va.pass = pass;
        va.force_replace = force_replace;
        va.u = u;
        va.v = v;
        va.pr = peak_ratio;
        va.cm = correlation_moment;
        if (alt_u && alt_v && pass->n_peaks > 1) {
            va.alt_u = alt_u;
            va.alt_v = alt_v;
        }
        va.outliers = outliers;
        // the counter is cleared by the chain's first kernel; the other
        // scratch cells are written before they are read
        va.count = reinterpret_cast<int *>(c->scratch);
        va.su = reinterpret_cast<float *>(c->scratch + kScratchCount);
        va.sv = va.su + kNoAllocCells;
        va.sflag = reinterpret_cast<uint8_t *>(va.sv + kNoAllocCells);
        va.mask = mask;
        launch_validate(va, c->stream);
    });
}

int hh_smooth_field(hh_ctx *c, int ny, int nx, const float *f, float *out)
{
    return guarded(c, [&] {
        if (!f || !out) throw Error(HH_EINVAL, "null argument");
        if (ny < 1 || nx < 1) throw Error(HH_EDIM, "empty grid");
        const size_t n = (size_t)ny * nx;
        float *df = upload(c, c->hk[10], f, n);
        float *dout = c->hk[11].ensure(n);
        launch_smooth(one_grid(ny, nx), df, dout, c->stream);
        download(c, out, (const float *)dout, n);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_run_piv(hh_ctx *c, const hh_pass *passes, int npasses, int predictor_smoothing, hh_field *result)
{
    return guarded(c, [&] {
        if (!passes) throw Error(HH_EINVAL, "null pass schedule");
        need_pair(c);
        engine_single_layout(c);
        c->have_result = false;
        engine_run(c, passes, npasses, predictor_smoothing);
        if (result) engine_fetch(c, 0, result);
    });
}

int hh_fetch_result(hh_ctx *c, hh_field *result)
{
    return guarded(c, [&] {
        if (!result) throw Error(HH_EINVAL, "null result");
        if (c->eng.have_result)
            engine_fetch(c, 0, result);
        else
            fetch(c, result);
    });
}

int hh_set_mask(hh_ctx *c, const uint8_t *mask, int nr, int nc, double mask_threshold)
{
    return guarded(c, [&] { engine_set_mask(c, mask, nr, nc, mask_threshold); });
}

int hh_set_pairs(hh_ctx *c, const float *const *A, const float *const *B, const int32_t *nr, const int32_t *nc,
                 int npairs)
{
    return guarded(c, [&] { engine_set_pairs(c, A, B, nr, nc, npairs); });
}

int hh_run_piv_batch(hh_ctx *c, const hh_pass *passes, int npasses, int predictor_smoothing, hh_field *results)
{
    return guarded(c, [&] {
        if (!passes) throw Error(HH_EINVAL, "null pass schedule");
        if (c->eng.lay.npairs < 1 || !c->eng.lay.A) throw Error(HH_ESTATE, "no batch: call hh_set_pairs first");
        c->have_result = false;
        engine_run(c, passes, npasses, predictor_smoothing);
        if (results)
            for (int q = 0; q < c->eng.lay.npairs; ++q) engine_fetch(c, q, &results[q]);
    });
}

int hh_fetch_pair(hh_ctx *c, int q, hh_field *result)
{
    return guarded(c, [&] {
        if (!result) throw Error(HH_EINVAL, "null result");
        engine_fetch(c, q, result);
    });
}

int hh_pair_grid(hh_ctx *c, int q, int *ny, int *nx)
{
    return guarded(c, [&] {
        if (!ny || !nx) throw Error(HH_EINVAL, "null argument");
        if (c->eng.plan.empty()) throw Error(HH_ESTATE, "no schedule has been planned");
        if (q < 0 || q >= c->eng.lay.npairs) throw Error(HH_EINVAL, "pair index out of range");
        *ny = c->eng.plan.back()->ny[q];
        *nx = c->eng.plan.back()->nx[q];
    });
}

int hh_set_graphs(hh_ctx *c, int enable)
{
    return guarded(c, [&] {
        engine_set_graphs(c, enable != 0);
    });
}

int hh_set_autotune(hh_ctx *c, int enable)
{
    return guarded(c, [&] {
        if (enable < 0 || enable > 2) throw Error(HH_EINVAL, "autotune level must be 0, 1 or 2");
        const int level = enable == 0 ? c->tune.level : enable;
        if (!enable || level != c->tune.level) {
            c->tune.entries.clear();
            c->eng.key.clear();  // re-plan with the new choices
        }
        c->tune.enabled = enable != 0;
        c->tune.level = level;
    });
}

int hh_tune_set_family(hh_ctx *c, const char *key, int family)
{
    return guarded(c, [&] {
        if (!key) throw Error(HH_EINVAL, "null key");
        if (family < 0 || family > 2) throw Error(HH_EINVAL, "family must be 0, 1 or 2");
        c->tune.entries[key].family = family;
        c->eng.key.clear();
    });
}

int hh_tune_report(hh_ctx *c, char *buf, size_t len, size_t *needed)
{
    return guarded(c, [&] {
        const std::string r = engine_tune_report(c);
        if (needed) *needed = r.size() + 1;
        if (buf && len) {
            const size_t n = std::min(len - 1, r.size());
            std::memcpy(buf, r.data(), n);
            buf[n] = 0;
        }
    });
}

int hh_tune_set(hh_ctx *c, const char *key, int waves_per_simd, int linear_order)
{
    return guarded(c, [&] {
        if (!key) throw Error(HH_EINVAL, "null key");
        if (waves_per_simd < 0 || waves_per_simd > 8) throw Error(HH_EINVAL, "waves_per_simd must be 0-8");
        TuneEntry &e = c->tune.entries[key];  // keeps a preloaded family
        e.waves_per_simd = waves_per_simd;
        e.linear_order = linear_order != 0;
        e.ms.clear();
        c->eng.key.clear();  // re-plan with the choice
    });
}

int hh_graph_replays(hh_ctx *c, int64_t *replays)
{
    return guarded(c, [&] {
        if (replays) *replays = c->eng.graph_replays;
    });
}

int hh_ensemble_begin(hh_ctx *c, const hh_pass *pass, const int32_t *jobs, int njobs, int uncertainty)
{
    return guarded(c, [&] {
        if (!pass) throw Error(HH_EINVAL, "null pass");
        if (njobs < 0 || (njobs > 0 && !jobs)) throw Error(HH_EINVAL, "invalid job list");
        for (int k = 0; k < njobs; ++k)
            if (jobs[4 * k] < 1 || jobs[4 * k + 1] < 1 || jobs[4 * k + 2] < 1 || jobs[4 * k + 3] < 1)
                throw Error(HH_EDIM, "job indices are 1-based");
        c->ens.hjobs.assign(jobs, jobs + (size_t)4 * njobs);
        const int32_t *dj = upload(c, c->ensjobs, jobs, (size_t)4 * njobs);
        ens_begin(c, *pass, dj, njobs, uncertainty != 0);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_ensemble_add(hh_ctx *c, int src)
{
    return guarded(c, [&] {
        if (!c->ens.active) throw Error(HH_ESTATE, "no ensemble: call hh_ensemble_begin first");
        const float *dA = src_image(c, src, true), *dB = src_image(c, src, false);
        for (int k = 0; k < c->ens.njobs; ++k)
            check_job_area(c, c->ens.pass, c->ens.hjobs[4 * k + 2], c->ens.hjobs[4 * k + 3]);
        const hh_pass &ep = c->ens.pass;
        const HookSource hs = hook_source(c, ep, c->ens.hjobs.data(), c->ens.njobs, dA, dB, c->ens.jobs);
        const HookSource us = ep.search_area_size[0] == ep.window_size[0]
                                  ? hs
                                  : hook_source(c, ep, c->ens.hjobs.data(), c->ens.njobs, dA, dB, c->ens.jobs, true, 1);
        ens_add(c, dA, dB, &hs, &us);
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_ensemble_analyze(hh_ctx *c, int ny, int nx, float *u, float *v, float *peak_ratio, float *correlation_moment,
                        float *alt_u, float *alt_v, float *unc_u, float *unc_v)
{
    return guarded(c, [&] {
        auto &e = c->ens;
        if (!e.active) throw Error(HH_ESTATE, "no ensemble: call hh_ensemble_begin first");
        if (!u || !v || !peak_ratio || !correlation_moment) throw Error(HH_EINVAL, "null output grid");
        for (int k = 0; k < e.njobs; ++k)
            if (e.hjobs[4 * k] > ny || e.hjobs[4 * k + 1] > nx) throw Error(HH_EDIM, "job grid index out of range");
        if (e.njobs == 0) return;
        const size_t ng = (size_t)ny * nx;
        const int nalt = e.pass.n_peaks - 1;
        float *du = upload(c, c->hk[0], u, ng), *dv = upload(c, c->hk[1], v, ng);
        float *dpr = upload(c, c->hk[2], peak_ratio, ng), *dcm = upload(c, c->hk[3], correlation_moment, ng);
        float *dau = nullptr, *dav = nullptr;
        if (alt_u && alt_v && nalt > 0) {
            dau = upload(c, c->hk[4], alt_u, ng * nalt);
            dav = upload(c, c->hk[5], alt_v, ng * nalt);
        }
        float *duu = nullptr, *duv = nullptr;
        if (e.unc) {
            duu = upload(c, c->hk[7], unc_u, ng);
            duv = upload(c, c->hk[8], unc_v, ng);
        }
        ens_analyze(c, ny, nx, du, dv, dpr, dcm, dau, dav, duu, duv);
        download(c, u, du, ng);
        download(c, v, dv, ng);
        download(c, peak_ratio, dpr, ng);
        download(c, correlation_moment, dcm, ng);
        if (dau) {
            download(c, alt_u, dau, ng * nalt);
            download(c, alt_v, dav, ng * nalt);
        }
        if (e.unc) {
            download(c, unc_u, duu, ng);
            download(c, unc_v, duv, ng);
        }
        HH_HIP(hipStreamSynchronize(c->stream));
    });
}

int hh_ensemble_planes(hh_ctx *c, double *planes)
{
    return guarded(c, [&] {
        auto &e = c->ens;
        if (!e.active) throw Error(HH_ESTATE, "no ensemble: call hh_ensemble_begin first");
        if (!planes) throw Error(HH_EINVAL, "null argument");
        const int M = e.cfg.M;
        std::vector<double> h(e.per * e.njobs);
        download(c, h.data(), (const double *)c->ensacc.p, h.size());
        HH_HIP(hipStreamSynchronize(c->stream));
        const int lo = e.cfg.box_lo, nb = e.cfg.box_hi - e.cfg.box_lo + 1;
        for (int k = 0; k < e.njobs; ++k) {
            double *dst = planes + (size_t)M * M * k;
            const double *src = h.data() + e.per * k;
            if (!e.box) {
                std::memcpy(dst, src, sizeof(double) * M * M);
                continue;
            }
            std::fill(dst, dst + (size_t)M * M, 0.0);
            for (int j = 0; j < nb; ++j)
                for (int i = 0; i < nb; ++i) dst[(lo + i) + (size_t)(lo + j) * M] = src[i + (size_t)j * nb];
        }
    });
}

int hh_run_piv_ensemble(hh_ctx *c, const float *const *A, const float *const *B, int npairs, int nr, int nc,
                        const hh_pass *passes, int npasses, int predictor_smoothing, hh_field *result)
{
    return guarded(c, [&] {
        if (!passes) throw Error(HH_EINVAL, "null pass schedule");
        run_ensemble(c, A, B, npairs, nr, nc, passes, npasses, predictor_smoothing);
        if (result) fetch(c, result);
    });
}

int hh_set_energy_sampling(hh_ctx *c, int sampling)
{
    return guarded(c, [&] {
        if (sampling != HH_SAMPLE_CUBIC && sampling != HH_SAMPLE_BILINEAR)
            throw Error(HH_EINVAL, "sampling must be HH_SAMPLE_CUBIC or HH_SAMPLE_BILINEAR");
        c->energy_bilinear = sampling == HH_SAMPLE_BILINEAR;
    });
}

int hh_energy(hh_ctx *c, const hh_pass *pass, const float *py, int gny, const float *px, int gnx, const float *pu,
              const float *pv, double *energy)
{
    return guarded(c, [&] {
        if (!pass || !energy) throw Error(HH_EINVAL, "null argument");
        EnergySetup s = energy_setup(c, *pass, py, gny, px, gnx);
        *energy = energy_eval(c, s, pu, pv, nullptr, nullptr);
    });
}

int hh_grad_energy(hh_ctx *c, const hh_pass *pass, const float *py, int gny, const float *px, int gnx, const float *pu,
                   const float *pv, double *energy, float *grad_u, float *grad_v)
{
    return guarded(c, [&] {
        if (!pass || !grad_u || !grad_v) throw Error(HH_EINVAL, "null argument");
        EnergySetup s = energy_setup(c, *pass, py, gny, px, gnx);
        const double E = energy_eval(c, s, pu, pv, grad_u, grad_v);
        if (energy) *energy = E;
    });
}

int hh_refine_nodes(hh_ctx *c, const hh_pass *pass, const float *py, int gny, const float *px, int gnx, float *pu,
                    float *pv, int niter, double *energies)
{
    return guarded(c, [&] {
        if (!pass) throw Error(HH_EINVAL, "null argument");
        if (niter < 0) throw Error(HH_EINVAL, "niter must be non-negative");
        EnergySetup s = energy_setup(c, *pass, py, gny, px, gnx);
        refine_nodes(c, s, pu, pv, niter, energies);
    });
}

void *hh_stream(hh_ctx *c) { return c ? (void *)c->stream : nullptr; }

int hh_synchronize(hh_ctx *c)
{
    return guarded(c, [&] { HH_HIP(hipStreamSynchronize(c->stream)); });
}

int hh_set_timing(hh_ctx *c, int enable)
{
    return guarded(c, [&] {
        drain_timing(c);
        c->timing = enable != 0;
        c->timing_family.clear();
        c->ktime.clear();
    });
}

int hh_set_timing_only(hh_ctx *c, const char *family, int window)
{
    return guarded(c, [&] {
        drain_timing(c);
        c->timing = true;
        c->timing_family = family ? family : "";
        c->timing_w = window;
        c->ktime.clear();
    });
}

int hh_kernel_time(hh_ctx *c, const char *family, int window, double *ms_total, int64_t *launches)
{
    return guarded(c, [&] {
        if (!family || !ms_total || !launches) throw Error(HH_EINVAL, "null argument");
        drain_timing(c);
        auto it = c->ktime.find({std::string(family), window});
        *ms_total = it == c->ktime.end() ? 0.0 : it->second.first;
        *launches = it == c->ktime.end() ? 0 : it->second.second;
    });
}

}  // extern "C"
