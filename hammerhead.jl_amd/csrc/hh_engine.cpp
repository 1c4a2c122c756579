// hh_engine.cpp — run_piv on the device for one pair or for a batch of pairs
// of uneven sizes.
//
// The orchestration follows run_piv / piv_pass (src/pipeline.jl:361-449,
// 655-832) per pair: each pass builds its predictor from the previous
// result (build_predictor, src/pipeline.jl:454-464), every sweep deforms the
// prefiltered pair by it, correlates all windows, adds the predictor back,
// validates and replaces, iterating passes rebuild their own predictor
// between sweeps, and the Wieneke uncertainty runs once on the final warp. A
// batch is run_piv_sequence (src/io.jl:229-247) without the serial loop: the
// pairs are packed into one column-major image atlas (shelves of equal
// width, stacked along the rows), so one correlation launch per sweep covers
// the windows of every pair — window origins are atlas coordinates and each
// job's grid index is its cell in the concatenated vector grids. Deformation,
// predictor node values, validation and smoothing are one launch each over
// all pairs through per-pair tables (no neighbourhood crosses a pair).
//
// With a fixed sweep count in every pass (max_iterations = 1 or
// convergence_tol = 0) the whole schedule is a pure sequence of launches on
// the context stream: it is captured once into a hipGraph and replayed.
// Otherwise each pair stops iterating on its own q95 field change
// (field_change, src/pipeline.jl:842-866, read back after every sweep) and
// later sweeps run on the pairs that have not converged.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "hh_ctx.h"

namespace hh {
namespace host {

namespace {

// ---- elementwise kernels over the enumerated cells of a Grids ---------------

__global__ void k_gfill(float *__restrict__ p, float *__restrict__ p2, Grids g, float v)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.total) return;
    const int k = cell_of(g, e).b;
    p[k] = v;
    if (p2) p2[k] = v;
}

__global__ void k_gcopy2(const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ x,
                         float *__restrict__ y, Grids g)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.total) return;
    const int k = cell_of(g, e).b;
    x[k] = a[k];
    y[k] = b[k];
}

// keep_measured (src/pipeline.jl:723,768-770,826-829): outliers get their
// last measured value back.
__global__ void k_grestore(float *__restrict__ u, float *__restrict__ v, const float *__restrict__ mu,
                           const float *__restrict__ mv, const uint8_t *__restrict__ out, Grids g)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.total) return;
    const int k = cell_of(g, e).b;
    if (out[k]) {
        u[k] = mu[k];
        v[k] = mv[k];
    }
}

// max(|du|, |dv|) per cell for field_change (src/pipeline.jl:849-866);
// NaN/finite flips give +Inf, NaN in both sweeps gives NaN (dropped on host).
__global__ void k_gchange(const float *__restrict__ u, const float *__restrict__ v, const float *__restrict__ pu,
                          const float *__restrict__ pv, Grids g, float *__restrict__ out)
{
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.total) return;
    const int k = cell_of(g, e).b;
    const float du = fabsf(u[k] - pu[k]), dv = fabsf(v[k] - pv[k]);
    if (isnan(du) || isnan(dv)) {
        const bool same = (isnan(u[k]) == isnan(pu[k])) && (isnan(v[k]) == isnan(pv[k]));
        out[k] = same ? NAN : INFINITY;
        return;
    }
    out[k] = du > dv ? du : dv;
}

dim3 cells_grid(const Grids &g) { return dim3((unsigned)((g.total + 255) / 256)); }

// Two fields set to v in one launch (p2 nullable).
void gfill(hh_ctx *c, float *p, float *p2, const Grids &g, float v)
{
    if (g.total <= 0) return;
    hipLaunchKernelGGL(k_gfill, cells_grid(g), dim3(256), 0, c->stream, p, p2, g, v);
    HH_HIP(hipGetLastError());
}

void gcopy2(hh_ctx *c, const float *a, const float *b, float *x, float *y, const Grids &g)
{
    if (g.total <= 0) return;
    hipLaunchKernelGGL(k_gcopy2, cells_grid(g), dim3(256), 0, c->stream, a, b, x, y, g);
    HH_HIP(hipGetLastError());
}

// ---- layout ---------------------------------------------------------------

// Shelf packing: pairs sorted by width (then height), each shelf is a run of
// atlas columns as wide as its pairs, pairs stacked along the rows up to the
// tallest pair's height.
void pack_layout(Layout &L, const int32_t *nr, const int32_t *nc, int npairs)
{
    L.npairs = npairs;
    L.nr.assign(nr, nr + npairs);
    L.nc.assign(nc, nc + npairs);
    L.roff.assign(npairs, 0);
    L.coff.assign(npairs, 0);
    L.ld = *std::max_element(nr, nr + npairs);
    std::vector<int> order(npairs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return nc[a] != nc[b] ? nc[a] > nc[b] : nr[a] > nr[b];
    });
    long long col = 0;
    int shelf_w = -1, row = 0;
    for (int q : order) {
        if (nc[q] != shelf_w || row + nr[q] > L.ld) {
            if (shelf_w > 0) col += shelf_w;
            shelf_w = nc[q];
            row = 0;
        }
        L.roff[q] = row;
        L.coff[q] = (int)col;
        row += nr[q];
    }
    col += shelf_w;
    L.ncols = col;
    L.coef_off.assign(npairs, 0);
    size_t off = 0;
    for (int q = 0; q < npairs; ++q) {
        L.coef_off[q] = off;
        off += (size_t)(nr[q] + 2) * (nc[q] + 2);
    }
    L.coef_total = off;
}

uint64_t g_layout_gen = 0;

void drop_graph(hh_ctx *c)
{
    Engine &E = c->eng;
    if (E.graph) (void)hipGraphExecDestroy(E.graph);
    E.graph = nullptr;
}

void invalidate_plan(hh_ctx *c)
{
    Engine &E = c->eng;
    drop_graph(c);
    E.graph_failed = false;
    E.graph_replays = 0;
    E.key.clear();
    E.plan.clear();
}

// Knots of a pass grid axis: first centre and spacing, in the float
// arithmetic of the PassGrid coordinates (start + half, difference of the
// first two).
void knots(int r0, int step, float half, int n, float &q0, float &qsp)
{
    const float a = (float)r0 + half;
    q0 = a;
    qsp = n > 1 ? ((float)(r0 + step) + half) - a : 1.f;
}

template <typename T>
T *upload_sync(hh_ctx *c, DBuf<T> &b, const std::vector<T> &h)
{
    T *p = b.ensure(std::max<size_t>(h.size(), 1));
    if (!h.empty()) HH_HIP(hipMemcpyAsync(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->stream));
    return p;
}

// Deformation table of the pairs in `pairs` with the predictor held by the
// grids of `pred` (per pair ny, nx, cell offset and knots).
std::vector<DeformPair> deform_table(hh_ctx *c, const std::vector<int> &pairs, const PassPlan &pred, int &blocks)
{
    Engine &E = c->eng;
    const Layout &L = E.lay;
    const float *cA = E.coefs.p, *cB = E.coefs.p + L.coef_total;
    std::vector<DeformPair> t;
    blocks = 0;
    for (int q : pairs) {
        DeformPair d{};
        d.cA = cA + L.coef_off[q];
        d.cB = cB + L.coef_off[q];
        d.pu = E.pu.p + pred.off[q];
        d.pv = E.pv.p + pred.off[q];
        d.nr = L.nr[q];
        d.nc = L.nc[q];
        d.gny = pred.ny[q];
        d.gnx = pred.nx[q];
        knots(pred.r0, pred.rstep, pred.half, d.gny, d.y0, d.ysp);
        knots(pred.c0, pred.cstep, pred.half, d.gnx, d.x0, d.xsp);
        d.out_off = (long long)L.roff[q] + (long long)L.coff[q] * L.ld;
        d.blk0 = blocks;
        blocks += deform_strips(d.nr) * d.nc;
        t.push_back(d);
    }
    return t;
}

// Tables of pass P over the pairs in `pairs` (jobs, windows with masked
// pixels, segments, counters and the deformation tables); prev is the
// previous pass (predictor source of the first sweep) or null.
void build_tables(hh_ctx *c, PassPlan &P, const PassPlan *prev, const std::vector<int> &pairs, const HostJobs &H,
                  PassTables &T)
{
    std::vector<int32_t> jobs, mjobs, tjobs, ujobs;
    std::vector<GridSeg> segs;
    const int w = P.p.window_size[0], sz = P.p.search_area_size[0], mg = (sz - w) / 2;
    int start = 0;
    for (size_t s = 0; s < pairs.size(); ++s) {
        const int q = pairs[s];
        const int n = P.ny[q] * P.nx[q];
        jobs.insert(jobs.end(), H.jobs.begin() + 4 * (size_t)H.off[q],
                    H.jobs.begin() + 4 * ((size_t)H.off[q] + H.cnt[q]));
        for (int t = 0; t < H.pcnt[q]; ++t) {
            const int32_t *e = &H.part[5 * ((size_t)H.poff[q] + t)];
            const int k = (int)(tjobs.size() / 4);
            tjobs.insert(tjobs.end(), {e[0] + 1, 1, 1 + mg, k * sz + 1 + mg});
            ujobs.insert(ujobs.end(), {e[0] + 1, 1, 1, k * w + 1});
            mjobs.insert(mjobs.end(), {e[1], e[2], e[3], e[4]});
        }
        segs.push_back(GridSeg{start, P.off[q], P.ny[q], P.nx[q], (int)s, {0, 0, 0}});
        start += n;
    }
    T.njobs = (int)(jobs.size() / 4);
    T.npart = (int)(tjobs.size() / 4);
    upload_sync(c, T.jobs, jobs);
    upload_sync(c, T.mjobs, mjobs);
    upload_sync(c, T.tjobs, tjobs);
    upload_sync(c, T.ujobs, ujobs);
    upload_sync(c, T.segs, segs);
    T.grids = Grids{T.segs.p, (int)segs.size(), 0, 0, start, P.total};
    // one grid starting at buffer cell 0: no table lookups in the kernels
    if (segs.size() == 1 && segs[0].off == 0) T.grids = Grids{nullptr, 1, segs[0].ny, segs[0].nx, start, P.total};
    T.count.ensure(std::max<size_t>(segs.size(), 1));
    int b1 = 0, b2 = 0;
    if (c->eng.deforms) {
        if (prev) {
            const std::vector<DeformPair> t = deform_table(c, pairs, *prev, b1);
            upload_sync(c, T.dprev, t);
            T.hprev = t.size() == 1 ? t[0] : DeformPair{};
        }
        const std::vector<DeformPair> t = deform_table(c, pairs, P, b2);
        upload_sync(c, T.dself, t);
        T.hself = t.size() == 1 ? t[0] : DeformPair{};
    }
    T.blocks = std::max(b1, b2);  // every pair has the same pixel count in both tables
    HH_HIP(hipStreamSynchronize(c->stream));
}

void serialize_pass(std::vector<uint8_t> &k, const hh_pass &p)
{
    auto put = [&](const void *v, size_t n) {
        const uint8_t *b = (const uint8_t *)v;
        k.insert(k.end(), b, b + n);
    };
    put(p.window_size, sizeof p.window_size);
    put(p.search_area_size, sizeof p.search_area_size);
    put(p.overlap, sizeof p.overlap);
    put(&p.correlation_method, 4);
    put(&p.padding, 4);
    put(&p.apodization, 4);
    put(&p.subpixel_method, 4);
    put(&p.n_peaks, 4);
    put(&p.peak_finder, 4);
    put(&p.uncertainty, 4);
    put(&p.uod_enable, 4);
    put(&p.uod_threshold, 8);
    put(&p.uod_neighborhood, 4);
    put(&p.replace_outliers, 4);
    put(&p.min_peak_ratio, 8);
    put(&p.max_iterations, 4);
    put(&p.convergence_tol, 8);
    put(&p.n_validators, 4);
    for (int q = 0; q < p.n_validators && q < HH_MAX_VALIDATORS; ++q) {
        put(&p.validators[q].kind, 4);
        put(&p.validators[q].neighborhood, 4);
        put(&p.validators[q].a, 8);
        put(&p.validators[q].b, 8);
    }
}

// ---- autotuner -----------------------------------------------------------
// The launch geometry of a pass's correlation kernel: for the persistent
// MFMA kernel the resident waves per SIMD (1-4, i.e. how many windows each
// wave walks) and the window order (contiguous bands per XCD, so the
// overlapping windows of neighbouring blocks share one L2, or plain launch
// order); for the register FFT kernels the window order. Every candidate is
// timed (best of 3, HIP events on the context stream) on the pass's own job
// list over the pair as loaded, once per (family, configuration, window
// count) per context. Each window's arithmetic does not depend on the
// geometry, so the choice never changes a result.
CorrConfig tune_corr(hh_ctx *c, const PassPlan &P, const CorrConfig &base, bool uq)
{
    Engine &E = c->eng;
    const char *fam = correlate_family(base, uq);
    const std::string f0(fam);
    const bool mfma = f0.rfind("mfma", 0) == 0, reg = f0.rfind("reg", 0) == 0;
    CorrConfig cfg = base;
    const bool families = c->tune.level >= 2;
    if (!c->tune.enabled || !(mfma || reg || families) || P.all.njobs <= 0) return cfg;
    const std::string key = f0 + (uq && !mfma ? "+uq" : "") + " w" + std::to_string(base.w) + " M" +
                            std::to_string(base.M) + " K" + std::to_string(base.K) + " m" +
                            std::to_string(base.method) + " n" + std::to_string(P.all.njobs);
    auto it = c->tune.entries.find(key);
    if (it == c->tune.entries.end()) {
        // candidates (family, waves per SIMD, window order): the form's own
        // family first, then, at level 2, the register FFT and the LDS
        // kernels wherever they differ from it
        struct Cand {
            int family, w, order;
        };
        std::vector<Cand> cand;
        std::vector<int> fams = {0};
        if (families) {
            CorrConfig t = base;
            t.family = 1;
            const std::string f1 = correlate_family(t, uq);
            t.family = 2;
            const std::string f2 = correlate_family(t, uq);
            if (f1 != f0) fams.push_back(1);
            if (f2 != f0 && f2 != f1) fams.push_back(2);
        }
        for (int fm : fams) {
            CorrConfig t = base;
            t.family = fm;
            const std::string f = correlate_family(t, uq);
            for (int order = 0; order < 2; ++order) {
                if (f == "mfma64")
                    for (int w : {2, 1}) cand.push_back({fm, w, order});
                else if (f == "mfma96")
                    cand.push_back({fm, 1, order});  // one wave per SIMD (LDS and registers)
                else if (f == "mfma32+uq") {
                    // the statistics kernel runs beside on the same CUs: few
                    // correlation waves leave it SIMD slots, and only XCD
                    // bands are offered, since back-to-back timing here
                    // favours launch order (about 5 %) where the schedule
                    // measures it 2 % slower (0.571 against 0.563 ms on the
                    // 4 MP pair)
                    if (order == 0)
                        for (int w : {1, 2, 3, 4}) cand.push_back({fm, w, 0});
                } else if (f.rfind("mfma", 0) == 0)
                    for (int w : {3, 2, 4, 1}) cand.push_back({fm, w, order});
                else if (f.rfind("reg", 0) == 0)
                    cand.push_back({fm, 0, order});
                else if (order == 0)
                    cand.push_back({fm, 0, 0});  // LDS / DFT kernels: one block per window, no geometry
            }
        }
        CorrOut out{};
        out.u = E.u.p;
        out.v = E.v.p;
        out.pr = E.pr.p;
        out.cm = E.cm.p;
        out.alt_u = base.n_peaks > 1 ? E.altu.p : nullptr;
        out.alt_v = base.n_peaks > 1 ? E.altv.p : nullptr;
        out.ny = P.total;
        out.nx = 1;
        if (uq) {
            out.unc_u = E.uu.p;
            out.unc_v = E.uv.p;
        }
        const ConfigBufs &cb = get_config(c, P.p);
        hipEvent_t a = get_event(c), b = get_event(c);
        TuneEntry best;
        // candidates are timed round-robin, each minimum over rounds 1-4
        // (round 0 warms caches): the clocks still ramp while tuning runs,
        // and timing each candidate's launches back to back favoured the
        // later candidates by 3-5 %
        std::vector<float> ms_min(cand.size(), 1e30f);
        for (int rep = 0; rep < 5; ++rep)
            for (size_t i = 0; i < cand.size(); ++i) {
                CorrConfig t = base;
                t.family = cand[i].family;
                t.waves_per_simd = cand[i].w;
                t.linear_order = cand[i].order;
                HH_HIP(hipEventRecord(a, c->stream));
                const bool fused = launch_correlate(t, E.lay.A, E.lay.B, E.lay.ld, 0, P.all.jobs.p, P.all.njobs, out,
                                                    c->stream, &c->side);
                // a family that does not fuse the statistics is timed with
                // the standalone launch the engine then makes
                if (uq && !fused)
                    launch_uq(base.w, cb.apod.p, E.lay.A, E.lay.B, E.lay.ld, P.all.jobs.p, P.all.njobs, P.total,
                              E.uu.p, E.uv.p, nullptr, false, c->stream);
                HH_HIP(hipEventRecord(b, c->stream));
                HH_HIP(hipEventSynchronize(b));
                float ms = 0.f;
                HH_HIP(hipEventElapsedTime(&ms, a, b));
                if (rep > 0) ms_min[i] = std::min(ms_min[i], ms);
            }
        float best_score = 1e30f;
        for (size_t i = 0; i < cand.size(); ++i) {
            const Cand &cd = cand[i];
            CorrConfig t = base;
            t.family = cd.family;
            t.waves_per_simd = cd.w;
            t.linear_order = cd.order;
            const std::string f = correlate_family(t, uq);
            std::string name = f.rfind("mfma", 0) == 0 ? "waves_per_simd=" + std::to_string(cd.w) + " " : "";
            if (f.rfind("mfma", 0) == 0 || f.rfind("reg", 0) == 0)
                name += cd.order ? "order=linear" : "order=xcd_bands";
            if (families) name = "family=" + f + (name.empty() ? "" : " " + name);
            best.ms.push_back({name, ms_min[i]});
            // a later candidate must win by more than timing noise (2 %);
            // the first listed is the form's own family in XCD bands. Launch
            // order deals neighbouring windows to different XCDs, so each L2
            // fetches most of the pair (2x the HBM bytes of XCD bands for the
            // 16-px matrix-core kernel, 5x for the statistics kernel), and
            // back-to-back launches here time it more favourably than the
            // schedule does: it must win by 5 % more
            const float score = ms_min[i] * (cd.order ? 1.05f : 1.f);
            if (score < 0.98f * best_score) {
                best_score = score;
                best.family = cd.family;
                best.waves_per_simd = cd.w;
                best.linear_order = cd.order;
            }
        }
        c->event_pool.push_back(a);
        c->event_pool.push_back(b);
        it = c->tune.entries.emplace(key, best).first;
    }
    cfg.family = it->second.family;
    cfg.waves_per_simd = it->second.waves_per_simd;
    cfg.linear_order = it->second.linear_order;
    return cfg;
}

// Masked-pixel counts of rectangles from a summed-area table of the mask
// (mod 2^32: every count asked for is below 2^32, so the wrapped prefix
// sums still difference exactly).
struct MaskCounts {
    std::vector<uint32_t> S;
    int h = 0;
    void build(const std::vector<uint8_t> &m, int nr, int nc)
    {
        h = nr + 1;
        S.assign((size_t)(nr + 1) * (nc + 1), 0u);
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nr; ++i)
                S[(i + 1) + (size_t)(j + 1) * h] = (uint32_t)m[i + (size_t)j * nr] + S[i + (size_t)(j + 1) * h] +
                                                   S[(i + 1) + (size_t)j * h] - S[i + (size_t)j * h];
    }
    // rows r0 .. r0 + nr - 1, columns c0 .. c0 + nc - 1 (0-based)
    uint32_t rect(int r0, int c0, int nr, int nc) const
    {
        const int r1 = r0 + nr, c1 = c0 + nc;
        return S[r1 + (size_t)c1 * h] - S[r0 + (size_t)c1 * h] - S[r1 + (size_t)c0 * h] + S[r0 + (size_t)c0 * h];
    }
};

void plan(hh_ctx *c, const hh_pass *passes, int np, int smoothing)
{
    Engine &E = c->eng;
    const Layout &L = E.lay;
    invalidate_plan(c);
    E.smoothing = smoothing;
    const bool masked = !E.hmask.empty();
    MaskCounts mc;
    if (masked) {
        for (int q = 0; q < L.npairs; ++q)
            if (L.nr[q] != E.mnr || L.nc[q] != E.mnc)
                throw Error(HH_EDIM, "mask must have the same size as the images, got (" + std::to_string(E.mnr) +
                                         ", " + std::to_string(E.mnc) + ") for images of (" +
                                         std::to_string(L.nr[q]) + ", " + std::to_string(L.nc[q]) + ")");
        mc.build(E.hmask, E.mnr, E.mnc);
    }
    E.maxtile = 0;
    E.deforms = np > 1;
    for (int k = 0; k < np; ++k) E.deforms |= passes[k].max_iterations > 1;
    std::vector<int> all(L.npairs);
    std::iota(all.begin(), all.end(), 0);
    E.maxcells = 0;
    E.maxalt = 0;
    E.hjobs.assign(np, {});
    for (int k = 0; k < np; ++k) {
        auto P = std::make_unique<PassPlan>();
        P->p = passes[k];
        P->cfg = make_corr(c, passes[k]);
        P->cfg_uq = P->cfg;
        const hh_pass &p = passes[k];
        const int w = p.window_size[0];
        // first window origin: past the search margin (pass_grid,
        // src/pipeline.jl:529-538)
        P->r0 = 1 + (p.search_area_size[0] - w) / 2;
        P->rstep = w - p.overlap[0];
        P->c0 = 1 + (p.search_area_size[1] - p.window_size[1]) / 2;
        P->cstep = p.window_size[1] - p.overlap[1];
        P->half = (float)((w - 1) / 2.0);
        P->fixed = p.max_iterations == 1 || p.convergence_tol == 0;
        HostJobs &H = E.hjobs[k];
        // per window: 0 clean, 1 with masked pixels, 2 grid-masked (no job)
        std::vector<uint8_t> kind;
        for (int q = 0; q < L.npairs; ++q) {
            PassGrid g = make_grid(p, L.nr[q], L.nc[q]);
            const size_t ng = (size_t)g.ny * g.nx;
            if (ng == 0) throw Error(HH_EINVAL, "the pass grid is empty");
            if ((long long)P->total + (long long)ng > (1LL << 30)) throw Error(HH_EINVAL, "too many vectors in one batch");
            if (masked && kind.empty()) {
                // pass_grid (src/pipeline.jl:542-549): the masked fractions of
                // the window and of its search area
                const int wr = p.window_size[0], wc = p.window_size[1];
                const int sr = p.search_area_size[0], sc = p.search_area_size[1];
                const int mr = (sr - wr) / 2, mcm = (sc - wc) / 2;
                kind.assign(ng, 0);
                P->gm.assign(ng, 0);
                for (size_t j = 0; j < ng; ++j) {
                    const int r0 = g.jobs[4 * j + 2] - 1, c0 = g.jobs[4 * j + 3] - 1;
                    const uint32_t nm = mc.rect(r0, c0, wr, wc);
                    const uint32_t ns = mc.rect(r0 - mr, c0 - mcm, sr, sc);
                    const bool gm = (double)nm / (double)(wr * wc) >= E.mthr || (double)ns / (double)(sr * sc) >= E.mthr;
                    P->gm[j] = gm;
                    P->anyg |= gm;
                    kind[j] = gm ? 2 : ((nm > 0 || ns > 0) ? 1 : 0);
                }
            }
            P->ny.push_back(g.ny);
            P->nx.push_back(g.nx);
            P->off.push_back(P->total);
            H.off.push_back((int)(H.jobs.size() / 4));
            H.poff.push_back((int)(H.part.size() / 5));
            for (size_t j = 0; j < ng; ++j) {
                const int kj = kind.empty() ? 0 : kind[j];
                const int rs = g.jobs[4 * j + 2], cs = g.jobs[4 * j + 3];
                if (kj == 0) {
                    H.jobs.push_back(P->total + (int32_t)j + 1);  // cell of column-major job order
                    H.jobs.push_back(1);
                    H.jobs.push_back(rs + L.roff[q]);
                    H.jobs.push_back(cs + L.coff[q]);
                } else if (kj == 1) {
                    H.part.insert(H.part.end(), {P->total + (int32_t)j, rs - 1 + L.roff[q], cs - 1 + L.coff[q], rs - 1,
                                                 cs - 1});
                }
            }
            H.cnt.push_back((int)(H.jobs.size() / 4) - H.off.back());
            H.pcnt.push_back((int)(H.part.size() / 5) - H.poff.back());
            P->total += (int)ng;
        }
        {
            const size_t npart = H.part.size() / 5, sz = (size_t)p.search_area_size[0];
            if (npart * sz >= (size_t)1 << 31) throw Error(HH_EINVAL, "too many windows with masked pixels");
            E.maxtile = std::max(E.maxtile, npart * sz * sz);
        }
        E.maxcells = std::max(E.maxcells, (size_t)P->total);
        E.maxalt = std::max(E.maxalt, (size_t)P->total * std::max(p.n_peaks - 1, 0));
        correlate_prepare(P->cfg, c->stream);
        E.plan.push_back(std::move(P));
    }
    // buffers first (the tables hold pointers into them)
    for (DBuf<float> *b : {&E.u, &E.v, &E.pr, &E.cm, &E.uu, &E.uv, &E.u0, &E.v0, &E.su, &E.sv, &E.mu, &E.mv,
                           &E.prevu, &E.prevv, &E.pu, &E.pv, &E.ch})
        b->ensure(std::max<size_t>(E.maxcells, 1));
    E.zero.ensure(std::max<size_t>(E.maxcells, 1));
    HH_HIP(hipMemsetAsync(E.zero.p, 0, sizeof(float) * E.zero.n, c->stream));
    E.altu.ensure(std::max<size_t>(E.maxalt, 1));
    E.altv.ensure(std::max<size_t>(E.maxalt, 1));
    E.out.ensure(std::max<size_t>(E.maxcells, 1));
    E.sflag.ensure(std::max<size_t>(E.maxcells, 1));
    if (E.maxtile) E.tiles.ensure(2 * E.maxtile);
    E.mcount.ensure((size_t)L.npairs);
    for (int k = 0; k < np; ++k) {
        PassPlan &P = *E.plan[k];
        if (!P.anyg) continue;
        std::vector<uint8_t> all((size_t)P.total);
        for (int q = 0; q < L.npairs; ++q) std::copy(P.gm.begin(), P.gm.end(), all.begin() + P.off[q]);
        upload_sync(c, P.gmask, all);
        HH_HIP(hipStreamSynchronize(c->stream));  // `all` ends with this iteration
    }
    if (E.deforms) {
        E.coefs.ensure(2 * L.coef_total);
        E.warp.ensure(2 * (size_t)L.ld * (size_t)L.ncols);
        size_t t = 0;
        for (int q = 0; q < L.npairs; ++q) t = std::max(t, (size_t)2 * (L.nr[q] + 2) * L.nc[q]);
        E.ptmp.ensure(t);
    }
    for (int k = 0; k < np; ++k)
        build_tables(c, *E.plan[k], k > 0 ? E.plan[k - 1].get() : nullptr, all, E.hjobs[k], E.plan[k]->all);
    for (int k = 0; k < np; ++k) {
        PassPlan &P = *E.plan[k];
        const bool unc = P.p.uncertainty && k == np - 1 && P.fixed;
        const bool plain = P.p.max_iterations > 1 || !unc;
        if (plain) P.cfg = tune_corr(c, P, P.cfg, false);
        P.cfg_uq = unc ? tune_corr(c, P, P.cfg_uq, true) : P.cfg;
    }
}

// build_predictor (src/pipeline.jl:454-464) from the field of pass P: its
// masked cells filled from the unmasked ones (into the validation scratch),
// then smoothed.
void predictor(hh_ctx *c, const Grids &g, const PassPlan &P)
{
    Engine &E = c->eng;
    const float *u = E.u.p, *v = E.v.p;
    if (P.anyg) {
        launch_fill_masked(g, E.u.p, E.v.p, P.gmask.p, E.mcount.p, E.su.p, E.sv.p, c->stream);
        u = E.su.p;
        v = E.sv.p;
    }
    if (E.smoothing) {
        launch_smooth2(g, u, v, E.pu.p, E.pv.p, c->stream);
    } else {
        gcopy2(c, u, v, E.pu.p, E.pv.p, g);
    }
}

// The launch sequence of the whole schedule. With `dynamic` false every
// pass has a fixed sweep count and nothing here touches the host (the
// sequence may be captured into a graph).
void record(hh_ctx *c, bool dynamic)
{
    Engine &E = c->eng;
    const Layout &L = E.lay;
    const int np = (int)E.plan.size();
    hipStream_t s = c->stream;
    float *cA = E.coefs.p, *cB = E.coefs.p + L.coef_total;
    float *wA = E.warp.p, *wB = E.warp.p + (size_t)L.ld * L.ncols;
    // The prefilter only feeds the first deformation: unless it is being
    // timed, it runs on the side stream beside the first pass (which reads
    // the raw pair) and is joined right before that deformation.
    bool pre_pending = false;
    if (E.deforms) {
        auto prefilter = [&](hipStream_t st) {
            for (int q = 0; q < L.npairs; ++q) {
                const size_t o = (size_t)L.roff[q] + (size_t)L.coff[q] * L.ld;
                launch_prefilter_pair(L.A + o, L.B + o, L.ld, L.nr[q], L.nc[q], cA + L.coef_off[q],
                                      cB + L.coef_off[q], E.ptmp.p, st);
            }
        };
        const bool timing_pre = c->timing && (c->timing_family.empty() || c->timing_family == "prefilter");
        if (timing_pre) {
            timed(c, "prefilter", 0, [&] { prefilter(s); });
        } else {
            HH_HIP(hipEventRecord(c->side.fork, s));
            HH_HIP(hipStreamWaitEvent(c->side.stream, c->side.fork, 0));
            prefilter(c->side.stream);
            HH_HIP(hipEventRecord(c->pre_join, c->side.stream));
            pre_pending = true;
        }
    }
    // the statistics of the final sweep run on the side stream beside the
    // correlation, the validation and the rest of the schedule (they only
    // read that sweep's warp, which nothing overwrites afterwards), joined
    // at the end
    SideStream side = c->side;
    side.defer_join = true;
    bool uq_pending = false;
    bool have_pred = false, pred_self = false;
    bool pred_fused = false;  // the previous pass's last replacement launch smoothed the predictor
    for (int k = 0; k < np; ++k) {
        PassPlan &P = *E.plan[k];
        const hh_pass &p = P.p;
        const bool force_replace = k < np - 1;
        if (k > 0) {
            if (!pred_fused) predictor(c, E.plan[k - 1]->all.grids, *E.plan[k - 1]);
            have_pred = true;
            pred_self = false;
        }
        pred_fused = false;
        const int maxiter = p.max_iterations;
        const bool unc = p.uncertainty && !force_replace;
        const bool keep_measured = maxiter > 1 && !(force_replace || p.replace_outliers);
        const int nalt = p.n_peaks - 1;
        // uncertainty grids: NaN unless this pass computes them; a fused
        // computation writes every cell with a window, so only the final
        // pass's grids need the fill when the statistics are not fused (or
        // not requested) or cells are masked
        const bool last = k == np - 1;
        if (last && (!(unc && P.fixed) || P.anyg)) gfill(c, E.uu.p, E.uv.p, P.all.grids, NAN);
        float *tA = E.tiles.p, *tB = E.tiles.p + E.maxtile;
        PassTables *T = &P.all;
        std::vector<int> active(L.npairs);
        std::iota(active.begin(), active.end(), 0);
        bool uq_fused = false;
        const float *srcA = L.A, *srcB = L.B;
        for (int it = 1; it <= maxiter; ++it) {
            const DeformPair *dt = pred_self ? T->dself.p : T->dprev.p;
            if (have_pred) {
                DeformTab tab{};
                tab.pairs = T->grids.nseg == 1 ? nullptr : dt;  // one pair: passed by value
                tab.npairs = T->grids.nseg;
                tab.nblocks = T->blocks;
                tab.one = pred_self ? T->hself : T->hprev;
                if (pre_pending) {
                    HH_HIP(hipStreamWaitEvent(s, c->pre_join, 0));
                    pre_pending = false;
                }
                // the pass's node values ride along as extra blocks
                tab.nv = node_values_job(T->grids, dt, P.r0, P.rstep, P.c0, P.cstep, P.half, E.u0.p, E.v0.p);
                timed(c, "deform", 0, [&] { launch_deform_batch(tab, L.ld, wA, wB, s); });
                srcA = wA;
                srcB = wB;
            } else {
                srcA = L.A;
                srcB = L.B;
            }
            CorrOut out{};
            out.u = E.u.p;
            out.v = E.v.p;
            out.pr = E.pr.p;
            out.cm = E.cm.p;
            out.alt_u = nalt > 0 ? E.altu.p : nullptr;
            out.alt_v = nalt > 0 ? E.altv.p : nullptr;
            // the first sweep has no predictor: zeros (a buffer nothing
            // writes, so no fill per run)
            out.u0 = have_pred ? E.u0.p : E.zero.p;
            out.v0 = have_pred ? E.v0.p : E.zero.p;
            out.ny = P.total;  // jobs carry (cell + 1, 1)
            out.nx = 1;
            // the uncertainty of an iterating pass is taken once on the final
            // sweep's warp; with a fixed sweep count that sweep is known in
            // advance and the statistics are fused into its correlation
            if (unc && it == maxiter && P.fixed) {
                out.unc_u = E.uu.p;
                out.unc_v = E.uv.p;
            }
            const CorrConfig &cc = out.unc_u ? P.cfg_uq : P.cfg;
            bool fused = true;
            timed(c, out.unc_u ? "correlate_uq" : "correlate", P.cfg.w, [&] {
                // the deferred join is only safe for the final sweep of the
                // final pass (no later deformation overwrites its warp)
                const bool timing_uq = c->timing && (c->timing_family.empty() || c->timing_family == "correlate_uq");
                const bool defer = out.unc_u && last && !timing_uq;
                fused = launch_correlate(cc, srcA, srcB, L.ld, 0, T->jobs.p, T->njobs, out, s,
                                         defer ? &side : &c->side);
                uq_pending = uq_pending || defer;
            });
            if (T->npart > 0) {
                // windows with masked pixels, as centred tiles
                timed(c, "mask_tiles", P.cfg.w, [&] {
                    launch_mask_tiles(srcA, srcB, L.ld, E.dmask.p, E.mnr, T->mjobs.p, T->npart, P.cfg.w, P.cfg.s, tA,
                                      tB, s);
                });
                timed(c, "correlate_masked", P.cfg.w, [&] {
                    const bool f = launch_correlate(cc, tA, tB, P.cfg.s, 0, T->tjobs.p, T->npart, out, s);
                    fused = fused && f;
                });
            }
            uq_fused = fused && out.unc_u != nullptr;
            if (P.anyg) launch_mask_nan(T->grids, P.gmask.p, E.u.p, E.v.p, E.pr.p, E.cm.p, s);
            if (keep_measured) gcopy2(c, E.u.p, E.v.p, E.mu.p, E.mv.p, T->grids);
            ValidateArgs va{};
            va.grids = T->grids;
            va.pass = &p;
            va.force_replace = force_replace || maxiter > 1;
            va.u = E.u.p;
            va.v = E.v.p;
            va.pr = E.pr.p;
            va.cm = E.cm.p;
            va.alt_u = out.alt_u;
            va.alt_v = out.alt_v;
            va.outliers = E.out.p;
            va.su = E.su.p;
            va.sv = E.sv.p;
            va.sflag = E.sflag.p;
            va.count = T->count.p;
            va.mask = P.anyg ? P.gmask.p : nullptr;
            // the next sweep's (or the next pass's) predictor comes out of
            // the replacement launch (fixed sweep counts: no pair drops out
            // in between, and the last sweep is known in advance)
            const bool fuse_pred = (it < maxiter || !last) && P.fixed && E.smoothing && !P.anyg;
            va.smooth_u = fuse_pred ? E.pu.p : nullptr;
            va.smooth_v = fuse_pred ? E.pv.p : nullptr;
            bool smoothed = false;
            timed(c, "validate", P.cfg.w, [&] { smoothed = launch_validate(va, s); });
            if (it == maxiter) {
                pred_fused = smoothed;
                break;
            }
            if (!P.fixed) {
                if (!dynamic) throw Error(HH_ESTATE, "internal: iterating pass without a fixed sweep count");
                if (it > 1) {
                    // field_change per pair; converged pairs stop here
                    hipLaunchKernelGGL(k_gchange, cells_grid(T->grids), dim3(256), 0, s, E.u.p, E.v.p, E.prevu.p,
                                       E.prevv.p, T->grids, E.ch.p);
                    HH_HIP(hipGetLastError());
                    std::vector<float> h(P.total);
                    HH_HIP(hipMemcpyAsync(h.data(), E.ch.p, sizeof(float) * P.total, hipMemcpyDeviceToHost, s));
                    HH_HIP(hipStreamSynchronize(s));
                    std::vector<int> still;
                    for (int q : active) {
                        std::vector<float> buf;
                        buf.reserve((size_t)P.ny[q] * P.nx[q]);
                        for (int j = 0; j < P.ny[q] * P.nx[q]; ++j) {
                            const float x = h[P.off[q] + j];
                            if (!std::isnan(x)) buf.push_back(x);
                        }
                        const float change = buf.empty() ? 0.f : quantile95(buf);
                        if (!(change < (float)p.convergence_tol)) still.push_back(q);
                    }
                    if (still.empty()) break;
                    if (still.size() != active.size()) {
                        active = still;
                        build_tables(c, P, nullptr, active, E.hjobs[k], E.sub);
                        T = &E.sub;
                    }
                }
                gcopy2(c, E.u.p, E.v.p, E.prevu.p, E.prevv.p, T->grids);
            }
            if (!smoothed) predictor(c, T->grids, P);
            have_pred = true;
            pred_self = true;
        }
        if (unc && !uq_fused) {
            const ConfigBufs &cb = get_config(c, p);
            timed(c, "uq", P.cfg.w, [&] {
                launch_uq(P.cfg.w, cb.apod.p, srcA, srcB, L.ld, P.all.jobs.p, P.all.njobs, P.total, E.uu.p, E.uv.p,
                          nullptr, false, s);
                if (P.all.npart > 0) {
                    // every pair's masked windows from its final warp (the
                    // tiles of the last sweep may cover a subset of pairs)
                    // window tiles of both frames: the uncertainty's load
                    launch_mask_tiles(srcA, srcB, L.ld, E.dmask.p, E.mnr, P.all.mjobs.p, P.all.npart, P.cfg.w,
                                      P.cfg.w, tA, tB, s);
                    launch_uq(P.cfg.w, cb.apod.p, tA, tB, P.cfg.w, P.all.ujobs.p, P.all.npart, P.total, E.uu.p,
                              E.uv.p, nullptr, false, s);
                }
            });
        }
        if (keep_measured)
            hipLaunchKernelGGL(k_grestore, cells_grid(P.all.grids), dim3(256), 0, s, E.u.p, E.v.p, E.mu.p, E.mv.p,
                               E.out.p, P.all.grids);
    }
    if (pre_pending) HH_HIP(hipStreamWaitEvent(s, c->pre_join, 0));
    if (uq_pending) HH_HIP(hipStreamWaitEvent(s, c->side.join, 0));
    HH_HIP(hipGetLastError());
}

}  // namespace

void engine_single_layout(hh_ctx *c)
{
    Layout &L = c->eng.lay;
    if (L.npairs == 1 && L.A == c->A && L.B == c->B && L.nr[0] == c->nr && L.nc[0] == c->nc && L.ld == c->nr)
        return;
    const int32_t nr = c->nr, nc = c->nc;
    pack_layout(L, &nr, &nc, 1);
    L.A = c->A;
    L.B = c->B;
    L.gen = ++g_layout_gen;
    invalidate_plan(c);
    c->eng.have_result = false;
}

void engine_set_pairs(hh_ctx *c, const float *const *A, const float *const *B, const int32_t *nr, const int32_t *nc,
                      int npairs)
{
    if (npairs < 1) throw Error(HH_EINVAL, "pairs must not be empty");
    if (!A || !B || !nr || !nc) throw Error(HH_EINVAL, "null pair list");
    for (int q = 0; q < npairs; ++q) {
        if (!A[q] || !B[q]) throw Error(HH_EINVAL, "null image in the pair list");
        if (nr[q] < 2 || nc[q] < 2) throw Error(HH_EDIM, "images must be at least 2 x 2");
    }
    Engine &E = c->eng;
    E.have_result = false;
    Layout &L = E.lay;
    // the same pair sizes in the same order as the resident batch (a stream
    // of equally shaped batches): the frames go into the same atlas places
    // and the plan and captured graph stay valid
    const bool same = L.A && L.npairs == npairs && std::equal(nr, nr + npairs, L.nr.begin()) &&
                      std::equal(nc, nc + npairs, L.nc.begin()) && L.A == E.atlas.p;
    if (!same) {
        invalidate_plan(c);
        pack_layout(L, nr, nc, npairs);
        if ((double)L.ld * (double)L.ncols > 4.0e9) throw Error(HH_EINVAL, "batch atlas exceeds 2^32 pixels");
    }
    const size_t per = (size_t)L.ld * (size_t)L.ncols;
    float *at = E.atlas.ensure(2 * per);
    // each frame crosses PCIe as one contiguous copy: straight into the
    // atlas when its column length is the atlas's, else into a staging
    // buffer and from there into its atlas place by a device-side 2D copy
    // (a pitched copy from pageable host memory goes column by column)
    size_t maxpx = 0;
    for (int q = 0; q < npairs; ++q)
        if (nr[q] != L.ld) maxpx = std::max(maxpx, (size_t)nr[q] * nc[q]);
    float *stage = maxpx ? E.stage.ensure(maxpx) : nullptr;
    for (int q = 0; q < npairs; ++q) {
        float *dA = at + (size_t)L.roff[q] + (size_t)L.coff[q] * L.ld;
        const size_t npx = (size_t)nr[q] * nc[q];
        for (int f = 0; f < 2; ++f) {
            const float *src = f ? B[q] : A[q];
            float *dst = dA + (f ? per : 0);
            if (nr[q] == L.ld) {
                HH_HIP(hipMemcpyAsync(dst, src, sizeof(float) * npx, hipMemcpyHostToDevice, c->stream));
            } else {
                HH_HIP(hipMemcpyAsync(stage, src, sizeof(float) * npx, hipMemcpyHostToDevice, c->stream));
                HH_HIP(hipMemcpy2DAsync(dst, sizeof(float) * L.ld, stage, sizeof(float) * nr[q], sizeof(float) * nr[q],
                                        nc[q], hipMemcpyDeviceToDevice, c->stream));
            }
        }
    }
    HH_HIP(hipStreamSynchronize(c->stream));
    L.A = at;
    L.B = at + per;
    if (!same) L.gen = ++g_layout_gen;
    // the single-pair state no longer describes the context's images
    c->A = c->B = nullptr;
    c->coefs_valid = c->warp_valid = c->have_result = false;
}

void engine_run(hh_ctx *c, const hh_pass *passes, int np, int smoothing)
{
    Engine &E = c->eng;
    if (E.lay.npairs < 1 || !E.lay.A) throw Error(HH_ESTATE, "no image pair: call hh_set_pair or hh_set_pairs first");
    if (np < 1) throw Error(HH_EINVAL, "at least one pass is required");
    for (int k = 0; k < np; ++k) check_pass_device(passes[k], true);
    std::vector<uint8_t> key;
    auto put = [&](const void *v, size_t n) {
        const uint8_t *b = (const uint8_t *)v;
        key.insert(key.end(), b, b + n);
    };
    put(&E.lay.gen, sizeof E.lay.gen);
    put(&E.mgen, sizeof E.mgen);
    put(&c->form, sizeof c->form);
    put(&smoothing, sizeof smoothing);
    put(&np, sizeof np);
    for (int k = 0; k < np; ++k) serialize_pass(key, passes[k]);
    if (key != E.key) {
        E.have_result = false;
        plan(c, passes, np, smoothing);
        E.key = key;
    }
    bool dynamic = false;
    for (auto &P : E.plan) dynamic |= !P->fixed;
    E.have_result = false;
    c->have_result = false;
    // per-kernel event timing (hh_set_timing) runs the sequence eagerly: event
    // records captured into a graph cannot be read back
    if (!dynamic && !c->timing && E.use_graphs && !E.graph_failed) {
        if (!E.graph) {
            hipGraph_t g = nullptr;
            HH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            bool ok = true;
            std::string why;
            try {
                record(c, false);
            } catch (const Error &e) {
                ok = false;
                why = e.msg;
            }
            const hipError_t ee = hipStreamEndCapture(c->stream, &g);
            hipGraphExec_t ex = nullptr;
            if (ok && ee == hipSuccess && g && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess) {
                E.graph = ex;
            } else {
                E.graph_failed = true;
                (void)hipGetLastError();
            }
            if (g) (void)hipGraphDestroy(g);
            if (!ok && why.find("internal") != std::string::npos) throw Error(HH_ESTATE, why);
        }
        if (E.graph) {
            HH_HIP(hipGraphLaunch(E.graph, c->stream));
            ++E.graph_replays;
            E.have_result = true;
            return;
        }
    }
    record(c, dynamic);
    E.have_result = true;
}

void engine_fetch(hh_ctx *c, int q, hh_field *r)
{
    Engine &E = c->eng;
    if (!E.have_result || E.plan.empty()) throw Error(HH_ESTATE, "no result: call hh_run_piv first");
    if (q < 0 || q >= E.lay.npairs) throw Error(HH_EINVAL, "pair index out of range");
    const PassPlan &P = *E.plan.back();
    const int ny = P.ny[q], nx = P.nx[q];
    if (r->ny != ny || r->nx != nx)
        throw Error(HH_EDIM, "result grid is " + std::to_string(ny) + "x" + std::to_string(nx) + ", got " +
                                 std::to_string(r->ny) + "x" + std::to_string(r->nx));
    const size_t ng = (size_t)ny * nx, o = (size_t)P.off[q];
    auto d2h = [&](void *dst, const void *src, size_t bytes) {
        if (dst) HH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    };
    d2h(r->u, E.u.p + o, sizeof(float) * ng);
    d2h(r->v, E.v.p + o, sizeof(float) * ng);
    d2h(r->peak_ratio, E.pr.p + o, sizeof(float) * ng);
    d2h(r->correlation_moment, E.cm.p + o, sizeof(float) * ng);
    d2h(r->uncertainty_u, E.uu.p + o, sizeof(float) * ng);
    d2h(r->uncertainty_v, E.uv.p + o, sizeof(float) * ng);
    d2h(r->outliers, E.out.p + o, ng);
    HH_HIP(hipStreamSynchronize(c->stream));
    if (r->x)
        for (int j = 0; j < nx; ++j) r->x[j] = (float)(P.c0 + j * P.cstep) + P.half;
    if (r->y)
        for (int i = 0; i < ny; ++i) r->y[i] = (float)(P.r0 + i * P.rstep) + P.half;
    if (r->mask) {
        if (P.anyg)
            std::memcpy(r->mask, P.gm.data(), ng);
        else
            std::memset(r->mask, 0, ng);
    }
}

void engine_set_mask(hh_ctx *c, const uint8_t *mask, int nr, int nc, double threshold)
{
    Engine &E = c->eng;
    std::vector<uint8_t> m;
    if (mask) {
        if (nr < 1 || nc < 1) throw Error(HH_EDIM, "mask must not be empty");
        if (!(threshold > 0 && threshold <= 1)) {
            char buf[96];
            std::snprintf(buf, sizeof buf, "mask_threshold must be in (0, 1], got %g", threshold);
            throw Error(HH_EINVAL, buf);
        }
        m.resize((size_t)nr * nc);
        bool any = false;
        for (size_t k = 0; k < m.size(); ++k) any |= (m[k] = mask[k] != 0);
        if (!any) m.clear();  // nothing masked: the unmasked schedule gives the same result
    }
    if (m.empty()) {
        if (E.hmask.empty()) return;
        E.hmask.clear();
        E.mnr = E.mnc = 0;
    } else {
        if (m == E.hmask && nr == E.mnr && nc == E.mnc && threshold == E.mthr) return;
        uint8_t *d = E.dmask.ensure(m.size());
        HH_HIP(hipMemcpyAsync(d, m.data(), m.size(), hipMemcpyHostToDevice, c->stream));
        HH_HIP(hipStreamSynchronize(c->stream));
        E.hmask = std::move(m);
        E.mnr = nr;
        E.mnc = nc;
        E.mthr = threshold;
    }
    ++E.mgen;
    E.have_result = false;
}

void engine_release(hh_ctx *c) { invalidate_plan(c); }

std::string engine_tune_report(hh_ctx *c)
{
    std::string j = "{";
    bool first = true;
    for (auto &[key, e] : c->tune.entries) {
        j += (first ? "\"" : ", \"") + key + "\": {\"choice\": {\"waves_per_simd\": " +
             std::to_string(e.waves_per_simd) + ", \"linear_order\": " + std::to_string(e.linear_order) +
             ", \"family\": " + std::to_string(e.family) + "}, \"ms\": {";
        for (size_t q = 0; q < e.ms.size(); ++q) {
            char buf[64];
            std::snprintf(buf, sizeof buf, "%.4f", e.ms[q].second);
            j += (q ? ", \"" : "\"") + e.ms[q].first + "\": " + buf;
        }
        j += "}}";
        first = false;
    }
    return j + "}";
}

void engine_set_graphs(hh_ctx *c, bool enable)
{
    c->eng.use_graphs = enable;
    if (!enable) drop_graph(c);
}

}  // namespace host
}  // namespace hh
