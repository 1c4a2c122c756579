// k_validate.hip — vector validation, peak substitution, local-median
// replacement and predictor smoothing on the vector grid.
//
// validate_and_replace! (src/pipeline.jl:617-645):
//   1. universal_outlier_detection (src/quality.jl:244-293): neighbours of a
//      (2n+1)^2 box in column-major order, centre and excluded cells left
//      out; medians (Statistics.median!: NaN if any NaN, mean of the middle
//      pair as a/2 + b/2) and residuals in T; the normalized residual
//      T / (Float64(median) + epsilon) in Float64, flagged when > threshold;
//   2. PeakRatioValidator when min_peak_ratio > 1, then the extra validators
//      (src/quality.jl:479-503);
//   3. substitute_alternatives! (src/quality.jl:307-351): Float64 median and
//      MAD of the valid neighbours of a snapshot; first alternative within
//      threshold wins and is unflagged;
//   4. replace_vectors! (src/quality.jl:85-114): Float64 median of the valid
//      cells of growing squares until >= 3, from a snapshot, stored as T.
// smooth_field (src/quality.jl:123-136): separable [1 2 1]/4 with
// replicated edges, in T, same operation order.
// All kernels read a snapshot and write a separate buffer, so results do not
// depend on thread order. Floating-point contraction is off: flags are
// bit-exact functions of the inputs.

#include "hh_internal.h"

#include <algorithm>
#include <cfloat>

namespace hh {

namespace {

#pragma clang fp contract(off)

// Sort of N values in registers by Batcher's merge exchange (Knuth, TAOCP
// vol. 3, 5.2.2 algorithm M): ceil(log2 N) merge rounds of comparators
// (i, i + d) on the indices with (i & p) == r, 127 comparators for N = 24
// against 288 for transposition, each one min and one max. The loops have
// constant bounds and unroll into straight-line code; outputs nobody reads
// (the upper half, for a median) are pruned by the compiler. Values are
// ordered as by <; callers detect NaN separately, so the NaN rules of
// min / max never reach a result.
template <int N>
constexpr int ceil_log2()
{
    int t = 0;
    while ((1 << t) < N) ++t;
    return t;
}

__device__ __forceinline__ float vmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double vmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double vmax(double a, double b) { return fmax(a, b); }

template <typename T, int N, int P, int D, int R>
__device__ __forceinline__ void merge_pass(T (&x)[N])
{
#pragma unroll
    for (int i = 0; i + D < N; ++i)
        if ((i & P) == R) {
            const T a = x[i], b = x[i + D];
            x[i] = vmin(a, b);
            x[i + D] = vmax(a, b);
        }
}

// passes (d = q - p, r = p) for q = 2^QI, 2^(QI-1), ..., 2p
template <typename T, int N, int PI, int QI>
__device__ __forceinline__ void merge_tail(T (&x)[N])
{
    if constexpr (QI > PI) {
        merge_pass<T, N, (1 << PI), (1 << QI) - (1 << PI), (1 << PI)>(x);
        merge_tail<T, N, PI, QI - 1>(x);
    }
}

template <typename T, int N, int PI>
__device__ __forceinline__ void merge_round(T (&x)[N])
{
    if constexpr (PI >= 0) {
        constexpr int t = ceil_log2<N>();
        merge_pass<T, N, (1 << PI), (1 << PI), 0>(x);
        merge_tail<T, N, PI, t - 1>(x);
        merge_round<T, N, PI - 1>(x);
    }
}

template <typename T, int N>
__device__ __forceinline__ void sort_net(T (&x)[N])
{
    merge_round<T, N, ceil_log2<N>() - 1>(x);
}

// Sorts a V-shaped sequence: the first entries non-increasing, then
// non-decreasing, the rest +inf padding (the residuals |x_i - m| of an
// ascending x with its padding). Such a sequence is bitonic, and a bitonic
// merge (half-cleaners of widths P/2, P/4, ..., 1, P the next power of two)
// sorts it; the comparators that would reach past N only meet the implied
// +inf padding and are identities. For N = 24: 67 min/max operations on the
// outputs a median reads, against 221 for the full sorting network.
template <typename T, int N, int D>
__device__ __forceinline__ void half_cleaners(T (&x)[N])
{
    if constexpr (D >= 1) {
#pragma unroll
        for (int i = 0; i + D < N; ++i)
            if ((i / D) % 2 == 0) {
                const T a = x[i], b = x[i + D];
                x[i] = vmin(a, b);
                x[i + D] = vmax(a, b);
            }
        half_cleaners<T, N, D / 2>(x);
    }
}

template <typename T, int N>
__device__ __forceinline__ void sort_vshaped(T (&x)[N])
{
    half_cleaners<T, N, (1 << ceil_log2<N>()) / 2>(x);
}

// median! of the first n of N sorted-with-padding registers (padding = +inf
// placed after the valid values, n <= N), in type R (the two middle values
// converted exactly, then halved and added in R). hasnan forces NaN.
template <typename R, typename T, int N>
__device__ __forceinline__ R median_sorted_as(const T (&x)[N], int n, bool hasnan)
{
    if (hasnan) return (R)NAN;
    R lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i <= N / 2; ++i) {
        if (i == (n - 1) / 2) lo = (R)x[i];
        if (i == n / 2) hi = (R)x[i];
    }
    if (n & 1) return lo;
    return lo / (R)2 + hi / (R)2;
}

template <typename T, int N>
__device__ __forceinline__ T median_sorted(const T (&x)[N], int n, bool hasnan)
{
    return median_sorted_as<T, T, N>(x, n, hasnan);
}

template <int NB>
// init: the first validator of a validation, which writes every cell's flag
// (so no separate clear is needed) and zeroes the per-segment counters;
// otherwise flagged cells are ORed in.
__global__ __launch_bounds__(256) void k_uod(const float *__restrict__ u, const float *__restrict__ v, Grids g,
                                             double thr, double eps, const uint8_t *__restrict__ exclude,
                                             uint8_t *__restrict__ flags, int init, int *__restrict__ count)
{
    constexpr int W = 2 * NB + 1;
    constexpr int N = W * W - 1;  // neighbours; the box centre is linear index N/2
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    if (init && ee < g.nseg) count[ee] = 0;
    if (ee >= g.total) return;
    const Cell k = cell_of(g, ee);
    const int e = k.b, sgoff = k.sg.off, ny = k.sg.ny, nx = k.sg.nx, r = k.r, c = k.c;
    if (init) flags[e] = 0;
    if (exclude && exclude[e]) return;
    const GridSeg sg = k.sg;
    const float *__restrict__ ug = u + sg.off;
    const float *__restrict__ vg = v + sg.off;
    const uint8_t *__restrict__ xg = exclude ? exclude + sgoff : nullptr;
    float su[N], sv[N];
    int n = 0;
    bool nanu = false, nanv = false;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const int lin = q < N / 2 ? q : q + 1;
        const int dc = lin / W - NB, dr = lin % W - NB;
        const int r2 = r + dr, c2 = c + dc;
        // clamped addresses: every load is in bounds and unconditional
        const int rc = min(max(r2, 0), ny - 1), cc = min(max(c2, 0), nx - 1);
        const int at = rc + cc * ny;
        const bool ok = r2 == rc && c2 == cc && !(xg && xg[at]);
        su[q] = ok ? ug[at] : INFINITY;
        sv[q] = ok ? vg[at] : INFINITY;
        nanu |= ok && isnan(su[q]);
        nanv |= ok && isnan(sv[q]);
        n += ok;
    }
    if (n == 0) return;
    sort_net<float, N>(su);
    sort_net<float, N>(sv);
    const float umed = median_sorted<float, N>(su, n, nanu);
    const float vmed = median_sorted<float, N>(sv, n, nanv);
    // residuals of the sorted values: the first n are the valid values'
    // (the same multiset as over the neighbours), the padding stays +inf,
    // and the sequence is V-shaped. A valid residual is NaN exactly when the
    // median is not finite (NaN among the values, or an infinite median,
    // which some valid value equals), and then the MAD is NaN whatever the
    // other entries hold.
    const bool nanru = !isfinite(umed), nanrv = !isfinite(vmed);
#pragma unroll
    for (int q = 0; q < N; ++q) {
        su[q] = fabsf(su[q] - umed);
        sv[q] = fabsf(sv[q] - vmed);
    }
    sort_vshaped<float, N>(su);
    sort_vshaped<float, N>(sv);
    const float madu = median_sorted<float, N>(su, n, nanru);
    const float madv = median_sorted<float, N>(sv, n, nanrv);
    const float du = fabsf(u[e] - umed), dv = fabsf(v[e] - vmed);
    const double un = (double)du / ((double)madu + eps);
    const double vn = (double)dv / ((double)madv + eps);
    if (un > thr || vn > thr) flags[e] = 1;
}

// Local validators (src/quality.jl:479-496): OR into flags.
__global__ __launch_bounds__(256) void k_local_validators(const float *__restrict__ u, const float *__restrict__ v,
                                                          const float *__restrict__ pr, const float *__restrict__ cm,
                                                          Grids g, int kind, double a, double b,
                                                          uint8_t *__restrict__ flags)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    if (ee >= g.total) return;
    const int e = cell_of(g, ee).b;
    bool bad = false;
    if (kind == HH_V_PEAK_RATIO)
        bad = !((double)pr[e] >= a);
    else if (kind == HH_V_MOMENT)
        bad = !((double)cm[e] <= a);
    else if (kind == HH_V_VELMAG) {
        const double mag = (double)hypotf(u[e], v[e]);
        bad = !(a <= mag && mag <= b);
    }
    if (bad) flags[e] = 1;
}

// k-th smallest (1-based) of the valid entries of a box by rank counting.
template <typename Get>
__device__ double kth_box(const Get &get, int r0, int r1, int c0, int c1, int k)
{
    for (int c = c0; c <= c1; ++c)
        for (int r = r0; r <= r1; ++r) {
            double x;
            if (!get(r, c, x)) continue;
            int less = 0, leq = 0;
            for (int c2 = c0; c2 <= c1; ++c2)
                for (int r2 = r0; r2 <= r1; ++r2) {
                    double y;
                    if (!get(r2, c2, y)) continue;
                    less += y < x;
                    leq += y <= x;
                }
            if (less < k && k <= leq) return x;
        }
    return NAN;
}

template <typename Get>
__device__ double median_box(const Get &get, int r0, int r1, int c0, int c1, int n)
{
    // NaN if any valid entry is NaN
    for (int c = c0; c <= c1; ++c)
        for (int r = r0; r <= r1; ++r) {
            double x;
            if (get(r, c, x) && isnan(x)) return NAN;
        }
    if (n & 1) return kth_box(get, r0, r1, c0, c1, (n + 1) / 2);
    const double a = kth_box(get, r0, r1, c0, c1, n / 2);
    const double b = kth_box(get, r0, r1, c0, c1, n / 2 + 1);
    return a / 2 + b / 2;
}

// substitute_alternatives! (src/quality.jl:307-351). The valid neighbours
// of a flagged cell are gathered into registers (+inf padding) and sorted
// by a network, so the Float64 median and MAD cost no memory round trips.
// One cell of substitute_alternatives!; returns the cell's final flag.
template <int NB>
__device__ __forceinline__ uint8_t substitute_cell(const float *__restrict__ uref, const float *__restrict__ vref,
                                                   const uint8_t *__restrict__ flags_in, const Grids &g,
                                                   const Cell &k, double thr, const float *__restrict__ alt_u,
                                                   const float *__restrict__ alt_v, int nalt,
                                                   float *__restrict__ u_out, float *__restrict__ v_out,
                                                   uint8_t *__restrict__ flags_out, const uint8_t *__restrict__ mask)
{
    constexpr int W = 2 * NB + 1;
    constexpr int N = W * W - 1;
    const int e = k.b, ny = k.sg.ny, nx = k.sg.nx, r = k.r, c = k.c;
    const GridSeg sg = k.sg;
    u_out[e] = uref[e];
    v_out[e] = vref[e];
    flags_out[e] = flags_in[e];
    if (!flags_in[e]) return 0;  // masked cells carry no flag
    const float *__restrict__ ug = uref + sg.off;
    const float *__restrict__ vg = vref + sg.off;
    const uint8_t *__restrict__ fg = flags_in + sg.off;
    const uint8_t *__restrict__ mg = mask ? mask + sg.off : nullptr;
    float su[N], sv[N];
    int n = 0;
    bool nanu = false, nanv = false;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const int lin = q < N / 2 ? q : q + 1;
        const int dc = lin / W - NB, dr = lin % W - NB;
        const int r2 = r + dr, c2 = c + dc;
        // clamped addresses: the flag, mask and vector loads of every
        // neighbour are unconditional and in flight together (a load behind
        // its flag test would wait for the flag)
        const int rc = min(max(r2, 0), ny - 1), cc = min(max(c2, 0), nx - 1);
        const int at = rc + cc * ny;
        const uint8_t fq = fg[at], mq = mg ? mg[at] : (uint8_t)0;
        const float uq = ug[at], vq = vg[at];
        const bool ok = r2 == rc && c2 == cc && !fq && !mq;
        su[q] = ok ? uq : INFINITY;
        sv[q] = ok ? vq : INFINITY;
        nanu |= ok && isnan(uq);
        nanv |= ok && isnan(vq);
        n += ok;
    }
    if (n < 3) return 1;
    // the values sort as floats (their Float64 images order the same); the
    // medians and residuals are formed in Float64 as the reference does
    sort_net<float, N>(su);
    sort_net<float, N>(sv);
    const double mu = median_sorted_as<double, float, N>(su, n, nanu);
    const double mv = median_sorted_as<double, float, N>(sv, n, nanv);
    // residuals of the sorted values (the valid multiset, then +inf padding,
    // V-shaped; NaN among the valid ones exactly when the median is not
    // finite, as in k_uod)
    double ru[N], rv[N];
    const bool nanru = !isfinite(mu), nanrv = !isfinite(mv);
#pragma unroll
    for (int q = 0; q < N; ++q) {
        ru[q] = fabs((double)su[q] - mu);
        rv[q] = fabs((double)sv[q] - mv);
    }
    sort_vshaped<double, N>(ru);
    sort_vshaped<double, N>(rv);
    const double madu = median_sorted<double, N>(ru, n, nanru);
    const double madv = median_sorted<double, N>(rv, n, nanrv);
    const size_t ng = (size_t)g.stride;  // alternatives: stride x nalt
    for (int q = 0; q < nalt; ++q) {
        const float au = alt_u[e + q * ng], av = alt_v[e + q * ng];
        if (!(isfinite(au) && isfinite(av))) break;
        if (fabs((double)au - mu) / (madu + 0.1) <= thr && fabs((double)av - mv) / (madv + 0.1) <= thr) {
            u_out[e] = au;
            v_out[e] = av;
            flags_out[e] = 0;
            return 0;
        }
    }
    return 1;
}

// A counter is read only as "fewer than 3 valid cells", so it saturates: an
// add whose counter already reads >= 3 is skipped (the final value is >= 3
// exactly when the segment has >= 3 valid cells, and exact below that),
// which keeps a whole grid's blocks from queueing atomics on one address.
__device__ __forceinline__ void add_count(int *__restrict__ count, int seg, int x)
{
    if (__hip_atomic_load(count + seg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 3) atomicAdd(count + seg, x);
}

// Adds the valid cells of a block to their segments' counters: one atomic
// per block when the block lies in one segment (always, for a single grid),
// else one per wave or per lane. Every thread of the block calls it.
__device__ __forceinline__ void add_valid(const Grids &g, int ee, bool valid, int *__restrict__ count)
{
    __shared__ int part[4], segw[4];
    const Cell k = cell_of(g, ee < g.total ? ee : g.total - 1);
    const int si = k.si, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = ee < g.total && valid ? 1 : 0;
    const int s0 = __shfl(si, 0);
    const bool uni = __all(si == s0);
    int wsum = x;
    for (int o = 32; o > 0; o >>= 1) wsum += __shfl_xor(wsum, o);
    if (lane == 0) {
        part[wave] = wsum;
        segw[wave] = uni ? s0 : -1;
    }
    __syncthreads();
    const int nw = (int)(blockDim.x >> 6);
    bool block_uni = true;
    for (int w = 0; w < nw; ++w) block_uni &= segw[w] == segw[0] && segw[w] >= 0;
    if (block_uni) {
        if (threadIdx.x == 0) {
            int t = 0;
            for (int w = 0; w < nw; ++w) t += part[w];
            if (t) add_count(count, segw[0], t);
        }
    } else if (uni) {
        if (lane == 0 && wsum) add_count(count, s0, wsum);
    } else if (x) {
        add_count(count, si, 1);
    }
}

// Number of valid cells per segment: with fewer than 3 in a grid no vector
// of that grid can be replaced and k_replace returns at once.
// mask (nullable) marks further invalid cells.
__global__ __launch_bounds__(256) void k_count_valid(const uint8_t *__restrict__ invalid,
                                                     const uint8_t *__restrict__ mask, Grids g,
                                                     int *__restrict__ count)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    if (ee < g.total) {
        const int b = cell_of(g, ee).b;
        valid = !invalid[b] && !(mask && mask[b]);
    }
    add_valid(g, ee, valid, count);
}

// substitute_alternatives! over the enumerated cells; with count non-null the
// valid cells of the result (unflagged and unmasked) are also counted per
// segment (for k_replace). mask (nullable): cells that are neither
// neighbours nor substituted.
template <int NB>
__global__ __launch_bounds__(256) void k_substitute(const float *__restrict__ uref, const float *__restrict__ vref,
                                                    const uint8_t *__restrict__ flags_in, Grids g, double thr,
                                                    const float *__restrict__ alt_u, const float *__restrict__ alt_v,
                                                    int nalt, float *__restrict__ u_out, float *__restrict__ v_out,
                                                    uint8_t *__restrict__ flags_out, int *__restrict__ count,
                                                    const uint8_t *__restrict__ mask)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    if (ee < g.total) {
        const Cell k = cell_of(g, ee);
        valid = substitute_cell<NB>(uref, vref, flags_in, g, k, thr, alt_u, alt_v, nalt, u_out, v_out, flags_out,
                                    mask) == 0 &&
                !(mask && mask[k.b]);
    }
    if (count) add_valid(g, ee, valid, count);
}

// The value replace_vectors! gives an invalid, unmasked cell (r, c) of a
// segment (0-based): the median of the valid vectors of the smallest box
// around it (radius 1, 2, ...) holding at least 3, or its own vector when the
// segment has fewer than 3 valid cells (nval).
__device__ __forceinline__ float2 replace_invalid(const float *__restrict__ uref, const float *__restrict__ vref,
                                                  const uint8_t *__restrict__ invalid,
                                                  const uint8_t *__restrict__ mask, int nval, int ny, int nx, int r,
                                                  int c)
{
    const int e = r + c * ny;
    auto bad = [&](int at) { return invalid[at] || (mask && mask[at]); };
    {
        // radius 1 (the common case) from registers; every neighbour's flag,
        // mask and vector are loaded together at clamped addresses (a load
        // behind a test would wait for it)
        double nu[9], nv[9];
        int n = 0;
        bool nanu = false, nanv = false;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int r2 = r + q % 3 - 1, c2 = c + q / 3 - 1;
            const int rc = min(max(r2, 0), ny - 1), cc = min(max(c2, 0), nx - 1);
            const int at = rc + cc * ny;
            const bool bq = invalid[at] || (mask && mask[at]);
            const float uq = uref[at], vq = vref[at];
            const bool ok = r2 == rc && c2 == cc && !bq;
            nu[q] = ok ? (double)uq : INFINITY;
            nv[q] = ok ? (double)vq : INFINITY;
            nanu |= ok && isnan(nu[q]);
            nanv |= ok && isnan(nv[q]);
            n += ok;
        }
        if (nval < 3) return make_float2(uref[e], vref[e]);
        if (n >= 3) {
            sort_net<double, 9>(nu);
            sort_net<double, 9>(nv);
            return make_float2((float)median_sorted<double, 9>(nu, n, nanu),
                               (float)median_sorted<double, 9>(nv, n, nanv));
        }
    }
    const int maxr = max(ny, nx) - 1;
    for (int rad = 2; rad <= maxr; ++rad) {
        const int r0 = max(0, r - rad), r1 = min(ny - 1, r + rad), c0 = max(0, c - rad), c1 = min(nx - 1, c + rad);
        int n = 0;
        for (int c2 = c0; c2 <= c1; ++c2)
            for (int r2 = r0; r2 <= r1; ++r2) n += !bad(r2 + c2 * ny);
        if (n < 3) continue;
        auto gu = [&](int r2, int c2, double &x) {
            if (bad(r2 + c2 * ny)) return false;
            x = (double)uref[r2 + c2 * ny];
            return true;
        };
        auto gv = [&](int r2, int c2, double &x) {
            if (bad(r2 + c2 * ny)) return false;
            x = (double)vref[r2 + c2 * ny];
            return true;
        };
        return make_float2((float)median_box(gu, r0, r1, c0, c1, n), (float)median_box(gv, r0, r1, c0, c1, n));
    }
    return make_float2(uref[e], vref[e]);
}

// The same as a call (the fused kernel below reaches it from nine places).
__device__ __noinline__ float2 replace_invalid_call(const float *__restrict__ uref, const float *__restrict__ vref,
                                                    const uint8_t *__restrict__ invalid, int nval, int ny, int nx,
                                                    int r, int c)
{
    return replace_invalid(uref, vref, invalid, nullptr, nval, ny, nx, r, c);
}

// flags_out (nullable) receives a copy of invalid. mask (nullable): cells
// that are invalid as well and end as NaN (validate_and_replace!,
// src/pipeline.jl:636-642: the invalid set is outliers .| mask, masked
// vectors are NaN afterwards).
__global__ __launch_bounds__(256) void k_replace(const float *__restrict__ uref_all, const float *__restrict__ vref_all,
                                                 const uint8_t *__restrict__ invalid_all, Grids g,
                                                 const int *__restrict__ nvalid, float *__restrict__ u_out,
                                                 float *__restrict__ v_out, uint8_t *__restrict__ flags_out,
                                                 const uint8_t *__restrict__ mask_all)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    if (ee >= g.total) return;
    const Cell k = cell_of(g, ee);
    const int e = k.b;
    const GridSeg sg = k.sg;
    u_out[e] = uref_all[e];
    v_out[e] = vref_all[e];
    if (flags_out) flags_out[e] = invalid_all[e];
    if (mask_all && mask_all[e]) {
        u_out[e] = v_out[e] = NAN;
        return;
    }
    if (!invalid_all[e]) return;
    const float2 f = replace_invalid(uref_all + sg.off, vref_all + sg.off, invalid_all + sg.off,
                                     mask_all ? mask_all + sg.off : nullptr, nvalid[k.si], sg.ny, sg.nx, k.r, k.c);
    u_out[e] = f.x;
    v_out[e] = f.y;
}

// The (1, 2, 1) / 4 smoothing of build_predictor (src/pipeline.jl:454-464,
// src/quality.jl:123-136) at one cell from its 3 x 3 neighbourhood f[dc][dr]
// (edges replicated).
__device__ __forceinline__ float smooth3x3(const float (&f)[3][3])
{
    const float q = 0.25f;
    auto tmp = [&](int dr) { return q * ((f[0][dr] + 2.f * f[1][dr]) + f[2][dr]); };
    return q * ((tmp(0) + 2.f * tmp(1)) + tmp(2));
}

// k_replace (no mask) followed by the predictor's smoothing of the replaced
// field in one launch: each cell also forms the replaced values of its 3 x 3
// neighbourhood (a replacement only reads the input field, so neighbours'
// results need no exchange; flagged cells are few, and the lanes of a wave
// that need a neighbour's median run the call together) and writes the
// smoothed pair to (pu, pv).
__global__ __launch_bounds__(256) void k_replace_smooth(const float *__restrict__ uref_all,
                                                        const float *__restrict__ vref_all,
                                                        const uint8_t *__restrict__ invalid_all, Grids g,
                                                        const int *__restrict__ nvalid, float *__restrict__ u_out,
                                                        float *__restrict__ v_out, uint8_t *__restrict__ flags_out,
                                                        float *__restrict__ pu, float *__restrict__ pv)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    if (ee >= g.total) return;
    const Cell k = cell_of(g, ee);
    const int e = k.b, ny = k.sg.ny, nx = k.sg.nx, r = k.r, c = k.c;
    const float *__restrict__ ur = uref_all + k.sg.off;
    const float *__restrict__ vr = vref_all + k.sg.off;
    const uint8_t *__restrict__ inv = invalid_all + k.sg.off;
    const int nval = nvalid[k.si];
    if (flags_out) flags_out[e] = invalid_all[e];
    // the nine flags and vectors in flight together, then the (rare)
    // replacements
    float fu[3][3], fv[3][3];
    bool fl[3][3];
#pragma unroll
    for (int dc = 0; dc < 3; ++dc)
#pragma unroll
        for (int dr = 0; dr < 3; ++dr) {
            const int at = min(max(r + dr - 1, 0), ny - 1) + min(max(c + dc - 1, 0), nx - 1) * ny;
            fl[dc][dr] = inv[at];
            fu[dc][dr] = ur[at];
            fv[dc][dr] = vr[at];
        }
#pragma unroll
    for (int dc = 0; dc < 3; ++dc)
#pragma unroll
        for (int dr = 0; dr < 3; ++dr)
            if (fl[dc][dr]) {
                const float2 f = replace_invalid_call(ur, vr, inv, nval, ny, nx, min(max(r + dr - 1, 0), ny - 1),
                                                      min(max(c + dc - 1, 0), nx - 1));
                fu[dc][dr] = f.x;
                fv[dc][dr] = f.y;
            }
    u_out[e] = fu[1][1];
    v_out[e] = fv[1][1];
    pu[e] = smooth3x3(fu);
    pv[e] = smooth3x3(fv);
}

__global__ __launch_bounds__(256) void k_smooth(const float *__restrict__ f_all, const float *__restrict__ f2_all,
                                                Grids g, float *__restrict__ out, float *__restrict__ out2)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    if (ee >= g.total) return;
    const Cell k = cell_of(g, ee);
    const int e = k.b, ny = k.sg.ny, nx = k.sg.nx, r = k.r, c = k.c;
    const int cl = c > 0 ? c - 1 : 0, cr = c < nx - 1 ? c + 1 : nx - 1;
    const int rl = r > 0 ? r - 1 : 0, rr = r < ny - 1 ? r + 1 : ny - 1;
    auto sm = [&](const float *__restrict__ f) {
        const int cs[3] = {cl, c, cr}, rs[3] = {rl, r, rr};
        float n3[3][3];
#pragma unroll
        for (int dc = 0; dc < 3; ++dc)
#pragma unroll
            for (int dr = 0; dr < 3; ++dr) n3[dc][dr] = f[rs[dr] + cs[dc] * ny];
        return smooth3x3(n3);
    };
    out[e] = sm(f_all + k.sg.off);
    if (f2_all) out2[e] = sm(f2_all + k.sg.off);
}

__global__ void k_fill_u8(uint8_t *p, int n, uint8_t v)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = v;
}

// Clear the flags and the valid-cell counters (one per segment) of one
// validation.
__global__ void k_val_init(uint8_t *flags, Grids g, int *count)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    if (ee < g.total) flags[cell_of(g, ee).b] = 0;
    if (ee < g.nseg) count[ee] = 0;
}

// Masked cells leave the outlier set (outliers .&= .!mask,
// src/pipeline.jl:628).
__global__ void k_unflag_masked(uint8_t *__restrict__ flags, const uint8_t *__restrict__ mask, Grids g)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    if (ee >= g.total) return;
    const int e = cell_of(g, ee).b;
    if (mask[e]) flags[e] = 0;
}

// Masked cells of a sweep hold NaN in u, v, peak ratio and correlation moment
// (src/pipeline.jl:763-766).
__global__ void k_mask_nan(const uint8_t *__restrict__ mask, Grids g, float *__restrict__ u, float *__restrict__ v,
                           float *__restrict__ pr, float *__restrict__ cm)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    if (ee >= g.total) return;
    const int e = cell_of(g, ee).b;
    if (mask[e]) u[e] = v[e] = pr[e] = cm[e] = NAN;
}

__global__ void k_zero_count(int *count, int n)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) count[e] = 0;
}

// Move a result computed into scratch back to the chain buffers.
__global__ void k_copy_back(const float *__restrict__ su, const float *__restrict__ sv,
                            const uint8_t *__restrict__ sf, float *__restrict__ u, float *__restrict__ v,
                            uint8_t *__restrict__ f, Grids g)
{
    const int ee = blockIdx.x * blockDim.x + threadIdx.x;
    if (ee >= g.total) return;
    const int e = cell_of(g, ee).b;
    u[e] = su[e];
    v[e] = sv[e];
    if (sf) f[e] = sf[e];
}

template <int NB>
void uod_launch(const ValidateArgs &a, double thr, double eps, bool init, hipStream_t s)
{
    const int n = std::max(a.grids.total, a.grids.nseg);
    hipLaunchKernelGGL(k_uod<NB>, dim3((n + 255) / 256), dim3(256), 0, s, a.u, a.v, a.grids, thr, eps, a.mask,
                       a.outliers, init ? 1 : 0, a.count);
}

void uod(const ValidateArgs &a, int nb, double thr, double eps, bool init, hipStream_t s)
{
    switch (nb) {
    case 1: uod_launch<1>(a, thr, eps, init, s); break;
    case 2: uod_launch<2>(a, thr, eps, init, s); break;
    case 3: uod_launch<3>(a, thr, eps, init, s); break;
    default: throw Error(HH_EUNSUPPORTED, "UOD neighborhood sizes 1, 2 and 3 are supported on the device");
    }
}

}  // namespace

// Launches: the first UOD writes every flag and clears the counters (else a
// clear kernel does), the substitution also counts the valid cells the
// replacement needs (else a count kernel does), then the replacement (with
// the predictor's smoothing when a.smooth_u is set and there is no mask).
bool launch_validate(const ValidateArgs &a, hipStream_t s)
{
    const hh_pass &p = *a.pass;
    const int n = a.grids.total;
    if (n <= 0) return false;
    const int ntile = std::max(n, a.grids.nseg);
    const dim3 g((n + 255) / 256), b(256);
    if (p.uod_enable)
        uod(a, p.uod_neighborhood, p.uod_threshold, 0.1, true, s);
    else
        hipLaunchKernelGGL(k_val_init, dim3((ntile + 255) / 256), b, 0, s, a.outliers, a.grids, a.count);
    if (p.min_peak_ratio > 1)
        hipLaunchKernelGGL(k_local_validators, g, b, 0, s, a.u, a.v, a.pr, a.cm, a.grids, (int)HH_V_PEAK_RATIO,
                           p.min_peak_ratio, 0.0, a.outliers);
    for (int q = 0; q < p.n_validators; ++q) {
        const hh_validator &vd = p.validators[q];
        if (vd.kind == HH_V_UOD)
            uod(a, vd.neighborhood, vd.a, vd.b, false, s);
        else
            hipLaunchKernelGGL(k_local_validators, g, b, 0, s, a.u, a.v, a.pr, a.cm, a.grids, (int)vd.kind, vd.a, vd.b,
                               a.outliers);
    }
    if (a.mask) hipLaunchKernelGGL(k_unflag_masked, g, b, 0, s, a.outliers, a.mask, a.grids);
    // Stages alternate between the chain buffers and the scratch buffers; a
    // result left in scratch is copied back once at the end.
    const bool sub = a.alt_u && p.n_peaks > 1;
    const bool rep = a.force_replace || p.replace_outliers;
    int *cnt = rep ? a.count : nullptr;  // counters zeroed above
    if (sub) {
        switch (p.uod_neighborhood) {
        case 1:
            hipLaunchKernelGGL(k_substitute<1>, g, b, 0, s, a.u, a.v, a.outliers, a.grids, p.uod_threshold,
                               a.alt_u, a.alt_v, p.n_peaks - 1, a.su, a.sv, a.sflag, cnt, a.mask);
            break;
        case 2:
            hipLaunchKernelGGL(k_substitute<2>, g, b, 0, s, a.u, a.v, a.outliers, a.grids, p.uod_threshold,
                               a.alt_u, a.alt_v, p.n_peaks - 1, a.su, a.sv, a.sflag, cnt, a.mask);
            break;
        default:
            hipLaunchKernelGGL(k_substitute<3>, g, b, 0, s, a.u, a.v, a.outliers, a.grids, p.uod_threshold,
                               a.alt_u, a.alt_v, p.n_peaks - 1, a.su, a.sv, a.sflag, cnt, a.mask);
            break;
        }
    }
    const bool smooth = rep && !a.mask && a.smooth_u && a.smooth_v;
    if (rep) {
        if (sub) {  // substituted field in scratch: replace back into the chain
            if (smooth)
                hipLaunchKernelGGL(k_replace_smooth, g, b, 0, s, a.su, a.sv, a.sflag, a.grids,
                                   (const int *)a.count, a.u, a.v, a.outliers, a.smooth_u, a.smooth_v);
            else
                hipLaunchKernelGGL(k_replace, g, b, 0, s, a.su, a.sv, a.sflag, a.grids, (const int *)a.count, a.u,
                                   a.v, a.outliers, a.mask);
        } else {
            hipLaunchKernelGGL(k_count_valid, g, b, 0, s, a.outliers, a.mask, a.grids, a.count);
            if (smooth)
                hipLaunchKernelGGL(k_replace_smooth, g, b, 0, s, a.u, a.v, a.outliers, a.grids,
                                   (const int *)a.count, a.su, a.sv, (uint8_t *)nullptr, a.smooth_u, a.smooth_v);
            else
                hipLaunchKernelGGL(k_replace, g, b, 0, s, a.u, a.v, a.outliers, a.grids, (const int *)a.count,
                                   a.su, a.sv, (uint8_t *)nullptr, a.mask);
        }
    }
    if (sub != rep)  // exactly one stage ran: its result is in scratch
        hipLaunchKernelGGL(k_copy_back, g, b, 0, s, a.su, a.sv, sub ? a.sflag : (const uint8_t *)nullptr, a.u, a.v,
                           a.outliers, a.grids);
    HH_HIP(hipGetLastError());
    return smooth;
}

void launch_mask_nan(const Grids &g, const uint8_t *mask, float *u, float *v, float *pr, float *cm, hipStream_t s)
{
    if (g.total <= 0) return;
    hipLaunchKernelGGL(k_mask_nan, dim3((g.total + 255) / 256), dim3(256), 0, s, mask, g, u, v, pr, cm);
    HH_HIP(hipGetLastError());
}

void launch_fill_masked(const Grids &g, const float *u, const float *v, const uint8_t *mask, int *count, float *ou,
                        float *ov, hipStream_t s)
{
    const int n = g.total;
    if (n <= 0) return;
    const dim3 gr((n + 255) / 256), b(256);
    hipLaunchKernelGGL(k_zero_count, dim3((g.nseg + 255) / 256), b, 0, s, count, g.nseg);
    hipLaunchKernelGGL(k_count_valid, gr, b, 0, s, mask, (const uint8_t *)nullptr, g, count);
    hipLaunchKernelGGL(k_replace, gr, b, 0, s, u, v, mask, g, (const int *)count, ou, ov, (uint8_t *)nullptr,
                       (const uint8_t *)nullptr);
    HH_HIP(hipGetLastError());
}

void launch_smooth(const Grids &g, const float *f, float *out, hipStream_t s)
{
    const int n = g.total;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_smooth, dim3((n + 255) / 256), dim3(256), 0, s, f, (const float *)nullptr, g, out,
                       (float *)nullptr);
    HH_HIP(hipGetLastError());
}

void launch_smooth2(const Grids &g, const float *f, const float *f2, float *out, float *out2, hipStream_t s)
{
    const int n = g.total;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_smooth, dim3((n + 255) / 256), dim3(256), 0, s, f, f2, g, out, out2);
    HH_HIP(hipGetLastError());
}

}  // namespace hh
