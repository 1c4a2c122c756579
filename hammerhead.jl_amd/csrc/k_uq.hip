// k_uq.hip — Wieneke (2015) correlation-statistics uncertainty per window.
//
// uncertainty_sweep! / accumulate_uncertainty! / uq_component! /
// finalize_uncertainty (src/pipeline.jl:873-899, src/uncertainty.jl:42-179):
// windows apod .* (x - mean) of the SAME window-sized origin in A and B; per
// component (v on the transposed windows) dC = a0 b1 - a1 b0 in T, C0, C+, C-
// as Float64 sums of T products, [1 2 1]/4 smoothing along the shift with
// clamped ends, the Float64 window mean subtracted, S00 and the 40 half-plane
// covariance sums of UQ_OFFSETS (src/uncertainty.jl:23-28); then the ring
// truncation at 0.05 S00 and eq. (9). One wave64 per window; the 44 sums are
// accumulated in Float64 lanes and reduced across the wave.

#include "hh_internal.h"

namespace hh {

namespace {

#pragma clang fp contract(off)

constexpr int kNoff = 40;
constexpr int kMaxW = 64;

struct Offsets {
    int8_t dr[kNoff], dc[kNoff];
    int ring_start[5];
};

Offsets make_offsets()
{
    Offsets o;
    int tmp[kNoff][2], n = 0;
    for (int dc = 0; dc <= 4; ++dc)
        for (int dr = (dc == 0 ? 1 : -4); dr <= 4; ++dr) {
            tmp[n][0] = dr;
            tmp[n][1] = dc;
            ++n;
        }
    int k = 0;
    for (int ring = 1; ring <= 4; ++ring) {
        o.ring_start[ring - 1] = k;
        for (int i = 0; i < n; ++i) {
            const int a = std::abs(tmp[i][0]), b = std::abs(tmp[i][1]);
            if ((a > b ? a : b) == ring) {
                o.dr[k] = (int8_t)tmp[i][0];
                o.dc[k] = (int8_t)tmp[i][1];
                ++k;
            }
        }
    }
    o.ring_start[4] = k;
    return o;
}

__device__ __forceinline__ double wave_sum(double x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ float wave_sumf(float x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Statistics of one component: shift along the second index of the
// (nr x ncl) view; view(A, r, c) = A[r*rs + c*cs]. dC/dCs: float scratch of
// nr*(ncl-1); d: double scratch of the same size. Lane k < 41 accumulates
// lag k (0 = S00, k = offset k-1) over the whole field: one Float64
// accumulator per lane, no cross-lane reduction of the 41 sums.
__device__ void uq_component(double *st, const float *wA, const float *wB, int nr, int ncl, int rs, int cs, float *dC,
                             float *dCs, double *d, const Offsets &off)
{
    const int lane = threadIdx.x & 63;
    const int m = ncl - 1;
    const int ne = nr * m;
    double C0 = 0.0, Cp = 0.0, Cm = 0.0;
    for (int e = lane; e < ne; e += 64) {
        const int r = e % nr, c = e / nr;
        const float a0 = wA[r * rs + c * cs], a1 = wA[r * rs + (c + 1) * cs];
        const float b0 = wB[r * rs + c * cs], b1 = wB[r * rs + (c + 1) * cs];
        const float p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0;
        C0 += (double)p00;
        Cp += (double)p01;
        Cm += (double)p10;
        dC[e] = p01 - p10;
    }
    __syncthreads();
    const float q = 0.25f;
    double mu = 0.0;
    for (int e = lane; e < ne; e += 64) {
        const int r = e % nr, c = e / nr;
        const int cl = c > 0 ? c - 1 : 0, cr = c < m - 1 ? c + 1 : m - 1;
        const float v = q * ((dC[r + cl * nr] + 2.f * dC[r + c * nr]) + dC[r + cr * nr]);
        dCs[e] = v;
        mu += (double)v;
    }
    C0 = wave_sum(C0);
    Cp = wave_sum(Cp);
    Cm = wave_sum(Cm);
    mu = wave_sum(mu) / (double)ne;
    __syncthreads();
    for (int e = lane; e < ne; e += 64) d[e] = (double)dCs[e] - mu;
    __syncthreads();
    double S = 0.0;
    if (lane <= kNoff) {
        const int dr = lane == 0 ? 0 : off.dr[lane - 1];
        const int dc = lane == 0 ? 0 : off.dc[lane - 1];
        const int c_lo = dc < 0 ? -dc : 0, c_hi = dc > 0 ? m - dc : m;  // c + dc in [0, m)
        const int r_lo = dr < 0 ? -dr : 0, r_hi = dr > 0 ? nr - dr : nr;
        const int shift = dr + dc * nr;
        for (int c = c_lo; c < c_hi; ++c) {
            const double *col = d + c * nr;
            for (int r = r_lo; r < r_hi; ++r) S += col[r] * col[r + shift];
        }
    }
    if (lane == 0) {
        st[0] += C0;
        st[1] += Cp;
        st[2] += Cm;
    }
    if (lane <= kNoff) st[3 + lane] += S;
    __syncthreads();
}

// 16-px windows (a 16 x 15 field per component): lane i takes column pair
// (c, c + dc) number i of the 65 pairs with dc = 0..4, holds both columns
// in registers and forms the 9 row-lag products (dr = -4..4; 0..4 for
// dc = 0) without memory traffic; the per-pair partial sums are then added
// over c for each of the 41 (dr, dc) statistics. Same sums as the generic
// path, in a different Float64 addition order.
__device__ void uq_component16(double *st, const float *wA, const float *wB, int rs, int cs, float *dC, float *dCs,
                               double *d, double *part, const Offsets &off)
{
    constexpr int nr = 16, m = 15, ne = nr * m;
    const int lane = threadIdx.x & 63;
    double C0 = 0.0, Cp = 0.0, Cm = 0.0;
    for (int e = lane; e < ne; e += 64) {
        const int r = e % nr, c = e / nr;
        const float a0 = wA[r * rs + c * cs], a1 = wA[r * rs + (c + 1) * cs];
        const float b0 = wB[r * rs + c * cs], b1 = wB[r * rs + (c + 1) * cs];
        const float p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0;
        C0 += (double)p00;
        Cp += (double)p01;
        Cm += (double)p10;
        dC[e] = p01 - p10;
    }
    __syncthreads();
    const float q = 0.25f;
    double mu = 0.0;
    for (int e = lane; e < ne; e += 64) {
        const int r = e % nr, c = e / nr;
        const int cl = c > 0 ? c - 1 : 0, cr = c < m - 1 ? c + 1 : m - 1;
        const float v = q * ((dC[r + cl * nr] + 2.f * dC[r + c * nr]) + dC[r + cr * nr]);
        dCs[e] = v;
        mu += (double)v;
    }
    C0 = wave_sum(C0);
    Cp = wave_sum(Cp);
    Cm = wave_sum(Cm);
    mu = wave_sum(mu) / (double)ne;
    __syncthreads();
    for (int e = lane; e < ne; e += 64) d[e] = (double)dCs[e] - mu;
    __syncthreads();
    // column pairs: dc = 0..4 has 15 - dc pairs (65 in all)
    for (int item = lane; item < 65; item += 64) {
        int dc = 0, c = item;
        while (c >= m - dc) {
            c -= m - dc;
            ++dc;
        }
        double a[nr], bb[nr];
#pragma unroll
        for (int r = 0; r < nr; ++r) {
            a[r] = d[r + c * nr];
            bb[r] = d[r + (c + dc) * nr];
        }
#pragma unroll
        for (int dr = -4; dr <= 4; ++dr) {
            double sum = 0.0;
#pragma unroll
            for (int r = 0; r < nr; ++r)
                if (r + dr >= 0 && r + dr < nr) sum += a[r] * bb[r + dr];
            part[item * 9 + dr + 4] = sum;
        }
    }
    __syncthreads();
    double S = 0.0;
    if (lane <= kNoff) {
        const int dr = lane == 0 ? 0 : off.dr[lane - 1];
        const int dc = lane == 0 ? 0 : off.dc[lane - 1];
        int base = 0;
        for (int k = 0; k < dc; ++k) base += m - k;
        for (int c = 0; c < m - dc; ++c) S += part[(base + c) * 9 + dr + 4];
    }
    if (lane == 0) {
        st[0] += C0;
        st[1] += Cp;
        st[2] += Cm;
    }
    if (lane <= kNoff) st[3 + lane] += S;
    __syncthreads();
}

__device__ float uq_finalize(const double *s, const Offsets &off)
{
    const double C0 = s[0], Cp = s[1], Cm = s[2], S00 = s[3];
    double s2 = S00;
    for (int ring = 0; ring < 4; ++ring) {
        double mx = -INFINITY;
        for (int k = off.ring_start[ring]; k < off.ring_start[ring + 1]; ++k) mx = s[4 + k] > mx ? s[4 + k] : mx;
        if (mx < 0.05 * S00) break;
        for (int k = off.ring_start[ring]; k < off.ring_start[ring + 1]; ++k) s2 += 2.0 * s[4 + k];
    }
    const double sdc = sqrt(s2 > 0.0 ? s2 : 0.0);
    const double Cpm = (Cp + Cm) / 2.0;
    const double lo = Cpm - sdc / 2.0, hi = Cpm + sdc / 2.0;
    if (!(C0 > 0.0 && lo > 0.0)) return NAN;
    const double denom = 2.0 * log(C0) - log(hi) - log(lo);
    if (!(denom > 0.0)) return NAN;
    return (float)((log(hi) - log(lo)) / (2.0 * denom));
}

template <int W>
__global__ __launch_bounds__(64) void k_uq(int w, const float *__restrict__ apod, const float *__restrict__ A,
                                           const float *__restrict__ B, int nr, const int32_t *__restrict__ jobs,
                                           int ny, float *__restrict__ unc_u, float *__restrict__ unc_v,
                                           double *__restrict__ stats, int accumulate, Offsets off)
{
    __shared__ float wA[W * W], wB[W * W];
    __shared__ float dC[W * W], dCs[W * W];
    __shared__ double dd[W * W];
    __shared__ double st[2 * HH_UQ_NSTATS];
    const int lane = threadIdx.x;
    const int job = blockIdx.x;
    const int gi = jobs[4 * job], gj = jobs[4 * job + 1], rs = jobs[4 * job + 2], cs = jobs[4 * job + 3];
    const float *Ab = A + (size_t)(cs - 1) * nr + (rs - 1);
    const float *Bb = B + (size_t)(cs - 1) * nr + (rs - 1);
    float sa = 0.f, sb = 0.f;
    for (int e = lane; e < w * w; e += 64) {
        const int i = e % w, j = e / w;
        const float a = Ab[i + (size_t)j * nr], b = Bb[i + (size_t)j * nr];
        wA[e] = a;
        wB[e] = b;
        sa += a;
        sb += b;
    }
    sa = wave_sumf(sa);
    sb = wave_sumf(sb);
    const float meanA = sa / (float)(w * w), meanB = sb / (float)(w * w);
    for (int k = lane; k < 2 * HH_UQ_NSTATS; k += 64) st[k] = 0.0;
    __syncthreads();
    for (int e = lane; e < w * w; e += 64) {
        const float a = apod[e];
        wA[e] = a * (wA[e] - meanA);
        wB[e] = a * (wB[e] - meanB);
    }
    __syncthreads();
    // u: shift along columns (element (r, c) at r + c*w); v: transposed views
    if constexpr (W == 16) {
                __shared__ double part[65 * 9];
        if (w == 16) {
            uq_component16(st, wA, wB, 1, 16, dC, dCs, dd, part, off);
            uq_component16(st + HH_UQ_NSTATS, wA, wB, 16, 1, dC, dCs, dd, part, off);
        } else {
            uq_component(st, wA, wB, w, w, 1, w, dC, dCs, dd, off);
            uq_component(st + HH_UQ_NSTATS, wA, wB, w, w, w, 1, dC, dCs, dd, off);
        }
    } else {
        uq_component(st, wA, wB, w, w, 1, w, dC, dCs, dd, off);
        uq_component(st + HH_UQ_NSTATS, wA, wB, w, w, w, 1, dC, dCs, dd, off);
    }
    if (lane == 0) {
        const size_t g = (size_t)(gi - 1) + (size_t)(gj - 1) * ny;
        if (unc_u) unc_u[g] = uq_finalize(st, off);
        if (unc_v) unc_v[g] = uq_finalize(st + HH_UQ_NSTATS, off);
    }
    if (stats)
        for (int k = lane; k < 2 * HH_UQ_NSTATS; k += 64) {
            double *d = stats + (size_t)2 * HH_UQ_NSTATS * job + k;
            *d = accumulate ? *d + st[k] : st[k];
        }
}

__global__ void k_uq_finalize(const double *__restrict__ stats, const int32_t *__restrict__ jobs, int njobs, int ny,
                              float *__restrict__ unc_u, float *__restrict__ unc_v, Offsets off)
{
    const int job = blockIdx.x * blockDim.x + threadIdx.x;
    if (job >= njobs) return;
    const size_t g = (size_t)(jobs[4 * job] - 1) + (size_t)(jobs[4 * job + 1] - 1) * ny;
    const double *s = stats + (size_t)2 * HH_UQ_NSTATS * job;
    if (unc_u) unc_u[g] = uq_finalize(s, off);
    if (unc_v) unc_v[g] = uq_finalize(s + HH_UQ_NSTATS, off);
}

}  // namespace

void launch_uq(int w, const float *apod, const float *A, const float *B, int nr, const int32_t *jobs, int njobs,
               int ny, float *unc_u, float *unc_v, double *stats, bool accumulate, hipStream_t s)
{
    if (njobs <= 0) return;
    if (w > kMaxW) throw Error(HH_EUNSUPPORTED, "uncertainty supports windows up to 64 px on the device");
    static const Offsets off = make_offsets();
    if (w <= 16)
        hipLaunchKernelGGL(k_uq<16>, dim3(njobs), dim3(64), 0, s, w, apod, A, B, nr, jobs, ny, unc_u, unc_v, stats,
                           (int)accumulate, off);
    else if (w <= 32)
        hipLaunchKernelGGL(k_uq<32>, dim3(njobs), dim3(64), 0, s, w, apod, A, B, nr, jobs, ny, unc_u, unc_v, stats,
                           (int)accumulate, off);
    else
        hipLaunchKernelGGL(k_uq<64>, dim3(njobs), dim3(64), 0, s, w, apod, A, B, nr, jobs, ny, unc_u, unc_v, stats,
                           (int)accumulate, off);
    HH_HIP(hipGetLastError());
}

void launch_uq_finalize(const double *stats, const int32_t *jobs, int njobs, int ny, float *unc_u, float *unc_v,
                        hipStream_t s)
{
    if (njobs <= 0) return;
    static const Offsets off = make_offsets();
    hipLaunchKernelGGL(k_uq_finalize, dim3((njobs + 63) / 64), dim3(64), 0, s, stats, jobs, njobs, ny, unc_u, unc_v,
                       off);
    HH_HIP(hipGetLastError());
}

}  // namespace hh
