// k_correlate.hip — fused per-window correlation for gfx950 (FP32 radix-2 FFT form).
//
// One workgroup (256 lanes = 4 wave64) per interrogation window runs the whole
// per-window chain of process_windows! (src/pipeline.jl:972-1027) in LDS:
//   load_windows! (src/correlators.jl:199-207): mean, apodize, zero-pad;
//   correlation_plane! (src/correlators.jl:326-341): A + iB packed forward
//     2D FFT (radix-2 DIF, bit-reversed output), Hermitian separation into
//     F{A}, F{B} and the fused cross (or phase) spectrum (spectrum!,
//     src/correlators.jl:266-279), inverse 2D FFT (radix-2 DIT from
//     bit-reversed input), fftshift + magnitude + overlap gain
//     (src/correlators.jl:282-289,338);
//   analyze_plane! (src/correlators.jl:459-468): regional-max top-K
//     (is_local_maximum, src/localmax.jl:3-19; find_peaks_regionalmax!,
//     src/correlators.jl:426-454) or the exclusion finder
//     (src/correlators.jl:390-421), gauss3/gauss9 subpixel fit
//     (src/correlators.jl:499-564), peak ratio, correlation moment
//     (src/quality.jl:49-74) and gauss3 alternatives
//     (src/pipeline.jl:1009-1017).
// Only the per-window scalars leave the chip.

#include "hh_internal.h"

#include <mutex>

#include <cfloat>
#include <climits>

namespace hh {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

template <int M>
__device__ __forceinline__ int bitrev(int x)
{
    constexpr int L = __builtin_ctz(M);
    return (int)(__builtin_bitreverse32((unsigned)x) >> (32 - L));
}

// Block reduction of two floats (sum).
__device__ __forceinline__ float2 block_sum2(float a, float b, float *red)
{
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[2 * wid] = a;
        red[2 * wid + 1] = b;
    }
    __syncthreads();
    float sa = 0.f, sb = 0.f;
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) {
        sa += red[2 * k];
        sb += red[2 * k + 1];
    }
    __syncthreads();
    return make_float2(sa, sb);
}

// Peak candidate ordering: larger value first, then smaller column-major
// index (the reference's first-wins insertion, src/correlators.jl:432-446).
struct Cand {
    float v;
    int idx;  // column-major linear index, INT_MAX = none
};
__device__ __forceinline__ bool better(const Cand &a, const Cand &b)
{
    if (a.idx == INT_MAX) return false;
    if (b.idx == INT_MAX) return true;
    return a.v > b.v || (a.v == b.v && a.idx < b.idx);
}

__device__ __forceinline__ Cand block_best(Cand c, Cand *sc)
{
    for (int o = 32; o > 0; o >>= 1) {
        Cand d;
        d.v = __shfl_xor(c.v, o);
        d.idx = __shfl_xor(c.idx, o);
        if (better(d, c)) c = d;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sc[wid] = c;
    __syncthreads();
    Cand b = sc[0];
#pragma unroll
    for (int k = 1; k < kThreads / 64; ++k)
        if (better(sc[k], b)) b = sc[k];
    __syncthreads();
    return b;
}

// R accessor: the fftshifted plane lives in the .y lanes of the LDS plane.
template <int M>
struct PlaneView {
    const float2 *pl;
    __device__ __forceinline__ float operator()(int r, int c) const { return pl[r + c * (M + 1)].y; }  // 0-based
};

template <int M>
__device__ bool is_local_max(const PlaneView<M> &R, int r, int c)
{
    const float I0 = R(r, c);
    if (isnan(I0)) return false;
    const int lin0 = r + c * M;
    const int jlo = c > 0 ? c - 1 : 0, jhi = c < M - 1 ? c + 1 : M - 1;
    const int ilo = r > 0 ? r - 1 : 0, ihi = r < M - 1 ? r + 1 : M - 1;
    for (int jj = jlo; jj <= jhi; ++jj)
        for (int ii = ilo; ii <= ihi; ++ii) {
            if (ii == r && jj == c) continue;
            const float In = R(ii, jj);
            if (ii + jj * M < lin0) {
                if (In > I0) return false;
            } else {
                if (In >= I0) return false;
            }
        }
    return true;
}

// subpixel_gauss3 on 1-based (i, j); returns refined (row, col).
template <int M>
__device__ float2 gauss3(const PlaneView<M> &R, int i, int j)
{
    float di = 0.f, dj = 0.f;
    const float I0 = R(i - 1, j - 1);
    if (1 < i && i < M) {
        const float Im = R(i - 2, j - 1), Ip = R(i, j - 1);
        if (Im > 0.f && I0 > 0.f && Ip > 0.f) {
            const float denom = logf(Im) - 2.f * logf(I0) + logf(Ip);
            di = denom != 0.f ? (logf(Im) - logf(Ip)) / (2.f * denom) : 0.f;
        }
    }
    if (1 < j && j < M) {
        const float Im = R(i - 1, j - 2), Ip = R(i - 1, j);
        if (Im > 0.f && I0 > 0.f && Ip > 0.f) {
            const float denom = logf(Im) - 2.f * logf(I0) + logf(Ip);
            dj = denom != 0.f ? (logf(Im) - logf(Ip)) / (2.f * denom) : 0.f;
        }
    }
    return make_float2((float)i + di, (float)j + dj);
}

// subpixel_gauss9 (src/correlators.jl:534-564).
template <int M>
__device__ float2 gauss9(const PlaneView<M> &R, int pr, int pc)
{
    if (!(1 < pr && pr < M && 1 < pc && pc < M)) return gauss3<M>(R, pr, pc);
    float sL = 0.f, sxL = 0.f, syL = 0.f, sxxL = 0.f, syyL = 0.f, sxyL = 0.f;
    for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) {
            const float val = R(pr + di - 1, pc + dj - 1);
            if (!(val > 0.f)) return gauss3<M>(R, pr, pc);
            const float L = logf(val);
            sL += L;
            sxL += (float)dj * L;
            syL += (float)di * L;
            sxxL += (float)(dj * dj) * L;
            syyL += (float)(di * di) * L;
            sxyL += (float)(di * dj) * L;
        }
    const float a1 = sxL / 6.f, a2 = syL / 6.f, a3 = sxxL / 2.f - sL / 3.f, a4 = sxyL / 4.f,
                a5 = syyL / 2.f - sL / 3.f;
    const float D = 4.f * a3 * a5 - a4 * a4;
    if (!(a3 < 0.f && D > 0.f)) return gauss3<M>(R, pr, pc);
    const float dx = (a4 * a2 - 2.f * a5 * a1) / D;
    const float dy = (a4 * a1 - 2.f * a3 * a2) / D;
    if (!(fabsf(dx) <= 1.f && fabsf(dy) <= 1.f)) return gauss3<M>(R, pr, pc);
    return make_float2((float)pr + dy, (float)pc + dx);
}

// calculate_correlation_moment, 3x3, Float64 sums (src/quality.jl:49-74).
template <int M>
__device__ float moment(const PlaneView<M> &R, float pr, float pc)
{
    if (!(isfinite(pr) && isfinite(pc))) return NAN;
    const int r0 = (int)rintf(pr), c0 = (int)rintf(pc);
    const int rlo = max(1, r0 - 1), rhi = min(M, r0 + 1);
    const int clo = max(1, c0 - 1), chi = min(M, c0 + 1);
    if (rlo > rhi || clo > chi) return NAN;
    double sumC = 0.0, sdx = 0.0, sdy = 0.0;
    for (int c = clo; c <= chi; ++c)
        for (int r = rlo; r <= rhi; ++r) {
            double w = (double)R(r - 1, c - 1);
            w = w > 0.0 ? w : 0.0;
            const float dx = (float)c - pc, dy = (float)r - pr;
            sumC += w;
            sdx += (double)(dx * dx) * w;
            sdy += (double)(dy * dy) * w;
        }
    if (sumC < 1e-9) return NAN;
    return (float)sqrt((sdx + sdy) / sumC);
}

template <int M>
__device__ void analyze_lds(const CorrConfig &cfg, const PlaneView<M> &R, int job, int gi, int gj,
                            const CorrOut &out, Cand *scand, Cand *ssel);

// load_windows! into an M x (M + 1)-pitched complex plane: frame A in .x,
// frame B in .y, zero elsewhere. Equal areas (src/correlators.jl:199-207):
// both apodized windows at the origin; an enlarged search area (:230-262):
// the apodized window at offset mg = (s - w) / 2, the search area
// mean-subtracted and not apodized. Means are block sums in T divided by
// the pixel counts. Ends with a block barrier. Returns 1, or 0 when either
// frame's loaded window is identically zero: the reference's separate
// transforms then give an exactly zero plane, while the packed A + iB
// transform leaks round-off of the other frame into it, so the caller
// scales the plane by the returned factor.
template <int M>
__device__ float load_windows_lds(const CorrConfig &cfg, const float *__restrict__ A, const float *__restrict__ B,
                                  int nr, int rs, int cs, float2 *pl, float *red)
{
#pragma clang fp contract(off)
    constexpr int P = M + 1;
    const int tid = threadIdx.x;
    const int w = cfg.w, sa_ = cfg.s, mg = (sa_ - w) / 2;
    const float *Ab = A + (size_t)(cs - 1) * nr + (rs - 1);
    // frame B: the search area, margins mg around the window (mg = 0: the window)
    const float *Bb = B + (size_t)(cs - 1 - mg) * nr + (rs - 1 - mg);
    float sa = 0.f, sb = 0.f;
    for (int e = tid; e < w * w; e += kThreads) {
        const int i = e % w, j = e / w;
        sa += Ab[i + (size_t)j * nr];
    }
    for (int e = tid; e < sa_ * sa_; e += kThreads) {
        const int i = e % sa_, j = e / sa_;
        sb += Bb[i + (size_t)j * nr];
    }
    const float2 s = block_sum2(sa, sb, red);
    const float meanA = s.x / (float)(w * w), meanB = s.y / (float)(sa_ * sa_);

    for (int e = tid; e < M * P; e += kThreads) pl[e] = make_float2(0.f, 0.f);
    __syncthreads();
    bool nza = false, nzb = false;
    if (mg == 0) {
        for (int e = tid; e < w * w; e += kThreads) {
            const int i = e % w, j = e / w;
            const float a = cfg.apod[e];
            const float2 v = make_float2(a * (Ab[i + (size_t)j * nr] - meanA), a * (Bb[i + (size_t)j * nr] - meanB));
            nza |= v.x != 0.f;
            nzb |= v.y != 0.f;
            pl[i + j * P] = v;
        }
    } else {
        for (int e = tid; e < sa_ * sa_; e += kThreads) {
            const int i = e % sa_, j = e / sa_;
            const int iw = i - mg, jw = j - mg;
            const bool inw = iw >= 0 && iw < w && jw >= 0 && jw < w;
            const float xa = inw ? cfg.apod[iw + jw * w] * (Ab[iw + (size_t)jw * nr] - meanA) : 0.f;
            const float xb = Bb[i + (size_t)j * nr] - meanB;
            nza |= xa != 0.f;
            nzb |= xb != 0.f;
            pl[i + j * P] = make_float2(xa, xb);
        }
    }
    const int anya = __syncthreads_or(nza);
    const int anyb = __syncthreads_or(nzb);
    return anya && anyb ? 1.f : 0.f;
}

template <int M>
__global__ __launch_bounds__(kThreads) void k_corr_lds(CorrConfig cfg, const float2 *__restrict__ tw,
                                                        const float *__restrict__ A, const float *__restrict__ B,
                                                        int nr, const int32_t *__restrict__ jobs, CorrOut out)
{
#pragma clang fp contract(off)
    constexpr int P = M + 1;
    constexpr int H = M / 2;
    __shared__ float2 pl[M * P];
    __shared__ float2 stw[M];
    __shared__ float red[2 * (kThreads / 64)];
    __shared__ Cand scand[kThreads / 64];
    __shared__ Cand ssel[kMaxPeaks];

    const int tid = threadIdx.x;
    const int job = blockIdx.x;
    const int gi = jobs[4 * job], gj = jobs[4 * job + 1], rs = jobs[4 * job + 2], cs = jobs[4 * job + 3];
    const int sa_ = cfg.s, mg = (sa_ - cfg.w) / 2;

    for (int k = tid; k < M; k += kThreads) stw[k] = tw[k];
    const float keep = load_windows_lds<M>(cfg, A, B, nr, rs, cs, pl, red);

    // ---- forward DIF along dim 1 (only the nonzero columns) ----
    for (int len = M; len >= 2; len >>= 1) {
        const int half = len >> 1, step = M / len;
        for (int b = tid; b < sa_ * H; b += kThreads) {
            const int j = b / H, t = b % H;
            const int i0 = (t / half) * len + (t % half), i1 = i0 + half;
            const float2 x0 = pl[i0 + j * P], x1 = pl[i1 + j * P];
            pl[i0 + j * P] = cadd(x0, x1);
            pl[i1 + j * P] = cmul(csub(x0, x1), stw[(t % half) * step]);
        }
        __syncthreads();
    }
    // ---- forward DIF along dim 2 (all rows) ----
    for (int len = M; len >= 2; len >>= 1) {
        const int half = len >> 1, step = M / len;
        for (int b = tid; b < M * H; b += kThreads) {
            const int r = b % M, t = b / M;
            const int j0 = (t / half) * len + (t % half), j1 = j0 + half;
            const float2 x0 = pl[r + j0 * P], x1 = pl[r + j1 * P];
            pl[r + j0 * P] = cadd(x0, x1);
            pl[r + j1 * P] = cmul(csub(x0, x1), stw[(t % half) * step]);
        }
        __syncthreads();
    }

    // ---- Hermitian separation + cross/phase spectrum, in place ----
    for (int e = tid; e < M * M; e += kThreads) {
        const int p1 = e % M, p2 = e / M;
        const int k1 = bitrev<M>(p1), k2 = bitrev<M>(p2);
        const int q1 = bitrev<M>((M - k1) & (M - 1)), q2 = bitrev<M>((M - k2) & (M - 1));
        const int e2 = q1 + q2 * M;
        if (e2 < e) continue;
        const float2 Z = pl[p1 + p2 * P], Zm = pl[q1 + q2 * P];
        // F{A} = (Z + conj(Zm))/2, F{B} = (Z - conj(Zm))/(2i)
        const float2 FA = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
        const float2 FB = make_float2(0.5f * (Z.y + Zm.y), -0.5f * (Z.x - Zm.x));
        float2 Pk = cmulc(FB, FA);  // conj(FA) * FB
        if (cfg.method == HH_PHASE) {
            const float den = hypotf(Pk.x, Pk.y) + FLT_EPSILON;
            const int f1 = k1, f2 = k2;
            const float wv = cfg.wfilt[f1 + f2 * M];
            Pk = make_float2(wv * Pk.x / den, wv * Pk.y / den);
        }
        pl[p1 + p2 * P] = Pk;
        if (e2 != e) pl[q1 + q2 * P] = make_float2(Pk.x, -Pk.y);
    }
    __syncthreads();

    // ---- inverse DIT along dim 2 (bit-reversed in, natural out) ----
    for (int len = 2; len <= M; len <<= 1) {
        const int half = len >> 1, step = M / len;
        for (int b = tid; b < M * H; b += kThreads) {
            const int r = b % M, t = b / M;
            const int j0 = (t / half) * len + (t % half), j1 = j0 + half;
            const float2 x0 = pl[r + j0 * P];
            const float2 x1 = cmulc(pl[r + j1 * P], stw[(t % half) * step]);
            pl[r + j0 * P] = cadd(x0, x1);
            pl[r + j1 * P] = csub(x0, x1);
        }
        __syncthreads();
    }
    // ---- inverse DIT along dim 1 ----
    for (int len = 2; len <= M; len <<= 1) {
        const int half = len >> 1, step = M / len;
        for (int b = tid; b < M * H; b += kThreads) {
            const int j = b / H, t = b % H;
            const int i0 = (t / half) * len + (t % half), i1 = i0 + half;
            const float2 x0 = pl[i0 + j * P];
            const float2 x1 = cmulc(pl[i1 + j * P], stw[(t % half) * step]);
            pl[i0 + j * P] = cadd(x0, x1);
            pl[i1 + j * P] = csub(x0, x1);
        }
        __syncthreads();
    }

    // ---- fftshift |.| * gain into the .y lanes (the .x lanes stay live);
    // restrict_search_lags! (src/correlators.jl:347-355): with a search
    // margin only lags |l| <= mg survive ----
    const float scale = 1.0f / (float)(M * M);
    for (int e = tid; e < M * M; e += kThreads) {
        const int i = e % M, j = e / M;
        const int ip = (i + H) & (M - 1), jp = (j + H) & (M - 1);
        float val = fabsf(pl[i + j * P].x * scale) * keep;
        if (cfg.gain) val *= cfg.gain[ip + jp * M];
        if (mg > 0 && (abs(ip - H) > mg || abs(jp - H) > mg)) val = 0.f;
        pl[ip + jp * P].y = val;
    }
    __syncthreads();
    PlaneView<M> R{pl};

    if (out.planes) {
        float *dst = out.planes + (size_t)job * M * M;
        for (int e = tid; e < M * M; e += kThreads) dst[e] = R(e % M, e / M);
    }
    if (out.acc) {
        // accumulate_planes!: add the plane to the window's FP64 sum
        double *dst = out.acc + (size_t)job * M * M;
        for (int e = tid; e < M * M; e += kThreads) dst[e] += (double)R(e % M, e / M);
        return;
    }
    analyze_lds<M>(cfg, R, job, gi, gj, out, scand, ssel);
}

// analyze_plane! (src/correlators.jl:459-468) of the plane in LDS and the
// per-window outputs of process_windows! (src/pipeline.jl:972-1027); all
// threads of the block take part.
template <int M>
__device__ void analyze_lds(const CorrConfig &cfg, const PlaneView<M> &R, int job, int gi, int gj,
                            const CorrOut &out, Cand *scand, Cand *ssel)
{
#pragma clang fp contract(off)
    constexpr int H = M / 2;
    const int tid = threadIdx.x;

    // ---- peak finding ----
    const int K = cfg.K;
    int nsel = 0;
    if (cfg.finder == HH_REGIONALMAX) {
        // per-lane top-K of local maxima in scan order
        Cand loc[kMaxPeaks];
        int nl = 0;
        for (int e = tid; e < M * M; e += kThreads) {
            const int r = e % M, c = e / M;
            if (!is_local_max<M>(R, r, c)) continue;
            Cand cd{R(r, c), e};
            if (nl == K && !better(cd, loc[K - 1])) continue;
            int pos = nl < K ? nl : K - 1;
            while (pos > 0 && better(cd, loc[pos - 1])) {
                loc[pos] = loc[pos - 1];
                --pos;
            }
            loc[pos] = cd;
            if (nl < K) ++nl;
        }
        int head = 0;
        for (int m = 0; m < K; ++m) {
            Cand mine = head < nl ? loc[head] : Cand{0.f, INT_MAX};
            const Cand b = block_best(mine, scand);
            if (b.idx == INT_MAX) break;
            if (mine.idx == b.idx) ++head;
            if (tid == 0) ssel[m] = b;
            ++nsel;
        }
    } else {
        // exclusion finder: K scans, excluding +/-2 boxes around earlier peaks
        for (int m = 0; m < K; ++m) {
            Cand best{-INFINITY, INT_MAX};
            for (int e = tid; e < M * M; e += kThreads) {
                const int r = e % M, c = e / M;
                bool ex = false;
                for (int q = 0; q < m; ++q) {
                    const int qi = ssel[q].idx % M, qj = ssel[q].idx / M;
                    if (abs(r - qi) <= 2 && abs(c - qj) <= 2) {
                        ex = true;
                        break;
                    }
                }
                if (ex) continue;
                const float v = R(r, c);
                if (best.idx == INT_MAX ? (v > -INFINITY) : (v > best.v)) best = Cand{v, e};
            }
            const Cand b = block_best(best, scand);
            if (b.idx == INT_MAX) break;
            if (m > 0 && b.v <= 0.f) break;
            if (tid == 0) ssel[m] = b;
            __syncthreads();
            ++nsel;
        }
    }
    __syncthreads();

    if (tid != 0) return;
    int found = 0;
    if (nsel > 0) {
        found = 1;
        if (cfg.finder == HH_REGIONALMAX)
            while (found < nsel && ssel[found].v > 0.f) ++found;
        else
            found = nsel;
    }
    const size_t g = (size_t)(gi - 1) + (size_t)(gj - 1) * out.ny;
    const size_t ng = (size_t)out.ny * out.nx;
    if (out.found) out.found[job] = found;
    if (out.locs)
        for (int m = 0; m < K; ++m) {
            out.locs[2 * (K * job + m)] = m < found ? ssel[m].idx % M + 1 : 0;
            out.locs[2 * (K * job + m) + 1] = m < found ? ssel[m].idx / M + 1 : 0;
            out.vals[K * job + m] = m < found ? ssel[m].v : NAN;
        }
    const float u0 = out.u0 ? out.u0[g] : 0.f, v0 = out.v0 ? out.v0[g] : 0.f;
    const int cc = H + 1;
    float du, dv, ratio, mom;
    if (found == 0) {
        du = dv = ratio = mom = NAN;
    } else {
        const int pr = ssel[0].idx % M + 1, pc = ssel[0].idx / M + 1;
        float2 ref;
        if (cfg.subpixel == HH_GAUSS9)
            ref = gauss9<M>(R, pr, pc);
        else if (cfg.subpixel == HH_SUBPIX_NONE)
            ref = make_float2((float)pr, (float)pc);
        else
            ref = gauss3<M>(R, pr, pc);
        du = ref.y - (float)cc;
        dv = ref.x - (float)cc;
        const float v1 = ssel[0].v;
        ratio = found >= 2 ? v1 / ssel[1].v : (v1 > 0.f ? INFINITY : NAN);
        mom = moment<M>(R, ref.x, ref.y);
    }
    if (out.u0) {
        out.u[g] = u0 + du;
        out.v[g] = v0 + dv;
    } else {
        out.u[g] = du;
        out.v[g] = dv;
    }
    out.pr[g] = ratio;
    out.cm[g] = mom;
    if (out.alt_u) {
        for (int m = 2; m <= cfg.n_peaks; ++m) {
            float au = NAN, av = NAN;
            if (m <= found) {
                const float2 a = gauss3<M>(R, ssel[m - 1].idx % M + 1, ssel[m - 1].idx / M + 1);
                au = a.y - (float)cc;
                av = a.x - (float)cc;
                if (out.u0) {
                    au = u0 + au;
                    av = v0 + av;
                }
            } else if (!out.u0) {
                continue;  // hook semantics: absent alternatives are left to the caller's NaN fill
            }
            out.alt_u[g + (size_t)(m - 2) * ng] = au;
            out.alt_v[g + (size_t)(m - 2) * ng] = av;
        }
    }
}


// ---- direct-DFT correlation for FFT sides that are not powers of two ----
// Search areas such as 24 or 48 px (M = s, or 2s when padded) have no radix-2
// transform. The window load, spectrum, magnitude, gain, lag restriction and
// analysis are those of k_corr_lds; each 1D transform is a direct O(M^2)
// DFT with the twiddle W^(k n mod M) of a table of M entries. Outputs of a
// pass are accumulated in registers (at most ceil(M^2 / 256) per thread),
// then written back after a barrier, so one LDS plane suffices. Only the
// nonzero input lines are transformed forward, and only the output lines
// whose lags survive restrict_search_lags! inversely.

// Lags kept along one axis, as plane indices before the fftshift: all M
// when mg = 0, else 0..mg and M - mg..M - 1 (2 mg + 1 of them).
__device__ __forceinline__ int kept_count(int M, int mg) { return mg == 0 ? M : 2 * mg + 1; }
__device__ __forceinline__ int kept_index(int M, int mg, int t) { return (mg == 0 || t <= mg) ? t : M - (2 * mg + 1) + t; }

// One DFT pass over the plane. Along dim 1 (DIM = 1) output (k, line) sums
// input rows n < nin of column `line`; along dim 2, output (line, k) sums
// input columns n < nin of row `line`. Output lines: nl of them, line index
// from lmap; outputs per line: nk, output index from kmap. INV: e^{+...}.
template <int M, int DIM, bool INV, typename LMap, typename KMap>
__device__ __forceinline__ void dft_pass(float2 *pl, const float2 *tw, int nin, int nl, int nk, const LMap &lmap,
                                         const KMap &kmap)
{
#pragma clang fp contract(off)
    constexpr int P = M + 1;
    constexpr int NO = (M * M + kThreads - 1) / kThreads;
    const int tid = threadIdx.x;
    const int nout = nl * nk;
    float2 acc[NO];
#pragma unroll
    for (int q = 0; q < NO; ++q) {
        const int e = tid + q * kThreads;
        // four interleaved partial sums of fused multiply-adds: the rounding
        // error grows with nin / 4 terms instead of nin
        float2 part[4] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f),
                          make_float2(0.f, 0.f)};
        if (e < nout) {
            const int k = kmap(e % nk), line = lmap(e / nk);
            int idx = 0;  // k n mod M
            for (int n = 0; n < nin; ++n) {
                const float2 x = DIM == 1 ? pl[n + line * P] : pl[line + n * P];
                const float2 t = tw[idx];
                const float ti = INV ? -t.y : t.y;
                float2 &sm = part[n & 3];
                sm.x = __builtin_fmaf(x.x, t.x, __builtin_fmaf(-x.y, ti, sm.x));
                sm.y = __builtin_fmaf(x.x, ti, __builtin_fmaf(x.y, t.x, sm.y));
                idx += k;
                if (idx >= M) idx -= M;
            }
        }
        acc[q] = make_float2((part[0].x + part[1].x) + (part[2].x + part[3].x),
                             (part[0].y + part[1].y) + (part[2].y + part[3].y));
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NO; ++q) {
        const int e = tid + q * kThreads;
        if (e < nout) {
            const int k = kmap(e % nk), line = lmap(e / nk);
            if (DIM == 1)
                pl[k + line * P] = acc[q];
            else
                pl[line + k * P] = acc[q];
        }
    }
    __syncthreads();
}

template <int M>
__global__ __launch_bounds__(kThreads) void k_corr_dft(CorrConfig cfg, const float2 *__restrict__ tw,
                                                        const float *__restrict__ A, const float *__restrict__ B,
                                                        int nr, const int32_t *__restrict__ jobs, CorrOut out)
{
#pragma clang fp contract(off)
    constexpr int P = M + 1;
    constexpr int H = M / 2;
    __shared__ float2 pl[M * P];
    __shared__ float2 stw[M];
    __shared__ float red[2 * (kThreads / 64)];
    __shared__ Cand scand[kThreads / 64];
    __shared__ Cand ssel[kMaxPeaks];

    const int tid = threadIdx.x;
    const int job = blockIdx.x;
    const int gi = jobs[4 * job], gj = jobs[4 * job + 1], rs = jobs[4 * job + 2], cs = jobs[4 * job + 3];
    const int S = cfg.s, mg = (S - cfg.w) / 2;

    for (int k = tid; k < M; k += kThreads) stw[k] = tw[k];
    const float keep = load_windows_lds<M>(cfg, A, B, nr, rs, cs, pl, red);

    const auto ident = [](int t) { return t; };
    const int nkept = kept_count(M, mg);
    const auto kept = [&](int t) { return kept_index(M, mg, t); };
    // forward: dim 1 over the S nonzero columns (rows n < S), then dim 2
    // over all rows (columns n < S)
    dft_pass<M, 1, false>(pl, stw, S, S, M, ident, ident);
    dft_pass<M, 2, false>(pl, stw, S, M, M, ident, ident);

    // ---- Hermitian separation + cross/phase spectrum, natural order ----
    for (int e = tid; e < M * M; e += kThreads) {
        const int k1 = e % M, k2 = e / M;
        const int q1 = (M - k1) % M, q2 = (M - k2) % M;
        const int e2 = q1 + q2 * M;
        if (e2 < e) continue;
        const float2 Z = pl[k1 + k2 * P], Zm = pl[q1 + q2 * P];
        const float2 FA = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
        const float2 FB = make_float2(0.5f * (Z.y + Zm.y), -0.5f * (Z.x - Zm.x));
        float2 Pk = cmulc(FB, FA);  // conj(FA) * FB
        if (cfg.method == HH_PHASE) {
            const float den = hypotf(Pk.x, Pk.y) + FLT_EPSILON;
            const float wv = cfg.wfilt[k1 + k2 * M];
            Pk = make_float2(wv * Pk.x / den, wv * Pk.y / den);
        }
        pl[k1 + k2 * P] = Pk;
        if (e2 != e) pl[q1 + q2 * P] = make_float2(Pk.x, -Pk.y);
    }
    __syncthreads();

    // inverse: dim 2 into the kept columns (all rows), then dim 1 into the
    // kept rows of the kept columns
    dft_pass<M, 2, true>(pl, stw, M, M, nkept, ident, kept);
    dft_pass<M, 1, true>(pl, stw, M, nkept, nkept, kept, kept);

    // ---- fftshift |.| * gain into the .y lanes; lags beyond the margin 0 ----
    const float scale = 1.0f / (float)(M * M);
    for (int e = tid; e < M * M; e += kThreads) {
        const int i = e % M, j = e / M;
        const int ip = (i + H) % M, jp = (j + H) % M;
        float val = 0.f;
        if (mg == 0 || (abs(ip - H) <= mg && abs(jp - H) <= mg)) {
            val = fabsf(pl[i + j * P].x * scale) * keep;
            if (cfg.gain) val *= cfg.gain[ip + jp * M];
        }
        pl[ip + jp * P].y = val;
    }
    __syncthreads();
    PlaneView<M> R{pl};

    if (out.planes) {
        float *dst = out.planes + (size_t)job * M * M;
        for (int e = tid; e < M * M; e += kThreads) dst[e] = R(e % M, e / M);
    }
    if (out.acc) {
        double *dst = out.acc + (size_t)job * M * M;
        for (int e = tid; e < M * M; e += kThreads) dst[e] += (double)R(e % M, e / M);
        return;
    }
    analyze_lds<M>(cfg, R, job, gi, gj, out, scand, ssel);
}

// ensemble_analyze! (src/ensemble.jl:284-313): one block per window; the
// FP64 plane sum is rounded to float and analyzed like a single plane.
template <int M>
__global__ __launch_bounds__(kThreads) void k_plane_analyze(CorrConfig cfg, const double *__restrict__ acc, int box,
                                                            const int32_t *__restrict__ jobs, CorrOut out)
{
    constexpr int P = M + 1;
    __shared__ float2 pl[M * P];
    __shared__ Cand scand[kThreads / 64];
    __shared__ Cand ssel[kMaxPeaks];
    const int tid = threadIdx.x;
    const int job = blockIdx.x;
    const int gi = jobs[4 * job], gj = jobs[4 * job + 1];
    const int lo = cfg.box_lo, nb = cfg.box_hi - cfg.box_lo + 1;
    const double *src = acc + (size_t)job * (box ? nb * nb : M * M);
    for (int e = tid; e < M * M; e += kThreads) {
        const int i = e % M, j = e / M;
        float val;
        if (box) {
            const int bi = i - lo, bj = j - lo;
            val = (bi >= 0 && bi < nb && bj >= 0 && bj < nb) ? (float)src[bi + bj * nb] : 0.f;
        } else {
            val = (float)src[e];
        }
        pl[i + j * P] = make_float2(0.f, val);
    }
    __syncthreads();
    PlaneView<M> R{pl};
    analyze_lds<M>(cfg, R, job, gi, gj, out, scand, ssel);
}

}  // namespace

static float2 *twiddle_table(int M, hipStream_t s)
{
    // e^{-2 pi i k / M}, k < M, built once per size in double precision
    static std::mutex mu;
    static float2 *tabs[kMaxM + 1] = {nullptr};
    std::lock_guard<std::mutex> lock(mu);
    if (!tabs[M]) {
        std::vector<float2> h(M);
        for (int k = 0; k < M; ++k) {
            const double a = -2.0 * M_PI * (double)k / (double)M;
            h[k] = make_float2((float)cos(a), (float)sin(a));
        }
        HH_HIP(hipMalloc(&tabs[M], sizeof(float2) * M));
        HH_HIP(hipMemcpyAsync(tabs[M], h.data(), sizeof(float2) * M, hipMemcpyHostToDevice, s));
        HH_HIP(hipStreamSynchronize(s));
    }
    return tabs[M];
}

const char *fast_family(const CorrConfig &cfg);
bool mfma_covers(const CorrConfig &cfg);
bool launch_uq16_mfma(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs, int njobs,
                      const CorrOut &out, hipStream_t s);

const char *correlate_family(const CorrConfig &cfg, bool uq)
{
    if (cfg.family == 0 && cfg.form == HH_FORM_MFMA16 && mfma_covers(cfg) && !(uq && cfg.w != 16))
        return cfg.M == 64 ? "mfma64" : (uq ? "mfma32+uq" : "mfma32");
    if (cfg.family == 0 && cfg.form == HH_FORM_MFMA16 && mfma96_covers(cfg) && !uq) return "mfma96";
    const char *f = cfg.family == 2 ? nullptr : fast_family(cfg);
    return f ? f : (cfg.M == 24 || cfg.M == 48 || cfg.M == 96 ? "dft" : "lds");
}

void correlate_prepare(const CorrConfig &cfg, hipStream_t s)
{
    (void)device_cu_count();
    if (cfg.form == HH_FORM_MFMA16 && mfma96_covers(cfg)) mfma96_prepare(s);
    if (cfg.M >= 2 && cfg.M <= kMaxM) (void)twiddle_table(cfg.M, s);
}

bool launch_correlate(const CorrConfig &cfg, const float *A, const float *B, int nr, int nc, const int32_t *jobs,
                      int njobs, const CorrOut &out, hipStream_t s, const SideStream *side)
{
    (void)nc;
    if (njobs <= 0) return out.unc_u != nullptr;
    bool uq_done = false;
    const bool mfma_ok = cfg.family == 0 && cfg.form == HH_FORM_MFMA16, fast_ok = cfg.family != 2;
    if (mfma_ok && out.unc_u && out.unc_v && !out.planes && !out.acc && mfma_covers(cfg) && cfg.w == 16) {
        // the sweep with uncertainty: matrix-core correlation, then the
        // statistics of the same windows with their covariance sums on the
        // matrix cores
        // matrix cores (the statistics kernel is latency-bound at its
        // occupancy, the correlation VALU-bound: on a side stream they
        // overlap); both only read the images and write disjoint outputs
        CorrOut o = out;
        o.unc_u = o.unc_v = nullptr;
        if (side) {
            HH_HIP(hipEventRecord(side->fork, s));
            HH_HIP(hipStreamWaitEvent(side->stream, side->fork, 0));
            const bool ok = launch_correlate_mfma(cfg, A, B, nr, jobs, njobs, o, s) &&
                            launch_uq16_mfma(cfg, A, B, nr, jobs, njobs, out, side->stream);
            HH_HIP(hipEventRecord(side->join, side->stream));
            if (!side->defer_join) HH_HIP(hipStreamWaitEvent(s, side->join, 0));
            if (ok) return true;
        } else if (launch_correlate_mfma(cfg, A, B, nr, jobs, njobs, o, s) &&
                   launch_uq16_mfma(cfg, A, B, nr, jobs, njobs, out, s)) {
            return true;
        }
    }
    if (mfma_ok && launch_correlate_mfma(cfg, A, B, nr, jobs, njobs, out, s)) return false;
    if (mfma_ok && launch_correlate_mfma96(cfg, A, B, nr, jobs, njobs, out, s)) return false;
    if (fast_ok && launch_correlate_fast(cfg, A, B, nr, jobs, njobs, out, s, &uq_done)) return uq_done;
    const float2 *tw = twiddle_table(cfg.M, s);
    switch (cfg.M) {
    case 16: hipLaunchKernelGGL(k_corr_lds<16>, dim3(njobs), dim3(kThreads), 0, s, cfg, tw, A, B, nr, jobs, out); break;
    case 32: hipLaunchKernelGGL(k_corr_lds<32>, dim3(njobs), dim3(kThreads), 0, s, cfg, tw, A, B, nr, jobs, out); break;
    case 64: hipLaunchKernelGGL(k_corr_lds<64>, dim3(njobs), dim3(kThreads), 0, s, cfg, tw, A, B, nr, jobs, out); break;
    case 128: hipLaunchKernelGGL(k_corr_lds<128>, dim3(njobs), dim3(kThreads), 0, s, cfg, tw, A, B, nr, jobs, out); break;
    case 24: hipLaunchKernelGGL(k_corr_dft<24>, dim3(njobs), dim3(kThreads), 0, s, cfg, tw, A, B, nr, jobs, out); break;
    case 48: hipLaunchKernelGGL(k_corr_dft<48>, dim3(njobs), dim3(kThreads), 0, s, cfg, tw, A, B, nr, jobs, out); break;
    case 96: hipLaunchKernelGGL(k_corr_dft<96>, dim3(njobs), dim3(kThreads), 0, s, cfg, tw, A, B, nr, jobs, out); break;
    default: throw Error(HH_EUNSUPPORTED, "FFT size must be 16, 24, 32, 48, 64, 96 or 128");
    }
    HH_HIP(hipGetLastError());
    return false;
}

void launch_plane_analyze(const CorrConfig &cfg, const double *acc, bool box, const int32_t *jobs, int njobs,
                          const CorrOut &out, hipStream_t s)
{
    if (njobs <= 0) return;
    const int b = box ? 1 : 0;
    switch (cfg.M) {
    case 16: hipLaunchKernelGGL(k_plane_analyze<16>, dim3(njobs), dim3(kThreads), 0, s, cfg, acc, b, jobs, out); break;
    case 32: hipLaunchKernelGGL(k_plane_analyze<32>, dim3(njobs), dim3(kThreads), 0, s, cfg, acc, b, jobs, out); break;
    case 64: hipLaunchKernelGGL(k_plane_analyze<64>, dim3(njobs), dim3(kThreads), 0, s, cfg, acc, b, jobs, out); break;
    case 128:
        hipLaunchKernelGGL(k_plane_analyze<128>, dim3(njobs), dim3(kThreads), 0, s, cfg, acc, b, jobs, out);
        break;
    case 24: hipLaunchKernelGGL(k_plane_analyze<24>, dim3(njobs), dim3(kThreads), 0, s, cfg, acc, b, jobs, out); break;
    case 48: hipLaunchKernelGGL(k_plane_analyze<48>, dim3(njobs), dim3(kThreads), 0, s, cfg, acc, b, jobs, out); break;
    case 96: hipLaunchKernelGGL(k_plane_analyze<96>, dim3(njobs), dim3(kThreads), 0, s, cfg, acc, b, jobs, out); break;
    default: throw Error(HH_EUNSUPPORTED, "FFT size must be 16, 24, 32, 48, 64, 96 or 128");
    }
    HH_HIP(hipGetLastError());
}

}  // namespace hh
