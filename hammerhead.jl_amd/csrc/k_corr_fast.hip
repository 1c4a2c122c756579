// k_corr_fast.hip — register-resident FFT correlation for 16-px windows
// padded to 32 x 32 (the dominant pass of a 64 -> 32 -> 16 schedule).
//
// Same per-window semantics as k_corr_lds (correlation_plane! +
// analyze_plane!, src/correlators.jl:326-341,459-468), restructured for
// gfx950 wave64: 16 lanes per window, 4 windows per wave, every line of the
// 2D transform held in registers with compile-time twiddles:
//   1. lane i loads window row i of A and B (coalesced), window means by
//      cross-lane sums, apodized A + iB packed as one complex row;
//   2. 32-point row DFT of 16 nonzero inputs (two 16-point FFTs); the
//      Hermitian split gives F_row{A}, F_row{B} for k2 = 0..16 (the other
//      half is conjugate-symmetric);
//   3. transpose through LDS; lane u transforms column k2 = u (lane 0 packs
//      the real columns k2 = 0 and 16), forms conj(F{A}) F{B} (or the
//      filtered phase spectrum) and inverse-transforms along dim 1;
//   4. only the rows and columns of the overlap-gain support box are kept:
//      outside it the plane is exactly zero (gain = 0, src/correlators.jl:
//      170-181), so the final complex-to-real inverse along dim 2 runs on
//      the box rows, two rows per complex transform;
//   5. regional-max top-K over the box, plus the plane corner (M, M), the
//      one zero-valued point that is always a local maximum of a plane
//      that vanishes outside the box — reproducing find_peaks_regionalmax!
//      on the full plane exactly; then gauss3/gauss9, ratio, moment and
//      alternatives on lane 0 of the window.
// Blocks are remapped so that each XCD processes a contiguous range of
// windows: overlapping windows then share one L2.

#include "hh_internal.h"

#include <cfloat>
#include <climits>
#include <utility>

namespace hh {

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double c_sin(double x)
{
    double term = x, sum = x;
    for (int n = 1; n < 32; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double c_cos(double x)
{
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 32; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
// 2 pi t / n reduced to [-pi, pi]
constexpr double angle(int t, int n)
{
    t %= n;
    if (2 * t > n) t -= n;
    return 2.0 * kPi * (double)t / (double)n;
}
constexpr int cbitrev(int x, int n)
{
    int r = 0;
    for (int b = 1; b < n; b <<= 1) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

template <typename F, int... I>
__device__ __forceinline__ void sfor_(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F &&f)
{
    sfor_(f, std::make_integer_sequence<int, N>{});
}

// b * e^{-+2 pi i T / N} (INV: +)
template <int N, int T, bool INV>
__device__ __forceinline__ float2 twmul(float2 b)
{
    constexpr int t = ((T % N) + N) % N;
    if constexpr (t == 0) {
        return b;
    } else if constexpr (4 * t == N) {
        return INV ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
    } else if constexpr (2 * t == N) {
        return make_float2(-b.x, -b.y);
    } else if constexpr (4 * t == 3 * N) {
        return INV ? make_float2(b.y, -b.x) : make_float2(-b.y, b.x);
    } else {
        constexpr float c = (float)c_cos(angle(t, N));
        constexpr float s = (float)(INV ? c_sin(angle(t, N)) : -c_sin(angle(t, N)));
        return make_float2(b.x * c - b.y * s, b.x * s + b.y * c);
    }
}

template <int N, bool INV, int LEN>
__device__ __forceinline__ void fft_stage(float2 (&y)[N])
{
    if constexpr (LEN <= N) {
        constexpr int H = LEN / 2;
        sfor<N / LEN>([&](auto blk) {
            sfor<H>([&](auto k) {
                constexpr int i0 = decltype(blk)::value * LEN + decltype(k)::value;
                constexpr int T = decltype(k)::value * (N / LEN);
                const float2 a = y[i0];
                const float2 b = twmul<N, T, INV>(y[i0 + H]);
                y[i0] = make_float2(a.x + b.x, a.y + b.y);
                y[i0 + H] = make_float2(a.x - b.x, a.y - b.y);
            });
        });
        fft_stage<N, INV, LEN * 2>(y);
    }
}

// In-register N-point DFT (radix-2 DIT), unnormalized; INV: e^{+...}.
template <int N, bool INV>
__device__ __forceinline__ void fft(float2 (&x)[N])
{
    float2 y[N];
    sfor<N>([&](auto i) {
        constexpr int src = cbitrev(decltype(i)::value, N);
        y[decltype(i)::value] = x[src];
    });
    fft_stage<N, INV, 2>(y);
    sfor<N>([&](auto i) { x[decltype(i)::value] = y[decltype(i)::value]; });
}

// 32-point DFT of 16 nonzero inputs (x[16..31] = 0): even outputs are the
// 16-point DFT of x, odd outputs the 16-point DFT of x_j e^{-2 pi i j/32}.
__device__ __forceinline__ void dft32_half(const float2 (&x)[16], float2 (&X)[32])
{
    float2 e[16], o[16];
    sfor<16>([&](auto j) {
        constexpr int J = decltype(j)::value;
        e[J] = x[J];
        o[J] = twmul<32, J, false>(x[J]);
    });
    fft<16, false>(e);
    fft<16, false>(o);
    sfor<16>([&](auto m) {
        constexpr int Mi = decltype(m)::value;
        X[2 * Mi] = e[Mi];
        X[2 * Mi + 1] = o[Mi];
    });
}

__device__ __forceinline__ float2 conj_mul(float2 a, float2 b)  // conj(a) * b
{
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ float2 phase_norm(float2 p, float w)
{
    const float den = hypotf(p.x, p.y) + FLT_EPSILON;
    return make_float2(w * p.x / den, w * p.y / den);
}

struct Cand {
    float v;
    int lin;  // column-major plane index, INT_MAX = none
};
__device__ __forceinline__ bool better(float av, int al, float bv, int bl)
{
    if (al == INT_MAX) return false;
    if (bl == INT_MAX) return true;
    return av > bv || (av == bv && al < bl);
}

// Box plane accessor: plane (0-based, fftshifted) -> value; 0 outside box.
template <int NB, int LO>
struct BoxView {
    const float *rb;
    __device__ __forceinline__ float operator()(int ip, int jp) const
    {
        const int bi = ip - LO, bj = jp - LO;
        return (bi >= 0 && bi < NB && bj >= 0 && bj < NB) ? rb[bi + bj * NB] : 0.f;
    }
};

template <typename V>
__device__ float2 gauss3_b(const V &R, int i, int j)  // 1-based, M = 32
{
    constexpr int M = 32;
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

template <typename V>
__device__ float2 gauss9_b(const V &R, int pr, int pc)
{
    constexpr int M = 32;
    if (!(1 < pr && pr < M && 1 < pc && pc < M)) return gauss3_b(R, pr, pc);
    float sL = 0.f, sxL = 0.f, syL = 0.f, sxxL = 0.f, syyL = 0.f, sxyL = 0.f;
    for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) {
            const float val = R(pr + di - 1, pc + dj - 1);
            if (!(val > 0.f)) return gauss3_b(R, pr, pc);
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
    if (!(a3 < 0.f && D > 0.f)) return gauss3_b(R, pr, pc);
    const float dx = (a4 * a2 - 2.f * a5 * a1) / D;
    const float dy = (a4 * a1 - 2.f * a3 * a2) / D;
    if (!(fabsf(dx) <= 1.f && fabsf(dy) <= 1.f)) return gauss3_b(R, pr, pc);
    return make_float2((float)pr + dy, (float)pc + dx);
}

template <typename V>
__device__ float moment_b(const V &R, float pr, float pc)
{
    constexpr int M = 32;
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

// HW: half-width of the analysis box (box = [16-HW, 16+HW]^2, which must
// contain the gain support plus a one-pixel zero margin).
template <int K, bool PHASE, int HW>
__global__ __launch_bounds__(64) void k_corr32(CorrConfig cfg, const float *__restrict__ A,
                                               const float *__restrict__ B, int nr,
                                               const int32_t *__restrict__ jobs, int njobs, int nblocks, CorrOut out)
{
#pragma clang fp contract(off)
    constexpr int M = 32;
    constexpr int NB = 2 * HW + 1;
    constexpr int LO = M / 2 - HW;
    constexpr int TSZ = 17 * 16;        // float4 per window for the transpose
    constexpr int NPAIR = (NB + 1) / 2;
    static_assert(NB * 17 * 2 + NB * NB <= TSZ * 4, "box does not fit the LDS slice");
    __shared__ float4 lds[4 * TSZ];

    const int lane = threadIdx.x;
    const int grp = lane >> 4;
    const int t = lane & 15;
    // XCD-aware order: hardware dispatches block b to XCD b % 8; give each XCD
    // a contiguous range of tiles
    const int b = blockIdx.x;
    const int xcd = b & 7, slot = b >> 3;
    const int q8 = nblocks >> 3, r8 = nblocks & 7;
    const int tile = xcd * q8 + min(xcd, r8) + slot;
    const int win = tile * 4 + grp;
    const bool active = win < njobs;
    const int jb = active ? win : njobs - 1;
    const int gi = jobs[4 * jb], gj = jobs[4 * jb + 1], rs = jobs[4 * jb + 2], cs = jobs[4 * jb + 3];

    // ---- 1. load row t of both windows; means; apodize; pack A + iB ----
    const float *pa = A + (size_t)(cs - 1) * nr + (rs - 1 + t);
    const float *pb = B + (size_t)(cs - 1) * nr + (rs - 1 + t);
    float ra[16], rb[16];
    float sa = 0.f, sb = 0.f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        ra[j] = pa[(size_t)j * nr];
        rb[j] = pb[(size_t)j * nr];
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        sa += ra[j];
        sb += rb[j];
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
        sa += __shfl_xor(sa, o);
        sb += __shfl_xor(sb, o);
    }
    const float meanA = sa / 256.f, meanB = sb / 256.f;
    float2 x[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float a = cfg.apod[t + 16 * j];
        x[j] = make_float2(a * (ra[j] - meanA), a * (rb[j] - meanB));
    }

    // ---- 2. row DFT (dim 2) and Hermitian split, k2 = 0..16 ----
    float4 *T = lds + grp * TSZ;
    {
        float2 Y[32];
        dft32_half(x, Y);
        sfor<17>([&](auto kk) {
            constexpr int k2 = decltype(kk)::value;
            constexpr int km = (32 - k2) & 31;
            const float2 z = Y[k2], zm = Y[km];
            T[k2 * 16 + t] = make_float4(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y), 0.5f * (z.y + zm.y),
                                         -0.5f * (z.x - zm.x));
        });
    }
    __syncthreads();

    // ---- 3. column DFT (dim 1) of column k2 = t; spectrum; inverse dim 1 ----
    float2 P[32];
    {
        float2 fa[16], fb[16];
        // lane 0 packs the real columns k2 = 0 and 16 as re + i im
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const float4 v = T[t * 16 + i], v16 = T[16 * 16 + i];
            fa[i] = t != 0 ? make_float2(v.x, v.y) : make_float2(v.x, v16.x);
            fb[i] = t != 0 ? make_float2(v.z, v.w) : make_float2(v.z, v16.z);
        }
        float2 FA[32], FB[32];
        dft32_half(fa, FA);
        dft32_half(fb, FB);
        // conjugate pairs (k1, 32 - k1) together, so FA/FB die as P fills;
        // lane 0 separates its packed real columns k2 = 0 and 16 and
        // repacks their (Hermitian-in-k1) spectra as P0 + i P16
        const bool unit0 = t == 0;
        sfor<17>([&](auto kk) {
            constexpr int k1 = decltype(kk)::value;
            constexpr int km = (32 - k1) & 31;
            const float2 ga = FA[k1], gam = FA[km], gb = FB[k1], gbm = FB[km];
            // general column k2 = t
            float2 p = conj_mul(ga, gb), pm = conj_mul(gam, gbm);
            // packed columns 0 and 16
            const float2 a0 = make_float2(0.5f * (ga.x + gam.x), 0.5f * (ga.y - gam.y));
            const float2 a16 = make_float2(0.5f * (ga.y + gam.y), -0.5f * (ga.x - gam.x));
            const float2 b0 = make_float2(0.5f * (gb.x + gbm.x), 0.5f * (gb.y - gbm.y));
            const float2 b16 = make_float2(0.5f * (gb.y + gbm.y), -0.5f * (gb.x - gbm.x));
            float2 p0 = conj_mul(a0, b0), p16 = conj_mul(a16, b16);
            if constexpr (PHASE) {
                const float w = cfg.wfilt[k1 + t * M], wm = cfg.wfilt[km + t * M];
                p = phase_norm(p, w);
                pm = phase_norm(pm, wm);
                p0 = phase_norm(p0, cfg.wfilt[k1]);
                p16 = phase_norm(p16, cfg.wfilt[k1 + 16 * M]);
            }
            // P0(km) = conj(P0(k1)), P16(km) = conj(P16(k1))
            const float2 u1 = make_float2(p0.x - p16.y, p0.y + p16.x);
            const float2 um = make_float2(p0.x + p16.y, -p0.y + p16.x);
            P[k1] = unit0 ? u1 : p;
            if constexpr (km != k1) P[km] = unit0 ? um : pm;
        });
        fft<32, true>(P);
    }
    __syncthreads();  // every lane holds its column; the LDS slice is reused

    // keep the box rows: Q(l1, k2), row q <-> l1 = q - HW (mod 32)
    float2 *Qs = reinterpret_cast<float2 *>(T);      // NB x 17
    float *Rb = reinterpret_cast<float *>(T) + NB * 17 * 2;  // NB x NB, column-major
    sfor<NB>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        constexpr int l1 = (32 - HW + q) & 31;
        Qs[q * 17 + t] = P[l1];
    });
    __syncthreads();

    // ---- 4. complex-to-real inverse along dim 2, two box rows per lane ----
    if (t < NPAIR) {
        const int qa = 2 * t, qb = 2 * t + 1;
        const bool hasb = qb < NB;
        const float2 *Qa = Qs + qa * 17;
        const float2 *Qb = Qs + (hasb ? qb : qa) * 17;
        const float zb = hasb ? 1.f : 0.f;
        float2 c[32];
        {
            const float2 a0 = Qa[0], b0 = Qb[0];  // (Q(.,0), Q(.,16)), real
            c[0] = make_float2(a0.x, zb * b0.x);
            c[16] = make_float2(a0.y, zb * b0.y);
        }
        sfor<15>([&](auto kk) {
            constexpr int k2 = decltype(kk)::value + 1;
            const float2 qa_ = Qa[k2];
            const float2 qb_ = make_float2(zb * Qb[k2].x, zb * Qb[k2].y);
            c[k2] = make_float2(qa_.x - qb_.y, qa_.y + qb_.x);           // Qa + i Qb
            c[32 - k2] = make_float2(qa_.x + qb_.y, -qa_.y + qb_.x);     // conj(Qa) + i conj(Qb)
        });
        fft<32, true>(c);
        const float scale = 1.0f / 1024.0f;
        const int ipa = LO + qa, ipb = LO + qb;
        sfor<NB>([&](auto qq) {
            constexpr int q2 = decltype(qq)::value;
            constexpr int l2 = (32 - HW + q2) & 31;
            constexpr int jp = LO + q2;
            Rb[qa + q2 * NB] = fabsf(c[l2].x * scale) * cfg.gain[ipa + jp * M];
            if (hasb) Rb[qb + q2 * NB] = fabsf(c[l2].y * scale) * cfg.gain[ipb + jp * M];
        });
    }
    __syncthreads();

    // ---- 5. regional-max top-K over the box ----
    const BoxView<NB, LO> R{Rb};
    float tv[K];
    int tl[K];
#pragma unroll
    for (int m = 0; m < K; ++m) {
        tv[m] = 0.f;
        tl[m] = INT_MAX;
    }
    for (int p = t; p < NB * NB; p += 16) {
        const int bi = p % NB, bj = p / NB;
        const float I0 = Rb[p];
        if (!(I0 > 0.f)) continue;  // zero-valued box points are never maxima
        bool ok = true;
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
            for (int di = -1; di <= 1; ++di) {
                if (di == 0 && dj == 0) continue;
                const int ni = bi + di, nj = bj + dj;
                const float In = (ni >= 0 && ni < NB && nj >= 0 && nj < NB) ? Rb[ni + nj * NB] : 0.f;
                const bool earlier = dj < 0 || (dj == 0 && di < 0);
                ok &= earlier ? !(In > I0) : !(In >= I0);
            }
        if (!ok) continue;
        const int lin = (LO + bi) + (LO + bj) * M;
        // insert (strictly better moves up; scan order is increasing lin)
        float cv = I0;
        int cl = lin;
#pragma unroll
        for (int m = 0; m < K; ++m) {
            if (better(cv, cl, tv[m], tl[m])) {
                const float sv = tv[m];
                const int sl = tl[m];
                tv[m] = cv;
                tl[m] = cl;
                cv = sv;
                cl = sl;
            }
        }
    }
    // merge the 16 lanes' lists: K rounds of group argmax
    float sv[K];
    int sl[K];
    int nsel = 0;
#pragma unroll
    for (int m = 0; m < K; ++m) {
        float bv = tv[0];
        int bl = tl[0];
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            const float ov = __shfl_xor(bv, o);
            const int ol = __shfl_xor(bl, o);
            if (better(ov, ol, bv, bl)) {
                bv = ov;
                bl = ol;
            }
        }
        sv[m] = bv;
        sl[m] = bl;
        nsel += bl != INT_MAX;
        if (tl[0] == bl && bl != INT_MAX) {  // pop the winner's head
#pragma unroll
            for (int q = 0; q + 1 < K; ++q) {
                tv[q] = tv[q + 1];
                tl[q] = tl[q + 1];
            }
            tv[K - 1] = 0.f;
            tl[K - 1] = INT_MAX;
        }
    }
    if (t != 0 || !active) return;
    // the plane corner (M, M): zero-valued, always a regional maximum
#pragma unroll
    for (int m = 0; m < K; ++m)
        if (m == nsel) {
            sv[m] = 0.f;
            sl[m] = M * M - 1;
        }
    nsel = nsel < K ? nsel + 1 : nsel;
    int found = 1;
#pragma unroll
    for (int m = 1; m < K; ++m)
        if (found == m && m < nsel && sv[m] > 0.f) found = m + 1;

    const size_t g = (size_t)(gi - 1) + (size_t)(gj - 1) * out.ny;
    const size_t ng = (size_t)out.ny * out.nx;
    if (out.found) out.found[jb] = found;
    if (out.locs) {
#pragma unroll
        for (int m = 0; m < K; ++m) {
            out.locs[2 * (K * jb + m)] = m < found ? sl[m] % M + 1 : 0;
            out.locs[2 * (K * jb + m) + 1] = m < found ? sl[m] / M + 1 : 0;
            out.vals[K * jb + m] = m < found ? sv[m] : NAN;
        }
    }
    const int pr = sl[0] % M + 1, pc = sl[0] / M + 1;
    float2 ref;
    if (cfg.subpixel == HH_GAUSS9)
        ref = gauss9_b(R, pr, pc);
    else if (cfg.subpixel == HH_SUBPIX_NONE)
        ref = make_float2((float)pr, (float)pc);
    else
        ref = gauss3_b(R, pr, pc);
    const float cc = (float)(M / 2 + 1);
    const float du = ref.y - cc, dv = ref.x - cc;
    const float v1 = sv[0];
    const float ratio = found >= 2 ? v1 / sv[1] : (v1 > 0.f ? INFINITY : NAN);
    const float mom = moment_b(R, ref.x, ref.y);
    const float u0 = out.u0 ? out.u0[g] : 0.f, v0 = out.v0 ? out.v0[g] : 0.f;
    out.u[g] = out.u0 ? u0 + du : du;
    out.v[g] = out.v0 ? v0 + dv : dv;
    out.pr[g] = ratio;
    out.cm[g] = mom;
    if (out.alt_u) {
#pragma unroll
        for (int m = 2; m <= K; ++m) {
            if (m > cfg.n_peaks) continue;
            float au = NAN, av = NAN;
            if (m <= found) {
                const float2 a = gauss3_b(R, sl[m - 1] % M + 1, sl[m - 1] / M + 1);
                au = a.y - cc;
                av = a.x - cc;
                if (out.u0) {
                    au = u0 + au;
                    av = v0 + av;
                }
            } else if (!out.u0) {
                continue;
            }
            out.alt_u[g + (size_t)(m - 2) * ng] = au;
            out.alt_v[g + (size_t)(m - 2) * ng] = av;
        }
    }
}

template <int K, bool PHASE, int HW>
void launch32(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs, int njobs,
              const CorrOut &out, hipStream_t s)
{
    const int nblocks = (njobs + 3) / 4;
    hipLaunchKernelGGL((k_corr32<K, PHASE, HW>), dim3(nblocks), dim3(64), 0, s, cfg, A, B, nr, jobs, njobs, nblocks,
                       out);
}

template <int K, bool PHASE>
bool launch32_hw(const CorrConfig &cfg, int hw, const float *A, const float *B, int nr, const int32_t *jobs,
                 int njobs, const CorrOut &out, hipStream_t s)
{
    switch (hw) {
    case 7: launch32<K, PHASE, 7>(cfg, A, B, nr, jobs, njobs, out, s); return true;
    case 9: launch32<K, PHASE, 9>(cfg, A, B, nr, jobs, njobs, out, s); return true;
    default: return false;
    }
}

}  // namespace

bool launch_correlate_fast(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs,
                           int njobs, const CorrOut &out, hipStream_t s)
{
    if (cfg.M != 32 || cfg.w != 16 || !cfg.padded || !cfg.gain || cfg.finder != HH_REGIONALMAX || cfg.K > 3 ||
        out.planes)
        return false;
    // the box must be symmetric about the zero lag with a zero margin
    const int hw = 16 - cfg.box_lo;
    if (cfg.box_hi != 16 + hw || cfg.box_lo < 1 || cfg.box_hi > 29) return false;
    if (njobs <= 0) return true;
    bool ok;
    if (cfg.method == HH_PHASE)
        ok = cfg.K == 2 ? launch32_hw<2, true>(cfg, hw, A, B, nr, jobs, njobs, out, s)
                        : launch32_hw<3, true>(cfg, hw, A, B, nr, jobs, njobs, out, s);
    else
        ok = cfg.K == 2 ? launch32_hw<2, false>(cfg, hw, A, B, nr, jobs, njobs, out, s)
                        : launch32_hw<3, false>(cfg, hw, A, B, nr, jobs, njobs, out, s);
    if (ok) HH_HIP(hipGetLastError());
    return ok;
}

}  // namespace hh
