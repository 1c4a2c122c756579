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
#include "box_peaks.h"

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

// ---- cross-lane combination of the L sub-transforms of one line ----------
// Lanes of a line are adjacent (quad-aligned), so exchanges are DPP
// quad permutations: full-rate VALU moves, no LDS.
constexpr int kQX1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kQX2 = 0x4E;   // [2,3,0,1]
constexpr int kQX3 = 0x1B;   // [3,2,1,0]
constexpr int kQNEG4 = 0xB4; // [0,1,3,2]

template <int CTRL>
__device__ __forceinline__ float qp(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ float2 qp2(float2 v)
{
    return make_float2(qp<CTRL>(v.x), qp<CTRL>(v.y));
}

// value of the lane holding the mirrored block L-1-m (index 32-k, k >= 1)
template <int L>
__device__ __forceinline__ float2 mirror(float2 v)
{
    if constexpr (L == 1) return v;
    else if constexpr (L == 2) return qp2<kQX1>(v);
    else return qp2<kQX3>(v);
}
// value of the lane holding block (-m) mod L (index 0)
template <int L>
__device__ __forceinline__ float2 neg0(float2 v)
{
    if constexpr (L == 4) return qp2<kQNEG4>(v);
    else return v;
}
// block held by lane l after a forward cross-lane transform (bit reversal)
template <int L>
__device__ __forceinline__ int block_of(int l)
{
    if constexpr (L == 4) return ((l & 1) << 1) | ((l >> 1) & 1);
    else return l;
}

__device__ __forceinline__ float2 cadd2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub2(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot4(float2 b)
{
    return INV ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
}

// Per-lane constants of the cross-lane steps: the butterfly signs of the
// lane's two index bits and whether it is the quad's lane 3 (the one whose
// partial result takes the +-i twiddle), plus its row of four-step twiddles.
struct XLane {
    float sg1, sg2;  // -1 where bit 0 / bit 1 of the sub-lane is set, else +1
    bool r3;
    const float2 *tw;  // W_M^{-q k}, k = 0..31, q = sub-lane (LDS)
};

// Four-step twiddles W_M^{-q k} = (cos, -sin)(2 pi q k / M) for q < 4, k < 32,
// from the same constexpr series as the compile-time twiddles. For M = 64
// (two lanes per line) the quarter turns are exact, as in twmul; for M = 128
// they are the rounded series values, as the per-lane constant selects had.
struct TwTab {
    float c[4][32], s[4][32];
};
template <int M>
constexpr TwTab make_twtab()
{
    TwTab t{};
    for (int q = 0; q < 4; ++q)
        for (int k = 0; k < 32; ++k) {
            const int r = (q * k) % M;
            if (M == 64 && r % (M / 4) == 0) {
                const int h = r / (M / 4);  // 0: 1, 1: -i, 2: -1, 3: +i
                t.c[q][k] = h == 0 ? 1.f : (h == 2 ? -1.f : 0.f);
                t.s[q][k] = h == 1 ? -1.f : (h == 3 ? 1.f : 0.f);
            } else {
                t.c[q][k] = (float)c_cos(angle(q * k, M));
                t.s[q][k] = -(float)c_sin(angle(q * k, M));
            }
        }
    return t;
}
__constant__ TwTab kTw64 = make_twtab<64>();
__constant__ TwTab kTw128 = make_twtab<128>();

// Forward L-point DFT across the lanes of a line, decimation in frequency:
// lane l holds input q = l, returns output block block_of(l). The butterflies
// o +- v are single FMAs with the lane's sign (exact: one rounding, as the
// add or subtract).
template <int L>
__device__ __forceinline__ float2 xdft_fwd(float2 v, const XLane &xl)
{
    if constexpr (L == 1) {
        return v;
    } else if constexpr (L == 2) {
        const float2 o = qp2<kQX1>(v);
        return make_float2(__builtin_fmaf(xl.sg1, v.x, o.x), __builtin_fmaf(xl.sg1, v.y, o.y));
    } else {
        float2 o = qp2<kQX2>(v);
        float2 s = make_float2(__builtin_fmaf(xl.sg2, v.x, o.x), __builtin_fmaf(xl.sg2, v.y, o.y));
        if (xl.r3) s = rot4<false>(s);
        o = qp2<kQX1>(s);
        return make_float2(__builtin_fmaf(xl.sg1, s.x, o.x), __builtin_fmaf(xl.sg1, s.y, o.y));
    }
}
// Inverse L-point DFT across lanes, decimation in time: lane l holds input
// block block_of(l), returns output q = l. Lane 3 applies the +i twiddle to
// its span-1 result before the span-2 exchange, so both spans are plain
// sign butterflies.
template <int L>
__device__ __forceinline__ float2 xdft_inv(float2 v, const XLane &xl)
{
    if constexpr (L == 1) {
        return v;
    } else if constexpr (L == 2) {
        const float2 o = qp2<kQX1>(v);
        return make_float2(__builtin_fmaf(xl.sg1, v.x, o.x), __builtin_fmaf(xl.sg1, v.y, o.y));
    } else {
        float2 o = qp2<kQX1>(v);
        float2 s = make_float2(__builtin_fmaf(xl.sg1, v.x, o.x), __builtin_fmaf(xl.sg1, v.y, o.y));
        if (xl.r3) s = rot4<true>(s);
        o = qp2<kQX2>(s);
        return make_float2(__builtin_fmaf(xl.sg2, s.x, o.x), __builtin_fmaf(xl.sg2, s.y, o.y));
    }
}

// four-step twiddle W_M^{-+q k} for the lane's q, compile-time k: one LDS read
template <int L, int K, bool INV>
__device__ __forceinline__ float2 twq(float2 v, const XLane &xl)
{
    if constexpr (L == 1 || K == 0) {
        return v;
    } else {
        const float2 w = xl.tw[K];
        const float sn = INV ? -w.y : w.y;
        return make_float2(v.x * w.x - v.y * sn, v.x * sn + v.y * w.x);
    }
}

// Forward M-point DFT of a line whose lane q holds x[L n + q], n < 16
// (x[L n + q] = 0 for n >= 16): returns X[k + 32 block_of(l)] in X[k].
template <int L>
__device__ __forceinline__ void line_fwd(const float2 (&x)[16], float2 (&X)[32], const XLane &xl)
{
    dft32_half(x, X);
    sfor<32>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        X[k] = xdft_fwd<L>(twq<L, k, false>(X[k], xl), xl);
    });
}
// Inverse M-point DFT: lane l holds X[k + 32 block_of(l)] in X[k]; returns
// x[L n + l] in X[n] (unnormalized).
template <int L>
__device__ __forceinline__ void line_inv(float2 (&X)[32], const XLane &xl)
{
    sfor<32>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        X[k] = twq<L, k, true>(xdft_inv<L>(X[k], xl), xl);
    });
    fft<32, true>(X);
}

// ---- fused Wieneke statistics for 16-px windows (one row per lane) -------
// DPP row operations act within rows of 16 lanes: exactly one window.
template <int S>
__device__ __forceinline__ int rowsh_i(int v)  // lane r gets lane r + S of its row, 0 outside
{
    if constexpr (S == 0) return v;
    else if constexpr (S > 0) return __builtin_amdgcn_mov_dpp(v, 0x100 + S, 0xF, 0xF, true);
    else return __builtin_amdgcn_mov_dpp(v, 0x110 - S, 0xF, 0xF, true);
}
template <int S>
__device__ __forceinline__ float rowsh(float v)
{
    return __int_as_float(rowsh_i<S>(__float_as_int(v)));
}
template <int S>
__device__ __forceinline__ double rowsh_d(double v)
{
    const int lo = rowsh_i<S>(__double2loint(v)), hi = rowsh_i<S>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
template <int R>
__device__ __forceinline__ double rowror_d(double v)  // rotate within the row by R
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x120 + R, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x120 + R, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double row_sum_d(double v)  // sum over the 16 lanes, in every lane
{
    v += rowror_d<8>(v);
    v += rowror_d<4>(v);
    v += rowror_d<2>(v);
    v += rowror_d<1>(v);
    return v;
}

// Accumulators of one component: C0, C+, C-, S00 and per UQ ring
// (max(|dr|, |dc|) = 1..4) the maximum and the sum of its covariance sums:
// everything finalize_uncertainty (src/uncertainty.jl:162-179) reads.
struct UQAcc {
    double C0, Cp, Cm, S00;
    double rmax[4], rsum[4];
};

__device__ __forceinline__ float uq_sigma(const UQAcc &a)
{
    double s2 = a.S00;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (a.rmax[q] < 0.05 * a.S00) break;
        s2 += 2.0 * a.rsum[q];
    }
    const double sdc = sqrt(s2 > 0.0 ? s2 : 0.0);
    const double Cpm = (a.Cp + a.Cm) / 2.0;
    const double lo = Cpm - sdc / 2.0, hi = Cpm + sdc / 2.0;
    if (!(a.C0 > 0.0 && lo > 0.0)) return NAN;
    const double denom = 2.0 * log(a.C0) - log(hi) - log(lo);
    if (!(denom > 0.0)) return NAN;
    return (float)((log(hi) - log(lo)) / (2.0 * denom));
}

__device__ __forceinline__ void uq_init(UQAcc &a)
{
    a.C0 = a.Cp = a.Cm = a.S00 = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a.rmax[q] = -INFINITY;
        a.rsum[q] = 0.0;
    }
}

// Covariance sums of one component whose centered, smoothed products sit one
// row per lane (e[c], c < NC): lane shift s (row r with row r + s of the same
// 16-lane row, by DPP, zero past the window) times in-lane offset o, with
// s = 0, o = 0..4 (o = 0 is S00) and s = 1..4, o = -4..4. Up to the point
// reflection S(-a, -b) = S(a, b) this is the half-plane offset set of the
// ring sums (the ring of an offset is max(|s|, |o|)). The cross-lane totals
// go through LDS: each lane writes its row partials of a group of lags (16 rows x kUQP floats
// per window, the row pitch chosen so 16 lanes hit distinct banks), then lane
// r adds the 16 row partials of lags r and r + 16 and folds its totals into
// per-ring maximum and sum, which DPP reductions finish over the window:
// about a third of the instructions of a 16-lane DPP reduction per lag. Lags
// are numbered s = 0, o = 0..4 (k = o; k = 0 is S00), then s = 1..4,
// o = -4..4 (k = 5 + 9 (s - 1) + o + 4), in two groups (s <= 2: 23 lags;
// s >= 3: 18) so the scratch is 1600 bytes per window. The partials are FP32
// dot products, so holding them as FP32 loses nothing; the totals are FP64.
constexpr int kUQP = 25;

__device__ __forceinline__ void lds_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int uq_ring(int k)
{
    if (k < 5) return k;
    const int s = 1 + (k - 5) / 9, o = (k - 5) % 9 - 4;
    const int ao = o < 0 ? -o : o;
    return s > ao ? s : ao;
}

__device__ __forceinline__ double row_max_d(double v)  // a NaN never replaces the running maximum
{
    double o = rowror_d<8>(v);
    v = o > v ? o : v;
    o = rowror_d<4>(v);
    v = o > v ? o : v;
    o = rowror_d<2>(v);
    v = o > v ? o : v;
    o = rowror_d<1>(v);
    v = o > v ? o : v;
    return v;
}

// Totals of the N lags of one group (global numbers kbase + 0..N-1).
template <int N>
__device__ __forceinline__ void uq_collect(const float *red, int r, int kbase, double (&rmax)[4], double (&rsum)[4],
                                           double &s00)
{
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int k = r + 16 * j;
        if (k < N) {
            double tot = 0.0;
#pragma unroll
            for (int q = 0; q < 16; ++q) tot += (double)red[q * kUQP + k];
            const int kg = kbase + k;
            const int ring = uq_ring(kg);
            if (kg == 0) s00 = tot;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (ring == q + 1) {
                    rmax[q] = tot > rmax[q] ? tot : rmax[q];
                    rsum[q] += tot;
                }
        }
    }
}

// red: this window's 16 x kUQP scratch; r: the lane's row (0..15). Each
// row's partial of a lag is a dot product of at most 16 terms, formed in
// FP32 from the centred values rounded to FP32 (relative error ~1e-6 of the
// row's |terms|, far inside the statistics' use) and taken to FP64 before
// every cross-row sum; FP32 products halve the cost of the FP64 ones.
template <int NC>
__device__ __forceinline__ void uq_lags_lds(const double (&ed)[NC], UQAcc &acc, float *red, int r)
{
    float e[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) e[c] = (float)ed[c];
    double rmax[4], rsum[4], s00 = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        rmax[q] = -INFINITY;
        rsum[q] = 0.0;
    }
    float *row = red + r * kUQP;
    auto shifted = [&](auto ss, int base) {
        constexpr int s = decltype(ss)::value;
        float p[9];
#pragma unroll
        for (int oi = 0; oi < 9; ++oi) p[oi] = 0.f;
#pragma unroll
        for (int cs = 0; cs < NC; ++cs) {
            const float sh = rowsh<s>(e[cs]);
#pragma unroll
            for (int oi = 8; oi >= 0; --oi) {
                const int o = oi - 4, c = cs - o;
                if (c >= 0 && c < NC) p[oi] = __builtin_fmaf(e[c], sh, p[oi]);
            }
        }
#pragma unroll
        for (int oi = 0; oi < 9; ++oi) row[base + oi] = p[oi];
    };
    // group 1: s = 0, 1, 2
#pragma unroll
    for (int o = 0; o < 5; ++o) {
        float p = 0.f;
#pragma unroll
        for (int c = 0; c + o < NC; ++c) p = __builtin_fmaf(e[c], e[c + o], p);
        row[o] = p;
    }
    shifted(std::integral_constant<int, 1>{}, 5);
    shifted(std::integral_constant<int, 2>{}, 14);
    lds_wave_sync();
    uq_collect<23>(red, r, 0, rmax, rsum, s00);
    lds_wave_sync();
    // group 2: s = 3, 4
    shifted(std::integral_constant<int, 3>{}, 0);
    shifted(std::integral_constant<int, 4>{}, 9);
    lds_wave_sync();
    uq_collect<18>(red, r, 23, rmax, rsum, s00);
    lds_wave_sync();
    acc.S00 = row_sum_d(s00);  // held by lane 0 alone: the sum is a broadcast
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        acc.rmax[q] = row_max_d(rmax[q]);
        acc.rsum[q] = row_sum_d(rsum[q]);
    }
}

// uq_component! (src/uncertainty.jl:104-151) of both components of a 16-px
// window whose apodized, mean-subtracted rows sit one per lane (x[j] =
// (wA[r, j], wB[r, j]), r = lane of the row). u: shift along columns
// (in-lane); v: the transposed windows, shift along rows (across lanes).
__device__ void uq_fused16(const float2 (&x)[16], int r, float &sig_u, float &sig_v, float *red)
{
#pragma clang fp contract(off)
    const float q = 0.25f;
    UQAcc accu, accv;
    // ---- u: d(r, c), c = 0..14 in registers; lags dr across lanes ----
    {
        UQAcc &acc = accu;
        uq_init(acc);
        float dC[15];
        double c0 = 0.0, cp = 0.0, cm = 0.0;
#pragma unroll
        for (int c = 0; c < 15; ++c) {
            const float a0 = x[c].x, a1 = x[c + 1].x, b0 = x[c].y, b1 = x[c + 1].y;
            const float p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0;
            c0 += (double)p00;
            cp += (double)p01;
            cm += (double)p10;
            dC[c] = p01 - p10;
        }
        double d[15];
        double mu = 0.0;
#pragma unroll
        for (int c = 0; c < 15; ++c) {
            const int cl = c > 0 ? c - 1 : 0, cr = c < 14 ? c + 1 : 14;
            const float v = q * ((dC[cl] + 2.f * dC[c]) + dC[cr]);
            d[c] = (double)v;
            mu += (double)v;
        }
        acc.C0 = row_sum_d(c0);
        acc.Cp = row_sum_d(cp);
        acc.Cm = row_sum_d(cm);
        mu = row_sum_d(mu) / 240.0;
#pragma unroll
        for (int c = 0; c < 15; ++c) d[c] -= mu;
        uq_lags_lds<15>(d, acc, red, r);
    }
    // ---- v: transposed windows: d(j; i), i = row = lane 0..14, j = 0..15 ----
    {
        UQAcc &acc = accv;
        uq_init(acc);
        const bool live = r < 15;
        float dC[16];
        double c0 = 0.0, cp = 0.0, cm = 0.0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float a0 = x[j].x, b0 = x[j].y;
            const float a1 = rowsh<1>(x[j].x), b1 = rowsh<1>(x[j].y);
            const float p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0;
            if (live) {
                c0 += (double)p00;
                cp += (double)p01;
                cm += (double)p10;
            }
            dC[j] = live ? p01 - p10 : 0.f;
        }
        double d[16];
        double mu = 0.0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float up = rowsh<-1>(dC[j]), dn = rowsh<1>(dC[j]);
            const float lft = r > 0 ? up : dC[j];
            const float rgt = r < 14 ? dn : dC[j];
            const float v = q * ((lft + 2.f * dC[j]) + rgt);
            d[j] = live ? (double)v : 0.0;
            mu += d[j];
        }
        acc.C0 = row_sum_d(c0);
        acc.Cp = row_sum_d(cp);
        acc.Cm = row_sum_d(cm);
        mu = row_sum_d(mu) / 240.0;
#pragma unroll
        for (int j = 0; j < 16; ++j) d[j] = live ? d[j] - mu : 0.0;
        uq_lags_lds<16>(d, acc, red, r);
    }
    // every lane of the window holds both sets of statistics: even lanes
    // finalize u, odd lanes v, and lane r takes v from lane r + 1
    const bool odd = r & 1;
    UQAcc a;
    a.C0 = odd ? accv.C0 : accu.C0;
    a.Cp = odd ? accv.Cp : accu.Cp;
    a.Cm = odd ? accv.Cm : accu.Cm;
    a.S00 = odd ? accv.S00 : accu.S00;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a.rmax[q] = odd ? accv.rmax[q] : accu.rmax[q];
        a.rsum[q] = odd ? accv.rsum[q] : accu.rsum[q];
    }
    const float sg = uq_sigma(a);
    sig_u = sg;             // valid on even lanes (lane 0 writes)
    sig_v = rowsh<1>(sg);
}

// ---- Wieneke statistics with the covariance sums on the matrix cores ------
// The uncertainty of a 16-px sweep whose correlation runs on the matrix cores
// (the FFT form fuses uq_fused16 into its correlation kernel instead).
// For one component with centred, smoothed values e(r, k) (u: r = window row,
// k = column 0..14; v: r = window column, k = row 0..14) the 41 half-plane
// covariance sums S(dr, dk) = sum e(r, k) e(r + dr, k + dk) (dk = 1..4 with
// dr = -4..4, and dk = 0 with dr = 0..4, dr = dk = 0 being S00) are diagonal
// sums of five 16 x 16 products G_dk(r, r') = sum_k e(r, k) e(r', k + dk):
// S(dr, dk) = sum_r G_dk(r, r + dr). An entry of G is the row partial that
// uq_lags_lds forms as an FP32 dot product of at most 15 FP32 products; here
// v_mfma_f32_16x16x16_f16 forms it from e split into FP16 hi + lo parts
// (hi*hi + hi*lo + lo*hi, FP32 accumulation, ~22-bit operands), after a
// power-of-two scale that puts max |e| of the (window, component) in
// [1/2, 1) and is removed exactly afterwards. The diagonal sums over r are
// FP32 row by row (each entry is already an FP32 dot product) and FP64 over
// the lag.
// One wave per block, 4 windows per wave. The fields are formed as in
// uq_fused16 (16 lanes per window, one window row per lane; C0, C+, C- sums
// of FP32 products, FP32 within a row and FP64 across rows) and go to LDS as
// FP32, centred by the FP64 window mean rounded to FP32: first the u fields
// of the four windows, then, in the same place, their v fields. Each
// (window, component) takes its 15 products; the entries of the nine
// diagonals |r' - r| <= 4 (r' >= r for the symmetric G_0) go to LDS diagonal by diagonal, so that lane
// k < 41 reads the 16 entries of lag k with four 16-byte loads; the lanes
// (window, component) finish the rings and sigma.
// LDS strides are chosen so that the field writes, the operand reads and the
// diagonal stores and loads spread over the banks. One component's fields at
// a time keep the LDS at 9.2 KB per wave, so four waves fit a SIMD. FP32
// products on v_mfma_f32_16x16x4_f32 (the same precision as the VALU row
// partials) ran no faster than the VALU, since they do not co-execute with
// VALU work and G is twice the band of lags.
constexpr int kUE = 20;              // e row pitch: k = 0..14, zero pads 15..19
constexpr int kUW = 320;             // window stride of the fields (16 * kUE)
constexpr int kUD = 20;              // diagonal pitch (16 entries r = 0..15)
constexpr int kUDK = 180;            // per dk >= 1: diagonals dr + 4 = 0..8
constexpr int kUD0 = 5 * kUD;        // dk = 0: diagonals dr = 0..4 (G_0 is symmetric)
constexpr int kUL = 41;              // lags: S00 and the 40 half-plane offsets

// Lag k: k = 0..4 -> (dr, dk) = (k, 0); k = 5..40 -> dk = 1 + (k - 5) / 9,
// dr = (k - 5) % 9 - 4. Ring max(|dr|, dk) (0 for S00).
__host__ __device__ constexpr int uqm_dk(int k) { return k < 5 ? 0 : 1 + (k - 5) / 9; }
__host__ __device__ constexpr int uqm_dr(int k) { return k < 5 ? k : (k - 5) % 9 - 4; }
// offset of diagonal (dk, dr) in the diagonal buffer
__host__ __device__ constexpr int uqm_diag(int dk, int dr) { return dk == 0 ? dr * kUD : kUD0 + (dk - 1) * kUDK + (dr + 4) * kUD; }
__host__ __device__ constexpr int uqm_ring(int k)
{
    const int a = uqm_dr(k) < 0 ? -uqm_dr(k) : uqm_dr(k);
    return a > uqm_dk(k) ? a : uqm_dk(k);
}

typedef float f4v __attribute__((ext_vector_type(4)));
typedef _Float16 h4v __attribute__((ext_vector_type(4)));
typedef _Float16 h2u __attribute__((ext_vector_type(2)));
typedef float f2u __attribute__((ext_vector_type(2)));

struct Split4 {
    h4v hi, lo;
};

// x = hi + lo with hi = f16(x), lo = f16(x - hi): one packed conversion per
// two values, and per value one mixed-precision FMA x + hi * (-1) (exact in
// FP32 inside the FMA) written rounded into its half of lo.
__device__ __forceinline__ Split4 split4(f4v x, float neg1)
{
    Split4 s;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const f2u xv = {x[2 * p], x[2 * p + 1]};
        const h2u hi = __builtin_convertvector(xv, h2u);
        const unsigned hp = __builtin_bit_cast(unsigned, hi);
        unsigned lp;
        asm("v_fma_mixlo_f16 %0, %1, %2, %3 op_sel_hi:[1,0,0]\n\t"
            "v_fma_mixhi_f16 %0, %1, %2, %4 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
            : "=&v"(lp)
            : "v"(hp), "v"(neg1), "v"(xv[0]), "v"(xv[1]));
        const h2u lo = __builtin_bit_cast(h2u, lp);
        s.hi[2 * p] = hi[0];
        s.hi[2 * p + 1] = hi[1];
        s.lo[2 * p] = lo[0];
        s.lo[2 * p + 1] = lo[1];
    }
    return s;
}

__device__ __forceinline__ f4v mfma3_16(const Split4 &a, const Split4 &b, f4v acc)
{
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.lo, b.hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.lo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.hi, acc, 0, 0, 0);
    return acc;
}

// max |e| over the 16 lanes of the window -> exponent p with max |e| 2^p in
// [1/2, 1) (0 for an all-zero field)
__device__ __forceinline__ int uqm_scale_exp(float amx)
{
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) amx = fmaxf(amx, __shfl_xor(amx, o));
    int ex = 0;
    (void)frexpf(amx, &ex);
    return amx > 0.f ? -ex : 0;
}

// The u and v fields of uq_component! (src/uncertainty.jl:104-151) for the
// matrix-core kernel: C0, C+, C- as in uq_fused16; e = the smoothed dC values
// minus the FP64 window mean rounded to FP32. The reference adds each
// FP32 product (and each smoothed value of the mean) to a Float64 sum; here a
// lane first adds its row's 15 or 16 terms in FP32 (relative error ~1e-6 of
// the row's terms) and the rows are summed in FP64, one conversion per row
// instead of one per term.
__device__ __forceinline__ void uqm_field_u(const float2 (&x)[16], double (&C)[3], float (&e)[15])
{
#pragma clang fp contract(off)
    const float q = 0.25f;
    float dC[15];
    float c0 = 0.f, cp = 0.f, cm = 0.f;
#pragma unroll
    for (int c = 0; c < 15; ++c) {
        const float a0 = x[c].x, a1 = x[c + 1].x, b0 = x[c].y, b1 = x[c + 1].y;
        const float p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0;
        c0 += p00;
        cp += p01;
        cm += p10;
        dC[c] = p01 - p10;
    }
    float mu = 0.f;
#pragma unroll
    for (int c = 0; c < 15; ++c) {
        const int cl = c > 0 ? c - 1 : 0, cr = c < 14 ? c + 1 : 14;
        e[c] = q * ((dC[cl] + 2.f * dC[c]) + dC[cr]);
        mu += e[c];
    }
    C[0] = row_sum_d((double)c0);
    C[1] = row_sum_d((double)cp);
    C[2] = row_sum_d((double)cm);
    const float mf = (float)(row_sum_d((double)mu) / 240.0);
#pragma unroll
    for (int c = 0; c < 15; ++c) e[c] -= mf;
}

// lane r = i holds e(i; j), j = 0..15, for i = 0..14 (lane 15 holds zeros)
__device__ __forceinline__ void uqm_field_v(const float2 (&x)[16], int r, double (&C)[3], float (&e)[16])
{
#pragma clang fp contract(off)
    const float q = 0.25f;
    const bool live = r < 15;
    float dC[16];
    float c0 = 0.f, cp = 0.f, cm = 0.f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float a0 = x[j].x, b0 = x[j].y;
        const float a1 = rowsh<1>(x[j].x), b1 = rowsh<1>(x[j].y);
        const float p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0;
        c0 += p00;
        cp += p01;
        cm += p10;
        dC[j] = live ? p01 - p10 : 0.f;
    }
    float mu = 0.f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float up = rowsh<-1>(dC[j]), dn = rowsh<1>(dC[j]);
        const float lft = r > 0 ? up : dC[j];
        const float rgt = r < 14 ? dn : dC[j];
        e[j] = live ? q * ((lft + 2.f * dC[j]) + rgt) : 0.f;
        mu += e[j];
    }
    C[0] = row_sum_d(live ? (double)c0 : 0.0);
    C[1] = row_sum_d(live ? (double)cp : 0.0);
    C[2] = row_sum_d(live ? (double)cm : 0.0);
    const float mf = (float)(row_sum_d((double)mu) / 240.0);
#pragma unroll
    for (int j = 0; j < 16; ++j) e[j] = live ? e[j] - mf : 0.f;
}

__global__ __launch_bounds__(64) void k_uq16_mfma(CorrConfig cfg, const float *__restrict__ A,
                                                  const float *__restrict__ B, int nr,
                                                  const int32_t *__restrict__ jobs, int njobs, int nblocks,
                                                  CorrOut out)
{
#pragma clang fp contract(off)
    // fields of one component: window w's e(r, k) at w * kUW + r * kUE + k
    __shared__ __align__(16) float ef[4 * kUW];
    // lag totals per (window, component), with the field scale 2^(2 pex)
    __shared__ float tot[4][2][kUL];
    __shared__ __align__(16) float gd[kUD0 + 4 * kUDK];  // G_dk(r, r + dr) at uqm_diag(dk, dr) + r
    __shared__ int pex[4][2];                            // field scale exponents
    static_assert(kUW >= 16 * kUE && kUW % 4 == 0, "field layout");
    const int tid = threadIdx.x, line = tid & 15, slot_w = tid >> 4;
    // diagonal entries past the ends of G (r + dr outside 0..15) are never
    // stored: zero once
    for (int i = tid; i < kUD0 + 4 * kUDK; i += 64) gd[i] = 0.f;
    const int b = blockIdx.x;
    const int xcd = b & 7, slot = b >> 3;
    const int q8 = nblocks >> 3, r8 = nblocks & 7;
    const int tile = cfg.linear_order ? b : xcd * q8 + min(xcd, r8) + slot;
    const int win = tile * 4 + slot_w;
    const bool active = win < njobs;
    const int jb = active ? win : njobs - 1;
    const int gi = jobs[4 * jb], gj = jobs[4 * jb + 1], rs = jobs[4 * jb + 2], cs = jobs[4 * jb + 3];
    const float *pa = A + (size_t)(cs - 1) * nr + (rs - 1);
    const float *pb = B + (size_t)(cs - 1) * nr + (rs - 1);
    float ra[16], rb[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        ra[j] = pa[line + (size_t)j * nr];
        rb[j] = pb[line + (size_t)j * nr];
    }
    float sa = 0.f, sb = 0.f;
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
        const float a = cfg.apod[line + 16 * j];
        x[j] = make_float2(a * (ra[j] - meanA), a * (rb[j] - meanB));
    }
    float *ew = ef + slot_w * kUW;
    // operands of v_mfma_f32_16x16x16_f16: lane l gives A(i = l % 16, k =
    // 4 (l / 16) + 0..3) and B(k = 4 (l / 16) + 0..3, j = l % 16); it
    // receives D(4 (l / 16) + v, l % 16)
    const int l = tid, lq = l >> 4, lj = l & 15;
    const int eo = lj * kUE + 4 * lq;
    const float neg1 = -1.f;
    // lag of this lane for the diagonal sums (lanes 41-63 read lag 0's)
    const int lk = l < kUL ? l : 0;
    const int gb = uqm_diag(uqm_dk(lk), uqm_dr(lk));
    double Cf[3];  // C0, C+, C- of this lane's component in the finalize (u: line 0, v: line 1)
#pragma unroll 1
    for (int comp = 0; comp < 2; ++comp) {
        double C[3];
        int p;
        if (comp == 0) {
            float e[15];
            uqm_field_u(x, C, e);
            float amx = 0.f;
#pragma unroll
            for (int c = 0; c < 15; ++c) amx = fmaxf(amx, fabsf(e[c]));
            p = uqm_scale_exp(amx);
            float *er = ew + line * kUE;  // r = line, k = c
#pragma unroll
            for (int c = 0; c < 15; ++c) er[c] = ldexpf(e[c], p);
#pragma unroll
            for (int c = 15; c < kUE; ++c) er[c] = 0.f;
        } else {
            // the u fields are consumed (the loop below ended with a sync)
            float e[16];
            uqm_field_v(x, line, C, e);
            float amx = 0.f;
#pragma unroll
            for (int j = 0; j < 16; ++j) amx = fmaxf(amx, fabsf(e[j]));
            p = uqm_scale_exp(amx);
            // r = window column j, k = row i = line (line 15: zeros)
#pragma unroll
            for (int j = 0; j < 16; ++j) ew[j * kUE + line] = ldexpf(e[j], p);
#pragma unroll
            for (int k = 16; k < kUE; ++k) ew[line * kUE + k] = 0.f;
        }
        if (line == comp) {
#pragma unroll
            for (int i = 0; i < 3; ++i) Cf[i] = C[i];
        }
        if (line == 0) pex[slot_w][comp] = p;
        lds_wave_sync();
#pragma unroll 1
        for (int w = 0; w < 4; ++w) {
            const float *e = ef + w * kUW;
            // k = 4 lq + 0..3 (A) and 4 lq + 4..7: two aligned 16-byte loads;
            // the B operand of dk takes k + dk = 4 lq + dk .. 4 lq + dk + 3
            // from the two split quads (the field's zero pads end each row)
            const f4v *er = reinterpret_cast<const f4v *>(e + eo);
            const Split4 q0 = split4(er[0], neg1), q1 = split4(er[1], neg1);
            f4v g[5];
#pragma unroll
            for (int dk = 0; dk < 5; ++dk) {
                Split4 bq;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    bq.hi[j] = j + dk < 4 ? q0.hi[j + dk] : q1.hi[j + dk - 4];
                    bq.lo[j] = j + dk < 4 ? q0.lo[j + dk] : q1.lo[j + dk - 4];
                }
                g[dk] = mfma3_16(q0, bq, (f4v){0.f, 0.f, 0.f, 0.f});
            }
            // entries (r, r') = (4 lq + v, lj) with |r' - r| <= 4, by diagonal
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 4 * lq + v, dr = lj - r;
                if (dr >= -4 && dr <= 4) {
                    if (dr >= 0) gd[uqm_diag(0, dr) + r] = g[0][v];
#pragma unroll
                    for (int dk = 1; dk < 5; ++dk) gd[uqm_diag(dk, dr) + r] = g[dk][v];
                }
            }
            lds_wave_sync();
            // the 16 entries of the lag's diagonal, added pairwise in FP32
            // (each entry is already an FP32 dot product; the error of the
            // sum stays ~1e-7 of S00), then taken to FP64
            const f4v *dp = reinterpret_cast<const f4v *>(gd + gb);
            float s4[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const f4v q = dp[i];
                s4[i] = (q[0] + q[1]) + (q[2] + q[3]);
            }
            if (l < kUL) tot[w][comp][l] = (s4[0] + s4[1]) + (s4[2] + s4[3]);
            lds_wave_sync();
        }
    }
    // lanes (window, component): rings and sigma (finalize_uncertainty,
    // src/uncertainty.jl:162-179); a NaN never replaces a ring maximum
    if (line < 2) {
        UQAcc acc;
        uq_init(acc);
        const float *s = tot[slot_w][line];
        const int sc = -2 * pex[slot_w][line];  // the lag totals without the field scale (exact)
        acc.C0 = Cf[0];
        acc.Cp = Cf[1];
        acc.Cm = Cf[2];
        acc.S00 = ldexp((double)s[0], sc);
        // the rings over the scaled totals, the scale removed once per ring
        // (a power of two: the FP64 sums and maxima scale exactly)
#pragma unroll
        for (int k = 1; k < kUL; ++k) {
            const int q = uqm_ring(k) - 1;
            const double t = (double)s[k];
            acc.rmax[q] = fmax(acc.rmax[q], t);  // a NaN t keeps the maximum
            acc.rsum[q] += t;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc.rmax[q] = ldexp(acc.rmax[q], sc);
            acc.rsum[q] = ldexp(acc.rsum[q], sc);
        }
        const float sg = uq_sigma(acc);
        if (active) {
            const size_t g = (size_t)(gi - 1) + (size_t)(gj - 1) * out.ny;
            (line ? out.unc_v : out.unc_u)[g] = sg;
        }
    }
}


// Correlation of w = 16 L windows padded to M = 32 L. L lanes per line of the
// 2D transform, G = 16 L^2 lanes per window. HW: half-width of the analysis
// box [M/2 - HW, M/2 + HW]^2, which holds the gain support plus a zero margin.
template <int L, int K, bool PHASE, int HW, bool UQ>
__global__ __launch_bounds__(L == 4 ? 256 : 64) __attribute__((amdgpu_waves_per_eu(L >= 2 ? 2 : 1))) void k_corrL(CorrConfig cfg, const float *__restrict__ A,
                                                           const float *__restrict__ B, int nr,
                                                           const int32_t *__restrict__ jobs, int njobs, int nblocks,
                                                           CorrOut out)
{
#pragma clang fp contract(off)
    constexpr int M = 32 * L;
    constexpr int H = M / 2;
    constexpr int W = 16 * L;                // window side = number of lines
    constexpr int G = 16 * L * L;            // lanes per window
    constexpr int WPB = G >= 64 ? 1 : 64 / G;  // windows per block
    constexpr int NT = G >= 64 ? G : 64;     // threads per block
    constexpr int NB = 2 * HW + 1;
    constexpr int LO = H - HW;
    constexpr int U = W + 1;                 // unit pitch of the kept rows
    // row-frequency spectra: column k2 of the window's lines at pitch TP =
    // W + 1 float4, so the column reads of step 3 (lanes of different units,
    // same row) and the row writes of step 2 fall in distinct LDS banks
    constexpr int TP = W + 1;
    constexpr int TSZ = (H + 1) * TP;        // float4 per window
    constexpr int NPAIR = (NB + 1) / 2;
    constexpr int NWAVE = NT / 64;
    static_assert(NB * U * 2 + NB * NB <= TSZ * 4, "box does not fit the LDS slice");
    static_assert(2 * W * (W + 1) <= TSZ * 4, "staging does not fit the LDS slice");
    __shared__ float4 lds[WPB * TSZ];
    __shared__ float red[2 * NWAVE];
    __shared__ float mv[NWAVE];
    __shared__ int ml[NWAVE];
    __shared__ float2 twl[4][33];  // four-step twiddles, rows padded against bank conflicts

    const int tid = threadIdx.x;
    const int slot_w = tid / G;               // window slot in the block
    const int t = tid % G;                    // lane within the window
    const int line = t / L, l = t % L;        // line (row / unit) and sub-lane
    if constexpr (L > 1) {
        const TwTab &tt = L == 2 ? kTw64 : kTw128;
        for (int e = tid; e < 4 * 32; e += NT) twl[e >> 5][e & 31] = make_float2(tt.c[e >> 5][e & 31], tt.s[e >> 5][e & 31]);
        __syncthreads();
    }
    const XLane xl{(l & 1) ? -1.f : 1.f, (l & 2) ? -1.f : 1.f, (l & 3) == 3, &twl[l][0]};
    const int wave = tid >> 6;
    // XCD-aware order: block b runs on XCD b % 8; give each XCD a contiguous
    // range of tiles so overlapping windows share its L2
    const int b = blockIdx.x;
    const int xcd = b & 7, slot = b >> 3;
    const int q8 = nblocks >> 3, r8 = nblocks & 7;
    const int tile = cfg.linear_order ? b : xcd * q8 + min(xcd, r8) + slot;
    const int win = tile * WPB + slot_w;
    const bool active = win < njobs;
    const int jb = active ? win : njobs - 1;
    const int gi = jobs[4 * jb], gj = jobs[4 * jb + 1], rs = jobs[4 * jb + 2], cs = jobs[4 * jb + 3];
    float4 *T = lds + slot_w * TSZ;

    // ---- 1. window rows -> means -> apodized A + iB ----
    float2 x[16];
    {
        const float *pa = A + (size_t)(cs - 1) * nr + (rs - 1);
        const float *pb = B + (size_t)(cs - 1) * nr + (rs - 1);
        float sa = 0.f, sb = 0.f;
        if constexpr (L == 1) {
            float ra[16], rb[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                ra[j] = pa[line + (size_t)j * nr];
                rb[j] = pb[line + (size_t)j * nr];
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
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const float a = cfg.apod[line + 16 * j];
                x[j] = make_float2(a * (ra[j] - meanA), a * (rb[j] - meanB));
            }
        } else {
            // stage the window through LDS (coalesced global reads)
            float *SA = reinterpret_cast<float *>(T);
            float *SB = SA + W * (W + 1);
            for (int e = t; e < W * W; e += G) {
                const int i = e % W, j = e / W;
                const float va = pa[i + (size_t)j * nr], vb = pb[i + (size_t)j * nr];
                SA[i + j * (W + 1)] = va;
                SB[i + j * (W + 1)] = vb;
                sa += va;
                sb += vb;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                sa += __shfl_xor(sa, o);
                sb += __shfl_xor(sb, o);
            }
            if constexpr (NWAVE > 1) {
                if ((tid & 63) == 0) {
                    red[2 * wave] = sa;
                    red[2 * wave + 1] = sb;
                }
                __syncthreads();
                sa = 0.f;
                sb = 0.f;
#pragma unroll
                for (int k = 0; k < NWAVE; ++k) {
                    sa += red[2 * k];
                    sb += red[2 * k + 1];
                }
            } else {
                __syncthreads();
            }
            const float meanA = sa / (float)(W * W), meanB = sb / (float)(W * W);
#pragma unroll
            for (int n = 0; n < 16; ++n) {
                const int j = L * n + l;
                const float a = cfg.apod[line + W * j];
                x[n] = make_float2(a * (SA[line + j * (W + 1)] - meanA), a * (SB[line + j * (W + 1)] - meanB));
            }
            __syncthreads();  // staging is overwritten by the transpose below
        }
    }

    // A frame whose loaded window is identically zero has an exactly zero
    // plane in the reference's separate transforms; the packed A + iB
    // transform would leak the other frame's round-off into it, so the box
    // is scaled by keep (0 for such a window, else 1).
    float keep;
    {
        bool nza = false, nzb = false;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            nza |= x[j].x != 0.f;
            nzb |= x[j].y != 0.f;
        }
        if constexpr (G >= 64) {
            const int anya = __syncthreads_or(nza);
            const int anyb = __syncthreads_or(nzb);
            keep = anya && anyb ? 1.f : 0.f;
        } else {
            const uint64_t grp = ((1ull << G) - 1) << ((tid & 63) / G * G);
            const uint64_t ba = __ballot(nza), bb = __ballot(nzb);
            keep = (ba & grp) && (bb & grp) ? 1.f : 0.f;
        }
    }

    // fused uncertainty: the final sweep's windows are the ones the
    // post-loop uncertainty_sweep! reads (src/pipeline.jl:810-824)
    float sig_u = NAN, sig_v = NAN;
    // (the window's LDS slice is free until the transpose below)
    if constexpr (UQ && L == 1) uq_fused16(x, line, sig_u, sig_v, reinterpret_cast<float *>(T));

    // ---- 2. row DFT (dim 2) and Hermitian split, k2 = 0..H ----
    const int mblk = block_of<L>(l);
    {
        float2 Y[32];
        line_fwd<L>(x, Y, xl);
        sfor<32>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const int k2 = k + 32 * mblk;
            const float2 z = Y[k];
            float2 zm;
            if constexpr (k == 0)
                zm = neg0<L>(Y[0]);
            else
                zm = mirror<L>(Y[32 - k]);
            if (k2 <= H)
                T[k2 * TP + line] = make_float4(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y), 0.5f * (z.y + zm.y),
                                               -0.5f * (z.x - zm.x));
        });
    }
    __syncthreads();

    // ---- 3. column DFT (dim 1) of unit u = line; spectrum; inverse dim 1 ----
    float2 P[32];
    {
        const int u = line;
        float2 fa[16], fb[16];
        // unit 0 packs the real columns k2 = 0 and H as re + i im
#pragma unroll
        for (int n = 0; n < 16; ++n) {
            const int i = L * n + l;
            const float4 v = T[u * TP + i], vh = T[H * TP + i];
            fa[n] = u != 0 ? make_float2(v.x, v.y) : make_float2(v.x, vh.x);
            fb[n] = u != 0 ? make_float2(v.z, v.w) : make_float2(v.z, vh.z);
        }
        float2 FA[32], FB[32];
        line_fwd<L>(fa, FA, xl);
        line_fwd<L>(fb, FB, xl);
        const bool unit0 = u == 0;
        sfor<17>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            constexpr int km = (32 - k) & 31;
            // FA(-k1) for k1 = k + 32 m and for k1' = km + 32 m
            float2 gam, gam2, gbm, gbm2;
            if constexpr (k == 0) {
                gam = neg0<L>(FA[0]);
                gbm = neg0<L>(FB[0]);
            } else {
                gam = mirror<L>(FA[km]);
                gbm = mirror<L>(FB[km]);
            }
            gam2 = mirror<L>(FA[k]);
            gbm2 = mirror<L>(FB[k]);
            auto spec = [&](float2 ga, float2 gm, float2 gb, float2 gbmm, int k1) {
                float2 p = conj_mul(ga, gb);
                const float2 a0 = make_float2(0.5f * (ga.x + gm.x), 0.5f * (ga.y - gm.y));
                const float2 aH = make_float2(0.5f * (ga.y + gm.y), -0.5f * (ga.x - gm.x));
                const float2 b0 = make_float2(0.5f * (gb.x + gbmm.x), 0.5f * (gb.y - gbmm.y));
                const float2 bH = make_float2(0.5f * (gb.y + gbmm.y), -0.5f * (gb.x - gbmm.x));
                float2 p0 = conj_mul(a0, b0), pH = conj_mul(aH, bH);
                if constexpr (PHASE) {
                    p = phase_norm(p, cfg.wfilt[k1 + u * M]);
                    p0 = phase_norm(p0, cfg.wfilt[k1]);
                    pH = phase_norm(pH, cfg.wfilt[k1 + H * M]);
                }
                return unit0 ? make_float2(p0.x - pH.y, p0.y + pH.x) : p;  // unit 0: P0 + i PH
            };
            P[k] = spec(FA[k], gam, FB[k], gbm, k + 32 * mblk);
            if constexpr (km != k) P[km] = spec(FA[km], gam2, FB[km], gbm2, km + 32 * mblk);
        });
        line_inv<L>(P, xl);  // P[n] = Q(L n + l, u)
    }
    __syncthreads();  // the LDS slice is reused for the kept rows

    float2 *Qs = reinterpret_cast<float2 *>(T);              // NB x U
    float *Rb = reinterpret_cast<float *>(T) + NB * U * 2;    // NB x NB, column-major
    sfor<32>([&](auto nn) {
        constexpr int n = decltype(nn)::value;
        const int l1 = L * n + l;
        const int ip = (l1 + H) & (M - 1);
        if (ip >= LO && ip < LO + NB) Qs[(ip - LO) * U + line] = P[n];
    });
    __syncthreads();

    // ---- 4. complex-to-real inverse along dim 2, two box rows per line ----
    if (t < NPAIR * L) {
        const int pair = line;
        const int qa = 2 * pair, qb = 2 * pair + 1;
        const bool hasb = qb < NB;
        const float2 *Qa = Qs + qa * U;
        const float2 *Qb = Qs + (hasb ? qb : qa) * U;
        const float zb = hasb ? 1.f : 0.f;
        float2 c[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const int k2 = k + 32 * mblk;
            const int idx = k2 <= H ? k2 : M - k2;
            const bool real0 = idx == 0 || idx == H;  // packed real columns
            const float2 a = Qa[real0 ? 0 : idx], bq = Qb[real0 ? 0 : idx];
            float2 cv;
            if (real0) {
                const float ar = idx == 0 ? a.x : a.y, br_ = idx == 0 ? bq.x : bq.y;
                cv = make_float2(ar, zb * br_);
            } else if (k2 < H) {
                cv = make_float2(a.x - zb * bq.y, a.y + zb * bq.x);  // Qa + i Qb
            } else {
                cv = make_float2(a.x + zb * bq.y, -a.y + zb * bq.x);  // conj(Qa) + i conj(Qb)
            }
            c[k] = cv;
        }
        line_inv<L>(c, xl);  // c[n] = r(qa) + i r(qb) at l2 = L n + l
        const float scale = keep / (float)(M * M);
        const int ipa = LO + qa, ipb = LO + qb;
        sfor<32>([&](auto nn) {
            constexpr int n = decltype(nn)::value;
            const int l2 = L * n + l;
            const int jp = (l2 + H) & (M - 1);
            if (jp >= LO && jp < LO + NB) {
                const int q2 = jp - LO;
                Rb[qa + q2 * NB] = fabsf(c[n].x * scale) * cfg.gain[ipa + jp * M];
                if (hasb) Rb[qb + q2 * NB] = fabsf(c[n].y * scale) * cfg.gain[ipb + jp * M];
            }
        });
    }
    __syncthreads();

    if (out.acc) {
        // accumulate_planes!: add the box to the window's FP64 plane sum
        if (active) {
            double *dst = out.acc + (size_t)jb * (NB * NB);
            for (int p = t; p < NB * NB; p += G) dst[p] += (double)Rb[p];
        }
        return;
    }

    analyze_box<M, K, NB, LO, G, NWAVE, UQ>(cfg, out, Rb, t, tid, mv, ml, active, jb, gi, gj, sig_u, sig_v);
}

template <int L, int K, bool PHASE, int HW, bool UQ>
void launchL(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs, int njobs,
             const CorrOut &out, hipStream_t s)
{
    constexpr int G = 16 * L * L;
    constexpr int WPB = G >= 64 ? 1 : 64 / G;
    constexpr int NT = G >= 64 ? G : 64;
    const int nblocks = (njobs + WPB - 1) / WPB;
    hipLaunchKernelGGL((k_corrL<L, K, PHASE, HW, UQ>), dim3(nblocks), dim3(NT), 0, s, cfg, A, B, nr, jobs, njobs,
                       nblocks, out);
}

template <int L, int K, bool PHASE, bool UQ>
bool launch_hw(const CorrConfig &cfg, int hw, const float *A, const float *B, int nr, const int32_t *jobs, int njobs,
               const CorrOut &out, hipStream_t s)
{
    // half-widths of the Gaussian and of the uniform apodization support
    constexpr int HG = L == 1 ? 7 : (L == 2 ? 13 : 25);
    constexpr int HN = L == 1 ? 9 : (L == 2 ? 17 : 33);
    if (hw == HG) {
        launchL<L, K, PHASE, HG, UQ>(cfg, A, B, nr, jobs, njobs, out, s);
        return true;
    }
    if (hw == HN) {
        launchL<L, K, PHASE, HN, UQ>(cfg, A, B, nr, jobs, njobs, out, s);
        return true;
    }
    return false;
}

template <int L, bool UQ>
bool launch_L(const CorrConfig &cfg, int hw, const float *A, const float *B, int nr, const int32_t *jobs, int njobs,
              const CorrOut &out, hipStream_t s)
{
    if (cfg.method == HH_PHASE)
        return cfg.K == 2 ? launch_hw<L, 2, true, UQ>(cfg, hw, A, B, nr, jobs, njobs, out, s)
                          : launch_hw<L, 3, true, UQ>(cfg, hw, A, B, nr, jobs, njobs, out, s);
    return cfg.K == 2 ? launch_hw<L, 2, false, UQ>(cfg, hw, A, B, nr, jobs, njobs, out, s)
                      : launch_hw<L, 3, false, UQ>(cfg, hw, A, B, nr, jobs, njobs, out, s);
}

}  // namespace

// Half-width of the analysis box when the register kernels cover cfg, else -1.
static int fast_halfwidth(const CorrConfig &cfg)
{
    if (!cfg.padded || !cfg.gain || cfg.finder != HH_REGIONALMAX || cfg.K > 3) return -1;
    if (!(cfg.M == 32 || cfg.M == 64 || cfg.M == 128) || cfg.w * 2 != cfg.M) return -1;
    // the box must be symmetric about the zero lag with a zero margin, and be
    // one of the two compiled widths (Gaussian or uniform apodization support)
    const int H = cfg.M / 2;
    const int hw = H - cfg.box_lo;
    if (cfg.box_hi != H + hw || cfg.box_lo < 1 || cfg.box_hi > cfg.M - 3) return -1;
    const int L = cfg.M / 32;
    const int hg = L == 1 ? 7 : (L == 2 ? 13 : 25), hn = L == 1 ? 9 : (L == 2 ? 17 : 33);
    return hw == hg || hw == hn ? hw : -1;
}

bool correlate_box_layout(const CorrConfig &cfg) { return fast_halfwidth(cfg) >= 0; }

bool launch_uq16_mfma(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs, int njobs,
                      const CorrOut &out, hipStream_t s)
{
    if (cfg.w != 16 || !out.unc_u || !out.unc_v) return false;
    if (njobs <= 0) return true;
    const int nblocks = (njobs + 3) / 4;
    hipLaunchKernelGGL(k_uq16_mfma, dim3(nblocks), dim3(64), 0, s, cfg, A, B, nr, jobs, njobs, nblocks, out);
    HH_HIP(hipGetLastError());
    return true;
}

const char *fast_family(const CorrConfig &cfg)
{
    if (fast_halfwidth(cfg) < 0) return nullptr;
    return cfg.M == 32 ? "regL1" : (cfg.M == 64 ? "regL2" : "regL4");
}

bool launch_correlate_fast(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs,
                           int njobs, const CorrOut &out, hipStream_t s, bool *uq_done)
{
    *uq_done = false;
    if (out.planes) return false;
    const int hw = fast_halfwidth(cfg);
    if (hw < 0) return false;
    const bool uq = out.unc_u != nullptr && out.unc_v != nullptr && cfg.M == 32 && !out.acc;
    if (njobs <= 0) {
        *uq_done = uq;
        return true;
    }
    bool ok = false;
    switch (cfg.M) {
    case 32:
        ok = uq ? launch_L<1, true>(cfg, hw, A, B, nr, jobs, njobs, out, s)
                : launch_L<1, false>(cfg, hw, A, B, nr, jobs, njobs, out, s);
        break;
    case 64: ok = launch_L<2, false>(cfg, hw, A, B, nr, jobs, njobs, out, s); break;
    case 128: ok = launch_L<4, false>(cfg, hw, A, B, nr, jobs, njobs, out, s); break;
    }
    if (ok) {
        HH_HIP(hipGetLastError());
        *uq_done = uq;
    }
    return ok;
}

}  // namespace hh
