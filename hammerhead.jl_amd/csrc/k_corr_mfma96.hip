// k_corr_mfma96.hip — correlation of 64-px windows as split-FP16 DFT products
// on the gfx950 matrix cores (HH_FORM_MFMA16, cross correlation).
//
// Same per-window semantics as the FP32 kernels (correlation_plane! +
// analyze_plane!, src/correlators.jl:326-341,459-468) for 64 x 64 windows
// with the reference's padding to 128: the plane the analysis reads is the
// linear correlation of the two apodized, mean-subtracted windows on the box
// |lag| <= 25 (outside it the overlap gain is exactly 0). The padded 128-point
// transforms give that linear correlation at every lag; so does the circular
// correlation of any length N with N - 25 > 63, because the linear
// correlation of two 64-sample supports vanishes beyond |lag| = 63 and no
// alias then reaches the box. This kernel uses N = 96 = 3 matrix tiles of 32
// in both dimensions (the 1/N^2 of the inverse is the 1/128^2 of the
// reference on the same lags). The phase filter is defined on the 128-point
// spectrum, so phase correlation stays on the FFT kernels.
//
// The stages follow the 32-px kernel (k_corr_mfma.hip), one wave per window:
//   1. D1[I][rt][J] = X_I (rows r of tile rt, K = c in 4 steps) x F1[.][J]:
//      the column DFT, parts of tile J = Re k2 = 16J..16J+15 (rows 0-15) and
//      Im of the same k2 (rows 16-31; the Im slot of k2 = 0 carries the real
//      Nyquist column k2 = 48), for all three tiles J at once; split as the
//      stage-2 operands and held for the window (192 registers, partly AGPRs);
//   2. per (k1 tile nt, J): D1^T (K = r in 4 steps) x [cos | -sin](r, k1):
//      Z(k1, k2), the cross spectrum P, and
//   3. at once its contribution to Q[Re | Im][lt] += P (K = Re/Im of k2) x
//      table: the inverse along k2 for the 51 box columns l2 (2 column tiles
//      lt), Hermitian weight 2 in the tables; k2 = 0 and 48 enter as
//      H = Re P - Im P in the K slots of k2 = 0;
//   4. once Q of tile nt is complete, R[lt1][lt2] += [2cos | -2sin](l1, k1) x
//      [Q_Re; Q_Im](k1, l2) (K = the 32 k1 of the tile in 2 steps): the tables
//      of stage 3 as the A operand (rows l1), Q as B, R accumulated whole.
// The separable Gaussian apodization is folded into the tables (its column
// factor into stage 1, its row factor into stage 2), so the pixels enter as
// (p - mean) 2^e. The inverse tables of stage 3 and the A operands of stage 4
// are one set of 2cos / -2sin tables of (k, l) (k = 16 m + the K slot order,
// l = the box lags by lane), so the LDS holds the stage-2 tables (24 split
// 32 x 16 operands), that set (24) and the k2 = 0 variants of stage 3 (8):
// 112 KB, beside one 51 x 51 analysis box per wave; the stage-1 tables (12)
// are read from a device buffer (24 KB, L2-resident). One block of 4 waves
// per CU (one wave per SIMD with the whole register file); persistent waves
// walk the windows at a stride (wave_windows).

#include "hh_internal.h"
#include "box_peaks.h"
#include "mfma_split.h"

#include <cmath>
#include <mutex>

namespace hh {

namespace {

using namespace mfma;

constexpr int kN96 = 96;
constexpr int kHW96 = 25;             // box half-width (Gaussian apodization of 64-px windows)
constexpr int kNB96 = 2 * kHW96 + 1;  // 51 box rows and columns
constexpr int kS0_96 = -5;            // input scale: |x| < 2^-5 (D1 < 2, Z < 2^7, P < 2^14)
constexpr int kS4_96 = -8;            // stage-4 operand scale (96 K slots of |P| < 2^14, weight 2)

struct Trig96 {
    float c[kN96], s[kN96];  // cos, sin of 2 pi t / 96
    // the Gaussian apodization of a 64-px window is separable,
    // exp(-((i - 32.5)^2 + (j - 32.5)^2) / (2 * 16^2)) = ap[i - 1] ap[j - 1]
    // (apodization_window, src/correlators.jl:18-25): its row factor goes
    // into the stage-2 tables and its column factor into the stage-1 tables
    float ap[64];
};

enum {
    T96_WC = 0,    // [kk][nt]: 4 x 3
    T96_WS = 12,   // [kk][nt]
    T96_TT = 24,   // [m][C | S][lt]: 6 x 2 x 2
    T96_J0 = 48,   // [src Re | Im][C | S][lt]: 2 x 2 x 2
    T96_N = 56,
    F96_N = 12     // stage-1 tables [kk][J]: 4 x 3, in the device buffer
};

// Lag l of lane n in box column tile lt (0..50 valid) as an index 0..50, or -1.
__device__ __forceinline__ int box_idx(int lt, int n)
{
    const int b = 32 * lt + n;
    return b < kNB96 ? b : -1;
}

// LDS operand table q for lane l (slot s of half h holds K index rmap(h, s)
// except for the stage-1 tables, whose K is the natural 8 h + s).
__device__ void build_table96(int q, int l, const Trig96 &tg, Split &out)
{
    const int h = l >> 5, n = l & 31;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        float f = 0.f;
        if (q < T96_TT) {  // Wc / Ws[kk][nt]: K = r = 16 kk + rmap, column k1 = 32 nt + n
            const bool sn = q >= T96_WS;
            const int i = q - (sn ? T96_WS : T96_WC);
            const int kk = i / 3, nt = i % 3;
            const int r = 16 * kk + rmap(h, s);
            const int t = ((32 * nt + n) * r) % kN96;
            f = tg.ap[r] * (sn ? -tg.s[t] : tg.c[t]);
        } else {
            int m, cs, lt, src = -1;
            if (q < T96_J0) {
                const int i = q - T96_TT;
                m = i >> 2;
                cs = (i >> 1) & 1;
                lt = i & 1;
            } else {
                const int i = q - T96_J0;
                m = 0;
                src = i >> 2;
                cs = (i >> 1) & 1;
                lt = i & 1;
            }
            const int b = box_idx(lt, n);
            if (b >= 0) {
                const int lag = b - kHW96;
                const int k = 16 * m + rmap(h, s);
                const int t = ((k * lag) % kN96 + kN96) % kN96;
                if (src >= 0 && k == 0) {
                    // the K slots of k2 = 0 carry H(k1, 0) (Re slots) and
                    // H(k1, 48) (Im slots), H = Re P - Im P: contributions
                    // H (1 - i) and H (1 - i) (-1)^l2 to Q
                    const float sg = (lag & 1) ? -1.f : 1.f;
                    f = src == 0 ? 1.f : (cs == 0 ? -sg : sg);
                } else {
                    f = cs == 0 ? 2.f * tg.c[t] : -2.f * tg.s[t];
                }
            }
        }
        put(out, s, f);
    }
}

// Stage-1 tables F1[kk][J] (K = c = 16 kk + 8 h + s, column = part n of tile
// J) times the column apodization, written once per device into the buffer
// the kernel reads.
__global__ __launch_bounds__(64) void k_build_f1_96(Trig96 tg, Split *f1)
{
    const int q = blockIdx.x, l = threadIdx.x;
    const int h = l >> 5, n = l & 31;
    const int kk = q / 3, J = q % 3;
    Split out;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int c = 16 * kk + 8 * h + s;
        float f;
        if (n < 16) {
            f = tg.c[((16 * J + n) * c) % kN96];
        } else {
            const int k2 = 16 * J + n - 16;
            f = k2 == 0 ? ((c & 1) ? -1.f : 1.f) : -tg.s[(k2 * c) % kN96];
        }
        put(out, s, tg.ap[c] * f);
    }
    f1[q * 64 + l] = out;
}

__device__ __forceinline__ Split neg_split(const Split &a)
{
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 m = {0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    Split r;
    r.hi = __builtin_bit_cast(h8, __builtin_bit_cast(u4, a.hi) ^ m);
    r.lo = __builtin_bit_cast(h8, __builtin_bit_cast(u4, a.lo) ^ m);
    return r;
}

// sum and maximum over all 64 lanes of the wave
__device__ __forceinline__ float wave_sum(float v)
{
    v += rowror_f<8>(v);
    v += rowror_f<4>(v);
    v += rowror_f<2>(v);
    v += rowror_f<1>(v);
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}
__device__ __forceinline__ float wave_max(float v)
{
    v = fmaxf(v, rowror_f<8>(v));
    v = fmaxf(v, rowror_f<4>(v));
    v = fmaxf(v, rowror_f<2>(v));
    v = fmaxf(v, rowror_f<1>(v));
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}

template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_corr_mfma96(
    CorrConfig cfg, Trig96 tg, const Split *__restrict__ f1, const float *__restrict__ A,
    const float *__restrict__ B, int nr, const int32_t *__restrict__ jobs, int njobs, CorrOut out)
{
#pragma clang fp contract(off)
    constexpr int M = 128, H = M / 2, LO = H - kHW96;
    __shared__ Split tabs[T96_N][64];
    __shared__ float rbox[4][kNB96 * kNB96 + 1];
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const int h = l >> 5, n = l & 31;
    for (int q = wave; q < T96_N; q += 4) {
        Split t;
        build_table96(q, l, tg, t);
        tabs[q][l] = t;
    }
    // the output pointers, read from LDS where the box is analyzed: held in
    // scalar registers through the loop they spill to VGPR lanes
    __shared__ CorrOut sout[1];
    if (tid == 0) sout[0] = out;
    __syncthreads();
    const f16v zero = {};
    const float neg1 = opaque_neg1();

    const WaveWindows ww = wave_windows(cfg.linear_order, gridDim.x, blockIdx.x, 4, wave, njobs);
    if (ww.n <= 0) return;
    const int4 *jobs4 = reinterpret_cast<const int4 *>(jobs);
    float *Rb = rbox[wave];
    for (int k = 0; k < ww.n; ++k) {
        const int jb = ww.base + k * ww.stride;
        const int4 jw = jobs4[jb];
        int z;
        asm volatile("v_mov_b32 %0, 0" : "=v"(z));
        const int tl = l + z;  // an opaque lane index keeps the table reads in the loop
        // ---- 0. means of both windows and the scales that put them into
        // |x| < 2^kS0 (lane (h, n) holds rows 32 rt + n, columns 16 kk + 8 h + s) ----
        const float *pI[2];
        float meanI[2];
        int eI[2];
#pragma unroll
        for (int I = 0; I < 2; ++I) {
            pI[I] = (I ? B : A) + (size_t)(jw.w - 1 + 8 * h) * nr + (jw.z - 1 + n);
            float px[64];
            float sum = 0.f;
#pragma unroll
            for (int q = 0; q < 64; ++q) {
                const int rt = q >> 5, kk = (q >> 3) & 3, s = q & 7;
                px[q] = pI[I][(size_t)(16 * kk + s) * nr + 32 * rt];
                sum += px[q];
            }
            const float mean = wave_sum(sum) / 4096.f;
            float amx = 0.f;
#pragma unroll
            for (int q = 0; q < 64; ++q) amx = fmaxf(amx, fabsf(px[q] - mean));
            const float bound = wave_max(amx);  // >= |x| (the apodization, <= 1, is in the tables)
            int ex = 0;
            (void)frexpf(bound, &ex);
            meanI[I] = mean;
            eI[I] = bound > 0.f ? kS0_96 - ex : 0;
        }
        const int es = eI[0] + eI[1];
        // ---- 1. column DFT of both windows for every tile J, split as the
        // stage-2 operand (K = r = 16 kk + rmap): held for the whole window
        // (some of it in AGPRs) instead of reloading the pixels per tile ----
        Split a2[2][3][4];  // [I][J][kk]
#pragma unroll
        for (int I = 0; I < 2; ++I) {
            const float *p = pI[I];
            // (p - mean) 2^e0 as one FMA: the scale is a power of two, so
            // p 2^e0 - mean 2^e0 rounds exactly as p - mean does
            const float sc = ldexpf(1.f, eI[I]), msc = -meanI[I] * sc;
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) {
                Split a1[4];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                    for (int s = 0; s < 8; s += 2)
                        put2(a1[kk], s, __builtin_fmaf(p[(size_t)(16 * kk + s) * nr + 32 * rt], sc, msc),
                             __builtin_fmaf(p[(size_t)(16 * kk + s + 1) * nr + 32 * rt], sc, msc), neg1);
#pragma unroll
                for (int J = 0; J < 3; ++J) {
                    f16v d1 = mfma3(a1[0], f1[(0 * 3 + J) * 64 + tl], zero);
#pragma unroll
                    for (int kk = 1; kk < 4; ++kk) d1 = mfma3(a1[kk], f1[(kk * 3 + J) * 64 + tl], d1);
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int s = 0; s < 8; s += 2)
                            put2(a2[I][J][2 * rt + q], s, d1[8 * q + s], d1[8 * q + s + 1], neg1);
                }
            }
        }
        f16v r4[2][2];  // R[lt1][lt2]
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) r4[a][b] = zero;
#pragma unroll
        for (int nt = 0; nt < 3; ++nt) {
            f16v q3[2][2];  // Q[Re | Im][lt] of k1 tile nt: rows k1, columns l2
#pragma unroll
            for (int ri = 0; ri < 2; ++ri)
#pragma unroll
                for (int lt = 0; lt < 2; ++lt) q3[ri][lt] = zero;
#pragma unroll
            for (int J = 0; J < 3; ++J) {
                const bool j0 = J == 0;
                // ---- 2. row DFT: Z(k1 = 32 nt + n, k2 = 16 J + rmap(h, v)) ----
                float zr[2][9], zi[2][9];
#pragma unroll
                for (int I = 0; I < 2; ++I) {
                    f16v dc = mfma3(a2[I][J][0], tabs[T96_WC + 0 * 3 + nt][tl], zero);
                    f16v ds = mfma3(a2[I][J][0], tabs[T96_WS + 0 * 3 + nt][tl], zero);
#pragma unroll
                    for (int kk = 1; kk < 4; ++kk) {
                        dc = mfma3(a2[I][J][kk], tabs[T96_WC + kk * 3 + nt][tl], dc);
                        ds = mfma3(a2[I][J][kk], tabs[T96_WS + kk * 3 + nt][tl], ds);
                    }
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
                        zr[I][v] = dc[v] - ds[v + 8];
                        zi[I][v] = dc[v + 8] + ds[v];
                    }
                    if (j0 && h == 0) {  // k2 = 0 (no Im part) and the Nyquist column in its slot
                        zr[I][0] = dc[0];
                        zi[I][0] = ds[0];
                    }
                    zr[I][8] = dc[8];
                    zi[I][8] = ds[8];
                }
                // ---- 3. spectrum; inverse along k2 into Q ----
                float sp_r[9], sp_i[9];
#pragma unroll
                for (int v = 0; v < 9; ++v) {
                    sp_r[v] = __builtin_fmaf(zr[0][v], zr[1][v], zi[0][v] * zi[1][v]);
                    sp_i[v] = __builtin_fmaf(zr[0][v], zi[1][v], -(zi[0][v] * zr[1][v]));
                }
                if (j0 && h == 0) {
                    const float h0 = sp_r[0] - sp_i[0], h48 = sp_r[8] - sp_i[8];
                    sp_r[0] = h0;
                    sp_i[0] = h48;
                }
                Split a3r, a3i;
#pragma unroll
                for (int s = 0; s < 8; s += 2) {
                    put2(a3r, s, sp_r[s], sp_r[s + 1], neg1);
                    put2(a3i, s, sp_i[s], sp_i[s + 1], neg1);
                }
                const Split na3r = neg_split(a3r);
                // Q_Re += P_Re C + P_Im S, Q_Im += (-P_Re) S + P_Im C, with the
                // k2 = 0 variants of the tables for tile 0
#pragma unroll
                for (int lt = 0; lt < 2; ++lt) {
                    const int cr = j0 ? T96_J0 + 0 + lt : T96_TT + J * 4 + 0 + lt;  // src Re, C
                    const int sr = j0 ? T96_J0 + 2 + lt : T96_TT + J * 4 + 2 + lt;  // src Re, S
                    const int ci = j0 ? T96_J0 + 4 + lt : T96_TT + J * 4 + 0 + lt;  // src Im, C
                    const int si = j0 ? T96_J0 + 6 + lt : T96_TT + J * 4 + 2 + lt;  // src Im, S
                    q3[0][lt] = mfma3(a3r, tabs[cr][tl], q3[0][lt]);
                    q3[0][lt] = mfma3(a3i, tabs[si][tl], q3[0][lt]);
                    q3[1][lt] = mfma3(na3r, tabs[sr][tl], q3[1][lt]);
                    q3[1][lt] = mfma3(a3i, tabs[ci][tl], q3[1][lt]);
                }
            }
            // ---- 4. R[lt1][lt2] += [2cos | -2sin](l1, k1) x [Q_Re; Q_Im](k1, l2)
            // over the k1 of tile nt ----
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                Split bq[2][2];  // [Re | Im][lt2]
#pragma unroll
                for (int ri = 0; ri < 2; ++ri)
#pragma unroll
                    for (int lt = 0; lt < 2; ++lt)
#pragma unroll
                        for (int s = 0; s < 8; s += 2)
                            put2(bq[ri][lt], s, ldexpf(q3[ri][lt][8 * kk + s], kS4_96),
                                 ldexpf(q3[ri][lt][8 * kk + s + 1], kS4_96), neg1);
                const int m = 2 * nt + kk;
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const Split &tc = tabs[T96_TT + m * 4 + 0 + a][tl];
                    const Split &ts = tabs[T96_TT + m * 4 + 2 + a][tl];
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        r4[a][b] = mfma3(tc, bq[0][b], r4[a][b]);
                        r4[a][b] = mfma3(ts, bq[1][b], r4[a][b]);
                    }
                }
            }
        }
        // the tables carry the Hermitian weight 2 throughout: R / 2, and the
        // inverse normalization 1 / 96^2
        const int etot = es + kS4_96;
        const float scale = ldexpf(0.5f / (float)(kN96 * kN96), -etot);
        // lane column l2 = 32 b + n - 25, rows l1 = 32 a + rmap(h, v) - 25
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int cb = 32 * b + n;
            if (cb < kNB96) {
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int ra = 32 * a + rmap(h, v);
                        if (ra < kNB96)
                            Rb[ra + cb * kNB96] =
                                fabsf(r4[a][b][v] * scale) * cfg.gain[(LO + ra) + (LO + cb) * M];
                    }
            }
        }
        wave_sync();
        float sg = NAN;
        int zo;
        asm volatile("v_mov_b32 %0, 0" : "=v"(zo));
        analyze_box<M, K, kNB96, LO, 64, 1, false>(cfg, sout[zo], Rb, l, l, nullptr, nullptr, true, jb, jw.x, jw.y,
                                                   sg, sg);
        wave_sync();
    }
}

}  // namespace

// covered: 64-px windows padded to 128 with the Gaussian support box, cross
bool mfma96_covers(const CorrConfig &cfg)
{
    if (!cfg.padded || !cfg.gain || cfg.finder != HH_REGIONALMAX || cfg.K > 3) return false;
    if (cfg.method != HH_CROSS || cfg.M != 128 || cfg.w != 64 || cfg.s != 64) return false;
    return cfg.box_lo == 64 - kHW96 && cfg.box_hi == 64 + kHW96;
}

static std::mutex g_f1_mu;
static std::vector<Split *> g_f1;  // per device

static const Trig96 &trig96()
{
    static const Trig96 t = [] {
        Trig96 r;
        for (int k = 0; k < kN96; ++k) {
            r.c[k] = (float)std::cos(2.0 * M_PI * k / (double)kN96);
            r.s[k] = (float)std::sin(2.0 * M_PI * k / (double)kN96);
        }
        for (int i = 0; i < 64; ++i) {
            const double d = (i + 1) - 32.5;
            r.ap[i] = (float)std::exp(-d * d / (2.0 * 16.0 * 16.0));
        }
        return r;
    }();
    return t;
}

// The stage-1 table buffer of the current device, built on first use (a
// device allocation: called while planning, outside any stream capture).
void mfma96_prepare(hipStream_t s)
{
    int dev = 0;
    HH_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_f1_mu);
    if ((int)g_f1.size() <= dev) g_f1.resize(dev + 1, nullptr);
    if (g_f1[dev]) return;
    Split *p = nullptr;
    HH_HIP(hipMalloc(&p, sizeof(Split) * 64 * F96_N));
    hipLaunchKernelGGL(k_build_f1_96, dim3(F96_N), dim3(64), 0, s, trig96(), p);
    HH_HIP(hipGetLastError());
    HH_HIP(hipStreamSynchronize(s));
    g_f1[dev] = p;
}

bool launch_correlate_mfma96(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs,
                             int njobs, const CorrOut &out, hipStream_t s)
{
    if (!mfma96_covers(cfg)) return false;
    if (out.planes || out.acc || out.unc_u) return false;
    if (njobs <= 0) return true;
    int dev = 0;
    HH_HIP(hipGetDevice(&dev));
    Split *f1 = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_f1_mu);
        if (dev < (int)g_f1.size()) f1 = g_f1[dev];
    }
    if (!f1) {
        // first use outside a prepared schedule (never inside a capture: a
        // schedule prepares its passes before recording)
        mfma96_prepare(s);
        std::lock_guard<std::mutex> lk(g_f1_mu);
        f1 = g_f1[dev];
    }
    const int nblocks = std::max(1, std::min((njobs + 3) / 4, device_cu_count()));
    if (cfg.K == 2)
        hipLaunchKernelGGL(k_corr_mfma96<2>, dim3(nblocks), dim3(256), 0, s, cfg, trig96(), f1, A, B, nr, jobs,
                           njobs, out);
    else
        hipLaunchKernelGGL(k_corr_mfma96<3>, dim3(nblocks), dim3(256), 0, s, cfg, trig96(), f1, A, B, nr, jobs,
                           njobs, out);
    HH_HIP(hipGetLastError());
    return true;
}

}  // namespace hh
