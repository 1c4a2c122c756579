// k_corr_mfma.hip — correlation of 16-px windows padded to 32 x 32 as
// dense DFT products on the gfx950 matrix cores (HH_FORM_MFMA16).
//
// Same per-window semantics as the FP32 kernels (correlation_plane! +
// analyze_plane!, src/correlators.jl:326-341,459-468). The padded DFT of a
// 16 x 16 window X is F32 X F32^T restricted to the 16 nonzero rows and
// columns, and the correlation plane is only needed on the analysis box
// (lags -7..7; outside it the overlap gain is exactly 0), so the whole
// correlation is four small real GEMM stages per window on
// v_mfma_f32_32x32x16_f16, each operand split into FP16 hi + lo parts
// (hi*hi + hi*lo + lo*hi, FP32 accumulation, ~22-bit operands):
//   1. [XA; XB] (32 x 16, rows = (image, r), K = c) x F1 (16 x 32): the
//      column DFT of both windows, columns = Re k2 = 0..16, Im k2 = 1..15;
//   2. per image, T1^T (rows = those 32 parts, K = r) x [cos | -sin](r, k1):
//      the row DFT, all k1 = 0..31; Re/Im parts of one k2 land in the same
//      lane and combine into Z(k1, k2), k2 = 0..16 (Hermitian half);
//   3. the spectrum conj(ZA) ZB (or the filtered phase spectrum) per lane,
//      then P (rows k1, K = Re/Im of k2) x table: the inverse along k2 for
//      the 15 box columns l2, Hermitian weights folded into the table. The
//      columns k2 = 0 and 16 are Hermitian in k1, so each enters as the one
//      real number H = Re P - Im P (in the K slots of Re and Im k2 = 0)
//      against table columns that make it H (1 - i) e^(i th k2 l2): summed
//      over k1 in stage 4 that gives sum H (cos + sin) = sum (Re P cos -
//      Im P sin), the same contribution, and K = 32 fits two K steps;
//   4. Q^T (rows = (l2, Re/Im), K = k1) x [cos | sin](k1, l1), the cos
//      columns of the 15 box rows l1 in lanes 0-15 and their sin columns in
//      lanes 16-31 of one product: the inverse along k1; the Re rows of the
//      cos lanes minus the Im rows of the sin lanes (one lane exchange) give R.
// Every stage feeds the next from registers: an MFMA result holds one
// column per lane and 16 rows in a fixed pattern, which is exactly an
// operand of the next product with K permuted the same way in the constant
// table, so no data moves between lanes or through LDS until the box is
// written for the peak search. Operands are scaled by powers of two to sit
// in FP16 range (a per-window scale at the input, fixed ones after) and the
// scale is removed exactly before the magnitude and gain.
// One wave per window; persistent waves walk the windows at a stride (see
// wave_windows), the constant operand tables staged once per block in LDS.

#include "hh_internal.h"
#include "box_peaks.h"
#include "mfma_split.h"

#include <cfloat>
#include <cmath>

namespace hh {

namespace {

using namespace mfma;

struct Trig {
    float c[32], s[32];  // cos, sin of 2 pi t / 32
    // the separable Gaussian apodization of a 16-px window (apodization_window,
    // src/correlators.jl:18-25: exp(-((i - 8.5)^2 + (j - 8.5)^2) / (2 * 4^2))
    // = ap[i - 1] ap[j - 1]): the column factor goes into the stage-1 table,
    // the row factor into the stage-2 tables
    float ap[16];
};

constexpr int kHW = 7;            // box half-width (Gaussian apodization)
constexpr int kNB = 2 * kHW + 1;  // 15 box rows and columns
// Operand scales. The input is scaled to |x| < 2^-1, so the stage-2 operand
// (16-term sums) stays below 2^3, Z below 2^7 and the cross spectrum below
// 2^14 (H = Re P - Im P below 2^14.5): stages 2 and 3 need no rescaling.
// The phase spectrum (|P| <= 1) is
// raised to 2^12. The stage-3 tables (weights <= 2) carry 2^kS4, so that the
// stage-3 result enters stage 4 without a rescale: by Parseval the cross
// spectrum of one k1 sums to at most sum |Z_A|^2 = 1024 * 256 * 2^-2 = 2^16
// in magnitude, which bounds the result by 2 * 2^16 * 2^-4 = 2^13 (phase:
// 17 bins of 2^12 give 2^13.1), inside FP16. The smallest table entries
// still split to within 2^-25 absolute.
constexpr int kS3P = 12;  // stage-3 operand scale (phase spectrum)
constexpr int kS4 = -4;   // stage-3 table scale

// W4: registers capped at 128 per lane for four resident waves per SIMD
// (a few bytes spill) in blocks of 8 waves (two blocks per CU share their
// LDS tables), else the natural 168 (three waves) in blocks of 4
template <int K, bool PHASE, bool W4>
__global__ __launch_bounds__(W4 ? 512 : 256) __attribute__((amdgpu_waves_per_eu(W4 ? 4 : 1))) void k_corr_mfma32(
    CorrConfig cfg, Trig tg, const float *__restrict__ A, const float *__restrict__ B, int nr,
    const int32_t *__restrict__ jobs, int njobs, CorrOut out)
{
#pragma clang fp contract(off)
    constexpr int M = 32, H = 16, LO = H - kHW;
    // windows analyzed together: 8, 8 lanes each (16 lanes each would share
    // the peak search's fixed costs per group, the candidate merges, fits,
    // moment and output writes, among half as many windows)
    constexpr int WPB = W4 ? 8 : 4;  // waves per block
    constexpr int GW = 8, GL = 64 / GW;
    __shared__ float rbox[WPB][GW][kNB * kNB + 1];  // [wave][window of the group]
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const int h = l >> 5, n = l & 31;

    // ---- constant operand tables: one copy per block in LDS, read per use
    // (registers are left to the accumulators) ----
    __shared__ Split tabs[7][64];
    if (wave == 0) {
        Split F1, WrT, WiT, T3[2], CS4[2];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int c = 8 * h + s;  // stage 1: K = c, column n
            float f;
            if (n < 16)
                f = tg.c[(c * n) & 31];
            else if (n == 16)
                f = (c & 1) ? -1.f : 1.f;
            else
                f = -tg.s[(c * (n - 16)) & 31];
            put(F1, s, tg.ap[c] * f);
            const int r = rmap(h, s);  // stage 2: K = r, column k1 = n
            put(WrT, s, tg.ap[r] * tg.c[(n * r) & 31]);
            put(WiT, s, tg.ap[r] * -tg.s[(n * r) & 31]);
            // stage 3: K = (Re | Im of k2 = rmap(h, s)), column n = (b = n & 15,
            // Re for n < 16 | Im), l2 = b - HW; the K slots of k2 = 0 carry
            // H(k1, 0) (Re table) and H(k1, 16) (Im table)
            const int b = n & 15, im = n >> 4;
            const bool vb = b < kNB;
            const int l2 = b - kHW;
            const int k2 = rmap(h, s);
            const int t = (k2 * l2) & 31;
            const float sgn16 = tg.c[(16 * l2) & 31];  // (-1)^l2 = e^(i pi l2)
            float t3r, t3i;
            if (k2 == 0) {
                t3r = im ? -1.f : 1.f;        // H(k1, 0) (1 - i)
                t3i = im ? -sgn16 : sgn16;    // H(k1, 16) (1 - i) (-1)^l2
            } else {
                t3r = im ? 2.f * tg.s[t] : 2.f * tg.c[t];   // 2 Re P e^(i th k2 l2)
                t3i = im ? 2.f * tg.c[t] : -2.f * tg.s[t];  // 2 Im P ...
            }
            // scaled by 2^kS4: the stage-3 result then enters stage 4 as is
            put(T3[0], s, ldexpf(vb ? t3r : 0.f, kS4));
            put(T3[1], s, ldexpf(vb ? t3i : 0.f, kS4));
            // stage 4: K = k1 = 16 q + rmap(h, s), column n = (a = n & 15,
            // cos for n < 16 | sin), l1 = a - HW
            const int a = n & 15, sn = n >> 4;
            const bool va = a < kNB;
            const int l1 = a - kHW;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int k1 = 16 * q + r;
                const int t4 = (k1 * l1) & 31;
                put(CS4[q], s, va ? (sn ? tg.s[t4] : tg.c[t4]) : 0.f);
            }
        }
        tabs[0][l] = F1;
        tabs[1][l] = WrT;
        tabs[2][l] = WiT;
        tabs[3][l] = T3[0];
        tabs[4][l] = T3[1];
        tabs[5][l] = CS4[0];
        tabs[6][l] = CS4[1];
    }
    // the output pointers are read from LDS where the boxes are analyzed:
    // held in scalar registers through the loop they spill to VGPR lanes and
    // cost a v_readlane per use
    __shared__ CorrOut sout[1];
    if (tid == 0) sout[0] = out;
    __syncthreads();
    const f16v zero = {};
    const float neg1 = opaque_neg1();
    const int img = n >> 4, row = n & 15;
    const float *src = img ? B : A;
    // overlap gain of this lane's box points (row n, columns rmap(h, v))
    float gain8[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const int bc = rmap(h, v);
        gain8[v] = (n < kNB && bc < kNB) ? cfg.gain[(LO + n) + (LO + bc) * M] : 0.f;
    }

    const WaveWindows ww = wave_windows(cfg.linear_order, gridDim.x, blockIdx.x, WPB, wave, njobs);
    if (ww.n <= 0) return;
    // software pipeline: the pixels of window k + 1 and the job of k + 2 are
    // in flight while window k is processed
    const int4 *jobs4 = reinterpret_cast<const int4 *>(jobs);
    int4 jcur = jobs4[ww.base];
    int4 jnext = jobs4[ww.base + min(1, ww.n - 1) * ww.stride];
    float pnext[8];
    {
        const float *p = src + (size_t)(jcur.w - 1 + 8 * h) * nr + (jcur.z - 1 + row);
#pragma unroll
        for (int s = 0; s < 8; ++s) pnext[s] = p[(size_t)s * nr];
    }
    // windows in groups of GW: the boxes of a group are analyzed together,
    // GL lanes per window
    for (int g = 0; g < ww.n; g += GW) {
        const int ng = min(GW, ww.n - g);
        for (int j = 0; j < ng; ++j) {
        const int k = g + j;
        float *Rb = rbox[wave][j];
        // an opaque zero keeps the table reads inside the loop
        int z;
        asm volatile("v_mov_b32 %0, 0" : "=v"(z));
        const int tl = l + z;
        float px[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) px[s] = pnext[s];
        {
            const float *p = src + (size_t)(jnext.w - 1 + 8 * h) * nr + (jnext.z - 1 + row);
#pragma unroll
            for (int s = 0; s < 8; ++s) pnext[s] = p[(size_t)s * nr];
        }
        jcur = jnext;
        jnext = jobs4[ww.base + min(k + 2, ww.n - 1) * ww.stride];
        // ---- 0. window: lane (image, row, 8 columns), mean, scale (the
        // apodization is in the tables) ----
        float sum = px[0];
#pragma unroll
        for (int s = 1; s < 8; ++s) sum += px[s];
        const float mean = img_sum(sum) / 256.f;
        float d[8], amx = 0.f;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            d[s] = px[s] - mean;
            amx = fmaxf(amx, fabsf(d[s]));
        }
        // max |px - mean| scaled into [2^-2, 2^-1)
        const float bound = img_max(amx);
        int ex = 0;
        (void)frexpf(bound, &ex);
        const int e0 = bound > 0.f ? -1 - ex : 0;
        Split a1;
#pragma unroll
        for (int s = 0; s < 8; s += 2) put2(a1, s, ldexpf(d[s], e0), ldexpf(d[s + 1], e0), neg1);
        const int eA = __shfl(e0, 0), eB = __shfl(e0, 16);

        // ---- 1. column DFT of both windows ----
        const f16v d1 = mfma3(a1, tabs[0][tl], zero);

        // ---- 2. row DFT per image; Z(k1 = n, k2 = rmap(h, v)), Z(k1, 16) ----
        float zr[2][9], zi[2][9];
#pragma unroll
        for (int I = 0; I < 2; ++I) {
            Split a2;
#pragma unroll
            for (int s = 0; s < 8; s += 2) put2(a2, s, d1[8 * I + s], d1[8 * I + s + 1], neg1);
            const f16v dr = mfma3(a2, tabs[1][tl], zero);
            const f16v di = mfma3(a2, tabs[2][tl], zero);
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                // k2 >= 1: Re = Wr Tr - Wi Ti, Im = Wr Ti + Wi Tr
                zr[I][v] = dr[v] - di[v + 8];
                zi[I][v] = dr[v + 8] + di[v];
            }
            // k2 = 0 (slot 0 of the first half) has no Im part; its Im slot
            // carries the real column k2 = 16
            if (h == 0) {
                zr[I][0] = dr[0];
                zi[I][0] = di[0];
            }
            zr[I][8] = dr[8];
            zi[I][8] = di[8];
        }

        // ---- 3. spectrum and inverse along k2 ----
        Split a3[2];
        const int es = eA + eB;  // P scale: P_true 2^es
        float sp_r[9], sp_i[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) {
            float pr = __builtin_fmaf(zr[0][v], zr[1][v], zi[0][v] * zi[1][v]);
            float pi = __builtin_fmaf(zr[0][v], zi[1][v], -(zi[0][v] * zr[1][v]));
            if constexpr (PHASE) {
                const int k2 = v < 8 ? rmap(h, v) : 16;
                pr = ldexpf(pr, -es);
                pi = ldexpf(pi, -es);
                const float den = hypotf(pr, pi) + FLT_EPSILON;
                const float wv = cfg.wfilt[n + k2 * M];
                pr = ldexpf(wv * pr / den, kS3P);
                pi = ldexpf(wv * pi / den, kS3P);
            }
            sp_r[v] = pr;
            sp_i[v] = pi;
        }
        // k2 = 0 and 16 (lanes h = 0, slots 0 and 8): H = Re P - Im P into
        // the K slots of Re and Im k2 = 0
        if (h == 0) {
            const float h0 = sp_r[0] - sp_i[0], h16 = sp_r[8] - sp_i[8];
            sp_r[0] = h0;
            sp_i[0] = h16;
        }
#pragma unroll
        for (int s = 0; s < 8; s += 2) {
            put2(a3[0], s, sp_r[s], sp_r[s + 1], neg1);
            put2(a3[1], s, sp_i[s], sp_i[s + 1], neg1);
        }
        f16v d3 = mfma3(a3[0], tabs[3][tl], zero);
        d3 = mfma3(a3[1], tabs[4][tl], d3);

        // ---- 4. inverse along k1 for the box rows ----
        Split a4[2];
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int s = 0; s < 8; s += 2)
                put2(a4[q], s, d3[8 * q + s], d3[8 * q + s + 1], neg1);
        f16v o4 = mfma3(a4[0], tabs[5][tl], zero);
        o4 = mfma3(a4[1], tabs[6][tl], o4);

        // R(l1 = n - HW, l2 = b - HW) = (Re row b) x cos - (Im row b) x sin:
        // the cos products sit in lanes n < 16, the sin products of the same
        // l1 in lane n + 16 of the same half
        float sim[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) sim[v] = __shfl_xor(o4[v + 8], 16);
        const int etot = (PHASE ? kS3P : es) + kS4;
        const float scale = ldexpf(1.f / (float)(M * M), -etot);
        if (n < kNB) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int b = rmap(h, v);
                if (b < kNB) {
                    const float rv = (o4[v] - sim[v]) * scale;
                    Rb[n + b * kNB] = fabsf(rv) * gain8[v];
                }
            }
        }
        }
        wave_sync();
        const int slot = l / GL;
        const int jb = ww.base + (g + min(slot, ng - 1)) * ww.stride;
        const int4 jj = jobs4[jb];
        float sg = NAN;
        int zo;
        asm volatile("v_mov_b32 %0, 0" : "=v"(zo));
        analyze_box<M, K, kNB, LO, GL, 1, false>(cfg, sout[zo], rbox[wave][slot], l % GL, l, nullptr, nullptr,
                                                  slot < ng, jb, jj.x, jj.y, sg, sg);
        wave_sync();
    }
}

// ============================================================================
// 32-px windows padded to 64: the same four real GEMM stages, on 2 x 2 tiles
// of 32 x 32 per product. One wave per window:
//   1. D1[I][J] = X_I (rows r, K = c in 2 steps) x F1[.][J]: image I's
//      column DFT, parts of tile J = Re k2 = 16J..16J+15 (rows 0-15) and Im of
//      the same k2 (rows 16-31; the Im slot of k2 = 0 carries Re k2 = 32);
//   2. per (J, k1 tile nt): D2 = D1[I][J]^T (K = r) x [cos | -sin](r, k1),
//      Z(k1, k2) for 8 k2 per lane, the cross (or phase) spectrum P, and
//   3. at once its contribution to Q_Re[nt], Q_Im[nt] += P (K = Re/Im of
//      k2) x table, the inverse along k2 for the 27 box columns l2 (one
//      column tile each for the Re and the Im parts of Q) with the Hermitian
//      weights in the table, so the spectrum is never held whole; k2 = 0 and
//      the Nyquist column k2 = 32 enter as H = Re P - Im P in the K slots of
//      k2 = 0, as in the 16-px kernel;
//   4. R = [cos | -sin](l1, k1) x [Q_Re; Q_Im] (K = k1 in 4 steps for each
//      half): the tables as the A operand and the stage-3 results as B, so
//      R accumulates whole in one tile.
// Constant operand tables (28 split 32 x 16 operands, 57 KB) live in LDS,
// one copy per block; boxes of 4 windows are analyzed together, 16 lanes
// (two box columns each) per window.
struct Trig64 {
    float c[64], s[64];  // cos, sin of 2 pi t / 64
    // the separable Gaussian apodization of a 32-px window (exp(-((i - 16.5)^2
    // + (j - 16.5)^2) / (2 * 8^2)) = ap[i - 1] ap[j - 1]), folded into the
    // stage-1 (column factor) and stage-2 (row factor) tables
    float ap[32];
};

constexpr int kHW64 = 13;                 // box half-width (Gaussian apodization)
constexpr int kNB64 = 2 * kHW64 + 1;      // 27 box rows and columns
constexpr int kS4_64 = -8;                // stage-4 operand scale (66 terms of |P| < 2^14)
enum {
    T64_F1 = 0,    // [kk][J]
    T64_WC = 4,    // [kk][nt]
    T64_WS = 8,    // [kk][nt]
    T64_T3 = 12,   // [J][source Re | Im][result Re | Im]
    T64_C4 = 20,   // [nt][kk]
    T64_S4 = 24,   // [nt][kk] (-sin)
    T64_N = 28
};

// Operand table q for lane l (slot s of half h holds K index k(h, s)).
__device__ void build_table64(int q, int l, const Trig64 &tg, Split &out)
{
    const int h = l >> 5, n = l & 31;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        float f = 0.f;
        if (q < T64_WC) {  // F1[kk][J]: K = c, column = part n of tile J
            const int kk = (q - T64_F1) >> 1, J = (q - T64_F1) & 1;
            const int c = 16 * kk + 8 * h + s;
            if (n < 16) {
                f = tg.c[((16 * J + n) * c) & 63];
            } else {
                const int k2 = 16 * J + n - 16;
                f = k2 == 0 ? ((c & 1) ? -1.f : 1.f) : -tg.s[(k2 * c) & 63];
            }
            f *= tg.ap[c];
        } else if (q < T64_T3) {  // Wc / Ws[kk][nt]: K = r, column k1
            const bool sn = q >= T64_WS;
            const int i = q - (sn ? T64_WS : T64_WC);
            const int kk = i >> 1, nt = i & 1;
            const int r = 16 * kk + rmap(h, s);
            const int t = ((32 * nt + n) * r) & 63;
            f = tg.ap[r] * (sn ? -tg.s[t] : tg.c[t]);
        } else if (q < T64_C4) {  // T3[J][src][dst]: K = k2, column l2 = b - HW, b = n
            const int i = q - T64_T3;
            const int J = i >> 2, imk = (i >> 1) & 1, imc = i & 1;
            const int k2 = 16 * J + rmap(h, s);
            const int b = n;
            if (b < kNB64) {
                const int t = (k2 * (b - kHW64)) & 63;
                if (k2 == 0) {
                    // the K slots of k2 = 0 carry H(k1, 0) (Re table) and
                    // H(k1, 32) (Im table), H = Re P - Im P: columns of
                    // H (1 - i) and H (1 - i) (-1)^l2
                    const float sg = imk ? tg.c[(32 * (b - kHW64)) & 63] : 1.f;
                    f = imc ? -sg : sg;
                } else {
                    f = imk ? (imc ? 2.f * tg.c[t] : -2.f * tg.s[t]) : (imc ? 2.f * tg.s[t] : 2.f * tg.c[t]);
                }
            }
        } else {  // C4 / S4[nt][kk]: the A operand of stage 4, row l1 = n - HW, K = k1
            const bool sn = q >= T64_S4;
            const int i = q - (sn ? T64_S4 : T64_C4);
            const int nt = i >> 1, kk = i & 1;
            if (n < kNB64) {
                const int t = ((32 * nt + 16 * kk + rmap(h, s)) * (n - kHW64)) & 63;
                f = sn ? -tg.s[t] : tg.c[t];
            }
        }
        put(out, s, f);
    }
}

// WV waves per block, one block per CU: WV = 8 keeps two waves per SIMD
// within 256 registers (some spill), WV = 4 one wave per SIMD with the whole
// register file (no spill); the autotuner picks.
template <int K, bool PHASE, int WV>
__global__ __launch_bounds__(64 * WV) __attribute__((amdgpu_waves_per_eu(WV / 4, WV / 4))) void k_corr_mfma64(
    CorrConfig cfg, Trig64 tg, const float *__restrict__ A, const float *__restrict__ B, int nr,
    const int32_t *__restrict__ jobs, int njobs, CorrOut out)
{
#pragma clang fp contract(off)
    constexpr int M = 64, H = 32, LO = H - kHW64;
    __shared__ Split tabs[T64_N][64];
    // boxes of 4 windows analyzed together, 16 lanes (two box columns each) per window
    __shared__ float rbox[WV][4][kNB64 * kNB64 + 1];
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const int h = l >> 5, n = l & 31;
    for (int q = wave; q < T64_N; q += WV) {
        Split t;
        build_table64(q, l, tg, t);
        tabs[q][l] = t;
    }
    __shared__ CorrOut sout[1];  // output pointers, read where used (as in k_corr_mfma32)
    if (tid == 0) sout[0] = out;
    __syncthreads();
    const f16v zero = {};
    const float neg1 = opaque_neg1();

    const WaveWindows ww = wave_windows(cfg.linear_order, gridDim.x, blockIdx.x, WV, wave, njobs);
    if (ww.n <= 0) return;
    const int4 *jobs4 = reinterpret_cast<const int4 *>(jobs);
    for (int g = 0; g < ww.n; g += 4) {
        const int ng = min(4, ww.n - g);
        for (int j = 0; j < ng; ++j) {
            const int4 jw = jobs4[ww.base + (g + j) * ww.stride];
            float *Rb = rbox[wave][j];
            int z;
            asm volatile("v_mov_b32 %0, 0" : "=v"(z));
            const int tl = l + z;
            // ---- 0. both windows: lane (h, n) holds row n, columns 16 kk + 8 h + s
            float px[2][16];
#pragma unroll
            for (int I = 0; I < 2; ++I) {
                const float *p = (I ? B : A) + (size_t)(jw.w - 1 + 8 * h) * nr + (jw.z - 1 + n);
#pragma unroll
                for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                    for (int s = 0; s < 8; ++s) px[I][8 * kk + s] = p[(size_t)(16 * kk + s) * nr];
            }
            Split a1[2][2];
            int eI[2];
#pragma unroll
            for (int I = 0; I < 2; ++I) {
                float sum = 0.f;
#pragma unroll
                for (int s = 0; s < 16; ++s) sum += px[I][s];
                sum = img_sum(sum);           // DPP row (16 lanes) and lane + 32
                sum += __shfl_xor(sum, 16);   // and the other row of the half: all 64 lanes
                const float mean = sum / 1024.f;
                float amx = 0.f;
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    px[I][s] -= mean;
                    amx = fmaxf(amx, fabsf(px[I][s]));
                }
                amx = img_max(amx);
                const float bound = fmaxf(amx, __shfl_xor(amx, 16));  // >= |x| (apodization <= 1)
                int ex = 0;
                (void)frexpf(bound, &ex);
                const int e0 = bound > 0.f ? -3 - ex : 0;  // |x| < 2^-3
                eI[I] = e0;
                // the apodization is in the tables
#pragma unroll
                for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                    for (int s = 0; s < 8; s += 2)
                        put2(a1[I][kk], s, ldexpf(px[I][8 * kk + s], e0), ldexpf(px[I][8 * kk + s + 1], e0), neg1);
            }
            const int es = eI[0] + eI[1];
            f16v q3[2][2];  // Q[nt][Re | Im]: rows k1 of tile nt, columns l2
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int ri = 0; ri < 2; ++ri) q3[nt][ri] = zero;
#pragma unroll
            for (int J = 0; J < 2; ++J) {
                // ---- 1. column DFT of both windows, parts of tile J ----
                Split a2[2][2];
#pragma unroll
                for (int I = 0; I < 2; ++I) {
                    f16v d1 = mfma3(a1[I][0], tabs[T64_F1 + 0 * 2 + J][tl], zero);
                    d1 = mfma3(a1[I][1], tabs[T64_F1 + 1 * 2 + J][tl], d1);
#pragma unroll
                    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                        for (int s = 0; s < 8; s += 2) put2(a2[I][kk], s, d1[8 * kk + s], d1[8 * kk + s + 1], neg1);
                }
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) {
                    // ---- 2. row DFT: Z(k1 = 32 nt + n, k2 = 16 J + rmap(h, v)) ----
                    float zr[2][9], zi[2][9];
#pragma unroll
                    for (int I = 0; I < 2; ++I) {
                        // one frame at a time keeps two products live, not four
                        f16v dc = mfma3(a2[I][0], tabs[T64_WC + 0 * 2 + nt][tl], zero);
                        dc = mfma3(a2[I][1], tabs[T64_WC + 1 * 2 + nt][tl], dc);
                        f16v ds = mfma3(a2[I][0], tabs[T64_WS + 0 * 2 + nt][tl], zero);
                        ds = mfma3(a2[I][1], tabs[T64_WS + 1 * 2 + nt][tl], ds);
#pragma unroll
                        for (int v = 0; v < 8; ++v) {
                            zr[I][v] = dc[v] - ds[v + 8];
                            zi[I][v] = dc[v + 8] + ds[v];
                        }
                        if (J == 0 && h == 0) {  // k2 = 0 (no Im part) and k2 = 32 in its slot
                            zr[I][0] = dc[0];
                            zi[I][0] = ds[0];
                        }
                        zr[I][8] = dc[8];
                        zi[I][8] = ds[8];
                    }
                    // ---- 3. spectrum; inverse along k2 into Q[nt][.] ----
                    float sp_r[9], sp_i[9];
#pragma unroll
                    for (int v = 0; v < 9; ++v) {
                        float pr = __builtin_fmaf(zr[0][v], zr[1][v], zi[0][v] * zi[1][v]);
                        float pi = __builtin_fmaf(zr[0][v], zi[1][v], -(zi[0][v] * zr[1][v]));
                        if constexpr (PHASE) {
                            const int k2 = v < 8 ? 16 * J + rmap(h, v) : 32;
                            pr = ldexpf(pr, -es);
                            pi = ldexpf(pi, -es);
                            const float den = hypotf(pr, pi) + FLT_EPSILON;
                            const float wv = cfg.wfilt[(32 * nt + n) + k2 * M];
                            pr = ldexpf(wv * pr / den, kS3P);
                            pi = ldexpf(wv * pi / den, kS3P);
                        }
                        sp_r[v] = pr;
                        sp_i[v] = pi;
                    }
                    // k2 = 0 and the Nyquist column k2 = 32 (tile 0, half 0,
                    // slots 0 and 8), Hermitian in k1: H = Re P - Im P into the
                    // K slots of Re and Im k2 = 0
                    if (J == 0 && h == 0) {
                        const float h0 = sp_r[0] - sp_i[0], h32 = sp_r[8] - sp_i[8];
                        sp_r[0] = h0;
                        sp_i[0] = h32;
                    }
                    Split a3r, a3i;
#pragma unroll
                    for (int s = 0; s < 8; s += 2) {
                        put2(a3r, s, sp_r[s], sp_r[s + 1], neg1);
                        put2(a3i, s, sp_i[s], sp_i[s + 1], neg1);
                    }
#pragma unroll
                    for (int ri = 0; ri < 2; ++ri) {
                        q3[nt][ri] = mfma3(a3r, tabs[T64_T3 + J * 4 + 0 * 2 + ri][tl], q3[nt][ri]);
                        q3[nt][ri] = mfma3(a3i, tabs[T64_T3 + J * 4 + 1 * 2 + ri][tl], q3[nt][ri]);
                    }
                }
            }
            // ---- 4. R = [cos | -sin](l1, k1) x [Q_Re; Q_Im](k1, l2): the
            // tables as A (rows l1), the stage-3 results as B (columns l2,
            // K = k1 in the slot order of their result rows) ----
            const int etot = (PHASE ? kS3P : es) + kS4_64;
            const float scale = ldexpf(1.f / (float)(M * M), -etot);
            f16v r4 = zero;
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int kk = 0; kk < 2; ++kk) {
                    Split br, bi;
#pragma unroll
                    for (int s = 0; s < 8; s += 2) {
                        put2(br, s, ldexpf(q3[nt][0][8 * kk + s], kS4_64), ldexpf(q3[nt][0][8 * kk + s + 1], kS4_64),
                             neg1);
                        put2(bi, s, ldexpf(q3[nt][1][8 * kk + s], kS4_64), ldexpf(q3[nt][1][8 * kk + s + 1], kS4_64),
                             neg1);
                    }
                    r4 = mfma3(tabs[T64_C4 + nt * 2 + kk][tl], br, r4);
                    r4 = mfma3(tabs[T64_S4 + nt * 2 + kk][tl], bi, r4);
                }
            // lane column n = l2 index, rows l1 = rmap(h, v)
            if (n < kNB64) {
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int a = rmap(h, v);
                    if (a < kNB64)
                        Rb[a + n * kNB64] = fabsf(r4[v] * scale) * cfg.gain[(LO + a) + (LO + n) * M];
                }
            }
        }
        wave_sync();
        const int slot = l >> 4;
        const int jb = ww.base + (g + min(slot, ng - 1)) * ww.stride;
        const int4 jj = jobs4[jb];
        float sg = NAN;
        int zo;
        asm volatile("v_mov_b32 %0, 0" : "=v"(zo));
        analyze_box<M, K, kNB64, LO, 16, 1, false>(cfg, sout[zo], rbox[wave][slot], l & 15, l, nullptr, nullptr,
                                                    slot < ng, jb, jj.x, jj.y, sg, sg);
        wave_sync();
    }
}

}  // namespace

// covered: 16- and 32-px windows padded to 32 / 64 with the Gaussian support box
bool mfma_covers(const CorrConfig &cfg)
{
    if (!cfg.padded || !cfg.gain || cfg.finder != HH_REGIONALMAX || cfg.K > 3) return false;
    if (cfg.M == 32 && cfg.w == 16) return cfg.box_lo == 16 - kHW && cfg.box_hi == 16 + kHW;
    if (cfg.M == 64 && cfg.w == 32) return cfg.box_lo == 32 - kHW64 && cfg.box_hi == 32 + kHW64;
    return false;
}

int device_cu_count()
{
    // the first call happens while planning a schedule, outside any stream
    // capture
    static const int n = [] {
        int dev = 0, v = 256;
        if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev);
        return v > 0 ? v : 256;
    }();
    return n;
}

bool launch_correlate_mfma(const CorrConfig &cfg, const float *A, const float *B, int nr, const int32_t *jobs,
                           int njobs, const CorrOut &out, hipStream_t s)
{
    if (!mfma_covers(cfg)) return false;
    if (out.planes || out.acc || out.unc_u) return false;
    if (njobs <= 0) return true;
    const int ncu = device_cu_count();
    if (cfg.M == 64) {
        static const Trig64 tg64 = [] {
            Trig64 t;
            for (int k = 0; k < 64; ++k) {
                t.c[k] = (float)std::cos(2.0 * M_PI * k / 64.0);
                t.s[k] = (float)std::sin(2.0 * M_PI * k / 64.0);
            }
            for (int i = 0; i < 32; ++i) {
                const double d = (i + 1) - 16.5;
                t.ap[i] = (float)std::exp(-d * d / (2.0 * 8.0 * 8.0));
            }
            return t;
        }();
        // one block per CU (tables and boxes take 84-108 KB of LDS):
        // waves_per_simd 1 -> 4-wave blocks, 2 (default) -> 8-wave blocks;
        // persistent waves walk the windows at a stride (wave_windows)
        const int wps = cfg.waves_per_simd == 1 ? 1 : 2;
        const int wv = 4 * wps;
        const int nblocks = std::max(1, std::min((njobs + wv - 1) / wv, ncu));
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(nblocks), dim3(64 * wv), 0, s, cfg, tg64, A, B, nr, jobs, njobs, out);
        };
        const bool ph = cfg.method == HH_PHASE, k2 = cfg.K == 2;
        if (wps == 1) {
            if (ph)
                k2 ? go(k_corr_mfma64<2, true, 4>) : go(k_corr_mfma64<3, true, 4>);
            else
                k2 ? go(k_corr_mfma64<2, false, 4>) : go(k_corr_mfma64<3, false, 4>);
        } else {
            if (ph)
                k2 ? go(k_corr_mfma64<2, true, 8>) : go(k_corr_mfma64<3, true, 8>);
            else
                k2 ? go(k_corr_mfma64<2, false, 8>) : go(k_corr_mfma64<3, false, 8>);
        }
        HH_HIP(hipGetLastError());
        return true;
    }
    static Trig tg = [] {
        Trig t;
        for (int k = 0; k < 32; ++k) {
            t.c[k] = (float)std::cos(2.0 * M_PI * k / 32.0);
            t.s[k] = (float)std::sin(2.0 * M_PI * k / 32.0);
        }
        for (int i = 0; i < 16; ++i) {
            const double d = (i + 1) - 8.5;
            t.ap[i] = (float)std::exp(-d * d / (2.0 * 4.0 * 4.0));
        }
        return t;
    }();
    const int wps = cfg.waves_per_simd > 0 ? cfg.waves_per_simd : 3;
    const int wpb = wps >= 4 ? 8 : 4;  // waves per block of the two builds
    const int waves = std::max(1, std::min((njobs + 3) / 4 * 4, ncu * 4 * wps));
    const int nblocks = (waves + wpb - 1) / wpb;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nblocks), dim3(64 * wpb), 0, s, cfg, tg, A, B, nr, jobs, njobs, out);
    };
    const bool ph = cfg.method == HH_PHASE, k2 = cfg.K == 2;
    if (wps >= 4) {
        if (ph)
            k2 ? go(k_corr_mfma32<2, true, true>) : go(k_corr_mfma32<3, true, true>);
        else
            k2 ? go(k_corr_mfma32<2, false, true>) : go(k_corr_mfma32<3, false, true>);
    } else {
        if (ph)
            k2 ? go(k_corr_mfma32<2, true, false>) : go(k_corr_mfma32<3, true, false>);
        else
            k2 ? go(k_corr_mfma32<2, false, false>) : go(k_corr_mfma32<3, false, false>);
    }
    HH_HIP(hipGetLastError());
    return true;
}

}  // namespace hh
