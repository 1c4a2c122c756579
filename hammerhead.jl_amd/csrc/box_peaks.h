// box_peaks.h — peak search and peak analysis on the analysis box of a
// correlation plane, shared by the register FFT and the MFMA DFT kernels.
//
// The box holds the overlap-gain support plus a zero margin; outside it the
// plane is exactly zero. find_peaks_regionalmax! (src/correlators.jl:
// 426-454, src/localmax.jl:3-19) over the full plane is reproduced by the
// regional maxima of the box plus the plane corner (M, M), the one
// zero-valued point that is always a local maximum; analyze_plane!
// (src/correlators.jl:459-468) follows on lane 0 of the window.
#pragma once

#include <climits>
#include <type_traits>

#include "hh_internal.h"

namespace hh {

__device__ __forceinline__ bool better(float av, int al, float bv, int bl)
{
    if (al == INT_MAX) return false;
    if (bl == INT_MAX) return true;
    return av > bv || (av == bv && al < bl);
}

// The best candidate over the G lanes of a window (groups of G lanes aligned
// to G), in the order of better(): the largest value (positive floats order
// as their bit patterns; an empty slot counts as 0), then, among the lanes
// holding it, the smallest location. Two 32-bit reductions instead of one
// over 64-bit keys.
template <int G>
__device__ __forceinline__ void cand_best(float v, int l, float &bv, int &bl)
{
    unsigned vb = l == INT_MAX ? 0u : __float_as_uint(v);
    auto red = [](unsigned x, auto op) {
        if constexpr (G == 8) {
            // within each half of a DPP row: i <-> 7 - i, then xor 2 and
            // xor 1 inside the quads (a maximum or minimum in every lane)
            x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, true));  // row_half_mirror
            x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, true));   // quad_perm [2,3,0,1]
            x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, true));   // quad_perm [1,0,3,2]
        } else if constexpr (G >= 16) {
            x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, true));  // row_ror:8
            x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x124, 0xF, 0xF, true));  // row_ror:4
            x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x122, 0xF, 0xF, true));  // row_ror:2
            x = op(x, (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x121, 0xF, 0xF, true));  // row_ror:1
            if constexpr (G >= 32) x = op(x, (unsigned)__shfl_xor((int)x, 16));
            if constexpr (G >= 64) x = op(x, (unsigned)__shfl_xor((int)x, 32));
        } else {
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) x = op(x, (unsigned)__shfl_xor((int)x, o));
        }
        return x;
    };
    const unsigned vm = red(vb, [](unsigned a, unsigned b) { return a > b ? a : b; });
    const unsigned lc = (vm != 0u && vb == vm) ? (unsigned)l : 0xFFFFFFFFu;
    const unsigned lm = red(lc, [](unsigned a, unsigned b) { return a < b ? a : b; });
    bv = vm ? __uint_as_float(vm) : 0.f;
    bl = vm ? (int)lm : INT_MAX;
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

template <int M, typename V>
__device__ float2 gauss3_b(const V &R, int i, int j)  // 1-based
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

template <int M, typename V>
__device__ float2 gauss9_b(const V &R, int pr, int pc)
{
    if (!(1 < pr && pr < M && 1 < pc && pc < M)) return gauss3_b<M>(R, pr, pc);
    float sL = 0.f, sxL = 0.f, syL = 0.f, sxxL = 0.f, syyL = 0.f, sxyL = 0.f;
    for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) {
            const float val = R(pr + di - 1, pc + dj - 1);
            if (!(val > 0.f)) return gauss3_b<M>(R, pr, pc);
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
    if (!(a3 < 0.f && D > 0.f)) return gauss3_b<M>(R, pr, pc);
    const float dx = (a4 * a2 - 2.f * a5 * a1) / D;
    const float dy = (a4 * a1 - 2.f * a3 * a2) / D;
    if (!(fabsf(dx) <= 1.f && fabsf(dy) <= 1.f)) return gauss3_b<M>(R, pr, pc);
    return make_float2((float)pr + dy, (float)pc + dx);
}

template <int M, typename V>
__device__ float moment_b(const V &R, float pr, float pc)
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

template <int R>
__device__ __forceinline__ double row_ror_d(double v)  // DPP row_ror:R of a double
{
    const unsigned long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x120 + R, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x120 + R, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double row_sum16_d(double v)  // over the 16 lanes of a DPP row, in every lane
{
    v += row_ror_d<8>(v);
    v += row_ror_d<4>(v);
    v += row_ror_d<2>(v);
    v += row_ror_d<1>(v);
    return v;
}

// moment_b with its (at most 9) points spread over lanes t = 0..8 of the
// window's first DPP row (the window's lanes start on a 16-lane boundary):
// lane t takes column c0 - 1 + t / 3, row r0 - 1 + t % 3, the three Float64
// sums are reduced by row rotations. Every lane of the row calls it (no
// early exit before the rotations); the result is valid on all of them.
template <int M, typename V>
__device__ __forceinline__ float moment_row(const V &R, float pr, float pc, int t)
{
    const bool fin = isfinite(pr) && isfinite(pc);
    const int r0 = fin ? (int)rintf(pr) : 1, c0 = fin ? (int)rintf(pc) : 1;
    const int rlo = max(1, r0 - 1), rhi = min(M, r0 + 1);
    const int clo = max(1, c0 - 1), chi = min(M, c0 + 1);
    const int c = c0 - 1 + t / 3, r = r0 - 1 + t % 3;
    double w = 0.0, tx = 0.0, ty = 0.0;
    if (fin && t < 9 && c >= clo && c <= chi && r >= rlo && r <= rhi) {
        w = (double)R(r - 1, c - 1);
        w = w > 0.0 ? w : 0.0;
        const float dx = (float)c - pc, dy = (float)r - pr;
        tx = (double)(dx * dx) * w;
        ty = (double)(dy * dy) * w;
    }
    const double sumC = row_sum16_d(w), sdx = row_sum16_d(tx), sdy = row_sum16_d(ty);
    if (!fin || rlo > rhi || clo > chi || sumC < 1e-9) return NAN;
    return (float)sqrt((sdx + sdy) / sumC);
}

template <int S>
__device__ __forceinline__ double row_shl_d(double v)  // lane i gets lane i + S of its DPP row (0 past it)
{
    const unsigned long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x100 + S, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x100 + S, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// moment_row for windows of 8 lanes (two per DPP row): lane t takes point t
// and lane 0 also point 8, which it adds first; the pairs (xor 4, 2, 1) then
// give lane 0 the sums in the same order as moment_row's rotations over 16
// lanes (whose lanes 9-15 add exact zeros), so the moment is bit-identical.
// The result is valid on lane 0 of the window.
template <int M, typename V>
__device__ __forceinline__ float moment_row8(const V &R, float pr, float pc, int t)
{
    const bool fin = isfinite(pr) && isfinite(pc);
    const int r0 = fin ? (int)rintf(pr) : 1, c0 = fin ? (int)rintf(pc) : 1;
    const int rlo = max(1, r0 - 1), rhi = min(M, r0 + 1);
    const int clo = max(1, c0 - 1), chi = min(M, c0 + 1);
    double w = 0.0, tx = 0.0, ty = 0.0;
    auto point = [&](int q) {  // point q of the 3 x 3 neighbourhood, added in
        const int c = c0 - 1 + q / 3, r = r0 - 1 + q % 3;
        if (fin && c >= clo && c <= chi && r >= rlo && r <= rhi) {
            double wq = (double)R(r - 1, c - 1);
            wq = wq > 0.0 ? wq : 0.0;
            const float dx = (float)c - pc, dy = (float)r - pr;
            w += wq;
            tx += (double)(dx * dx) * wq;
            ty += (double)(dy * dy) * wq;
        }
    };
    point(t);
    if (t == 0) point(8);
    w += row_shl_d<4>(w);
    tx += row_shl_d<4>(tx);
    ty += row_shl_d<4>(ty);
    w += row_shl_d<2>(w);
    tx += row_shl_d<2>(tx);
    ty += row_shl_d<2>(ty);
    w += row_shl_d<1>(w);
    tx += row_shl_d<1>(tx);
    ty += row_shl_d<1>(ty);
    if (!fin || rlo > rhi || clo > chi || w < 1e-9) return NAN;
    return (float)sqrt((tx + ty) / w);
}

// Regional-max top-K over the box Rb (NB x NB column-major, plane rows and
// columns LO .. LO + NB - 1, fftshifted) by the G lanes t = 0 .. G - 1 of
// one window (NWAVE waves per window, exchanging through mv/ml), then the
// per-window outputs of process_windows! (src/pipeline.jl:972-1027).
template <int M, int K, int NB, int LO, int G, int NWAVE, bool UQ>
__device__ __forceinline__ void analyze_box(const CorrConfig &cfg, const CorrOut &out, const float *Rb, int t, int tid,
                                            float *mv, int *ml, bool active, int jb, int gi, int gj, float sig_u,
                                            float sig_v)
{
#pragma clang fp contract(off)
    constexpr int H = M / 2;
    const int wave = tid >> 6;
    (void)wave;
    // ---- 5. regional-max top-K over the box ----
    float tv[K];
    int tl[K];
#pragma unroll
    for (int m = 0; m < K; ++m) {
        tv[m] = 0.f;
        tl[m] = INT_MAX;
    }
    // A lane meets its candidates in increasing location order, so a new
    // candidate goes before the first entry of smaller value (an equal value
    // arrived earlier, at the smaller location, and stays ahead) and every
    // entry from there on moves down one slot; an empty slot (0, INT_MAX)
    // loses to any candidate (> 0).
    auto insert = [&](float cv, int cl) {
        bool moved = false;
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const bool c = moved || cv > tv[m];
            moved = c;
            const float sv_ = tv[m];
            const int sl_ = tl[m];
            tv[m] = c ? cv : sv_;
            tl[m] = c ? cl : sl_;
            cv = c ? sv_ : cv;
            cl = c ? sl_ : cl;
        }
    };
    if constexpr ((G == 16 && NB <= 16) || (G == 32 && NB < 32) || (G == 64 && NB < 64)) {
        // lane t walks box column t down the rows with its 3 x 3
        // neighbourhood in registers (left and right columns from lanes
        // t -/+ 1, zero past the box), one LDS read per point; neighbours
        // earlier in column order (left column, row above) disqualify when
        // strictly greater, later ones when >=. 16 lanes: DPP row shifts
        // within the window's row; 32 and 64 lanes: wave shifts, whose reads
        // across a window boundary (or past lane 63) land on lanes past the
        // box, which hold zero
        const bool col = t < NB;
        auto left = [](float v) {  // lane t gets lane t - 1 (0 at t = 0)
            constexpr int ctrl = G == 16 ? 0x111 : 0x138;  // row_shr:1 / wave_shr:1
            return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, true));
        };
        auto right = [](float v) {  // lane t gets lane t + 1
            constexpr int ctrl = G == 16 ? 0x101 : 0x130;  // row_shl:1 / wave_shl:1
            return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, true));
        };
        // V: the lane's own column over rows bi-1..bi+1; a neighbour lane's V
        // is that lane's three rows, so left(V) / right(V) are the left and
        // right columns of the 3 x 3 neighbourhood. Points that pass are
        // marked in a bit mask and inserted afterwards, lowest row first
        // (increasing location), so the insertion runs once per regional
        // maximum instead of once per point.
        static_assert(NB <= 64, "one mask bit per box row");
        using Mask = std::conditional_t<(NB > 32), unsigned long long, unsigned>;
        Mask okm = 0;
        float cp = 0.f;
        float cc = col ? Rb[t * NB] : 0.f;
        for (int bi = 0; bi < NB; ++bi) {
            const float cn = (col && bi + 1 < NB) ? Rb[bi + 1 + t * NB] : 0.f;
            const float V = fmaxf(fmaxf(cp, cc), cn);
            const float I0 = cc;
            // fmaxf skips NaN neighbours exactly as !(NaN > I0) and
            // !(NaN >= I0) pass them
            const float me = fmaxf(left(V), cp);    // earlier: disqualify when >
            const float ml_ = fmaxf(right(V), cn);  // later: disqualify when >=
            const bool ok = I0 > 0.f && !(me > I0) && !(ml_ >= I0);
            okm |= (Mask)(ok ? 1u : 0u) << bi;
            cp = cc;
            cc = cn;
        }
        while (okm) {
            const int bi = NB > 32 ? __builtin_ctzll((unsigned long long)okm) : __builtin_ctz((unsigned)okm);
            okm &= okm - 1u;
            insert(Rb[bi + t * NB], (LO + bi) + (LO + t) * M);
        }
    } else if constexpr ((G == 8 && NB <= 15) || (G == 16 && NB <= 31)) {
        // lane t walks box columns 2t and 2t + 1 down the rows: the left
        // column of 2t is lane t - 1's second column, the right column of
        // 2t + 1 lane t + 1's first (DPP row shifts). With 16 lanes a window
        // is one DPP row (shifts past its ends read 0). With 8, two windows
        // share a row; a shift across the boundary reads column 15 of the
        // window on the left or column 0 of the one on the right only for
        // lanes whose own column there is past the box (NB <= 15): column
        // 15 is zero, and lane 7's first column (14) takes its right column
        // from its own second column. Marked maxima are inserted column by
        // column, lowest row first (increasing location); a lane's top K of
        // its two columns hold every candidate of the window's top K that
        // lies there, so the merge gives the same result as one column per
        // lane.
        auto left = [](float v) {
            return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x111, 0xF, 0xF, true));
        };
        auto right = [](float v) {
            return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x101, 0xF, 0xF, true));
        };
        const int ca = 2 * t, cb = 2 * t + 1;
        const bool va = ca < NB, vb = cb < NB;
        unsigned oka = 0, okb = 0;
        float pa = 0.f, pb = 0.f;
        float a0 = va ? Rb[ca * NB] : 0.f, b0 = vb ? Rb[cb * NB] : 0.f;
        for (int bi = 0; bi < NB; ++bi) {
            const float an = (va && bi + 1 < NB) ? Rb[bi + 1 + ca * NB] : 0.f;
            const float bn = (vb && bi + 1 < NB) ? Rb[bi + 1 + cb * NB] : 0.f;
            const float Va = fmaxf(fmaxf(pa, a0), an), Vb = fmaxf(fmaxf(pb, b0), bn);
            // column 2t: left = lane t - 1's Vb, right = own Vb
            const float mea = fmaxf(left(Vb), pa), mla = fmaxf(Vb, an);
            // column 2t + 1: left = own Va, right = lane t + 1's Va
            const float meb = fmaxf(Va, pb), mlb = fmaxf(right(Va), bn);
            const bool oa = a0 > 0.f && !(mea > a0) && !(mla >= a0);
            const bool ob = b0 > 0.f && !(meb > b0) && !(mlb >= b0);
            oka |= (oa ? 1u : 0u) << bi;
            okb |= (ob ? 1u : 0u) << bi;
            pa = a0;
            a0 = an;
            pb = b0;
            b0 = bn;
        }
        while (oka) {
            const int bi = __builtin_ctz(oka);
            oka &= oka - 1u;
            insert(Rb[bi + ca * NB], (LO + bi) + (LO + ca) * M);
        }
        while (okb) {
            const int bi = __builtin_ctz(okb);
            okb &= okb - 1u;
            insert(Rb[bi + cb * NB], (LO + bi) + (LO + cb) * M);
        }
    } else {
        for (int p = t; p < NB * NB; p += G) {
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
            insert(I0, (LO + bi) + (LO + bj) * M);
        }
    }
    float sv[K];
    int sl[K];
    int nsel = 0;
#pragma unroll
    for (int m = 0; m < K; ++m) {
        float bv;
        int bl;
        cand_best<G>(tv[0], tl[0], bv, bl);
        if constexpr (NWAVE > 1) {
            if ((tid & 63) == 0) {
                mv[wave] = bv;
                ml[wave] = bl;
            }
            __syncthreads();
#pragma unroll
            for (int w2 = 0; w2 < NWAVE; ++w2)
                if (better(mv[w2], ml[w2], bv, bl)) {
                    bv = mv[w2];
                    bl = ml[w2];
                }
            __syncthreads();
        }
        sv[m] = bv;
        sl[m] = bl;
        nsel += bl != INT_MAX;
        if (tl[0] == bl && bl != INT_MAX) {
#pragma unroll
            for (int q = 0; q + 1 < K; ++q) {
                tv[q] = tv[q + 1];
                tl[q] = tl[q + 1];
            }
            tv[K - 1] = 0.f;
            tl[K - 1] = INT_MAX;
        }
    }
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

    const BoxView<NB, LO> R{Rb};
    // subpixel fits of the K candidates side by side: lane t of the window
    // fits candidate t (lane 0 with the pass's method, the others gauss3 as
    // the alternatives use), and lane 0 gathers them
    float2 fit = make_float2(NAN, NAN);
    if constexpr (G >= K) {
        int slt = sl[0];
#pragma unroll
        for (int m = 1; m < K; ++m)
            if (t == m) slt = sl[m];
        if (t < K && t < found) {
            const int fr = slt % M + 1, fc = slt / M + 1;
            if (t == 0 && cfg.subpixel == HH_GAUSS9)
                fit = gauss9_b<M>(R, fr, fc);
            else if (t == 0 && cfg.subpixel == HH_SUBPIX_NONE)
                fit = make_float2((float)fr, (float)fc);
            else
                fit = gauss3_b<M>(R, fr, fc);
        }
    }
    float2 fits[K];
#pragma unroll
    for (int m = 0; m < K; ++m) {
        if constexpr (G >= K) {
            const int src = (tid & 63) - t + m;  // lane of candidate m in this wave
            fits[m] = make_float2(__shfl(fit.x, src), __shfl(fit.y, src));
        } else {
            fits[m] = make_float2(NAN, NAN);
        }
    }
    float mom_row = NAN;
    if constexpr (G >= 16)
        mom_row = moment_row<M>(R, fits[0].x, fits[0].y, t);
    else if constexpr (G == 8)
        mom_row = moment_row8<M>(R, fits[0].x, fits[0].y, t);
    if (t != 0 || !active) return;
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
    float2 ref;
    if constexpr (G >= K) {
        ref = fits[0];
    } else {
        const int pr = sl[0] % M + 1, pc = sl[0] / M + 1;
        if (cfg.subpixel == HH_GAUSS9)
            ref = gauss9_b<M>(R, pr, pc);
        else if (cfg.subpixel == HH_SUBPIX_NONE)
            ref = make_float2((float)pr, (float)pc);
        else
            ref = gauss3_b<M>(R, pr, pc);
    }
    const float cc = (float)(H + 1);
    const float du = ref.y - cc, dv = ref.x - cc;
    const float v1 = sv[0];
    const float ratio = found >= 2 ? v1 / sv[1] : (v1 > 0.f ? INFINITY : NAN);
    float mom;
    if constexpr (G >= 8)
        mom = mom_row;
    else
        mom = moment_b<M>(R, ref.x, ref.y);
    const float u0 = out.u0 ? out.u0[g] : 0.f, v0 = out.v0 ? out.v0[g] : 0.f;
    out.u[g] = out.u0 ? u0 + du : du;
    out.v[g] = out.v0 ? v0 + dv : dv;
    out.pr[g] = ratio;
    out.cm[g] = mom;
    if constexpr (UQ) {
        out.unc_u[g] = sig_u;
        out.unc_v[g] = sig_v;
    }
    if (out.alt_u) {
#pragma unroll
        for (int m = 2; m <= K; ++m) {
            if (m > cfg.n_peaks) continue;
            float au = NAN, av = NAN;
            if (m <= found) {
                float2 a;
                if constexpr (G >= K)
                    a = fits[m - 1];
                else
                    a = gauss3_b<M>(R, sl[m - 1] % M + 1, sl[m - 1] / M + 1);
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

}  // namespace hh
