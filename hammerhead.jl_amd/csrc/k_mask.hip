// k_mask.hip — the windows of a pass that hold masked pixels.
//
// load_windows! (src/correlators.jl:214-262) and load_uncertainty_windows!
// (src/uncertainty.jl:51-79) give a window (and a search area) with masked
// pixels the mean of its valid pixels and enter its masked pixels as zero.
// Such windows are few (the rim of a masked region) and are copied, centred,
// into a tile atlas: one s x s column-major tile per window and frame, tile t
// at column t s of a matrix with leading dimension s. The frame-A tile holds
// the w x w window at offset (s - w) / 2, the frame-B tile the s x s search
// area; valid pixels hold x - mean_valid (each region its own mean), masked
// pixels and the rest 0. The correlation and uncertainty kernels then run
// over the tiles as ordinary windows: the mean they subtract from a centred
// region is zero to rounding, so what they transform are the masked windows.
// With s = w (equal areas, or the uncertainty's window-sized second frame)
// both tiles are windows.

#include "hh_internal.h"

namespace hh {

namespace {

// One block of 256 threads per window; the sums run in FP64 (wave
// reductions then one per wave), the tile writes are column-coalesced.
__global__ __launch_bounds__(256) void k_mask_tiles(const float *__restrict__ A, const float *__restrict__ B, int ld,
                                                    const uint8_t *__restrict__ mask, int mld,
                                                    const int4 *__restrict__ mjobs, int w, int s,
                                                    float *__restrict__ tA, float *__restrict__ tB)
{
    const int t = blockIdx.x;
    const int4 j = mjobs[t];
    const int mg = (s - w) / 2, npx = s * s;
    // tile pixel (i, c) <- search-area pixel (i, c) of B; the window covers
    // tile rows / columns mg .. mg + w - 1
    auto inwin = [&](int i, int c) { return i >= mg && i < mg + w && c >= mg && c < mg + w; };
    double sa = 0.0, sb = 0.0;
    int na = 0, nb = 0;
    for (int p = threadIdx.x; p < npx; p += 256) {
        const int i = p % s, c = p / s;
        if (mask[(size_t)(j.z - mg + i) + (size_t)(j.w - mg + c) * mld]) continue;
        const size_t g = (size_t)(j.x - mg + i) + (size_t)(j.y - mg + c) * ld;
        sb += (double)B[g];
        ++nb;
        if (inwin(i, c)) {
            sa += (double)A[g];
            ++na;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        sa += __shfl_xor(sa, o);
        sb += __shfl_xor(sb, o);
        na += __shfl_xor(na, o);
        nb += __shfl_xor(nb, o);
    }
    __shared__ double ps[2][4];
    __shared__ int pn[2][4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        ps[0][wave] = sa;
        ps[1][wave] = sb;
        pn[0][wave] = na;
        pn[1][wave] = nb;
    }
    __syncthreads();
    double SA = 0.0, SB = 0.0;
    int NA = 0, NB = 0;
    for (int q = 0; q < 4; ++q) {
        SA += ps[0][q];
        SB += ps[1][q];
        NA += pn[0][q];
        NB += pn[1][q];
    }
    const float ma = NA > 0 ? (float)(SA / NA) : 0.f, mb = NB > 0 ? (float)(SB / NB) : 0.f;
    float *__restrict__ oa = tA + (size_t)t * npx;
    float *__restrict__ ob = tB + (size_t)t * npx;
    for (int p = threadIdx.x; p < npx; p += 256) {
        const int i = p % s, c = p / s;
        const bool m = mask[(size_t)(j.z - mg + i) + (size_t)(j.w - mg + c) * mld];
        const size_t g = (size_t)(j.x - mg + i) + (size_t)(j.y - mg + c) * ld;
        oa[p] = (m || !inwin(i, c)) ? 0.f : A[g] - ma;
        ob[p] = m ? 0.f : B[g] - mb;
    }
}

}  // namespace

void launch_mask_tiles(const float *A, const float *B, int ld, const uint8_t *mask, int mld, const int32_t *mjobs,
                       int nw, int w, int sz, float *tA, float *tB, hipStream_t s)
{
    if (nw <= 0) return;
    hipLaunchKernelGGL(k_mask_tiles, dim3((unsigned)nw), dim3(256), 0, s, A, B, ld, mask, mld,
                       reinterpret_cast<const int4 *>(mjobs), w, sz, tA, tB);
    HH_HIP(hipGetLastError());
}

}  // namespace hh
