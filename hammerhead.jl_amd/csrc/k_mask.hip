// k_mask.hip — the windows of a pass that hold masked pixels.
//
// load_windows! (src/correlators.jl:214-227) and load_uncertainty_windows!
// (src/uncertainty.jl:51-79) give a window with masked pixels the mean of
// its valid pixels and enter its masked pixels as zero. Such windows are few
// (the rim of a masked region) and are copied, centred, into a tile atlas:
// one w x w column-major tile per window, tile t at column t w of a matrix
// with leading dimension w, valid pixels holding x - mean_valid and masked
// pixels 0. The correlation and uncertainty kernels then run over the tiles
// as ordinary windows: the mean they subtract from a centred tile is zero to
// rounding, so the windows they transform are the masked windows.

#include "hh_internal.h"

namespace hh {

namespace {

// One block of 256 threads per window; the sums run in FP64 (wave
// reductions then one per wave), the tile writes are column-coalesced.
__global__ __launch_bounds__(256) void k_mask_tiles(const float *__restrict__ A, const float *__restrict__ B, int ld,
                                                    const uint8_t *__restrict__ mask, int mld,
                                                    const int4 *__restrict__ mjobs, int w, float *__restrict__ tA,
                                                    float *__restrict__ tB)
{
    const int t = blockIdx.x;
    const int4 j = mjobs[t];
    const int npx = w * w;
    double sa = 0.0, sb = 0.0;
    int n = 0;
    for (int p = threadIdx.x; p < npx; p += 256) {
        const int i = p % w, c = p / w;
        if (mask[(size_t)(j.z + i) + (size_t)(j.w + c) * mld]) continue;
        const size_t g = (size_t)(j.x + i) + (size_t)(j.y + c) * ld;
        sa += (double)A[g];
        sb += (double)B[g];
        ++n;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sa += __shfl_xor(sa, o);
        sb += __shfl_xor(sb, o);
        n += __shfl_xor(n, o);
    }
    __shared__ double ps[2][4];
    __shared__ int pn[4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        ps[0][wave] = sa;
        ps[1][wave] = sb;
        pn[wave] = n;
    }
    __syncthreads();
    double SA = 0.0, SB = 0.0;
    int N = 0;
    for (int q = 0; q < 4; ++q) {
        SA += ps[0][q];
        SB += ps[1][q];
        N += pn[q];
    }
    const float ma = N > 0 ? (float)(SA / N) : 0.f, mb = N > 0 ? (float)(SB / N) : 0.f;
    float *__restrict__ oa = tA + (size_t)t * npx;
    float *__restrict__ ob = tB + (size_t)t * npx;
    for (int p = threadIdx.x; p < npx; p += 256) {
        const int i = p % w, c = p / w;
        const bool m = mask[(size_t)(j.z + i) + (size_t)(j.w + c) * mld];
        const size_t g = (size_t)(j.x + i) + (size_t)(j.y + c) * ld;
        oa[p] = m ? 0.f : A[g] - ma;
        ob[p] = m ? 0.f : B[g] - mb;
    }
}

}  // namespace

void launch_mask_tiles(const float *A, const float *B, int ld, const uint8_t *mask, int mld, const int32_t *mjobs,
                       int nw, int w, float *tA, float *tB, hipStream_t s)
{
    if (nw <= 0) return;
    hipLaunchKernelGGL(k_mask_tiles, dim3((unsigned)nw), dim3(256), 0, s, A, B, ld, mask, mld,
                       reinterpret_cast<const int4 *>(mjobs), w, tA, tB);
    HH_HIP(hipGetLastError());
}

}  // namespace hh
