// k_dewarp.hip — dewarp! (src/dewarp.jl:233-248): camera frames resampled
// onto a world-plane grid at the dewarper's precomputed source coordinates
// (ImageDewarper, src/dewarp.jl:196-221), with the frame's cubic B-spline
// interpolant (BSpline(Cubic(Line(OnGrid()))), zero outside [1, n]); nodes
// out of view are 0.
//
// The map holds one int2 per output pixel: the 1-based source row and
// column in 16.16 fixed point (INT32_MIN marks a node out of view). The
// integer part is the sample's base index and the low 16 bits its fraction,
// so every node resolves its source position to 2^-16 px whatever the
// image size (a float32 coordinate resolves ulp(2048) = 2.4e-4 px), and the
// map costs 8 bytes per pixel like two float32 coordinates. One thread per
// output pixel: the map entry is read once and both frames of a pair are
// sampled from it (16-tap gathers of the prefiltered coefficients, which
// overlapping neighbours share through L2); consecutive threads take
// consecutive pixels of an output column.

#include "hh_internal.h"
#include "interp.h"

namespace hh {

namespace {

__global__ __launch_bounds__(256) void k_dewarp(const int2 *__restrict__ map, int n, const float *__restrict__ cA,
                                                const float *__restrict__ cB, int nr, int nc,
                                                float *__restrict__ oA, float *__restrict__ oB)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int2 m = map[e];
    float va = 0.f, vb = 0.f;
    if (m.x != INT32_MIN) {
        const int ry = m.x >> 16, rx = m.y >> 16;
        const float fy = (float)(m.x & 0xFFFF) * (1.f / 65536.f);
        const float fx = (float)(m.y & 0xFFFF) * (1.f / 65536.f);
        va = cubic_at(cA, nr, nc, ry, fy, rx, fx);
        if (oB) vb = cubic_at(cB, nr, nc, ry, fy, rx, fx);
    }
    oA[e] = va;
    if (oB) oB[e] = vb;
}

}  // namespace

void launch_dewarp(const int2 *map, int ny, int nx, const float *cA, const float *cB, int nr, int nc, float *oA,
                   float *oB, hipStream_t s)
{
    const int n = ny * nx;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_dewarp, dim3((n + 255) / 256), dim3(256), 0, s, map, n, cA, cB, nr, nc, oA, oB);
    HH_HIP(hipGetLastError());
}

}  // namespace hh
