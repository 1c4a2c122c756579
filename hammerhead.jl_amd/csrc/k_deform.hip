// k_deform.hip — cubic B-spline prefilter and symmetric image deformation.
//
// Prefilter. Interpolations' BSpline(Cubic(Line(OnGrid()))) coefficients
// (image_interpolant, src/pipeline.jl:909-915) solve, per axis, the (n+2)
// system (c[k-1] + 4c[k] + c[k+1])/6 = y[k], k = 1..n, with the natural
// boundary rows c[0] - 2c[1] + c[2] = 0 = c[n-1] - 2c[n] + c[n+1]. The
// solution equals the infinite-line inverse filter h[d] = sqrt(3) z^|d|,
// z = sqrt(3) - 2 (the root of z^2 + 4z + 1 = 0), applied to the data
// extended point-antisymmetrically about both end samples
// (y[1-d] = 2y[1] - y[1+d], y[n+d] = 2y[n] - y[n-d]): that extension makes
// the filtered sequence satisfy the boundary rows exactly. |z|^17 < 2e-10,
// so 33 taps reproduce the exact solve to far below float32 rounding. The
// filter is separable and local, hence one coalesced pass per axis instead
// of a serial tridiagonal sweep per line.
//
// Deformation (deform_images / _ka_deform!, src/pipeline.jl:936-965,
// src/ka_backend.jl:293-312): per output pixel (r, c), the predictor
// displacement d is the bilinear (Gridded(Linear()), Flat()) value of the
// coarse vector grid, and warpA(r, c) = A(r - dv/2, c - du/2),
// warpB(r, c) = B(r + dv/2, c + du/2), sampled with the cubic B-spline
// weights of _pk_cubic (src/ka_backend.jl:265-291), zero outside [1, n].
// The sample's base index and fraction come from the pixel index and the
// half displacement separately (cubic_at), so the position is as exact as
// the displacement itself rather than rounded to float32 at the pixel's
// magnitude.

#include "hh_internal.h"
#include "interp.h"

namespace hh {

namespace {

constexpr int kTaps = 16;  // half-width of the prefilter

struct Taps {
    float h[kTaps + 1];
};

__host__ Taps make_taps()
{
    Taps t;
    const double z = std::sqrt(3.0) - 2.0;
    double p = 1.0;
    for (int d = 0; d <= kTaps; ++d) {
        t.h[d] = (float)(std::sqrt(3.0) * p);
        p *= z;
    }
    return t;
}

// Value of the antisymmetrically extended line at logical index j (1-based
// data 1..n). get(k) reads data sample k in [1, n].
template <typename Get>
__device__ __forceinline__ float ext_value(const Get &get, int j, int n)
{
    // reflect until inside; sign/offset bookkeeping: value = off + sgn * y[j]
    float off = 0.f, sgn = 1.f;
    for (int it = 0; it < 8 && (j < 1 || j > n); ++it) {
        if (j < 1) {
            off += sgn * 2.f * get(1);
            sgn = -sgn;
            j = 2 - j;
        } else {
            off += sgn * 2.f * get(n);
            sgn = -sgn;
            j = 2 * n - j;
        }
    }
    if (j < 1) j = 1;
    if (j > n) j = n;
    return off + sgn * get(j);
}

// Pass 1: along dim 1. out is (nr+2) x nc, logical row k at out[k + j*(nr+2)].
// Block (x, j): logical rows [256 x, 256 x + 256) of column j; the extended
// segment with its 16-sample halos is staged once in LDS.
__global__ __launch_bounds__(256) void k_prefilter_rows(const float *__restrict__ imgA,
                                                         const float *__restrict__ imgB, int ld, int nr, int nc,
                                                         float *__restrict__ outA, float *__restrict__ outB,
                                                         Taps taps)
{
    __shared__ float seg[256 + 2 * kTaps];
    const float *__restrict__ img = blockIdx.z ? imgB : imgA;
    float *__restrict__ out = blockIdx.z ? outB : outA;
    const int N1 = nr + 2;
    const int j = blockIdx.y, k0 = blockIdx.x * 256;
    const float *col = img + (size_t)j * ld;
    auto get = [&](int q) { return col[q - 1]; };
    for (int t = threadIdx.x; t < 256 + 2 * kTaps; t += 256) seg[t] = ext_value(get, k0 - kTaps + t, nr);
    __syncthreads();
    const int k = k0 + threadIdx.x;
    if (k >= N1) return;
    const float *x = seg + threadIdx.x + kTaps;
    float s = taps.h[0] * x[0];
#pragma unroll
    for (int d = 1; d <= kTaps; ++d) s += taps.h[d] * (x[-d] + x[d]);
    out[k + (size_t)j * N1] = s;
}

// Pass 2: along dim 2 over the (nr+2) rows of tmp; coef is (nr+2) x (nc+2).
// Block (x, y): rows [64 x, 64 x + 64), logical columns [64 y, 64 y + 64),
// staged with the column halos in LDS (coalesced along the rows).
__global__ __launch_bounds__(256) void k_prefilter_cols(const float *__restrict__ tmpA,
                                                         const float *__restrict__ tmpB, int nr, int nc,
                                                         float *__restrict__ coefA, float *__restrict__ coefB,
                                                         Taps taps)
{
    __shared__ float tile[64 + 2 * kTaps][65];
    const float *__restrict__ tmp = blockIdx.z ? tmpB : tmpA;
    float *__restrict__ coef = blockIdx.z ? coefB : coefA;
    const int N1 = nr + 2, N2 = nc + 2;
    const int i0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    const int ti = threadIdx.x & 63, tg = threadIdx.x >> 6;
    const int i = i0 + ti;
    const bool row_ok = i < N1;
    auto get = [&](int q) { return tmp[(size_t)(row_ok ? i : 0) + (size_t)(q - 1) * N1]; };
    for (int t = tg; t < 64 + 2 * kTaps; t += 4) tile[t][ti] = row_ok ? ext_value(get, k0 - kTaps + t, nc) : 0.f;
    __syncthreads();
    if (!row_ok) return;
    for (int t = tg; t < 64; t += 4) {
        const int k = k0 + t;
        if (k >= N2) break;
        float s = taps.h[0] * tile[t + kTaps][ti];
#pragma unroll
        for (int d = 1; d <= kTaps; ++d) s += taps.h[d] * (tile[t + kTaps - d][ti] + tile[t + kTaps + d][ti]);
        coef[i + (size_t)k * N1] = s;
    }
}

// Predictor node value of enumerated cell ee of nv.g.
__device__ __forceinline__ void node_values_cell(const NodeValuesJob &nv, int ee)
{
    if (ee >= nv.g.total) return;
    const Cell k = cell_of(nv.g, ee);
    const int e = k.b, i = k.r, j = k.c;
    const DeformPair &P = nv.pairs[k.sg.pair];
    const float y = (float)(nv.r0 + i * nv.rstep) + nv.half, x = (float)(nv.c0 + j * nv.cstep) + nv.half;
    nv.u[e] = bilinear(P.pu, P.gny, P.gnx, P.y0, P.ysp, P.x0, P.xsp, y, x);
    nv.v[e] = bilinear(P.pv, P.gny, P.gnx, P.y0, P.ysp, P.x0, P.xsp, y, x);
}

// One warped pixel of each image per thread. Blocks [blk0, blk0 + n) of a
// batched launch belong to one pair (found by bisection of the pair table);
// within a pair, block b covers rows 256 (b mod S) + 1 .. of column b / S + 1
// (S = deform_strips(nr) blocks per column), so the pixel's column and row
// come from the block index (uniform) and the thread index, with no per-lane
// division.
__global__ __launch_bounds__(256) void k_deform(DeformTab tab, int ld, float *__restrict__ wA,
                                                float *__restrict__ wB)
{
    if ((int)blockIdx.x >= tab.nblocks) {  // the node-values blocks (uniform branch)
        node_values_cell(tab.nv, ((int)blockIdx.x - tab.nblocks) * 256 + (int)threadIdx.x);
        return;
    }
    int q = 0;
    if (tab.pairs) {
        int lo = 0, hi = tab.npairs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tab.pairs[mid].blk0 <= (int)blockIdx.x)
                lo = mid;
            else
                hi = mid - 1;
        }
        q = lo;
    }
    // a copy, selected with uniform control flow: the fields stay in scalar
    // registers (a reference to either source would force per-lane loads)
    DeformPair P;
    if (tab.pairs)
        P = tab.pairs[q];
    else
        P = tab.one;
    const int nr = P.nr, nc = P.nc;
    const int S = deform_strips(nr), b = (int)blockIdx.x - P.blk0;
    const int c = b / S + 1, r = (b % S) * 256 + (int)threadIdx.x + 1;
    if (r > nr || c > nc) return;
    const float R = (float)r, Cc = (float)c;
    int iy, ix;
    float ly, lx;
    lin_axis(P.y0, P.ysp, P.gny, R, iy, ly);
    lin_axis(P.x0, P.xsp, P.gnx, Cc, ix, lx);
    const float *__restrict__ cA = P.cA;
    const float *__restrict__ cB = P.cB;
    const float du2 = bilinear_at(P.pu, P.gny, P.gnx, iy, ix, ly, lx) / 2.f;
    const float dv2 = bilinear_at(P.pv, P.gny, P.gnx, iy, ix, ly, lx) / 2.f;
    const size_t o = (size_t)P.out_off + (size_t)(r - 1) + (size_t)(c - 1) * ld;
    // both gathers before either store: all 32 coefficient loads in flight
    const float va = cubic_at(cA, nr, nc, r, -dv2, c, -du2);
    const float vb = cubic_at(cB, nr, nc, r, dv2, c, du2);
    wA[o] = va;
    wB[o] = vb;
}

__global__ __launch_bounds__(256) void k_node_values(const float *__restrict__ pu, const float *__restrict__ pv,
                                                     int gny, int gnx, float y0, float ysp, float x0, float xsp,
                                                     const float *__restrict__ y, int ny,
                                                     const float *__restrict__ x, int nx, float *__restrict__ u,
                                                     float *__restrict__ v)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ny * nx) return;
    const int i = e % ny, j = e / ny;
    u[e] = bilinear(pu, gny, gnx, y0, ysp, x0, xsp, y[i], x[j]);
    v[e] = bilinear(pv, gny, gnx, y0, ysp, x0, xsp, y[i], x[j]);
}

// predictor_node_values for the cells of several grids: segment s holds
// the window centres of pair seg.pair.
__global__ __launch_bounds__(256) void k_node_values_batch(NodeValuesJob nv)
{
    node_values_cell(nv, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

}  // namespace

void launch_prefilter_pair(const float *imgA, const float *imgB, int ld, int nr, int nc, float *coefA, float *coefB,
                           float *tmp, hipStream_t s)
{
    static const Taps taps = make_taps();
    float *tA = tmp, *tB = tmp + (size_t)(nr + 2) * nc;
    hipLaunchKernelGGL(k_prefilter_rows, dim3((unsigned)((nr + 2 + 255) / 256), (unsigned)nc, 2), dim3(256), 0, s,
                       imgA, imgB, ld, nr, nc, tA, tB, taps);
    hipLaunchKernelGGL(k_prefilter_cols, dim3((unsigned)((nr + 2 + 63) / 64), (unsigned)((nc + 2 + 63) / 64), 2),
                       dim3(256), 0, s, tA, tB, nr, nc, coefA, coefB, taps);
    HH_HIP(hipGetLastError());
}

void launch_deform(const float *cA, const float *cB, int nr, int nc, const float *pu, const float *pv, int gny,
                   int gnx, float y0, float ysp, float x0, float xsp, float *wA, float *wB, hipStream_t s)
{
    DeformTab t{};
    t.pairs = nullptr;
    t.npairs = 1;
    t.nblocks = deform_strips(nr) * nc;
    t.one = DeformPair{cA, cB, pu, pv, nr, nc, gny, gnx, y0, ysp, x0, xsp, 0, 0, 0};
    launch_deform_batch(t, nr, wA, wB, s);
}

void launch_deform_batch(const DeformTab &tab, int ld, float *wA, float *wB, hipStream_t s)
{
    const int nb = tab.nblocks + (tab.nv.g.total > 0 ? tab.nv.nblocks : 0);
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_deform, dim3((unsigned)nb), dim3(256), 0, s, tab, ld, wA, wB);
    HH_HIP(hipGetLastError());
}

NodeValuesJob node_values_job(const Grids &g, const DeformPair *pairs, int r0, int rstep, int c0, int cstep,
                              float half, float *u, float *v)
{
    return NodeValuesJob{g, pairs, r0, rstep, c0, cstep, half, u, v, (g.total + 255) / 256};
}

void launch_node_values_batch(const Grids &g, const DeformPair *pairs, int r0, int rstep, int c0, int cstep,
                              float half, float *u, float *v, hipStream_t s)
{
    if (g.total <= 0) return;
    const NodeValuesJob nv = node_values_job(g, pairs, r0, rstep, c0, cstep, half, u, v);
    hipLaunchKernelGGL(k_node_values_batch, dim3((unsigned)nv.nblocks), dim3(256), 0, s, nv);
    HH_HIP(hipGetLastError());
}

void launch_node_values(const float *pu, const float *pv, int gny, int gnx, float y0, float ysp, float x0, float xsp,
                        const float *y, int ny, const float *x, int nx, float *u, float *v, hipStream_t s)
{
    const int n = ny * nx;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_node_values, dim3((n + 255) / 256), dim3(256), 0, s, pu, pv, gny, gnx, y0, ysp, x0, xsp, y,
                       ny, x, nx, u, v);
    HH_HIP(hipGetLastError());
}

}  // namespace hh
