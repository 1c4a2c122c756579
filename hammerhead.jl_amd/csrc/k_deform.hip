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
__global__ __launch_bounds__(256) void k_prefilter_rows(const float *__restrict__ img, int nr, int nc,
                                                         float *__restrict__ out, Taps taps)
{
    const int N1 = nr + 2;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N1 * nc) return;
    const int k = (int)(e % N1);
    const int j = (int)(e / N1);
    const float *col = img + (size_t)j * nr;
    auto get = [&](int q) { return col[q - 1]; };
    float s = taps.h[0] * ext_value(get, k, nr);
#pragma unroll
    for (int d = 1; d <= kTaps; ++d) s += taps.h[d] * (ext_value(get, k - d, nr) + ext_value(get, k + d, nr));
    out[e] = s;
}

// Pass 2: along dim 2 over the (nr+2) rows of tmp; coef is (nr+2) x (nc+2).
__global__ __launch_bounds__(256) void k_prefilter_cols(const float *__restrict__ tmp, int nr, int nc,
                                                         float *__restrict__ coef, Taps taps)
{
    const int N1 = nr + 2, N2 = nc + 2;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N1 * N2) return;
    const int i = (int)(e % N1);
    const int k = (int)(e / N1);
    auto get = [&](int q) { return tmp[i + (size_t)(q - 1) * N1]; };
    float s = taps.h[0] * ext_value(get, k, nc);
#pragma unroll
    for (int d = 1; d <= kTaps; ++d) s += taps.h[d] * (ext_value(get, k - d, nc) + ext_value(get, k + d, nc));
    coef[e] = s;
}

__global__ __launch_bounds__(256) void k_deform(const float *__restrict__ cA, const float *__restrict__ cB, int nr,
                                                int nc, const float *__restrict__ pu, const float *__restrict__ pv,
                                                int gny, int gnx, float y0, float ysp, float x0, float xsp,
                                                float *__restrict__ wA, float *__restrict__ wB)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nr * nc) return;
    const int r = (int)(e % nr) + 1, c = (int)(e / nr) + 1;
    const float R = (float)r, Cc = (float)c;
    const float du2 = bilinear(pu, gny, gnx, y0, ysp, x0, xsp, R, Cc) / 2.f;
    const float dv2 = bilinear(pv, gny, gnx, y0, ysp, x0, xsp, R, Cc) / 2.f;
    wA[e] = cubic(cA, nr, nc, R - dv2, Cc - du2);
    wB[e] = cubic(cB, nr, nc, R + dv2, Cc + du2);
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

}  // namespace

void launch_prefilter(const float *img, int nr, int nc, float *coef, float *tmp, const float *, int,
                      hipStream_t s)
{
    static const Taps taps = make_taps();
    const size_t n1 = (size_t)(nr + 2) * nc, n2 = (size_t)(nr + 2) * (nc + 2);
    hipLaunchKernelGGL(k_prefilter_rows, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, s, img, nr, nc, tmp, taps);
    hipLaunchKernelGGL(k_prefilter_cols, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, tmp, nr, nc, coef, taps);
    HH_HIP(hipGetLastError());
}

void launch_deform(const float *cA, const float *cB, int nr, int nc, const float *pu, const float *pv, int gny,
                   int gnx, float y0, float ysp, float x0, float xsp, float *wA, float *wB, hipStream_t s)
{
    const size_t n = (size_t)nr * nc;
    hipLaunchKernelGGL(k_deform, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cA, cB, nr, nc, pu, pv, gny,
                       gnx, y0, ysp, x0, xsp, wA, wB);
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
