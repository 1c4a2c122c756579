// interp.h — device interpolation helpers shared by the deformation and the
// adjoint kernels: the bilinear Flat() predictor of the node grid and the
// cubic B-spline sample of prefiltered coefficients.
#pragma once

#include <hip/hip_runtime.h>

namespace hh {

// _pk_lin_axis (src/ka_backend.jl:257-263), 0-based interval index.
__device__ __forceinline__ void lin_axis(float q0, float qstep, int n, float q, int &i, float &l)
{
    if (n == 1) {
        i = 0;
        l = 0.f;
        return;
    }
    float t = (q - q0) / qstep;
    // written so that a NaN position clamps to the first knot instead of
    // reaching the float-to-int conversion
    t = t >= 0.f ? t : 0.f;
    t = t > (float)(n - 1) ? (float)(n - 1) : t;
    int ii = (int)floorf(t);
    if (ii > n - 2) ii = n - 2;
    i = ii;
    l = t - (float)ii;
}

// Bilinear value of grid p at interval (iy, ix) with fractions (ly, lx).
__device__ __forceinline__ float bilinear_at(const float *__restrict__ p, int gny, int gnx, int iy, int ix, float ly,
                                             float lx)
{
    const int iy2 = gny == 1 ? iy : iy + 1, ix2 = gnx == 1 ? ix : ix + 1;
    const float a = p[iy + ix * gny] + lx * (p[iy + ix2 * gny] - p[iy + ix * gny]);
    const float b = p[iy2 + ix * gny] + lx * (p[iy2 + ix2 * gny] - p[iy2 + ix * gny]);
    return a + ly * (b - a);
}

__device__ __forceinline__ float bilinear(const float *__restrict__ p, int gny, int gnx, float y0, float ysp,
                                          float x0, float xsp, float y, float x)
{
    int iy, ix;
    float ly, lx;
    lin_axis(y0, ysp, gny, y, iy, ly);
    lin_axis(x0, xsp, gnx, x, ix, lx);
    return bilinear_at(p, gny, gnx, iy, ix, ly, lx);
}

__device__ __forceinline__ void cubw(float d, float w[4])
{
    const float d2 = d * d, d3 = d2 * d, e = 1.f - d;
    w[0] = e * e * e / 6.f;
    w[1] = (4.f - 6.f * d2 + 3.f * d3) / 6.f;
    w[2] = (1.f + 3.f * d + 3.f * d2 - 3.f * d3) / 6.f;
    w[3] = d3 / 6.f;
}

// The cubic B-spline weights times 6 (the separable sum of a sample is
// scaled by 1/36 once, a multiply, instead of dividing each of the 8 weights
// by 6; within an ulp of the divided form).
__device__ __forceinline__ void cubw6(float d, float w[4])
{
    const float d2 = d * d, d3 = d2 * d, e = 1.f - d;
    w[0] = e * e * e;
    w[1] = 4.f - 6.f * d2 + 3.f * d3;
    w[2] = 1.f + 3.f * d + 3.f * d2 - 3.f * d3;
    w[3] = d3;
}

__device__ __forceinline__ float cubic(const float *__restrict__ C, int nr, int nc, float yy, float xx)
{
    // negated in-domain test: a NaN coordinate is outside (sample 0), never
    // an index
    if (!(yy >= 1.f && yy <= (float)nr && xx >= 1.f && xx <= (float)nc)) return 0.f;
    int iy = (int)floorf(yy);
    if (iy > nr - 1) iy = nr - 1;
    int ix = (int)floorf(xx);
    if (ix > nc - 1) ix = nc - 1;
    float wy[4], wx[4];
    cubw6(yy - (float)iy, wy);
    cubw6(xx - (float)ix, wx);
    const int N1 = nr + 2;
    const float *col0 = C + (size_t)(ix - 1) * N1 + (iy - 1);  // one 64-bit product, then column steps
    float s = 0.f;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const float *col = col0 + n * N1;
        s += wx[n] * (wy[0] * col[0] + wy[1] * col[1] + wy[2] * col[2] + wy[3] * col[3]);
    }
    return s * (1.f / 36.f);
}

// One axis of a sample position given as an integer pixel k plus a small
// offset o (the half displacement): the base index and fraction of k + o
// without forming k + o in float32, whose rounding (ulp(2048) = 2.4e-4 px)
// would move the sample. Returns false outside [1, n] (or for a non-finite
// offset); otherwise i in [1, n-1] and d in [0, 1] as the spline needs.
__device__ __forceinline__ bool split_axis(int k, float o, int n, int &i, float &d)
{
    if (!(fabsf(o) < 1.0e6f)) return false;
    const float f = floorf(o);
    int ii = k + (int)f;
    float dd = o - f;  // exact for |o| >= 1; may round up to 1 for tiny negative o
    if (dd >= 1.f) {
        ii += 1;
        dd = 0.f;
    }
    if (ii < 1 || ii > n || (ii == n && dd > 0.f)) return false;
    if (ii > n - 1) {  // the position n itself: base n - 1, fraction 1
        ii = n - 1;
        dd += 1.f;
    }
    i = ii;
    d = dd;
    return true;
}

// cubic() at (r + oy, c + ox) for integer r, c and small offsets, with the
// fractions taken from the offsets exactly (split_axis).
__device__ __forceinline__ float cubic_at(const float *__restrict__ C, int nr, int nc, int r, float oy, int c,
                                          float ox)
{
    int iy, ix;
    float fy, fx;
    if (!split_axis(r, oy, nr, iy, fy) || !split_axis(c, ox, nc, ix, fx)) return 0.f;
    float wy[4], wx[4];
    cubw6(fy, wy);
    cubw6(fx, wx);
    const int N1 = nr + 2;
    const float *col0 = C + (size_t)(ix - 1) * N1 + (iy - 1);  // one 64-bit product, then column steps
    float s = 0.f;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const float *col = col0 + n * N1;
        s += wx[n] * (wy[0] * col[0] + wy[1] * col[1] + wy[2] * col[2] + wy[3] * col[3]);
    }
    return s * (1.f / 36.f);
}

// Bilinear sample of an nr x nc image at (r + oy, c + ox) and its partial
// derivatives (sy along dim 1, sx along dim 2), positions split as in
// cubic_at; zero with zero slope outside [1, nr] x [1, nc]. The slope is that
// of the interval the sample's base index selects (base clamped to n - 1).
__device__ __forceinline__ float bilinear_d_at(const float *__restrict__ I, int nr, int nc, int r, float oy, int c,
                                               float ox, float &sy, float &sx)
{
    int iy, ix;
    float fy, fx;
    if (!split_axis(r, oy, nr, iy, fy) || !split_axis(c, ox, nc, ix, fx)) {
        sy = sx = 0.f;
        return 0.f;
    }
    const float *q = I + (size_t)(iy - 1) + (size_t)(ix - 1) * nr;
    const float a = q[0], b = q[nr], cc = q[1], d = q[nr + 1];  // (iy,ix) (iy,ix+1) (iy+1,ix) (iy+1,ix+1)
    sy = (1.f - fx) * (cc - a) + fx * (d - b);
    sx = (1.f - fy) * (b - a) + fy * (d - cc);
    return (1.f - fy) * ((1.f - fx) * a + fx * b) + fy * ((1.f - fx) * cc + fx * d);
}

// Derivative of the cubic B-spline weights of cubw with respect to d.
__device__ __forceinline__ void cubw_d(float d, float g[4])
{
    const float e = 1.f - d;
    g[0] = -e * e / 2.f;
    g[1] = -2.f * d + 1.5f * d * d;
    g[2] = 0.5f + d - 1.5f * d * d;
    g[3] = d * d / 2.f;
}

// cubic_d at (r + oy, c + ox) for integer r, c (positions split as in
// cubic_at).
__device__ __forceinline__ float cubic_d_at(const float *__restrict__ C, int nr, int nc, int r, float oy, int c,
                                            float ox, float &sy, float &sx);

// cubic() and its partial derivatives (sy along dim 1, sx along dim 2);
// zero with zero slope outside [1, nr] x [1, nc].
__device__ __forceinline__ float cubic_d(const float *__restrict__ C, int nr, int nc, float yy, float xx, float &sy,
                                         float &sx)
{
    if (!(yy >= 1.f && yy <= (float)nr && xx >= 1.f && xx <= (float)nc)) {
        sy = sx = 0.f;
        return 0.f;
    }
    int iy = (int)floorf(yy);
    if (iy > nr - 1) iy = nr - 1;
    int ix = (int)floorf(xx);
    if (ix > nc - 1) ix = nc - 1;
    float wy[4], wx[4], gy[4], gx[4];
    cubw(yy - (float)iy, wy);
    cubw(xx - (float)ix, wx);
    cubw_d(yy - (float)iy, gy);
    cubw_d(xx - (float)ix, gx);
    const int N1 = nr + 2;
    float s = 0.f, s_y = 0.f, s_x = 0.f;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const float *col = C + (size_t)(ix - 1 + n) * N1 + (iy - 1);
        const float c0 = col[0], c1 = col[1], c2 = col[2], c3 = col[3];
        const float cv = wy[0] * c0 + wy[1] * c1 + wy[2] * c2 + wy[3] * c3;
        const float cd = gy[0] * c0 + gy[1] * c1 + gy[2] * c2 + gy[3] * c3;
        s += wx[n] * cv;
        s_y += wx[n] * cd;
        s_x += gx[n] * cv;
    }
    sy = s_y;
    sx = s_x;
    return s;
}

__device__ __forceinline__ float cubic_d_at(const float *__restrict__ C, int nr, int nc, int r, float oy, int c,
                                            float ox, float &sy, float &sx)
{
    int iy, ix;
    float dy, dx;
    if (!split_axis(r, oy, nr, iy, dy) || !split_axis(c, ox, nc, ix, dx)) {
        sy = sx = 0.f;
        return 0.f;
    }
    float wy[4], wx[4], gy[4], gx[4];
    cubw(dy, wy);
    cubw(dx, wx);
    cubw_d(dy, gy);
    cubw_d(dx, gx);
    const int N1 = nr + 2;
    float s = 0.f, s_y = 0.f, s_x = 0.f;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const float *col = C + (size_t)(ix - 1 + n) * N1 + (iy - 1);
        const float c0 = col[0], c1 = col[1], c2 = col[2], c3 = col[3];
        const float cv = wy[0] * c0 + wy[1] * c1 + wy[2] * c2 + wy[3] * c3;
        const float cd = gy[0] * c0 + gy[1] * c1 + gy[2] * c2 + gy[3] * c3;
        s += wx[n] * cv;
        s_y += wx[n] * cd;
        s_x += gx[n] * cv;
    }
    sy = s_y;
    sx = s_x;
    return s;
}

}  // namespace hh
