// k_adjoint.hip — windowed correlation energy of a node field and its
// hand-written adjoint.
//
// The reference has no gradient path (its nearest relative is the
// non-gradient correlate_deformable, src/correlators.jl:620-638). The
// objective goes through the symmetric deformation of the WIDIM chain
// (apply_predictor / deform_images, src/pipeline.jl:575-600,936-965): pixel
// displacements are the bilinear Flat() values of the node field (pu, pv),
// A^(r,c) = A(r - v/2, c - u/2) and B^(r,c) = B(r + v/2, c + u/2) are
// sampled from the cubic B-spline interpolants or bilinearly from the images
// (zero outside [1, n] either way), and over the windows of a pass
// (pass_grid, apodization p)
//     E = - sum_w sum_{x in w} p_x (A^_x - mA_w) (B^_x - mB_w).
// With N = w^2, PA_w = sum p (A^ - mA), PB_w = sum p (B^ - mB):
//     dE/dA^_x = sum_{w contains x} (PB_w / N - p_x (B^_x - mB_w))
//     dE/dB^_x = sum_{w contains x} (PA_w / N - p_x (A^_x - mA_w))
//     dE/du_x = (dE/dB^ dB/dx - dE/dA^ dA/dx) / 2 (and v with d/dy)
//     dE/dpu_n = sum_x w_n(x) dE/du_x (bilinear node weights).
// Nothing per window is stored: the window pass recomputes its warped
// samples to produce (mA, mB, PA/N, PB/N, E_w), and the pixel pass
// recomputes every pixel's samples and slopes, walks the windows that
// contain it, and reduces the pixel gradients of one node cell per block
// into that cell's four corner nodes. A node gathers the corners of its
// adjacent cells, so every sum has a fixed order.

#include "hh_internal.h"
#include "interp.h"

namespace hh {

namespace {

__device__ __forceinline__ float wave_sum(float x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ void node_uv(const EnergyArgs &g, float r, float c, float &u, float &v)
{
    u = bilinear(g.pu, g.gny, g.gnx, g.y0, g.ysp, g.x0, g.xsp, r, c);
    v = bilinear(g.pv, g.gny, g.gnx, g.y0, g.ysp, g.x0, g.xsp, r, c);
}

// The warped samples A^(r, c) (s = -1) or B^(r, c) (s = +1) and their
// slopes, by the configured sampling (g.bilinear is uniform per launch).
__device__ __forceinline__ float warp_sample(const EnergyArgs &g, bool frame_b, int r, int c, float u, float v,
                                             float &sy, float &sx)
{
    const float h = frame_b ? 0.5f : -0.5f;
    if (g.bilinear) return bilinear_d_at(frame_b ? g.B : g.A, g.nr, g.nc, r, h * v, c, h * u, sy, sx);
    return cubic_d_at(frame_b ? g.cB : g.cA, g.nr, g.nc, r, h * v, c, h * u, sy, sx);
}

// One wave per window: recompute the warped window, then its statistics.
template <int W>
__global__ __launch_bounds__(64) void k_energy_windows(EnergyArgs g, float4 *__restrict__ stats,
                                                       float *__restrict__ ew)
{
    constexpr int N = W * W;
    constexpr int PPL = N / 64;
    const int lane = threadIdx.x;
    const int win = blockIdx.x;
    const int ki = win % g.ny, kj = win / g.ny;
    const int rs = g.r0 + ki * g.st_r, cs = g.c0 + kj * g.st_c;
    float a[PPL], b[PPL];
    float sa = 0.f, sb = 0.f;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        const int e = lane + 64 * q;
        const float r = (float)(rs + e % W), c = (float)(cs + e / W);
        float u, v, sy, sx;
        node_uv(g, r, c, u, v);
        a[q] = warp_sample(g, false, rs + e % W, cs + e / W, u, v, sy, sx);
        b[q] = warp_sample(g, true, rs + e % W, cs + e / W, u, v, sy, sx);
        sa += a[q];
        sb += b[q];
    }
    const float mA = wave_sum(sa) / (float)N, mB = wave_sum(sb) / (float)N;
    float pa = 0.f, pb = 0.f, en = 0.f;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        const float p = g.apod[lane + 64 * q];
        const float da = a[q] - mA, db = b[q] - mB;
        pa += p * da;
        pb += p * db;
        en -= p * da * db;
    }
    pa = wave_sum(pa);
    pb = wave_sum(pb);
    en = wave_sum(en);
    if (lane == 0) {
        stats[win] = make_float4(mA, mB, pa / (float)N, pb / (float)N);
        ew[win] = en;
    }
}

// One block per node cell (cy, cx): the pixels whose bilinear interval is
// that cell, each reduced into the cell's four corner nodes.
__global__ __launch_bounds__(256) void k_energy_adjoint(EnergyArgs g, const float4 *__restrict__ stats,
                                                        const int2 *__restrict__ rows, const int2 *__restrict__ cols,
                                                        int ncy, float *__restrict__ part)
{
    __shared__ float red[4][8];
    const int tid = threadIdx.x;
    const int cell = blockIdx.x;
    const int cy = cell % ncy, cx = cell / ncy;
    const int2 rr = rows[cy], cr = cols[cx];
    const int nrow = rr.y - rr.x + 1, ncol = cr.y - cr.x + 1;
    float acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.f;
    for (int idx = tid; idx < nrow * ncol; idx += 256) {
        const int r = rr.x + idx % nrow, c = cr.x + idx / nrow;
        int iy, ix;
        float ly, lx;
        lin_axis(g.y0, g.ysp, g.gny, (float)r, iy, ly);
        lin_axis(g.x0, g.xsp, g.gnx, (float)c, ix, lx);
        if (iy != cy || ix != cx) continue;
        float u, v;
        node_uv(g, (float)r, (float)c, u, v);
        float ay, ax, by, bx;
        const float ah = warp_sample(g, false, r, c, u, v, ay, ax);
        const float bh = warp_sample(g, true, r, c, u, v, by, bx);
        // windows containing (r, c): starts r0 + k st within [r - w + 1, r]
        const int kilo = max(0, (r - g.w + 1 - g.r0 + g.st_r - 1) / g.st_r);
        const int kihi = min(g.ny - 1, (r - g.r0) / g.st_r);
        const int kjlo = max(0, (c - g.w + 1 - g.c0 + g.st_c - 1) / g.st_c);
        const int kjhi = min(g.nx - 1, (c - g.c0) / g.st_c);
        float GA = 0.f, GB = 0.f;
        for (int kj = kjlo; kj <= kjhi; ++kj) {
            const int oc = c - (g.c0 + kj * g.st_c);
            for (int ki = kilo; ki <= kihi; ++ki) {
                const int orow = r - (g.r0 + ki * g.st_r);
                const float4 s = stats[ki + kj * g.ny];
                const float p = g.apod[orow + oc * g.w];
                GA += s.w - p * (bh - s.y);
                GB += s.z - p * (ah - s.x);
            }
        }
        const float gx = 0.5f * (GB * bx - GA * ax);
        const float gy = 0.5f * (GB * by - GA * ay);
        const float w00 = (1.f - ly) * (1.f - lx), w01 = (1.f - ly) * lx, w10 = ly * (1.f - lx), w11 = ly * lx;
        acc[0] += w00 * gx;
        acc[1] += w01 * gx;
        acc[2] += w10 * gx;
        acc[3] += w11 * gx;
        acc[4] += w00 * gy;
        acc[5] += w01 * gy;
        acc[6] += w10 * gy;
        acc[7] += w11 * gy;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = wave_sum(acc[k]);
    if ((tid & 63) == 0)
#pragma unroll
        for (int k = 0; k < 8; ++k) red[tid >> 6][k] = acc[k];
    __syncthreads();
    if (tid < 8) part[(size_t)cell * 8 + tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// Node (k, l) gathers the corners of the cells whose bilinear weights reach
// it: corner (dy, dx) of cell (cy, cx) is node (cy + dy, cx + dx) (dy, dx
// dropped along a one-node axis).
__global__ __launch_bounds__(256) void k_energy_gather(const float *__restrict__ part, int gny, int gnx, int ncy,
                                                       int ncx, float *__restrict__ gu, float *__restrict__ gv)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gny * gnx) return;
    const int k = n % gny, l = n / gny;
    float su = 0.f, sv = 0.f;
    for (int cx = max(0, l - 1); cx <= min(ncx - 1, l); ++cx)
        for (int cy = max(0, k - 1); cy <= min(ncy - 1, k); ++cy)
            for (int dx = 0; dx < 2; ++dx)
                for (int dy = 0; dy < 2; ++dy) {
                    const int ty = cy + (gny > 1 ? dy : 0), tx = cx + (gnx > 1 ? dx : 0);
                    if (ty != k || tx != l) continue;
                    const float *p = part + (size_t)(cy + cx * ncy) * 8;
                    su += p[dy * 2 + dx];
                    sv += p[4 + dy * 2 + dx];
                }
    gu[n] = su;
    gv[n] = sv;
}

// E = sum of the per-window energies, in a fixed order.
__global__ __launch_bounds__(256) void k_energy_sum(const float *__restrict__ ew, int n, double *__restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += (double)ew[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

void launch_energy(const EnergyArgs &g, float4 *stats, float *ew, double *esum, hipStream_t s)
{
    const int nw = g.ny * g.nx;
    if (nw <= 0) return;
    switch (g.w) {
    case 8: hipLaunchKernelGGL(k_energy_windows<8>, dim3(nw), dim3(64), 0, s, g, stats, ew); break;
    case 16: hipLaunchKernelGGL(k_energy_windows<16>, dim3(nw), dim3(64), 0, s, g, stats, ew); break;
    case 32: hipLaunchKernelGGL(k_energy_windows<32>, dim3(nw), dim3(64), 0, s, g, stats, ew); break;
    default: throw Error(HH_EUNSUPPORTED, "the correlation-energy adjoint supports 8, 16 and 32 px windows");
    }
    hipLaunchKernelGGL(k_energy_sum, dim3(1), dim3(256), 0, s, ew, nw, esum);
    HH_HIP(hipGetLastError());
}

void launch_energy_adjoint(const EnergyArgs &g, const float4 *stats, const int2 *rows, const int2 *cols, int ncy,
                           int ncx, float *part, float *gu, float *gv, hipStream_t s)
{
    const int ncell = ncy * ncx;
    if (ncell <= 0) return;
    hipLaunchKernelGGL(k_energy_adjoint, dim3(ncell), dim3(256), 0, s, g, stats, rows, cols, ncy, part);
    const int nn = g.gny * g.gnx;
    hipLaunchKernelGGL(k_energy_gather, dim3((nn + 255) / 256), dim3(256), 0, s, part, g.gny, g.gnx, ncy, ncx, gu,
                       gv);
    HH_HIP(hipGetLastError());
}

}  // namespace hh
