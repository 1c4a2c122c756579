// mfma_split.h — split-FP16 operands and wave helpers shared by the
// matrix-core DFT correlation kernels (v_mfma_f32_32x32x16_f16 with each
// operand split into FP16 hi + lo parts, FP32 accumulation).
#pragma once

#include <hip/hip_runtime.h>

namespace hh {
namespace mfma {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

struct Split {
    h8 hi, lo;
};

// Row of the 32 x 32 MFMA result held in register slot v of lane half h,
// and the K index of operand slot s in lane half h (same pattern).
__device__ __forceinline__ int rmap(int h, int s) { return 8 * (s >> 2) + 4 * h + (s & 3); }

__device__ __forceinline__ void put(Split &a, int s, float x)
{
    const _Float16 hi = (_Float16)x;
    a.hi[s] = hi;
    a.lo[s] = (_Float16)(x - (float)hi);
}

typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

// Slots s, s + 1 (s even) from two values: hi by one packed conversion;
// each lo = f16(x + hi * -1) by one mixed-precision FMA that reads hi as
// FP16 from its half of the packed register and writes the rounded
// residual straight into its half of lo (v_fma_mixlo/mixhi_f16): the
// residual x - f32(hi) is exact in FP32 inside the FMA, so lo is the same
// value an FP32 FMA followed by a conversion gives, in one instruction per
// value instead of one and a half.
__device__ __forceinline__ void put2(Split &a, int s, float x0, float x1, float neg1)
{
    const f2v x = {x0, x1};
    const h2v hi = __builtin_convertvector(x, h2v);
    const unsigned hp = __builtin_bit_cast(unsigned, hi);
    unsigned lp;
    asm("v_fma_mixlo_f16 %0, %1, %2, %3 op_sel_hi:[1,0,0]\n\t"
        "v_fma_mixhi_f16 %0, %1, %2, %4 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "=&v"(lp)
        : "v"(hp), "v"(neg1), "v"(x0), "v"(x1));
    const h2v lo = __builtin_bit_cast(h2v, lp);
    a.hi[s] = hi[0];
    a.hi[s + 1] = hi[1];
    a.lo[s] = lo[0];
    a.lo[s + 1] = lo[1];
}

// -1 from a register move the compiler cannot fold (read by ordinary VALU
// instructions only, which wait for it in hardware).
__device__ __forceinline__ float opaque_neg1()
{
    float v;
    asm volatile("v_mov_b32 %0, -1.0" : "=v"(v));
    return v;
}

__device__ __forceinline__ f16v mfma3(const Split &a, const Split &b, f16v acc)
{
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.lo, b.hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.hi, b.lo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.hi, b.hi, acc, 0, 0, 0);
    return acc;
}

// The 32 lanes of one image are DPP rows 0 and 2 (image A) or 1 and 3
// (image B): row rotations reduce within a row, one exchange across halves.
template <int R>
__device__ __forceinline__ float rowror_f(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x120 + R, 0xF, 0xF, true));
}

// Sum and maximum over the 32 lanes of one image: row rotations reduce
// within a DPP row, one exchange across halves.
__device__ __forceinline__ float img_sum(float v)
{
    v += rowror_f<8>(v);
    v += rowror_f<4>(v);
    v += rowror_f<2>(v);
    v += rowror_f<1>(v);
    return v + __shfl_xor(v, 32);
}
// The maximum of non-negative, non-NaN values (the callers pass max |x|):
// their bit patterns order as the values, so an unsigned maximum with the
// DPP operand fused (no NaN canonicalization of fmaxf) gives the same result.
__device__ __forceinline__ float img_max(float v)
{
    unsigned u = __float_as_uint(v);
    u = max(u, (unsigned)__builtin_amdgcn_mov_dpp((int)u, 0x128, 0xF, 0xF, true));  // row_ror:8
    u = max(u, (unsigned)__builtin_amdgcn_mov_dpp((int)u, 0x124, 0xF, 0xF, true));  // row_ror:4
    u = max(u, (unsigned)__builtin_amdgcn_mov_dpp((int)u, 0x122, 0xF, 0xF, true));  // row_ror:2
    u = max(u, (unsigned)__builtin_amdgcn_mov_dpp((int)u, 0x121, 0xF, 0xF, true));  // row_ror:1
    return __uint_as_float(max(u, (unsigned)__shfl_xor((int)u, 32)));
}

// LDS hand-off between the lanes of one wave (the waves of a block run
// different numbers of windows, so no block barrier inside the loop).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The windows a persistent wave takes: base + k * stride, k = 0 .. n - 1
// (wv waves per block). XCD bands: block b runs on XCD b % 8; XCD x takes a
// contiguous band of the windows in proportion to its blocks, and its waves
// take every (waves on the XCD)-th window of the band, wave i the windows
// i, i + stride, ...: at any moment the waves of one XCD work on adjacent
// windows, so their 75 % overlap is fetched into that XCD's L2 once (with a
// contiguous range per wave the XCD's waves read as many separate regions,
// and the overlap of neighbouring ranges is evicted before it is reused).
// Launch order: the same striding over all windows and all waves.
struct WaveWindows {
    int base, stride, n;
};
__device__ __forceinline__ WaveWindows wave_windows(int linear, int nb, int b, int wv, int wave, int njobs)
{
    WaveWindows r;
    int end = njobs;
    if (linear) {
        r.base = b * wv + wave;
        r.stride = nb * wv;
    } else {
        const int xcd = b & 7, q8 = nb >> 3, r8 = nb & 7;
        const int before = xcd * q8 + min(xcd, r8);  // blocks on the lower XCDs
        const int mine = q8 + (xcd < r8 ? 1 : 0);    // blocks on this XCD
        const int xs = (int)((long long)njobs * before / nb);
        end = (int)((long long)njobs * (before + mine) / nb);
        r.base = xs + (b >> 3) * wv + wave;
        r.stride = mine * wv;
    }
    r.n = r.base < end ? (end - r.base + r.stride - 1) / r.stride : 0;
    return r;
}

}  // namespace mfma
}  // namespace hh
