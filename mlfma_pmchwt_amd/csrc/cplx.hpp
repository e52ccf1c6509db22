// Complex arithmetic on double2 / float2 (x = re, y = im) and wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>

namespace mlfma {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// acc += a * b
__device__ __forceinline__ void cfma(double2& acc, double2 a, double2 b) {
    acc.x += a.x * b.x - a.y * b.y;
    acc.y += a.x * b.y + a.y * b.x;
}
// acc -= a * b
__device__ __forceinline__ void cfms(double2& acc, double2 a, double2 b) {
    acc.x -= a.x * b.x - a.y * b.y;
    acc.y -= a.x * b.y + a.y * b.x;
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// FP32 complex (float2), same operations; accumulation stays in FP32.
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ void cfma(float2& acc, float2 a, float2 b) {
    acc.x += a.x * b.x - a.y * b.y;
    acc.y += a.x * b.y + a.y * b.x;
}
__device__ __forceinline__ void cfms(float2& acc, float2 a, float2 b) {
    acc.x -= a.x * b.x - a.y * b.y;
    acc.y -= a.x * b.y + a.y * b.x;
}

// acc += a * b as four chained fused multiply-adds (one rounding per partial
// product instead of the mul / fma / add triple the expression form compiles
// to): the translation kernels' inner loop.
__device__ __forceinline__ void cmac(float2& acc, float2 a, float2 b) {
    acc.x = __builtin_fmaf(a.x, b.x, acc.x);
    acc.y = __builtin_fmaf(a.x, b.y, acc.y);
    acc.x = __builtin_fmaf(-a.y, b.y, acc.x);
    acc.y = __builtin_fmaf(a.y, b.x, acc.y);
}
__device__ __forceinline__ void cmac(double2& acc, double2 a, double2 b) {
    acc.x = __builtin_fma(a.x, b.x, acc.x);
    acc.y = __builtin_fma(a.x, b.y, acc.y);
    acc.x = __builtin_fma(-a.y, b.y, acc.x);
    acc.y = __builtin_fma(a.y, b.x, acc.y);
}

// Blocks are dispatched round-robin over the 8 XCDs of the MI355X (block b on
// XCD b % 8).  Chunked order for a grid of units x ng groups: chunk c (units
// c cs .. c cs + cs - 1 with all their groups) runs on XCD c % 8, its blocks
// consecutive there, so the units of a chunk (e.g. the row ranges of one
// leaf, or the sample groups of one leaf) and the groups of a unit share
// their inputs through that XCD's L2, while consecutive chunks still spread
// over all XCDs.  The grid is xcd_chunk_grid blocks; a block whose unit is
// past the unit count has no work.
__device__ __forceinline__ void xcd_chunk(int b, int cs, int ng, int& unit, int& group) {
    const int x = b & 7, s = b >> 3;
    const int per = cs * ng;
    const int k = s / per, j = s - k * per;
    const int u = j / ng;
    unit = (k * 8 + x) * cs + u;
    group = j - u * ng;
}
inline unsigned xcd_chunk_grid(int units, int cs, int ng) {
    return (unsigned)((units + 8 * cs - 1) / (8 * cs)) * 8u * (unsigned)cs * (unsigned)ng;
}

// Conversions between the FP64 solver vectors and the far-field precision.
__device__ __forceinline__ void narrow(double2 a, double2& b) { b = a; }
__device__ __forceinline__ void narrow(double2 a, float2& b) { b = make_float2((float)a.x, (float)a.y); }
__device__ __forceinline__ double2 widen(double2 a) { return a; }
__device__ __forceinline__ double2 widen(float2 a) { return make_double2(a.x, a.y); }

// Butterfly sum over the 64 lanes of a wavefront; every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double2 wave_sum(double2 v) {
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float2 wave_sum(float2 v) {
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    return v;
}

}  // namespace mlfma
