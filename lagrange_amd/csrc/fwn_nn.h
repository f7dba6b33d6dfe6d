// Shared device helpers of the learned surrogate: bf16 conversion, the gfx950 BF16 MFMA tile,
// the multi-resolution hash-grid lookup and the counter-based RNG.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fwn {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int kHidden = 128;   // MLP width
constexpr int kLevelsMax = 16; // hash-grid levels
constexpr int kFeat = 2;       // features per level

// fp32 -> bf16 (round to nearest even), as raw bits; NaN stays NaN
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// v_mfma_f32_32x32x16_bf16: acc(32x32) += A(32x16) * B(16x32).
// Lane l supplies A[m = l % 32][k = 8 (l / 32) + e] and B[k = 8 (l / 32) + e][n = l % 32], e = 0..7;
// it receives D[m][n] for n = l % 32, m = 8 (r / 4) + 4 (l / 32) + (r % 4), r = 0..15.
__device__ __forceinline__ f32x16 mfma32(const u16x8& a, const u16x8& b, const f32x16& c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}
__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// ---- hash grid (multi-resolution, trilinear, Teschner-style spatial hash on fine levels) ---
struct GridLevel {
    int32_t res;   // cells per axis; vertices 0..res
    int32_t dense; // 1: index = x + y (res+1) + z (res+1)^2
    float rf, rf1; // res and res - 1 as floats (kernel arguments: no conversion per point)
};

__device__ __forceinline__ uint32_t grid_index(const GridLevel& g, uint32_t x, uint32_t y, uint32_t z, uint32_t mask)
{
    if (g.dense) {
        const uint32_t n = (uint32_t)g.res + 1u;
        return x + n * (y + n * z);
    }
    // hashed levels: the vertices in 4 x 4 x 4 blocks, a block's 64 entries contiguous (256
    // bytes) and only the block coordinates hashed, so the 8 corners of a cell touch ~2.4
    // 128-byte lines instead of ~4 (the gathers are bound by L2 requests)
    const uint32_t b = ((x >> 2) ^ ((y >> 2) * 2654435761u) ^ ((z >> 2) * 805459861u)) & (mask >> 6);
    return (b << 6) | ((z & 3u) << 4) | ((y & 3u) << 2) | (x & 3u);
}

// Cell and trilinear weights of p (in [0,1]^3, clamped) on level g.
struct Cell {
    uint32_t x, y, z;
    float fx, fy, fz;
};
__device__ __forceinline__ Cell grid_cell(const GridLevel& g, float px, float py, float pz)
{
    Cell c;
    const float r = g.rf;
    float sx = fminf(fmaxf(px, 0.0f), 1.0f) * r, sy = fminf(fmaxf(py, 0.0f), 1.0f) * r,
          sz = fminf(fmaxf(pz, 0.0f), 1.0f) * r;
    float ix = fminf(floorf(sx), g.rf1), iy = fminf(floorf(sy), g.rf1), iz = fminf(floorf(sz), g.rf1);
    c.x = (uint32_t)ix;
    c.y = (uint32_t)iy;
    c.z = (uint32_t)iz;
    c.fx = sx - ix;
    c.fy = sy - iy;
    c.fz = sz - iz;
    return c;
}
// weight of corner k (bit 0: x+1, bit 1: y+1, bit 2: z+1)
__device__ __forceinline__ float corner_weight(const Cell& c, int k)
{
    const float wx = (k & 1) ? c.fx : 1.0f - c.fx;
    const float wy = (k & 2) ? c.fy : 1.0f - c.fy;
    const float wz = (k & 4) ? c.fz : 1.0f - c.fz;
    return wx * wy * wz;
}

// ---- counter-based RNG (SplitMix64 finaliser) -------------------------------------------
__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}
__host__ __device__ inline float rng_unit(uint64_t seed, uint64_t stream, uint64_t i)
{
    uint64_t base = mix64(seed * 0x100000001B3ull + stream * 0x5851F42D4C957F2Dull);
    uint64_t z = mix64(base + (i + 1ull) * 0x9E3779B97F4A7C15ull);
    return (float)(z >> 40) * (1.0f / 16777216.0f);
}

} // namespace nn
} // namespace fwn
