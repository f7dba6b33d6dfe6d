// Learned surrogate for far-from-surface winding-number queries (no reference counterpart).
//
// w(q) ~ MLP(hash_grid(q)): a multi-resolution hash-grid encoding (16 levels, 2^19 entries per
// level, 2 FP16 features) feeds a 32 -> 128 -> 128 -> 128 -> 128 -> 1 ReLU MLP whose hidden layers
// run on BF16 MFMA with FP32 accumulation.  It is trained per mesh on winding numbers generated
// by the exact engine (uniform in the box and near the surface) with Adam, all on the GPU; points
// inside the distance bound, which are always answered exactly, carry no loss:
//
//   k_sample, exact engine      training points and labels
//   k_encode_fwd / k_encode_bwd hash-grid lookup (trilinear over 8 corners) and its scatter-add
//                               gradient (FP32 atomics into the table gradient)
//   k_gemm<TA, TB, EPI>         MFMA GEMM (64x64 tiles, LDS staging, optional transposes, split-K,
//                               bias+ReLU / ReLU-mask epilogues) for every hidden layer forward,
//                               data gradient and weight gradient
//   k_out_fwd / k_out_bwd       128 -> 1 output layer (FP32 VALU)
//   k_adam                      Adam on FP32 master weights, refreshes the FP16 / BF16 copies
//
// Queries: a point falls back to the exact engine when its prediction lies within `band` of 0.5
// or when its cell of a distance-bound bit grid is marked; the grid marks every cell touched by a
// triangle's box (taken from the engine's sorted triangle array) grown by delta.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fwn_internal.h"
#include "fwn_nn.h"

namespace fwn {
using namespace nn;

namespace {

// ============================================================================================
// GEMM: C[M x N] = A[M x K] * B[K x N], bf16 storage, FP32 accumulation.
//   A(m, k) = TA ? A[k * lda + m] : A[m * lda + k];   B(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n]
// Workgroup = 4 waves = 64 x 64 tile, wave (wm, wn) owns a 32 x 32 MFMA accumulator.  Split-K:
// blockIdx.z handles K range [z * kchunk, (z + 1) * kchunk) and EPI_F32 writes partial z.
// ============================================================================================
constexpr int GT = 256;
constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;
enum { EPI_F32 = 0, EPI_BIAS_RELU_BF16 = 1, EPI_MASK_BF16 = 2 };

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(GT) void k_gemm(const uint16_t* __restrict__ A, int lda, const uint16_t* __restrict__ B,
                                             int ldb, int M, int N, int K, int kchunk, float* __restrict__ Cf,
                                             uint16_t* __restrict__ Cb, int ldc, const float* __restrict__ bias,
                                             const uint16_t* __restrict__ mask, int ldm)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[TM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
    const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
    f32x16 acc = {};
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        if (!TA) { // thread: row m, 8 consecutive k
            const int m = tid >> 2, kk = (tid & 3) * 8, gm = m0 + m;
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kk + e;
                As[m][kk + e] = (gm < M && gk < kend) ? A[(size_t)gm * lda + gk] : (uint16_t)0;
            }
        } else { // thread: column k, 8 consecutive m
            const int kk = tid >> 3, m = (tid & 7) * 8, gk = k0 + kk;
            for (int e = 0; e < 8; ++e) {
                const int gm = m0 + m + e;
                As[m + e][kk] = (gm < M && gk < kend) ? A[(size_t)gk * lda + gm] : (uint16_t)0;
            }
        }
        if (!TB) { // B(k, n) at k * ldb + n: thread k, 8 consecutive n
            const int kk = tid >> 3, n = (tid & 7) * 8, gk = k0 + kk;
            for (int e = 0; e < 8; ++e) {
                const int gn = n0 + n + e;
                Bs[n + e][kk] = (gn < N && gk < kend) ? B[(size_t)gk * ldb + gn] : (uint16_t)0;
            }
        } else { // B(k, n) at n * ldb + k: thread n, 8 consecutive k
            const int n = tid >> 2, kk = (tid & 3) * 8, gn = n0 + n;
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kk + e;
                Bs[n][kk + e] = (gn < N && gk < kend) ? B[(size_t)gn * ldb + gk] : (uint16_t)0;
            }
        }
        __syncthreads();
        for (int ks = 0; ks < TK; ks += 16) {
            const u16x8 a = *reinterpret_cast<const u16x8*>(&As[wm * 32 + (lane & 31)][ks + 8 * (lane >> 5)]);
            const u16x8 b = *reinterpret_cast<const u16x8*>(&Bs[wn * 32 + (lane & 31)][ks + 8 * (lane >> 5)]);
            acc = mfma32(a, b, acc);
        }
        __syncthreads();
    }
    const int n = n0 + wn * 32 + (lane & 31);
    if (n >= N) return;
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm * 32 + acc_row(r, lane);
        if (m >= M) continue;
        float v = acc[r];
        if (EPI == EPI_F32) {
            Cf[(size_t)blockIdx.z * (size_t)M * ldc + (size_t)m * ldc + n] = v;
        } else if (EPI == EPI_BIAS_RELU_BF16) {
            v = fmaxf(v + bias[n], 0.0f);
            Cb[(size_t)m * ldc + n] = f2bf(v);
        } else {
            v = (bf2f(mask[(size_t)m * ldm + n]) > 0.0f) ? v : 0.0f;
            Cb[(size_t)m * ldc + n] = f2bf(v);
        }
    }
}

__global__ void k_reduce_parts(const float* __restrict__ parts, int nparts, int64_t len, float* __restrict__ out)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    float s = 0.0f;
    for (int p = 0; p < nparts; ++p) s += parts[(size_t)p * len + j];
    out[j] = s;
}

// column sums of a bf16 [n x N] matrix, accumulated into out[N] (one atomic per column per block)
__global__ void k_colsum(const uint16_t* __restrict__ X, int64_t n, int N, int64_t rows_per_block,
                         float* __restrict__ out)
{
    const int c = threadIdx.x;
    if (c >= N) return;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
    float s = 0.0f;
    for (int64_t r = r0; r < r1; ++r) s += bf2f(X[(size_t)r * N + c]);
    atomicAdd(&out[c], s);
}

// ============================================================================================
// hash-grid encoding
// ============================================================================================
struct GridParams {
    GridLevel lv[kLevelsMax];
    int32_t levels;
    uint32_t mask;    // T - 1
    uint32_t T;       // entries per level
    float lo[3], inv[3]; // normalisation of points into [0, 1]^3
};

// one thread per point: 2 * levels features, stored as bf16 (MLP input) and optionally FP32
__global__ __launch_bounds__(256) void k_encode_fwd(const float* __restrict__ P, int64_t n, GridParams g,
                                                    const __half* __restrict__ table, uint16_t* __restrict__ X0,
                                                    float* __restrict__ F32)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float px = (P[3 * i] - g.lo[0]) * g.inv[0], py = (P[3 * i + 1] - g.lo[1]) * g.inv[1],
                pz = (P[3 * i + 2] - g.lo[2]) * g.inv[2];
    for (int l = 0; l < g.levels; ++l) {
        const Cell c = grid_cell(g.lv[l], px, py, pz);
        float f0 = 0.0f, f1 = 0.0f;
        for (int k = 0; k < 8; ++k) {
            const uint32_t idx = grid_index(g.lv[l], c.x + (k & 1), c.y + ((k >> 1) & 1), c.z + (k >> 2), g.mask);
            const __half2 v = reinterpret_cast<const __half2*>(table)[(size_t)l * g.T + idx];
            const float w = corner_weight(c, k);
            f0 = fmaf(w, __low2float(v), f0);
            f1 = fmaf(w, __high2float(v), f1);
        }
        X0[(size_t)i * 32 + 2 * l] = f2bf(f0);
        X0[(size_t)i * 32 + 2 * l + 1] = f2bf(f1);
        if (F32) {
            F32[(size_t)i * 32 + 2 * l] = f0;
            F32[(size_t)i * 32 + 2 * l + 1] = f1;
        }
    }
}

__global__ __launch_bounds__(256) void k_encode_bwd(const float* __restrict__ P, int64_t n, GridParams g,
                                                    const float* __restrict__ dX0, float* __restrict__ gtable)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float px = (P[3 * i] - g.lo[0]) * g.inv[0], py = (P[3 * i + 1] - g.lo[1]) * g.inv[1],
                pz = (P[3 * i + 2] - g.lo[2]) * g.inv[2];
    for (int l = 0; l < g.levels; ++l) {
        const float d0 = dX0[(size_t)i * 32 + 2 * l], d1 = dX0[(size_t)i * 32 + 2 * l + 1];
        if (d0 == 0.0f && d1 == 0.0f) continue;
        const Cell c = grid_cell(g.lv[l], px, py, pz);
        for (int k = 0; k < 8; ++k) {
            const uint32_t idx = grid_index(g.lv[l], c.x + (k & 1), c.y + ((k >> 1) & 1), c.z + (k >> 2), g.mask);
            const float w = corner_weight(c, k);
            float* dst = gtable + 2 * ((size_t)l * g.T + idx);
            atomicAdd(dst, w * d0);
            atomicAdd(dst + 1, w * d1);
        }
    }
}

// ============================================================================================
// output layer, loss, Adam, sampling, distance-bound grid
// ============================================================================================
__global__ __launch_bounds__(256) void k_out_fwd(const uint16_t* __restrict__ X4, int64_t n,
                                                 const float* __restrict__ w4, const float* __restrict__ b4,
                                                 float* __restrict__ pred)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint16_t* x = X4 + (size_t)i * kHidden;
    float s = 0.0f;
    for (int k = 0; k < kHidden; ++k) s = fmaf(w4[k], bf2f(x[k]), s);
    pred[i] = s + b4[0];
}

__device__ __forceinline__ bool near_cell(const uint32_t* __restrict__ bits, int R, float px, float py, float pz)
{
    // points outside the grid box count as near (the network was not trained there)
    if (!(px >= 0.0f && px <= 1.0f && py >= 0.0f && py <= 1.0f && pz >= 0.0f && pz <= 1.0f)) return true;
    const int x = min(R - 1, (int)(px * (float)R)), y = min(R - 1, (int)(py * (float)R)),
              z = min(R - 1, (int)(pz * (float)R));
    const uint64_t cell = ((uint64_t)z * R + y) * R + x;
    return (bits[cell >> 5] >> (cell & 31u)) & 1u;
}

// Mean squared error and its gradient.  With `bits` set, points in marked cells (always answered
// by the exact engine) carry no loss: the network's capacity goes to the region it serves.
__global__ __launch_bounds__(256) void k_loss_grad(const float* __restrict__ pred, const float* __restrict__ y,
                                                   int64_t n, float scale, const float* __restrict__ P, GridParams g,
                                                   int R, const uint32_t* __restrict__ bits,
                                                   float* __restrict__ dpred, float* __restrict__ loss)
{
    __shared__ float part[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float e = 0.0f;
    if (i < n) {
        e = pred[i] - y[i];
        if (bits) {
            const float px = (P[3 * i] - g.lo[0]) * g.inv[0], py = (P[3 * i + 1] - g.lo[1]) * g.inv[1],
                        pz = (P[3 * i + 2] - g.lo[2]) * g.inv[2];
            if (near_cell(bits, R, px, py, pz)) e = 0.0f;
        }
        dpred[i] = 2.0f * e * scale;
    }
    part[threadIdx.x] = e * e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0] * scale);
}

// dPre3 = (dpred w4^T) * (X4 > 0); gw4 += X4^T dpred; gb4 += sum dpred.  128 threads per block,
// thread k owns column k over a chunk of rows.
__global__ void k_out_bwd(const uint16_t* __restrict__ X4, int64_t n, int64_t rows_per_block,
                          const float* __restrict__ dpred, const float* __restrict__ w4, uint16_t* __restrict__ dP3,
                          float* __restrict__ gw4, float* __restrict__ gb4)
{
    const int k = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
    float gw = 0.0f, gb = 0.0f;
    for (int64_t r = r0; r < r1; ++r) {
        const float x = bf2f(X4[(size_t)r * kHidden + k]);
        const float d = dpred[r];
        gw = fmaf(d, x, gw);
        gb += d;
        dP3[(size_t)r * kHidden + k] = f2bf(x > 0.0f ? d * w4[k] : 0.0f);
    }
    atomicAdd(&gw4[k], gw);
    if (k == 0) atomicAdd(gb4, gb);
}

__global__ void k_adam(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                       uint16_t* __restrict__ bf_out, __half* __restrict__ h_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    p[i] = pi;
    g[i] = 0.0f;
    if (bf_out) bf_out[i] = f2bf(pi);
    if (h_out) h_out[i] = __float2half(pi);
}

__global__ void k_cast(const float* __restrict__ p, int64_t n, uint16_t* __restrict__ bf_out, __half* __restrict__ h_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (bf_out) bf_out[i] = f2bf(p[i]);
    if (h_out) h_out[i] = __float2half(p[i]);
}

// Training points: with probability near_frac a point near the surface (a uniformly chosen
// triangle, a uniform barycentric point, offset +-diag * 10^U(lo_exp, -1) along its normal),
// otherwise uniform in the box.
__global__ __launch_bounds__(256) void k_sample(int64_t n, uint64_t seed, int64_t offset, const float* __restrict__ box,
                                                const uint32_t* __restrict__ tris, int64_t ntris, float near_frac,
                                                float lo_exp, float* __restrict__ P)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const uint64_t i = (uint64_t)(j + offset);
    const float lo[3] = {box[0], box[1], box[2]}, hi[3] = {box[3], box[4], box[5]};
    if (ntris == 0 || rng_unit(seed, 9, i) >= near_frac) {
        for (int d = 0; d < 3; ++d) P[3 * j + d] = lo[d] + (hi[d] - lo[d]) * rng_unit(seed, 10 + d, i);
        return;
    }
    const float u0 = rng_unit(seed, 20, i), u1 = rng_unit(seed, 21, i), u2 = rng_unit(seed, 22, i),
                u3 = rng_unit(seed, 23, i), u4 = rng_unit(seed, 24, i);
    const int64_t t = min((int64_t)(u0 * (float)ntris), ntris - 1);
    const float* v = reinterpret_cast<const float*>(tris + (size_t)t * kTriWords);
    const float s1 = sqrtf(u1), a = 1.0f - s1, b = s1 * (1.0f - u2), c = s1 * u2;
    const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]}, e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    float nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const float len = sqrtf(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    const float inv = len > 0.0f ? 1.0f / len : 0.0f;
    float diag2 = 0.0f;
    for (int d = 0; d < 3; ++d) diag2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    const float off = sqrtf(diag2) * exp10f(lo_exp + (-1.0f - lo_exp) * u3) * (u4 < 0.5f ? -1.0f : 1.0f);
    for (int d = 0; d < 3; ++d) P[3 * j + d] = a * v[d] + b * v[3 + d] + c * v[6 + d] + off * nn[d] * inv;
}

__global__ void k_scale(const float* __restrict__ in, int64_t n, float s, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] * s;
}

// Distance-bound grid: every cell overlapped by a triangle's box grown by delta is marked.
__global__ __launch_bounds__(256) void k_mark_near(const uint32_t* __restrict__ tris, int64_t ntris, GridParams g,
                                                   int R, float delta, uint32_t* __restrict__ bits)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= ntris) return;
    const float* v = reinterpret_cast<const float*>(tris + (size_t)t * kTriWords);
    int c0[3], c1[3];
    for (int d = 0; d < 3; ++d) {
        const float lo = fminf(v[d], fminf(v[3 + d], v[6 + d])) - delta;
        const float hi = fmaxf(v[d], fmaxf(v[3 + d], v[6 + d])) + delta;
        c0[d] = max(0, min(R - 1, (int)floorf((lo - g.lo[d]) * g.inv[d] * (float)R)));
        c1[d] = max(0, min(R - 1, (int)floorf((hi - g.lo[d]) * g.inv[d] * (float)R)));
    }
    for (int z = c0[2]; z <= c1[2]; ++z)
        for (int y = c0[1]; y <= c1[1]; ++y)
            for (int x = c0[0]; x <= c1[0]; ++x) {
                const uint64_t cell = ((uint64_t)z * R + y) * R + x;
                atomicOr(&bits[cell >> 5], 1u << (cell & 31u));
            }
}

// fallback flag per point: prediction within band of 0.5, or a marked cell
__global__ __launch_bounds__(256) void k_fallback_flags(const float* __restrict__ P, const float* __restrict__ pred,
                                                        int64_t n, GridParams g, int R,
                                                        const uint32_t* __restrict__ bits, float band,
                                                        uint8_t* __restrict__ fb)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float px = (P[3 * i] - g.lo[0]) * g.inv[0], py = (P[3 * i + 1] - g.lo[1]) * g.inv[1],
                pz = (P[3 * i + 2] - g.lo[2]) * g.inv[2];
    fb[i] = (fabsf(pred[i] - 0.5f) < band || near_cell(bits, R, px, py, pz)) ? 1 : 0;
}

// ============================================================================================
// inference: one fused kernel (hash-grid encoding, distance-bound flag, 4-layer MFMA MLP)
// ============================================================================================
// Physical k index -> logical input feature: k-step s, lane half h, element e.  With this order
// the accumulator registers of one layer are the B fragments of the next (see k_sg_mlp).
__host__ __device__ inline int kin_of(int s, int h, int e) { return 32 * (s >> 1) + 8 * (2 * (s & 1) + (e >> 2)) + 4 * h + (e & 3); }

// A fragments of W^T for every (layer, output tile t, k-step s): lane (h = l / 32, m = l % 32)
// holds W[kin_of(s, h, e)][32 t + m], e = 0..7.  Fragment order: layer 0 (t, s) = t * 2 + s,
// then layers 1..3 at 8 + (layer - 1) * 32 + t * 8 + s.
__global__ void k_pack_frags(const uint16_t* __restrict__ mlp_bf, const int64_t* __restrict__ off_w,
                             uint16_t* __restrict__ frags)
{
    const int f = blockIdx.x, lane = threadIdx.x;
    int layer, t, s;
    if (f < 8) {
        layer = 0;
        t = f >> 1;
        s = f & 1;
    } else {
        layer = 1 + (f - 8) / 32;
        t = ((f - 8) % 32) >> 3;
        s = (f - 8) & 7;
    }
    const uint16_t* W = mlp_bf + off_w[layer];
    const int h = lane >> 5, m = lane & 31;
    for (int e = 0; e < 8; ++e)
        frags[((size_t)f * 64 + lane) * 8 + e] = W[(size_t)kin_of(s, h, e) * kHidden + 32 * t + m];
}

// waves per inference workgroup: two per SIMD (at most 256 registers each), so that one wave's
// table gathers overlap the other's MFMAs; the workgroup's waves share one LDS copy of the packed
// weights (106 KB, one workgroup per CU)
constexpr int IW = 8;
constexpr int64_t kSortMin = 1 << 16; // batches at least this large are encoded in Morton order
constexpr int kFrags = 8 + 3 * 32;    // A fragments of the four hidden layers

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
// two floats -> packed BF16 (a in the low half), round to nearest even: one v_cvt_pk_bf16_f32 on
// gfx950, the same values as f2bf for finite inputs
__device__ __forceinline__ uint32_t pack2(float a, float b)
{
    const f32x2_t x = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2_t));
}
typedef short s16x2_t __attribute__((ext_vector_type(2)));
// ReLU of two packed BF16 values as one packed 16-bit integer max with 0: BF16 bit patterns of
// negative values (and -0) are negative as int16, so max(pattern, 0) is +0 for them and leaves
// the rest unchanged; round-to-nearest-even keeps the sign, so relu_pk(pack2(a, b)) equals
// pack2(max(a, 0), max(b, 0)) for finite a, b
__device__ __forceinline__ uint32_t relu_pk(uint32_t w)
{
    const s16x2_t z = {0, 0};
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2_t, w), z));
}

// One hidden layer: for each 32-row output tile t, acc = sum_s A(layer, t, s) B[p][s], then bias +
// ReLU + BF16; registers 8q..8q+7 of tile t become fragment 2t+q of the next layer's input (Bn),
// or, for the last layer, feed the output-layer dot product part[p].  Only one tile's
// accumulators are live at a time; the scheduling barrier keeps the next tile's LDS loads from
// being hoisted above this tile's MFMAs.
template <int KS, int FBASE, bool LAST>
__device__ __forceinline__ void mlp_layer(const u16x8* __restrict__ sA, int lane, int h, const u16x8 (&B)[2][8],
                                          u16x8 (&Bn)[2][8], const float* __restrict__ bias,
                                          const float* __restrict__ w4, float (&part)[2])
{
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x16 acc[2] = {f32x16{}, f32x16{}};
#pragma unroll
        for (int s2 = 0; s2 < KS; ++s2) {
            const u16x8 a = sA[(FBASE + t * KS + s2) * 64 + lane];
            acc[0] = mfma32(a, B[0][s2], acc[0]);
            acc[1] = mfma32(a, B[1][s2], acc[1]);
        }
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            // bias + ReLU, rounded to BF16 in pairs (the next layer's input, or the output layer's)
            uint32_t w[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int r = 2 * i;
                const int m = 32 * t + 8 * (r >> 2) + 4 * h + (r & 3);
                w[i] = relu_pk(pack2(acc[p][r] + bias[m], acc[p][r + 1] + bias[m + 1]));
            }
            if (!LAST) {
#pragma unroll
                for (int q = 0; q < 2; ++q)
                    Bn[p][2 * t + q] =
                        __builtin_bit_cast(u16x8, make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]));
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int r = 2 * i;
                    const int m = 32 * t + 8 * (r >> 2) + 4 * h + (r & 3);
                    part[p] = fmaf(w4[m], __uint_as_float(w[i] << 16), part[p]);
                    part[p] = fmaf(w4[m + 1], __uint_as_float(w[i] & 0xFFFF0000u), part[p]);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Hash-grid features of one point, kLevelsMax levels: trilinear interpolation of the 8 corners'
// FP16 pairs, rounded to BF16 (f0 in the low half).  The corners of 4 levels (32 loads) are
// requested before any of them is used, so a wave waits for 4 memory round trips per point
// rather than 16 (the loads per wave in flight stay below the 63 the counter tracks).
__device__ __forceinline__ void encode_point(const GridParams& g, const __half2* __restrict__ tab, float px, float py,
                                             float pz, uint32_t (&f)[kLevelsMax])
{
#pragma unroll
    for (int l0 = 0; l0 < kLevelsMax; l0 += 4) {
        __half2 v[4][8];
#pragma unroll
        for (int dl = 0; dl < 4; ++dl) {
            const GridLevel& lv = g.lv[l0 + dl];
            const Cell c = grid_cell(lv, px, py, pz);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                v[dl][k] = tab[(size_t)(l0 + dl) * g.T +
                               grid_index(lv, c.x + (k & 1), c.y + ((k >> 1) & 1), c.z + (k >> 2), g.mask)];
        }
#pragma unroll
        for (int dl = 0; dl < 4; ++dl) {
            const Cell c = grid_cell(g.lv[l0 + dl], px, py, pz);
            float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const float w = corner_weight(c, k);
                f0 = fmaf(w, __low2float(v[dl][k]), f0);
                f1 = fmaf(w, __high2float(v[dl][k]), f1);
            }
            f[l0 + dl] = (uint32_t)f2bf(f0) | ((uint32_t)f2bf(f1) << 16);
        }
    }
}

// One wave = 64 points per tile as two 32-point MFMA column tiles (p = 0, 1).  Lane (h, n) owns
// point n of each column tile: it feeds B(k, n) of Y^T = W^T X^T and receives D(m, n) for
// m = 32 t + 8 (r / 4) + 4 h + r % 4.  Encoding first: lane l encodes row l of the tile (all 16
// levels, the level uniform across the wave) and writes the row's 16 feature pairs, its
// distance-bound flag and its point index to the wave's rows in LDS; the layer-0 B fragments
// (features kin_of(s, h, e) of rows n and 32 + n) are then read from there, so nothing but the
// packed weights and these rows sits in LDS and nothing is live across the MLP.  Row i of the
// tile order is point perm[i] (Morton order for large batches: neighbouring lanes gather from
// neighbouring cells); results are written to the point's own index.  Bias, ReLU and BF16
// rounding are applied in registers; the fallback flag is the distance-bound cell's bit or a
// prediction within `band` of 0.5.  Two waves per SIMD: one wave's gathers overlap the other's
// MFMAs.
constexpr int kRowWords = 20; // LDS row: 16 feature pairs, near flag, point index (16-byte aligned rows)

__global__ __launch_bounds__(IW * 64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_sg_infer(
    const float* __restrict__ P, const uint32_t* __restrict__ perm, int64_t n, GridParams g,
    const __half* __restrict__ table, int R, const uint32_t* __restrict__ bits, const uint16_t* __restrict__ frags,
    const float* __restrict__ mlp, int64_t off_b0, int64_t off_b1, int64_t off_b2, int64_t off_b3, int64_t off_w4,
    int64_t off_b4, float band, float* __restrict__ pred, uint8_t* __restrict__ fb)
{
    __shared__ u16x8 sA[kFrags * 64];
    __shared__ float sb[4][kHidden];
    __shared__ float sw4[kHidden];
    __shared__ __attribute__((aligned(16))) uint32_t sX[IW][64 * kRowWords];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < kFrags * 64; i += IW * 64) sA[i] = reinterpret_cast<const u16x8*>(frags)[i];
    for (int i = tid; i < kHidden; i += IW * 64) {
        sb[0][i] = mlp[off_b0 + i];
        sb[1][i] = mlp[off_b1 + i];
        sb[2][i] = mlp[off_b2 + i];
        sb[3][i] = mlp[off_b3 + i];
        sw4[i] = mlp[off_w4 + i];
    }
    __syncthreads();
    const float b4 = mlp[off_b4];
    const int h = lane >> 5, nn = lane & 31;
    const __half2* tab = reinterpret_cast<const __half2*>(table);
    uint32_t* rows = sX[wave];
    const int64_t ntiles = (n + 63) / 64;
    // XCD-aware tile order: the hardware hands workgroup b to XCD b % 8, so with a grid of a
    // multiple of 8 workgroups each XCD takes one contiguous eighth of the (Morton-ordered) tiles
    // and its waves walk it together: at any time an XCD's waves encode one compact region, whose
    // points share corners in that XCD's L2
    const int nx = (gridDim.x % 8 == 0) ? 8 : 1;
    const int64_t xb = (int64_t)(blockIdx.x % nx) * ntiles / nx, xe = (int64_t)(blockIdx.x % nx + 1) * ntiles / nx;
    const int64_t wstride = (int64_t)(gridDim.x / nx) * IW;
    for (int64_t tile = xb + (int64_t)(blockIdx.x / nx) * IW + wave; tile < xe; tile += wstride) {
        {
            const int64_t pt = tile * 64 + lane;
            const bool ok = pt < n;
            const uint32_t j = ok ? (perm ? perm[pt] : (uint32_t)pt) : 0u;
            const float px = (P[3 * (size_t)j] - g.lo[0]) * g.inv[0], py = (P[3 * (size_t)j + 1] - g.lo[1]) * g.inv[1],
                        pz = (P[3 * (size_t)j + 2] - g.lo[2]) * g.inv[2];
            uint32_t f[kLevelsMax];
            encode_point(g, tab, px, py, pz, f); // a row past n encodes point 0 and is zeroed
            uint4* row = reinterpret_cast<uint4*>(rows + lane * kRowWords);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                row[q] = ok ? make_uint4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]) : make_uint4(0u, 0u, 0u, 0u);
            row[4] = make_uint4(near_cell(bits, R, px, py, pz) ? 1u : 0u, j, 0u, 0u);
        }
        // LDS operations of one wave complete in order; the barrier keeps the compiler from
        // moving the row reads above the row writes
        __builtin_amdgcn_wave_barrier();
        u16x8 B[2][8];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const uint32_t* row = rows + (32 * p + nn) * kRowWords;
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2) {
                // features 16 s2 + 4 h + 0..3 and 16 s2 + 8 + 4 h + 0..3: levels 8 s2 + 2 h + {0, 1, 4, 5}
                const uint2 lo = *reinterpret_cast<const uint2*>(row + 8 * s2 + 2 * h);
                const uint2 hi = *reinterpret_cast<const uint2*>(row + 8 * s2 + 2 * h + 4);
                B[p][s2] = __builtin_bit_cast(u16x8, make_uint4(lo.x, lo.y, hi.x, hi.y));
            }
        }
        u16x8 C[2][8];
        float part[2] = {0.0f, 0.0f};
        mlp_layer<2, 0, false>(sA, lane, h, B, C, sb[0], sw4, part);
        mlp_layer<8, 8, false>(sA, lane, h, C, B, sb[1], sw4, part);
        mlp_layer<8, 40, false>(sA, lane, h, B, C, sb[2], sw4, part);
        mlp_layer<8, 72, true>(sA, lane, h, C, B, sb[3], sw4, part);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const float tot = part[p] + __shfl_xor(part[p], 32);
            const int64_t pt = tile * 64 + 32 * p + nn;
            if (h == 0 && pt < n) {
                const uint2 nj = *reinterpret_cast<const uint2*>(rows + (32 * p + nn) * kRowWords + 16);
                const float y = tot + b4;
                pred[nj.y] = y;
                fb[nj.y] = (nj.x || fabsf(y - 0.5f) < band) ? 1 : 0;
            }
        }
        __builtin_amdgcn_wave_barrier(); // the next tile's rows are written after these reads
    }
}

} // namespace

// ============================================================================================
// host object
// ============================================================================================
struct Surrogate {
    Engine* exact = nullptr; // not owned; must outlive the surrogate (fwn.h)
    int device = 0;
    fwn_surrogate_options opt{};
    GridParams grid{};
    int64_t n_table = 0;  // floats in the table (L * T * 2)
    int64_t n_mlp = 0;    // floats in the packed MLP
    int64_t off_w[5] = {}, off_b[5] = {};
    int R = 0;
    float delta = 0.0f;
    bool trained = false;
    float last_loss = 0.0f;
    double train_ms = 0.0;

    DBuf table, table_h, g_table, m_table, v_table;
    DBuf mlp, mlp_bf, g_mlp, m_mlp, v_mlp;
    DBuf bits, frags, offw;
    bool frags_ok = false;
    // activations / scratch (grown on demand)
    DBuf P, y, X0, Xh[4], dP, dQ, dX0, pred, dpred, parts, loss, fb;
    int adam_t = 0;
    int n_cu = 0;
    hipEvent_t ev[6] = {};
    double encode_ms = 0.0, mlp_ms = 0.0, fallback_ms = 0.0, query_ms = 0.0;
};

namespace {

void sg_check(const fwn_surrogate_options& o)
{
    if (o.levels < 1 || o.levels > kLevelsMax) fail(FWN_INVALID_ARGUMENT, "levels must be in [1, 16]");
    if (o.log2_table < 10 || o.log2_table > 24) fail(FWN_INVALID_ARGUMENT, "log2_table must be in [10, 24]");
    if (o.base_res < 2 || o.max_res < o.base_res) fail(FWN_INVALID_ARGUMENT, "resolutions must satisfy 2 <= base <= max");
    if (o.hidden_layers != 4) fail(FWN_INVALID_ARGUMENT, "the MLP has 4 hidden layers of width 128");
    if (o.batch < 256 || o.batch % 256) fail(FWN_INVALID_ARGUMENT, "batch must be a positive multiple of 256");
    if (o.steps < 0 || o.n_samples < o.batch) fail(FWN_INVALID_ARGUMENT, "need steps >= 0 and n_samples >= batch");
    if (!(o.lr > 0.0f) || !(o.band >= 0.0f) || !(o.band < 0.5f)) fail(FWN_INVALID_ARGUMENT, "bad lr or band");
    if (o.grid_res < 8 || o.grid_res > 2048) fail(FWN_INVALID_ARGUMENT, "grid_res must be in [8, 2048]");
    if (!(o.dist_rel >= 0.0f) || !(o.near_frac >= 0.0f && o.near_frac <= 1.0f))
        fail(FWN_INVALID_ARGUMENT, "need dist_rel >= 0 and near_frac in [0, 1]");
}

inline int gb(int64_t n, int t) { return grid_blocks(n, t); }

// Y = relu(X W + b) for one hidden layer (row-major activations)
void layer_fwd(Surrogate& S, hipStream_t s, const uint16_t* X, int K, int l, uint16_t* Y, int64_t n)
{
    const uint16_t* W = S.mlp_bf.as<uint16_t>() + S.off_w[l];
    const float* b = S.mlp.as<float>() + S.off_b[l];
    dim3 grid((unsigned)((n + TM - 1) / TM), (unsigned)((kHidden + TN - 1) / TN), 1);
    hipLaunchKernelGGL((k_gemm<false, false, EPI_BIAS_RELU_BF16>), grid, dim3(GT), 0, s, X, K, W, kHidden, (int)n,
                       kHidden, K, K, (float*)nullptr, Y, kHidden, b, (const uint16_t*)nullptr, 0);
    FWN_LAUNCH_CHECK("k_gemm fwd");
}

void forward(Surrogate& S, hipStream_t s, const float* P, int64_t n, bool keep_f32)
{
    S.X0.reserve(sizeof(uint16_t) * 32 * n);
    for (int l = 0; l < 4; ++l) S.Xh[l].reserve(sizeof(uint16_t) * kHidden * n);
    S.pred.reserve(sizeof(float) * n);
    if (keep_f32) S.dX0.reserve(sizeof(float) * 32 * n);
    hipLaunchKernelGGL(k_encode_fwd, dim3(gb(n, 256)), dim3(256), 0, s, P, n, S.grid, S.table_h.as<__half>(),
                       S.X0.as<uint16_t>(), keep_f32 ? S.dX0.as<float>() : (float*)nullptr);
    FWN_LAUNCH_CHECK("k_encode_fwd");
    layer_fwd(S, s, S.X0.as<uint16_t>(), 32, 0, S.Xh[0].as<uint16_t>(), n);
    for (int l = 1; l < 4; ++l) layer_fwd(S, s, S.Xh[l - 1].as<uint16_t>(), kHidden, l, S.Xh[l].as<uint16_t>(), n);
    hipLaunchKernelGGL(k_out_fwd, dim3(gb(n, 256)), dim3(256), 0, s, S.Xh[3].as<uint16_t>(), n,
                       S.mlp.as<float>() + S.off_w[4], S.mlp.as<float>() + S.off_b[4], S.pred.as<float>());
    FWN_LAUNCH_CHECK("k_out_fwd");
}

// gradient of the weights of layer l: gW = X^T dP (split over rows), gb = colsum(dP)
void layer_wgrad(Surrogate& S, hipStream_t s, const uint16_t* X, int K, const uint16_t* dP, int l, int64_t n)
{
    const int64_t chunk = 1024; // split-K: enough workgroups to fill the chip at batch 65536
    const int nparts = (int)((n + chunk - 1) / chunk);
    S.parts.reserve(sizeof(float) * (size_t)nparts * K * kHidden);
    dim3 grid((unsigned)((K + TM - 1) / TM), (unsigned)((kHidden + TN - 1) / TN), (unsigned)nparts);
    hipLaunchKernelGGL((k_gemm<true, false, EPI_F32>), grid, dim3(GT), 0, s, X, K, dP, kHidden, K, kHidden, (int)n,
                       (int)chunk, S.parts.as<float>(), (uint16_t*)nullptr, kHidden, (const float*)nullptr,
                       (const uint16_t*)nullptr, 0);
    FWN_LAUNCH_CHECK("k_gemm wgrad");
    // add the reduced partials into the gradient buffer
    float* gW = S.g_mlp.as<float>() + S.off_w[l];
    hipLaunchKernelGGL(k_reduce_parts, dim3(gb((int64_t)K * kHidden, 256)), dim3(256), 0, s, S.parts.as<float>(), nparts,
                       (int64_t)K * kHidden, gW);
    FWN_LAUNCH_CHECK("k_reduce_parts");
    const int64_t rows = 64;
    hipLaunchKernelGGL(k_colsum, dim3(gb(n, (int)rows)), dim3(kHidden), 0, s, dP, n, kHidden, rows,
                       S.g_mlp.as<float>() + S.off_b[l]);
    FWN_LAUNCH_CHECK("k_colsum");
}

// one forward + backward over n points (targets y), gradients accumulated into g_table / g_mlp;
// the MLP gradient buffer is overwritten for the weights (reduced partials) and added for biases.
// With mask_near, points in cells of the distance-bound grid carry no loss.
float forward_backward(Surrogate& S, hipStream_t s, const float* P, const float* y, int64_t n, bool mask_near)
{
    forward(S, s, P, n, false);
    S.dpred.reserve(sizeof(float) * n);
    S.loss.reserve(sizeof(float));
    S.dP.reserve(sizeof(uint16_t) * kHidden * n);
    S.dQ.reserve(sizeof(uint16_t) * kHidden * n);
    S.dX0.reserve(sizeof(float) * 32 * n);
    FWN_HIP(hipMemsetAsync(S.loss.p, 0, sizeof(float), s));
    hipLaunchKernelGGL(k_loss_grad, dim3(gb(n, 256)), dim3(256), 0, s, S.pred.as<float>(), y, n, 1.0f / (float)n, P,
                       S.grid, S.R, mask_near ? S.bits.as<uint32_t>() : (const uint32_t*)nullptr,
                       S.dpred.as<float>(), S.loss.as<float>());
    FWN_LAUNCH_CHECK("k_loss_grad");
    const int64_t rows = 64;
    hipLaunchKernelGGL(k_out_bwd, dim3(gb(n, (int)rows)), dim3(kHidden), 0, s, S.Xh[3].as<uint16_t>(), n, rows,
                       S.dpred.as<float>(), S.mlp.as<float>() + S.off_w[4], S.dP.as<uint16_t>(),
                       S.g_mlp.as<float>() + S.off_w[4], S.g_mlp.as<float>() + S.off_b[4]);
    FWN_LAUNCH_CHECK("k_out_bwd");
    uint16_t* dcur = S.dP.as<uint16_t>();
    uint16_t* dnext = S.dQ.as<uint16_t>();
    for (int l = 3; l >= 1; --l) {
        // dP_l is the gradient at the pre-activation of layer l (input Xh[l-1])
        layer_wgrad(S, s, S.Xh[l - 1].as<uint16_t>(), kHidden, dcur, l, n);
        const uint16_t* W = S.mlp_bf.as<uint16_t>() + S.off_w[l];
        dim3 grid((unsigned)((n + TM - 1) / TM), (unsigned)((kHidden + TN - 1) / TN), 1);
        hipLaunchKernelGGL((k_gemm<false, true, EPI_MASK_BF16>), grid, dim3(GT), 0, s, dcur, kHidden, W, kHidden,
                           (int)n, kHidden, kHidden, kHidden, (float*)nullptr, dnext, kHidden, (const float*)nullptr,
                           S.Xh[l - 1].as<uint16_t>(), kHidden);
        FWN_LAUNCH_CHECK("k_gemm dgrad");
        std::swap(dcur, dnext);
    }
    layer_wgrad(S, s, S.X0.as<uint16_t>(), 32, dcur, 0, n);
    {
        const uint16_t* W0 = S.mlp_bf.as<uint16_t>() + S.off_w[0];
        dim3 grid((unsigned)((n + TM - 1) / TM), 1, 1);
        hipLaunchKernelGGL((k_gemm<false, true, EPI_F32>), grid, dim3(GT), 0, s, dcur, kHidden, W0, kHidden, (int)n, 32,
                           kHidden, kHidden, S.dX0.as<float>(), (uint16_t*)nullptr, 32, (const float*)nullptr,
                           (const uint16_t*)nullptr, 0);
        FWN_LAUNCH_CHECK("k_gemm dX0");
    }
    hipLaunchKernelGGL(k_encode_bwd, dim3(gb(n, 256)), dim3(256), 0, s, P, n, S.grid, S.dX0.as<float>(),
                       S.g_table.as<float>());
    FWN_LAUNCH_CHECK("k_encode_bwd");
    float h_loss = 0.0f;
    FWN_HIP(hipMemcpyAsync(&h_loss, S.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
    return h_loss;
}

void refresh_copies(Surrogate& S, hipStream_t s)
{
    hipLaunchKernelGGL(k_cast, dim3(gb(S.n_table, 256)), dim3(256), 0, s, S.table.as<float>(), S.n_table,
                       (uint16_t*)nullptr, S.table_h.as<__half>());
    hipLaunchKernelGGL(k_cast, dim3(gb(S.n_mlp, 256)), dim3(256), 0, s, S.mlp.as<float>(), S.n_mlp,
                       S.mlp_bf.as<uint16_t>(), (__half*)nullptr);
    FWN_LAUNCH_CHECK("k_cast");
}

void adam_step(Surrogate& S, hipStream_t s, float lr)
{
    ++S.adam_t;
    const float b1 = 0.9f, b2 = 0.99f, eps = 1e-10f;
    const float c1 = 1.0f / (1.0f - powf(b1, (float)S.adam_t)), c2 = 1.0f / (1.0f - powf(b2, (float)S.adam_t));
    hipLaunchKernelGGL(k_adam, dim3(gb(S.n_table, 256)), dim3(256), 0, s, S.table.as<float>(), S.g_table.as<float>(),
                       S.m_table.as<float>(), S.v_table.as<float>(), S.n_table, lr, b1, b2, eps, c1, c2,
                       (uint16_t*)nullptr, S.table_h.as<__half>());
    hipLaunchKernelGGL(k_adam, dim3(gb(S.n_mlp, 256)), dim3(256), 0, s, S.mlp.as<float>(), S.g_mlp.as<float>(),
                       S.m_mlp.as<float>(), S.v_mlp.as<float>(), S.n_mlp, lr, b1, b2, eps, c1, c2,
                       S.mlp_bf.as<uint16_t>(), (__half*)nullptr);
    FWN_LAUNCH_CHECK("k_adam");
}

void zero(DBuf& b, int64_t bytes, hipStream_t s) { FWN_HIP(hipMemsetAsync(b.p, 0, bytes, s)); }

} // namespace

void surrogate_init(Surrogate& S, Engine* exact, const fwn_surrogate_options& o)
{
    sg_check(o);
    if (exact->n_meshes != 1) fail(FWN_INVALID_ARGUMENT, "the surrogate is built on a single-mesh engine");
    S.exact = exact;
    S.device = exact->device;
    S.opt = o;
    hipStream_t s = exact->stream;
    GridParams& g = S.grid;
    g.levels = o.levels;
    g.T = 1u << o.log2_table;
    g.mask = g.T - 1u;
    const double growth = (o.levels > 1) ? std::exp(std::log((double)o.max_res / o.base_res) / (o.levels - 1)) : 1.0;
    for (int l = 0; l < o.levels; ++l) {
        const int res = (int)std::floor(o.base_res * std::pow(growth, l) + 0.5);
        g.lv[l].res = res;
        g.lv[l].rf = (float)res;
        g.lv[l].rf1 = (float)res - 1.0f;
        const double nv = (double)(res + 1) * (res + 1) * (res + 1);
        g.lv[l].dense = nv <= (double)g.T ? 1 : 0;
    }
    float mi[8];
    FWN_HIP(hipMemcpy(mi, exact->mesh_info.p, sizeof(mi), hipMemcpyDeviceToHost));
    float diag2 = 0.0f;
    for (int d = 0; d < 3; ++d) {
        g.lo[d] = mi[d];
        g.inv[d] = mi[3 + d];
        const float ext = mi[3 + d] > 0.0f ? 1.0f / mi[3 + d] : 0.0f;
        diag2 += ext * ext;
    }
    S.R = o.grid_res;
    S.delta = o.dist_rel * std::sqrt(diag2);
    S.n_table = (int64_t)o.levels * g.T * kFeat;
    int64_t off = 0;
    const int K[4] = {2 * o.levels, kHidden, kHidden, kHidden};
    if (K[0] != 32) fail(FWN_INVALID_ARGUMENT, "the MLP input is 32 features (16 levels x 2)");
    for (int l = 0; l < 4; ++l) {
        S.off_w[l] = off;
        off += (int64_t)K[l] * kHidden;
        S.off_b[l] = off;
        off += kHidden;
    }
    S.off_w[4] = off;
    off += kHidden;
    S.off_b[4] = off;
    off += 1;
    S.n_mlp = off;

    S.table.alloc(sizeof(float) * S.n_table);
    S.table_h.alloc(sizeof(__half) * S.n_table);
    S.g_table.alloc(sizeof(float) * S.n_table);
    S.m_table.alloc(sizeof(float) * S.n_table);
    S.v_table.alloc(sizeof(float) * S.n_table);
    S.mlp.alloc(sizeof(float) * S.n_mlp);
    S.mlp_bf.alloc(sizeof(uint16_t) * S.n_mlp);
    S.g_mlp.alloc(sizeof(float) * S.n_mlp);
    S.m_mlp.alloc(sizeof(float) * S.n_mlp);
    S.v_mlp.alloc(sizeof(float) * S.n_mlp);
    // deterministic initialisation on the host: table U(-1e-4, 1e-4); weights U(-sqrt(6/K), sqrt(6/K));
    // biases 0; output bias 0.5
    std::vector<float> h(S.n_table);
    for (int64_t i = 0; i < S.n_table; ++i) h[i] = (rng_unit(o.seed, 1, (uint64_t)i) * 2.0f - 1.0f) * 1e-4f;
    FWN_HIP(hipMemcpy(S.table.p, h.data(), sizeof(float) * S.n_table, hipMemcpyHostToDevice));
    std::vector<float> w(S.n_mlp, 0.0f);
    for (int l = 0; l < 5; ++l) {
        const int kin = l < 4 ? K[l] : kHidden, nout = l < 4 ? kHidden : 1;
        const float a = std::sqrt(6.0f / kin);
        for (int64_t i = 0; i < (int64_t)kin * nout; ++i)
            w[S.off_w[l] + i] = (rng_unit(o.seed, 2 + l, (uint64_t)i) * 2.0f - 1.0f) * a;
    }
    w[S.off_b[4]] = 0.5f;
    FWN_HIP(hipMemcpy(S.mlp.p, w.data(), sizeof(float) * S.n_mlp, hipMemcpyHostToDevice));
    zero(S.g_table, sizeof(float) * S.n_table, s);
    zero(S.m_table, sizeof(float) * S.n_table, s);
    zero(S.v_table, sizeof(float) * S.n_table, s);
    zero(S.g_mlp, sizeof(float) * S.n_mlp, s);
    zero(S.m_mlp, sizeof(float) * S.n_mlp, s);
    zero(S.v_mlp, sizeof(float) * S.n_mlp, s);
    refresh_copies(S, s);
    // distance-bound grid
    const uint64_t cells = (uint64_t)S.R * S.R * S.R;
    S.bits.alloc(sizeof(uint32_t) * ((cells + 31) / 32));
    zero(S.bits, sizeof(uint32_t) * ((cells + 31) / 32), s);
    if (exact->n_tris > 0) {
        hipLaunchKernelGGL(k_mark_near, dim3(gb(exact->n_tris, 256)), dim3(256), 0, s, exact->tris.as<uint32_t>(),
                           exact->n_tris, g, S.R, S.delta, S.bits.as<uint32_t>());
        FWN_LAUNCH_CHECK("k_mark_near");
    }
    FWN_HIP(hipStreamSynchronize(s));
}

void pack_frags(Surrogate& S, hipStream_t s)
{
    if (!S.offw.p) {
        S.offw.alloc(sizeof(int64_t) * 5);
        FWN_HIP(hipMemcpyAsync(S.offw.p, S.off_w, sizeof(int64_t) * 5, hipMemcpyHostToDevice, s));
    }
    S.frags.reserve(sizeof(uint16_t) * kFrags * 64 * 8);
    hipLaunchKernelGGL(k_pack_frags, dim3(kFrags), dim3(64), 0, s, S.mlp_bf.as<uint16_t>(), S.offw.as<int64_t>(),
                       S.frags.as<uint16_t>());
    FWN_LAUNCH_CHECK("k_pack_frags");
    S.frags_ok = true;
}

void surrogate_refresh(Surrogate& S)
{
    refresh_copies(S, S.exact->stream);
    S.frags_ok = false;
    FWN_HIP(hipStreamSynchronize(S.exact->stream));
}

float surrogate_gradients(Surrogate& S, const float* P, const float* y, int64_t n)
{
    hipStream_t s = S.exact->stream;
    zero(S.g_table, sizeof(float) * S.n_table, s);
    zero(S.g_mlp, sizeof(float) * S.n_mlp, s);
    return forward_backward(S, s, P, y, n, false);
}

void surrogate_train(Surrogate& S)
{
    Engine& E = *S.exact;
    hipStream_t s = E.stream;
    const int64_t ns = S.opt.n_samples;
    hipEvent_t t0, t1;
    FWN_HIP(hipEventCreate(&t0));
    FWN_HIP(hipEventCreate(&t1));
    FWN_HIP(hipEventRecord(t0, s));
    S.P.reserve(sizeof(float) * 3 * ns);
    S.y.reserve(sizeof(float) * ns);
    // sample box = the mesh's vertex box
    float box[6];
    for (int d = 0; d < 3; ++d) {
        box[d] = S.grid.lo[d];
        box[3 + d] = S.grid.lo[d] + (S.grid.inv[d] > 0.0f ? 1.0f / S.grid.inv[d] : 0.0f);
    }
    DBuf dbox;
    dbox.alloc(sizeof(box));
    FWN_HIP(hipMemcpyAsync(dbox.p, box, sizeof(box), hipMemcpyHostToDevice, s));
    // near-surface offsets start at half the distance bound: closer points are answered exactly
    const float lo_exp = std::min(-1.5f, std::log10(std::max(S.opt.dist_rel, 1e-7f)) - 0.3f);
    hipLaunchKernelGGL(k_sample, dim3(gb(ns, 256)), dim3(256), 0, s, ns, S.opt.seed, (int64_t)0, dbox.as<float>(),
                       E.tris.as<uint32_t>(), E.n_tris, S.opt.near_frac, lo_exp, S.P.as<float>());
    FWN_LAUNCH_CHECK("k_sample");
    // labels from the exact engine: w = omega / (4 pi)
    run_queries(E, S.P.as<float>(), nullptr, ns, S.y.as<float>(), nullptr, nullptr, true);
    hipLaunchKernelGGL(k_scale, dim3(gb(ns, 256)), dim3(256), 0, s, S.y.as<float>(), ns,
                       (float)(1.0 / (4.0 * 3.14159265358979323846)), S.y.as<float>());
    FWN_LAUNCH_CHECK("k_scale");
    const int64_t B = S.opt.batch;
    const int64_t nb = ns / B;
    float loss = 0.0f;
    for (int step = 0; step < S.opt.steps; ++step) {
        const int64_t b = step % nb;
        // learning rate: constant, then x0.1 decay over the last third
        const float frac = (float)step / (float)std::max(1, S.opt.steps);
        const float lr = S.opt.lr * (frac < 2.0f / 3.0f ? 1.0f : std::pow(0.1f, (frac - 2.0f / 3.0f) * 3.0f));
        loss = forward_backward(S, s, S.P.as<float>() + 3 * b * B, S.y.as<float>() + b * B, B, true);
        adam_step(S, s, lr);
    }
    FWN_HIP(hipEventRecord(t1, s));
    FWN_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    FWN_HIP(hipEventElapsedTime(&ms, t0, t1));
    FWN_HIP(hipEventDestroy(t0));
    FWN_HIP(hipEventDestroy(t1));
    S.train_ms = ms;
    S.last_loss = loss;
    S.trained = true;
    S.frags_ok = false;
}

// network prediction for n points (device pointers) with the inference kernels; fallback flags
// (distance bound or prediction within band of 0.5) when fb_out is not NULL
void surrogate_predict(Surrogate& S, const float* P, int64_t n, float* pred_out, uint8_t* fb_out)
{
    hipStream_t s = S.exact->stream;
    if (!S.frags_ok) pack_frags(S, s);
    uint8_t* fb = fb_out;
    if (!fb) {
        S.fb.reserve(n);
        fb = S.fb.as<uint8_t>();
    }
    if (!S.ev[0])
        for (auto& ev : S.ev) FWN_HIP(hipEventCreate(&ev));
    FWN_HIP(hipEventRecord(S.ev[0], s));
    // large batches run in Morton order (the exact engine's point sort); results go to input order
    const uint32_t* perm = n >= kSortMin ? sort_points_single(*S.exact, P, n) : nullptr;
    FWN_HIP(hipEventRecord(S.ev[1], s));
    if (S.n_cu == 0) {
        hipDeviceProp_t prop;
        FWN_HIP(hipGetDeviceProperties(&prop, S.exact->device));
        S.n_cu = prop.multiProcessorCount;
    }
    const int64_t ntiles = (n + 63) / 64;
    const int64_t blocks = std::min<int64_t>((ntiles + IW - 1) / IW, (int64_t)S.n_cu);
    hipLaunchKernelGGL(k_sg_infer, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(IW * 64), 0, s, P, perm, n,
                       S.grid, S.table_h.as<__half>(), S.R, S.bits.as<uint32_t>(), S.frags.as<uint16_t>(),
                       S.mlp.as<float>(), S.off_b[0], S.off_b[1], S.off_b[2], S.off_b[3], S.off_w[4], S.off_b[4],
                       S.opt.band, pred_out, fb);
    FWN_LAUNCH_CHECK("k_sg_infer");
    FWN_HIP(hipEventRecord(S.ev[2], s));
    FWN_HIP(hipEventSynchronize(S.ev[2]));
    float a = 0.0f, b = 0.0f;
    FWN_HIP(hipEventElapsedTime(&a, S.ev[0], S.ev[1]));
    FWN_HIP(hipEventElapsedTime(&b, S.ev[1], S.ev[2]));
    S.encode_ms = a; // Morton sort of the points
    S.mlp_ms = b;    // fused encoding + MLP
}

} // namespace fwn

// ============================================================================================
// surrogate query: prediction, fallback compaction, exact engine on the fallback points
// ============================================================================================
namespace fwn {
namespace {

__global__ void k_flags_i32(const uint8_t* __restrict__ fb, int64_t n, int32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = fb[i];
}

__global__ void k_gather_fallback(const float* __restrict__ P, const uint8_t* __restrict__ fb,
                                  const int32_t* __restrict__ pos, int64_t n, float* __restrict__ Pf)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !fb[i]) return;
    const int64_t k = pos[i];
    Pf[3 * k] = P[3 * i];
    Pf[3 * k + 1] = P[3 * i + 1];
    Pf[3 * k + 2] = P[3 * i + 2];
}

// w = prediction, or omega / (4 pi) of the exact engine for fallback points; is_inside as the
// reference computes it for exact points (omega / (4 pi) > 0.5 in double), w > 0.5 otherwise
__global__ void k_merge(const float* __restrict__ pred, const uint8_t* __restrict__ fb, const int32_t* __restrict__ pos,
                        const float* __restrict__ omf, int64_t n, float* __restrict__ w, uint8_t* __restrict__ inside)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fb[i]) {
        const float om = omf[pos[i]];
        const double wd = (double)om / (4.0 * 3.14159265358979323846);
        w[i] = (float)wd;
        if (inside) inside[i] = wd > 0.5 ? 1 : 0;
    } else {
        const float p = pred[i];
        w[i] = p;
        if (inside) inside[i] = p > 0.5f ? 1 : 0;
    }
}

} // namespace

struct SurrogateQueryScratch {
    DBuf P, w, ins, pred, fb, cnt, pos, Pf, omf, scan_tmp;
};

void surrogate_query(Surrogate& S, SurrogateQueryScratch& Q, const float* P, int64_t n, float* w, uint8_t* inside,
                     bool dev, int64_t* n_fallback)
{
    Engine& E = *S.exact;
    hipStream_t s = E.stream;
    const float* dP = P;
    float* dW = w;
    uint8_t* dIn = inside;
    if (!dev) {
        Q.P.reserve(sizeof(float) * 3 * n);
        Q.w.reserve(sizeof(float) * n);
        FWN_HIP(hipMemcpyAsync(Q.P.p, P, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
        dP = Q.P.as<float>();
        dW = Q.w.as<float>();
        if (inside) {
            Q.ins.reserve(n);
            dIn = Q.ins.as<uint8_t>();
        }
    }
    Q.pred.reserve(sizeof(float) * n);
    Q.fb.reserve(n);
    Q.cnt.reserve(sizeof(int32_t) * n);
    Q.pos.reserve(sizeof(int32_t) * (n + 1));
    hipEvent_t q0, q1, q2;
    FWN_HIP(hipEventCreate(&q0));
    FWN_HIP(hipEventCreate(&q1));
    FWN_HIP(hipEventCreate(&q2));
    FWN_HIP(hipEventRecord(q0, s));
    surrogate_predict(S, dP, n, Q.pred.as<float>(), Q.fb.as<uint8_t>());
    FWN_HIP(hipEventRecord(q1, s));
    hipLaunchKernelGGL(k_flags_i32, dim3(grid_blocks(n, 256)), dim3(256), 0, s, Q.fb.as<uint8_t>(), n, Q.cnt.as<int32_t>());
    exclusive_scan_i32(Q.cnt.as<int32_t>(), Q.pos.as<int32_t>(), n, Q.scan_tmp, s);
    int32_t nf = 0;
    FWN_HIP(hipMemcpyAsync(&nf, Q.pos.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
    Q.Pf.reserve(sizeof(float) * 3 * std::max<int64_t>(nf, 1));
    Q.omf.reserve(sizeof(float) * std::max<int64_t>(nf, 1));
    if (nf > 0) {
        hipLaunchKernelGGL(k_gather_fallback, dim3(grid_blocks(n, 256)), dim3(256), 0, s, dP, Q.fb.as<uint8_t>(),
                           Q.pos.as<int32_t>(), n, Q.Pf.as<float>());
        FWN_LAUNCH_CHECK("k_gather_fallback");
        run_queries(E, Q.Pf.as<float>(), nullptr, nf, Q.omf.as<float>(), nullptr, nullptr, true);
    }
    hipLaunchKernelGGL(k_merge, dim3(grid_blocks(n, 256)), dim3(256), 0, s, Q.pred.as<float>(), Q.fb.as<uint8_t>(),
                       Q.pos.as<int32_t>(), Q.omf.as<float>(), n, dW, dIn);
    FWN_LAUNCH_CHECK("k_merge");
    FWN_HIP(hipEventRecord(q2, s));
    if (!dev) {
        FWN_HIP(hipMemcpyAsync(w, dW, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        if (inside) FWN_HIP(hipMemcpyAsync(inside, dIn, n, hipMemcpyDeviceToHost, s));
    }
    FWN_HIP(hipStreamSynchronize(s));
    float fa = 0.0f, fq = 0.0f;
    FWN_HIP(hipEventElapsedTime(&fa, q1, q2));
    FWN_HIP(hipEventElapsedTime(&fq, q0, q2));
    S.fallback_ms = fa;
    S.query_ms = fq;
    FWN_HIP(hipEventDestroy(q0));
    FWN_HIP(hipEventDestroy(q1));
    FWN_HIP(hipEventDestroy(q2));
    if (n_fallback) *n_fallback = nf;
}

} // namespace fwn

// ============================================================================================
// C ABI
// ============================================================================================
struct fwn_surrogate {
    fwn::Surrogate S;
    fwn::SurrogateQueryScratch Q;
};

namespace {
fwn_surrogate* SG(fwn_surrogate* h)
{
    if (!h) fwn::fail(FWN_INVALID_ARGUMENT, "surrogate is NULL");
    return h;
}
} // namespace

#pragma GCC visibility push(default)
extern "C" {

int fwn_surrogate_options_default(fwn_surrogate_options* o)
{
    return fwn::guarded("fwn_surrogate_options_default", [&] {
        if (!o) fwn::fail(FWN_INVALID_ARGUMENT, "opt is NULL");
        std::memset(o, 0, sizeof(*o));
        o->levels = 16;
        o->log2_table = 19;
        o->base_res = 16;
        o->max_res = 2048;
        o->hidden_layers = 4;
        o->batch = 65536;
        o->steps = 3000;
        o->grid_res = 1024;
        o->lr = 5e-3f;
        o->band = 0.2f;
        o->dist_rel = 2e-3f;
        o->near_frac = 0.5f;
        o->n_samples = 20000000;
        o->seed = 5;
    });
}

int fwn_surrogate_create(fwn_surrogate** out, fwn_engine* exact, const fwn_surrogate_options* opt)
{
    return fwn::guarded("fwn_surrogate_create", [&] {
        if (!out || !exact) fwn::fail(FWN_INVALID_ARGUMENT, "NULL argument");
        *out = nullptr;
        fwn::Engine* E = reinterpret_cast<fwn::Engine*>(exact);
        fwn_surrogate_options o;
        fwn_surrogate_options_default(&o);
        if (opt) o = *opt;
        std::unique_ptr<fwn_surrogate> h(new fwn_surrogate());
        std::lock_guard<std::mutex> lk(E->mu);
        fwn::DeviceGuard g(E->device);
        fwn::surrogate_init(h->S, E, o);
        *out = h.release();
    });
}

int fwn_surrogate_destroy(fwn_surrogate* h)
{
    return fwn::guarded("fwn_surrogate_destroy", [&] {
        if (!h) return;
        // only the device id is used here: the engine may already be gone
        (void)hipSetDevice(h->S.device);
        (void)hipDeviceSynchronize();
        for (auto& ev : h->S.ev)
            if (ev) (void)hipEventDestroy(ev);
        delete h;
    });
}

int fwn_surrogate_train(fwn_surrogate* h)
{
    return fwn::guarded("fwn_surrogate_train", [&] {
        fwn::Surrogate& S = SG(h)->S;
        std::lock_guard<std::mutex> lk(S.exact->mu);
        fwn::DeviceGuard g(S.exact->device);
        fwn::surrogate_train(S);
    });
}

int fwn_surrogate_info_get(const fwn_surrogate* h, fwn_surrogate_info* info)
{
    return fwn::guarded("fwn_surrogate_info_get", [&] {
        if (!h || !info) fwn::fail(FWN_INVALID_ARGUMENT, "NULL argument");
        const fwn::Surrogate& S = h->S;
        std::memset(info, 0, sizeof(*info));
        info->levels = S.grid.levels;
        for (int l = 0; l < S.grid.levels; ++l) {
            info->res[l] = S.grid.lv[l].res;
            info->dense[l] = S.grid.lv[l].dense;
        }
        info->table_floats = S.n_table;
        info->mlp_floats = S.n_mlp;
        info->delta = S.delta;
        info->last_loss = S.last_loss;
        info->train_ms = S.train_ms;
        info->encode_ms = S.encode_ms;
        info->mlp_ms = S.mlp_ms;
        info->fallback_ms = S.fallback_ms;
        info->query_ms = S.query_ms;
    });
}

int fwn_surrogate_query(fwn_surrogate* h, const float* Q, int64_t nQ, float* w_out, uint8_t* inside_out, int32_t flags,
                        int64_t* n_fallback_out)
{
    return fwn::guarded("fwn_surrogate_query", [&] {
        fwn_surrogate* sg = SG(h);
        if (nQ < 0) fwn::fail(FWN_INVALID_ARGUMENT, "negative point count");
        if (nQ == 0) {
            if (n_fallback_out) *n_fallback_out = 0;
            return;
        }
        if (!Q || !w_out) fwn::fail(FWN_INVALID_ARGUMENT, "NULL points or output");
        if (nQ > (int64_t)INT32_MAX) fwn::fail(FWN_LIMIT_EXCEEDED, "more than 2^31-1 points in one call");
        std::lock_guard<std::mutex> lk(sg->S.exact->mu);
        fwn::DeviceGuard g(sg->S.exact->device);
        fwn::surrogate_query(sg->S, sg->Q, Q, nQ, w_out, inside_out, (flags & FWN_DEVICE_POINTERS) != 0, n_fallback_out);
    });
}

int fwn_surrogate_predict(fwn_surrogate* h, const float* Q, int64_t nQ, float* pred_out, uint8_t* fallback_out,
                          int32_t flags)
{
    return fwn::guarded("fwn_surrogate_predict", [&] {
        fwn_surrogate* sg = SG(h);
        if (nQ < 0) fwn::fail(FWN_INVALID_ARGUMENT, "negative point count");
        if (nQ == 0) return;
        if (!Q || !pred_out) fwn::fail(FWN_INVALID_ARGUMENT, "NULL points or output");
        fwn::Surrogate& S = sg->S;
        std::lock_guard<std::mutex> lk(S.exact->mu);
        fwn::DeviceGuard g(S.exact->device);
        hipStream_t s = S.exact->stream;
        const bool dev = (flags & FWN_DEVICE_POINTERS) != 0;
        fwn::SurrogateQueryScratch& W = sg->Q;
        const float* dP = Q;
        float* dPred = pred_out;
        uint8_t* dFb = fallback_out;
        if (!dev) {
            W.P.reserve(sizeof(float) * 3 * nQ);
            W.pred.reserve(sizeof(float) * nQ);
            W.fb.reserve(nQ);
            FWN_HIP(hipMemcpyAsync(W.P.p, Q, sizeof(float) * 3 * nQ, hipMemcpyHostToDevice, s));
            dP = W.P.as<float>();
            dPred = W.pred.as<float>();
            dFb = fallback_out ? W.fb.as<uint8_t>() : nullptr;
        }
        fwn::surrogate_predict(S, dP, nQ, dPred, dFb);
        if (!dev) {
            FWN_HIP(hipMemcpyAsync(pred_out, dPred, sizeof(float) * nQ, hipMemcpyDeviceToHost, s));
            if (fallback_out) FWN_HIP(hipMemcpyAsync(fallback_out, dFb, nQ, hipMemcpyDeviceToHost, s));
        }
        FWN_HIP(hipStreamSynchronize(s));
    });
}

int fwn_surrogate_params(fwn_surrogate* h, float* table_out, float* mlp_out, const float* table_in, const float* mlp_in)
{
    return fwn::guarded("fwn_surrogate_params", [&] {
        fwn::Surrogate& S = SG(h)->S;
        std::lock_guard<std::mutex> lk(S.exact->mu);
        fwn::DeviceGuard g(S.exact->device);
        if (table_out) FWN_HIP(hipMemcpy(table_out, S.table.p, sizeof(float) * S.n_table, hipMemcpyDeviceToHost));
        if (mlp_out) FWN_HIP(hipMemcpy(mlp_out, S.mlp.p, sizeof(float) * S.n_mlp, hipMemcpyDeviceToHost));
        if (table_in) FWN_HIP(hipMemcpy(S.table.p, table_in, sizeof(float) * S.n_table, hipMemcpyHostToDevice));
        if (mlp_in) FWN_HIP(hipMemcpy(S.mlp.p, mlp_in, sizeof(float) * S.n_mlp, hipMemcpyHostToDevice));
        if (table_in || mlp_in) {
            fwn::surrogate_refresh(S);
        }
    });
}

int fwn_surrogate_gradients(fwn_surrogate* h, const float* Q, const float* y, int64_t n, float* loss_out,
                            float* pred_out, float* g_table_out, float* g_mlp_out)
{
    return fwn::guarded("fwn_surrogate_gradients", [&] {
        fwn::Surrogate& S = SG(h)->S;
        if (n <= 0 || !Q || !y) fwn::fail(FWN_INVALID_ARGUMENT, "need points and targets");
        std::lock_guard<std::mutex> lk(S.exact->mu);
        fwn::DeviceGuard g(S.exact->device);
        hipStream_t s = S.exact->stream;
        fwn::DBuf dP, dy;
        dP.alloc(sizeof(float) * 3 * n);
        dy.alloc(sizeof(float) * n);
        FWN_HIP(hipMemcpyAsync(dP.p, Q, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
        FWN_HIP(hipMemcpyAsync(dy.p, y, sizeof(float) * n, hipMemcpyHostToDevice, s));
        const float loss = fwn::surrogate_gradients(S, dP.as<float>(), dy.as<float>(), n);
        if (loss_out) *loss_out = loss;
        if (pred_out) FWN_HIP(hipMemcpy(pred_out, S.pred.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        if (g_table_out) FWN_HIP(hipMemcpy(g_table_out, S.g_table.p, sizeof(float) * S.n_table, hipMemcpyDeviceToHost));
        if (g_mlp_out) FWN_HIP(hipMemcpy(g_mlp_out, S.g_mlp.p, sizeof(float) * S.n_mlp, hipMemcpyDeviceToHost));
    });
}

} // extern "C"
#pragma GCC visibility pop
