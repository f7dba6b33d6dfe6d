// GPU tree build for the fast winding number engine (gfx950).
//
// Replaces UT_SolidAngle<float,float>::init, called by the reference at
// modules/winding/src/FastWindingNumber.cpp:57 after the float/int copy of :39-52.  That
// dependency builds a 4-ary box-area BVH on the CPU and accumulates order-2 moments per child;
// here the same per-node quantities are produced by a Morton-code LBVH built on the GPU:
//
//   k_tri_setup     centroids, per-mesh centroid/vertex boxes (wave-reduced atomics), input checks
//   k_box_final     per-mesh Morton box, query box, FP64 near-field distance tau
//   k_tri_keys      key = mesh << 32 | 30-bit Morton code of the centroid
//   radix sort      (fwn_sort.hip)
//   k_karras        binary radix tree over the sorted keys (Karras 2012), ties broken by index
//   k_find_roots    the node covering exactly one mesh's triangles is that mesh's root
//   k_leaf_records  sorted triangle records + FP32 per-triangle moments
//   k_node_depth, k_depth_bucket, k_merge_level
//                   moments level by level from the deepest internal nodes up: each node merges
//                   its two children, FP32 for small nodes and FP64 for nodes with
//                   >= fp64_moment_min triangles (kernel boundaries order the levels)
//   k_bottom_up     the same merges by an arrival-counter walk to the roots, for trees deeper
//                   than kMaxLevels
//   k_collapse_*    level by level, each wide node takes the binary node's children and
//                   repeatedly opens the largest non-leaf slot until it has tree_width slots
//   k_write_slots   FP64 records -> the 32-word FP32 slot format of fwn.h, plus each slot's box
//                   (6 floats, rounded outward) for distance queries
//
// Several meshes are built together: their keys carry the mesh id above the Morton bits, so
// each mesh is one subtree of a single radix tree and every kernel runs once for all meshes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fwn_internal.h"

namespace fwn {

namespace {

constexpr int BT = 256; // threads per block for build kernels

__device__ __forceinline__ int mesh_of(const int64_t* off, int M, int64_t t)
{
    // largest m with off[m] <= t (off has M+1 entries, non-decreasing)
    int lo = 0, hi = M - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_init(int32_t* __restrict__ boxes, int M, int32_t* __restrict__ err)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    // 12 ints per mesh: cen lo[3] hi[3], vert lo[3] hi[3] (order-preserving ints)
    if (i < 12 * M) {
        int k = i % 12;
        bool is_lo = (k % 6) < 3;
        boxes[i] = is_lo ? INT32_MAX : INT32_MIN;
    }
    if (i == 0) *err = 0;
}

__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Per-triangle centroid (and box, for the SAH builder), per-mesh centroid and vertex bounds.
// The bounds are reduced per wave, then across the workgroup's 16 waves in LDS, so a
// single-mesh block issues 12 atomics; a wave that straddles a mesh boundary falls back to
// per-lane atomics.
constexpr int TS = 1024; // threads per k_tri_setup block
__global__ __launch_bounds__(TS) void k_tri_setup(const float* __restrict__ V, const int32_t* __restrict__ F,
                                                  int64_t nF, const int64_t* __restrict__ f_off,
                                                  const int64_t* __restrict__ v_off, int M,
                                                  float* __restrict__ cen, int32_t* __restrict__ boxes,
                                                  float* __restrict__ tbox, int32_t* __restrict__ err)
{
    __shared__ int32_t part[TS / 64][13]; // per wave: mesh (-1: nothing to add), 12 bound words
    int64_t t = (int64_t)blockIdx.x * TS + threadIdx.x;
    bool valid = t < nF;
    int m = valid ? mesh_of(f_off, M, t) : 0;
    float c[3] = {0, 0, 0}, lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (valid) {
        int64_t nv = v_off[m + 1] - v_off[m];
        float v[3][3];
        for (int k = 0; k < 3; ++k) {
            int32_t idx = F[3 * t + k];
            if (idx < 0 || (int64_t)idx >= nv) {
                atomicOr(err, 1);
                idx = 0;
                if (nv == 0) {
                    v[k][0] = v[k][1] = v[k][2] = 0.0f;
                    continue;
                }
            }
            int64_t g = v_off[m] + idx;
            v[k][0] = V[3 * g];
            v[k][1] = V[3 * g + 1];
            v[k][2] = V[3 * g + 2];
        }
        for (int d = 0; d < 3; ++d) {
            c[d] = ((v[0][d] + v[1][d]) + v[2][d]) * (1.0f / 3.0f);
            lo[d] = fminf(v[0][d], fminf(v[1][d], v[2][d]));
            hi[d] = fmaxf(v[0][d], fmaxf(v[1][d], v[2][d]));
            if (!isfinite(v[0][d]) || !isfinite(v[1][d]) || !isfinite(v[2][d])) atomicOr(err, 2);
            cen[3 * t + d] = c[d];
            if (tbox) {
                tbox[6 * t + d] = lo[d];
                tbox[6 * t + 3 + d] = hi[d];
            }
        }
    }
    float clo[3], chi[3];
    for (int d = 0; d < 3; ++d) {
        clo[d] = valid ? c[d] : INFINITY;
        chi[d] = valid ? c[d] : -INFINITY;
    }
    const int w = threadIdx.x >> 6;
    int m0 = __builtin_amdgcn_readfirstlane(m);
    const bool any = __ballot(valid) != 0ull;
    bool same = __ballot(valid && m != m0) == 0ull;
    if (same) {
        for (int d = 0; d < 3; ++d) {
            clo[d] = wave_min(clo[d]);
            chi[d] = wave_max(chi[d]);
            lo[d] = wave_min(lo[d]);
            hi[d] = wave_max(hi[d]);
        }
        if ((threadIdx.x & 63) == 0) {
            part[w][0] = any ? m0 : -1;
            for (int d = 0; d < 3; ++d) {
                part[w][1 + d] = f2ord(clo[d]);
                part[w][4 + d] = f2ord(chi[d]);
                part[w][7 + d] = f2ord(lo[d]);
                part[w][10 + d] = f2ord(hi[d]);
            }
        }
    } else {
        if ((threadIdx.x & 63) == 0) part[w][0] = -1;
    }
    __syncthreads();
    if (threadIdx.x < 12) { // word k of the bounds: min for k % 6 < 3, max otherwise
        const int k = threadIdx.x;
        const bool is_min = (k % 6) < 3;
        int cur = -1;
        int32_t acc = 0;
        for (int ww = 0; ww < TS / 64; ++ww) {
            const int mw = part[ww][0];
            if (mw < 0) continue;
            const int32_t v = part[ww][1 + k];
            if (mw != cur) {
                if (cur >= 0) {
                    if (is_min) atomicMin(boxes + 12 * cur + k, acc);
                    else atomicMax(boxes + 12 * cur + k, acc);
                }
                cur = mw;
                acc = v;
            } else {
                acc = is_min ? min(acc, v) : max(acc, v);
            }
        }
        if (cur >= 0) {
            if (is_min) atomicMin(boxes + 12 * cur + k, acc);
            else atomicMax(boxes + 12 * cur + k, acc);
        }
    }
    if (!same && valid) {
        int32_t* b = boxes + 12 * m;
        for (int d = 0; d < 3; ++d) {
            atomicMin(b + d, f2ord(clo[d]));
            atomicMax(b + 3 + d, f2ord(chi[d]));
            atomicMin(b + 6 + d, f2ord(lo[d]));
            atomicMax(b + 9 + d, f2ord(hi[d]));
        }
    }
}

// per mesh: cinfo = centroid lo[3], inv[3]; minfo = query lo[3], inv[3], tau, 0
__global__ void k_box_final(const int32_t* __restrict__ boxes, int M, float rel, float* __restrict__ cinfo,
                            float* __restrict__ minfo)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const int32_t* b = boxes + 12 * m;
    bool empty = b[0] == INT32_MAX;
    float diag2 = 0.0f;
    for (int d = 0; d < 3; ++d) {
        float clo = empty ? 0.0f : ord2f(b[d]), chi = empty ? 0.0f : ord2f(b[3 + d]);
        float vlo = empty ? 0.0f : ord2f(b[6 + d]), vhi = empty ? 0.0f : ord2f(b[9 + d]);
        float ce = chi - clo, ve = vhi - vlo;
        cinfo[6 * m + d] = clo;
        cinfo[6 * m + 3 + d] = (ce > 0.0f) ? 1.0f / ce : 0.0f;
        minfo[8 * m + d] = vlo;
        minfo[8 * m + 3 + d] = (ve > 0.0f) ? 1.0f / ve : 0.0f;
        diag2 += ve * ve;
    }
    minfo[8 * m + 6] = rel * sqrtf(diag2);
    minfo[8 * m + 7] = 0.0f;
}

__global__ __launch_bounds__(BT) void k_tri_keys(const float* __restrict__ cen, int64_t nF,
                                                 const int64_t* __restrict__ f_off, int M,
                                                 const float* __restrict__ cinfo, uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals)
{
    int64_t t = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (t >= nF) return;
    int m = mesh_of(f_off, M, t);
    const float* ci = cinfo + 6 * m;
    uint32_t mc = morton30(cen[3 * t], cen[3 * t + 1], cen[3 * t + 2], ci, ci + 3);
    keys[t] = ((uint64_t)(uint32_t)m << 32) | mc;
    vals[t] = (uint32_t)t;
}

__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int64_t n, int64_t i, int64_t j)
{
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

// unified node ids: internal i in [0, n-1), leaf j -> (n-1) + j
__global__ __launch_bounds__(BT) void k_karras(const uint64_t* __restrict__ keys, int64_t n,
                                               int32_t* __restrict__ left, int32_t* __restrict__ right,
                                               int32_t* __restrict__ parent, int32_t* __restrict__ first,
                                               int32_t* __restrict__ last)
{
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int64_t lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int64_t s = 0;
    int64_t t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    int32_t lc = (lo == gamma) ? (int32_t)(n - 1 + gamma) : (int32_t)gamma;
    int32_t rc = (hi == gamma + 1) ? (int32_t)(n - 1 + gamma + 1) : (int32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    parent[lc] = (int32_t)i;
    parent[rc] = (int32_t)i;
    first[i] = (int32_t)lo;
    last[i] = (int32_t)hi;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(BT) void k_find_roots(const int32_t* __restrict__ first, const int32_t* __restrict__ last,
                                                   int64_t n, const uint64_t* __restrict__ keys,
                                                   const int64_t* __restrict__ f_off, int M,
                                                   int32_t* __restrict__ root_bin, uint8_t* __restrict__ is_root)
{
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n - 1) {
        int m = (int)(keys[first[i]] >> 32);
        if (first[i] == f_off[m] && (int64_t)last[i] == f_off[m + 1] - 1) {
            root_bin[m] = (int32_t)i;
            is_root[i] = 1;
        }
    }
    if (i < M && f_off[i + 1] - f_off[i] == 1) {
        int64_t u = n - 1 + f_off[i];
        root_bin[i] = (int32_t)u;
        is_root[u] = 1;
    }
}

// ---- moments ------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void tri_record(const float (&v)[3][3], double* __restrict__ r)
{
    T e1[3], e2[3], c[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = (T)v[1][d] - (T)v[0][d];
        e2[d] = (T)v[2][d] - (T)v[0][d];
        c[d] = (((T)v[0][d] + (T)v[1][d]) + (T)v[2][d]) * (T)(1.0 / 3.0);
    }
    T N[3] = {(T)0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), (T)0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
              (T)0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
    T A = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    T C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 3; ++k) {
        T e[3] = {(T)v[k][0] - c[0], (T)v[k][1] - c[1], (T)v[k][2] - c[2]};
        for (int j = 0; j < 3; ++j)
            for (int jj = 0; jj < 3; ++jj) C[j][jj] += e[j] * e[jj] * (T)(1.0 / 12.0);
    }
    r[R_A] = A;
    for (int d = 0; d < 3; ++d) {
        r[R_P + d] = c[d];
        r[R_N + d] = N[d];
        r[R_LO + d] = fminf(v[0][d], fminf(v[1][d], v[2][d]));
        r[R_HI + d] = fmaxf(v[0][d], fmaxf(v[1][d], v[2][d]));
    }
    for (int k = 0; k < 9; ++k) r[R_M1 + k] = 0.0;
    const int SI[10][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {0, 0, 1}, {0, 0, 2},
                           {1, 1, 0}, {1, 1, 2}, {2, 2, 0}, {2, 2, 1}, {0, 1, 2}};
    for (int s = 0; s < 10; ++s) {
        int a = SI[s][0], b = SI[s][1], cc = SI[s][2];
        r[R_S + s] = (C[a][b] * N[cc] + C[b][cc] * N[a] + C[cc][a] * N[b]) * (T)(1.0 / 3.0);
    }
}

// parent record from two child records, each about its own centroid (parallel-axis shift)
template <typename T>
__device__ void merge2(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ o)
{
    const double* ch[2] = {a, b};
    T A = (T)a[R_A] + (T)b[R_A];
    T P[3];
    for (int d = 0; d < 3; ++d)
        P[d] = (A > (T)0) ? ((T)a[R_A] * (T)a[R_P + d] + (T)b[R_A] * (T)b[R_P + d]) / A
                          : ((T)a[R_P + d] + (T)b[R_P + d]) * (T)0.5;
    T N[3] = {0, 0, 0}, M1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, S[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int SI[10][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {0, 0, 1}, {0, 0, 2},
                           {1, 1, 0}, {1, 1, 2}, {2, 2, 0}, {2, 2, 1}, {0, 1, 2}};
    for (int c = 0; c < 2; ++c) {
        const double* x = ch[c];
        T dl[3] = {(T)x[R_P] - P[0], (T)x[R_P + 1] - P[1], (T)x[R_P + 2] - P[2]};
        T xn[3] = {(T)x[R_N], (T)x[R_N + 1], (T)x[R_N + 2]};
        T xm[9];
        for (int k = 0; k < 9; ++k) xm[k] = (T)x[R_M1 + k];
        for (int i = 0; i < 3; ++i) N[i] += xn[i];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) M1[3 * j + i] += xm[3 * j + i] + dl[j] * xn[i];
        for (int s = 0; s < 10; ++s) {
            int p = SI[s][0], q = SI[s][1], r = SI[s][2];
            // Y_jk,i = d_j M1_ki + d_k M1_ji + d_j d_k N_i ; Sym(Y)_pqr = (Y_pq,r + Y_qr,p + Y_rp,q) / 3
            T y1 = dl[p] * xm[3 * q + r] + dl[q] * xm[3 * p + r] + dl[p] * dl[q] * xn[r];
            T y2 = dl[q] * xm[3 * r + p] + dl[r] * xm[3 * q + p] + dl[q] * dl[r] * xn[p];
            T y3 = dl[r] * xm[3 * p + q] + dl[p] * xm[3 * r + q] + dl[r] * dl[p] * xn[q];
            S[s] += (T)x[R_S + s] + (y1 + y2 + y3) * (T)(1.0 / 3.0);
        }
    }
    o[R_A] = A;
    for (int d = 0; d < 3; ++d) {
        o[R_P + d] = P[d];
        o[R_N + d] = N[d];
        o[R_LO + d] = fmin(a[R_LO + d], b[R_LO + d]);
        o[R_HI + d] = fmax(a[R_HI + d], b[R_HI + d]);
    }
    for (int k = 0; k < 9; ++k) o[R_M1 + k] = M1[k];
    for (int s = 0; s < 10; ++s) o[R_S + s] = S[s];
}

__global__ __launch_bounds__(BT) void k_leaf_records(const uint32_t* __restrict__ vals, int64_t n,
                                                     const float* __restrict__ V, const int32_t* __restrict__ F,
                                                     const int64_t* __restrict__ f_off,
                                                     const int64_t* __restrict__ v_off, int M,
                                                     uint32_t* __restrict__ tris, double* __restrict__ rec)
{
    int64_t j = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (j >= n) return;
    int64_t t = vals[j];
    int m = mesh_of(f_off, M, t);
    int64_t nv = v_off[m + 1] - v_off[m];
    float v[3][3];
    for (int k = 0; k < 3; ++k) {
        int32_t idx = F[3 * t + k];
        if (idx < 0 || (int64_t)idx >= nv) idx = 0; // reported by k_tri_setup
        int64_t g = v_off[m] + idx;
        for (int d = 0; d < 3; ++d) v[k][d] = (nv > 0) ? V[3 * g + d] : 0.0f;
    }
    uint32_t* tw = tris + j * kTriWords;
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) tw[3 * k + d] = __float_as_uint(v[k][d]);
    tw[FWN_T_FACET] = (uint32_t)(t - f_off[m]);
    tw[FWN_T_NEAR2] = 0u; // the leaf data of a leaf's first triangle is written with the slots
    tw[FWN_T_PLANE] = 0u;
    tri_record<float>(v, rec + (n - 1 + j) * kRecWords);
}

__global__ __launch_bounds__(BT) void k_bottom_up(const int32_t* __restrict__ parent,
                                                  const int32_t* __restrict__ left,
                                                  const int32_t* __restrict__ right,
                                                  const int32_t* __restrict__ first, const int32_t* __restrict__ last,
                                                  const uint8_t* __restrict__ is_root, int64_t n,
                                                  int32_t fp64_min, double* rec, int32_t* visits)
{
    int64_t j = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (j >= n) return;
    int64_t u = n - 1 + j;
    while (!is_root[u]) {
        int32_t p = parent[u];
        if (p < 0) return;
        __threadfence();
        if (atomicAdd(&visits[p], 1) == 0) return; // the sibling's thread finishes this node
        __threadfence();
        const double* a = rec + (int64_t)left[p] * kRecWords;
        const double* b = rec + (int64_t)right[p] * kRecWords;
        if (last[p] - first[p] + 1 >= fp64_min) merge2<double>(a, b, rec + (int64_t)p * kRecWords);
        else merge2<float>(a, b, rec + (int64_t)p * kRecWords);
        u = p;
    }
}

// ---- level-synchronous moments ------------------------------------------------------------------
constexpr int kMaxLevels = 512; // deeper trees (pathological splits) use the arrival-counter walk

// depth of each internal node below its mesh root; -1 for ids outside every mesh tree (unused
// ids, nodes above the mesh roots) and for nodes deeper than kMaxLevels (then *deep is set).
// Per-depth counts are gathered in LDS and added to hist once per block and depth.
__global__ __launch_bounds__(BT) void k_node_depth(const int32_t* __restrict__ parent,
                                                   const uint8_t* __restrict__ is_root, int64_t n_int,
                                                   int32_t* __restrict__ depth, int32_t* __restrict__ hist,
                                                   int32_t* __restrict__ deep)
{
    __shared__ int32_t cnt[kMaxLevels];
    for (int k = threadIdx.x; k < kMaxLevels; k += BT) cnt[k] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n_int) {
        int d = 0;
        int32_t u = (int32_t)i;
        for (;;) { // at most kMaxLevels steps
            if (is_root[u]) break;
            const int32_t p = parent[u];
            if (p < 0) {
                d = -1;
                break;
            }
            if (++d >= kMaxLevels) {
                atomicOr(deep, 1);
                d = -1;
                break;
            }
            u = p;
        }
        depth[i] = d;
        if (d >= 0) atomicAdd(&cnt[d], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxLevels; k += BT)
        if (cnt[k] > 0) atomicAdd(&hist[k], cnt[k]);
}

// node ids grouped by depth: level d occupies order[off[d], off[d] + hist[d]); each block
// reserves its range per depth with one atomic on cursor[d] (order within a level is free)
__global__ __launch_bounds__(BT) void k_depth_bucket(const int32_t* __restrict__ depth, int64_t n_int,
                                                     const int32_t* __restrict__ off, int32_t* __restrict__ cursor,
                                                     int32_t* __restrict__ order)
{
    __shared__ int32_t cnt[kMaxLevels];
    __shared__ int32_t base[kMaxLevels];
    for (int k = threadIdx.x; k < kMaxLevels; k += BT) cnt[k] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    const int d = (i < n_int) ? depth[i] : -1;
    int r = 0;
    if (d >= 0) r = atomicAdd(&cnt[d], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxLevels; k += BT)
        if (cnt[k] > 0) base[k] = off[k] + atomicAdd(&cursor[k], cnt[k]);
    __syncthreads();
    if (d >= 0) order[base[d] + r] = (int32_t)i;
}

__global__ __launch_bounds__(BT) void k_merge_level(const int32_t* __restrict__ nodes, int32_t cnt,
                                                    const int32_t* __restrict__ left,
                                                    const int32_t* __restrict__ right,
                                                    const int32_t* __restrict__ first,
                                                    const int32_t* __restrict__ last, int32_t fp64_min,
                                                    double* rec)
{
    const int i = blockIdx.x * BT + threadIdx.x;
    if (i >= cnt) return;
    const int32_t p = nodes[i];
    const double* a = rec + (int64_t)left[p] * kRecWords;
    const double* b = rec + (int64_t)right[p] * kRecWords;
    if (last[p] - first[p] + 1 >= fp64_min) merge2<double>(a, b, rec + (int64_t)p * kRecWords);
    else merge2<float>(a, b, rec + (int64_t)p * kRecWords);
}

// ---- collapse into wide nodes -----------------------------------------------------------------
__device__ __forceinline__ int32_t node_count(int32_t u, int64_t n, const int32_t* first, const int32_t* last)
{
    return (u >= n - 1) ? 1 : last[u] - first[u] + 1;
}

__global__ __launch_bounds__(BT) void k_collapse_level(const int32_t* __restrict__ level_bin, int32_t cnt,
                                                       int32_t gbase, const int32_t* __restrict__ left,
                                                       const int32_t* __restrict__ right,
                                                       const int32_t* __restrict__ first,
                                                       const int32_t* __restrict__ last, int64_t n, int leaf_max,
                                                       int width, int rounds, int32_t* __restrict__ wide_slot_bin,
                                                       int32_t* __restrict__ wide_ns, int32_t* __restrict__ newcount)
{
    int i = blockIdx.x * BT + threadIdx.x;
    if (i >= cnt) return;
    int32_t b = level_bin[i];
    int32_t sl[kWidth], sc[kWidth];
    int ns;
    if (b >= n - 1 || node_count(b, n, first, last) <= leaf_max) {
        sl[0] = b;
        sc[0] = node_count(b, n, first, last);
        ns = 1;
    } else {
        sl[0] = left[b];
        sl[1] = right[b];
        sc[0] = node_count(sl[0], n, first, last);
        sc[1] = node_count(sl[1], n, first, last);
        ns = 2;
        // rounds > 0: take whole binary levels (every non-leaf slot opened, rounds - 1 times)
        for (int r = 1; r < rounds; ++r) {
            int32_t nl[kWidth], nc[kWidth];
            int nn = 0;
            for (int s = 0; s < ns; ++s) {
                int32_t u = sl[s];
                if (u < n - 1 && sc[s] > leaf_max) {
                    nl[nn] = left[u];
                    nc[nn] = node_count(nl[nn], n, first, last);
                    ++nn;
                    nl[nn] = right[u];
                    nc[nn] = node_count(nl[nn], n, first, last);
                    ++nn;
                } else {
                    nl[nn] = u;
                    nc[nn] = sc[s];
                    ++nn;
                }
            }
            for (int s = 0; s < nn; ++s) {
                sl[s] = nl[s];
                sc[s] = nc[s];
            }
            ns = nn;
        }
        // rounds == 0: repeatedly open the largest non-leaf slot until the node has width slots
        while (rounds == 0 && ns < width) {
            int best = -1, bc = leaf_max;
            for (int s = 0; s < ns; ++s)
                if (sc[s] > bc) {
                    bc = sc[s];
                    best = s;
                }
            if (best < 0) break;
            int32_t u = sl[best];
            for (int s = ns; s > best + 1; --s) {
                sl[s] = sl[s - 1];
                sc[s] = sc[s - 1];
            }
            sl[best] = left[u];
            sl[best + 1] = right[u];
            sc[best] = node_count(sl[best], n, first, last);
            sc[best + 1] = node_count(sl[best + 1], n, first, last);
            ++ns;
        }
    }
    int g = gbase + i;
    int nc = 0;
    for (int s = 0; s < kWidth; ++s) {
        wide_slot_bin[(int64_t)g * kWidth + s] = (s < ns) ? sl[s] : -1;
        if (s < ns && sc[s] > leaf_max) ++nc;
    }
    wide_ns[g] = ns;
    newcount[i] = nc;
}

__global__ __launch_bounds__(BT) void k_collapse_assign(int32_t cnt, int32_t gbase, int32_t next_base,
                                                        const int32_t* __restrict__ wide_slot_bin,
                                                        const int32_t* __restrict__ first,
                                                        const int32_t* __restrict__ last, int64_t n, int leaf_max,
                                                        const int32_t* __restrict__ offs,
                                                        int32_t* __restrict__ child_wide,
                                                        int32_t* __restrict__ level_next)
{
    int i = blockIdx.x * BT + threadIdx.x;
    if (i >= cnt) return;
    int g = gbase + i;
    int k = offs[i];
    for (int s = 0; s < kWidth; ++s) {
        int32_t u = wide_slot_bin[(int64_t)g * kWidth + s];
        int32_t cw = -1;
        if (u >= 0 && node_count(u, n, first, last) > leaf_max) {
            cw = next_base + k;
            level_next[k] = u;
            ++k;
        }
        child_wide[(int64_t)g * kWidth + s] = cw;
    }
}

__global__ __launch_bounds__(BT) void k_write_slots(int64_t W, int width, const int32_t* __restrict__ wide_slot_bin,
                                                    const int32_t* __restrict__ wide_ns,
                                                    const int32_t* __restrict__ child_wide,
                                                    const double* __restrict__ rec,
                                                    const int32_t* __restrict__ first,
                                                    const int32_t* __restrict__ last, int64_t n, float beta2,
                                                    const int64_t* __restrict__ f_off, int M,
                                                    const float* __restrict__ minfo, uint32_t* __restrict__ slots,
                                                    float* __restrict__ boxes, uint32_t* __restrict__ tris)
{
    int64_t idx = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (idx >= W * width) return;
    int64_t g = idx / width;
    int s = (int)(idx % width);
    uint32_t* w = slots + idx * kSlotWords;
    float* bx = boxes + idx * 6;
    int64_t src = g * kWidth + s; // collapse scratch keeps kWidth entries per node
    int32_t u = wide_slot_bin[src];
    if (s >= wide_ns[g] || u < 0) {
        for (int k = 0; k < FWN_W_LINK; ++k) w[k] = 0u;
        w[FWN_W_LINK] = FWN_LINK_EMPTY;
        for (int k = 0; k < 6; ++k) bx[k] = 0.0f;
        return;
    }
    const double* r = rec + (int64_t)u * kRecWords;
    // the slot's box for distance queries, rounded outward so it still contains the triangles
    for (int d = 0; d < 3; ++d) {
        bx[d] = __double2float_rd(r[R_LO + d]);
        bx[3 + d] = __double2float_ru(r[R_HI + d]);
    }
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        double a = fabs(r[R_P + d] - r[R_LO + d]), b = fabs(r[R_HI + d] - r[R_P + d]);
        double e = a > b ? a : b;
        r2 += e * e;
    }
    const float r2f = nextafterf((float)r2, INFINITY); // rounded up: still a bound
    float f[FWN_W_LINK];
    f[FWN_W_PX] = (float)r[R_P];
    f[FWN_W_PY] = (float)r[R_P + 1];
    f[FWN_W_PZ] = (float)r[R_P + 2];
    f[FWN_W_THR] = beta2 * r2f; // the query's far-field threshold
    f[FWN_W_NX] = (float)r[R_N];
    f[FWN_W_NY] = (float)r[R_N + 1];
    f[FWN_W_NZ] = (float)r[R_N + 2];
    // tr and C are folded into the homogeneous quadratic and cubic: tr = tr |d|^2 / |d|^2 adds
    // tr to the diagonal of A, (C.d) = (C.d) |d|^2 / |d|^2 adds C_i to the cubic monomials
    // d_i d_j d_j; the far field then needs no constant terms
    const double* M1 = r + R_M1;
    const double tr = M1[0] + M1[4] + M1[8];
    float* A = f + FWN_W_A;
    A[0] = (float)(-3.0 * M1[0] + tr);
    A[1] = (float)(-3.0 * M1[4] + tr);
    A[2] = (float)(-3.0 * M1[8] + tr);
    A[3] = (float)(-3.0 * (M1[1] + M1[3]));
    A[4] = (float)(-3.0 * (M1[5] + M1[7]));
    A[5] = (float)(-3.0 * (M1[6] + M1[2]));
    const double* S = r + R_S;
    const double sc[10] = {7.5, 7.5, 7.5, 22.5, 22.5, 22.5, 22.5, 22.5, 22.5, 45.0};
    const double cx = -4.5 * (S[0] + S[5] + S[7]);
    const double cy = -4.5 * (S[3] + S[1] + S[8]);
    const double cz = -4.5 * (S[4] + S[6] + S[2]);
    // monomials: xxx yyy zzz xxy xxz yyx yyz zzx zzy xyz
    const double fold[10] = {cx, cy, cz, cy, cz, cx, cz, cx, cy, 0.0};
    for (int k = 0; k < 10; ++k) f[FWN_W_B + k] = (float)(sc[k] * S[k] + fold[k]);
    for (int k = 0; k < FWN_W_LINK; ++k) w[k] = __float_as_uint(f[k]);
    // a leaf slot is any slot without a wide child: a single triangle or a bucket of at most
    // leaf_max triangles (an internal binary node); both cover a contiguous sorted range
    const int32_t cw = child_wide[src];
    if (cw >= 0) {
        w[FWN_W_LINK] = ((uint32_t)cw << 4) | (uint32_t)wide_ns[cw];
        return;
    }
    const bool bin_leaf = u >= n - 1;
    const int64_t tf = bin_leaf ? (u - (n - 1)) : first[u];
    const int32_t tc = bin_leaf ? 1 : last[u] - first[u] + 1;
    w[FWN_W_LINK] = FWN_LINK_LEAF | ((uint32_t)tf << 3) | (uint32_t)(tc - 1);
    // the leaf's FP64 near-field rule, in its first triangle: q is within tau of the leaf's
    // bounding sphere when |d|^2 < (sqrt(r2) + tau)^2 and, for a single triangle, within tau of
    // its plane when (d.N)^2 < tau^2 |N|^2
    const float tau = minfo[8 * mesh_of(f_off, M, tf) + 6];
    const float rr = sqrtf(r2f) + tau;
    const float nn = fmaf(f[FWN_W_NX], f[FWN_W_NX], fmaf(f[FWN_W_NY], f[FWN_W_NY], f[FWN_W_NZ] * f[FWN_W_NZ]));
    uint32_t* tw = tris + tf * kTriWords;
    tw[FWN_T_NEAR2] = __float_as_uint(rr * rr);
    tw[FWN_T_PLANE] = __float_as_uint((tau * tau) * nn);
}

} // namespace

int grid_blocks(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > INT32_MAX) fail(FWN_LIMIT_EXCEEDED, "grid too large");
    return (int)b;
}

void build_engine(Engine& e, int32_t M, const float* V_host, const int64_t* v_off, const int32_t* F_host,
                  const int64_t* f_off)
{
    hipStream_t s = e.stream;
    const int64_t nV = v_off[M];
    const int64_t nF = f_off[M];
    // a leaf slot's link holds its first triangle in 28 bits
    if (nF >= ((int64_t)1 << 28) - 1) fail(FWN_LIMIT_EXCEEDED, "2^28 - 1 or more triangles");
    e.n_meshes = M;
    e.n_tris = nF;
    e.h_f_off.assign(f_off, f_off + M + 1);
    const int leaf_max = e.opt.leaf_max;
    const int width = e.opt.tree_width;
    e.width = width;

    hipEvent_t b0, b1;
    FWN_HIP(hipEventCreate(&b0));
    FWN_HIP(hipEventCreate(&b1));
    FWN_HIP(hipEventRecord(b0, s));

    DBuf dV, dF, dvo, dfo, boxes, err, cinfo, cen, keys, vals, keys2, vals2, hist, rowsum;
    dV.alloc(sizeof(float) * 3 * (size_t)std::max<int64_t>(nV, 1));
    dF.alloc(sizeof(int32_t) * 3 * (size_t)std::max<int64_t>(nF, 1));
    dvo.alloc(sizeof(int64_t) * (M + 1));
    dfo.alloc(sizeof(int64_t) * (M + 1));
    if (nV) FWN_HIP(hipMemcpyAsync(dV.p, V_host, sizeof(float) * 3 * nV, hipMemcpyHostToDevice, s));
    if (nF) FWN_HIP(hipMemcpyAsync(dF.p, F_host, sizeof(int32_t) * 3 * nF, hipMemcpyHostToDevice, s));
    FWN_HIP(hipMemcpyAsync(dvo.p, v_off, sizeof(int64_t) * (M + 1), hipMemcpyHostToDevice, s));
    FWN_HIP(hipMemcpyAsync(dfo.p, f_off, sizeof(int64_t) * (M + 1), hipMemcpyHostToDevice, s));

    boxes.alloc(sizeof(int32_t) * 12 * M);
    err.alloc(sizeof(int32_t));
    cinfo.alloc(sizeof(float) * 6 * M);
    e.mesh_info.alloc(sizeof(float) * 8 * M);
    hipLaunchKernelGGL(k_init, dim3(grid_blocks(12 * M, 256)), dim3(256), 0, s, boxes.as<int32_t>(), M,
                       err.as<int32_t>());
    FWN_LAUNCH_CHECK("k_init");

    const bool sah = e.opt.tree_builder == FWN_TREE_SAH;
    DBuf tbox;
    cen.alloc(sizeof(float) * 3 * (size_t)std::max<int64_t>(nF, 1));
    if (sah) tbox.alloc(sizeof(float) * 6 * (size_t)std::max<int64_t>(nF, 1));
    if (nF) {
        hipLaunchKernelGGL(k_tri_setup, dim3(grid_blocks(nF, TS)), dim3(TS), 0, s, dV.as<float>(), dF.as<int32_t>(),
                           nF, dfo.as<int64_t>(), dvo.as<int64_t>(), M, cen.as<float>(), boxes.as<int32_t>(),
                           sah ? tbox.as<float>() : nullptr, err.as<int32_t>());
        FWN_LAUNCH_CHECK("k_tri_setup");
    }
    hipLaunchKernelGGL(k_box_final, dim3(grid_blocks(M, 64)), dim3(64), 0, s, boxes.as<int32_t>(), M,
                       e.opt.near_fp64_rel, cinfo.as<float>(), e.mesh_info.as<float>());
    FWN_LAUNCH_CHECK("k_box_final");
    int32_t h_err = 0;
    FWN_HIP(hipMemcpyAsync(&h_err, err.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
    if (h_err & 1) fail(FWN_INVALID_ARGUMENT, "facet index out of range of its mesh's vertices");
    if (h_err & 2) fail(FWN_INVALID_ARGUMENT, "non-finite vertex position");

    e.h_tau.assign(M, 0.0f);
    {
        std::vector<float> mi(8 * (size_t)M);
        FWN_HIP(hipMemcpyAsync(mi.data(), e.mesh_info.p, sizeof(float) * 8 * M, hipMemcpyDeviceToHost, s));
        FWN_HIP(hipStreamSynchronize(s));
        for (int m = 0; m < M; ++m) e.h_tau[m] = mi[8 * m + 6];
    }

    e.tris.alloc(sizeof(uint32_t) * kTriWords * (size_t)std::max<int64_t>(nF, 1));
    e.roots.alloc(sizeof(int32_t) * 2 * M);
    e.h_roots.assign(2 * (size_t)M, -1);
    for (int m = 0; m < M; ++m) e.h_roots[2 * m + 1] = 0;

    if (nF == 0) {
        e.slots.alloc(sizeof(uint32_t) * width * kSlotWords);
        e.boxes.alloc(sizeof(float) * width * 6);
        e.n_nodes = 0;
        e.max_depth = 0;
        FWN_HIP(hipMemcpyAsync(e.roots.p, e.h_roots.data(), sizeof(int32_t) * 2 * M, hipMemcpyHostToDevice, s));
        FWN_HIP(hipStreamSynchronize(s));
        FWN_HIP(hipEventDestroy(b0));
        FWN_HIP(hipEventDestroy(b1));
        return;
    }

    // ---- binary tree topology: Morton-order radix tree, or top-down SAH splits ----
    const int64_t n = nF;
    const int64_t nnodes = 2 * n - 1;
    DBuf left, right, parent, first, last, isroot, rootbin, rec, visits;
    left.alloc(sizeof(int32_t) * std::max<int64_t>(n - 1, 1));
    right.alloc(sizeof(int32_t) * std::max<int64_t>(n - 1, 1));
    first.alloc(sizeof(int32_t) * std::max<int64_t>(n - 1, 1));
    last.alloc(sizeof(int32_t) * std::max<int64_t>(n - 1, 1));
    parent.alloc(sizeof(int32_t) * nnodes);
    isroot.alloc(nnodes);
    rootbin.alloc(sizeof(int32_t) * M);
    FWN_HIP(hipMemsetAsync(isroot.p, 0, nnodes, s));
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_blocks(nnodes, 256)), dim3(256), 0, s, parent.as<int32_t>(), nnodes, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_blocks(M, 256)), dim3(256), 0, s, rootbin.as<int32_t>(), (int64_t)M, -1);
    FWN_LAUNCH_CHECK("k_fill_i32");
    vals.alloc(sizeof(uint32_t) * nF);
    uint32_t* vp = vals.as<uint32_t>();
    if (sah) {
        build_topology_sah(cen.as<float>(), tbox.as<float>(), nF, f_off, M, reinterpret_cast<int32_t*>(vp),
                           left.as<int32_t>(), right.as<int32_t>(), parent.as<int32_t>(), first.as<int32_t>(),
                           last.as<int32_t>(), isroot.as<uint8_t>(), rootbin.as<int32_t>(), s);
        tbox.release();
    } else {
        keys.alloc(sizeof(uint64_t) * nF);
        keys2.alloc(sizeof(uint64_t) * nF);
        vals2.alloc(sizeof(uint32_t) * nF);
        hipLaunchKernelGGL(k_tri_keys, dim3(grid_blocks(nF, BT)), dim3(BT), 0, s, cen.as<float>(), nF,
                           dfo.as<int64_t>(), M, cinfo.as<float>(), keys.as<uint64_t>(), vp);
        FWN_LAUNCH_CHECK("k_tri_keys");
        int mesh_bits = 0;
        while ((1 << mesh_bits) < M) ++mesh_bits;
        int nbits = (M > 1) ? 32 + mesh_bits : 30;
        uint64_t* kp = keys.as<uint64_t>();
        uint64_t* kp2 = keys2.as<uint64_t>();
        uint32_t* vp2 = vals2.as<uint32_t>();
        radix_sort_pairs(kp, vp, kp2, vp2, nF, nbits, hist, rowsum, s);
        if (n > 1) {
            hipLaunchKernelGGL(k_karras, dim3(grid_blocks(n - 1, BT)), dim3(BT), 0, s, kp, n, left.as<int32_t>(),
                               right.as<int32_t>(), parent.as<int32_t>(), first.as<int32_t>(), last.as<int32_t>());
            FWN_LAUNCH_CHECK("k_karras");
        }
        hipLaunchKernelGGL(k_find_roots, dim3(grid_blocks(std::max<int64_t>(n - 1, M), BT)), dim3(BT), 0, s,
                           first.as<int32_t>(), last.as<int32_t>(), n, kp, dfo.as<int64_t>(), M,
                           rootbin.as<int32_t>(), isroot.as<uint8_t>());
        FWN_LAUNCH_CHECK("k_find_roots");
    }

    // ---- moments ----
    rec.alloc(sizeof(double) * kRecWords * nnodes);
    hipLaunchKernelGGL(k_leaf_records, dim3(grid_blocks(n, BT)), dim3(BT), 0, s, vp, n, dV.as<float>(),
                       dF.as<int32_t>(), dfo.as<int64_t>(), dvo.as<int64_t>(), M, e.tris.as<uint32_t>(),
                       rec.as<double>());
    FWN_LAUNCH_CHECK("k_leaf_records");
    if (n > 1) {
        const int64_t n_int = n - 1;
        DBuf depth, lvl_hist, lvl_off, order;
        depth.alloc(sizeof(int32_t) * n_int);
        lvl_hist.alloc(sizeof(int32_t) * (2 * kMaxLevels + 1)); // hist[kMaxLevels], cursor[kMaxLevels], deep
        lvl_off.alloc(sizeof(int32_t) * kMaxLevels);
        order.alloc(sizeof(int32_t) * n_int);
        int32_t* hist_d = lvl_hist.as<int32_t>();
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_blocks(2 * kMaxLevels + 1, 256)), dim3(256), 0, s, hist_d,
                           (int64_t)(2 * kMaxLevels + 1), 0);
        hipLaunchKernelGGL(k_node_depth, dim3(grid_blocks(n_int, BT)), dim3(BT), 0, s, parent.as<int32_t>(),
                           isroot.as<uint8_t>(), n_int, depth.as<int32_t>(), hist_d, hist_d + 2 * kMaxLevels);
        FWN_LAUNCH_CHECK("k_node_depth");
        std::vector<int32_t> hh(kMaxLevels + 1);
        FWN_HIP(hipMemcpyAsync(hh.data(), hist_d, sizeof(int32_t) * kMaxLevels, hipMemcpyDeviceToHost, s));
        FWN_HIP(hipMemcpyAsync(&hh[kMaxLevels], hist_d + 2 * kMaxLevels, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        FWN_HIP(hipStreamSynchronize(s));
        if (hh[kMaxLevels] == 0 && e.opt.moment_pass == FWN_MOMENTS_LEVELS) {
            std::vector<int32_t> off(kMaxLevels, 0);
            int levels = 0;
            for (int d = 0, acc = 0; d < kMaxLevels; ++d) {
                off[d] = acc;
                acc += hh[d];
                if (hh[d] > 0) levels = d + 1;
            }
            FWN_HIP(hipMemcpyAsync(lvl_off.p, off.data(), sizeof(int32_t) * kMaxLevels, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_depth_bucket, dim3(grid_blocks(n_int, BT)), dim3(BT), 0, s, depth.as<int32_t>(),
                               n_int, lvl_off.as<int32_t>(), hist_d + kMaxLevels, order.as<int32_t>());
            FWN_LAUNCH_CHECK("k_depth_bucket");
            for (int d = levels - 1; d >= 0; --d) {
                if (hh[d] == 0) continue;
                hipLaunchKernelGGL(k_merge_level, dim3(grid_blocks(hh[d], BT)), dim3(BT), 0, s,
                                   order.as<int32_t>() + off[d], hh[d], left.as<int32_t>(), right.as<int32_t>(),
                                   first.as<int32_t>(), last.as<int32_t>(), e.opt.fp64_moment_min, rec.as<double>());
                FWN_LAUNCH_CHECK("k_merge_level");
            }
            // the host vectors above are read by copies already waited for; off by the bucket kernel
            FWN_HIP(hipStreamSynchronize(s));
        } else {
            visits.alloc(sizeof(int32_t) * n_int);
            hipLaunchKernelGGL(k_fill_i32, dim3(grid_blocks(n_int, 256)), dim3(256), 0, s, visits.as<int32_t>(),
                               n_int, 0);
            hipLaunchKernelGGL(k_bottom_up, dim3(grid_blocks(n, BT)), dim3(BT), 0, s, parent.as<int32_t>(),
                               left.as<int32_t>(), right.as<int32_t>(), first.as<int32_t>(), last.as<int32_t>(),
                               isroot.as<uint8_t>(), n, e.opt.fp64_moment_min, rec.as<double>(),
                               visits.as<int32_t>());
            FWN_LAUNCH_CHECK("k_bottom_up");
        }
    }

    // ---- collapse ----
    std::vector<int32_t> rb(M);
    FWN_HIP(hipMemcpyAsync(rb.data(), rootbin.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> level0;
    std::vector<int32_t> mesh_wide(M, -1);
    for (int m = 0; m < M; ++m) {
        if (f_off[m + 1] > f_off[m]) {
            if (rb[m] < 0) fail(FWN_HIP_ERROR, "tree build: mesh root not found");
            mesh_wide[m] = (int32_t)level0.size();
            level0.push_back(rb[m]);
        }
    }
    DBuf wsb, wns, cw, lvl[2], newc, offs, scan_tmp;
    // every wide node is a distinct binary node with more than leaf_max triangles, or a mesh root
    int64_t wcap = n + M;
    wsb.alloc(sizeof(int32_t) * kWidth * wcap);
    cw.alloc(sizeof(int32_t) * kWidth * wcap);
    wns.alloc(sizeof(int32_t) * wcap);
    lvl[0].alloc(sizeof(int32_t) * wcap);
    lvl[1].alloc(sizeof(int32_t) * wcap);
    newc.alloc(sizeof(int32_t) * wcap);
    offs.alloc(sizeof(int32_t) * (wcap + 1));
    FWN_HIP(hipMemcpyAsync(lvl[0].p, level0.data(), sizeof(int32_t) * level0.size(), hipMemcpyHostToDevice, s));
    int rounds = 0; // binary levels per wide node for the SAH layout: log2(width)
    while ((1 << rounds) < width) ++rounds;
    int32_t cnt = (int32_t)level0.size(), gbase = 0, depth = 0;
    int cur = 0;
    while (cnt > 0) {
        hipLaunchKernelGGL(k_collapse_level, dim3(grid_blocks(cnt, BT)), dim3(BT), 0, s, lvl[cur].as<int32_t>(), cnt,
                           gbase, left.as<int32_t>(), right.as<int32_t>(), first.as<int32_t>(), last.as<int32_t>(),
                           n, leaf_max, width, sah ? rounds : 0, wsb.as<int32_t>(), wns.as<int32_t>(),
                           newc.as<int32_t>());
        FWN_LAUNCH_CHECK("k_collapse_level");
        exclusive_scan_i32(newc.as<int32_t>(), offs.as<int32_t>(), cnt, scan_tmp, s);
        int32_t next_base = gbase + cnt;
        hipLaunchKernelGGL(k_collapse_assign, dim3(grid_blocks(cnt, BT)), dim3(BT), 0, s, cnt, gbase, next_base,
                           wsb.as<int32_t>(), first.as<int32_t>(), last.as<int32_t>(), n, leaf_max,
                           offs.as<int32_t>(), cw.as<int32_t>(), lvl[cur ^ 1].as<int32_t>());
        FWN_LAUNCH_CHECK("k_collapse_assign");
        int32_t total = 0;
        FWN_HIP(hipMemcpyAsync(&total, offs.as<int32_t>() + cnt, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        FWN_HIP(hipStreamSynchronize(s));
        ++depth;
        gbase = next_base;
        cnt = total;
        cur ^= 1;
        if (gbase + (int64_t)cnt > wcap) fail(FWN_HIP_ERROR, "tree build: wide node bound exceeded");
    }
    const int64_t W = gbase;
    e.n_nodes = W;
    e.max_depth = depth;
    e.slots.alloc(sizeof(uint32_t) * width * kSlotWords * (size_t)std::max<int64_t>(W, 1));
    e.boxes.alloc(sizeof(float) * width * 6 * (size_t)std::max<int64_t>(W, 1));
    if (W >= ((int64_t)1 << 27)) fail(FWN_LIMIT_EXCEEDED, "more than 2^27 tree nodes");
    hipLaunchKernelGGL(k_write_slots, dim3(grid_blocks(W * width, BT)), dim3(BT), 0, s, W, width, wsb.as<int32_t>(),
                       wns.as<int32_t>(), cw.as<int32_t>(), rec.as<double>(), first.as<int32_t>(), last.as<int32_t>(),
                       n, e.opt.accuracy_scale * e.opt.accuracy_scale, dfo.as<int64_t>(), M,
                       e.mesh_info.as<float>(), e.slots.as<uint32_t>(), e.boxes.as<float>(), e.tris.as<uint32_t>());
    FWN_LAUNCH_CHECK("k_write_slots");
    std::vector<int32_t> ns_root(level0.size());
    if (!level0.empty())
        FWN_HIP(hipMemcpyAsync(ns_root.data(), wns.p, sizeof(int32_t) * level0.size(), hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
    for (int m = 0; m < M; ++m) {
        if (mesh_wide[m] >= 0) {
            e.h_roots[2 * m] = mesh_wide[m];
            e.h_roots[2 * m + 1] = ns_root[mesh_wide[m]];
        }
    }
    FWN_HIP(hipMemcpyAsync(e.roots.p, e.h_roots.data(), sizeof(int32_t) * 2 * M, hipMemcpyHostToDevice, s));
    FWN_HIP(hipEventRecord(b1, s));
    FWN_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    FWN_HIP(hipEventElapsedTime(&ms, b0, b1));
    e.build_ms = ms;
    FWN_HIP(hipEventDestroy(b0));
    FWN_HIP(hipEventDestroy(b1));
}

} // namespace fwn
