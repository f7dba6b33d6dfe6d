// Top-down SAH topology for the tree build (FWN_TREE_SAH).
//
// Every binary split follows one rule, the binned box-area split of the CPU restatement of the
// reference's 4-ary tree build (the tree the reference builds comes from the dependency's
// UT_SolidAngle::init, called at modules/winding/src/FastWindingNumber.cpp:57):
//   * centroid bounds of the segment's triangles; the axis of the longest extent (x before y
//     before z on ties); a segment whose centroids coincide splits at its middle position;
//   * 16 equal bins along that axis (bin = int((c - lo) * (16 / extent)), clamped), each
//     accumulating its triangle count and the union of its triangles' boxes;
//   * a sweep over the 15 bin boundaries minimising area(left) * n_left + area(right) * n_right,
//     with area = dx*dy + dy*dz + dz*dx of float extents, in double (first minimum wins; a
//     segment with no boundary that has triangles on both sides splits at its middle);
//   * a stable partition: triangles of bins <= best first, each side in its previous order.
// Segments are split down to single triangles.  Centroid and box bounds are exact min/max, so
// they do not depend on the order in which threads combine partial results, and the arithmetic
// is the sequential rule's, operation for operation (no contraction into FMAs): the tree equals
// that of a sequential top-down build.
//
// Segments with more than kSahSmall triangles are split level by level; the triangles of a level
// are cut into chunks of kSahChunk (one workgroup per chunk, partial bounds and bins combined
// with atomics, a chunk-offset scan for the partition).  A segment of at most kSahSmall
// triangles is finished by one wave that holds its centroids, boxes and order in LDS and splits
// it recursively.
//
// Output: the binary tree layout of the LBVH path (internal nodes [0, n-1), leaf j -> n-1+j for
// the triangle at sorted position j; first/last = the sorted range an internal node covers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "fwn_internal.h"

#pragma clang fp contract(off)

namespace fwn {

namespace {

constexpr int kSahBins = 16;
constexpr int kSahSmall = 128;  // segments up to this size are finished by one wave in LDS
constexpr int kSahChunk = 2048; // triangles per workgroup in the level-synchronous phase
constexpr int kSahThreads = 256;
// per-segment accumulator (ints): centroid lo[3] hi[3] (order-preserving), bin counts[16],
// bin boxes [16][lo3 hi3] (order-preserving)
constexpr int kAccCb = 0, kAccCnt = 6, kAccBox = 6 + kSahBins, kAccWords = 6 + kSahBins + 6 * kSahBins;

__device__ __forceinline__ double sah_area(const float lo[3], const float hi[3])
{
    double dx = (double)(hi[0] - lo[0]), dy = (double)(hi[1] - lo[1]), dz = (double)(hi[2] - lo[2]);
    if (dx < 0.0) return 0.0;
    return (dx * dy + dy * dz) + dz * dx;
}

// best bin boundary (bins <= result go left) or -1; *nleft = triangles in bins <= result.
// cnt(b), blo(b, d), bhi(b, d): per-bin counts and boxes (+inf / -inf for empty bins)
template <class Cnt, class Lo, class Hi>
__device__ __forceinline__ int sah_sweep(Cnt cnt, Lo blo, Hi bhi, int64_t* nleft)
{
    double rightA[kSahBins];
    int64_t rightN[kSahBins];
    float rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
    int64_t rn = 0;
#pragma unroll
    for (int b = kSahBins - 1; b >= 1; --b) {
        for (int d = 0; d < 3; ++d) {
            const float l = blo(b, d), h = bhi(b, d);
            if (l < rl[d]) rl[d] = l;
            if (h > rh[d]) rh[d] = h;
        }
        rn += cnt(b);
        rightA[b] = sah_area(rl, rh);
        rightN[b] = rn;
    }
    rightA[0] = 0.0;
    rightN[0] = 0;
    float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
    int64_t ln = 0, bestn = 0;
    double best = INFINITY;
    int bestb = -1;
#pragma unroll
    for (int b = 0; b < kSahBins - 1; ++b) {
        for (int d = 0; d < 3; ++d) {
            const float l = blo(b, d), h = bhi(b, d);
            if (l < ll[d]) ll[d] = l;
            if (h > lh[d]) lh[d] = h;
        }
        ln += cnt(b);
        if (ln == 0 || rightN[b + 1] == 0) continue;
        double cost = sah_area(ll, lh) * (double)ln + rightA[b + 1] * (double)rightN[b + 1];
        if (cost < best) {
            best = cost;
            bestb = b;
            bestn = ln;
        }
    }
    *nleft = bestn;
    return bestb;
}

// axis of the longest centroid extent (first on ties) and that extent
__device__ __forceinline__ int sah_axis(const float clo[3], const float chi[3], float* ext)
{
    int a = 0;
    if (chi[1] - clo[1] > chi[a] - clo[a]) a = 1;
    if (chi[2] - clo[2] > chi[a] - clo[a]) a = 2;
    *ext = chi[a] - clo[a];
    return a;
}

__device__ __forceinline__ int sah_bin(float c, float clo, float sc)
{
    int b = (int)((c - clo) * sc);
    return b < 0 ? 0 : (b >= kSahBins ? kSahBins - 1 : b);
}

__device__ __forceinline__ float wmin(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wmax(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__global__ void k_iota(int32_t* __restrict__ p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// ---- level-synchronous phase -----------------------------------------------------------------
// Level arrays (int32, packed by the host): seg_lo[S], seg_hi[S], seg_fc[S] (first chunk),
// ch_seg[C], ch_lo[C], ch_hi[C].
struct Level {
    const int32_t* seg_lo;
    const int32_t* seg_hi;
    const int32_t* seg_fc;
    const int32_t* ch_seg;
    const int32_t* ch_lo;
    const int32_t* ch_hi;
};

__global__ void k_acc_init(int32_t* __restrict__ acc, int64_t nseg)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg * kAccWords) return;
    int k = (int)(i % kAccWords);
    int32_t v;
    if (k < kAccCnt) v = (k < 3) ? f2ord(INFINITY) : f2ord(-INFINITY);
    else if (k < kAccBox) v = 0;
    else v = ((k - kAccBox) % 6 < 3) ? f2ord(INFINITY) : f2ord(-INFINITY);
    acc[i] = v;
}

__global__ __launch_bounds__(kSahThreads) void k_level_bounds(Level L, const int32_t* __restrict__ items,
                                                              const float* __restrict__ cen,
                                                              int32_t* __restrict__ acc)
{
    const int c = blockIdx.x;
    const int sg = L.ch_seg[c];
    const int a = L.ch_lo[c], b = L.ch_hi[c];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = a + (int)threadIdx.x; i < b; i += kSahThreads) {
        int64_t t = items[i];
        for (int d = 0; d < 3; ++d) {
            float v = cen[3 * t + d];
            lo[d] = fminf(lo[d], v);
            hi[d] = fmaxf(hi[d], v);
        }
    }
    for (int d = 0; d < 3; ++d) {
        lo[d] = wmin(lo[d]);
        hi[d] = wmax(hi[d]);
    }
    if ((threadIdx.x & 63) == 0) {
        int32_t* g = acc + (int64_t)sg * kAccWords + kAccCb;
        for (int d = 0; d < 3; ++d) {
            atomicMin(g + d, f2ord(lo[d]));
            atomicMax(g + 3 + d, f2ord(hi[d]));
        }
    }
}

// per segment: axis, bin origin and scale from the centroid bounds (mode 0: middle split)
__global__ void k_level_axis(int64_t nseg, const int32_t* __restrict__ acc, int32_t* __restrict__ seg_ax,
                             float* __restrict__ seg_clo, float* __restrict__ seg_sc)
{
    int64_t sg = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= nseg) return;
    const int32_t* g = acc + sg * kAccWords + kAccCb;
    float clo[3], chi[3];
    for (int d = 0; d < 3; ++d) {
        clo[d] = ord2f(g[d]);
        chi[d] = ord2f(g[3 + d]);
    }
    float ext;
    int ax = sah_axis(clo, chi, &ext);
    if (ext > 0.0f) {
        seg_ax[sg] = ax;
        seg_clo[sg] = clo[ax];
        seg_sc[sg] = __fdiv_rn((float)kSahBins, ext);
    } else {
        seg_ax[sg] = -1;
        seg_clo[sg] = 0.0f;
        seg_sc[sg] = 0.0f;
    }
}

__global__ __launch_bounds__(kSahThreads) void k_level_bins(Level L, const int32_t* __restrict__ items,
                                                            const float* __restrict__ cen,
                                                            const float* __restrict__ tbox,
                                                            const int32_t* __restrict__ seg_ax,
                                                            const float* __restrict__ seg_clo,
                                                            const float* __restrict__ seg_sc,
                                                            int32_t* __restrict__ acc)
{
    __shared__ int32_t lcnt[kSahBins];
    __shared__ int32_t lbox[6 * kSahBins];
    const int c = blockIdx.x;
    const int sg = L.ch_seg[c];
    const int ax = seg_ax[sg];
    if (ax < 0) return; // uniform over the block
    const int a = L.ch_lo[c], b = L.ch_hi[c];
    const float clo = seg_clo[sg], sc = seg_sc[sg];
    const int tid = threadIdx.x;
    if (tid < kSahBins) lcnt[tid] = 0;
    if (tid < 6 * kSahBins) lbox[tid] = ((tid % 6) < 3) ? f2ord(INFINITY) : f2ord(-INFINITY);
    __syncthreads();
    for (int i = a + tid; i < b; i += kSahThreads) {
        int64_t t = items[i];
        int bin = sah_bin(cen[3 * t + ax], clo, sc);
        atomicAdd(&lcnt[bin], 1);
        for (int d = 0; d < 3; ++d) {
            atomicMin(&lbox[6 * bin + d], f2ord(tbox[6 * t + d]));
            atomicMax(&lbox[6 * bin + 3 + d], f2ord(tbox[6 * t + 3 + d]));
        }
    }
    __syncthreads();
    int32_t* g = acc + (int64_t)sg * kAccWords;
    if (tid < kSahBins && lcnt[tid] > 0) atomicAdd(g + kAccCnt + tid, lcnt[tid]);
    if (tid < 6 * kSahBins && lcnt[tid / 6] > 0) {
        if ((tid % 6) < 3) atomicMin(g + kAccBox + tid, lbox[tid]);
        else atomicMax(g + kAccBox + tid, lbox[tid]);
    }
}

// per segment: best boundary (seg_best, -1 = middle split) and split position (seg_mid)
__global__ void k_level_split(Level L, int64_t nseg, const int32_t* __restrict__ acc,
                              const int32_t* __restrict__ seg_ax, int32_t* __restrict__ seg_best,
                              int32_t* __restrict__ seg_mid)
{
    int64_t sg = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= nseg) return;
    const int32_t lo = L.seg_lo[sg], hi = L.seg_hi[sg];
    int32_t mid = lo + (hi - lo) / 2;
    int best = -1;
    if (seg_ax[sg] >= 0) {
        const int32_t* g = acc + sg * kAccWords;
        int64_t nl = 0;
        best = sah_sweep([&](int b) { return g[kAccCnt + b]; },
                         [&](int b, int d) { return ord2f(g[kAccBox + 6 * b + d]); },
                         [&](int b, int d) { return ord2f(g[kAccBox + 6 * b + 3 + d]); }, &nl);
        if (best >= 0) mid = lo + (int32_t)nl;
    }
    seg_best[sg] = best;
    seg_mid[sg] = mid;
}

__device__ __forceinline__ bool goes_left(const float* __restrict__ cen, int64_t t, int ax, float clo, float sc,
                                          int best)
{
    return sah_bin(cen[3 * t + ax], clo, sc) <= best;
}

// block-wide count of x over the block's threads (kSahThreads = 4 waves)
__device__ __forceinline__ int block_count(bool x, int32_t* red)
{
    unsigned long long m = __ballot(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = __popcll(m);
    __syncthreads();
    int tot = 0;
    for (int k = 0; k < kSahThreads / 64; ++k) tot += red[k];
    __syncthreads();
    return tot;
}

__global__ __launch_bounds__(kSahThreads) void k_level_count(Level L, const int32_t* __restrict__ items,
                                                             const float* __restrict__ cen,
                                                             const int32_t* __restrict__ seg_ax,
                                                             const float* __restrict__ seg_clo,
                                                             const float* __restrict__ seg_sc,
                                                             const int32_t* __restrict__ seg_best,
                                                             int32_t* __restrict__ ch_left)
{
    __shared__ int32_t red[kSahThreads / 64];
    const int c = blockIdx.x;
    const int sg = L.ch_seg[c];
    const int best = seg_best[sg];
    if (best < 0) {
        if (threadIdx.x == 0) ch_left[c] = 0;
        return;
    }
    const int ax = seg_ax[sg];
    const float clo = seg_clo[sg], sc = seg_sc[sg];
    const int a = L.ch_lo[c], b = L.ch_hi[c];
    int tot = 0;
    for (int base = a; base < b; base += kSahThreads) {
        int i = base + (int)threadIdx.x;
        bool f = i < b && goes_left(cen, items[i], ax, clo, sc, best);
        tot += block_count(f, red);
    }
    if (threadIdx.x == 0) ch_left[c] = tot;
}

__global__ __launch_bounds__(kSahThreads) void k_level_scatter(Level L, const int32_t* __restrict__ items,
                                                               const float* __restrict__ cen,
                                                               const int32_t* __restrict__ seg_ax,
                                                               const float* __restrict__ seg_clo,
                                                               const float* __restrict__ seg_sc,
                                                               const int32_t* __restrict__ seg_best,
                                                               const int32_t* __restrict__ seg_mid,
                                                               const int32_t* __restrict__ ch_off,
                                                               int32_t* __restrict__ out)
{
    __shared__ int32_t wl[kSahThreads / 64];
    const int c = blockIdx.x;
    const int sg = L.ch_seg[c];
    const int a = L.ch_lo[c], b = L.ch_hi[c];
    const int best = seg_best[sg];
    if (best < 0) {
        for (int i = a + (int)threadIdx.x; i < b; i += kSahThreads) out[i] = items[i];
        return;
    }
    const int ax = seg_ax[sg];
    const float clo = seg_clo[sg], sc = seg_sc[sg];
    const int32_t lo = L.seg_lo[sg], mid = seg_mid[sg];
    const int32_t relL = ch_off[c] - ch_off[L.seg_fc[sg]];
    const int32_t relR = (a - lo) - relL;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int32_t runL = 0, runR = 0;
    for (int base = a; base < b; base += kSahThreads) {
        int i = base + (int)threadIdx.x;
        bool valid = i < b;
        int32_t t = valid ? items[i] : 0;
        bool f = valid && goes_left(cen, t, ax, clo, sc, best);
        unsigned long long m = __ballot(f);
        if (lane == 0) wl[w] = __popcll(m);
        __syncthreads();
        int woff = 0, tot = 0;
        for (int k = 0; k < kSahThreads / 64; ++k) {
            if (k < w) woff += wl[k];
            tot += wl[k];
        }
        __syncthreads();
        int pre = woff + __popcll(m & lt); // left triangles before this one in the round
        int j = (int)threadIdx.x;          // position in the round (all earlier ones are valid)
        if (valid) out[f ? lo + relL + runL + pre : mid + relR + runR + (j - pre)] = t;
        int nv = std::min(kSahThreads, b - base);
        runL += tot;
        runR += nv - tot;
    }
}

__global__ __launch_bounds__(kSahThreads) void k_level_copy(Level L, const int32_t* __restrict__ in,
                                                            int32_t* __restrict__ out)
{
    const int c = blockIdx.x;
    for (int i = L.ch_lo[c] + (int)threadIdx.x; i < L.ch_hi[c]; i += kSahThreads) out[i] = in[i];
}

// rows of 5: node, left, right, first, last
__global__ void k_set_rows(const int32_t* __restrict__ rows, int64_t nrows, int32_t* __restrict__ left,
                           int32_t* __restrict__ right, int32_t* __restrict__ parent, int32_t* __restrict__ first,
                           int32_t* __restrict__ last)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int32_t* x = rows + 5 * r;
    left[x[0]] = x[1];
    right[x[0]] = x[2];
    first[x[0]] = x[3];
    last[x[0]] = x[4];
    parent[x[1]] = x[0];
    parent[x[2]] = x[0];
}

// ---- one wave per small segment ---------------------------------------------------------------
// small[4 * k]: lo, hi, node id, first id for the segment's other internal nodes
__global__ __launch_bounds__(64) void k_small(const int32_t* __restrict__ small, int64_t n,
                                              const float* __restrict__ cen, const float* __restrict__ tbox,
                                              int32_t* __restrict__ items, int32_t* __restrict__ left,
                                              int32_t* __restrict__ right, int32_t* __restrict__ parent,
                                              int32_t* __restrict__ first, int32_t* __restrict__ last)
{
    __shared__ int32_t tri[kSahSmall];
    __shared__ float c3[3][kSahSmall];
    __shared__ float bl[3][kSahSmall];
    __shared__ float bh[3][kSahSmall];
    __shared__ int32_t perm[kSahSmall];
    __shared__ int32_t perm2[kSahSmall];
    __shared__ int32_t stk[kSahSmall][3];
    __shared__ int32_t lcnt[kSahBins];
    __shared__ int32_t lbox[6 * kSahBins];

    const int lane = threadIdx.x;
    const int32_t* sm = small + 4 * (int64_t)blockIdx.x;
    const int32_t lo = sm[0], s = sm[1] - sm[0];
    int32_t next_id = sm[3];
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    for (int k = lane; k < s; k += 64) {
        int64_t t = items[lo + k];
        tri[k] = (int32_t)t;
        for (int d = 0; d < 3; ++d) {
            c3[d][k] = cen[3 * t + d];
            bl[d][k] = tbox[6 * t + d];
            bh[d][k] = tbox[6 * t + 3 + d];
        }
        perm[k] = k;
    }
    if (lane == 0) {
        stk[0][0] = 0;
        stk[0][1] = s;
        stk[0][2] = sm[2];
    }
    int sp = 1;
    __syncthreads();
    while (sp > 0) {
        --sp;
        const int a = stk[sp][0], b = stk[sp][1], nd = stk[sp][2];
        __syncthreads();
        // centroid bounds
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int p = a + lane; p < b; p += 64) {
            int k = perm[p];
            for (int d = 0; d < 3; ++d) {
                clo[d] = fminf(clo[d], c3[d][k]);
                chi[d] = fmaxf(chi[d], c3[d][k]);
            }
        }
        for (int d = 0; d < 3; ++d) {
            clo[d] = wmin(clo[d]);
            chi[d] = wmax(chi[d]);
        }
        float ext;
        const int ax = sah_axis(clo, chi, &ext);
        int mid = a + (b - a) / 2;
        if (ext > 0.0f) {
            const float sc = __fdiv_rn((float)kSahBins, ext), cl = clo[ax];
            if (lane < kSahBins) lcnt[lane] = 0;
            for (int k = lane; k < 6 * kSahBins; k += 64) lbox[k] = ((k % 6) < 3) ? f2ord(INFINITY) : f2ord(-INFINITY);
            __syncthreads();
            for (int p = a + lane; p < b; p += 64) {
                int k = perm[p];
                int bin = sah_bin(c3[ax][k], cl, sc);
                atomicAdd(&lcnt[bin], 1);
                for (int d = 0; d < 3; ++d) {
                    atomicMin(&lbox[6 * bin + d], f2ord(bl[d][k]));
                    atomicMax(&lbox[6 * bin + 3 + d], f2ord(bh[d][k]));
                }
            }
            __syncthreads();
            int64_t nl = 0;
            const int best = sah_sweep([&](int q) { return lcnt[q]; },
                                       [&](int q, int d) { return ord2f(lbox[6 * q + d]); },
                                       [&](int q, int d) { return ord2f(lbox[6 * q + 3 + d]); }, &nl);
            if (best >= 0) {
                mid = a + (int)nl;
                int runL = 0, runR = 0;
                for (int base = a; base < b; base += 64) {
                    int p = base + lane;
                    bool valid = p < b;
                    int k = valid ? perm[p] : 0;
                    bool f = valid && sah_bin(c3[ax][k], cl, sc) <= best;
                    unsigned long long m = __ballot(f);
                    int pre = __popcll(m & lt);
                    if (valid) perm2[f ? a + runL + pre : mid + runR + (lane - pre)] = k;
                    int tot = __popcll(m);
                    int nv = std::min(64, b - base);
                    runL += tot;
                    runR += nv - tot;
                }
                __syncthreads();
                for (int p = a + lane; p < b; p += 64) perm[p] = perm2[p];
            }
            __syncthreads();
        }
        const int lsz = mid - a, rsz = b - mid;
        const int32_t lid = (lsz == 1) ? (int32_t)(n - 1 + lo + a) : next_id++;
        const int32_t rid = (rsz == 1) ? (int32_t)(n - 1 + lo + mid) : next_id++;
        if (lane == 0) {
            left[nd] = lid;
            right[nd] = rid;
            first[nd] = lo + a;
            last[nd] = lo + b - 1;
            parent[lid] = nd;
            parent[rid] = nd;
            if (rsz >= 2) {
                stk[sp][0] = mid;
                stk[sp][1] = b;
                stk[sp][2] = rid;
            }
            if (lsz >= 2) {
                stk[sp + (rsz >= 2)][0] = a;
                stk[sp + (rsz >= 2)][1] = mid;
                stk[sp + (rsz >= 2)][2] = lid;
            }
        }
        sp += (rsz >= 2) + (lsz >= 2);
        __syncthreads();
    }
    for (int p = lane; p < s; p += 64) items[lo + p] = tri[perm[p]];
}

__global__ void k_set_roots(const int32_t* __restrict__ root_bin, int M, uint8_t* __restrict__ is_root)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < M && root_bin[m] >= 0) is_root[root_bin[m]] = 1;
}

int blocks_for(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return (int)std::max<int64_t>(b, 1);
}

} // namespace

void build_topology_sah(const float* cen, const float* tbox, int64_t nF, const int64_t* f_off, int M,
                        int32_t* items, int32_t* left, int32_t* right, int32_t* parent, int32_t* first,
                        int32_t* last, uint8_t* is_root, int32_t* root_bin, hipStream_t s)
{
    const int64_t n = nF;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, s, items, n);
    FWN_LAUNCH_CHECK("k_iota");

    struct Seg {
        int32_t lo, hi, node;
    };
    std::vector<int32_t> roots(M, -1);
    std::vector<Seg> large, small;
    int64_t next_id = 0;
    for (int m = 0; m < M; ++m) {
        const int64_t nm = f_off[m + 1] - f_off[m];
        if (nm == 1) roots[m] = (int32_t)(n - 1 + f_off[m]);
        if (nm < 2) continue;
        const int32_t id = (int32_t)next_id++;
        roots[m] = id;
        (nm > kSahSmall ? large : small).push_back({(int32_t)f_off[m], (int32_t)f_off[m + 1], id});
    }

    std::vector<int32_t> rows; // node, left, right, first, last of the level-synchronous phase
    std::vector<int32_t> hlev, hmid;
    DBuf dlev, dacc, dax, dclo, dsc, dbest, dmid, dchl, dcho, dtmp, scan_tmp, drows;
    if (!large.empty()) dtmp.alloc(sizeof(int32_t) * (size_t)n);
    while (!large.empty()) {
        const int64_t S = (int64_t)large.size();
        // chunks of each segment, in segment order
        std::vector<int32_t> ch_seg, ch_lo, ch_hi, seg_fc(S);
        for (int64_t g = 0; g < S; ++g) {
            seg_fc[g] = (int32_t)ch_seg.size();
            for (int64_t a = large[g].lo; a < large[g].hi; a += kSahChunk) {
                ch_seg.push_back((int32_t)g);
                ch_lo.push_back((int32_t)a);
                ch_hi.push_back((int32_t)std::min<int64_t>(a + kSahChunk, large[g].hi));
            }
        }
        const int64_t C = (int64_t)ch_seg.size();
        hlev.assign(3 * S + 3 * C, 0);
        for (int64_t g = 0; g < S; ++g) {
            hlev[g] = large[g].lo;
            hlev[S + g] = large[g].hi;
            hlev[2 * S + g] = seg_fc[g];
        }
        std::copy(ch_seg.begin(), ch_seg.end(), hlev.begin() + 3 * S);
        std::copy(ch_lo.begin(), ch_lo.end(), hlev.begin() + 3 * S + C);
        std::copy(ch_hi.begin(), ch_hi.end(), hlev.begin() + 3 * S + 2 * C);
        dlev.reserve(sizeof(int32_t) * hlev.size());
        FWN_HIP(hipMemcpyAsync(dlev.p, hlev.data(), sizeof(int32_t) * hlev.size(), hipMemcpyHostToDevice, s));
        const int32_t* lv = dlev.as<int32_t>();
        Level L{lv, lv + S, lv + 2 * S, lv + 3 * S, lv + 3 * S + C, lv + 3 * S + 2 * C};

        dacc.reserve(sizeof(int32_t) * kAccWords * (size_t)S);
        dax.reserve(sizeof(int32_t) * S);
        dclo.reserve(sizeof(float) * S);
        dsc.reserve(sizeof(float) * S);
        dbest.reserve(sizeof(int32_t) * S);
        dmid.reserve(sizeof(int32_t) * S);
        dchl.reserve(sizeof(int32_t) * C);
        dcho.reserve(sizeof(int32_t) * (C + 1));
        int32_t* acc = dacc.as<int32_t>();

        hipLaunchKernelGGL(k_acc_init, dim3(blocks_for(S * kAccWords, 256)), dim3(256), 0, s, acc, S);
        FWN_LAUNCH_CHECK("k_acc_init");
        hipLaunchKernelGGL(k_level_bounds, dim3((unsigned)C), dim3(kSahThreads), 0, s, L, items, cen, acc);
        FWN_LAUNCH_CHECK("k_level_bounds");
        hipLaunchKernelGGL(k_level_axis, dim3(blocks_for(S, 64)), dim3(64), 0, s, S, acc, dax.as<int32_t>(),
                           dclo.as<float>(), dsc.as<float>());
        FWN_LAUNCH_CHECK("k_level_axis");
        hipLaunchKernelGGL(k_level_bins, dim3((unsigned)C), dim3(kSahThreads), 0, s, L, items, cen, tbox,
                           dax.as<int32_t>(), dclo.as<float>(), dsc.as<float>(), acc);
        FWN_LAUNCH_CHECK("k_level_bins");
        hipLaunchKernelGGL(k_level_split, dim3(blocks_for(S, 64)), dim3(64), 0, s, L, S, acc, dax.as<int32_t>(),
                           dbest.as<int32_t>(), dmid.as<int32_t>());
        FWN_LAUNCH_CHECK("k_level_split");
        hipLaunchKernelGGL(k_level_count, dim3((unsigned)C), dim3(kSahThreads), 0, s, L, items, cen,
                           dax.as<int32_t>(), dclo.as<float>(), dsc.as<float>(), dbest.as<int32_t>(),
                           dchl.as<int32_t>());
        FWN_LAUNCH_CHECK("k_level_count");
        exclusive_scan_i32(dchl.as<int32_t>(), dcho.as<int32_t>(), C, scan_tmp, s);
        hipLaunchKernelGGL(k_level_scatter, dim3((unsigned)C), dim3(kSahThreads), 0, s, L, items, cen,
                           dax.as<int32_t>(), dclo.as<float>(), dsc.as<float>(), dbest.as<int32_t>(),
                           dmid.as<int32_t>(), dcho.as<int32_t>(), dtmp.as<int32_t>());
        FWN_LAUNCH_CHECK("k_level_scatter");
        hipLaunchKernelGGL(k_level_copy, dim3((unsigned)C), dim3(kSahThreads), 0, s, L, dtmp.as<int32_t>(), items);
        FWN_LAUNCH_CHECK("k_level_copy");
        hmid.resize(S);
        FWN_HIP(hipMemcpyAsync(hmid.data(), dmid.p, sizeof(int32_t) * S, hipMemcpyDeviceToHost, s));
        FWN_HIP(hipStreamSynchronize(s));

        std::vector<Seg> next;
        for (int64_t g = 0; g < S; ++g) {
            const Seg sg = large[g];
            const int32_t mid = hmid[g];
            if (mid <= sg.lo || mid >= sg.hi) fail(FWN_HIP_ERROR, "tree build: SAH split outside its segment");
            int32_t ids[2];
            const int32_t clo[2] = {sg.lo, mid}, chi[2] = {mid, sg.hi};
            for (int h = 0; h < 2; ++h) {
                const int32_t sz = chi[h] - clo[h];
                if (sz == 1) {
                    ids[h] = (int32_t)(n - 1 + clo[h]);
                } else {
                    ids[h] = (int32_t)next_id++;
                    (sz > kSahSmall ? next : small).push_back({clo[h], chi[h], ids[h]});
                }
            }
            rows.insert(rows.end(), {sg.node, ids[0], ids[1], sg.lo, sg.hi - 1});
        }
        large.swap(next);
    }
    if (!rows.empty()) {
        const int64_t nr = (int64_t)rows.size() / 5;
        drows.alloc(sizeof(int32_t) * rows.size());
        FWN_HIP(hipMemcpyAsync(drows.p, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_set_rows, dim3(blocks_for(nr, 256)), dim3(256), 0, s, drows.as<int32_t>(), nr, left,
                           right, parent, first, last);
        FWN_LAUNCH_CHECK("k_set_rows");
    }
    std::vector<int32_t> hsmall;
    hsmall.reserve(4 * small.size());
    for (const Seg& sg : small) {
        hsmall.insert(hsmall.end(), {sg.lo, sg.hi, sg.node, (int32_t)next_id});
        next_id += (sg.hi - sg.lo) - 2;
    }
    if (next_id > n - 1) fail(FWN_HIP_ERROR, "tree build: SAH node count exceeds n - 1");
    DBuf dsmall, droots;
    if (!small.empty()) {
        dsmall.alloc(sizeof(int32_t) * hsmall.size());
        FWN_HIP(hipMemcpyAsync(dsmall.p, hsmall.data(), sizeof(int32_t) * hsmall.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_small, dim3((unsigned)small.size()), dim3(64), 0, s, dsmall.as<int32_t>(), n, cen, tbox,
                           items, left, right, parent, first, last);
        FWN_LAUNCH_CHECK("k_small");
    }
    droots.alloc(sizeof(int32_t) * std::max(M, 1));
    FWN_HIP(hipMemcpyAsync(droots.p, roots.data(), sizeof(int32_t) * M, hipMemcpyHostToDevice, s));
    FWN_HIP(hipMemcpyAsync(root_bin, droots.p, sizeof(int32_t) * M, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_set_roots, dim3(blocks_for(M, 64)), dim3(64), 0, s, droots.as<int32_t>(), M, is_root);
    FWN_LAUNCH_CHECK("k_set_roots");
    // the host arrays and the small-segment table must outlive the launches that read them
    FWN_HIP(hipStreamSynchronize(s));
}

} // namespace fwn
