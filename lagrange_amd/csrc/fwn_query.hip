// Winding number queries on gfx950.
//
// Replaces UT_SolidAngle<float,float>::computeSolidAngle, called per point by the reference at
// modules/winding/src/FastWindingNumber.cpp:66 (is_inside) and :75 (solid_angle), and the
// per-voxel loop of modules/volume/src/mesh_to_volume.cpp:167-183 that issues those calls.
//
// Points are sorted by mesh and 30-bit Morton code, then each wave64 takes 64 consecutive
// points (one per lane) and walks its mesh's wide tree once for all of them:
//  * the wave pops a (node, lane mask) pair from a wave-private stack held in three VGPRs
//    (entry i lives in lane i%64 of bank i/64; pushes are v_cndmask, pops v_readlane), so
//    node ids and masks stay scalar and slot/triangle loads are scalar (s_load) broadcasts;
//  * for every slot each active lane tests |q - p|^2 > beta^2 r2: far lanes add the order-2
//    expansion, near lanes either take the exact triangle solid angles (leaf slot) or stay
//    active in the child pushed with their mask;
//  * a lane's own sequence of tests, far terms and triangles is exactly the single-point
//    traversal (children pushed in slot order, popped last-in-first-out), so each point gets
//    the same result the per-point reference loop would give on this tree.
// Triangles closer than tau (near_fp64_rel x bbox diagonal) to the leaf's bounding sphere
// use the FP64 van Oosterom-Strackee formula.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <thread>

#include <algorithm>
#include <cmath>

#include "fwn_internal.h"

namespace fwn {

namespace {

constexpr int QT = 256; // key kernel threads

__device__ __forceinline__ int mesh_of64(const int64_t* off, int M, int64_t t)
{
    int lo = 0, hi = M - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// a value equal in every lane of the wave, made visibly uniform (scalar registers)
__device__ __forceinline__ int64_t wave_uniform64(int64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// 32-bit point keys: mesh id in the top mesh_bits bits, then the leading bits of the point's
// 30-bit Morton code in its mesh's box.  Only the 24 top bits are sorted: a wave needs its 64
// points close together, not in exact curve order, and every point's result is independent of
// which points share its wave.
__global__ __launch_bounds__(QT) void k_query_keys(const float* __restrict__ Q, int64_t nQ,
                                                   const int64_t* __restrict__ q_off, int M, int mesh_bits,
                                                   const float* __restrict__ minfo, uint32_t* __restrict__ keys)
{
    int64_t i = (int64_t)blockIdx.x * QT + threadIdx.x;
    if (i >= nQ) return;
    int m = 0;
    if (M > 1) {
        // the wave's first and last points (wave-uniform: scalar loads); only a wave that spans a
        // mesh boundary searches per lane
        const int64_t w0 = wave_uniform64(i - (threadIdx.x & 63));
        const int m0 = mesh_of64(q_off, M, w0), m1 = mesh_of64(q_off, M, wave_uniform64(min(w0 + 63, nQ - 1)));
        m = (m0 == m1) ? m0 : mesh_of64(q_off, M, i);
    }
    const float* mi = minfo + 8 * m;
    uint32_t mc = morton30(Q[3 * i], Q[3 * i + 1], Q[3 * i + 2], mi, mi + 3);
    const int mbits = 32 - mesh_bits; // bits left for the Morton code
    uint32_t key = (mbits >= 30) ? mc : (mc >> (30 - mbits));
    if (mesh_bits > 0) key |= (uint32_t)m << mbits;
    keys[i] = key; // the sort's first pass takes the values to be the positions i
}

// exact triangle solid angle, FP32, normalised-vector van Oosterom-Strackee form
__device__ __forceinline__ float tri_f32(const float* __restrict__ t, float qx, float qy, float qz)
{
    float ax = t[0] - qx, ay = t[1] - qy, az = t[2] - qz;
    float bx = t[3] - qx, by = t[4] - qy, bz = t[5] - qz;
    float cx = t[6] - qx, cy = t[7] - qy, cz = t[8] - qz;
    float la2 = fmaf(ax, ax, fmaf(ay, ay, az * az));
    float lb2 = fmaf(bx, bx, fmaf(by, by, bz * bz));
    float lc2 = fmaf(cx, cx, fmaf(cy, cy, cz * cz));
    if (la2 == 0.0f || lb2 == 0.0f || lc2 == 0.0f) return 0.0f;
    float ia = __builtin_amdgcn_rsqf(la2), ib = __builtin_amdgcn_rsqf(lb2), ic = __builtin_amdgcn_rsqf(lc2);
    ax *= ia; ay *= ia; az *= ia;
    bx *= ib; by *= ib; bz *= ib;
    cx *= ic; cy *= ic; cz *= ic;
    float ux = bx - ax, uy = by - ay, uz = bz - az;
    float vx = cx - ax, vy = cy - ay, vz = cz - az;
    float nx = fmaf(uy, vz, -uz * vy);
    float ny = fmaf(uz, vx, -ux * vz);
    float nz = fmaf(ux, vy, -uy * vx);
    float num = fmaf(ax, nx, fmaf(ay, ny, az * nz));
    if (num == 0.0f) return 0.0f;
    float ab = fmaf(ax, bx, fmaf(ay, by, az * bz));
    float ac = fmaf(ax, cx, fmaf(ay, cy, az * cz));
    float bc = fmaf(bx, cx, fmaf(by, cy, bz * cz));
    float den = 1.0f + ab + ac + bc;
    return 2.0f * atan2f(num, den);
}

// exact triangle solid angle, FP64, 2 atan2(a.(b x c), |a||b||c| + (a.b)|c| + (b.c)|a| + (c.a)|b|)
__device__ __forceinline__ double tri_f64(const float* __restrict__ t, float qx, float qy, float qz)
{
    double q0 = qx, q1 = qy, q2 = qz;
    double ax = (double)t[0] - q0, ay = (double)t[1] - q1, az = (double)t[2] - q2;
    double bx = (double)t[3] - q0, by = (double)t[4] - q1, bz = (double)t[5] - q2;
    double cx = (double)t[6] - q0, cy = (double)t[7] - q1, cz = (double)t[8] - q2;
    double la = sqrt(ax * ax + ay * ay + az * az);
    double lb = sqrt(bx * bx + by * by + bz * bz);
    double lc = sqrt(cx * cx + cy * cy + cz * cz);
    if (la == 0.0 || lb == 0.0 || lc == 0.0) return 0.0;
    double num = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    if (num == 0.0) return 0.0;
    double den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (bx * cx + by * cy + bz * cz) * la +
                 (cx * ax + cy * ay + cz * az) * lb;
    return 2.0 * atan2(num, den);
}

// One slot's 24 words held in SGPRs (a 16-dword and an 8-dword scalar load issued together at
// the top of the slot).  The other waves of the SIMD (8 resident) hide the load latency.
typedef uint32_t v16u __attribute__((ext_vector_type(16)));
typedef uint32_t v8u __attribute__((ext_vector_type(8)));
typedef v16u v16u_a4 __attribute__((aligned(4))); // slots are 96-byte records: 4-byte alignment
typedef v8u v8u_a4 __attribute__((aligned(4)));
struct SlotRegs {
    v16u a;
    v8u b;
};

__device__ __forceinline__ SlotRegs load_slot(const uint32_t* __restrict__ p)
{
    return SlotRegs{*reinterpret_cast<const v16u_a4*>(p), *reinterpret_cast<const v8u_a4*>(p + 16)};
}

__device__ __forceinline__ uint32_t su(const SlotRegs& r, int k) { return k < 16 ? r.a[k] : r.b[k - 16]; }
__device__ __forceinline__ float sw(const SlotRegs& r, int k) { return __uint_as_float(su(r, k)); }

// order-2 far field of one slot, split into t and inv3 (the caller adds fmaf(t, inv3, omega)):
//   omega += inv3 (o0 + inv2 (q1 + inv2 q2)),  inv = 1 / |d|,
// o0 = d.N, q1 = d^T A d, q2 = B(d, d, d).  The build folds the constant terms (tr M1 of o0 and
// C.d of q1) into A and B through |d|^2 / |d|^2, so every product has one coefficient operand
// and no coefficient needs a move into a vector register; q1 and q2 are evaluated in Horner form
// over the packed coefficients (no monomials needed).  W(k) reads slot word k: scalar registers
// in the batched traversal, per-lane registers in the single-point kernel (the same operations,
// so the same bits).
template <class W>
__device__ __forceinline__ void far_field_terms(const W& w, float dx, float dy, float dz, float d2, float& t,
                                                float& inv3)
{
    const float inv = __builtin_amdgcn_rsqf(d2);
    const float inv2 = inv * inv;
    inv3 = inv2 * inv;
    const float o0 = fmaf(dx, w(FWN_W_NX), fmaf(dy, w(FWN_W_NY), dz * w(FWN_W_NZ)));
    const float zz = dz * dz; // shared with the slot test's |d|^2
    // A: xx yy zz xy yz zx;  q1 = A_zz zz + x (A_xx x + A_xy y + A_zx z) + y (A_yy y + A_yz z)
    const float a1 = fmaf(w(FWN_W_A), dx, fmaf(w(FWN_W_A + 3), dy, w(FWN_W_A + 5) * dz));
    const float b1 = fmaf(w(FWN_W_A + 1), dy, w(FWN_W_A + 4) * dz);
    const float q1 = fmaf(dy, b1, fmaf(dx, a1, w(FWN_W_A + 2) * zz));
    // B: xxx yyy zzz xxy xxz yyx yyz zzx zzy xyz;
    // q2 = zz (B_zzx x + B_zzy y + B_zzz z) + x (x (B_xxx x + B_xxy y + B_xxz z) + y (B_yyx y + B_xyz z))
    //      + y y (B_yyy y + B_yyz z)
    const float L = fmaf(w(FWN_W_B + 7), dx, fmaf(w(FWN_W_B + 8), dy, w(FWN_W_B + 2) * dz));
    const float P = fmaf(w(FWN_W_B), dx, fmaf(w(FWN_W_B + 3), dy, w(FWN_W_B + 4) * dz));
    const float Qv = fmaf(w(FWN_W_B + 5), dy, w(FWN_W_B + 9) * dz);
    const float X = fmaf(dy, Qv, dx * P);
    const float R = fmaf(w(FWN_W_B + 1), dy, w(FWN_W_B + 6) * dz);
    const float q2 = fmaf(dy * dy, R, fmaf(dx, X, zz * L));
    t = fmaf(inv2, fmaf(inv2, q2, q1), o0);
}

struct SlotWords {
    const SlotRegs& r;
    __device__ __forceinline__ float operator()(int k) const { return sw(r, k); }
};

__device__ __forceinline__ float far_field_acc(const SlotRegs& r, float dx, float dy, float dz, float d2, float om)
{
    float t, inv3;
    far_field_terms(SlotWords{r}, dx, dy, dz, d2, t, inv3);
    return fmaf(t, inv3, om);
}

// Deferred FP64 items are written in chunks of kItemChunk slots: a wave claims a chunk with one
// atomic on the chunk counter and fills it over several leaves, so the shared counter sees one
// atomic per 64 items instead of one per leaf event.  Unfilled slots of a chunk are written as
// empty items (count 0), which the FP64 kernels skip.
constexpr uint32_t kItemChunk = 64;
struct ItemChunk {
    uint32_t base = 0u;         // first item slot of the wave's open chunk
    uint32_t fill = kItemChunk; // slots used in it (kItemChunk: no chunk open)
};

__device__ __forceinline__ void pad_chunk(const ItemChunk& ch, int lane, uint32_t* __restrict__ items,
                                          uint32_t item_cap)
{
    if ((uint32_t)lane >= ch.fill) {
        const uint32_t k = ch.base + (uint32_t)lane;
        if (k < item_cap) {
            items[3 * (size_t)k] = 0u;
            items[3 * (size_t)k + 1] = 0u;
            items[3 * (size_t)k + 2] = 0u;
        }
    }
}

// One batch of up to bpts <= 64 points (one per lane) of one mesh, walked through that mesh's
// tree (bpts is 64 except in small calls, which spread their points over more waves).  INL:
// FP64 near leaves are evaluated in place (each leaf's FP64 sum in 2^-40 fixed point, the
// point's total added once at the end: the value the deferred pass gives) instead of being
// deferred to the work list; for calls that use no work list.
template <bool STATS, int NB, bool INL = false>
__device__ __forceinline__ void traverse_batch(
    int64_t batch, const uint32_t* __restrict__ slots, int width, const uint32_t* __restrict__ tris,
    const int32_t* __restrict__ roots, const float* __restrict__ minfo, const float* __restrict__ Q,
    const uint32_t* __restrict__ perm, const int64_t* __restrict__ q_off, const int64_t* __restrict__ b_off,
    const int32_t* __restrict__ mesh_order, int M, int bpts, float beta2, float* __restrict__ omega,
    uint8_t* __restrict__ inside, int32_t* __restrict__ stats, uint32_t* __restrict__ items,
    uint32_t* __restrict__ n_chunks, uint32_t item_cap, int32_t* __restrict__ flag, ItemChunk& ch)
{
    const int lane = threadIdx.x & 63;
    // batches are numbered over the meshes in mesh_order; b_off is the prefix over that order
    const int pos = mesh_of64(b_off, M, batch);
    const int m = mesh_order[pos];
    const int64_t start = q_off[m] + (batch - b_off[pos]) * bpts;
    const int64_t cnt64 = q_off[m + 1] - start;
    const int cnt = cnt64 < bpts ? (int)cnt64 : bpts;
    const bool valid = lane < cnt;
    const uint32_t jpos = (uint32_t)(start + lane); // sorted position of this lane's point
    const uint32_t orig = valid ? (perm ? perm[jpos] : jpos) : 0u;
    const float qx = valid ? Q[3 * (int64_t)orig] : 0.0f;
    const float qy = valid ? Q[3 * (int64_t)orig + 1] : 0.0f;
    const float qz = valid ? Q[3 * (int64_t)orig + 2] : 0.0f;
    const int32_t root = roots[2 * m];
    const int32_t root_ns = roots[2 * m + 1];
    const float tau = minfo[8 * m + 6];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    float om = 0.0f;
    unsigned long long acc64 = 0ull; // INL: FP64 near leaves in 2^-40 fixed point
    int32_t n_test = 0, n_far = 0, n_t32 = 0, n_t64 = 0;
    if (root >= 0) {
        WaveStack<NB> S;
        S.push(0, lane, ((uint32_t)root << 4) | (uint32_t)root_ns, __ballot(valid));
        int sp = 1;
        bool overflow = false;
        while (sp > 0) {
            --sp;
            uint32_t e;
            uint64_t mk;
            S.pop(sp, e, mk);
            const bool act = __builtin_amdgcn_inverse_ballot_w64(mk); // this lane's bit of mk, as a mask
            const uint32_t* nb = slots + (size_t)(e >> 4) * (size_t)width * kSlotWords;
            const int ns = (int)(e & 15u);
            if (ns == 0) continue;
            // one slot: test, far term, leaf triangles or push of the child
            auto slot = [&](const SlotRegs& cur) __attribute__((always_inline)) {
                const float dx = sw(cur, FWN_W_PX) - qx, dy = sw(cur, FWN_W_PY) - qy, dz = sw(cur, FWN_W_PZ) - qz;
                const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const bool far = d2 > sw(cur, FWN_W_THR); // beta^2 r^2, written by the build
                // near lanes: active and not far (scalar mask arithmetic on the compare's ballot,
                // taken before the divergent far-field branch)
                const uint64_t nearm = mk & ~__ballot(far);
                if (STATS) n_test += act ? 1 : 0;
                if (act && far) {
                    om = far_field_acc(cur, dx, dy, dz, d2, om);
                    if (STATS) ++n_far;
                }
                const uint32_t link = su(cur, FWN_W_LINK);
                if (nearm == 0ull) return;
                // written above the stack top for every slot with near lanes, kept (sp advanced)
                // only for an internal slot: the writelanes then sit in a branch-free block and
                // the stack registers need no copies.  The position written is below the
                // capacity (the host keeps a one-entry margin for it), so the entries below stay
                // intact; an overflow is checked once per node.  The link of an internal slot is
                // the stack entry: (child << 4) | its slot count
                S.push(sp, lane, link, nearm);
                const bool leaf = (int32_t)link < 0; // FWN_LINK_LEAF is the sign bit
                sp += leaf ? 0 : 1;
                if (leaf) {
                    const int32_t first = (int32_t)((link & ~FWN_LINK_LEAF) >> 3);
                    const int32_t tcnt = (int32_t)(link & 7u) + 1;
                    const bool nl = __builtin_amdgcn_inverse_ballot_w64(nearm);
                    // FP64 for points within tau of the leaf's bounding sphere and, for a single
                    // triangle, within tau of its plane (p lies on that plane); both thresholds are
                    // stored in the leaf's first triangle
                    const float* t0 = reinterpret_cast<const float*>(tris + (size_t)first * kTriWords);
                    const float near2 = t0[FWN_T_NEAR2], plane = t0[FWN_T_PLANE];
                    const float nd = fmaf(dx, sw(cur, FWN_W_NX), fmaf(dy, sw(cur, FWN_W_NY), dz * sw(cur, FWN_W_NZ)));
                    const bool use64 = nl && tau > 0.0f && d2 < near2 && (tcnt > 1 || nd * nd < plane);
                    const bool use32 = nl && !use64;
                    if (__ballot(use32) != 0ull) {
                        for (int32_t t = first; t < first + tcnt; ++t) {
                            const float* tw = reinterpret_cast<const float*>(tris + (size_t)t * kTriWords);
                            if (use32) {
                                om += tri_f32(tw, qx, qy, qz);
                                if (STATS) ++n_t32;
                            }
                        }
                    }
                    const uint64_t m64 = __ballot(use64);
                    if (INL && m64 != 0ull) {
                        if (use64) {
                            double s64 = 0.0;
                            for (int32_t t = first; t < first + tcnt; ++t)
                                s64 += tri_f64(reinterpret_cast<const float*>(tris + (size_t)t * kTriWords), qx, qy, qz);
                            acc64 += (unsigned long long)__double2ll_rn(s64 * kFixScale);
                            if (STATS) n_t64 += tcnt;
                        }
                    } else if (m64 != 0ull) {
                        // deferred: (point, leaf) items, evaluated in FP64 by k_near64
                        const uint32_t c = (uint32_t)__popcll(m64);
                        if (ch.fill + c > kItemChunk) {
                            pad_chunk(ch, lane, items, item_cap);
                            uint32_t nc = 0u;
                            if (lane == 0) nc = atomicAdd(n_chunks, 1u);
                            nc = __builtin_amdgcn_readfirstlane(nc);
                            ch.base = min(nc, 0xFFFFFFFFu / kItemChunk) * kItemChunk;
                            ch.fill = 0u;
                        }
                        if (use64) {
                            const uint32_t k = ch.base + ch.fill + (uint32_t)__popcll(m64 & lt_mask);
                            if (k < item_cap) {
                                items[3 * (size_t)k] = jpos;
                                items[3 * (size_t)k + 1] = (uint32_t)first;
                                items[3 * (size_t)k + 2] = (uint32_t)tcnt;
                            }
                            if (STATS) n_t64 += tcnt;
                        }
                        ch.fill += c;
                    }
                }
            };
            const uint32_t* sl = nb;
            for (int s = 0; s < ns; ++s, sl += kSlotWords) slot(load_slot(sl));
            // the host only launches trees whose stack bound 1 + (width - 1) depth fits; a push
            // past the capacity is reported
            overflow |= sp > WaveStack<NB>::kCap;
        }
        if (overflow && lane == 0) atomicOr(flag, 1);
    }
    if (INL && acc64 != 0ull) om = (float)((double)om + (double)(long long)acc64 * (1.0 / kFixScale));
    if (valid) {
        omega[orig] = om;
        if (inside) inside[orig] = ((double)om / (4.0 * 3.14159265358979323846) > 0.5) ? 1 : 0;
        if (STATS) {
            stats[4 * (int64_t)orig + 0] = n_test;
            stats[4 * (int64_t)orig + 1] = n_far;
            stats[4 * (int64_t)orig + 2] = n_t32;
            stats[4 * (int64_t)orig + 3] = n_t64;
        }
    }
}

// Grid launch: one wave per batch, XCD-aware block order.  Persistent launch: a grid sized to
// the device's resident waves; each wave takes batches from a work counter until none is left
// (the batch numbering puts the meshes with the most triangles first).
template <bool STATS, int WAVES, bool PERSIST, int NB, bool INL = false>
__global__ __launch_bounds__(WAVES * 64) void k_traverse(
    const uint32_t* __restrict__ slots, int width, const uint32_t* __restrict__ tris,
    const int32_t* __restrict__ roots, const float* __restrict__ minfo, const float* __restrict__ Q,
    const uint32_t* __restrict__ perm, const int64_t* __restrict__ q_off, const int64_t* __restrict__ b_off,
    const int32_t* __restrict__ mesh_order, int M, int64_t n_batches, int bpts, float beta2, float* __restrict__ omega,
    uint8_t* __restrict__ inside, int32_t* __restrict__ stats, uint32_t* __restrict__ items,
    uint32_t* __restrict__ n_chunks, uint32_t item_cap, int32_t* __restrict__ flag, uint32_t* __restrict__ work)
{
    ItemChunk ch;
    if (!PERSIST) {
        const int64_t batch = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * WAVES + (threadIdx.x >> 6);
        if (batch >= n_batches) return; // whole wave
        traverse_batch<STATS, NB, INL>(batch, slots, width, tris, roots, minfo, Q, perm, q_off, b_off, mesh_order, M, bpts,
                                  beta2, omega, inside, stats, items, n_chunks, item_cap, flag, ch);
    } else {
        for (;;) {
            uint32_t b = 0u;
            if ((threadIdx.x & 63) == 0) b = atomicAdd(work, 1u);
            b = __builtin_amdgcn_readfirstlane(b);
            if ((int64_t)b >= n_batches) break; // every wave leaves once the counter passes the end
            traverse_batch<STATS, NB, INL>((int64_t)b, slots, width, tris, roots, minfo, Q, perm, q_off, b_off, mesh_order, M,
                                      bpts, beta2, omega, inside, stats, items, n_chunks, item_cap, flag, ch);
        }
    }
    if (!INL) pad_chunk(ch, threadIdx.x & 63, items, item_cap); // close the wave's open chunk
}

// FP64 solid angles of the deferred (point, leaf) items, summed per point in 2^-40 fixed
// point: integer addition is associative, so the per-point sum does not depend on the order
// in which items are processed.
__global__ __launch_bounds__(256) void k_near64(const uint32_t* __restrict__ items,
                                                const uint32_t* __restrict__ n_chunks, uint32_t item_cap,
                                                const uint32_t* __restrict__ tris, const float* __restrict__ Q,
                                                const uint32_t* __restrict__ perm,
                                                unsigned long long* __restrict__ acc)
{
    const uint32_t n = (uint32_t)min((uint64_t)*n_chunks * kItemChunk, (uint64_t)item_cap);
    for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < n; k += gridDim.x * 256u) {
        const uint32_t j = items[3 * (size_t)k], first = items[3 * (size_t)k + 1], cnt = items[3 * (size_t)k + 2];
        if (cnt == 0u) continue; // unfilled slot of a chunk
        const uint32_t orig = perm ? perm[j] : j;
        const float qx = Q[3 * (size_t)orig], qy = Q[3 * (size_t)orig + 1], qz = Q[3 * (size_t)orig + 2];
        double s = 0.0;
        for (uint32_t t = first; t < first + cnt; ++t)
            s += tri_f64(reinterpret_cast<const float*>(tris + (size_t)t * kTriWords), qx, qy, qz);
        atomicAdd(&acc[j], (unsigned long long)__double2ll_rn(s * kFixScale));
    }
}

// omega = fp32 sum + FP64 near-field sum, once per point: the item that takes the accumulator
// (atomic exchange with 0) does the update, which also leaves the accumulator zeroed.
__global__ __launch_bounds__(256) void k_near64_combine(const uint32_t* __restrict__ items,
                                                        const uint32_t* __restrict__ n_chunks, uint32_t item_cap,
                                                        const uint32_t* __restrict__ perm,
                                                        unsigned long long* __restrict__ acc,
                                                        float* __restrict__ omega, uint8_t* __restrict__ inside)
{
    const uint32_t n = (uint32_t)min((uint64_t)*n_chunks * kItemChunk, (uint64_t)item_cap);
    for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < n; k += gridDim.x * 256u) {
        if (items[3 * (size_t)k + 2] == 0u) continue; // unfilled slot of a chunk
        const uint32_t j = items[3 * (size_t)k];
        const long long v = (long long)atomicExch(&acc[j], 0ull);
        if (v == 0) continue;
        const uint32_t orig = perm ? perm[j] : j;
        const float om = (float)((double)omega[orig] + (double)v * (1.0 / kFixScale));
        omega[orig] = om;
        if (inside) inside[orig] = ((double)om / (4.0 * 3.14159265358979323846) > 0.5) ? 1 : 0;
    }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ __launch_bounds__(256) void k_bruteforce(const uint32_t* __restrict__ tris, int64_t first, int64_t cnt,
                                                    const float* __restrict__ Q, int64_t nQ,
                                                    double* __restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    bool valid = i < nQ;
    float qx = valid ? Q[3 * i] : 0.0f, qy = valid ? Q[3 * i + 1] : 0.0f, qz = valid ? Q[3 * i + 2] : 0.0f;
    double s = 0.0;
    for (int64_t t = first; t < first + cnt; ++t)
        s += tri_f64(reinterpret_cast<const float*>(tris + t * kTriWords), qx, qy, qz);
    if (valid) out[i] = s;
}

// ---------------------------------------------------------------------------------------
// Single-point calls (the reference's per-point solid_angle / is_inside) and other host calls of
// a few points run k_point_bfs (below); k_point answers the points whose breadth-first walk does
// not fit in its LDS records.  k_point: one wave per point, the lanes of the wave taking the
// slots of a node in parallel.  A point walks the same tree in the same order as in the batched
// traversal (children pushed in slot order, popped last-in-first-out), so the walk is a chain of
// dependent node loads.  Here each node is tested and expanded in one parallel step (lane s
// takes slot s), and with PF every pushed child's node is fetched at push time straight into the
// LDS buffer of its stack position (global_load_lds): the child popped next waits for one load
// round trip, its siblings are already there when their turn comes.  The terms are then added
// in the batched traversal's order (slot order, the same FP32 operations: far terms as
// fmaf(t, inv3, omega), near FP32 triangles one by one) and the FP64 near leaves in 2^-40 fixed
// point, so a point gets bit for bit the value a batched call gives it.  Points are read from
// and results written to pinned host memory.
struct LaneWords {
    const float* w;
    __device__ __forceinline__ float operator()(int k) const { return w[k]; }
};

__device__ __forceinline__ float readlane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// calls of up to kPointArgs points pass them in the kernel arguments (no read of host memory
// across the bus before the first node)
constexpr int kPointArgs = 16;
struct PointArgs {
    float q[kPointArgs][3];
};

// one node (width x 96 bytes) into LDS: lane l < width * 6 copies 16 bytes
__device__ __forceinline__ void fetch_node(const uint32_t* __restrict__ nb, uint4* lds, int width, int lane)
{
    if (lane < width * (kSlotWords / 4))
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint4*>(nb) + lane, lds, 16, 0, 0);
}

template <int NB, bool PF>
__global__ __launch_bounds__(64) void k_point(const uint32_t* __restrict__ slots, int width,
                                              const uint32_t* __restrict__ tris, int32_t root, int32_t root_ns,
                                              float tau, const float* __restrict__ Q, PointArgs pa,
                                              float* __restrict__ omega, uint8_t* __restrict__ inside,
                                              int32_t* __restrict__ flag, int lds_nodes, uint8_t* __restrict__ done)
{
    extern __shared__ uint4 nodebuf[]; // PF: one node per stack position (width * 6 uint4 each), lds_nodes of them
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (done && *(volatile const uint8_t*)(done + i) != 0) return; // answered by k_point_bfs
    const int nwords = width * (kSlotWords / 4); // uint4 per node
    const float qx = Q ? Q[3 * (size_t)i] : pa.q[i][0];
    const float qy = Q ? Q[3 * (size_t)i + 1] : pa.q[i][1];
    const float qz = Q ? Q[3 * (size_t)i + 2] : pa.q[i][2];
    float om = 0.0f;
    unsigned long long acc = 0ull; // FP64 near leaves, 2^-40 fixed point
    if (root >= 0) {
        WaveStack<NB> S;
        S.push(0, lane, ((uint32_t)root << 4) | (uint32_t)root_ns, 0ull);
        if (PF) fetch_node(slots + (size_t)root * (size_t)width * kSlotWords, nodebuf, width, lane);
        int sp = 1;
        bool overflow = false;
        while (sp > 0) {
            --sp;
            uint32_t e;
            uint64_t unused;
            S.pop(sp, e, unused);
            const int ns = (int)(e & 15u);
            if (ns == 0) continue;
            // parallel phase: lane s < ns reads, tests and (when far) expands slot s
            float w[FWN_W_LINK];
            uint32_t link = FWN_LINK_EMPTY;
            float dx = 0.0f, dy = 0.0f, dz = 0.0f, d2 = 0.0f, t = 0.0f, inv3 = 0.0f;
            bool far = false;
            if (PF) asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); // this node's fetch has landed
            if (PF && sp >= lds_nodes) overflow = true; // never with the host's stack bound
            if (lane < ns && !(PF && sp >= lds_nodes)) {
                const uint4* src = PF ? nodebuf + (size_t)sp * nwords + lane * (kSlotWords / 4)
                                      : reinterpret_cast<const uint4*>(slots + ((size_t)(e >> 4) * (size_t)width + lane) * kSlotWords);
                uint32_t u[kSlotWords];
#pragma unroll
                for (int k = 0; k < kSlotWords / 4; ++k) {
                    const uint4 v = src[k];
                    u[4 * k] = v.x;
                    u[4 * k + 1] = v.y;
                    u[4 * k + 2] = v.z;
                    u[4 * k + 3] = v.w;
                }
#pragma unroll
                for (int k = 0; k < FWN_W_LINK; ++k) w[k] = __uint_as_float(u[k]);
                link = u[FWN_W_LINK];
                dx = w[FWN_W_PX] - qx;
                dy = w[FWN_W_PY] - qy;
                dz = w[FWN_W_PZ] - qz;
                d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                far = d2 > w[FWN_W_THR];
                // evaluated on every lane (used only where far): the lanes run in parallel, and
                // an unconditional use keeps all six loads of the slot in one round trip
                far_field_terms(LaneWords{w}, dx, dy, dz, d2, t, inv3);
            }
            const uint64_t farm = __ballot(far);
            // near leaves: the FP64 rule's thresholds from the leaf's first triangle (loaded before the
            // children's fetches, so that its wait does not include them)
            const bool nleaf = lane < ns && !far && (link & FWN_LINK_LEAF) != 0u;
            bool use64 = false;
            if (nleaf && tau > 0.0f) {
                const int32_t first = (int32_t)((link & ~FWN_LINK_LEAF) >> 3);
                const int32_t tcnt = (int32_t)(link & 7u) + 1;
                // words 10-11 of a 48-byte record: one 8-byte load
                const float2 np2 = *reinterpret_cast<const float2*>(tris + (size_t)first * kTriWords + FWN_T_NEAR2);
                const float nd = fmaf(dx, w[FWN_W_NX], fmaf(dy, w[FWN_W_NY], dz * w[FWN_W_NZ]));
                use64 = (d2 < np2.x) & ((tcnt > 1) | (nd * nd < np2.y)); // no short circuit: one load
            }
            const uint64_t m64 = __ballot(use64);
            // children first, in slot order, each fetched as it is pushed (before the serial sums)
            const uint64_t pushm = __ballot(lane < ns && !far && (link & FWN_LINK_LEAF) == 0u);
            for (int s2 = 0; s2 < ns; ++s2) {
                if (!((pushm >> s2) & 1ull)) continue;
                const uint32_t ls = (uint32_t)__builtin_amdgcn_readlane((int)link, s2);
                overflow |= sp >= WaveStack<NB>::kCap;
                S.push(sp, lane, ls, 0ull);
                if (PF && sp < lds_nodes)
                    fetch_node(slots + (size_t)(ls >> 4) * (size_t)width * kSlotWords, nodebuf + (size_t)sp * nwords,
                               width, lane);
                ++sp;
            }
            // serial phase: the terms in slot order
            for (int s2 = 0; s2 < ns; ++s2) {
                if ((farm >> s2) & 1ull) {
                    om = fmaf(readlane_f(t, s2), readlane_f(inv3, s2), om);
                } else if (!((pushm >> s2) & 1ull)) {
                    const uint32_t ls = (uint32_t)__builtin_amdgcn_readlane((int)link, s2);
                    const int32_t first = (int32_t)((ls & ~FWN_LINK_LEAF) >> 3);
                    const int32_t tcnt = (int32_t)(ls & 7u) + 1;
                    if ((m64 >> s2) & 1ull) {
                        double s64 = 0.0;
                        for (int32_t k = first; k < first + tcnt; ++k)
                            s64 += tri_f64(reinterpret_cast<const float*>(tris + (size_t)k * kTriWords), qx, qy, qz);
                        acc += (unsigned long long)__double2ll_rn(s64 * kFixScale);
                    } else {
                        for (int32_t k = first; k < first + tcnt; ++k)
                            om += tri_f32(reinterpret_cast<const float*>(tris + (size_t)k * kTriWords), qx, qy, qz);
                    }
                }
            }
        }
        if (overflow && lane == 0) flag[0] = 1;
    }
    const long long v = (long long)acc;
    if (v != 0) om = (float)((double)om + (double)v * (1.0 / kFixScale));
    if (lane == 0) {
        omega[i] = om;
        if (inside) inside[i] = ((double)om / (4.0 * 3.14159265358979323846) > 0.5) ? 1 : 0;
        if (done) { // the host waits on this byte: the results and the overflow flag go first
            __threadfence_system();
            *(volatile uint8_t*)(done + i) = 1;
        }
    }
}

// Breadth-first single-point walk (host calls of a few points; the per-point API).  The walk of
// k_point is a chain of dependent node loads; here one wave per point takes the tree a level at
// a time, every slot of up to 64 / width of the level's nodes in its own lane, so the point waits
// for one load round trip per level step instead of one per node.  Each visited node gets a
// record in LDS: per slot its kind (far, near internal, near FP32 leaf, near FP64 leaf), the far
// term's t and inv3, the child's record or the FP32 leaf's triangle values.  FP64 near leaves
// are evaluated on the way and summed in 2^-40 fixed point (integer addition: any order).
//
// The FP32 terms must then be added in the batched traversal's order: a node's own terms in slot
// order (far terms, then the triangles of near leaves), followed by the subtrees of its near
// children in reverse slot order (they were pushed in slot order and are popped last first).  So
// two more level-parallel passes over the records compute, bottom up, every record's number of
// terms T (kept in its parent's slot entry, beside the other slots' own term counts, so that a
// pass reads one word per slot) and, top down, where its terms start in that order (a slot's own terms after the
// node's earlier own terms; a child's subtree after the node's own terms and the subtrees of the
// children in later slots), writing each term's pair (a, b) to that position of a list.  One
// chain of fmaf over the list gives the sum: far terms as fmaf(t, inv3, omega) and triangle
// values as fmaf(1, v, omega) = omega + v (exact product, one rounding), the same operations on
// the same values as the batched traversal, so a point gets bit for bit the value a batched call
// gives it.  A walk that needs more than max_rec records, max_tv triangle values or kBfsLevels
// levels writes no result and marks its done byte 2 (the host answers that point with k_point);
// every other point marks it 1.
constexpr uint32_t kRecLeaf64 = 0u, kRecFar = 1u, kRecInner = 2u, kRecLeaf32 = 3u;
constexpr int kBfsLevels = 63; // level bounds live in the lanes of one register (lane L: first record of level L)
// slot entry: kind in bits 0-1; a near internal slot: child record << 2; an FP32 leaf:
// (triangles - 1) << 2 | first value << 5

// sums over aligned groups of w lanes (w = 2, 4 or 8, all lanes active; ls = lane % w): inclusive
// prefix and inclusive suffix by DPP row shifts (a group lies inside one 16-lane row, and the
// lanes a shift brings in from outside the group are masked off)
template <int D>
__device__ __forceinline__ int dpp_shr(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x110 + D, 0xF, 0xF, true); }
template <int D>
__device__ __forceinline__ int dpp_shl(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x100 + D, 0xF, 0xF, true); }
__device__ __forceinline__ int seg_prefix(int x, int ls, int w)
{
    int y = dpp_shr<1>(x);
    x += ls >= 1 ? y : 0;
    if (w > 2) {
        y = dpp_shr<2>(x);
        x += ls >= 2 ? y : 0;
    }
    if (w > 4) {
        y = dpp_shr<4>(x);
        x += ls >= 4 ? y : 0;
    }
    return x;
}
__device__ __forceinline__ int seg_suffix(int x, int ls, int w)
{
    int y = dpp_shl<1>(x);
    x += ls + 1 < w ? y : 0;
    if (w > 2) {
        y = dpp_shl<2>(x);
        x += ls + 2 < w ? y : 0;
    }
    if (w > 4) {
        y = dpp_shl<4>(x);
        x += ls + 4 < w ? y : 0;
    }
    return x;
}

__global__ __launch_bounds__(64) void k_point_bfs(const uint32_t* __restrict__ slots, int width,
                                                  const uint32_t* __restrict__ tris, int32_t root, int32_t root_ns,
                                                  float tau, const float* __restrict__ Q, PointArgs pa, int max_rec,
                                                  int max_tv, float* __restrict__ omega, uint8_t* __restrict__ inside,
                                                  uint8_t* __restrict__ done)
{
    // LDS: the list of term pairs (at most one per slot entry and FP32 value); per record its
    // node link, its parent's slot entry and its start; per slot entry its kind word (kind 0 for
    // a node's unused slots), t, inv3 and its term count (own terms; T of the child for a near
    // internal slot); the FP32 triangle values
    extern __shared__ float2 terms[];
    const int nsl = max_rec * width; // slot entries
    uint32_t* rec_node = reinterpret_cast<uint32_t*>(terms + nsl + max_tv);
    int* rec_par = reinterpret_cast<int*>(rec_node + max_rec);
    int* rec_S = rec_par + max_rec;
    uint32_t* rec_k = reinterpret_cast<uint32_t*>(rec_S + max_rec);
    float* rec_a = reinterpret_cast<float*>(rec_k + nsl);
    float* rec_b = rec_a + nsl;
    int* slot_n = reinterpret_cast<int*>(rec_b + nsl);
    float* trival = reinterpret_cast<float*>(slot_n + nsl);
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const float qx = Q ? Q[3 * (size_t)i] : pa.q[i][0];
    const float qy = Q ? Q[3 * (size_t)i + 1] : pa.q[i][1];
    const float qz = Q ? Q[3 * (size_t)i + 2] : pa.q[i][2];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int per = 64 / width;       // nodes per step
    const int ln = lane / width;      // this lane's node within the step
    const int ls = lane - ln * width; // and slot
    float om = 0.0f;
    unsigned long long acc_lane = 0ull; // this lane's FP64 near leaves, 2^-40 fixed point
    bool redo = false;
    int lvl = 0; // lane L: the first record of level L (level L holds records [lvl_L, lvl_L+1))
    if (root >= 0) {
        if (lane == 0) {
            rec_node[0] = ((uint32_t)root << 4) | (uint32_t)root_ns;
            rec_S[0] = 0;
            rec_par[0] = -1;
        }
        __syncthreads();
        // pass 1: the walk, a level at a time
        int b = 0, e = 1, nfree = 1, tvfree = 0, nlev = 0;
        while (b < e) {
            if (nlev == kBfsLevels) {
                redo = true;
                break;
            }
            for (int c = b; c < e; c += per) {
                const int r = c + ln;
                const bool in_r = ln < per && r < e;
                const uint32_t ne = in_r ? rec_node[r] : 0u;
                const bool act = in_r && ls < (int)(ne & 15u);
                float w[FWN_W_LINK];
                uint32_t link = FWN_LINK_EMPTY;
                float dx = 0.0f, dy = 0.0f, dz = 0.0f, d2 = 0.0f, t = 0.0f, inv3 = 0.0f;
                bool far = false;
                if (act) {
                    const uint4* src =
                        reinterpret_cast<const uint4*>(slots + ((size_t)(ne >> 4) * (size_t)width + ls) * kSlotWords);
                    uint32_t u[kSlotWords];
#pragma unroll
                    for (int k = 0; k < kSlotWords / 4; ++k) {
                        const uint4 v = src[k];
                        u[4 * k] = v.x;
                        u[4 * k + 1] = v.y;
                        u[4 * k + 2] = v.z;
                        u[4 * k + 3] = v.w;
                    }
#pragma unroll
                    for (int k = 0; k < FWN_W_LINK; ++k) w[k] = __uint_as_float(u[k]);
                    link = u[FWN_W_LINK];
                    dx = w[FWN_W_PX] - qx;
                    dy = w[FWN_W_PY] - qy;
                    dz = w[FWN_W_PZ] - qz;
                    d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    far = d2 > w[FWN_W_THR];
                    far_field_terms(LaneWords{w}, dx, dy, dz, d2, t, inv3);
                }
                const bool inner = act && !far && (int32_t)link >= 0;
                const bool nleaf = act && !far && (int32_t)link < 0;
                // a near leaf: its first triangle's record (vertices and the FP64 rule's
                // thresholds) in one round trip
                int32_t first = 0, tcnt = 0;
                float tw[kTriWords];
                bool use64 = false;
                if (nleaf) {
                    first = (int32_t)((link & ~FWN_LINK_LEAF) >> 3);
                    tcnt = (int32_t)(link & 7u) + 1;
                    const uint4* tp = reinterpret_cast<const uint4*>(tris + (size_t)first * kTriWords);
#pragma unroll
                    for (int k = 0; k < kTriWords / 4; ++k) {
                        const uint4 v = tp[k];
                        tw[4 * k] = __uint_as_float(v.x);
                        tw[4 * k + 1] = __uint_as_float(v.y);
                        tw[4 * k + 2] = __uint_as_float(v.z);
                        tw[4 * k + 3] = __uint_as_float(v.w);
                    }
                    const float nd = fmaf(dx, w[FWN_W_NX], fmaf(dy, w[FWN_W_NY], dz * w[FWN_W_NZ]));
                    use64 = tau > 0.0f && d2 < tw[FWN_T_NEAR2] && (tcnt > 1 || nd * nd < tw[FWN_T_PLANE]);
                }
                const bool use32 = nleaf && !use64;
                // records of the near children and places for the FP32 values, in lane order
                const uint64_t im = __ballot(inner);
                const int cnt = __popcll(im);
                const int c32 = use32 ? tcnt : 0; // 1..8: four bits
                int pre = 0, tot = 0;
#pragma unroll
                for (int bit = 0; bit < 4; ++bit) {
                    const uint64_t bm = __ballot((c32 >> bit) & 1);
                    pre += __popcll(bm & lt_mask) << bit;
                    tot += __popcll(bm) << bit;
                }
                if (nfree + cnt > max_rec || tvfree + tot > max_tv) {
                    redo = true;
                    break;
                }
                const int child = nfree + __popcll(im & lt_mask);
                if (inner) {
                    rec_node[child] = link;
                    rec_par[child] = r * width + ls;
                }
                if (use32) {
                    float* tv = trival + tvfree + pre;
                    tv[0] = tri_f32(tw, qx, qy, qz);
                    for (int32_t k = 1; k < tcnt; ++k)
                        tv[k] = tri_f32(reinterpret_cast<const float*>(tris + (size_t)(first + k) * kTriWords), qx, qy, qz);
                }
                if (use64) {
                    double s64 = tri_f64(tw, qx, qy, qz);
                    for (int32_t k = 1; k < tcnt; ++k)
                        s64 += tri_f64(reinterpret_cast<const float*>(tris + (size_t)(first + k) * kTriWords), qx, qy, qz);
                    acc_lane += (unsigned long long)__double2ll_rn(s64 * kFixScale);
                }
                if (in_r) {
                    const int k = r * width + ls;
                    rec_k[k] = !act ? kRecLeaf64 : far ? kRecFar
                                   : (inner ? (kRecInner | ((uint32_t)child << 2))
                                            : (use32 ? (kRecLeaf32 | ((uint32_t)(tcnt - 1) << 2) |
                                                        ((uint32_t)(tvfree + pre) << 5))
                                                     : kRecLeaf64));
                    rec_a[k] = t;
                    rec_b[k] = inv3;
                    slot_n[k] = (act && far) ? 1 : (use32 ? (int)tcnt : 0); // internal: set by pass 2
                }
                nfree += cnt;
                tvfree += tot;
            }
            if (redo) break;
            ++nlev;
            if (lane == nlev) lvl = e; // the level just walked is [b, e); the next, [e, nfree)
            __syncthreads(); // the level's records before the next level reads them
            b = e;
            e = nfree;
        }
        if (!redo) {
            // pass 2, bottom up: T, the terms of a record's subtree, into its parent's slot entry
            int t_root = 0;
            for (int L = nlev - 1; L >= 0; --L) {
                const int lb = __builtin_amdgcn_readlane(lvl, L), le = __builtin_amdgcn_readlane(lvl, L + 1);
                for (int c = lb; c < le; c += per) {
                    const int r = c + ln;
                    const bool in_r = ln < per && r < le;
                    int n = 0, par = -1;
                    if (in_r) {
                        n = slot_n[r * width + ls];
                        par = rec_par[r];
                    }
                    n = seg_prefix(n, ls, width) + seg_suffix(n, ls, width) - n; // the group's total
                    if (in_r && ls == 0) {
                        if (par >= 0) slot_n[par] = n;
                        else t_root = n;
                    }
                }
                __syncthreads();
            }
            t_root = __builtin_amdgcn_readfirstlane(t_root); // lane 0 holds the root's record
            // pass 3, top down: where each record's terms start; each term's location in the list
            for (int L = 0; L < nlev; ++L) {
                const int lb = __builtin_amdgcn_readlane(lvl, L), le = __builtin_amdgcn_readlane(lvl, L + 1);
                for (int c = lb; c < le; c += per) {
                    const int r = c + ln;
                    const bool in_r = ln < per && r < le;
                    const int k = r * width + ls;
                    uint32_t kk = 0u;
                    int sr = 0, nk = 0;
                    float ta = 0.0f, tb = 0.0f;
                    if (in_r) { // one round trip: every word of the slot and the record's start
                        kk = rec_k[k];
                        sr = rec_S[r];
                        ta = rec_a[k];
                        tb = rec_b[k];
                        nk = slot_n[k];
                    }
                    const uint32_t kind = kk & 3u;
                    const int own = kind == kRecInner ? 0 : nk;
                    const int sub = kind == kRecInner ? nk : 0;
                    const int own_incl = seg_prefix(own, ls, width);
                    const int own_all = own_incl + seg_suffix(own, ls, width) - own; // the node's own terms
                    const int sub_incl = seg_suffix(sub, ls, width);
                    if (kind == kRecFar) {
                        terms[sr + own_incl - 1] = make_float2(ta, tb);
                    } else if (kind == kRecLeaf32) {
                        const int p0 = sr + own_incl - own;
                        const float* tv = trival + (kk >> 5);
                        for (int j = 0; j < own; ++j) terms[p0 + j] = make_float2(1.0f, tv[j]);
                    } else if (kind == kRecInner) {
                        rec_S[kk >> 2] = sr + own_all + sub_incl - sub;
                    }
                }
                __syncthreads();
            }
            // the sum, in list order: one fmaf per term (every lane runs the chain on broadcast
            // reads, 16 reads in flight ahead of their fmaf)
            const int nt = min(t_root, nsl + max_tv); // at most one term per slot entry and value
            int j = 0;
            for (; j + 16 <= nt; j += 16) {
                float2 v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) v[u] = terms[j + u];
#pragma unroll
                for (int u = 0; u < 16; ++u) om = fmaf(v[u].x, v[u].y, om);
            }
            for (; j < nt; ++j) {
                const float2 v = terms[j];
                om = fmaf(v.x, v.y, om);
            }
        }
    }
    // the wave's FP64 sum
    unsigned long long acc = 0ull;
    for (uint64_t m = __ballot(acc_lane != 0ull); m != 0ull; m &= m - 1ull) {
        const int l = (int)__builtin_ctzll(m);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)acc_lane, l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(acc_lane >> 32), l);
        acc += ((unsigned long long)hi << 32) | lo;
    }
    const long long v = (long long)acc;
    if (v != 0) om = (float)((double)om + (double)v * (1.0 / kFixScale));
    if (lane == 0) {
        if (!redo) {
            omega[i] = om;
            if (inside) inside[i] = ((double)om / (4.0 * 3.14159265358979323846) > 0.5) ? 1 : 0;
        }
        __threadfence_system(); // the host waits on this byte: the results go first
        *(volatile uint8_t*)(done + i) = redo ? 2 : 1;
    }
}

__global__ void k_zero3(int32_t* a)
{
    a[0] = 0;
    a[1] = 0;
    a[2] = 0;
}

template <bool ST, int WV, bool PE, int NB>
void* traverse_fn()
{
    return reinterpret_cast<void*>(&k_traverse<ST, WV, PE, NB>);
}

// banks: traversal stack banks the tree needs (1: 64 entries, 3: kStack entries)
template <typename... Args>
void launch_traverse(bool stats, int waves, bool persist, int banks, unsigned nblk, hipStream_t s, Args... args)
{
#define FWN_TRAV(ST, WV, PE)                                                                                  \
    do {                                                                                                      \
        if (banks == 1) hipLaunchKernelGGL((k_traverse<ST, WV, PE, 1>), dim3(nblk), dim3(WV * 64), 0, s, args...); \
        else hipLaunchKernelGGL((k_traverse<ST, WV, PE, 3>), dim3(nblk), dim3(WV * 64), 0, s, args...);        \
    } while (0)
    if (stats) {
        FWN_TRAV(true, 1, false);
        return;
    }
    switch (waves * 2 + (persist ? 1 : 0)) {
    case 2: FWN_TRAV(false, 1, false); break;
    case 3: FWN_TRAV(false, 1, true); break;
    case 4: FWN_TRAV(false, 2, false); break;
    case 5: FWN_TRAV(false, 2, true); break;
    case 8: FWN_TRAV(false, 4, false); break;
    case 9: FWN_TRAV(false, 4, true); break;
    case 16: FWN_TRAV(false, 8, false); break;
    case 17: FWN_TRAV(false, 8, true); break;
    default: fail(FWN_INVALID_ARGUMENT, "unknown traversal launch shape (waves per block 1, 2, 4 or 8)");
    }
#undef FWN_TRAV
}

// resident workgroups per CU of a traversal variant (persistent grid size)
int resident_blocks(bool persist, int waves, int banks)
{
    const void* f = nullptr;
    const bool one = banks == 1;
    switch (waves * 2 + (persist ? 1 : 0)) {
    case 3: f = one ? traverse_fn<false, 1, true, 1>() : traverse_fn<false, 1, true, 3>(); break;
    case 5: f = one ? traverse_fn<false, 2, true, 1>() : traverse_fn<false, 2, true, 3>(); break;
    case 9: f = one ? traverse_fn<false, 4, true, 1>() : traverse_fn<false, 4, true, 3>(); break;
    case 17: f = one ? traverse_fn<false, 8, true, 1>() : traverse_fn<false, 8, true, 3>(); break;
    default: fail(FWN_INVALID_ARGUMENT, "unknown persistent launch shape (waves per block 1, 2, 4 or 8)");
    }
    int nb = 0;
    FWN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, waves * 64, 0));
    return nb > 0 ? nb : 1;
}

// Calls with at most this many points skip the Morton sort (12 small launches that would cost
// more than the coherence they buy) and move host data through the pinned staging buffer.
constexpr int64_t kSmallBatch = 2048;

char* host_stage(Engine& e, size_t bytes)
{
    if (e.h_stage_bytes < bytes) {
        if (e.h_stage) FWN_HIP(hipHostFree(e.h_stage));
        e.h_stage = nullptr;
        e.h_stage_bytes = 0;
        const size_t nb = std::max<size_t>(bytes, 1 << 16);
        FWN_HIP(hipHostMalloc(&e.h_stage, nb, hipHostMallocDefault));
        e.h_stage_bytes = nb;
    }
    return static_cast<char*>(e.h_stage);
}

// Host calls of a single-mesh engine larger than one chunk are pipelined: chunk k's points are
// copied (several host threads) into a pinned buffer and sent to the device on one copy stream
// while chunk k-1 is answered on the engine's stream and chunk k-2's results come back on a
// second copy stream and are copied out to the caller's arrays.  Each chunk is an independent
// Morton-sorted query (a point's result does not depend on the other points of its call).
constexpr int64_t kPipeChunk = 1 << 21; // points per chunk

void parallel_copy(void* dst, const void* src, size_t bytes)
{
    const size_t piece = (size_t)4 << 20; // at least 4 MB per thread
    const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes / piece));
    if (nt <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes + nt - 1) / nt;
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) {
        const size_t a = (size_t)t * part, b = std::min(bytes, a + part);
        if (a < b) th.emplace_back([=] { std::memcpy((char*)dst + a, (const char*)src + a, b - a); });
    }
    std::memcpy(dst, src, std::min(bytes, part));
    for (auto& x : th) x.join();
}

void run_queries_pipelined(Engine& e, const float* Q, int64_t nQ, float* omega, uint8_t* inside, int banks)
{
    hipStream_t s = e.stream;
    const int64_t C = kPipeChunk;
    const int64_t nch = (nQ + C - 1) / C;
    const size_t in_b = sizeof(float) * 3 * (size_t)C, om_b = sizeof(float) * (size_t)C, i8_b = (size_t)C;
    const size_t slot_b = in_b + om_b + i8_b;
    if (e.h_pipe_bytes < 2 * slot_b) {
        if (e.h_pipe) FWN_HIP(hipHostFree(e.h_pipe));
        e.h_pipe = nullptr;
        e.h_pipe_bytes = 0;
        FWN_HIP(hipHostMalloc(&e.h_pipe, 2 * slot_b, hipHostMallocDefault));
        e.h_pipe_bytes = 2 * slot_b;
    }
    if (!e.cs_in) FWN_HIP(hipStreamCreateWithFlags(&e.cs_in, hipStreamNonBlocking));
    if (!e.cs_out) FWN_HIP(hipStreamCreateWithFlags(&e.cs_out, hipStreamNonBlocking));
    if (!e.pev[0])
        for (auto& ev : e.pev) FWN_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    // timing: the call spans the first upload to the last download; each chunk's traversal has
    // its own event pair (the tuner and fwn_last_timing read these)
    if (!e.ev[0])
        for (int k = 0; k < 4; ++k) FWN_HIP(hipEventCreate(&e.ev[k]));
    while ((int64_t)e.tev.size() < 2 * nch) {
        hipEvent_t ev = nullptr;
        FWN_HIP(hipEventCreate(&ev));
        e.tev.push_back(ev);
    }
    char* pin = static_cast<char*>(e.h_pipe);
    auto pin_in = [&](int sl) { return reinterpret_cast<float*>(pin + sl * slot_b); };
    auto pin_om = [&](int sl) { return reinterpret_cast<float*>(pin + sl * slot_b + in_b); };
    auto pin_i8 = [&](int sl) { return reinterpret_cast<uint8_t*>(pin + sl * slot_b + in_b + om_b); };

    e.q_in.reserve(sizeof(float) * 3 * nQ);
    e.q_om.reserve(sizeof(float) * nQ);
    if (inside) e.q_in8.reserve(nQ);
    float* dIn3 = e.q_in.as<float>();
    float* dOm = e.q_om.as<float>();
    uint8_t* dIn = inside ? e.q_in8.as<uint8_t>() : nullptr;
    for (int k = 0; k < 2; ++k) {
        e.q_keys[k].reserve(sizeof(uint32_t) * C);
        e.q_vals[k].reserve(sizeof(uint32_t) * C);
    }
    if (e.acc_cap < C) {
        e.q_acc.alloc(sizeof(unsigned long long) * C);
        e.acc_cap = C;
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_blocks(C, 256)), dim3(256), 0, s, e.q_acc.as<unsigned long long>(),
                           C);
        FWN_LAUNCH_CHECK("k_fill_u64");
    }
    if (e.item_cap < std::max<int64_t>(C, 1 << 20)) {
        e.item_cap = std::max<int64_t>(C, 1 << 20);
        e.q_items.alloc(sizeof(uint32_t) * 3 * (size_t)e.item_cap);
    }
    // per chunk: q_off {0, n}, b_off {0, batches}, mesh order {0} (5 int64 words)
    std::vector<int64_t> meta(5 * nch, 0);
    for (int64_t k = 0; k < nch; ++k) {
        const int64_t ck = std::min(C, nQ - k * C);
        meta[5 * k + 1] = ck;
        meta[5 * k + 3] = (ck + 63) / 64;
    }
    e.q_pmeta.reserve(sizeof(int64_t) * meta.size());
    FWN_HIP(hipMemcpyAsync(e.q_pmeta.p, meta.data(), sizeof(int64_t) * meta.size(), hipMemcpyHostToDevice, s));
    e.q_pflags.reserve(sizeof(int32_t) * 4 * nch);
    const int waves = e.opt.waves_per_block > 0 ? e.opt.waves_per_block : 1;
    const bool persist = e.opt.launch_mode == FWN_LAUNCH_PERSISTENT;
    if (persist && e.n_cu == 0) {
        hipDeviceProp_t prop;
        FWN_HIP(hipGetDeviceProperties(&prop, e.device));
        e.n_cu = prop.multiProcessorCount;
    }
    const float beta2 = e.opt.accuracy_scale * e.opt.accuracy_scale;
    const uint32_t cap = (uint32_t)e.item_cap;

    for (int64_t k = 0; k <= nch; ++k) {
        if (k < nch) {
            const int sl = (int)(k & 1);
            const int64_t off = k * C, ck = std::min(C, nQ - off);
            if (k >= 2) FWN_HIP(hipEventSynchronize(e.pev[sl])); // chunk k-2's upload has read the buffer
            parallel_copy(pin_in(sl), Q + 3 * off, sizeof(float) * 3 * (size_t)ck);
            if (k == 0) FWN_HIP(hipEventRecord(e.ev[0], e.cs_in));
            FWN_HIP(hipMemcpyAsync(dIn3 + 3 * off, pin_in(sl), sizeof(float) * 3 * (size_t)ck, hipMemcpyHostToDevice,
                                   e.cs_in));
            FWN_HIP(hipEventRecord(e.pev[sl], e.cs_in));
            FWN_HIP(hipStreamWaitEvent(s, e.pev[sl], 0));
            // chunk k on the engine's stream: sort, traversal, FP64 pass (no host wait)
            const float* dQ = dIn3 + 3 * off;
            const int64_t* d_qoff = e.q_pmeta.as<int64_t>() + 5 * k;
            const int64_t* d_boff = d_qoff + 2;
            const int32_t* d_order = reinterpret_cast<const int32_t*>(d_qoff + 4);
            uint32_t* kp = e.q_keys[0].as<uint32_t>();
            uint32_t* vp = e.q_vals[0].as<uint32_t>();
            uint32_t* kp2 = e.q_keys[1].as<uint32_t>();
            uint32_t* vp2 = e.q_vals[1].as<uint32_t>();
            hipLaunchKernelGGL(k_query_keys, dim3(grid_blocks(ck, QT)), dim3(QT), 0, s, dQ, ck, d_qoff, 1, 0,
                               e.mesh_info.as<float>(), kp);
            FWN_LAUNCH_CHECK("k_query_keys");
            radix_sort_pairs32(kp, vp, kp2, vp2, ck, 8, 32, true, e.q_hist, e.q_rowsum, s);
            const int64_t nb = (ck + 63) / 64;
            int64_t nblk = (nb + waves - 1) / waves;
            if (persist) nblk = std::min<int64_t>(nblk, (int64_t)e.n_cu * resident_blocks(true, waves, banks));
            int32_t* dflag = e.q_pflags.as<int32_t>() + 4 * k;
            uint32_t* dnitems = reinterpret_cast<uint32_t*>(dflag + 1);
            uint32_t* dwork = reinterpret_cast<uint32_t*>(dflag + 2);
            hipLaunchKernelGGL(k_zero3, dim3(1), dim3(1), 0, s, dflag);
            FWN_HIP(hipEventRecord(e.tev[2 * k], s));
            launch_traverse(false, waves, persist, banks, (unsigned)nblk, s, e.slots.as<uint32_t>(), e.width,
                            e.tris.as<uint32_t>(), e.roots.as<int32_t>(), e.mesh_info.as<float>(), dQ, vp, d_qoff,
                            d_boff, d_order, 1, nb, 64, beta2, dOm + off, dIn ? dIn + off : nullptr, (int32_t*)nullptr,
                            e.q_items.as<uint32_t>(), dnitems, cap, dflag, dwork);
            FWN_LAUNCH_CHECK("k_traverse");
            FWN_HIP(hipEventRecord(e.tev[2 * k + 1], s));
            hipLaunchKernelGGL(k_near64, dim3(2048), dim3(256), 0, s, e.q_items.as<uint32_t>(), dnitems, cap,
                               e.tris.as<uint32_t>(), dQ, vp, e.q_acc.as<unsigned long long>());
            FWN_LAUNCH_CHECK("k_near64");
            hipLaunchKernelGGL(k_near64_combine, dim3(2048), dim3(256), 0, s, e.q_items.as<uint32_t>(), dnitems,
                               cap, vp, e.q_acc.as<unsigned long long>(), dOm + off, dIn ? dIn + off : nullptr);
            FWN_LAUNCH_CHECK("k_near64_combine");
            FWN_HIP(hipEventRecord(e.pev[2 + sl], s));
            // chunk k's results back to the pinned buffer on the other copy stream; chunk k-2's
            // results were copied out of that buffer in the previous iteration
            FWN_HIP(hipStreamWaitEvent(e.cs_out, e.pev[2 + sl], 0));
            FWN_HIP(hipMemcpyAsync(pin_om(sl), dOm + off, sizeof(float) * (size_t)ck, hipMemcpyDeviceToHost, e.cs_out));
            if (inside)
                FWN_HIP(hipMemcpyAsync(pin_i8(sl), dIn + off, (size_t)ck, hipMemcpyDeviceToHost, e.cs_out));
            FWN_HIP(hipEventRecord(e.pev[4 + sl], e.cs_out));
            if (k == nch - 1) FWN_HIP(hipEventRecord(e.ev[3], e.cs_out));
        }
        if (k >= 1) { // chunk k-1 to the caller's arrays
            const int64_t j = k - 1;
            const int sl = (int)(j & 1);
            const int64_t off = j * C, cj = std::min(C, nQ - off);
            FWN_HIP(hipEventSynchronize(e.pev[4 + sl]));
            parallel_copy(omega + off, pin_om(sl), sizeof(float) * (size_t)cj);
            if (inside) parallel_copy(inside + off, pin_i8(sl), (size_t)cj);
        }
    }
    std::vector<int32_t> fl(4 * nch);
    FWN_HIP(hipMemcpyAsync(fl.data(), e.q_pflags.p, sizeof(int32_t) * 4 * nch, hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
    FWN_HIP(hipStreamSynchronize(e.cs_in));
    FWN_HIP(hipStreamSynchronize(e.cs_out));
    float t_all = 0.0f;
    double t_tr = 0.0;
    FWN_HIP(hipEventElapsedTime(&t_all, e.ev[0], e.ev[3]));
    for (int64_t k = 0; k < nch; ++k) {
        float t = 0.0f;
        FWN_HIP(hipEventElapsedTime(&t, e.tev[2 * k], e.tev[2 * k + 1]));
        t_tr += t;
    }
    e.last_traverse_ms = t_tr / (double)nch; // average over the chunks' launches
    e.last_total_ms = t_all;
    for (int64_t k = 0; k < nch; ++k) {
        if (fl[4 * k]) fail(FWN_LIMIT_EXCEEDED, "traversal stack overflow");
        const int64_t need = (int64_t)(uint32_t)fl[4 * k + 1] * kItemChunk;
        if (need > (int64_t)cap) { // the chunk's FP64 work list overflowed its capacity in the
                                   // pipeline (a redo may have grown it since): answer it again
            const int64_t off = k * C, ck = std::min(C, nQ - off);
            const double tr = e.last_traverse_ms, tot = e.last_total_ms;
            run_queries(e, Q + 3 * off, nullptr, ck, omega + off, inside ? inside + off : nullptr, nullptr, false);
            e.last_traverse_ms = tr;
            e.last_total_ms = tot + e.last_total_ms; // the redo is part of this call's device time
        }
    }
}

// Host calls of at most kSmallBatch points on a single-mesh engine: one k_point_bfs launch
// (points in its arguments up to kPointArgs, else read from the pinned stage) writes the results
// into the pinned stage and a done byte per point; the points it could not hold in LDS (done
// byte 2) go to one k_point launch.  k_point's LDS node buffers (one per stack position the tree
// can need) are used when they fit in kPointLds bytes.  These calls record no device time.
constexpr size_t kPointLds = 32768;
constexpr size_t kPointBfsLds = 65536; // k_point_bfs: records, FP32 triangle values, term list
constexpr int kPointBfsValues = 1024;

void run_points(Engine& e, const float* Q, int64_t nQ, float* omega, uint8_t* inside, int banks)
{
    hipStream_t s = e.stream;
    const size_t pts_b = (sizeof(float) * 3 * (size_t)nQ + 15) & ~(size_t)15;
    const size_t done_b = ((size_t)nQ + 15) & ~(size_t)15;
    // stage: overflow flag (16 B), a done byte per point, points, omega, inside
    const size_t out_b = 16 + done_b + pts_b + (sizeof(float) + 1) * (size_t)nQ;
    char* hs = host_stage(e, out_b);
    int32_t* hflag = reinterpret_cast<int32_t*>(hs);
    uint8_t* hdone = reinterpret_cast<uint8_t*>(hs + 16);
    float* hq = reinterpret_cast<float*>(hs + 16 + done_b);
    float* hom = reinterpret_cast<float*>(hs + 16 + done_b + pts_b);
    uint8_t* hin = reinterpret_cast<uint8_t*>(hs + 16 + done_b + pts_b + sizeof(float) * (size_t)nQ);
    *hflag = 0;
    std::memset(hdone, 0, (size_t)nQ);
    PointArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    const float* dq = hq; // points read from the pinned stage, or from the arguments (nullptr)
    if (nQ <= kPointArgs) {
        std::memcpy(pa.q, Q, sizeof(float) * 3 * (size_t)nQ);
        dq = nullptr;
    } else {
        std::memcpy(hq, Q, sizeof(float) * 3 * (size_t)nQ);
    }
    const int32_t root = e.h_roots[0], root_ns = e.h_roots[1];
    const float tau = e.h_tau[0];
    const dim3 grid((unsigned)nQ), block(64);
    // the results are waited for by polling the done bytes the kernels write after them: no
    // stream synchronisation in the round trip.  Every 256 polls the stream is asked whether it
    // has finished, so a failed launch ends the wait.
    auto wait_done = [&]() {
        const volatile uint8_t* vd = hdone;
        bool polled = false;
        for (uint32_t spin = 1;; ++spin) {
            int64_t k = 0;
            while (k < nQ && vd[k] != 0) ++k;
            if (k == nQ) {
                polled = true;
                break;
            }
            if ((spin & 255u) == 0u && hipStreamQuery(s) != hipErrorNotReady) break;
        }
        std::atomic_thread_fence(std::memory_order_acquire);
        if (!polled) FWN_HIP(hipStreamSynchronize(s));
    };
    // breadth-first walk, records in LDS
    const int max_tv = kPointBfsValues;
    const int max_rec = (int)((kPointBfsLds / sizeof(uint32_t) - 3 * (size_t)max_tv) / (size_t)(3 + 6 * e.width));
    hipLaunchKernelGGL(k_point_bfs, grid, block, kPointBfsLds, s, e.slots.as<uint32_t>(), e.width,
                       e.tris.as<uint32_t>(), root, root_ns, tau, dq, pa, max_rec, max_tv, hom, hin, hdone);
    FWN_LAUNCH_CHECK("k_point_bfs");
    wait_done();
    // points whose walk needed more records than LDS holds: the depth-first kernel, which skips
    // the points already answered (done byte 1)
    int64_t n_redo = 0;
    for (int64_t k = 0; k < nQ; ++k)
        if (hdone[k] == 2) {
            hdone[k] = 0;
            ++n_redo;
        }
    if (n_redo > 0) {
        const size_t need = (size_t)(1 + (e.width - 1) * std::max(e.max_depth, 1)); // stack entries
        const size_t lds = need * (size_t)e.width * kSlotWords * sizeof(uint32_t);
        const bool pf = banks == 1 && lds <= kPointLds;
        if (pf)
            hipLaunchKernelGGL((k_point<1, true>), grid, block, lds, s, e.slots.as<uint32_t>(), e.width,
                               e.tris.as<uint32_t>(), root, root_ns, tau, dq, pa, hom, hin, hflag, (int)need, hdone);
        else if (banks == 1)
            hipLaunchKernelGGL((k_point<1, false>), grid, block, 0, s, e.slots.as<uint32_t>(), e.width,
                               e.tris.as<uint32_t>(), root, root_ns, tau, dq, pa, hom, hin, hflag, 0, hdone);
        else
            hipLaunchKernelGGL((k_point<3, false>), grid, block, 0, s, e.slots.as<uint32_t>(), e.width,
                               e.tris.as<uint32_t>(), root, root_ns, tau, dq, pa, hom, hin, hflag, 0, hdone);
        FWN_LAUNCH_CHECK("k_point");
        wait_done();
    }
    if (*hflag) fail(FWN_LIMIT_EXCEEDED, "traversal stack overflow");
    std::memcpy(omega, hom, sizeof(float) * (size_t)nQ);
    if (inside) std::memcpy(inside, hin, (size_t)nQ);
    e.last_traverse_ms = 0.0;
    e.last_total_ms = 0.0;
}

} // namespace

void launch_point_keys(Engine& e, const float* dQ, int64_t nQ, uint32_t* keys, hipStream_t s)
{
    hipLaunchKernelGGL(k_query_keys, dim3(grid_blocks(nQ, QT)), dim3(QT), 0, s, dQ, nQ, (const int64_t*)nullptr, 1, 0,
                       e.mesh_info.as<float>(), keys);
    FWN_LAUNCH_CHECK("k_query_keys");
}

void launch_traverse_inplace(Engine& e, int banks, const float* dQ, const uint32_t* perm, const int64_t* meta,
                             int64_t nQ, float* omega, uint8_t* inside, int32_t* flag, hipStream_t s)
{
    const int64_t nb = (nQ + 63) / 64;
    if (nb > INT32_MAX) fail(FWN_LIMIT_EXCEEDED, "too many point batches");
    const float beta2 = e.opt.accuracy_scale * e.opt.accuracy_scale;
    const int64_t* q_off = meta;
    const int64_t* b_off = meta + 2;
    const int32_t* order = reinterpret_cast<const int32_t*>(meta + 4);
#define FWN_TRAV_INL(NB)                                                                                            \
    hipLaunchKernelGGL((k_traverse<false, 1, false, NB, true>), dim3((unsigned)nb), dim3(64), 0, s,                \
                       e.slots.as<uint32_t>(), e.width, e.tris.as<uint32_t>(), e.roots.as<int32_t>(),              \
                       e.mesh_info.as<float>(), dQ, perm, q_off, b_off, order, 1, nb, 64, beta2, omega, inside,    \
                       (int32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u, flag, (uint32_t*)nullptr)
    if (banks == 1) FWN_TRAV_INL(1);
    else FWN_TRAV_INL(3);
#undef FWN_TRAV_INL
    FWN_LAUNCH_CHECK("k_traverse");
}

void run_queries(Engine& e, const float* Q, const int64_t* q_off_host, int64_t nQ, float* omega, uint8_t* inside,
                 int32_t* stats, bool dev)
{
    hipStream_t s = e.stream;
    const int M = e.n_meshes;
    std::vector<int64_t> qoff(M + 1);
    if (q_off_host) {
        for (int m = 0; m <= M; ++m) qoff[m] = q_off_host[m];
        if (qoff[0] != 0) fail(FWN_INVALID_ARGUMENT, "q_off[0] must be 0");
        for (int m = 0; m < M; ++m)
            if (qoff[m + 1] < qoff[m]) fail(FWN_INVALID_ARGUMENT, "q_off must be non-decreasing");
        nQ = qoff[M];
    } else {
        if (M != 1) fail(FWN_INVALID_ARGUMENT, "per-mesh point offsets are required for a multi-mesh engine");
        qoff[0] = 0;
        qoff[1] = nQ;
    }
    if (nQ < 0) fail(FWN_INVALID_ARGUMENT, "negative point count");
    if (nQ == 0) return;
    if (nQ > (int64_t)UINT32_MAX / 2) fail(FWN_LIMIT_EXCEEDED, "more than 2^31 points in one call");
    const int banks = stack_banks(e.width, e.max_depth);
    if (banks == 0) fail(FWN_LIMIT_EXCEEDED, "tree deeper than the traversal stack supports");
    if (!dev && !stats && M == 1 && nQ > kPipeChunk) {
        run_queries_pipelined(e, Q, nQ, omega, inside, banks);
        return;
    }
    if (!dev && !stats && M == 1 && nQ <= kSmallBatch) {
        run_points(e, Q, nQ, omega, inside, banks);
        return;
    }

    // batches of 64 points, numbered over the meshes in decreasing triangle count so that a
    // persistent launch starts the deepest trees first
    std::vector<int32_t> order(M);
    for (int m = 0; m < M; ++m) order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return e.h_f_off[a + 1] - e.h_f_off[a] > e.h_f_off[b + 1] - e.h_f_off[b];
    });
    // points per wave: 64, or fewer for small calls so that their waves spread over the device
    // (a point's result does not depend on which points share its wave)
    const bool small = nQ <= kSmallBatch && !stats;
    const int bpts = small ? (int)std::min<int64_t>(64, std::max<int64_t>(1, (nQ + 1023) / 1024)) : 64;
    std::vector<int64_t> boff(M + 1, 0);
    for (int k = 0; k < M; ++k) boff[k + 1] = boff[k] + (qoff[order[k] + 1] - qoff[order[k]] + bpts - 1) / bpts;
    const int64_t nb = boff[M];

    const bool stage_io = small && !dev; // points in and results out through the pinned stage
    // metadata (and the points of a small host batch) in one staged copy:
    // q_off[M+1] (int64), b_off[M+1] (int64), mesh order[M] (int32), then the points at 16 B
    const size_t meta_b = (sizeof(int64_t) * 2 * (M + 1) + sizeof(int32_t) * M + 15) & ~(size_t)15;
    const size_t in_b = meta_b + (stage_io ? sizeof(float) * 3 * (size_t)nQ : 0);
    const size_t out_b = stage_io ? 16 + (sizeof(float) + 1) * (size_t)nQ : 0; // flags, omega, inside
    char* hs = host_stage(e, std::max(in_b, out_b));
    std::memcpy(hs, qoff.data(), sizeof(int64_t) * (M + 1));
    std::memcpy(hs + sizeof(int64_t) * (M + 1), boff.data(), sizeof(int64_t) * (M + 1));
    std::memcpy(hs + sizeof(int64_t) * 2 * (M + 1), order.data(), sizeof(int32_t) * M);
    if (stage_io) std::memcpy(hs + meta_b, Q, sizeof(float) * 3 * (size_t)nQ);
    const void* meta_before = e.q_meta.p;
    e.q_meta.reserve(in_b);
    // repeated calls of one shape (same meshes, point counts and batch size) find the metadata
    // already on the device and skip the upload
    const bool meta_on_device = !stage_io && e.q_meta.p == meta_before && e.meta_dev.size() == meta_b &&
                                std::memcmp(e.meta_dev.data(), hs, meta_b) == 0;
    if (!meta_on_device) {
        FWN_HIP(hipMemcpyAsync(e.q_meta.p, hs, in_b, hipMemcpyHostToDevice, s));
        e.meta_dev.assign(hs, hs + meta_b);
    }
    const int64_t* d_qoff = e.q_meta.as<int64_t>();
    const int64_t* d_boff = d_qoff + (M + 1);
    const int32_t* d_order = reinterpret_cast<const int32_t*>(d_qoff + 2 * (M + 1));

    const float* dQ = Q;
    float* dOm = omega;
    uint8_t* dIn = inside;
    int32_t* dSt = stats;
    if (!dev) {
        e.q_om.reserve(sizeof(float) * nQ);
        if (stage_io) {
            dQ = reinterpret_cast<const float*>(e.q_meta.as<char>() + meta_b);
        } else {
            e.q_in.reserve(sizeof(float) * 3 * nQ);
            FWN_HIP(hipMemcpyAsync(e.q_in.p, Q, sizeof(float) * 3 * nQ, hipMemcpyHostToDevice, s));
            dQ = e.q_in.as<float>();
        }
        dOm = e.q_om.as<float>();
        if (inside) {
            e.q_in8.reserve(nQ);
            dIn = e.q_in8.as<uint8_t>();
        }
        if (stats) {
            e.q_stats.reserve(sizeof(int32_t) * 4 * nQ);
            dSt = e.q_stats.as<int32_t>();
        }
    }
    e.q_flag.reserve(3 * sizeof(int32_t));

    for (int k = 0; k < 2; ++k) {
        e.q_keys[k].reserve(sizeof(uint32_t) * nQ);
        e.q_vals[k].reserve(sizeof(uint32_t) * nQ);
    }
    if (!e.ev[0])
        for (int k = 0; k < 4; ++k) FWN_HIP(hipEventCreate(&e.ev[k]));
    // per-point FP64 accumulators (sorted positions), kept zero between calls
    if (e.acc_cap < nQ) {
        e.q_acc.alloc(sizeof(unsigned long long) * nQ);
        e.acc_cap = nQ;
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_blocks(nQ, 256)), dim3(256), 0, s,
                           e.q_acc.as<unsigned long long>(), nQ);
        FWN_LAUNCH_CHECK("k_fill_u64");
    }
    if (e.item_cap < std::max<int64_t>(nQ, 1 << 20)) {
        e.item_cap = std::min<int64_t>(std::max<int64_t>(nQ, 1 << 20), (int64_t)UINT32_MAX / 4);
        e.q_items.alloc(sizeof(uint32_t) * 3 * (size_t)e.item_cap);
    }

    FWN_HIP(hipEventRecord(e.ev[0], s));
    // sorted position -> input index (nullptr: small calls keep the input order)
    uint32_t* vp = nullptr;
    if (!small) {
        uint32_t* kp = e.q_keys[0].as<uint32_t>();
        uint32_t* kp2 = e.q_keys[1].as<uint32_t>();
        uint32_t* vp2 = e.q_vals[1].as<uint32_t>();
        vp = e.q_vals[0].as<uint32_t>();
        int mesh_bits = 0;
        while ((1 << mesh_bits) < M) ++mesh_bits;
        if (mesh_bits > 24) fail(FWN_LIMIT_EXCEEDED, "more than 2^24 meshes in one engine");
        hipLaunchKernelGGL(k_query_keys, dim3(grid_blocks(nQ, QT)), dim3(QT), 0, s, dQ, nQ, d_qoff, M, mesh_bits,
                           e.mesh_info.as<float>(), kp);
        FWN_LAUNCH_CHECK("k_query_keys");
        radix_sort_pairs32(kp, vp, kp2, vp2, nQ, 8, 32, true, e.q_hist, e.q_rowsum, s);
    }

    const int waves = e.opt.waves_per_block > 0 ? e.opt.waves_per_block : 1;
    const bool persist = (e.opt.launch_mode == FWN_LAUNCH_PERSISTENT) && !stats;
    int64_t nblk = (nb + waves - 1) / waves;
    if (persist) {
        if (e.n_cu == 0) {
            hipDeviceProp_t prop;
            FWN_HIP(hipGetDeviceProperties(&prop, e.device));
            e.n_cu = prop.multiProcessorCount;
        }
        nblk = std::min<int64_t>(nblk, (int64_t)e.n_cu * resident_blocks(true, waves, banks));
    }
    if (nblk > INT32_MAX) fail(FWN_LIMIT_EXCEEDED, "too many point batches");
    const float beta2 = e.opt.accuracy_scale * e.opt.accuracy_scale;
    int32_t* dflag = e.q_flag.as<int32_t>();
    uint32_t* dnitems = reinterpret_cast<uint32_t*>(dflag + 1); // chunks of FP64 items claimed
    uint32_t* dwork = reinterpret_cast<uint32_t*>(dflag + 2);
    for (int attempt = 0;; ++attempt) {
        hipLaunchKernelGGL(k_zero3, dim3(1), dim3(1), 0, s, dflag);
        FWN_HIP(hipEventRecord(e.ev[1], s));
        const uint32_t cap = (uint32_t)e.item_cap;
        launch_traverse(stats != nullptr, waves, persist, banks, (unsigned)nblk, s, e.slots.as<uint32_t>(), e.width,
                        e.tris.as<uint32_t>(), e.roots.as<int32_t>(), e.mesh_info.as<float>(), dQ, vp, d_qoff,
                        d_boff, d_order, M, nb, bpts, beta2, dOm, dIn, dSt, e.q_items.as<uint32_t>(), dnitems, cap,
                        dflag, dwork);
        FWN_LAUNCH_CHECK("k_traverse");
        FWN_HIP(hipEventRecord(e.ev[2], s));
        const unsigned g64 = small ? 16u : 2048u; // grid-stride loops: any grid size covers every item
        hipLaunchKernelGGL(k_near64, dim3(g64), dim3(256), 0, s, e.q_items.as<uint32_t>(), dnitems, cap,
                           e.tris.as<uint32_t>(), dQ, vp, e.q_acc.as<unsigned long long>());
        FWN_LAUNCH_CHECK("k_near64");
        hipLaunchKernelGGL(k_near64_combine, dim3(g64), dim3(256), 0, s, e.q_items.as<uint32_t>(), dnitems, cap, vp,
                           e.q_acc.as<unsigned long long>(), dOm, dIn);
        FWN_LAUNCH_CHECK("k_near64_combine");
        FWN_HIP(hipEventRecord(e.ev[3], s));
        int32_t h_fl[2] = {0, 0};
        if (stage_io) { // flags and results in one wait
            FWN_HIP(hipMemcpyAsync(hs, dflag, sizeof(h_fl), hipMemcpyDeviceToHost, s));
            FWN_HIP(hipMemcpyAsync(hs + 16, dOm, sizeof(float) * nQ, hipMemcpyDeviceToHost, s));
            if (inside) FWN_HIP(hipMemcpyAsync(hs + 16 + sizeof(float) * nQ, dIn, nQ, hipMemcpyDeviceToHost, s));
            FWN_HIP(hipStreamSynchronize(s));
            std::memcpy(h_fl, hs, sizeof(h_fl));
        } else { // through the pinned stage too (its metadata bytes were uploaded before, in order)
            FWN_HIP(hipMemcpyAsync(hs, dflag, sizeof(h_fl), hipMemcpyDeviceToHost, s));
            FWN_HIP(hipStreamSynchronize(s));
            std::memcpy(h_fl, hs, sizeof(h_fl));
        }
        if (h_fl[0]) fail(FWN_LIMIT_EXCEEDED, "traversal stack overflow");
        const int64_t need = (int64_t)(uint32_t)h_fl[1] * kItemChunk; // claimed chunks x slots
        if (need <= e.item_cap) break;
        if (attempt > 0 || need > (int64_t)UINT32_MAX / 4)
            fail(FWN_LIMIT_EXCEEDED, "FP64 near-field work list too large");
        e.item_cap = need + need / 4;
        e.q_items.alloc(sizeof(uint32_t) * 3 * (size_t)e.item_cap);
    }
    if (stage_io) {
        std::memcpy(omega, hs + 16, sizeof(float) * nQ);
        if (inside) std::memcpy(inside, hs + 16 + sizeof(float) * nQ, nQ);
    } else if (!dev) {
        FWN_HIP(hipMemcpyAsync(omega, dOm, sizeof(float) * nQ, hipMemcpyDeviceToHost, s));
        if (inside) FWN_HIP(hipMemcpyAsync(inside, dIn, nQ, hipMemcpyDeviceToHost, s));
        if (stats) FWN_HIP(hipMemcpyAsync(stats, dSt, sizeof(int32_t) * 4 * nQ, hipMemcpyDeviceToHost, s));
    }
    FWN_HIP(hipStreamSynchronize(s));
    float t_tr = 0.0f, t_all = 0.0f;
    FWN_HIP(hipEventElapsedTime(&t_tr, e.ev[1], e.ev[2]));
    FWN_HIP(hipEventElapsedTime(&t_all, e.ev[0], e.ev[3]));
    e.last_traverse_ms = t_tr;
    e.last_total_ms = t_all;
}

const uint32_t* sort_points_single(Engine& e, const float* dQ, int64_t nQ)
{
    if (e.n_meshes != 1) fail(FWN_INVALID_ARGUMENT, "a single-mesh engine is required");
    if (nQ > (int64_t)UINT32_MAX / 2) fail(FWN_LIMIT_EXCEEDED, "more than 2^31 points in one call");
    hipStream_t s = e.stream;
    for (int k = 0; k < 2; ++k) {
        e.q_keys[k].reserve(sizeof(uint32_t) * nQ);
        e.q_vals[k].reserve(sizeof(uint32_t) * nQ);
    }
    uint32_t* kp = e.q_keys[0].as<uint32_t>();
    uint32_t* vp = e.q_vals[0].as<uint32_t>();
    uint32_t* kp2 = e.q_keys[1].as<uint32_t>();
    uint32_t* vp2 = e.q_vals[1].as<uint32_t>();
    hipLaunchKernelGGL(k_query_keys, dim3(grid_blocks(nQ, QT)), dim3(QT), 0, s, dQ, nQ, (const int64_t*)nullptr, 1, 0,
                       e.mesh_info.as<float>(), kp);
    FWN_LAUNCH_CHECK("k_query_keys");
    radix_sort_pairs32(kp, vp, kp2, vp2, nQ, 8, 32, true, e.q_hist, e.q_rowsum, s);
    return vp;
}

void run_bruteforce(Engine& e, const float* Q, int64_t nQ, double* out, bool dev)
{
    hipStream_t s = e.stream;
    if (nQ <= 0) return;
    const float* dQ = Q;
    double* dOut = out;
    if (!dev) {
        e.q_in.reserve(sizeof(float) * 3 * nQ);
        e.q_bf.reserve(sizeof(double) * nQ);
        FWN_HIP(hipMemcpyAsync(e.q_in.p, Q, sizeof(float) * 3 * nQ, hipMemcpyHostToDevice, s));
        dQ = e.q_in.as<float>();
        dOut = e.q_bf.as<double>();
    }
    // mesh 0's triangles are the first f_off[1] entries of the sorted triangle array
    const int64_t cnt = e.h_f_off.size() > 1 ? e.h_f_off[1] : 0;
    hipLaunchKernelGGL(k_bruteforce, dim3(grid_blocks(nQ, 256)), dim3(256), 0, s, e.tris.as<uint32_t>(), (int64_t)0,
                       cnt, dQ, nQ, dOut);
    FWN_LAUNCH_CHECK("k_bruteforce");
    if (!dev) FWN_HIP(hipMemcpyAsync(out, dOut, sizeof(double) * nQ, hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
}

} // namespace fwn
