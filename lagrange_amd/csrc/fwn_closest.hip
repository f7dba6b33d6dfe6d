// Closest point on the mesh (unsigned distance) over the same tree as the winding-number
// engine: the query of lagrange::bvh::TriangleAABBTree::get_closest_point
// (modules/bvh/src/TriangleAABBTree.cpp:86-127), whose tree walk is AABB::get_closest_element
// (modules/bvh/src/AABB.cpp:187-250) and whose per-triangle distance is
// lagrange::point_triangle_squared_distance
// (modules/core/include/lagrange/utils/point_triangle_squared_distance.h), Eberly's region
// method evaluated in double.
//
// One wave64 walks the tree for 64 Morton-sorted points.  The wave stack holds (node, lane mask)
// entries as in the winding-number traversal; a lane takes part in a slot while the slot's box is
// nearer than the lane's best distance so far.  Leaf slots are evaluated before a node's children
// are pushed (to tighten the bounds early) and children are pushed farthest first, ordered by the
// box distance of each slot's first participating lane, so the nearest is walked next.  Inside a
// leaf, a triangle is evaluated only for lanes whose best distance its vertex box does not
// exceed.  Slot boxes come from the build (each slot's vertex box, rounded outward to float).
#include <hip/hip_runtime.h>

#include <cmath>

#include "fwn_internal.h"

namespace fwn {
namespace {

// Squared distance from q to the triangle (v0, v1, v2) and the parameters (s, t) of the closest
// point v0 + s (v1 - v0) + t (v2 - v0); every operation rounded as written (no contraction).
__device__ double tri_closest_f64(const float* __restrict__ tw, double qx, double qy, double qz, double& so,
                                  double& to)
{
#pragma clang fp contract(off)
    const double v0[3] = {tw[0], tw[1], tw[2]};
    const double v1[3] = {tw[3], tw[4], tw[5]};
    const double v2[3] = {tw[6], tw[7], tw[8]};
    const double q[3] = {qx, qy, qz};
    double df[3], e0[3], e1[3];
    for (int d = 0; d < 3; ++d) {
        df[d] = v0[d] - q[d];
        e0[d] = v1[d] - v0[d];
        e1[d] = v2[d] - v0[d];
    }
    const double a00 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    const double a01 = e0[0] * e1[0] + e0[1] * e1[1] + e0[2] * e1[2];
    const double a11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const double b0 = df[0] * e0[0] + df[1] * e0[1] + df[2] * e0[2];
    const double b1 = df[0] * e1[0] + df[1] * e1[1] + df[2] * e1[2];
    const double c = df[0] * df[0] + df[1] * df[1] + df[2] * df[2];
    const double det = fabs(a00 * a11 - a01 * a01);
    double s = a01 * b1 - a11 * b0;
    double t = a01 * b0 - a00 * b1;
    double r;
    if (det < 1e-30) {
        // degenerate triangle: nearest of the three edges (v0 v1), (v0 v2), (v1 v2)
        const double* P[3][2] = {{v0, v1}, {v0, v2}, {v1, v2}};
        r = INFINITY;
        for (int k = 0; k < 3; ++k) {
            const double* A = P[k][0];
            const double* B = P[k][1];
            double ab[3], aq[3];
            for (int d = 0; d < 3; ++d) {
                ab[d] = B[d] - A[d];
                aq[d] = q[d] - A[d];
            }
            const double l2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
            const double tp = aq[0] * ab[0] + aq[1] * ab[1] + aq[2] * ab[2];
            double u = 0.0;
            if (!(tp <= 0.0 || l2 == 0.0)) u = tp > l2 ? 1.0 : tp / l2;
            double dd = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double x = q[d] - (A[d] + u * ab[d]);
                dd += x * x;
            }
            if (dd < r) {
                r = dd;
                // express the point in (s, t) of the triangle: edge k interpolates its endpoints
                so = k == 0 ? u : (k == 1 ? 0.0 : 1.0 - u);
                to = k == 0 ? 0.0 : (k == 1 ? u : u);
            }
        }
        return r;
    }
    if (s + t <= det) {
        if (s < 0.0) {
            if (t < 0.0) { // region 4
                if (b0 < 0.0) {
                    t = 0.0;
                    if (-b0 >= a00) {
                        s = 1.0;
                        r = a00 + 2.0 * b0 + c;
                    } else {
                        s = -b0 / a00;
                        r = b0 * s + c;
                    }
                } else {
                    s = 0.0;
                    if (b1 >= 0.0) {
                        t = 0.0;
                        r = c;
                    } else if (-b1 >= a11) {
                        t = 1.0;
                        r = a11 + 2.0 * b1 + c;
                    } else {
                        t = -b1 / a11;
                        r = b1 * t + c;
                    }
                }
            } else { // region 3: edge s = 0
                s = 0.0;
                if (b1 >= 0.0) {
                    t = 0.0;
                    r = c;
                } else if (-b1 >= a11) {
                    t = 1.0;
                    r = a11 + 2.0 * b1 + c;
                } else {
                    t = -b1 / a11;
                    r = b1 * t + c;
                }
            }
        } else if (t < 0.0) { // region 5: edge t = 0
            t = 0.0;
            if (b0 >= 0.0) {
                s = 0.0;
                r = c;
            } else if (-b0 >= a00) {
                s = 1.0;
                r = a00 + 2.0 * b0 + c;
            } else {
                s = -b0 / a00;
                r = b0 * s + c;
            }
        } else { // region 0: interior
            const double inv = 1.0 / det;
            s = s * inv;
            t = t * inv;
            r = s * (a00 * s + a01 * t + 2.0 * b0) + t * (a01 * s + a11 * t + 2.0 * b1) + c;
        }
    } else {
        if (s < 0.0) { // region 2
            const double tmp0 = a01 + b0, tmp1 = a11 + b1;
            if (tmp1 > tmp0) {
                const double numer = tmp1 - tmp0, denom = a00 - 2.0 * a01 + a11;
                if (numer >= denom) {
                    s = 1.0;
                    t = 0.0;
                    r = a00 + 2.0 * b0 + c;
                } else {
                    s = numer / denom;
                    t = 1.0 - s;
                    r = s * (a00 * s + a01 * t + 2.0 * b0) + t * (a01 * s + a11 * t + 2.0 * b1) + c;
                }
            } else {
                s = 0.0;
                if (tmp1 <= 0.0) {
                    t = 1.0;
                    r = a11 + 2.0 * b1 + c;
                } else if (b1 >= 0.0) {
                    t = 0.0;
                    r = c;
                } else {
                    t = -b1 / a11;
                    r = b1 * t + c;
                }
            }
        } else if (t < 0.0) { // region 6
            const double tmp0 = a01 + b1, tmp1 = a00 + b0;
            if (tmp1 > tmp0) {
                const double numer = tmp1 - tmp0, denom = a00 - 2.0 * a01 + a11;
                if (numer >= denom) {
                    t = 1.0;
                    s = 0.0;
                    r = a11 + 2.0 * b1 + c;
                } else {
                    t = numer / denom;
                    s = 1.0 - t;
                    r = s * (a00 * s + a01 * t + 2.0 * b0) + t * (a01 * s + a11 * t + 2.0 * b1) + c;
                }
            } else {
                t = 0.0;
                if (tmp1 <= 0.0) {
                    s = 1.0;
                    r = a00 + 2.0 * b0 + c;
                } else if (b0 >= 0.0) {
                    s = 0.0;
                    r = c;
                } else {
                    s = -b0 / a00;
                    r = b0 * s + c;
                }
            }
        } else { // region 1: edge s + t = 1
            const double numer = a11 + b1 - a01 - b0;
            if (numer <= 0.0) {
                s = 0.0;
                t = 1.0;
                r = a11 + 2.0 * b1 + c;
            } else {
                const double denom = a00 - 2.0 * a01 + a11;
                if (numer >= denom) {
                    s = 1.0;
                    t = 0.0;
                    r = a00 + 2.0 * b0 + c;
                } else {
                    s = numer / denom;
                    t = 1.0 - s;
                    r = s * (a00 * s + a01 * t + 2.0 * b0) + t * (a01 * s + a11 * t + 2.0 * b1) + c;
                }
            }
        }
    }
    so = s;
    to = t;
    return r < 0.0 ? 0.0 : r;
}

// v0 + s (v1 - v0) + t (v2 - v0) in double, rounded as written
__device__ __forceinline__ void closest_on(const float* __restrict__ tw, double s, double t, double* cp)
{
#pragma clang fp contract(off)
    for (int d = 0; d < 3; ++d) {
        const double v0 = tw[d];
        cp[d] = v0 + s * ((double)tw[3 + d] - v0) + t * ((double)tw[6 + d] - v0);
    }
}

// squared distance from q to the box (lo[3], hi[3]); 0 inside
__device__ __forceinline__ double box_dist2(const float* __restrict__ b, double qx, double qy, double qz)
{
    const double q[3] = {qx, qy, qz};
    double r = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double lo = b[d], hi = b[3 + d];
        const double x = q[d] < lo ? lo - q[d] : (q[d] > hi ? q[d] - hi : 0.0);
        r = fma(x, x, r);
    }
    return r;
}

// Squared distance from q to the box of the triangle's vertices: a lower bound on the squared
// distance to the triangle
__device__ __forceinline__ double tri_box_dist2(const float* __restrict__ tw, double qx, double qy, double qz)
{
    float b[6];
    for (int d = 0; d < 3; ++d) {
        b[d] = fminf(fminf(tw[d], tw[3 + d]), tw[6 + d]);
        b[3 + d] = fmaxf(fmaxf(tw[d], tw[3 + d]), tw[6 + d]);
    }
    return box_dist2(b, qx, qy, qz);
}

template <int NB>
__global__ __launch_bounds__(64) void k_closest(const uint32_t* __restrict__ slots, const float* __restrict__ boxes,
                                                int width, const uint32_t* __restrict__ tris, int32_t root,
                                                int32_t root_ns, const float* __restrict__ Q,
                                                const uint32_t* __restrict__ perm, int64_t nQ,
                                                double* __restrict__ d2_out, int32_t* __restrict__ tri_out,
                                                double* __restrict__ cp_out, int32_t* __restrict__ flag)
{
    const int lane = threadIdx.x;
    const int64_t batch = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t jpos = batch * 64 + lane;
    const bool valid = jpos < nQ;
    const uint32_t orig = valid ? perm[jpos] : 0u;
    const double qx = valid ? (double)Q[3 * (int64_t)orig] : 0.0;
    const double qy = valid ? (double)Q[3 * (int64_t)orig + 1] : 0.0;
    const double qz = valid ? (double)Q[3 * (int64_t)orig + 2] : 0.0;
    double best = INFINITY, bs = 0.0, bt = 0.0;
    int32_t btri = -1;
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
            const int64_t node = e >> 4;
            const int ns = (int)(e & 15u);
            const uint32_t* nb = slots + (size_t)node * (size_t)width * kSlotWords;
            const float* nbx = boxes + (size_t)node * (size_t)width * 6;
            double db[kWidth];
            uint32_t link[kWidth]; // leaf slots have the top bit set
#pragma unroll
            for (int s = 0; s < kWidth; ++s) {
                db[s] = INFINITY;
                link[s] = 0u;
                if (s < ns) {
                    db[s] = box_dist2(nbx + 6 * s, qx, qy, qz);
                    link[s] = nb[s * kSlotWords + FWN_W_LINK];
                }
            }
            // leaf slots first
#pragma unroll
            for (int s = 0; s < kWidth; ++s) {
                if (s < ns && (link[s] & FWN_LINK_LEAF) != 0u) {
                    const bool go = act && db[s] < best;
                    if (__ballot(go) != 0ull) {
                        const int32_t first = (int32_t)((link[s] & ~FWN_LINK_LEAF) >> 3);
                        const int32_t cnt = (int32_t)(link[s] & 7u) + 1;
                        for (int32_t t = first; t < first + cnt; ++t) {
                            const float* tw = reinterpret_cast<const float*>(tris + (size_t)t * kTriWords);
                            // a triangle whose box is farther than the lane's best cannot improve it;
                            // the 1e-12 margin is far above the rounding of the bound and of the
                            // exact evaluation, so the nearest triangle (and the first of equally
                            // near ones) is always evaluated and results do not change
                            if (go && tri_box_dist2(tw, qx, qy, qz) <= best * (1.0 + 1e-12)) {
                                double ss, tt;
                                const double d = tri_closest_f64(tw, qx, qy, qz, ss, tt);
                                if (d < best) {
                                    best = d;
                                    btri = t;
                                    bs = ss;
                                    bt = tt;
                                }
                            }
                        }
                    }
                }
            }
            // inner slots: push farthest first so the nearest is walked next
            uint64_t pm[kWidth];
            float key[kWidth];
#pragma unroll
            for (int s = 0; s < kWidth; ++s) {
                const bool go = s < ns && (link[s] & FWN_LINK_LEAF) == 0u && act && db[s] < best;
                pm[s] = __ballot(go);
                // ordering key: the box distance seen by the slot's first participating lane (the
                // wave's points are close together; the order only steers the walk)
                key[s] = pm[s] ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int((float)db[s]),
                                                                          (int)__builtin_ctzll(pm[s])))
                               : -1.0f;
            }
            for (int k = 0; k < kWidth; ++k) {
                int sel = -1;
                float kmax = -1.0f;
#pragma unroll
                for (int s = 0; s < kWidth; ++s)
                    if (pm[s] != 0ull && key[s] >= kmax) {
                        kmax = key[s];
                        sel = s;
                    }
                if (sel < 0) break;
#pragma unroll
                for (int s = 0; s < kWidth; ++s) {
                    if (s == sel) {
                        if (sp < WaveStack<NB>::kCap) {
                            S.push(sp, lane, link[s], pm[s]); // (child << 4) | its slot count
                            ++sp;
                        } else {
                            overflow = true;
                        }
                        pm[s] = 0ull;
                    }
                }
            }
        }
        if (overflow && lane == 0) atomicOr(flag, 1);
    }
    if (valid) {
        d2_out[orig] = best;
        int32_t fid = -1;
        double cp[3] = {NAN, NAN, NAN};
        if (btri >= 0) {
            fid = (int32_t)tris[(size_t)btri * kTriWords + FWN_T_FACET];
            closest_on(reinterpret_cast<const float*>(tris + (size_t)btri * kTriWords), bs, bt, cp);
        }
        tri_out[orig] = fid;
        if (cp_out) {
            cp_out[3 * (int64_t)orig] = cp[0];
            cp_out[3 * (int64_t)orig + 1] = cp[1];
            cp_out[3 * (int64_t)orig + 2] = cp[2];
        }
    }
}

} // namespace

void run_closest(Engine& e, const float* Q, int64_t nQ, double* d2, int32_t* tri, double* cp, bool dev)
{
    if (e.n_meshes != 1) fail(FWN_INVALID_ARGUMENT, "closest point needs a single-mesh engine");
    if (nQ < 0) fail(FWN_INVALID_ARGUMENT, "negative point count");
    if (nQ == 0) return;
    if (!Q || !d2 || !tri) fail(FWN_INVALID_ARGUMENT, "NULL points or output");
    const int banks = stack_banks(e.width, e.max_depth);
    if (banks == 0) fail(FWN_LIMIT_EXCEEDED, "tree deeper than the traversal stack supports");
    hipStream_t s = e.stream;
    const float* dQ = Q;
    double* dD = d2;
    int32_t* dT = tri;
    double* dC = cp;
    if (!dev) {
        e.q_in.reserve(sizeof(float) * 3 * nQ);
        e.d_d2.reserve(sizeof(double) * nQ);
        e.d_tri.reserve(sizeof(int32_t) * nQ);
        FWN_HIP(hipMemcpyAsync(e.q_in.p, Q, sizeof(float) * 3 * nQ, hipMemcpyHostToDevice, s));
        dQ = e.q_in.as<float>();
        dD = e.d_d2.as<double>();
        dT = e.d_tri.as<int32_t>();
        if (cp) {
            e.d_cp.reserve(sizeof(double) * 3 * nQ);
            dC = e.d_cp.as<double>();
        }
    }
    const uint32_t* perm = sort_points_single(e, dQ, nQ);
    e.q_flag.reserve(3 * sizeof(int32_t));
    int32_t* dflag = e.q_flag.as<int32_t>();
    FWN_HIP(hipMemsetAsync(dflag, 0, sizeof(int32_t), s));
    const int64_t nb = (nQ + 63) / 64;
    if (nb > INT32_MAX) fail(FWN_LIMIT_EXCEEDED, "too many point batches");
    if (banks == 1)
        hipLaunchKernelGGL(k_closest<1>, dim3((unsigned)nb), dim3(64), 0, s, e.slots.as<uint32_t>(),
                           e.boxes.as<float>(), e.width, e.tris.as<uint32_t>(), e.h_roots[0], e.h_roots[1], dQ, perm,
                           nQ, dD, dT, dC, dflag);
    else
        hipLaunchKernelGGL(k_closest<3>, dim3((unsigned)nb), dim3(64), 0, s, e.slots.as<uint32_t>(),
                           e.boxes.as<float>(), e.width, e.tris.as<uint32_t>(), e.h_roots[0], e.h_roots[1], dQ, perm,
                           nQ, dD, dT, dC, dflag);
    FWN_LAUNCH_CHECK("k_closest");
    int32_t h_flag = 0;
    FWN_HIP(hipMemcpyAsync(&h_flag, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (!dev) {
        FWN_HIP(hipMemcpyAsync(d2, dD, sizeof(double) * nQ, hipMemcpyDeviceToHost, s));
        FWN_HIP(hipMemcpyAsync(tri, dT, sizeof(int32_t) * nQ, hipMemcpyDeviceToHost, s));
        if (cp) FWN_HIP(hipMemcpyAsync(cp, dC, sizeof(double) * 3 * nQ, hipMemcpyDeviceToHost, s));
    }
    FWN_HIP(hipStreamSynchronize(s));
    if (h_flag) fail(FWN_LIMIT_EXCEEDED, "traversal stack overflow");
}

} // namespace fwn
