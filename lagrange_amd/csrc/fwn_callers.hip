// Batched versions of the reference's two winding-number examples, with every query point made
// on the device:
//
// * orientation criterion (modules/winding/examples/fix_orientation.cpp:86-115): per facet of an
//   input mesh, normal n = normalized((b - a) x (c - a)), barycentre g = (a + b + c) / 3, probes
//   p = g + eps n and q = g - eps n, criterion = (Omega(p) - Omega(q)) / (4.f * pi) where pi is a
//   double (lagrange/internal/constants.h), so the float difference is divided in double.  The
//   float arithmetic follows the reference expression by expression, with no fused multiply-add.
// * inside sampling (modules/winding/examples/sample_points_in_mesh.cpp:59-76): uniform points in
//   a box, kept when is_inside.  The reference draws them sequentially from std::mt19937 through
//   uniform_real_distribution<float> ((hi - lo) * u + lo); here u comes from a counter-based
//   generator (SplitMix64 of seed and point index, 24-bit mantissa), so the points are the same
//   distribution, not the same sequence.  Kept points are compacted in sample order.
#include <hip/hip_runtime.h>

#include "fwn_internal.h"
#include "fwn_nn.h"

namespace fwn {
namespace {

__global__ __launch_bounds__(256) void k_orient_probes(const float* __restrict__ V, int64_t nV,
                                                       const int32_t* __restrict__ F, int64_t nF, float eps,
                                                       float* __restrict__ P, int32_t* __restrict__ err)
{
#pragma clang fp contract(off)
    const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (f >= nF) return;
    const int32_t i0 = F[3 * f], i1 = F[3 * f + 1], i2 = F[3 * f + 2];
    if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= nV || i1 >= nV || i2 >= nV) {
        atomicOr(err, 1);
        for (int d = 0; d < 6; ++d) P[6 * f + d] = 0.0f;
        return;
    }
    float a[3], b[3], c[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = V[3 * (int64_t)i0 + d];
        b[d] = V[3 * (int64_t)i1 + d];
        c[d] = V[3 * (int64_t)i2 + d];
    }
    const float u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const float z = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (z > 0.0f) {
        const float s = sqrtf(z);
        for (float& x : n) x = x / s;
    }
    for (int d = 0; d < 3; ++d) {
        const float g = (a[d] + b[d] + c[d]) / 3.0f;
        const float t = eps * n[d];
        P[6 * f + d] = g + t;
        P[6 * f + 3 + d] = g - t;
    }
}

__global__ __launch_bounds__(256) void k_orient_criterion(const float* __restrict__ om, int64_t nF,
                                                          double* __restrict__ crit)
{
    const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (f >= nF) return;
    const float d = om[2 * f] - om[2 * f + 1];
    crit[f] = (double)d / (4.0 * 3.14159265358979323846);
}

__global__ __launch_bounds__(256) void k_sample_box(int64_t n, uint64_t seed, float lx, float ly, float lz, float hx,
                                                    float hy, float hz, float* __restrict__ P)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float lo[3] = {lx, ly, lz}, hi[3] = {hx, hy, hz};
    for (int d = 0; d < 3; ++d) {
        const float u = nn::rng_unit(seed, 40 + d, (uint64_t)i);
        P[3 * i + d] = (hi[d] - lo[d]) * u + lo[d];
    }
}

__global__ __launch_bounds__(256) void k_flags_to_i32(const uint8_t* __restrict__ in, int64_t n,
                                                      int32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = in[i] ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_compact_points(const float* __restrict__ P, const uint8_t* __restrict__ keep,
                                                        const int32_t* __restrict__ pos, int64_t n,
                                                        float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const int64_t j = pos[i];
    out[3 * j] = P[3 * i];
    out[3 * j + 1] = P[3 * i + 1];
    out[3 * j + 2] = P[3 * i + 2];
}

} // namespace

void orientation_criterion(Engine& e, const float* V, int64_t nV, const int32_t* F, int64_t nF, float eps,
                           double* crit, bool dev)
{
    if (nV < 0 || nF < 0) fail(FWN_INVALID_ARGUMENT, "negative size");
    if (!(eps == eps) || eps < 0.0f || eps > 3.0e38f) fail(FWN_INVALID_ARGUMENT, "epsilon must be finite and >= 0");
    if (nF == 0) return;
    if (!V || !F || !crit) fail(FWN_INVALID_ARGUMENT, "NULL array");
    if (2 * nF > (int64_t)UINT32_MAX / 2) fail(FWN_LIMIT_EXCEEDED, "more than 2^30 facets in one call");
    hipStream_t s = e.stream;
    const float* dV = V;
    const int32_t* dF = F;
    if (!dev) {
        e.c_V.reserve(sizeof(float) * 3 * (size_t)nV);
        e.c_F.reserve(sizeof(int32_t) * 3 * (size_t)nF);
        FWN_HIP(hipMemcpyAsync(e.c_V.p, V, sizeof(float) * 3 * (size_t)nV, hipMemcpyHostToDevice, s));
        FWN_HIP(hipMemcpyAsync(e.c_F.p, F, sizeof(int32_t) * 3 * (size_t)nF, hipMemcpyHostToDevice, s));
        dV = e.c_V.as<float>();
        dF = e.c_F.as<int32_t>();
    }
    e.v_pts.reserve(sizeof(float) * 6 * (size_t)nF);
    e.v_om.reserve(sizeof(float) * 2 * (size_t)nF);
    e.c_err.reserve(sizeof(int32_t));
    FWN_HIP(hipMemsetAsync(e.c_err.p, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_orient_probes, dim3(grid_blocks(nF, 256)), dim3(256), 0, s, dV, nV, dF, nF, eps,
                       e.v_pts.as<float>(), e.c_err.as<int32_t>());
    FWN_LAUNCH_CHECK("k_orient_probes");
    int32_t err = 0;
    FWN_HIP(hipMemcpyAsync(&err, e.c_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
    if (err) fail(FWN_INVALID_ARGUMENT, "facet index out of range");
    run_queries(e, e.v_pts.as<float>(), nullptr, 2 * nF, e.v_om.as<float>(), nullptr, nullptr, true);
    double* dC = crit;
    if (!dev) {
        e.c_out.reserve(sizeof(double) * (size_t)nF);
        dC = e.c_out.as<double>();
    }
    hipLaunchKernelGGL(k_orient_criterion, dim3(grid_blocks(nF, 256)), dim3(256), 0, s, e.v_om.as<float>(), nF, dC);
    FWN_LAUNCH_CHECK("k_orient_criterion");
    if (!dev) FWN_HIP(hipMemcpyAsync(crit, dC, sizeof(double) * (size_t)nF, hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
}

void sample_inside(Engine& e, const float lo[3], const float hi[3], int64_t n, uint64_t seed, float* out,
                   int64_t* n_inside, bool dev)
{
    if (n < 0) fail(FWN_INVALID_ARGUMENT, "negative sample count");
    for (int d = 0; d < 3; ++d)
        if (!(lo[d] <= hi[d]) || !(hi[d] - lo[d] <= 3.0e38f)) fail(FWN_INVALID_ARGUMENT, "bad sampling box");
    *n_inside = 0;
    if (n == 0) return;
    if (!out) fail(FWN_INVALID_ARGUMENT, "NULL output");
    if (n > (int64_t)INT32_MAX) fail(FWN_LIMIT_EXCEEDED, "more than 2^31 samples in one call");
    hipStream_t s = e.stream;
    e.v_pts.reserve(sizeof(float) * 3 * (size_t)n);
    e.v_om.reserve(sizeof(float) * (size_t)n);
    e.v_in.reserve((size_t)n);
    hipLaunchKernelGGL(k_sample_box, dim3(grid_blocks(n, 256)), dim3(256), 0, s, n, seed, lo[0], lo[1], lo[2], hi[0],
                       hi[1], hi[2], e.v_pts.as<float>());
    FWN_LAUNCH_CHECK("k_sample_box");
    run_queries(e, e.v_pts.as<float>(), nullptr, n, e.v_om.as<float>(), e.v_in.as<uint8_t>(), nullptr, true);
    e.c_pos.reserve(sizeof(int32_t) * ((size_t)n + 1));
    hipLaunchKernelGGL(k_flags_to_i32, dim3(grid_blocks(n, 256)), dim3(256), 0, s, e.v_in.as<uint8_t>(), n,
                       e.c_pos.as<int32_t>());
    FWN_LAUNCH_CHECK("k_flags_to_i32");
    exclusive_scan_i32(e.c_pos.as<int32_t>(), e.c_pos.as<int32_t>(), n, e.c_scan, s);
    int32_t total = 0;
    FWN_HIP(hipMemcpyAsync(&total, e.c_pos.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    float* dOut = out;
    if (!dev) {
        e.c_outp.reserve(sizeof(float) * 3 * (size_t)n);
        dOut = e.c_outp.as<float>();
    }
    hipLaunchKernelGGL(k_compact_points, dim3(grid_blocks(n, 256)), dim3(256), 0, s, e.v_pts.as<float>(),
                       e.v_in.as<uint8_t>(), e.c_pos.as<int32_t>(), n, dOut);
    FWN_LAUNCH_CHECK("k_compact_points");
    FWN_HIP(hipStreamSynchronize(s));
    if (!dev && total > 0)
        FWN_HIP(hipMemcpy(out, dOut, sizeof(float) * 3 * (size_t)total, hipMemcpyDeviceToHost));
    *n_inside = total;
}

} // namespace fwn
