// Winding-number sign of voxels: the interior test that lagrange::volume::mesh_to_volume runs
// per voxel with Sign::WindingNumber (modules/volume/src/mesh_to_volume.cpp:156-183), batched.
//
// The reference maps voxel (i, j, k) to world space with OpenVDB's linear transform of scale
// voxel_size post-translated by voxel_size / 2 (mesh_to_volume.cpp:147-149), in double, then
// rounds each coordinate to float (:175-181) and calls is_inside.  Here the voxel centres are
// generated on the device with the same double arithmetic (a multiply and an add, each rounded,
// no fused multiply-add), rounded to float, and answered by the query path of the engine.  The
// caller never builds a host array of points.
#include <hip/hip_runtime.h>

#include <cstring>

#include "fwn_internal.h"

namespace fwn {
namespace {

__device__ __forceinline__ float voxel_coord(int64_t idx, double vs, double half)
{
#pragma clang fp contract(off)
    const double w = (double)idx * vs;
    return (float)(w + half);
}

// dense box: voxel t = x + dx (y + dy z) has index lo + (x, y, z)
__global__ __launch_bounds__(256) void k_voxel_box(int32_t lx, int32_t ly, int32_t lz, int64_t dx, int64_t dy,
                                                   int64_t n, double vs, float* __restrict__ P)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const int64_t x = t % dx, r = t / dx, y = r % dy, z = r / dy;
    const double half = vs * 0.5;
    P[3 * t + 0] = voxel_coord(lx + x, vs, half);
    P[3 * t + 1] = voxel_coord(ly + y, vs, half);
    P[3 * t + 2] = voxel_coord(lz + z, vs, half);
}

// explicit list of (i, j, k) int32 triples
__global__ __launch_bounds__(256) void k_voxel_list(const int32_t* __restrict__ ijk, int64_t n, double vs,
                                                    float* __restrict__ P)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= 3 * n) return;
    P[t] = voxel_coord(ijk[t], vs, vs * 0.5);
}

} // namespace

// Generate the centres into the engine's scratch and answer them; outputs host or device.
static void sign_centres(Engine& e, int64_t n, uint8_t* inside, float* omega, bool dev,
                         void (*gen)(Engine&, float*, const void*), const void* arg)
{
    hipStream_t s = e.stream;
    e.v_pts.reserve(sizeof(float) * 3 * n);
    gen(e, e.v_pts.as<float>(), arg);
    float* dOm = omega;
    uint8_t* dIn = inside;
    if (!dev || !omega) {
        e.v_om.reserve(sizeof(float) * n);
        dOm = e.v_om.as<float>();
    }
    if (!dev) {
        e.v_in.reserve(n);
        dIn = e.v_in.as<uint8_t>();
    }
    run_queries(e, e.v_pts.as<float>(), nullptr, n, dOm, dIn, nullptr, true);
    if (!dev) {
        FWN_HIP(hipMemcpyAsync(inside, dIn, n, hipMemcpyDeviceToHost, s));
        if (omega) FWN_HIP(hipMemcpyAsync(omega, dOm, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    }
    FWN_HIP(hipStreamSynchronize(s));
}

struct BoxArg {
    int32_t lo[3];
    int64_t dims[3];
    int64_t n;
    double vs;
};
struct ListArg {
    const int32_t* ijk;
    int64_t n;
    double vs;
    bool dev;
};

static void gen_box(Engine& e, float* P, const void* a)
{
    const BoxArg& b = *static_cast<const BoxArg*>(a);
    hipLaunchKernelGGL(k_voxel_box, dim3(grid_blocks(b.n, 256)), dim3(256), 0, e.stream, b.lo[0], b.lo[1], b.lo[2],
                       b.dims[0], b.dims[1], b.n, b.vs, P);
    FWN_LAUNCH_CHECK("k_voxel_box");
}

static void gen_list(Engine& e, float* P, const void* a)
{
    const ListArg& l = *static_cast<const ListArg*>(a);
    const int32_t* dijk = l.ijk;
    if (!l.dev) {
        e.v_ijk.reserve(sizeof(int32_t) * 3 * l.n);
        FWN_HIP(hipMemcpyAsync(e.v_ijk.p, l.ijk, sizeof(int32_t) * 3 * l.n, hipMemcpyHostToDevice, e.stream));
        dijk = e.v_ijk.as<int32_t>();
    }
    hipLaunchKernelGGL(k_voxel_list, dim3(grid_blocks(3 * l.n, 256)), dim3(256), 0, e.stream, dijk, l.n, l.vs, P);
    FWN_LAUNCH_CHECK("k_voxel_list");
}

static void check_voxel_size(double vs)
{
    if (!(vs > 0.0) || !(vs < 1e300)) fail(FWN_INVALID_ARGUMENT, "voxel_size must be a positive finite number");
}

void sign_voxel_box(Engine& e, const int32_t lo[3], const int32_t dims[3], double vs, uint8_t* inside, float* omega,
                    bool dev)
{
    check_voxel_size(vs);
    BoxArg b{};
    b.n = 1;
    for (int d = 0; d < 3; ++d) {
        if (dims[d] < 0) fail(FWN_INVALID_ARGUMENT, "negative box dimension");
        if ((int64_t)lo[d] + dims[d] - 1 > INT32_MAX) fail(FWN_INVALID_ARGUMENT, "voxel index beyond int32");
        b.lo[d] = lo[d];
        b.dims[d] = dims[d];
        b.n *= dims[d];
        if (b.n > (int64_t)UINT32_MAX / 2) fail(FWN_LIMIT_EXCEEDED, "more than 2^31 voxels in one call");
    }
    b.vs = vs;
    if (b.n == 0) return;
    if (!inside) fail(FWN_INVALID_ARGUMENT, "inside_out is NULL");
    sign_centres(e, b.n, inside, omega, dev, gen_box, &b);
}

void sign_voxel_list(Engine& e, const int32_t* ijk, int64_t n, double vs, uint8_t* inside, float* omega, bool dev)
{
    check_voxel_size(vs);
    if (n < 0) fail(FWN_INVALID_ARGUMENT, "negative voxel count");
    if (n == 0) return;
    if (!ijk || !inside) fail(FWN_INVALID_ARGUMENT, "NULL voxel list or inside_out");
    if (n > (int64_t)UINT32_MAX / 2) fail(FWN_LIMIT_EXCEEDED, "more than 2^31 voxels in one call");
    ListArg l{ijk, n, vs, dev};
    sign_centres(e, n, inside, omega, dev, gen_list, &l);
}

} // namespace fwn
