// Query entry point that allocates nothing (no counterpart in the reference).
//
// fwn_solid_angle_noalloc answers the same query as fwn_solid_angle with FWN_DEVICE_POINTERS on a
// single-mesh engine, bit for bit, for callers that cannot allocate device memory during a call.
// Its scratch lives in one file-scope device array of fixed size; the host reaches each device's
// copy through hipGetSymbolAddress, taken once per device with that device current.  The same
// kernels as the allocating path run: point keys, the 24-bit radix sort (histograms in the
// array), and the traversal with the FP64 near leaves evaluated in place (no work list), which
// gives the deferred pass's value exactly.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fwn_internal.h"

namespace fwn {
namespace {

constexpr int64_t kMaxPoints = FWN_NOALLOC_MAX_POINTS;
constexpr int64_t kHistWords = 256 * ((kMaxPoints + 4095) / 4096); // 256 digits x sort tiles of 4,096 keys
// byte offsets in the array: keys and order (twice, for the sort's ping-pong), sort histograms and
// digit sums, call metadata (q_off, b_off, mesh order), flags
constexpr size_t kOffKeys0 = 0;
constexpr size_t kOffKeys1 = kOffKeys0 + 4 * (size_t)kMaxPoints;
constexpr size_t kOffVals0 = kOffKeys1 + 4 * (size_t)kMaxPoints;
constexpr size_t kOffVals1 = kOffVals0 + 4 * (size_t)kMaxPoints;
constexpr size_t kOffHist = kOffVals1 + 4 * (size_t)kMaxPoints;
constexpr size_t kOffRows = kOffHist + 4 * (size_t)kHistWords;
constexpr size_t kOffMeta = kOffRows + 4 * 256;
constexpr size_t kOffFlag = kOffMeta + 8 * 8;
constexpr size_t kScratchBytes = kOffFlag + 16;

} // namespace

// The scratch.  In host code this name is a stand-in, not a device address: the host uses the
// per-device pointer that hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;            // one call at a time in the process: calls share each GPU's array
char* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device ordinal

// the calling device's copy of the array (the device is current)
char* device_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) fail(FWN_INVALID_ARGUMENT, "device ordinal out of range");
    if (!g_scratch[device]) {
        void* p = nullptr;
        FWN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(FWN_HIP_ERROR, "no device address for the query scratch");
        g_scratch[device] = static_cast<char*>(p);
    }
    return g_scratch[device];
}

// call metadata of one single-mesh query and a cleared flag (a kernel, so that the call makes no
// host-to-device copy)
__global__ void k_noalloc_meta(int64_t* __restrict__ meta, int32_t* __restrict__ flag, int64_t nQ)
{
    if (threadIdx.x != 0) return;
    meta[0] = 0;             // q_off
    meta[1] = nQ;
    meta[2] = 0;             // b_off
    meta[3] = (nQ + 63) / 64;
    meta[4] = 0;             // mesh order {0}
    flag[0] = 0;
}

} // namespace

void run_queries_noalloc(Engine& e, const float* dQ, int64_t nQ, float* dOm, uint8_t* dIn)
{
    if (e.n_meshes != 1) fail(FWN_INVALID_ARGUMENT, "a single-mesh engine is required");
    if (nQ < 0) fail(FWN_INVALID_ARGUMENT, "negative point count");
    if (nQ > kMaxPoints)
        fail(FWN_LIMIT_EXCEEDED, "more than FWN_NOALLOC_MAX_POINTS (2^20) points in one no-allocation call");
    if (nQ == 0) return;
    if (radix_sort_hist_words(kMaxPoints) > kHistWords) fail(FWN_HIP_ERROR, "sort scratch too small");
    const int banks = stack_banks(e.width, e.max_depth);
    if (banks == 0) fail(FWN_LIMIT_EXCEEDED, "tree deeper than the traversal stack supports");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    char* base = device_scratch(e.device);
    hipStream_t s = e.stream;
    uint32_t* kp = reinterpret_cast<uint32_t*>(base + kOffKeys0);
    uint32_t* kp2 = reinterpret_cast<uint32_t*>(base + kOffKeys1);
    uint32_t* vp = reinterpret_cast<uint32_t*>(base + kOffVals0);
    uint32_t* vp2 = reinterpret_cast<uint32_t*>(base + kOffVals1);
    int64_t* meta = reinterpret_cast<int64_t*>(base + kOffMeta);
    int32_t* flag = reinterpret_cast<int32_t*>(base + kOffFlag);
    hipLaunchKernelGGL(k_noalloc_meta, dim3(1), dim3(64), 0, s, meta, flag, nQ);
    FWN_LAUNCH_CHECK("k_noalloc_meta");
    launch_point_keys(e, dQ, nQ, kp, s);
    radix_sort_pairs32_scratch(kp, vp, kp2, vp2, nQ, 8, 32, true, reinterpret_cast<int32_t*>(base + kOffHist),
                               reinterpret_cast<int32_t*>(base + kOffRows), s);
    launch_traverse_inplace(e, banks, dQ, vp, meta, nQ, dOm, dIn, flag, s);
    int32_t h_flag = 0;
    FWN_HIP(hipMemcpyAsync(&h_flag, flag, sizeof(h_flag), hipMemcpyDeviceToHost, s));
    FWN_HIP(hipStreamSynchronize(s));
    if (h_flag) fail(FWN_LIMIT_EXCEEDED, "traversal stack overflow");
}

} // namespace fwn
