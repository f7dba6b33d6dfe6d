// Internal declarations of the MI355X fast winding number engine (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <atomic>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/fwn.h"

namespace fwn {

// ---------------------------------------------------------------------------------------
// error plumbing: C-ABI functions never throw; internal code throws fwn::Failure which the
// ABI layer turns into a status code + thread-local message.
// ---------------------------------------------------------------------------------------
struct Failure {
    int status;
    std::string msg;
};

void set_error(const std::string& msg);
[[noreturn]] void fail(int status, const std::string& msg);
void hip_check(hipError_t e, const char* what, const char* file, int line);

#define FWN_HIP(call) ::fwn::hip_check((call), #call, __FILE__, __LINE__)
#define FWN_LAUNCH_CHECK(name) ::fwn::hip_check(hipGetLastError(), name, __FILE__, __LINE__)

// ---------------------------------------------------------------------------------------
// constants shared by host and device code
// ---------------------------------------------------------------------------------------
constexpr int kWidth = FWN_WIDTH;          // maximum slots per node
constexpr int kSlotWords = FWN_SLOT_WORDS; // 32-bit words per slot
constexpr int kTriWords = FWN_TRI_WORDS;   // 32-bit words per stored triangle
constexpr int kRecWords = 32;              // doubles per binary-node moment record
constexpr int kStack = 192;                // per-wave traversal stack entries
constexpr double kFixScale = 1099511627776.0; // 2^40: fixed-point scale of FP64 near-field sums
constexpr int kWave = 64;

// binary-node moment record layout (doubles)
enum : int { R_A = 0, R_P = 1, R_N = 4, R_M1 = 7, R_S = 16, R_LO = 26, R_HI = 29 };

// RAII device buffer
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release();
    void alloc(size_t n);          // exact (re)allocation, contents undefined
    void reserve(size_t n);        // grow only
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// one single-point query waiting to be answered in a coalesced batch
struct PointRequest {
    float q[3];
    float omega = 0.0f;
    uint8_t inside = 0;
    int status = 0;
    std::string err;
    std::atomic<bool> done{false}; // set (release) after the fields above, under the engine's pt_mu
};

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    // concurrent single-point calls are answered together: a caller that finds no batch running
    // takes every pending request and answers them in one query
    std::mutex pt_mu;
    std::condition_variable pt_cv;
    std::vector<PointRequest*> pt_pending;
    bool pt_busy = false;
    size_t pt_conc = 1; // requests the last coalesced batch held (how many callers to expect)
    fwn_options opt{};

    int32_t n_meshes = 0;
    int32_t width = 4;         // slots per node
    int64_t n_tris = 0;
    int64_t n_nodes = 0;
    int32_t max_depth = 0;
    double build_ms = 0.0;

    // tree (resident)
    DBuf slots;      // n_nodes * width * kSlotWords u32
    DBuf boxes;      // n_nodes * width * 6 floats: each slot's box (lo, hi), for distance queries
    DBuf tris;       // n_tris * kTriWords u32
    DBuf roots;      // 2 int32 per mesh: wide root, occupied slots
    DBuf mesh_info;  // per mesh: 8 floats: qbox lo[3], qbox inv[3], tau, pad
    std::vector<int32_t> h_roots;
    std::vector<float> h_tau;
    std::vector<int64_t> h_f_off;

    // query scratch (grown on demand)
    DBuf q_keys[2], q_vals[2], q_hist, q_rowsum, q_in, q_om, q_in8, q_stats, q_flag;
    DBuf q_bf;
    DBuf v_pts, v_om, v_in, v_ijk; // query points generated on the device and their outputs
    DBuf c_V, c_F, c_err, c_out, c_pos, c_scan, c_outp; // orientation / sampling scratch
    DBuf d_d2, d_tri, d_cp;                               // closest-point outputs for host callers
    int n_cu = 0;
    // deferred FP64 near-field work: (sorted point, first triangle, count) items and per-point
    // fixed-point accumulators (zero between calls)
    DBuf q_items, q_acc;
    int64_t item_cap = 0;
    int64_t acc_cap = 0;

    // pinned host staging for small calls: metadata, points and results in one copy each way
    // (reused only after the call that filled it has synchronised its stream)
    void* h_stage = nullptr;
    size_t h_stage_bytes = 0;
    DBuf q_meta; // device copy of the staged metadata (and small host batches' points)
    std::vector<char> meta_dev; // the metadata bytes q_meta holds (compared before an upload)
    // pipelined host calls: pinned double buffers (points in, results out), a stream per copy
    // direction, per-chunk flags
    void* h_pipe = nullptr;
    size_t h_pipe_bytes = 0;
    hipStream_t cs_in = nullptr, cs_out = nullptr;
    hipEvent_t pev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}; // h2d[2], comp[2], d2h[2]
    DBuf q_pmeta, q_pflags;

    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> tev; // pipelined calls: a (start, end) pair around each chunk's traversal
    double last_traverse_ms = 0.0;
    double last_total_ms = 0.0;
};

// sorting / scanning primitives (fwn_sort.hip)
// Stable LSD radix sort of (keys, vals) pairs by key bits [0, nbits).  Result ends in keys/vals
// (swapping with the scratch buffers as needed; the pointers are updated).
void radix_sort_pairs(uint64_t*& keys, uint32_t*& vals, uint64_t*& keys_alt, uint32_t*& vals_alt, int64_t n,
                      int nbits, DBuf& hist, DBuf& rowsum, hipStream_t s);
// Same for 32-bit keys, sorting key bits [begin_bit, end_bit) only (low bits keep input order).
// vals_are_index: the input values are not read; value i is taken to be i (the sorted values are
// then the permutation), so the key kernel need not write them.
void radix_sort_pairs32(uint32_t*& keys, uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt, int64_t n,
                        int begin_bit, int end_bit, bool vals_are_index, DBuf& hist, DBuf& rowsum, hipStream_t s);
// The same over caller-provided histogram scratch (no allocation): hist holds
// radix_sort_hist_words(n) int32, rowsum 256 int32.
int64_t radix_sort_hist_words(int64_t n);
void radix_sort_pairs32_scratch(uint32_t*& keys, uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt, int64_t n,
                                int begin_bit, int end_bit, bool vals_are_index, int32_t* hist, int32_t* rowsum,
                                hipStream_t s);
// exclusive scan of int32 (in may alias out); returns nothing, total in out[n] if out has n+1 slots
void exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, DBuf& tmp, hipStream_t s);

// build (fwn_build.hip)
void build_engine(Engine& e, int32_t n_meshes, const float* V_host, const int64_t* v_off, const int32_t* F_host,
                  const int64_t* f_off);

// SAH topology (fwn_sah.hip): top-down binned box-area splits of each mesh's triangles down to
// single triangles, in the binary tree layout of the LBVH path: internal nodes [0, n-1), leaf j
// -> n-1+j for the triangle at sorted position j, `items` = sorted position -> triangle,
// left/right/first/last per internal node, parent per node (the caller fills -1 first),
// is_root (the caller fills 0) and root_bin per mesh.
void build_topology_sah(const float* cen, const float* tbox, int64_t nF, const int64_t* f_off, int M,
                        int32_t* items, int32_t* left, int32_t* right, int32_t* parent, int32_t* first,
                        int32_t* last, uint8_t* is_root, int32_t* root_bin, hipStream_t s);

// queries (fwn_query.hip)
void run_queries(Engine& e, const float* Q, const int64_t* q_off_host, int64_t nQ, float* omega, uint8_t* inside,
                 int32_t* stats, bool device_ptrs);
void run_bruteforce(Engine& e, const float* Q, int64_t nQ, double* out, bool device_ptrs);
// Pieces of a query that uses no engine scratch (single-mesh engine, device points): the 32-bit
// point keys, and the traversal in sorted order `perm` with FP64 near leaves evaluated in place
// (no work list); meta holds q_off {0, nQ}, b_off {0, batches} and the mesh order {0} (5 int64
// words), flag one int32 (set on stack overflow).  Results equal run_queries' bit for bit.
void launch_point_keys(Engine& e, const float* dQ, int64_t nQ, uint32_t* keys, hipStream_t s);
void launch_traverse_inplace(Engine& e, int banks, const float* dQ, const uint32_t* perm, const int64_t* meta,
                             int64_t nQ, float* omega, uint8_t* inside, int32_t* flag, hipStream_t s);
// fwn_solid_angle_noalloc (fwn_noalloc.hip): device points of a single-mesh engine through a fixed
// device scratch array; no allocation
void run_queries_noalloc(Engine& e, const float* dQ, int64_t nQ, float* dOm, uint8_t* dIn);
// Morton order of nQ device points of a single-mesh engine: returns the permutation (sorted
// position -> original index) in the engine's query scratch, valid until the next query call.
const uint32_t* sort_points_single(Engine& e, const float* dQ, int64_t nQ);

// closest point on the mesh of a single-mesh engine (fwn_closest.hip): squared distance (double),
// original facet index (-1 for an empty mesh) and, when cp is not NULL, the point (3 doubles)
void run_closest(Engine& e, const float* Q, int64_t nQ, double* d2, int32_t* tri, double* cp, bool device_ptrs);

// voxel signs (fwn_voxel.hip): centres float(voxel_size * ijk + voxel_size / 2) made on the device
void sign_voxel_box(Engine& e, const int32_t lo[3], const int32_t dims[3], double voxel_size, uint8_t* inside,
                    float* omega, bool device_ptrs);
void sign_voxel_list(Engine& e, const int32_t* ijk, int64_t n, double voxel_size, uint8_t* inside, float* omega,
                     bool device_ptrs);

// batched example callers (fwn_callers.hip)
void orientation_criterion(Engine& e, const float* V, int64_t nV, const int32_t* F, int64_t nF, float eps,
                           double* crit, bool device_ptrs);
void sample_inside(Engine& e, const float lo[3], const float hi[3], int64_t n, uint64_t seed, float* out,
                   int64_t* n_inside, bool device_ptrs);

int grid_blocks(int64_t n, int threads);

// validates an options struct (throws Failure)
void check_options(const fwn_options& o);

// Sets the current device for a scope and restores the previous one.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        FWN_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Runs body, turning failures into a status code and the thread-local error message.
template <class Body>
int guarded(const char* fn, Body&& body)
{
    try {
        body();
        return FWN_OK;
    } catch (const Failure& f) {
        set_error(std::string(fn) + ": " + f.msg);
        return f.status;
    } catch (const std::bad_alloc&) {
        set_error(std::string(fn) + ": host allocation failed");
        return FWN_OUT_OF_MEMORY;
    } catch (...) {
        set_error(std::string(fn) + ": unknown failure");
        return FWN_HIP_ERROR;
    }
}

} // namespace fwn

// ---------------------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------------------
namespace fwn {

__device__ __forceinline__ uint32_t expand10(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// (c - lo) * inv scaled to [0, 1023]; same float operations as the CPU check
__device__ __forceinline__ uint32_t quant10(float c, float lo, float inv)
{
    float t = (c - lo) * inv;
    float s = t * 1024.0f;
    if (!(s > 0.0f)) return 0u;
    if (s >= 1023.0f) return 1023u;
    return (uint32_t)s;
}

__device__ __forceinline__ uint32_t morton30(float x, float y, float z, const float* lo, const float* inv)
{
    return (expand10(quant10(x, lo[0], inv[0])) << 2) | (expand10(quant10(y, lo[1], inv[1])) << 1) |
           expand10(quant10(z, lo[2], inv[2]));
}

// order-preserving float <-> int maps for atomic min/max
__device__ __forceinline__ int32_t f2ord(float f)
{
    int32_t i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}
__host__ __device__ inline float ord2f(int32_t i)
{
    int32_t j = i >= 0 ? i : (i ^ 0x7FFFFFFF);
    float f;
    __builtin_memcpy(&f, &j, 4);
    return f;
}

// v_writelane_b32: returns `old` with lane `lane` replaced by the wave-uniform `v` (the LLVM
// intrinsic, which clang exposes under no builtin name)
extern "C" __device__ int fwn_llvm_writelane(int v, int lane, int old) __asm("llvm.amdgcn.writelane.i32");
__device__ __forceinline__ uint32_t wave_writelane(uint32_t v, int lane, uint32_t old)
{
    return (uint32_t)fwn_llvm_writelane((int)v, lane, (int)old);
}

// Wave-private traversal stack of (node entry, 64-bit lane mask), one entry per lane and bank:
// entry k lives in lane k % 64 of bank k / 64.  NB = 1 (64 entries, enough for trees with
// 1 + (width - 1) * depth <= 64) or 3 (kStack entries).  The entry and the mask are
// wave-uniform, so a push is one v_writelane per word (the written register is tied to the old
// one: no copies at the join of a conditional push) and a pop is one v_readlane per word; the
// bank is chosen by uniform branches, so nothing invites the optimiser to turn the banks into an
// indexed private array in scratch memory.  A push past the capacity writes no lane (callers
// report it as an overflow).
template <int NB>
struct WaveStack;

template <>
struct WaveStack<1> {
    static constexpr int kCap = 64;
    uint32_t n0 = 0, l0 = 0, h0 = 0;
    // sp < kCap is the caller's contract (the host launches this stack only for trees whose
    // bound fits; callers still flag sp >= kCap as an overflow): no guard here, so the push is
    // three writelanes with no branch around them
    __device__ __forceinline__ void push(int sp, int /*lane*/, uint32_t e, uint64_t mk)
    {
        n0 = wave_writelane(e, sp & 63, n0);
        l0 = wave_writelane((uint32_t)mk, sp & 63, l0);
        h0 = wave_writelane((uint32_t)(mk >> 32), sp & 63, h0);
    }
    __device__ __forceinline__ void pop(int sp, uint32_t& e, uint64_t& mk) const
    {
        const int l = sp & 63;
        e = __builtin_amdgcn_readlane(n0, l);
        const uint32_t lo = __builtin_amdgcn_readlane(l0, l), hi = __builtin_amdgcn_readlane(h0, l);
        mk = ((uint64_t)hi << 32) | lo;
    }
};

template <>
struct WaveStack<3> {
    static constexpr int kCap = kStack;
    uint32_t n0 = 0, n1 = 0, n2 = 0, l0 = 0, l1 = 0, l2 = 0, h0 = 0, h1 = 0, h2 = 0;
    __device__ __forceinline__ void push(int sp, int /*lane*/, uint32_t e, uint64_t mk)
    {
        const int l = sp & 63;
        const uint32_t mlo = (uint32_t)mk, mhi = (uint32_t)(mk >> 32);
        if (sp < 64) {
            n0 = wave_writelane(e, l, n0);
            l0 = wave_writelane(mlo, l, l0);
            h0 = wave_writelane(mhi, l, h0);
        } else if (sp < 128) {
            n1 = wave_writelane(e, l, n1);
            l1 = wave_writelane(mlo, l, l1);
            h1 = wave_writelane(mhi, l, h1);
        } else if (sp < kStack) {
            n2 = wave_writelane(e, l, n2);
            l2 = wave_writelane(mlo, l, l2);
            h2 = wave_writelane(mhi, l, h2);
        }
    }
    __device__ __forceinline__ void pop(int sp, uint32_t& e, uint64_t& mk) const
    {
        const int l = sp & 63;
        const uint32_t a0 = __builtin_amdgcn_readlane(n0, l), a1 = __builtin_amdgcn_readlane(n1, l),
                       a2 = __builtin_amdgcn_readlane(n2, l);
        const uint32_t c0 = __builtin_amdgcn_readlane(l0, l), c1 = __builtin_amdgcn_readlane(l1, l),
                       c2 = __builtin_amdgcn_readlane(l2, l);
        const uint32_t d0 = __builtin_amdgcn_readlane(h0, l), d1 = __builtin_amdgcn_readlane(h1, l),
                       d2 = __builtin_amdgcn_readlane(h2, l);
        const uint32_t lo = sp < 64 ? c0 : (sp < 128 ? c1 : c2);
        const uint32_t hi = sp < 64 ? d0 : (sp < 128 ? d1 : d2);
        e = sp < 64 ? a0 : (sp < 128 ? a1 : a2);
        mk = ((uint64_t)hi << 32) | lo;
    }
};

// stack banks a tree of this width and depth needs (1 or 3), 0 if it is too deep.  The bound
// 1 + (width - 1) depth on the live entries keeps one position free above the top: the traversal
// writes that position for leaf slots too (and only advances over it for internal ones).
inline int stack_banks(int width, int depth)
{
    const int need = 1 + (width - 1) * depth;
    return need < 64 ? 1 : (need < kStack ? 3 : 0);
}

// XCD-aware block remap: hardware hands consecutive workgroups to the 8 XCDs round-robin;
// this bijection gives each XCD one contiguous run of logical blocks (shared L2 reuse).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    uint32_t per = nb / 8u, rem = nb % 8u;
    uint32_t x = b % 8u, i = b / 8u;
    return (x < rem) ? x * (per + 1u) + i : rem * (per + 1u) + (x - rem) * per + i;
}

} // namespace fwn
