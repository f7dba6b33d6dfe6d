// C ABI of the engine (include/fwn.h).  Every entry point catches internal failures and
// returns a status code; fwn_last_error() gives the message of the last failure on the
// calling thread.  Argument checks mirror the reference's runtime asserts
// (modules/winding/src/FastWindingNumber.cpp:36-37, 91-96), which throw lagrange::Error.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <thread>

#include "fwn_internal.h"

namespace fwn {

namespace {
thread_local std::string g_last_error;
} // namespace

void check_options(const fwn_options& o)
{
    if (!(o.accuracy_scale > 0.0f) || !std::isfinite(o.accuracy_scale))
        fail(FWN_INVALID_ARGUMENT, "accuracy_scale must be positive and finite");
    if (!(o.near_fp64_rel >= 0.0f) || !std::isfinite(o.near_fp64_rel))
        fail(FWN_INVALID_ARGUMENT, "near_fp64_rel must be >= 0");
    if (o.leaf_max < 1 || o.leaf_max > 8) fail(FWN_INVALID_ARGUMENT, "leaf_max must be in [1, 8]");
    if (o.fp64_moment_min < 1) fail(FWN_INVALID_ARGUMENT, "fp64_moment_min must be >= 1");
    if (o.waves_per_block != 0 && o.waves_per_block != 1 && o.waves_per_block != 2 && o.waves_per_block != 4 &&
        o.waves_per_block != 8)
        fail(FWN_INVALID_ARGUMENT, "waves_per_block must be 0 (default), 1, 2, 4 or 8");
    if (o.launch_mode != FWN_LAUNCH_GRID && o.launch_mode != FWN_LAUNCH_PERSISTENT)
        fail(FWN_INVALID_ARGUMENT, "launch_mode must be FWN_LAUNCH_GRID or FWN_LAUNCH_PERSISTENT");
    if (o.tree_width != 2 && o.tree_width != 4 && o.tree_width != 8)
        fail(FWN_INVALID_ARGUMENT, "tree_width must be 2, 4 or 8");
    if (o.tree_builder != FWN_TREE_LBVH && o.tree_builder != FWN_TREE_SAH)
        fail(FWN_INVALID_ARGUMENT, "tree_builder must be FWN_TREE_LBVH or FWN_TREE_SAH");
    if (o.moment_pass != FWN_MOMENTS_LEVELS && o.moment_pass != FWN_MOMENTS_CLIMB)
        fail(FWN_INVALID_ARGUMENT, "moment_pass must be FWN_MOMENTS_LEVELS or FWN_MOMENTS_CLIMB");
}

namespace {

Engine* make_engine(const fwn_options* opt)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) fail(FWN_NO_DEVICE, "no HIP device available");
    fwn_options o;
    fwn_options_default(&o);
    if (opt) o = *opt;
    check_options(o);
    if (o.device < 0 || o.device >= ndev) fail(FWN_NO_DEVICE, "device ordinal out of range");
    Engine* e = new Engine();
    e->opt = o;
    e->device = o.device;
    return e;
}

} // namespace

// Single-point query (the reference's per-point solid_angle / is_inside).  Concurrent callers
// on one engine are coalesced: each request joins the pending list; a caller that finds no
// batch running takes the whole list, answers it with one batched query (results do not depend
// on which points share a batch) and wakes the others.  Throughput therefore grows with the
// number of calling threads instead of being one GPU round trip per point.  Callers waiting for
// a running batch poll their result for up to kPointSpinUs before they sleep, so they are back
// with their next request when the batch ends; a new batch first waits (at most
// kPointGatherUs) until as many requests are pending as the previous batch held.  A lone caller
// (previous batch of one) never waits, and a wait that times out lowers the expectation to what
// arrived.
constexpr int kPointGatherUs = 10;
constexpr int kPointSpinUs = 200;

void point_query(Engine& e, const float q[3], float* omega, uint8_t* inside)
{
    PointRequest r;
    std::memcpy(r.q, q, sizeof(r.q));
    std::unique_lock<std::mutex> lk(e.pt_mu);
    e.pt_pending.push_back(&r);
    while (!r.done.load(std::memory_order_acquire)) {
        if (e.pt_busy) {
            // a batch is running: poll for this request's result without the lock (a woken
            // sleeper would come back too late for the next batch), then sleep until it is done
            // or the engine is free
            lk.unlock();
            const auto t_end = std::chrono::steady_clock::now() + std::chrono::microseconds(kPointSpinUs);
            while (!r.done.load(std::memory_order_acquire) && std::chrono::steady_clock::now() < t_end)
                std::this_thread::yield();
            lk.lock();
            e.pt_cv.wait(lk, [&] { return r.done.load(std::memory_order_acquire) || !e.pt_busy; });
            continue;
        }
        e.pt_busy = true;
        if (e.pt_pending.size() < e.pt_conc) {
            const auto t_end = std::chrono::steady_clock::now() + std::chrono::microseconds(kPointGatherUs);
            while (e.pt_pending.size() < e.pt_conc && std::chrono::steady_clock::now() < t_end) {
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
        }
        std::vector<PointRequest*> batch;
        batch.swap(e.pt_pending);
        e.pt_conc = batch.size();
        lk.unlock();
        const size_t n = batch.size();
        int st = FWN_OK;
        std::string msg;
        std::vector<float> Q(3 * n), om(n, 0.0f);
        std::vector<uint8_t> in(n, 0);
        try {
            for (size_t k = 0; k < n; ++k) std::memcpy(&Q[3 * k], batch[k]->q, sizeof(batch[k]->q));
            std::lock_guard<std::mutex> g(e.mu);
            DeviceGuard dg(e.device);
            run_queries(e, Q.data(), nullptr, (int64_t)n, om.data(), in.data(), nullptr, false);
        } catch (const Failure& f) {
            st = f.status;
            msg = f.msg;
        } catch (const std::bad_alloc&) {
            st = FWN_OUT_OF_MEMORY;
            msg = "host allocation failed";
        } catch (...) {
            st = FWN_HIP_ERROR;
            msg = "unknown failure";
        }
        lk.lock();
        for (size_t k = 0; k < n; ++k) {
            batch[k]->omega = om[k];
            batch[k]->inside = in[k];
            batch[k]->status = st;
            batch[k]->err = msg;
            batch[k]->done.store(true, std::memory_order_release);
        }
        e.pt_busy = false;
        e.pt_cv.notify_all();
    }
    lk.unlock();
    if (r.status != FWN_OK) throw Failure{r.status, r.err};
    *omega = r.omega;
    *inside = r.inside;
}

namespace {

void destroy_engine(Engine* e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->cs_in) (void)hipStreamSynchronize(e->cs_in);
    if (e->cs_out) (void)hipStreamSynchronize(e->cs_out);
    for (auto& ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->h_stage) (void)hipHostFree(e->h_stage);
    if (e->h_pipe) (void)hipHostFree(e->h_pipe);
    for (auto& ev : e->pev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : e->tev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->cs_in) (void)hipStreamDestroy(e->cs_in);
    if (e->cs_out) (void)hipStreamDestroy(e->cs_out);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

void create_common(fwn_engine** out, int32_t M, const float* V, const int64_t* v_off, const int32_t* F,
                   const int64_t* f_off, const fwn_options* opt)
{
    if (!out) fail(FWN_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (M < 1) fail(FWN_INVALID_ARGUMENT, "at least one mesh is required");
    if (!v_off || !f_off) fail(FWN_INVALID_ARGUMENT, "offset arrays are NULL");
    if (v_off[0] != 0 || f_off[0] != 0) fail(FWN_INVALID_ARGUMENT, "offsets must start at 0");
    for (int m = 0; m < M; ++m)
        if (v_off[m + 1] < v_off[m] || f_off[m + 1] < f_off[m])
            fail(FWN_INVALID_ARGUMENT, "offsets must be non-decreasing");
    if (v_off[M] > 0 && !V) fail(FWN_INVALID_ARGUMENT, "vertex array is NULL");
    if (f_off[M] > 0 && !F) fail(FWN_INVALID_ARGUMENT, "facet array is NULL");
    for (int m = 0; m < M; ++m)
        if (v_off[m + 1] - v_off[m] > (int64_t)INT32_MAX)
            fail(FWN_LIMIT_EXCEEDED, "a mesh has more than 2^31-1 vertices (indices are int32)");
    Engine* e = make_engine(opt);
    try {
        DeviceGuard g(e->device);
        FWN_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        build_engine(*e, M, V, v_off, F, f_off);
        if (e->opt.tree_builder == FWN_TREE_SAH && stack_banks(e->width, e->max_depth) == 0) {
            // an SAH tree too deep for the traversal stack (extreme splits on extreme inputs):
            // the Morton tree's depth is bounded by its key bits
            e->opt.tree_builder = FWN_TREE_LBVH;
            build_engine(*e, M, V, v_off, F, f_off);
        }
    } catch (...) {
        destroy_engine(e);
        throw;
    }
    *out = reinterpret_cast<fwn_engine*>(e);
}

Engine* E(fwn_engine* h)
{
    if (!h) fail(FWN_INVALID_ARGUMENT, "engine is NULL");
    return reinterpret_cast<Engine*>(h);
}
const Engine* E(const fwn_engine* h)
{
    if (!h) fail(FWN_INVALID_ARGUMENT, "engine is NULL");
    return reinterpret_cast<const Engine*>(h);
}
} // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error_text() { return g_last_error.c_str(); }

void fail(int status, const std::string& msg) { throw Failure{status, msg}; }

void hip_check(hipError_t e, const char* what, const char* file, int line)
{
    if (e == hipSuccess) return;
    int st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? FWN_OUT_OF_MEMORY : FWN_HIP_ERROR;
    fail(st, std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) +
                 ")");
}

void DBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

void DBuf::alloc(size_t n)
{
    release();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        fail(FWN_OUT_OF_MEMORY, "device allocation of " + std::to_string(n) + " bytes failed");
    }
    bytes = n;
}

void DBuf::reserve(size_t n)
{
    if (n <= bytes && p) return;
    alloc(n + n / 4);
}

} // namespace fwn

using fwn::E;
using fwn::guarded;

#pragma GCC visibility push(default)
extern "C" {

int fwn_options_default(fwn_options* opt)
{
    if (!opt) {
        fwn::g_last_error = "fwn_options_default: opt is NULL";
        return FWN_INVALID_ARGUMENT;
    }
    std::memset(opt, 0, sizeof(*opt));
    opt->accuracy_scale = 2.0f;
    opt->near_fp64_rel = 1e-4f;
    opt->leaf_max = 1;
    opt->fp64_moment_min = 256;
    opt->device = 0;
    opt->waves_per_block = 1;
    opt->tree_width = 4;
    opt->launch_mode = FWN_LAUNCH_GRID;
    opt->tree_builder = FWN_TREE_SAH;
    opt->moment_pass = FWN_MOMENTS_LEVELS;
    return FWN_OK;
}

int fwn_create(fwn_engine** out, const float* V, int64_t nV, const int32_t* F, int64_t nF, const fwn_options* opt)
{
    return guarded("fwn_create", [&] {
        if (nV < 0 || nF < 0) fwn::fail(FWN_INVALID_ARGUMENT, "negative size");
        int64_t vo[2] = {0, nV}, fo[2] = {0, nF};
        fwn::create_common(out, 1, V, vo, F, fo, opt);
    });
}

int fwn_create_multi(fwn_engine** out, int32_t n_meshes, const float* V, const int64_t* v_off, const int32_t* F,
                     const int64_t* f_off, const fwn_options* opt)
{
    return guarded("fwn_create_multi", [&] { fwn::create_common(out, n_meshes, V, v_off, F, f_off, opt); });
}

int fwn_destroy(fwn_engine* h)
{
    return guarded("fwn_destroy", [&] { fwn::destroy_engine(reinterpret_cast<fwn::Engine*>(h)); });
}

int fwn_solid_angle(fwn_engine* h, const float* Q, int64_t nQ, float* omega_out, uint8_t* inside_out, int32_t flags)
{
    return guarded("fwn_solid_angle", [&] {
        fwn::Engine* e = E(h);
        if (nQ < 0) fwn::fail(FWN_INVALID_ARGUMENT, "negative point count");
        if (nQ > 0 && (!Q || !omega_out)) fwn::fail(FWN_INVALID_ARGUMENT, "NULL points or output");
        if (e->n_meshes != 1) fwn::fail(FWN_INVALID_ARGUMENT, "multi-mesh engine: use fwn_solid_angle_multi");
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::run_queries(*e, Q, nullptr, nQ, omega_out, inside_out, nullptr, (flags & FWN_DEVICE_POINTERS) != 0);
    });
}

int fwn_solid_angle_noalloc(fwn_engine* h, const float* Q, int64_t nQ, float* omega_out, uint8_t* inside_out)
{
    return guarded("fwn_solid_angle_noalloc", [&] {
        fwn::Engine* e = E(h);
        if (nQ > 0 && (!Q || !omega_out)) fwn::fail(FWN_INVALID_ARGUMENT, "NULL points or output");
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::run_queries_noalloc(*e, Q, nQ, omega_out, inside_out);
    });
}

int fwn_solid_angle_multi(fwn_engine* h, const float* Q, const int64_t* q_off, float* omega_out, uint8_t* inside_out,
                          int32_t flags)
{
    return guarded("fwn_solid_angle_multi", [&] {
        fwn::Engine* e = E(h);
        if (!q_off) fwn::fail(FWN_INVALID_ARGUMENT, "q_off is NULL");
        if (q_off[e->n_meshes] > 0 && (!Q || !omega_out)) fwn::fail(FWN_INVALID_ARGUMENT, "NULL points or output");
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::run_queries(*e, Q, q_off, 0, omega_out, inside_out, nullptr, (flags & FWN_DEVICE_POINTERS) != 0);
    });
}

int fwn_sign_voxel_box(fwn_engine* h, const int32_t lo[3], const int32_t dims[3], double voxel_size,
                       uint8_t* inside_out, float* omega_out, int32_t flags)
{
    return guarded("fwn_sign_voxel_box", [&] {
        fwn::Engine* e = E(h);
        if (!lo || !dims) fwn::fail(FWN_INVALID_ARGUMENT, "NULL lo or dims");
        if (e->n_meshes != 1) fwn::fail(FWN_INVALID_ARGUMENT, "voxel signs need a single-mesh engine");
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::sign_voxel_box(*e, lo, dims, voxel_size, inside_out, omega_out, (flags & FWN_DEVICE_POINTERS) != 0);
    });
}

int fwn_sign_voxels(fwn_engine* h, const int32_t* ijk, int64_t n, double voxel_size, uint8_t* inside_out,
                    float* omega_out, int32_t flags)
{
    return guarded("fwn_sign_voxels", [&] {
        fwn::Engine* e = E(h);
        if (e->n_meshes != 1) fwn::fail(FWN_INVALID_ARGUMENT, "voxel signs need a single-mesh engine");
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::sign_voxel_list(*e, ijk, n, voxel_size, inside_out, omega_out, (flags & FWN_DEVICE_POINTERS) != 0);
    });
}

int fwn_orientation_criterion(fwn_engine* h, const float* V, int64_t nV, const int32_t* F, int64_t nF,
                              float epsilon, double* criterion_out, int32_t flags)
{
    return guarded("fwn_orientation_criterion", [&] {
        fwn::Engine* e = E(h);
        if (e->n_meshes != 1) fwn::fail(FWN_INVALID_ARGUMENT, "the orientation criterion needs a single-mesh engine");
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::orientation_criterion(*e, V, nV, F, nF, epsilon, criterion_out, (flags & FWN_DEVICE_POINTERS) != 0);
    });
}

int fwn_sample_inside(fwn_engine* h, const float lo[3], const float hi[3], int64_t n_samples, uint64_t seed,
                      float* points_out, int64_t* n_inside_out, int32_t flags)
{
    return guarded("fwn_sample_inside", [&] {
        fwn::Engine* e = E(h);
        if (!lo || !hi || !n_inside_out) fwn::fail(FWN_INVALID_ARGUMENT, "NULL box or count");
        if (e->n_meshes != 1) fwn::fail(FWN_INVALID_ARGUMENT, "sampling needs a single-mesh engine");
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::sample_inside(*e, lo, hi, n_samples, seed, points_out, n_inside_out, (flags & FWN_DEVICE_POINTERS) != 0);
    });
}

int fwn_closest_point(fwn_engine* h, const float* Q, int64_t nQ, double* sq_dist_out, int32_t* facet_out,
                      double* closest_out, int32_t flags)
{
    return guarded("fwn_closest_point", [&] {
        fwn::Engine* e = E(h);
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::run_closest(*e, Q, nQ, sq_dist_out, facet_out, closest_out, (flags & FWN_DEVICE_POINTERS) != 0);
    });
}

int fwn_solid_angle_point(fwn_engine* h, const float q[3], float* omega_out)
{
    return guarded("fwn_solid_angle_point", [&] {
        fwn::Engine* e = E(h);
        if (!q || !omega_out) fwn::fail(FWN_INVALID_ARGUMENT, "NULL point or output");
        if (e->n_meshes != 1) fwn::fail(FWN_INVALID_ARGUMENT, "multi-mesh engine has no single-point query");
        uint8_t inside = 0;
        fwn::point_query(*e, q, omega_out, &inside);
    });
}

int fwn_is_inside_point(fwn_engine* h, const float q[3], uint8_t* inside_out)
{
    return guarded("fwn_is_inside_point", [&] {
        fwn::Engine* e = E(h);
        if (!q || !inside_out) fwn::fail(FWN_INVALID_ARGUMENT, "NULL point or output");
        if (e->n_meshes != 1) fwn::fail(FWN_INVALID_ARGUMENT, "multi-mesh engine has no single-point query");
        float om = 0.0f;
        fwn::point_query(*e, q, &om, inside_out);
    });
}

int fwn_tree_info_get(const fwn_engine* h, fwn_tree_info* info)
{
    return guarded("fwn_tree_info_get", [&] {
        const fwn::Engine* e = E(h);
        if (!info) fwn::fail(FWN_INVALID_ARGUMENT, "info is NULL");
        info->num_meshes = e->n_meshes;
        info->num_nodes = e->n_nodes;
        info->num_triangles = e->n_tris;
        info->width = e->width;
        info->max_depth = e->max_depth;
        info->build_ms = e->build_ms;
    });
}

int fwn_export_tree(const fwn_engine* h, uint32_t* slots, uint32_t* tris, int32_t* roots, float* tau)
{
    return guarded("fwn_export_tree", [&] {
        fwn::Engine* e = const_cast<fwn::Engine*>(E(h));
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        if (slots && e->n_nodes)
            FWN_HIP(hipMemcpy(slots, e->slots.p, sizeof(uint32_t) * e->width * fwn::kSlotWords * e->n_nodes,
                              hipMemcpyDeviceToHost));
        if (tris && e->n_tris)
            FWN_HIP(hipMemcpy(tris, e->tris.p, sizeof(uint32_t) * fwn::kTriWords * e->n_tris, hipMemcpyDeviceToHost));
        if (roots) std::memcpy(roots, e->h_roots.data(), sizeof(int32_t) * e->h_roots.size());
        if (tau) std::memcpy(tau, e->h_tau.data(), sizeof(float) * e->h_tau.size());
    });
}

int fwn_query_stats(fwn_engine* h, const float* Q, const int64_t* q_off, int64_t nQ, float* omega_out,
                    int32_t* stats_out, int32_t flags)
{
    return guarded("fwn_query_stats", [&] {
        fwn::Engine* e = E(h);
        if (!stats_out || !omega_out) fwn::fail(FWN_INVALID_ARGUMENT, "NULL output");
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::run_queries(*e, Q, q_off, nQ, omega_out, nullptr, stats_out, (flags & FWN_DEVICE_POINTERS) != 0);
    });
}

int fwn_bruteforce_f64(fwn_engine* h, const float* Q, int64_t nQ, double* omega_out, int32_t flags)
{
    return guarded("fwn_bruteforce_f64", [&] {
        fwn::Engine* e = E(h);
        if (nQ < 0) fwn::fail(FWN_INVALID_ARGUMENT, "negative point count");
        if (nQ > 0 && (!Q || !omega_out)) fwn::fail(FWN_INVALID_ARGUMENT, "NULL points or output");
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        fwn::run_bruteforce(*e, Q, nQ, omega_out, (flags & FWN_DEVICE_POINTERS) != 0);
    });
}

int fwn_set_launch(fwn_engine* h, int32_t waves_per_block, int32_t mode)
{
    return guarded("fwn_set_launch", [&] {
        fwn::Engine* e = E(h);
        fwn_options o = e->opt;
        o.waves_per_block = waves_per_block;
        o.launch_mode = mode;
        fwn::check_options(o);
        std::lock_guard<std::mutex> lk(e->mu);
        e->opt = o;
    });
}

int fwn_autotune(fwn_engine* h, const float* Q, const int64_t* q_off, int64_t nQ, int32_t flags,
                 const char* cache_path, int32_t* waves_out, int32_t* mode_out, double* best_ms_out)
{
    return guarded("fwn_autotune", [&] {
        fwn::Engine* e = E(h);
        std::lock_guard<std::mutex> lk(e->mu);
        fwn::DeviceGuard g(e->device);
        const bool dev = (flags & FWN_DEVICE_POINTERS) != 0;
        if (q_off) nQ = q_off[e->n_meshes];
        if (nQ <= 0 || !Q) fwn::fail(FWN_INVALID_ARGUMENT, "autotune needs points");
        hipDeviceProp_t prop;
        FWN_HIP(hipGetDeviceProperties(&prop, e->device));
        int nq_bucket = 0;
        while ((int64_t(1) << (nq_bucket + 1)) <= nQ) ++nq_bucket;
        std::ostringstream ks;
        ks << prop.gcnArchName << "|w" << e->width << "|l" << e->opt.leaf_max << "|m" << e->n_meshes << "|q" << nq_bucket;
        const std::string key = ks.str();
        int best_w = -1, best_m = -1;
        double best = 1e300;
        if (cache_path) {
            std::ifstream in(cache_path);
            std::string line;
            while (std::getline(in, line)) {
                std::istringstream ls(line);
                std::string k;
                int w = 0, m = -1;
                double ms = 0.0;
                if (!(ls >> k >> w >> m >> ms) || k != key) continue;
                // a line with a shape the engine does not launch (stale or edited file) is ignored
                const bool shape_ok = (w == 1 || w == 2 || w == 4 || w == 8) &&
                                      (m == FWN_LAUNCH_GRID || m == FWN_LAUNCH_PERSISTENT);
                if (shape_ok && std::isfinite(ms) && ms > 0.0) {
                    best_w = w;
                    best_m = m;
                    best = ms;
                }
            }
        }
        if (best_w < 0) {
            fwn::DBuf out;
            std::vector<float> hout;
            float* po = nullptr;
            if (dev) {
                out.alloc(sizeof(float) * nQ);
                po = out.as<float>();
            } else {
                hout.resize(nQ);
                po = hout.data();
            }
            const fwn_options saved = e->opt;
            const int waves[4] = {1, 2, 4, 8};
            for (int mode = FWN_LAUNCH_GRID; mode <= FWN_LAUNCH_PERSISTENT; ++mode)
                for (int w : waves) {
                    e->opt.waves_per_block = w;
                    e->opt.launch_mode = mode;
                    double t = 1e300;
                    for (int rep = 0; rep < 3; ++rep) { // first run warms caches; keep the best of the rest
                        fwn::run_queries(*e, Q, q_off, nQ, po, nullptr, nullptr, dev);
                        if (!(e->last_total_ms > 0.0)) {
                            e->opt = saved;
                            fwn::fail(FWN_HIP_ERROR, "autotune: the query call reported no device time");
                        }
                        if (rep > 0) t = std::min(t, e->last_total_ms);
                    }
                    if (t < best) {
                        best = t;
                        best_w = w;
                        best_m = mode;
                    }
                }
            e->opt = saved;
            if (cache_path) {
                std::ofstream o(cache_path, std::ios::app);
                o << key << " " << best_w << " " << best_m << " " << best << "\n";
            }
        }
        e->opt.waves_per_block = best_w;
        e->opt.launch_mode = best_m;
        if (waves_out) *waves_out = best_w;
        if (mode_out) *mode_out = best_m;
        if (best_ms_out) *best_ms_out = best;
    });
}

int fwn_create_tuned(fwn_engine** out, int32_t n_meshes, const float* V, const int64_t* v_off, const int32_t* F,
                     const int64_t* f_off, const fwn_options* opt, const float* Q, const int64_t* q_off, int64_t nQ,
                     int32_t flags, const int32_t* leaf_sizes, int32_t n_leaf_sizes, int32_t* leaf_out,
                     double* best_ms_out)
{
    return guarded("fwn_create_tuned", [&] {
        if (!out) fwn::fail(FWN_INVALID_ARGUMENT, "out is NULL");
        *out = nullptr;
        if (!leaf_sizes || n_leaf_sizes < 1) fwn::fail(FWN_INVALID_ARGUMENT, "no leaf sizes to try");
        fwn_options o;
        if (opt) o = *opt;
        else fwn_options_default(&o);
        fwn_engine* best = nullptr;
        double best_ms = 1e300;
        int32_t best_leaf = -1;
        for (int32_t k = 0; k < n_leaf_sizes; ++k) {
            o.leaf_max = leaf_sizes[k];
            fwn_engine* e = nullptr;
            int st = fwn_create_multi(&e, n_meshes, V, v_off, F, f_off, &o);
            if (st == FWN_OK) {
                int32_t w = 0, m = 0;
                double ms = 0.0;
                st = fwn_autotune(e, Q, q_off, nQ, flags, nullptr, &w, &m, &ms);
                if (st == FWN_OK && ms < best_ms) {
                    if (best) fwn_destroy(best);
                    best = e;
                    best_ms = ms;
                    best_leaf = o.leaf_max;
                    continue;
                }
                fwn_destroy(e);
            }
            if (st != FWN_OK) {
                const std::string msg = fwn::last_error_text();
                if (best) fwn_destroy(best);
                fwn::fail(st, msg);
            }
        }
        *out = best;
        if (leaf_out) *leaf_out = best_leaf;
        if (best_ms_out) *best_ms_out = best_ms;
    });
}

int fwn_last_timing(const fwn_engine* h, double* traverse_ms, double* total_ms)
{
    return guarded("fwn_last_timing", [&] {
        const fwn::Engine* e = E(h);
        if (traverse_ms) *traverse_ms = e->last_traverse_ms;
        if (total_ms) *total_ms = e->last_total_ms;
    });
}

const char* fwn_last_error(void) { return fwn::g_last_error.c_str(); }

} // extern "C"
#pragma GCC visibility pop
