// bench_point_calls: the reference's per-point winding benchmark
// (modules/winding/tests/test_fast_winding_number.cpp:85-110, "pimpl wrapper") run through the
// drop-in class: a lagrange::winding::FastWindingNumber built from a SurfaceMesh<float, uint32_t>,
// then num_samples serial is_inside calls on points drawn from std::mt19937 (default seed) with
// one std::uniform_real_distribution<float> per axis over the float bounding box (x, y, z per
// call).  Beside it:
//  * the same calls from T threads at once on one engine (thread t draws from std::mt19937(t + 1));
//    concurrent per-point calls are answered together in shared batches;
//  * the serial sequence's points in one batched is_inside / solid_angle call, whose results the
//    per-point calls must equal bit for bit;
//  * batched solid_angle calls of k = 1 ... 4096 of those points (per-call time), to place the
//    crossover between k per-point calls and one batched call.
// Prints one JSON object.  Times are wall-clock (steady_clock), median of the timed runs after one
// warm-up run.
//
//   bench_point_calls MESH [-n NUM_SAMPLES=10000] [-t THREADS=16] [-r RUNS=5]
#include <lagrange/utils/Error.h>
#include <lagrange/winding/FastWindingNumber.h>

#include <fwn_io.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double median(std::vector<double> v)
{
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
}

} // namespace

int main(int argc, char** argv)
{
    std::string path;
    size_t num_samples = 10000;
    int threads = 16, runs = 5;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-n" && i + 1 < argc) num_samples = std::strtoull(argv[++i], nullptr, 10);
        else if (a == "-t" && i + 1 < argc) threads = std::atoi(argv[++i]);
        else if (a == "-r" && i + 1 < argc) runs = std::atoi(argv[++i]);
        else path = a;
    }
    if (path.empty() || num_samples == 0 || threads < 1 || runs < 1) {
        std::fprintf(stderr, "usage: %s MESH [-n NUM_SAMPLES] [-t THREADS] [-r RUNS]\n", argv[0]);
        return 2;
    }
    try {
        fwn_mesh* m = nullptr;
        if (fwn_mesh_load(path.c_str(), FWN_IO_FORMAT_AUTO, 1, &m) != FWN_IO_OK)
            throw lagrange::Error(std::string("cannot load: ") + fwn_io_last_error());
        int64_t nv = 0, nf = 0, nc = 0;
        fwn_mesh_sizes(m, &nv, &nf, &nc);
        std::vector<double> V((size_t)nv * 3);
        std::vector<int32_t> cor((size_t)nc);
        fwn_mesh_copy(m, V.data(), nullptr, cor.data());
        fwn_mesh_free(m);
        lagrange::SurfaceMesh<float, uint32_t> mesh;
        std::array<float, 3> lo{0, 0, 0}, hi{0, 0, 0};
        for (int64_t i = 0; i < nv; ++i) {
            const float p[3] = {(float)V[3 * i], (float)V[3 * i + 1], (float)V[3 * i + 2]};
            mesh.add_vertex(p);
            for (int d = 0; d < 3; ++d) {
                lo[d] = i ? std::min(lo[d], p[d]) : p[d];
                hi[d] = i ? std::max(hi[d], p[d]) : p[d];
            }
        }
        for (int64_t f = 0; f < nf; ++f)
            mesh.add_triangle((uint32_t)cor[3 * f], (uint32_t)cor[3 * f + 1], (uint32_t)cor[3 * f + 2]);
        std::uniform_real_distribution<float> px(lo[0], hi[0]);
        std::uniform_real_distribution<float> py(lo[1], hi[1]);
        std::uniform_real_distribution<float> pz(lo[2], hi[2]);

        const double t_build = now_s();
        lagrange::winding::FastWindingNumber engine(mesh);
        const double build_s = now_s() - t_build;

        // serial: the reference benchmark's loop, one generator per timed run
        std::vector<double> serial;
        size_t num_insides = 0;
        for (int r = 0; r <= runs; ++r) {
            std::mt19937 gen;
            const double t0 = now_s();
            size_t k_in = 0;
            for (size_t k = 0; k < num_samples; ++k) k_in += engine.is_inside({px(gen), py(gen), pz(gen)});
            const double dt = now_s() - t0;
            if (r > 0) serial.push_back(dt);
            num_insides = k_in;
        }
        // the same points, one batched call, and per-point solid_angle against it
        std::vector<std::array<float, 3>> pts(num_samples);
        {
            std::mt19937 gen;
            for (auto& p : pts) p = {px(gen), py(gen), pz(gen)};
        }
        std::vector<float> om(num_samples);
        std::vector<char> ins(num_samples);
        std::vector<double> batched;
        for (int r = 0; r <= runs; ++r) {
            const double t0 = now_s();
            engine.solid_angle(pts.data(), num_samples, om.data());
            engine.is_inside(pts.data(), num_samples, reinterpret_cast<bool*>(ins.data()));
            const double dt = now_s() - t0;
            if (r > 0) batched.push_back(dt);
        }
        size_t batched_insides = 0, mismatches = 0;
        for (size_t k = 0; k < num_samples; ++k) {
            batched_insides += ins[k] != 0;
            const float o = engine.solid_angle(pts[k]);
            if (std::memcmp(&o, &om[k], sizeof(float)) != 0 || engine.is_inside(pts[k]) != (ins[k] != 0)) ++mismatches;
        }
        // threads: T callers, each num_samples calls
        std::vector<double> threaded;
        for (int r = 0; r <= runs; ++r) {
            std::vector<std::thread> th;
            std::vector<size_t> cnt(threads, 0);
            const double t0 = now_s();
            for (int t = 0; t < threads; ++t)
                th.emplace_back([&, t] {
                    std::mt19937 gen((unsigned)t + 1u);
                    size_t c = 0;
                    for (size_t k = 0; k < num_samples; ++k) c += engine.is_inside({px(gen), py(gen), pz(gen)});
                    cnt[t] = c;
                });
            for (auto& x : th) x.join();
            const double dt = now_s() - t0;
            if (r > 0) threaded.push_back(dt);
        }
        // batched calls of k points (the crossover with k per-point calls)
        std::string bk;
        for (size_t k : {1, 2, 4, 8, 16, 17, 32, 64, 128, 256, 1024, 4096}) {
            if (k > num_samples) break;
            std::vector<double> tk;
            for (int r = 0; r <= 4 * runs; ++r) {
                const double t0 = now_s();
                engine.solid_angle(pts.data(), k, om.data());
                const double dt = now_s() - t0;
                if (r > 0) tk.push_back(dt);
            }
            char buf[64];
            std::snprintf(buf, sizeof(buf), "%s\"%zu\": %.3f", bk.empty() ? "" : ", ", k, 1e6 * median(tk));
            bk += buf;
        }
        const double ts = median(serial), tb = median(batched), tt = median(threaded);
        std::printf("{\"mesh_triangles\": %lld, \"num_samples\": %zu, \"build_s\": %.6f, "
                    "\"serial_s\": %.6f, \"serial_calls_per_s\": %.1f, \"serial_us_per_call\": %.3f, "
                    "\"num_insides\": %zu, \"batched_s\": %.6f, \"batched_points_per_s\": %.1f, "
                    "\"batched_insides\": %zu, \"per_point_vs_batched_mismatches\": %zu, "
                    "\"threads\": %d, \"threaded_s\": %.6f, \"threaded_calls_per_s\": %.1f, \"runs\": %d, "
                    "\"batched_call_us\": {%s}}\n",
                    (long long)nf, num_samples, build_s, ts, num_samples / ts, 1e6 * ts / num_samples, num_insides,
                    tb, num_samples / tb, batched_insides, mismatches, threads, tt,
                    (double)threads * num_samples / tt, runs, bk.c_str());
        return mismatches == 0 && batched_insides == num_insides ? 0 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
