// Tests of lagrange::winding::FastWindingNumber (drop-in class over the GPU engine).
// Self-contained runner: prints one line per case and returns non-zero on failure.
// Geometry: the unit cube and octant triangle of the reference's pytest fixtures
// (modules/conftest.py:29-37, 108-144) and a subdivided octahedron projected to the sphere.
#include <lagrange/utils/Error.h>
#include <lagrange/winding/FastWindingNumber.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <functional>
#include <map>
#include <random>
#include <string>
#include <thread>
#include <utility>
#include <vector>

using lagrange::SurfaceMesh;
using lagrange::winding::FastWindingNumber;

static int g_fail = 0;
#define CHECK(cond)                                                                   \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::printf("  CHECK failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__);  \
            ++g_fail;                                                                 \
        }                                                                             \
    } while (0)

static const double kPi = 3.14159265358979323846;

template <typename S, typename I>
SurfaceMesh<S, I> cube()
{
    SurfaceMesh<S, I> m;
    const S v[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (auto& p : v) m.add_vertex(p);
    const I f[12][3] = {{0, 3, 2}, {2, 1, 0}, {4, 5, 6}, {6, 7, 4}, {1, 2, 6}, {6, 5, 1},
                        {4, 7, 3}, {3, 0, 4}, {2, 3, 7}, {7, 6, 2}, {0, 1, 5}, {5, 4, 0}};
    for (auto& t : f) m.add_triangle(t[0], t[1], t[2]);
    return m;
}

// octahedron with each face split into n^2 triangles, projected to the unit sphere
SurfaceMesh<double, uint32_t> sphere(int n)
{
    SurfaceMesh<double, uint32_t> m;
    const double P[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    const int F[8][3] = {{0, 2, 4}, {2, 1, 4}, {1, 3, 4}, {3, 0, 4}, {2, 0, 5}, {1, 2, 5}, {3, 1, 5}, {0, 3, 5}};
    std::map<std::array<long long, 3>, uint32_t> ids;
    auto vid = [&](const double* a, const double* b, const double* c, int i, int j) {
        double p[3];
        for (int d = 0; d < 3; ++d) p[d] = (a[d] * (n - i - j) + b[d] * i + c[d] * j) / n;
        double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (double& x : p) x /= l;
        std::array<long long, 3> key{std::llround(p[0] * 1e9), std::llround(p[1] * 1e9), std::llround(p[2] * 1e9)};
        auto it = ids.find(key);
        if (it != ids.end()) return it->second;
        uint32_t id = m.get_num_vertices();
        m.add_vertex(p);
        ids[key] = id;
        return id;
    };
    for (auto& f : F) {
        const double *a = P[f[0]], *b = P[f[1]], *c = P[f[2]];
        for (int i = 0; i < n; ++i)
            for (int j = 0; i + j < n; ++j) {
                m.add_triangle(vid(a, b, c, i, j), vid(a, b, c, i + 1, j), vid(a, b, c, i, j + 1));
                if (i + j < n - 1)
                    m.add_triangle(vid(a, b, c, i + 1, j), vid(a, b, c, i + 1, j + 1), vid(a, b, c, i, j + 1));
            }
    }
    return m;
}

template <typename S, typename I>
void cube_case()
{
    FastWindingNumber e(cube<S, I>());
    CHECK(std::fabs(e.solid_angle({0.5f, 0.5f, 0.5f}) - 4 * kPi) < 1e-5);
    CHECK(std::fabs(e.solid_angle({1.5f, 0.5f, 0.5f})) < 1e-5);
    CHECK(e.is_inside({0.25f, 0.5f, 0.75f}));
    CHECK(!e.is_inside({-0.1f, 0.5f, 0.5f}));
}

int main()
{
    std::vector<std::pair<std::string, std::function<void()>>> cases = {
        {"cube float/uint32", cube_case<float, uint32_t>},
        {"cube double/uint32", cube_case<double, uint32_t>},
        {"cube float/uint64", cube_case<float, uint64_t>},
        {"cube double/uint64", cube_case<double, uint64_t>},
        {"octant triangle", [] {
             SurfaceMesh<float, uint32_t> m;
             m.add_vertex({1, 0, 0});
             m.add_vertex({0, 1, 0});
             m.add_vertex({0, 0, 1});
             m.add_triangle(0, 1, 2);
             FastWindingNumber e(m);
             CHECK(std::fabs(e.solid_angle({0, 0, 0}) - kPi / 2) < 1e-6);
         }},
        {"rejects 2D and non-triangle meshes", [] {
             bool threw = false;
             try {
                 SurfaceMesh<double, uint32_t> m(2);
                 m.add_vertex({0, 0});
                 m.add_vertex({1, 0});
                 m.add_vertex({0, 1});
                 m.add_triangle(0, 1, 2);
                 FastWindingNumber e(m);
             } catch (const lagrange::Error& ex) {
                 // la_runtime_assert's message: the condition and the reference's reason text
                 const std::string w = ex.what();
                 threw = w.find("Assertion failed") != std::string::npos &&
                         w.find("Fast winding number engine only supports 3D meshes") != std::string::npos;
             }
             CHECK(threw);
             threw = false;
             try {
                 SurfaceMesh<double, uint32_t> m;
                 for (int i = 0; i < 4; ++i) m.add_vertex({double(i & 1), double(i >> 1), 0.0});
                 m.add_quad(0, 1, 3, 2);
                 FastWindingNumber e(m);
             } catch (const lagrange::Error& ex) {
                 const std::string w = ex.what();
                 threw = w.find("Assertion failed") != std::string::npos &&
                         w.find("Fast winding number engine only supports triangle meshes") != std::string::npos;
             }
             CHECK(threw);
         }},
        {"move-only semantics", [] {
             FastWindingNumber a(cube<float, uint32_t>());
             FastWindingNumber b(std::move(a));
             CHECK(b.is_inside({0.5f, 0.5f, 0.5f}));
             FastWindingNumber c;
             c = std::move(b);
             CHECK(c.is_inside({0.5f, 0.5f, 0.5f}));
             bool threw = false;
             try {
                 FastWindingNumber d;
                 d.is_inside({0, 0, 0});
             } catch (const lagrange::Error&) {
                 threw = true;
             }
             CHECK(threw);
         }},
        {"sphere: batched equals per-point, inside/outside", [] {
             FastWindingNumber e(sphere(24));
             std::vector<std::array<float, 3>> pts;
             for (int i = 0; i < 300; ++i) {
                 float t = i / 299.0f;
                 pts.push_back({-1.4f + 2.8f * t, 0.3f * std::sin(7.0f * t), 0.2f * std::cos(5.0f * t)});
             }
             std::vector<float> om(pts.size());
             std::vector<char> in(pts.size());
             e.solid_angle(pts.data(), pts.size(), om.data());
             e.is_inside(pts.data(), pts.size(), reinterpret_cast<bool*>(in.data()));
             for (size_t i = 0; i < pts.size(); ++i) {
                 CHECK(e.solid_angle(pts[i]) == om[i]);
                 CHECK(e.is_inside(pts[i]) == (in[i] != 0));
                 double r = std::sqrt(pts[i][0] * pts[i][0] + pts[i][1] * pts[i][1] + pts[i][2] * pts[i][2]);
                 if (r < 0.9) CHECK(in[i]);
                 if (r > 1.1) CHECK(!in[i]);
             }
         }},
        {"reference sampling loop: serial per-point is_inside equals one batched call", [] {
             // modules/winding/examples/sample_points_in_mesh.cpp:60-76 and the benchmark of
             // modules/winding/tests/test_fast_winding_number.cpp:99-107: std::mt19937 points
             // uniform in the bounding box, one is_inside call each
             FastWindingNumber e(sphere(16));
             std::uniform_real_distribution<float> px(-1.f, 1.f), py(-1.f, 1.f), pz(-1.f, 1.f);
             std::mt19937 gen;
             std::vector<std::array<float, 3>> pts(2000);
             for (auto& p : pts) p = {px(gen), py(gen), pz(gen)};
             std::vector<std::array<float, 3>> serial, batched;
             for (const auto& p : pts)
                 if (e.is_inside(p)) serial.push_back(p);
             std::vector<char> in(pts.size());
             e.is_inside(pts.data(), pts.size(), reinterpret_cast<bool*>(in.data()));
             for (size_t k = 0; k < pts.size(); ++k)
                 if (in[k]) batched.push_back(pts[k]);
             CHECK(serial == batched);
             CHECK(serial.size() > 600 && serial.size() < 1400);
         }},
        {"voxel signs: box and list equal per-point is_inside of the voxel centres", [] {
             // mesh_to_volume's interior test (modules/volume/src/mesh_to_volume.cpp:175-183):
             // is_inside(float(voxel_size * ijk + voxel_size / 2)) per voxel
             FastWindingNumber e(sphere(20));
             const double vs = 0.0731;
             const std::array<int32_t, 3> lo{-16, -15, -17}, dims{31, 29, 33};
             std::vector<char> box((size_t)dims[0] * dims[1] * dims[2]);
             e.is_inside_voxels(lo, dims, vs, reinterpret_cast<bool*>(box.data()));
             std::vector<std::array<int32_t, 3>> ijk;
             std::vector<std::array<float, 3>> centres;
             size_t t = 0, n_in = 0;
             for (int z = 0; z < dims[2]; ++z)
                 for (int y = 0; y < dims[1]; ++y)
                     for (int x = 0; x < dims[0]; ++x, ++t) {
                         std::array<int32_t, 3> v{lo[0] + x, lo[1] + y, lo[2] + z};
                         std::array<float, 3> c;
                         for (int d = 0; d < 3; ++d) {
                             volatile double w = v[d] * vs; // keep the multiply and add separate
                             c[d] = static_cast<float>(w + vs / 2.0);
                         }
                         if (t % 7 == 0) {
                             ijk.push_back(v);
                             centres.push_back(c);
                         }
                         n_in += box[t] != 0;
                         const double r = std::sqrt(double(c[0]) * c[0] + double(c[1]) * c[1] + double(c[2]) * c[2]);
                         if (r < 0.9) CHECK(box[t]);
                         if (r > 1.1) CHECK(!box[t]);
                     }
             CHECK(n_in > 0);
             std::vector<char> list(ijk.size()), ref(ijk.size());
             e.is_inside_voxels(ijk.data(), ijk.size(), vs, reinterpret_cast<bool*>(list.data()));
             e.is_inside(centres.data(), centres.size(), reinterpret_cast<bool*>(ref.data()));
             for (size_t i = 0; i < ijk.size(); ++i) {
                 CHECK(list[i] == ref[i]);
                 const size_t bt = (size_t)(ijk[i][0] - lo[0]) +
                                   (size_t)dims[0] * ((ijk[i][1] - lo[1]) + (size_t)dims[1] * (ijk[i][2] - lo[2]));
                 CHECK(box[bt] == ref[i]);
             }
             bool threw = false;
             try {
                 e.is_inside_voxels(lo, dims, -1.0, reinterpret_cast<bool*>(box.data()));
             } catch (const lagrange::Error&) {
                 threw = true;
             }
             CHECK(threw);
         }},
        {"closest point on the cube", [] {
             FastWindingNumber e(cube<float, uint32_t>());
             const std::array<float, 3> q[4] = {{0.5f, 0.5f, 2.0f}, {-1.0f, 0.5f, 0.5f}, {2.0f, 2.0f, 2.0f},
                                                {0.5f, 0.5f, 0.75f}};
             const double want[4] = {1.0, 1.0, 3.0, 0.0625};
             double d2[4];
             int32_t f[4];
             std::array<double, 3> cp[4];
             e.closest_point(q, 4, d2, f, cp);
             for (int i = 0; i < 4; ++i) {
                 CHECK(std::fabs(d2[i] - want[i]) < 1e-12);
                 CHECK(f[i] >= 0 && f[i] < 12);
             }
             CHECK(std::fabs(cp[0][2] - 1.0) < 1e-12 && std::fabs(cp[2][0] - 1.0) < 1e-12);
         }},
        {"engine options: accuracy scale, leaf buckets, Morton tree", [] {
             auto m = sphere(24);
             FastWindingNumber def(m), acc(m, {2.75f, 1, true}), fast(m, {2.0f, 8, false});
             // closed sphere: every variant is exact far from the surface and inside
             for (float z : {-0.5f, 0.0f, 0.5f}) {
                 CHECK(def.is_inside({0.1f, 0.2f, z}) && acc.is_inside({0.1f, 0.2f, z}) && fast.is_inside({0.1f, 0.2f, z}));
                 CHECK(!def.is_inside({1.5f, 0.2f, z}) && !acc.is_inside({1.5f, 0.2f, z}) && !fast.is_inside({1.5f, 0.2f, z}));
             }
             // near the surface (inside a closed sphere: exactly 4 pi) every variant stays within the
             // order-2 truncation bound, and the larger accuracy scale within 1e-3 in w
             double e_def = 0.0, e_acc = 0.0, e_fast = 0.0;
             for (int i = 0; i < 64; ++i) {
                 const float t = 0.1f * i, r = 0.97f;
                 const std::array<float, 3> p = {r * std::sin(t) * std::cos(2.3f * t), r * std::sin(t) * std::sin(2.3f * t),
                                                 r * std::cos(t)};
                 e_def = std::max(e_def, std::fabs(def.solid_angle(p) - 4 * kPi) / (4 * kPi));
                 e_acc = std::max(e_acc, std::fabs(acc.solid_angle(p) - 4 * kPi) / (4 * kPi));
                 e_fast = std::max(e_fast, std::fabs(fast.solid_angle(p) - 4 * kPi) / (4 * kPi));
             }
             std::printf("  max |dw| at 0.97 R: beta 2 %.2e, beta 2.75 %.2e, leaf 8 Morton %.2e\n", e_def, e_acc, e_fast);
             CHECK(e_def <= 4e-3 && e_fast <= 4e-3 && e_acc <= 1e-3);
             bool threw = false;
             try {
                 FastWindingNumber bad(m, {2.0f, 9, true});
             } catch (const lagrange::Error&) {
                 threw = true;
             }
             CHECK(threw);
         }},
        {"concurrent const queries", [] {
             FastWindingNumber e(sphere(16));
             std::vector<float> a(64), b(64);
             auto work = [&](std::vector<float>& out) {
                 for (int i = 0; i < 64; ++i) out[i] = e.solid_angle({0.02f * i - 0.6f, 0.1f, -0.2f});
             };
             std::thread t1(work, std::ref(a)), t2(work, std::ref(b));
             t1.join();
             t2.join();
             CHECK(a == b);
             // 16 threads of per-point calls (coalesced into batches by the engine) give the
             // values of one batched call
             const int T = 16, K = 200;
             std::vector<std::array<float, 3>> pts(T * K);
             for (int i = 0; i < T * K; ++i)
                 pts[i] = {0.9f * std::sin(0.37f * i), 0.9f * std::cos(0.61f * i), 0.8f * std::sin(0.13f * i + 1.0f)};
             std::vector<float> want(T * K), got(T * K);
             e.solid_angle(pts.data(), T * K, want.data());
             auto t0 = std::chrono::steady_clock::now();
             std::vector<std::thread> th;
             for (int t = 0; t < T; ++t)
                 th.emplace_back([&, t] {
                     for (int k = 0; k < K; ++k) got[t * K + k] = e.solid_angle(pts[t * K + k]);
                 });
             for (auto& x : th) x.join();
             const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
             CHECK(got == want);
             std::printf("  16 threads x %d per-point calls: %.0f points/s\n", K, T * K / s);
         }},
    };
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    for (auto& c : cases) {
        std::printf("RUN  %s\n", c.first.c_str());
        int before = g_fail;
        try {
            c.second();
        } catch (const std::exception& ex) {
            std::printf("  exception: %s\n", ex.what());
            ++g_fail;
        }
        std::printf("%s %s\n", g_fail == before ? "PASS" : "FAIL", c.first.c_str());
    }
    std::printf("%s\n", g_fail ? "FAILED" : "ALL PASSED");
    return g_fail ? 1 : 0;
}
