// sample_points_in_mesh: uniform samples in a mesh's bounding box kept when inside, written as
// .xyz, with the output of modules/winding/examples/sample_points_in_mesh.cpp:
//  * the samples are the reference's sequence: std::mt19937 with its default seed and one
//    std::uniform_real_distribution<float> per axis over the float bounding box, drawn x, y, z
//    per sample (:60-63, :69-71);
//  * the reference tests them one is_inside call at a time (:72-75); here all of them go to the
//    GPU in one batched is_inside call, and the kept points stay in sample order;
//  * each kept point is one line "x y z" in fmt's "{}" float form (:80-83).
//
//   sample_points_in_mesh INPUT [OUTPUT=output.xyz] [-n NUM_SAMPLES=10000]
#include <lagrange/utils/Error.h>
#include <lagrange/winding/FastWindingNumber.h>

#include <fwn_io.h>

#include "format_float.h"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
    std::vector<std::string> pos;
    size_t num_samples = 10000;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if ((a == "-n" || a == "--num-samples") && i + 1 < argc) num_samples = std::strtoull(argv[++i], nullptr, 10);
        else pos.push_back(a);
    }
    if (pos.empty() || pos.size() > 2) {
        std::fprintf(stderr, "usage: %s INPUT [OUTPUT.xyz] [-n NUM_SAMPLES]\n", argv[0]);
        return 2;
    }
    const std::string output = pos.size() == 2 ? pos[1] : "output.xyz";
    if (output.size() < 4 || output.substr(output.size() - 4) != ".xyz")
        std::fprintf(stderr, "Output file extension should be .xyz. '%s' was given.\n", output.c_str());
    try {
        fwn_mesh* m = nullptr;
        if (fwn_mesh_load(pos[0].c_str(), FWN_IO_FORMAT_AUTO, 1, &m) != FWN_IO_OK)
            throw lagrange::Error(std::string("cannot load: ") + fwn_io_last_error());
        int64_t nv = 0, nf = 0, nc = 0;
        fwn_mesh_sizes(m, &nv, &nf, &nc);
        std::vector<double> V((size_t)nv * 3);
        std::vector<int32_t> cor((size_t)nc);
        fwn_mesh_copy(m, V.data(), nullptr, cor.data());
        fwn_mesh_free(m);
        // the reference loads a SurfaceMesh<float, uint32_t>: positions and the box are float
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

        lagrange::winding::FastWindingNumber engine(mesh);

        std::mt19937 gen;
        std::vector<std::array<float, 3>> samples(num_samples);
        for (size_t k = 0; k < num_samples; ++k) samples[k] = {px(gen), py(gen), pz(gen)}; // x, y, z in order
        std::vector<char> inside(num_samples);
        engine.is_inside(samples.data(), num_samples, reinterpret_cast<bool*>(inside.data()));

        FILE* out = std::fopen(output.c_str(), "w");
        if (!out) throw lagrange::Error("cannot write " + output);
        size_t kept = 0;
        for (size_t k = 0; k < num_samples; ++k) {
            if (!inside[k]) continue;
            const auto& p = samples[k];
            std::fprintf(out, "%s %s %s\n", fwn_examples::format_float(p[0]).c_str(),
                         fwn_examples::format_float(p[1]).c_str(), fwn_examples::format_float(p[2]).c_str());
            ++kept;
        }
        std::fclose(out);
        std::printf("%zu of %zu samples inside\n", kept, num_samples);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
