// fix_orientation: flips the facets of an input mesh whose orientation disagrees with a reference
// mesh, as modules/winding/examples/fix_orientation.cpp does, with the per-facet probe pairs
// evaluated in one batched GPU call.
//
//   fix_orientation INPUT REFERENCE [OUTPUT] [-e EPSILON=1e-2] [-t THRESHOLD=0.8]
//                   [--criterion-json PATH]
#include <lagrange/utils/Error.h>
#include <lagrange/winding/FastWindingNumber.h>

#include <fwn_io.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

namespace {

struct Loaded {
    std::vector<std::array<float, 3>> V;
    std::vector<std::array<int32_t, 3>> F;
};

Loaded load(const std::string& path)
{
    fwn_mesh* m = nullptr;
    if (fwn_mesh_load(path.c_str(), FWN_IO_FORMAT_AUTO, 1, &m) != FWN_IO_OK)
        throw lagrange::Error("cannot load " + path + ": " + fwn_io_last_error());
    int64_t nv = 0, nf = 0, nc = 0;
    fwn_mesh_sizes(m, &nv, &nf, &nc);
    std::vector<double> V((size_t)nv * 3);
    std::vector<int64_t> off((size_t)nf + 1);
    std::vector<int32_t> cor((size_t)nc);
    fwn_mesh_copy(m, V.data(), off.data(), cor.data());
    fwn_mesh_free(m);
    Loaded out;
    out.V.resize((size_t)nv);
    for (int64_t i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) out.V[i][d] = static_cast<float>(V[3 * i + d]);
    out.F.resize((size_t)nf);
    for (int64_t f = 0; f < nf; ++f)
        for (int d = 0; d < 3; ++d) out.F[f][d] = cor[3 * f + d];
    return out;
}

lagrange::SurfaceMesh<float, uint32_t> to_mesh(const Loaded& l)
{
    lagrange::SurfaceMesh<float, uint32_t> m;
    for (const auto& p : l.V) m.add_vertex(p.data());
    for (const auto& t : l.F) m.add_triangle((uint32_t)t[0], (uint32_t)t[1], (uint32_t)t[2]);
    return m;
}

} // namespace

int main(int argc, char** argv)
{
    std::vector<std::string> pos;
    float epsilon = 1e-2f, threshold = 8e-1f;
    std::string json_path;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if ((a == "-e" || a == "--epsilon") && i + 1 < argc) epsilon = std::strtof(argv[++i], nullptr);
        else if ((a == "-t" || a == "--threshold") && i + 1 < argc) threshold = std::strtof(argv[++i], nullptr);
        else if (a == "--criterion-json" && i + 1 < argc) json_path = argv[++i];
        else pos.push_back(a);
    }
    if (pos.size() < 2 || pos.size() > 3) {
        std::fprintf(stderr, "usage: %s INPUT REFERENCE [OUTPUT] [-e EPS] [-t THRESHOLD] [--criterion-json PATH]\n",
                     argv[0]);
        return 2;
    }
    try {
        Loaded input = load(pos[0]);
        const lagrange::winding::FastWindingNumber engine(to_mesh(load(pos[1])));
        std::vector<double> crit(input.F.size());
        engine.orientation_criterion(input.V.data(), input.V.size(), input.F.data(), input.F.size(), epsilon,
                                     crit.data());
        size_t count_positive = 0, count_negative = 0;
        for (size_t f = 0; f < crit.size(); ++f) {
            if (crit[f] > threshold) {
                std::swap(input.F[f][0], input.F[f][1]);
                ++count_positive;
            }
            if (crit[f] < threshold) ++count_negative;
        }
        std::printf("pos %zu neg %zu tot %zu\n", count_positive, count_negative, crit.size());
        if (!json_path.empty()) {
            std::ofstream j(json_path);
            j << "{\"criterions\":[";
            char buf[40];
            for (size_t f = 0; f < crit.size(); ++f) {
                std::snprintf(buf, sizeof(buf), "%.17g", crit[f]);
                j << (f ? "," : "") << buf;
            }
            j << "],\"epsilon\":" << epsilon << ",\"threshold\":" << threshold << "}";
        }
        if (pos.size() == 3) {
            std::vector<double> V(input.V.size() * 3);
            for (size_t i = 0; i < input.V.size(); ++i)
                for (int d = 0; d < 3; ++d) V[3 * i + d] = input.V[i][d];
            std::vector<int64_t> off(input.F.size() + 1);
            for (size_t f = 0; f <= input.F.size(); ++f) off[f] = 3 * (int64_t)f;
            if (fwn_mesh_save(pos[2].c_str(), pos[2].size() > 4 && pos[2].substr(pos[2].size() - 4) == ".obj"
                                                  ? FWN_IO_FORMAT_OBJ
                                                  : FWN_IO_FORMAT_PLY_BINARY,
                              V.data(), (int64_t)input.V.size(), off.data(),
                              input.F.empty() ? nullptr : input.F[0].data(), (int64_t)input.F.size()) != FWN_IO_OK)
                throw lagrange::Error(std::string("cannot save: ") + fwn_io_last_error());
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
