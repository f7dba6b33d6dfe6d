// lagrange::winding::FastWindingNumber over the engine's C ABI.
//
// Mirrors modules/winding/src/FastWindingNumber.cpp: the constructor asserts a 3D triangle mesh
// (:91-96), copies positions to float and indices to int (:39-52) and builds the engine
// (:57); is_inside is solid_angle / (4 pi) > 0.5 in double (:66).  Here the build and every
// query run on the GPU through fwn_create / fwn_solid_angle*.  The includes are those a real
// Lagrange build resolves (SurfaceMeshTypes.h for the X macro, views.h, utils/assert.h,
// utils/Error.h, as the reference source's :13-18); the compat headers mirror that layout.
#include <lagrange/SurfaceMeshTypes.h>
#include <lagrange/utils/Error.h>
#include <lagrange/utils/assert.h>
#include <lagrange/views.h>
#include <lagrange/winding/FastWindingNumber.h>

#include <fwn.h>

#include <string>
#include <vector>

namespace lagrange {
namespace winding {

namespace {
[[noreturn]] void raise(const char* what, int status)
{
    throw Error(std::string(what) + " failed (status " + std::to_string(status) + "): " + fwn_last_error());
}
} // namespace

struct FastWindingNumber::Impl
{
    fwn_engine* engine = nullptr;

    Impl() = default;
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl()
    {
        if (engine) fwn_destroy(engine);
    }

    template <typename VView, typename FView>
    void initialize(const VView& vertices, const FView& facets, const fwn_options* opt = nullptr)
    {
        la_runtime_assert(vertices.cols() == 3);
        la_runtime_assert(facets.cols() == 3);
        std::vector<float> v(static_cast<size_t>(vertices.rows()) * 3);
        for (std::ptrdiff_t i = 0; i < (std::ptrdiff_t)vertices.rows(); ++i)
            for (int d = 0; d < 3; ++d) v[3 * i + d] = static_cast<float>(vertices(i, d));
        std::vector<int32_t> f(static_cast<size_t>(facets.rows()) * 3);
        for (std::ptrdiff_t i = 0; i < (std::ptrdiff_t)facets.rows(); ++i)
            for (int d = 0; d < 3; ++d) f[3 * i + d] = static_cast<int32_t>(facets(i, d));
        int st = fwn_create(&engine, v.data(), (int64_t)vertices.rows(), f.data(), (int64_t)facets.rows(), opt);
        if (st != FWN_OK) raise("fwn_create", st);
    }

};

// The engine of an initialised object; a default-constructed or moved-from object has no Impl
// (the reference leaves such calls undefined; here they throw).
#define FWN_ENGINE_OF(impl)                                                                    \
    ((impl) && (impl)->engine ? (impl)->engine                                                 \
                              : throw Error("FastWindingNumber is not initialised"))

template <typename Scalar, typename Index>
FastWindingNumber::FastWindingNumber(const SurfaceMesh<Scalar, Index>& mesh) : m_impl(make_value_ptr<Impl>())
{
    la_runtime_assert(mesh.get_dimension() == 3, "Fast winding number engine only supports 3D meshes");
    la_runtime_assert(mesh.is_triangle_mesh(), "Fast winding number engine only supports triangle meshes");
    m_impl->initialize(vertex_view(mesh), facet_view(mesh));
}

template <typename Scalar, typename Index>
FastWindingNumber::FastWindingNumber(const SurfaceMesh<Scalar, Index>& mesh, const FastWindingNumberOptions& options)
    : m_impl(make_value_ptr<Impl>())
{
    la_runtime_assert(mesh.get_dimension() == 3, "Fast winding number engine only supports 3D meshes");
    la_runtime_assert(mesh.is_triangle_mesh(), "Fast winding number engine only supports triangle meshes");
    fwn_options o;
    fwn_options_default(&o);
    o.accuracy_scale = options.accuracy_scale;
    o.leaf_max = options.leaf_max;
    o.tree_builder = options.sah_tree ? FWN_TREE_SAH : FWN_TREE_LBVH;
    m_impl->initialize(vertex_view(mesh), facet_view(mesh), &o);
}

FastWindingNumber::FastWindingNumber() = default;
FastWindingNumber::~FastWindingNumber() = default;
FastWindingNumber::FastWindingNumber(FastWindingNumber&& other) noexcept = default;
FastWindingNumber& FastWindingNumber::operator=(FastWindingNumber&& other) noexcept = default;

bool FastWindingNumber::is_inside(const std::array<float, 3>& pos) const
{
    uint8_t r = 0;
    int st = fwn_is_inside_point(FWN_ENGINE_OF(m_impl), pos.data(), &r);
    if (st != FWN_OK) raise("fwn_is_inside_point", st);
    return r != 0;
}

float FastWindingNumber::solid_angle(const std::array<float, 3>& pos) const
{
    float r = 0.0f;
    int st = fwn_solid_angle_point(FWN_ENGINE_OF(m_impl), pos.data(), &r);
    if (st != FWN_OK) raise("fwn_solid_angle_point", st);
    return r;
}

void FastWindingNumber::solid_angle(const std::array<float, 3>* pos, size_t n, float* out) const
{
    int st = fwn_solid_angle(FWN_ENGINE_OF(m_impl), pos ? pos->data() : nullptr, (int64_t)n, out, nullptr, 0);
    if (st != FWN_OK) raise("fwn_solid_angle", st);
}

void FastWindingNumber::is_inside(const std::array<float, 3>* pos, size_t n, bool* out) const
{
    fwn_engine* eng = FWN_ENGINE_OF(m_impl);
    std::vector<float> om(n);
    std::vector<uint8_t> in(n);
    int st = fwn_solid_angle(eng, pos ? pos->data() : nullptr, (int64_t)n, om.data(), in.data(), 0);
    if (st != FWN_OK) raise("fwn_solid_angle", st);
    for (size_t i = 0; i < n; ++i) out[i] = in[i] != 0;
}

void FastWindingNumber::is_inside_voxels(const std::array<int32_t, 3>& lo, const std::array<int32_t, 3>& dims,
                                         double voxel_size, bool* out) const
{
    fwn_engine* eng = FWN_ENGINE_OF(m_impl);
    size_t n = 1;
    for (int d = 0; d < 3; ++d) n *= dims[d] > 0 ? (size_t)dims[d] : 0;
    std::vector<uint8_t> in(n);
    int st = fwn_sign_voxel_box(eng, lo.data(), dims.data(), voxel_size, in.data(), nullptr, 0);
    if (st != FWN_OK) raise("fwn_sign_voxel_box", st);
    for (size_t i = 0; i < n; ++i) out[i] = in[i] != 0;
}

void FastWindingNumber::is_inside_voxels(const std::array<int32_t, 3>* ijk, size_t n, double voxel_size,
                                         bool* out) const
{
    fwn_engine* eng = FWN_ENGINE_OF(m_impl);
    std::vector<uint8_t> in(n);
    int st = fwn_sign_voxels(eng, ijk ? ijk->data() : nullptr, (int64_t)n, voxel_size, in.data(), nullptr, 0);
    if (st != FWN_OK) raise("fwn_sign_voxels", st);
    for (size_t i = 0; i < n; ++i) out[i] = in[i] != 0;
}

void FastWindingNumber::orientation_criterion(const std::array<float, 3>* V, size_t nv,
                                              const std::array<int32_t, 3>* F, size_t nf, float epsilon,
                                              double* out) const
{
    int st = fwn_orientation_criterion(FWN_ENGINE_OF(m_impl), V ? V->data() : nullptr, (int64_t)nv,
                                       F ? F->data() : nullptr, (int64_t)nf, epsilon, out, 0);
    if (st != FWN_OK) raise("fwn_orientation_criterion", st);
}

size_t FastWindingNumber::sample_inside(const std::array<float, 3>& lo, const std::array<float, 3>& hi, size_t n,
                                        uint64_t seed, std::array<float, 3>* out) const
{
    int64_t kept = 0;
    int st = fwn_sample_inside(FWN_ENGINE_OF(m_impl), lo.data(), hi.data(), (int64_t)n, seed,
                               out ? out->data() : nullptr, &kept, 0);
    if (st != FWN_OK) raise("fwn_sample_inside", st);
    return (size_t)kept;
}

void FastWindingNumber::closest_point(const std::array<float, 3>* pos, size_t n, double* sq_dist, int32_t* facet,
                                      std::array<double, 3>* closest) const
{
    int st = fwn_closest_point(FWN_ENGINE_OF(m_impl), pos ? pos->data() : nullptr, (int64_t)n, sq_dist, facet,
                               closest ? closest->data() : nullptr, 0);
    if (st != FWN_OK) raise("fwn_closest_point", st);
}

// float/double x uint32/uint64 (modules/core/include/lagrange/SurfaceMeshTypes.h:49-53)
#define LA_X_fast_winding_number(_, Scalar, Index)                                                     \
    template FastWindingNumber::FastWindingNumber(const SurfaceMesh<Scalar, Index>& mesh);            \
    template FastWindingNumber::FastWindingNumber(const SurfaceMesh<Scalar, Index>& mesh,             \
                                                  const FastWindingNumberOptions& options);
LA_SURFACE_MESH_X(fast_winding_number, 0)

} // namespace winding
} // namespace lagrange
