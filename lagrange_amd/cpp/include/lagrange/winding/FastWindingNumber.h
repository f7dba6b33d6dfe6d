// lagrange::winding::FastWindingNumber backed by the MI355X engine (include/fwn.h).
//
// The class keeps the reference declaration (modules/winding/include/lagrange/winding/
// FastWindingNumber.h:24-108): constructor from SurfaceMesh<Scalar, Index> (float/double x
// uint32/uint64), default constructor, move-only, const is_inside/solid_angle per point, PIMPL
// through value_ptr<Impl>.  The batched overloads below are additions; they take plain arrays of
// points and run one GPU launch for all of them.
#pragma once

#include <lagrange/SurfaceMesh.h>

#include <array>
#include <cstddef>
#include <cstdint>

namespace lagrange {

namespace winding {

///
/// Engine options (an addition; the reference engine always uses its defaults, which these are).
///
struct FastWindingNumberOptions
{
    /// beta of the far-field criterion |q - p|^2 > beta^2 r^2 (reference default 2; 2.75 keeps the
    /// winding number within 1e-3 of the exact sum on the benchmark meshes)
    float accuracy_scale = 2.0f;
    /// triangles per leaf bucket, 1..8 (1: the reference's single-triangle items)
    int leaf_max = 1;
    /// true: the reference's box-area tree (top-down SAH); false: Morton LBVH
    bool sah_tree = true;
};

///
/// Fast winding number computation for triangle soups.
///
class FastWindingNumber
{
public:
    ///
    /// Builds the acceleration structure of a triangle mesh (positions converted to float and
    /// indices to int, as the reference engine does).  Throws lagrange::Error if the mesh is not
    /// 3D, not a triangle mesh, or the GPU engine cannot be created.
    ///
    template <typename Scalar, typename Index>
    FastWindingNumber(const SurfaceMesh<Scalar, Index>& mesh);

    /// Same with explicit engine options (an addition to the reference API).
    template <typename Scalar, typename Index>
    FastWindingNumber(const SurfaceMesh<Scalar, Index>& mesh, const FastWindingNumberOptions& options);

    FastWindingNumber();
    ~FastWindingNumber();
    FastWindingNumber(FastWindingNumber&& other) noexcept;
    FastWindingNumber& operator=(FastWindingNumber&& other) noexcept;
    FastWindingNumber(const FastWindingNumber& other) = delete;
    FastWindingNumber& operator=(const FastWindingNumber& other) = delete;

    /// True if the winding number at pos exceeds 1/2 (solid angle / 4 pi > 0.5, in double).
    bool is_inside(const std::array<float, 3>& pos) const;

    /// Solid angle (steradians) at pos.
    float solid_angle(const std::array<float, 3>& pos) const;

    /// Batched solid angles of n points (row-major n x 3); out has n entries.
    void solid_angle(const std::array<float, 3>* pos, size_t n, float* out) const;

    /// Batched inside tests of n points; out has n entries.
    void is_inside(const std::array<float, 3>* pos, size_t n, bool* out) const;

    /// Inside tests of voxel centres float(voxel_size * ijk + voxel_size / 2), the per-voxel
    /// interior test of volume::mesh_to_volume with Sign::WindingNumber, for every voxel of the
    /// index box lo + [0, dims) (x fastest; out has dims[0] * dims[1] * dims[2] entries).  The
    /// centres are generated on the GPU.  voxel_size is absolute (> 0).
    void is_inside_voxels(const std::array<int32_t, 3>& lo, const std::array<int32_t, 3>& dims, double voxel_size,
                          bool* out) const;

    /// Same for n explicit voxel indices.
    void is_inside_voxels(const std::array<int32_t, 3>* ijk, size_t n, double voxel_size, bool* out) const;

    /// Orientation criterion of every facet of an input triangle mesh (V: nv points, F: nf
    /// triangles) against this mesh: (solid_angle(g + eps n) - solid_angle(g - eps n)) / (4 pi)
    /// at the barycentre g along the unit normal n, as the fix_orientation example computes per
    /// facet.  out has nf entries.
    void orientation_criterion(const std::array<float, 3>* V, size_t nv, const std::array<int32_t, 3>* F, size_t nf,
                               float epsilon, double* out) const;

    /// Draws n points uniform in the box [lo, hi] (counter-based generator of seed and sample
    /// index) and writes those inside the mesh to out (capacity n) in sample order, as the
    /// sample_points_in_mesh example does; returns how many were written.
    size_t sample_inside(const std::array<float, 3>& lo, const std::array<float, 3>& hi, size_t n, uint64_t seed,
                         std::array<float, 3>* out) const;

    /// Closest point on the mesh of n points (bvh::TriangleAABBTree::get_closest_point, batched):
    /// squared distance, facet index and (when closest is not null) the closest point.
    void closest_point(const std::array<float, 3>* pos, size_t n, double* sq_dist, int32_t* facet,
                       std::array<double, 3>* closest) const;

protected:
    /// Internal implementation.
    struct Impl;

    /// PIMPL to hide internal data structures.
    value_ptr<Impl> m_impl;
};

} // namespace winding
} // namespace lagrange
