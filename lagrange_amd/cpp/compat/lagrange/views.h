// Minimal stand-in for modules/core/include/lagrange/views.h, used only when the real Lagrange
// headers are not on the include path: read-only row-major views of a mesh's positions and facets
// (vertex_view / facet_view, views.h:347,375; modules/core/src/views.cpp:155-175).  The real views
// are Eigen::Map objects; this one offers the same rows(), cols() and operator()(i, j).
#pragma once

#include <lagrange/SurfaceMesh.h>

#include <cstddef>

namespace lagrange {

/// Row-major (rows x cols) read-only view.
template <typename T>
struct RowMajorView
{
    const T* data;
    std::ptrdiff_t nrows, ncols;
    std::ptrdiff_t rows() const { return nrows; }
    std::ptrdiff_t cols() const { return ncols; }
    const T& operator()(std::ptrdiff_t i, std::ptrdiff_t j) const { return data[i * ncols + j]; }
};

template <typename Scalar, typename Index>
RowMajorView<Scalar> vertex_view(const SurfaceMesh<Scalar, Index>& mesh)
{
    return {mesh.position_data(), (std::ptrdiff_t)mesh.get_num_vertices(), (std::ptrdiff_t)mesh.get_dimension()};
}

/// Facet view of a mesh whose facets all have the same size (as the reference's facet_view).
template <typename Scalar, typename Index>
RowMajorView<Index> facet_view(const SurfaceMesh<Scalar, Index>& mesh)
{
    const Index nf = mesh.get_num_facets();
    const Index k = nf ? mesh.get_facet_size(0) : 3;
    return {mesh.corner_data(), (std::ptrdiff_t)nf, (std::ptrdiff_t)k};
}

} // namespace lagrange
