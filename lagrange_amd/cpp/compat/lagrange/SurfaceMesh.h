// Minimal stand-in for lagrange::SurfaceMesh, used only when the real Lagrange headers are not on
// the include path.  It offers exactly what the winding engine's constructor consumes
// (modules/core/include/lagrange/SurfaceMesh.h:72-90 type constraints; get_dimension,
// is_triangle_mesh; row-major vertex/facet views as modules/core/src/views.cpp:155-175):
// positions stored row-major (num_vertices x dim), facets with a uniform or mixed vertex count.
#pragma once

#include <lagrange/utils/value_ptr.h>

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace lagrange {

/// Error raised on invalid input (counterpart of lagrange::Error, a std::runtime_error).
class Error : public std::runtime_error
{
public:
    using std::runtime_error::runtime_error;
};

template <typename Scalar_, typename Index_>
class SurfaceMesh
{
    static_assert(std::is_same_v<Scalar_, float> || std::is_same_v<Scalar_, double>, "Scalar must be float or double");
    static_assert(std::is_same_v<Index_, uint32_t> || std::is_same_v<Index_, uint64_t>,
                  "Index must be uint32_t or uint64_t");

public:
    using Scalar = Scalar_;
    using Index = Index_;

    explicit SurfaceMesh(Index dimension = 3) : m_dim(dimension) {}

    Index get_dimension() const { return m_dim; }
    Index get_num_vertices() const { return static_cast<Index>(m_pos.size() / m_dim); }
    Index get_num_facets() const { return static_cast<Index>(m_offsets.size() - 1); }
    Index get_facet_size(Index f) const { return static_cast<Index>(m_offsets[f + 1] - m_offsets[f]); }

    bool is_triangle_mesh() const
    {
        for (size_t f = 0; f + 1 < m_offsets.size(); ++f)
            if (m_offsets[f + 1] - m_offsets[f] != 3) return false;
        return true;
    }

    void add_vertex(const Scalar* p)
    {
        for (Index d = 0; d < m_dim; ++d) m_pos.push_back(p[d]);
    }
    void add_vertex(std::initializer_list<Scalar> p)
    {
        if (p.size() != m_dim) throw Error("add_vertex: wrong dimension");
        m_pos.insert(m_pos.end(), p.begin(), p.end());
    }
    void add_vertices(Index n, const Scalar* coords) { m_pos.insert(m_pos.end(), coords, coords + n * m_dim); }
    void add_triangle(Index a, Index b, Index c) { add_polygon({a, b, c}); }
    void add_quad(Index a, Index b, Index c, Index d) { add_polygon({a, b, c, d}); }
    void add_polygon(std::initializer_list<Index> v)
    {
        m_corners.insert(m_corners.end(), v.begin(), v.end());
        m_offsets.push_back(m_corners.size());
    }
    void add_triangles(Index n, const Index* tri)
    {
        for (Index f = 0; f < n; ++f) add_triangle(tri[3 * f], tri[3 * f + 1], tri[3 * f + 2]);
    }

    const Scalar* position_data() const { return m_pos.data(); }
    const Index* corner_data() const { return m_corners.data(); }

private:
    Index m_dim;
    std::vector<Scalar> m_pos;
    std::vector<Index> m_corners;
    std::vector<size_t> m_offsets{0};
};

/// Row-major (rows x cols) read-only view with Eigen-like rows()/cols()/operator()(i, j).
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
