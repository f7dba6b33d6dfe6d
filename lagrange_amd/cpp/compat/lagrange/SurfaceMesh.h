// Minimal stand-in for lagrange::SurfaceMesh, used only when the real Lagrange headers are not on
// the include path.  It offers exactly what the winding engine's constructor consumes
// (modules/core/include/lagrange/SurfaceMesh.h:72-90 type constraints; get_dimension,
// is_triangle_mesh): positions stored row-major (num_vertices x dim), facets with a uniform or
// mixed vertex count.  Like the real header (SurfaceMesh.h:14-18) it declares neither
// lagrange::Error (<lagrange/utils/Error.h>) nor the matrix views (<lagrange/views.h>).
#pragma once

#include <lagrange/utils/value_ptr.h>

#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace lagrange {

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
        if (p.size() != m_dim) throw std::invalid_argument("add_vertex: wrong dimension");
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

} // namespace lagrange
