// Minimal stand-in for modules/core/include/lagrange/SurfaceMeshTypes.h, used only when the real
// Lagrange headers are not on the include path: the X macro over the four SurfaceMesh
// instantiations (Scalar float/double x Index uint32_t/uint64_t, SurfaceMeshTypes.h:49-53).
#pragma once

#include <cstdint>

#define LA_SURFACE_MESH_X(mode, data)      \
    LA_X_##mode(data, float, uint32_t)     \
    LA_X_##mode(data, double, uint32_t)    \
    LA_X_##mode(data, float, uint64_t)     \
    LA_X_##mode(data, double, uint64_t)
