// Minimal stand-in for lagrange::Error (modules/core/include/lagrange/utils/Error.h:26), used only
// when the real Lagrange headers are not on the include path: the base exception of Lagrange
// functions, a std::runtime_error.
#pragma once

#include <stdexcept>

namespace lagrange {

struct Error : public std::runtime_error
{
    using std::runtime_error::runtime_error;
};

} // namespace lagrange
