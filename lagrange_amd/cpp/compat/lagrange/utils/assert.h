// Minimal stand-in for la_runtime_assert (modules/core/include/lagrange/utils/assert.h:173-176),
// used only when the real Lagrange headers are not on the include path.  As in the real header,
// this file does not declare lagrange::Error: the failure handler (compat/src/assert.cpp, in the
// real build modules/core/src/utils/assert.cpp:75-97) throws it, so code that throws Error itself
// must include <lagrange/utils/Error.h>.
#pragma once

#include <string_view>

namespace lagrange {

/// Throws lagrange::Error describing the failed condition; never returns.
[[noreturn]] bool assertion_failed(const char* function, const char* file, unsigned int line, const char* condition,
                                   std::string_view message);

} // namespace lagrange

#define LA_COMPAT_ASSERT_2(condition, message)                                                               \
    ((condition) ? (void)0                                                                                   \
                 : (void)::lagrange::assertion_failed(__func__, __FILE__, __LINE__, #condition, (message)))
#define LA_COMPAT_ASSERT_1(condition) LA_COMPAT_ASSERT_2(condition, "")
#define LA_COMPAT_ASSERT_PICK(_1, _2, NAME, ...) NAME

/// la_runtime_assert(condition) or la_runtime_assert(condition, message)
#ifndef la_runtime_assert
#define la_runtime_assert(...) LA_COMPAT_ASSERT_PICK(__VA_ARGS__, LA_COMPAT_ASSERT_2, LA_COMPAT_ASSERT_1, )(__VA_ARGS__)
#endif
