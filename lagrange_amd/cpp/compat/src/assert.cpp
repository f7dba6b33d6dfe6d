// Failure handler of the compat la_runtime_assert (compat/lagrange/utils/assert.h).  In a real
// Lagrange build lagrange::core provides it (modules/core/src/utils/assert.cpp:75-97); the message
// has the same shape: condition, optional reason, file, line and function.
#include <lagrange/utils/Error.h>
#include <lagrange/utils/assert.h>

#include <string>

namespace lagrange {

bool assertion_failed(const char* function, const char* file, unsigned int line, const char* condition,
                      std::string_view message)
{
    std::string msg = "Assertion failed: \"";
    msg += condition;
    msg += "\"";
    if (!message.empty()) {
        msg += ", reason: ";
        msg += message;
    }
    msg += "\n\tIn file: ";
    msg += file;
    msg += ", line " + std::to_string(line) + ";\n\tIn function: ";
    msg += function;
    msg += ";";
    throw Error(msg);
}

} // namespace lagrange
