// Shortest round-trip text of a float in the form fmt's "{}" writes it (the reference examples
// write their points with fmt::format("{}", fmt::join(p, " ")), e.g.
// modules/winding/examples/sample_points_in_mesh.cpp:80-83; Lagrange pins fmt 12.1.0).
//
// The digits are the shortest decimal string that reads back to the same float
// (std::to_chars); the layout follows fmt's rule for a value with decimal exponent x (d.ddd x
// 10^x): exponent form d[.ddd]e+XX / e-XX when x < -4 or x >= 7 (digits10 + 1 of float), else
// fixed form with no trailing ".0" (1 -> "1", 0.5 -> "0.5", 100000 -> "100000").
#pragma once

#include <charconv>
#include <cstdlib>
#include <string>

namespace fwn_examples {

inline std::string format_float(float v)
{
    char buf[64];
    // scientific shortest form: [-]d[.ddd]e(+|-)XX
    auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
    std::string s(buf, r.ptr);
    if (s == "inf" || s == "-inf" || s == "nan" || s == "-nan") return s;
    const bool neg = s[0] == '-';
    if (neg) s.erase(0, 1);
    const size_t e = s.find('e');
    std::string digits = s.substr(0, e);
    const int x = std::atoi(s.c_str() + e + 1);
    if (digits.size() > 1) digits.erase(1, 1); // drop the decimal point: d ddd
    std::string out;
    if (x < -4 || x >= 7) {
        out = digits.substr(0, 1);
        if (digits.size() > 1) out += "." + digits.substr(1);
        out += x < 0 ? "e-" : "e+";
        const int ax = x < 0 ? -x : x;
        if (ax < 10) out += "0";
        out += std::to_string(ax);
    } else if (x >= 0) {
        if ((int)digits.size() <= x + 1) {
            out = digits + std::string(x + 1 - digits.size(), '0');
        } else {
            out = digits.substr(0, x + 1) + "." + digits.substr(x + 1);
        }
    } else {
        out = "0." + std::string(-x - 1, '0') + digits;
    }
    return neg ? "-" + out : out;
}

} // namespace fwn_examples
