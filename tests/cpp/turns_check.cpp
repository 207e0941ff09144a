// Bounds the error of ucmc::sincos_turn / sincos_turn_any (the O(N) and XY
// proposals' sin and cos of 2 pi t) against double sin / cos: every 977th
// 32-bit draw t = k 2^-32 as the proposals use it, and angles in turns over
// [-2, 2].  Prints the largest absolute errors and exits non-zero above 2e-7.
#include <cmath>
#include <cstdint>
#include <cstdio>

#include "turns.hpp"

int main()
{
    double es = 0.0, ec = 0.0;
    for (std::uint64_t k = 0; k < (1ull << 32); k += 977) {
        const float t = static_cast<float>(k) * 0x1.0p-32f;
        float s, c;
        ucmc::sincos_turn(t, &s, &c);
        const double a = 2.0 * M_PI * static_cast<double>(t);
        es = std::fmax(es, std::fabs(s - std::sin(a)));
        ec = std::fmax(ec, std::fabs(c - std::cos(a)));
    }
    double ea = 0.0;
    for (int k = -200000; k <= 200000; ++k) {
        const float t = static_cast<float>(k) * 1e-5f;
        float s, c;
        ucmc::sincos_turn_any(t, &s, &c);
        const double a = 2.0 * M_PI * static_cast<double>(t);
        ea = std::fmax(ea, std::fmax(std::fabs(s - std::sin(a)), std::fabs(c - std::cos(a))));
    }
    std::printf("sincos_turn max error: sin %.3g cos %.3g; any-turn %.3g\n", es, ec, ea);
    return (es > 2e-7 || ec > 2e-7 || ea > 2e-7) ? 1 : 0;
}
