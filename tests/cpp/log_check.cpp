// Exhaustive check of wobtrig::log_cr (wobtoolbox_amd/csrc/wob_trig.hpp) against the C
// library's extended-precision (64-bit significand) log rounded to float, over every float in
// [2^-26, 2^12): the arguments of the 2D fundamental solution's logarithm, max(|r|, eps), in
// every scene this engine takes.  Prints the number of floats that differ and the number
// checked.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../wobtoolbox_amd/csrc/wob_trig.hpp"

int main() {
    float lo = std::ldexp(1.0f, -26), hi = std::ldexp(1.0f, 12);
    uint32_t a, b;
    std::memcpy(&a, &lo, 4);
    std::memcpy(&b, &hi, 4);
    long bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (long k = (long)a; k < (long)b; ++k) {
        const uint32_t u = (uint32_t)k;
        float x;
        std::memcpy(&x, &u, 4);
        if (wobtrig::log_cr(x) != (float)logl((long double)x)) ++bad;
    }
    std::printf("%ld %ld\n", bad, (long)b - (long)a);
    return 0;
}
