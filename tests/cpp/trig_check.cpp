// Exhaustive check of wobtrig::sincos_small (wobtoolbox_amd/csrc/wob_trig.hpp) against the C
// library's double sin and cos, rounded to float, over every angle the walk can draw: the
// (0, 1] uniforms are (k + 1) / 2^24, k < 2^24, and the arguments are
//   2D line direction   (double)theta * M_PI,        theta = u          (scene.cuh:475-476)
//   3D line direction   (double)phi * M_PI,          phi = 2.f * u      (scene.cuh:784-786)
//   Box-Muller angle    2.0 * M_PI * (double)u.
// Prints the number of float values that differ per argument family.
#include <cmath>
#include <cstdint>
#include <cstdio>

#include "../../wobtoolbox_amd/csrc/wob_trig.hpp"

int main() {
    long bad[3] = {0, 0, 0};
    const int N = 1 << 24;
#pragma omp parallel for schedule(static) reduction(+ : bad[:3])
    for (int k = 0; k < N; ++k) {
        const float u = (float)((uint32_t)k + 1u) * (1.0f / 16777216.0f);
        const double xs[3] = {(double)u * M_PI, (double)(2.f * u) * M_PI, 2.0 * M_PI * (double)u};
        for (int f = 0; f < 3; ++f) {
            double s, c;
            wobtrig::sincos_small(xs[f], s, c);
            if ((float)s != (float)std::sin(xs[f]) || (float)c != (float)std::cos(xs[f])) ++bad[f];
        }
    }
    std::printf("%ld %ld %ld\n", bad[0], bad[1], bad[2]);
    return 0;
}
