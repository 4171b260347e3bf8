// Host build of alea_amd/csrc/trig.h for tests/test_trig.py: max ulp error of
// sincos_bounded, and max absolute error of sincos_bounded_f32, against libm
// over a sweep of arguments.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include "../alea_amd/csrc/trig.h"
static double ulp_err(double a, double b) {
    if (a == b) return 0.0;
    const double u = std::nextafter(std::fabs(b), INFINITY) - std::fabs(b);
    return std::fabs(a - b) / (std::fabs(b) > 1e-300 ? u : 1e-300);
}
int main(int argc, char** argv) {
    const double lim = argc > 1 ? std::atof(argv[1]) : 1e4;
    uint64_t st = 0x9e3779b97f4a7c15ULL;
    double worst_s = 0, worst_c = 0, abs_err = 0;
    for (int i = 0; i < 4000000; ++i) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        const double x = ((st >> 11) * 0x1.0p-53 * 2.0 - 1.0) * lim;
        double s, c;
        alea_amd::sincos_bounded(x, &s, &c);
        const double rs = std::sin(x), rc = std::cos(x);
        abs_err = std::fmax(abs_err, std::fmax(std::fabs(s - rs), std::fabs(c - rc)));
        if (std::fabs(rs) > 1e-3) worst_s = std::fmax(worst_s, ulp_err(s, rs));
        if (std::fabs(rc) > 1e-3) worst_c = std::fmax(worst_c, ulp_err(c, rc));
    }
    // single-precision values (the mixed form's orbitals): absolute error
    double abs_f32 = 0;
    for (int i = 0; i < 4000000; ++i) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        const double x = ((st >> 11) * 0x1.0p-53 * 2.0 - 1.0) * lim;
        float s, c;
        alea_amd::sincos_bounded_f32(x, &s, &c);
        abs_f32 = std::fmax(abs_f32, std::fmax(std::fabs(s - std::sin(x)), std::fabs(c - std::cos(x))));
    }
    std::printf("%.3f %.3f %.3e %.3e\n", worst_s, worst_c, abs_err, abs_f32);
    return 0;
}
