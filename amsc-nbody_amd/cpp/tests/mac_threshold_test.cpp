// The FP64 walk's opening threshold (csrc/mac_threshold.hpp) on the host:
// for every tested crit, t = mac_threshold(crit) satisfies
// sqrt(t) <= crit < sqrt(next(t)), and (sqrt(d2) > crit) == (d2 > t) for d2
// in the ulp neighbourhoods of t and of crit^2 and for random d2.
// Prints the number of checks; exit status 1 on the first mismatch.
#define NBH_HOST_DEVICE
#include "../../csrc/mac_threshold.hpp"

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

int main()
{
    std::mt19937_64 rng(20261204);
    std::vector<double> crits = {0.0, 1.0, 2.0, 0.5, 3.0, 1e-300, 1e-160, 1e-155, 4.9e-324,
                                 1e150, 1e154, 1.3407807929942596e154, 1e300, DBL_MAX,
                                 INFINITY, -1.0};
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (int k = 0; k < 200000; ++k) {
        const double e = std::ldexp(1.0, static_cast<int>(u(rng) * 2000.0) - 1000);
        crits.push_back(e * (1.0 + u(rng)));
    }
    long checks = 0;
    for (const double c : crits) {
        const double t = nbh::mac_threshold(c);
        const auto ok = [&](double d2) {
            ++checks;
            return (std::sqrt(d2) > c) == (d2 > t);
        };
        if (c < 0.0) {
            if (!(t == -INFINITY) || !ok(0.0) || !ok(1.0)) {
                std::printf("FAIL negative crit\n");
                return 1;
            }
            continue;
        }
        if (c == INFINITY) {
            if (t != INFINITY || !ok(DBL_MAX) || !ok(INFINITY)) {
                std::printf("FAIL infinite crit\n");
                return 1;
            }
            continue;
        }
        if (std::sqrt(t) > c || (t < DBL_MAX && !(std::sqrt(std::nextafter(t, INFINITY)) > c))) {
            std::printf("FAIL threshold not maximal: crit %.17g t %.17g\n", c, t);
            return 1;
        }
        double probes[4] = {t, c * c, 0.0, 0.0};
        probes[2] = u(rng) * 4.0 * (c * c < DBL_MAX / 4 ? c * c : DBL_MAX / 4);
        probes[3] = std::nextafter(c * c, 0.0);
        for (double p : probes) {
            double d = p;
            for (int s = 0; s < 3; ++s)
                d = std::nextafter(d, 0.0);
            for (int s = 0; s < 7; ++s, d = std::nextafter(d, INFINITY)) {
                if (!ok(d)) {
                    std::printf("FAIL crit %.17g t %.17g d2 %.17g\n", c, t, d);
                    return 1;
                }
            }
        }
    }
    std::printf("mac_threshold: %ld checks passed\n", checks);
    return 0;
}
