// mac_threshold.hpp -- the reference's opening test sqrt(d2) > crit
// (barnes_hut.cu:273-294) as one comparison d2 > t, exactly.
//
// sqrt rounded to nearest is non-decreasing, so {d2 : sqrt(d2) > crit} is
// {d2 : d2 > t} with t the largest double whose rounded square root does not
// exceed crit.  t is found from crit * crit by stepping one ulp at a time
// (a correctly rounded square is within an ulp or two of t).  Then, for
// every d2, (sqrt(d2) > crit) == (d2 > t): crit = +inf gives t = +inf
// (nothing is accepted, as theta = 0 requires) and crit = NaN gives t = NaN.
// Pure C++ so that the host test compiles the same function with g++.
#pragma once

#include <cfloat>
#include <cmath>

#ifndef NBH_HOST_DEVICE
#define NBH_HOST_DEVICE __host__ __device__
#endif

namespace nbh {

NBH_HOST_DEVICE inline double mac_threshold(double crit)
{
    if (crit != crit)
        return crit;
    if (crit < 0.0)
        return -INFINITY; // every d2 >= 0 has sqrt(d2) >= 0 > crit
    if (crit == INFINITY)
        return INFINITY;
    double t = crit * crit;
    if (t == INFINITY)
        t = DBL_MAX;
    // each loop moves t by one ulp; a handful of steps always suffice, the
    // bound only guarantees termination
    for (int k = 0; k < 64 && sqrt(t) > crit; ++k)
        t = nextafter(t, 0.0);
    for (int k = 0; k < 64 && t < DBL_MAX && !(sqrt(nextafter(t, INFINITY)) > crit); ++k)
        t = nextafter(t, INFINITY);
    return t;
}

} // namespace nbh
