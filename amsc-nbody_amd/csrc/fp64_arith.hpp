// fp64_arith.hpp -- correctly rounded FP64 square root and reciprocal for
// operands well inside the normal range, for the FP64 traversal.
//
// The FP64 walk must reproduce sqrt(r2) and 1.0 / sqrt(r2) as IEEE-754
// round-to-nearest operations (the oracle's arithmetic).  The general
// lowering of both on gfx950 is a hardware estimate (v_rsq_f64 / v_rcp_f64)
// refined by Newton-Raphson steps with fused multiply-adds, wrapped in
// operand scaling for exponents near the ends of the range and in fix-ups
// for zeros, infinities and NaNs.  For operands whose exponents are far from
// both ends no scaling or fix-up step changes anything, so the refinement
// alone yields the same bits; these functions are that refinement, used only
// when in_fast_range(r2) holds (2^-600 <= r2 <= 2^600, so that sqrt(r2)
// lies in [2^-300, 2^300]).  The GPU test cpp/tests/fp64_arith_test.hip
// checks both against sqrt() and 1.0 / x bit for bit.
//
// Refinements (a: the operand, e: a hardware estimate):
//   sqrt:  y = rsq(a), g = a y, h = y / 2, r = 1/2 - h g, g += g r,
//          h += h r, then twice g += (a - g^2) h        (Goldschmidt steps)
//   1 / s: y = rcp(s), twice y += y (1 - s y), then y + y (1 - s y)
#pragma once

#include <hip/hip_runtime.h>

namespace nbh {

__device__ __forceinline__ bool in_fast_range(double r2)
{
    return r2 >= 0x1p-600 && r2 <= 0x1p600;
}

__device__ __forceinline__ double sqrt_rn_fast(double a)
{
    const double y = __builtin_amdgcn_rsq(a);
    double g = a * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, a);
    g = fma(d, h, g);
    d = fma(-g, g, a);
    g = fma(d, h, g);
    return g;
}

__device__ __forceinline__ double rcp_rn_fast(double s)
{
    double y = __builtin_amdgcn_rcp(s);
    double e = fma(-s, y, 1.0);
    y = fma(y, e, y);
    e = fma(-s, y, 1.0);
    y = fma(y, e, y);
    const double r = fma(-s, y, 1.0);
    return fma(r, y, y);
}

} // namespace nbh
