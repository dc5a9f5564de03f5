// The FP64 walk's reduced square root and reciprocal (csrc/fp64_arith.hpp)
// against the device's IEEE sqrt() and 1.0 / x, bit for bit, over
// 2^26 inputs per family:
//   - uniform random exponents in [-600, 600] and random mantissas;
//   - near-ties of the square root: the neighbours (+-3 ulp) of g * next(g)
//     for random g, within an ulp of the square of the midpoint between g
//     and its successor;
//   - near-ties of the reciprocal: the neighbours (+-3 ulp) of
//     1 / midpoint-products, and square roots of random operands (the walk
//     takes 1 / sqrt(r2)).
// Prints the counts and the first mismatch; exit status 1 on any mismatch.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../csrc/fp64_arith.hpp"

namespace {

__device__ __forceinline__ uint64_t mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// A double with a uniform exponent in [lo, hi] and a random mantissa.
__device__ __forceinline__ double random_double(uint64_t h, int lo, int hi)
{
    const int e = lo + static_cast<int>((h >> 52) % static_cast<uint64_t>(hi - lo + 1));
    const uint64_t mant = mix(h) & 0x000FFFFFFFFFFFFFull;
    const uint64_t bits = (static_cast<uint64_t>(e + 1023) << 52) | mant;
    return __longlong_as_double(static_cast<long long>(bits));
}

struct Report {
    unsigned long long checked;
    unsigned long long sqrt_bad;
    unsigned long long rcp_bad;
    unsigned long long skipped;
    double first_bad;
    int first_kind;
};

struct Counts {
    unsigned long long checked = 0, sqrt_bad = 0, rcp_bad = 0, skipped = 0;
};

__device__ void first_mismatch(Report* rep, double v, int kind)
{
    if (atomicCAS(&rep->first_kind, 0, kind) == 0)
        rep->first_bad = v;
}

__device__ void check_sqrt(double x, Counts& c, Report* rep)
{
    if (!nbh::in_fast_range(x)) {
        ++c.skipped;
        return;
    }
    const double a = nbh::sqrt_rn_fast(x);
    const double b = sqrt(x);
    if (__double_as_longlong(a) != __double_as_longlong(b)) {
        ++c.sqrt_bad;
        first_mismatch(rep, x, 1);
    }
    ++c.checked;
}

__device__ void check_rcp(double s, Counts& c, Report* rep)
{
    if (!(s >= 0x1p-300 && s <= 0x1p300)) {
        ++c.skipped;
        return;
    }
    const double a = nbh::rcp_rn_fast(s);
    const double b = 1.0 / s;
    if (__double_as_longlong(a) != __double_as_longlong(b)) {
        ++c.rcp_bad;
        first_mismatch(rep, s, 2);
    }
    ++c.checked;
}

__global__ void arith_kernel(uint64_t seed, unsigned per_thread, Report* rep)
{
    const uint64_t tid = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    Counts c;
    for (unsigned k = 0; k < per_thread; ++k) {
        const uint64_t h = mix(seed ^ mix(tid * 1024 + k));
        // uniform operands
        const double x = random_double(h, -600, 600);
        check_sqrt(x, c, rep);
        check_rcp(random_double(mix(h ^ 1), -300, 300), c, rep);
        check_rcp(sqrt(x), c, rep);
        // square-root near-ties: g * next(g) is within an ulp of the square
        // of the midpoint of g and next(g)
        const double g = random_double(mix(h ^ 2), -299, 299);
        double t = g * nextafter(g, 2.0 * g);
        for (int s = 0; s < 3; ++s)
            t = nextafter(t, 0.0);
        for (int s = 0; s < 7; ++s, t = nextafter(t, 1e308))
            check_sqrt(t, c, rep);
        // reciprocal near-ties: 1 / (q * next(q)) ~ the reciprocal of a
        // midpoint-square, and 1 / q's neighbours
        const double q = random_double(mix(h ^ 3), -149, 149);
        double u = 1.0 / (q * nextafter(q, 2.0 * q));
        double w = 1.0 / q;
        for (int s = 0; s < 3; ++s) {
            u = nextafter(u, 0.0);
            w = nextafter(w, 0.0);
        }
        for (int s = 0; s < 7; ++s, u = nextafter(u, 1e308), w = nextafter(w, 1e308)) {
            check_rcp(u, c, rep);
            check_rcp(w, c, rep);
        }
    }
    atomicAdd(&rep->checked, c.checked);
    atomicAdd(&rep->sqrt_bad, c.sqrt_bad);
    atomicAdd(&rep->rcp_bad, c.rcp_bad);
    atomicAdd(&rep->skipped, c.skipped);
}

} // namespace

int main()
{
    Report* d = nullptr;
    if (hipMalloc(&d, sizeof(Report)) != hipSuccess) {
        std::printf("hipMalloc failed\n");
        return 2;
    }
    if (hipMemset(d, 0, sizeof(Report)) != hipSuccess)
        return 2;
    const unsigned blocks = 4096, threads = 256, per = 64; // 2^26 per family
    arith_kernel<<<blocks, threads>>>(20261204ull, per, d);
    Report h{};
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&h, d, sizeof(Report), hipMemcpyDeviceToHost) != hipSuccess) {
        std::printf("kernel failed\n");
        return 2;
    }
    (void)hipFree(d);
    std::printf("fp64 arith: %llu checked, %llu skipped, sqrt mismatches %llu, "
                "reciprocal mismatches %llu\n",
                h.checked, h.skipped, h.sqrt_bad, h.rcp_bad);
    if (h.sqrt_bad || h.rcp_bad) {
        std::printf("first mismatch (%s) at %.17g\n", h.first_kind == 1 ? "sqrt" : "rcp",
                    h.first_bad);
        return 1;
    }
    std::printf("PASS\n");
    return 0;
}
