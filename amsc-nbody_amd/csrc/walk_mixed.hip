// walk_mixed.hip -- mixed-precision packet traversal.
//
// The opening test and every interaction run in FP32 on exact integer
// differences of 31-bit fixed-point coordinates over the bounding box, i.e.
// on coordinates relative to the target: the only rounding of a separation
// is the conversion of the exact integer difference, relative to the
// separation itself.  Node records hold G m / unit^2, so the sums are
// accelerations in physical units.
//
// Hot path.  Internal nodes, single-particle leaves and zero-mass nodes
// share one straight-line path: "accept if d^2 > crit^2", with crit^2 = -1
// for leaves and zero-mass nodes (always accepted; a zero mass adds exactly
// zero, and with positive softening so does the target itself).  A node's
// record is one 32-byte {node, link} pair fetched with a single scalar load
// at the top of the iteration.  Only multi-particle leaves, zero softening
// and error records take the slow path.  Single-child chains are collapsed
// into their lowest member when the layout is built, which removes visits
// without changing any lane's interaction list.
//
// Accumulation.  Each lane adds its interactions into FP32 partial sums and
// folds them into a Kahan-compensated sum every kFold visits, so the
// compensated total carries the rounding of at most kFold terms per fold.
#include "walk_common.hpp"

namespace nbh {

namespace {

constexpr int kFold = 32;

// One position of the mixed walk: the node (fixed-point COM, G m) and its
// link (crit^2, skip, kind | count << 8, first particle).
struct alignas(32) WalkRec {
    int4 node;
    int4 link;
};

// Kahan-compensated FP32 accumulator of one component.
struct Kahan {
    float s = 0.0f;
    float c = 0.0f;
    __device__ __forceinline__ void add(float v)
    {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const
    {
        return static_cast<double>(s) - static_cast<double>(c);
    }
};

struct Acc {
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    Kahan kx, ky, kz;
    __device__ __forceinline__ void fold()
    {
        kx.add(px);
        ky.add(py);
        kz.add(pz);
        px = py = pz = 0.0f;
    }
};

__device__ __forceinline__ void separation(const int4& s, const int4& t,
                                           float& dx, float& dy, float& dz,
                                           float& d2)
{
    dx = static_cast<float>(s.x - t.x);
    dy = static_cast<float>(s.y - t.y);
    dz = static_cast<float>(s.z - t.z);
    d2 = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
}

// Softened FP32 monopole; `f` is forced to zero where `use` is false.
__device__ __forceinline__ void interact(float dx, float dy, float dz, float d2,
                                         float gm, float eps2, bool use,
                                         Acc& a)
{
    const float r2 = d2 + eps2;
    const float inv = __builtin_amdgcn_rsqf(r2);
    float f = gm * (inv * inv * inv);
    f = use ? f : 0.0f;
    a.px = fmaf(f, dx, a.px);
    a.py = fmaf(f, dy, a.py);
    a.pz = fmaf(f, dz, a.pz);
}

template <bool CONTAINS, bool SOFT, bool COUNT>
__global__ __launch_bounds__(kWalkThreads) void walk_mixed_kernel(
    const WalkRec* __restrict__ wrec,
    const int4* __restrict__ wcenter, const int* __restrict__ d_wcount,
    const int4* __restrict__ rec, const uint32_t* __restrict__ idx, int n,
    int nblocks_padded, float eps2, double* __restrict__ ax_out,
    double* __restrict__ ay_out, double* __restrict__ az_out, int* err,
    unsigned long long* __restrict__ counters)
{
    constexpr int kIdle = 0x7FFFFFFF;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int packet = packet_of_block(nblocks_padded) * 4 + wave;
    const int t = packet * 64 + lane;
    if (packet * 64 >= n)
        return;
    const bool valid = t < n;
    const int wcount = *d_wcount;
    int4 tq = make_int4(0, 0, 0, 0);
    if (valid)
        tq = rec[t];
    Acc a;
    unsigned long long npp = 0, npn = 0, visits = 0;
    int next = valid ? 0 : kIdle;
    int code = 0;
    int fold = 0;
    int i = 0;
    // A lane's `next` is the first position it still has to see.  At
    // position i a lane is active when next <= i.  Accepting (or finishing a
    // leaf) sets next = skip; opening leaves next <= i, so the lane is active
    // again at i + 1.  Lanes waiting for a later position always wait for
    // skip[i] or beyond (subtrees nest), hence the wave moves to i + 1 when
    // some lane opened and to skip[i] otherwise.
    while (i < wcount) {
        const WalkRec r = wrec[i];
        const int4 lk = r.link;
        const int4 nd = r.node;
        const int skip = lk.y;
        const int kind = lk.z & 0xFF;
        const bool active = next <= i;
        if (COUNT)
            ++visits;
        // Straight-line MAC for every record: records of other kinds carry
        // crit^2 = +inf and are never accepted here.
        float dx, dy, dz, d2;
        separation(nd, tq, dx, dy, dz, d2);
        bool acc = active && d2 > __int_as_float(lk.x);
        if (CONTAINS && acc) {
            const int4 c = wcenter[i];
            const bool inside = abs(tq.x - c.x) <= c.w &&
                                abs(tq.y - c.y) <= c.w &&
                                abs(tq.z - c.z) <= c.w;
            acc = !inside;
        }
        if (__ballot(acc) != 0ull) {
            interact(dx, dy, dz, d2, __int_as_float(nd.w), eps2, acc, a);
            if (COUNT && acc && __int_as_float(nd.w) != 0.0f) {
                if ((lk.z >> 8) == 1) {
                    if (lk.w != t)
                        ++npp;
                } else {
                    ++npn;
                }
            }
        }
        next = acc ? skip : next;
        if (kind != KIND_INTERNAL) {
            if (kind == KIND_LEAF) {
                // multi-particle leaf, or a single particle without softening
                const int count = lk.z >> 8;
                const int first = lk.w;
                for (int k = first; k < first + count; ++k) {
                    const int4 s = rec[k];
                    float sx, sy, sz, s2;
                    separation(s, tq, sx, sy, sz, s2);
                    const float gm = __int_as_float(s.w);
                    bool use = active && k != t && gm != 0.0f;
                    if (!SOFT && use && s2 == 0.0f) {
                        code = 6;
                        use = false;
                    }
                    interact(sx, sy, sz, s2, gm, eps2, use, a);
                    if (COUNT && use)
                        ++npp;
                }
                next = active ? skip : next;
            } else if (kind == KIND_SKIP) {
                next = active ? skip : next;
            } else if (active) {
                code = (kind == KIND_CHILDLESS) ? 5 : kind - KIND_ERROR;
            }
            if (code != 0)
                next = kIdle;
        }
        if (++fold == kFold) {
            a.fold();
            fold = 0;
        }
        i = (__ballot(next <= i) != 0ull) ? i + 1 : skip;
    }
    a.fold();
    if (code != 0) {
        raise_error(err, code);
        return;
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) {
            npp += __shfl_xor(npp, o);
            npn += __shfl_xor(npn, o);
        }
        if (lane == 0) {
            atomicAdd(&counters[0], npp);
            atomicAdd(&counters[1], npn);
            atomicAdd(&counters[2], visits);
        }
        return;
    }
    if (valid) {
        const uint32_t o = idx[t];
        if (o >= static_cast<uint32_t>(n)) {
            raise_error(err, 3);
            return;
        }
        ax_out[o] = a.kx.value();
        ay_out[o] = a.ky.value();
        az_out[o] = a.kz.value();
    }
}

} // namespace

int walk_blocks_padded(int n);

void launch_walk_mixed(hipStream_t s, const int4* wpairs,
                       const int4* wcenter, const int* d_wcount,
                       const int4* rec, const uint32_t* idx, int n,
                       float eps2, bool contains, bool soft, double* ax,
                       double* ay, double* az, int* err,
                       unsigned long long* counters)
{
    const int nb = walk_blocks_padded(n);
    const auto* wrec = reinterpret_cast<const WalkRec*>(wpairs);
#define NBH_ARGS                                                                 \
    wrec, wcenter, d_wcount, rec, idx, n, nb, eps2, ax, ay, az, err,       \
        counters
#define NBH_LAUNCH(C, S, K)                                                      \
    walk_mixed_kernel<C, S, K><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS)
    const bool cnt = counters != nullptr;
    if (contains) {
        if (soft) {
            if (cnt) NBH_LAUNCH(true, true, true); else NBH_LAUNCH(true, true, false);
        } else {
            if (cnt) NBH_LAUNCH(true, false, true); else NBH_LAUNCH(true, false, false);
        }
    } else {
        if (soft) {
            if (cnt) NBH_LAUNCH(false, true, true); else NBH_LAUNCH(false, true, false);
        } else {
            if (cnt) NBH_LAUNCH(false, false, true); else NBH_LAUNCH(false, false, false);
        }
    }
#undef NBH_LAUNCH
#undef NBH_ARGS
    NBH_HIP(hipGetLastError());
}

} // namespace nbh
