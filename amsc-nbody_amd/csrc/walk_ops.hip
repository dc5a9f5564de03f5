// walk_ops.hip -- FP64 walks that apply another operation than the force to
// the interaction list of each target: the softened potential (energy
// reporting) and the force vector-Jacobian product (the adjoint).
//
// The walk is the one of the FP64 force traversal (same records, same
// opening test, same list per target), so the potential and the gradient
// belong to exactly the forces the step integrates.  Per interaction with a
// source of mass m at separation d = x_s - x_t, r2 = |d|^2 + eps^2:
//   a_t   += G m d r2^{-3/2}
//   phi_t -= G m r2^{-1/2}
//   g      = (da_t/dd)^T abar_t = G m (r2^{-3/2} abar - 3 r2^{-5/2} d (d . abar))
//   xbar_t -= g,  xbar_s += g
//   mbar_s += (da_t/dm)^T abar_t = G r2^{-3/2} (d . abar)
// A monopole source is a centre of mass, COM_A = sum_k m_k x_k / M_A with
// M_A = sum_k m_k, so its position cotangent reaches particle k as
// (m_k / M_A) xbar_A, and its mass cotangent as
// mbar_A + (x_k - COM_A) . xbar_A / M_A.  Node cotangents are pushed down
// the octree, accumulating xbar_A / M_A and mbar_A - COM_A . xbar_A / M_A
// from the root; each leaf particle adds m_k times the first sum to its
// position cotangent, and the second sum plus x_k . (first sum) to its mass
// cotangent.  Source cotangents are summed over the wavefront (every lane
// sees the same source) in batches: the lanes park their terms of a few
// source visits in LDS, the batch is summed transposed (each lane one
// component of one visit over a group of lanes, then across the group), and
// one atomic instruction adds the totals to the FP64 accumulators.  The
// record of each walk position is read into scalar registers, so that the
// walk's branches stay uniform.
#include "fp64_arith.hpp"
#include "walk_common.hpp"

namespace nbh {

namespace {

enum : int { OP_POTENTIAL = 0, OP_VJP = 1 };

struct OpArgs {
    double* phi;                   // potential, original order
    const double* abx;             // acceleration cotangent, original order
    const double* aby;
    const double* abz;
    double* tbx;                   // target terms, Morton order (plain stores)
    double* tby;
    double* tbz;
    double* pbx;                   // particle-source terms, Morton order
    double* pby;
    double* pbz;
    double* nbx;                   // node-source terms, by walk position
    double* nby;
    double* nbz;
    double* pm;                    // particle-source mass terms, Morton order
    double* nm;                    // node-source mass terms, by walk position
    // FP64 VJP split into subtree units (few packets): the split table
    // {s, then per subtree {first, end, -, -}}, the queue order (wavefront q
    // walks unit perm[q]; units numbered packet * s + subtree) and each
    // unit's visit count; null for whole-packet walks
    const int* split;
    const uint32_t* perm;
    uint32_t* pcost;
};


// The walk position is wave-uniform; records are read into scalar registers
// (one scalar load and its wait in one statement, so that no output register
// is written after the statement ends), which keeps every record-driven
// branch of the walk uniform.
typedef int vjp_int8v __attribute__((ext_vector_type(8)));
typedef int vjp_int16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ Walk64Rec load_record64(const Walk64Rec* w64, int i)
{
    const Walk64Rec* p = w64 + static_cast<size_t>(__builtin_amdgcn_readfirstlane(i));
    vjp_int16v v;
    asm volatile("s_load_dwordx16 %0, %1, 0x0\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=s"(v)
                 : "s"(p)
                 : "memory");
    return __builtin_bit_cast(Walk64Rec, v);
}

// The {node, link} pair of the mixed records at position i.
__device__ __forceinline__ void load_record32(const int4* wrec, int i, int4& nd, int4& lk)
{
    const int4* p = wrec + 2 * static_cast<size_t>(__builtin_amdgcn_readfirstlane(i));
    vjp_int8v v;
    asm volatile("s_load_dwordx8 %0, %1, 0x0\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=s"(v)
                 : "s"(p)
                 : "memory");
    nd = make_int4(v[0], v[1], v[2], v[3]);
    lk = make_int4(v[4], v[5], v[6], v[7]);
}

constexpr int kSrcBatch64 = 2;

// 1 / sqrt(r2) for the potential and the force VJP, which are checked
// against tolerances rather than bit for bit: the hardware estimate and two
// Newton steps y (3 - r2 y^2) / 2 (relative error near 1e-16) inside
// 2^-600 <= r2 <= 2^600, the full-precision library division outside.
__device__ __forceinline__ double rsqrt_fp64(double r2)
{
    if (!in_fast_range(r2))
        return 1.0 / sqrt(r2);
    double y = __builtin_amdgcn_rsq(r2);
    double h = 0.5 * r2;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

template <int OP>
struct Lane {
    double tx = 0.0, ty = 0.0, tz = 0.0;
    double bx = 0.0, by = 0.0, bz = 0.0; // abar of the target
    double phi = 0.0;
    double gx = 0.0, gy = 0.0, gz = 0.0; // xbar of the target
    int code = 0;

    // One source for every lane of the wave (`use` false on lanes that do
    // not interact).  Called from wave-uniform control flow.  The potential
    // skips zero masses; the product keeps them for the mass cotangent (a
    // is linear in every source mass), except at a singular separation,
    // where a zero mass adds nothing either way.
    __device__ __forceinline__ void source(bool use, double sx, double sy,
                                           double sz, double sm, double G,
                                           double eps2, double* dx_out,
                                           double* dy_out, double* dz_out,
                                           double* m_out, int index)
    {
        double cx = 0.0, cy = 0.0, cz = 0.0, cm = 0.0;
        if (use && (sm != 0.0 || OP == OP_VJP)) {
            const double dx = sx - tx;
            const double dy = sy - ty;
            const double dz = sz - tz;
            const double r2 = fma(dz, dz, fma(dx, dx, dy * dy)) + eps2;
            if (r2 <= 0.0) {
                if (sm != 0.0)
                    code = 6;
            } else {
                const double inv = rsqrt_fp64(r2);
                const double gm = G * sm;
                if (OP == OP_POTENTIAL) {
                    phi = fma(-gm, inv, phi);
                } else {
                    const double inv2 = inv * inv;
                    const double inv3 = inv2 * inv;
                    const double dot = fma(dz, bz, fma(dx, bx, dy * by));
                    const double c5 = 3.0 * dot * inv3 * inv2;
                    cx = gm * fma(inv3, bx, -c5 * dx);
                    cy = gm * fma(inv3, by, -c5 * dy);
                    cz = gm * fma(inv3, bz, -c5 * dz);
                    cm = G * dot * inv3;
                    gx -= cx;
                    gy -= cy;
                    gz -= cz;
                }
            }
        }
        if (OP == OP_VJP) {
            if (__ballot(cx != 0.0 || cy != 0.0 || cz != 0.0 || cm != 0.0) != 0ull)
                park(cx, cy, cz, cm, dx_out, index);
        }
    }

    // ---- the wave's batch of source terms (OP_VJP; wave-uniform calls):
    // the four terms of each lane for kSrcBatch64 source visits parked in
    // LDS, then summed transposed, as in the mixed product
    double4* slots = nullptr; // kSrcBatch64 x 64 lanes
    int2* meta = nullptr;     // per parked visit: source index, node flag
    int parked = 0;
    double* pbase = nullptr;
    double* nbase = nullptr;
    size_t pstride = 0, nstride = 0;

    __device__ __forceinline__ void park(double cx, double cy, double cz, double cm,
                                         const double* dx_out, int index)
    {
        const int l = threadIdx.x & 63;
        slots[parked * 64 + l] = make_double4(cx, cy, cz, cm);
        if (l == 0)
            meta[parked] = make_int2(index, dx_out == nbase ? 1 : 0);
        if (++parked == kSrcBatch64)
            flush_sources();
    }

    // Lane l: component (l >> 3) & 3 of parked visit l >> 5, summed over
    // lanes (l & 7) * 8 .. + 7 of the visit, then over the group of eight.
    __device__ __forceinline__ void flush_sources()
    {
        if (parked == 0)
            return;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int l = threadIdx.x & 63;
        const int v = l >> 5;
        const int c = (l >> 3) & 3;
        const int j = l & 7;
        double sum = 0.0;
        if (v < parked) {
            const double* f = reinterpret_cast<const double*>(slots + v * 64 + j * 8) + c;
#pragma unroll
            for (int k = 0; k < 8; ++k)
                sum += f[4 * k];
        }
        for (int o = 4; o > 0; o >>= 1)
            sum += __shfl_xor(sum, o, 8);
        if (j == 0 && v < parked) {
            const int2 md = meta[v];
            double* out = md.y ? nbase + static_cast<size_t>(c) * nstride
                               : pbase + static_cast<size_t>(c) * pstride;
            atomicAdd(out + md.x, sum);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        parked = 0;
    }
};

template <bool CONTAINS, int OP, bool SPLIT = false>
__global__ __launch_bounds__(kWalkThreads) void walk_op_kernel(
    const Walk64Rec* __restrict__ w64, const double4* __restrict__ wcenter,
    const int* __restrict__ d_wcount, const double4* __restrict__ rec,
    const uint32_t* __restrict__ idx, int n, int nblocks_padded, double eps2,
    double G, OpArgs a, int* err)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    // SPLIT (the VJP with a split table): a unit per wavefront
    constexpr bool split = OP == OP_VJP && SPLIT;
    int packet = packet_of_block(nblocks_padded) * 4 + wave;
    int unit = 0;
    int begin = 0;
    int end = *d_wcount;
    if (split) {
        // units by decreasing cost, dealt to the wavefronts in dispatch
        // order (a permutation by construction; the bound keeps every
        // access inside the units whatever it holds)
        const int q = static_cast<int>(blockIdx.x) * 4 + wave;
        const int subtrees = __builtin_amdgcn_readfirstlane(a.split[0]);
        const int total = (n + 63) / 64 * subtrees;
        if (q >= total)
            return;
        unit = a.perm ? __builtin_amdgcn_readfirstlane(static_cast<int>(
                            min(a.perm[q], static_cast<uint32_t>(total - 1))))
                      : q;
        packet = unit / subtrees;
        const int sub = unit - packet * subtrees;
        begin = __builtin_amdgcn_readfirstlane(a.split[1 + 4 * sub]);
        end = __builtin_amdgcn_readfirstlane(a.split[2 + 4 * sub]);
    }
    const int t = packet * 64 + lane;
    if (packet * 64 >= n)
        return;
    const bool valid = t < n;
    Lane<OP> L;
    __shared__ double4 s_slots64[OP == OP_VJP ? kWalkThreads / 64 : 1]
                                [OP == OP_VJP ? kSrcBatch64 * 64 : 1];
    __shared__ int2 s_meta64[OP == OP_VJP ? kWalkThreads / 64 : 1][kSrcBatch64];
    if (OP == OP_VJP) {
        L.slots = s_slots64[threadIdx.x >> 6];
        L.meta = s_meta64[threadIdx.x >> 6];
        L.pbase = a.pbx;
        L.nbase = a.nbx;
        L.pstride = static_cast<size_t>(a.pby - a.pbx);
        L.nstride = static_cast<size_t>(a.nby - a.nbx);
    }
    uint32_t o = 0;
    if (valid) {
        const double4 p = rec[t];
        L.tx = p.x;
        L.ty = p.y;
        L.tz = p.z;
        o = idx[t];
        if (o >= static_cast<uint32_t>(n)) {
            L.code = 3;
            o = 0;
        }
        if (OP == OP_VJP) {
            L.bx = a.abx[o];
            L.by = a.aby[o];
            L.bz = a.abz[o];
        }
    }
    int next = valid ? begin : 0x7FFFFFFF;
    int i = begin;
    unsigned visits = 0;
    while (i < end) {
        i = __builtin_amdgcn_readfirstlane(i); // wave-uniform
        if (split)
            ++visits;
        const Walk64Rec r = load_record64(w64, i);
        const int skip = r.skip;
        const int kind = r.kind;
        const bool active = next <= i && L.code == 0;
        bool open = false;
        // a record with thr = -inf is a one-particle leaf (count 1) or a
        // zero-mass node (count 0), taken without an opening test
        const bool fast = kind == KIND_INTERNAL && r.thr == -INFINITY;
        if (fast) {
            if (r.count == 1)
                L.source(active && r.first != t, r.x, r.y, r.z, r.m, G, eps2, a.pbx,
                         a.pby, a.pbz, a.pm, r.first);
            if (active)
                next = skip;
        } else if (kind == KIND_INTERNAL || kind == KIND_CHILDLESS) {
            const double dx = r.x - L.tx;
            const double dy = r.y - L.ty;
            const double dz = r.z - L.tz;
            const double d2 = fma(dz, dz, fma(dx, dx, dy * dy));
            // the opening test sqrt(d2) > crit as d2 > its exact threshold
            bool accept = active && d2 > r.thr;
            if (CONTAINS && accept) {
                const double4 c = wcenter[i];
                const bool inside = fabs(L.tx - c.x) <= c.w &&
                                    fabs(L.ty - c.y) <= c.w &&
                                    fabs(L.tz - c.z) <= c.w;
                accept = !inside;
            }
            if (accept) {
                next = skip;
            } else if (active) {
                if (kind == KIND_CHILDLESS) {
                    L.code = 5;
                } else {
                    open = true;
                    next = i + 1;
                }
            }
            L.source(accept, r.x, r.y, r.z, r.m, G, eps2, a.nbx, a.nby,
                     a.nbz, a.nm, i);
        } else if (kind == KIND_LEAF) {
            for (int k = r.first; k < r.first + r.count; ++k) {
                const double4 s = rec[k];
                L.source(active && k != t, s.x, s.y, s.z, s.w, G, eps2,
                         a.pbx, a.pby, a.pbz, a.pm, k);
            }
            if (active)
                next = skip;
        } else if (active) {
            L.code = kind - KIND_ERROR;
        }
        i = (__ballot(open) != 0ull) ? i + 1 : skip;
    }
    if (OP == OP_VJP)
        L.flush_sources();
    if (split && a.pcost && lane == 0)
        a.pcost[unit] = visits;
    if (L.code != 0) {
        raise_error(err, L.code);
        return;
    }
    if (!valid)
        return;
    if (OP == OP_POTENTIAL) {
        a.phi[o] = L.phi;
    } else if (split) {
        // the units of a packet add their target terms (zeroed first)
        atomicAdd(&a.tbx[t], L.gx);
        atomicAdd(&a.tby[t], L.gy);
        atomicAdd(&a.tbz[t], L.gz);
    } else {
        a.tbx[t] = L.gx;
        a.tby[t] = L.gy;
        a.tbz[t] = L.gz;
    }
}

// The split of the FP64 walk below its first record, which every target
// opens (as for the mixed walk's split): the subtrees that follow position
// 0, read off the skip links, as units {first, end}.  A first record that is
// not an internal node with an opening test, or links that do not tile the
// walk, give the single unit {0, wcount}.
__global__ void walk_split64_kernel(const Walk64Rec* __restrict__ w64,
                                    const int* __restrict__ d_wcount, int* __restrict__ split)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const int wc = *d_wcount;
    bool ok = wc > 1 && w64[0].kind == KIND_INTERNAL && w64[0].thr != -INFINITY;
    int k = 0;
    int p = 1;
    while (ok && p < wc) {
        const int e = w64[p].skip;
        if (k == kWalkSplitMax || e <= p || e > wc) {
            ok = false;
            break;
        }
        split[1 + 4 * k] = p;
        split[2 + 4 * k] = e;
        p = e;
        ++k;
    }
    if (!ok || k == 0) {
        split[0] = 1;
        split[1] = 0;
        split[2] = wc;
        return;
    }
    split[0] = k;
}

// Top-down over one level: u_v = u_parent + nbar(pre_v) / M_v and
// w_v = w_parent + mbar(pre_v) - COM_v . nbar(pre_v) / M_v; then every
// particle k of a leaf v adds m_k u_v to its position cotangent and
// w_v + x_k . u_v to its mass cotangent.
__global__ __launch_bounds__(256) void push_down_kernel(
    nbh_octree_view o, int L, const int2* __restrict__ pos,
    const int4* __restrict__ aux, int wcount,
    const double* __restrict__ nbx, const double* __restrict__ nby,
    const double* __restrict__ nbz, const double* __restrict__ nm,
    double* __restrict__ sx, double* __restrict__ sy, double* __restrict__ sz,
    double* __restrict__ sw, const double4* __restrict__ rec,
    double* __restrict__ pbx, double* __restrict__ pby, double* __restrict__ pbz,
    double* __restrict__ pm)
{
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int v = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
         v < o.n_nodes; v += stride) {
        if (o.level[v] != L)
            continue;
        double ux = 0.0, uy = 0.0, uz = 0.0, uw = 0.0;
        if (L > 0) {
            const int p = o.parent[v];
            ux = sx[p];
            uy = sy[p];
            uz = sz[p];
            uw = sw[p];
        }
        const int pre = pos[v].x;
        const double m = o.mass[v];
        // a node the mixed layout collapsed into its single child owns no
        // walk position (its child's record took it)
        const bool owns = !(aux && o.particle_count[v] <= 0 && aux[v].w != 0);
        if (owns && m > 0.0 && pre >= 0 && pre < wcount) {
            const double bx = nbx[pre] / m;
            const double by = nby[pre] / m;
            const double bz = nbz[pre] / m;
            ux += bx;
            uy += by;
            uz += bz;
            uw += nm[pre] - fma(o.com_z[v], bz, fma(o.com_x[v], bx, o.com_y[v] * by));
        }
        sx[v] = ux;
        sy[v] = uy;
        sz[v] = uz;
        sw[v] = uw;
        const int pc = o.particle_count[v];
        if (pc > 0) {
            const int f = o.first_particle[v];
            for (int k = f; k < f + pc; ++k) {
                const double4 r = rec[k];
                pbx[k] = fma(r.w, ux, pbx[k]);
                pby[k] = fma(r.w, uy, pby[k]);
                pbz[k] = fma(r.w, uz, pbz[k]);
                pm[k] += uw + fma(r.z, uz, fma(r.x, ux, r.y * uy));
            }
        }
    }
}

// xbar (original order) += target terms + source terms of Morton slot t;
// mbar (original order, optional) += mass terms.
__global__ __launch_bounds__(256) void combine_kernel(
    const uint32_t* __restrict__ idx, int n, const double* __restrict__ tbx,
    const double* __restrict__ tby, const double* __restrict__ tbz,
    const double* __restrict__ pbx, const double* __restrict__ pby,
    const double* __restrict__ pbz, const double* __restrict__ pm,
    double* __restrict__ gx, double* __restrict__ gy, double* __restrict__ gz,
    double* __restrict__ gm, int* err)
{
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= n)
        return;
    const uint32_t o = idx[t];
    if (o >= static_cast<uint32_t>(n)) {
        raise_error(err, 3);
        return;
    }
    gx[o] += tbx[t] + pbx[t];
    gy[o] += tby[t] + pby[t];
    gz[o] += tbz[t] + pbz[t];
    if (gm)
        gm[o] += pm[t];
}

// Per-block partial sums of 1/2 m v^2 and 1/2 m phi in a fixed order.
__global__ __launch_bounds__(256) void energy_partials_kernel(
    const double* __restrict__ m, const double* __restrict__ vx,
    const double* __restrict__ vy, const double* __restrict__ vz,
    const double* __restrict__ phi, size_t n, double* __restrict__ partials)
{
    __shared__ double sk[256];
    __shared__ double sp[256];
    double k = 0.0, p = 0.0;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < n; i += stride) {
        const double v2 = fma(vz[i], vz[i], fma(vx[i], vx[i], vy[i] * vy[i]));
        k = fma(0.5 * m[i], v2, k);
        p = fma(0.5 * m[i], phi[i], p);
    }
    sk[threadIdx.x] = k;
    sp[threadIdx.x] = p;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            sk[threadIdx.x] += sk[threadIdx.x + w];
            sp[threadIdx.x] += sp[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sk[0];
        partials[2 * blockIdx.x + 1] = sp[0];
    }
}

// ------------------------------------------------ mixed-precision VJP
// The force VJP along the mixed walk: the same records, opening test and
// interaction lists as the mixed force traversal (fixed-point separations,
// FP32 interactions), so the gradient belongs to the forces the mixed step
// integrates.  Per interaction, in the records' units (separations d in
// fixed-point units, source mass gm = m / m_max):
//   g  = gm (r^-3 ab - 3 r^-5 d (d . ab)),   cm = (d . ab) r^-3,
// with the target's abar scaled by a power of two common to the wave, so
// that the FP32 terms stay in range and the wave's source sums need no
// rescaling.  Physical units: x cotangents G m_max unit^3 g, mass
// cotangents G unit^2 cm (unit = fixed-point steps per length).  Target
// terms are summed per lane in FP32 partials folded into Kahan sums every
// 64 visits; each source's terms are summed over the wave in FP32 and added
// to FP64 accumulators (particles by Morton index, nodes by walk position).
// The source sums are batched: a wave parks the four terms of each lane for
// kSrcBatch source visits in LDS, then transposes them, so that one batch
// costs a few adds per lane and one atomic instruction, where a sum per
// visit took four six-step cross-lane reductions and four atomics.
struct MixedVjp {
    const int4* wrec;    // interleaved {node, link} pairs, 2 int4 per position
    const int4* wcenter; // cell centres (theta >= 1.15), else null
    const int* d_wcount;
    const int* split;    // split table of the walk's units, else null
    const int4* rec;     // particles, Morton order: fixed point, m / m_max
    const uint32_t* idx; // Morton -> original index
    const uint32_t* order; // target order, null for Morton order
    int n;
    int units;           // units per packet: kWalkSplitMax with a split table, else 1
    float eps2;          // softening^2 in fixed-point units
    double gunit;        // G m_max unit^3
    double munit;        // G unit^2
    size_t pstride;      // particle-source arrays: x, y, z and mass terms at
                         // pbx + {0, 1, 2, 3} pstride
    size_t nstride;      // node-source arrays likewise from nbx
    const uint32_t* perm; // split walks: wavefront q walks unit perm[q] (units
                          // numbered packet * subtrees + subtree), else null
};

constexpr int kSrcBatch = 4;

// A DPP move of v (lanes whose source lies outside the masks read zero).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f(float v)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWS, 0xF, false));
}

struct VjpKahan {
    float s = 0.0f, c = 0.0f;
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

struct VjpLane {
    int4 tq;
    int t;
    float bx, by, bz;       // abar / 2^e (e common to the wave)
    float px = 0.0f, py = 0.0f, pz = 0.0f; // target partials (-g)
    VjpKahan kx, ky, kz;
    int code = 0;

    // One source for the whole wave (use false on lanes that do not
    // interact); wave-uniform control flow.
    template <bool SOFT>
    __device__ __forceinline__ void source(bool use, const int4& sq, float eps2, bool node,
                                           int index)
    {
        const float dx = static_cast<float>(sq.x - tq.x);
        const float dy = static_cast<float>(sq.y - tq.y);
        const float dz = static_cast<float>(sq.z - tq.z);
        const float r2 = fmaf(dz, dz, fmaf(dx, dx, fmaf(dy, dy, eps2)));
        source_at<SOFT>(use, dx, dy, dz, r2, __int_as_float(sq.w), node, index);
    }

    // The same with the separation and softened r^2 already at hand (the
    // opening test of an internal node computed them).
    template <bool SOFT>
    __device__ __forceinline__ void source_at(bool use, float dx, float dy, float dz,
                                              float r2, float gm, bool node, int index)
    {
        float gx = 0.0f, gy = 0.0f, gz = 0.0f, cm = 0.0f;
        if (use) {
            if (!SOFT && r2 == 0.0f) {
                if (gm != 0.0f)
                    code = 6;
            } else {
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float inv2 = inv * inv;
                const float inv3 = inv2 * inv;
                const float dot = fmaf(dz, bz, fmaf(dx, bx, dy * by));
                const float t1 = gm * inv3;
                const float t2 = 3.0f * t1 * dot * inv2;
                gx = fmaf(t1, bx, -t2 * dx);
                gy = fmaf(t1, by, -t2 * dy);
                gz = fmaf(t1, bz, -t2 * dz);
                cm = dot * inv3;
                px -= gx;
                py -= gy;
                pz -= gz;
            }
        }
        if (__ballot(use) != 0ull)
            park(gx, gy, gz, cm, node, index);
    }

    // ---- the wave's batch of source terms (wave-uniform calls)
    float4* slots = nullptr; // kSrcBatch x 64 lanes
    int2* meta = nullptr;    // per parked visit: source index, node flag
    int parked = 0;
    double* pbase = nullptr;
    double* nbase = nullptr;
    size_t pstride = 0, nstride = 0;
    double unit_g = 0.0, unit_m = 0.0;

    __device__ __forceinline__ void park(float gx, float gy, float gz, float cm, bool node,
                                         int index)
    {
        const int l = threadIdx.x & 63;
        slots[parked * 64 + l] = make_float4(gx, gy, gz, cm);
        if (l == 0)
            meta[parked] = make_int2(index, node ? 1 : 0);
        if (++parked == kSrcBatch)
            flush_sources();
    }

    // Lane l sums component (l >> 2) & 3 of parked visit l >> 4 over the
    // 16 lanes of quarter l & 3, a quad sum adds the quarters, and the
    // first lane of each quad adds the total to its source's accumulator.
    __device__ __forceinline__ void flush_sources()
    {
        if (parked == 0)
            return;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int l = threadIdx.x & 63;
        const int v = l >> 4;
        const int c = (l >> 2) & 3;
        const int q = l & 3;
        float sum = 0.0f;
        if (v < parked) {
            const float* f = reinterpret_cast<const float*>(slots + v * 64 + q * 16) + c;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                sum += f[4 * j];
        }
        sum += dpp_f<0xB1, 0xF>(sum); // quad_perm [1, 0, 3, 2]
        sum += dpp_f<0x4E, 0xF>(sum); // quad_perm [2, 3, 0, 1]
        if (q == 0 && v < parked) {
            const int2 md = meta[v];
            double* out = md.y ? nbase + static_cast<size_t>(c) * nstride
                               : pbase + static_cast<size_t>(c) * pstride;
            atomicAdd(out + md.x, static_cast<double>(sum) * (c == 3 ? unit_m : unit_g));
        }
        // the slots are read before the next visits overwrite them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        parked = 0;
    }
};

// One unit per wave: a packet of 64 targets and, with a split table, one
// subtree of the walk below its first record (which every target opens), so
// that few packets still fill the device; target terms are added to FP64
// accumulators per unit.
template <bool CONTAINS, bool SOFT>
__global__ __launch_bounds__(kWalkThreads) void vjp_mixed_kernel(MixedVjp m, OpArgs a,
                                                                 int* err)
{
    const int lane = threadIdx.x & 63;
    const int q = static_cast<int>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int n = m.n;
    int packet = q / m.units;
    int sub = q - packet * m.units;
    if (m.perm && m.split) {
        // units in the order of their costs, dealt to the wavefronts in
        // dispatch order (a permutation by construction; the bound keeps
        // every access inside the units whatever it holds)
        const int subtrees = __builtin_amdgcn_readfirstlane(m.split[0]);
        const int total = (n + 63) / 64 * subtrees;
        if (q >= total)
            return;
        const int u = __builtin_amdgcn_readfirstlane(
            static_cast<int>(min(m.perm[q], static_cast<uint32_t>(total - 1))));
        packet = u / subtrees;
        sub = u - packet * subtrees;
    }
    if (packet * 64 >= n)
        return;
    int begin = 0;
    int end = __builtin_amdgcn_readfirstlane(*m.d_wcount);
    if (m.split) {
        if (sub >= __builtin_amdgcn_readfirstlane(m.split[0]))
            return;
        begin = __builtin_amdgcn_readfirstlane(m.split[1 + 4 * sub]);
        end = __builtin_amdgcn_readfirstlane(m.split[2 + 4 * sub]);
    }
    const int k = packet * 64 + lane;
    bool valid = k < n;
    VjpLane L;
    L.t = k;
    if (valid && m.order) {
        const uint32_t o = m.order[k];
        if (o < static_cast<uint32_t>(n)) {
            L.t = static_cast<int>(o);
        } else {
            valid = false;
            L.code = 3;
        }
    }
    uint32_t orig = 0;
    double bx = 0.0, by = 0.0, bz = 0.0;
    if (valid) {
        orig = m.idx[L.t];
        if (orig >= static_cast<uint32_t>(n)) {
            L.code = 3;
            valid = false;
        } else {
            bx = a.abx[orig];
            by = a.aby[orig];
            bz = a.abz[orig];
        }
    }
    L.tq = valid ? m.rec[L.t] : make_int4(0, 0, 0, 0);
    // a power of two at or above the wave's largest |abar| component
    double bmax = fmax(fabs(bx), fmax(fabs(by), fabs(bz)));
    for (int o = 32; o > 0; o >>= 1)
        bmax = fmax(bmax, __shfl_xor(bmax, o));
    int e = 0;
    if (bmax > 0.0 && bmax <= 1.7e308)
        (void)frexp(bmax, &e);
    L.bx = static_cast<float>(ldexp(bx, -e));
    L.by = static_cast<float>(ldexp(by, -e));
    L.bz = static_cast<float>(ldexp(bz, -e));
    const double ug = ldexp(m.gunit, e);
    const double um = ldexp(m.munit, e);
    __shared__ float4 s_slots[kWalkThreads / 64][kSrcBatch * 64];
    __shared__ int2 s_meta[kWalkThreads / 64][kSrcBatch];
    L.slots = s_slots[threadIdx.x >> 6];
    L.meta = s_meta[threadIdx.x >> 6];
    L.pbase = a.pbx;
    L.nbase = a.nbx;
    L.pstride = m.pstride;
    L.nstride = m.nstride;
    L.unit_g = ug;
    L.unit_m = um;
    int next = valid ? begin : 0x7FFFFFFF;
    int i = begin;
    unsigned cnt = 0;
    while (i < end) {
        i = __builtin_amdgcn_readfirstlane(i); // wave-uniform
        int4 nd, lk;
        load_record32(m.wrec, i, nd, lk);
        if (m.split && (lk.z & kMixedBoundary) != 0) {
            // a marked subtree start of a split layout: its own link words
            // (read as wave-uniform values, like the record)
            const int subtrees = __builtin_amdgcn_readfirstlane(m.split[0]);
            for (int q = 0; q < subtrees; ++q)
                if (__builtin_amdgcn_readfirstlane(m.split[1 + 4 * q]) == i) {
                    lk.x = __builtin_amdgcn_readfirstlane(m.split[3 + 4 * q]);
                    lk.z = __builtin_amdgcn_readfirstlane(m.split[4 + 4 * q]);
                }
        }
        const int skip = lk.y;
        const int kind = lk.z & 0x7F;
        const float crit2 = __int_as_float(lk.x);
        const bool active = next <= i && L.code == 0;
        bool open = false;
        if (kind == KIND_INTERNAL && crit2 == -1.0f) {
            // always taken: a one-particle leaf (count 1, its particle) or
            // a zero-mass node (count 0, nothing)
            if ((lk.z >> 8) == 1)
                L.source<SOFT>(active && lk.w != L.t, nd, m.eps2, false, lk.w);
            if (active)
                next = skip;
        } else if (kind == KIND_INTERNAL) {
            const float dx = static_cast<float>(nd.x - L.tq.x);
            const float dy = static_cast<float>(nd.y - L.tq.y);
            const float dz = static_cast<float>(nd.z - L.tq.z);
            const float r2 = fmaf(dz, dz, fmaf(dx, dx, fmaf(dy, dy, m.eps2)));
            bool accept = active && r2 > crit2;
            if (CONTAINS && accept) {
                const int4 c = m.wcenter[i];
                accept = !(abs(L.tq.x - c.x) <= c.w && abs(L.tq.y - c.y) <= c.w &&
                           abs(L.tq.z - c.z) <= c.w);
            }
            if (accept) {
                next = skip;
            } else if (active) {
                open = true;
                next = i + 1;
            }
            L.source_at<SOFT>(accept, dx, dy, dz, r2, __int_as_float(nd.w), true, i);
        } else if (kind == KIND_LEAF) {
            const int count = lk.z >> 8;
            const int first = lk.w;
            for (int q = first; q < first + count; ++q)
                L.source<SOFT>(active && q != L.t, m.rec[q], m.eps2, false, q);
            if (active)
                next = skip;
        } else if (kind == KIND_SKIP) {
            if (active)
                next = skip;
        } else if (active) {
            L.code = (kind == KIND_CHILDLESS) ? 5 : kind - KIND_ERROR;
        }
        if (L.code != 0)
            next = 0x7FFFFFFF;
        if (++cnt == 64) {
            L.kx.add(L.px);
            L.ky.add(L.py);
            L.kz.add(L.pz);
            L.px = L.py = L.pz = 0.0f;
            cnt = 0;
        }
        i = (__ballot(open) != 0ull) ? i + 1 : skip;
    }
    L.flush_sources();
    if (L.code != 0) {
        raise_error(err, L.code);
        return;
    }
    if (!valid)
        return;
    L.kx.add(L.px);
    L.ky.add(L.py);
    L.kz.add(L.pz);
    atomicAdd(a.tbx + L.t, L.kx.value() * ug);
    atomicAdd(a.tby + L.t, L.ky.value() * ug);
    atomicAdd(a.tbz + L.t, L.kz.value() * ug);
}

template <int OP>
void launch_op(hipStream_t s, const Walk64Rec* w64, const double4* wcenter,
               const int* d_wcount, const double4* rec, const uint32_t* idx, int n,
               double eps2, double G, bool contains, const OpArgs& a, int* err)
{
    const int nb = walk_blocks_padded(n);
    if (contains)
        walk_op_kernel<true, OP><<<nb, kWalkThreads, 0, s>>>(
            w64, wcenter, d_wcount, rec, idx, n, nb, eps2, G, a, err);
    else
        walk_op_kernel<false, OP><<<nb, kWalkThreads, 0, s>>>(
            w64, wcenter, d_wcount, rec, idx, n, nb, eps2, G, a, err);
    NBH_HIP(hipGetLastError());
}

} // namespace

int walk_blocks_padded(int n);

void launch_potential_fp64(hipStream_t s, const Walk64Rec* w64,
                           const double4* wcenter, const int* d_wcount,
                           const double4* rec, const uint32_t* idx, int n,
                           double eps2, double G, bool contains, double* phi,
                           int* err)
{
    OpArgs a{};
    a.phi = phi;
    launch_op<OP_POTENTIAL>(s, w64, wcenter, d_wcount, rec, idx, n, eps2, G, contains,
                            a, err);
}

void launch_energy_partials(hipStream_t s, const double* m, const double* vx,
                            const double* vy, const double* vz,
                            const double* phi, size_t n, double* partials,
                            int blocks)
{
    energy_partials_kernel<<<blocks, 256, 0, s>>>(m, vx, vy, vz, phi, n, partials);
    NBH_HIP(hipGetLastError());
}

void launch_force_vjp_fp64(hipStream_t s, const nbh_octree_view& o,
                           int max_level, const int2* pos,
                           const Walk64Rec* w64, const double4* wcenter,
                           const int* d_wcount, int wcount_host,
                           const double4* rec, const uint32_t* idx, int n,
                           double eps2, double G, bool contains,
                           const double* abx, const double* aby,
                           const double* abz, double* scratch, double* gx,
                           double* gy, double* gz, double* gm, int* split,
                           uint32_t* perm, uint32_t* pcost, int* err)
{
    // scratch: tb (3n) | pb (3n) | pm (n) | nb (3 wcount) | nm (wcount) |
    //          u (3 n_nodes) | w (n_nodes)
    const size_t N = static_cast<size_t>(n);
    const size_t W = static_cast<size_t>(wcount_host);
    const size_t M = static_cast<size_t>(o.n_nodes);
    double* tb = scratch;
    double* pb = tb + 3 * N;
    double* pm = pb + 3 * N;
    double* nbar = pm + N;
    double* nm = nbar + 3 * W;
    double* sv = nm + W;
    launch_fill_f64(s, pb, 0.0, 4 * N + 4 * W);
    OpArgs a{};
    a.abx = abx;
    a.aby = aby;
    a.abz = abz;
    a.tbx = tb;
    a.tby = tb + N;
    a.tbz = tb + 2 * N;
    a.pbx = pb;
    a.pby = pb + N;
    a.pbz = pb + 2 * N;
    a.nbx = nbar;
    a.nby = nbar + W;
    a.nbz = nbar + 2 * W;
    a.pm = pm;
    a.nm = nm;
    // the batched source sums address component c at base + c * stride
    if (a.pbz != a.pbx + 2 * N || a.pm != a.pbx + 3 * N || a.nbz != a.nbx + 2 * W ||
        a.nm != a.nbx + 3 * W)
        fail(NBH_LOGIC_ERROR, "force VJP source arrays are not laid out at fixed strides");
    if (split && perm && pcost && n > 0) {
        // few packets: units of one subtree below the first record each, by
        // decreasing visit count in the previous product
        const int packets = (n + 63) / 64;
        walk_split64_kernel<<<1, 64, 0, s>>>(w64, d_wcount, split);
        packet_schedule_kernel<<<8, kScheduleThreads, 0, s>>>(pcost, packets, split,
                                                             packets * kWalkSplitMax, perm);
        launch_fill_f64(s, tb, 0.0, 3 * N);
        NBH_HIP(hipGetLastError());
        a.split = split;
        a.perm = perm;
        a.pcost = pcost;
        const unsigned blocks = blocks_for(static_cast<size_t>(packets) * kWalkSplitMax, 4);
        if (contains)
            walk_op_kernel<true, OP_VJP, true><<<blocks, kWalkThreads, 0, s>>>(
                w64, wcenter, d_wcount, rec, idx, n, 0, eps2, G, a, err);
        else
            walk_op_kernel<false, OP_VJP, true><<<blocks, kWalkThreads, 0, s>>>(
                w64, wcenter, d_wcount, rec, idx, n, 0, eps2, G, a, err);
        NBH_HIP(hipGetLastError());
    } else {
        launch_op<OP_VJP>(s, w64, wcenter, d_wcount, rec, idx, n, eps2, G, contains, a,
                          err);
    }
    unsigned blocks = blocks_for(M, 256);
    if (blocks > 2048)
        blocks = 2048;
    for (int L = 0; L <= max_level; ++L) {
        push_down_kernel<<<blocks, 256, 0, s>>>(
            o, L, pos, nullptr, wcount_host, a.nbx, a.nby, a.nbz, a.nm, sv, sv + M,
            sv + 2 * M, sv + 3 * M, rec, a.pbx, a.pby, a.pbz, a.pm);
        NBH_HIP(hipGetLastError());
    }
    combine_kernel<<<blocks_for(N, 256), 256, 0, s>>>(idx, n, a.tbx, a.tby,
                                                     a.tbz, a.pbx, a.pby, a.pbz,
                                                     a.pm, gx, gy, gz, gm, err);
    NBH_HIP(hipGetLastError());
}

// The mixed walk's force VJP: the VJP walk over the mixed records, then the
// same FP64 push-down of the node cotangents and the same combination as
// the FP64 product (nodes the layout collapsed into their child skipped).
void launch_force_vjp_mixed(hipStream_t s, const nbh_octree_view& o, int max_level,
                            const int2* pos, const int4* aux, int4* wrec,
                            const int4* wcenter, const int* d_wcount, int* split,
                            int wcount_host, const int4* rec32, const double4* rec64,
                            const uint32_t* idx, const uint32_t* order, int n, float eps2,
                            double gunit, double munit, bool contains, bool soft,
                            const double* abx, const double* aby, const double* abz,
                            double* scratch, double* gx, double* gy, double* gz,
                            double* gm, uint32_t* perm, const uint32_t* cost, int* err)
{
    const size_t N = static_cast<size_t>(n);
    const size_t W = static_cast<size_t>(wcount_host);
    const size_t M = static_cast<size_t>(o.n_nodes);
    double* tb = scratch;
    double* pb = tb + 3 * N;
    double* pm = pb + 3 * N;
    double* nbar = pm + N;
    double* nm = nbar + 3 * W;
    double* sv = nm + W;
    launch_fill_f64(s, tb, 0.0, 7 * N + 4 * W);
    OpArgs a{};
    a.abx = abx;
    a.aby = aby;
    a.abz = abz;
    a.tbx = tb;
    a.tby = tb + N;
    a.tbz = tb + 2 * N;
    a.pbx = pb;
    a.pby = pb + N;
    a.pbz = pb + 2 * N;
    a.nbx = nbar;
    a.nby = nbar + W;
    a.nbz = nbar + 2 * W;
    a.pm = pm;
    a.nm = nm;
    // few packets: one unit per subtree below the first record
    if (split)
        launch_walk_split(s, wrec, d_wcount, split);
    const int units = split ? kWalkSplitMax : 1;
    // Split walks deal their units by decreasing cost: the force walk's
    // visit counts per unit, which the VJP walk's visits equal (the same
    // records, opening test and units).
    const bool ordered = split && perm && cost && n > 0;
    if (ordered) {
        const int packets = (n + 63) / 64;
        packet_schedule_kernel<<<8, kScheduleThreads, 0, s>>>(cost, packets, split,
                                                             packets * kWalkSplitMax, perm);
        NBH_HIP(hipGetLastError());
    }
    // the batched source sums address component c at base + c * stride
    if (a.pby != a.pbx + N || a.pbz != a.pbx + 2 * N || a.pm != a.pbx + 3 * N ||
        a.nby != a.nbx + W || a.nbz != a.nbx + 2 * W || a.nm != a.nbx + 3 * W)
        fail(NBH_LOGIC_ERROR, "force VJP source arrays are not laid out at fixed strides");
    const MixedVjp mv{wrec, wcenter, d_wcount, split, rec32, idx, order, n, units, eps2,
                      gunit, munit, N, W, ordered ? perm : nullptr};
    const unsigned blocks = blocks_for((N + 63) / 64 * units, 4);
    if (contains) {
        if (soft)
            vjp_mixed_kernel<true, true><<<blocks, kWalkThreads, 0, s>>>(mv, a, err);
        else
            vjp_mixed_kernel<true, false><<<blocks, kWalkThreads, 0, s>>>(mv, a, err);
    } else {
        if (soft)
            vjp_mixed_kernel<false, true><<<blocks, kWalkThreads, 0, s>>>(mv, a, err);
        else
            vjp_mixed_kernel<false, false><<<blocks, kWalkThreads, 0, s>>>(mv, a, err);
    }
    NBH_HIP(hipGetLastError());
    unsigned pblocks = blocks_for(M, 256);
    if (pblocks > 2048)
        pblocks = 2048;
    for (int L = 0; L <= max_level; ++L) {
        push_down_kernel<<<pblocks, 256, 0, s>>>(
            o, L, pos, aux, wcount_host, a.nbx, a.nby, a.nbz, a.nm, sv, sv + M,
            sv + 2 * M, sv + 3 * M, rec64, a.pbx, a.pby, a.pbz, a.pm);
        NBH_HIP(hipGetLastError());
    }
    combine_kernel<<<blocks_for(N, 256), 256, 0, s>>>(idx, n, a.tbx, a.tby, a.tbz, a.pbx,
                                                     a.pby, a.pbz, a.pm, gx, gy, gz, gm,
                                                     err);
    NBH_HIP(hipGetLastError());
}

size_t force_vjp_scratch_doubles(size_t n, size_t wcount, size_t n_nodes)
{
    return 7 * n + 4 * wcount + 4 * n_nodes;
}

} // namespace nbh
