// walk.hip -- the traversal: derived depth-first layout and the wavefront
// packet walk, FP64 and mixed precision.
//
// Layout.  The reference walks the octree per target with parent climbs and
// 8-slot sibling scans, visiting octants 7 -> 0 (barnes_hut.cu:90-162).  Here
// the same visiting order is laid out once per build: nodes are stored at
// their depth-first pre-order position (children 7 -> 0), each record
// carrying `skip` = pre + subtree size, the position that follows its
// subtree.  "Open" is then pre + 1 and "done with this subtree" is skip: no
// stack, no parent reads.
//
// Packets.  One wavefront owns 64 Morton-consecutive targets.  The wave
// walks the layout in increasing position.  A lane is active at position i
// when its own next position is <= i; active lanes evaluate the node exactly
// as the per-target walk would (leaf -> exact particles, zero mass -> skip,
// else MAC: accept -> monopole, reject -> open).  If any lane opens, the
// wave steps to i + 1, else to skip[i].  Every lane therefore sees precisely
// the node sequence of its own per-target walk, in the same order, while node
// records are fetched once per wave with scalar loads.
//
// Opening test.  The reference accepts a node when the target is outside the
// cell, d^2 > 0, theta > 0 and d > s/theta + delta (barnes_hut.cu:273-294).
// For theta < 2/sqrt(3) the containment test never decides: inside the cell
// d <= delta + sqrt(3) h < 2h/theta + delta.  The kernels therefore load
// cell geometry only when theta >= 1.15.
//
// Precision.  FP64 reproduces the reference arithmetic (and is bit-exact
// with the oracle).  Mixed evaluates the opening test and the interaction in
// FP32 on exact integer differences of 31-bit fixed-point coordinates over
// the bounding box -- coordinates relative to the target, so the rounding of
// each separation is relative to the separation itself -- and sums each
// target's acceleration with Kahan compensation.
#include "device_utils.hpp"
#include "nbh_internal.hpp"
#include "octree_device.hpp"

namespace nbh {

// Node kinds of the walk records.
enum : int {
    KIND_INTERNAL = 0,   // MAC decides accept / open
    KIND_LEAF = 1,       // exact particle interactions
    KIND_SKIP = 2,       // zero mass: contributes nothing, never opened
    KIND_CHILDLESS = 3,  // internal with mass but no children: open -> error 5
    KIND_ERROR = 100     // kind - 100 is the traversal error code
};

namespace {

// ------------------------------------------------------ pipeline climb
// Occupied leaves start; the last child to arrive at a parent combines the
// children (octant order 0..7, octree_physics.cu:155-244), computes the
// parent's cell geometry (octree_geometry.cu:56-102), its subtree size, the
// pre-order offsets of its children (octants 7 -> 0) and its opening radius
// side/theta + |com - centre| (barnes_hut.cu:277-291).
template <int BITS>
__global__ __launch_bounds__(256) void climb_kernel(
    nbh_octree_view o, const double4* __restrict__ rec, double theta,
    const double* __restrict__ box, int* __restrict__ size,
    int* __restrict__ local, double* __restrict__ crit, int* err)
{
    using Key = typename Layout<BITS>::Key;
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= o.n_nodes)
        return;
    const int count = o.particle_count[node];
    if (count <= 0)
        return;
    const double rcx = box[0], rcy = box[1], rcz = box[2], side = box[3];
    const Key* prefix = static_cast<const Key*>(o.prefix);
    const int f = o.first_particle[node];
    double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
    for (int k = 0; k < count; ++k) {
        const double4 p = rec[f + k];
        tm += p.w;
        wx = fma(p.w, p.x, wx);
        wy = fma(p.w, p.y, wy);
        wz = fma(p.w, p.z, wz);
    }
    double cx = 0.0, cy = 0.0, cz = 0.0;
    if (tm > 0.0) {
        const double inv = 1.0 / tm;
        cx = wx * inv;
        cy = wy * inv;
        cz = wz * inv;
    }
    o.mass[node] = tm;
    o.com_x[node] = cx;
    o.com_y[node] = cy;
    o.com_z[node] = cz;
    double gx, gy, gz, gh;
    int code = 0;
    if (!node_geometry<BITS>(prefix[node], o.level[node], rcx, rcy, rcz, side,
                             gx, gy, gz, gh, code)) {
        raise_error(err, 8);
        return;
    }
    o.center_x[node] = gx;
    o.center_y[node] = gy;
    o.center_z[node] = gz;
    o.half_size[node] = gh;
    size[node] = 1;
    int v = node;
    while (v != 0) {
        const int p = o.parent[v];
        if (p < 0 || p >= o.n_nodes) {
            raise_error(err, 5);
            return;
        }
        const int required = o.child_count[p];
        const int done = fetch_add_acq_rel(&o.pending_children[p], 1) + 1;
        if (done < required)
            return;
        if (done > required) {
            raise_error(err, 7);
            return;
        }
        int ch[8];
#pragma unroll
        for (int oc = 0; oc < 8; ++oc)
            ch[oc] = o.children[p * 8 + oc];
        tm = 0.0;
        wx = wy = wz = 0.0;
#pragma unroll
        for (int oc = 0; oc < 8; ++oc) {
            if (ch[oc] == -1)
                continue;
            const double cm = o.mass[ch[oc]];
            tm += cm;
            wx = fma(cm, o.com_x[ch[oc]], wx);
            wy = fma(cm, o.com_y[ch[oc]], wy);
            wz = fma(cm, o.com_z[ch[oc]], wz);
        }
        int off = 1;
#pragma unroll
        for (int oc = 7; oc >= 0; --oc) {
            if (ch[oc] == -1)
                continue;
            local[ch[oc]] = off;
            off += size[ch[oc]];
        }
        size[p] = off;
        cx = cy = cz = 0.0;
        if (tm > 0.0) {
            const double inv = 1.0 / tm;
            cx = wx * inv;
            cy = wy * inv;
            cz = wz * inv;
        }
        o.mass[p] = tm;
        o.com_x[p] = cx;
        o.com_y[p] = cy;
        o.com_z[p] = cz;
        if (!node_geometry<BITS>(prefix[p], o.level[p], rcx, rcy, rcz, side,
                                 gx, gy, gz, gh, code)) {
            raise_error(err, 8);
            return;
        }
        o.center_x[p] = gx;
        o.center_y[p] = gy;
        o.center_z[p] = gz;
        o.half_size[p] = gh;
        const double ox = cx - gx;
        const double oy = cy - gy;
        const double oz = cz - gz;
        const double o2 = fma(oz, oz, fma(ox, ox, oy * oy));
        crit[p] = (2.0 * gh) / theta + sqrt(o2);
        v = p;
    }
    local[0] = 0;
}

// ------------------------------------------------- generic tree sizes
// For octrees handed in through the C ABI: child ids must be valid (else
// code 1, barnes_hut.cu:101-105) and every child must name its parent (else
// code 4, barnes_hut.cu:150-154).  Writes the child counts to the scratch
// array o.child_count.
__global__ __launch_bounds__(256) void check_parents_kernel(nbh_octree_view o,
                                                            int* err)
{
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= o.n_nodes)
        return;
    int count = 0;
    for (int oc = 0; oc < 8; ++oc) {
        const int c = o.children[node * 8 + oc];
        if (c == -1)
            continue;
        if (c <= 0 || c >= o.n_nodes) {
            raise_error(err, 1);
            return;
        }
        if (o.parent[c] != node)
            raise_error(err, 4);
        ++count;
    }
    o.child_count[node] = count;
}

// Subtree sizes and pre-order offsets, started from every childless node.
__global__ __launch_bounds__(256) void generic_sizes_kernel(
    nbh_octree_view o, int* __restrict__ size, int* __restrict__ local,
    int* err)
{
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= o.n_nodes)
        return;
    if (o.child_count[node] != 0)
        return;
    size[node] = 1;
    int v = node;
    int guard = 0;
    while (v != 0) {
        const int p = o.parent[v];
        if (p < 0 || p >= o.n_nodes || ++guard > o.n_nodes) {
            // not connected to the root: never visited by the traversal
            return;
        }
        const int required = o.child_count[p];
        const int done = fetch_add_acq_rel(&o.pending_children[p], 1) + 1;
        if (done < required)
            return;
        if (done > required) {
            raise_error(err, 4);
            return;
        }
        int off = 1;
        for (int oc = 7; oc >= 0; --oc) {
            const int c = o.children[p * 8 + oc];
            if (c <= 0 || c >= o.n_nodes)
                continue;
            local[c] = off;
            off += size[c];
        }
        size[p] = off;
        v = p;
    }
    local[0] = 0;
}

// Pre-order position of a node: the sum of the sibling offsets along its
// path to the root.  Returns -1 when the path does not reach node 0.
__device__ __forceinline__ int preorder(const nbh_octree_view& o,
                                        const int* __restrict__ local, int v)
{
    int pre = 0;
    int guard = 0;
    while (v != 0) {
        pre += local[v];
        v = o.parent[v];
        if (v < 0 || ++guard > 64)
            return -1;
    }
    return pre;
}

// --------------------------------------------------------- walk records
// GENERIC: records of a caller-supplied octree, with the reference's
// per-node failure conditions encoded as error kinds and the opening radius
// computed here.  Otherwise: records of the pipeline's own octree.
template <bool GENERIC, bool MIXED>
__global__ __launch_bounds__(256) void walk_build_kernel(
    nbh_octree_view o, const int* __restrict__ size,
    const int* __restrict__ local, const double* __restrict__ crit_in,
    double theta, const double4* __restrict__ rec64,
    const int4* __restrict__ rec32, const double* __restrict__ box,
    double G, bool need_center, double4* __restrict__ w64node,
    double* __restrict__ w64crit, int4* __restrict__ w64link,
    double4* __restrict__ w64center, int4* __restrict__ w32node,
    int4* __restrict__ w32link, int4* __restrict__ w32center,
    int* __restrict__ wcount)
{
    const int v = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= o.n_nodes)
        return;
    const int pre = preorder(o, local, v);
    if (pre < 0 || pre >= o.n_nodes)
        return;
    if (v == 0)
        *wcount = size[0] < o.n_nodes ? size[0] : o.n_nodes;
    const int skip = pre + size[v];
    const int pc = o.particle_count[v];
    const double h = o.half_size[v];
    int kind;
    int first = 0, count = 0;
    if (GENERIC) {
        const int nch = o.child_count[v];
        if (!(h > 0.0)) {
            kind = KIND_ERROR + 8;
        } else if (pc > 0) {
            first = o.first_particle[v];
            count = pc;
            if (nch != 0)
                kind = KIND_ERROR + 7;
            else if (first < 0 || pc > o.n_particles ||
                     first > o.n_particles - pc)
                kind = KIND_ERROR + 2;
            else
                kind = KIND_LEAF;
        } else if (o.mass[v] == 0.0) {
            kind = KIND_SKIP;
        } else if (nch == 0) {
            kind = KIND_CHILDLESS;
        } else {
            kind = KIND_INTERNAL;
        }
    } else {
        if (pc > 0) {
            kind = KIND_LEAF;
            first = o.first_particle[v];
            count = pc;
        } else if (o.mass[v] == 0.0) {
            kind = KIND_SKIP;
        } else {
            kind = KIND_INTERNAL;
        }
    }
    const double m = o.mass[v];
    const double cx = o.com_x[v], cy = o.com_y[v], cz = o.com_z[v];
    double crit = 0.0;
    if (kind == KIND_INTERNAL || kind == KIND_CHILDLESS) {
        if (GENERIC) {
            const double ox = cx - o.center_x[v];
            const double oy = cy - o.center_y[v];
            const double oz = cz - o.center_z[v];
            const double o2 = fma(oz, oz, fma(ox, ox, oy * oy));
            crit = (2.0 * h) / theta + sqrt(o2);
        } else {
            crit = crit_in[v];
        }
    }
    if (!MIXED) {
        if (kind == KIND_LEAF && count == 1)
            w64node[pre] = rec64[first];
        else
            w64node[pre] = make_double4(cx, cy, cz, m);
        w64crit[pre] = crit;
        w64link[pre] = make_int4(skip, kind, first, count);
        if (need_center)
            w64center[pre] =
                make_double4(o.center_x[v], o.center_y[v], o.center_z[v], h);
    } else {
        const double side = box[3];
        const double scale = 2147483648.0 / side;
        const double ox = box[0] - side / 2.0;
        const double oy = box[1] - side / 2.0;
        const double oz = box[2] - side / 2.0;
        const double lim = 2147483647.0;
        if (kind == KIND_LEAF && count == 1) {
            w32node[pre] = rec32[first];
        } else {
            const int qx = static_cast<int>(fmin(fmax(rint((cx - ox) * scale), 0.0), lim));
            const int qy = static_cast<int>(fmin(fmax(rint((cy - oy) * scale), 0.0), lim));
            const int qz = static_cast<int>(fmin(fmax(rint((cz - oz) * scale), 0.0), lim));
            const float gm = static_cast<float>(G * m * scale * scale);
            w32node[pre] = make_int4(qx, qy, qz, __float_as_int(gm));
        }
        const double cq = crit * scale;
        const float crit2 = static_cast<float>(cq * cq);
        w32link[pre] = make_int4(__float_as_int(crit2), skip,
                                 kind | (count << 8), first);
        if (need_center) {
            const int qx = static_cast<int>(rint((o.center_x[v] - ox) * scale));
            const int qy = static_cast<int>(rint((o.center_y[v] - oy) * scale));
            const int qz = static_cast<int>(rint((o.center_z[v] - oz) * scale));
            const int qh = static_cast<int>(fmin(rint(h * scale), lim));
            w32center[pre] = make_int4(qx, qy, qz, qh);
        }
    }
}

// Every link starts as "error 4, skip to the end", so a position that the
// build leaves unwritten (a node off the root's tree) stops the walk with a
// topology error instead of reading an undefined record.
__global__ __launch_bounds__(256) void link_sentinel_kernel(int4* __restrict__ link,
                                                            int4 v, int n)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        link[i] = v;
}

// ------------------------------------------------------------ traversal
constexpr int kWalkThreads = 256; // 4 independent wavefronts per block

// Packets are Morton-consecutive; blocks are dealt to the 8 XCDs round
// robin, so block b takes packet range (b % 8) * per + b / 8: each XCD's L2
// then serves one contiguous eighth of space.
__device__ __forceinline__ int packet_of_block(int nblocks_padded)
{
    const int b = static_cast<int>(blockIdx.x);
    const int per = nblocks_padded / 8;
    return (b % 8) * per + b / 8;
}

// barnes_hut.cu:56-80 in FP64 with the contraction order of the oracle.
__device__ __forceinline__ void pm_fp64(double tx, double ty, double tz,
                                        double sx, double sy, double sz,
                                        double sm, double G, double eps2,
                                        double& ax, double& ay, double& az,
                                        bool& singular)
{
    if (sm == 0.0)
        return;
    const double dx = sx - tx;
    const double dy = sy - ty;
    const double dz = sz - tz;
    const double d2 = fma(dz, dz, fma(dx, dx, dy * dy)) + eps2;
    if (d2 <= 0.0) {
        singular = true;
        return;
    }
    const double inv = 1.0 / sqrt(d2);
    const double inv3 = inv * inv * inv;
    const double f = G * sm * inv3;
    ax = fma(f, dx, ax);
    ay = fma(f, dy, ay);
    az = fma(f, dz, az);
}

template <bool CONTAINS, bool COUNT>
__global__ __launch_bounds__(kWalkThreads) void walk_fp64_kernel(
    const double4* __restrict__ wnode, const double* __restrict__ wcrit,
    const int4* __restrict__ wlink, const double4* __restrict__ wcenter,
    const int* __restrict__ d_wcount, const double4* __restrict__ rec,
    const uint32_t* __restrict__ idx, int n, int nblocks_padded, double eps2,
    double G, double* __restrict__ ax_out, double* __restrict__ ay_out,
    double* __restrict__ az_out, int* err,
    unsigned long long* __restrict__ counters)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int packet = packet_of_block(nblocks_padded) * 4 + wave;
    const int t = packet * 64 + lane;
    if (packet * 64 >= n)
        return;
    const bool valid = t < n;
    const int wcount = *d_wcount;
    double tx = 0.0, ty = 0.0, tz = 0.0;
    if (valid) {
        const double4 p = rec[t];
        tx = p.x;
        ty = p.y;
        tz = p.z;
    }
    double ax = 0.0, ay = 0.0, az = 0.0;
    unsigned long long npp = 0, npn = 0;
    int next = valid ? 0 : 0x7FFFFFFF;
    int code = 0;
    int i = 0;
    while (i < wcount) {
        const int4 lk = wlink[i];
        const int skip = lk.x;
        const int kind = lk.y;
        bool open = false;
        if (next <= i && code == 0) {
            if (kind == KIND_INTERNAL || kind == KIND_CHILDLESS) {
                const double4 nd = wnode[i];
                const double dx = nd.x - tx;
                const double dy = nd.y - ty;
                const double dz = nd.z - tz;
                const double d2 = fma(dz, dz, fma(dx, dx, dy * dy));
                bool accept = sqrt(d2) > wcrit[i];
                if (CONTAINS && accept) {
                    const double4 c = wcenter[i];
                    const bool inside = fabs(tx - c.x) <= c.w &&
                                        fabs(ty - c.y) <= c.w &&
                                        fabs(tz - c.z) <= c.w;
                    accept = !inside;
                }
                if (accept) {
                    bool sing = false;
                    pm_fp64(tx, ty, tz, nd.x, nd.y, nd.z, nd.w, G, eps2, ax,
                            ay, az, sing);
                    if (sing)
                        code = 6;
                    if (COUNT)
                        ++npn;
                    next = skip;
                } else if (kind == KIND_CHILDLESS) {
                    code = 5;
                } else {
                    open = true;
                    next = i + 1;
                }
            } else if (kind == KIND_LEAF) {
                const int first = lk.z;
                const int count = lk.w;
                bool sing = false;
                if (count == 1) {
                    if (first != t) {
                        const double4 s = wnode[i];
                        pm_fp64(tx, ty, tz, s.x, s.y, s.z, s.w, G, eps2, ax, ay,
                                az, sing);
                        if (COUNT && s.w != 0.0)
                            ++npp;
                    }
                } else {
                    for (int k = first; k < first + count; ++k) {
                        if (k == t)
                            continue;
                        const double4 s = rec[k];
                        pm_fp64(tx, ty, tz, s.x, s.y, s.z, s.w, G, eps2, ax, ay,
                                az, sing);
                        if (COUNT && s.w != 0.0)
                            ++npp;
                    }
                }
                if (sing)
                    code = 6;
                next = skip;
            } else if (kind == KIND_SKIP) {
                next = skip;
            } else {
                code = kind - KIND_ERROR;
            }
        }
        i = (__ballot(open) != 0ull) ? i + 1 : skip;
    }
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
        }
        return;
    }
    if (valid) {
        const uint32_t o = idx[t];
        if (o >= static_cast<uint32_t>(n)) {
            raise_error(err, 3);
            return;
        }
        ax_out[o] = ax;
        ay_out[o] = ay;
        az_out[o] = az;
    }
}

// Kahan-compensated FP32 accumulation of one component.
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

// One softened FP32 interaction with separation (dx, dy, dz) and squared
// distance d2 (grid units); gm = G m / unit^2.
__device__ __forceinline__ void pm_fp32(float dx, float dy, float dz, float d2,
                                        float gm, float eps2, Kahan& ax,
                                        Kahan& ay, Kahan& az)
{
    const float r2 = d2 + eps2;
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float inv3 = inv * inv * inv;
    const float f = gm * inv3;
    ax.add(f * dx);
    ay.add(f * dy);
    az.add(f * dz);
}

// SOFT: eps > 0, so coincident points and the target itself contribute an
// exact zero and need no test.  Otherwise the self term is skipped by index
// and a zero separation with mass is singular (code 6).
template <bool CONTAINS, bool SOFT, bool COUNT>
__global__ __launch_bounds__(kWalkThreads) void walk_mixed_kernel(
    const int4* __restrict__ wnode, const int4* __restrict__ wlink,
    const int4* __restrict__ wcenter, const int* __restrict__ d_wcount,
    const int4* __restrict__ rec, const uint32_t* __restrict__ idx, int n,
    int nblocks_padded, float eps2, double* __restrict__ ax_out,
    double* __restrict__ ay_out, double* __restrict__ az_out, int* err,
    unsigned long long* __restrict__ counters)
{
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
    Kahan ax, ay, az;
    unsigned long long npp = 0, npn = 0;
    int next = valid ? 0 : 0x7FFFFFFF;
    int code = 0;
    int i = 0;
    while (i < wcount) {
        const int4 lk = wlink[i];
        const int skip = lk.y;
        const int kind = lk.z & 0xFF;
        bool open = false;
        if (next <= i && code == 0) {
            if (kind == KIND_INTERNAL || kind == KIND_CHILDLESS) {
                const int4 nd = wnode[i];
                const float dx = static_cast<float>(nd.x - tq.x);
                const float dy = static_cast<float>(nd.y - tq.y);
                const float dz = static_cast<float>(nd.z - tq.z);
                const float d2 = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
                bool accept = d2 > __int_as_float(lk.x);
                if (CONTAINS && accept) {
                    const int4 c = wcenter[i];
                    const bool inside = abs(tq.x - c.x) <= c.w &&
                                        abs(tq.y - c.y) <= c.w &&
                                        abs(tq.z - c.z) <= c.w;
                    accept = !inside;
                }
                if (accept) {
                    pm_fp32(dx, dy, dz, d2, __int_as_float(nd.w), eps2, ax, ay,
                            az);
                    if (COUNT)
                        ++npn;
                    next = skip;
                } else if (kind == KIND_CHILDLESS) {
                    code = 5;
                } else {
                    open = true;
                    next = i + 1;
                }
            } else if (kind == KIND_LEAF) {
                const int count = lk.z >> 8;
                const int first = lk.w;
                if (count == 1) {
                    const int4 s = wnode[i];
                    const float dx = static_cast<float>(s.x - tq.x);
                    const float dy = static_cast<float>(s.y - tq.y);
                    const float dz = static_cast<float>(s.z - tq.z);
                    const float d2 = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
                    const float gm = __int_as_float(s.w);
                    if (SOFT) {
                        pm_fp32(dx, dy, dz, d2, gm, eps2, ax, ay, az);
                    } else if (first != t && gm != 0.0f) {
                        if (d2 == 0.0f)
                            code = 6;
                        else
                            pm_fp32(dx, dy, dz, d2, gm, eps2, ax, ay, az);
                    }
                    if (COUNT && first != t && gm != 0.0f)
                        ++npp;
                } else {
                    for (int k = first; k < first + count; ++k) {
                        const int4 s = rec[k];
                        const float dx = static_cast<float>(s.x - tq.x);
                        const float dy = static_cast<float>(s.y - tq.y);
                        const float dz = static_cast<float>(s.z - tq.z);
                        const float d2 = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
                        const float gm = __int_as_float(s.w);
                        if (SOFT) {
                            pm_fp32(dx, dy, dz, d2, gm, eps2, ax, ay, az);
                        } else if (k != t && gm != 0.0f) {
                            if (d2 == 0.0f)
                                code = 6;
                            else
                                pm_fp32(dx, dy, dz, d2, gm, eps2, ax, ay, az);
                        }
                        if (COUNT && k != t && gm != 0.0f)
                            ++npp;
                    }
                }
                next = skip;
            } else if (kind == KIND_SKIP) {
                next = skip;
            } else {
                code = kind - KIND_ERROR;
            }
        }
        i = (__ballot(open) != 0ull) ? i + 1 : skip;
    }
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
        }
        return;
    }
    if (valid) {
        const uint32_t o = idx[t];
        if (o >= static_cast<uint32_t>(n)) {
            raise_error(err, 3);
            return;
        }
        ax_out[o] = ax.value();
        ay_out[o] = ay.value();
        az_out[o] = az.value();
    }
}

// Softened FP64 direct sum on a target list: one block per target, fixed
// reduction order.
__global__ __launch_bounds__(256) void direct_kernel(
    const double* __restrict__ m, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, size_t n,
    const uint32_t* __restrict__ targets, double eps2, double G,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az)
{
    __shared__ double s_red[3][4];
    const uint32_t t = targets[blockIdx.x];
    if (t >= n)
        return;
    const double tx = x[t], ty = y[t], tz = z[t];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    bool sing = false;
    for (size_t s = threadIdx.x; s < n; s += blockDim.x) {
        if (s == t)
            continue;
        pm_fp64(tx, ty, tz, x[s], y[s], z[s], m[s], G, eps2, a0, a1, a2, sing);
    }
    for (int o = 32; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o);
        a1 += __shfl_xor(a1, o);
        a2 += __shfl_xor(a2, o);
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[0][wave] = a0;
        s_red[1][wave] = a1;
        s_red[2][wave] = a2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ax[blockIdx.x] = (s_red[0][0] + s_red[0][1]) + (s_red[0][2] + s_red[0][3]);
        ay[blockIdx.x] = (s_red[1][0] + s_red[1][1]) + (s_red[1][2] + s_red[1][3]);
        az[blockIdx.x] = (s_red[2][0] + s_red[2][1]) + (s_red[2][2] + s_red[2][3]);
    }
}

} // namespace

// ---------------------------------------------------------------- hosts
template <int BITS>
void launch_climb(hipStream_t s, const nbh_octree_view& o, const double4* rec,
                  double theta, const double* box, int* size, int* local,
                  double* crit, int* err)
{
    climb_kernel<BITS><<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(
        o, rec, theta, box, size, local, crit, err);
    NBH_HIP(hipGetLastError());
}
template void launch_climb<10>(hipStream_t, const nbh_octree_view&, const double4*, double,
                               const double*, int*, int*, double*, int*);
template void launch_climb<21>(hipStream_t, const nbh_octree_view&, const double4*, double,
                               const double*, int*, int*, double*, int*);

void launch_generic_sizes(hipStream_t s, const nbh_octree_view& o, int* size,
                          int* local, int* err)
{
    check_parents_kernel<<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(o, err);
    NBH_HIP(hipGetLastError());
    generic_sizes_kernel<<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(o, size,
                                                                    local, err);
    NBH_HIP(hipGetLastError());
}

void launch_walk_build(hipStream_t s, const nbh_octree_view& o, bool generic,
                       bool mixed, const int* size, const int* local,
                       const double* crit, double theta, const double4* rec64,
                       const int4* rec32, const double* box, double G,
                       bool need_center, double4* w64node, double* w64crit,
                       int4* w64link, double4* w64center, int4* w32node,
                       int4* w32link, int4* w32center, int* wcount)
{
    const unsigned blocks = blocks_for(o.n_nodes, 256);
    const int M = o.n_nodes;
    if (mixed)
        link_sentinel_kernel<<<blocks, 256, 0, s>>>(
            w32link, make_int4(0, M, KIND_ERROR + 4, 0), M);
    else
        link_sentinel_kernel<<<blocks, 256, 0, s>>>(
            w64link, make_int4(M, KIND_ERROR + 4, 0, 0), M);
    NBH_HIP(hipGetLastError());
#define NBH_WALK_ARGS                                                          \
    o, size, local, crit, theta, rec64, rec32, box, G, need_center, w64node,     \
        w64crit, w64link, w64center, w32node, w32link, w32center, wcount
    if (generic && !mixed)
        walk_build_kernel<true, false><<<blocks, 256, 0, s>>>(NBH_WALK_ARGS);
    else if (generic && mixed)
        walk_build_kernel<true, true><<<blocks, 256, 0, s>>>(NBH_WALK_ARGS);
    else if (!mixed)
        walk_build_kernel<false, false><<<blocks, 256, 0, s>>>(NBH_WALK_ARGS);
    else
        walk_build_kernel<false, true><<<blocks, 256, 0, s>>>(NBH_WALK_ARGS);
#undef NBH_WALK_ARGS
    NBH_HIP(hipGetLastError());
}

// Blocks of 4 packets, padded to a multiple of 8 for the XCD mapping.
int walk_blocks_padded(int n)
{
    const int packets = (n + 63) / 64;
    const int blocks = (packets + 3) / 4;
    return (blocks + 7) / 8 * 8;
}

void launch_walk_fp64(hipStream_t s, const double4* wnode, const double* wcrit,
                      const int4* wlink, const double4* wcenter,
                      const int* d_wcount, const double4* rec,
                      const uint32_t* idx, int n, double eps2, double G,
                      bool contains, double* ax, double* ay, double* az,
                      int* err, unsigned long long* counters)
{
    const int nb = walk_blocks_padded(n);
#define NBH_ARGS                                                                 \
    wnode, wcrit, wlink, wcenter, d_wcount, rec, idx, n, nb, eps2, G, ax, ay, az,  \
        err, counters
    if (counters) {
        if (contains)
            walk_fp64_kernel<true, true><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);
        else
            walk_fp64_kernel<false, true><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);
    } else {
        if (contains)
            walk_fp64_kernel<true, false><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);
        else
            walk_fp64_kernel<false, false><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);
    }
#undef NBH_ARGS
    NBH_HIP(hipGetLastError());
}

void launch_walk_mixed(hipStream_t s, const int4* wnode, const int4* wlink,
                       const int4* wcenter, const int* d_wcount,
                       const int4* rec, const uint32_t* idx, int n,
                       float eps2, bool contains, bool soft, double* ax,
                       double* ay, double* az, int* err,
                       unsigned long long* counters)
{
    const int nb = walk_blocks_padded(n);
#define NBH_ARGS                                                                 \
    wnode, wlink, wcenter, d_wcount, rec, idx, n, nb, eps2, ax, ay, az, err,       \
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

void launch_direct(hipStream_t s, const double* m, const double* x,
                   const double* y, const double* z, size_t n,
                   const uint32_t* targets, size_t n_targets, double eps2,
                   double G, double* ax, double* ay, double* az)
{
    if (n_targets == 0)
        return;
    direct_kernel<<<static_cast<unsigned>(n_targets), 256, 0, s>>>(
        m, x, y, z, n, targets, eps2, G, ax, ay, az);
    NBH_HIP(hipGetLastError());
}

} // namespace nbh
