// karras.hip -- binary radix tree over sorted Morton keys (Karras 2012) and
// the legacy binary-tree mass / centre of mass.
//
// Tree construction follows tree_builder.cu:9-113: for internal node i the
// direction of its range comes from the longer common prefix with i+-1, the
// far end is found by exponential then binary search, and the split by a
// binary search on the common prefix with the first key.  delta(i, j) is
// clz(k_i ^ k_j) over the key storage width W (32 or 64); equal keys fall
// back to W + clz32(i ^ j); i == j gives 2W; out of range gives -1.
#include "device_utils.hpp"
#include "nbh_internal.hpp"

namespace nbh {

namespace {

template <typename K>
__device__ __forceinline__ int clz_key(K v)
{
    if constexpr (sizeof(K) == 4)
        return __clz(static_cast<int>(v));
    else
        return __clzll(static_cast<long long>(v));
}

template <typename K>
__device__ __forceinline__ int delta(const K* __restrict__ keys, int i, int j,
                                     int n)
{
    constexpr int W = static_cast<int>(sizeof(K) * 8);
    if (j < 0 || j >= n)
        return -1;
    if (i == j)
        return 2 * W;
    const K a = keys[i];
    const K b = keys[j];
    if (a != b)
        return clz_key<K>(a ^ b);
    return W + __clz(i ^ j);
}

template <typename K>
__global__ __launch_bounds__(256) void karras_kernel(
    const K* __restrict__ keys, const int* __restrict__ d_n, int n_host,
    int* __restrict__ left, int* __restrict__ right, int* __restrict__ parent,
    int* __restrict__ range_first, int* __restrict__ range_last,
    int* __restrict__ prefix_len, int* __restrict__ parent_prefix_len)
{
    const int n = d_n ? *d_n : n_host;
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i == 0 && d_n)
        parent[n > 1 ? n : 0] = -1; // the root has no parent
    if (i >= n - 1)
        return;
    // determineDirection / determineRange (tree_builder.cu:30-62)
    const int dr = delta(keys, i, i + 1, n);
    const int dl = delta(keys, i, i - 1, n);
    const int d = (dr > dl) ? 1 : -1;
    const int dmin = delta(keys, i, i - d, n);
    int length = 2;
    while (delta(keys, i, i + length * d, n) > dmin)
        length *= 2;
    int l = 0;
    for (int t = length / 2; t >= 1; t /= 2)
        if (delta(keys, i, i + (l + t) * d, n) > dmin)
            l += t;
    const int j = i + l * d;
    const int first = i < j ? i : j;
    const int last = i > j ? i : j;
    const int node_prefix = delta(keys, first, last, n);
    // findSplit (tree_builder.cu:64-84)
    int split = first;
    if (first < last) {
        int step = last - first;
        do {
            step = (step + 1) >> 1;
            const int c = split + step;
            if (c < last && delta(keys, first, c, n) > node_prefix)
                split = c;
        } while (step > 1);
    }
    const int lc = (split == first) ? split : n + split;
    const int rc = (split + 1 == last) ? split + 1 : n + split + 1;
    range_first[i] = first;
    range_last[i] = last;
    prefix_len[i] = node_prefix;
    left[i] = lc;
    right[i] = rc;
    parent[lc] = n + i;
    parent[rc] = n + i;
    if (parent_prefix_len) {
        // each child's parent prefix length, so that the radix -> octree
        // plan reads its parent's level sequentially
        parent_prefix_len[lc] = node_prefix;
        parent_prefix_len[rc] = node_prefix;
    }
}

// tree_builder.cu:115-180: leaf mass and centre of mass; one particle per
// leaf when first == nullptr, else the Morton group's particles.
__global__ __launch_bounds__(256) void tree_leaves_kernel(
    nbh_tree_view t, const int* __restrict__ first,
    const int* __restrict__ count, const uint32_t* __restrict__ idx,
    const double* __restrict__ m, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z)
{
    const int leaf = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (leaf >= t.n_leaves)
        return;
    if (!first) {
        const uint32_t p = idx[leaf];
        t.mass[leaf] = m[p];
        t.com_x[leaf] = x[p];
        t.com_y[leaf] = y[p];
        t.com_z[leaf] = z[p];
        return;
    }
    double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
    const int f = first[leaf];
    const int c = count[leaf];
    for (int k = 0; k < c; ++k) {
        const uint32_t p = idx[f + k];
        const double pm = m[p];
        tm += pm;
        wx = fma(pm, x[p], wx);
        wy = fma(pm, y[p], wy);
        wz = fma(pm, z[p], wz);
    }
    t.mass[leaf] = tm;
    if (tm > 0.0) {
        const double inv = 1.0 / tm;
        t.com_x[leaf] = wx * inv;
        t.com_y[leaf] = wy * inv;
        t.com_z[leaf] = wz * inv;
    } else {
        t.com_x[leaf] = 0.0;
        t.com_y[leaf] = 0.0;
        t.com_z[leaf] = 0.0;
    }
}

// tree_builder.cu:183-233: the second child to arrive at an internal node
// combines both; the arrival counter is an agent-scope acquire/release RMW so
// the winner sees its sibling's results from any XCD.
__global__ __launch_bounds__(256) void tree_climb_kernel(nbh_tree_view t)
{
    const int n = t.n_leaves;
    const int leaf = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (leaf >= n)
        return;
    int p = t.parent[leaf];
    while (p != -1) {
        const int pi = p - n;
        const int old = fetch_add_acq_rel(&t.visit_count[pi], 1);
        if (old == 0)
            return;
        const int l = t.left[pi];
        const int r = t.right[pi];
        const double ml = t.mass[l];
        const double mr = t.mass[r];
        const double tm = ml + mr;
        t.mass[p] = tm;
        if (tm > 0.0) {
            const double inv = 1.0 / tm;
            t.com_x[p] = fma(ml, t.com_x[l], mr * t.com_x[r]) * inv;
            t.com_y[p] = fma(ml, t.com_y[l], mr * t.com_y[r]) * inv;
            t.com_z[p] = fma(ml, t.com_z[l], mr * t.com_z[r]) * inv;
        } else {
            t.com_x[p] = 0.0;
            t.com_y[p] = 0.0;
            t.com_z[p] = 0.0;
        }
        p = t.parent[p];
    }
}

} // namespace

template <typename K>
void launch_karras(hipStream_t s, const K* keys, const int* d_nleaves,
                   int max_leaves, int* left, int* right, int* parent,
                   int* range_first, int* range_last, int* prefix_len,
                   int* parent_prefix_len)
{
    const int threads = max_leaves > 1 ? max_leaves - 1 : 1;
    karras_kernel<K><<<blocks_for(threads, 256), 256, 0, s>>>(
        keys, d_nleaves, max_leaves, left, right, parent, range_first,
        range_last, prefix_len, parent_prefix_len);
    NBH_HIP(hipGetLastError());
}

template void launch_karras<uint32_t>(hipStream_t, const uint32_t*, const int*, int, int*,
                                      int*, int*, int*, int*, int*, int*);
template void launch_karras<uint64_t>(hipStream_t, const uint64_t*, const int*, int, int*,
                                      int*, int*, int*, int*, int*, int*);

void launch_tree_com(hipStream_t s, const nbh_tree_view& t, const int* first,
                     const int* count, const uint32_t* idx, const double* m,
                     const double* x, const double* y, const double* z)
{
    tree_leaves_kernel<<<blocks_for(t.n_leaves, 256), 256, 0, s>>>(
        t, first, count, idx, m, x, y, z);
    NBH_HIP(hipGetLastError());
    if (t.n_leaves == 1)
        return;
    launch_fill_i32(s, t.visit_count, 0, t.n_leaves - 1);
    tree_climb_kernel<<<blocks_for(t.n_leaves, 256), 256, 0, s>>>(t);
    NBH_HIP(hipGetLastError());
}

} // namespace nbh
