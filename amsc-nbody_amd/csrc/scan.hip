// scan.hip -- single-pass decoupled look-back scans.
//
// Leaf groups (replaces cub::DeviceRunLengthEncode::Encode +
// cub::DeviceScan::ExclusiveSum, morton_leaf_groups.cu:58-103): one kernel
// scans the pair (head flag, position of the latest head) with (+, max);
// each element then knows its group id and, at a run's tail, the run length,
// so unique keys, first particles and counts are written in the same pass
// and the group count stays on the device.
//
// Radix -> octree plan (replaces computeRadixToOctreePlanKernel + its
// ExclusiveSum, radix_to_octree.cu:54-145,174-244): each element computes the
// octree level and edge node count of one radix node and the same kernel
// scans the counts; the octree node total stays on the device.
//
// Tiles take tickets in launch order; a tile publishes its aggregate, looks
// back over earlier tiles (64-bit status words: 2 flag bits + 62 payload
// bits, agent-scope atomics) and publishes its inclusive prefix.  The
// look-back is wave-wide: the 64 lanes of the first wave read 64 preceding
// status words at once and reduce them up to the nearest inclusive one, so a
// chain of aggregate-only predecessors costs one round trip per 64 tiles.
#include "device_utils.hpp"
#include "nbh_internal.hpp"

#include <type_traits>

namespace nbh {

namespace {

// Tiles of THREADS x 8 elements.  The inclusive prefix travels back from the
// first tile by about 64 tiles per look-back round trip, so the tile count
// bounds the scan's latency: large scans take 8192-element tiles (1024
// threads; 2048-element tiles are about 2x slower at 8M elements), small
// ones 2048-element tiles, which keep more of the device busy.
constexpr int kScanItems = 8;
constexpr int kScanThreadsLarge = 1024;
constexpr int kScanThreadsSmall = 256;
constexpr int kScanTileSmall = kScanThreadsSmall * kScanItems;
constexpr size_t kScanLargeFrom = 2u << 20; // elements

template <int THREADS>
constexpr int scan_tile()
{
    return THREADS * kScanItems;
}
constexpr unsigned long long kAgg = 1ull << 62;
constexpr unsigned long long kInc = 2ull << 62;
constexpr unsigned long long kPayload = (1ull << 62) - 1ull;

struct SumMax {
    int a; // number of heads so far
    int b; // position of the latest head so far
};

__device__ __forceinline__ SumMax combine(SumMax x, SumMax y)
{
    return {x.a + y.a, x.b > y.b ? x.b : y.b};
}
__device__ __forceinline__ long long combine(long long x, long long y)
{
    return x + y;
}
__device__ __forceinline__ SumMax shfl_up(SumMax v, int o)
{
    return {__shfl_up(v.a, o), __shfl_up(v.b, o)};
}
__device__ __forceinline__ long long shfl_up(long long v, int o)
{
    return __shfl_up(v, o);
}
__device__ __forceinline__ SumMax shfl_xor(SumMax v, int o)
{
    return {__shfl_xor(v.a, o), __shfl_xor(v.b, o)};
}
__device__ __forceinline__ long long shfl_xor(long long v, int o)
{
    return __shfl_xor(v, o);
}
__device__ __forceinline__ unsigned long long pack(SumMax v)
{
    return (static_cast<unsigned long long>(v.a) << 31) |
           static_cast<unsigned long long>(v.b);
}
__device__ __forceinline__ unsigned long long pack(long long v)
{
    return static_cast<unsigned long long>(v) & kPayload;
}
__device__ __forceinline__ void unpack(unsigned long long w, SumMax& v)
{
    v.a = static_cast<int>((w >> 31) & 0x7FFFFFFFull);
    v.b = static_cast<int>(w & 0x7FFFFFFFull);
}
__device__ __forceinline__ void unpack(unsigned long long w, long long& v)
{
    v = static_cast<long long>(w & kPayload);
}

// Scan of one tile whose thread-inclusive values are in v[0..kScanItems).
// On return v[] holds the global inclusive values.  `identity` must be the
// op's identity, packed as payload 0.  The first wave performs the
// look-back.
template <int THREADS, typename T>
__device__ void tile_scan(T (&v)[kScanItems], T identity, uint32_t tile,
                          unsigned long long* status)
{
    constexpr int kScanWaves = THREADS / 64;
    __shared__ T s_wave[kScanWaves];
    __shared__ T s_prefix;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    // thread-local inclusive
#pragma unroll
    for (int k = 1; k < kScanItems; ++k)
        v[k] = combine(v[k - 1], v[k]);
    const T agg = v[kScanItems - 1];
    // wave inclusive over thread aggregates
    T x = agg;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T y = shfl_up(x, o);
        if (lane >= o)
            x = combine(y, x);
    }
    if (lane == 63)
        s_wave[wave] = x;
    __syncthreads();
    T wave_prefix = identity;
    T tile_total = identity;
#pragma unroll
    for (int w = 0; w < kScanWaves; ++w) {
        if (w < wave)
            wave_prefix = combine(wave_prefix, s_wave[w]);
        tile_total = combine(tile_total, s_wave[w]);
    }
    // exclusive prefix of this thread inside the tile
    T excl = shfl_up(x, 1);
    if (lane == 0)
        excl = identity;
    excl = combine(wave_prefix, excl);
    if (wave == 0) {
        T prefix = identity;
        if (tile == 0) {
            if (lane == 0)
                store_agent(status, kInc | pack(tile_total));
        } else {
            if (lane == 0)
                store_agent(status + tile, kAgg | pack(tile_total));
            // Window of tiles j, j-1, ..., j-63 (lane k reads tile j-k).
            // Tiles below 0 read as inclusive identities (payload 0).
            int j = static_cast<int>(tile) - 1;
            while (true) {
                const int idx = j - lane;
                const unsigned long long w = idx >= 0 ? load_agent(status + idx) : kInc;
                const unsigned long long flag = w & ~kPayload;
                const unsigned long long inc = __ballot(flag == kInc);
                const unsigned long long pending = __ballot(flag == 0ull);
                // lanes 0..first inclusive lane (all lanes when none is)
                const int first = inc ? __ffsll(static_cast<long long>(inc)) - 1 : 63;
                const unsigned long long need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
                if ((pending & need) != 0ull) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                T val;
                unpack(w, val);
                if (lane > first)
                    val = identity;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1)
                    val = combine(val, shfl_xor(val, o));
                prefix = combine(val, prefix);
                if (inc != 0ull)
                    break;
                j -= 64;
            }
            if (lane == 0)
                store_agent(status + tile, kInc | pack(combine(prefix, tile_total)));
        }
        if (lane == 0)
            s_prefix = prefix;
    }
    __syncthreads();
    const T before = combine(s_prefix, excl);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        v[k] = combine(before, v[k]);
}

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void leaf_groups_kernel(
    const K* __restrict__ keys, int n, K* __restrict__ unique,
    int* __restrict__ first, int* __restrict__ count, int* __restrict__ ngroups,
    unsigned long long* status, unsigned* ticket)
{
    __shared__ uint32_t s_tile;
    if (threadIdx.x == 0)
        s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const int base = static_cast<int>(tile) * scan_tile<THREADS>() + threadIdx.x * kScanItems;
    K k[kScanItems + 2];
    // k[0] = key[base-1], k[1..ITEMS] = keys, k[ITEMS+1] = key[base+ITEMS]
#pragma unroll
    for (int j = 0; j < kScanItems + 2; ++j) {
        const int i = base - 1 + j;
        k[j] = (i >= 0 && i < n) ? keys[i] : K(0);
    }
    SumMax v[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int i = base + j;
        const bool head = i < n && (i == 0 || k[j + 1] != k[j]);
        v[j] = {head ? 1 : 0, head ? i : 0};
    }
    tile_scan<THREADS, SumMax>(v, SumMax{0, 0}, tile, status);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int i = base + j;
        if (i >= n)
            break;
        const int g = v[j].a - 1;
        if (g < 0 || g >= n)
            break;
        const bool head = (i == 0 || k[j + 1] != k[j]);
        const bool tail = (i == n - 1 || k[j + 1] != k[j + 2]);
        if (head) {
            unique[g] = k[j + 1];
            first[g] = i;
        }
        if (tail)
            count[g] = i + 1 - v[j].b;
        if (i == n - 1)
            *ngroups = v[j].a;
    }
}

// radix_to_octree.cu:54-68 (spatial prefix, level) and :70-145 (per-node
// plan with error codes), fused with the exclusive scan of edge counts.
template <int BITS, int THREADS>
__global__ __launch_bounds__(THREADS) void plan_scan_kernel(
    const int* __restrict__ d_nleaves, const int* __restrict__ tparent,
    const int* __restrict__ prefix_len,
    const typename Layout<BITS>::Key* __restrict__ unique,
    int* __restrict__ level, int* __restrict__ count, int* __restrict__ offset,
    int* __restrict__ d_noctree, int* err, unsigned long long* status,
    unsigned* ticket, int* __restrict__ level_count,
    const int* __restrict__ parent_plen)
{
    using L = Layout<BITS>;
    __shared__ uint32_t s_tile;
    __shared__ int s_lc[kMaxLevels];
    if (threadIdx.x == 0)
        s_tile = atomicAdd(ticket, 1u);
    if (threadIdx.x < kMaxLevels)
        s_lc[threadIdx.x] = 0;
    __syncthreads();
    const int nl = *d_nleaves;
    const int R = 2 * nl - 1;
    const uint32_t tile = s_tile;
    if (static_cast<long long>(tile) * scan_tile<THREADS>() >= R)
        return; // beyond the last tile: nobody looks back at it
    const int base = static_cast<int>(tile) * scan_tile<THREADS>() + threadIdx.x * kScanItems;
    auto level_of = [&](int node) {
        const int sp = (node < nl) ? L::spatial_bits
                                   : prefix_len[node - nl] - L::padding_bits;
        return sp / 3;
    };
    long long v[kScanItems];
    int lv[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int node = base + j;
        v[j] = 0;
        lv[j] = 0;
        if (node >= R)
            continue;
        bool ok = true;
        if (node < nl) {
            if ((unique[node] & ~L::spatial_mask) != 0) {
                raise_error(err, 1);
                ok = false;
            }
        } else {
            const int sp = prefix_len[node - nl] - L::padding_bits;
            if (sp < 0 || sp >= L::spatial_bits) {
                raise_error(err, 4);
                ok = false;
            }
        }
        int parent_level = 0;
        if (ok) {
            const int p = tparent[node];
            if (p != -1) {
                if (p < nl || p >= R) {
                    raise_error(err, 2);
                    ok = false;
                } else {
                    // the parent's prefix length, from the Karras build when
                    // it recorded one per child (a sequential read)
                    parent_level = parent_plen
                                       ? (parent_plen[node] - L::padding_bits) / 3
                                       : level_of(p);
                }
            }
        }
        if (ok) {
            const int l = (node < nl) ? parent_level + 1 : level_of(node);
            if (l < 0 || l > L::bits || l < parent_level) {
                raise_error(err, 3);
            } else {
                lv[j] = l;
                v[j] = l - parent_level;
            }
        }
    }
    long long c[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        c[j] = v[j];
    // Internal nodes per octree level: internal radix node r owns the chain
    // of octree nodes at levels level[r] - count[r] + 1 .. level[r] (a radix
    // leaf owns one octree leaf).  Counted per tile in LDS, one global add
    // per level and tile (the counters start at zero).
    if (level_count) {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const int node = base + j;
            const int cj = (node >= nl && node < R) ? static_cast<int>(c[j]) : 0;
            for (int k = 0; __ballot(k < cj) != 0ull; ++k) {
                const int lk = lv[j] - cj + 1 + k;
                (void)wave_claim(s_lc, lk, k < cj && lk >= 0 && lk < kMaxLevels);
            }
        }
    }
    tile_scan<THREADS, long long>(v, 0ll, tile, status);
    // tile_scan's barriers order the LDS counts before these reads
    if (level_count && threadIdx.x < kMaxLevels && s_lc[threadIdx.x] != 0)
        atomicAdd(&level_count[threadIdx.x * kLevelCountStride], s_lc[threadIdx.x]);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int node = base + j;
        if (node >= R)
            break;
        level[node] = lv[j];
        count[node] = static_cast<int>(c[j]);
        const long long ex = v[j] - c[j];
        offset[node] = static_cast<int>(ex);
        if (node == R - 1) {
            const long long total = v[j];
            *d_noctree = (total >= 0x7FFFFFFFll) ? -1 : 1 + static_cast<int>(total);
        }
    }
}

} // namespace

size_t scan_status_words(size_t n)
{
    return (n + kScanTileSmall - 1) / kScanTileSmall + 64;
}

// status: scan_status_words(n) zeroed 64-bit words; the ticket lives in the
// last of them.
template <typename K>
void launch_leaf_groups(hipStream_t s, const K* keys, int n, K* unique,
                        int* first, int* count, int* d_ngroups,
                        unsigned long long* status)
{
    const auto go = [&](auto threads) {
        constexpr int T = decltype(threads)::value;
        const size_t tiles = (static_cast<size_t>(n) + scan_tile<T>() - 1) / scan_tile<T>();
        unsigned* ticket = reinterpret_cast<unsigned*>(status + tiles + 32);
        leaf_groups_kernel<K, T><<<static_cast<unsigned>(tiles), T, 0, s>>>(
            keys, n, unique, first, count, d_ngroups, status, ticket);
    };
    if (static_cast<size_t>(n) >= kScanLargeFrom)
        go(std::integral_constant<int, kScanThreadsLarge>{});
    else
        go(std::integral_constant<int, kScanThreadsSmall>{});
    NBH_HIP(hipGetLastError());
}

// status: scan_status_words(2*max_leaves) zeroed words.
template <int BITS>
void launch_plan_scan(hipStream_t s, const int* d_nleaves, int max_leaves,
                      const int* tparent, const int* prefix_len,
                      const typename Layout<BITS>::Key* unique, int* level,
                      int* count, int* offset, int* d_noctree, int* err,
                      unsigned long long* status, int* level_count,
                      const int* parent_plen)
{
    const size_t rmax = 2 * static_cast<size_t>(max_leaves);
    const auto go = [&](auto threads) {
        constexpr int T = decltype(threads)::value;
        const size_t tiles = (rmax + scan_tile<T>() - 1) / scan_tile<T>();
        unsigned* ticket = reinterpret_cast<unsigned*>(status + tiles + 32);
        plan_scan_kernel<BITS, T><<<static_cast<unsigned>(tiles), T, 0, s>>>(
            d_nleaves, tparent, prefix_len, unique, level, count, offset, d_noctree, err,
            status, ticket, level_count, parent_plen);
    };
    if (rmax >= kScanLargeFrom)
        go(std::integral_constant<int, kScanThreadsLarge>{});
    else
        go(std::integral_constant<int, kScanThreadsSmall>{});
    NBH_HIP(hipGetLastError());
}

template void launch_leaf_groups<uint32_t>(hipStream_t, const uint32_t*, int, uint32_t*,
                                           int*, int*, int*, unsigned long long*);
template void launch_leaf_groups<uint64_t>(hipStream_t, const uint64_t*, int, uint64_t*,
                                           int*, int*, int*, unsigned long long*);
template void launch_plan_scan<10>(hipStream_t, const int*, int, const int*, const int*,
                                   const uint32_t*, int*, int*, int*, int*, int*,
                                   unsigned long long*, int*, const int*);
template void launch_plan_scan<21>(hipStream_t, const int*, int, const int*, const int*,
                                   const uint64_t*, int*, int*, int*, int*, int*,
                                   unsigned long long*, int*, const int*);

} // namespace nbh
