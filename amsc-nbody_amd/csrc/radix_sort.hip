// radix_sort.hip -- stable LSD radix sort of (key, index) pairs, onesweep
// style, for wave64.
//
// Replaces thrust::stable_sort_by_key (morton.cu:221-224).  One histogram
// pass counts the digits of every pass at once (fused into the Morton encode
// in the pipeline); each sort pass is then a single kernel over tiles:
//   1. a tile takes a ticket (dynamic tile id, so every tile it waits on is
//      already resident: forward progress without co-residency assumptions),
//   2. loads its keys in wave-striped order and ranks them stably with a
//      wave64 match (8 ballots per 8-bit digit) plus mbcnt; a per-wave LDS
//      counter per digit gives each key its rank among equal digits,
//   3. publishes its digit counts and resolves its global digit offsets by a
//      decoupled look-back over the preceding tiles (agent-scope atomics on
//      32-bit status words: 2 flag bits + 30 count bits),
//   4. scatters the tile into LDS in digit order and writes it out, so equal
//      digits leave the tile in contiguous runs.
// Ranking follows (wave, item, lane) order, which is input order; every pass
// is therefore stable and so is the whole sort.
#include "device_utils.hpp"
#include "nbh_internal.hpp"

namespace nbh {

namespace {

constexpr int kSortThreads = 256;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagInclusive = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;
constexpr int kLookback = 4; // predecessor statuses read per round trip

template <typename K>
struct SortTile;
template <>
struct SortTile<uint32_t> {
    static constexpr int items = 16;
};
// 4096-key tiles for both widths: fewer tiles mean fewer tickets and shorter
// look-backs (8 items: +14 % time, 12: +10 %, 20-24: within 1 % at 4M keys).
template <>
struct SortTile<uint64_t> {
    static constexpr int items = 16;
};

template <typename K>
constexpr int tile_size()
{
    return kSortThreads * SortTile<K>::items;
}
// Up to kSmallSortKeys keys the tiles hold 1024 keys (4 per thread), so
// that a pass spreads over more workgroups: few 4096-key tiles leave the
// pass the latency of one tile's look-back chain (SI blob step: 100K 1.170
// -> 1.145 ms, 250K 1.37 -> 1.355 ms; at 1M the small tiles cost 0.11 ms).
constexpr int kSmallSortItems = 4;
constexpr size_t kSmallSortKeys = size_t(1) << 18;
inline size_t sort_tile_keys(size_t n)
{
    return n <= kSmallSortKeys ? static_cast<size_t>(kSortThreads * kSmallSortItems)
                               : static_cast<size_t>(tile_size<uint64_t>());
}
// The look-back status is sized from the uint64 tile count.
static_assert(tile_size<uint32_t>() >= tile_size<uint64_t>(),
              "uint32 tiles must not outnumber uint64 tiles");

template <typename K>
__global__ __launch_bounds__(256) void histogram_kernel(const K* __restrict__ keys,
                                                        size_t n, int passes,
                                                        uint32_t* __restrict__ hist)
{
    __shared__ uint32_t s_hist[8 * 256];
    for (int j = threadIdx.x; j < passes * 256; j += blockDim.x)
        s_hist[j] = 0;
    __syncthreads();
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < n; i += stride) {
        const K k = keys[i];
        for (int p = 0; p < passes; ++p)
            atomicAdd(&s_hist[p * 256 + static_cast<uint32_t>((k >> (8 * p)) & 0xFFu)], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < passes * 256; j += blockDim.x) {
        const uint32_t c = s_hist[j];
        if (c)
            atomicAdd(&hist[j], c);
    }
}

// In-place exclusive scan of each pass's 256 digit counts (one block per
// pass).
__global__ __launch_bounds__(256) void digit_offsets_kernel(uint32_t* hist)
{
    __shared__ unsigned tmp[4];
    uint32_t* h = hist + blockIdx.x * 256;
    const unsigned c = h[threadIdx.x];
    unsigned total;
    const unsigned ex = block256_exclusive_sum(c, tmp, total);
    h[threadIdx.x] = ex;
}

template <typename K, int ITEMS>
__global__ __launch_bounds__(256) void onesweep_kernel(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin,
    K* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, int shift,
    const uint32_t* __restrict__ digit_base, uint32_t* status,
    uint32_t* ticket)
{
    constexpr int TILE = kSortThreads * ITEMS;
    __shared__ K s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_wave[4][256];
    __shared__ uint32_t s_tile_off[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ unsigned s_tmp[4];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (tid == 0)
        s_tile = atomicAdd(ticket, 1u);
#pragma unroll
    for (int w = 0; w < 4; ++w)
        s_wave[w][tid] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * TILE;

    K key[ITEMS];
    uint32_t val[ITEMS];
    uint32_t rank[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const uint32_t pos = base + wave * (64 * ITEMS) + it * 64 + lane;
        const bool valid = pos < n;
        key[it] = valid ? kin[pos] : K(0);
        val[it] = valid ? vin[pos] : 0u;
    }
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const uint32_t pos = base + wave * (64 * ITEMS) + it * 64 + lane;
        const bool valid = pos < n;
        const uint32_t d = static_cast<uint32_t>((key[it] >> shift) & 0xFFu);
        const unsigned long long peers = match_digit8(d, valid);
        const unsigned below = lanes_below(peers);
        uint32_t b = 0;
        if (valid)
            b = s_wave[wave][d];
        // All lanes have read the counter (one LDS read instruction for the
        // wave) before the group's lowest lane advances it.
        if (valid && below == 0)
            s_wave[wave][d] = b + static_cast<uint32_t>(__popcll(peers));
        rank[it] = b + below;
    }
    __syncthreads();

    // Per digit (thread tid == digit): exclusive offsets of the 4 waves and
    // the tile count.
    const int d = tid;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t c = s_wave[w][d];
        s_wave[w][d] = run;
        run += c;
    }
    const uint32_t count = run;
    uint32_t* my_status = status + static_cast<size_t>(tile) * 256 + d;
    if (tile == 0)
        store_agent(my_status, kFlagInclusive | count);
    else
        store_agent(my_status, kFlagAggregate | count);

    unsigned total;
    const uint32_t tile_off = block256_exclusive_sum(count, s_tmp, total);
    s_tile_off[d] = tile_off;

    uint32_t prefix = 0;
    if (tile > 0) {
        // Look back kLookback tiles per round trip: the statuses of tiles
        // j, j-1, ... are read together and consumed in order up to the
        // first inclusive one (done) or the first one not yet published
        // (read again from there).  Tiles below 0 read as inclusive zeros.
        int j = static_cast<int>(tile) - 1;
        bool done = false;
        while (!done) {
            uint32_t sw[kLookback];
#pragma unroll
            for (int q = 0; q < kLookback; ++q)
                sw[q] = (j - q >= 0)
                            ? load_agent(status + static_cast<size_t>(j - q) * 256 + d)
                            : kFlagInclusive;
            int used = 0;
            bool stalled = false;
#pragma unroll
            for (int q = 0; q < kLookback; ++q) {
                if (done || stalled)
                    continue;
                const uint32_t flag = sw[q] & ~kValueMask;
                if (flag == 0u) {
                    stalled = true;
                    continue;
                }
                prefix += sw[q] & kValueMask;
                ++used;
                if (flag == kFlagInclusive)
                    done = true;
            }
            j -= used;
            if (!done && used == 0)
                __builtin_amdgcn_s_sleep(1);
        }
        store_agent(my_status, kFlagInclusive | (prefix + count));
    }
    s_gbase[d] = digit_base[d] + prefix - tile_off;
    __syncthreads();

#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const uint32_t pos = base + wave * (64 * ITEMS) + it * 64 + lane;
        if (pos < n) {
            const uint32_t dd = static_cast<uint32_t>((key[it] >> shift) & 0xFFu);
            const uint32_t loc = s_tile_off[dd] + s_wave[wave][dd] + rank[it];
            s_keys[loc] = key[it];
            s_vals[loc] = val[it];
        }
    }
    __syncthreads();
    const uint32_t tile_n = (n - base < static_cast<uint32_t>(TILE)) ? n - base : TILE;
    for (uint32_t j = tid; j < tile_n; j += kSortThreads) {
        const K k = s_keys[j];
        const uint32_t dd = static_cast<uint32_t>((k >> shift) & 0xFFu);
        const uint32_t p = s_gbase[dd] + j;
        kout[p] = k;
        vout[p] = s_vals[j];
    }
}

} // namespace

// Words of look-back status + tickets needed for `passes` passes over n keys
// (uint32 and uint64 tiles hold the same number of keys).
size_t sort_status_words(size_t n, int passes)
{
    const size_t tk = sort_tile_keys(n);
    const size_t tiles = (n + tk - 1) / tk;
    return static_cast<size_t>(passes) * (tiles * 256 + 64);
}

template <typename K>
void launch_histogram(hipStream_t s, const K* keys, size_t n, int passes,
                      uint32_t* hist)
{
    unsigned blocks = blocks_for(n, 256 * 8);
    if (blocks > 2048)
        blocks = 2048;
    histogram_kernel<K><<<blocks, 256, 0, s>>>(keys, n, passes, hist);
    NBH_HIP(hipGetLastError());
}

// hist: passes*256 digit counts (consumed: turned into exclusive offsets).
// status: sort_status_words(n, passes) zeroed words.
template <typename K>
void launch_radix_passes(hipStream_t s, K* keys, uint32_t* vals, K* keys_alt,
                         uint32_t* vals_alt, size_t n, int passes,
                         uint32_t* hist, uint32_t* status)
{
    if (n > kValueMask)
        fail(NBH_OVERFLOW_ERROR, "radix sort supports at most 2^30-1 keys");
    if (passes % 2 != 0)
        fail(NBH_INVALID_ARGUMENT, "radix sort needs an even pass count");
    digit_offsets_kernel<<<passes, 256, 0, s>>>(hist);
    NBH_HIP(hipGetLastError());
    const size_t tk = sort_tile_keys(n);
    const size_t tiles = (n + tk - 1) / tk;
    const size_t per_pass = tiles * 256 + 64;
    const bool small = tk != static_cast<size_t>(tile_size<uint64_t>());
    K* src_k = keys;
    uint32_t* src_v = vals;
    K* dst_k = keys_alt;
    uint32_t* dst_v = vals_alt;
    for (int p = 0; p < passes; ++p) {
        uint32_t* st = status + static_cast<size_t>(p) * per_pass;
        uint32_t* ticket = st + (per_pass - 64);
        if (small)
            onesweep_kernel<K, kSmallSortItems>
                <<<static_cast<unsigned>(tiles), kSortThreads, 0, s>>>(
                    src_k, src_v, dst_k, dst_v, static_cast<uint32_t>(n), 8 * p,
                    hist + p * 256, st, ticket);
        else
            onesweep_kernel<K, SortTile<K>::items>
                <<<static_cast<unsigned>(tiles), kSortThreads, 0, s>>>(
                    src_k, src_v, dst_k, dst_v, static_cast<uint32_t>(n), 8 * p,
                    hist + p * 256, st, ticket);
        NBH_HIP(hipGetLastError());
        K* tk = src_k;
        src_k = dst_k;
        dst_k = tk;
        uint32_t* tv = src_v;
        src_v = dst_v;
        dst_v = tv;
    }
}

template void launch_histogram<uint32_t>(hipStream_t, const uint32_t*, size_t, int, uint32_t*);
template void launch_histogram<uint64_t>(hipStream_t, const uint64_t*, size_t, int, uint32_t*);
template void launch_radix_passes<uint32_t>(hipStream_t, uint32_t*, uint32_t*, uint32_t*,
                                            uint32_t*, size_t, int, uint32_t*, uint32_t*);
template void launch_radix_passes<uint64_t>(hipStream_t, uint64_t*, uint32_t*, uint64_t*,
                                            uint32_t*, size_t, int, uint32_t*, uint32_t*);

} // namespace nbh
