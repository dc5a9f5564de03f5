// nbody_oracle.cpp -- CPU oracle for the Barnes-Hut step pipeline.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain FP64 C++, the
// algorithm of the reference's CUDA path (sytrams/AMSC-Nbody,
// gpu/cuda_kernels/*.cu and src/simulation.cpp) so that the HIP library can
// be checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it, and only as the checker.  The product library
// never links or calls it.
//
// Every function cites the reference lines it follows.  Where the reference
// leaves FMA contraction to nvcc (it does not disable --fmad), this file fixes
// the contraction explicitly with std::fma; the HIP library uses the same
// explicit order, so FP64 results of the two agree bit for bit.  The chosen
// contractions are:
//   a*b + c            -> fma(a, b, c)           (accumulations, leapfrog)
//   a*b + c*d          -> fma(a, b, c*d)
//   dx*dx+dy*dy+dz*dz  -> fma(dz, dz, fma(dx, dx, dy*dy))
//
// Key width is a template parameter: 10 bits per axis -> uint32_t keys (the
// reference layout), 21 bits per axis -> uint64_t keys.
//
// Parity pins (tests/test_oracle.py): the reference KATs from
// tests/unit/cuda/*.cu and tests/integration/cuda/*.cu, std::stable_sort and
// std::minmax_element equivalence, brute-force recursive octrees, the
// direct-sum limit at theta = 0, and closed forms for the integrator.
// Frame capture (tests/test_frames.py): hand-computed sample indices, exact
// rational arithmetic for the bounds and for every quantized value, and the
// endpoint values of the reference's frame test.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Status values shared with include/nbody_amd.h.
enum {
    OK = 0,
    INVALID_ARGUMENT = -1,
    LOGIC_ERROR = -2,
    RUNTIME_ERROR = -3,
    DOMAIN_ERROR = -4,
    OVERFLOW_ERROR = -5,
};

thread_local std::string g_message;
thread_local int g_device_code = 0;

struct OracleError {
    int status;
    int code;
    std::string message;
};

[[noreturn]] void fail(int status, const std::string& message, int code = 0)
{
    throw OracleError{status, code, message};
}

template <typename F>
int guarded(F&& f)
{
    g_device_code = 0;
    try {
        f();
        return OK;
    } catch (const OracleError& e) {
        g_message = e.message;
        g_device_code = e.code;
        return e.status;
    } catch (const std::exception& e) {
        g_message = e.what();
        return RUNTIME_ERROR;
    }
}

// Key layout: src/radix_to_octree.hpp:10-19 for 10 bits per axis; the same
// constants generalised for 21 bits per axis.
template <int BITS>
struct Layout;

template <>
struct Layout<10> {
    using Key = std::uint32_t;
    static constexpr int bits = 10;
    static constexpr int spatial_bits = 30;
    static constexpr int storage_bits = 32;
    static constexpr int padding_bits = 2;
    static constexpr Key spatial_mask = 0x3FFFFFFFu;
};

template <>
struct Layout<21> {
    using Key = std::uint64_t;
    static constexpr int bits = 21;
    static constexpr int spatial_bits = 63;
    static constexpr int storage_bits = 64;
    static constexpr int padding_bits = 1;
    static constexpr Key spatial_mask = 0x7FFFFFFFFFFFFFFFull;
};

// ------------------------------------------------------------------ bbox
// globalbounding.cu:26-56 (per-particle box, component-wise min/max,
// non-finite flag) and :94-144 (centre, padded side, side 0 -> 1).
void bbox(const double* x, const double* y, const double* z, std::size_t n,
          double out[4])
{
    if (n == 0)
        fail(INVALID_ARGUMENT,
             "Cannot compute a bounding box for an empty particle set.");
    const double inf = std::numeric_limits<double>::infinity();
    double xmin = inf, xmax = -inf, ymin = inf, ymax = -inf, zmin = inf,
           zmax = -inf;
    bool nonfinite = false;
    for (std::size_t i = 0; i < n; ++i) {
        nonfinite = nonfinite || !(std::isfinite(x[i]) &&
                                   std::isfinite(y[i]) && std::isfinite(z[i]));
        xmin = std::fmin(xmin, x[i]);
        xmax = std::fmax(xmax, x[i]);
        ymin = std::fmin(ymin, y[i]);
        ymax = std::fmax(ymax, y[i]);
        zmin = std::fmin(zmin, z[i]);
        zmax = std::fmax(zmax, z[i]);
    }
    if (nonfinite)
        fail(DOMAIN_ERROR,
             "Particle positions must be finite to compute a bounding box.");
    out[0] = (xmin + xmax) * 0.5;
    out[1] = (ymin + ymax) * 0.5;
    out[2] = (zmin + zmax) * 0.5;
    double side = 1.1 * std::max({xmax - xmin, ymax - ymin, zmax - zmin});
    if (side == 0.0)
        side = 1.0;
    out[3] = side;
}

// ---------------------------------------------------------------- Morton
// morton.cu:119-139: 10-bit spread with the reference's masks.
std::uint32_t expand10(std::uint32_t v)
{
    std::uint32_t x = v & 0x000003FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// The same spread for 21 bits: bit k of v moves to bit 3k.  Written as a
// plain loop so that it is independent of the magic-mask form used on the
// GPU.
std::uint64_t expand21(std::uint32_t v)
{
    std::uint64_t out = 0;
    for (int k = 0; k < 21; ++k)
        out |= static_cast<std::uint64_t>((v >> k) & 1u) << (3 * k);
    return out;
}

template <int BITS>
typename Layout<BITS>::Key interleave(std::uint32_t qx, std::uint32_t qy,
                                      std::uint32_t qz)
{
    using Key = typename Layout<BITS>::Key;
    if constexpr (BITS == 10)
        return static_cast<Key>((expand10(qx) << 2) | (expand10(qy) << 1) |
                                expand10(qz));
    else
        return static_cast<Key>((expand21(qx) << 2) | (expand21(qy) << 1) |
                                expand21(qz));
}

// morton.cu:98-117: normalise (a - (c - s/2)) / s, scale by the grid size,
// clamp to [0, maxQ], truncate.
template <int BITS>
std::uint32_t quantise(double a, double centre, double side)
{
    constexpr std::uint32_t grid = 1u << BITS;
    constexpr std::uint32_t maxq = grid - 1u;
    const double normalised = (a - (centre - side / 2.0)) / side;
    const double scaled = normalised * static_cast<double>(grid);
    const double clamped =
        std::fmin(std::fmax(scaled, 0.0), static_cast<double>(maxq));
    return static_cast<std::uint32_t>(clamped);
}

// morton.cu:160-225 plus the input checks of morton.cu:34-60 and the bbox
// check of morton.cu:169-170.  The stable sort is std::stable_sort on
// (key, index) pairs, as the reference's own host oracle does
// (tests/unit/cuda/morton_test.cu:165-206).
template <int BITS>
void morton(const double* x, const double* y, const double* z, std::size_t n,
            const double box[4], typename Layout<BITS>::Key* keys,
            std::uint32_t* indices)
{
    using Key = typename Layout<BITS>::Key;
    if (n == 0)
        fail(INVALID_ARGUMENT,
             "Cannot compute Morton keys for an empty particle set");
    if (!x || !y || !z)
        fail(INVALID_ARGUMENT, "Morton coordinate pointers must not be null");
    if (n > std::numeric_limits<std::uint32_t>::max())
        fail(OVERFLOW_ERROR,
             "Particle count exceeds the Morton index representation");
    if (!std::isfinite(box[0]) || !std::isfinite(box[1]) ||
        !std::isfinite(box[2]) || !std::isfinite(box[3]) || box[3] <= 0.0)
        fail(INVALID_ARGUMENT, "Morton bounding box is not valid");
    std::vector<std::pair<Key, std::uint32_t>> pairs(n);
#pragma omp parallel for schedule(static)
    for (std::int64_t i = 0; i < static_cast<std::int64_t>(n); ++i) {
        const std::uint32_t qx = quantise<BITS>(x[i], box[0], box[3]);
        const std::uint32_t qy = quantise<BITS>(y[i], box[1], box[3]);
        const std::uint32_t qz = quantise<BITS>(z[i], box[2], box[3]);
        pairs[i] = {interleave<BITS>(qx, qy, qz),
                    static_cast<std::uint32_t>(i)};
    }
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const auto& a, const auto& b) {
                         return a.first < b.first;
                     });
    for (std::size_t i = 0; i < n; ++i) {
        keys[i] = pairs[i].first;
        indices[i] = pairs[i].second;
    }
}

// ------------------------------------------------------------ leaf groups
// morton_leaf_groups.cu:58-103: run-length encode + exclusive scan.
template <typename Key>
int leaf_groups(const Key* sorted, int n, int capacity, Key* unique,
                int* first, int* count)
{
    if (n <= 0)
        fail(INVALID_ARGUMENT, "buildMortonLeafGroups requires nParticles >= 1");
    if (n > capacity)
        fail(INVALID_ARGUMENT, "nParticles exceeds MortonLeafGroups capacity");
    int g = 0;
    for (int i = 0; i < n; ++i) {
        if (i == 0 || sorted[i] != sorted[i - 1]) {
            unique[g] = sorted[i];
            count[g] = 1;
            ++g;
        } else {
            ++count[g - 1];
        }
    }
    int run = 0;
    for (int k = 0; k < g; ++k) {
        first[k] = run;
        run += count[k];
    }
    return g;
}

// ------------------------------------------------------------ Karras tree
// tree_builder.cu:9-113.  delta uses the key storage width W: clz_W(a ^ b),
// ties W + clz32(i ^ j), i == j -> 2W, out of range -1.
int clz32(std::uint32_t v) { return v == 0 ? 32 : __builtin_clz(v); }
int clzk(std::uint32_t v) { return clz32(v); }
int clzk(std::uint64_t v) { return v == 0 ? 64 : __builtin_clzll(v); }

template <typename Key>
int lcp(const Key* keys, int i, int j, int n)
{
    constexpr int W = static_cast<int>(sizeof(Key) * 8);
    if (j < 0 || j >= n)
        return -1;
    if (i == j)
        return 2 * W;
    const Key a = keys[i];
    const Key b = keys[j];
    if (a != b)
        return clzk(static_cast<Key>(a ^ b));
    return W + clz32(static_cast<std::uint32_t>(i ^ j));
}

template <typename Key>
void determine_range(const Key* keys, int i, int n, int& first, int& last)
{
    const int dr = lcp(keys, i, i + 1, n);
    const int dl = lcp(keys, i, i - 1, n);
    const int d = (dr > dl) ? +1 : -1;
    const int dmin = lcp(keys, i, i - d, n);
    int length = 2;
    while (lcp(keys, i, i + length * d, n) > dmin)
        length *= 2;
    int l = 0;
    for (int t = length / 2; t >= 1; t /= 2)
        if (lcp(keys, i, i + (l + t) * d, n) > dmin)
            l += t;
    const int j = i + l * d;
    first = std::min(i, j);
    last = std::max(i, j);
}

template <typename Key>
int find_split(const Key* keys, int first, int last, int n)
{
    if (first >= last)
        return first;
    const int node_prefix = lcp(keys, first, last, n);
    int split = first;
    int step = last - first;
    do {
        step = (step + 1) >> 1;
        const int c = split + step;
        if (c < last && lcp(keys, first, c, n) > node_prefix)
            split = c;
    } while (step > 1);
    return split;
}

// tree_builder.cu:303-352 (buildTree with its memset of parent to -1).
template <typename Key>
void tree_build(const Key* keys, int n, int* left, int* right, int* parent,
                int* range_first, int* range_last, int* prefix_length)
{
    if (n <= 0)
        fail(INVALID_ARGUMENT, "buildTree requires N >= 1");
    for (int k = 0; k < 2 * n - 1; ++k)
        parent[k] = -1;
    if (n == 1)
        return;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n - 1; ++i) {
        int first, last;
        determine_range(keys, i, n, first, last);
        range_first[i] = first;
        range_last[i] = last;
        prefix_length[i] = lcp(keys, first, last, n);
        const int split = find_split(keys, first, last, n);
        const int lc = (split == first) ? split : n + split;
        const int rc = (split + 1 == last) ? split + 1 : n + split + 1;
        left[i] = lc;
        right[i] = rc;
        parent[lc] = n + i;
        parent[rc] = n + i;
    }
}

// tree_builder.cu:115-233 (legacy binary-tree centre of mass).  Leaves take
// one particle (first == nullptr) or one Morton group; internal nodes combine
// left and right: (mL*xL + mR*xR) * (1/(mL+mR)), with the inner products
// contracted as fma(mL, xL, mR*xR).
void tree_com(int n, const int* left, const int* right, const int* parent,
              const int* first, const int* count, const std::uint32_t* idx,
              const double* m, const double* x, const double* y,
              const double* z, double* mass, double* cx, double* cy,
              double* cz, int* visit)
{
    for (int leaf = 0; leaf < n; ++leaf) {
        if (!first) {
            const std::uint32_t p = idx[leaf];
            mass[leaf] = m[p];
            cx[leaf] = x[p];
            cy[leaf] = y[p];
            cz[leaf] = z[p];
            continue;
        }
        double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
        for (int k = 0; k < count[leaf]; ++k) {
            const std::uint32_t p = idx[first[leaf] + k];
            tm += m[p];
            wx = std::fma(m[p], x[p], wx);
            wy = std::fma(m[p], y[p], wy);
            wz = std::fma(m[p], z[p], wz);
        }
        mass[leaf] = tm;
        if (tm > 0.0) {
            const double inv = 1.0 / tm;
            cx[leaf] = wx * inv;
            cy[leaf] = wy * inv;
            cz[leaf] = wz * inv;
        } else {
            cx[leaf] = cy[leaf] = cz[leaf] = 0.0;
        }
    }
    if (n == 1)
        return;
    for (int i = 0; i < n - 1; ++i)
        visit[i] = 0;
    // Arrival semantics of computeCentersOfMassKernel: the second arriver of
    // each internal node combines its children; process leaves in index
    // order, which yields the same values as any arrival order.
    for (int leaf = 0; leaf < n; ++leaf) {
        int p = parent[leaf];
        while (p != -1) {
            const int pi = p - n;
            const int old = visit[pi]++;
            if (old == 0)
                break;
            const int l = left[pi];
            const int r = right[pi];
            const double ml = mass[l];
            const double mr = mass[r];
            const double tm = ml + mr;
            mass[p] = tm;
            if (tm > 0.0) {
                const double inv = 1.0 / tm;
                cx[p] = std::fma(ml, cx[l], mr * cx[r]) * inv;
                cy[p] = std::fma(ml, cy[l], mr * cy[r]) * inv;
                cz[p] = std::fma(ml, cz[l], mr * cz[r]) * inv;
            } else {
                cx[p] = cy[p] = cz[p] = 0.0;
            }
            p = parent[p];
        }
    }
}

// --------------------------------------------------- radix -> octree plan
// radix_to_octree.cu:54-145 (per-node plan with error codes 1-4) and
// :174-244 (exclusive scan, node total).
template <int BITS>
void plan(int n_leaves, const int* parent, const int* prefix_length,
          const typename Layout<BITS>::Key* unique, int* level, int* count,
          int* offset, int* n_octree)
{
    using L = Layout<BITS>;
    if (n_leaves <= 0)
        fail(INVALID_ARGUMENT, "The radix tree has no leaves");
    const int R = 2 * n_leaves - 1;
    auto level_of = [&](int node) {
        const int sp = (node < n_leaves)
                           ? L::spatial_bits
                           : prefix_length[node - n_leaves] - L::padding_bits;
        return sp / 3;
    };
    int error = 0;
    auto raise = [&](int code) {
        if (error == 0)
            error = code;
    };
    for (int node = 0; node < R; ++node) {
        auto zero = [&]() {
            level[node] = 0;
            count[node] = 0;
        };
        if (node < n_leaves) {
            if ((unique[node] & ~L::spatial_mask) != 0) {
                raise(1);
                zero();
                continue;
            }
        } else {
            const int sp = prefix_length[node - n_leaves] - L::padding_bits;
            if (sp < 0 || sp >= L::spatial_bits) {
                raise(4);
                zero();
                continue;
            }
        }
        const int p = parent[node];
        int parent_level = 0;
        if (p != -1) {
            if (p < n_leaves || p >= R) {
                raise(2);
                zero();
                continue;
            }
            parent_level = level_of(p);
        }
        const int lv = (node < n_leaves) ? parent_level + 1 : level_of(node);
        if (lv < 0 || lv > L::bits || lv < parent_level) {
            raise(3);
            zero();
            continue;
        }
        level[node] = lv;
        count[node] = lv - parent_level;
    }
    if (error != 0) {
        static const char* msg[] = {
            "", "A Morton key uses bits outside the configured spatial layout",
            "A radix node has an invalid parent",
            "A radix child has an octree level smaller than its parent",
            "A radix internal node has an invalid prefix length"};
        fail(RUNTIME_ERROR, msg[error], error);
    }
    long long run = 0;
    for (int node = 0; node < R; ++node) {
        offset[node] = static_cast<int>(run);
        run += count[node];
    }
    if (run < 0 || run >= std::numeric_limits<int>::max())
        fail(OVERFLOW_ERROR,
             "The computed octree node count does not fit in int");
    *n_octree = 1 + static_cast<int>(run);
}

// ------------------------------------------------------ octree topology
// octree_builder.cu:49-224 (materialisation) and :254-362 (initialisation:
// children/parent/level/firstParticle = -1, prefix/particleCount = 0, root
// level 0 prefix 0).
template <int BITS>
typename Layout<BITS>::Key prefix_at_level(typename Layout<BITS>::Key key,
                                           int level)
{
    using L = Layout<BITS>;
    using Key = typename L::Key;
    if (level == 0)
        return 0;
    const int remaining = L::spatial_bits - level * 3;
    const Key mask = L::spatial_mask & (~Key(0) << remaining);
    return key & mask;
}

template <int BITS>
int octant_at_level(typename Layout<BITS>::Key key, int level)
{
    const int shift = Layout<BITS>::spatial_bits - level * 3;
    return static_cast<int>((key >> shift) & 0x7u);
}

template <int BITS>
void octree(int n_nodes, int n_leaves, const int* tparent,
            const int* range_first, const typename Layout<BITS>::Key* unique,
            const int* gfirst, const int* gcount, const int* level_plan,
            const int* count_plan, const int* offset_plan, int* children,
            int* parent, int* level, typename Layout<BITS>::Key* prefix,
            int* first, int* pcount)
{
    using L = Layout<BITS>;
    using Key = typename L::Key;
    const int R = 2 * n_leaves - 1;
    for (int k = 0; k < n_nodes * 8; ++k)
        children[k] = -1;
    for (int k = 0; k < n_nodes; ++k) {
        parent[k] = -1;
        level[k] = -1;
        prefix[k] = 0;
        first[k] = -1;
        pcount[k] = 0;
    }
    level[0] = 0;
    prefix[0] = 0;
    int error = 0;
    auto raise = [&](int code) {
        if (error == 0)
            error = code;
    };
    for (int node = 0; node < R; ++node) {
        const int count = count_plan[node];
        const int lv = level_plan[node];
        if (count < 0 || lv < 0 || lv > L::bits) {
            raise(1);
            continue;
        }
        if (count == 0) {
            if (node < n_leaves)
                raise(5);
            continue;
        }
        const int parent_level = lv - count;
        if (parent_level < 0) {
            raise(1);
            continue;
        }
        Key key;
        if (node < n_leaves) {
            key = unique[node];
        } else {
            const int f = range_first[node - n_leaves];
            if (f < 0 || f >= n_leaves) {
                raise(7);
                continue;
            }
            key = unique[f];
        }
        int parent_node = 0;
        int anc = tparent[node];
        int iterations = 0;
        bool bad = false;
        while (anc != -1) {
            if (anc < 0 || anc >= R) {
                raise(2);
                bad = true;
                break;
            }
            const int ac = count_plan[anc];
            if (ac > 0) {
                if (level_plan[anc] != parent_level) {
                    raise(3);
                    bad = true;
                    break;
                }
                parent_node = 1 + offset_plan[anc] + ac - 1;
                break;
            }
            anc = tparent[anc];
            if (++iterations > R) {
                raise(2);
                bad = true;
                break;
            }
        }
        if (bad)
            continue;
        if (anc == -1 && parent_level != 0) {
            raise(3);
            continue;
        }
        const int first_node = 1 + offset_plan[node];
        for (int local = 0; local < count; ++local) {
            const int nd = first_node + local;
            if (nd <= 0 || nd >= n_nodes) {
                raise(6);
                bad = true;
                break;
            }
            const int l = parent_level + local + 1;
            const int p = (local == 0) ? parent_node : nd - 1;
            level[nd] = l;
            prefix[nd] = prefix_at_level<BITS>(key, l);
            parent[nd] = p;
            const int oct = octant_at_level<BITS>(key, l);
            int& slot = children[p * 8 + oct];
            if (slot == -1)
                slot = nd;
            else if (slot != nd) {
                raise(4);
                bad = true;
                break;
            }
        }
        if (bad)
            continue;
        if (node < n_leaves) {
            const int leaf = first_node + count - 1;
            if (level[leaf] <= 0 || level[leaf] > L::bits) {
                raise(5);
                continue;
            }
            first[leaf] = gfirst[node];
            pcount[leaf] = gcount[node];
        }
    }
    if (error != 0) {
        static const char* msg[] = {
            "",
            "Invalid radix edge metadata",
            "Invalid radix ancestor while constructing octree",
            "Could not find the correct octree parent level",
            "Two octree nodes attempted to occupy the same parent octant",
            "A radix leaf terminates at an invalid octree level",
            "Generated octree node ID is outside the allocated range",
            "Invalid representative Morton key"};
        fail(RUNTIME_ERROR, msg[error], error);
    }
}

// ------------------------------------------------------ octree geometry
// octree_geometry.cu:39-102: validate level/prefix, then walk the prefix
// digits from the root: half /= 2, centre += +-half by octant bits 4/2/1.
template <int BITS>
void geometry(int n_nodes, const int* level, const typename Layout<BITS>::Key* prefix,
              const double box[4], double* cx, double* cy, double* cz,
              double* half)
{
    using L = Layout<BITS>;
    using Key = typename L::Key;
    int error = 0;
    for (int node = 0; node < n_nodes; ++node) {
        const int lv = level[node];
        if (lv < 0 || lv > L::bits) {
            if (error == 0)
                error = 1;
            continue;
        }
        const Key p = prefix[node];
        bool valid = (p & ~L::spatial_mask) == 0;
        if (valid) {
            if (lv == 0)
                valid = (p == 0);
            else {
                const int remaining = L::spatial_bits - lv * 3;
                const Key mask = L::spatial_mask & (~Key(0) << remaining);
                valid = (p & ~mask) == 0;
            }
        }
        if (!valid) {
            if (error == 0)
                error = 2;
            continue;
        }
        double x = box[0], y = box[1], z = box[2], h = box[3] * 0.5;
        for (int depth = 1; depth <= lv; ++depth) {
            const int shift = L::spatial_bits - depth * 3;
            const int oct = static_cast<int>((p >> shift) & 0x7u);
            const double ch = h * 0.5;
            x += (oct & 4) ? ch : -ch;
            y += (oct & 2) ? ch : -ch;
            z += (oct & 1) ? ch : -ch;
            h = ch;
        }
        cx[node] = x;
        cy[node] = y;
        cz[node] = z;
        half[node] = h;
    }
    if (error == 1)
        fail(RUNTIME_ERROR, "Octree node has an invalid level", 1);
    if (error == 2)
        fail(RUNTIME_ERROR, "Octree node has an invalid Morton prefix", 2);
}

// -------------------------------------------------------- mass and COM
// octree_physics.cu:64-245 and :283-337.  Three phases with an error check
// after each (count children, initialise occupied leaves, bottom-up
// propagation with arrival counters, last arriver combines octants 0..7).
void mass_com(int n_nodes, int n_particles, const int* children,
              const int* parent, const int* first, const int* pcount,
              const std::uint32_t* idx, const double* m, const double* x,
              const double* y, const double* z, double* mass, double* cx,
              double* cy, double* cz, int* pending, int* child_count)
{
    for (int k = 0; k < n_nodes; ++k) {
        mass[k] = cx[k] = cy[k] = cz[k] = 0.0;
        pending[k] = 0;
        child_count[k] = 0;
    }
    static const char* msg[] = {
        "",
        "Octree contains an invalid child ID",
        "Occupied octree leaf has invalid particle range",
        "Sorted particle index is outside the particle array",
        "Occupied octree leaf has children",
        "Invalid octree parent during bottom-up propagation",
        "Invalid octree child count",
        "An octree parent was signalled too many times"};
    int error = 0;
    auto raise = [&](int code) {
        if (error == 0)
            error = code;
    };
    // countOctreeChildrenKernel (:64-93)
    for (int node = 0; node < n_nodes; ++node) {
        int c = 0;
        bool bad = false;
        for (int o = 0; o < 8; ++o) {
            const int ch = children[node * 8 + o];
            if (ch == -1)
                continue;
            if (ch <= 0 || ch >= n_nodes) {
                raise(1);
                bad = true;
                break;
            }
            ++c;
        }
        if (bad)
            continue;
        child_count[node] = c;
        if (pcount[node] > 0 && c != 0)
            raise(4);
    }
    if (error)
        fail(RUNTIME_ERROR, msg[error], error);
    // initializeOctreeLeavesKernel (:95-154)
    for (int node = 0; node < n_nodes; ++node) {
        const int count = pcount[node];
        if (count <= 0)
            continue;
        const int f = first[node];
        if (f < 0 || f + count > n_particles) {
            raise(2);
            continue;
        }
        double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
        bool bad = false;
        for (int k = 0; k < count; ++k) {
            const std::uint32_t p = idx[f + k];
            if (p >= static_cast<std::uint32_t>(n_particles)) {
                raise(3);
                bad = true;
                break;
            }
            tm += m[p];
            wx = std::fma(m[p], x[p], wx);
            wy = std::fma(m[p], y[p], wy);
            wz = std::fma(m[p], z[p], wz);
        }
        if (bad)
            continue;
        mass[node] = tm;
        if (tm > 0.0) {
            const double inv = 1.0 / tm;
            cx[node] = wx * inv;
            cy[node] = wy * inv;
            cz[node] = wz * inv;
        }
    }
    if (error)
        fail(RUNTIME_ERROR, msg[error], error);
    // propagateOctreeMassKernel (:156-245); completion is decided by arrival
    // counts, so walking the occupied leaves in index order reproduces it.
    for (int leaf = 0; leaf < n_nodes; ++leaf) {
        if (pcount[leaf] <= 0)
            continue;
        if (child_count[leaf] != 0) {
            raise(4);
            continue;
        }
        int p = parent[leaf];
        while (p != -1) {
            if (p < 0 || p >= n_nodes) {
                raise(5);
                break;
            }
            const int required = child_count[p];
            if (required <= 0 || required > 8) {
                raise(6);
                break;
            }
            const int done = ++pending[p];
            if (done < required)
                break;
            if (done > required) {
                raise(7);
                break;
            }
            double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
            for (int o = 0; o < 8; ++o) {
                const int ch = children[p * 8 + o];
                if (ch == -1)
                    continue;
                const double cm = mass[ch];
                tm += cm;
                wx = std::fma(cm, cx[ch], wx);
                wy = std::fma(cm, cy[ch], wy);
                wz = std::fma(cm, cz[ch], wz);
            }
            mass[p] = tm;
            if (tm > 0.0) {
                const double inv = 1.0 / tm;
                cx[p] = wx * inv;
                cy[p] = wy * inv;
                cz[p] = wz * inv;
            } else {
                cx[p] = cy[p] = cz[p] = 0.0;
            }
            p = parent[p];
        }
    }
    if (error)
        fail(RUNTIME_ERROR, msg[error], error);
}

// ------------------------------------------------------------ Barnes-Hut
struct OctreeHost {
    int n_particles, n_nodes, root;
    const int *children, *parent, *first, *pcount;
    const double *mass, *cx, *cy, *cz, *ctx, *cty, *ctz, *half;
};

// barnes_hut.cu:56-80 (softened monopole; zero mass skipped; d2 <= 0 is
// singular).
inline bool point_mass(double tx, double ty, double tz, double sx, double sy,
                       double sz, double sm, double G, double eps2,
                       double& ax, double& ay, double& az)
{
    if (sm == 0.0)
        return true;
    const double dx = sx - tx;
    const double dy = sy - ty;
    const double dz = sz - tz;
    const double d2 = std::fma(dz, dz, std::fma(dx, dx, dy * dy)) + eps2;
    if (d2 <= 0.0)
        return false;
    const double inv = 1.0 / std::sqrt(d2);
    const double inv3 = inv * inv * inv;
    const double f = G * sm * inv3;
    ax = std::fma(f, dx, ax);
    ay = std::fma(f, dy, ay);
    az = std::fma(f, dz, az);
    return true;
}

// barnes_hut.cu:90-162: first child in octant order 7 -> 0; next node climbs
// parents looking for the next sibling in the same order.
int first_child(const OctreeHost& t, int node, int& err)
{
    for (int o = 7; o >= 0; --o) {
        const int c = t.children[node * 8 + o];
        if (c == -1)
            continue;
        if (c <= 0 || c >= t.n_nodes) {
            err = 1;
            return -2;
        }
        return c;
    }
    err = 5;
    return -2;
}

int next_node(const OctreeHost& t, int node, int& err)
{
    while (node != t.root) {
        const int p = t.parent[node];
        if (p < 0 || p >= t.n_nodes) {
            err = 4;
            return -2;
        }
        bool found = false;
        for (int o = 7; o >= 0; --o) {
            const int c = t.children[p * 8 + o];
            if (c == -1)
                continue;
            if (c <= 0 || c >= t.n_nodes) {
                err = 1;
                return -2;
            }
            if (found)
                return c;
            if (c == node)
                found = true;
        }
        if (!found) {
            err = 4;
            return -2;
        }
        node = p;
    }
    return -1;
}

// barnes_hut.cu:164-322: one target's stackless traversal.  Returns the
// device error code (0 on success).  Counts leaf (particle-particle) and
// accepted-node interactions when the counters are non-null.
int bh_target(const OctreeHost& t, const std::uint32_t* idx, const double* m,
              const double* x, const double* y, const double* z,
              int sorted_target, bool morton, double theta, double eps2,
              double G, double out[3], std::uint64_t* npp, std::uint64_t* npn)
{
    const std::uint32_t ti = idx[sorted_target];
    if (ti >= static_cast<std::uint32_t>(t.n_particles))
        return 3;
    const int target = static_cast<int>(ti);
    const int td = morton ? sorted_target : target;
    const double tx = x[td], ty = y[td], tz = z[td];
    double ax = 0.0, ay = 0.0, az = 0.0;
    int node = t.root;
    int err = 0;
    while (node >= 0) {
        const double h = t.half[node];
        if (h <= 0.0)
            return 8;
        const int lc = t.pcount[node];
        if (lc > 0) {
            for (int o = 0; o < 8; ++o)
                if (t.children[node * 8 + o] != -1)
                    return 7;
            const int f = t.first[node];
            if (f < 0 || lc > t.n_particles || f > t.n_particles - lc)
                return 2;
            for (int k = 0; k < lc; ++k) {
                const int sp = f + k;
                const std::uint32_t src = idx[sp];
                if (src >= static_cast<std::uint32_t>(t.n_particles))
                    return 3;
                if (src == static_cast<std::uint32_t>(target))
                    continue;
                const int sd = morton ? sp : static_cast<int>(src);
                if (!point_mass(tx, ty, tz, x[sd], y[sd], z[sd], m[sd], G,
                                eps2, ax, ay, az))
                    return 6;
                if (npp && m[sd] != 0.0)
                    ++*npp;
            }
            node = next_node(t, node, err);
            if (node == -2)
                return err;
            continue;
        }
        const double nm = t.mass[node];
        if (nm == 0.0) {
            node = next_node(t, node, err);
            if (node == -2)
                return err;
            continue;
        }
        const double dx = t.cx[node] - tx;
        const double dy = t.cy[node] - ty;
        const double dz = t.cz[node] - tz;
        const double d2 = std::fma(dz, dz, std::fma(dx, dx, dy * dy));
        const double side = 2.0 * h;
        const bool contains = std::fabs(tx - t.ctx[node]) <= h &&
                              std::fabs(ty - t.cty[node]) <= h &&
                              std::fabs(tz - t.ctz[node]) <= h;
        bool accept = false;
        if (!contains && d2 > 0.0 && theta > 0.0) {
            const double ox = t.cx[node] - t.ctx[node];
            const double oy = t.cy[node] - t.cty[node];
            const double oz = t.cz[node] - t.ctz[node];
            const double o2 = std::fma(oz, oz, std::fma(ox, ox, oy * oy));
            const double crit = side / theta + std::sqrt(o2);
            accept = std::sqrt(d2) > crit;
        }
        if (accept) {
            if (!point_mass(tx, ty, tz, t.cx[node], t.cy[node], t.cz[node], nm,
                            G, eps2, ax, ay, az))
                return 6;
            if (npn)
                ++*npn;
            node = next_node(t, node, err);
            if (node == -2)
                return err;
            continue;
        }
        node = first_child(t, node, err);
        if (node == -2)
            return err;
    }
    out[0] = ax;
    out[1] = ay;
    out[2] = az;
    return 0;
}

const char* bh_message(int code)
{
    static const char* msg[] = {
        "",
        "Barnes-Hut traversal found an invalid child node",
        "Barnes-Hut traversal found an invalid leaf particle range",
        "Sorted particle index is outside the particle array",
        "Barnes-Hut traversal found inconsistent parent-child topology",
        "Barnes-Hut traversal found an empty terminal node",
        "Singular gravitational interaction; use positive softening",
        "Occupied octree leaf unexpectedly has children",
        "Octree node has invalid geometry"};
    return (code >= 1 && code <= 8) ? msg[code]
                                    : "Unknown Barnes-Hut traversal error";
}

void check_bh_params(double theta, double softening, double G)
{
    if (!std::isfinite(theta) || theta < 0.0)
        fail(INVALID_ARGUMENT,
             "Barnes-Hut theta must be finite and non-negative");
    if (!std::isfinite(softening) || softening < 0.0)
        fail(INVALID_ARGUMENT,
             "Barnes-Hut softening must be finite and non-negative");
    if (!std::isfinite(G) || G <= 0.0)
        fail(INVALID_ARGUMENT, "Gravitational constant must be finite and positive");
}

// barnes_hut.cu:325-391 over a subset [t_begin, t_end) of sorted targets
// (the whole set for parity checks; a bounded sample for CPU timing).
void bh_accel(const OctreeHost& t, const std::uint32_t* idx, const double* m,
              const double* x, const double* y, const double* z, double* ax,
              double* ay, double* az, double theta, double softening, double G,
              bool morton, int t_begin, int t_end, std::uint64_t* npp,
              std::uint64_t* npn)
{
    check_bh_params(theta, softening, G);
    const double eps2 = softening * softening;
    int error = 0;
    std::uint64_t pp = 0, pn = 0;
    const bool counting = npp != nullptr;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : pp, pn)
    for (int s = t_begin; s < t_end; ++s) {
        double out[3] = {0.0, 0.0, 0.0};
        std::uint64_t lpp = 0, lpn = 0;
        const int e = bh_target(t, idx, m, x, y, z, s, morton, theta, eps2, G,
                                out, counting ? &lpp : nullptr,
                                counting ? &lpn : nullptr);
        pp += lpp;
        pn += lpn;
        if (e != 0) {
#pragma omp critical
            if (error == 0)
                error = e;
            continue;
        }
        const std::uint32_t ti = idx[s];
        ax[ti] = out[0];
        ay[ti] = out[1];
        az[ti] = out[2];
    }
    if (error)
        fail(RUNTIME_ERROR, bh_message(error), error);
    if (npp) {
        *npp = pp;
        *npn = pn;
    }
}

// ---------------------------------------------------------- integrator
// integrator.cu:51-106 with the validation of :21-46 and :109-145.
void kick_drift(double* x, double* y, double* z, double* vx, double* vy,
                double* vz, const double* ax, const double* ay,
                const double* az, std::size_t n, double dt)
{
    if (!std::isfinite(dt) || dt <= 0.0)
        fail(INVALID_ARGUMENT, "Leapfrog timestep must be finite and positive");
    const double h = 0.5 * dt;
#pragma omp parallel for schedule(static)
    for (std::int64_t i = 0; i < static_cast<std::int64_t>(n); ++i) {
        vx[i] = std::fma(h, ax[i], vx[i]);
        vy[i] = std::fma(h, ay[i], vy[i]);
        vz[i] = std::fma(h, az[i], vz[i]);
        x[i] = std::fma(dt, vx[i], x[i]);
        y[i] = std::fma(dt, vy[i], y[i]);
        z[i] = std::fma(dt, vz[i], z[i]);
    }
}

void kick(double* vx, double* vy, double* vz, const double* ax,
          const double* ay, const double* az, std::size_t n, double dt)
{
    if (!std::isfinite(dt) || dt <= 0.0)
        fail(INVALID_ARGUMENT, "Leapfrog timestep must be finite and positive");
    const double h = 0.5 * dt;
#pragma omp parallel for schedule(static)
    for (std::int64_t i = 0; i < static_cast<std::int64_t>(n); ++i) {
        vx[i] = std::fma(h, ax[i], vx[i]);
        vy[i] = std::fma(h, ay[i], vy[i]);
        vz[i] = std::fma(h, az[i], vz[i]);
    }
}

// ------------------------------------------------------------ simulation
// src/simulation.cpp:19-225: the staged rebuild (bbox, Morton, groups,
// Karras, plan, octree, geometry, COM), the Morton-order gather
// (simulation.cu:34-58) and the traversal in Morton order; the leapfrog
// step is KD, rebuild, force, K, swap.
template <int BITS>
struct Sim {
    using Key = typename Layout<BITS>::Key;
    std::size_t n = 0;
    double dt = 0, theta = 0, eps = 0, G = 0;
    std::vector<double> m, x, y, z, vx, vy, vz;
    std::vector<double> ax, ay, az, nax, nay, naz;
    double time = 0;
    std::size_t steps = 0;
    bool ready = false;
    // last structure
    std::vector<Key> keys;
    std::vector<std::uint32_t> idx;
    std::vector<Key> uniq;
    std::vector<int> gfirst, gcount;
    int ngroups = 0, nnodes = 0;
    std::vector<int> left, right, tparent, rfirst, rlast, plen;
    std::vector<int> plevel, pcount_, poffset;
    std::vector<int> children, parent, level, first, pcount, pending, ccount;
    std::vector<Key> prefix;
    std::vector<double> mass, cx, cy, cz, ctx, cty, ctz, half;
    std::vector<double> sm, sx, sy, sz;
    double box[4];

    void rebuild()
    {
        bbox(x.data(), y.data(), z.data(), n, box);
        keys.resize(n);
        idx.resize(n);
        morton<BITS>(x.data(), y.data(), z.data(), n, box, keys.data(),
                     idx.data());
        uniq.resize(n);
        gfirst.resize(n);
        gcount.resize(n);
        ngroups = leaf_groups<Key>(keys.data(), static_cast<int>(n),
                                   static_cast<int>(n), uniq.data(),
                                   gfirst.data(), gcount.data());
        const int G1 = std::max(ngroups - 1, 1);
        left.assign(G1, 0);
        right.assign(G1, 0);
        rfirst.assign(G1, 0);
        rlast.assign(G1, 0);
        plen.assign(G1, 0);
        tparent.assign(2 * ngroups - 1, -1);
        tree_build<Key>(uniq.data(), ngroups, left.data(), right.data(),
                        tparent.data(), rfirst.data(), rlast.data(),
                        plen.data());
        const int R = 2 * ngroups - 1;
        plevel.assign(R, 0);
        pcount_.assign(R, 0);
        poffset.assign(R, 0);
        plan<BITS>(ngroups, tparent.data(), plen.data(), uniq.data(),
                   plevel.data(), pcount_.data(), poffset.data(), &nnodes);
        children.assign(8 * nnodes, -1);
        parent.assign(nnodes, -1);
        level.assign(nnodes, -1);
        prefix.assign(nnodes, 0);
        first.assign(nnodes, -1);
        pcount.assign(nnodes, 0);
        octree<BITS>(nnodes, ngroups, tparent.data(), rfirst.data(),
                     uniq.data(), gfirst.data(), gcount.data(), plevel.data(),
                     pcount_.data(), poffset.data(), children.data(),
                     parent.data(), level.data(), prefix.data(), first.data(),
                     pcount.data());
        ctx.assign(nnodes, 0);
        cty.assign(nnodes, 0);
        ctz.assign(nnodes, 0);
        half.assign(nnodes, 0);
        geometry<BITS>(nnodes, level.data(), prefix.data(), box, ctx.data(),
                       cty.data(), ctz.data(), half.data());
        mass.assign(nnodes, 0);
        cx.assign(nnodes, 0);
        cy.assign(nnodes, 0);
        cz.assign(nnodes, 0);
        pending.assign(nnodes, 0);
        ccount.assign(nnodes, 0);
        mass_com(nnodes, static_cast<int>(n), children.data(), parent.data(),
                 first.data(), pcount.data(), idx.data(), m.data(), x.data(),
                 y.data(), z.data(), mass.data(), cx.data(), cy.data(),
                 cz.data(), pending.data(), ccount.data());
    }

    OctreeHost view() const
    {
        return OctreeHost{static_cast<int>(n), nnodes, 0,
                          children.data(), parent.data(), first.data(),
                          pcount.data(), mass.data(), cx.data(), cy.data(),
                          cz.data(), ctx.data(), cty.data(), ctz.data(),
                          half.data()};
    }

    void forces(int t_begin, int t_end)
    {
        sm.resize(n);
        sx.resize(n);
        sy.resize(n);
        sz.resize(n);
        for (std::size_t i = 0; i < n; ++i) {
            const std::uint32_t p = idx[i];
            sm[i] = m[p];
            sx[i] = x[p];
            sy[i] = y[p];
            sz[i] = z[p];
        }
        bh_accel(view(), idx.data(), sm.data(), sx.data(), sy.data(),
                 sz.data(), nax.data(), nay.data(), naz.data(), theta, eps, G,
                 true, t_begin, t_end, nullptr, nullptr);
    }

    void initialize()
    {
        if (ready)
            fail(LOGIC_ERROR, "Simulation can only be initialized once");
        ax.assign(n, 0);
        ay.assign(n, 0);
        az.assign(n, 0);
        nax.assign(n, 0);
        nay.assign(n, 0);
        naz.assign(n, 0);
        rebuild();
        forces(0, static_cast<int>(n));
        std::swap(ax, nax);
        std::swap(ay, nay);
        std::swap(az, naz);
        ready = true;
    }

    void step()
    {
        if (!ready)
            fail(LOGIC_ERROR, "Simulation must be initialized before stepping");
        kick_drift(x.data(), y.data(), z.data(), vx.data(), vy.data(),
                   vz.data(), ax.data(), ay.data(), az.data(), n, dt);
        rebuild();
        forces(0, static_cast<int>(n));
        kick(vx.data(), vy.data(), vz.data(), nax.data(), nay.data(),
             naz.data(), n, dt);
        std::swap(ax, nax);
        std::swap(ay, nay);
        std::swap(az, naz);
        time += dt;
        ++steps;
    }
};

struct SimHandle {
    int bits;
    Sim<10> s10;
    Sim<21> s21;
};

template <typename F>
auto with_sim(SimHandle* h, F&& f)
{
    if (h->bits == 10)
        return f(h->s10);
    return f(h->s21);
}

// ---------------------------------------------------------------- frames
// src/streaming/device_capture.cu:27-137 (gatherPositionBounds,
// quantizePosition, gatherQuantizedPositionSamples) and :139-153
// (gatherParticleTypeSamples), with the host reduction of write() at
// :252-294.  Sample j of s reads particle j * n / s (integer division).
// Bounds are the per-axis min/max over the sampled positions; the reference
// reduces with fmin/fmax, which leave the choice between -0 and +0 open, and
// this restatement fixes it: -0 orders below +0.  quantizePosition's
// `normalized * 65535 + 0.5` is contracted to one fma by nvcc's default
// --fmad, so the fma is written out here.
std::size_t frame_source(std::size_t j, std::size_t n, std::size_t s)
{
    return (j * n) / s;
}

bool frame_less(double a, double b)
{
    return a < b || (a == b && std::signbit(a) && !std::signbit(b));
}

void frame_bounds(const double* x, const double* y, const double* z, std::size_t n,
                  std::size_t s, double* bounds)
{
    const double inf = std::numeric_limits<double>::infinity();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    for (std::size_t j = 0; j < s; ++j) {
        const std::size_t i = frame_source(j, n, s);
        const double p[3] = {x[i], y[i], z[i]};
        if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]))
            fail(RUNTIME_ERROR, "Cannot quantize a graphical frame with non-finite positions");
        for (int c = 0; c < 3; ++c) {
            if (frame_less(p[c], lo[c]))
                lo[c] = p[c];
            if (frame_less(hi[c], p[c]))
                hi[c] = p[c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        bounds[c] = lo[c];
        bounds[3 + c] = hi[c];
    }
}

std::uint16_t frame_quantize(double p, double lo, double hi)
{
    if (hi == lo)
        return 0;
    double t = (p - lo) / (hi - lo);
    t = std::fmin(1.0, std::fmax(0.0, t));
    return static_cast<std::uint16_t>(std::fma(t, 65535.0, 0.5));
}

} // namespace

// =====================================================================
// C ABI of the oracle (host pointers only), loaded by the tests with ctypes.
// =====================================================================
extern "C" {

const char* orc_last_error_message(void) { return g_message.c_str(); }
int orc_last_device_code(void) { return g_device_code; }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_bbox(const double* x, const double* y, const double* z,
             std::size_t n, double* out)
{
    return guarded([&] { bbox(x, y, z, n, out); });
}

int orc_morton_keys(const double* x, const double* y, const double* z,
                    std::size_t n, const double* box, int bits, void* keys,
                    std::uint32_t* idx)
{
    return guarded([&] {
        if (bits == 10)
            morton<10>(x, y, z, n, box, static_cast<std::uint32_t*>(keys),
                       idx);
        else if (bits == 21)
            morton<21>(x, y, z, n, box, static_cast<std::uint64_t*>(keys),
                       idx);
        else
            fail(INVALID_ARGUMENT, "bits_per_axis must be 10 or 21");
    });
}

int orc_leaf_groups(const void* sorted, int key_bytes, int n, int capacity,
                    void* unique, int* first, int* count, int* ngroups)
{
    return guarded([&] {
        if (key_bytes == 4)
            *ngroups = leaf_groups<std::uint32_t>(
                static_cast<const std::uint32_t*>(sorted), n, capacity,
                static_cast<std::uint32_t*>(unique), first, count);
        else
            *ngroups = leaf_groups<std::uint64_t>(
                static_cast<const std::uint64_t*>(sorted), n, capacity,
                static_cast<std::uint64_t*>(unique), first, count);
    });
}

int orc_tree_build(const void* keys, int key_bytes, int n, int* left,
                   int* right, int* parent, int* range_first, int* range_last,
                   int* prefix_length)
{
    return guarded([&] {
        if (key_bytes == 4)
            tree_build<std::uint32_t>(static_cast<const std::uint32_t*>(keys),
                                      n, left, right, parent, range_first,
                                      range_last, prefix_length);
        else
            tree_build<std::uint64_t>(static_cast<const std::uint64_t*>(keys),
                                      n, left, right, parent, range_first,
                                      range_last, prefix_length);
    });
}

int orc_tree_com(int n, const int* left, const int* right, const int* parent,
                 const int* first, const int* count, const std::uint32_t* idx,
                 const double* m, const double* x, const double* y,
                 const double* z, double* mass, double* cx, double* cy,
                 double* cz, int* visit)
{
    return guarded([&] {
        tree_com(n, left, right, parent, first, count, idx, m, x, y, z, mass,
                 cx, cy, cz, visit);
    });
}

int orc_plan(int n_leaves, const int* parent, const int* prefix_length,
             const void* unique, int bits, int* level, int* count,
             int* offset, int* n_octree)
{
    return guarded([&] {
        if (bits == 10)
            plan<10>(n_leaves, parent, prefix_length,
                     static_cast<const std::uint32_t*>(unique), level, count,
                     offset, n_octree);
        else
            plan<21>(n_leaves, parent, prefix_length,
                     static_cast<const std::uint64_t*>(unique), level, count,
                     offset, n_octree);
    });
}

int orc_octree(int n_nodes, int n_leaves, const int* tparent,
               const int* range_first, const void* unique, const int* gfirst,
               const int* gcount, const int* plevel, const int* pcount,
               const int* poffset, int bits, int* children, int* parent,
               int* level, void* prefix, int* first, int* count)
{
    return guarded([&] {
        if (bits == 10)
            octree<10>(n_nodes, n_leaves, tparent, range_first,
                       static_cast<const std::uint32_t*>(unique), gfirst,
                       gcount, plevel, pcount, poffset, children, parent,
                       level, static_cast<std::uint32_t*>(prefix), first,
                       count);
        else
            octree<21>(n_nodes, n_leaves, tparent, range_first,
                       static_cast<const std::uint64_t*>(unique), gfirst,
                       gcount, plevel, pcount, poffset, children, parent,
                       level, static_cast<std::uint64_t*>(prefix), first,
                       count);
    });
}

int orc_geometry(int n_nodes, const int* level, const void* prefix, int bits,
                 const double* box, double* cx, double* cy, double* cz,
                 double* half)
{
    return guarded([&] {
        if (bits == 10)
            geometry<10>(n_nodes, level,
                         static_cast<const std::uint32_t*>(prefix), box, cx,
                         cy, cz, half);
        else
            geometry<21>(n_nodes, level,
                         static_cast<const std::uint64_t*>(prefix), box, cx,
                         cy, cz, half);
    });
}

int orc_mass_com(int n_nodes, int n_particles, const int* children,
                 const int* parent, const int* first, const int* pcount,
                 const std::uint32_t* idx, const double* m, const double* x,
                 const double* y, const double* z, double* mass, double* cx,
                 double* cy, double* cz, int* pending, int* child_count)
{
    return guarded([&] {
        mass_com(n_nodes, n_particles, children, parent, first, pcount, idx,
                 m, x, y, z, mass, cx, cy, cz, pending, child_count);
    });
}

int orc_bh_accel(int n_particles, int n_nodes, int root, const int* children,
                 const int* parent, const int* first, const int* pcount,
                 const double* mass, const double* cx, const double* cy,
                 const double* cz, const double* ctx, const double* cty,
                 const double* ctz, const double* half,
                 const std::uint32_t* idx, const double* m, const double* x,
                 const double* y, const double* z, double* ax, double* ay,
                 double* az, double theta, double softening, double G,
                 int morton_order, int t_begin, int t_end, std::uint64_t* npp,
                 std::uint64_t* npn)
{
    return guarded([&] {
        const OctreeHost t{n_particles, n_nodes, root, children, parent,
                           first, pcount, mass, cx, cy, cz, ctx, cty, ctz,
                           half};
        bh_accel(t, idx, m, x, y, z, ax, ay, az, theta, softening, G,
                 morton_order != 0, t_begin, t_end, npp, npn);
    });
}

int orc_kick_drift(double* x, double* y, double* z, double* vx, double* vy,
                   double* vz, const double* ax, const double* ay,
                   const double* az, std::size_t n, double dt)
{
    return guarded(
        [&] { kick_drift(x, y, z, vx, vy, vz, ax, ay, az, n, dt); });
}

int orc_kick(double* vx, double* vy, double* vz, const double* ax,
             const double* ay, const double* az, std::size_t n, double dt)
{
    return guarded([&] { kick(vx, vy, vz, ax, ay, az, n, dt); });
}

// Softened FP64 direct sum (tests/integration/cuda/barnes_hut_accuracy_test.cu:
// 101-172) on the listed targets, with the same arithmetic as point_mass.
int orc_direct(const double* m, const double* x, const double* y,
               const double* z, std::size_t n, const std::uint32_t* targets,
               std::size_t n_targets, double softening, double G, double* ax,
               double* ay, double* az)
{
    return guarded([&] {
        const double eps2 = softening * softening;
#pragma omp parallel for schedule(dynamic, 16)
        for (std::int64_t k = 0; k < static_cast<std::int64_t>(n_targets);
             ++k) {
            const std::size_t t = targets ? targets[k] : k;
            double a[3] = {0, 0, 0};
            for (std::size_t s = 0; s < n; ++s) {
                if (s == t)
                    continue;
                point_mass(x[t], y[t], z[t], x[s], y[s], z[s], m[s], G, eps2,
                           a[0], a[1], a[2]);
            }
            ax[k] = a[0];
            ay[k] = a[1];
            az[k] = a[2];
        }
    });
}

// Inputs of the reference accuracy test, generated with the same libstdc++
// engines and distributions (barnes_hut_accuracy_test.cu:294-397).
int orc_accuracy_inputs(int clustered, std::size_t n, double* m, double* x,
                        double* y, double* z)
{
    return guarded([&] {
        if (!clustered) {
            std::mt19937_64 g(0xBADC0FFEEULL);
            std::uniform_real_distribution<double> pos(-1.0, 1.0);
            std::uniform_real_distribution<double> mass(0.5, 2.0);
            for (std::size_t i = 0; i < n; ++i) {
                m[i] = mass(g);
                x[i] = pos(g);
                y[i] = pos(g);
                z[i] = pos(g);
            }
        } else {
            std::mt19937_64 g(0xC1A57E5ULL);
            std::normal_distribution<double> off(0.0, 0.08);
            std::uniform_real_distribution<double> mass(0.5, 2.0);
            const double c[4][3] = {{-0.70, -0.65, -0.60},
                                    {0.70, 0.65, 0.60},
                                    {-0.60, 0.70, -0.55},
                                    {0.65, -0.70, 0.70}};
            for (std::size_t i = 0; i < n; ++i) {
                const double* cc = c[i % 4];
                m[i] = mass(g);
                x[i] = cc[0] + off(g);
                y[i] = cc[1] + off(g);
                z[i] = cc[2] + off(g);
            }
        }
    });
}

// Random Karras inputs of tests/unit/cuda/tree_builder_test.cu:604-637 and
// sparse_octree_test.cu:523-548 (std::mt19937(0xC0FFEE), sorted, with
// duplicates planted every `stride` keys starting at `start`).
int orc_random_keys(std::size_t n, std::uint32_t mask, int start, int stride,
                    std::uint32_t* keys)
{
    return guarded([&] {
        std::mt19937 g(0xC0FFEEu);
        for (std::size_t i = 0; i < n; ++i)
            keys[i] = static_cast<std::uint32_t>(g()) & mask;
        std::sort(keys, keys + n);
        for (std::size_t i = start; i < n; i += stride)
            keys[i] = keys[i - 1];
    });
}

// CPU simulation (the reference's Simulation loop on the oracle).
void* orc_sim_create(int bits, std::size_t n, const double* m,
                     const double* x, const double* y, const double* z,
                     const double* vx, const double* vy, const double* vz,
                     double dt, double theta, double softening, double G)
{
    auto* h = new SimHandle();
    h->bits = bits;
    with_sim(h, [&](auto& s) {
        s.n = n;
        s.dt = dt;
        s.theta = theta;
        s.eps = softening;
        s.G = G;
        s.m.assign(m, m + n);
        s.x.assign(x, x + n);
        s.y.assign(y, y + n);
        s.z.assign(z, z + n);
        s.vx.assign(vx, vx + n);
        s.vy.assign(vy, vy + n);
        s.vz.assign(vz, vz + n);
        return 0;
    });
    return h;
}

void orc_sim_destroy(void* h) { delete static_cast<SimHandle*>(h); }

int orc_sim_initialize(void* h)
{
    return guarded([&] {
        with_sim(static_cast<SimHandle*>(h), [](auto& s) {
            s.initialize();
            return 0;
        });
    });
}

int orc_sim_step(void* h, int steps)
{
    return guarded([&] {
        with_sim(static_cast<SimHandle*>(h), [&](auto& s) {
            for (int k = 0; k < steps; ++k)
                s.step();
            return 0;
        });
    });
}

// Rebuild the structures for the current positions and evaluate the
// forces on the sorted targets [t_begin, t_end) only (bounded CPU timing).
int orc_sim_partial_forces(void* h, int t_begin, int t_end)
{
    return guarded([&] {
        with_sim(static_cast<SimHandle*>(h), [&](auto& s) {
            s.ax.assign(s.n, 0);
            s.ay.assign(s.n, 0);
            s.az.assign(s.n, 0);
            s.nax.assign(s.n, 0);
            s.nay.assign(s.n, 0);
            s.naz.assign(s.n, 0);
            s.rebuild();
            s.forces(t_begin, t_end);
            return 0;
        });
    });
}

int orc_sim_rebuild(void* h)
{
    return guarded([&] {
        with_sim(static_cast<SimHandle*>(h), [&](auto& s) {
            s.rebuild();
            return 0;
        });
    });
}

// Copies of the current state / last structure; any pointer may be null.
int orc_sim_state(void* h, double* x, double* y, double* z, double* vx,
                  double* vy, double* vz, double* ax, double* ay, double* az)
{
    return guarded([&] {
        with_sim(static_cast<SimHandle*>(h), [&](auto& s) {
            auto cp = [&](double* dst, const std::vector<double>& src) {
                if (dst)
                    std::memcpy(dst, src.data(), src.size() * sizeof(double));
            };
            cp(x, s.x);
            cp(y, s.y);
            cp(z, s.z);
            cp(vx, s.vx);
            cp(vy, s.vy);
            cp(vz, s.vz);
            cp(ax, s.ax);
            cp(ay, s.ay);
            cp(az, s.az);
            return 0;
        });
    });
}

int orc_sim_sizes(void* h, int* n_groups, int* n_nodes)
{
    return guarded([&] {
        with_sim(static_cast<SimHandle*>(h), [&](auto& s) {
            *n_groups = s.ngroups;
            *n_nodes = s.nnodes;
            return 0;
        });
    });
}

int orc_sim_structure(void* h, void* keys, std::uint32_t* idx, int* children,
                      int* parent, int* first, int* count, double* mass,
                      double* cx, double* cy, double* cz)
{
    return guarded([&] {
        with_sim(static_cast<SimHandle*>(h), [&](auto& s) {
            using Key = typename std::decay_t<decltype(s)>::Key;
            if (keys)
                std::memcpy(keys, s.keys.data(), s.n * sizeof(Key));
            if (idx)
                std::memcpy(idx, s.idx.data(), s.n * sizeof(std::uint32_t));
            const std::size_t M = static_cast<std::size_t>(s.nnodes);
            if (children)
                std::memcpy(children, s.children.data(), 8 * M * sizeof(int));
            if (parent)
                std::memcpy(parent, s.parent.data(), M * sizeof(int));
            if (first)
                std::memcpy(first, s.first.data(), M * sizeof(int));
            if (count)
                std::memcpy(count, s.pcount.data(), M * sizeof(int));
            if (mass)
                std::memcpy(mass, s.mass.data(), M * sizeof(double));
            if (cx)
                std::memcpy(cx, s.cx.data(), M * sizeof(double));
            if (cy)
                std::memcpy(cy, s.cy.data(), M * sizeof(double));
            if (cz)
                std::memcpy(cz, s.cz.data(), M * sizeof(double));
            return 0;
        });
    });
}

// Frame capture (device_capture.cu:27-153,244-313): bounds[6] and the
// 3*s interleaved quantized positions of the sample; types (optional) are
// the sampled type codes, any above 4 (ParticleType::Asteroid) an error.
int orc_frame_capture(const double* x, const double* y, const double* z,
                      const std::uint8_t* type, std::size_t n, std::size_t s,
                      double* bounds, std::uint16_t* q, std::uint8_t* types)
{
    return guarded([&] {
        if (n == 0 || s == 0 || s > n)
            fail(INVALID_ARGUMENT, "frame sample count must be in [1, particle count]");
        if (type && types) {
            for (std::size_t j = 0; j < s; ++j) {
                types[j] = type[frame_source(j, n, s)];
                if (types[j] > 4)
                    fail(RUNTIME_ERROR, "Cannot capture a frame with invalid particle types");
            }
        }
        if (!x)
            return;
        frame_bounds(x, y, z, n, s, bounds);
        for (std::size_t j = 0; j < s; ++j) {
            const std::size_t i = frame_source(j, n, s);
            q[3 * j + 0] = frame_quantize(x[i], bounds[0], bounds[3]);
            q[3 * j + 1] = frame_quantize(y[i], bounds[1], bounds[4]);
            q[3 * j + 2] = frame_quantize(z[i], bounds[2], bounds[5]);
        }
    });
}

} // extern "C"
