// hilbert.cpp -- state tables of the 3-D Hilbert curve over the octree.
//
// The mixed traversal groups its targets into packets along a Hilbert curve
// through the particles' 21-bit grid positions: such packets are more
// compact in space than Morton-consecutive ones, so the union of a packet's
// interaction lists (its wave visits) is smaller, while every target's own
// interaction list stays the same.  The octree's cells are contiguous along
// the curve, so the order follows from a walk of the octree that visits the
// children of a cell in curve order; that order depends only on the cell's
// orientation state.
//
// The curve is the one of J. Skilling, "Programming the Hilbert curve",
// AIP Conf. Proc. 707 (2004): the axes-to-transpose transform followed by
// reading the transposed bits from the most significant level down, x, y,
// z.  The tables are derived from that transform once, at first use: a
// cell's state is identified by the curve ranks of its eight child octants,
// the states are found breadth-first from the root, and every derived rank
// is checked against the transform itself along 256 random paths.
#include <cstdint>
#include <map>
#include <random>
#include <stdexcept>
#include <vector>

#include "hilbert.hpp"

namespace nbh {

namespace {

constexpr int kBits = 21;

// Skilling's axes-to-transpose transform on `kBits`-bit coordinates, then
// the index bits read level by level, x then y then z.
uint64_t hilbert_index(uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t X[3] = {x, y, z};
    const uint32_t M = 1u << (kBits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; ++i)
        X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q)
            t ^= Q - 1;
    for (int i = 0; i < 3; ++i)
        X[i] ^= t;
    uint64_t key = 0;
    for (int b = kBits - 1; b >= 0; --b)
        key = (key << 3) | (uint64_t((X[0] >> b) & 1u) << 2) |
              (uint64_t((X[1] >> b) & 1u) << 1) | uint64_t((X[2] >> b) & 1u);
    return key;
}

// Curve ranks of the eight child octants (x = bit 2, y = bit 1, z = bit 0)
// of the cell reached from the root by `path`.
void child_ranks(const std::vector<int>& path, uint8_t out[8])
{
    const int L = static_cast<int>(path.size());
    uint32_t base[3] = {0, 0, 0};
    for (int l = 0; l < L; ++l) {
        const int s = kBits - 1 - l;
        base[0] |= uint32_t((path[l] >> 2) & 1) << s;
        base[1] |= uint32_t((path[l] >> 1) & 1) << s;
        base[2] |= uint32_t(path[l] & 1) << s;
    }
    const int s = kBits - 1 - L;
    for (int oc = 0; oc < 8; ++oc) {
        const uint64_t key = hilbert_index(base[0] | (uint32_t((oc >> 2) & 1) << s),
                                           base[1] | (uint32_t((oc >> 1) & 1) << s),
                                           base[2] | (uint32_t(oc & 1) << s));
        out[oc] = static_cast<uint8_t>((key >> (3 * (kBits - 1 - L))) & 7u);
    }
}

HilbertTables derive()
{
    HilbertTables tab{};
    std::map<std::vector<uint8_t>, int> ids;
    std::vector<std::vector<int>> rep; // a path to a cell of each state
    auto id_of = [&](const std::vector<int>& path) {
        uint8_t r[8];
        child_ranks(path, r);
        const std::vector<uint8_t> sig(r, r + 8);
        auto it = ids.find(sig);
        if (it != ids.end())
            return it->second;
        const int id = static_cast<int>(rep.size());
        if (id >= kHilbertStates)
            throw std::logic_error("Hilbert curve: more orientation states than expected");
        ids.emplace(sig, id);
        rep.push_back(path);
        for (int oc = 0; oc < 8; ++oc)
            tab.rank[id][oc] = r[oc];
        return id;
    };
    if (id_of({}) != 0)
        throw std::logic_error("Hilbert curve: root state");
    for (size_t k = 0; k < rep.size(); ++k) {
        for (int oc = 0; oc < 8; ++oc) {
            std::vector<int> child = rep[k];
            child.push_back(oc);
            tab.next[k][oc] = static_cast<uint8_t>(id_of(child));
        }
    }
    // The state machine must reproduce the transform's digit at every level
    // of random paths.
    std::mt19937 gen(12345);
    for (int trial = 0; trial < 256; ++trial) {
        std::vector<int> path;
        int st = 0;
        for (int l = 0; l < kBits - 1; ++l) {
            uint8_t r[8];
            child_ranks(path, r);
            const int oc = static_cast<int>(gen() & 7u);
            for (int c = 0; c < 8; ++c)
                if (r[c] != tab.rank[st][c])
                    throw std::logic_error("Hilbert curve: inconsistent state tables");
            st = tab.next[st][oc];
            path.push_back(oc);
        }
    }
    return tab;
}

} // namespace

const HilbertTables& hilbert_tables()
{
    static const HilbertTables tab = derive();
    return tab;
}

} // namespace nbh
