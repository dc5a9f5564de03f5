// hilbert.hpp -- orientation-state tables of the 3-D Hilbert curve
// (hilbert.cpp), shared by the host and the walk-layout kernels.
#pragma once

#include <cstdint>

namespace nbh {

constexpr int kHilbertStates = 24;

// For a cell in orientation state s: rank[s][oc] is the curve rank (0-7) of
// child octant oc (x = bit 2, y = bit 1, z = bit 0) among the cell's
// children, next[s][oc] that child's state.  The root is in state 0.
struct HilbertTables {
    uint8_t rank[kHilbertStates][8];
    uint8_t next[kHilbertStates][8];
};

// Derived once from the curve's index transform and checked against it;
// throws std::logic_error if the derivation fails its check.
const HilbertTables& hilbert_tables();

} // namespace nbh
