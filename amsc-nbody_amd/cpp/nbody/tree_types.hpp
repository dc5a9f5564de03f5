// tree_types.hpp -- the binary radix tree (src/tree_types.hpp:5-108): leaves
// 0..nLeaves-1, internal node i stored at nLeaves + i, root nLeaves.
#pragma once

#include "device_buffer.hpp"

struct TreeView {
    int nLeaves = 0;
    int nInternalNodes = 0;
    int* left = nullptr;
    int* right = nullptr;
    int* parent = nullptr;
    int* rangeFirst = nullptr;
    int* rangeLast = nullptr;
    int* prefixLength = nullptr;
    double* mass = nullptr;
    double* comX = nullptr;
    double* comY = nullptr;
    double* comZ = nullptr;
    int* visitCount = nullptr;
};

struct ConstTreeView {
    int nLeaves = 0;
    int nInternalNodes = 0;
    const int* left = nullptr;
    const int* right = nullptr;
    const int* parent = nullptr;
    const int* rangeFirst = nullptr;
    const int* rangeLast = nullptr;
    const int* prefixLength = nullptr;
    const double* mass = nullptr;
    const double* comX = nullptr;
    const double* comY = nullptr;
    const double* comZ = nullptr;
    const int* visitCount = nullptr;
};

struct Tree {
    int nLeaves = 0;
    int nInternalNodes = 0;

    DeviceBuffer<int> left;
    DeviceBuffer<int> right;
    DeviceBuffer<int> parent;
    DeviceBuffer<int> rangeFirst;
    DeviceBuffer<int> rangeLast;
    DeviceBuffer<int> prefixLength;
    DeviceBuffer<double> mass;
    DeviceBuffer<double> comX;
    DeviceBuffer<double> comY;
    DeviceBuffer<double> comZ;
    DeviceBuffer<int> visitCount;

    [[nodiscard]] TreeView view() noexcept
    {
        return {nLeaves, nInternalNodes, left.data(), right.data(), parent.data(),
                rangeFirst.data(), rangeLast.data(), prefixLength.data(), mass.data(),
                comX.data(), comY.data(), comZ.data(), visitCount.data()};
    }

    [[nodiscard]] ConstTreeView view() const noexcept
    {
        return {nLeaves, nInternalNodes, left.data(), right.data(), parent.data(),
                rangeFirst.data(), rangeLast.data(), prefixLength.data(), mass.data(),
                comX.data(), comY.data(), comZ.data(), visitCount.data()};
    }
};

namespace nbody_amd {

// The C ABI's view of a Tree (member names of the C struct).
template <class TreeT>
inline nbh_tree_view tree_abi_view(TreeT& tree)
{
    nbh_tree_view v{};
    v.n_leaves = tree.nLeaves;
    v.n_internal_nodes = tree.nInternalNodes;
    v.left = const_cast<int*>(tree.left.data());
    v.right = const_cast<int*>(tree.right.data());
    v.parent = const_cast<int*>(tree.parent.data());
    v.range_first = const_cast<int*>(tree.rangeFirst.data());
    v.range_last = const_cast<int*>(tree.rangeLast.data());
    v.prefix_length = const_cast<int*>(tree.prefixLength.data());
    v.mass = const_cast<double*>(tree.mass.data());
    v.com_x = const_cast<double*>(tree.comX.data());
    v.com_y = const_cast<double*>(tree.comY.data());
    v.com_z = const_cast<double*>(tree.comZ.data());
    v.visit_count = const_cast<int*>(tree.visitCount.data());
    return v;
}

} // namespace nbody_amd
