// The density-refined octree in its finished table form, shared by the host builder
// (host_model.cpp) and the device builder (engine.hip), and the neighbour-list sort both use.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/skirt_mi355x.h"

// A finished octree: nodes in creation order (children of a node are 8 consecutive ids, the
// order DensityTreePolicy::constructTree creates them), leaves numbered in node order
// (TreeSpatialGrid.cpp:37-49), and every node's six per-wall neighbour lists in one CSR
// (empty for non-leaf nodes), each sorted by decreasing overlap area (TreeNode::sortNeighbors).
struct skg_octree
{
    std::vector<double> bounds;   // 6 per node: xmin, ymin, zmin, xmax, ymax, zmax
    std::vector<int32_t> first;   // first child id, -1 for a leaf
    std::vector<int32_t> cellOf;  // cell index of a leaf, -1 otherwise
    std::vector<int64_t> nbOff;   // 6 * nodes + 1 offsets into nbId, wall-major within a node
    std::vector<int32_t> nbId;    // neighbour node ids
    int64_t ncells = 0;
};

namespace skg
{
    // sampler(boxes (6 per node), stream ids, n, rho): the mean normalised density of each box
    // over the builder's num_samples random points, from the Philox stream (seed, id, tree tag)
    using OctreeSampler = std::function<int(const double*, const uint64_t*, int64_t, double*)>;

    int octree_build_with(const double* box, const skg_disk_model* model, int32_t min_level, int32_t max_level,
                          double max_dust_fraction, int32_t num_samples, uint64_t seed, int32_t threads,
                          skg_octree** out, const OctreeSampler& sampler);

    // Sorts each leaf's per-wall lists that hold at least min_len entries by decreasing overlap
    // area with std::sort, the reference's sort (TreeNode.cpp:200-207); shorter lists are left
    // as they are.  std::sort sorts ranges of at most 16 elements by insertion, which keeps
    // equal elements in their order, so a device sort of the short lists by stable insertion
    // gives the same lists.
    void octree_sort_lists(skg_octree& t, int64_t min_len, int32_t threads);

    // the overlap area of two adjacent nodes' faces on wall w of the first (TreeNode.cpp:145-196)
    __host__ __device__ inline double octree_overlap(const double* base, const double* nb, int wall)
    {
        const int u = wall < 2 ? 1 : 0, w = wall < 4 ? 2 : 1;
        const double du = (base[u + 3] < nb[u + 3] ? base[u + 3] : nb[u + 3]) - (base[u] > nb[u] ? base[u] : nb[u]);
        const double dw = (base[w + 3] < nb[w + 3] ? base[w + 3] : nb[w + 3]) - (base[w] > nb[w] ? base[w] : nb[w]);
        return (du > 0. ? du : 0.) * (dw > 0. ? dw : 0.);
    }

    constexpr uint32_t kTagTree = 0x54524545u;  // stream tag of the tree refinement samples
}
