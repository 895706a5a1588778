// The density-refined octree builder with a pluggable sampler for the per-level density
// estimates: the host builder samples on the CPU, skg_octree_build_gpu on the engine's GPU.
#pragma once

#include <cstdint>
#include <functional>

#include "../../include/skirt_mi355x.h"

namespace skg
{
    // sampler(boxes (6 per node), stream ids, n, rho): the mean normalised density of each box
    // over the builder's num_samples random points, from the Philox stream (seed, id, tree tag)
    using OctreeSampler = std::function<int(const double*, const uint64_t*, int64_t, double*)>;

    int octree_build_with(const double* box, const skg_disk_model* model, int32_t min_level, int32_t max_level,
                          double max_dust_fraction, int32_t num_samples, uint64_t seed, int32_t threads,
                          skg_octree** out, const OctreeSampler& sampler);

    constexpr uint32_t kTagTree = 0x54524545u;  // stream tag of the tree refinement samples
}
