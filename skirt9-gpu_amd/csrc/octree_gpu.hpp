// Density-refined octree construction on the device, level by level, giving the host
// builder's tree (DensityTreePolicy::constructTree, DensityTreePolicy.cpp:180-241, with the
// neighbour maintenance of OctTreeNode::addNeighbors, OctTreeNode.cpp:46-136).
//
// One level: (1) a flag per node of the level (below the minimum level always; below the
// maximum level when the sampled mass fraction exceeds the threshold); (2) a scan of the flags
// gives each flagged node its 8 children at consecutive ids after the level, in node order --
// the ids the host's in-order subdivisions hand out; (3) the neighbour lists of every leaf are
// rebuilt from the previous level's lists.
//
// Step (3) restates the result of the host's sequential relinking as a function of the lists
// before the level.  Subdividing node e replaces e, in the list of each neighbour, by the
// children of e that touch that neighbour (in child order), appended at the list's end; the
// subdivisions of a level happen in node order.  So after a level:
//   * a leaf that was not subdivided keeps its unsubdivided neighbours in their order, followed,
//     for each subdivided neighbour in increasing id, by that neighbour's children that touch
//     it;
//   * a new child keeps, on a wall inside its parent, only its sibling; on a wall on its
//     parent's face, the parent's neighbours it touches by the same rule (the parent's list as
//     it stood when the parent was subdivided, the neighbours subdivided before the parent
//     already replaced, then the later ones replaced in its own list, which is the same order).
// "Touches" is the host's comparison of the neighbour's bounds with the subdivided node's
// split point, closed on both sides (OctTreeNode.cpp:99-135).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "disk_model.hpp"
#include "octree_build.hpp"
#include "philox.hpp"

namespace skg
{
    // the two axes spanning a wall perpendicular to `axis` (0 = x: BACK/FRONT, 1 = y: LEFT/RIGHT,
    // 2 = z: BOTTOM/TOP), in increasing order
    __device__ __forceinline__ void tree_span(int axis, int& u, int& w)
    {
        u = axis == 0 ? 1 : 0;
        w = axis == 2 ? 1 : 2;
    }

    // the relink test of OctTreeNode::addNeighbors: does the box tb reach the quadrant of the
    // child with bits (bu, bw) along u and w of a node split at S (closed comparisons)
    __device__ __forceinline__ bool tree_quad(const double* tb, const double* S, int bu, int bw, int u, int w)
    {
        const bool tu = bu == 0 ? tb[u] <= S[u] : tb[u + 3] >= S[u];
        const bool tw = bw == 0 ? tb[w] <= S[w] : tb[w + 3] >= S[w];
        return tu && tw;
    }

    // (1) subdivision flag of node lbeg + i of a level: cnt[i + 1] = flag, cnt[0] = 0
    __global__ void tree_flags(DiskModel md, const double* __restrict__ B, int64_t lbeg, int64_t n, int level,
                               int minLevel, int maxLevel, double frac, int samples, uint64_t seed, int64_t* cnt)
    {
        const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i == 0) cnt[0] = 0;
        if (i >= n) return;
        int64_t f = 0;
        if (level < minLevel)
            f = 1;
        else if (level < maxLevel)
        {
            // DensityTreePolicy::needsSubdivide, maxDustFraction criterion: the mean normalised
            // density over `samples` uniform points of the box (stream id = node id) x volume
            const double* b = B + 6 * (lbeg + i);
            Philox rng;
            rng.init(seed, static_cast<uint64_t>(lbeg + i), kTagTree);
            double sum = 0.;
            for (int q = 0; q < samples; ++q)
            {
                const double x = b[0] + rng.uniform() * (b[3] - b[0]);
                const double y = b[1] + rng.uniform() * (b[4] - b[1]);
                const double z = b[2] + rng.uniform() * (b[5] - b[2]);
                sum += md.geom(x, y, z);
            }
            const double rho = sum / samples;
            const double vol = (b[3] - b[0]) * (b[4] - b[1]) * (b[5] - b[2]);
            f = rho * vol > frac ? 1 : 0;
        }
        cnt[i + 1] = f;
    }

    // (2) the children of the flagged nodes: off holds the scanned flags (off[i] flagged nodes
    // before node lbeg + i); children boxes as OctTreeNode::createChildren
    __global__ void tree_children(double* __restrict__ B, int32_t* __restrict__ F, int32_t* __restrict__ P, int64_t lbeg,
                                  int64_t lend, int64_t n, const int64_t* __restrict__ off)
    {
        const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i >= n || off[i + 1] == off[i]) return;
        const int64_t p = lbeg + i;
        const int64_t c = lend + 8 * off[i];
        double b[6];
        for (int q = 0; q < 6; ++q) b[q] = B[6 * p + q];
        const double xc = 0.5 * (b[0] + b[3]), yc = 0.5 * (b[1] + b[4]), zc = 0.5 * (b[2] + b[5]);
        F[p] = static_cast<int32_t>(c);
        for (int k = 0; k < 8; ++k)
        {
            double* o = B + 6 * (c + k);
            o[0] = (k & 1) ? xc : b[0];
            o[1] = (k & 2) ? yc : b[1];
            o[2] = (k & 4) ? zc : b[2];
            o[3] = (k & 1) ? b[3] : xc;
            o[4] = (k & 2) ? b[4] : yc;
            o[5] = (k & 4) ? b[5] : zc;
            F[c + k] = -1;
            P[c + k] = static_cast<int32_t>(p);
        }
    }

    // (3) the lists after the level, one thread per (node, wall) of the nodes [0, nn): with
    // kFill false the counts into cnt[t + 1] (cnt[0] = 0), with kFill true the entries at the
    // scanned offsets.  Nodes below lend existed before the level (old lists), the others are
    // the level's new children.
    template<bool kFill>
    __global__ void tree_lists(const double* __restrict__ B, const int32_t* __restrict__ F, const int32_t* __restrict__ P,
                               int64_t lend, int64_t nn, const int64_t* __restrict__ oldOff,
                               const int32_t* __restrict__ oldId, int64_t* cnt, int32_t* __restrict__ newId)
    {
        const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (!kFill && t == 0) cnt[0] = 0;
        if (t >= 6 * nn) return;
        const int64_t n = t / 6;
        const int wall = static_cast<int>(t % 6);
        const int axis = wall >> 1, side = wall & 1;
        int u, w;
        tree_span(axis, u, w);
        const int64_t o = kFill ? cnt[t] : 0;
        int64_t c = 0;
        auto emit = [&](int64_t id) {
            if (kFill) newId[o + c] = static_cast<int32_t>(id);
            ++c;
        };
        // the children of subdivided neighbour e on its face toward wall `wall` of the box tb,
        // in child order, that reach tb
        const double* tb = B + 6 * n;
        auto children = [&](int32_t e) {
            const double* Se = B + 6 * static_cast<int64_t>(F[e]) + 3;  // upper corner of child 0
            for (int k = 0; k < 8; ++k)
                if (((k >> axis) & 1) == 1 - side && tree_quad(tb, Se, (k >> u) & 1, (k >> w) & 1, u, w))
                    emit(static_cast<int64_t>(F[e]) + k);
        };
        if (F[n] < 0)
        {
            int64_t lo, hi;
            const double* Sp = nullptr;
            int bu = 0, bw = 0;
            if (n < lend)
            {
                lo = oldOff[t];
                hi = oldOff[t + 1];
            }
            else
            {
                const int64_t p = P[n];
                const int k = static_cast<int>((n - lend) & 7);
                if (((k >> axis) & 1) != side)
                {
                    emit(n + ((k ^ (1 << axis)) - k));  // the sibling across an inner wall
                    lo = hi = 0;
                }
                else
                {
                    lo = oldOff[6 * p + wall];
                    hi = oldOff[6 * p + wall + 1];
                    Sp = B + 6 * static_cast<int64_t>(F[p]) + 3;
                    bu = (k >> u) & 1;
                    bw = (k >> w) & 1;
                }
            }
            // the parent's neighbours that reach this child's quadrant (all, for an old leaf)
            auto reaches = [&](int32_t e) { return !Sp || tree_quad(B + 6 * static_cast<int64_t>(e), Sp, bu, bw, u, w); };
            for (int64_t q = lo; q < hi; ++q)
            {
                const int32_t e = oldId[q];
                if (F[e] < 0 && reaches(e)) emit(e);
            }
            // subdivided neighbours in increasing id, the order of the level's subdivisions
            int32_t last = -1;
            while (true)
            {
                int32_t next = INT32_MAX;
                for (int64_t q = lo; q < hi; ++q)
                {
                    const int32_t e = oldId[q];
                    if (F[e] >= 0 && e > last && e < next && reaches(e)) next = e;
                }
                if (next == INT32_MAX) break;
                children(next);
                last = next;
            }
        }
        if (!kFill) cnt[t + 1] = c;
    }

    // the leaves' lists of 2..16 entries sorted by decreasing overlap with a stable insertion
    // sort, which is what std::sort does with ranges of at most 16 elements; longer lists are
    // sorted on the host with std::sort
    __global__ void tree_sort_short(const double* __restrict__ B, const int32_t* __restrict__ F, int64_t nn,
                                    const int64_t* __restrict__ off, int32_t* id)
    {
        const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (t >= 6 * nn) return;
        const int64_t n = t / 6;
        const int wall = static_cast<int>(t % 6);
        const int64_t lo = off[t], len = off[t + 1] - lo;
        if (F[n] >= 0 || len < 2 || len > 16) return;
        const double* base = B + 6 * n;
        double key[16];
        int32_t v[16];
        for (int q = 0; q < len; ++q)
        {
            v[q] = id[lo + q];
            key[q] = octree_overlap(base, B + 6 * static_cast<int64_t>(v[q]), wall);
        }
        for (int i = 1; i < len; ++i)
        {
            const double kv = key[i];
            const int32_t vv = v[i];
            int j = i;
            while (j > 0 && kv > key[j - 1])
            {
                key[j] = key[j - 1];
                v[j] = v[j - 1];
                --j;
            }
            key[j] = kv;
            v[j] = vv;
        }
        for (int q = 0; q < len; ++q) id[lo + q] = v[q];
    }
}
