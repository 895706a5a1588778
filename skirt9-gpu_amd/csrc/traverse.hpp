// Device-side path segment generators for the MI355X engine.
//
// The generators reproduce SKIRT's PathSegmentGenerator state machine bit for bit in FP64
// (this file is compiled with -ffp-contract=off so no multiply-add is fused):
//   moveInside          PathSegmentGenerator.cpp:11-112
//   Cartesian next()    CartesianSpatialGrid.cpp:77-153
//   octree next()       TreeSpatialGrid.cpp:143-228, TreeNode.cpp:66-76/103-112, OctTreeNode.cpp:37-42
// One lane owns one path; the generator state lives in registers.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

namespace skg
{
    enum GridType : int
    {
        kCartesian = 0,
        kOctree = 1
    };

    // One 128-byte record per octree leaf (indexed by cell): everything a cell crossing
    // reads -- the leaf's FP64 bounds, its number density, the first (largest-overlap)
    // neighbour on each wall and the start of each wall's full neighbour list -- sits in one
    // cache line, so the common crossing (the first neighbour contains the new position)
    // touches a single line.
    struct alignas(128) OctLeaf
    {
        double lo[3], hi[3];  // bounds
        double dens;          // number density of the cell
        int32_t node;         // node id in the full tree
        int32_t off_end;      // end of this leaf's neighbour lists in nbc
        int32_t first[6];     // first neighbour cell per wall (BACK..TOP), -1 if none
        int32_t off[6];       // start of each wall's neighbour list in nbc
        int32_t pad[4];
    };
    static_assert(sizeof(OctLeaf) == 128, "leaf record must be one 128-byte line");

    // Read-only grid tables resident in HBM.  Octree: node-indexed bounds, first child
    // (children are 8 consecutive ids) and cell index for the top-down search, plus the
    // cell-indexed leaf records and their neighbour lists (cell ids, sorted by decreasing
    // overlap as TreeNode::sortNeighbors).
    struct GridView
    {
        double box[6];  // xmin, ymin, zmin, xmax, ymax, zmax of the domain
        double eps;     // 1e-12 * domain diagonal
        const double* xb;
        const double* yb;
        const double* zb;
        int nx, ny, nz;
        const double* dens;  // Cartesian: number density per cell
        const double* bounds;
        const int32_t* child_first;
        const int32_t* cell_of;
        const OctLeaf* leaf;
        const int32_t* nbc;
    };

    enum GenState : int
    {
        kUnknown = 0,
        kInside = 1,
        kOutside = 2
    };

    __device__ __forceinline__ int locate_clip(const double* xv, int n, double x)
    {
        // NR::locateClip over n borders (NR.hpp:130-144, :177-183)
        if (x < xv[0]) return 0;
        int jl = -1, ju = n - 1;
        while (ju - jl > 1)
        {
            int jm = (ju + jl) >> 1;
            if (x < xv[jm])
                ju = jm;
            else
                jl = jm;
        }
        return jl;
    }

    __device__ __forceinline__ bool node_contains(const double* b, double x, double y, double z)
    {
        return x >= b[0] && x <= b[3] && y >= b[1] && y <= b[4] && z >= b[2] && z <= b[5];
    }

    // top-down search for the leaf containing r; -1 outside the root (closed box)
    __device__ __forceinline__ int leaf_child(const GridView& g, double x, double y, double z)
    {
        if (!node_contains(g.bounds, x, y, z)) return -1;
        int node = 0;
        int c0;
        while ((c0 = g.child_first[node]) >= 0)
        {
            const double* b = g.bounds + 6 * static_cast<int64_t>(c0);
            int l = (x < b[3] ? 0 : 1) + (y < b[4] ? 0 : 2) + (z < b[5] ? 0 : 4);
            node = c0 + l;
        }
        return node;
    }

    template<int GT> struct Gen
    {
        double rx, ry, rz, kx, ky, kz;
        int state;
        int i, j, k;  // Cartesian cell indices; octree: i = current cell
        double b0, b1, b2, b3, b4, b5, bn;  // octree: bounds and density of the current cell

        __device__ __forceinline__ void load_leaf(const GridView& g, int c)
        {
            const OctLeaf& L = g.leaf[c];
            b0 = L.lo[0];
            b1 = L.lo[1];
            b2 = L.lo[2];
            b3 = L.hi[0];
            b4 = L.hi[1];
            b5 = L.hi[2];
            bn = L.dens;
        }

        __device__ __forceinline__ int leaf_cell(const GridView& g, double x, double y, double z) const
        {
            int node = leaf_child(g, x, y, z);
            return node >= 0 ? g.cell_of[node] : -1;
        }

        __device__ __forceinline__ void start(double r0, double r1, double r2, double k0, double k1, double k2)
        {
            rx = r0;
            ry = r1;
            rz = r2;
            kx = k0;
            ky = k1;
            kz = k2;
            state = kUnknown;
        }

        __device__ __forceinline__ bool move_inside(const GridView& g, double& cumds)
        {
            const double eps = g.eps;
            state = kOutside;
            cumds = 0.;
            if (rx <= g.box[0])
            {
                if (kx <= 0.0) return false;
                double d = (g.box[0] - rx) / kx;
                rx = g.box[0] + eps;
                ry += ky * d;
                rz += kz * d;
                cumds += d;
            }
            else if (rx >= g.box[3])
            {
                if (kx >= 0.0) return false;
                double d = (g.box[3] - rx) / kx;
                rx = g.box[3] - eps;
                ry += ky * d;
                rz += kz * d;
                cumds += d;
            }
            if (ry <= g.box[1])
            {
                if (ky <= 0.0) return false;
                double d = (g.box[1] - ry) / ky;
                rx += kx * d;
                ry = g.box[1] + eps;
                rz += kz * d;
                cumds += d;
            }
            else if (ry >= g.box[4])
            {
                if (ky >= 0.0) return false;
                double d = (g.box[4] - ry) / ky;
                rx += kx * d;
                ry = g.box[4] - eps;
                rz += kz * d;
                cumds += d;
            }
            if (rz <= g.box[2])
            {
                if (kz <= 0.0) return false;
                double d = (g.box[2] - rz) / kz;
                rx += kx * d;
                ry += ky * d;
                rz = g.box[2] + eps;
                cumds += d;
            }
            else if (rz >= g.box[5])
            {
                if (kz >= 0.0) return false;
                double d = (g.box[5] - rz) / kz;
                rx += kx * d;
                ry += ky * d;
                rz = g.box[5] - eps;
                cumds += d;
            }
            if (!(rx >= g.box[0] && rx <= g.box[3] && ry >= g.box[1] && ry <= g.box[4] && rz >= g.box[2]
                  && rz <= g.box[5]))
                return false;
            state = kInside;
            return true;
        }

        // Returns false when the path has no more segments; otherwise sets the cell m, the
        // segment length ds and the cell's number density n (0 for m = -1).
        __device__ __forceinline__ bool next(const GridView& g, int& m, double& ds, double& n)
        {
            if (state == kOutside) return false;
            if (state == kUnknown)
            {
                double cumds;
                if (!move_inside(g, cumds)) return false;
                if (GT == kCartesian)
                {
                    i = locate_clip(g.xb, g.nx + 1, rx);
                    j = locate_clip(g.yb, g.ny + 1, ry);
                    k = locate_clip(g.zb, g.nz + 1, rz);
                }
                else
                {
                    i = leaf_cell(g, rx, ry, rz);
                    if (i < 0)
                    {
                        state = kOutside;
                        return false;
                    }
                    load_leaf(g, i);
                }
                if (cumds > 0.)
                {
                    m = -1;
                    ds = cumds;
                    n = 0.;
                    return true;
                }
            }
            if (GT == kCartesian)
            {
                int mm = k + g.nz * j + g.nz * g.ny * i;
                double xE = (kx < 0.0) ? g.xb[i] : g.xb[i + 1];
                double yE = (ky < 0.0) ? g.yb[j] : g.yb[j + 1];
                double zE = (kz < 0.0) ? g.zb[k] : g.zb[k + 1];
                double dsx = (fabs(kx) > 1e-15) ? (xE - rx) / kx : DBL_MAX;
                double dsy = (fabs(ky) > 1e-15) ? (yE - ry) / ky : DBL_MAX;
                double dsz = (fabs(kz) > 1e-15) ? (zE - rz) / kz : DBL_MAX;
                m = mm;
                n = g.dens ? g.dens[mm] : 0.;
                if (dsx <= dsy && dsx <= dsz)
                {
                    ds = dsx;
                    rx = xE;
                    ry += ky * dsx;
                    rz += kz * dsx;
                    i += (kx < 0.0) ? -1 : 1;
                    if (i >= g.nx || i < 0) state = kOutside;
                }
                else if (dsy < dsx && dsy <= dsz)
                {
                    ds = dsy;
                    ry = yE;
                    rx += kx * dsy;
                    rz += kz * dsy;
                    j += (ky < 0.0) ? -1 : 1;
                    if (j >= g.ny || j < 0) state = kOutside;
                }
                else
                {
                    ds = dsz;
                    rz = zE;
                    rx += kx * dsz;
                    ry += ky * dsz;
                    k += (kz < 0.0) ? -1 : 1;
                    if (k >= g.nz || k < 0) state = kOutside;
                }
                return true;
            }
            else
            {
                const int cur = i;
                double xnext = (kx < 0.0) ? b0 : b3;
                double ynext = (ky < 0.0) ? b1 : b4;
                double znext = (kz < 0.0) ? b2 : b5;
                double dsx = (fabs(kx) > 1e-15) ? (xnext - rx) / kx : DBL_MAX;
                double dsy = (fabs(ky) > 1e-15) ? (ynext - ry) / ky : DBL_MAX;
                double dsz = (fabs(kz) > 1e-15) ? (znext - rz) / kz : DBL_MAX;
                double d;
                int wall;
                if (dsx <= dsy && dsx <= dsz)
                {
                    d = dsx;
                    wall = (kx < 0.0) ? 0 : 1;
                }
                else if (dsy <= dsx && dsy <= dsz)
                {
                    d = dsy;
                    wall = (ky < 0.0) ? 2 : 3;
                }
                else
                {
                    d = dsz;
                    wall = (kz < 0.0) ? 4 : 5;
                }
                double step = d + g.eps;
                rx += kx * step;
                ry += ky * step;
                rz += kz * step;
                m = cur;
                ds = d;
                n = bn;
                // first neighbour on the crossed wall whose closed box contains r
                const OctLeaf& R = g.leaf[cur];
                int nxt = -1;
                int f = R.first[wall];
                if (f >= 0)
                {
                    const OctLeaf& F = g.leaf[f];
                    if (rx >= F.lo[0] && rx <= F.hi[0] && ry >= F.lo[1] && ry <= F.hi[1] && rz >= F.lo[2] && rz <= F.hi[2])
                        nxt = f;
                    else
                    {
                        int e0 = R.off[wall] + 1;
                        int e1 = wall < 5 ? R.off[wall + 1] : R.off_end;
                        for (int q = e0; q < e1; ++q)
                        {
                            int c = g.nbc[q];
                            const OctLeaf& C = g.leaf[c];
                            if (rx >= C.lo[0] && rx <= C.hi[0] && ry >= C.lo[1] && ry <= C.hi[1] && rz >= C.lo[2]
                                && rz <= C.hi[2])
                            {
                                nxt = c;
                                break;
                            }
                        }
                    }
                }
                if (nxt < 0) nxt = leaf_cell(g, rx, ry, rz);
                if (nxt == cur)
                {
                    rx = nextafter(rx, (kx < 0.) ? -DBL_MAX : DBL_MAX);
                    ry = nextafter(ry, (ky < 0.) ? -DBL_MAX : DBL_MAX);
                    rz = nextafter(rz, (kz < 0.) ? -DBL_MAX : DBL_MAX);
                    nxt = leaf_cell(g, rx, ry, rz);
                }
                if (nxt < 0 || nxt == cur)
                    state = kOutside;
                else
                    load_leaf(g, nxt);
                i = nxt;
                return true;
            }
        }
    };
}
