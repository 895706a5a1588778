// Device-side path segment generators for the MI355X engine.
//
// The generators reproduce SKIRT's PathSegmentGenerator state machine bit for bit in FP64
// (this file is compiled with -ffp-contract=off so no multiply-add is fused):
//   moveInside          PathSegmentGenerator.cpp:11-112
//   Cartesian next()    CartesianSpatialGrid.cpp:77-153
//   octree next()       TreeSpatialGrid.cpp:143-228, TreeNode.cpp:66-76/103-112, OctTreeNode.cpp:37-42
//   Voronoi next()      VoronoiMeshSnapshot.cpp:1094-1213, cellIndex :1037-1071
//   tetra next()        TetraMeshSpatialGrid.cpp:697-820, findEnteringFace :87-108, contains :110-133,
//                       cellIndex :615-624
// One lane owns one path; the generator state lives in registers.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstddef>
#include <cstdint>

namespace skg
{
    enum GridType : int
    {
        kCartesian = 0,
        kOctree = 1,
        kVoronoi = 2,
        kTetra = 3
    };

    // Voronoi cell record (64 bytes, indexed by cell): the site, the cell's number density and
    // its run in the entry table (begin: the run's header slot; count: its neighbour entries).
    // Read at a path's start; a crossing reads only the entry table.
    struct alignas(64) VorCell
    {
        double x, y, z;
        double dens;
        int32_t begin, count;
        int32_t pad[6];
    };
    static_assert(sizeof(VorCell) == 64, "Voronoi cell record must be 64 bytes");

    // The entry table holds one run per cell: a header slot (id kVorHeader, x = the cell's
    // number density, run = the number of neighbour entries) followed by one entry (32 bytes)
    // per neighbour: the neighbour's site, its id and the header slot of its own run.  A cell's
    // exit search streams through its run, and the entry it leaves through gives the next
    // cell's site and run, so a crossing makes one dependent read (the run) instead of two (the
    // cell record, then its neighbours).  Domain walls have ids -1..-6 (xmin, xmax, ymin, ymax,
    // zmin, zmax) and run -1.  The table ends with 3 slots of padding (a run is read from its
    // header in groups of four slots).
    struct alignas(32) VorNb
    {
        double x, y, z;
        int32_t id;
        int32_t run;
    };
    static_assert(sizeof(VorNb) == 32, "Voronoi neighbour entry must be 32 bytes");
    constexpr int32_t kVorHeader = -100;

    // Tetrahedron record (256 bytes, two cache lines, indexed by cell): vertex coordinates,
    // density and face links in the first line, the outward unit face normals in the second.
    struct alignas(128) TetCell
    {
        double v[4][3];     // vertices in the reference's per-tetra order
        double dens;        // number density
        int32_t nb[4];      // neighbour tetra across face f (opposite vertex f), -1 on the hull
        int8_t nface[4];    // the neighbour's index of the shared face
        int32_t pad0;
        double n[4][3];     // outward unit normal of face f
        double pad1[4];
    };
    static_assert(sizeof(TetCell) == 256, "tetra record must be 256 bytes");
    static_assert(offsetof(TetCell, n) == 128, "tetra normals start the second line");

    // Consecutive "no exit found" steps after which a Voronoi or tetra path is ended.  The
    // reference steps by eps without limit; a path needs that many only when it has left the
    // valid geometry (such as a non-finite position).
    constexpr int kMaxNoExit = 4096;

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

    // An octree node in the form the top-down search reads (one 32-byte record per level): the
    // split point (upper corner of the first child, TreeNode::leafChild's test point) and the
    // first child id, or for a leaf (child -1) its cell index.  The records of the top levels
    // are also staged in LDS.
    struct TopNode
    {
        double sx, sy, sz;
        int32_t child;
        int32_t cell;
    };
    static_assert(sizeof(TopNode) == 32, "top node must be 32 bytes");
#ifndef SKG_TOP_NODES
#define SKG_TOP_NODES 1024
#endif
    constexpr int kTopNodes = SKG_TOP_NODES;  // 1024: 32 KB of LDS, levels 0-3 of a full octree and part of level 4

    // Read-only grid tables resident in HBM.  Octree: the node records of the top-down search
    // (TopNode, children are 8 consecutive ids), the cell-indexed leaf records and their
    // neighbour lists (cell ids, sorted by decreasing overlap as TreeNode::sortNeighbors).
    struct GridView
    {
        double box[6];  // xmin, ymin, zmin, xmax, ymax, zmax of the domain
        double eps;     // 1e-12 * domain diagonal
        const double* xb;
        const double* yb;
        const double* zb;
        int nx, ny, nz;
        const double* dens;  // Cartesian: number density per cell
        const OctLeaf* leaf;
        const int32_t* nbc;
        const TopNode* top;  // every node in TopNode form, indexed by node id
        int topN;            // nodes with id < topN are staged in LDS (0: none, non-octree grids)
        // Voronoi: cell records, neighbour entries and the block search lists (CSR over vbn^3
        // blocks of cell ids, in the reference's order)
        const VorCell* vcell;
        const VorNb* vnb;
        int vbn;
        const int32_t* vbl_off;
        const int32_t* vbl_id;
        // tetra: records and a block search grid (tbn^3 blocks, borders per axis, CSR of tetra
        // ids in increasing order)
        const TetCell* tcell;
        int tbn;
        const double* tbx;
        const double* tby;
        const double* tbz;
        const int32_t* tbl_off;
        const int32_t* tbl_id;
    };

    // Box::cellIndices (Box.hpp:162-167) for the Voronoi block grid
    __device__ __forceinline__ int vor_block(const GridView& g, double x, double y, double z)
    {
        const int n = g.vbn;
        int i = static_cast<int>(n * (x - g.box[0]) / (g.box[3] - g.box[0]));
        int j = static_cast<int>(n * (y - g.box[1]) / (g.box[4] - g.box[1]));
        int k = static_cast<int>(n * (z - g.box[2]) / (g.box[5] - g.box[2]));
        i = max(0, min(n - 1, i));
        j = max(0, min(n - 1, j));
        k = max(0, min(n - 1, k));
        return i * n * n + j * n + k;
    }

    // VoronoiMeshSnapshot::cellIndex: nearest site among the block's candidates, squared
    // distance (r - site).norm2(), strict '<' (the first of equal distances wins)
    __device__ __forceinline__ int vor_cell_index(const GridView& g, double x, double y, double z)
    {
        if (!(x >= g.box[0] && x <= g.box[3] && y >= g.box[1] && y <= g.box[4] && z >= g.box[2] && z <= g.box[5]))
            return -1;
        const int b = vor_block(g, x, y, z);
        int m = -1;
        double best = DBL_MAX;
        for (int q = g.vbl_off[b], e = g.vbl_off[b + 1]; q < e; ++q)
        {
            const int id = g.vbl_id[q];
            const VorCell& c = g.vcell[id];
            double dx = x - c.x, dy = y - c.y, dz = z - c.z;
            double d = dx * dx + dy * dy + dz * dz;
            if (d < best)
            {
                best = d;
                m = id;
            }
        }
        return m;
    }

    // True when the point lies in the domain and is nearer to the site of cell h than to each
    // of h's neighbouring sites by a relative margin far above the rounding of the squared
    // distances.  Inside the domain, the cell is the set of points nearer to its site than to
    // the sites across its faces, so h's site is then the unique nearest site: the cell that
    // VoronoiMeshSnapshot::cellIndex (and vor_cell_index) returns.
    __device__ __forceinline__ bool vor_nearest_strict(const GridView& g, int h, double x, double y, double z)
    {
        if (!(x >= g.box[0] && x <= g.box[3] && y >= g.box[1] && y <= g.box[4] && z >= g.box[2] && z <= g.box[5]))
            return false;
        const VorCell& C = g.vcell[h];
        const double hx = x - C.x, hy = y - C.y, hz = z - C.z;
        const double dh = hx * hx + hy * hy + hz * hz;
        for (int q = C.begin + 1, e = C.begin + 1 + C.count; q < e; ++q)
        {
            const VorNb E = g.vnb[q];
            if (E.id < 0) continue;
            const double ex = x - E.x, ey = y - E.y, ez = z - E.z;
            const double dj = ex * ex + ey * ey + ez * ez;
            if (!(dj - dh > 1e-12 * (dj + dh))) return false;
        }
        return true;
    }

    // number of borders <= v, minus one, clipped to a block index
    __device__ __forceinline__ int tet_axis_block(const double* bv, int n, double v)
    {
        int lo = 0, hi = n + 1;  // count of borders <= v lies in [lo, hi]
        while (lo < hi)
        {
            int mid = (lo + hi) >> 1;
            if (bv[mid] <= v)
                lo = mid + 1;
            else
                hi = mid;
        }
        return max(0, min(n - 1, lo - 1));
    }

    // Tetra::contains: bounding box of the vertices, then the face planes (outward normals;
    // faces 0..2 through vertex 3, face 3 through vertex 0)
    __device__ __forceinline__ bool tet_contains(const TetCell& T, double x, double y, double z)
    {
        double lx = DBL_MAX, ly = DBL_MAX, lz = DBL_MAX, hx = -DBL_MAX, hy = -DBL_MAX, hz = -DBL_MAX;
        for (int t = 0; t < 4; ++t)
        {
            lx = fmin(lx, T.v[t][0]);
            ly = fmin(ly, T.v[t][1]);
            lz = fmin(lz, T.v[t][2]);
            hx = fmax(hx, T.v[t][0]);
            hy = fmax(hy, T.v[t][1]);
            hz = fmax(hz, T.v[t][2]);
        }
        if (!(x >= lx && x <= hx && y >= ly && y <= hy && z >= lz && z <= hz)) return false;
        for (int f = 0; f <= 2; ++f)
            if ((T.v[3][0] - x) * T.n[f][0] + (T.v[3][1] - y) * T.n[f][1] + (T.v[3][2] - z) * T.n[f][2] < 0) return false;
        if ((T.v[0][0] - x) * T.n[3][0] + (T.v[0][1] - y) * T.n[3][1] + (T.v[0][2] - z) * T.n[3][2] < 0.) return false;
        return true;
    }

    // True when the point lies inside the tetra by a margin far above the rounding of the
    // face-plane tests: then no other tetra's tet_contains accepts it, so it is the one
    // tetra TetraMeshSpatialGrid::cellIndex returns.  The margin is 1e-13 of the summed
    // distances to the four vertices (a bound on the extent the rounding scales with).
    __device__ __forceinline__ bool tet_inside_strict(const TetCell& T, double x, double y, double z)
    {
        double scale = 0.;
        for (int t = 0; t < 4; ++t) scale += fabs(T.v[t][0] - x) + fabs(T.v[t][1] - y) + fabs(T.v[t][2] - z);
        const double tol = 1e-13 * scale;
        for (int f = 0; f <= 2; ++f)
            if (!((T.v[3][0] - x) * T.n[f][0] + (T.v[3][1] - y) * T.n[f][1] + (T.v[3][2] - z) * T.n[f][2] > tol)) return false;
        return (T.v[0][0] - x) * T.n[3][0] + (T.v[0][1] - y) * T.n[3][1] + (T.v[0][2] - z) * T.n[3][2] > tol;
    }

    // TetraMeshSpatialGrid::cellIndex: the lowest-index tetra containing the point (the
    // reference's search lists hold, in increasing index order, every tetra whose bounding
    // box meets the point's block, so its first hit is the lowest index; the block lists here
    // have the same property)
    __device__ __forceinline__ int tet_cell_index(const GridView& g, double x, double y, double z)
    {
        const int n = g.tbn;
        int i = tet_axis_block(g.tbx, n, x), j = tet_axis_block(g.tby, n, y), k = tet_axis_block(g.tbz, n, z);
        int b = (i * n + j) * n + k;
        for (int q = g.tbl_off[b], e = g.tbl_off[b + 1]; q < e; ++q)
        {
            const int id = g.tbl_id[q];
            if (tet_contains(g.tcell[id], x, y, z)) return id;
        }
        return -1;
    }

    // TetraMeshSpatialGrid.cpp:44-50
    __device__ __forceinline__ void clockwise_vertices(int face, int& c0, int& c1, int& c2)
    {
        c0 = (face + 3) & 3;
        c1 = (face + 2) & 3;
        c2 = (face + 1) & 3;
        if ((face & 1) == 0)
        {
            int t = c0;
            c0 = c2;
            c2 = t;
        }
    }

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

    // The top of the octree staged in LDS for the top-down search every path start makes:
    // for node ids below GridView::topN, the split point (upper corner of the first child) and
    // the first child id.  Every kernel that walks an octree fills it with stage_top() first;
    // with topN = 0 the search reads only global memory.
    __shared__ TopNode skg_top_lds[kTopNodes];

    __device__ __forceinline__ void stage_top(const GridView& g)
    {
        for (int t = threadIdx.x; t < g.topN; t += blockDim.x) skg_top_lds[t] = g.top[t];
        __syncthreads();
    }

    // a ? x : b ? y : z on values (selecting among members by reference would make the
    // compiler select their addresses and keep the generator in scratch memory)
    template<typename T> __device__ __forceinline__ T pick3(bool a, bool b, T x, T y, T z)
    {
        const T yz = b ? y : z;
        return a ? x : yz;
    }

    // What the octree generator's next advance() call loads and does with the record.  Every
    // call makes at most one dependent global load (a leaf record, a search node or a
    // neighbour-list entry), so lanes of a wavefront in different phases share one memory
    // round trip per call instead of the wave waiting for each lane's chain of loads.
    enum OctStep : int
    {
        kOsStart = 0,     // enter the domain, then the start hint's record (or the search)
        kOsCross = 1,     // the crossing of the current leaf, then the wall's first neighbour
        kOsListStart = 2, // the current leaf's record again: the wall's neighbour-list range
        kOsListId = 3,    // a neighbour-list entry (cell id)
        kOsListRec = 4,   // that neighbour's record
        kOsNode = 5,      // a search node below the LDS-staged levels
        kOsLeaf = 6,      // the record of the leaf the search found
        kOsEnd = 7,       // no more segments
        kOsSearch = 8     // (a step's action only) the search from the root after a load of no use:
                          // a start without hint, a wall on the domain boundary
    };

    template<int GT> struct Gen
    {
        double rx, ry, rz, kx, ky, kz;
        int state;
        int i, j, k;  // Cartesian cell indices; octree, Voronoi, tetra: i = current cell
                      // (tetra: j = entering face, -1 if unknown; Voronoi: j = the header slot
                      // of the cell's run in the entry table)
        double sx, sy, sz;  // Voronoi: the current cell's site
        double b0, b1, b2, bn;  // octree: the current cell's wall ahead on each axis, its density
        int f0, f1, f2;         // octree: the first neighbour across each of those walls
        int hint;  // octree, Voronoi, tetra: a cell that probably contains the start point (-1:
                   // none); after the first step, the start cell
        // octree step machine (advance): the next step, the crossed wall, a candidate cell,
        // the neighbour-list cursor and end, the search node, and the flags for the start
        // search and the nextafter retry (TreeSpatialGrid.cpp:210-215)
        int ost, wall, cand, q, qe, node;
        bool locating, na;
        float rkx, rky, rkz;  // FP32 reciprocals of the direction, for the exit-axis estimates

        __device__ __forceinline__ void start(double r0, double r1, double r2, double k0, double k1, double k2,
                                              int start_hint = -1)
        {
            hint = start_hint;
            ost = kOsStart;
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

        // ---------------------------------------------------------------- octree step machine
        // The record fields a step reads, from the 112 bytes it loads (an OctLeaf, a TopNode or
        // neighbour-list entries; the tables are padded so that 112 bytes past any entry stay
        // inside them)
        struct RawRec
        {
            // the 112 bytes as 14 doubles in registers (plain scalars: no member is selected by
            // reference, which would leave the record in scratch memory)
            double q0, q1, q2, q3, q4, q5, q6, q7, q8, q9, q10, q11, q12, q13;
            __device__ __forceinline__ void load(const char* addr)
            {
                const double* p = reinterpret_cast<const double*>(addr);
                q0 = p[0];
                q1 = p[1];
                q2 = p[2];
                q3 = p[3];
                q4 = p[4];
                q5 = p[5];
                q6 = p[6];
                q7 = p[7];
                q8 = p[8];
                q9 = p[9];
                q10 = p[10];
                q11 = p[11];
                q12 = p[12];
                q13 = p[13];
                // the words only the neighbour-list states use (node / off_end and the walls'
                // list starts) are marked as needed here: the compiler would otherwise issue
                // their loads inside those states, after the wait for the first words, and the
                // step would read the line twice in a row (101.6 -> 99.6 ms per 2^22-packet
                // launch on the spiral disk at tauV 50 with all of it issued at once)
                asm volatile("" : "+v"(q7), "+v"(q11), "+v"(q12), "+v"(q13));
            }
            // double k (bytes 8k..8k+7), k a compile-time constant
            __device__ __forceinline__ double d(int k) const
            {
                switch (k)
                {
                    case 0: return q0;
                    case 1: return q1;
                    case 2: return q2;
                    case 3: return q3;
                    case 4: return q4;
                    case 5: return q5;
                    case 6: return q6;
                    case 7: return q7;
                    case 8: return q8;
                    case 9: return q9;
                    case 10: return q10;
                    case 11: return q11;
                    case 12: return q12;
                    default: return q13;
                }
            }
            // int k (bytes 4k..4k+3), k a compile-time constant
            __device__ __forceinline__ int i32(int k) const
            {
                const double v = d(k >> 1);
                return (k & 1) ? __double2hiint(v) : __double2loint(v);
            }
            // the entry r (run time) of the first four ints
            __device__ __forceinline__ int first4(int r) const
            {
                const int e0 = __double2loint(q0), e1 = __double2hiint(q0), e2 = __double2loint(q1),
                          e3 = __double2hiint(q1);
                const int e23 = r == 2 ? e2 : e3;
                const int e123 = r == 1 ? e1 : e23;
                return r == 0 ? e0 : e123;
            }
            // OctLeaf::off[wl] (ints 22..27) and the end of the wall's list (the next wall's
            // start, or off_end, int 15, after the last wall), wl at run time
            __device__ __forceinline__ int off_at(int wl) const
            {
                const int o0 = __double2loint(q11), o1 = __double2hiint(q11), o2 = __double2loint(q12),
                          o3 = __double2hiint(q12), o4 = __double2loint(q13), o5 = __double2hiint(q13);
                return pick6(wl, o0, o1, o2, o3, o4, o5);
            }
            __device__ __forceinline__ int off_end_at(int wl) const
            {
                const int o1 = __double2hiint(q11), o2 = __double2loint(q12), o3 = __double2hiint(q12),
                          o4 = __double2loint(q13), o5 = __double2hiint(q13), oe = __double2hiint(q7);
                return pick6(wl, o1, o2, o3, o4, o5, oe);
            }
            __device__ __forceinline__ static int pick6(int w, int a, int b, int c, int d, int e, int f)
            {
                const int ef = w == 4 ? e : f;
                const int def = w == 3 ? d : ef;
                const int cdef = w == 2 ? c : def;
                const int bcdef = w == 1 ? b : cdef;
                return w == 0 ? a : bcdef;
            }
        };
        // OctLeaf: lo/hi doubles 0..5, dens double 6, node int 14, off_end int 15, first ints
        // 16..21, off ints 22..27.  TopNode: split point doubles 0..2, child int 6, cell int 7.
        static_assert(offsetof(OctLeaf, dens) == 48 && offsetof(OctLeaf, off_end) == 60 && offsetof(OctLeaf, first) == 64
                          && offsetof(OctLeaf, off) == 88,
                      "leaf record layout the step machine reads");
        static_assert(offsetof(TopNode, child) == 24 && offsetof(TopNode, cell) == 28, "node layout the step machine reads");

        // the leaf c becomes the current cell: of its record, the crossing needs only the wall
        // ahead on each axis (the direction does not change along a path) -- its coordinate
        // (TreeSpatialGrid.cpp:157-159) and its first neighbour -- and the density
        __device__ __forceinline__ void adopt(const RawRec& R, int c)
        {
            i = c;
            b0 = (kx < 0.0) ? R.d(0) : R.d(3);
            b1 = (ky < 0.0) ? R.d(1) : R.d(4);
            b2 = (kz < 0.0) ? R.d(2) : R.d(5);
            bn = R.d(6);
            f0 = (kx < 0.0) ? R.i32(16) : R.i32(17);
            f1 = (ky < 0.0) ? R.i32(18) : R.i32(19);
            f2 = (kz < 0.0) ? R.i32(20) : R.i32(21);
            ost = kOsCross;
        }

        __device__ __forceinline__ static bool closed_contains(const RawRec& R, double x, double y, double z)
        {
            return x >= R.d(0) && x <= R.d(3) && y >= R.d(1) && y <= R.d(4) && z >= R.d(2) && z <= R.d(5);
        }

        // TreeNode::leafChild (TreeSpatialGrid.cpp:190-219, OctTreeNode.cpp:37-42) from node nd
        // through the LDS-staged levels, with no global load: the cell of the leaf containing r
        // (-1 outside the root's closed box), or -2 with the first node below the staged levels
        // in `node`, whose record a later step loads
        __device__ __forceinline__ int search_lds(const GridView& g, int nd)
        {
            if (nd == 0
                && !(rx >= g.box[0] && rx <= g.box[3] && ry >= g.box[1] && ry <= g.box[4] && rz >= g.box[2]
                     && rz <= g.box[5]))
                return -1;
            while (nd < g.topN)
            {
                const TopNode t = skg_top_lds[nd];
                if (t.child < 0) return t.cell;
                nd = t.child + (rx < t.sx ? 0 : 1) + (ry < t.sy ? 0 : 2) + (rz < t.sz ? 0 : 4);
            }
            node = nd;
            return -2;
        }

        // a search result c (-2: the search continues at `node`).  Starting a path, c is the
        // start cell.  After a crossing, c is the next cell unless it is the current one: then
        // the position moves by one ulp toward the direction and the search is made once more;
        // outside, or still the current cell, the path ends (TreeSpatialGrid.cpp:204-218).
        __device__ __forceinline__ void resolve(const GridView& g, int c)
        {
            if (c == -2)
            {
                ost = kOsNode;
                return;
            }
            if (!locating && c == i && !na)
            {
                na = true;
                rx = nextafter(rx, (kx < 0.) ? -DBL_MAX : DBL_MAX);
                ry = nextafter(ry, (ky < 0.) ? -DBL_MAX : DBL_MAX);
                rz = nextafter(rz, (kz < 0.) ? -DBL_MAX : DBL_MAX);
                c = search_lds(g, 0);
                if (c == -2)
                {
                    ost = kOsNode;
                    return;
                }
            }
            // the same members are written on both outcomes (as values), so that the compiler
            // keeps them in registers
            const bool end = c < 0 || (!locating && c == i);
            ost = end ? kOsEnd : kOsLeaf;
            state = end ? kOutside : state;
            cand = end ? cand : c;
        }

        // One step of the octree generator (TreeSpatialGrid::MySegmentGenerator::next,
        // TreeSpatialGrid.cpp:143-228, split so that a step makes at most one dependent global
        // load): 1 with the segment (m, ds, n) when the step completed one, 2 with the path's last
        // segment (the path ends after it), 0 when it only progressed (a start, a neighbour-list
        // entry, a search level), -1 at the end of the path.  The sequence of segments equals
        // next()'s.
        __device__ __forceinline__ int advance_octree(const GridView& g, int& m, double& ds, double& n)
        {
            int out = 0;
            int act = ost;
            const char* addr = nullptr;
            // phase 1: the step's own computation and the address of the record it needs
            if (ost == kOsEnd) return -1;
            if (ost == kOsStart)
            {
                double cumds;
                if (!move_inside(g, cumds))
                {
                    ost = kOsEnd;
                    return -1;
                }
                locating = true;
                na = false;
                rkx = 1.f / static_cast<float>(kx);
                rky = 1.f / static_cast<float>(ky);
                rkz = 1.f / static_cast<float>(kz);
                if (cumds > 0.)
                {
                    m = -1;
                    ds = cumds;
                    n = 0.;
                    out = 1;
                }
                act = hint >= 0 ? kOsStart : kOsSearch;  // the hint's record follows, or the search
            }
            else if (ost == kOsCross)
            {
                const int cur = i;
                // The exit distance along each axis to the wall ahead (b0..b2, from adopt):
                // (wall - r) / k, or DBL_MAX for a component within 1e-15 of zero, and the
                // smallest wins with the ties of TreeSpatialGrid.cpp:157-181 (x if dsx <= dsy &&
                // dsx <= dsz, y if dsy <= dsx && dsy <= dsz, else z).  FP32 estimates (relative
                // error below 3e-7) pick the axis when the two smallest differ by more than 1e-6
                // of their size: the exact FP64 quotients then order the same way and differ,
                // so only the winner's is computed (one FP64 division instead of three).  Close
                // estimates, infinities and NaNs take the three exact quotients.
                const double dx = b0 - rx, dy = b1 - ry, dz = b2 - rz;
                const bool tx = fabs(kx) > 1e-15, ty = fabs(ky) > 1e-15, tz = fabs(kz) > 1e-15;
                const float ex = tx ? static_cast<float>(dx) * rkx : INFINITY;
                const float ey = ty ? static_cast<float>(dy) * rky : INFINITY;
                const float ez = tz ? static_cast<float>(dz) * rkz : INFINITY;
                const bool fx = ex <= ey && ex <= ez;
                const bool fy = !fx && ey <= ez;
                const float e1 = pick3(fx, fy, ex, ey, ez);
                const float e2 = pick3(fx, fy, fminf(ey, ez), fminf(ex, ez), fminf(ex, ey));
                const bool clear = e2 - e1 > 1e-6f * (fabsf(e1) + fabsf(e2));  // false for inf / NaN
                bool ax, ay;
                double d;
                if (clear)
                {
                    ax = fx;
                    ay = fy;
                    d = pick3(ax, ay, dx, dy, dz) / pick3(ax, ay, kx, ky, kz);
                }
                else
                {
                    const double dsx = tx ? dx / kx : DBL_MAX;
                    const double dsy = ty ? dy / ky : DBL_MAX;
                    const double dsz = tz ? dz / kz : DBL_MAX;
                    ax = dsx <= dsy && dsx <= dsz;
                    ay = !ax && dsy <= dsx && dsy <= dsz;
                    d = pick3(ax, ay, dsx, dsy, dsz);
                }
                wall = pick3(ax, ay, (kx < 0.0) ? 0 : 1, (ky < 0.0) ? 2 : 3, (kz < 0.0) ? 4 : 5);
                const int f = pick3(ax, ay, f0, f1, f2);
                double step = d + g.eps;
                rx += kx * step;
                ry += ky * step;
                rz += kz * step;
                m = cur;
                ds = d;
                n = bn;
                out = 1;
                locating = false;
                na = false;
                cand = f;
                act = f >= 0 ? kOsCross : kOsSearch;  // the first neighbour's record follows, or the search
                // a wall without neighbours is on the domain's boundary: outside the root's closed
                // box the search finds no leaf (TreeNode::leafChild), and the path ends here
                if (f < 0 && !(rx >= g.box[0] && rx <= g.box[3] && ry >= g.box[1] && ry <= g.box[4] && rz >= g.box[2]
                               && rz <= g.box[5]))
                {
                    ost = kOsEnd;
                    state = kOutside;
                    return 2;
                }
            }
            // one search for all lanes that need it, after the step's load (a start without hint
            // and a wall on the domain boundary load a record they do not use, so that the search
            // is written once, in phase 3)
            switch (act)
            {
                case kOsSearch: addr = reinterpret_cast<const char*>(g.leaf); break;
                case kOsStart: addr = reinterpret_cast<const char*>(g.leaf + hint); break;
                case kOsCross:
                case kOsListRec:
                case kOsLeaf: addr = reinterpret_cast<const char*>(g.leaf + cand); break;
                case kOsListStart: addr = reinterpret_cast<const char*>(g.leaf + i); break;
                case kOsListId: addr = reinterpret_cast<const char*>(g.nbc + (q & ~3)); break;
                default: addr = reinterpret_cast<const char*>(g.top + node); break;  // kOsNode
            }
            // phase 2: the one load of this step
            RawRec R;
            R.load(addr);
            // phase 3: what the record decides.  Each state ends by adopting the loaded record
            // as cell `take`, by resolving a search from node `from` or a known cell `known`, or
            // by a change of state; the adoption and the search are written once for all
            // states, so a wave whose lanes are in different states executes each once.
            const bool closed = closed_contains(R, rx, ry, rz);
            int take = -1, from = -1, known = -1;
            switch (act)
            {
                case kOsStart:
                {
                    // the start hint, when the point lies in its box the way the top-down search
                    // assigns points to leaves: lo <= r < hi on each axis (a point on a split
                    // plane goes to the upper child), hi included at the domain's upper boundary.
                    // Every split plane of the leaf's ancestors then sends leafChild's descent
                    // toward it, so it is the leaf the search returns.
                    const bool inx = rx >= R.d(0) && (rx < R.d(3) || R.d(3) == g.box[3]);
                    const bool iny = ry >= R.d(1) && (ry < R.d(4) || R.d(4) == g.box[4]);
                    const bool inz = rz >= R.d(2) && (rz < R.d(5) || R.d(5) == g.box[5]);
                    if (inx && iny && inz)
                        take = hint;
                    else
                        from = 0;
                    break;
                }
                case kOsCross:
                case kOsListRec:
                    // a neighbour on the crossed wall (the first, largest-overlap one, then the
                    // rest of the list in order), when its closed box contains the new position
                    if (closed)
                    {
                        if (cand != i)
                            take = cand;
                        else
                            known = cand;
                    }
                    else if (act == kOsCross)
                        ost = kOsListStart;
                    else if (++q < qe)
                        ost = kOsListId;
                    else
                        from = 0;
                    break;
                case kOsListStart:
                    q = R.off_at(wall) + 1;
                    qe = R.off_end_at(wall);
                    if (q < qe)
                        ost = kOsListId;
                    else
                        from = 0;
                    break;
                case kOsListId:
                    cand = R.first4(q & 3);
                    ost = kOsListRec;
                    break;
                case kOsSearch:
                    from = 0;
                    break;
                case kOsNode:
                {
                    const int child = R.i32(6);
                    if (child < 0)
                        known = R.i32(7);
                    else
                        from = child + (rx < R.d(0) ? 0 : 1) + (ry < R.d(1) ? 0 : 2) + (rz < R.d(2) ? 0 : 4);
                    break;
                }
                default:  // kOsLeaf
                    take = cand;
                    break;
            }
            if (take >= 0)
            {
                adopt(R, take);
                if (locating) hint = take;
            }
            else if (from >= 0)
                resolve(g, search_lds(g, from));
            else if (known >= 0)
                resolve(g, known);
            if (ost == kOsEnd) return out ? 2 : -1;
            return out;
        }

        // one advance() per call for the octree (next() for the other grids)
        __device__ __forceinline__ int advance(const GridView& g, int& m, double& ds, double& n)
        {
            if constexpr (GT == kOctree)
                return advance_octree(g, m, ds, n);
            else
                return next(g, m, ds, n) ? 1 : -1;
        }

        // Returns false when the path has no more segments; otherwise sets the cell m, the
        // segment length ds and the cell's number density n (0 for m = -1).
        __device__ __forceinline__ bool next(const GridView& g, int& m, double& ds, double& n)
        {
            if constexpr (GT == kOctree)
            {
                int r;
                while ((r = advance_octree(g, m, ds, n)) == 0) {}
                return r > 0;
            }
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
                else if (GT == kVoronoi)
                {
                    if (hint >= 0 && vor_nearest_strict(g, hint, rx, ry, rz))
                        i = hint;
                    else
                        i = vor_cell_index(g, rx, ry, rz);
                    hint = i;
                    if (i < 0)
                    {
                        state = kOutside;
                        return false;
                    }
                    vor_enter(g, i);
                }
                else
                {
                    if (hint >= 0 && tet_inside_strict(g.tcell[hint], rx, ry, rz))
                        i = hint;
                    else
                        i = tet_cell_index(g, rx, ry, rz);
                    j = -1;
                    hint = i;
                    if (i < 0)
                    {
                        state = kOutside;
                        return false;
                    }
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
            else if constexpr (GT == kVoronoi)
                return next_voronoi(g, m, ds, n);
            else if constexpr (GT == kTetra)
                return next_tetra(g, m, ds, n);
            return false;  // the octree steps through advance_octree
        }

        // the cell c becomes the current cell, from its record: its site and its run
        __device__ __forceinline__ void vor_enter(const GridView& g, int c)
        {
            const VorCell& C = g.vcell[c];
            j = C.begin;
            sx = C.x;
            sy = C.y;
            sz = C.z;
        }

        // exit through the nearest bisector plane or domain wall (VoronoiMeshSnapshot.cpp:1119-1210)
        __device__ __forceinline__ bool next_voronoi(const GridView& g, int& m, double& ds, double& n)
        {
            for (int tries = 0; tries < kMaxNoExit; ++tries)
            {
                const double px0 = sx, py0 = sy, pz0 = sz;
                double sq = DBL_MAX;
                int mq = -99;
                double nsx = 0., nsy = 0., nsz = 0.;  // the winning entry's site and run
                int nrun = -1;
                // exit distance through one neighbour entry; the smallest positive one wins,
                // strict '<' (the first of equal distances wins)
                auto consider = [&](const VorNb E) {
                    const int mi = E.id;
                    double si = 0;
                    if (mi >= 0)
                    {
                        double nx = E.x - px0, ny = E.y - py0, nz = E.z - pz0;
                        double ndotk = nx * kx + ny * ky + nz * kz;
                        if (ndotk > 0)
                        {
                            double qx = 0.5 * (E.x + px0), qy = 0.5 * (E.y + py0), qz = 0.5 * (E.z + pz0);
                            si = (nx * (qx - rx) + ny * (qy - ry) + nz * (qz - rz)) / ndotk;
                        }
                    }
                    else
                    {
                        // walls -1..-6: (box face - r_axis) / k_axis
                        const int w = -1 - mi;
                        const int ax = w >> 1;
                        const double face = g.box[(w & 1) ? 3 + ax : ax];
                        // register copies, selected by axis: a select among the members
                        // themselves becomes an indexed load from a scratch copy of them, which
                        // is then stored at every crossing (1412 -> 1085 ms per 2^22-packet launch
                        // on the 1e6-site clumpy medium without it)
                        double tx = rx, ty = ry, tz = rz, ux = kx, uy = ky, uz = kz;
                        asm volatile("" : "+v"(tx), "+v"(ty), "+v"(tz), "+v"(ux), "+v"(uy), "+v"(uz));
                        const double ra = ax == 0 ? tx : (ax == 1 ? ty : tz);
                        const double ka = ax == 0 ? ux : (ax == 1 ? uy : uz);
                        if (w < 6) si = (face - ra) / ka;
                    }
                    if (si > 0 && si < sq)
                    {
                        sq = si;
                        mq = mi;
                        nsx = E.x;
                        nsy = E.y;
                        nsz = E.z;
                        nrun = E.run;
                    }
                };
                // the run: its header and first three entries in one group, then the entries
                // four at a time so that their loads are in flight together
                const int hb = j;
                const VorNb H = g.vnb[hb], E1 = g.vnb[hb + 1], E2 = g.vnb[hb + 2], E3 = g.vnb[hb + 3];
                const double dens = H.x;
                const int e = hb + 1 + H.run;
                if (hb + 1 < e) consider(E1);
                if (hb + 2 < e) consider(E2);
                if (hb + 3 < e) consider(E3);
                int q = hb + 4;
                for (; q + 4 <= e; q += 4)
                {
                    const VorNb F0 = g.vnb[q], F1 = g.vnb[q + 1], F2 = g.vnb[q + 2], F3 = g.vnb[q + 3];
                    consider(F0);
                    consider(F1);
                    consider(F2);
                    consider(F3);
                }
                for (; q < e; ++q) consider(g.vnb[q]);
                if (mq == -99)
                {
                    rx += kx * g.eps;
                    ry += ky * g.eps;
                    rz += kz * g.eps;
                    i = vor_cell_index(g, rx, ry, rz);
                    if (i < 0)
                    {
                        state = kOutside;
                        return false;
                    }
                    vor_enter(g, i);
                    continue;
                }
                const double step = sq + g.eps;
                rx += kx * step;
                ry += ky * step;
                rz += kz * step;
                m = i;
                ds = sq;
                n = dens;
                i = mq;
                j = nrun;
                sx = nsx;
                sy = nsy;
                sz = nsz;
                if (mq < 0) state = kOutside;
                return true;
            }
            state = kOutside;
            return false;
        }

        // Tetra::findEnteringFace: the first of the six edges with a nonzero Pluecker product
        __device__ __forceinline__ int entering_face(const TetCell& T) const
        {
            int e = 0;
            for (int t1 = 0; t1 < 3; t1++)
                for (int t2 = t1 + 1; t2 < 4; t2++)
                {
                    double wx = rx - T.v[t1][0], wy = ry - T.v[t1][1], wz = rz - T.v[t1][2];
                    double mx = ky * wz - kz * wy, my = kz * wx - kx * wz, mz = kx * wy - ky * wx;
                    double prod = mx * (T.v[t2][0] - T.v[t1][0]) + my * (T.v[t2][1] - T.v[t1][1])
                                  + mz * (T.v[t2][2] - T.v[t1][2]);
                    if (prod != 0.)
                    {
                        // etable {{3,2},{1,3},{2,1},{3,0},{0,2},{1,0}} (TetraMeshSpatialGrid.cpp:91)
                        constexpr int neg = 3 | (1 << 2) | (2 << 4) | (3 << 6) | (0 << 8) | (1 << 10);
                        constexpr int pos = 2 | (3 << 2) | (1 << 4) | (0 << 6) | (2 << 8) | (0 << 10);
                        return ((prod < 0 ? neg : pos) >> (2 * e)) & 3;
                    }
                    e++;
                }
            return -1;
        }

        // Pluecker exit-face decision with the plane-intersection fallback
        // (TetraMeshSpatialGrid.cpp:719-819); the final segment before the hull is not returned.
        // Ends the path where the reference would not terminate: no entering face (all six
        // products zero) or a non-finite position.
        __device__ __forceinline__ bool next_tetra(const GridView& g, int& m, double& ds, double& n)
        {
            for (int tries = 0; tries < kMaxNoExit; ++tries)
            {
                if (!(isfinite(rx) && isfinite(ry) && isfinite(rz))) break;
                const TetCell& T = g.tcell[i];
                if (j < 0) j = entering_face(T);
                if (j < 0) break;
                const int ef = j;
                const double vx = T.v[ef][0], vy = T.v[ef][1], vz = T.v[ef][2];
                double wx = rx - vx, wy = ry - vy, wz = rz - vz;
                double mx = ky * wz - kz * wy, my = kz * wx - kx * wz, mz = kx * wy - ky * wx;
                int c0, c1, c2;
                clockwise_vertices(ef, c0, c1, c2);
                double prod0 = mx * (vx - T.v[c0][0]) + my * (vy - T.v[c0][1]) + mz * (vz - T.v[c0][2]);
                int clock0 = prod0 < 0;
                int ci = clock0 ? c1 : c2;
                double prodi = mx * (vx - T.v[ci][0]) + my * (vy - T.v[ci][1]) + mz * (vz - T.v[ci][2]);
                int cclocki = prodi >= 0;
                int leaving = -1;
                double dsl = DBL_MAX;
                if (prod0 == 0. || prodi == 0.)
                {
                    const int cv[3] = {c0, c1, c2};
                    for (int q = 0; q < 3; ++q)
                    {
                        const int f = cv[q];
                        double ndotk = T.n[f][0] * kx + T.n[f][1] * ky + T.n[f][2] * kz;
                        if (ndotk > 0)
                        {
                            double dq = (T.n[f][0] * (vx - rx) + T.n[f][1] * (vy - ry) + T.n[f][2] * (vz - rz)) / ndotk;
                            if (dq < dsl)
                            {
                                dsl = dq;
                                leaving = f;
                            }
                        }
                    }
                }
                else
                {
                    // dtable {{1,0},{0,2}} indexes the clockwise vertices
                    leaving = clock0 ? (cclocki ? c2 : c0) : (cclocki ? c0 : c1);
                    double ndotk = T.n[leaving][0] * kx + T.n[leaving][1] * ky + T.n[leaving][2] * kz;
                    dsl = (T.n[leaving][0] * (vx - rx) + T.n[leaving][1] * (vy - ry) + T.n[leaving][2] * (vz - rz)) / ndotk;
                }
                if (leaving == -1 || dsl < g.eps)
                {
                    rx += kx * g.eps;
                    ry += ky * g.eps;
                    rz += kz * g.eps;
                    i = tet_cell_index(g, rx, ry, rz);
                    if (i < 0)
                    {
                        state = kOutside;
                        return false;
                    }
                    j = -1;
                    continue;
                }
                rx += kx * dsl;
                ry += ky * dsl;
                rz += kz * dsl;
                const int nxt = T.nb[leaving];
                if (nxt < 0)
                {
                    state = kOutside;
                    return false;
                }
                m = i;
                ds = dsl;
                n = T.dens;
                j = T.nface[leaving];
                i = nxt;
                return true;
            }
            state = kOutside;
            return false;
        }
    };
}
