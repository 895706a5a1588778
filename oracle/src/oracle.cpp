// FP64 CPU oracle for the SKIRT photon-packet hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in our own code, the reference CPU
// algorithm for grid traversal, optical depth, forced scattering, Henyey-Greenstein
// sampling, peel-off and frame detection.  It is the checker for the HIP engine and the
// "port" CPU baseline; nothing in the product links or calls it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Each function cites the reference file:line it follows (paths relative to the
// reference repository root).  Random numbers come from a counter-based Philox4x32-10
// stream per photon-packet history (the reference uses a thread-local mt19937_64,
// Random.cpp:20-56); the distributions drawn from it follow the reference exactly, so
// images agree with the reference statistically and with the HIP engine draw-for-draw.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off, so no FMA contraction, as in the
// reference's x86-64 build).

#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace
{
    // ---------------------------------------------------------------- Philox4x32-10
    // Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11).
    struct U4
    {
        uint32_t v[4];
    };

    inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo)
    {
        uint64_t p = static_cast<uint64_t>(a) * b;
        hi = static_cast<uint32_t>(p >> 32);
        lo = static_cast<uint32_t>(p);
    }

    U4 philox(U4 c, uint32_t k0, uint32_t k1)
    {
        for (int r = 0; r < 10; ++r)
        {
            if (r)
            {
                k0 += 0x9E3779B9u;
                k1 += 0xBB67AE85u;
            }
            uint32_t hi0, lo0, hi1, lo1;
            mulhilo(0xD2511F53u, c.v[0], hi0, lo0);
            mulhilo(0xCD9E8D57u, c.v[2], hi1, lo1);
            U4 n;
            n.v[0] = hi1 ^ c.v[1] ^ k0;
            n.v[1] = lo1;
            n.v[2] = hi0 ^ c.v[3] ^ k1;
            n.v[3] = lo0;
            c = n;
        }
        return c;
    }

    // Uniform deviates on (0,1) for one history: draw i uses block i/2 of the stream
    // (counter = history lo, history hi, block, tag) and the word pair (2*(i%2), 2*(i%2)+1);
    // u = (52-bit integer + 0.5) * 2^-52, never 0 or 1 (as Random.cpp:24-27 excludes both).
    struct Stream
    {
        uint32_t k0, k1, h0, h1, tag;
        uint32_t draw = 0;
        uint32_t block = 0xFFFFFFFFu;
        U4 cache{};

        Stream(uint64_t seed, uint64_t history, uint32_t tg)
            : k0(static_cast<uint32_t>(seed)), k1(static_cast<uint32_t>(seed >> 32)),
              h0(static_cast<uint32_t>(history)), h1(static_cast<uint32_t>(history >> 32)), tag(tg)
        {}

        double uniform()
        {
            uint32_t b = draw >> 1;
            if (b != block)
            {
                cache = philox(U4{{h0, h1, b, tag}}, k0, k1);
                block = b;
            }
            int w = (draw & 1) * 2;
            ++draw;
            uint64_t a = (static_cast<uint64_t>(cache.v[w]) << 20) ^ (cache.v[w + 1] >> 12);  // 52 bits
            return (static_cast<double>(a) + 0.5) * (1.0 / 4503599627370496.0);
        }
    };

    // ---------------------------------------------------------------- geometry helpers
    struct Box
    {
        double xmin, ymin, zmin, xmax, ymax, zmax;
        // Box.hpp:99-102 (closed box)
        bool contains(double x, double y, double z) const
        {
            return x >= xmin && x <= xmax && y >= ymin && y <= ymax && z >= zmin && z <= zmax;
        }
        // Box.hpp:125-129
        double diagonal() const
        {
            return std::sqrt((xmax - xmin) * (xmax - xmin) + (ymax - ymin) * (ymax - ymin)
                             + (zmax - zmin) * (zmax - zmin));
        }
    };

    // NR.hpp:130-144 (locateBasicImpl) and :177-183 (locateClip)
    inline int locateClip(const double* xv, int n, double x)
    {
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

    enum class State
    {
        Unknown,
        Inside,
        Outside
    };

    // PathSegmentGenerator state and helpers (PathSegmentGenerator.hpp:43-200)
    struct GenBase
    {
        State state = State::Unknown;
        double rx, ry, rz, kx, ky, kz;
        int m = -1;
        double ds = 0.;

        void start(const double* r, const double* k)
        {
            state = State::Unknown;
            rx = r[0];
            ry = r[1];
            rz = r[2];
            kx = k[0];
            ky = k[1];
            kz = k[2];
        }
        void propagater(double s)
        {
            rx += kx * s;
            ry += ky * s;
            rz += kz * s;
        }
        // PathSegmentGenerator.hpp:148-153
        void propagateToNextAfter()
        {
            rx = std::nextafter(rx, (kx < 0.) ? -DBL_MAX : DBL_MAX);
            ry = std::nextafter(ry, (ky < 0.) ? -DBL_MAX : DBL_MAX);
            rz = std::nextafter(rz, (kz < 0.) ? -DBL_MAX : DBL_MAX);
        }
        // PathSegmentGenerator.cpp:11-112
        bool moveInside(const Box& box, double eps)
        {
            m = -1;
            ds = 0.;
            state = State::Outside;
            double cumds = 0.;
            if (rx <= box.xmin)
            {
                if (kx <= 0.0) return false;
                double d = (box.xmin - rx) / kx;
                rx = box.xmin + eps;
                ry += ky * d;
                rz += kz * d;
                cumds += d;
            }
            else if (rx >= box.xmax)
            {
                if (kx >= 0.0) return false;
                double d = (box.xmax - rx) / kx;
                rx = box.xmax - eps;
                ry += ky * d;
                rz += kz * d;
                cumds += d;
            }
            if (ry <= box.ymin)
            {
                if (ky <= 0.0) return false;
                double d = (box.ymin - ry) / ky;
                rx += kx * d;
                ry = box.ymin + eps;
                rz += kz * d;
                cumds += d;
            }
            else if (ry >= box.ymax)
            {
                if (ky >= 0.0) return false;
                double d = (box.ymax - ry) / ky;
                rx += kx * d;
                ry = box.ymax - eps;
                rz += kz * d;
                cumds += d;
            }
            if (rz <= box.zmin)
            {
                if (kz <= 0.0) return false;
                double d = (box.zmin - rz) / kz;
                rx += kx * d;
                ry += ky * d;
                rz = box.zmin + eps;
                cumds += d;
            }
            else if (rz >= box.zmax)
            {
                if (kz >= 0.0) return false;
                double d = (box.zmax - rz) / kz;
                rx += kx * d;
                ry += ky * d;
                rz = box.zmax - eps;
                cumds += d;
            }
            if (!box.contains(rx, ry, rz)) return false;
            m = -1;
            ds = cumds;
            state = State::Inside;
            return true;
        }
    };

    // ---------------------------------------------------------------- grids
    struct Grid
    {
        int type = 0;  // 0 Cartesian, 1 octree
        Box extent{};
        double eps = 0.;
        // Cartesian
        const double *xv = nullptr, *yv = nullptr, *zv = nullptr;
        int Nx = 0, Ny = 0, Nz = 0;
        // octree: node bounds (6 per node), first child id (-1 for leaves), cell index per node,
        // per-node per-wall neighbour lists (CSR over 6*nnodes entries of node ids)
        const double* nb = nullptr;
        const int32_t* childFirst = nullptr;
        const int32_t* cellOf = nullptr;
        const int64_t* nbOff = nullptr;
        const int32_t* nbId = nullptr;

        // Voronoi (VoronoiMeshSnapshot.hpp:338-349): site coordinates, neighbour lists (walls
        // as -1..-6) and the block search lists over a vbn^3 block grid
        const double* sites = nullptr;
        const int32_t* vnbBegin = nullptr;
        const int32_t* vnbCount = nullptr;
        const int32_t* vnbIndex = nullptr;
        int vbn = 0;
        const int32_t* vblBegin = nullptr;
        const int32_t* vblCount = nullptr;
        const int32_t* vblIndex = nullptr;
        // tetra (TetraMeshSpatialGrid.hpp:293-303): vertices, per-tetra vertex indices, face
        // neighbours and outward unit normals; the neighbour's face index and a block search
        // grid are derived here
        const double* verts = nullptr;
        int64_t ntets = 0;
        const int32_t* tetVerts = nullptr;
        const int32_t* faceNb = nullptr;
        const double* faceNormal = nullptr;
        std::vector<int8_t> faceNbFace;
        int tbn = 0;
        std::vector<double> tbx, tby, tbz;
        std::vector<int64_t> tblOff;
        std::vector<int32_t> tblIds;

        // CartesianSpatialGrid.cpp:200-203
        int index(int i, int j, int k) const { return k + Nz * j + Nz * Ny * i; }

        // Box::cellIndices, Box.hpp:162-167
        void blockIndices(int& i, int& j, int& k, double x, double y, double z, int n) const
        {
            i = std::max(0, std::min(n - 1, static_cast<int>(n * (x - extent.xmin) / (extent.xmax - extent.xmin))));
            j = std::max(0, std::min(n - 1, static_cast<int>(n * (y - extent.ymin) / (extent.ymax - extent.ymin))));
            k = std::max(0, std::min(n - 1, static_cast<int>(n * (z - extent.zmin) / (extent.zmax - extent.zmin))));
        }

        // VoronoiMeshSnapshot::cellIndex, VoronoiMeshSnapshot.cpp:1037-1071: the nearest site
        // among the block's candidates (squared distance (r - site).norm2(), strict '<', first
        // wins).  The reference searches blocks with more than 9 candidates with a k-d tree,
        // which returns the same site except for exact distance ties.
        int vorCellIndex(double x, double y, double z) const
        {
            if (!extent.contains(x, y, z)) return -1;
            int i, j, k;
            blockIndices(i, j, k, x, y, z, vbn);
            int b = i * vbn * vbn + j * vbn + k;
            int m = -1;
            double mdist = DBL_MAX;
            for (int q = vblBegin[b], e = vblBegin[b] + vblCount[b]; q < e; ++q)
            {
                int id = vblIndex[q];
                const double* p = sites + 3 * static_cast<int64_t>(id);
                double dx = x - p[0], dy = y - p[1], dz = z - p[2];
                double d = dx * dx + dy * dy + dz * dz;
                if (d < mdist)
                {
                    m = id;
                    mdist = d;
                }
            }
            return m;
        }

        const double* tetVertex(int64_t m, int t) const { return verts + 3 * static_cast<int64_t>(tetVerts[4 * m + t]); }
        const double* tetNormal(int64_t m, int f) const { return faceNormal + 12 * m + 3 * f; }

        // Tetra::contains, TetraMeshSpatialGrid.cpp:110-133 (bounding box first, then the four
        // face planes with the outward normals; a point on a face is contained)
        bool tetContains(int64_t m, double x, double y, double z) const
        {
            double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
            for (int t = 0; t < 4; ++t)
            {
                const double* v = tetVertex(m, t);
                for (int a = 0; a < 3; ++a)
                {
                    lo[a] = std::min(lo[a], v[a]);
                    hi[a] = std::max(hi[a], v[a]);
                }
            }
            if (!(x >= lo[0] && x <= hi[0] && y >= lo[1] && y <= hi[1] && z >= lo[2] && z <= hi[2])) return false;
            const double* v = tetVertex(m, 3);
            for (int f = 0; f <= 2; ++f)
            {
                const double* n = tetNormal(m, f);
                if ((v[0] - x) * n[0] + (v[1] - y) * n[1] + (v[2] - z) * n[2] < 0) return false;
            }
            v = tetVertex(m, 0);
            const double* n = tetNormal(m, 3);
            if ((v[0] - x) * n[0] + (v[1] - y) * n[1] + (v[2] - z) * n[2] < 0.) return false;
            return true;
        }

        // TetraMeshSpatialGrid::cellIndex, TetraMeshSpatialGrid.cpp:615-624: the first candidate
        // of the point's search block that contains it.  The reference's BoxSearch lists hold
        // every tetra whose (closed) bounding box meets the block in increasing index order, so
        // the answer is the lowest-index tetra containing the point; the uniform block grid
        // used here has the same property.
        int tetCellIndex(double x, double y, double z) const
        {
            if (tbn == 0) return -1;
            int i = static_cast<int>(std::upper_bound(tbx.begin(), tbx.end(), x) - tbx.begin()) - 1;
            int j = static_cast<int>(std::upper_bound(tby.begin(), tby.end(), y) - tby.begin()) - 1;
            int k = static_cast<int>(std::upper_bound(tbz.begin(), tbz.end(), z) - tbz.begin()) - 1;
            i = std::max(0, std::min(tbn - 1, i));
            j = std::max(0, std::min(tbn - 1, j));
            k = std::max(0, std::min(tbn - 1, k));
            int64_t b = (static_cast<int64_t>(i) * tbn + j) * tbn + k;
            for (int64_t q = tblOff[b]; q < tblOff[b + 1]; ++q)
                if (tetContains(tblIds[q], x, y, z)) return tblIds[q];
            return -1;
        }

        void buildTetraSearch()
        {
            // neighbour face index: the face of the neighbour that points back (TetraMeshSpatialGrid.cpp:508-520)
            faceNbFace.assign(4 * ntets, -1);
            for (int64_t m = 0; m < ntets; ++m)
                for (int f = 0; f < 4; ++f)
                {
                    int nt = faceNb[4 * m + f];
                    if (nt != -1)
                        for (int fn = 0; fn < 4; ++fn)
                            if (faceNb[4 * static_cast<int64_t>(nt) + fn] == m)
                            {
                                faceNbFace[4 * m + f] = static_cast<int8_t>(fn);
                                break;
                            }
                }
            tbn = std::max(1, static_cast<int>(std::cbrt(static_cast<double>(ntets)) / 2));
            tbx.resize(tbn + 1);
            tby.resize(tbn + 1);
            tbz.resize(tbn + 1);
            for (int i = 0; i <= tbn; ++i)
            {
                tbx[i] = extent.xmin + (extent.xmax - extent.xmin) * i / tbn;
                tby[i] = extent.ymin + (extent.ymax - extent.ymin) * i / tbn;
                tbz[i] = extent.zmin + (extent.zmax - extent.zmin) * i / tbn;
            }
            std::vector<std::vector<int32_t>> lists(static_cast<size_t>(tbn) * tbn * tbn);
            auto lower = [](const std::vector<double>& g, double v) {
                // first block whose upper border is >= v (closed overlap)
                int i = static_cast<int>(std::lower_bound(g.begin() + 1, g.end(), v) - (g.begin() + 1));
                return std::min(i, static_cast<int>(g.size()) - 2);
            };
            auto upper = [](const std::vector<double>& g, double v) {
                // last block whose lower border is <= v
                int i = static_cast<int>(std::upper_bound(g.begin(), g.end() - 1, v) - g.begin()) - 1;
                return std::max(0, i);
            };
            for (int64_t m = 0; m < ntets; ++m)
            {
                double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
                for (int t = 0; t < 4; ++t)
                    for (int a = 0; a < 3; ++a)
                    {
                        lo[a] = std::min(lo[a], tetVertex(m, t)[a]);
                        hi[a] = std::max(hi[a], tetVertex(m, t)[a]);
                    }
                int i1 = lower(tbx, lo[0]), i2 = upper(tbx, hi[0]);
                int j1 = lower(tby, lo[1]), j2 = upper(tby, hi[1]);
                int k1 = lower(tbz, lo[2]), k2 = upper(tbz, hi[2]);
                for (int i = i1; i <= i2; ++i)
                    for (int j = j1; j <= j2; ++j)
                        for (int k = k1; k <= k2; ++k)
                            lists[(static_cast<size_t>(i) * tbn + j) * tbn + k].push_back(static_cast<int32_t>(m));
            }
            tblOff.assign(lists.size() + 1, 0);
            for (size_t b = 0; b < lists.size(); ++b) tblOff[b + 1] = tblOff[b] + static_cast<int64_t>(lists[b].size());
            tblIds.clear();
            for (auto& l : lists) tblIds.insert(tblIds.end(), l.begin(), l.end());
        }

        const double* nodeBounds(int n) const { return nb + 6 * static_cast<int64_t>(n); }
        bool nodeContains(int n, double x, double y, double z) const
        {
            const double* b = nodeBounds(n);
            return x >= b[0] && x <= b[3] && y >= b[1] && y <= b[4] && z >= b[2] && z <= b[5];
        }
        // TreeNode.cpp:66-76 + OctTreeNode.cpp:37-42 (strict '<' against the split point,
        // which is the upper corner of child 0)
        int leafChild(int node, double x, double y, double z) const
        {
            if (!nodeContains(node, x, y, z)) return -1;
            while (childFirst[node] >= 0)
            {
                int c0 = childFirst[node];
                const double* b = nodeBounds(c0);
                int l = (x < b[3] ? 0 : 1) + (y < b[4] ? 0 : 2) + (z < b[5] ? 0 : 4);
                node = c0 + l;
            }
            return node;
        }
        // TreeNode.cpp:103-112
        int neighbor(int node, int wall, double x, double y, double z) const
        {
            int64_t b = nbOff[6 * static_cast<int64_t>(node) + wall], e = nbOff[6 * static_cast<int64_t>(node) + wall + 1];
            for (int64_t i = b; i < e; ++i)
                if (nodeContains(nbId[i], x, y, z)) return nbId[i];
            return -1;
        }
    };

    struct Generator : GenBase
    {
        const Grid* g;
        int i = -1, j = -1, k = -1;
        int node = -1;
        explicit Generator(const Grid* grid) : g(grid) {}

        int mr = -1;            // Voronoi / tetra: current cell
        int enteringFace = -1;  // tetra: entering face of the current cell

        bool next()
        {
            switch (g->type)
            {
                case 0: return nextCartesian();
                case 1: return nextTree();
                case 2: return nextVoronoi();
                default: return nextTetra();
            }
        }

        // VoronoiMeshSnapshot::MySegmentGenerator::next, VoronoiMeshSnapshot.cpp:1094-1213
        bool nextVoronoi()
        {
            switch (state)
            {
                case State::Unknown:
                {
                    if (!moveInside(g->extent, g->eps)) return false;
                    mr = g->vorCellIndex(rx, ry, rz);
                    if (ds > 0.) return true;
                }
                // fall through
                case State::Inside:
                {
                    while (true)
                    {
                        const double* pr = g->sites + 3 * static_cast<int64_t>(mr);
                        double sq = DBL_MAX;
                        const int NO_INDEX = -99;
                        int mq = NO_INDEX;
                        for (int q = g->vnbBegin[mr], e = g->vnbBegin[mr] + g->vnbCount[mr]; q < e; ++q)
                        {
                            int mi = g->vnbIndex[q];
                            double si = 0;
                            if (mi >= 0)
                            {
                                const double* pi = g->sites + 3 * static_cast<int64_t>(mi);
                                double nx = pi[0] - pr[0], ny = pi[1] - pr[1], nz = pi[2] - pr[2];
                                double ndotk = nx * kx + ny * ky + nz * kz;
                                if (ndotk > 0)
                                {
                                    double px = 0.5 * (pi[0] + pr[0]), py = 0.5 * (pi[1] + pr[1]), pz = 0.5 * (pi[2] + pr[2]);
                                    si = (nx * (px - rx) + ny * (py - ry) + nz * (pz - rz)) / ndotk;
                                }
                            }
                            else
                            {
                                switch (mi)
                                {
                                    case -1: si = (g->extent.xmin - rx) / kx; break;
                                    case -2: si = (g->extent.xmax - rx) / kx; break;
                                    case -3: si = (g->extent.ymin - ry) / ky; break;
                                    case -4: si = (g->extent.ymax - ry) / ky; break;
                                    case -5: si = (g->extent.zmin - rz) / kz; break;
                                    case -6: si = (g->extent.zmax - rz) / kz; break;
                                    default: break;
                                }
                            }
                            if (si > 0 && si < sq)
                            {
                                sq = si;
                                mq = mi;
                            }
                        }
                        if (mq == NO_INDEX)
                        {
                            propagater(g->eps);
                            mr = g->vorCellIndex(rx, ry, rz);
                            if (mr < 0)
                            {
                                state = State::Outside;
                                return false;
                            }
                        }
                        else
                        {
                            propagater(sq + g->eps);
                            m = mr;
                            ds = sq;
                            mr = mq;
                            if (mr < 0) state = State::Outside;
                            return true;
                        }
                    }
                }
                case State::Outside: break;
            }
            return false;
        }

        // TetraMeshSpatialGrid.cpp:44-50
        static void clockwiseVertices(int face, int* cv)
        {
            cv[0] = (face + 3) % 4;
            cv[1] = (face + 2) % 4;
            cv[2] = (face + 1) % 4;
            if (face % 2 == 0) std::swap(cv[0], cv[2]);
        }

        // Tetra::findEnteringFace, TetraMeshSpatialGrid.cpp:87-108
        int findEnteringFace(int64_t t) const
        {
            static constexpr int etable[6][2] = {{3, 2}, {1, 3}, {2, 1}, {3, 0}, {0, 2}, {1, 0}};
            int e = 0;
            for (int t1 = 0; t1 < 3; t1++)
                for (int t2 = t1 + 1; t2 < 4; t2++)
                {
                    const double* a = g->tetVertex(t, t1);
                    const double* b = g->tetVertex(t, t2);
                    double wx = rx - a[0], wy = ry - a[1], wz = rz - a[2];
                    double mx = ky * wz - kz * wy, my = kz * wx - kx * wz, mz = kx * wy - ky * wx;
                    double prod = mx * (b[0] - a[0]) + my * (b[1] - a[1]) + mz * (b[2] - a[2]);
                    if (prod != 0.) return prod < 0 ? etable[e][0] : etable[e][1];
                    e++;
                }
            return -1;
        }

        // TetraMeshSpatialGrid::MySegmentGenerator::next, TetraMeshSpatialGrid.cpp:697-820.
        // The final segment, in the last tetrahedron before the path leaves the convex hull, is
        // set but not returned, as in the reference.  Two cases in which the reference does not
        // terminate end the path here instead: no entering face found (all six Pluecker products
        // zero; the reference then indexes vertex -1) and a position that is no longer finite (a
        // 0/0 exit distance, after which the reference walks on forever).
        bool nextTetra()
        {
            if (state == State::Unknown)
            {
                if (!moveInside(g->extent, g->eps)) return false;
                mr = g->tetCellIndex(rx, ry, rz);
                enteringFace = -1;
                if (mr == -1) return false;
                if (ds > 0.) return true;
            }
            if (state == State::Inside)
            {
                while (true)
                {
                    if (!(std::isfinite(rx) && std::isfinite(ry) && std::isfinite(rz)))
                    {
                        state = State::Outside;
                        return false;
                    }
                    int leavingFace = -1;
                    double dsl = DBL_MAX;
                    if (enteringFace == -1) enteringFace = findEnteringFace(mr);
                    if (enteringFace == -1)
                    {
                        state = State::Outside;
                        return false;
                    }
                    const double* v = g->tetVertex(mr, enteringFace);
                    double wx = rx - v[0], wy = ry - v[1], wz = rz - v[2];
                    double mx = ky * wz - kz * wy, my = kz * wx - kx * wz, mz = kx * wy - ky * wx;
                    int cv[3];
                    clockwiseVertices(enteringFace, cv);
                    const double* c0 = g->tetVertex(mr, cv[0]);
                    double prod0 = mx * (v[0] - c0[0]) + my * (v[1] - c0[1]) + mz * (v[2] - c0[2]);
                    int clock0 = prod0 < 0;
                    int i = clock0 ? 1 : 2;
                    const double* ci = g->tetVertex(mr, cv[i]);
                    double prodi = mx * (v[0] - ci[0]) + my * (v[1] - ci[1]) + mz * (v[2] - ci[2]);
                    int cclocki = prodi >= 0;
                    if (prod0 == 0. || prodi == 0.)
                    {
                        for (int q = 0; q < 3; ++q)
                        {
                            int face = cv[q];
                            const double* n = g->tetNormal(mr, face);
                            double ndotk = n[0] * kx + n[1] * ky + n[2] * kz;
                            if (ndotk > 0)
                            {
                                double dq = (n[0] * (v[0] - rx) + n[1] * (v[1] - ry) + n[2] * (v[2] - rz)) / ndotk;
                                if (dq < dsl)
                                {
                                    dsl = dq;
                                    leavingFace = face;
                                }
                            }
                        }
                    }
                    else
                    {
                        static constexpr int dtable[2][2] = {{1, 0}, {0, 2}};
                        leavingFace = cv[dtable[clock0][cclocki]];
                        const double* n = g->tetNormal(mr, leavingFace);
                        double ndotk = n[0] * kx + n[1] * ky + n[2] * kz;
                        dsl = (n[0] * (v[0] - rx) + n[1] * (v[1] - ry) + n[2] * (v[2] - rz)) / ndotk;
                    }
                    if (leavingFace == -1 || dsl < g->eps)
                    {
                        propagater(g->eps);
                        mr = g->tetCellIndex(rx, ry, rz);
                        if (mr < 0)
                        {
                            state = State::Outside;
                            return false;
                        }
                        enteringFace = -1;
                    }
                    else
                    {
                        propagater(dsl);
                        m = mr;
                        ds = dsl;
                        int64_t fi = 4 * static_cast<int64_t>(mr) + leavingFace;
                        mr = g->faceNb[fi];
                        if (mr < 0)
                        {
                            state = State::Outside;
                            return false;
                        }
                        enteringFace = g->faceNbFace[fi];
                        return true;
                    }
                }
            }
            return false;
        }

        // CartesianSpatialGrid.cpp:77-153
        bool nextCartesian()
        {
            switch (state)
            {
                case State::Unknown:
                {
                    if (!moveInside(g->extent, 1e-12 * g->extent.diagonal())) return false;
                    i = locateClip(g->xv, g->Nx + 1, rx);
                    j = locateClip(g->yv, g->Ny + 1, ry);
                    k = locateClip(g->zv, g->Nz + 1, rz);
                    if (ds > 0.) return true;
                }
                // fall through
                case State::Inside:
                {
                    int mm = g->index(i, j, k);
                    double xE = (kx < 0.0) ? g->xv[i] : g->xv[i + 1];
                    double yE = (ky < 0.0) ? g->yv[j] : g->yv[j + 1];
                    double zE = (kz < 0.0) ? g->zv[k] : g->zv[k + 1];
                    double dsx = (std::fabs(kx) > 1e-15) ? (xE - rx) / kx : DBL_MAX;
                    double dsy = (std::fabs(ky) > 1e-15) ? (yE - ry) / ky : DBL_MAX;
                    double dsz = (std::fabs(kz) > 1e-15) ? (zE - rz) / kz : DBL_MAX;
                    if (dsx <= dsy && dsx <= dsz)
                    {
                        m = mm;
                        ds = dsx;
                        rx = xE;
                        ry += ky * dsx;
                        rz += kz * dsx;
                        i += (kx < 0.0) ? -1 : 1;
                        if (i >= g->Nx || i < 0) state = State::Outside;
                    }
                    else if (dsy < dsx && dsy <= dsz)
                    {
                        m = mm;
                        ds = dsy;
                        ry = yE;
                        rx += kx * dsy;
                        rz += kz * dsy;
                        j += (ky < 0.0) ? -1 : 1;
                        if (j >= g->Ny || j < 0) state = State::Outside;
                    }
                    else
                    {
                        m = mm;
                        ds = dsz;
                        rz = zE;
                        rx += kx * dsz;
                        ry += ky * dsz;
                        k += (kz < 0.0) ? -1 : 1;
                        if (k >= g->Nz || k < 0) state = State::Outside;
                    }
                    return true;
                }
                case State::Outside: break;
            }
            return false;
        }

        // TreeSpatialGrid.cpp:143-228
        bool nextTree()
        {
            switch (state)
            {
                case State::Unknown:
                {
                    if (!moveInside(g->extent, g->eps)) return false;
                    node = g->leafChild(0, rx, ry, rz);
                    if (ds > 0.) return true;
                }
                // fall through
                case State::Inside:
                {
                    const double* b = g->nodeBounds(node);
                    double xnext = (kx < 0.0) ? b[0] : b[3];
                    double ynext = (ky < 0.0) ? b[1] : b[4];
                    double znext = (kz < 0.0) ? b[2] : b[5];
                    double dsx = (std::fabs(kx) > 1e-15) ? (xnext - rx) / kx : DBL_MAX;
                    double dsy = (std::fabs(ky) > 1e-15) ? (ynext - ry) / ky : DBL_MAX;
                    double dsz = (std::fabs(kz) > 1e-15) ? (znext - rz) / kz : DBL_MAX;
                    double d;
                    int wall;  // BACK=0, FRONT, LEFT, RIGHT, BOTTOM, TOP (TreeNode.hpp:104)
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
                    propagater(d + g->eps);
                    m = g->cellOf[node];
                    ds = d;
                    int old = node;
                    node = g->neighbor(node, wall, rx, ry, rz);
                    if (node < 0) node = g->leafChild(0, rx, ry, rz);
                    if (node == old)
                    {
                        propagateToNextAfter();
                        node = g->leafChild(0, rx, ry, rz);
                    }
                    if (node < 0 || node == old) state = State::Outside;
                    return true;
                }
                case State::Outside: break;
            }
            return false;
        }
    };

    // ---------------------------------------------------------------- HG phase function
    constexpr double kGLarge = 0.95;                  // DustMix.cpp:36-39
    constexpr double kDelta = 4. * M_PI / 180.;

    // DustMix.cpp:395-425
    double valueHG(double g, double c)
    {
        double t = 1. + g * g - 2. * g * c;
        return (1. - g) * (1. + g) / std::sqrt(t * t * t);
    }
    double integralHG(double g, double ca, double cb)
    {
        double ta = std::sqrt(1. + g * g - 2. * g * ca);
        double tb = std::sqrt(1. + g * g - 2. * g * cb);
        double f1 = (1. - g) * (1. + g) / g;
        double f2 = (tb - ta) / (tb * ta);
        return f1 * f2;
    }
    double meanHG(double g, double c)
    {
        double theta = std::acos(c);
        double ca = std::cos(theta - kDelta);
        double cb = std::cos(theta + kDelta);
        if (theta < kDelta) return (integralHG(g, 1., ca) + integralHG(g, 1., cb)) / (2. - ca - cb);
        if (theta > M_PI - kDelta) return (integralHG(g, ca, -1.) + integralHG(g, cb, -1.)) / (2. + ca + cb);
        return integralHG(g, ca, cb) / (ca - cb);
    }
    // DustMix.cpp:430-445
    double peelPhase(double g, double c) { return std::fabs(g) > kGLarge ? meanHG(g, c) : valueHG(g, c); }

    // Direction(theta, phi), Direction.cpp:11-36
    void directionFromAngles(double theta, double phi, double* k)
    {
        const double eps = 1e-8;
        if (theta <= eps)
        {
            k[0] = 0;
            k[1] = 0;
            k[2] = 1;
        }
        else if (theta >= M_PI - eps)
        {
            k[0] = 0;
            k[1] = 0;
            k[2] = -1;
        }
        else
        {
            double st = std::sin(theta);
            k[0] = st * std::cos(phi);
            k[1] = st * std::sin(phi);
            k[2] = std::cos(theta);
        }
    }
    // Random::direction(), Random.cpp:121-126
    void isotropic(Stream& rng, double* k)
    {
        double theta = std::acos(2.0 * rng.uniform() - 1.0);
        double phi = 2.0 * M_PI * rng.uniform();
        directionFromAngles(theta, phi, k);
    }
    // Random::direction(bfk, costheta), Random.cpp:130-164
    void rotate(Stream& rng, const double* k, double costheta, double* kn)
    {
        double phi = 2.0 * M_PI * rng.uniform();
        double cosphi = std::cos(phi);
        double sinphi = std::sin(phi);
        double sintheta = std::sqrt(std::fabs((1.0 - costheta) * (1.0 + costheta)));
        double kx = k[0], ky = k[1], kz = k[2];
        if (kz > 0.99999)
        {
            kn[0] = cosphi * sintheta;
            kn[1] = sinphi * sintheta;
            kn[2] = costheta;
        }
        else if (kz < -0.99999)
        {
            kn[0] = cosphi * sintheta;
            kn[1] = sinphi * sintheta;
            kn[2] = -costheta;
        }
        else
        {
            double root = std::sqrt((1.0 - kz) * (1.0 + kz));
            kn[0] = sintheta / root * (-kx * kz * cosphi + ky * sinphi) + kx * costheta;
            kn[1] = -sintheta / root * (ky * kz * cosphi + kx * sinphi) + ky * costheta;
            kn[2] = root * sintheta * cosphi + kz * costheta;
        }
    }
    // DustMix::performScattering HG branch, DustMix.cpp:496-512
    void scatterHG(Stream& rng, double g, const double* k, double* kn)
    {
        if (std::fabs(g) < 1e-6)
            isotropic(rng, kn);
        else
        {
            double f = ((1.0 - g) * (1.0 + g)) / (1.0 - g + 2.0 * g * rng.uniform());
            double costheta = (1.0 + g * g - f * f) / (2.0 * g);
            rotate(rng, k, costheta, kn);
        }
    }
    // Random::exponCutoff, Random.cpp:105-117
    double exponCutoff(Stream& rng, double xmax)
    {
        if (xmax == 0.0) return 0.0;
        if (xmax < 1e-10) return rng.uniform() * xmax;
        double x = -std::log(1.0 - rng.uniform() * (1.0 - std::exp(-xmax)));
        while (x > xmax) x = -std::log(1.0 - rng.uniform() * (1.0 - std::exp(-xmax)));
        return x;
    }
}

// ==================================================================== C interface
extern "C" {

struct orc_grid
{
    int32_t type;  // 0 Cartesian, 1 octree, 2 Voronoi, 3 tetra
    int32_t nx, ny, nz;
    const double *xb, *yb, *zb;  // Cartesian borders (nx+1, ny+1, nz+1)
    int64_t nnodes;
    const double* bounds;        // octree node bounds, 6 per node (xmin,ymin,zmin,xmax,ymax,zmax)
    const int32_t* child_first;  // first child id (children are consecutive) or -1
    const int32_t* cell_index;   // cell index per node, -1 for non-leaves
    const int64_t* nb_off;       // 6*nnodes+1 offsets into nb_id
    const int32_t* nb_id;        // neighbour node ids, per node per wall, sorted
    double box[6];               // domain extent
    // Voronoi (type 2)
    const double* sites;         // 3 per site
    const int32_t* vnb_begin;
    const int32_t* vnb_count;
    const int32_t* vnb_index;    // neighbour site ids, walls as -1..-6
    int32_t block_n;
    const int32_t* block_begin;  // block_n^3
    const int32_t* block_count;
    const int32_t* block_index;
    // tetra (type 3)
    const double* verts;         // 3 per vertex
    int64_t ntets;
    const int32_t* tet_verts;    // 4 per tetra
    const int32_t* face_nb;      // 4 per tetra, -1 on the hull
    const double* face_normal;   // 12 per tetra
};

struct orc_instrument
{
    double inclination, azimuth, roll;  // radians
    int32_t nx, ny;
    double fovx, fovy, cx, cy;          // metres
    int32_t sed;                        // SEDInstrument (nx = ny = 1) with an aperture radius (m)
    double radius;
    int32_t persp;                      // PerspectiveInstrument: transform (4x4 row-major,
    double M[16];                       // row vector times M), eye and pixel size
    double eye[3];
    double psize;
};

struct orc_sim
{
    const orc_grid* grid;
    const double* density;  // number density per cell
    int32_t nlambda;
    const double* lambda;   // wavelength per bin (m)
    const double* sext;     // extinction cross section per bin
    const double* ssca;     // scattering cross section per bin
    const double* g;        // HG asymmetry parameter per bin
    const double* lpack;    // launch luminosity per packet per bin
    double src[3];
    double xi;              // path length bias
    double min_weight_reduction;
    int32_t min_scatt_events;
    int32_t ninstr;
    const orc_instrument* instr;
    uint64_t seed;
    int64_t ncells;         // cells of the grid (radiation-field table rows)
    double* rf;             // radiation field ncells x nlambda to add into, or NULL
    int64_t nent;           // imported particles (0: the point source at src)
    int32_t ent_kernel;     // 0 uniform, 1 cubic spline
    const double* ent;      // per particle: x, y, z, smoothing length
    const int64_t* ent_first;  // first history index of each particle (nent + 1)
    const double* ent_lpack;   // launch luminosity per packet, nent x nlambda
    const double* ent_X;       // cubic spline: cumulative mass at u = k / 400, k = 0..400
};

// version tag used by the loader
int orc_version() { return 6; }

// the logarithmic mean of the radiation-field storage, exposed for the known-answer tests
double orc_lnmean(double x1, double x2);

// Philox4x32-10 block, exposed for the known-answer tests
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
    U4 r = philox(U4{{ctr[0], ctr[1], ctr[2], ctr[3]}}, key[0], key[1]);
    std::memcpy(out, r.v, 16);
}

void orc_uniforms(uint64_t seed, uint64_t history, uint32_t tag, int64_t n, double* out)
{
    Stream s(seed, history, tag);
    for (int64_t i = 0; i < n; ++i) out[i] = s.uniform();
}

static Grid makeGrid(const orc_grid* og)
{
    Grid g;
    g.type = og->type;
    g.extent = Box{og->box[0], og->box[1], og->box[2], og->box[3], og->box[4], og->box[5]};
    // TreeSpatialGrid.cpp:28, VoronoiMeshSnapshot.cpp:396, TetraMeshSpatialGrid.cpp:239
    g.eps = 1e-12 * g.extent.diagonal();
    g.xv = og->xb;
    g.yv = og->yb;
    g.zv = og->zb;
    g.Nx = og->nx;
    g.Ny = og->ny;
    g.Nz = og->nz;
    g.nb = og->bounds;
    g.childFirst = og->child_first;
    g.cellOf = og->cell_index;
    g.nbOff = og->nb_off;
    g.nbId = og->nb_id;
    if (g.type == 2)
    {
        g.sites = og->sites;
        g.vnbBegin = og->vnb_begin;
        g.vnbCount = og->vnb_count;
        g.vnbIndex = og->vnb_index;
        g.vbn = og->block_n;
        g.vblBegin = og->block_begin;
        g.vblCount = og->block_count;
        g.vblIndex = og->block_index;
    }
    if (g.type == 3)
    {
        g.verts = og->verts;
        g.ntets = og->ntets;
        g.tetVerts = og->tet_verts;
        g.faceNb = og->face_nb;
        g.faceNormal = og->face_normal;
        g.buildTetraSearch();
    }
    return g;
}

// Raw generator output (m, ds) for each ray, including empty segments, as produced by
// SpatialGrid::createPathSegmentGenerator()->next().  Returns the total number of segments,
// or -1 if cap is exceeded.
int64_t orc_trace(const orc_grid* og, const double* rays, int64_t n, int64_t cap, int64_t* off, int32_t* mout,
                  double* dsout)
{
    Grid g = makeGrid(og);
    Generator gen(&g);
    int64_t t = 0;
    off[0] = 0;
    for (int64_t i = 0; i < n; ++i)
    {
        gen.start(rays + 6 * i, rays + 6 * i + 3);
        while (gen.next())
        {
            if (t >= cap) return -1;
            mout[t] = gen.m;
            dsout[t] = gen.ds;
            ++t;
        }
        off[i + 1] = t;
    }
    return t;
}

// Cumulative extinction optical depth per stored segment (MediumSystem.cpp:1785-1810 with
// SpatialGridPath::addSegment dropping ds<=0, SpatialGridPath.cpp:37-45) and the observer
// optical depth of MediumSystem::getExtinctionOpticalDepth(pp, inf) (MediumSystem.cpp:2328-2376)
// for luminosity L (early-out at log(L)+745).
int64_t orc_tau(const orc_grid* og, const double* density, double section, const double* rays, int64_t n,
                double L, int64_t cap, int64_t* toff, double* tau, double* s, int32_t* mseg, double* tauobs)
{
    Grid g = makeGrid(og);
    Generator gen(&g);
    int64_t t = 0;
    toff[0] = 0;
    double taumax = std::log(L) + 745;
    for (int64_t i = 0; i < n; ++i)
    {
        gen.start(rays + 6 * i, rays + 6 * i + 3);
        double tc = 0., sc = 0.;
        while (gen.next())
        {
            if (gen.ds > 0.)
            {
                if (t >= cap) return -1;
                sc += gen.ds;
                if (gen.m >= 0) tc += section * density[gen.m] * gen.ds;
                tau[t] = tc;
                s[t] = sc;
                mseg[t] = gen.m;
                ++t;
            }
        }
        toff[i + 1] = t;
        gen.start(rays + 6 * i, rays + 6 * i + 3);
        double to = 0.;
        while (gen.next())
        {
            if (gen.m >= 0)
            {
                to += section * density[gen.m] * gen.ds;
                if (to >= taumax)
                {
                    to = INFINITY;
                    break;
                }
            }
        }
        tauobs[i] = to;
    }
    return t;
}

// SpatialGrid::cellIndex for each point (Voronoi and tetra grids; -1 outside)
void orc_cell_index(const orc_grid* og, const double* pts, int64_t n, int32_t* out)
{
    Grid g = makeGrid(og);
    for (int64_t i = 0; i < n; ++i)
    {
        const double* p = pts + 3 * i;
        out[i] = g.type == 2 ? g.vorCellIndex(p[0], p[1], p[2]) : g.type == 3 ? g.tetCellIndex(p[0], p[1], p[2]) : -1;
    }
}

// HG helpers for the tests
double orc_phase_hg(double g, double c) { return peelPhase(g, c); }

void orc_scatter_hg(uint64_t seed, uint64_t history, double g, const double* k, int64_t n, double* out)
{
    Stream rng(seed, history, 0);
    for (int64_t i = 0; i < n; ++i) scatterHG(rng, g, k, out + 3 * i);
}

double orc_expon_cutoff(uint64_t seed, uint64_t history, double xmax, int64_t n, double* out)
{
    Stream rng(seed, history, 0);
    for (int64_t i = 0; i < n; ++i) out[i] = exponCutoff(rng, xmax);
    return 0.;
}

// -------------------------------------------------------------------- the life cycle
namespace
{
    struct Frame
    {
        // FrameInstrument.cpp:12-33 (setup) and :117-136 (pixelOnDetector)
        double costheta, sintheta, cosphi, sinphi, cosomega, sinomega;
        int Nxp, Nyp;
        double xpmin, xpsiz, ypmin, ypsiz;
        double kobs[3];
        bool sameAsPrev;
        bool sed;
        double radius2;
        bool persp;
        double M[16], eye[3], psize;

        // PerspectiveInstrument::detect (PerspectiveInstrument.cpp:103-116) with
        // HomogeneousTransform::transform: pixel (or -1) and depth zp
        int perspPixel(const double* r, double& zp) const
        {
            double x = r[0], y = r[1], z = r[2];
            double xp = x * M[0] + y * M[4] + z * M[8] + 1. * M[12];
            double yp = x * M[1] + y * M[5] + z * M[9] + 1. * M[13];
            zp = x * M[2] + y * M[6] + z * M[10] + 1. * M[14];
            double wp = x * M[3] + y * M[7] + z * M[11] + 1. * M[15];
            int i = static_cast<int>(xp / wp);
            int j = static_cast<int>(yp / wp);
            if (i >= 0 && i < Nxp && j >= 0 && j < Nyp && zp > psize / 10.) return i + Nxp * j;
            return -1;
        }

        int pixel(const double* r) const
        {
            double x = r[0], y = r[1], z = r[2];
            if (sed)
            {
                // SEDInstrument::detect with ApertureInstrument::isInsideAperture (ApertureInstrument.cpp:24-43)
                if (radius2)
                {
                    double xpp = -sinphi * x + cosphi * y;
                    double ypp = -cosphi * costheta * x - sinphi * costheta * y + sintheta * z;
                    if (xpp * xpp + ypp * ypp > radius2) return -1;
                }
                return 0;
            }
            double xpp = -sinphi * x + cosphi * y;
            double ypp = -cosphi * costheta * x - sinphi * costheta * y + sintheta * z;
            double xp = cosomega * xpp - sinomega * ypp;
            double yp = sinomega * xpp + cosomega * ypp;
            int i = static_cast<int>(std::floor((xp - xpmin) / xpsiz));
            int j = static_cast<int>(std::floor((yp - ypmin) / ypsiz));
            if (i < 0 || i >= Nxp || j < 0 || j >= Nyp) return -1;
            return i + Nxp * j;
        }
    };

    struct Seg
    {
        int m;
        double s, tau, ds;
    };

    // SpecialFunctions::lnmean(x1, x2, lnx1, lnx2), SpecialFunctions.cpp:860-880
    double lnMean(double x1, double x2, double lnx1, double lnx2)
    {
        if (x1 > x2)
        {
            std::swap(x1, x2);
            std::swap(lnx1, lnx2);
        }
        if (x1 <= 0) return 0.;
        double x = x2 / x1 - 1.;
        if (x < 1e-3)
            return x1 / (1. - 1. / 2. * x + 1. / 3. * x * x - 1. / 4. * x * x * x + 1. / 5. * x * x * x * x
                         - 1. / 6. * x * x * x * x * x);
        return (x2 - x1) / (lnx2 - lnx1);
    }

    struct Sim
    {
        Grid grid;
        const orc_sim* p;
        std::vector<Frame> frames;
        int64_t npix_total = 0;
        std::vector<int64_t> frameOff;

        // MediumSystem::getExtinctionOpticalDepth(pp, distance) for a distant observer
        double observerTau(const double* r, const double* k, double section, double L, Generator& gen,
                           double distance = INFINITY) const
        {
            if (L <= 0) return INFINITY;
            double taumax = std::log(L) + 745;
            gen.start(r, k);
            double tau = 0.;
            double s = 0.;
            while (gen.next())
            {
                if (gen.m >= 0)
                {
                    tau += section * p->density[gen.m] * gen.ds;
                    if (tau >= taumax) return INFINITY;
                }
                s += gen.ds;
                if (s > distance) break;
            }
            return tau;
        }

        // peel-off toward every instrument: MonteCarloSimulation.cpp:698-715 (emission) and
        // :2034-2092 (scattering), FrameInstrument::detect, FluxRecorder::detect (:372-418)
        void peel(const double* r, const double* k, double W, int ell, bool scattering, double* frame,
                  Generator& gen) const
        {
            double lambda = p->lambda[ell];
            double section = p->sext[ell];
            double tauCached = 0.;
            bool haveTau = false;
            double Wp = W;
            for (size_t f = 0; f < frames.size(); ++f)
            {
                const Frame& fr = frames[f];
                if (fr.persp)
                {
                    // PerspectiveInstrument::bfkobs (:93-101) and FluxRecorder::detect for a
                    // local instrument (FluxRecorder.cpp:372-405)
                    double zp;
                    int l = fr.perspPixel(r, zp);
                    double d[3] = {fr.eye[0] - r[0], fr.eye[1] - r[1], fr.eye[2] - r[2]};
                    double D = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    double ko[3] = {d[0] / D, d[1] / D, d[2] / D};
                    if (!fr.sameAsPrev)
                    {
                        haveTau = false;
                        Wp = W;
                        if (scattering) Wp = W * peelPhase(p->g[ell], k[0] * ko[0] + k[1] * ko[1] + k[2] * ko[2]);
                    }
                    if (l < 0) continue;
                    double L = Wp / lambda;
                    if (!haveTau)
                    {
                        tauCached = observerTau(r, ko, section, L, gen, zp);
                        haveTau = true;
                    }
                    frame[frameOff[f] + l + static_cast<int64_t>(ell) * fr.Nxp * fr.Nyp] += L / (zp * zp) * std::exp(-tauCached);
                    continue;
                }
                if (!fr.sameAsPrev)
                {
                    haveTau = false;
                    Wp = W;
                    if (scattering)
                    {
                        double c = k[0] * fr.kobs[0] + k[1] * fr.kobs[1] + k[2] * fr.kobs[2];
                        Wp = W * peelPhase(p->g[ell], c);
                    }
                }
                int l = fr.pixel(r);
                if (l < 0) continue;
                double L = Wp / lambda;
                if (!haveTau)
                {
                    tauCached = observerTau(r, fr.kobs, section, L, gen);
                    haveTau = true;
                }
                frame[frameOff[f] + l + static_cast<int64_t>(ell) * fr.Nxp * fr.Nyp] += L * std::exp(-tauCached);
            }
        }

        // MonteCarloSimulation::performLifeCycle forced-scattering branch (:626-665),
        // simulateForcedPropagation (:1492-1544), MediumSystem::simulateScattering (:1607-1653)
        void history(uint64_t h, double* frame, double* rf, std::vector<Seg>& segs, Generator& gen,
                     int64_t* counters) const
        {
            Stream rng(p->seed, h, 0);
            // launch: uniform wavelength bin (OligoWavelengthDistribution.cpp:34-38), isotropic direction
            int ell = static_cast<int>(rng.uniform() * p->nlambda);
            if (ell >= p->nlambda) ell = p->nlambda - 1;
            double lambda = p->lambda[ell];
            double W = p->lpack[ell] * lambda;  // PhotonPacket.cpp:18-22
            double r[3] = {p->src[0], p->src[1], p->src[2]};
            if (p->nent > 0)
            {
                // ImportedSource::launch (ImportedSource.cpp:466-471): the particle whose history
                // range holds h, as std::upper_bound(Iv, h) - 1
                const int64_t m = std::upper_bound(p->ent_first, p->ent_first + p->nent + 1, static_cast<int64_t>(h))
                                  - p->ent_first - 1;
                W = p->ent_lpack[m * p->nlambda + ell] * lambda;
                // ParticleSnapshot::generatePosition (ParticleSnapshot.cpp:275-286) with the
                // kernel's generateRadius (CubicSplineSmoothingKernel.cpp:78-85,
                // UniformSmoothingKernel.cpp:27-31)
                const double X = rng.uniform();
                double u;
                if (p->ent_kernel == 1)
                {
                    int kk = static_cast<int>(std::upper_bound(p->ent_X, p->ent_X + 401, X) - p->ent_X) - 1;
                    kk = std::max(0, std::min(399, kk));  // NR::locateClip
                    const double f = (X - p->ent_X[kk]) / (p->ent_X[kk + 1] - p->ent_X[kk]);
                    u = (kk + f) / 400;
                }
                else
                    u = std::pow(X, 1.0 / 3.0);
                double d[3];
                isotropic(rng, d);
                const double hm = p->ent[4 * m + 3];
                for (int q = 0; q < 3; ++q) r[q] = p->ent[4 * m + q] + d[q] * u * hm;
            }
            double k[3];
            isotropic(rng, k);
            if (!(W / lambda > 0)) return;
            peel(r, k, W, ell, false, frame, gen);
            counters[1]++;

            double Lthreshold = (W / lambda) / p->min_weight_reduction;
            int nscatt = 0;
            double section = p->sext[ell];
            double xi = p->xi;
            while (true)
            {
                // setExtinctionOpticalDepths: complete path and cumulative tau
                segs.clear();
                gen.start(r, k);
                double tau = 0., s = 0.;
                while (gen.next())
                {
                    if (gen.ds > 0.)
                    {
                        s += gen.ds;
                        if (gen.m >= 0) tau += section * p->density[gen.m] * gen.ds;
                        segs.push_back(Seg{gen.m, s, tau, gen.ds});
                    }
                }
                counters[2]++;
                // MonteCarloSimulation::storeRadiationField (:888-927), oligochromatic branch:
                // the bin of the radiation-field grid is the packet's bin
                if (rf)
                {
                    const double L = W / lambda;
                    double lnExtBeg = 0., extBeg = 1.;
                    for (const Seg& sg : segs)
                    {
                        double lnExtEnd = -sg.tau;
                        double extEnd = std::exp(lnExtEnd);
                        if (sg.m >= 0)
                            rf[static_cast<int64_t>(sg.m) * p->nlambda + ell] +=
                                L * lnMean(extEnd, extBeg, lnExtEnd, lnExtBeg) * sg.ds;
                        lnExtBeg = lnExtEnd;
                        extBeg = extEnd;
                    }
                }
                double taupath = segs.empty() ? 0. : segs.back().tau;
                if (taupath <= 0.)
                {
                    W *= 0.;
                    break;
                }
                double tint;
                if (xi == 0.)
                    tint = exponCutoff(rng, taupath);
                else
                {
                    tint = rng.uniform() < xi ? rng.uniform() * taupath : exponCutoff(rng, taupath);
                    double pp = -std::exp(-tint) / std::expm1(-taupath);
                    double q = (1.0 - xi) * pp + xi / taupath;
                    W *= pp / q;
                }
                // SpatialGridPath::findInteractionPoint, SpatialGridPath.cpp:161-203
                auto it = std::upper_bound(segs.begin(), segs.end(), tint,
                                           [](double t, const Seg& sg) { return t < sg.tau; });
                int m;
                double sint;
                if (it == segs.begin())
                {
                    m = it->m;
                    sint = 0. + ((tint - 0.) / (it->tau - 0.)) * (it->s - 0.);
                }
                else if (it != segs.end())
                {
                    m = it->m;
                    sint = (it - 1)->s + ((tint - (it - 1)->tau) / (it->tau - (it - 1)->tau)) * (it->s - (it - 1)->s);
                }
                else
                {
                    m = (it - 1)->m;
                    sint = (it - 1)->s;
                }
                // MediumSystem::albedoForScattering, MediumSystem.cpp:793-836 (kext>0 ? ksca/kext : 0)
                double n = m >= 0 ? p->density[m] : 0.;
                double ksca = n > 0. ? n * p->ssca[ell] : 0.;
                double kext = n > 0. ? n * p->sext[ell] : 0.;
                double albedo = kext > 0. ? ksca / kext : 0.;
                W *= -std::expm1(-taupath) * albedo;
                r[0] += sint * k[0];
                r[1] += sint * k[1];
                r[2] += sint * k[2];
                double L = W / lambda;
                if (L <= 0 || (L <= Lthreshold && nscatt >= p->min_scatt_events)) break;
                peel(r, k, W, ell, true, frame, gen);
                counters[3]++;
                double kn[3];
                scatterHG(rng, p->g[ell], k, kn);
                k[0] = kn[0];
                k[1] = kn[1];
                k[2] = kn[2];
                nscatt++;
            }
            counters[0]++;
        }
    };
}

double orc_lnmean(double x1, double x2) { return lnMean(x1, x2, std::log(x1), std::log(x2)); }

int64_t orc_frame_size(const orc_sim* p)
{
    int64_t t = 0;
    for (int f = 0; f < p->ninstr; ++f) t += static_cast<int64_t>(p->instr[f].nx) * p->instr[f].ny * p->nlambda;
    return t;
}

// Simulates histories [first, first+count) with `threads` worker threads and adds the
// recorded luminosities into frame (per instrument: nlambda x ny x nx, row index
// l + ell*nx*ny with l = i + nx*j, as FluxRecorder.cpp:501).  counters (optional, 4 entries):
// histories, emission peel-offs, forced paths, scattering events.
int orc_simulate(const orc_sim* p, uint64_t first, uint64_t count, int threads, double* frame, int64_t* counters)
{
    Sim sim;
    sim.grid = makeGrid(p->grid);
    sim.p = p;
    for (int f = 0; f < p->ninstr; ++f)
    {
        const orc_instrument& in = p->instr[f];
        Frame fr;
        fr.costheta = std::cos(in.inclination);
        fr.sintheta = std::sin(in.inclination);
        fr.cosphi = std::cos(in.azimuth);
        fr.sinphi = std::sin(in.azimuth);
        fr.cosomega = std::cos(in.roll);
        fr.sinomega = std::sin(in.roll);
        fr.Nxp = in.nx;
        fr.Nyp = in.ny;
        fr.xpmin = in.cx - 0.5 * in.fovx;
        fr.xpsiz = in.fovx / in.nx;
        fr.ypmin = in.cy - 0.5 * in.fovy;
        fr.ypsiz = in.fovy / in.ny;
        directionFromAngles(in.inclination, in.azimuth, fr.kobs);  // DistantInstrument.cpp:47
        // DistantInstrument.cpp:55-63
        fr.sed = in.sed != 0;
        fr.radius2 = in.radius * in.radius;
        fr.persp = in.persp != 0;
        std::memcpy(fr.M, in.M, sizeof(fr.M));
        std::memcpy(fr.eye, in.eye, sizeof(fr.eye));
        fr.psize = in.psize;
        if (f > 0 && (in.persp || p->instr[f - 1].persp))
        {
            const orc_instrument& q = p->instr[f - 1];
            fr.sameAsPrev = in.persp && q.persp && std::memcmp(in.M, q.M, sizeof(in.M)) == 0
                            && std::memcmp(in.eye, q.eye, sizeof(in.eye)) == 0 && in.psize == q.psize;
            sim.frameOff.push_back(sim.npix_total);
            sim.npix_total += static_cast<int64_t>(in.nx) * in.ny * p->nlambda;
            sim.frames.push_back(fr);
            continue;
        }
        fr.sameAsPrev = f > 0 && in.inclination == p->instr[f - 1].inclination
                        && in.azimuth == p->instr[f - 1].azimuth && in.roll == p->instr[f - 1].roll;
        sim.frameOff.push_back(sim.npix_total);
        sim.npix_total += static_cast<int64_t>(in.nx) * in.ny * p->nlambda;
        sim.frames.push_back(fr);
    }
    if (threads < 1) threads = 1;
    std::vector<std::vector<double>> frames(threads, std::vector<double>(sim.npix_total, 0.));
    const int64_t nrf = p->rf ? p->ncells * p->nlambda : 0;
    std::vector<std::vector<double>> rfs(threads, std::vector<double>(nrf, 0.));
    std::vector<std::array<int64_t, 4>> cnt(threads, std::array<int64_t, 4>{0, 0, 0, 0});
    std::atomic<uint64_t> next{0};
    const uint64_t chunk = 256;
    auto work = [&](int t) {
        Generator gen(&sim.grid);
        std::vector<Seg> segs;
        segs.reserve(1024);
        while (true)
        {
            uint64_t b = next.fetch_add(chunk);
            if (b >= count) break;
            uint64_t e = std::min(count, b + chunk);
            for (uint64_t i = b; i < e; ++i)
                sim.history(first + i, frames[t].data(), nrf ? rfs[t].data() : nullptr, segs, gen, cnt[t].data());
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (int t = 0; t < threads; ++t)
        for (int64_t i = 0; i < sim.npix_total; ++i) frame[i] += frames[t][i];
    for (int t = 0; t < threads; ++t)
        for (int64_t i = 0; i < nrf; ++i) p->rf[i] += rfs[t][i];
    if (counters)
        for (int t = 0; t < threads; ++t)
            for (int c = 0; c < 4; ++c) counters[c] += cnt[t][c];
    return 0;
}

}  // extern "C"
