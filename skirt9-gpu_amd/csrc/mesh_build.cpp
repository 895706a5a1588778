// Host-side construction of Voronoi grids: the tessellation's neighbour lists, cell extents
// and the block search lists that VoronoiMeshSpatialGrid hands to the traversal
// (VoronoiMeshSnapshot.cpp:491-751 builds them with Voro++, :783-831 the block lists).
//
// Each cell is computed on its own as the domain box clipped by the bisector half-spaces of
// nearby sites, taken in order of increasing distance from a bucket grid; the clipping stops
// once no unvisited site can be closer than twice the cell's farthest vertex (the same
// "security radius" argument Voro++ uses).  The faces that remain give the neighbour list
// (neighbour site ids, domain walls as -1..-6 = xmin, xmax, ymin, ymax, zmin, zmax), and the
// vertices give the cell's bounding box for the block lists.
//
// Site preparation follows VoronoiMeshSnapshot::buildMesh (without relaxation): sites
// outside the domain are dropped, the rest are sorted by x and a site closer than eps to a
// later one is dropped, so that cell indices are positions in the sorted list.

#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "disk_model.hpp"
#include "../../include/skirt_mi355x.h"
#include "philox.hpp"

struct skg_voronoi
{
    double box[6];
    double eps;
    std::vector<double> sites;  // 3 per cell
    std::vector<int32_t> nbBegin, nbCount, nbIndex;
    std::vector<double> lo, hi;  // cell extents, 3 per cell
    std::vector<double> volume;
    int32_t blockN = 0;
    std::vector<int32_t> blockBegin, blockCount, blockIndex;
};

namespace
{
    template<typename F> void parallelChunks(int64_t n, int threads, F&& f)
    {
        threads = std::max(1, threads);
        std::atomic<int64_t> next{0};
        const int64_t chunk = 256;
        auto work = [&]() {
            while (true)
            {
                int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                f(b, std::min(n, b + chunk));
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < threads; ++t) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
    }

    // Convex polyhedron as vertices plus faces (vertex cycles, each with the id of its plane).
    struct Poly
    {
        std::vector<double> v;  // 3 per vertex
        struct Face
        {
            int id;
            std::vector<int> cyc;
        };
        std::vector<Face> faces;

        void box(const double* b)
        {
            v.clear();
            for (int i = 0; i < 8; ++i)
            {
                v.push_back((i & 1) ? b[3] : b[0]);
                v.push_back((i & 2) ? b[4] : b[1]);
                v.push_back((i & 4) ? b[5] : b[2]);
            }
            // corner i = x bit 1, y bit 2, z bit 4; cycles in any consistent order
            faces = {{-1, {0, 2, 6, 4}}, {-2, {1, 5, 7, 3}}, {-3, {0, 4, 5, 1}},
                     {-4, {2, 3, 7, 6}}, {-5, {0, 1, 3, 2}}, {-6, {4, 6, 7, 5}}};
        }

        double maxDist2(const double* p) const
        {
            double r = 0;
            for (size_t i = 0; i < v.size(); i += 3)
            {
                double dx = v[i] - p[0], dy = v[i + 1] - p[1], dz = v[i + 2] - p[2];
                r = std::max(r, dx * dx + dy * dy + dz * dz);
            }
            return r;
        }

        // Keeps the part with n.x <= c; the new face gets plane id `id`.  tol is the distance
        // below which a vertex counts as on the plane.  Returns true if the plane cut the cell.
        bool clip(const double* n, double c, int id, double tol, std::vector<double>& d, std::vector<int>& remap)
        {
            const int nv = static_cast<int>(v.size() / 3);
            const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            d.resize(nv);
            bool any = false;
            for (int i = 0; i < nv; ++i)
            {
                d[i] = (n[0] * v[3 * i] + n[1] * v[3 * i + 1] + n[2] * v[3 * i + 2] - c) / nn;
                if (d[i] > tol) any = true;
            }
            if (!any) return false;
            // new vertex list: kept vertices, then one new vertex per cut edge
            remap.assign(nv, -1);
            std::vector<double> nvtx;
            for (int i = 0; i < nv; ++i)
                if (d[i] <= tol)
                {
                    remap[i] = static_cast<int>(nvtx.size() / 3);
                    nvtx.insert(nvtx.end(), v.begin() + 3 * i, v.begin() + 3 * i + 3);
                }
            std::vector<std::pair<uint64_t, int>> cut;  // edge key -> new vertex
            auto edgeVertex = [&](int a, int b) {
                uint64_t key = (static_cast<uint64_t>(std::min(a, b)) << 32) | static_cast<uint32_t>(std::max(a, b));
                for (auto& e : cut)
                    if (e.first == key) return e.second;
                double t = d[a] / (d[a] - d[b]);
                int idx = static_cast<int>(nvtx.size() / 3);
                for (int q = 0; q < 3; ++q) nvtx.push_back(v[3 * a + q] + t * (v[3 * b + q] - v[3 * a + q]));
                cut.emplace_back(key, idx);
                return idx;
            };
            std::vector<Face> nf;
            std::vector<int> onPlane;
            for (auto& f : faces)
            {
                Face g{f.id, {}};
                const int m = static_cast<int>(f.cyc.size());
                for (int q = 0; q < m; ++q)
                {
                    int a = f.cyc[q], b = f.cyc[(q + 1) % m];
                    bool ina = d[a] <= tol, inb = d[b] <= tol;
                    if (ina) g.cyc.push_back(remap[a]);
                    if (ina != inb)
                    {
                        // crossing: only when one end is strictly outside and the other strictly inside
                        if (d[a] < -tol || d[b] < -tol) g.cyc.push_back(edgeVertex(a, b));
                    }
                }
                // drop repeated consecutive vertices
                std::vector<int> c2;
                for (int x : g.cyc)
                    if (c2.empty() || c2.back() != x) c2.push_back(x);
                while (c2.size() > 1 && c2.front() == c2.back()) c2.pop_back();
                g.cyc.swap(c2);
                if (g.cyc.size() >= 3) nf.push_back(std::move(g));
            }
            // the new face: every kept vertex on the plane and every cut vertex, ordered by angle
            std::vector<int> poly;
            for (int i = 0; i < nv; ++i)
                if (remap[i] >= 0 && d[i] >= -tol) poly.push_back(remap[i]);
            for (auto& e : cut) poly.push_back(e.second);
            std::sort(poly.begin(), poly.end());
            poly.erase(std::unique(poly.begin(), poly.end()), poly.end());
            v.swap(nvtx);
            faces.swap(nf);
            if (poly.size() >= 3)
            {
                double cx = 0, cy = 0, cz = 0;
                for (int i : poly)
                {
                    cx += v[3 * i];
                    cy += v[3 * i + 1];
                    cz += v[3 * i + 2];
                }
                cx /= poly.size();
                cy /= poly.size();
                cz /= poly.size();
                // in-plane basis
                double ux, uy, uz;
                if (std::fabs(n[0]) < 0.9 * nn)
                {
                    ux = 0;
                    uy = n[2];
                    uz = -n[1];
                }
                else
                {
                    ux = -n[2];
                    uy = 0;
                    uz = n[0];
                }
                double un = std::sqrt(ux * ux + uy * uy + uz * uz);
                ux /= un;
                uy /= un;
                uz /= un;
                double wx = (n[1] * uz - n[2] * uy) / nn, wy = (n[2] * ux - n[0] * uz) / nn, wz = (n[0] * uy - n[1] * ux) / nn;
                std::vector<std::pair<double, int>> ang;
                for (int i : poly)
                {
                    double px = v[3 * i] - cx, py = v[3 * i + 1] - cy, pz = v[3 * i + 2] - cz;
                    ang.emplace_back(std::atan2(px * wx + py * wy + pz * wz, px * ux + py * uy + pz * uz), i);
                }
                std::sort(ang.begin(), ang.end());
                Face g{id, {}};
                for (auto& a : ang) g.cyc.push_back(a.second);
                faces.push_back(std::move(g));
            }
            // drop vertices no face uses any more
            std::vector<int> used(v.size() / 3, 0);
            for (auto& f : faces)
                for (int x : f.cyc) used[x] = 1;
            std::vector<int> ren(v.size() / 3, -1);
            std::vector<double> vv;
            for (size_t i = 0; i < used.size(); ++i)
                if (used[i])
                {
                    ren[i] = static_cast<int>(vv.size() / 3);
                    vv.insert(vv.end(), v.begin() + 3 * i, v.begin() + 3 * i + 3);
                }
            for (auto& f : faces)
                for (int& x : f.cyc) x = ren[x];
            v.swap(vv);
            return true;
        }

        // face area (for discarding degenerate faces) and polyhedron volume
        double faceArea(const Face& f) const
        {
            double ax = 0, ay = 0, az = 0;
            const double* p0 = &v[3 * f.cyc[0]];
            for (size_t q = 1; q + 1 < f.cyc.size(); ++q)
            {
                const double* p1 = &v[3 * f.cyc[q]];
                const double* p2 = &v[3 * f.cyc[q + 1]];
                double e1x = p1[0] - p0[0], e1y = p1[1] - p0[1], e1z = p1[2] - p0[2];
                double e2x = p2[0] - p0[0], e2y = p2[1] - p0[1], e2z = p2[2] - p0[2];
                ax += e1y * e2z - e1z * e2y;
                ay += e1z * e2x - e1x * e2z;
                az += e1x * e2y - e1y * e2x;
            }
            return 0.5 * std::sqrt(ax * ax + ay * ay + az * az);
        }

        double volume(const double* p) const
        {
            double vol = 0;
            for (auto& f : faces)
            {
                const double* p0 = &v[3 * f.cyc[0]];
                for (size_t q = 1; q + 1 < f.cyc.size(); ++q)
                {
                    const double* p1 = &v[3 * f.cyc[q]];
                    const double* p2 = &v[3 * f.cyc[q + 1]];
                    double ax = p0[0] - p[0], ay = p0[1] - p[1], az = p0[2] - p[2];
                    double bx = p1[0] - p[0], by = p1[1] - p[1], bz = p1[2] - p[2];
                    double cx = p2[0] - p[0], cy = p2[1] - p[1], cz = p2[2] - p[2];
                    vol += std::fabs(ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx)) / 6.0;
                }
            }
            return vol;
        }
    };
}

extern "C" {

int skg_voronoi_build(int64_t nsites, const double* sites, const double* box, int32_t threads, skg_voronoi** out)
{
    if (!out || nsites < 1 || nsites > INT32_MAX || !sites || !box || !(box[0] < box[3] && box[1] < box[4] && box[2] < box[5]))
        return SKG_ERR_ARGUMENT;
    *out = nullptr;
    try
    {
        auto V = std::make_unique<skg_voronoi>();
        std::memcpy(V->box, box, sizeof(V->box));
        double dx = box[3] - box[0], dy = box[4] - box[1], dz = box[5] - box[2];
        V->eps = 1e-12 * std::sqrt(dx * dx + dy * dy + dz * dz);  // VoronoiMeshSnapshot.cpp:396
        const double eps = V->eps;
        // VoronoiMeshSnapshot::buildMesh: drop sites outside, sort by x, drop a site within eps of a later one
        std::vector<std::array<double, 3>> s;
        s.reserve(nsites);
        for (int64_t i = 0; i < nsites; ++i)
        {
            const double* p = sites + 3 * i;
            if (p[0] >= box[0] && p[0] <= box[3] && p[1] >= box[1] && p[1] <= box[4] && p[2] >= box[2] && p[2] <= box[5])
                s.push_back({p[0], p[1], p[2]});
        }
        std::sort(s.begin(), s.end(), [](const std::array<double, 3>& a, const std::array<double, 3>& b) { return a[0] < b[0]; });
        {
            std::vector<char> drop(s.size(), 0);
            for (size_t m = 0; m < s.size(); ++m)
                for (size_t j = m + 1; j < s.size() && s[j][0] - s[m][0] < eps; ++j)
                {
                    double ex = s[j][0] - s[m][0], ey = s[j][1] - s[m][1], ez = s[j][2] - s[m][2];
                    if (ex * ex + ey * ey + ez * ez < eps * eps)
                    {
                        drop[m] = 1;
                        break;
                    }
                }
            size_t w = 0;
            for (size_t m = 0; m < s.size(); ++m)
                if (!drop[m]) s[w++] = s[m];
            s.resize(w);
        }
        const int64_t n = static_cast<int64_t>(s.size());
        if (n < 1) return SKG_ERR_ARGUMENT;
        V->sites.resize(3 * n);
        for (int64_t i = 0; i < n; ++i)
            for (int q = 0; q < 3; ++q) V->sites[3 * i + q] = s[i][q];

        // bucket grid, about two sites per bucket
        int nbk[3];
        const double ext[3] = {dx, dy, dz};
        const double per = std::cbrt(static_cast<double>(n) / 2.0 / (dx * dy * dz));
        for (int a = 0; a < 3; ++a) nbk[a] = std::max(1, std::min(1024, static_cast<int>(ext[a] * per)));
        auto bucketOf = [&](const double* p, int a) {
            int i = static_cast<int>((p[a] - box[a]) / ext[a] * nbk[a]);
            return std::max(0, std::min(nbk[a] - 1, i));
        };
        const int64_t nb3 = static_cast<int64_t>(nbk[0]) * nbk[1] * nbk[2];
        std::vector<int64_t> boff(nb3 + 1, 0);
        std::vector<int32_t> bid(n);
        std::vector<int64_t> key(n);
        for (int64_t i = 0; i < n; ++i)
        {
            const double* p = &V->sites[3 * i];
            key[i] = (static_cast<int64_t>(bucketOf(p, 0)) * nbk[1] + bucketOf(p, 1)) * nbk[2] + bucketOf(p, 2);
            boff[key[i] + 1]++;
        }
        for (int64_t b = 0; b < nb3; ++b) boff[b + 1] += boff[b];
        {
            std::vector<int64_t> fill(boff.begin(), boff.end() - 1);
            for (int64_t i = 0; i < n; ++i) bid[fill[key[i]]++] = static_cast<int32_t>(i);
        }
        const double bw[3] = {dx / nbk[0], dy / nbk[1], dz / nbk[2]};
        const double hmin = std::min(bw[0], std::min(bw[1], bw[2]));

        std::vector<std::vector<int32_t>> nbs(n);
        V->lo.resize(3 * n);
        V->hi.resize(3 * n);
        V->volume.resize(n);
        std::atomic<int> failed{0};
        parallelChunks(n, threads, [&](int64_t b0, int64_t b1) {
            Poly P;
            std::vector<double> dtmp;
            std::vector<int> rtmp;
            std::vector<std::pair<double, int32_t>> cand;
            for (int64_t i = b0; i < b1; ++i)
            {
                const double* p = &V->sites[3 * i];
                P.box(box);
                int ci[3] = {bucketOf(p, 0), bucketOf(p, 1), bucketOf(p, 2)};
                double r2 = P.maxDist2(p);
                const int maxShell = std::max(nbk[0], std::max(nbk[1], nbk[2]));
                for (int sh = 0; sh <= maxShell; ++sh)
                {
                    // every site outside shells 0..sh-1 is at least (sh - 1) * hmin away
                    double reach = (sh - 1) * hmin;
                    if (sh > 0 && reach > 0 && reach * reach >= 4.0 * r2) break;
                    cand.clear();
                    for (int a = ci[0] - sh; a <= ci[0] + sh; ++a)
                    {
                        if (a < 0 || a >= nbk[0]) continue;
                        for (int b = ci[1] - sh; b <= ci[1] + sh; ++b)
                        {
                            if (b < 0 || b >= nbk[1]) continue;
                            for (int c = ci[2] - sh; c <= ci[2] + sh; ++c)
                            {
                                if (c < 0 || c >= nbk[2]) continue;
                                if (std::max(std::abs(a - ci[0]), std::max(std::abs(b - ci[1]), std::abs(c - ci[2]))) != sh)
                                    continue;
                                int64_t bk = (static_cast<int64_t>(a) * nbk[1] + b) * nbk[2] + c;
                                for (int64_t q = boff[bk]; q < boff[bk + 1]; ++q)
                                {
                                    int32_t j = bid[q];
                                    if (j == i) continue;
                                    const double* pj = &V->sites[3 * j];
                                    double ex = pj[0] - p[0], ey = pj[1] - p[1], ez = pj[2] - p[2];
                                    cand.emplace_back(ex * ex + ey * ey + ez * ez, j);
                                }
                            }
                        }
                    }
                    std::sort(cand.begin(), cand.end());
                    for (auto& cj : cand)
                    {
                        if (cj.first >= 4.0 * r2) break;
                        const double* pj = &V->sites[3 * cj.second];
                        double nrm[3] = {pj[0] - p[0], pj[1] - p[1], pj[2] - p[2]};
                        double c = nrm[0] * 0.5 * (pj[0] + p[0]) + nrm[1] * 0.5 * (pj[1] + p[1]) + nrm[2] * 0.5 * (pj[2] + p[2]);
                        double tol = 1e-10 * std::sqrt(r2);
                        if (P.clip(nrm, c, cj.second, tol, dtmp, rtmp)) r2 = P.maxDist2(p);
                    }
                }
                // neighbour list: faces with a non-degenerate area
                double scale = r2;
                std::vector<int32_t>& L = nbs[i];
                for (auto& f : P.faces)
                    if (P.faceArea(f) > 1e-14 * scale) L.push_back(f.id);
                if (L.empty()) failed.store(1);
                double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
                for (size_t q = 0; q < P.v.size(); q += 3)
                    for (int a = 0; a < 3; ++a)
                    {
                        lo[a] = std::min(lo[a], P.v[q + a]);
                        hi[a] = std::max(hi[a], P.v[q + a]);
                    }
                for (int a = 0; a < 3; ++a)
                {
                    V->lo[3 * i + a] = lo[a];
                    V->hi[3 * i + a] = hi[a];
                }
                V->volume[i] = P.volume(p);
            }
        });
        if (failed.load()) return SKG_ERR_STATE;
        // make the neighbour relation mutual (a face seen from one side only is a sliver
        // below the clipping tolerance on the other side): keep a pair only if both see it
        std::vector<int64_t> nbo(n + 1, 0);
        {
            std::vector<std::vector<int32_t>> keep(n);
            for (int64_t i = 0; i < n; ++i)
                for (int32_t j : nbs[i])
                {
                    if (j < 0)
                    {
                        keep[i].push_back(j);
                        continue;
                    }
                    const auto& o = nbs[j];
                    if (std::find(o.begin(), o.end(), static_cast<int32_t>(i)) != o.end()) keep[i].push_back(j);
                }
            nbs.swap(keep);
        }
        V->nbBegin.resize(n);
        V->nbCount.resize(n);
        for (int64_t i = 0; i < n; ++i)
        {
            if (V->nbIndex.size() + nbs[i].size() > static_cast<size_t>(INT32_MAX)) return SKG_ERR_CAPACITY;
            V->nbBegin[i] = static_cast<int32_t>(V->nbIndex.size());
            V->nbCount[i] = static_cast<int32_t>(nbs[i].size());
            V->nbIndex.insert(V->nbIndex.end(), nbs[i].begin(), nbs[i].end());
        }
        // block lists: nb = max(3, min(250, cbrt(n))) blocks per axis; a cell is listed in every
        // block its extent (widened by eps) overlaps (VoronoiMeshSnapshot.cpp:544, 783-831)
        const int nbb = std::max(3, std::min(250, static_cast<int>(std::cbrt(static_cast<double>(n)))));
        V->blockN = nbb;
        auto cellIdx = [&](double x, int a) {
            // Box::cellIndices (Box.hpp:162-167)
            int i = static_cast<int>(nbb * (x - box[a]) / (box[3 + a] - box[a]));
            return std::max(0, std::min(nbb - 1, i));
        };
        const int64_t nbl = static_cast<int64_t>(nbb) * nbb * nbb;
        std::vector<std::vector<int32_t>> lists(nbl);
        for (int64_t m = 0; m < n; ++m)
        {
            int i1 = cellIdx(V->lo[3 * m] - eps, 0), j1 = cellIdx(V->lo[3 * m + 1] - eps, 1), k1 = cellIdx(V->lo[3 * m + 2] - eps, 2);
            int i2 = cellIdx(V->hi[3 * m] + eps, 0), j2 = cellIdx(V->hi[3 * m + 1] + eps, 1), k2 = cellIdx(V->hi[3 * m + 2] + eps, 2);
            for (int i = i1; i <= i2; i++)
                for (int j = j1; j <= j2; j++)
                    for (int k = k1; k <= k2; k++) lists[(static_cast<int64_t>(i) * nbb + j) * nbb + k].push_back(static_cast<int32_t>(m));
        }
        V->blockBegin.resize(nbl);
        V->blockCount.resize(nbl);
        for (int64_t b = 0; b < nbl; ++b)
        {
            if (V->blockIndex.size() + lists[b].size() > static_cast<size_t>(INT32_MAX)) return SKG_ERR_CAPACITY;
            V->blockBegin[b] = static_cast<int32_t>(V->blockIndex.size());
            V->blockCount[b] = static_cast<int32_t>(lists[b].size());
            V->blockIndex.insert(V->blockIndex.end(), lists[b].begin(), lists[b].end());
        }
        *out = V.release();
        return SKG_OK;
    }
    catch (const std::bad_alloc&)
    {
        return SKG_ERR_CAPACITY;
    }
}

int skg_voronoi_info(const skg_voronoi* v, int64_t* ncells, int64_t* nneighbors, int32_t* block_n, int64_t* nblockrefs)
{
    if (!v) return SKG_ERR_ARGUMENT;
    if (ncells) *ncells = static_cast<int64_t>(v->sites.size() / 3);
    if (nneighbors) *nneighbors = static_cast<int64_t>(v->nbIndex.size());
    if (block_n) *block_n = v->blockN;
    if (nblockrefs) *nblockrefs = static_cast<int64_t>(v->blockIndex.size());
    return SKG_OK;
}

int skg_voronoi_export(const skg_voronoi* v, double* sites, int32_t* nb_begin, int32_t* nb_count, int32_t* nb_index,
                       int32_t* block_begin, int32_t* block_count, int32_t* block_index, double* volume)
{
    if (!v) return SKG_ERR_ARGUMENT;
    const size_t n = v->sites.size() / 3;
    if (sites) std::memcpy(sites, v->sites.data(), 3 * n * sizeof(double));
    if (nb_begin) std::memcpy(nb_begin, v->nbBegin.data(), n * sizeof(int32_t));
    if (nb_count) std::memcpy(nb_count, v->nbCount.data(), n * sizeof(int32_t));
    if (nb_index) std::memcpy(nb_index, v->nbIndex.data(), v->nbIndex.size() * sizeof(int32_t));
    if (block_begin) std::memcpy(block_begin, v->blockBegin.data(), v->blockBegin.size() * sizeof(int32_t));
    if (block_count) std::memcpy(block_count, v->blockCount.data(), v->blockCount.size() * sizeof(int32_t));
    if (block_index) std::memcpy(block_index, v->blockIndex.data(), v->blockIndex.size() * sizeof(int32_t));
    if (volume) std::memcpy(volume, v->volume.data(), n * sizeof(double));
    return SKG_OK;
}

int skg_voronoi_upload(skg_engine* e, const skg_voronoi* v)
{
    if (!e || !v) return SKG_ERR_ARGUMENT;
    return skg_grid_voronoi(e, static_cast<int64_t>(v->sites.size() / 3), v->sites.data(), v->nbBegin.data(),
                            v->nbCount.data(), v->nbIndex.data(), v->blockN, v->blockBegin.data(), v->blockCount.data(),
                            v->blockIndex.data(), v->box);
}

// Mean of a lattice field (nx*ny*nz values over the Voronoi domain, index (i*ny + j)*nz + k,
// nearest lattice cell) at `samples` uniform random points of each cell, drawn in the cell's
// extent and kept when the cell's site is the nearest among the cell and its neighbours
// (VoronoiMeshSnapshot::generatePosition, :1005-1023); per-cell Philox streams.
int skg_sample_density_voronoi(const skg_voronoi* v, const double* lattice, int32_t nx, int32_t ny, int32_t nz,
                               int32_t samples, uint64_t seed, int32_t threads, double* out)
{
    if (!v || !lattice || nx < 1 || ny < 1 || nz < 1 || samples < 1 || !out) return SKG_ERR_ARGUMENT;
    const int64_t n = static_cast<int64_t>(v->sites.size() / 3);
    const double* b = v->box;
    parallelChunks(n, threads, [&](int64_t c0, int64_t c1) {
        for (int64_t m = c0; m < c1; ++m)
        {
            skg::Philox rng;
            rng.init(seed, static_cast<uint64_t>(m), 0x564f524fu);
            const double* lo = &v->lo[3 * m];
            const double* hi = &v->hi[3 * m];
            const double* p = &v->sites[3 * m];
            double sum = 0;
            int got = 0;
            for (int tries = 0; got < samples && tries < 100 * samples; ++tries)
            {
                double x = lo[0] + (hi[0] - lo[0]) * rng.uniform();
                double y = lo[1] + (hi[1] - lo[1]) * rng.uniform();
                double z = lo[2] + (hi[2] - lo[2]) * rng.uniform();
                double d0 = (x - p[0]) * (x - p[0]) + (y - p[1]) * (y - p[1]) + (z - p[2]) * (z - p[2]);
                bool inside = true;
                for (int32_t q = v->nbBegin[m]; q < v->nbBegin[m] + v->nbCount[m] && inside; ++q)
                {
                    int32_t j = v->nbIndex[q];
                    if (j < 0) continue;
                    const double* s = &v->sites[3 * static_cast<int64_t>(j)];
                    if ((x - s[0]) * (x - s[0]) + (y - s[1]) * (y - s[1]) + (z - s[2]) * (z - s[2]) < d0) inside = false;
                }
                if (!inside) continue;
                int i = std::max(0, std::min(nx - 1, static_cast<int>((x - b[0]) / (b[3] - b[0]) * nx)));
                int j = std::max(0, std::min(ny - 1, static_cast<int>((y - b[1]) / (b[4] - b[1]) * ny)));
                int k = std::max(0, std::min(nz - 1, static_cast<int>((z - b[2]) / (b[5] - b[2]) * nz)));
                sum += lattice[(static_cast<int64_t>(i) * ny + j) * nz + k];
                ++got;
            }
            out[m] = got ? sum / got : 0.;
        }
    });
    return SKG_OK;
}

int skg_sample_density_voronoi_model(const skg_voronoi* v, const skg_disk_model* model, int32_t samples, uint64_t seed,
                                     int32_t threads, double* out)
{
    if (!v || !model || samples < 1 || !out) return SKG_ERR_ARGUMENT;
    const skg::DiskModel md(*model);
    const int64_t n = static_cast<int64_t>(v->sites.size() / 3);
    parallelChunks(n, threads, [&](int64_t c0, int64_t c1) {
        for (int64_t m = c0; m < c1; ++m)
        {
            // positions uniform in the cell: uniform in its extent, kept when the site is the
            // nearest of the cell's neighbours (VoronoiMeshSnapshot::generatePosition, :1005-1023)
            skg::Philox rng;
            rng.init(seed, static_cast<uint64_t>(m), 0x564f524du);
            const double* lo = &v->lo[3 * m];
            const double* hi = &v->hi[3 * m];
            const double* p = &v->sites[3 * m];
            double sum = 0;
            int got = 0;
            for (int tries = 0; got < samples && tries < 100 * samples; ++tries)
            {
                double x = lo[0] + (hi[0] - lo[0]) * rng.uniform();
                double y = lo[1] + (hi[1] - lo[1]) * rng.uniform();
                double z = lo[2] + (hi[2] - lo[2]) * rng.uniform();
                double d0 = (x - p[0]) * (x - p[0]) + (y - p[1]) * (y - p[1]) + (z - p[2]) * (z - p[2]);
                bool inside = true;
                for (int32_t q = v->nbBegin[m]; q < v->nbBegin[m] + v->nbCount[m] && inside; ++q)
                {
                    int32_t j = v->nbIndex[q];
                    if (j < 0) continue;
                    const double* s = &v->sites[3 * static_cast<int64_t>(j)];
                    if ((x - s[0]) * (x - s[0]) + (y - s[1]) * (y - s[1]) + (z - s[2]) * (z - s[2]) < d0) inside = false;
                }
                if (!inside) continue;
                sum += md.geom(x, y, z);
                ++got;
            }
            out[m] = got ? md.nnorm * (sum / got) : 0.;
        }
    });
    return SKG_OK;
}

void skg_voronoi_free(skg_voronoi* v) { delete v; }

}  // extern "C"


// ==================================================================== Delaunay tetrahedra
//
// The tetrahedral grid of TetraMeshSpatialGrid (TetraMeshSpatialGrid.cpp:300-452: sampled
// vertices plus the 8 domain corners, outside points dropped, sorted by x, near duplicates
// dropped, then TetGen's Delaunay tetrahedralisation) built here by Bowyer-Watson insertion
// with exact predicates: the vertices are mapped to an integer lattice of 2^23 steps along the
// longest axis (only for the predicates; the output keeps the input coordinates), where the orientation
// and in-sphere determinants are exact in 128-bit integers.  Starting from a Delaunay
// triangulation of the domain cube, each vertex (in Morton order) is located by a visibility
// walk and its cavity -- the tetrahedra whose circumsphere strictly contains it -- replaced
// by the fan from the vertex to the cavity boundary.  Output tetrahedra are positively
// oriented (det[b-a, c-a, d-a] > 0, the reference's convention) with neighbour f opposite
// vertex f and outward face normals computed as storeTetrahedra does.

struct skg_tetra
{
    double box[6];
    std::vector<double> verts;    // 3 per vertex
    std::vector<int32_t> tv;      // 4 per tetra
    std::vector<int32_t> nb;      // 4 per tetra, -1 on the hull
    std::vector<double> normals;  // 12 per tetra
};

namespace
{
    using i128 = __int128;
    struct LP
    {
        int64_t x, y, z;
    };

    inline i128 orient3(const LP& a, const LP& b, const LP& c, const LP& d)
    {
        const i128 bx = b.x - a.x, by = b.y - a.y, bz = b.z - a.z;
        const i128 cx = c.x - a.x, cy = c.y - a.y, cz = c.z - a.z;
        const i128 dx = d.x - a.x, dy = d.y - a.y, dz = d.z - a.z;
        return bx * (cy * dz - cz * dy) - by * (cx * dz - cz * dx) + bz * (cx * dy - cy * dx);
    }

    // < 0 when e lies strictly inside the sphere through a, b, c, d of a positively oriented tetra
    inline i128 insphere(const LP& a, const LP& b, const LP& c, const LP& d, const LP& e)
    {
        const LP* q[4] = {&a, &b, &c, &d};
        i128 m[4][4];
        for (int r = 0; r < 4; ++r)
        {
            const i128 x = q[r]->x - e.x, y = q[r]->y - e.y, z = q[r]->z - e.z;
            m[r][0] = x;
            m[r][1] = y;
            m[r][2] = z;
            m[r][3] = x * x + y * y + z * z;
        }
        // Laplace expansion along the last column with 3x3 minors
        auto det3 = [&](int r0, int r1, int r2) {
            return m[r0][0] * (m[r1][1] * m[r2][2] - m[r1][2] * m[r2][1]) -
                   m[r0][1] * (m[r1][0] * m[r2][2] - m[r1][2] * m[r2][0]) +
                   m[r0][2] * (m[r1][0] * m[r2][1] - m[r1][1] * m[r2][0]);
        };
        return -m[0][3] * det3(1, 2, 3) + m[1][3] * det3(0, 2, 3) - m[2][3] * det3(0, 1, 3) + m[3][3] * det3(0, 1, 2);
    }

    struct Tet
    {
        int32_t v[4];
        int32_t n[4];  // neighbour across the face opposite v[i]
        bool alive;
    };

    uint64_t morton3(uint64_t x, uint64_t y, uint64_t z)
    {
        auto spread = [](uint64_t a) {
            a &= 0x1fffff;
            a = (a | a << 32) & 0x1f00000000ffffULL;
            a = (a | a << 16) & 0x1f0000ff0000ffULL;
            a = (a | a << 8) & 0x100f00f00f00f00fULL;
            a = (a | a << 4) & 0x10c30c30c30c30c3ULL;
            a = (a | a << 2) & 0x1249249249249249ULL;
            return a;
        };
        return spread(x) | spread(y) << 1 | spread(z) << 2;
    }
}

extern "C" {

int skg_tetra_build(int64_t npoints, const double* points, const double* box, skg_tetra** out)
{
    if (!out || npoints < 0 || (npoints && !points) || !box || !(box[0] < box[3] && box[1] < box[4] && box[2] < box[5]))
        return SKG_ERR_ARGUMENT;
    *out = nullptr;
    try
    {
        auto M = std::make_unique<skg_tetra>();
        std::memcpy(M->box, box, sizeof(M->box));
        const double dx = box[3] - box[0], dy = box[4] - box[1], dz = box[5] - box[2];
        const double eps = 1e-12 * std::sqrt(dx * dx + dy * dy + dz * dz);  // TetraMeshSpatialGrid.cpp:239
        // TetraMeshSpatialGrid::addCorners and removeInvalid (:300-380)
        std::vector<std::array<double, 3>> v;
        v.reserve(npoints + 8);
        for (int64_t i = 0; i < npoints; ++i) v.push_back({points[3 * i], points[3 * i + 1], points[3 * i + 2]});
        for (int c = 0; c < 8; ++c)
            v.push_back({(c & 4) ? box[3] : box[0], (c & 2) ? box[4] : box[1], (c & 1) ? box[5] : box[2]});
        v.erase(std::remove_if(v.begin(), v.end(),
                               [&](const std::array<double, 3>& p) {
                                   return !(p[0] >= box[0] && p[0] <= box[3] && p[1] >= box[1] && p[1] <= box[4]
                                            && p[2] >= box[2] && p[2] <= box[5]);
                               }),
                v.end());
        std::sort(v.begin(), v.end(), [](const std::array<double, 3>& a, const std::array<double, 3>& b) { return a[0] < b[0]; });
        {
            const int n0 = static_cast<int>(v.size());
            std::vector<int> toRemove;
            for (int i = 0; i < n0; ++i)
                for (int j = i + 1; j < n0 && v[j][0] - v[i][0] < eps; ++j)
                {
                    double ex = v[j][0] - v[i][0], ey = v[j][1] - v[i][1], ez = v[j][2] - v[i][2];
                    if (ex * ex + ey * ey + ez * ez < eps * eps)
                    {
                        toRemove.push_back(i);
                        break;
                    }
                }
            for (auto it = toRemove.rbegin(); it != toRemove.rend(); ++it) v.erase(v.begin() + *it);
        }
        const int nv = static_cast<int>(v.size());
        // lattice coordinates for the predicates, one step h for all axes (the in-sphere test is
        // not invariant under anisotropic scaling); 2^23 steps along the longest axis keep the
        // in-sphere determinant below 2^127; interior points stay off the domain faces
        const double ext = std::max(dx, std::max(dy, dz));
        const double h = ext / static_cast<double>(int64_t(1) << 23);
        int64_t Sa[3];
        for (int a = 0; a < 3; ++a) Sa[a] = std::max<int64_t>(2, std::llround((box[3 + a] - box[a]) / h));
        std::vector<LP> L(nv);
        std::vector<int> corner(8, -1);
        for (int i = 0; i < nv; ++i)
        {
            int64_t q[3];
            bool onCorner = true;
            for (int a = 0; a < 3; ++a)
            {
                q[a] = static_cast<int64_t>(std::llround((v[i][a] - box[a]) / h));
                if (v[i][a] != box[a] && v[i][a] != box[3 + a]) onCorner = false;
            }
            if (onCorner)
            {
                int c = (v[i][0] == box[3] ? 4 : 0) | (v[i][1] == box[4] ? 2 : 0) | (v[i][2] == box[5] ? 1 : 0);
                corner[c] = i;
                for (int a = 0; a < 3; ++a) q[a] = (v[i][a] == box[a]) ? 0 : Sa[a];
            }
            else
                for (int a = 0; a < 3; ++a) q[a] = std::max<int64_t>(1, std::min<int64_t>(Sa[a] - 1, q[a]));
            L[i] = LP{q[0], q[1], q[2]};
        }
        for (int c = 0; c < 8; ++c)
            if (corner[c] < 0) return SKG_ERR_STATE;
        // the cube as 6 positively oriented tetrahedra around the diagonal corner 0 - corner 7
        std::vector<Tet> T;
        T.reserve(static_cast<size_t>(7.5 * nv) + 16);
        {
            const int c0 = corner[0], c7 = corner[7];
            const int ring[6] = {corner[1], corner[3], corner[2], corner[6], corner[4], corner[5]};
            for (int k = 0; k < 6; ++k)
            {
                Tet t;
                t.v[0] = c0;
                t.v[1] = c7;
                t.v[2] = ring[k];
                t.v[3] = ring[(k + 1) % 6];
                if (orient3(L[t.v[0]], L[t.v[1]], L[t.v[2]], L[t.v[3]]) < 0) std::swap(t.v[2], t.v[3]);
                for (int f = 0; f < 4; ++f) t.n[f] = -1;
                t.alive = true;
                T.push_back(t);
            }
        }
        // adjacency by shared faces (sorted vertex triples)
        auto faceKey = [](int a, int b, int c) {
            int x[3] = {a, b, c};
            std::sort(x, x + 3);
            return (static_cast<uint64_t>(x[0]) << 42) | (static_cast<uint64_t>(x[1]) << 21) | static_cast<uint64_t>(x[2]);
        };
        auto link = [&](std::vector<int>& ids) {
            std::vector<std::pair<uint64_t, int>> fk;
            for (int id : ids)
                for (int f = 0; f < 4; ++f)
                {
                    const Tet& t = T[id];
                    int o[3], m = 0;
                    for (int q = 0; q < 4; ++q)
                        if (q != f) o[m++] = t.v[q];
                    fk.emplace_back(faceKey(o[0], o[1], o[2]), id * 4 + f);
                }
            std::sort(fk.begin(), fk.end());
            for (size_t q = 0; q + 1 < fk.size(); ++q)
                if (fk[q].first == fk[q + 1].first)
                {
                    int a = fk[q].second, b = fk[q + 1].second;
                    T[a / 4].n[a % 4] = b / 4;
                    T[b / 4].n[b % 4] = a / 4;
                }
        };
        {
            std::vector<int> ids{0, 1, 2, 3, 4, 5};
            link(ids);
        }
        // insertion order: Morton order of the lattice points, corners excluded
        std::vector<std::pair<uint64_t, int>> order;
        for (int i = 0; i < nv; ++i)
        {
            bool isCorner = false;
            for (int c = 0; c < 8; ++c) isCorner |= corner[c] == i;
            if (!isCorner) order.emplace_back(morton3(L[i].x, L[i].y, L[i].z), i);
        }
        std::sort(order.begin(), order.end());
        std::vector<int32_t> freeList, cavity, stamp, newIds;
        std::vector<int32_t> mark;
        int32_t last = 0;
        int32_t epoch = 0;
        struct BFace
        {
            int a, b, c, outer, outerFace;
        };
        std::vector<BFace> bfaces;
        for (auto& oi : order)
        {
            const int p = oi.second;
            const LP& P = L[p];
            // visibility walk
            int cur = last;
            if (!T[cur].alive)
                for (cur = 0; !T[cur].alive; ++cur) {}
            for (int guard = 0;; ++guard)
            {
                if (guard > 100 * (nv + 10)) return SKG_ERR_STATE;
                const Tet& t = T[cur];
                int next = -1;
                for (int f = 0; f < 4 && next < 0; ++f)
                {
                    LP q[4] = {L[t.v[0]], L[t.v[1]], L[t.v[2]], L[t.v[3]]};
                    q[f] = P;
                    if (orient3(q[0], q[1], q[2], q[3]) < 0) next = t.n[f];
                }
                if (next < 0) break;  // inside (or on the boundary of) cur
                cur = next;
            }
            // cavity: tetrahedra whose circumsphere strictly contains p, grown from cur
            ++epoch;
            if (mark.size() < T.size()) mark.resize(T.size() * 2, 0);
            cavity.clear();
            cavity.push_back(cur);
            mark[cur] = epoch;
            bfaces.clear();
            for (size_t q = 0; q < cavity.size(); ++q)
            {
                const int id = cavity[q];
                for (int f = 0; f < 4; ++f)
                {
                    const int nb = T[id].n[f];
                    if (nb >= 0 && mark[nb] == epoch) continue;
                    bool inside = false;
                    if (nb >= 0)
                    {
                        const Tet& u = T[nb];
                        inside = insphere(L[u.v[0]], L[u.v[1]], L[u.v[2]], L[u.v[3]], P) < 0;
                    }
                    if (inside)
                    {
                        mark[nb] = epoch;
                        cavity.push_back(nb);
                    }
                    else
                    {
                        int o[3], m = 0;
                        for (int r = 0; r < 4; ++r)
                            if (r != f) o[m++] = T[id].v[r];
                        int outerFace = -1;
                        if (nb >= 0)
                            for (int r = 0; r < 4; ++r)
                                if (T[nb].n[r] == id) outerFace = r;
                        bfaces.push_back({o[0], o[1], o[2], nb, outerFace});
                    }
                }
            }
            // a boundary face already counted from the other side cannot occur: the neighbour
            // across it is either in the cavity (skipped) or outside (counted once)
            for (int id : cavity)
            {
                T[id].alive = false;
                freeList.push_back(id);
            }
            newIds.clear();
            for (const BFace& bf : bfaces)
            {
                Tet t;
                t.v[0] = bf.a;
                t.v[1] = bf.b;
                t.v[2] = bf.c;
                t.v[3] = p;
                i128 o = orient3(L[t.v[0]], L[t.v[1]], L[t.v[2]], L[t.v[3]]);
                if (o == 0) return SKG_ERR_STATE;  // degenerate (coplanar) insertion
                if (o < 0) std::swap(t.v[0], t.v[1]);
                for (int f = 0; f < 4; ++f) t.n[f] = -1;
                t.n[3] = bf.outer;  // face opposite p is the cavity boundary face
                t.alive = true;
                int id;
                if (!freeList.empty())
                {
                    id = freeList.back();
                    freeList.pop_back();
                    T[id] = t;
                }
                else
                {
                    id = static_cast<int>(T.size());
                    T.push_back(t);
                }
                if (bf.outer >= 0) T[bf.outer].n[bf.outerFace] = id;
                newIds.push_back(id);
            }
            // link the new tetrahedra among themselves (faces through p)
            {
                std::vector<std::pair<uint64_t, int>> fk;
                fk.reserve(newIds.size() * 3);
                for (int id : newIds)
                    for (int f = 0; f < 3; ++f)
                    {
                        const Tet& t = T[id];
                        int o[3], m = 0;
                        for (int q = 0; q < 4; ++q)
                            if (q != f) o[m++] = t.v[q];
                        fk.emplace_back(faceKey(o[0], o[1], o[2]), id * 4 + f);
                    }
                std::sort(fk.begin(), fk.end());
                for (size_t q = 0; q + 1 < fk.size(); ++q)
                    if (fk[q].first == fk[q + 1].first)
                    {
                        int a = fk[q].second, b = fk[q + 1].second;
                        T[a / 4].n[a % 4] = b / 4;
                        T[b / 4].n[b % 4] = a / 4;
                        ++q;
                    }
            }
            last = newIds.empty() ? last : newIds.back();
        }
        // compact, then faces and normals as storeTetrahedra (TetraMeshSpatialGrid.cpp:500-540)
        std::vector<int32_t> newId(T.size(), -1);
        int32_t nt = 0;
        for (size_t i = 0; i < T.size(); ++i)
            if (T[i].alive) newId[i] = nt++;
        M->verts.resize(3 * static_cast<size_t>(nv));
        for (int i = 0; i < nv; ++i)
            for (int a = 0; a < 3; ++a) M->verts[3 * i + a] = v[i][a];
        M->tv.resize(4 * static_cast<size_t>(nt));
        M->nb.resize(4 * static_cast<size_t>(nt));
        M->normals.resize(12 * static_cast<size_t>(nt));
        for (size_t i = 0; i < T.size(); ++i)
        {
            if (!T[i].alive) continue;
            const int32_t m = newId[i];
            for (int q = 0; q < 4; ++q)
            {
                M->tv[4 * m + q] = T[i].v[q];
                M->nb[4 * m + q] = T[i].n[q] >= 0 ? newId[T[i].n[q]] : -1;
            }
            for (int f = 0; f < 4; ++f)
            {
                int cv[3] = {(f + 3) % 4, (f + 2) % 4, (f + 1) % 4};
                if (f % 2 == 0) std::swap(cv[0], cv[2]);
                const double* v0 = &M->verts[3 * static_cast<size_t>(T[i].v[cv[0]])];
                const double* v1 = &M->verts[3 * static_cast<size_t>(T[i].v[cv[1]])];
                const double* v2 = &M->verts[3 * static_cast<size_t>(T[i].v[cv[2]])];
                double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
                double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
                double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
                double nn = std::sqrt(nx * nx + ny * ny + nz * nz);
                M->normals[12 * m + 3 * f] = nx / nn;
                M->normals[12 * m + 3 * f + 1] = ny / nn;
                M->normals[12 * m + 3 * f + 2] = nz / nn;
            }
        }
        *out = M.release();
        return SKG_OK;
    }
    catch (const std::bad_alloc&)
    {
        return SKG_ERR_CAPACITY;
    }
}

int skg_tetra_info(const skg_tetra* t, int64_t* nverts, int64_t* ntets)
{
    if (!t) return SKG_ERR_ARGUMENT;
    if (nverts) *nverts = static_cast<int64_t>(t->verts.size() / 3);
    if (ntets) *ntets = static_cast<int64_t>(t->tv.size() / 4);
    return SKG_OK;
}

int skg_tetra_export(const skg_tetra* t, double* verts, int32_t* tet_verts, int32_t* face_nb, double* face_normal)
{
    if (!t) return SKG_ERR_ARGUMENT;
    if (verts) std::memcpy(verts, t->verts.data(), t->verts.size() * sizeof(double));
    if (tet_verts) std::memcpy(tet_verts, t->tv.data(), t->tv.size() * sizeof(int32_t));
    if (face_nb) std::memcpy(face_nb, t->nb.data(), t->nb.size() * sizeof(int32_t));
    if (face_normal) std::memcpy(face_normal, t->normals.data(), t->normals.size() * sizeof(double));
    return SKG_OK;
}

int skg_tetra_upload(skg_engine* e, const skg_tetra* t)
{
    if (!e || !t) return SKG_ERR_ARGUMENT;
    return skg_grid_tetra(e, static_cast<int64_t>(t->verts.size() / 3), t->verts.data(),
                          static_cast<int64_t>(t->tv.size() / 4), t->tv.data(), t->nb.data(), t->normals.data(), t->box);
}

void skg_tetra_free(skg_tetra* t) { delete t; }

}  // extern "C"
