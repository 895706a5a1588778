// Host-side set-up of the medium the life cycle reads: the density model, the
// density-refined octree with its per-wall neighbour lists, cell-averaged number densities
// and the dust tables at the simulation wavelengths.  These mirror the reference set-up
// that precedes the photon loop (TreeSpatialGrid::setupSelfAfter, DensityTreePolicy,
// MediumSystem::setupSelfAfter, TabulatedDustMix) so that the engine can be driven end to
// end without the reference.
//
// Random sampling (tree refinement, cell densities) uses per-node / per-cell Philox streams,
// so results do not depend on the number of threads.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "../../include/skirt_mi355x.h"
#include "disk_model.hpp"
#include "octree_build.hpp"
#include "philox.hpp"

namespace
{
    constexpr uint32_t kTagDensity = 0x44454e53u;  // stream tag for cell density samples

    using Model = skg::DiskModel;

    template<typename F> void parallelFor(int64_t n, int threads, F&& f)
    {
        if (threads <= 1 || n < 4096)
        {
            f(0, n);
            return;
        }
        std::atomic<int64_t> next{0};
        const int64_t chunk = std::max<int64_t>(1024, n / (threads * 16));
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
        for (auto& th : pool) th.join();
    }

    int defaultThreads(int threads)
    {
        if (threads > 0) return threads;
        unsigned hc = std::thread::hardware_concurrency();
        return hc ? static_cast<int>(std::min(hc, 32u)) : 4;
    }

    // mean normalised density over n random points in the box (Random::position(box): x, y, z
    // drawn in that order, x = xmin + u (xmax - xmin))
    double sampleMean(const Model& md, const double* b, int n, uint64_t seed, uint64_t id, uint32_t tag)
    {
        skg::Philox rng;
        rng.init(seed, id, tag);
        double sum = 0.;
        for (int i = 0; i < n; ++i)
        {
            double x = b[0] + rng.uniform() * (b[3] - b[0]);
            double y = b[1] + rng.uniform() * (b[4] - b[1]);
            double z = b[2] + rng.uniform() * (b[5] - b[2]);
            sum += md.geom(x, y, z);
        }
        return sum / n;
    }
}

// ------------------------------------------------------------------------------- octree
namespace
{
    // a node while the tree is built: bounds, first child, level and the neighbour lists that
    // subdivisions relink (TreeNode's members)
    struct Node
    {
        double b[6];
        int32_t first = -1;  // first child id
        int32_t level = 0;
        std::vector<int32_t> nb[6];
    };
    enum Wall
    {
        BACK = 0,
        FRONT,
        LEFT,
        RIGHT,
        BOTTOM,
        TOP
    };

    // TreeNode::makeNeighbors / deleteNeighbor (TreeNode.cpp:116-141)
    inline void makeNeighbors(std::vector<Node>& v, int wall1, int32_t n1, int32_t n2)
    {
        static const int complement[] = {FRONT, BACK, RIGHT, LEFT, TOP, BOTTOM};
        v[n1].nb[wall1].push_back(n2);
        v[n2].nb[complement[wall1]].push_back(n1);
    }
    inline void deleteNeighbor(std::vector<Node>& v, int32_t node, int wall, int32_t other)
    {
        auto& l = v[node].nb[wall];
        auto it = std::find(l.begin(), l.end(), other);
        if (it != l.end()) l.erase(it);
    }

    // OctTreeNode::createChildren + TreeNode::subdivide + OctTreeNode::addNeighbors
    // (OctTreeNode.cpp:22-136): the eight children get consecutive ids, sibling links are made
    // first, then each wall's neighbours are relinked to the children they touch.
    void subdivide(std::vector<Node>& v, int32_t p)
    {
        const int32_t c = static_cast<int32_t>(v.size());
        double b[6];
        std::memcpy(b, v[p].b, sizeof(b));
        double xc = 0.5 * (b[0] + b[3]), yc = 0.5 * (b[1] + b[4]), zc = 0.5 * (b[2] + b[5]);
        const double boxes[8][6] = {{b[0], b[1], b[2], xc, yc, zc}, {xc, b[1], b[2], b[3], yc, zc},
                                    {b[0], yc, b[2], xc, b[4], zc}, {xc, yc, b[2], b[3], b[4], zc},
                                    {b[0], b[1], zc, xc, yc, b[5]}, {xc, b[1], zc, b[3], yc, b[5]},
                                    {b[0], yc, zc, xc, b[4], b[5]}, {xc, yc, zc, b[3], b[4], b[5]}};
        const int32_t lvl = v[p].level + 1;
        v.resize(v.size() + 8);
        for (int l = 0; l < 8; ++l)
        {
            std::memcpy(v[c + l].b, boxes[l], sizeof(boxes[l]));
            v[c + l].level = lvl;
        }
        v[p].first = c;
        makeNeighbors(v, FRONT, c + 0, c + 1);
        makeNeighbors(v, RIGHT, c + 0, c + 2);
        makeNeighbors(v, TOP, c + 0, c + 4);
        makeNeighbors(v, RIGHT, c + 1, c + 3);
        makeNeighbors(v, TOP, c + 1, c + 5);
        makeNeighbors(v, FRONT, c + 2, c + 3);
        makeNeighbors(v, TOP, c + 2, c + 6);
        makeNeighbors(v, TOP, c + 3, c + 7);
        makeNeighbors(v, FRONT, c + 4, c + 5);
        makeNeighbors(v, RIGHT, c + 4, c + 6);
        makeNeighbors(v, RIGHT, c + 5, c + 7);
        makeNeighbors(v, FRONT, c + 6, c + 7);
        // the split point is the upper corner of child 0
        const double X = v[c].b[3], Y = v[c].b[4], Z = v[c].b[5];
        auto relink = [&](int wall, int opposite, int a0, int a1, int a2, int a3, int u, int w) {
            // u/w: which two coordinates (0=x,1=y,2=z) span the wall
            const double U = (u == 0 ? X : u == 1 ? Y : Z), Wc = (w == 0 ? X : w == 1 ? Y : Z);
            std::vector<int32_t> list = v[p].nb[wall];
            for (int32_t nbr : list)
            {
                deleteNeighbor(v, nbr, opposite, p);
                const double* nb = v[nbr].b;
                double umin = nb[u], umax = nb[u + 3], wmin = nb[w], wmax = nb[w + 3];
                if (umin <= U && wmin <= Wc) makeNeighbors(v, opposite, nbr, c + a0);
                if (umax >= U && wmin <= Wc) makeNeighbors(v, opposite, nbr, c + a1);
                if (umin <= U && wmax >= Wc) makeNeighbors(v, opposite, nbr, c + a2);
                if (umax >= U && wmax >= Wc) makeNeighbors(v, opposite, nbr, c + a3);
            }
        };
        relink(BACK, FRONT, 0, 2, 4, 6, 1, 2);
        relink(FRONT, BACK, 1, 3, 5, 7, 1, 2);
        relink(LEFT, RIGHT, 0, 1, 4, 5, 0, 2);
        relink(RIGHT, LEFT, 2, 3, 6, 7, 0, 2);
        relink(BOTTOM, TOP, 0, 1, 2, 3, 0, 1);
        relink(TOP, BOTTOM, 4, 5, 6, 7, 0, 1);
        // the subdivided node's own lists are never read again
        for (auto& l : v[p].nb) std::vector<int32_t>().swap(l);
    }

    // the finished table form: leaves numbered in node order, the lists in one CSR, then
    // sorted by decreasing overlap (TreeNode::sortNeighbors)
    void finishTree(std::vector<Node>& v, skg_octree& t, int threads)
    {
        const size_t n = v.size();
        t.bounds.resize(6 * n);
        t.first.resize(n);
        t.cellOf.assign(n, -1);
        t.nbOff.resize(6 * n + 1);
        int64_t total = 0;
        int32_t m = 0;
        for (size_t l = 0; l < n; ++l)
        {
            std::memcpy(&t.bounds[6 * l], v[l].b, 6 * sizeof(double));
            t.first[l] = v[l].first;
            if (v[l].first < 0) t.cellOf[l] = m++;
            for (int w = 0; w < 6; ++w)
            {
                t.nbOff[6 * l + w] = total;
                total += static_cast<int64_t>(v[l].nb[w].size());
            }
        }
        t.nbOff[6 * n] = total;
        t.ncells = m;
        t.nbId.resize(static_cast<size_t>(total));
        for (size_t l = 0; l < n; ++l)
        {
            for (int w = 0; w < 6; ++w)
                std::copy(v[l].nb[w].begin(), v[l].nb[w].end(), t.nbId.begin() + t.nbOff[6 * l + w]);
            for (auto& q : v[l].nb) std::vector<int32_t>().swap(q);
        }
        std::vector<Node>().swap(v);
        skg::octree_sort_lists(t, 2, threads);
    }
}

// the builder itself (C++ linkage; the host and GPU entry points pass their sampler)
int skg::octree_build_with(const double* box, const skg_disk_model* model, int32_t min_level, int32_t max_level,
                           double max_dust_fraction, int32_t num_samples, uint64_t seed, int32_t threads,
                           skg_octree** out, const OctreeSampler& sampler)
{
    if (!box || !model || !out || min_level < 0 || max_level < min_level || max_level > 20 || num_samples < 1
        || !(box[3] > box[0] && box[4] > box[1] && box[5] > box[2]))
        return SKG_ERR_ARGUMENT;
    try
    {
        threads = defaultThreads(threads);
        auto t = std::make_unique<skg_octree>();
        std::vector<Node> v;
        v.reserve(1 << 20);
        v.emplace_back();
        std::memcpy(v[0].b, box, 6 * sizeof(double));
        // DensityTreePolicy::constructTree (DensityTreePolicy.cpp:180-241): evaluate all nodes of a
        // level, then subdivide the flagged ones in node order
        size_t lbeg = 0, lend = 1;
        while (lend != lbeg)
        {
            const size_t n = lend - lbeg;
            std::vector<uint8_t> divide(n, 0);
            // DensityTreePolicy::needsSubdivide (:82-170), maxDustFraction criterion: the nodes
            // between the minimum and maximum level are sampled (stream id = node id)
            std::vector<double> boxes;
            std::vector<uint64_t> ids;
            std::vector<size_t> at;
            for (size_t i = 0; i < n; ++i)
            {
                const Node& nd = v[lbeg + i];
                if (nd.level < min_level)
                    divide[i] = 1;
                else if (nd.level < max_level)
                {
                    boxes.insert(boxes.end(), nd.b, nd.b + 6);
                    ids.push_back(lbeg + i);
                    at.push_back(i);
                }
            }
            if (!ids.empty())
            {
                std::vector<double> rho(ids.size());
                int st = sampler(boxes.data(), ids.data(), static_cast<int64_t>(ids.size()), rho.data());
                if (st != SKG_OK) return st;
                for (size_t q = 0; q < ids.size(); ++q)
                {
                    const double* b = boxes.data() + 6 * q;
                    double vol = (b[3] - b[0]) * (b[4] - b[1]) * (b[5] - b[2]);
                    divide[at[q]] = rho[q] * vol > max_dust_fraction ? 1 : 0;
                }
            }
            for (size_t i = 0; i < n; ++i)
                if (divide[i]) subdivide(v, static_cast<int32_t>(lbeg + i));
            lbeg = lend;
            lend = v.size();
            if (v.size() > static_cast<size_t>(INT32_MAX / 8)) return SKG_ERR_CAPACITY;
        }
        finishTree(v, *t, threads);
        *out = t.release();
        return SKG_OK;
    }
    catch (const std::bad_alloc&)
    {
        return SKG_ERR_CAPACITY;
    }
}


void skg::octree_sort_lists(skg_octree& t, int64_t min_len, int32_t threads)
{
    const int64_t n = static_cast<int64_t>(t.first.size());
    const double* B = t.bounds.data();
    parallelFor(n, defaultThreads(threads), [&](int64_t b, int64_t e) {
        for (int64_t l = b; l < e; ++l)
        {
            if (t.first[l] >= 0) continue;
            const double* base = B + 6 * l;
            for (int w = 0; w < 6; ++w)
            {
                int32_t* lo = t.nbId.data() + t.nbOff[6 * l + w];
                int32_t* hi = t.nbId.data() + t.nbOff[6 * l + w + 1];
                if (hi - lo < min_len) continue;
                std::sort(lo, hi, [&](int32_t a, int32_t c) {
                    return octree_overlap(base, B + 6 * static_cast<int64_t>(a), w)
                           > octree_overlap(base, B + 6 * static_cast<int64_t>(c), w);
                });
            }
        }
    });
}

extern "C" {

double skg_disk_density(const skg_disk_model* model, double x, double y, double z)
{
    if (!model) return 0.;
    Model md(*model);
    return md.nnorm * md.geom(x, y, z);
}

int skg_octree_build(const double* box, const skg_disk_model* model, int32_t min_level, int32_t max_level,
                     double max_dust_fraction, int32_t num_samples, uint64_t seed, int32_t threads, skg_octree** out)
{
    if (!model) return SKG_ERR_ARGUMENT;
    Model md(*model);
    const int nthreads = defaultThreads(threads);
    // the host sampler: the mean normalised density of each box over num_samples points
    return skg::octree_build_with(box, model, min_level, max_level, max_dust_fraction, num_samples, seed, threads, out,
                                  [&](const double* boxes, const uint64_t* ids, int64_t n, double* rho) {
                                      parallelFor(n, nthreads, [&](int64_t b, int64_t e) {
                                          for (int64_t i = b; i < e; ++i)
                                              rho[i] = sampleMean(md, boxes + 6 * i, num_samples, seed, ids[i], skg::kTagTree);
                                      });
                                      return SKG_OK;
                                  });
}

int skg_octree_info(const skg_octree* t, int64_t* nnodes, int64_t* ncells, int64_t* nneighbors)
{
    if (!t) return SKG_ERR_ARGUMENT;
    if (nnodes) *nnodes = static_cast<int64_t>(t->first.size());
    if (ncells) *ncells = t->ncells;
    if (nneighbors) *nneighbors = static_cast<int64_t>(t->nbId.size());
    return SKG_OK;
}

int skg_octree_export(const skg_octree* t, double* bounds, int32_t* child_first, int32_t* cell_index, int64_t* nb_off,
                      int32_t* nb_id)
{
    if (!t || !bounds || !child_first || !cell_index || !nb_off || (!nb_id && !t->nbId.empty())) return SKG_ERR_ARGUMENT;
    std::copy(t->bounds.begin(), t->bounds.end(), bounds);
    std::copy(t->first.begin(), t->first.end(), child_first);
    std::copy(t->cellOf.begin(), t->cellOf.end(), cell_index);
    std::copy(t->nbOff.begin(), t->nbOff.end(), nb_off);
    std::copy(t->nbId.begin(), t->nbId.end(), nb_id);
    return SKG_OK;
}

void skg_octree_free(skg_octree* t) { delete t; }

// Replays a given topology: nodes are visited in creation order and node l is subdivided when
// divide[l] != 0 -- the order DensityTreePolicy::constructTree subdivides in (level by level,
// node order within a level), so the node ids, bounds and neighbour lists are those of the
// grid the flags were taken from.
int skg_octree_from_flags(const double* box, int64_t nflags, const uint8_t* divide, skg_octree** out)
{
    if (!box || !out || nflags < 1 || !divide) return SKG_ERR_ARGUMENT;
    try
    {
        auto t = std::make_unique<skg_octree>();
        std::vector<Node> v;
        v.emplace_back();
        std::memcpy(v[0].b, box, 6 * sizeof(double));
        for (int64_t l = 0; l < static_cast<int64_t>(v.size()); ++l)
        {
            if (l >= nflags) return SKG_ERR_ARGUMENT;
            if (divide[l]) subdivide(v, static_cast<int32_t>(l));
        }
        if (static_cast<int64_t>(v.size()) != nflags) return SKG_ERR_ARGUMENT;
        finishTree(v, *t, 0);
        *out = t.release();
        return SKG_OK;
    }
    catch (const std::bad_alloc&)
    {
        return SKG_ERR_CAPACITY;
    }
}

int skg_sample_density_cartesian(const double* xb, int32_t nx, const double* yb, int32_t ny, const double* zb,
                                 int32_t nz, const skg_disk_model* model, int32_t num_samples, uint64_t seed,
                                 int32_t threads, double* out)
{
    if (!xb || !yb || !zb || nx < 1 || ny < 1 || nz < 1 || !model || num_samples < 1 || !out) return SKG_ERR_ARGUMENT;
    Model md(*model);
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    parallelFor(n, defaultThreads(threads), [&](int64_t b, int64_t e) {
        for (int64_t m = b; m < e; ++m)
        {
            // cell index k + nz*j + nz*ny*i (CartesianSpatialGrid.cpp:200-203)
            int64_t i = m / (static_cast<int64_t>(nz) * ny), j = (m / nz) % ny, k = m % nz;
            double bx[6] = {xb[i], yb[j], zb[k], xb[i + 1], yb[j + 1], zb[k + 1]};
            out[m] = md.nnorm * sampleMean(md, bx, num_samples, seed, static_cast<uint64_t>(m), kTagDensity);
        }
    });
    return SKG_OK;
}

int skg_sample_density_octree(const skg_octree* t, const skg_disk_model* model, int32_t num_samples, uint64_t seed,
                              int32_t threads, double* out)
{
    if (!t || !model || num_samples < 1 || !out) return SKG_ERR_ARGUMENT;
    Model md(*model);
    parallelFor(static_cast<int64_t>(t->first.size()), defaultThreads(threads), [&](int64_t b, int64_t e) {
        for (int64_t l = b; l < e; ++l)
        {
            int32_t m = t->cellOf[l];
            if (m < 0) continue;
            out[m] = md.nnorm * sampleMean(md, &t->bounds[6 * l], num_samples, seed, static_cast<uint64_t>(m), kTagDensity);
        }
    });
    return SKG_OK;
}

// tetrahedral cells from their vertex tables: num_samples points per tetra drawn uniformly by
// the folding method of Rocchini & Cignoni (J. Graphics Tools 5(4), 2000): a point of the unit
// cube is folded into the corner simplex s + t + u <= 1 and used as barycentric weights
int skg_sample_density_tetra(int64_t ntets, const double* verts, const int32_t* tet_verts, const skg_disk_model* model,
                             int32_t num_samples, uint64_t seed, int32_t threads, double* out)
{
    if (ntets < 0 || (ntets > 0 && (!verts || !tet_verts || !out)) || !model || num_samples < 1) return SKG_ERR_ARGUMENT;
    Model md(*model);
    parallelFor(ntets, defaultThreads(threads), [&](int64_t b, int64_t e) {
        for (int64_t m = b; m < e; ++m)
        {
            const double* v[4];
            for (int q = 0; q < 4; ++q) v[q] = verts + 3 * static_cast<int64_t>(tet_verts[4 * m + q]);
            skg::Philox rng;
            rng.init(seed, static_cast<uint64_t>(m), kTagDensity);
            double sum = 0.;
            for (int i = 0; i < num_samples; ++i)
            {
                double s = rng.uniform(), t = rng.uniform(), u = rng.uniform();
                if (s + t > 1.)
                {
                    s = 1. - s;
                    t = 1. - t;
                }
                if (t + u > 1.)
                {
                    double tt = u;
                    u = 1. - s - t;
                    t = 1. - tt;
                }
                else if (s + t + u > 1.)
                {
                    double tt = u;
                    u = s + t + u - 1.;
                    s = 1. - t - tt;
                }
                double a = 1. - s - t - u;
                double x = a * v[0][0] + s * v[1][0] + t * v[2][0] + u * v[3][0];
                double y = a * v[0][1] + s * v[1][1] + t * v[2][1] + u * v[3][1];
                double z = a * v[0][2] + s * v[1][2] + t * v[2][2] + u * v[3][2];
                sum += md.geom(x, y, z);
            }
            out[m] = md.nnorm * sum / num_samples;
        }
    });
    return SKG_OK;
}

// TabulatedDustMix::getOpticalProperties + NR::clampedResample (TabulatedDustMix.cpp:12-43,
// NR.hpp:338-434): sigma_abs = mu kappa (1-a) and sigma_sca = mu kappa a resampled log-log,
// g resampled log-lin, values clamped beyond the table; sigma_ext = sigma_abs + sigma_sca.
int skg_dust_meanlist(int32_t nin, const double* lam_in, const double* kappa_in, const double* albedo_in,
                      const double* g_in, double mu, int32_t nout, const double* lam_out, double* sext, double* ssca,
                      double* g)
{
    if (nin < 1 || nout < 1 || !lam_in || !kappa_in || !albedo_in || !g_in || !lam_out || !sext || !ssca || !g)
        return SKG_ERR_ARGUMENT;
    std::vector<double> L(lam_in, lam_in + nin), K(kappa_in, kappa_in + nin), A(albedo_in, albedo_in + nin),
        G(g_in, g_in + nin);
    if (nin > 1 && L[0] > L[nin - 1])
    {
        std::reverse(L.begin(), L.end());
        std::reverse(K.begin(), K.end());
        std::reverse(A.begin(), A.end());
        std::reverse(G.begin(), G.end());
    }
    std::vector<double> sabs(nin), ssc(nin);
    for (int i = 0; i < nin; ++i)
    {
        sabs[i] = mu * K[i] * (1. - A[i]);
        ssc[i] = mu * K[i] * A[i];
    }
    auto locate = [&](double x) {
        // NR::locate
        if (x == L[nin - 1]) return nin - 2;
        int jl = -1, ju = nin;
        while (ju - jl > 1)
        {
            int jm = (ju + jl) >> 1;
            if (x < L[jm])
                ju = jm;
            else
                jl = jm;
        }
        return jl;
    };
    auto loglog = [](double x, double x1, double x2, double f1, double f2) {
        if (f1 <= 0 || f2 <= 0)
        {
            if (x == x1) return f1;
            if (x == x2) return f2;
            return 0.;
        }
        return f1 * std::exp(std::log(x / x1) / std::log(x2 / x1) * (std::log(f2 / f1)));
    };
    auto loglin = [](double x, double x1, double x2, double f1, double f2) {
        if (x1 <= 0 || x2 <= 0) return 0.;
        return f1 + std::log(x / x1) / std::log(x2 / x1) * (f2 - f1);
    };
    auto clamped = [&](double x, const std::vector<double>& y, auto interp) {
        if (nin == 1) return y[0];
        int i = locate(x);
        if (i < 0) return y[0];
        if (i >= nin - 1) return y[nin - 1];
        return interp(x, L[i], L[i + 1], y[i], y[i + 1]);
    };
    for (int l = 0; l < nout; ++l)
    {
        double a = clamped(lam_out[l], sabs, loglog);
        double s = clamped(lam_out[l], ssc, loglog);
        double gg = clamped(lam_out[l], G, loglin);
        if (std::fabs(gg) > 0.999999) gg = std::copysign(0.999999, gg);  // DustMix.cpp:131-146
        ssca[l] = s;
        sext[l] = a + s;
        g[l] = gg;
    }
    return SKG_OK;
}

int skg_octree_upload(skg_engine* e, const skg_octree* t)
{
    if (!e || !t) return SKG_ERR_ARGUMENT;
    int64_t nn = 0, nc = 0, nnb = 0;
    skg_octree_info(t, &nn, &nc, &nnb);
    std::vector<double> bounds(6 * nn);
    std::vector<int32_t> cf(nn), ci(nn), nbi(std::max<int64_t>(nnb, 1));
    std::vector<int64_t> nbo(6 * nn + 1);
    skg_octree_export(t, bounds.data(), cf.data(), ci.data(), nbo.data(), nbi.data());
    return skg_grid_octree(e, nn, bounds.data(), cf.data(), ci.data(), nbo.data(), nbi.data());
}

}  // extern "C"
