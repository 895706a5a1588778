// skirt_ref -- fixture and timing driver linked against the reference SKIRT CPU libraries.
//
// Test infrastructure only: this program is built from the reference sources by
// oracle/ref/Makefile into oracle/_ref/ and is used (a) to write golden vectors under
// tests/golden/ and (b) as the "reference" CPU baseline.  Nothing in the product links it.
//
// Modes
//   skirt_ref run   <ski> <outdir> <threads>          run the simulation as the reference CLI does
//                                                     (SkirtCommandLineHandler.cpp:315-420)
//   skirt_ref paths <ski> <rays.bin> <out.bin> <lambda_m>
//                                                     set up the hierarchy, then trace each ray with
//                                                     SpatialGrid::createPathSegmentGenerator()->next()
//                                                     and compute cumulative / observer optical depths
//                                                     with MediumSystem (MediumSystem.cpp:1785, :2328)
//   skirt_ref grid  <ski> <out.bin>                   dump per-cell number densities and, for tree
//                                                     grids, node bounds, children and the per-wall
//                                                     neighbour lists (replayed with TreeNode/OctTreeNode);
//                                                     for Voronoi and tetrahedral grids their traversal tables
//   skirt_ref cells <ski> <points.bin> <out.bin>      SpatialGrid::cellIndex for each point
//   skirt_ref dust  <ski> <out.bin> <ncos>            dump sectionExt/Sca, asymmpar and peel-off phase
//                                                     values (DustMix.cpp:395-445) at the simulation
//                                                     wavelengths for ncos cosines in [-1,1]
//
// Binary files are little-endian; layouts are documented at each writer below.

#include "Direction.hpp"
#include "FatalError.hpp"
#include "PathSegmentGenerator.hpp"
#include "DustMix.hpp"
#include "FileLog.hpp"
#include "FilePaths.hpp"
#include "MediumSystem.hpp"
#include "MonteCarloSimulation.hpp"
#include "OctTreeNode.hpp"
#include "ParallelFactory.hpp"
#include "PhotonPacket.hpp"
#include "Position.hpp"
#include "ProcessManager.hpp"
#include "SimulationItemRegistry.hpp"
#include "SourceSystem.hpp"
#include "SpatialGrid.hpp"
#include "StringUtils.hpp"
#include "System.hpp"
#include "TreeSpatialGrid.hpp"
#include "TetraMeshSpatialGrid.hpp"
#include "VoronoiMeshSpatialGrid.hpp"
#include "Configuration.hpp"
#include "XmlHierarchyCreator.hpp"
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>

namespace
{
    std::unique_ptr<Item> loadSki(const string& skipath, const string& outpath, int threads)
    {
        auto schema = SimulationItemRegistry::getSchemaDef();
        auto topitem = XmlHierarchyCreator::readFile(schema, skipath);
        auto simulation = dynamic_cast<MonteCarloSimulation*>(topitem.get());
        if (!simulation) throw std::runtime_error("ski file does not hold a MonteCarloSimulation");
        simulation->filePaths()->setOutputPrefix(StringUtils::filenameBase(skipath));
        simulation->filePaths()->setInputPath(StringUtils::dirPath(skipath));
        simulation->filePaths()->setOutputPath(outpath);
        simulation->parallelFactory()->setMaxThreadCount(threads);
        return topitem;
    }

    template<typename T> void put(std::ofstream& out, const T& v) { out.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
    template<typename T> void putv(std::ofstream& out, const std::vector<T>& v)
    {
        if (!v.empty()) out.write(reinterpret_cast<const char*>(v.data()), sizeof(T) * v.size());
    }

    int doRun(const string& ski, const string& outdir, int threads)
    {
        auto top = loadSki(ski, outdir, threads);
        auto simulation = dynamic_cast<MonteCarloSimulation*>(top.get());
        FileLog* log = new FileLog();
        simulation->log()->setLinkedLog(log);
        log->setup();
        auto t0 = std::chrono::steady_clock::now();
        simulation->setupAndRun();
        auto t1 = std::chrono::steady_clock::now();
        std::printf("{\"wall_s\": %.6f, \"threads\": %d}\n", std::chrono::duration<double>(t1 - t0).count(), threads);
        return 0;
    }

    // rays.bin:  int64 n;  n x {rx, ry, rz, kx, ky, kz} float64
    // out.bin :  "SKPT" int32(1) int64 n
    //            int64 off[n+1]; int32 m[off[n]]; float64 ds[off[n]]      raw next() output, incl. empty segments
    //            int64 toff[n+1]; float64 tau[toff[n]]                    cumulative tau per stored (ds>0) segment
    //            float64 tauobs[n]                                        getExtinctionOpticalDepth(pp, inf)
    int doPaths(const string& ski, const string& raysPath, const string& outPath, double lambda)
    {
        auto top = loadSki(ski, "/tmp", 1);
        auto simulation = dynamic_cast<MonteCarloSimulation*>(top.get());
        simulation->log()->setLowestLevel(Log::Level::Warning);
        simulation->setup();
        MediumSystem* ms = simulation->find<MediumSystem>();
        SpatialGrid* grid = ms->grid();

        std::ifstream in(raysPath, std::ios::binary);
        int64_t n = 0;
        in.read(reinterpret_cast<char*>(&n), 8);
        std::vector<double> rays(6 * n);
        in.read(reinterpret_cast<char*>(rays.data()), 8 * rays.size());

        std::vector<int64_t> off{0}, toff{0};
        std::vector<int32_t> mv;
        std::vector<double> dsv, tauv, tauobs;
        auto gen = grid->createPathSegmentGenerator();
        for (int64_t i = 0; i != n; ++i)
        {
            const double* r = &rays[6 * i];
            Position bfr(r[0], r[1], r[2]);
            Direction bfk(r[3], r[4], r[5], false);
            PhotonPacket pp;
            pp.launch(static_cast<size_t>(i), lambda, 1.0, bfr, bfk);
            gen->start(&pp);
            while (gen->next())
            {
                mv.push_back(gen->m());
                dsv.push_back(gen->ds());
            }
            off.push_back(mv.size());
            ms->setExtinctionOpticalDepths(&pp);
            for (const auto& seg : pp.segments()) tauv.push_back(seg.tauExtOrSca());
            toff.push_back(tauv.size());
            PhotonPacket pq;
            pq.launch(static_cast<size_t>(i), lambda, 1.0, bfr, bfk);
            tauobs.push_back(ms->getExtinctionOpticalDepth(&pq, std::numeric_limits<double>::infinity()));
        }
        std::ofstream out(outPath, std::ios::binary);
        out.write("SKPT", 4);
        put<int32_t>(out, 1);
        put<int64_t>(out, n);
        putv(out, off);
        putv(out, mv);
        putv(out, dsv);
        putv(out, toff);
        putv(out, tauv);
        putv(out, tauobs);
        return 0;
    }

    // out.bin: "SKGD" int32(1) int64 ncells; float64 density[ncells]
    //          int32 kind (0 other, 1 tree, 2 Voronoi, 3 tetra); if tree:
    //            int64 nnodes; float64 bounds[6*nnodes]; int32 childBegin[nnodes]; int32 childCount[nnodes];
    //            int64 nchild; int32 childIndex[nchild]; int32 cellIndex[nnodes];
    //            int64 nboff; int64 nbOff[6*nnodes+1]; int32 nbId[nbOff[end]]   (replayed, sorted neighbours)
    //            int32 boundsMatch (1 if the replayed node bounds equal the grid's, bit for bit)
    //          if Voronoi (VoronoiMeshSnapshot.hpp:338-349):
    //            float64 box[6]; float64 eps; int64 nsites; float64 sites[3*nsites];
    //            int32 nbBegin[nsites]; int32 nbCount[nsites]; int64 nnb; int32 nbIndex[nnb];
    //            int32 blockN; int32 blockBegin[N^3]; int32 blockCount[N^3]; int64 nbl; int32 blockIndex[nbl]
    //          if tetra (TetraMeshSpatialGrid.hpp:293-303):
    //            float64 box[6]; float64 eps; int64 nverts; float64 verts[3*nverts]; int64 ntets;
    //            int32 tetVerts[4*ntets]; int32 faceAnchor[4*ntets]; int32 faceNeighbor[4*ntets];
    //            float64 faceNormal[12*ntets]; float64 centroid[3*ntets]
    int doGrid(const string& ski, const string& outPath)
    {
        auto top = loadSki(ski, "/tmp", 1);
        auto simulation = dynamic_cast<MonteCarloSimulation*>(top.get());
        simulation->log()->setLowestLevel(Log::Level::Warning);
        simulation->setup();
        MediumSystem* ms = simulation->find<MediumSystem>();
        SpatialGrid* grid = ms->grid();
        int64_t ncells = ms->numCells();
        std::vector<double> dens(ncells);
        for (int64_t m = 0; m != ncells; ++m) dens[m] = ms->numberDensity(static_cast<int>(m), 0);

        std::ofstream out(outPath, std::ios::binary);
        out.write("SKGD", 4);
        put<int32_t>(out, 1);
        put<int64_t>(out, ncells);
        putv(out, dens);
        auto tree = dynamic_cast<TreeSpatialGrid*>(grid);
        auto vor = dynamic_cast<VoronoiMeshSpatialGrid*>(grid);
        auto tet = dynamic_cast<TetraMeshSpatialGrid*>(grid);
        put<int32_t>(out, tree ? 1 : vor ? 2 : tet ? 3 : 0);
        if (vor || tet)
        {
            auto box = dynamic_cast<BoxSpatialGrid*>(grid)->extent();
            std::vector<double> bv{box.xmin(), box.ymin(), box.zmin(), box.xmax(), box.ymax(), box.zmax()};
            putv(out, bv);
        }
        if (vor)
        {
            put<double>(out, vor->traversalEpsilon());
            const auto& sites = vor->traversalSiteCoordinates();
            put<int64_t>(out, static_cast<int64_t>(sites.size() / 3));
            putv(out, sites);
            putv(out, vor->traversalNeighborBegin());
            putv(out, vor->traversalNeighborCount());
            put<int64_t>(out, static_cast<int64_t>(vor->traversalNeighborIndex().size()));
            putv(out, vor->traversalNeighborIndex());
            put<int32_t>(out, vor->traversalBlockGridN());
            putv(out, vor->traversalBlockBegin());
            putv(out, vor->traversalBlockCount());
            put<int64_t>(out, static_cast<int64_t>(vor->traversalBlockIndex().size()));
            putv(out, vor->traversalBlockIndex());
        }
        if (tet)
        {
            put<double>(out, tet->traversalEpsilon());
            const auto& verts = tet->traversalVertexCoordinates();
            put<int64_t>(out, static_cast<int64_t>(verts.size() / 3));
            putv(out, verts);
            put<int64_t>(out, static_cast<int64_t>(tet->traversalTetraVertexIndices().size() / 4));
            putv(out, tet->traversalTetraVertexIndices());
            putv(out, tet->traversalFaceAnchorIndices());
            putv(out, tet->traversalFaceNeighborIndices());
            putv(out, tet->traversalFaceNormals());
            putv(out, tet->traversalCentroids());
        }
        if (tree)
        {
            const auto& bounds = tree->traversalNodeBounds();
            const auto& cb = tree->traversalChildBegin();
            const auto& cc = tree->traversalChildCount();
            const auto& ci = tree->traversalChildIndex();
            const auto& cell = tree->traversalCellIndex();
            int64_t nnodes = cb.size();

            // replay the level-by-level subdivision (node IDs are assigned in creation order)
            std::vector<TreeNode*> nodev{new OctTreeNode(Box(bounds[0], bounds[1], bounds[2], bounds[3], bounds[4], bounds[5]))};
            for (int64_t l = 0; l < static_cast<int64_t>(nodev.size()); ++l)
                if (cc[l] > 0) nodev[l]->subdivide(nodev);
            for (auto node : nodev) node->sortNeighbors();
            int32_t match = static_cast<int64_t>(nodev.size()) == nnodes;
            for (int64_t l = 0; match && l != nnodes; ++l)
            {
                const TreeNode* nd = nodev[l];
                const double* b = &bounds[6 * l];
                if (nd->xmin() != b[0] || nd->ymin() != b[1] || nd->zmin() != b[2] || nd->xmax() != b[3]
                    || nd->ymax() != b[4] || nd->zmax() != b[5] || nd->id() != l)
                    match = 0;
            }
            std::vector<int64_t> nbOff{0};
            std::vector<int32_t> nbId;
            for (int64_t l = 0; l != nnodes; ++l)
                for (int w = 0; w != 6; ++w)
                {
                    for (auto nb : nodev[l]->neighbors(static_cast<TreeNode::Wall>(w))) nbId.push_back(nb->id());
                    nbOff.push_back(nbId.size());
                }
            put<int64_t>(out, nnodes);
            putv(out, bounds);
            putv(out, cb);
            putv(out, cc);
            put<int64_t>(out, static_cast<int64_t>(ci.size()));
            putv(out, ci);
            putv(out, cell);
            put<int64_t>(out, static_cast<int64_t>(nbOff.size()));
            putv(out, nbOff);
            putv(out, nbId);
            put<int32_t>(out, match);
            for (auto node : nodev) delete node;
        }
        return 0;
    }

    // points.bin: int64 n; n x {x, y, z} float64
    // out.bin   : "SKCI" int32(1) int64 n; int32 cell[n]
    int doCells(const string& ski, const string& pointsPath, const string& outPath)
    {
        auto top = loadSki(ski, "/tmp", 1);
        auto simulation = dynamic_cast<MonteCarloSimulation*>(top.get());
        simulation->log()->setLowestLevel(Log::Level::Warning);
        simulation->setup();
        SpatialGrid* grid = simulation->find<MediumSystem>()->grid();
        std::ifstream in(pointsPath, std::ios::binary);
        int64_t n = 0;
        in.read(reinterpret_cast<char*>(&n), 8);
        std::vector<double> pts(3 * n);
        in.read(reinterpret_cast<char*>(pts.data()), 8 * pts.size());
        std::vector<int32_t> cell(n);
        for (int64_t i = 0; i != n; ++i) cell[i] = grid->cellIndex(Position(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
        std::ofstream out(outPath, std::ios::binary);
        out.write("SKCI", 4);
        put<int32_t>(out, 1);
        put<int64_t>(out, n);
        putv(out, cell);
        return 0;
    }

    // out.bin: "SKDM" int32(1) int64 nlambda int64 ncos
    //          float64 lambda[nl]; float64 sext[nl]; float64 ssca[nl]; float64 g[nl]
    //          float64 cos[ncos]; float64 phase[nl*ncos]   (row-major: lambda, then cosine)
    int doDust(const string& ski, const string& outPath, int ncos)
    {
        auto top = loadSki(ski, "/tmp", 1);
        auto simulation = dynamic_cast<MonteCarloSimulation*>(top.get());
        simulation->log()->setLowestLevel(Log::Level::Warning);
        simulation->setup();
        MediumSystem* ms = simulation->find<MediumSystem>();
        auto mix = dynamic_cast<const DustMix*>(ms->mix(0, 0));
        if (!mix) throw std::runtime_error("first medium is not a dust mix");
        auto config = simulation->find<Configuration>();
        std::vector<double> lam;
        for (double l : config->simulationWavelengths()) lam.push_back(l);
        std::sort(lam.begin(), lam.end());
        int64_t nl = lam.size();
        std::vector<double> sext, ssca, g, cosv, phase;
        for (double l : lam)
        {
            sext.push_back(mix->sectionExt(l));
            ssca.push_back(mix->sectionSca(l));
            g.push_back(mix->asymmpar(l));
        }
        for (int i = 0; i != ncos; ++i) cosv.push_back(-1. + 2. * i / (ncos - 1));
        for (double l : lam)
            for (double c : cosv)
            {
                // incoming direction along +z; observer direction at polar cosine c
                double s = std::sqrt(std::max(0., (1. - c) * (1. + c)));
                PhotonPacket pp;
                pp.launch(0, l, 1.0, Position(0, 0, 0), Direction(0, 0, 1, false));
                double I = 0, Q = 0, U = 0, V = 0, lambda = l;
                mix->peeloffScattering(I, Q, U, V, lambda, Direction(s, 0, c, false), Direction(), nullptr, &pp);
                phase.push_back(I);
            }
        std::ofstream out(outPath, std::ios::binary);
        out.write("SKDM", 4);
        put<int32_t>(out, 1);
        put<int64_t>(out, nl);
        put<int64_t>(out, static_cast<int64_t>(ncos));
        putv(out, lam);
        putv(out, sext);
        putv(out, ssca);
        putv(out, g);
        putv(out, cosv);
        putv(out, phase);
        return 0;
    }
}

int main(int argc, char** argv)
{
    ProcessManager pm(&argc, &argv);
    System system(argc, argv);
    SimulationItemRegistry registry("v9.0-ref", "9");
    if (argc < 2)
    {
        std::fprintf(stderr, "usage: skirt_ref run|paths|grid|cells|dust ...\n");
        return 2;
    }
    string mode = argv[1];
    try
    {
        if (mode == "run" && argc == 5) return doRun(argv[2], argv[3], std::atoi(argv[4]));
        if (mode == "paths" && argc == 6) return doPaths(argv[2], argv[3], argv[4], std::atof(argv[5]));
        if (mode == "grid" && argc == 4) return doGrid(argv[2], argv[3]);
        if (mode == "cells" && argc == 5) return doCells(argv[2], argv[3], argv[4]);
        if (mode == "dust" && argc == 5) return doDust(argv[2], argv[3], std::atoi(argv[4]));
    }
    catch (FatalError& error)
    {
        for (string line : error.message()) std::fprintf(stderr, "%s\n", line.c_str());
        return 1;
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "bad arguments for mode %s\n", mode.c_str());
    return 2;
}
