/* skirt_mi355x.h -- C-ABI of the MI355X photon-packet engine for SKIRT.
 *
 * The engine replaces SKIRT's CPU photon-packet loop behind the seams the reference
 * offloads through `class GpuAcceleration` (SKIRT/core/GpuAcceleration.hpp:30-607):
 *   - spatial grid upload + path segments   (SpatialGrid::createPathSegmentGenerator,
 *                                            SKIRT/core/SpatialGrid.hpp:65-69)
 *   - medium state / dust tables            (MediumSystem setup, MediumSystem.cpp:234-514)
 *   - optical depths along paths            (MediumSystem::setExtinctionOpticalDepths :1785,
 *                                            MediumSystem::getExtinctionOpticalDepth :2328)
 *   - the forced-scattering life cycle      (MonteCarloSimulation::performLifeCycle :614-694)
 *   - frame instruments                     (FrameInstrument::detect, FrameInstrument.cpp:37-41;
 *                                            FluxRecorder::detect, FluxRecorder.cpp:372-536)
 *
 * Conventions (kept from the reference boundary, GpuAcceleration.hpp:27-29):
 *   - every entry point returns SKG_OK (0) or a negative status; on any non-OK status the
 *     caller runs its CPU code path, and skg_last_error() says why;
 *   - the caller owns all host buffers; outputs are written only on SKG_OK;
 *   - one engine object drives one GPU; calls on one engine are serialized internally.
 * Plain C types only: pointers, sizes, doubles.  Units are SI (metres, radians).
 */
#ifndef SKIRT_MI355X_H
#define SKIRT_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SKG_OK 0
#define SKG_ERR_ARGUMENT -1   /* invalid argument or inconsistent sizes */
#define SKG_ERR_DEVICE -2     /* HIP runtime error (allocation, launch, copy) */
#define SKG_ERR_CAPACITY -3   /* an output buffer is too small */
#define SKG_ERR_STATE -4      /* a required upload (grid, medium, ...) is missing */

#define SKG_GRID_CARTESIAN 0
#define SKG_GRID_OCTREE 1
#define SKG_GRID_VORONOI 2
#define SKG_GRID_TETRA 3

#define SKG_MODE_PIPELINE 0   /* wavefront pipeline: event kernel + persistent trace kernel, hipGraph */
#define SKG_MODE_PERSISTENT 1 /* one persistent kernel running whole packets per lane */

#define SKG_FP64 0            /* optical depth and path length sums in FP64 */
#define SKG_FP32_KAHAN 1      /* the same sums in FP32 with compensated summation */

typedef struct skg_engine skg_engine;

/* Creates an engine on HIP device `device`.  Replaces GpuAcceleration's runtime set-up
 * (GpuAcceleration.cpp:10716-10831).  *out is set only on success. */
int skg_create(int device, skg_engine** out);
void skg_destroy(skg_engine* e);
/* Text of the last error on this engine (or of the last failed skg_create when e is NULL);
 * the counterpart of GpuAcceleration::status() (GpuAcceleration.hpp:40-41). */
const char* skg_last_error(const skg_engine* e);

/* ---- spatial grids ------------------------------------------------------------------ */

/* Cartesian grid from its cell borders (nx+1, ny+1, nz+1 increasing values), cell index
 * k + nz*j + nz*ny*i.  Replaces CartesianSpatialGrid::setupSelfAfter (CartesianSpatialGrid.cpp:14-25)
 * for the path; traversal follows CartesianSpatialGrid.cpp:77-153. */
int skg_grid_cartesian(skg_engine* e, const double* xb, int32_t nx, const double* yb, int32_t ny,
                       const double* zb, int32_t nz);

/* Octree from the flattened node table of TreeSpatialGrid (TreeSpatialGrid.cpp:37-70):
 * bounds (6 per node: xmin,ymin,zmin,xmax,ymax,zmax; node 0 is the root = domain),
 * child_first (id of the first of 8 consecutive children, -1 for a leaf, OctTreeNode.cpp:22-33),
 * cell_index (-1 for non-leaves), and per node per wall (BACK,FRONT,LEFT,RIGHT,BOTTOM,TOP)
 * the neighbour lists of TreeNode (TreeNode.cpp:103-112, sorted as TreeNode::sortNeighbors)
 * as CSR: nb_off[6*nnodes+1], nb_id[nb_off[6*nnodes]].  Traversal follows TreeSpatialGrid.cpp:143-228. */
int skg_grid_octree(skg_engine* e, int64_t nnodes, const double* bounds, const int32_t* child_first,
                    const int32_t* cell_index, const int64_t* nb_off, const int32_t* nb_id);

/* Voronoi grid from the traversal tables of VoronoiMeshSpatialGrid (VoronoiMeshSnapshot.hpp:338-349):
 * sites (3 per cell), per-cell neighbour lists nb_begin/nb_count into nb_index (neighbour cell
 * ids; domain walls as -1..-6 = xmin,xmax,ymin,ymax,zmin,zmax), the block search grid
 * (block_n^3 blocks, block_begin/block_count into block_index, VoronoiMeshSnapshot.cpp:783-831)
 * and the domain box (xmin,ymin,zmin,xmax,ymax,zmax).  cellIndex follows
 * VoronoiMeshSnapshot.cpp:1037-1071 and traversal VoronoiMeshSnapshot.cpp:1094-1213. */
int skg_grid_voronoi(skg_engine* e, int64_t nsites, const double* sites, const int32_t* nb_begin,
                     const int32_t* nb_count, const int32_t* nb_index, int32_t block_n, const int32_t* block_begin,
                     const int32_t* block_count, const int32_t* block_index, const double* box);

/* Tetrahedral grid from the traversal tables of TetraMeshSpatialGrid (TetraMeshSpatialGrid.hpp:293-303):
 * vertices (3 per vertex), tet_verts (4 vertex ids per tetra), face_nb (4 per tetra: neighbour
 * across the face opposite vertex f, -1 on the hull), face_normal (12 per tetra: outward unit
 * normals) and the domain box.  cellIndex follows TetraMeshSpatialGrid.cpp:615-624 and
 * traversal TetraMeshSpatialGrid.cpp:697-820, including the final segment before the hull
 * not being returned. */
int skg_grid_tetra(skg_engine* e, int64_t nverts, const double* verts, int64_t ntets, const int32_t* tet_verts,
                   const int32_t* face_nb, const double* face_normal, const double* box);

/* Number of cells of the current grid (SpatialGrid::numCells). */
int64_t skg_num_cells(const skg_engine* e);

/* ---- medium ------------------------------------------------------------------------- */

/* Number density per cell for the single dust medium (MediumState number density,
 * MediumState.hpp:222-266; sampled as MediumSystem.cpp:401-504). */
int skg_medium(skg_engine* e, int64_t ncells, const double* number_density);

/* Dust properties at the simulation's wavelength bins: wavelength (m), extinction and
 * scattering cross sections (m2) and HG asymmetry parameter, as DustMix::sectionExt/
 * sectionSca/asymmpar return them at those wavelengths (DustMix.cpp:319-343). */
int skg_dust(skg_engine* e, int32_t nlambda, const double* lambda, const double* sext, const double* ssca,
             const double* g);

/* ---- path service (the SpatialGrid seam) ---------------------------------------------- */

/* For n rays {rx,ry,rz,kx,ky,kz} (host memory), returns the raw path segments produced by
 * SpatialGrid::createPathSegmentGenerator()->next() (including empty segments, m = -1):
 * off[n+1] offsets, m[total], ds[total].  cap = capacity of m and ds; *total receives the
 * number of segments (also when it exceeds cap, with SKG_ERR_CAPACITY). */
int skg_trace_paths(skg_engine* e, const double* rays, int64_t n, int64_t cap, int64_t* off, int32_t* m,
                    double* ds, int64_t* total);

/* skg_trace_paths without any device allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  rays (n x 6), off (n + 1), m and ds (cap) are DEVICE memory
 * owned by the caller; *total (host) receives the number of segments, also when it exceeds cap
 * (SKG_ERR_CAPACITY).  The offsets are scanned on the device with a fixed 128 KB scratch array
 * per device that the library holds until the process ends: at most 16,777,216 rays per call
 * (larger n fails with SKG_ERR_CAPACITY), and calls that overlap on one GPU race on it.
 * Synchronous on the engine's stream. */
int skg_trace_paths_noalloc(skg_engine* e, const double* rays, int64_t n, int64_t cap, int64_t* off, int32_t* m,
                            double* ds, int64_t* total);

/* Optical depths at wavelength bin `ell` for n rays (host memory):
 *   tau_path[n]  total extinction optical depth of the complete path (the last cumulative
 *                value of MediumSystem::setExtinctionOpticalDepths, MediumSystem.cpp:1785-1810)
 *   tau_obs[n]   MediumSystem::getExtinctionOpticalDepth(pp, inf) for a packet of luminosity L
 *                (early-out +inf at log(L)+745, MediumSystem.cpp:2328-2376)
 * and, when toff is not NULL, the cumulative optical depth after every stored (ds>0) segment
 * (tau_cum at toff[i]..toff[i+1]; toff must be the stored-segment offsets of these rays).
 * fp_mode selects SKG_FP64 or SKG_FP32_KAHAN. */
int skg_trace_tau(skg_engine* e, const double* rays, int64_t n, int32_t ell, double L, int32_t fp_mode,
                  double* tau_path, double* tau_obs, const int64_t* toff, double* tau_cum);

/* ---- instruments -------------------------------------------------------------------- */

/* Adds a distant frame instrument (FrameInstrument.cpp:12-33, DistantInstrument.cpp:13-63);
 * returns its index (>= 0) or a negative status.  At most 1024 instruments. */
int skg_frame_add(skg_engine* e, double inclination, double azimuth, double roll, int32_t nx, int32_t ny,
                  double fovx, double fovy, double cx, double cy);
/* Adds an SED instrument (SEDInstrument, SEDInstrument.cpp:14-25): one value per wavelength
 * bin, the luminosity peeled off toward the direction (inclination, azimuth) from positions
 * whose projection on the sky lies within `radius` (m) of the centre (ApertureInstrument.cpp:
 * 24-43; 0 = no aperture).  It shares the instrument list and indices with the frames
 * (skg_frame_read returns its nlambda values); returns its index or a negative status. */
int skg_sed_add(skg_engine* e, double inclination, double azimuth, double radius);
/* Adds a perspective instrument (PerspectiveInstrument.cpp:22-116): transform is the 4x4
 * world -> (pixel x, pixel y, depth, w) matrix of PerspectiveInstrument::setupSelfBefore in
 * HomogeneousTransform's row-vector convention (row-major, p' = (x, y, z, 1) M), eye the eye
 * position, pixel_size the viewport pixel size.  A packet at r peels off toward the eye; it is
 * recorded in pixel (int(x'/w'), int(y'/w')) when the depth z' exceeds pixel_size / 10, with
 * luminosity / z'^2 attenuated over a path length z'.  Both engines; not combined with
 * instances or the guide.  Returns the index or a negative status. */
int skg_perspective_add(skg_engine* e, const double* transform, const double* eye, double pixel_size, int32_t nx,
                        int32_t ny);
int skg_frame_remove_all(skg_engine* e);
/* Zeroes all frame accumulators. */
int skg_frame_clear(skg_engine* e);
/* Copies the recorded luminosities of instrument `index` (nlambda*ny*nx doubles, index
 * i + nx*j + nx*ny*ell as FluxRecorder.cpp:501) to host memory; the counterpart of
 * FluxRecorder::flush (FluxRecorder.cpp:856-872) before calibration. */
int skg_frame_read(skg_engine* e, int32_t index, double* out, int64_t count);

/* ---- sources and the life cycle (the MonteCarloSimulation seam) ------------------------ */

/* Isotropic point source at (x,y,z) (PointSource.cpp:32-42).  lpack[nlambda] is the launch
 * luminosity of one packet at each wavelength bin (SourceSystem::launch + NormalizedSource::launch
 * with the oligochromatic bias weight, SourceSystem.cpp:101-115, NormalizedSource.cpp:73-110);
 * the bin of each packet is drawn uniformly (OligoWavelengthDistribution.cpp:34-38). */
int skg_source_point(skg_engine* e, double x, double y, double z, const double* lpack);

/* PhotonPacketOptions (pathLengthBias, minWeightReduction, minScattEvents,
 * PhotonPacketOptions.hpp:71-86), the precision mode and the random seed. */
/* Replaces the point source by n imported particles (ImportedSource with a ParticleSnapshot,
 * ImportedSource.cpp:290-322, 466-471, 653-733; ParticleSnapshot.cpp:275-286): xyzh holds
 * x, y, z and the smoothing length h of each particle (m), lbin its luminosity in each
 * wavelength bin (n x nlambda, W), ltot its luminosity over the source wavelength range (the
 * launch weights (1 - source_bias) ltot_m / sum + source_bias / n split the history indices
 * [0, num_packets) into consecutive ranges, one per particle, as prepareForLaunch does).  A
 * packet takes a uniform bin, the luminosity lbin[m][l] * nlambda / (num_packets * W_m), a
 * position r_m + k u h_m (k isotropic; u from the kernel: 0 uniform, u = X^(1/3); 1 cubic
 * spline, the 400-interval table inversion of CubicSplineSmoothingKernel.cpp:78-85) and an
 * isotropic direction.  Persistent engine; not combined with instances or the guide.  n = 0
 * restores the point source. */
int skg_source_particles(skg_engine* e, int64_t n, const double* xyzh, const double* lbin, const double* ltot,
                         double source_bias, int32_t kernel, int64_t num_packets);
int skg_options(skg_engine* e, double path_length_bias, double min_weight_reduction, int32_t min_scatt_events,
                int32_t fp_mode, uint64_t seed);

/* Runs the forced-scattering life cycle of histories [first, first+count) with peel-off to
 * every instrument, accumulating into the frames on the device (MonteCarloSimulation.cpp:614-694,
 * :1492-1544, :2034-2092).  Synchronous. */
int skg_run_packets(skg_engine* e, uint64_t first_history, uint64_t count);

/* Batched independent instances of the configured simulation in one launch (no counterpart
 * in the reference, which runs one simulation per process; its per-simulation inputs are the
 * ones varied here).  Instance i has packets[i] histories with their own random stream
 * (seeds[i]; history j of the instance draws as history j of a single run with that seed),
 * optical depths scaled by tau_scale[i] (sections sext*tau_scale, ssca*tau_scale, as a
 * medium normalised to tau_scale times the optical depth), launch luminosities
 * lpack[i*nlambda .. +nlambda) as skg_source_point's, and peel-off to the frames
 * [frame_begin[i], frame_begin[i]+frame_count[i]) only.  While instances are set,
 * skg_run_packets(first, count) runs histories [first, first+count) of the concatenation of
 * the instances (sum of packets[i] in all), in the persistent engine; n = 0 removes them.
 * Each instance's images equal those of a single run with its parameters. */
int skg_instances(skg_engine* e, int32_t n, const uint64_t* packets, const uint64_t* seeds, const double* tau_scale,
                  const double* lpack, const int32_t* frame_begin, const int32_t* frame_count);

/* Radiation-field storage of the primary emission (MonteCarloSimulation::storeRadiationField,
 * MonteCarloSimulation.cpp:888-927, with MediumSystem::storeRadiationField :2488-2494): while
 * enabled, every forced path adds L * lnmean(e^-tau_end, e^-tau_begin) * ds of each of its
 * stored segments in a cell (SpecialFunctions::lnmean, SpecialFunctions.cpp:860-880) to the
 * table entry m * nlambda + ell (the layout of MediumSystem's _rf1(m, ell)), L being the
 * packet's luminosity at the start of the path and ell its wavelength bin (the oligochromatic
 * radiation-field grid, Configuration.cpp:243).  FP64 accumulation on the device, persistent
 * engine, not combined with instances or the guide.  skg_rf_enable allocates and clears the
 * table for the current grid and wavelengths (on = 0 frees it); skg_rf_read copies it out
 * (count >= ncells * nlambda). */
int skg_rf_enable(skg_engine* e, int32_t on);
int skg_rf_clear(skg_engine* e);
int skg_rf_read(skg_engine* e, double* out, int64_t count);

/* Same, asynchronous on the engine's stream; skg_sync waits for completion. */
int skg_run_packets_async(skg_engine* e, uint64_t first_history, uint64_t count);
int skg_sync(skg_engine* e);

/* Device counters accumulated since the last skg_frame_clear: [0] histories, [1] peel-off
 * traversals, [2] forced-path traversals (first pass), [3] scattering events,
 * [4] segment crossings, [5] FP64 detector atomics issued (after the wavefront combines
 * deposits into the same pixel), [6] second-pass traversals, [7] crossings of the second
 * passes (persistent engine; 0 in the pipeline engine). */
int skg_counters(skg_engine* e, int64_t* out8);
/* The first n (1..16) counters: [0..7] as skg_counters, [8] detector deposits (before the
 * wavefront combines them; [8] / [5] is the combining factor), [9..15] zero. */
int skg_counters_n(skg_engine* e, int64_t* out, int32_t n);

/* Launch shape of the persistent life-cycle kernel: workgroups (0 = automatic), threads per
 * workgroup (64..1024, multiple of 64) and the number of lanes of a wavefront that must be
 * waiting before a wave-wide transition round runs (1..64).  Chosen by the autotuner. */
int skg_tune(skg_engine* e, int32_t workgroups, int32_t threads, int32_t transition_threshold);

/* Autotuner (no counterpart in the reference, which has no launch parameters): times a trial
 * run of `trial_packets` histories (from first_history on; the detector contents, the stored
 * radiation field and the counters are restored afterwards) for each candidate launch
 * configuration -- persistent
 * engine: transition threshold 12/16/20/24/32/48/64 x 4/8 workgroups per CU; pipeline engine: 2^20/2^21/
 * 2^22 packet slots x 8/16 graph iterations x 4/8 trace workgroups per CU -- then times the
 * three fastest again on 4 x trial_packets histories (the end of a launch weighs less there;
 * skipped when batched instances hold too few histories) and keeps the fastest of those.
 * engines: bit 0 persistent, bit 1 pipeline (batched instances, particle sources and
 * radiation-field storage: persistent only).
 * best6 (may be NULL) receives {mode, threshold, workgroups, pool_slots, trace_workgroups,
 * graph_iterations}; best_rate the packets/s of its last timed run.  Not to be called concurrently with other
 * calls on the same engine. */
int skg_autotune(skg_engine* e, uint64_t first_history, uint64_t trial_packets, int32_t engines, int64_t* best6,
                 double* best_rate);

/* Duration in milliseconds of the last life-cycle run (all its kernels), measured with HIP
 * events on the engine's stream. */
double skg_last_kernel_ms(const skg_engine* e);

/* Engine mode and pipeline shape: SKG_MODE_PIPELINE (default) or SKG_MODE_PERSISTENT; the
 * number of packet slots in flight, workgroups of the trace kernel (0 = 8 per compute unit)
 * and pipeline iterations per captured hipGraph (even, 2..256).  Chosen by the autotuner. */
int skg_tune_pipeline(skg_engine* e, int32_t mode, int64_t pool_slots, int32_t trace_workgroups,
                      int32_t graph_iterations);

/* Variance diagnostics (persistent engine): with key != 0, every history continues on the
 * random stream of Philox tag `key` from its first scattering on, so that runs with
 * different keys share each history's emission, emission peel-off, first interaction and
 * first scattering peel-off and continue independently.  The covariance of two such runs'
 * batch means is the variance that no choice of scattering directions can remove (the
 * ceiling of any direction guide); key 0 (the default) restores the single stream.  No
 * reference counterpart (the reference's Random is one thread-local stream). */
int skg_continuation_key(skg_engine* e, uint32_t key);

/* Upper bound on the pipeline's request queue (entries per ping-pong buffer; 0, the default:
 * slots x (1 + observer groups), which holds every request of an iteration).  A smaller queue
 * bounds its device memory; a run whose requests of one iteration do not fit stops and
 * fails with SKG_ERR_DEVICE instead of losing deposits.  No reference counterpart: the
 * reference keeps no request queue (its device buffers are high-water scratch,
 * Docs/SKIRTGPU.md:101). */
int skg_tune_queue(skg_engine* e, int64_t max_requests);

/* Of the last pipeline run: summed duration of the trace kernels (HIP events on the engine's
 * stream), the number of trace launches and of pipeline iterations. */
int skg_last_run_stats(const skg_engine* e, double* trace_ms, int64_t* trace_launches, int64_t* iterations);

/* ---- medium set-up on the host (mirrors the reference set-up before the photon loop) --- */

/* Exponential disk (ExpDiskGeometry, no truncation) optionally decorated with a logarithmic
 * spiral perturbation (SpiralStructureGeometryDecorator), normalised to a face-on optical
 * depth tau_v along Z for a dust of extinction cross section sext_v at the normalisation
 * wavelength (OpticalDepthMaterialNormalization).  Lengths in metres, angles in radians. */
typedef struct skg_disk_model
{
    double scale_length, scale_height;
    int32_t spiral_arms;   /* 0: no spiral perturbation */
    int32_t spiral_index;  /* N of the sin^(2N) arm profile */
    double pitch_angle, radius_zero, phase_zero, perturbation_weight;
    double tau_v, sext_v;
} skg_disk_model;

typedef struct skg_octree skg_octree;

/* Number density of the model at a point (GeometricMedium::numberDensity). */
double skg_disk_density(const skg_disk_model* model, double x, double y, double z);

/* Density-refined octree over box[6] (PolicyTreeSpatialGrid + DensityTreePolicy::constructTree,
 * DensityTreePolicy.cpp:180-241, maxDustFraction criterion of needsSubdivide :82-170), with the
 * neighbour lists maintained as OctTreeNode::addNeighbors (OctTreeNode.cpp:46-136) and sorted as
 * TreeNode::sortNeighbors (TreeNode.cpp:200-207).  num_samples random points per node; threads
 * 0 = all cores.  Results do not depend on the thread count. */
int skg_octree_build(const double* box, const skg_disk_model* model, int32_t min_level, int32_t max_level,
                     double max_dust_fraction, int32_t num_samples, uint64_t seed, int32_t threads,
                     skg_octree** out);
/* The octree of a given topology: node l (in creation order) is subdivided when divide[l] != 0,
 * the order DensityTreePolicy::constructTree subdivides in, so ids, bounds and neighbour lists
 * reproduce the grid the flags describe (e.g. flags from TreeSpatialGrid's child counts). */
int skg_octree_from_flags(const double* box, int64_t nflags, const uint8_t* divide, skg_octree** out);
int skg_octree_info(const skg_octree* t, int64_t* nnodes, int64_t* ncells, int64_t* nneighbors);
/* Flattened tables in the layout skg_grid_octree takes (non-leaf nodes have empty lists). */
int skg_octree_export(const skg_octree* t, double* bounds, int32_t* child_first, int32_t* cell_index,
                      int64_t* nb_off, int32_t* nb_id);
/* skg_grid_octree with the tables of t. */
int skg_octree_upload(skg_engine* e, const skg_octree* t);
void skg_octree_free(skg_octree* t);

/* Voronoi grid built on the host from a site list: VoronoiMeshSnapshot::buildMesh without
 * relaxation (VoronoiMeshSnapshot.cpp:491-751: sites outside the box dropped, sorted by x,
 * near duplicates within eps dropped) with the tessellation computed by clipping each cell
 * with the bisector planes of its nearby sites (the reference uses Voro++), then the block
 * search lists of VoronoiMeshSnapshot::buildSearchPerBlock (:783-831).  Cell m is the m-th
 * retained site in x order. */
typedef struct skg_voronoi skg_voronoi;
int skg_voronoi_build(int64_t nsites, const double* sites, const double* box, int32_t threads, skg_voronoi** out);
int skg_voronoi_info(const skg_voronoi* v, int64_t* ncells, int64_t* nneighbors, int32_t* block_n, int64_t* nblockrefs);
/* Tables in the layout skg_grid_voronoi takes, plus each cell's volume; any pointer may be NULL. */
int skg_voronoi_export(const skg_voronoi* v, double* sites, int32_t* nb_begin, int32_t* nb_count, int32_t* nb_index,
                       int32_t* block_begin, int32_t* block_count, int32_t* block_index, double* volume);
/* skg_grid_voronoi with the tables of v. */
int skg_voronoi_upload(skg_engine* e, const skg_voronoi* v);
/* Cell-averaged lattice field (nx*ny*nz values over the domain, index (i*ny+j)*nz+k, nearest
 * lattice cell) at `samples` random points per cell, drawn as VoronoiMeshSnapshot::generatePosition
 * does (:1005-1023: uniform in the cell's extent, kept if the site is the nearest). */
int skg_sample_density_voronoi(const skg_voronoi* v, const double* lattice, int32_t nx, int32_t ny, int32_t nz,
                               int32_t samples, uint64_t seed, int32_t threads, double* out);
/* The same for the density model: number densities averaged over `samples` positions per
 * cell drawn the same way. */
int skg_sample_density_voronoi_model(const skg_voronoi* v, const skg_disk_model* model, int32_t samples, uint64_t seed,
                                     int32_t threads, double* out);
void skg_voronoi_free(skg_voronoi* v);

/* Tetrahedral grid built on the host: TetraMeshSpatialGrid's vertex preparation (the 8 domain
 * corners added, points outside dropped, sorted by x, near duplicates within eps dropped;
 * TetraMeshSpatialGrid.cpp:300-380) and a Delaunay tetrahedralisation (the reference uses
 * TetGen, :386-452) by Bowyer-Watson insertion with exact predicates on a lattice of 2^23 steps;
 * tetrahedra positively oriented, neighbour f opposite vertex f, outward unit normals as
 * storeTetrahedra (:455-545). */
typedef struct skg_tetra skg_tetra;
int skg_tetra_build(int64_t npoints, const double* points, const double* box, skg_tetra** out);
int skg_tetra_info(const skg_tetra* t, int64_t* nverts, int64_t* ntets);
/* Tables in the layout skg_grid_tetra takes; any pointer may be NULL. */
int skg_tetra_export(const skg_tetra* t, double* verts, int32_t* tet_verts, int32_t* face_nb, double* face_normal);
/* skg_grid_tetra with the tables of t. */
int skg_tetra_upload(skg_engine* e, const skg_tetra* t);
void skg_tetra_free(skg_tetra* t);

/* Cell-averaged number densities from num_samples random points per cell
 * (MediumSystem.cpp:401-504, PropertySampler :129-155). */
int skg_sample_density_cartesian(const double* xb, int32_t nx, const double* yb, int32_t ny, const double* zb,
                                 int32_t nz, const skg_disk_model* model, int32_t num_samples, uint64_t seed,
                                 int32_t threads, double* out);
int skg_sample_density_octree(const skg_octree* t, const skg_disk_model* model, int32_t num_samples,
                              uint64_t seed, int32_t threads, double* out);
/* skg_octree_build on the engine's GPU, level by level (DensityTreePolicy::constructTree,
 * DensityTreePolicy.cpp:180-241): the per-node density estimates (the same random positions
 * per node), the subdivision flags, the children's ids and bounds and every leaf's neighbour
 * lists (the lists OctTreeNode::addNeighbors, OctTreeNode.cpp:46-136, leaves after the
 * level's subdivisions in node order) are computed on the device; lists of more than 16
 * entries are sorted on the host with std::sort, the others on the device.  The tree is
 * returned on the host in the same form as skg_octree_build's. */
int skg_octree_build_gpu(skg_engine* e, const double* box, const skg_disk_model* model, int32_t min_level,
                         int32_t max_level, double max_dust_fraction, int32_t num_samples, uint64_t seed,
                         int32_t threads, skg_octree** out);
/* The same on the GPU for the engine's resident Cartesian, octree or tetrahedral grid: the
 * same random positions per cell as the host samplers above (Philox stream per cell), the
 * model evaluated on the device (results equal to the host's to FP64 rounding of the
 * transcendental functions).  out receives ncells values (count >= ncells). */
int skg_sample_density(skg_engine* e, const skg_disk_model* model, int32_t num_samples, uint64_t seed, double* out,
                       int64_t count);
/* The same for a tetrahedral grid given by its tables (as skg_grid_tetra's), positions drawn
 * uniformly in each tetra as Tetra::generatePosition does (TetraMeshSpatialGrid.cpp:164-173). */
int skg_sample_density_tetra(int64_t ntets, const double* verts, const int32_t* tet_verts, const skg_disk_model* model,
                             int32_t num_samples, uint64_t seed, int32_t threads, double* out);

/* Dust cross sections and asymmetry parameters at the simulation wavelengths from a
 * MeanListDustMix table (TabulatedDustMix::getOpticalProperties, TabulatedDustMix.cpp:12-43;
 * NR::clampedResample log-log for cross sections, log-lin for g; g clamped to +-0.999999). */
int skg_dust_meanlist(int32_t nin, const double* lambda_in, const double* kappa_in, const double* albedo_in,
                      const double* g_in, double mass_per_hydrogen, int32_t nout, const double* lambda_out,
                      double* sext, double* ssca, double* g);

/* ---- learned path guide (no counterpart in the reference) --------------------------------
 * A 9 -> 64 -> 64 -> 64 -> 40 MLP (BF16 MFMA, FP32 accumulation) mapping a scattering event
 * (position / half_size, incoming direction, normalised log wavelength, |r|, |z|) to an 8-lobe
 * von Mises-Fisher mixture of outgoing directions.  In guided mode each scattering samples the
 * defensive mixture (1 - mix) HG + mix guide and multiplies the packet weight by
 * Phi_HG / p_mix, which keeps every estimator unbiased; peel-off is unchanged (it uses Phi_HG).
 * The guide replaces the direction sampling of DustMix::performScattering (DustMix.cpp:496-512). */

#define SKG_GUIDE_PARAMS 11560

/* Parameters in the order W0 (64x9, row-major), b0, W1 (64x64), b1, W2 (64x64), b2,
 * W3 (40x64), b3.  Setting them resets the optimiser state. */
int skg_guide_params(skg_engine* e, const float* theta, int32_t n);
int skg_guide_read_params(skg_engine* e, float* theta, int32_t n);
/* Raw network outputs (n x 40) for n feature rows (n x 9), host memory. */
int skg_guide_infer(skg_engine* e, const float* features, int64_t n, float* out);
/* The life cycle's FP32 guide mixture for n rows of raw network outputs (n x 40 floats): a
 * direction drawn from row i's mixture with the Philox stream (seed, i) into dirs (n x 3), the
 * mixture density per steradian there into pdf (n), and, when at (n x 3 unit vectors) is
 * given, the density at at[i] into pdf_at (n).  No reference counterpart (the guide is new).
 * Currently switched off: returns SKG_ERR_STATE (its first GPU run ended in an abort whose
 * cause has not been found). */
int skg_guide_sample(skg_engine* e, const float* outputs, int64_t n, uint64_t seed, const double* at, double* dirs,
                     double* pdf, double* pdf_at);
/* Gradient of sum_i -w_i ln q(k_i | x_i) with respect to the parameters for n records of 16
 * floats (9 features, direction k (3), weight w, 3 unused), and the loss. */
int skg_guide_gradient(skg_engine* e, const float* records, int64_t n, float* grad, double* loss);
/* mode 0: plain HG; 1: guided (defensive mixture with probability mix of sampling the guide;
 * both engines, with the same results for the same histories); 2: plain HG that records
 * training data at every scattering (pipeline engine only).  half_size, lambda_center (m)
 * and lambda_scale set the feature normalisation. */
int skg_guide_mode(skg_engine* e, int32_t mode, double mix, double half_size, double lambda_center,
                   double lambda_scale);
/* Allocates (or resets) the training-record buffer for mode-2 runs. */
int skg_guide_collect(skg_engine* e, int64_t capacity);
/* Training target of mode-2 runs: the weight of a detector deposit in the records'
 * contributions.  mode 0: the deposited luminosity itself (default).  mode 1: luminosity /
 * max(F, floor) per frame entry (pixel, wavelength), F being the image accumulated so far (run
 * plain packets first) and floor = floor_frac x the median of the positive entries of the
 * entry's (frame, wavelength) plane: every pixel's relative error then counts alike, in
 * place of the frame total.  mode 2: the caller's weights, n = one per frame entry in the
 * layout of skg_frame_read.  No reference counterpart (the guide is new). */
int skg_guide_target(skg_engine* e, int32_t mode, double floor_frac, const double* weights, int64_t n);
/* Number of records of the last mode-2 run; copies up to cap of them (16 floats each; the
 * weight is the detector luminosity the packet deposited after the event) when out != NULL. */
int skg_guide_records(skg_engine* e, int64_t* n, float* out, int64_t cap);
/* Adam on the collected records: steps minibatches of `batch` records (multiple of 64). */
int skg_guide_train(skg_engine* e, int32_t steps, int32_t batch, float lr, double* loss_first, double* loss_last);

#ifdef __cplusplus
}
#endif

#endif /* SKIRT_MI355X_H */
