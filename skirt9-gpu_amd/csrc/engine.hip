// MI355X photon-packet engine: HIP kernels and the C-ABI declared in include/skirt_mi355x.h.
//
// Kernels
//   trace_count / trace_write   path service: raw (m, ds) segments per ray (SpatialGrid seam)
//   trace_tau                   total / observer / cumulative optical depths per ray
//   lifecycle                   persistent forced-scattering life cycle with peel-off
//                               (persistent.hpp)
//   pipe_event / pipe_trace /   wavefront pipeline of the same life cycle (pipeline.hpp)
//   pipe_detect
//   guide_*                     path-guide network on the matrix cores (guide.hpp)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/skirt_mi355x.h"
#include "disk_model.hpp"
#include "lifecycle.hpp"
#include "octree_build.hpp"
#include "octree_gpu.hpp"
#include "persistent.hpp"
#include "pipeline.hpp"
#include "traverse.hpp"

namespace skg
{
    // ------------------------------------------------------------------ path service kernels
    template<int GT> __global__ void trace_count(GridView g, const double* __restrict__ rays, int64_t n, int64_t* counts)
    {
        int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        stage_top(g);
        if (t >= n) return;
        const double* r = rays + 6 * t;
        Gen<GT> gen;
        gen.start(r[0], r[1], r[2], r[3], r[4], r[5]);
        int m;
        double ds, nd;
        int64_t c = 0, cs = 0;
        while (gen.next(g, m, ds, nd))
        {
            ++c;
            if (ds > 0.) ++cs;
        }
        counts[2 * t] = c;
        counts[2 * t + 1] = cs;
    }

    template<int GT>
    __global__ void trace_write(GridView g, const double* __restrict__ rays, int64_t n, const int64_t* off, int32_t* mo,
                                double* dso)
    {
        int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        stage_top(g);
        if (t >= n) return;
        const double* r = rays + 6 * t;
        Gen<GT> gen;
        gen.start(r[0], r[1], r[2], r[3], r[4], r[5]);
        int m;
        double ds, nd;
        int64_t o = off[t];
        while (gen.next(g, m, ds, nd))
        {
            mo[o] = m;
            dso[o] = ds;
            ++o;
        }
    }

    // ------------------------------------------------------------------ path service without allocation
    // skg_trace_paths_noalloc: the segment counts go straight into the caller's offset array
    // and are turned into offsets by a three-kernel scan whose block sums live in `scratch`.
    constexpr int kScanItems = 1024;          // offsets per scan block (256 threads x 4)
    constexpr int kScanMaxBlocks = 16384;     // 16.7 M rays per call at most
    // the scan's block sums and, after the last one, the total; shared by every call on a device
    static __device__ __align__(256) int64_t scratch[kScanMaxBlocks + 1];

    template<int GT> __global__ void trace_count_off(GridView g, const double* __restrict__ rays, int64_t n, int64_t* off)
    {
        int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        stage_top(g);
        if (t == 0) off[0] = 0;
        if (t >= n) return;
        const double* r = rays + 6 * t;
        Gen<GT> gen;
        gen.start(r[0], r[1], r[2], r[3], r[4], r[5]);
        int m;
        double ds, nd;
        int64_t c = 0;
        while (gen.next(g, m, ds, nd)) ++c;
        off[t + 1] = c;
    }

    // inclusive scan of off[1..n] within blocks of kScanItems; block b's sum into sums[b]
    __global__ void __launch_bounds__(256) scan_block_sums(int64_t* off, int64_t n, int64_t* sums)
    {
        __shared__ int64_t part[256];
        const int64_t base = 1 + static_cast<int64_t>(blockIdx.x) * kScanItems + 4 * threadIdx.x;
        int64_t v[4], acc = 0;
        for (int q = 0; q < 4; ++q)
        {
            const int64_t i = base + q;
            v[q] = i <= n ? off[i] : 0;
            acc += v[q];
        }
        part[threadIdx.x] = acc;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1)
        {
            const int64_t x = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
            __syncthreads();
            part[threadIdx.x] += x;
            __syncthreads();
        }
        int64_t run = part[threadIdx.x] - acc;  // exclusive prefix of this thread's items
        for (int q = 0; q < 4; ++q)
        {
            run += v[q];
            const int64_t i = base + q;
            if (i <= n) off[i] = run;
        }
        if (threadIdx.x == 255) sums[blockIdx.x] = part[255];
    }

    // exclusive scan of the nb block sums in place (one workgroup of 1024 threads, up to 16
    // sums each); the total after them
    __global__ void __launch_bounds__(1024) scan_sums(int64_t* sums, int nb)
    {
        __shared__ int64_t part[1024];
        const int per = (nb + 1023) / 1024;
        const int b0 = threadIdx.x * per;
        int64_t acc = 0;
        for (int q = 0; q < per; ++q)
            if (b0 + q < nb) acc += sums[b0 + q];
        part[threadIdx.x] = acc;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1)
        {
            const int64_t x = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
            __syncthreads();
            part[threadIdx.x] += x;
            __syncthreads();
        }
        int64_t run = part[threadIdx.x] - acc;
        for (int q = 0; q < per; ++q)
        {
            if (b0 + q >= nb) break;
            const int64_t v = sums[b0 + q];
            sums[b0 + q] = run;
            run += v;
        }
        if (threadIdx.x == 1023) sums[nb] = part[1023];
    }

    // adds each block's exclusive prefix to its offsets
    __global__ void __launch_bounds__(256) add_block_offsets(int64_t* off, int64_t n, const int64_t* sums)
    {
        const int64_t base = 1 + static_cast<int64_t>(blockIdx.x) * kScanItems + 4 * threadIdx.x;
        const int64_t add = sums[blockIdx.x];
        for (int q = 0; q < 4; ++q)
        {
            const int64_t i = base + q;
            if (i <= n) off[i] += add;
        }
    }

    // copies the cell densities into the cell records (octree leaves, Voronoi cells, tetrahedra)
    template<typename Rec> __global__ void fill_record_density(Rec* rec, const double* dens, int64_t n)
    {
        int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        if (i < n) rec[i].dens = dens[i];
    }

    // and into the header slot of each Voronoi cell's run in the entry table
    __global__ void fill_vor_run_density(VorNb* nb, const VorCell* cells, const double* dens, int64_t n)
    {
        int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        if (i < n) nb[cells[i].begin].x = dens[i];
    }

    template<int GT, int FP>
    __global__ void trace_tau(GridView g, const double* __restrict__ dens, const float* __restrict__ dens32,
                              double section, double taumax, const double* __restrict__ rays, int64_t n,
                              double* tau_path, double* tau_obs, const int64_t* toff, double* tau_cum)
    {
        int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        stage_top(g);
        if (t >= n) return;
        const double* r = rays + 6 * t;
        Gen<GT> gen;
        int m;
        double ds, nd;
        // complete path, cumulative tau over stored (ds > 0) segments
        gen.start(r[0], r[1], r[2], r[3], r[4], r[5]);
        double tau = 0.;
        Kahan32 k32;
        k32.reset();
        const float s32 = static_cast<float>(section);
        int64_t o = toff ? toff[t] : 0;
        const int64_t oend = toff ? toff[t + 1] : 0;
        while (gen.next(g, m, ds, nd))
        {
            if (ds > 0.)
            {
                if (m >= 0)
                {
                    if (FP == SKG_FP64)
                        tau += section * nd * ds;
                    else
                        k32.add(s32 * static_cast<float>(nd) * static_cast<float>(ds));
                }
                if (tau_cum && o < oend) tau_cum[o++] = (FP == SKG_FP64) ? tau : static_cast<double>(k32.sum);
            }
        }
        tau_path[t] = (FP == SKG_FP64) ? tau : static_cast<double>(k32.sum);
        // observer optical depth with early-out (every segment from next() counts)
        gen.start(r[0], r[1], r[2], r[3], r[4], r[5]);
        tau = 0.;
        k32.reset();
        double res = 0.;
        bool inf = false;
        while (gen.next(g, m, ds, nd))
        {
            if (m >= 0)
            {
                if (FP == SKG_FP64)
                {
                    tau += section * nd * ds;
                    if (tau >= taumax)
                    {
                        inf = true;
                        break;
                    }
                }
                else
                {
                    k32.add(s32 * static_cast<float>(nd) * static_cast<float>(ds));
                    if (static_cast<double>(k32.sum) >= taumax)
                    {
                        inf = true;
                        break;
                    }
                }
            }
        }
        res = (FP == SKG_FP64) ? tau : static_cast<double>(k32.sum);
        tau_obs[t] = inf ? INFINITY : res;
    }

    // ------------------------------------------------------------------ cell densities
    constexpr uint32_t kTagDensity = 0x44454e53u;  // the host samplers' stream tag for cell densities

    // Cell-averaged number density of the disk model in every cell of the resident grid from
    // n random points per cell, drawn from the same Philox stream (seed, cell, tag) and placed
    // the same way as the host samplers do (uniform in the Cartesian / octree cell box, x then
    // y then z; folded barycentric weights in a tetra), as MediumSystem's PropertySampler
    // (MediumSystem.cpp:129-155) averages a medium over random positions in each cell.
    template<int GT>
    __global__ void sample_density_kernel(GridView g, DiskModel md, int n, uint64_t seed, int64_t ncells, double* out)
    {
        const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (m >= ncells) return;
        Philox rng;
        rng.init(seed, static_cast<uint64_t>(m), kTagDensity);
        double sum = 0.;
        if constexpr (GT == kTetra)
        {
            const TetCell& T = g.tcell[m];
            for (int q = 0; q < n; ++q)
            {
                double s = rng.uniform(), t = rng.uniform(), u = rng.uniform();
                if (s + t > 1.)
                {
                    s = 1. - s;
                    t = 1. - t;
                }
                if (t + u > 1.)
                {
                    const double tt = u;
                    u = 1. - s - t;
                    t = 1. - tt;
                }
                else if (s + t + u > 1.)
                {
                    const double tt = u;
                    u = s + t + u - 1.;
                    s = 1. - t - tt;
                }
                const double a = 1. - s - t - u;
                const double x = a * T.v[0][0] + s * T.v[1][0] + t * T.v[2][0] + u * T.v[3][0];
                const double y = a * T.v[0][1] + s * T.v[1][1] + t * T.v[2][1] + u * T.v[3][1];
                const double z = a * T.v[0][2] + s * T.v[1][2] + t * T.v[2][2] + u * T.v[3][2];
                sum += md.geom(x, y, z);
            }
            out[m] = md.nnorm * sum / n;
            return;
        }
        double b[6];
        if constexpr (GT == kCartesian)
        {
            // cell index k + nz*j + nz*ny*i (CartesianSpatialGrid.cpp:200-203)
            const int64_t i = m / (static_cast<int64_t>(g.nz) * g.ny), j = (m / g.nz) % g.ny, k = m % g.nz;
            b[0] = g.xb[i];
            b[1] = g.yb[j];
            b[2] = g.zb[k];
            b[3] = g.xb[i + 1];
            b[4] = g.yb[j + 1];
            b[5] = g.zb[k + 1];
        }
        else
        {
            const OctLeaf& L = g.leaf[m];
            for (int a = 0; a < 3; ++a)
            {
                b[a] = L.lo[a];
                b[3 + a] = L.hi[a];
            }
        }
        for (int q = 0; q < n; ++q)
        {
            const double x = b[0] + rng.uniform() * (b[3] - b[0]);
            const double y = b[1] + rng.uniform() * (b[4] - b[1]);
            const double z = b[2] + rng.uniform() * (b[5] - b[2]);
            sum += md.geom(x, y, z);
        }
        out[m] = md.nnorm * (sum / n);
    }

}

namespace skg
{
    // TreeNode::leafChild from the root over the node records (TreeSpatialGrid.cpp:190-219): the
    // cell of the leaf holding (x, y, z), -1 outside the root's closed box (one thread)
    __global__ void octree_locate(GridView g, double x, double y, double z, int* out)
    {
        if (!(x >= g.box[0] && x <= g.box[3] && y >= g.box[1] && y <= g.box[4] && z >= g.box[2] && z <= g.box[5]))
        {
            *out = -1;
            return;
        }
        int nd = 0;
        while (true)
        {
            const TopNode t = g.top[nd];
            if (t.child < 0)
            {
                *out = t.cell;
                return;
            }
            nd = t.child + (x < t.sx ? 0 : 1) + (y < t.sy ? 0 : 2) + (z < t.sz ? 0 : 4);
        }
    }
}

// ==================================================================== host side
using namespace skg;

// The device address of `scratch` on each device, looked up once per device at the first
// engine creation on it (hipGetSymbolAddress with that device current); host code uses this
// address, never the array's name, wherever it needs the array's memory.
namespace
{
    std::mutex g_scratchMu;
    std::vector<int64_t*> g_scratch;  // indexed by device; null until looked up

    hipError_t scratchAddress(int device, int64_t** out)
    {
        std::lock_guard<std::mutex> lk(g_scratchMu);
        if (static_cast<int>(g_scratch.size()) <= device) g_scratch.resize(device + 1, nullptr);
        if (!g_scratch[device])
        {
            void* p = nullptr;
            hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (st != hipSuccess) return st;
            g_scratch[device] = static_cast<int64_t*>(p);
        }
        *out = g_scratch[device];
        return hipSuccess;
    }
}

struct skg_engine
{
    int device = 0;
    int64_t* scan = nullptr;  // this device's `scratch` (the no-allocation path service)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    std::mutex mu;
    int cus = 0;
    // grid
    int gridType = -1;
    GridView grid{};
    int64_t ncells = 0;
    std::vector<void*> gridBufs;
    // medium
    double* dens = nullptr;
    float* dens32 = nullptr;
    int64_t densCells = 0;
    // dust
    int nlambda = 0;
    double *lambda = nullptr, *sext = nullptr, *ssca = nullptr, *gpar = nullptr, *lpack = nullptr;
    std::vector<double> hLambda, hSext, hSsca, hG, hLpack;
    int lpackCount = 0;
    // source / options
    double src[3] = {0, 0, 0};
    bool haveSource = false;
    int srcHint = -1;           // octree: the leaf holding src (start hint of new packets)
    bool srcHintValid = false;  // srcHint belongs to the current grid and source
    double xi = 0.5, mwr = 1e4;
    int minScatt = 0;
    int fpMode = SKG_FP64;
    uint64_t seed = 0;
    // instruments
    struct FrameDesc
    {
        double inc, az, roll;
        int nx, ny;
        double fovx, fovy, cx, cy;
        int sed = 0;
        double radius = 0.;
        int persp = 0;
        double M[16] = {};
        double eye[3] = {};
        double psize = 0.;
    };
    std::vector<FrameDesc> frames;
    double* frameBuf = nullptr;
    FrameGeom* frameGeom = nullptr;
    int64_t frameSize = 0;
    bool frameDirty = true;
    unsigned long long* queue = nullptr;
    unsigned long long* counters = nullptr;
    // radiation field (ncells x nlambda FP64 table) and the shape it was allocated for
    double* rf = nullptr;
    int64_t rfCells = 0;
    int rfLambda = 0;
    // batched instances
    std::vector<Instance> hInst;
    Instance* inst = nullptr;
    double* ilpack = nullptr;
    uint64_t instTotal = 0;
    // particle source (0 entities: the point source)
    int64_t nent = 0;
    int entLambda = 0;  // number of wavelength bins the per-packet luminosities were made for
    int entKernel = 1;
    double* ent = nullptr;
    int64_t* entFirst = nullptr;
    double* entLpack = nullptr;
    double* entX = nullptr;
    // launch shape
    int workgroups = 0, threads = 256, threshold = 32;
    float lastMs = 0.f;
    // wavefront pipeline
    int mode = SKG_MODE_PIPELINE;
    unsigned long long* ctl = nullptr;  // 8 control words followed by the kNumCounters event counters
    unsigned long long* hostCtl = nullptr;
    Pool pool{};
    void* poolMem = nullptr;
    Request* q[2] = {nullptr, nullptr};
    double* qtau[2] = {nullptr, nullptr};
    int64_t qcap = 0;
    int64_t queueLimit = 0;  // 0: the request queue holds slots x (1 + observer groups) entries
    int64_t poolSlots = 1 << 21;
    int traceBlocks = 0;
    int graphIters = 8;
    std::vector<hipEvent_t> traceEv;
    double lastTraceMs = 0.;
    int64_t lastTraceLaunches = 0;
    int64_t lastIterations = 0;
    // path guide
    int guideMode = 0;
    double guideMix = 0.5;
    uint32_t contKey = 0;  // continuation stream tag from the first scattering on (0: off)
    GuideFrame gframe{1.0, 0.0, 1.0};
    float* gTheta = nullptr;  // FP32 parameters, Adam moments and gradient (4 x kGuideParams)
    uint16_t* gBf = nullptr;  // packed BF16 tiles (W and W^T per layer)
    float* gBias = nullptr;   // padded FP32 biases
    GuideNet net{};
    int adamStep = 0;
    float* rec = nullptr;
    uint32_t* recHist = nullptr;
    double* recA = nullptr;
    unsigned long long* recCount = nullptr;
    double* recSum = nullptr;
    int64_t recCap = 0;
    int64_t recN = 0;
    double* contrib = nullptr;
    int64_t* recLast = nullptr;
    int64_t slotCapGuide = 0;
    double* finalLum = nullptr;
    int64_t finalCap = 0;
    // training target: per frame entry (pixel, wavelength) the weight of a deposit in the
    // records' contributions (skg_guide_target); none = the deposited luminosity itself
    double* pixW = nullptr;
    int64_t pixWSize = 0;
};

namespace
{
    thread_local std::string t_createError;

    int fail(skg_engine* e, int code, const std::string& msg)
    {
        if (e) e->err = msg;
        return code;
    }

#define HIPCHK(e, call)                                                                                   \
    do                                                                                                    \
    {                                                                                                     \
        hipError_t _st = (call);                                                                          \
        if (_st != hipSuccess)                                                                            \
            return fail((e), SKG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(_st));         \
    } while (0)

    template<typename T> int upload(skg_engine* e, T** dst, const T* src, size_t n)
    {
        if (*dst)
        {
            hipFree(*dst);
            *dst = nullptr;
        }
        if (n == 0) return SKG_OK;
        HIPCHK(e, hipMalloc(reinterpret_cast<void**>(dst), n * sizeof(T)));
        HIPCHK(e, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
        return SKG_OK;
    }

    void freeGrid(skg_engine* e)
    {
        for (void* p : e->gridBufs) hipFree(p);
        e->gridBufs.clear();
        e->srcHintValid = false;
        e->gridType = -1;
        e->ncells = 0;
    }

    int refreshLeafDensity(skg_engine* e)
    {
        if (e->gridType == SKG_GRID_CARTESIAN || e->gridType < 0 || !e->dens || e->densCells < e->ncells) return SKG_OK;
        int64_t n = e->ncells;
        dim3 grid(static_cast<unsigned>((n + 255) / 256)), block(256);
        if (e->gridType == SKG_GRID_OCTREE)
            hipLaunchKernelGGL(fill_record_density<OctLeaf>, grid, block, 0, e->stream, const_cast<OctLeaf*>(e->grid.leaf), e->dens, n);
        else if (e->gridType == SKG_GRID_VORONOI)
        {
            hipLaunchKernelGGL(fill_record_density<VorCell>, grid, block, 0, e->stream, const_cast<VorCell*>(e->grid.vcell), e->dens, n);
            hipLaunchKernelGGL(fill_vor_run_density, grid, block, 0, e->stream, const_cast<VorNb*>(e->grid.vnb), e->grid.vcell,
                               e->dens, n);
        }
        else
            hipLaunchKernelGGL(fill_record_density<TetCell>, grid, block, 0, e->stream, const_cast<TetCell*>(e->grid.tcell), e->dens, n);
        HIPCHK(e, hipGetLastError());
        HIPCHK(e, hipStreamSynchronize(e->stream));
        return SKG_OK;
    }

    template<typename T> int gridUpload(skg_engine* e, const T* src, size_t n, const T** out)
    {
        T* d = nullptr;
        int st = upload(e, &d, src, n);
        if (st) return st;
        e->gridBufs.push_back(d);
        *out = d;
        return SKG_OK;
    }

    // instrument f shares the observer (direction, shared optical depth) of instrument f - 1:
    // the same inclination, azimuth and roll (FrameInstrument / SEDInstrument), or, for
    // perspective instruments, the same viewport (PerspectiveInstrument::
    // determineSameObserverAsPreceding, PerspectiveInstrument.cpp:83-91)
    bool sameObserverAsPrev(const skg_engine* e, size_t f)
    {
        if (f == 0) return false;
        const auto& d = e->frames[f];
        const auto& q = e->frames[f - 1];
        if (d.persp || q.persp)
            return d.persp && q.persp && std::memcmp(d.M, q.M, sizeof(d.M)) == 0
                   && std::memcmp(d.eye, q.eye, sizeof(d.eye)) == 0 && d.psize == q.psize;
        return d.inc == q.inc && d.az == q.az && d.roll == q.roll;
    }

    void fillFrameGeom(const skg_engine* e, std::vector<FrameGeom>& out)
    {
        int64_t off = 0;
        out.assign(e->frames.size(), FrameGeom{});
        for (size_t f = 0; f < e->frames.size(); ++f)
        {
            const auto& d = e->frames[f];
            FrameGeom& g = out[f];
            g.costheta = std::cos(d.inc);
            g.sintheta = std::sin(d.inc);
            g.cosphi = std::cos(d.az);
            g.sinphi = std::sin(d.az);
            g.cosomega = std::cos(d.roll);
            g.sinomega = std::sin(d.roll);
            g.nx = d.nx;
            g.ny = d.ny;
            g.xpmin = d.cx - 0.5 * d.fovx;
            g.xpsiz = d.fovx / d.nx;
            g.ypmin = d.cy - 0.5 * d.fovy;
            g.ypsiz = d.fovy / d.ny;
            // Direction(theta, phi), Direction.cpp:11-36
            const double eps = 1e-8;
            if (d.inc <= eps)
            {
                g.kobs[0] = 0;
                g.kobs[1] = 0;
                g.kobs[2] = 1;
            }
            else if (d.inc >= M_PI - eps)
            {
                g.kobs[0] = 0;
                g.kobs[1] = 0;
                g.kobs[2] = -1;
            }
            else
            {
                double st = std::sin(d.inc);
                g.kobs[0] = st * std::cos(d.az);
                g.kobs[1] = st * std::sin(d.az);
                g.kobs[2] = std::cos(d.inc);
            }
            g.same_as_prev = sameObserverAsPrev(e, f);
            g.sed = d.sed;
            g.radius2 = d.radius * d.radius;
            g.persp = d.persp;
            std::memcpy(g.M, d.M, sizeof(g.M));
            std::memcpy(g.eye, d.eye, sizeof(g.eye));
            g.psize = d.psize;
            g.offset = off;
            off += static_cast<int64_t>(d.nx) * d.ny * e->nlambda;
        }
    }

    int ensureFrames(skg_engine* e)
    {
        if (!e->frameDirty) return SKG_OK;
        int64_t total = 0;
        for (auto& f : e->frames) total += static_cast<int64_t>(f.nx) * f.ny * std::max(1, e->nlambda);
        if (e->frameBuf) hipFree(e->frameBuf);
        e->frameBuf = nullptr;
        e->frameSize = total;
        if (total) HIPCHK(e, hipMalloc(&e->frameBuf, total * sizeof(double)));
        if (total) HIPCHK(e, hipMemset(e->frameBuf, 0, total * sizeof(double)));
        HIPCHK(e, hipMemset(e->counters, 0, kNumCounters * sizeof(unsigned long long)));
        std::vector<FrameGeom> geo;
        fillFrameGeom(e, geo);
        if (!geo.empty()) HIPCHK(e, hipMemcpy(e->frameGeom, geo.data(), geo.size() * sizeof(FrameGeom), hipMemcpyHostToDevice));
        e->frameDirty = false;
        return SKG_OK;
    }

    template<int GT, int FP>
    hipError_t launchLifecycle(const SimParams& p, bool guided, int blocks, int threads, hipStream_t s)
    {
        // batched instances, guided sampling and radiation-field storage take their own
        // instantiations; the plain run carries none of them
        if (p.ninst)
            hipLaunchKernelGGL((lifecycle<GT, FP, kBatched>), dim3(blocks), dim3(threads), 0, s, p);
        else if (p.rf)
            hipLaunchKernelGGL((lifecycle<GT, FP, kStoreRF>), dim3(blocks), dim3(threads), 0, s, p);
        else if (p.persp)
            hipLaunchKernelGGL((lifecycle<GT, FP, kPersp>), dim3(blocks), dim3(threads), 0, s, p);
        else if (p.nent)
            hipLaunchKernelGGL((lifecycle<GT, FP, kParticles>), dim3(blocks), dim3(threads), 0, s, p);
        else if (guided)
            hipLaunchKernelGGL((lifecycle<GT, FP, kGuided>), dim3(blocks), dim3(threads), 0, s, p);
        else
            hipLaunchKernelGGL((lifecycle<GT, FP, kPlain>), dim3(blocks), dim3(threads), 0, s, p);
        return hipGetLastError();
    }
}

namespace
{
    int checkTraceState(skg_engine* e)
    {
        if (e->gridType < 0) return fail(e, SKG_ERR_STATE, "no grid uploaded");
        return SKG_OK;
    }

    template<typename F> int dispatchGrid(skg_engine* e, F&& f)
    {
        switch (e->gridType)
        {
            case SKG_GRID_CARTESIAN: return f(std::integral_constant<int, kCartesian>());
            case SKG_GRID_OCTREE: return f(std::integral_constant<int, kOctree>());
            case SKG_GRID_VORONOI: return f(std::integral_constant<int, kVoronoi>());
            default: return f(std::integral_constant<int, kTetra>());
        }
    }

    // block search lists for a tetrahedral grid: tbn^3 equal blocks over the domain; a tetra is
    // listed (in increasing index order) in every block its closed bounding box meets
    void buildTetraBlocks(const std::vector<TetCell>& T, const double* box, int& tbn, std::vector<double>& borders,
                          std::vector<int32_t>& off, std::vector<int32_t>& ids)
    {
        const int64_t nt = static_cast<int64_t>(T.size());
        tbn = std::max(1, std::min(256, static_cast<int>(std::cbrt(static_cast<double>(nt)) / 2)));
        borders.resize(3 * static_cast<size_t>(tbn + 1));
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i <= tbn; ++i)
                borders[a * (tbn + 1) + i] = box[a] + (box[3 + a] - box[a]) * i / tbn;
        auto lower = [&](int a, double v) {  // first block whose upper border is >= v
            const double* b = &borders[a * (tbn + 1)];
            int i = static_cast<int>(std::lower_bound(b + 1, b + tbn + 1, v) - (b + 1));
            return std::min(i, tbn - 1);
        };
        auto upper = [&](int a, double v) {  // last block whose lower border is <= v
            const double* b = &borders[a * (tbn + 1)];
            int i = static_cast<int>(std::upper_bound(b, b + tbn, v) - b) - 1;
            return std::max(0, i);
        };
        const int64_t nb3 = static_cast<int64_t>(tbn) * tbn * tbn;
        std::vector<int32_t> cnt(nb3 + 1, 0);
        std::vector<int> rng(6 * nt);
        for (int64_t m = 0; m < nt; ++m)
        {
            for (int a = 0; a < 3; ++a)
            {
                double lo = DBL_MAX, hi = -DBL_MAX;
                for (int t = 0; t < 4; ++t)
                {
                    lo = std::min(lo, T[m].v[t][a]);
                    hi = std::max(hi, T[m].v[t][a]);
                }
                rng[6 * m + 2 * a] = lower(a, lo);
                rng[6 * m + 2 * a + 1] = upper(a, hi);
            }
            for (int i = rng[6 * m]; i <= rng[6 * m + 1]; ++i)
                for (int j = rng[6 * m + 2]; j <= rng[6 * m + 3]; ++j)
                    for (int k = rng[6 * m + 4]; k <= rng[6 * m + 5]; ++k) cnt[(static_cast<int64_t>(i) * tbn + j) * tbn + k + 1]++;
        }
        off.assign(nb3 + 1, 0);
        for (int64_t b = 0; b < nb3; ++b) off[b + 1] = off[b] + cnt[b + 1];
        ids.assign(off[nb3], 0);
        std::vector<int32_t> fill(off.begin(), off.end() - 1);
        for (int64_t m = 0; m < nt; ++m)
            for (int i = rng[6 * m]; i <= rng[6 * m + 1]; ++i)
                for (int j = rng[6 * m + 2]; j <= rng[6 * m + 3]; ++j)
                    for (int k = rng[6 * m + 4]; k <= rng[6 * m + 5]; ++k)
                        ids[fill[(static_cast<int64_t>(i) * tbn + j) * tbn + k]++] = static_cast<int32_t>(m);
    }
}

namespace
{
    // a device buffer that grows on demand (contents not kept) and is freed with its owner
    struct DevBuf
    {
        void* p = nullptr;
        size_t bytes = 0;
        DevBuf() = default;
        DevBuf(const DevBuf&) = delete;
        DevBuf& operator=(const DevBuf&) = delete;
        ~DevBuf()
        {
            if (p) hipFree(p);
        }
        hipError_t ensure(size_t n)
        {
            if (n <= bytes) return hipSuccess;
            if (p) hipFree(p);
            p = nullptr;
            bytes = 0;
            hipError_t st = hipMalloc(&p, n);
            if (st == hipSuccess) bytes = n;
            return st;
        }
        template<typename T> T* as() const { return static_cast<T*>(p); }
    };

    // turns counts in off[1..n] (off[0] = 0) into offsets on the engine's stream; the total
    // into *total
    hipError_t scanCounts(hipStream_t s, int64_t* off, int64_t n, DevBuf& sums, int64_t* total)
    {
        const int64_t nb = (n + kScanItems - 1) / kScanItems;
        hipError_t st = sums.ensure(static_cast<size_t>(nb + 1) * sizeof(int64_t));
        if (st != hipSuccess) return st;
        hipLaunchKernelGGL(scan_block_sums, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, off, n, sums.as<int64_t>());
        hipLaunchKernelGGL(scan_sums, dim3(1), dim3(1024), 0, s, sums.as<int64_t>(), static_cast<int>(nb));
        hipLaunchKernelGGL(add_block_offsets, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, off, n, sums.as<int64_t>());
        st = hipGetLastError();
        if (st == hipSuccess) st = hipMemcpyAsync(total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s);
        if (st == hipSuccess) st = hipStreamSynchronize(s);
        return st;
    }

    unsigned blocksFor(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }
}

extern "C" {

int skg_create(int device, skg_engine** out)
{
    if (!out)
    {
        t_createError = "skg_create: null output pointer";
        return SKG_ERR_ARGUMENT;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    {
        t_createError = "skg_create: no HIP device " + std::to_string(device);
        return SKG_ERR_DEVICE;
    }
    auto* e = new skg_engine();
    e->device = device;
    hipError_t st = hipSetDevice(device);
    if (st == hipSuccess) st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (st == hipSuccess) st = hipEventCreate(&e->ev0);
    if (st == hipSuccess) st = hipEventCreate(&e->ev1);
    if (st == hipSuccess) st = hipMalloc(&e->queue, sizeof(unsigned long long));
    if (st == hipSuccess) st = hipMalloc(&e->ctl, (kCtlCnt + kNumCounters) * sizeof(unsigned long long));
    if (st == hipSuccess) st = hipMemset(e->ctl, 0, (kCtlCnt + kNumCounters) * sizeof(unsigned long long));
    if (st == hipSuccess) e->counters = e->ctl + kCtlCnt;
    if (st == hipSuccess) st = hipHostMalloc(&e->hostCtl, 16 * sizeof(unsigned long long), hipHostMallocDefault);
    if (st == hipSuccess) st = hipMalloc(&e->frameGeom, kMaxFrames * sizeof(FrameGeom));
    if (st == hipSuccess) st = scratchAddress(device, &e->scan);
    hipDeviceProp_t prop;
    if (st == hipSuccess) st = hipGetDeviceProperties(&prop, device);
    if (st != hipSuccess)
    {
        t_createError = std::string("skg_create: ") + hipGetErrorString(st);
        delete e;
        return SKG_ERR_DEVICE;
    }
    e->cus = prop.multiProcessorCount;
    *out = e;
    return SKG_OK;
}

void skg_destroy(skg_engine* e)
{
    if (!e) return;
    hipSetDevice(e->device);
    hipStreamSynchronize(e->stream);
    freeGrid(e);
    for (void* p : {static_cast<void*>(e->dens), static_cast<void*>(e->dens32), static_cast<void*>(e->lambda),
                    static_cast<void*>(e->sext), static_cast<void*>(e->ssca), static_cast<void*>(e->gpar),
                    static_cast<void*>(e->lpack), static_cast<void*>(e->frameBuf), static_cast<void*>(e->frameGeom), static_cast<void*>(e->queue),
                    static_cast<void*>(e->ctl), e->poolMem, static_cast<void*>(e->q[0]), static_cast<void*>(e->q[1]),
                    static_cast<void*>(e->qtau[0]), static_cast<void*>(e->qtau[1]), static_cast<void*>(e->gTheta),
                    static_cast<void*>(e->gBf), static_cast<void*>(e->gBias), static_cast<void*>(e->rec),
                    static_cast<void*>(e->recHist), static_cast<void*>(e->recA), static_cast<void*>(e->recCount),
                    static_cast<void*>(e->recSum), static_cast<void*>(e->contrib), static_cast<void*>(e->recLast),
                    static_cast<void*>(e->finalLum), static_cast<void*>(e->pixW), static_cast<void*>(e->inst), static_cast<void*>(e->ilpack),
                    static_cast<void*>(e->rf), static_cast<void*>(e->ent), static_cast<void*>(e->entFirst),
                    static_cast<void*>(e->entLpack), static_cast<void*>(e->entX)})
        if (p) hipFree(p);
    if (e->hostCtl) hipHostFree(e->hostCtl);
    for (auto ev : e->traceEv) hipEventDestroy(ev);
    hipEventDestroy(e->ev0);
    hipEventDestroy(e->ev1);
    hipStreamDestroy(e->stream);
    delete e;
}

const char* skg_last_error(const skg_engine* e) { return e ? e->err.c_str() : t_createError.c_str(); }

int skg_grid_cartesian(skg_engine* e, const double* xb, int32_t nx, const double* yb, int32_t ny, const double* zb,
                       int32_t nz)
{
    if (!e || !xb || !yb || !zb || nx < 1 || ny < 1 || nz < 1) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_cartesian: bad arguments");
    for (auto [b, n] : {std::pair<const double*, int>{xb, nx}, {yb, ny}, {zb, nz}})
        for (int i = 0; i < n; ++i)
            if (!(b[i] < b[i + 1])) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_cartesian: borders must increase");
    if (static_cast<int64_t>(nx) * ny * nz > INT32_MAX) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_cartesian: too many cells");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    freeGrid(e);
    GridView g{};
    int st;
    if ((st = gridUpload(e, xb, nx + 1, &g.xb)) || (st = gridUpload(e, yb, ny + 1, &g.yb))
        || (st = gridUpload(e, zb, nz + 1, &g.zb)))
        return st;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.box[0] = xb[0];
    g.box[1] = yb[0];
    g.box[2] = zb[0];
    g.box[3] = xb[nx];
    g.box[4] = yb[ny];
    g.box[5] = zb[nz];
    // Box::diagonal (Box.hpp:125-129); eps as CartesianSpatialGrid.cpp:92
    double dx = g.box[3] - g.box[0], dy = g.box[4] - g.box[1], dz = g.box[5] - g.box[2];
    g.eps = 1e-12 * std::sqrt(dx * dx + dy * dy + dz * dz);
    e->grid = g;
    e->gridType = SKG_GRID_CARTESIAN;
    e->ncells = static_cast<int64_t>(nx) * ny * nz;
    return SKG_OK;
}

int skg_grid_octree(skg_engine* e, int64_t nnodes, const double* bounds, const int32_t* child_first,
                    const int32_t* cell_index, const int64_t* nb_off, const int32_t* nb_id)
{
    if (!e || nnodes < 1 || !bounds || !child_first || !cell_index || !nb_off || (!nb_id && nb_off[6 * nnodes] > 0))
        return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: bad arguments");
    if (nnodes > INT32_MAX / 6 || nb_off[6 * nnodes] > INT32_MAX)
        return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: table too large for 32-bit offsets");
    int64_t ncells = 0;
    for (int64_t n = 0; n < nnodes; ++n)
    {
        int32_t c = child_first[n];
        if (c >= 0 && (c + 8 > nnodes || c <= n)) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: bad child index");
        if (c < 0)
        {
            if (cell_index[n] < 0) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: leaf without cell index");
            ncells = std::max<int64_t>(ncells, cell_index[n] + 1);
        }
        else if (cell_index[n] >= 0)
            return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: non-leaf node with a cell index");
    }
    if (ncells > INT32_MAX) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: too many cells");
    // cell-indexed leaf records with the neighbour lists translated to cell ids
    std::vector<OctLeaf> leaves(ncells);
    std::vector<int32_t> nbc;
    nbc.reserve(static_cast<size_t>(nb_off[6 * nnodes]));
    std::vector<char> seen(ncells, 0);
    for (int64_t n = 0; n < nnodes; ++n)
    {
        if (child_first[n] >= 0) continue;
        int32_t m = cell_index[n];
        if (seen[m]) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: duplicate cell index");
        seen[m] = 1;
        OctLeaf& L = leaves[m];
        std::memset(&L, 0, sizeof(L));
        for (int q = 0; q < 3; ++q)
        {
            L.lo[q] = bounds[6 * n + q];
            L.hi[q] = bounds[6 * n + 3 + q];
        }
        L.node = static_cast<int32_t>(n);
        for (int w = 0; w < 6; ++w)
        {
            int64_t b0 = nb_off[6 * n + w], b1 = nb_off[6 * n + w + 1];
            if (b1 < b0) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: offsets must not decrease");
            L.off[w] = static_cast<int32_t>(nbc.size());
            L.first[w] = -1;
            for (int64_t q = b0; q < b1; ++q)
            {
                int32_t id = nb_id[q];
                if (id < 0 || id >= nnodes || cell_index[id] < 0)
                    return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: a leaf's neighbour must be a leaf");
                if (q == b0) L.first[w] = cell_index[id];
                nbc.push_back(cell_index[id]);
            }
        }
        L.off_end = static_cast<int32_t>(nbc.size());
    }
    for (int64_t m = 0; m < ncells; ++m)
        if (!seen[m]) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_octree: cell indices must be 0..ncells-1");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    freeGrid(e);
    GridView g{};
    int st;
    if ((st = gridUpload(e, leaves.data(), leaves.size(), &g.leaf))) return st;
    // the octree step machine loads 112 bytes from any entry: 28 ints of padding
    nbc.resize(nbc.size() + 28, 0);
    if ((st = gridUpload(e, nbc.data(), nbc.size(), &g.nbc))) return st;
    // every node in the search's form; the kernels stage ids below kTopNodes in LDS (the
    // level-by-level construction numbers the top levels first)
    {
        std::vector<TopNode> top(nnodes + 3);  // + 96 bytes: the step machine loads 112 from an entry
        for (int64_t n = 0; n < nnodes; ++n)
        {
            const int32_t c = child_first[n];
            top[n].child = c;
            top[n].cell = c >= 0 ? -1 : cell_index[n];
            top[n].sx = c >= 0 ? bounds[6 * static_cast<int64_t>(c) + 3] : 0.;
            top[n].sy = c >= 0 ? bounds[6 * static_cast<int64_t>(c) + 4] : 0.;
            top[n].sz = c >= 0 ? bounds[6 * static_cast<int64_t>(c) + 5] : 0.;
        }
        for (int64_t n = nnodes; n < nnodes + 3; ++n) top[n] = TopNode{0., 0., 0., -1, 0};
        if ((st = gridUpload(e, top.data(), top.size(), &g.top))) return st;
        g.topN = static_cast<int>(std::min<int64_t>(nnodes, kTopNodes));
    }
    for (int q = 0; q < 6; ++q) g.box[q] = bounds[q];
    double dx = g.box[3] - g.box[0], dy = g.box[4] - g.box[1], dz = g.box[5] - g.box[2];
    g.eps = 1e-12 * std::sqrt(dx * dx + dy * dy + dz * dz);  // TreeSpatialGrid.cpp:28
    e->grid = g;
    e->gridType = SKG_GRID_OCTREE;
    e->ncells = ncells;
    return refreshLeafDensity(e);
}

int skg_grid_voronoi(skg_engine* e, int64_t nsites, const double* sites, const int32_t* nb_begin, const int32_t* nb_count,
                     const int32_t* nb_index, int32_t block_n, const int32_t* block_begin, const int32_t* block_count,
                     const int32_t* block_index, const double* box)
{
    if (!e || nsites < 1 || nsites > INT32_MAX || !sites || !nb_begin || !nb_count || !nb_index || block_n < 1
        || block_n > 1024 || !block_begin || !block_count || !block_index || !box)
        return fail(e, SKG_ERR_ARGUMENT, "skg_grid_voronoi: bad arguments");
    if (!(box[0] < box[3] && box[1] < box[4] && box[2] < box[5])) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_voronoi: empty box");
    std::vector<VorCell> cells(nsites);
    // one run per cell in the entry table: a header slot, then the neighbour entries
    int64_t slots = 0;
    for (int64_t m = 0; m < nsites; ++m)
    {
        if (nb_begin[m] < 0 || nb_count[m] < 0) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_voronoi: bad neighbour range");
        VorCell& c = cells[m];
        std::memset(&c, 0, sizeof(c));
        c.x = sites[3 * m];
        c.y = sites[3 * m + 1];
        c.z = sites[3 * m + 2];
        c.begin = static_cast<int32_t>(std::min<int64_t>(slots, INT32_MAX));
        c.count = nb_count[m];
        slots += 1 + static_cast<int64_t>(nb_count[m]);
        if (slots + 3 > static_cast<int64_t>(INT32_MAX)) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_voronoi: too many neighbours");
    }
    std::vector<VorNb> nbs(static_cast<size_t>(slots + 3));
    std::memset(nbs.data(), 0, nbs.size() * sizeof(VorNb));
    for (int64_t m = 0; m < nsites; ++m)
    {
        const VorCell& c = cells[m];
        VorNb& H = nbs[c.begin];
        H.id = kVorHeader;
        H.run = c.count;  // the density (x) is filled with the medium
        for (int32_t q = 0; q < nb_count[m]; ++q)
        {
            int32_t id = nb_index[static_cast<int64_t>(nb_begin[m]) + q];
            if (id < -6 || id >= nsites) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_voronoi: neighbour id out of range");
            VorNb& E = nbs[static_cast<size_t>(c.begin) + 1 + q];
            E.id = id;
            E.run = id >= 0 ? cells[id].begin : -1;
            if (id >= 0)
            {
                E.x = sites[3 * static_cast<int64_t>(id)];
                E.y = sites[3 * static_cast<int64_t>(id) + 1];
                E.z = sites[3 * static_cast<int64_t>(id) + 2];
            }
        }
    }
    for (int q = 0; q < 3; ++q) nbs[static_cast<size_t>(slots) + q].id = kVorHeader;  // padding
    const int64_t nb3 = static_cast<int64_t>(block_n) * block_n * block_n;
    std::vector<int32_t> boff(nb3 + 1, 0), bids;
    for (int64_t b = 0; b < nb3; ++b)
    {
        if (block_begin[b] < 0 || block_count[b] < 0) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_voronoi: bad block range");
        for (int32_t q = block_begin[b]; q < block_begin[b] + block_count[b]; ++q)
        {
            if (block_index[q] < 0 || block_index[q] >= nsites) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_voronoi: block id out of range");
            bids.push_back(block_index[q]);
        }
        if (bids.size() > static_cast<size_t>(INT32_MAX)) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_voronoi: block lists too long");
        boff[b + 1] = static_cast<int32_t>(bids.size());
    }
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    freeGrid(e);
    GridView g{};
    int st;
    if ((st = gridUpload(e, cells.data(), cells.size(), &g.vcell)) || (st = gridUpload(e, boff.data(), boff.size(), &g.vbl_off)))
        return st;
    if ((st = gridUpload(e, nbs.data(), nbs.size(), &g.vnb))) return st;
    if (!bids.empty() && (st = gridUpload(e, bids.data(), bids.size(), &g.vbl_id))) return st;
    g.vbn = block_n;
    for (int q = 0; q < 6; ++q) g.box[q] = box[q];
    double dx = box[3] - box[0], dy = box[4] - box[1], dz = box[5] - box[2];
    g.eps = 1e-12 * std::sqrt(dx * dx + dy * dy + dz * dz);  // VoronoiMeshSnapshot.cpp:396
    e->grid = g;
    e->gridType = SKG_GRID_VORONOI;
    e->ncells = nsites;
    return refreshLeafDensity(e);
}

int skg_grid_tetra(skg_engine* e, int64_t nverts, const double* verts, int64_t ntets, const int32_t* tet_verts,
                   const int32_t* face_nb, const double* face_normal, const double* box)
{
    if (!e || nverts < 4 || ntets < 1 || ntets > INT32_MAX || !verts || !tet_verts || !face_nb || !face_normal || !box)
        return fail(e, SKG_ERR_ARGUMENT, "skg_grid_tetra: bad arguments");
    if (!(box[0] < box[3] && box[1] < box[4] && box[2] < box[5])) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_tetra: empty box");
    std::vector<TetCell> T(ntets);
    for (int64_t m = 0; m < ntets; ++m)
    {
        TetCell& c = T[m];
        std::memset(&c, 0, sizeof(c));
        for (int t = 0; t < 4; ++t)
        {
            int32_t vi = tet_verts[4 * m + t];
            if (vi < 0 || vi >= nverts) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_tetra: vertex id out of range");
            for (int a = 0; a < 3; ++a)
            {
                c.v[t][a] = verts[3 * static_cast<int64_t>(vi) + a];
                c.n[t][a] = face_normal[12 * m + 3 * t + a];
            }
            int32_t nt = face_nb[4 * m + t];
            if (nt < -1 || nt >= ntets) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_tetra: neighbour id out of range");
            c.nb[t] = nt;
            c.nface[t] = -1;
            // the neighbour's face that points back (TetraMeshSpatialGrid.cpp:508-520)
            if (nt >= 0)
                for (int fn = 0; fn < 4; ++fn)
                    if (face_nb[4 * static_cast<int64_t>(nt) + fn] == m)
                    {
                        c.nface[t] = static_cast<int8_t>(fn);
                        break;
                    }
            if (nt >= 0 && c.nface[t] < 0) return fail(e, SKG_ERR_ARGUMENT, "skg_grid_tetra: neighbour relation not mutual");
        }
    }
    int tbn = 0;
    std::vector<double> borders;
    std::vector<int32_t> off, ids;
    buildTetraBlocks(T, box, tbn, borders, off, ids);
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    freeGrid(e);
    GridView g{};
    int st;
    const double* dB = nullptr;
    if ((st = gridUpload(e, T.data(), T.size(), &g.tcell)) || (st = gridUpload(e, borders.data(), borders.size(), &dB))
        || (st = gridUpload(e, off.data(), off.size(), &g.tbl_off)) || (st = gridUpload(e, ids.data(), ids.size(), &g.tbl_id)))
        return st;
    g.tbn = tbn;
    g.tbx = dB;
    g.tby = dB + (tbn + 1);
    g.tbz = dB + 2 * (tbn + 1);
    for (int q = 0; q < 6; ++q) g.box[q] = box[q];
    double dx = box[3] - box[0], dy = box[4] - box[1], dz = box[5] - box[2];
    g.eps = 1e-12 * std::sqrt(dx * dx + dy * dy + dz * dz);  // TetraMeshSpatialGrid.cpp:239
    e->grid = g;
    e->gridType = SKG_GRID_TETRA;
    e->ncells = ntets;
    return refreshLeafDensity(e);
}

int64_t skg_num_cells(const skg_engine* e) { return e ? e->ncells : 0; }

int skg_medium(skg_engine* e, int64_t ncells, const double* number_density)
{
    if (!e || ncells < 1 || !number_density) return fail(e, SKG_ERR_ARGUMENT, "skg_medium: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    std::vector<float> d32(number_density, number_density + ncells);
    int st;
    if ((st = upload(e, &e->dens, number_density, ncells)) || (st = upload(e, &e->dens32, d32.data(), ncells))) return st;
    e->densCells = ncells;
    return refreshLeafDensity(e);
}

int skg_dust(skg_engine* e, int32_t nlambda, const double* lambda, const double* sext, const double* ssca,
             const double* g)
{
    if (!e || nlambda < 1 || !lambda || !sext || !ssca || !g) return fail(e, SKG_ERR_ARGUMENT, "skg_dust: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    e->hLambda.assign(lambda, lambda + nlambda);
    e->hSext.assign(sext, sext + nlambda);
    e->hSsca.assign(ssca, ssca + nlambda);
    e->hG.assign(g, g + nlambda);
    int st;
    if ((st = upload(e, &e->lambda, lambda, nlambda)) || (st = upload(e, &e->sext, sext, nlambda))
        || (st = upload(e, &e->ssca, ssca, nlambda)) || (st = upload(e, &e->gpar, g, nlambda)))
        return st;
    if (e->nlambda != nlambda) e->frameDirty = true;
    e->nlambda = nlambda;
    return SKG_OK;
}


int skg_trace_paths(skg_engine* e, const double* rays, int64_t n, int64_t cap, int64_t* off, int32_t* m, double* ds,
                    int64_t* total)
{
    if (!e || n < 0 || !off || (n && !rays)) return fail(e, SKG_ERR_ARGUMENT, "skg_trace_paths: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    int st = checkTraceState(e);
    if (st) return st;
    off[0] = 0;
    if (total) *total = 0;
    if (n == 0) return SKG_OK;
    double* dRays = nullptr;
    int64_t* dCnt = nullptr;
    HIPCHK(e, hipMalloc(&dRays, n * 6 * sizeof(double)));
    HIPCHK(e, hipMalloc(&dCnt, n * 2 * sizeof(int64_t)));
    HIPCHK(e, hipMemcpy(dRays, rays, n * 6 * sizeof(double), hipMemcpyHostToDevice));
    int blocks = static_cast<int>((n + 255) / 256);
    GridView g = e->grid;
    g.dens = e->dens;
    dispatchGrid(e, [&](auto gt) {
        hipLaunchKernelGGL((trace_count<decltype(gt)::value>), dim3(blocks), dim3(256), 0, e->stream, g, dRays, n, dCnt);
        return 0;
    });
    std::vector<int64_t> cnt(2 * n);
    HIPCHK(e, hipMemcpyAsync(cnt.data(), dCnt, n * 2 * sizeof(int64_t), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    std::vector<int64_t> hoff(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) hoff[i + 1] = hoff[i] + cnt[2 * i];
    int64_t t = hoff[n];
    if (total) *total = t;
    if (t > cap || (t && (!m || !ds)))
    {
        hipFree(dRays);
        hipFree(dCnt);
        return fail(e, SKG_ERR_CAPACITY, "skg_trace_paths: capacity " + std::to_string(cap) + " < " + std::to_string(t));
    }
    int32_t* dM = nullptr;
    double* dDs = nullptr;
    int64_t* dOff = nullptr;
    if (t)
    {
        HIPCHK(e, hipMalloc(&dM, t * sizeof(int32_t)));
        HIPCHK(e, hipMalloc(&dDs, t * sizeof(double)));
    }
    HIPCHK(e, hipMalloc(&dOff, (n + 1) * sizeof(int64_t)));
    HIPCHK(e, hipMemcpy(dOff, hoff.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (t)
    {
        dispatchGrid(e, [&](auto gt) {
            hipLaunchKernelGGL((trace_write<decltype(gt)::value>), dim3(blocks), dim3(256), 0, e->stream, g, dRays, n, dOff,
                               dM, dDs);
            return 0;
        });
        HIPCHK(e, hipGetLastError());
        HIPCHK(e, hipMemcpyAsync(m, dM, t * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
        HIPCHK(e, hipMemcpyAsync(ds, dDs, t * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    std::memcpy(off, hoff.data(), (n + 1) * sizeof(int64_t));
    hipFree(dRays);
    hipFree(dCnt);
    hipFree(dOff);
    if (dM) hipFree(dM);
    if (dDs) hipFree(dDs);
    return SKG_OK;
}

int skg_trace_paths_noalloc(skg_engine* e, const double* rays, int64_t n, int64_t cap, int64_t* off, int32_t* m,
                            double* ds, int64_t* total)
{
    if (!e) return SKG_ERR_ARGUMENT;
    const int64_t nb = (n + kScanItems - 1) / kScanItems;
    if (nb > kScanMaxBlocks)
        return fail(e, SKG_ERR_CAPACITY, "skg_trace_paths_noalloc: at most " +
                                             std::to_string(static_cast<int64_t>(kScanMaxBlocks) * kScanItems) +
                                             " rays per call");
    if (n < 0 || cap < 0 || (n && (!rays || !off)))
        return fail(e, SKG_ERR_ARGUMENT, "skg_trace_paths_noalloc: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    int st = checkTraceState(e);
    if (st) return st;
    if (total) *total = 0;
    if (n == 0) return SKG_OK;
    GridView g = e->grid;
    g.dens = e->dens;
    const int blocks = static_cast<int>((n + 255) / 256);
    dispatchGrid(e, [&](auto gt) {
        hipLaunchKernelGGL((trace_count_off<decltype(gt)::value>), dim3(blocks), dim3(256), 0, e->stream, g, rays, n, off);
        return 0;
    });
    hipLaunchKernelGGL(scan_block_sums, dim3(static_cast<unsigned>(nb)), dim3(256), 0, e->stream, off, n, e->scan);
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(1024), 0, e->stream, e->scan, static_cast<int>(nb));
    hipLaunchKernelGGL(add_block_offsets, dim3(static_cast<unsigned>(nb)), dim3(256), 0, e->stream, off, n, e->scan);
    HIPCHK(e, hipGetLastError());
    int64_t t = 0;
    HIPCHK(e, hipMemcpyAsync(&t, e->scan + nb, sizeof(int64_t), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if (total) *total = t;
    if (t > cap || (t && (!m || !ds)))
        return fail(e, SKG_ERR_CAPACITY, "skg_trace_paths_noalloc: capacity " + std::to_string(cap) + " < " + std::to_string(t));
    if (t)
    {
        dispatchGrid(e, [&](auto gt) {
            hipLaunchKernelGGL((trace_write<decltype(gt)::value>), dim3(blocks), dim3(256), 0, e->stream, g, rays, n, off, m, ds);
            return 0;
        });
        HIPCHK(e, hipGetLastError());
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    return SKG_OK;
}

int skg_trace_tau(skg_engine* e, const double* rays, int64_t n, int32_t ell, double L, int32_t fp_mode,
                  double* tau_path, double* tau_obs, const int64_t* toff, double* tau_cum)
{
    if (!e || n < 0 || (n && (!rays || !tau_path || !tau_obs)) || (toff && !tau_cum) || !(fp_mode == SKG_FP64 || fp_mode == SKG_FP32_KAHAN))
        return fail(e, SKG_ERR_ARGUMENT, "skg_trace_tau: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    int st = checkTraceState(e);
    if (st) return st;
    if (!e->dens || e->densCells < e->ncells) return fail(e, SKG_ERR_STATE, "skg_trace_tau: no medium for this grid");
    if (ell < 0 || ell >= e->nlambda) return fail(e, SKG_ERR_ARGUMENT, "skg_trace_tau: wavelength bin out of range");
    if (n == 0) return SKG_OK;
    int64_t ncum = toff ? toff[n] : 0;
    double *dRays = nullptr, *dTp = nullptr, *dTo = nullptr, *dCum = nullptr;
    int64_t* dToff = nullptr;
    HIPCHK(e, hipMalloc(&dRays, n * 6 * sizeof(double)));
    HIPCHK(e, hipMalloc(&dTp, n * sizeof(double)));
    HIPCHK(e, hipMalloc(&dTo, n * sizeof(double)));
    HIPCHK(e, hipMemcpy(dRays, rays, n * 6 * sizeof(double), hipMemcpyHostToDevice));
    if (toff)
    {
        HIPCHK(e, hipMalloc(&dToff, (n + 1) * sizeof(int64_t)));
        HIPCHK(e, hipMemcpy(dToff, toff, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
        if (ncum) HIPCHK(e, hipMalloc(&dCum, ncum * sizeof(double)));
    }
    double section = e->hSext[ell];
    double taumax = L > 0 ? std::log(L) + 745 : -INFINITY;
    int blocks = static_cast<int>((n + 255) / 256);
    GridView g = e->grid;
    g.dens = e->dens;
    dispatchGrid(e, [&](auto gt) {
        if (fp_mode == SKG_FP64)
            hipLaunchKernelGGL((trace_tau<decltype(gt)::value, SKG_FP64>), dim3(blocks), dim3(256), 0, e->stream, g, e->dens,
                               e->dens32, section, taumax, dRays, n, dTp, dTo, dToff, dCum);
        else
            hipLaunchKernelGGL((trace_tau<decltype(gt)::value, SKG_FP32_KAHAN>), dim3(blocks), dim3(256), 0, e->stream, g,
                               e->dens, e->dens32, section, taumax, dRays, n, dTp, dTo, dToff, dCum);
        return 0;
    });
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemcpyAsync(tau_path, dTp, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipMemcpyAsync(tau_obs, dTo, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (ncum) HIPCHK(e, hipMemcpyAsync(tau_cum, dCum, ncum * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if (L <= 0)
        for (int64_t i = 0; i < n; ++i) tau_obs[i] = INFINITY;
    hipFree(dRays);
    hipFree(dTp);
    hipFree(dTo);
    if (dToff) hipFree(dToff);
    if (dCum) hipFree(dCum);
    return SKG_OK;
}

int skg_frame_add(skg_engine* e, double inclination, double azimuth, double roll, int32_t nx, int32_t ny, double fovx,
                  double fovy, double cx, double cy)
{
    if (!e || nx < 1 || ny < 1 || !(fovx > 0) || !(fovy > 0) || !std::isfinite(fovx) || !std::isfinite(fovy)
        || !std::isfinite(azimuth) || !std::isfinite(roll) || !std::isfinite(cx) || !std::isfinite(cy))
        return fail(e, SKG_ERR_ARGUMENT, "skg_frame_add: bad arguments");
    if (static_cast<int64_t>(nx) * ny > INT32_MAX) return fail(e, SKG_ERR_ARGUMENT, "skg_frame_add: too many pixels");
    if (inclination < -1e-8 || inclination > M_PI + 1e-8) return fail(e, SKG_ERR_ARGUMENT, "skg_frame_add: inclination outside [0, pi]");
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->frames.size() >= static_cast<size_t>(kMaxFrames)) return fail(e, SKG_ERR_CAPACITY, "skg_frame_add: at most 1024 instruments");
    e->frames.push_back({inclination, azimuth, roll, nx, ny, fovx, fovy, cx, cy});
    e->frameDirty = true;
    return static_cast<int>(e->frames.size()) - 1;
}

int skg_sed_add(skg_engine* e, double inclination, double azimuth, double radius)
{
    if (!e || !(radius >= 0) || !std::isfinite(radius) || !std::isfinite(azimuth))
        return fail(e, SKG_ERR_ARGUMENT, "skg_sed_add: bad arguments");
    if (inclination < -1e-8 || inclination > M_PI + 1e-8) return fail(e, SKG_ERR_ARGUMENT, "skg_sed_add: inclination outside [0, pi]");
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->frames.size() >= static_cast<size_t>(kMaxFrames)) return fail(e, SKG_ERR_CAPACITY, "skg_sed_add: at most 1024 instruments");
    skg_engine::FrameDesc d{inclination, azimuth, 0.0, 1, 1, 1.0, 1.0, 0.0, 0.0};
    d.sed = 1;
    d.radius = radius;
    e->frames.push_back(d);
    e->frameDirty = true;
    return static_cast<int>(e->frames.size()) - 1;
}

int skg_perspective_add(skg_engine* e, const double* transform, const double* eye, double pixel_size, int32_t nx,
                        int32_t ny)
{
    if (!e || !transform || !eye || !(pixel_size > 0) || !std::isfinite(pixel_size) || nx < 1 || ny < 1)
        return fail(e, SKG_ERR_ARGUMENT, "skg_perspective_add: bad arguments");
    for (int q = 0; q < 16; ++q)
        if (!std::isfinite(transform[q])) return fail(e, SKG_ERR_ARGUMENT, "skg_perspective_add: transform not finite");
    for (int q = 0; q < 3; ++q)
        if (!std::isfinite(eye[q])) return fail(e, SKG_ERR_ARGUMENT, "skg_perspective_add: eye not finite");
    if (static_cast<int64_t>(nx) * ny > INT32_MAX) return fail(e, SKG_ERR_ARGUMENT, "skg_perspective_add: too many pixels");
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->frames.size() >= static_cast<size_t>(kMaxFrames)) return fail(e, SKG_ERR_CAPACITY, "skg_perspective_add: at most 1024 instruments");
    skg_engine::FrameDesc d{0.0, 0.0, 0.0, nx, ny, 1.0, 1.0, 0.0, 0.0};
    d.persp = 1;
    std::memcpy(d.M, transform, sizeof(d.M));
    std::memcpy(d.eye, eye, sizeof(d.eye));
    d.psize = pixel_size;
    e->frames.push_back(d);
    e->frameDirty = true;
    return static_cast<int>(e->frames.size()) - 1;
}

int skg_frame_remove_all(skg_engine* e)
{
    if (!e) return SKG_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(e->mu);
    e->frames.clear();
    e->frameDirty = true;
    return SKG_OK;
}

int skg_frame_clear(skg_engine* e)
{
    if (!e) return SKG_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    int st = ensureFrames(e);
    if (st) return st;
    if (e->frameSize) HIPCHK(e, hipMemsetAsync(e->frameBuf, 0, e->frameSize * sizeof(double), e->stream));
    HIPCHK(e, hipMemsetAsync(e->counters, 0, kNumCounters * sizeof(unsigned long long), e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SKG_OK;
}

int skg_frame_read(skg_engine* e, int32_t index, double* out, int64_t count)
{
    if (!e || !out || index < 0) return fail(e, SKG_ERR_ARGUMENT, "skg_frame_read: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (index >= static_cast<int>(e->frames.size())) return fail(e, SKG_ERR_ARGUMENT, "skg_frame_read: no such instrument");
    int st = ensureFrames(e);
    if (st) return st;
    int64_t off = 0;
    for (int f = 0; f < index; ++f) off += static_cast<int64_t>(e->frames[f].nx) * e->frames[f].ny * e->nlambda;
    int64_t sz = static_cast<int64_t>(e->frames[index].nx) * e->frames[index].ny * e->nlambda;
    if (count < sz) return fail(e, SKG_ERR_CAPACITY, "skg_frame_read: output too small");
    HIPCHK(e, hipMemcpyAsync(out, e->frameBuf + off, sz * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SKG_OK;
}

int skg_source_point(skg_engine* e, double x, double y, double z, const double* lpack)
{
    if (!e || !lpack) return fail(e, SKG_ERR_ARGUMENT, "skg_source_point: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (e->nlambda < 1) return fail(e, SKG_ERR_STATE, "skg_source_point: upload the dust wavelengths first");
    e->src[0] = x;
    e->src[1] = y;
    e->src[2] = z;
    e->srcHintValid = false;
    e->hLpack.assign(lpack, lpack + e->nlambda);
    int st = upload(e, &e->lpack, lpack, e->nlambda);
    if (st) return st;
    e->lpackCount = e->nlambda;
    e->haveSource = true;
    return SKG_OK;
}

int skg_source_particles(skg_engine* e, int64_t n, const double* xyzh, const double* lbin, const double* ltot,
                         double source_bias, int32_t kernel, int64_t num_packets)
{
    if (!e) return SKG_ERR_ARGUMENT;
    if (n == 0)
    {
        // back to the point source
        std::lock_guard<std::mutex> lk(e->mu);
        e->nent = 0;
        return SKG_OK;
    }
    if (n < 0 || n > INT32_MAX || !xyzh || !lbin || !ltot || !(source_bias >= 0 && source_bias <= 1)
        || !(kernel == 0 || kernel == 1) || num_packets < 1)
        return fail(e, SKG_ERR_ARGUMENT, "skg_source_particles: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    const int nl = e->nlambda;
    if (nl < 1) return fail(e, SKG_ERR_STATE, "skg_source_particles: upload the dust wavelengths first");
    double sum = 0.;
    for (int64_t m = 0; m < n; ++m)
    {
        for (int q = 0; q < 4; ++q)
            if (!std::isfinite(xyzh[4 * m + q])) return fail(e, SKG_ERR_ARGUMENT, "skg_source_particles: position not finite");
        if (!(xyzh[4 * m + 3] >= 0)) return fail(e, SKG_ERR_ARGUMENT, "skg_source_particles: negative smoothing length");
        if (!(ltot[m] >= 0) || !std::isfinite(ltot[m])) return fail(e, SKG_ERR_ARGUMENT, "skg_source_particles: bad luminosity");
        for (int l = 0; l < nl; ++l)
            if (!(lbin[m * nl + l] >= 0) || !std::isfinite(lbin[m * nl + l]))
                return fail(e, SKG_ERR_ARGUMENT, "skg_source_particles: bad bin luminosity");
        sum += ltot[m];
    }
    if (!(sum > 0)) return fail(e, SKG_ERR_ARGUMENT, "skg_source_particles: zero total luminosity");
    // ImportedSource::prepareForLaunch (ImportedSource.cpp:290-322): launch weights mixing the
    // luminosity fractions with a uniform share, and the first history index of each entity
    std::vector<double> Wv(n);
    for (int64_t m = 0; m < n; ++m) Wv[m] = (1 - source_bias) * (ltot[m] / sum) + source_bias / n;
    std::vector<int64_t> Iv(n + 1);
    Iv[0] = 0;
    double W = 0.;
    for (int64_t m = 1; m < n; ++m)
    {
        W += Wv[m - 1];
        Iv[m] = std::min<int64_t>(num_packets, static_cast<int64_t>(std::round(W * num_packets)));
    }
    Iv[n] = num_packets;
    // per packet of entity m in bin l: L_src / N * (L_m / L_src) / W_m (the selection weight
    // ws) * (L_m,l / L_m) * nlambda (the oligochromatic bias weight s / b with a uniform bin
    // choice, ImportedSource.cpp:670-705); entities without luminosity launch nothing
    std::vector<double> lp(static_cast<size_t>(n) * nl);
    for (int64_t m = 0; m < n; ++m)
        for (int l = 0; l < nl; ++l)
            lp[m * nl + l] = ltot[m] > 0 ? lbin[m * nl + l] * nl / (static_cast<double>(num_packets) * Wv[m]) : 0.;
    // CubicSplineSmoothingKernel::setupSelfBefore (CubicSplineSmoothingKernel.cpp:22-40)
    std::vector<double> X(kSplineN + 1);
    const double du = 1.0 / kSplineN;
    for (int k = 0; k <= kSplineN; ++k)
    {
        double u = k * du;
        double u2 = u * u;
        double u3 = u * u2;
        if (u < 0.5)
            X[k] = u3 * (32.0 / 3.0 - 192.0 / 5.0 * u2 + 32.0 * u3);
        else
            X[k] = -1.0 / 15.0 + u3 * (64.0 / 3.0 - 48.0 * u + 192.0 / 5.0 * u2 - 32.0 / 3.0 * u3);
    }
    int st;
    if ((st = upload(e, &e->ent, xyzh, static_cast<size_t>(4 * n)))) return st;
    if ((st = upload(e, &e->entFirst, Iv.data(), Iv.size()))) return st;
    if ((st = upload(e, &e->entLpack, lp.data(), lp.size()))) return st;
    if ((st = upload(e, &e->entX, X.data(), X.size()))) return st;
    e->nent = n;
    e->entLambda = nl;
    e->entKernel = kernel;
    return SKG_OK;
}

int skg_options(skg_engine* e, double path_length_bias, double min_weight_reduction, int32_t min_scatt_events,
                int32_t fp_mode, uint64_t seed)
{
    if (!e || path_length_bias < 0 || path_length_bias > 1 || !(min_weight_reduction > 0) || min_scatt_events < 0
        || !(fp_mode == SKG_FP64 || fp_mode == SKG_FP32_KAHAN))
        return fail(e, SKG_ERR_ARGUMENT, "skg_options: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    e->xi = path_length_bias;
    e->mwr = min_weight_reduction;
    e->minScatt = min_scatt_events;
    e->fpMode = fp_mode;
    e->seed = seed;
    return SKG_OK;
}

int skg_instances(skg_engine* e, int32_t n, const uint64_t* packets, const uint64_t* seeds, const double* tau_scale,
                  const double* lpack, const int32_t* frame_begin, const int32_t* frame_count)
{
    if (!e || n < 0 || (n > 0 && (!packets || !seeds || !tau_scale || !lpack || !frame_begin || !frame_count)))
        return fail(e, SKG_ERR_ARGUMENT, "skg_instances: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (n > 0 && e->nlambda < 1) return fail(e, SKG_ERR_STATE, "skg_instances: upload the dust wavelengths first");
    std::vector<Instance> h(n);
    uint64_t total = 0;
    for (int i = 0; i < n; ++i)
    {
        if (!(tau_scale[i] >= 0) || !std::isfinite(tau_scale[i]) || frame_begin[i] < 0 || frame_count[i] < 0
            || frame_begin[i] + frame_count[i] > kMaxFrames || packets[i] > (uint64_t(1) << 52) - total)
            return fail(e, SKG_ERR_ARGUMENT, "skg_instances: bad parameters of instance " + std::to_string(i));
        h[i] = Instance{total, seeds[i], tau_scale[i], frame_begin[i], frame_begin[i] + frame_count[i]};
        total += packets[i];
    }
    for (void** b : {reinterpret_cast<void**>(&e->inst), reinterpret_cast<void**>(&e->ilpack)})
        if (*b)
        {
            hipFree(*b);
            *b = nullptr;
        }
    e->hInst.clear();
    e->instTotal = 0;
    if (n == 0) return SKG_OK;
    int st = upload(e, &e->inst, h.data(), h.size());
    if (!st) st = upload(e, &e->ilpack, lpack, static_cast<size_t>(n) * e->nlambda);
    if (st) return st;
    e->hInst = std::move(h);
    e->instTotal = total;
    return SKG_OK;
}

int skg_sample_density(skg_engine* e, const skg_disk_model* model, int32_t num_samples, uint64_t seed, double* out,
                       int64_t count)
{
    if (!e || !model || num_samples < 1 || !out) return fail(e, SKG_ERR_ARGUMENT, "skg_sample_density: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (e->gridType < 0) return fail(e, SKG_ERR_STATE, "skg_sample_density: no grid");
    if (e->gridType == SKG_GRID_VORONOI)
        return fail(e, SKG_ERR_STATE, "skg_sample_density: Voronoi cells are sampled on the host");
    if (count < e->ncells) return fail(e, SKG_ERR_CAPACITY, "skg_sample_density: the output holds fewer values than cells");
    const DiskModel md(*model);
    double* d = nullptr;
    HIPCHK(e, hipMalloc(&d, static_cast<size_t>(e->ncells) * sizeof(double)));
    dim3 grid(static_cast<unsigned>((e->ncells + 255) / 256)), block(256);
    if (e->gridType == SKG_GRID_CARTESIAN)
        hipLaunchKernelGGL(sample_density_kernel<kCartesian>, grid, block, 0, e->stream, e->grid, md, num_samples, seed, e->ncells, d);
    else if (e->gridType == SKG_GRID_OCTREE)
        hipLaunchKernelGGL(sample_density_kernel<kOctree>, grid, block, 0, e->stream, e->grid, md, num_samples, seed, e->ncells, d);
    else
        hipLaunchKernelGGL(sample_density_kernel<kTetra>, grid, block, 0, e->stream, e->grid, md, num_samples, seed, e->ncells, d);
    hipError_t st = hipGetLastError();
    if (st == hipSuccess) st = hipMemcpyAsync(out, d, static_cast<size_t>(e->ncells) * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    hipFree(d);
    if (st != hipSuccess) return fail(e, SKG_ERR_DEVICE, std::string("skg_sample_density: ") + hipGetErrorString(st));
    return SKG_OK;
}

int skg_octree_build_gpu(skg_engine* e, const double* box, const skg_disk_model* model, int32_t min_level,
                         int32_t max_level, double max_dust_fraction, int32_t num_samples, uint64_t seed,
                         int32_t threads, skg_octree** out)
{
    if (!e || !box || !model || !out || min_level < 0 || max_level < min_level || max_level > 20 || num_samples < 1
        || !(box[3] > box[0] && box[4] > box[1] && box[5] > box[2]))
        return fail(e, SKG_ERR_ARGUMENT, "skg_octree_build_gpu: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    const DiskModel md(*model);
    hipStream_t s = e->stream;
    // node tables (bounds, first child, parent), grown by doubling with their contents kept
    DevBuf B, F, P, flags, sums, offA, idA, offB, idB;
    int64_t cap = 0;
    auto grow = [&](int64_t need) -> hipError_t {
        if (need <= cap) return hipSuccess;
        const int64_t nc = std::max<int64_t>(need, std::max<int64_t>(2 * cap, 1 << 16));
        DevBuf b2, f2, p2;
        hipError_t st = b2.ensure(static_cast<size_t>(nc) * 6 * sizeof(double));
        if (st == hipSuccess) st = f2.ensure(static_cast<size_t>(nc) * sizeof(int32_t));
        if (st == hipSuccess) st = p2.ensure(static_cast<size_t>(nc) * sizeof(int32_t));
        if (st == hipSuccess && cap)
        {
            st = hipMemcpyAsync(b2.p, B.p, static_cast<size_t>(cap) * 6 * sizeof(double), hipMemcpyDeviceToDevice, s);
            if (st == hipSuccess) st = hipMemcpyAsync(f2.p, F.p, static_cast<size_t>(cap) * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
            if (st == hipSuccess) st = hipMemcpyAsync(p2.p, P.p, static_cast<size_t>(cap) * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
            if (st == hipSuccess) st = hipStreamSynchronize(s);
        }
        if (st != hipSuccess) return st;
        std::swap(B.p, b2.p);
        std::swap(B.bytes, b2.bytes);
        std::swap(F.p, f2.p);
        std::swap(F.bytes, f2.bytes);
        std::swap(P.p, p2.p);
        std::swap(P.bytes, p2.bytes);
        cap = nc;
        return hipSuccess;
    };
    HIPCHK(e, grow(1));
    const int32_t none = -1;
    HIPCHK(e, hipMemcpyAsync(B.p, box, 6 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(e, hipMemcpyAsync(F.p, &none, sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPCHK(e, hipMemcpyAsync(P.p, &none, sizeof(int32_t), hipMemcpyHostToDevice, s));
    // the root's six empty lists
    DevBuf* oldOff = &offA;
    DevBuf* oldId = &idA;
    DevBuf* newOff = &offB;
    DevBuf* newId = &idB;
    HIPCHK(e, oldOff->ensure(7 * sizeof(int64_t)));
    HIPCHK(e, oldId->ensure(sizeof(int32_t)));
    HIPCHK(e, hipMemsetAsync(oldOff->p, 0, 7 * sizeof(int64_t), s));
    int64_t lbeg = 0, lend = 1, entries = 0;
    for (int level = 0;; ++level)
    {
        // (1) flags of the level's nodes and (2) their children
        const int64_t n = lend - lbeg;
        HIPCHK(e, flags.ensure(static_cast<size_t>(n + 1) * sizeof(int64_t)));
        hipLaunchKernelGGL(tree_flags, dim3(blocksFor(n)), dim3(256), 0, s, md, B.as<double>(), lbeg, n, level, min_level,
                           max_level, max_dust_fraction, num_samples, seed, flags.as<int64_t>());
        HIPCHK(e, hipGetLastError());
        int64_t nflag = 0;
        HIPCHK(e, scanCounts(s, flags.as<int64_t>(), n, sums, &nflag));
        if (nflag == 0) break;
        const int64_t nn = lend + 8 * nflag;
        if (nn > INT32_MAX / 8) return fail(e, SKG_ERR_CAPACITY, "skg_octree_build_gpu: too many nodes");
        HIPCHK(e, grow(nn));
        hipLaunchKernelGGL(tree_children, dim3(blocksFor(n)), dim3(256), 0, s, B.as<double>(), F.as<int32_t>(),
                           P.as<int32_t>(), lbeg, lend, n, flags.as<const int64_t>());
        HIPCHK(e, hipGetLastError());
        // (3) the lists of all nodes after the level
        HIPCHK(e, newOff->ensure(static_cast<size_t>(6 * nn + 1) * sizeof(int64_t)));
        hipLaunchKernelGGL(tree_lists<false>, dim3(blocksFor(6 * nn)), dim3(256), 0, s, B.as<const double>(),
                           F.as<const int32_t>(), P.as<const int32_t>(), lend, nn, oldOff->as<const int64_t>(),
                           oldId->as<const int32_t>(), newOff->as<int64_t>(), static_cast<int32_t*>(nullptr));
        HIPCHK(e, hipGetLastError());
        HIPCHK(e, scanCounts(s, newOff->as<int64_t>(), 6 * nn, sums, &entries));
        if (entries > INT32_MAX) return fail(e, SKG_ERR_CAPACITY, "skg_octree_build_gpu: too many neighbour entries");
        HIPCHK(e, newId->ensure(static_cast<size_t>(std::max<int64_t>(entries, 1)) * sizeof(int32_t)));
        hipLaunchKernelGGL(tree_lists<true>, dim3(blocksFor(6 * nn)), dim3(256), 0, s, B.as<const double>(),
                           F.as<const int32_t>(), P.as<const int32_t>(), lend, nn, oldOff->as<const int64_t>(),
                           oldId->as<const int32_t>(), newOff->as<int64_t>(), newId->as<int32_t>());
        HIPCHK(e, hipGetLastError());
        std::swap(oldOff, newOff);
        std::swap(oldId, newId);
        lbeg = lend;
        lend = nn;
    }
    const int64_t nn = lend;
    hipLaunchKernelGGL(tree_sort_short, dim3(blocksFor(6 * nn)), dim3(256), 0, s, B.as<const double>(),
                       F.as<const int32_t>(), nn, oldOff->as<const int64_t>(), oldId->as<int32_t>());
    HIPCHK(e, hipGetLastError());
    try
    {
        auto t = std::make_unique<skg_octree>();
        t->bounds.resize(6 * nn);
        t->first.resize(nn);
        t->nbOff.resize(6 * nn + 1);
        t->nbId.resize(entries);
        HIPCHK(e, hipMemcpyAsync(t->bounds.data(), B.p, t->bounds.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(e, hipMemcpyAsync(t->first.data(), F.p, t->first.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPCHK(e, hipMemcpyAsync(t->nbOff.data(), oldOff->p, t->nbOff.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        if (entries)
            HIPCHK(e, hipMemcpyAsync(t->nbId.data(), oldId->p, t->nbId.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPCHK(e, hipStreamSynchronize(s));
        // leaves numbered in node order (TreeSpatialGrid.cpp:37-49)
        t->cellOf.assign(nn, -1);
        int32_t m = 0;
        for (int64_t l = 0; l < nn; ++l)
            if (t->first[l] < 0) t->cellOf[l] = m++;
        t->ncells = m;
        skg::octree_sort_lists(*t, 17, threads);
        *out = t.release();
    }
    catch (const std::bad_alloc&)
    {
        return fail(e, SKG_ERR_CAPACITY, "skg_octree_build_gpu: out of host memory");
    }
    return SKG_OK;
}

int skg_rf_enable(skg_engine* e, int32_t on)
{
    if (!e) return SKG_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (e->rf) hipFree(e->rf);
    e->rf = nullptr;
    e->rfCells = 0;
    e->rfLambda = 0;
    if (!on) return SKG_OK;
    if (e->gridType < 0 || e->nlambda < 1) return fail(e, SKG_ERR_STATE, "skg_rf_enable: upload the grid and the dust first");
    const size_t n = static_cast<size_t>(e->ncells) * e->nlambda;
    HIPCHK(e, hipMalloc(&e->rf, n * sizeof(double)));
    HIPCHK(e, hipMemset(e->rf, 0, n * sizeof(double)));
    e->rfCells = e->ncells;
    e->rfLambda = e->nlambda;
    return SKG_OK;
}

int skg_rf_clear(skg_engine* e)
{
    if (!e) return SKG_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (!e->rf) return fail(e, SKG_ERR_STATE, "skg_rf_clear: radiation-field storage is not enabled");
    HIPCHK(e, hipMemsetAsync(e->rf, 0, static_cast<size_t>(e->rfCells) * e->rfLambda * sizeof(double), e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SKG_OK;
}

int skg_rf_read(skg_engine* e, double* out, int64_t count)
{
    if (!e || !out) return SKG_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (!e->rf) return fail(e, SKG_ERR_STATE, "skg_rf_read: radiation-field storage is not enabled");
    const int64_t n = e->rfCells * e->rfLambda;
    if (count < n) return fail(e, SKG_ERR_CAPACITY, "skg_rf_read: the output holds fewer than ncells x nlambda values");
    HIPCHK(e, hipMemcpyAsync(out, e->rf, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SKG_OK;
}

int skg_tune(skg_engine* e, int32_t workgroups, int32_t threads, int32_t transition_threshold)
{
    if (!e || workgroups < 0 || threads < 64 || threads > 256 || threads % 64 || transition_threshold < 1
        || transition_threshold > 64)
        return fail(e, SKG_ERR_ARGUMENT, "skg_tune: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    e->workgroups = workgroups;
    e->threads = threads;
    e->threshold = transition_threshold;
    return SKG_OK;
}

} // extern "C"

static int runPersistent(skg_engine* e, uint64_t first_history, uint64_t count)
{
    SimParams p{};
    p.grid = e->grid;
    p.grid.dens = e->dens;
    p.dens = e->dens;
    p.dens32 = e->dens32;
    p.nlambda = e->nlambda;
    p.lambda = e->lambda;
    p.sext = e->sext;
    p.ssca = e->ssca;
    p.gpar = e->gpar;
    p.lpack = e->lpack;
    for (int q = 0; q < 3; ++q) p.src[q] = e->src[q];
    p.xi = e->xi;
    p.min_weight_reduction = e->mwr;
    p.min_scatt = e->minScatt;
    p.threshold = e->threshold;
    p.seed = e->seed;
    p.first = first_history;
    p.count = count;
    p.queue = e->queue;
    p.counters = e->counters;
    p.frame = e->frameBuf;
    p.frames = e->frameGeom;
    p.nframes = static_cast<int>(e->frames.size());
    p.ninst = static_cast<int>(e->hInst.size());
    p.inst = e->inst;
    p.ilpack = e->ilpack;
    p.net = e->net;
    p.gframe = e->gframe;
    p.guideMix = e->guideMix;
    p.contKey = e->contKey;
    // the leaf of the point source, the start hint of every new packet (octree): found once
    // per grid and source by the top-down search, which the hint check then reproduces
    if (e->gridType == SKG_GRID_OCTREE && !e->srcHintValid)
    {
        int* d = nullptr;
        HIPCHK(e, hipMalloc(&d, sizeof(int)));
        (void)hipGetLastError();  // a launch error below is then this launch's own
        hipLaunchKernelGGL(octree_locate, dim3(1), dim3(1), 0, e->stream, e->grid, e->src[0], e->src[1], e->src[2], d);
        hipError_t st = hipGetLastError();
        if (st == hipSuccess) st = hipMemcpyAsync(&e->srcHint, d, sizeof(int), hipMemcpyDeviceToHost, e->stream);
        if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
        hipFree(d);
        if (st != hipSuccess) return fail(e, SKG_ERR_DEVICE, std::string("source cell: ") + hipGetErrorString(st));
        e->srcHintValid = true;
    }
    p.srcHint = e->gridType == SKG_GRID_OCTREE ? e->srcHint : -1;
    p.rf = e->rf;
    p.persp = 0;
    for (const auto& f : e->frames) p.persp |= f.persp;
    p.nent = static_cast<int>(e->nent);
    p.entKernel = e->entKernel;
    p.ent = e->ent;
    p.entFirst = e->entFirst;
    p.entLpack = e->entLpack;
    p.entX = e->entX;
    if (p.nent && (!e->hInst.empty() || e->guideMode != 0))
        return fail(e, SKG_ERR_STATE, "skg_run_packets: a particle source runs without instances and the guide");
    if (p.nent && e->entLambda != e->nlambda)
        return fail(e, SKG_ERR_STATE,
                    "skg_run_packets: the particle luminosities were uploaded for another number of wavelengths");
    if (p.persp && (!e->hInst.empty() || e->guideMode != 0))
        return fail(e, SKG_ERR_STATE, "skg_run_packets: perspective instruments run without instances and the guide");
    if (e->rf && (e->rfCells != e->ncells || e->rfLambda != e->nlambda))
        return fail(e, SKG_ERR_STATE, "skg_run_packets: the radiation-field table does not match the grid / wavelengths");
    if (e->rf && (!e->hInst.empty() || e->guideMode != 0))
        return fail(e, SKG_ERR_STATE, "skg_run_packets: radiation-field storage runs without instances and without the guide");
    if (e->guideMode == 2)
        return fail(e, SKG_ERR_STATE, "skg_run_packets: training records are collected by the pipeline engine");
    const bool guided = e->guideMode == 1;
    if (guided && !e->gTheta) return fail(e, SKG_ERR_STATE, "guided run without guide parameters");
    HIPCHK(e, hipMemsetAsync(e->queue, 0, sizeof(unsigned long long), e->stream));
    (void)hipGetLastError();
    int threads = e->threads;
    int blocks = e->workgroups > 0 ? e->workgroups : e->cus * 8;
    // never more lanes than packets (rounded up to whole workgroups)
    int64_t needBlocks = static_cast<int64_t>((count + threads - 1) / threads);
    blocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(blocks, needBlocks)));
    HIPCHK(e, hipEventRecord(e->ev0, e->stream));
    hipError_t err;
    dispatchGrid(e, [&](auto gt) {
        constexpr int GT = decltype(gt)::value;
        err = e->fpMode == SKG_FP64 ? launchLifecycle<GT, SKG_FP64>(p, guided, blocks, threads, e->stream)
                                    : launchLifecycle<GT, SKG_FP32_KAHAN>(p, guided, blocks, threads, e->stream);
        return 0;
    });
    if (err != hipSuccess) return fail(e, SKG_ERR_DEVICE, std::string("lifecycle launch: ") + hipGetErrorString(err));
    HIPCHK(e, hipEventRecord(e->ev1, e->stream));
    return SKG_OK;
}


namespace
{
    int ensurePool(skg_engine* e, int64_t slots, int64_t qcap)
    {
        if (slots <= e->pool.capacity && qcap <= e->qcap) return SKG_OK;
        if (e->poolMem) hipFree(e->poolMem);
        if (e->q[0]) hipFree(e->q[0]);
        if (e->q[1]) hipFree(e->q[1]);
        if (e->qtau[0]) hipFree(e->qtau[0]);
        if (e->qtau[1]) hipFree(e->qtau[1]);
        e->poolMem = nullptr;
        e->q[0] = e->q[1] = nullptr;
        e->qtau[0] = e->qtau[1] = nullptr;
        e->pool = Pool{};
        e->qcap = 0;
        // 13 eight-byte and 5 four-byte arrays per slot
        size_t bytes = static_cast<size_t>(slots) * (13 * 8 + 5 * 4) + 18 * 256;
        HIPCHK(e, hipMalloc(&e->poolMem, bytes));
        char* b = static_cast<char*>(e->poolMem);
        auto take = [&](size_t sz) {
            char* r = b;
            b += (static_cast<size_t>(slots) * sz + 255) / 256 * 256;
            return r;
        };
        Pool& P = e->pool;
        P.rx = reinterpret_cast<double*>(take(8));
        P.ry = reinterpret_cast<double*>(take(8));
        P.rz = reinterpret_cast<double*>(take(8));
        P.kx = reinterpret_cast<double*>(take(8));
        P.ky = reinterpret_cast<double*>(take(8));
        P.kz = reinterpret_cast<double*>(take(8));
        P.W = reinterpret_cast<double*>(take(8));
        P.lthr = reinterpret_cast<double*>(take(8));
        P.taupath = reinterpret_cast<double*>(take(8));
        P.tint = reinterpret_cast<double*>(take(8));
        P.res_tau = reinterpret_cast<double*>(take(8));
        P.res_s = reinterpret_cast<double*>(take(8));
        P.hist = reinterpret_cast<uint64_t*>(take(8));
        P.draw = reinterpret_cast<uint32_t*>(take(4));
        P.res_m = reinterpret_cast<int32_t*>(take(4));
        P.ell = reinterpret_cast<int32_t*>(take(4));
        P.nscatt = reinterpret_cast<int32_t*>(take(4));
        P.state = reinterpret_cast<int32_t*>(take(4));
        P.capacity = slots;
        HIPCHK(e, hipMalloc(&e->q[0], qcap * sizeof(Request)));
        HIPCHK(e, hipMalloc(&e->q[1], qcap * sizeof(Request)));
        HIPCHK(e, hipMalloc(&e->qtau[0], qcap * sizeof(double)));
        HIPCHK(e, hipMalloc(&e->qtau[1], qcap * sizeof(double)));
        e->qcap = qcap;
        return SKG_OK;
    }

    template<int GT, int FP> int capturePipeline(skg_engine* e, const PipeParams& p, int eventBlocks, int traceBlocks,
                                                 int iters, hipGraphExec_t* exec)
    {
        const int detectBlocks = e->cus * 4;
        hipGraph_t graph = nullptr;
        HIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < iters; ++i)
        {
            int par = i & 1;
            if (p.guideMode == 1)
                hipLaunchKernelGGL((pipe_event<GT, 1>), dim3(eventBlocks), dim3(256), 0, e->stream, p, par);
            else if (p.guideMode == 2)
                hipLaunchKernelGGL((pipe_event<GT, 2>), dim3(eventBlocks), dim3(256), 0, e->stream, p, par);
            else
                hipLaunchKernelGGL((pipe_event<GT, 0>), dim3(eventBlocks), dim3(256), 0, e->stream, p, par);
            // external event nodes: recorded on every replay, so each trace launch is timed
            hipError_t r1 = hipEventRecordWithFlags(e->traceEv[2 * i], e->stream, hipEventRecordExternal);
            hipLaunchKernelGGL((pipe_trace<GT, FP>), dim3(traceBlocks), dim3(256), 0, e->stream, p, par);
            hipError_t r2 = hipEventRecordWithFlags(e->traceEv[2 * i + 1], e->stream, hipEventRecordExternal);
            if (r1 != hipSuccess || r2 != hipSuccess)
            {
                hipGraph_t g = nullptr;
                hipStreamEndCapture(e->stream, &g);
                if (g) hipGraphDestroy(g);
                (void)hipGetLastError();
                return fail(e, SKG_ERR_DEVICE, "pipeline capture: event record node");
            }
            hipLaunchKernelGGL(pipe_detect, dim3(detectBlocks), dim3(256), 0, e->stream, p, par);
        }
        hipError_t st = hipStreamEndCapture(e->stream, &graph);
        if (st != hipSuccess) return fail(e, SKG_ERR_DEVICE, std::string("pipeline capture: ") + hipGetErrorString(st));
        st = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (st != hipSuccess) return fail(e, SKG_ERR_DEVICE, std::string("pipeline instantiate: ") + hipGetErrorString(st));
        return SKG_OK;
    }

    int runPipeline(skg_engine* e, uint64_t first_history, uint64_t count)
    {
        int ngroups = 0;
        // one peel-off request per observer group per event: the device's grouping rule
        for (size_t f = 0; f < e->frames.size(); ++f)
            if (!sameObserverAsPrev(e, f)) ++ngroups;
        int64_t slots = std::max<int64_t>(256, std::min<int64_t>(e->poolSlots, static_cast<int64_t>(count)));
        slots = (slots + 255) / 256 * 256;
        int64_t qcap = slots * (1 + ngroups);  // every slot's requests of one iteration
        if (e->queueLimit > 0) qcap = std::min(qcap, e->queueLimit);
        int st = ensurePool(e, slots, qcap);
        if (st) return st;
        PipeParams p{};
        p.grid = e->grid;
        p.grid.dens = e->dens;
    p.grid.dens = e->dens;
        p.pool = e->pool;
        p.pool.capacity = slots;
        p.dens = e->dens;
        p.dens32 = e->dens32;
        p.nlambda = e->nlambda;
        p.nframes = static_cast<int>(e->frames.size());
        p.lambda = e->lambda;
        p.sext = e->sext;
        p.ssca = e->ssca;
        p.gpar = e->gpar;
        p.lpack = e->lpack;
        for (int q = 0; q < 3; ++q) p.src[q] = e->src[q];
        p.xi = e->xi;
        p.min_weight_reduction = e->mwr;
        p.min_scatt = e->minScatt;
        p.seed = e->seed;
        p.first = first_history;
        p.count = count;
        p.frame = e->frameBuf;
        p.frames = e->frameGeom;
        p.q[0] = e->q[0];
        p.q[1] = e->q[1];
        p.qtau[0] = e->qtau[0];
        p.qtau[1] = e->qtau[1];
        p.qcap = qcap;
        p.ctl = e->ctl;
        p.guideMode = e->guideMode;
        p.guideMix = e->guideMix;
        p.net = e->net;
        p.gframe = e->gframe;
        if (e->guideMode == 1 && !e->gTheta) return fail(e, SKG_ERR_STATE, "guided run without guide parameters");
        if (e->guideMode == 2)
        {
            if (!e->rec) return fail(e, SKG_ERR_STATE, "collect run without a record buffer (skg_guide_collect)");
            if (e->slotCapGuide < slots)
            {
                if (e->contrib) hipFree(e->contrib);
                if (e->recLast) hipFree(e->recLast);
                e->contrib = nullptr;
                e->recLast = nullptr;
                HIPCHK(e, hipMalloc(&e->contrib, slots * sizeof(double)));
                HIPCHK(e, hipMalloc(&e->recLast, slots * sizeof(int64_t)));
                e->slotCapGuide = slots;
            }
            if (e->finalCap < static_cast<int64_t>(count))
            {
                if (e->finalLum) hipFree(e->finalLum);
                e->finalLum = nullptr;
                HIPCHK(e, hipMalloc(&e->finalLum, count * sizeof(double)));
                e->finalCap = static_cast<int64_t>(count);
            }
            HIPCHK(e, hipMemsetAsync(e->recCount, 0, sizeof(unsigned long long), e->stream));
            p.rec = e->rec;
            p.recHist = e->recHist;
            p.recA = e->recA;
            p.recCount = e->recCount;
            p.recCap = e->recCap;
            p.contrib = e->contrib;
            p.recLast = e->recLast;
            p.finalLum = e->finalLum;
            if (e->pixW && e->pixWSize != e->frameSize)
                return fail(e, SKG_ERR_STATE, "collect run: the guide target's pixel weights do not match the frames");
            p.pixW = e->pixW;
        }
        const int iters = std::max(2, e->graphIters & ~1);
        while (static_cast<int>(e->traceEv.size()) < 2 * iters)
        {
            hipEvent_t ev;
            HIPCHK(e, hipEventCreate(&ev));
            e->traceEv.push_back(ev);
        }
        const int eventBlocks = static_cast<int>((slots + 255) / 256);
        const int traceBlocks = e->traceBlocks > 0 ? e->traceBlocks : e->cus * 8;
        hipGraphExec_t exec = nullptr;
        st = dispatchGrid(e, [&](auto gt) {
            constexpr int GT = decltype(gt)::value;
            return e->fpMode == SKG_FP64 ? capturePipeline<GT, SKG_FP64>(e, p, eventBlocks, traceBlocks, iters, &exec)
                                         : capturePipeline<GT, SKG_FP32_KAHAN>(e, p, eventBlocks, traceBlocks, iters, &exec);
        });
        if (st) return st;
        HIPCHK(e, hipEventRecord(e->ev0, e->stream));
        hipLaunchKernelGGL(pipe_reset, dim3(eventBlocks), dim3(256), 0, e->stream, p.pool, e->ctl);
        double traceMs = 0.;
        int64_t launches = 0;
        int64_t rounds = 0;
        // every packet needs at most ~2 (1 + scatterings) iterations; the bound only stops a
        // runaway loop (e.g. non-terminating weights), never a regular run
        const int64_t maxRounds = 1 + (static_cast<int64_t>(count) / slots + 1) * ((1 << 21) / iters + 64);
        hipError_t hst = hipSuccess;
        while (true)
        {
            hst = hipGraphLaunch(exec, e->stream);
            if (hst != hipSuccess) break;
            ++rounds;
            hst = hipMemcpyAsync(e->hostCtl, e->ctl, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, e->stream);
            if (hst != hipSuccess) break;
            hst = hipStreamSynchronize(e->stream);
            if (hst != hipSuccess) break;
            for (int i = 0; i < iters; ++i)
            {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, e->traceEv[2 * i], e->traceEv[2 * i + 1]) == hipSuccess) traceMs += ms;
            }
            (void)hipGetLastError();  // a failed timing query must not leave a sticky error
            launches += iters;
            unsigned long long active = e->hostCtl[kCtlActive + ((iters - 1) & 1)];
            if (active == 0 || e->hostCtl[kCtlOverflow]) break;
            if (rounds > maxRounds)
            {
                hipGraphExecDestroy(exec);
                return fail(e, SKG_ERR_DEVICE, "pipeline did not drain");
            }
        }
        hipGraphExecDestroy(exec);
        if (hst != hipSuccess) return fail(e, SKG_ERR_DEVICE, std::string("pipeline: ") + hipGetErrorString(hst));
        if (e->hostCtl[kCtlOverflow])
            return fail(e, SKG_ERR_DEVICE, "pipeline: the request queue overflowed (requests were dropped)");
        HIPCHK(e, hipEventRecord(e->ev1, e->stream));
        if (e->guideMode == 2)
        {
            unsigned long long nrec = 0;
            HIPCHK(e, hipMemcpy(&nrec, e->recCount, sizeof(nrec), hipMemcpyDeviceToHost));
            // each collect run writes its records from index 0 (the record counter is cleared at
            // its start): all of them take this run's totals
            int64_t n = std::min<int64_t>(static_cast<int64_t>(nrec), e->recCap);
            if (n > 0)
                hipLaunchKernelGGL(guide_finalize_records, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                                   e->stream, e->rec, e->recHist, e->recA, e->finalLum, n);
            HIPCHK(e, hipStreamSynchronize(e->stream));
            e->recN = n;
        }
        e->lastTraceMs = traceMs;
        e->lastTraceLaunches = launches;
        e->lastIterations = rounds * iters;
        return SKG_OK;
    }
}

int skg_run_packets_async(skg_engine* e, uint64_t first_history, uint64_t count)
{
    if (!e) return SKG_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (e->gridType < 0) return fail(e, SKG_ERR_STATE, "skg_run_packets: no grid");
    if (!e->dens || e->densCells < e->ncells) return fail(e, SKG_ERR_STATE, "skg_run_packets: no medium for this grid");
    if (e->nlambda < 1 || !e->haveSource) return fail(e, SKG_ERR_STATE, "skg_run_packets: no dust/source");
    if (e->lpackCount != e->nlambda)
        return fail(e, SKG_ERR_STATE, "skg_run_packets: source luminosities do not match the dust wavelength bins");
    int st = ensureFrames(e);
    if (st) return st;
    if (!e->hInst.empty())
    {
        if (e->mode != SKG_MODE_PERSISTENT || e->guideMode != 0)
            return fail(e, SKG_ERR_STATE, "skg_run_packets: batched instances run in the persistent engine without the guide");
        if (first_history > e->instTotal || count > e->instTotal - first_history)
            return fail(e, SKG_ERR_ARGUMENT, "skg_run_packets: histories beyond the instances' total");
        for (const Instance& I : e->hInst)
            if (I.fend > static_cast<int>(e->frames.size()))
                return fail(e, SKG_ERR_STATE, "skg_run_packets: an instance names a frame that does not exist");
    }
    if (count == 0) return SKG_OK;
    if (e->mode == SKG_MODE_PIPELINE && e->contKey)
        return fail(e, SKG_ERR_STATE, "skg_run_packets: continuation streams run in the persistent engine");
    if (e->mode == SKG_MODE_PIPELINE && e->rf)
        return fail(e, SKG_ERR_STATE, "skg_run_packets: radiation-field storage runs in the persistent engine");
    if (e->mode == SKG_MODE_PIPELINE && e->nent)
        return fail(e, SKG_ERR_STATE, "skg_run_packets: a particle source runs in the persistent engine");
    if (e->mode == SKG_MODE_PIPELINE) return runPipeline(e, first_history, count);
    return runPersistent(e, first_history, count);
}

int skg_tune_pipeline(skg_engine* e, int32_t mode, int64_t pool_slots, int32_t trace_workgroups, int32_t graph_iterations)
{
    if (!e || !(mode == SKG_MODE_PIPELINE || mode == SKG_MODE_PERSISTENT) || pool_slots < 256 || pool_slots > (1LL << 30)
        || trace_workgroups < 0 || graph_iterations < 2 || graph_iterations > 256)
        return fail(e, SKG_ERR_ARGUMENT, "skg_tune_pipeline: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    e->mode = mode;
    e->poolSlots = pool_slots;
    e->traceBlocks = trace_workgroups;
    e->graphIters = graph_iterations;
    return SKG_OK;
}

int skg_continuation_key(skg_engine* e, uint32_t key)
{
    if (!e) return SKG_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(e->mu);
    e->contKey = key;
    return SKG_OK;
}

int skg_tune_queue(skg_engine* e, int64_t max_requests)
{
    if (!e || max_requests < 0) return fail(e, SKG_ERR_ARGUMENT, "skg_tune_queue: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    e->queueLimit = max_requests;
    return SKG_OK;
}

int skg_last_run_stats(const skg_engine* e, double* trace_ms, int64_t* trace_launches, int64_t* iterations)
{
    if (!e) return SKG_ERR_ARGUMENT;
    if (trace_ms) *trace_ms = e->lastTraceMs;
    if (trace_launches) *trace_launches = e->lastTraceLaunches;
    if (iterations) *iterations = e->lastIterations;
    return SKG_OK;
}

extern "C" {


namespace
{
    int ensureGuide(skg_engine* e)
    {
        if (e->gTheta) return SKG_OK;
        HIPCHK(e, hipMalloc(&e->gTheta, 4 * kGuideParams * sizeof(float)));
        HIPCHK(e, hipMemset(e->gTheta, 0, 4 * kGuideParams * sizeof(float)));
        const size_t tiles = 64 * 16 + 64 * 64 + 64 * 64 + 48 * 64;
        HIPCHK(e, hipMalloc(&e->gBf, 2 * tiles * sizeof(uint16_t)));
        HIPCHK(e, hipMalloc(&e->gBias, (64 * 3 + 48) * sizeof(float)));
        uint16_t* w = e->gBf;
        uint16_t* t = e->gBf + tiles;
        const size_t sz[4] = {64 * 16, 64 * 64, 64 * 64, 48 * 64};
        for (int l = 0; l < 4; ++l)
        {
            e->net.W[l] = w;
            e->net.WT[l] = t;
            w += sz[l];
            t += sz[l];
        }
        e->net.b[0] = e->gBias;
        e->net.b[1] = e->gBias + 64;
        e->net.b[2] = e->gBias + 128;
        e->net.b[3] = e->gBias + 192;
        return SKG_OK;
    }

    int packGuide(skg_engine* e)
    {
        const int total = 64 * 16 + 64 * 64 + 64 * 64 + 48 * 64;
        hipLaunchKernelGGL(guide_pack_kernel, dim3((total + 255) / 256), dim3(256), 0, e->stream, e->gTheta,
                           const_cast<uint16_t*>(e->net.W[0]), const_cast<uint16_t*>(e->net.W[1]),
                           const_cast<uint16_t*>(e->net.W[2]), const_cast<uint16_t*>(e->net.W[3]),
                           const_cast<uint16_t*>(e->net.WT[0]), const_cast<uint16_t*>(e->net.WT[1]),
                           const_cast<uint16_t*>(e->net.WT[2]), const_cast<uint16_t*>(e->net.WT[3]),
                           const_cast<float*>(e->net.b[0]), const_cast<float*>(e->net.b[1]),
                           const_cast<float*>(e->net.b[2]), const_cast<float*>(e->net.b[3]));
        HIPCHK(e, hipGetLastError());
        return SKG_OK;
    }

    GuideTilesW guideTilesW(skg_engine* e)
    {
        GuideTilesW t;
        for (int l = 0; l < 4; ++l)
        {
            t.W[l] = const_cast<uint16_t*>(e->net.W[l]);
            t.WT[l] = const_cast<uint16_t*>(e->net.WT[l]);
            t.b[l] = const_cast<float*>(e->net.b[l]);
        }
        return t;
    }
}

int skg_guide_params(skg_engine* e, const float* theta, int32_t n)
{
    if (!e || !theta || n != kGuideParams) return fail(e, SKG_ERR_ARGUMENT, "skg_guide_params: expected 11560 values");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    int st = ensureGuide(e);
    if (st) return st;
    HIPCHK(e, hipMemcpy(e->gTheta, theta, n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(e, hipMemset(e->gTheta + kGuideParams, 0, 3 * kGuideParams * sizeof(float)));
    e->adamStep = 0;
    st = packGuide(e);
    if (st) return st;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SKG_OK;
}

int skg_guide_read_params(skg_engine* e, float* theta, int32_t n)
{
    if (!e || !theta || n != kGuideParams) return fail(e, SKG_ERR_ARGUMENT, "skg_guide_read_params: expected 11560 values");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (!e->gTheta) return fail(e, SKG_ERR_STATE, "skg_guide_read_params: no guide");
    HIPCHK(e, hipMemcpy(theta, e->gTheta, n * sizeof(float), hipMemcpyDeviceToHost));
    return SKG_OK;
}

int skg_guide_infer(skg_engine* e, const float* features, int64_t n, float* out)
{
    if (!e || n < 0 || (n && (!features || !out))) return fail(e, SKG_ERR_ARGUMENT, "skg_guide_infer: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (!e->gTheta) return fail(e, SKG_ERR_STATE, "skg_guide_infer: no guide parameters");
    if (n == 0) return SKG_OK;
    float *dF = nullptr, *dO = nullptr;
    HIPCHK(e, hipMalloc(&dF, n * kGuideIn * sizeof(float)));
    HIPCHK(e, hipMalloc(&dO, n * kGuideOut * sizeof(float)));
    HIPCHK(e, hipMemcpy(dF, features, n * kGuideIn * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(guide_infer_kernel, dim3(static_cast<unsigned>((n + 63) / 64)), dim3(64), 0, e->stream, e->net, dF,
                       n, dO);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemcpyAsync(out, dO, n * kGuideOut * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    hipFree(dF);
    hipFree(dO);
    return SKG_OK;
}

int skg_guide_sample(skg_engine* e, const float* outputs, int64_t n, uint64_t seed, const double* at, double* dirs,
                     double* pdf, double* pdf_at)
{
    // switched off: the first GPU run of this entry point ended in an abort whose cause has
    // not been found; the guided life cycle checks the same sampler (test_gpu_guide.py)
    (void)outputs;
    (void)n;
    (void)seed;
    (void)at;
    (void)dirs;
    (void)pdf;
    (void)pdf_at;
    return fail(e, SKG_ERR_STATE, "skg_guide_sample: switched off (an unexplained abort on its first GPU run)");
}

int skg_guide_gradient(skg_engine* e, const float* records, int64_t n, float* grad, double* loss)
{
    if (!e || n < 1 || !records || !grad) return fail(e, SKG_ERR_ARGUMENT, "skg_guide_gradient: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (!e->gTheta) return fail(e, SKG_ERR_STATE, "skg_guide_gradient: no guide parameters");
    const int64_t nb = (n + kGradSamples - 1) / kGradSamples;
    if (nb > INT32_MAX) return fail(e, SKG_ERR_ARGUMENT, "skg_guide_gradient: too many records");
    DevBuf recs, lossBuf, part;  // freed on every return
    HIPCHK(e, recs.ensure(static_cast<size_t>(n) * 16 * sizeof(float)));
    HIPCHK(e, lossBuf.ensure(sizeof(double)));
    HIPCHK(e, part.ensure(static_cast<size_t>(nb) * kGuideParams * sizeof(float)));
    float* dR = recs.as<float>();
    double* dL = lossBuf.as<double>();
    HIPCHK(e, hipMemcpyAsync(dR, records, n * 16 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemsetAsync(dL, 0, sizeof(double), e->stream));
    float* g = e->gTheta + 3 * kGuideParams;
    hipLaunchKernelGGL(guide_grad_part, dim3(static_cast<unsigned>(nb)), dim3(64 * kGradWaves), 0, e->stream, e->net, dR,
                       16, n, 1.0f, part.as<float>(), dL);
    HIPCHK(e, hipGetLastError());
    hipLaunchKernelGGL(guide_reduce_step, dim3((kGuideParams + kReduceParams - 1) / kReduceParams),
                       dim3(kReduceParams * kReduceSlices), 0, e->stream, part.as<float>(), static_cast<int>(nb), g,
                       nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, guideTilesW(e));
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemcpyAsync(grad, g, kGuideParams * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    double l = 0.;
    HIPCHK(e, hipMemcpyAsync(&l, dL, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if (loss) *loss = l;
    return SKG_OK;
}

int skg_guide_mode(skg_engine* e, int32_t mode, double mix, double half_size, double lambda_center, double lambda_scale)
{
    if (!e || mode < 0 || mode > 2 || !(mix >= 0 && mix < 1) || !(half_size > 0) || !(lambda_scale > 0))
        return fail(e, SKG_ERR_ARGUMENT, "skg_guide_mode: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    e->guideMode = mode;
    e->guideMix = mix;
    e->gframe.inv_half = 1.0 / half_size;
    e->gframe.lam_c = std::log(lambda_center);
    e->gframe.lam_inv_s = 1.0 / lambda_scale;
    return SKG_OK;
}

int skg_guide_collect(skg_engine* e, int64_t capacity)
{
    if (!e || capacity < 0) return fail(e, SKG_ERR_ARGUMENT, "skg_guide_collect: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (capacity > e->recCap)
    {
        for (void* q : {static_cast<void*>(e->rec), static_cast<void*>(e->recHist), static_cast<void*>(e->recA)})
            if (q) hipFree(q);
        e->rec = nullptr;
        e->recHist = nullptr;
        e->recA = nullptr;
        HIPCHK(e, hipMalloc(&e->rec, capacity * 16 * sizeof(float)));
        HIPCHK(e, hipMalloc(&e->recHist, capacity * sizeof(uint32_t)));
        HIPCHK(e, hipMalloc(&e->recA, capacity * sizeof(double)));
        e->recCap = capacity;
    }
    if (!e->recCount) HIPCHK(e, hipMalloc(&e->recCount, sizeof(unsigned long long)));
    if (!e->recSum) HIPCHK(e, hipMalloc(&e->recSum, sizeof(double)));
    HIPCHK(e, hipMemset(e->recCount, 0, sizeof(unsigned long long)));
    e->recN = 0;
    return SKG_OK;
}

int skg_guide_target(skg_engine* e, int32_t mode, double floor_frac, const double* weights, int64_t n)
{
    if (!e || mode < 0 || mode > 2 || (mode == 1 && !(floor_frac > 0)) || (mode == 2 && (!weights || n < 1)))
        return fail(e, SKG_ERR_ARGUMENT, "skg_guide_target: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (mode == 0)
    {
        if (e->pixW) hipFree(e->pixW);
        e->pixW = nullptr;
        e->pixWSize = 0;
        return SKG_OK;
    }
    int st = ensureFrames(e);
    if (st) return st;
    const int64_t size = e->frameSize;
    if (size < 1) return fail(e, SKG_ERR_STATE, "skg_guide_target: no frames");
    std::vector<double> w(static_cast<size_t>(size));
    if (mode == 2)
    {
        if (n != size) return fail(e, SKG_ERR_ARGUMENT, "skg_guide_target: weights must have one value per frame entry");
        for (int64_t i = 0; i < size; ++i)
        {
            if (!(weights[i] >= 0) || !std::isfinite(weights[i]))
                return fail(e, SKG_ERR_ARGUMENT, "skg_guide_target: weights must be finite and non-negative");
            w[i] = weights[i];
        }
    }
    else
    {
        // 1 / max(F, floor) per entry of the image accumulated so far; the floor is floor_frac
        // times the median of the positive entries of the entry's (frame, wavelength) plane
        std::vector<double> F(static_cast<size_t>(size));
        HIPCHK(e, hipStreamSynchronize(e->stream));
        HIPCHK(e, hipMemcpy(F.data(), e->frameBuf, size * sizeof(double), hipMemcpyDeviceToHost));
        int64_t off = 0;
        for (auto& f : e->frames)
        {
            const int64_t plane = static_cast<int64_t>(f.nx) * f.ny;
            for (int ell = 0; ell < std::max(1, e->nlambda); ++ell, off += plane)
            {
                std::vector<double> pos;
                for (int64_t i = 0; i < plane; ++i)
                    if (F[off + i] > 0) pos.push_back(F[off + i]);
                if (pos.empty()) return fail(e, SKG_ERR_STATE, "skg_guide_target: a frame plane holds no flux yet");
                std::nth_element(pos.begin(), pos.begin() + pos.size() / 2, pos.end());
                const double fl = floor_frac * pos[pos.size() / 2];
                for (int64_t i = 0; i < plane; ++i) w[off + i] = 1.0 / std::max(F[off + i], fl);
            }
        }
    }
    if (e->pixWSize != size)
    {
        if (e->pixW) hipFree(e->pixW);
        e->pixW = nullptr;
        e->pixWSize = 0;
        HIPCHK(e, hipMalloc(&e->pixW, size * sizeof(double)));
        e->pixWSize = size;
    }
    HIPCHK(e, hipMemcpy(e->pixW, w.data(), size * sizeof(double), hipMemcpyHostToDevice));
    return SKG_OK;
}

int skg_guide_records(skg_engine* e, int64_t* n, float* out, int64_t cap)
{
    if (!e || !n) return fail(e, SKG_ERR_ARGUMENT, "skg_guide_records: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    *n = e->recN;
    if (out && cap > 0 && e->recN > 0)
    {
        int64_t m = std::min(cap, e->recN);
        HIPCHK(e, hipMemcpy(out, e->rec, m * 16 * sizeof(float), hipMemcpyDeviceToHost));
    }
    return SKG_OK;
}

int skg_guide_train(skg_engine* e, int32_t steps, int32_t batch, float lr, double* loss_first, double* loss_last)
{
    if (!e || steps < 1 || batch < 64 || batch % 64 || !(lr > 0)) return fail(e, SKG_ERR_ARGUMENT, "skg_guide_train: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (!e->gTheta) return fail(e, SKG_ERR_STATE, "skg_guide_train: no guide parameters");
    if (e->recN < batch) return fail(e, SKG_ERR_STATE, "skg_guide_train: fewer records than one batch");
    // mean training weight over all records, so that a batch's weights sum to about one
    double sum = 0.;
    {
        double* dSum = nullptr;
        HIPCHK(e, hipMalloc(&dSum, sizeof(double)));
        hipError_t st = hipMemsetAsync(dSum, 0, sizeof(double), e->stream);
        if (st == hipSuccess)
        {
            hipLaunchKernelGGL(guide_weight_sum, dim3(1024), dim3(256), 0, e->stream, e->rec, e->recN, dSum);
            st = hipGetLastError();
        }
        if (st == hipSuccess) st = hipMemcpyAsync(&sum, dSum, sizeof(double), hipMemcpyDeviceToHost, e->stream);
        if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
        hipFree(dSum);
        if (st != hipSuccess) return fail(e, SKG_ERR_DEVICE, std::string("skg_guide_train: ") + hipGetErrorString(st));
    }
    if (!(sum > 0)) return fail(e, SKG_ERR_STATE, "skg_guide_train: the records carry no detector luminosity");
    const float cscale = static_cast<float>(static_cast<double>(e->recN) / (static_cast<double>(batch) * sum));
    // batch s takes every nb-th record starting at s mod nb, so that each batch samples the
    // whole collection (records are stored in the order the histories produced them)
    const int64_t nb = e->recN / batch;
    if (16 * nb > INT32_MAX) return fail(e, SKG_ERR_CAPACITY, "skg_guide_train: too many records per batch stride");
    // one partial gradient per workgroup of kGradSamples records, summed by the Adam step
    const int nwg = batch / kGradSamples;
    DevBuf part, loss;
    HIPCHK(e, part.ensure(static_cast<size_t>(nwg) * kGuideParams * sizeof(float)));
    HIPCHK(e, loss.ensure(sizeof(double)));
    double* dL = loss.as<double>();
    float* th = e->gTheta;
    float* m1 = e->gTheta + kGuideParams;
    float* m2 = e->gTheta + 2 * kGuideParams;
    const GuideTilesW tw = guideTilesW(e);
    for (int s = 0; s < steps; ++s)
    {
        const int64_t first = s % nb;
        bool track = s == 0 || s == steps - 1;
        if (track) HIPCHK(e, hipMemsetAsync(dL, 0, sizeof(double), e->stream));
        hipLaunchKernelGGL(guide_grad_part, dim3(nwg), dim3(64 * kGradWaves), 0, e->stream, e->net, e->rec + first * 16,
                           static_cast<int>(16 * nb), static_cast<int64_t>(batch), cscale, part.as<float>(),
                           track ? dL : nullptr);
        ++e->adamStep;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, static_cast<float>(e->adamStep));
        const float c2 = 1.f - std::pow(b2, static_cast<float>(e->adamStep));
        hipLaunchKernelGGL(guide_reduce_step, dim3((kGuideParams + kReduceParams - 1) / kReduceParams),
                           dim3(kReduceParams * kReduceSlices), 0, e->stream, part.as<float>(), nwg, nullptr, th, m1,
                           m2, lr, b1, b2, 1e-8f, c1, c2, tw);
        HIPCHK(e, hipGetLastError());
        if (track)
        {
            double l = 0.;
            HIPCHK(e, hipMemcpyAsync(&l, dL, sizeof(double), hipMemcpyDeviceToHost, e->stream));
            HIPCHK(e, hipStreamSynchronize(e->stream));
            if (s == 0 && loss_first) *loss_first = l;
            if (s == steps - 1 && loss_last) *loss_last = l;
        }
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return SKG_OK;
}

int skg_sync(skg_engine* e)
{
    if (!e) return SKG_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e->ev0, e->ev1) == hipSuccess) e->lastMs = ms;
    return SKG_OK;
}

int skg_run_packets(skg_engine* e, uint64_t first_history, uint64_t count)
{
    int st = skg_run_packets_async(e, first_history, count);
    if (st) return st;
    return skg_sync(e);
}

int skg_autotune(skg_engine* e, uint64_t first_history, uint64_t trial_packets, int32_t engines, int64_t* best6,
                 double* best_rate)
{
    if (!e || trial_packets < 1024 || !(engines >= 1 && engines <= 3))
        return fail(e, SKG_ERR_ARGUMENT, "skg_autotune: bad arguments");
    if (e->guideMode == 2) return fail(e, SKG_ERR_STATE, "skg_autotune: not in guide collect mode");
    // instances, particle sources and radiation-field storage run in the persistent engine only
    bool persistentOnly = !e->hInst.empty() || e->nent > 0 || e->rf != nullptr;
    if (persistentOnly) engines &= 1;
    if (!engines) return fail(e, SKG_ERR_ARGUMENT, "skg_autotune: no engine supports this configuration");
    if (!e->hInst.empty() && e->instTotal < trial_packets + trial_packets / 4)
        return fail(e, SKG_ERR_ARGUMENT, "skg_autotune: fewer instance histories than one trial");
    int st = SKG_OK;
    // keep the caller's detector contents, stored radiation field and counters
    double* saved = nullptr;
    double* savedRf = nullptr;
    const size_t rfSize = e->rf ? static_cast<size_t>(e->rfCells) * e->rfLambda : 0;
    unsigned long long savedCnt[kNumCounters];
    {
        std::lock_guard<std::mutex> lk(e->mu);
        hipSetDevice(e->device);
        if ((st = ensureFrames(e))) return st;
        if (e->frameSize) HIPCHK(e, hipMalloc(&saved, e->frameSize * sizeof(double)));
        if (saved) HIPCHK(e, hipMemcpy(saved, e->frameBuf, e->frameSize * sizeof(double), hipMemcpyDeviceToDevice));
        if (rfSize) HIPCHK(e, hipMalloc(&savedRf, rfSize * sizeof(double)));
        if (savedRf) HIPCHK(e, hipMemcpy(savedRf, e->rf, rfSize * sizeof(double), hipMemcpyDeviceToDevice));
        HIPCHK(e, hipMemcpy(savedCnt, e->counters, sizeof(savedCnt), hipMemcpyDeviceToHost));
    }
    struct Cand
    {
        int mode, threshold, workgroups;
        int64_t pool;
        int traceWgs, iters;
    };
    std::vector<Cand> cands;
    const int cus = e->cus;
    // transition thresholds: the octree and Voronoi kernels are fastest around 16-24 waiting
    // lanes, the Cartesian and tetra kernels with rounds only once every lane of the wave waits
    // (64); the candidates cover both
    if (engines & 1)
        for (int thr : {12, 16, 20, 24, 32, 48, 64})
            for (int wg : {4 * cus, 8 * cus}) cands.push_back({SKG_MODE_PERSISTENT, thr, wg, e->poolSlots, e->traceBlocks, e->graphIters});
    if (engines & 2)
        for (int64_t pool : {int64_t(1) << 20, int64_t(1) << 21, int64_t(1) << 22})
            for (int it : {8, 16})
                for (int twg : {4 * cus, 8 * cus}) cands.push_back({SKG_MODE_PIPELINE, e->threshold, e->workgroups, pool, twg, it});
    double bestRate = -1.;
    Cand best = cands[0];
    Cand caller;  // the tuning in force before the trials, restored when a trial fails
    {
        std::lock_guard<std::mutex> lk(e->mu);
        caller = {e->mode, e->threshold, e->workgroups, e->poolSlots, e->traceBlocks, e->graphIters};
    }
    auto applyTuning = [e](const Cand& c) {
        std::lock_guard<std::mutex> lk(e->mu);
        e->mode = c.mode;
        e->threshold = c.threshold;
        e->workgroups = c.workgroups;
        e->poolSlots = c.pool;
        e->traceBlocks = c.traceWgs;
        e->graphIters = c.iters;
    };
    uint64_t h = first_history;
    // one warm-up run (pool allocation, graph capture), then the timed run of `packets`
    auto trial = [&](const Cand& c, uint64_t packets, double& rate) {
        applyTuning(c);
        if (!e->hInst.empty() && h + trial_packets / 4 + packets > e->instTotal) h = 0;
        int r = skg_run_packets(e, h, trial_packets / 4);
        if (r) return r;
        h += trial_packets / 4;
        auto t0 = std::chrono::steady_clock::now();
        if ((r = skg_run_packets(e, h, packets))) return r;
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        h += packets;
        rate = static_cast<double>(packets) / dt;
        return SKG_OK;
    };
    std::vector<std::pair<double, size_t>> ranked;
    for (size_t ci = 0; ci < cands.size(); ++ci)
    {
        double rate = 0.;
        if ((st = trial(cands[ci], trial_packets, rate))) break;
        ranked.push_back({rate, ci});
        if (rate > bestRate)
        {
            bestRate = rate;
            best = cands[ci];
        }
    }
    // The end of a launch, when the last waves finish while other units idle, weighs more in a
    // short trial than in a long run: the three fastest candidates are timed again on four
    // times the trial (where the instances allow it), and the fastest of those is kept.
    const uint64_t longTrial = 4 * trial_packets;
    if (!st && ranked.size() > 1 && (e->hInst.empty() || e->instTotal >= longTrial + trial_packets / 4))
    {
        std::sort(ranked.begin(), ranked.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        double bestLong = -1.;
        for (size_t r = 0; r < ranked.size() && r < 3; ++r)
        {
            double rate = 0.;
            if ((st = trial(cands[ranked[r].second], longTrial, rate))) break;
            if (rate > bestLong)
            {
                bestLong = rate;
                best = cands[ranked[r].second];
                bestRate = rate;
            }
        }
    }
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    if (saved)
    {
        hipMemcpy(e->frameBuf, saved, e->frameSize * sizeof(double), hipMemcpyDeviceToDevice);
        hipFree(saved);
    }
    if (savedRf)
    {
        hipMemcpy(e->rf, savedRf, rfSize * sizeof(double), hipMemcpyDeviceToDevice);
        hipFree(savedRf);
    }
    hipMemcpy(e->counters, savedCnt, sizeof(savedCnt), hipMemcpyHostToDevice);
    const Cand& keep = st ? caller : best;
    e->mode = keep.mode;
    e->threshold = keep.threshold;
    e->workgroups = keep.workgroups;
    e->poolSlots = keep.pool;
    e->traceBlocks = keep.traceWgs;
    e->graphIters = keep.iters;
    if (st) return st;
    if (best6)
    {
        best6[0] = best.mode;
        best6[1] = best.threshold;
        best6[2] = best.workgroups;
        best6[3] = best.pool;
        best6[4] = best.traceWgs;
        best6[5] = best.iters;
    }
    if (best_rate) *best_rate = bestRate;
    return SKG_OK;
}

int skg_counters(skg_engine* e, int64_t* out8) { return skg_counters_n(e, out8, 8); }

int skg_counters_n(skg_engine* e, int64_t* out, int32_t n)
{
    if (!e || !out || n < 1 || n > kNumCounters) return fail(e, SKG_ERR_ARGUMENT, "skg_counters_n: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipSetDevice(e->device);
    unsigned long long c[kNumCounters];
    HIPCHK(e, hipMemcpyAsync(c, e->counters, sizeof(c), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (int i = 0; i < n; ++i) out[i] = static_cast<int64_t>(c[i]);
    return SKG_OK;
}

double skg_last_kernel_ms(const skg_engine* e) { return e ? e->lastMs : 0.0; }

}  // extern "C"
