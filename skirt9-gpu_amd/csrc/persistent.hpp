// Persistent forced-scattering life-cycle kernel (the benchmarked engine): one photon packet
// per lane at a time, run as a small state machine -- peel-off traversal, forced path first
// pass, second pass to the interaction point, scattering.  Every loop iteration advances each
// busy lane by one cell crossing; lanes whose traversal ended wait until enough lanes of the
// wavefront are waiting (the transition threshold), then the wave runs one transition round
// together, so the divergent event code is amortised over many lanes.  Lanes whose packet
// terminated refill from a global history counter with one atomic per wavefront (ballot +
// mbcnt prefix), which stream-compacts packets of very uneven path length without a
// separate compaction pass.  The order of operations within each packet is the reference's
// (MonteCarloSimulation.cpp:614-694, :1492-1544, :2034-2092).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/skirt_mi355x.h"
#include "guide.hpp"
#include "lifecycle.hpp"
#include "pipeline.hpp"
#include "traverse.hpp"

namespace skg
{
    // ------------------------------------------------------------------ the life cycle
    struct SimParams
    {
        GridView grid;
        const double* dens;
        const float* dens32;
        int nlambda;
        int nframes;
        const double* lambda;
        const double* sext;
        const double* ssca;
        const double* gpar;
        const double* lpack;
        double src[3];
        double xi, min_weight_reduction;
        int min_scatt;
        int threshold;
        uint64_t seed;
        uint64_t first, count;
        unsigned long long* queue;     // next history offset (work queue head)
        unsigned long long* counters;  // kNumCounters event counters
        double* frame;
        const FrameGeom* frames;  // device copy of the instrument geometry
        const Instance* inst;     // batched instances (ninst > 0) or none
        int ninst;
        const double* ilpack;     // launch luminosities per instance (ninst x nlambda)
        GuideNet net;             // guided variant: the network, its input frame and the mixture
        GuideFrame gframe;
        double guideMix;
        uint32_t contKey;         // nonzero: histories continue on Philox tag contKey from their first
                                  // scattering on (same emission and first interaction, independent
                                  // continuations)
        int32_t srcHint;          // octree: the leaf of the point source (the start hint of new packets), -1: none
        double* rf;               // radiation-field variant: ncells x nlambda table
        int persp;                // some instruments are perspective instruments
        int nent;                 // particle-source variant: number of entities (0: the point source)
        int entKernel;            // smoothing kernel: 0 uniform, 1 cubic spline
        const double* ent;        // per entity: x, y, z, smoothing length h
        const int64_t* entFirst;  // first history index of each entity (nent + 1 entries)
        const double* entLpack;   // launch luminosity per packet, nent x nlambda
        const double* entX;       // cubic spline: cumulative mass at u = k / kSplineN, k = 0..kSplineN
    };

    // kernel variants of the life cycle
    enum Variant : int
    {
        kPlain = 0,      // one run
        kBatched = 1,    // batched independent instances
        kGuided = 2,     // scattering directions from the defensive guide / HG mixture
        kStoreRF = 3,    // one run that also stores the radiation field of the forced paths (perspective
                         // instruments allowed)
        kPersp = 4,      // one run with perspective instruments among the instruments
        kParticles = 5   // one run whose source is a set of imported particles
    };
    constexpr int kSplineN = 400;  // CubicSplineSmoothingKernel's table size (CubicSplineSmoothingKernel.cpp:18)

    // SpecialFunctions::lnmean(x1, x2, lnx1, lnx2) (SpecialFunctions.cpp:860-880): the
    // logarithmic mean (x2 - x1) / ln(x2 / x1), by its series for nearly equal arguments
    __device__ __forceinline__ double ln_mean(double x1, double x2, double lnx1, double lnx2)
    {
        if (x1 > x2)
        {
            const double t = x1, lt = lnx1;
            x1 = x2;
            x2 = t;
            lnx1 = lnx2;
            lnx2 = lt;
        }
        if (x1 <= 0) return 0.;
        const double x = x2 / x1 - 1.;
        if (x < 1e-3)
        {
            const double c3 = 1. / 3., c5 = 1. / 5., c6 = 1. / 6.;
            return x1 / (1. - 0.5 * x + c3 * x * x - 0.25 * x * x * x + c5 * x * x * x * x - c6 * x * x * x * x * x);
        }
        return (x2 - x1) / (lnx2 - lnx1);
    }

    enum Phase : int
    {
        kFetch = 0,
        kPeel = 1,
        kFwd1 = 2,
        kFwd2 = 3,
        kDone = 4,
        kScatter = 5  // guided variant: peel-offs done, waiting for the wave's guided scattering
    };

#ifndef SKG_GUIDE_LANES
#define SKG_GUIDE_LANES 16
#endif
    constexpr int kGuideLanes = SKG_GUIDE_LANES;  // guided variant: lanes per network evaluation round
    constexpr int kMaxScatterings = 1 << 20;  // guard against non-terminating histories (NaN weights)
    constexpr int kMaxSteps = 1 << 26;        // guard against non-terminating traversals

    // Packet state that only the event transitions read and write.  Between transitions it
    // lives in LDS as a structure of arrays over the workgroup's lanes (conflict-free, one
    // 8-byte or 4-byte word per lane and field), so the crossing loop keeps only the
    // traversal state in registers.
    enum ColdF64 : int
    {
        cRx, cRy, cRz, cKx, cKy, cKz, cW, cLam, cLthr, cTaupath, cTauCached, cWp, kColdF64
    };
    enum ColdU32 : int
    {
        cHist0, cHist1, cDraw, cEll, cNscatt, cF, cPix, cFlags, cInst, cHint, kColdU32
    };
    constexpr int kLcThreads = 256;
    constexpr int kColdBytes = kLcThreads * (8 * kColdF64 + 4 * kColdU32);  // 34 KB per workgroup

    struct Packet
    {
        Philox rng;
        double rx, ry, rz, kx, ky, kz;
        double W, lam, Lthr, taupath, tauCached, Wp;
        int ell, nscatt, f, pix, inst;
        int hint;  // a cell that probably holds the packet's position (-1: none)
        bool scat, haveTau;

        template<bool INST>
        __device__ __forceinline__ void load(const double* cd, const uint32_t* cu, int t, const SimParams& p)
        {
            rx = cd[cRx * kLcThreads + t];
            ry = cd[cRy * kLcThreads + t];
            rz = cd[cRz * kLcThreads + t];
            kx = cd[cKx * kLcThreads + t];
            ky = cd[cKy * kLcThreads + t];
            kz = cd[cKz * kLcThreads + t];
            W = cd[cW * kLcThreads + t];
            lam = cd[cLam * kLcThreads + t];
            Lthr = cd[cLthr * kLcThreads + t];
            taupath = cd[cTaupath * kLcThreads + t];
            tauCached = cd[cTauCached * kLcThreads + t];
            Wp = cd[cWp * kLcThreads + t];
            uint64_t h = static_cast<uint64_t>(cu[cHist0 * kLcThreads + t])
                         | (static_cast<uint64_t>(cu[cHist1 * kLcThreads + t]) << 32);
            inst = static_cast<int>(cu[cInst * kLcThreads + t]);
            hint = static_cast<int>(cu[cHint * kLcThreads + t]);
            nscatt = static_cast<int>(cu[cNscatt * kLcThreads + t]);
            rng.init(INST ? p.inst[inst].seed : p.seed, h, p.contKey != 0u && nscatt > 0 ? p.contKey : 0u);
            rng.draw = cu[cDraw * kLcThreads + t];
            ell = static_cast<int>(cu[cEll * kLcThreads + t]);
            f = static_cast<int>(cu[cF * kLcThreads + t]);
            pix = static_cast<int>(cu[cPix * kLcThreads + t]);
            uint32_t fl = cu[cFlags * kLcThreads + t];
            scat = fl & 1u;
            haveTau = (fl >> 1) & 1u;
        }

        __device__ __forceinline__ void store(double* cd, uint32_t* cu, int t) const
        {
            cd[cRx * kLcThreads + t] = rx;
            cd[cRy * kLcThreads + t] = ry;
            cd[cRz * kLcThreads + t] = rz;
            cd[cKx * kLcThreads + t] = kx;
            cd[cKy * kLcThreads + t] = ky;
            cd[cKz * kLcThreads + t] = kz;
            cd[cW * kLcThreads + t] = W;
            cd[cLam * kLcThreads + t] = lam;
            cd[cLthr * kLcThreads + t] = Lthr;
            cd[cTaupath * kLcThreads + t] = taupath;
            cd[cTauCached * kLcThreads + t] = tauCached;
            cd[cWp * kLcThreads + t] = Wp;
            cu[cHist0 * kLcThreads + t] = rng.h0;
            cu[cHist1 * kLcThreads + t] = rng.h1;
            cu[cDraw * kLcThreads + t] = rng.draw;
            cu[cEll * kLcThreads + t] = static_cast<uint32_t>(ell);
            cu[cNscatt * kLcThreads + t] = static_cast<uint32_t>(nscatt);
            cu[cF * kLcThreads + t] = static_cast<uint32_t>(f);
            cu[cPix * kLcThreads + t] = static_cast<uint32_t>(pix);
            cu[cFlags * kLcThreads + t] = (scat ? 1u : 0u) | (haveTau ? 2u : 0u);
            cu[cInst * kLcThreads + t] = static_cast<uint32_t>(inst);
            cu[cHint * kLcThreads + t] = static_cast<uint32_t>(hint);
        }
    };

    // traversal state of one lane: the generator, the running sums and the phase
    template<int GT, int FP> struct Trav
    {
        Gen<GT> gen;
        int phase;
        bool ended;
        int steps;
        double tau, s, tprev, sprev;  // FP64 sums
        Kahan32 tau32, s32;           // FP32 sums
        int lastm;
        double section;
        float section32;
        double taumax;  // peel-off: early-out optical depth
        double tint;    // second forced pass: interaction optical depth
        int mint;
        double sint;
        int ell;        // radiation-field variant: the packet's bin, luminosity and the
        double rfL;     // extinction factor at the start of the current segment
        double rfExt;
        double slim;    // perspective variant: path length after which a peel-off ends

        __device__ __forceinline__ double tau_value() const
        {
            return FP == SKG_FP64 ? tau : static_cast<double>(tau32.sum);
        }
        __device__ __forceinline__ double s_value() const
        {
            return FP == SKG_FP64 ? s : static_cast<double>(s32.sum);
        }

        __device__ __forceinline__ void start(const Packet& pk, double dx, double dy, double dz, int ph)
        {
            gen.start(pk.rx, pk.ry, pk.rz, dx, dy, dz, pk.hint);
            phase = ph;
            ended = false;
            steps = 0;
            tau = 0.;
            s = 0.;
            tprev = 0.;
            sprev = 0.;
            tau32.reset();
            s32.reset();
            lastm = -1;
            // radiation-field variant: the luminosity at the start of a forced path and the
            // extinction factor before its first segment (unused, and removed, elsewhere)
            ell = pk.ell;
            rfL = pk.W / pk.lam;
            rfExt = 1.;
            slim = INFINITY;
        }

        // one cell crossing of the current traversal; with RF, the first forced pass also
        // stores the radiation field (MonteCarloSimulation::storeRadiationField, :909-925):
        // L * lnmean(e^-tau_end, e^-tau_begin) * ds per stored segment in a cell
        template<bool RF, bool PERSP>
        __device__ __forceinline__ void step(const GridView& grid, unsigned int* lds, double* rf, int nl)
        {
            int m;
            double ds, nd;
            const int r = gen.advance(grid, m, ds, nd);
            if (r == 0) return;  // an octree step that loaded a search node or list entry
            // the traversal ends at the path's end (r < 0, or r == 2 after its last segment), at
            // the early-out optical depth of a peel-off, past a perspective peel-off's distance, or
            // at a second pass's interaction; the crossing counters are added once, where it ends
            bool fin = r < 0 || ++steps > kMaxSteps;
            if (!fin)
            {
                if (phase == kPeel)
                {
                    if (m >= 0)
                    {
                        if (FP == SKG_FP64)
                        {
                            tau += section * nd * ds;
                            if (tau >= taumax)
                            {
                                tau = INFINITY;
                                fin = true;
                            }
                        }
                        else
                        {
                            tau32.add(section32 * static_cast<float>(nd) * static_cast<float>(ds));
                            if (static_cast<double>(tau32.sum) >= taumax)
                            {
                                tau32.sum = INFINITY;
                                fin = true;
                            }
                        }
                    }
                    if constexpr (PERSP)
                    {
                        // MediumSystem::getExtinctionOpticalDepth(pp, distance) (MediumSystem.cpp:
                        // 2376-2382): the segment that passes the distance still counts
                        s += ds;
                        if (s > slim) fin = true;
                    }
                }
                else if (ds > 0.)
                {
                    if (FP == SKG_FP64)
                    {
                        tprev = tau;
                        sprev = s;
                        s += ds;
                        if (m >= 0) tau += section * nd * ds;
                    }
                    else
                    {
                        tprev = static_cast<double>(tau32.sum);
                        sprev = static_cast<double>(s32.sum);
                        s32.add(static_cast<float>(ds));
                        if (m >= 0) tau32.add(section32 * static_cast<float>(nd) * static_cast<float>(ds));
                    }
                    lastm = m;
                    if constexpr (RF)
                    {
                        if (phase == kFwd1)
                        {
                            const double lnEnd = -tau_value();
                            const double extEnd = exp(lnEnd);
                            if (m >= 0)
                            {
                                const double em = ln_mean(extEnd, rfExt, lnEnd, -tprev);
                                unsafeAtomicAdd(rf + static_cast<int64_t>(m) * nl + ell, rfL * em * ds);
                            }
                            rfExt = extEnd;
                        }
                    }
                    // second pass: stop at the first segment whose exit optical depth exceeds tau_int
                    if (phase == kFwd2 && tint < tau_value())
                    {
                        mint = m;
                        double tv = tau_value(), sv = s_value();
                        sint = sprev + ((tint - tprev) / (tv - tprev)) * (sv - sprev);
                        fin = true;
                    }
                }
            }
            fin = fin || r == 2;
            if (fin)
            {
                ended = true;
                atomicAdd(lds + 4, static_cast<unsigned int>(steps));
                if (phase == kFwd2) atomicAdd(lds + 7, static_cast<unsigned int>(steps));
            }
        }
    };

    // A lane's detector deposit of the current transition round, held until the wave
    // combines the round's deposits (flush_deposits); off < 0: none
    struct Pending
    {
        int64_t off;
        double v;
    };

    // The deposits of one transition round into the frames (FluxRecorder::detect,
    // FluxRecorder.cpp:414-418, with LockFree::add, LockFree.hpp:23-36), combined across the
    // wavefront: the deposits into the pixel of the first depositing lane are summed by a
    // shuffle reduction over the whole (converged) wave into one FP64 atomic, the others are
    // added directly.  Counter 5 counts the atomics issued.
    __device__ __forceinline__ void flush_deposits(double* frame, Pending& pd, unsigned int* lds)
    {
        const bool has = pd.off >= 0;
        const unsigned long long m = __ballot(has);
        if (!m) return;
        const int leader = __ffsll(static_cast<long long>(m)) - 1;
        const int64_t key = __shfl(pd.off, leader);
        const bool same = has && pd.off == key;
        double v = same ? pd.v : 0.;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (__lane_id() == leader)
        {
            unsafeAtomicAdd(frame + key, v);
            atomicAdd(lds + 5, 1u);
        }
        if (has && !same)
        {
            unsafeAtomicAdd(frame + pd.off, pd.v);
            atomicAdd(lds + 5, 1u);
        }
        pd.off = -1;
    }

    // The guided scattering direction (defensive mixture of the guide's lobes and HG) and its
    // weight factor Phi_HG / ((1 - mix) Phi_HG + mix q), from the network outputs o.  Not
    // inlined: its registers are then not added to those of the traversal state that is live
    // around the call in the persistent kernel.
    __device__ __attribute__((noinline)) double guided_direction(const float* o, double g, double mix, Philox& rng,
                                                                 double& kx, double& ky, double& kz)
    {
        LobesF Lb;
        guide_lobes_f(o, Lb);
        const double kxo = kx, kyo = ky, kzo = kz;
        const bool useGuide = rng.uniform() < mix;
        if (useGuide)
            guide_sample_f(Lb, rng, kx, ky, kz);
        else
            scatter_hg(rng, g, kx, ky, kz);
        const double c = kxo * kx + kyo * ky + kzo * kz;
        const double phi = (fabs(g) < 1e-6 ? 1.0 : value_hg(g, c)) / (4 * M_PI);
        const double q = guide_pdf_f(Lb, kx, ky, kz);
        return phi / ((1.0 - mix) * phi + mix * q);
    }

    // The event transitions of the life cycle, on a packet loaded from LDS.
    template<int GT, int FP, int V> struct Events
    {
        static constexpr bool INST = V == kBatched;
        const SimParams& p;
        Trav<GT, FP>& T;
        Packet& pk;
        unsigned int* lds;
        Pending& pd;

        __device__ __forceinline__ void finish()
        {
            atomicAdd(lds + 0, 1u);
            T.phase = kFetch;
            T.ended = false;
        }

        // L exp(-tau) into the packet's pixel: held for the wave's combined flush at the end of
        // the round (a lane's second deposit in one round adds its first one directly)
        __device__ __forceinline__ void deposit(const FrameGeom& fg, double L, double t)
        {
            double v = L * exp(-t);
            if (pd.off >= 0)
            {
                unsafeAtomicAdd(p.frame + pd.off, pd.v);
                atomicAdd(lds + 5, 1u);
            }
            pd.off = fg.offset + pk.pix + static_cast<int64_t>(pk.ell) * fg.nx * fg.ny;
            pd.v = v;
            atomicAdd(lds + 8, 1u);
        }

        // MonteCarloSimulation::performLifeCycle launch + SourceSystem::launch; in a batched
        // launch, history h is history h - offset of its instance
        __device__ __forceinline__ void launch(uint64_t h)
        {
            uint64_t seed = p.seed;
            const double* lp = p.lpack;
            pk.inst = 0;
            if constexpr (INST)
            {
                int lo = 0, hi = p.ninst - 1;  // last instance whose offset is <= h
                while (lo < hi)
                {
                    int mid = (lo + hi + 1) >> 1;
                    if (p.inst[mid].offset <= h)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                pk.inst = lo;
                h -= p.inst[lo].offset;
                seed = p.inst[lo].seed;
                lp = p.ilpack + static_cast<int64_t>(lo) * p.nlambda;
            }
            pk.rng.init(seed, h);
            pk.ell = static_cast<int>(pk.rng.uniform() * p.nlambda);
            if (pk.ell >= p.nlambda) pk.ell = p.nlambda - 1;
            pk.lam = p.lambda[pk.ell];
            T.section = INST ? p.sext[pk.ell] * p.inst[pk.inst].scale : p.sext[pk.ell];
            T.section32 = static_cast<float>(T.section);
            // the particle launch: its own variant, and a run-time branch of the radiation-field
            // and perspective variants
            bool particles = false;
            if constexpr (V == kParticles)
                particles = true;
            else if constexpr (V == kStoreRF || V == kPersp)
                particles = p.nent > 0;
            if (particles)
            {
                // ImportedSource::launch: the entity whose history index range holds h
                // (ImportedSource.cpp:466-471), its luminosity for the bin (:653-705, the
                // oligochromatic bias weight folded into entLpack), a position from its smoothing
                // kernel (ParticleSnapshot::generatePosition, ParticleSnapshot.cpp:275-286)
                int lo = 0, hi = p.nent - 1;
                while (lo < hi)
                {
                    int mid = (lo + hi + 1) >> 1;
                    if (p.entFirst[mid] <= static_cast<int64_t>(h))
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                const int m = lo;
                pk.W = p.entLpack[static_cast<int64_t>(m) * p.nlambda + pk.ell] * pk.lam;
                const double X = pk.rng.uniform();
                double u;
                if (p.entKernel == 1)
                {
                    // CubicSplineSmoothingKernel::generateRadius (:78-85): NR::locateClip on the table
                    int a = 0, b = kSplineN;
                    while (b - a > 1)
                    {
                        int mid = (a + b) >> 1;
                        if (X >= p.entX[mid])
                            a = mid;
                        else
                            b = mid;
                    }
                    const double f = (X - p.entX[a]) / (p.entX[a + 1] - p.entX[a]);
                    u = (a + f) / kSplineN;
                }
                else
                    u = pow(X, 1.0 / 3.0);  // UniformSmoothingKernel::generateRadius (:27-31)
                double dx, dy, dz;
                isotropic(pk.rng, dx, dy, dz);
                const double hm = p.ent[4 * m + 3];
                pk.rx = p.ent[4 * m + 0] + dx * u * hm;
                pk.ry = p.ent[4 * m + 1] + dy * u * hm;
                pk.rz = p.ent[4 * m + 2] + dz * u * hm;
            }
            else
            {
                pk.W = lp[pk.ell] * pk.lam;
                pk.rx = p.src[0];
                pk.ry = p.src[1];
                pk.rz = p.src[2];
            }
            const int startHint = V == kParticles ? -1 : p.srcHint;
            isotropic(pk.rng, pk.kx, pk.ky, pk.kz);
            pk.nscatt = 0;
            pk.taupath = 0.;
            pk.tauCached = 0.;
            pk.Wp = 0.;
            pk.pix = -1;
            pk.hint = startHint;
            pk.haveTau = false;
            if (!(pk.W / pk.lam > 0))
            {
                finish();
                return;
            }
            pk.Lthr = (pk.W / pk.lam) / p.min_weight_reduction;
            begin_peel(false);
        }

        __device__ __forceinline__ void begin_peel(bool scattering)
        {
            pk.scat = scattering;
            pk.f = (INST ? p.inst[pk.inst].fbeg : 0) - 1;
            next_peel();
        }

        // peel-off toward each instrument in turn (MonteCarloSimulation.cpp:698-715, :2034-2092;
        // FluxRecorder::detect skips the optical depth for packets outside the frame, :375)
        __device__ __forceinline__ void next_peel()
        {
            int fb = 0, fe = p.nframes;
            if constexpr (INST)
            {
                fb = p.inst[pk.inst].fbeg;
                fe = p.inst[pk.inst].fend;
            }
            while (++pk.f < fe)
            {
                const FrameGeom& fg = p.frames[pk.f];
                if constexpr (V == kPersp || V == kStoreRF)
                {
                    if (fg.persp)
                    {
                        // PerspectiveInstrument: toward the eye from this position
                        // (PerspectiveInstrument.cpp:93-101), luminosity / zp^2 and the
                        // optical depth over the path length zp (FluxRecorder.cpp:372-405)
                        double zp;
                        pk.pix = fg.persp_pixel(pk.rx, pk.ry, pk.rz, zp);
                        const double dx = fg.eye[0] - pk.rx, dy = fg.eye[1] - pk.ry, dz = fg.eye[2] - pk.rz;
                        const double D = sqrt(dx * dx + dy * dy + dz * dz);
                        const double ox = dx / D, oy = dy / D, oz = dz / D;
                        if (!fg.same_as_prev || pk.f == fb)
                        {
                            pk.haveTau = false;
                            pk.Wp = pk.W;
                            if (pk.scat) pk.Wp = pk.W * peel_phase(p.gpar[pk.ell], pk.kx * ox + pk.ky * oy + pk.kz * oz);
                        }
                        if (pk.pix < 0) continue;
                        const double L = pk.Wp / pk.lam;
                        if (pk.haveTau)
                        {
                            deposit(fg, L / (zp * zp), pk.tauCached);
                            continue;
                        }
                        if (L <= 0)
                        {
                            pk.haveTau = true;
                            pk.tauCached = INFINITY;
                            continue;
                        }
                        T.taumax = log(L) + 745;
                        atomicAdd(lds + 1, 1u);
                        T.start(pk, ox, oy, oz, kPeel);
                        T.slim = zp;
                        return;
                    }
                }
                if (!fg.same_as_prev || pk.f == fb)
                {
                    pk.haveTau = false;
                    pk.Wp = pk.W;
                    if (pk.scat)
                    {
                        double c = pk.kx * fg.kobs[0] + pk.ky * fg.kobs[1] + pk.kz * fg.kobs[2];
                        pk.Wp = pk.W * peel_phase(p.gpar[pk.ell], c);
                    }
                }
                pk.pix = fg.pixel(pk.rx, pk.ry, pk.rz);
                if (pk.pix < 0) continue;
                double L = pk.Wp / pk.lam;
                if (pk.haveTau)
                {
                    deposit(fg, L, pk.tauCached);
                    continue;
                }
                if (L <= 0)
                {
                    pk.haveTau = true;
                    pk.tauCached = INFINITY;
                    continue;
                }
                T.taumax = log(L) + 745;
                atomicAdd(lds + 1, 1u);
                T.start(pk, fg.kobs[0], fg.kobs[1], fg.kobs[2], kPeel);
                return;
            }
            // all instruments done: scatter (after a scattering peel-off) and start a forced path
            if (pk.scat)
            {
                if constexpr (V == kGuided)
                {
                    // the wave's guided scattering round follows (guided_scatter)
                    T.phase = kScatter;
                    T.ended = false;
                    return;
                }
                if (p.contKey != 0u && pk.nscatt == 0) pk.rng.retag(p.contKey);
                scatter_hg(pk.rng, p.gpar[pk.ell], pk.kx, pk.ky, pk.kz);
                pk.nscatt++;
                atomicAdd(lds + 3, 1u);
            }
            atomicAdd(lds + 2, 1u);
            T.start(pk, pk.kx, pk.ky, pk.kz, kFwd1);
        }

        // Guided scattering with the network outputs o of this packet's event: a direction from
        // the defensive mixture (1 - mix) HG + mix guide, and the weight factor
        // Phi_HG / p_mix that keeps the estimator unbiased.  The termination threshold takes the
        // same factor, so the rule (MonteCarloSimulation.cpp:643-660) applies to the weight
        // without the guide's factors (packets the guide down-weights are not cut early).
        __device__ __forceinline__ void guided_scatter(const float* o)
        {
            if (p.contKey != 0u && pk.nscatt == 0) pk.rng.retag(p.contKey);
            const double ratio = guided_direction(o, p.gpar[pk.ell], p.guideMix, pk.rng, pk.kx, pk.ky, pk.kz);
            pk.W *= ratio;
            pk.Lthr *= ratio;
            pk.nscatt++;
            atomicAdd(lds + 3, 1u);
            atomicAdd(lds + 2, 1u);
            T.start(pk, pk.kx, pk.ky, pk.kz, kFwd1);
        }

        // event handling once a traversal ended
        __device__ __forceinline__ void transition()
        {
            // every traversal starts at the packet's position: its start cell is the hint for
            // the next traversal from there (octree, Voronoi and tetrahedral grids)
            if constexpr (GT != kCartesian)
                if (T.gen.hint >= 0) pk.hint = T.gen.hint;
            if (T.phase == kPeel)
            {
                pk.haveTau = true;
                pk.tauCached = T.tau_value();
                const FrameGeom& fg = p.frames[pk.f];
                double L = pk.Wp / pk.lam;
                if constexpr (V == kPersp || V == kStoreRF)
                {
                    if (fg.persp)
                    {
                        double zp;
                        fg.persp_pixel(pk.rx, pk.ry, pk.rz, zp);
                        L /= zp * zp;
                    }
                }
                deposit(fg, L, pk.tauCached);
                next_peel();
            }
            else if (T.phase == kFwd1)
            {
                // MonteCarloSimulation::simulateForcedPropagation (:1492-1544)
                pk.taupath = T.tau_value();
                if (pk.taupath <= 0.)
                {
                    pk.W *= 0.;
                    finish();
                    return;
                }
                double xi = p.xi;
                double tint;
                if (xi == 0.)
                    tint = expon_cutoff(pk.rng, pk.taupath);
                else
                {
                    tint = pk.rng.uniform() < xi ? pk.rng.uniform() * pk.taupath : expon_cutoff(pk.rng, pk.taupath);
                    double pp = -exp(-tint) / expm1(-pk.taupath);
                    double q = (1.0 - xi) * pp + xi / pk.taupath;
                    pk.W *= pp / q;
                }
                atomicAdd(lds + 6, 1u);
                T.start(pk, pk.kx, pk.ky, pk.kz, kFwd2);
                T.tint = tint;
                T.mint = -1;
            }
            else if (T.phase == kFwd2)
            {
                // SpatialGridPath::findInteractionPoint (SpatialGridPath.cpp:161-203): beyond the
                // last segment, use the last segment's cell and exit distance
                if (T.mint < 0 && !(T.tint < T.tau_value()))
                {
                    T.mint = T.lastm;
                    T.sint = T.s_value();
                }
                int m = T.mint;
                double n = m >= 0 ? p.dens[m] : 0.;
                double ss = INST ? p.ssca[pk.ell] * p.inst[pk.inst].scale : p.ssca[pk.ell];
                double ksca = n > 0. ? n * ss : 0.;
                double kext = n > 0. ? n * T.section : 0.;
                double albedo = kext > 0. ? ksca / kext : 0.;
                pk.W *= -expm1(-pk.taupath) * albedo;
                if constexpr (GT != kCartesian)
                    pk.hint = m;  // the interaction point lies in the cell of its segment
                pk.rx += T.sint * pk.kx;
                pk.ry += T.sint * pk.ky;
                pk.rz += T.sint * pk.kz;
                double L = pk.W / pk.lam;
                if (L <= 0 || (L <= pk.Lthr && pk.nscatt >= p.min_scatt) || pk.nscatt >= kMaxScatterings)
                {
                    finish();
                    return;
                }
                begin_peel(true);
            }
        }
    };

    __device__ __forceinline__ int lane_id() { return __lane_id(); }

// Minimum waves per SIMD of the persistent kernel on Cartesian and octree grids.  Two waves
// give it 256 VGPRs: the event transitions (FP64 transcendentals on a loaded packet) need far
// more than the crossing loop, and on the octree at three or four waves the register spills to
// scratch cost more than the extra latency hiding gains.
#ifndef SKG_LIFECYCLE_WAVES
#define SKG_LIFECYCLE_WAVES 2
#endif
#ifndef SKG_LIFECYCLE_WAVES_CART
#define SKG_LIFECYCLE_WAVES_CART 2
#endif

    // Minimum waves per SIMD per grid type.  The Voronoi and tetra kernels hold less
    // traversal state and stage no octree nodes in LDS (35 KB per workgroup): measured on the
    // clumpy Voronoi workload 1152 / 1086 / 1061 ms per 2^22-packet launch at 2 / 3 / 4 waves
    // (each at its best transition threshold), on the batched tetra workload 6.72 / 5.80 /
    // 5.27 s.  The octree kernel loses with a
    // third wave (101.4 vs 142.5 ms per 2^22-packet launch on the spiral disk at tauV 50, the
    // top-node table halved to fit three workgroups per CU: 168 VGPRs spill the traversal
    // state), and the Cartesian one gains nothing (58.2 / 61.3 / 61.1 ms at 2 / 3 / 4 waves).
    template<int GT> constexpr int lifecycle_waves()
    {
        return GT == kTetra ? 4 : GT == kVoronoi ? 4 : GT == kCartesian ? SKG_LIFECYCLE_WAVES_CART : SKG_LIFECYCLE_WAVES;
    }

    // The kernel's parameter block where it lies in the kernarg segment, behind a pointer the
    // compiler cannot see through: fields read through the returned reference are loaded again
    // (scalar loads through the constant cache) instead of being held in registers across the
    // whole loop, which keeps the register pressure of the uniform values down.
    __device__ __forceinline__ const SimParams& kernel_params()
    {
        typedef const __attribute__((address_space(4))) SimParams* KernargPtr;
        KernargPtr q = (KernargPtr)__builtin_amdgcn_kernarg_segment_ptr();
        asm volatile("" : "+s"(q));
        return *(const SimParams*)q;
    }

    template<int GT, int FP, int V>
    __global__ void __launch_bounds__(kLcThreads, lifecycle_waves<GT>()) lifecycle(SimParams p)
    {
        constexpr bool INST = V == kBatched;
        __shared__ unsigned int wgcnt[kNumCounters];
        __shared__ double coldD[kColdF64 * kLcThreads];
        __shared__ uint32_t coldU[kColdU32 * kLcThreads];
        if (threadIdx.x < kNumCounters) wgcnt[threadIdx.x] = 0;
        if constexpr (GT == kOctree)
            stage_top(p.grid);  // other grids allocate no top-node LDS (stage_top ends with a barrier)
        else
            __syncthreads();  // the counters are zero before any wave adds to them
        const int t = threadIdx.x;
        Trav<GT, FP> T;
        T.phase = kFetch;
        T.ended = false;
        Pending pd{-1, 0.};
        const int lane = lane_id();
        while (true)
        {
            // advance every busy lane by one cell crossing
            bool busy = (T.phase == kPeel || T.phase == kFwd1 || T.phase == kFwd2) && !T.ended;
            const SimParams& sp = kernel_params();
            if (busy) T.template step<V == kStoreRF, V == kPersp || V == kStoreRF>(sp.grid, wgcnt, sp.rf, sp.nlambda);

            // wave-wide transition round once enough lanes wait for an event (or, guided, for the
            // network: lanes waiting to scatter call a round only once there are kGuideLanes of
            // them, so that they do not call transition rounds with few events to handle)
            const unsigned long long waiting = __ballot(T.ended || T.phase == kFetch);
            const unsigned long long scatter = V == kGuided ? __ballot(T.phase == kScatter) : 0ull;
            const unsigned long long active = __ballot(T.phase != kDone);
            if (active == 0) break;
            if (__popcll(waiting) >= sp.threshold || (waiting | scatter) == active
                || (V == kGuided && __popcll(scatter) >= kGuideLanes))
            {
                const SimParams& kp = kernel_params();
                if (T.ended)
                {
                    Packet pk;
                    pk.load<INST>(coldD, coldU, t, kp);
                    Events<GT, FP, V>{kp, T, pk, wgcnt, pd}.transition();
                    pk.store(coldD, coldU, t);
                }
                if constexpr (V == kGuided)
                {
                    // one network evaluation for the lanes of the wave that wait to scatter, once
                    // there are kGuideLanes of them or no lane of the wave is still traversing
                    const bool sc = T.phase == kScatter;
                    const unsigned long long scm = __ballot(sc);
                    const unsigned long long trav =
                        __ballot((T.phase == kPeel || T.phase == kFwd1 || T.phase == kFwd2) && !T.ended);
                    if (scm && (__popcll(scm) >= kGuideLanes || trav == 0))
                    {
                        Packet pk;
                        float feat[kGuideIn];
                        for (int q = 0; q < kGuideIn; ++q) feat[q] = 0.f;
                        if (sc)
                        {
                            pk.load<false>(coldD, coldU, t, kp);
                            guide_features(kp.gframe, pk.rx, pk.ry, pk.rz, pk.kx, pk.ky, pk.kz, pk.lam, feat);
                        }
                        float o[kGuideOut];
                        guide_wave_infer_regs(kp.net, feat, sc, o);
                        if (sc)
                        {
                            Events<GT, FP, V>{kp, T, pk, wgcnt, pd}.guided_scatter(o);
                            pk.store(coldD, coldU, t);
                        }
                    }
                }
                unsigned long long need = __ballot(T.phase == kFetch);
                if (need)
                {
                    int cntNeed = __popcll(need);
                    int leader = __ffsll(static_cast<long long>(need)) - 1;
                    unsigned long long base = 0;
                    if (lane == leader) base = atomicAdd(kp.queue, static_cast<unsigned long long>(cntNeed));
                    base = __shfl(base, leader);
                    if (T.phase == kFetch)
                    {
                        unsigned long long below = need & ((1ull << lane) - 1ull);
                        unsigned long long mine = base + __popcll(below);
                        if (mine < kp.count)
                        {
                            Packet pk;
                            Events<GT, FP, V>{kp, T, pk, wgcnt, pd}.launch(kp.first + mine);
                            pk.store(coldD, coldU, t);
                        }
                        else
                            T.phase = kDone;
                    }
                }
                flush_deposits(kp.frame, pd, wgcnt);  // the wave is converged here
            }
        }
        // workgroup event counters to the global counters
        __syncthreads();
        if (threadIdx.x < kNumCounters && wgcnt[threadIdx.x])
            atomicAdd(p.counters + threadIdx.x, static_cast<unsigned long long>(wgcnt[threadIdx.x]));
    }
}
