// Wavefront-style life-cycle pipeline.
//
// The photon packets in flight live in a pool of slots (structure-of-arrays in HBM).  One
// pipeline iteration is two kernels:
//
//   event   one thread per slot: consumes the traversal results of the previous trace launch
//           and advances the packet's life cycle (launch, forced-propagation sampling, the
//           interaction point, weight updates, HG scattering), then appends the traversals
//           the packet needs next to a request queue -- one per distinct observer for
//           peel-off, plus the packet's own forced path (first pass: total optical depth;
//           second pass: locate the interaction point).  Appends are stream-compacted with
//           one atomic per wavefront (ballot + mbcnt prefix).  Terminated slots refill from
//           the global history counter in the same way.
//   trace   persistent kernel: each lane takes one request at a time from the queue
//           (wave-batched work stealing) and walks the grid cell by cell; peel-off requests
//           deposit L exp(-tau) into the frames at the end, forced-path requests write their
//           result back to the slot.  Its register footprint is small, so many wavefronts
//           hide the latency of the dependent grid reads.
//
// Iterations are captured in a hipGraph; the host replays it until no slot is active.
// The order of operations within each packet is the reference's
// (MonteCarloSimulation.cpp:614-694, :1492-1544, :2034-2092); only the interleaving of
// independent packets differs, which changes nothing but the order of detector additions.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/skirt_mi355x.h"
#include "guide.hpp"
#include "lifecycle.hpp"
#include "traverse.hpp"

namespace skg
{
    enum SlotState : int
    {
        kSlotEmpty = 0,    // needs a new history
        kSlotWaitFwd1 = 1, // waiting for the first pass (total optical depth)
        kSlotWaitFwd2 = 2, // waiting for the second pass (interaction point)
        kSlotDone = 3      // no histories left
    };

    enum RequestKind : int
    {
        kReqPeel = 0,
        kReqFwd1 = 1,
        kReqFwd2 = 2
    };

    struct Request
    {
        int32_t slot;
        int32_t kind;
        int32_t frame;  // peel-off: first instrument of the observer group
        int32_t pad;
        double L;       // peel-off: luminosity to deposit
        double limit;   // peel-off: optical depth beyond which the contribution vanishes;
                        // second pass: the interaction optical depth
    };

    struct Pool
    {
        double *rx, *ry, *rz, *kx, *ky, *kz;
        double *W, *lthr, *taupath, *tint;
        double *res_tau, *res_s;
        int32_t* res_m;
        uint64_t* hist;
        uint32_t* draw;
        int32_t *ell, *nscatt, *state;
        int64_t capacity;
    };

    struct PipeParams
    {
        GridView grid;
        Pool pool;
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
        int pad;
        uint64_t seed;
        uint64_t first, count;
        double* frame;
        const FrameGeom* frames;
        // queues: two request buffers used alternately; ctl holds the control words
        Request* q[2];
        double* qtau[2];  // peel-off results (optical depth) per request
        int64_t qcap;
        unsigned long long* ctl;  // [0] history counter, [1..2] queue counts, [3..4] queue heads,
                                  // [5..6] active slots per parity, [8..15] event counters
        // path guide: 0 plain HG, 1 guided (mixture with HG), 2 plain HG collecting training records
        int guideMode;
        double guideMix;  // probability of sampling the guide instead of HG
        GuideNet net;
        GuideFrame gframe;
        // training records (collect mode): 16 floats each (9 features, direction out, weight),
        // the history offset and the packet's accumulated detector luminosity at the record
        float* rec;
        uint32_t* recHist;
        double* recA;
        unsigned long long* recCount;
        int64_t recCap;
        double* contrib;     // per slot: luminosity its packet has deposited so far
        int64_t* recLast;    // per slot: record awaiting its luminosity mark, or -1
        double* finalLum;    // per history offset: total luminosity deposited by the packet
        const double* pixW;  // per frame entry: weight of a deposit in contrib, or null (weight 1)
    };

    constexpr int kCtlHist = 0;
    constexpr int kCtlCount = 1;
    constexpr int kCtlHead = 3;
    constexpr int kCtlActive = 5;
    constexpr int kCtlOverflow = 7;  // requests that did not fit the queue (the run then fails)
    constexpr int kCtlCnt = 8;
    // event counters (both engines): 0 histories, 1 peel-off traversals, 2 first forced passes,
    // 3 scatterings, 4 cell crossings, 5 FP64 detector atomics issued, 6 second forced passes,
    // 7 crossings of second passes, 8 detector deposits (before the wavefront combines them),
    // 9..15 unused
    constexpr int kNumCounters = 16;

    // wave-aggregated append of `want` entries per lane; returns the lane's first index
    __device__ __forceinline__ unsigned long long wave_reserve(unsigned long long* counter, unsigned want)
    {
        // prefix sum of `want` across the wavefront with one atomic by the first active lane
        unsigned long long mask = __ballot(want > 0);
        if (!mask) return 0;
        const int lane = __lane_id();
        unsigned incl = want;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            unsigned v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        unsigned total = __shfl(incl, 63);
        unsigned long long base = 0;
        int leader = __ffsll(static_cast<long long>(mask)) - 1;
        if (lane == leader) base = atomicAdd(counter, static_cast<unsigned long long>(total));
        base = __shfl(base, leader);
        return base + (incl - want);
    }

    // True when instrument f leads an observer group (the instruments that share its
    // observer) and some instrument of the group sees the position, so that the group gets
    // one peel-off request (FluxRecorder::detect skips the optical depth when no instrument
    // of the group sees the packet)
    __device__ __forceinline__ bool group_sees(const PipeParams& p, int f, double x, double y, double z)
    {
        if (p.frames[f].same_as_prev) return false;
        for (int f2 = f; f2 < p.nframes && (f2 == f || p.frames[f2].same_as_prev); ++f2)
            if (p.frames[f2].any_pixel(x, y, z) >= 0) return true;
        return false;
    }

    // ------------------------------------------------------------------ event kernel
    template<int GT, int GM> __global__ void __launch_bounds__(256) pipe_event(PipeParams p, int out)
    {
        const int lane = __lane_id();
        const int64_t slot = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        const bool valid = slot < p.pool.capacity;
        // reset the queue the following trace launch does not read (the one this kernel's
        // predecessor trace consumed) and this parity's active counter
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            p.ctl[kCtlCount + (out ^ 1)] = 0;
            p.ctl[kCtlHead + (out ^ 1)] = 0;
            p.ctl[kCtlActive + (out ^ 1)] = 0;
        }
        Pool& P = p.pool;
        int st = valid ? P.state[slot] : kSlotDone;
        Philox rng;
        double rx = 0, ry = 0, rz = 0, kx = 0, ky = 0, kz = 0, W = 0;
        int ell = 0, nscatt = 0;
        bool emitFwd = false;
        int emitFwdKind = kReqFwd1;
        // peel-off requests: one per observer group that sees the packet (counted here; the
        // groups are found again when the requests are written, so any number of instruments
        // works); the luminosity is recomputed then (from the direction before scattering,
        // kept in kxo..kzo)
        unsigned npeelReq = 0;
        bool peelScat = false;
        double kxo = 0, kyo = 0, kzo = 0;
        double Wpeel = 0.;  // weight of this event's peel-offs (before the guide's factor)
        bool terminated = false;
        unsigned cntHist = 0, cntScat = 0, cntFwd2 = 0;

        if (st == kSlotWaitFwd1 || st == kSlotWaitFwd2)
        {
            rng.init(p.seed, P.hist[slot]);
            rng.draw = P.draw[slot];
            ell = P.ell[slot];
            W = P.W[slot];
            nscatt = P.nscatt[slot];
            rx = P.rx[slot];
            ry = P.ry[slot];
            rz = P.rz[slot];
            kx = P.kx[slot];
            ky = P.ky[slot];
            kz = P.kz[slot];
        }
        if (st == kSlotWaitFwd1)
        {
            if (GM == 2)
            {
                int64_t rl = p.recLast[slot];
                if (rl >= 0)
                {
                    p.recA[rl] = p.contrib[slot];
                    p.recLast[slot] = -1;
                }
            }
            // MonteCarloSimulation::simulateForcedPropagation (:1492-1544), sampling part
            double taupath = P.res_tau[slot];
            if (taupath <= 0.)
            {
                terminated = true;
            }
            else
            {
                double xi = p.xi, tint;
                if (xi == 0.)
                    tint = expon_cutoff(rng, taupath);
                else
                {
                    tint = rng.uniform() < xi ? rng.uniform() * taupath : expon_cutoff(rng, taupath);
                    double pp = -exp(-tint) / expm1(-taupath);
                    double q = (1.0 - xi) * pp + xi / taupath;
                    W *= pp / q;
                }
                P.taupath[slot] = taupath;
                P.tint[slot] = tint;
                emitFwd = true;
                emitFwdKind = kReqFwd2;
                cntFwd2++;
            }
        }
        else if (st == kSlotWaitFwd2)
        {
            // interaction point, albedo weight, propagation, termination rule, peel-off, scattering
            double taupath = P.taupath[slot];
            int m = P.res_m[slot];
            double sint = P.res_s[slot];
            double n = m >= 0 ? p.dens[m] : 0.;
            double ksca = n > 0. ? n * p.ssca[ell] : 0.;
            double kext = n > 0. ? n * p.sext[ell] : 0.;
            double albedo = kext > 0. ? ksca / kext : 0.;
            W *= -expm1(-taupath) * albedo;
            rx += sint * kx;
            ry += sint * ky;
            rz += sint * kz;
            double lam = p.lambda[ell];
            double L = W / lam;
            if (L <= 0 || (L <= P.lthr[slot] && nscatt >= p.min_scatt) || nscatt >= (1 << 20))
            {
                terminated = true;
            }
            else
            {
                // scattering peel-off toward each distinct observer (FluxRecorder::detect skips the
                // optical depth when no instrument of the group sees the packet)
                for (int f = 0; f < p.nframes; ++f)
                    if (group_sees(p, f, rx, ry, rz)) ++npeelReq;
                peelScat = true;
                Wpeel = W;
                kxo = kx;
                kyo = ky;
                kzo = kz;
                emitFwd = true;
                emitFwdKind = kReqFwd1;
            }
        }

        // scattering: HG (DustMix::performScattering) or, in guided mode, the defensive mixture
        // (1 - mix) HG + mix guide with the weight Phi_HG / p_mix that keeps the estimator unbiased
        const bool doScatter = peelScat;
        if constexpr (GM == 1)
        {
            bool useGuide = false;
            if (doScatter) useGuide = rng.uniform() < p.guideMix;
            __shared__ uint16_t gIn[4][64 * 16];
            __shared__ float gOut[4][64 * 48];
            const int wv = threadIdx.x >> 6;
            float feat[kGuideIn];
            guide_features(p.gframe, rx, ry, rz, kx, ky, kz, p.lambda[ell], feat);
            float o[kGuideOut];
            // every thread of the block takes part (uniform control flow around the MFMA)
            guide_wave_infer(p.net, feat, o, gIn[wv], gOut[wv]);
            if (doScatter)
            {
                LobesF Lb;
                guide_lobes_f(o, Lb);
                const double g = p.gpar[ell];
                if (useGuide)
                    guide_sample_f(Lb, rng, kx, ky, kz);
                else
                    scatter_hg(rng, g, kx, ky, kz);
                double c = kxo * kx + kyo * ky + kzo * kz;
                double phi = (fabs(g) < 1e-6 ? 1.0 : value_hg(g, c)) / (4 * M_PI);
                double q = guide_pdf_f(Lb, kx, ky, kz);
                const double ratio = phi / ((1.0 - p.guideMix) * phi + p.guideMix * q);
                W *= ratio;
                // the termination rule (MonteCarloSimulation.cpp:643-660) must see the weight
                // without the guide's factors, or packets the guide down-weights would be cut
                // early and their later deposits lost (a bias); scaling the threshold with the
                // same factor applies the rule to W / (product of the factors)
                P.lthr[slot] *= ratio;
            }
        }
        else if (doScatter)
        {
            scatter_hg(rng, p.gpar[ell], kx, ky, kz);
        }
        if (doScatter)
        {
            nscatt++;
            cntScat++;
            if (GM == 2)
            {
                // training record: features of the event, the sampled direction, and the mark of
                // the luminosity deposited so far (set at this slot's next event)
                unsigned long long ri = atomicAdd(p.recCount, 1ull);
                if (ri < static_cast<unsigned long long>(p.recCap))
                {
                    float feat[kGuideIn];
                    guide_features(p.gframe, rx, ry, rz, kxo, kyo, kzo, p.lambda[ell], feat);
                    float* r = p.rec + ri * 16;
                    for (int q = 0; q < kGuideIn; ++q) r[q] = feat[q];
                    r[9] = static_cast<float>(kx);
                    r[10] = static_cast<float>(ky);
                    r[11] = static_cast<float>(kz);
                    r[12] = 0.f;
                    p.recHist[ri] = static_cast<uint32_t>(P.hist[slot] - p.first);
                    p.recLast[slot] = static_cast<int64_t>(ri);
                }
            }
        }
        if (terminated)
        {
            cntHist++;
            if (GM == 2) p.finalLum[P.hist[slot] - p.first] = p.contrib[slot];
            st = kSlotEmpty;
        }

        // refill empty slots from the history counter (one atomic per wavefront)
        bool need = valid && st == kSlotEmpty;
        unsigned long long h = wave_reserve(p.ctl + kCtlHist, need ? 1u : 0u);
        if (need)
        {
            if (h < p.count)
            {
                // launch: SourceSystem::launch, PhotonPacket::launch, isotropic point source
                rng.init(p.seed, p.first + h);
                ell = static_cast<int>(rng.uniform() * p.nlambda);
                if (ell >= p.nlambda) ell = p.nlambda - 1;
                double lam = p.lambda[ell];
                W = p.lpack[ell] * lam;
                rx = p.src[0];
                ry = p.src[1];
                rz = p.src[2];
                isotropic(rng, kx, ky, kz);
                nscatt = 0;
                if (!(W / lam > 0))
                {
                    // zero-luminosity packet: done immediately
                    cntHist++;
                    if (GM == 2) p.finalLum[h] = 0.;
                    st = kSlotEmpty;
                }
                else
                {
                    P.hist[slot] = p.first + h;
                    P.lthr[slot] = (W / lam) / p.min_weight_reduction;
                    if (GM == 2)
                    {
                        p.contrib[slot] = 0.;
                        p.recLast[slot] = -1;
                    }
                    for (int f = 0; f < p.nframes; ++f)
                        if (group_sees(p, f, rx, ry, rz)) ++npeelReq;
                    emitFwd = true;
                    emitFwdKind = kReqFwd1;
                }
            }
            else
            {
                st = kSlotDone;
            }
        }

        // write the packet state back and append its traversal requests
        if (emitFwd)
        {
            P.state[slot] = emitFwdKind == kReqFwd1 ? kSlotWaitFwd1 : kSlotWaitFwd2;
            P.draw[slot] = rng.draw;
            P.ell[slot] = ell;
            P.W[slot] = W;
            P.nscatt[slot] = nscatt;
            P.rx[slot] = rx;
            P.ry[slot] = ry;
            P.rz[slot] = rz;
            P.kx[slot] = kx;
            P.ky[slot] = ky;
            P.kz[slot] = kz;
        }
        else if (valid)
        {
            P.state[slot] = st;
        }
        unsigned want = (emitFwd ? 1u : 0u) + npeelReq;
        unsigned long long base = wave_reserve(p.ctl + kCtlCount + out, want);
        if (want)
        {
            Request* q = p.q[out];
            unsigned long long o = base;
            const double lam = p.lambda[ell];
            for (int f = 0; f < p.nframes && npeelReq; ++f)
            {
                if (!group_sees(p, f, rx, ry, rz)) continue;
                double Wp = W;
                const FrameGeom& fg = p.frames[f];
                // a perspective instrument is seen toward its eye (PerspectiveInstrument.cpp:93-101)
                double ox = fg.kobs[0], oy = fg.kobs[1], oz = fg.kobs[2], zp = 1.;
                if (fg.persp)
                {
                    fg.persp_pixel(rx, ry, rz, zp);
                    const double dx = fg.eye[0] - rx, dy = fg.eye[1] - ry, dz = fg.eye[2] - rz;
                    const double D = sqrt(dx * dx + dy * dy + dz * dz);
                    ox = dx / D;
                    oy = dy / D;
                    oz = dz / D;
                }
                if (peelScat)
                {
                    // the peel-off belongs to the event, before the new direction is drawn: the
                    // guide's factor of that direction applies to the continuing packet only
                    double c = kxo * ox + kyo * oy + kzo * oz;
                    Wp = Wpeel * peel_phase(p.gpar[ell], c);
                }
                if (o < static_cast<unsigned long long>(p.qcap))
                {
                    Request r;
                    r.slot = static_cast<int32_t>(slot);
                    r.kind = kReqPeel;
                    r.frame = f;
                    r.pad = 0;
                    r.L = Wp / lam;
                    r.limit = log(r.L) + 745;  // MediumSystem.cpp:2335
                    if (fg.persp) r.L /= zp * zp;  // local recorder (FluxRecorder.cpp:387)
                    q[o] = r;
                }
                ++o;
            }
            if (emitFwd && o < static_cast<unsigned long long>(p.qcap))
            {
                Request r;
                r.slot = static_cast<int32_t>(slot);
                r.kind = emitFwdKind;
                r.frame = -1;
                r.pad = 0;
                r.L = 0.;
                r.limit = emitFwdKind == kReqFwd2 ? P.tint[slot] : 0.;
                q[o] = r;
            }
            else if (emitFwd)
            {
                // the packet's traversal does not fit the queue: the packet stops here (it would
                // otherwise continue from results of no traversal), and the run fails with the
                // overflow count
                P.state[slot] = kSlotDone;
            }
            if (emitFwd) ++o;
            if (o > static_cast<unsigned long long>(p.qcap))
                atomicAdd(p.ctl + kCtlOverflow, o - max(base, static_cast<unsigned long long>(p.qcap)));
        }
        // active slots and event counters, aggregated per wavefront
        unsigned long long act = __ballot(emitFwd);
        unsigned long long f1 = __popcll(__ballot(emitFwd && emitFwdKind == kReqFwd1));
        unsigned long long nh = 0, ns = 0, nf = 0, npeel = 0;
        {
            unsigned a = cntHist, b = cntScat, c = cntFwd2, d = npeelReq;
            for (int o = 32; o > 0; o >>= 1)
            {
                a += __shfl_xor(a, o);
                b += __shfl_xor(b, o);
                c += __shfl_xor(c, o);
                d += __shfl_xor(d, o);
            }
            nh = a;
            ns = b;
            nf = c;
            npeel = d;
        }
        if (lane == 0)
        {
            if (act) atomicAdd(p.ctl + kCtlActive + out, static_cast<unsigned long long>(__popcll(act)));
            if (nh) atomicAdd(p.ctl + kCtlCnt + 0, nh);
            if (npeel) atomicAdd(p.ctl + kCtlCnt + 1, npeel);
            if (ns) atomicAdd(p.ctl + kCtlCnt + 3, ns);
            if (nf) atomicAdd(p.ctl + kCtlCnt + 6, nf);
            if (f1) atomicAdd(p.ctl + kCtlCnt + 2, f1);
        }
        (void)lane;
    }

    // ------------------------------------------------------------------ trace kernel
    constexpr unsigned kTraceChunk = 256;  // request indices a wavefront reserves at once
    constexpr int kTraceRefill = 8;        // idle lanes that trigger a refill round
#ifndef SKG_TRACE_WAVES
#define SKG_TRACE_WAVES 4  // minimum waves per SIMD of the trace kernel
#endif
    template<int GT, int FP> __global__ void __launch_bounds__(256, SKG_TRACE_WAVES) pipe_trace(PipeParams p, int in)
    {
        stage_top(p.grid);
        const int lane = __lane_id();
        const unsigned long long count = p.ctl[kCtlCount + in];
        const Request* q = p.q[in];
        const Pool& P = p.pool;
        Gen<GT> gen;
        bool busy = false;
        bool done = false;
        int kind = 0, slot = 0, steps = 0, lastm = -1;
        unsigned long long req = 0;
        double tau = 0., s = 0., tprev = 0., sprev = 0., limit = 0.;
        double slim = INFINITY;  // perspective peel-off: path length after which it ends
        float tau32 = 0.f, ctau = 0.f, s32 = 0.f, cs = 0.f;
        double section = 0.;
        float section32 = 0.f;
        unsigned long long segs = 0;
        // the wavefront's reserved run of request indices [cnext, cend), refilled with one
        // atomic per kTraceChunk requests; idle lanes are refilled once kTraceRefill of them
        // wait (or when no lane is busy), so most crossings run without a refill round
        unsigned long long cnext = 0, cend = 0;
        while (true)
        {
            bool idle = !busy && !done;
            const unsigned long long want = __ballot(idle);
            const int nidle = __popcll(want);
            if (want && (nidle >= kTraceRefill || __ballot(busy) == 0))
            {
                const unsigned long long avail = cend - cnext;
                unsigned long long nb = 0;
                if (static_cast<unsigned long long>(nidle) > avail)
                {
                    const int leader = __ffsll(static_cast<long long>(want)) - 1;
                    if (lane == leader) nb = atomicAdd(p.ctl + kCtlHead + in, static_cast<unsigned long long>(kTraceChunk));
                    nb = __shfl(nb, leader);
                }
                const unsigned idx = __popcll(want & ((1ull << lane) - 1ull));
                const unsigned long long base = idx < avail ? cnext + idx : nb + (idx - avail);
                if (static_cast<unsigned long long>(nidle) > avail)
                {
                    cnext = nb + (nidle - avail);
                    cend = nb + kTraceChunk;
                }
                else
                    cnext += nidle;
                if (idle)
                {
                    if (base < count && base < static_cast<unsigned long long>(p.qcap))
                    {
                        const Request& r = q[base];
                        req = base;
                        slot = r.slot;
                        kind = r.kind;
                        limit = r.limit;
                        int ell = P.ell[slot];
                        section = p.sext[ell];
                        section32 = static_cast<float>(section);
                        double dx, dy, dz;
                        slim = INFINITY;
                        if (kind == kReqPeel)
                        {
                            const FrameGeom& fg = p.frames[r.frame];
                            dx = fg.kobs[0];
                            dy = fg.kobs[1];
                            dz = fg.kobs[2];
                            if (fg.persp)
                            {
                                // toward the eye, over the path length zp (FluxRecorder.cpp:372-405)
                                const double rx0 = P.rx[slot], ry0 = P.ry[slot], rz0 = P.rz[slot];
                                fg.persp_pixel(rx0, ry0, rz0, slim);
                                const double ex = fg.eye[0] - rx0, ey = fg.eye[1] - ry0, ez = fg.eye[2] - rz0;
                                const double D = sqrt(ex * ex + ey * ey + ez * ez);
                                dx = ex / D;
                                dy = ey / D;
                                dz = ez / D;
                            }
                        }
                        else
                        {
                            dx = P.kx[slot];
                            dy = P.ky[slot];
                            dz = P.kz[slot];
                        }
                        gen.start(P.rx[slot], P.ry[slot], P.rz[slot], dx, dy, dz);
                        tau = s = tprev = sprev = 0.;
                        tau32 = ctau = s32 = cs = 0.f;
                        lastm = -1;
                        steps = 0;
                        busy = true;
                    }
                    else
                        done = true;
                }
            }
            if (__ballot(busy) == 0) break;
            if (!busy) continue;

            // one cell crossing
            int m;
            double ds, nd;
            bool more = gen.next(p.grid, m, ds, nd) && ++steps <= (1 << 26);
            bool finished = !more;
            if (more)
            {
                ++segs;
                if (kind == kReqPeel)
                {
                    if (m >= 0)
                    {
                        if (FP == SKG_FP64)
                        {
                            tau += section * nd * ds;
                            if (tau >= limit)
                            {
                                tau = INFINITY;
                                finished = true;
                            }
                        }
                        else
                        {
                            float y = section32 * static_cast<float>(nd) * static_cast<float>(ds) - ctau;
                            float t = tau32 + y;
                            ctau = (t - tau32) - y;
                            tau32 = t;
                            if (static_cast<double>(tau32) >= limit)
                            {
                                tau32 = INFINITY;
                                finished = true;
                            }
                        }
                    }
                    // MediumSystem::getExtinctionOpticalDepth(pp, distance) (MediumSystem.cpp:
                    // 2376-2382): the segment that passes the distance still counts
                    s += ds;
                    if (s > slim) finished = true;
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
                        tprev = static_cast<double>(tau32);
                        sprev = static_cast<double>(s32);
                        float y = static_cast<float>(ds) - cs;
                        float t = s32 + y;
                        cs = (t - s32) - y;
                        s32 = t;
                        if (m >= 0)
                        {
                            float y2 = section32 * static_cast<float>(nd) * static_cast<float>(ds) - ctau;
                            float t2 = tau32 + y2;
                            ctau = (t2 - tau32) - y2;
                            tau32 = t2;
                        }
                    }
                    lastm = m;
                    if (kind == kReqFwd2)
                    {
                        double tv = FP == SKG_FP64 ? tau : static_cast<double>(tau32);
                        if (limit < tv)
                        {
                            // SpatialGridPath::findInteractionPoint: first segment whose exit optical
                            // depth exceeds tau_int, linear interpolation of the distance
                            double sv = FP == SKG_FP64 ? s : static_cast<double>(s32);
                            P.res_m[slot] = m;
                            P.res_s[slot] = sprev + ((limit - tprev) / (tv - tprev)) * (sv - sprev);
                            busy = false;
                            continue;
                        }
                    }
                }
            }
            if (finished)
            {
                double tv = FP == SKG_FP64 ? tau : static_cast<double>(tau32);
                if (kind == kReqPeel)
                {
                    p.qtau[in][req] = tv;
                }
                else if (kind == kReqFwd1)
                {
                    P.res_tau[slot] = tv;
                }
                else
                {
                    // beyond the last segment: the last segment's cell and exit distance
                    P.res_m[slot] = lastm;
                    P.res_s[slot] = FP == SKG_FP64 ? s : static_cast<double>(s32);
                }
                busy = false;
            }
        }
        // segment counter, one atomic per wavefront
        for (int o = 32; o > 0; o >>= 1) segs += __shfl_xor(segs, o);
        if (lane == 0 && segs) atomicAdd(p.ctl + kCtlCnt + 4, segs);
    }

    // Adds v at base[key] for every lane with key >= 0, called by all 64 lanes of the wave.
    // The lanes whose key equals the first such lane's combine into one atomic (a wave of
    // emission peel-offs from a point source all land in the source's pixel); the others add
    // on their own.
    __device__ __forceinline__ void wave_add(double* base, int64_t key, double v, unsigned long long& atomics)
    {
        const unsigned long long has = __ballot(key >= 0);
        if (!has) return;
        const int leader = __ffsll(static_cast<long long>(has)) - 1;
        const int64_t lk = __shfl(key, leader);
        const bool same = key == lk;
        double sum = same ? v : 0.;
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (__lane_id() == leader)
        {
            unsafeAtomicAdd(base + lk, sum);
            ++atomics;
        }
        else if (key >= 0 && !same)
        {
            unsafeAtomicAdd(base + key, v);
            ++atomics;
        }
    }

    // ------------------------------------------------------------------ detect kernel
    // FluxRecorder::detect (FluxRecorder.cpp:372-418) for the peel-off requests of one trace
    // launch: L exp(-tau) into every instrument of the observer group that sees the packet.
    // The loop runs with whole wavefronts so that the first instrument's deposits can be
    // combined across the wave before the atomics.
    __global__ void __launch_bounds__(256) pipe_detect(PipeParams p, int in)
    {
        const unsigned long long count = p.ctl[kCtlCount + in];
        const Pool& P = p.pool;
        unsigned long long atomics = 0, deposits = 0;
        const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
        const unsigned long long n = count < static_cast<unsigned long long>(p.qcap) ? count : p.qcap;
        for (unsigned long long i0 = blockIdx.x * static_cast<unsigned long long>(blockDim.x); i0 < n; i0 += stride)
        {
            const unsigned long long i = i0 + threadIdx.x;
            bool valid = i < n;
            int kind = valid ? p.q[in][i].kind : -1;
            valid = valid && kind == kReqPeel;
            int64_t key = -1;
            double v = 0.;
            int slot = 0, f0 = 0, ell = 0;
            double x = 0., y = 0., z = 0.;
            if (valid)
            {
                const Request& r = p.q[in][i];
                v = r.L * exp(-p.qtau[in][i]);
                slot = r.slot;
                f0 = r.frame;
                x = P.rx[slot];
                y = P.ry[slot];
                z = P.rz[slot];
                ell = P.ell[slot];
                const FrameGeom& fg = p.frames[f0];
                int pix = fg.any_pixel(x, y, z);
                if (pix >= 0) key = fg.offset + pix + static_cast<int64_t>(ell) * fg.nx * fg.ny;
            }
            wave_add(p.frame, key, v, atomics);
            if (key >= 0) ++deposits;
            if (valid)
            {
                double dep = key >= 0 ? v * (p.pixW ? p.pixW[key] : 1.) : 0.;
                // further instruments of the same observer (same direction, shared optical depth)
                for (int f = f0 + 1; f < p.nframes && p.frames[f].same_as_prev; ++f)
                {
                    const FrameGeom& fg = p.frames[f];
                    int pix = fg.any_pixel(x, y, z);
                    if (pix < 0) continue;
                    const int64_t k = fg.offset + pix + static_cast<int64_t>(ell) * fg.nx * fg.ny;
                    unsafeAtomicAdd(p.frame + k, v);
                    dep += v * (p.pixW ? p.pixW[k] : 1.);
                    ++atomics;
                    ++deposits;
                }
                if (p.guideMode == 2 && dep > 0.) atomicAdd(p.contrib + slot, dep);
            }
        }
        for (int o = 32; o > 0; o >>= 1)
        {
            atomics += __shfl_xor(atomics, o);
            deposits += __shfl_xor(deposits, o);
        }
        if (__lane_id() == 0 && atomics) atomicAdd(p.ctl + kCtlCnt + 5, atomics);
        if (__lane_id() == 0 && deposits) atomicAdd(p.ctl + kCtlCnt + 8, deposits);
    }

    // training weights of the collected records: the luminosity the packet deposited after the
    // record's scattering event (total at termination minus the mark at the event)
    __global__ void guide_finalize_records(float* rec, const uint32_t* recHist, const double* recA, const double* finalLum,
                                           int64_t n)
    {
        int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        if (i >= n) return;
        double c = finalLum[recHist[i]] - recA[i];
        rec[i * 16 + 12] = static_cast<float>(c > 0. ? c : 0.);
    }

    // resets all control words and marks every slot empty
    __global__ void pipe_reset(Pool pool, unsigned long long* ctl)
    {
        int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        if (i < kCtlCnt) ctl[i] = 0;
        if (i < pool.capacity) pool.state[i] = kSlotEmpty;
    }
}
