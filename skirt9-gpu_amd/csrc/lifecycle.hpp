// Device pieces of the forced-scattering photon-packet life cycle.
//
// Random numbers come from one Philox4x32-10 stream per history (philox.hpp).
// Distributions follow SKIRT's Random (Random.cpp:98-164) and DustMix HG code
// (DustMix.cpp:395-512); peel-off detection follows FrameInstrument.cpp:117-136.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "philox.hpp"

namespace skg
{
    // Direction(theta, phi), Direction.cpp:11-36
    __device__ __forceinline__ void direction_from_angles(double theta, double phi, double& kx, double& ky, double& kz)
    {
        const double eps = 1e-8;
        if (theta <= eps)
        {
            kx = 0;
            ky = 0;
            kz = 1;
        }
        else if (theta >= M_PI - eps)
        {
            kx = 0;
            ky = 0;
            kz = -1;
        }
        else
        {
            double st = sin(theta);
            kx = st * cos(phi);
            ky = st * sin(phi);
            kz = cos(theta);
        }
    }

    // Random::direction(), Random.cpp:121-126
    __device__ __forceinline__ void isotropic(Philox& rng, double& kx, double& ky, double& kz)
    {
        double theta = acos(2.0 * rng.uniform() - 1.0);
        double phi = 2.0 * M_PI * rng.uniform();
        direction_from_angles(theta, phi, kx, ky, kz);
    }

    // Random::direction(bfk, costheta), Random.cpp:130-164
    __device__ __forceinline__ void rotate(Philox& rng, double& kx, double& ky, double& kz, double costheta)
    {
        double phi = 2.0 * M_PI * rng.uniform();
        double cosphi = cos(phi);
        double sinphi = sin(phi);
        double sintheta = sqrt(fabs((1.0 - costheta) * (1.0 + costheta)));
        double x = kx, y = ky, z = kz;
        if (z > 0.99999)
        {
            kx = cosphi * sintheta;
            ky = sinphi * sintheta;
            kz = costheta;
        }
        else if (z < -0.99999)
        {
            kx = cosphi * sintheta;
            ky = sinphi * sintheta;
            kz = -costheta;
        }
        else
        {
            double root = sqrt((1.0 - z) * (1.0 + z));
            kx = sintheta / root * (-x * z * cosphi + y * sinphi) + x * costheta;
            ky = -sintheta / root * (y * z * cosphi + x * sinphi) + y * costheta;
            kz = root * sintheta * cosphi + z * costheta;
        }
    }

    // DustMix::performScattering, HG branch (DustMix.cpp:496-512)
    __device__ __forceinline__ void scatter_hg(Philox& rng, double g, double& kx, double& ky, double& kz)
    {
        if (fabs(g) < 1e-6)
            isotropic(rng, kx, ky, kz);
        else
        {
            double f = ((1.0 - g) * (1.0 + g)) / (1.0 - g + 2.0 * g * rng.uniform());
            double costheta = (1.0 + g * g - f * f) / (2.0 * g);
            rotate(rng, kx, ky, kz, costheta);
        }
    }

    // Random::exponCutoff (Random.cpp:105-117)
    __device__ __forceinline__ double expon_cutoff(Philox& rng, double xmax)
    {
        if (xmax == 0.0) return 0.0;
        if (xmax < 1e-10) return rng.uniform() * xmax;
        double x = -log(1.0 - rng.uniform() * (1.0 - exp(-xmax)));
        while (x > xmax) x = -log(1.0 - rng.uniform() * (1.0 - exp(-xmax)));
        return x;
    }

    // HG phase function value and its 4-degree average for |g| > 0.95 (DustMix.cpp:395-445)
    __device__ __forceinline__ double value_hg(double g, double c)
    {
        double t = 1. + g * g - 2. * g * c;
        return (1. - g) * (1. + g) / sqrt(t * t * t);
    }
    __device__ __forceinline__ double integral_hg(double g, double ca, double cb)
    {
        double ta = sqrt(1. + g * g - 2. * g * ca);
        double tb = sqrt(1. + g * g - 2. * g * cb);
        double f1 = (1. - g) * (1. + g) / g;
        double f2 = (tb - ta) / (tb * ta);
        return f1 * f2;
    }
    __device__ __forceinline__ double peel_phase(double g, double c)
    {
        if (!(fabs(g) > 0.95)) return value_hg(g, c);
        const double delta = 4. * M_PI / 180.;
        double theta = acos(c);
        double ca = cos(theta - delta);
        double cb = cos(theta + delta);
        if (theta < delta) return (integral_hg(g, 1., ca) + integral_hg(g, 1., cb)) / (2. - ca - cb);
        if (theta > M_PI - delta) return (integral_hg(g, ca, -1.) + integral_hg(g, cb, -1.)) / (2. + ca + cb);
        return integral_hg(g, ca, cb) / (ca - cb);
    }

    // distant frame instrument geometry (FrameInstrument.cpp:12-33, DistantInstrument.cpp:40-48)
    struct FrameGeom
    {
        double costheta, sintheta, cosphi, sinphi, cosomega, sinomega;
        double xpmin, xpsiz, ypmin, ypsiz;
        double kobs[3];
        int nx, ny;
        int same_as_prev;
        int sed;          // SEDInstrument: one pixel, inside the aperture radius (0: no aperture)
        int64_t offset;   // first element of this instrument in the frame buffer
        double radius2;
        // PerspectiveInstrument: world -> (pixel x, pixel y, depth, w) transform (row vector
        // times M, row-major), the eye and the pixel size
        int persp;
        int pad2;
        double M[16];
        double eye[3];
        double psize;

        // PerspectiveInstrument::detect (PerspectiveInstrument.cpp:103-116): the pixel (-1 if
        // off the viewport or too close) and the depth zp, the distance the flux and the
        // optical depth are taken over
        __device__ __forceinline__ int persp_pixel(double x, double y, double z, double& zp) const
        {
            const double xp = x * M[0] + y * M[4] + z * M[8] + M[12];
            const double yp = x * M[1] + y * M[5] + z * M[9] + M[13];
            zp = x * M[2] + y * M[6] + z * M[10] + M[14];
            const double wp = x * M[3] + y * M[7] + z * M[11] + M[15];
            const int i = static_cast<int>(xp / wp);
            const int j = static_cast<int>(yp / wp);
            if (i >= 0 && i < nx && j >= 0 && j < ny && zp > psize / 10.) return i + nx * j;
            return -1;
        }

        // the pixel of either kind of instrument (perspective: persp_pixel without its depth)
        __device__ __forceinline__ int any_pixel(double x, double y, double z) const
        {
            if (persp)
            {
                double zp;
                return persp_pixel(x, y, z, zp);
            }
            return pixel(x, y, z);
        }

        // FrameInstrument::pixelOnDetector (FrameInstrument.cpp:117-136); for an SED instrument
        // ApertureInstrument::isInsideAperture (ApertureInstrument.cpp:24-43: the projected
        // radius without the roll) and pixel 0
        __device__ __forceinline__ int pixel(double x, double y, double z) const
        {
            if (sed)
            {
                if (radius2 != 0.)
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
            int i = static_cast<int>(floor((xp - xpmin) / xpsiz));
            int j = static_cast<int>(floor((yp - ypmin) / ypsiz));
            if (i < 0 || i >= nx || j < 0 || j >= ny) return -1;
            return i + nx * j;
        }
    };

    constexpr int kMaxFrames = 1024;

    // one instance of a batched launch: its histories are [offset, next offset) of the launch
    struct Instance
    {
        uint64_t offset;
        uint64_t seed;
        double scale;  // optical-depth scale: sections multiplied by it
        int32_t fbeg, fend;
    };

    // compensated (Kahan) FP32 accumulator used in SKG_FP32_KAHAN mode
    struct Kahan32
    {
        float sum, comp;
        __device__ __forceinline__ void reset()
        {
            sum = 0.f;
            comp = 0.f;
        }
        __device__ __forceinline__ void add(float x)
        {
            float y = x - comp;
            float t = sum + y;
            comp = (t - sum) - y;
            sum = t;
        }
    };
}
