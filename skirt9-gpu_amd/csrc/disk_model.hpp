// The exponential-disk density model (optionally with a logarithmic spiral perturbation)
// normalised to a face-on optical depth: set up on the host, evaluated on the host (set-up,
// octree refinement, cell densities) and on the device (cell densities of a resident grid).
// ExpDiskGeometry (ExpDiskGeometry.cpp), SpiralStructureGeometryDecorator
// (SpiralStructureGeometryDecorator.cpp) and OpticalDepthMaterialNormalization.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

#include "../../include/skirt_mi355x.h"

namespace skg
{
    struct DiskModel
    {
        double hR, hz, rho0;
        bool spiral;
        double m, tanp, R0, phi0, w, cn;
        int N;
        double nnorm;  // number density per unit normalised geometric density

        explicit DiskModel(const skg_disk_model& d)
        {
            hR = d.scale_length;
            hz = d.scale_height;
            // ExpDiskGeometry::setupSelfBefore (no truncation): rho0 = 1 / (2 pi hR^2 * 2 hz)
            double intphi = 2.0 * M_PI;
            double intz = 2.0 * hz;
            double intR = hR * hR * (1.0 - 0.0);
            rho0 = 1.0 / (intR * intphi * intz);
            spiral = d.spiral_arms > 0;
            m = d.spiral_arms;
            N = d.spiral_index;
            tanp = std::tan(d.pitch_angle);
            R0 = d.radius_zero;
            phi0 = d.phase_zero;
            w = d.perturbation_weight;
            // SpiralStructureGeometryDecorator::setupSelfBefore: cn = sqrt(pi) Gamma(N+1) / Gamma(N+1/2)
            cn = std::sqrt(M_PI) * std::tgamma(N + 1.0) / std::tgamma(N + 0.5);
            // OpticalDepthMaterialNormalization along Z: number = tau / (sigma * SigmaZ), SigmaZ = 2 rho0 hz
            double sigmaZ = 2.0 * rho0 * hz;
            nnorm = d.tau_v / d.sext_v / sigmaZ;
        }

        // normalised geometric density (integrates to one)
        __host__ __device__ double geom(double x, double y, double z) const
        {
            double R = sqrt(x * x + y * y);
            double v = rho0 * exp(-R / hR) * exp(-fabs(z) / hz);
            if (spiral)
            {
                double phi = atan2(y, x);
                double gamma = log(R / R0) / tanp + phi0 + 0.5 * M_PI / m;
                v *= (1.0 - w) + w * cn * pow(sin(0.5 * m * (gamma - phi)), 2 * N);
            }
            return v;
        }
    };
}
