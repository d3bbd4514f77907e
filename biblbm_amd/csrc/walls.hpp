// Analytic boundaries of the fsi module, shared by host and device code:
// a pipe along x and two parallel plates normal to y
// (src/fsi/Boundary.h:132-186,269-273 and 305-365), in FP64 with the
// reference's expressions.  Contraction into fused multiply-adds is off so
// that host and device evaluate them the same way.
#pragma once

#include "common.hpp"

namespace bl {

__host__ __device__ inline double bsqr(double x) { return x * x; }

__host__ __device__ inline bool bnd_contains(const bl_boundary& b, const double* p, double margin)
{
#pragma clang fp contract(off)
    if (b.kind == BL_BOUNDARY_PIPE) return (bsqr(p[1] - b.p[0]) + bsqr(p[2] - b.p[1])) < bsqr(b.p[2] - margin);
    return p[1] > (b.p[0] + margin) && p[1] < (b.p[1] - margin);
}

__host__ __device__ inline bool bnd_less_than(const bl_boundary& b, const double* p, double dist)
{
#pragma clang fp contract(off)
    if (b.kind == BL_BOUNDARY_PIPE) return (bsqr(p[1] - b.p[0]) + bsqr(p[2] - b.p[1])) > bsqr(b.p[2] - dist);
    return (p[1] + dist) > b.p[1] || (p[1] - dist) < b.p[0];
}

__host__ __device__ inline double bnd_distance(const bl_boundary& b, const double* p)
{
#pragma clang fp contract(off)
    if (b.kind == BL_BOUNDARY_PIPE) return b.p[2] - sqrt(bsqr(p[1] - b.p[0]) + bsqr(p[2] - b.p[1]));
    return (p[1] < 0.5 * (b.p[0] + b.p[1])) ? p[1] - b.p[0] : b.p[1] - p[1];
}

// First crossing of the segment x0 -> x1, as the parameter t in [0, 1]
// (pipe) or (0, 1) (plates).
__host__ __device__ inline bool bnd_trace_ray(const bl_boundary& b, const double* x0, const double* x1, double& t)
{
#pragma clang fp contract(off)
    if (b.kind == BL_BOUNDARY_PIPE) {
        const double y1 = x0[1] - b.p[0], y2 = x1[1] - b.p[0], z1 = x0[2] - b.p[1], z2 = x1[2] - b.p[1];
        const double a = (y2 - y1) * (y2 - y1) + (z2 - z1) * (z2 - z1);
        const double bb = y1 * (y2 - y1) + z1 * (z2 - z1);
        const double c = y1 * y1 + z1 * z1 - b.p[2] * b.p[2];
        const double disc = bb * bb - c * a;
        if (disc < 0. || fabs(a) < 10 * 2.220446049250313e-16) return false;
        const double bOverA = -bb / a;
        const double sqrtDisc = sqrt(disc) / a;
        t = bOverA - sqrtDisc;
        if (t >= 0. && t <= 1.) return true;
        t = bOverA + sqrtDisc;
        return t >= 0. && t <= 1.;
    }
    t = (b.p[1] - x0[1]) / (x1[1] - x0[1]);
    if (t > 0 && t < 1) return true;
    t = (b.p[0] - x0[1]) / (x1[1] - x0[1]);
    return t > 0 && t < 1;
}

__host__ __device__ inline void bnd_normal(const bl_boundary& b, const double* p, double* n)
{
#pragma clang fp contract(off)
    if (b.kind == BL_BOUNDARY_PIPE) {
        const double d1 = p[1] - b.p[0], d2 = p[2] - b.p[1];
        const double nn = sqrt(0.0 * 0.0 + d1 * d1 + d2 * d2);
        n[0] = 0.0 / nn;
        n[1] = d1 / nn;
        n[2] = d2 / nn;
        return;
    }
    n[0] = 0;
    n[1] = (p[1] < 0.5 * (b.p[0] + b.p[1])) ? -1.0 : 1.0;
    n[2] = 0;
}

}  // namespace bl
