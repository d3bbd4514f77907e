// Interaction forces of the fsi module, shared by host and device code: the
// Morse and spring potentials (src/fsi/Potentials.h:28-72), the wall
// interaction decorator (src/fsi/ForceDecorator.hh:27-35) and the geometry
// of the collision handler's cell grid (src/fsi/Grid.hh:30-97 over the box
// of src/fsi/ImmersedBoundaryDynamics_forces.hh:336-342), in FP64 with the
// reference's expressions and no contraction into fused multiply-adds.
#pragma once

#include "common.hpp"
#include "walls.hpp"

namespace bl {

constexpr int kMaxWallDeco = 4;
constexpr int kMaxPPForce = 4;

// A vertex of the collision handler's cell lists: position, owner and the
// next larger vertex id in its cell (-1 at the end).
struct alignas(32) PPRec {
    double x, y, z;
    int obj, succ;
};

__host__ __device__ inline double pot_force(const bl_potential& u, double dist)
{
#pragma clang fp contract(off)
    if (u.kind == BL_POTENTIAL_MORSE) {
        if (dist < 0) dist = 0;
        const double two_beta_strength = 2 * u.p[1] * u.p[2];
        const double x = exp(u.p[1] * (u.p[0] - dist));
        return two_beta_strength * x * (x - 1);
    }
    return (u.p[0] - dist) * u.p[1];
}

__host__ __device__ inline double pot_cutoff(const bl_potential& u) { return u.p[0]; }

// The wall-interaction decorators of a boundary, in the order added.
struct DecoSet {
    int n;
    bl_boundary b[kMaxWallDeco];
    bl_potential u[kMaxWallDeco];
};

// force -= F(d) n of every decorator, on a force that starts at f.
__host__ __device__ inline void deco_apply(const DecoSet& s, const double* p, double f[3])
{
#pragma clang fp contract(off)
    for (int k = 0; k < s.n; ++k) {
        const double d = bnd_distance(s.b[k], p);
        if (d <= pot_cutoff(s.u[k])) {
            const double F = pot_force(s.u[k], d);
            double n[3];
            bnd_normal(s.b[k], p, n);
            f[0] -= F * n[0];
            f[1] -= F * n[1];
            f[2] -= F * n[2];
        }
    }
}

// Cell grid of the collision handler over the lattice box [0, n-1]^3:
// bulk cells 1..num per axis and a one-cell envelope; cell index
// k + (num_z+2) (j + (num_y+2) i).
struct CollGrid {
    double bb_lo[3], cs[3], x0g[3], lo[3], hi[3], L[3], Linv[3];
    int num[3];
    int per[3];
    double cut, cut2;
};

inline CollGrid coll_grid(int nx, int ny, int nz, const int periodic[3], double cutoff)
{
#pragma clang fp contract(off)
    CollGrid g{};
    const int n[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d) {
        const double lo = 0 - 0.5, hi = (double)(n[d] - 1) + 0.5 - 2.220446049250313e-16;
        g.bb_lo[d] = lo;
        g.num[d] = (int)std::floor((hi - lo) / 1.0);
        g.cs[d] = (hi - lo) / (double)g.num[d];
        g.x0g[d] = lo - g.cs[d];
        g.lo[d] = lo - cutoff;
        g.hi[d] = hi + cutoff;
        g.L[d] = (double)n[d];
        g.Linv[d] = 1. / g.L[d];
        g.per[d] = periodic[d] != 0;
    }
    g.cut = cutoff;
    g.cut2 = cutoff * cutoff;
    return g;
}

inline long coll_ncell(const CollGrid& g) { return (long)(g.num[0] + 2) * (g.num[1] + 2) * (g.num[2] + 2); }

// Cell of a node, or -1 if it does not take part: outside the box enlarged
// by the cut-off (the reference's strict containment test), or past the
// grid's envelope (where the reference would index beyond its grid, which
// only cut-offs above one cell allow).
__host__ __device__ inline int coll_cell(const CollGrid& g, const double* p)
{
#pragma clang fp contract(off)
    int ix[3];
    for (int d = 0; d < 3; ++d)
        if (!(p[d] > g.lo[d] && p[d] < g.hi[d])) return -1;
    for (int d = 0; d < 3; ++d) {
        const double q = floor((p[d] - g.x0g[d]) / g.cs[d]);
        if (q < 0 || q > (double)(g.num[d] + 1)) return -1;
        ix[d] = (int)q;
    }
    return ix[2] + (g.num[2] + 2) * (ix[1] + (g.num[1] + 2) * ix[0]);
}

// Boundary::does_intersect(rect, margin) (src/fsi/Boundary.h:227-242,
// 316-320) for the handler's wall forces.
__host__ __device__ inline double bnd_clamp(double v, double lo, double hi)
{
    if (v >= hi) return hi;
    if (v <= lo) return lo;
    return v;
}

__host__ __device__ inline bool bnd_intersects_rect(const bl_boundary& b, const double r[6], double margin)
{
#pragma clang fp contract(off)
    if (b.kind == BL_BOUNDARY_PIPE) {
        const double y0 = b.p[0], z0 = b.p[1], R = b.p[2];
        const double tmp = bsqr(y0 - bnd_clamp(y0, r[2], r[3])) + bsqr(z0 - bnd_clamp(z0, r[4], r[5]));
        if (tmp <= bsqr(R)) {
            const double q00[3] = {0, r[2], r[4]}, q01[3] = {0, r[2], r[5]}, q10[3] = {0, r[3], r[4]},
                         q11[3] = {0, r[3], r[5]};
            return !(bnd_contains(b, q00, margin) && bnd_contains(b, q01, margin) && bnd_contains(b, q10, margin) &&
                     bnd_contains(b, q11, margin));
        }
        return false;
    }
    return (r[2] < (b.p[0] + margin) && r[3] > (b.p[0] + margin)) ||
           (r[2] < (b.p[1] - margin) && r[3] > (b.p[1] - margin));
}

}  // namespace bl
