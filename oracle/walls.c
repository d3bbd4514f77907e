/*
 * oracle/walls.c -- FP64 CPU restatement of the reference's curved-wall
 * pieces: analytic boundaries, the Guo-Zheng-Shi extrapolation wall, and the
 * immersed boundary's near-wall kernel blending and vertex reflection.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).  Never linked into
 * the product library.
 *
 * Follows, operation by operation:
 *   PipeBoundary                   src/fsi/Boundary.h:132-186,269-273
 *   ParallelPlatesBoundary         src/fsi/Boundary.h:305-365
 *   wall-node instantiation        src/fsi/GuoBoundaryCondition.hh:62-135
 *   Guo extrapolation (depth 2)    src/fsi/GuoBoundaryCondition.hh:138-199
 *   near-wall interpolation blend  src/fsi/ImmersedBoundaryDynamics_interpolation.hh:182-230
 *   near-wall spreading blend      src/fsi/ImmersedBoundaryDynamics_forces.hh:193-237
 *   node classification            src/fsi/ImmersedBoundaryDynamics.hh:346-357
 *   spreading order (bulk, then boundary nodes)
 *                                  src/fsi/ImmersedBoundaryDynamics_forces.hh:259-277
 *   vertex reflection              src/fsi/DeformableParticle3D.hh:26-50
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

static const int C[19][3] = {
    {0, 0, 0},
    {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
    {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
    {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};

#define T0 (1.0 / 3.0)
#define T1 (1.0 / 18.0)
#define T4 (1.0 / 36.0)
static const double TW[19] = {T0, T1, T1, T1, T4, T4, T4, T4, T4, T4,
                              T1, T1, T1, T4, T4, T4, T4, T4, T4};

static inline double sqr(double x) { return x * x; }

/* ------------------------------------------------------------ boundaries */
int or_boundary_contains(const or_boundary *b, const double p[3], double margin)
{
    if (b->kind == OR_BOUNDARY_PIPE)
        return (sqr(p[1] - b->p[0]) + sqr(p[2] - b->p[1])) < sqr(b->p[2] - margin);
    return p[1] > (b->p[0] + margin) && p[1] < (b->p[1] - margin);
}

int or_boundary_less_than(const or_boundary *b, const double p[3], double dist)
{
    if (b->kind == OR_BOUNDARY_PIPE)
        return (sqr(p[1] - b->p[0]) + sqr(p[2] - b->p[1])) > sqr(b->p[2] - dist);
    return (p[1] + dist) > b->p[1] || (p[1] - dist) < b->p[0];
}

double or_boundary_distance(const or_boundary *b, const double p[3])
{
    if (b->kind == OR_BOUNDARY_PIPE)
        return b->p[2] - sqrt(sqr(p[1] - b->p[0]) + sqr(p[2] - b->p[1]));
    return (p[1] < 0.5 * (b->p[0] + b->p[1])) ? p[1] - b->p[0] : b->p[1] - p[1];
}

int or_boundary_trace_ray(const or_boundary *b, const double x0[3], const double x1[3], double *t)
{
    if (b->kind == OR_BOUNDARY_PIPE) {
        const double y1 = x0[1] - b->p[0], y2 = x1[1] - b->p[0], z1 = x0[2] - b->p[1], z2 = x1[2] - b->p[1];
        const double a = (y2 - y1) * (y2 - y1) + (z2 - z1) * (z2 - z1);
        const double bb = y1 * (y2 - y1) + z1 * (z2 - z1);
        const double c = y1 * y1 + z1 * z1 - b->p[2] * b->p[2];
        const double disc = bb * bb - c * a;
        if (disc < 0. || fabs(a) < 10 * 2.220446049250313e-16) return 0;
        const double bOverA = -bb / a;
        const double sqrtDisc = sqrt(disc) / a;
        *t = bOverA - sqrtDisc;
        if (*t >= 0. && *t <= 1.) return 1;
        *t = bOverA + sqrtDisc;
        if (*t >= 0. && *t <= 1.) return 1;
        return 0;
    }
    *t = (b->p[1] - x0[1]) / (x1[1] - x0[1]);
    if (*t > 0 && *t < 1) return 1;
    *t = (b->p[0] - x0[1]) / (x1[1] - x0[1]);
    if (*t > 0 && *t < 1) return 1;
    return 0;
}

void or_boundary_normal(const or_boundary *b, const double p[3], double n[3])
{
    if (b->kind == OR_BOUNDARY_PIPE) {
        const double d[3] = {0, p[1] - b->p[0], p[2] - b->p[1]};
        const double nn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        n[0] = d[0] / nn;
        n[1] = d[1] / nn;
        n[2] = d[2] / nn;
        return;
    }
    n[0] = 0;
    n[1] = (p[1] < 0.5 * (b->p[0] + b->p[1])) ? -1.0 : 1.0;
    n[2] = 0;
}

/* ------------------------------------------------------- Guo wall nodes */
long or_guo_wall_nodes(int nx, int ny, int nz, const or_boundary *b, long cap, long *cell, int *dir,
                       double *delta, double *wall_pos)
{
    long n = 0;
    double norms[19];
    for (int i = 1; i < 19; ++i) norms[i] = sqrt((double)(C[i][0] * C[i][0] + C[i][1] * C[i][1] + C[i][2] * C[i][2]));
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const double pos[3] = {x, y, z};
                if (or_boundary_contains(b, pos, 0)) continue;
                double cos_min = 2, delta_min = -1, wall_min[3] = {0, 0, 0};
                int dir_min = 0, success = 0;
                for (int i = 1; i < 19; ++i) {
                    const double tip[3] = {pos[0] + C[i][0], pos[1] + C[i][1], pos[2] + C[i][2]};
                    double t = -1;
                    if (!or_boundary_trace_ray(b, pos, tip, &t)) continue;
                    const double d2[3] = {C[i][0] / norms[i], C[i][1] / norms[i], C[i][2] / norms[i]};
                    /* the reference steps along the normalised direction */
                    const double wp[3] = {pos[0] + d2[0] * t, pos[1] + d2[1] * t, pos[2] + d2[2] * t};
                    double nrm[3];
                    or_boundary_normal(b, wp, nrm);
                    const double cosang = nrm[0] * d2[0] + nrm[1] * d2[1] + nrm[2] * d2[2];
                    if (cosang < cos_min) {
                        dir_min = i;
                        cos_min = cosang;
                        delta_min = 1 - t;
                        memcpy(wall_min, wp, sizeof(wp));
                    }
                    success = 1;
                }
                if (!success) continue;
                if (n < cap) {
                    cell[n] = ((long)x * ny + y) * nz + z;
                    dir[n] = dir_min;
                    delta[n] = delta_min;
                    memcpy(wall_pos + 3 * n, wall_min, sizeof(wall_min));
                }
                ++n;
            }
    return n;
}

static inline double rho_bar(const double *f)
{
    return f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] +
           f[10] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
}

static inline void get_j(const double *f, double j[3])
{
    double xm = f[1] + f[4] + f[5] + f[6] + f[7];
    double xp = f[10] + f[13] + f[14] + f[15] + f[16];
    double ym = f[2] + f[4] + f[8] + f[9] + f[14];
    double yp = f[5] + f[11] + f[13] + f[17] + f[18];
    double zm = f[3] + f[6] + f[8] + f[16] + f[18];
    double zp = f[7] + f[9] + f[12] + f[15] + f[17];
    j[0] = xp - xm;
    j[1] = yp - ym;
    j[2] = zp - zm;
}

static inline double feq_i(int i, double rhoBar, const double j[3], double jSqr)
{
    const double invRho = 1.0 / (rhoBar + 1.0);
    const double c_j = C[i][0] * j[0] + C[i][1] * j[1] + C[i][2] * j[2];
    return TW[i] * (rhoBar + 3. * c_j + invRho * (4.5 * c_j * c_j - 1.5 * jSqr));
}

static inline long wrapped(int nx, int ny, int nz, const int periodic[3], int x, int y, int z)
{
    int q[3] = {x, y, z};
    const int n[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d)
        if (periodic[d]) q[d] = ((q[d] % n[d]) + n[d]) % n[d];
    return ((long)q[0] * ny + q[1]) * nz + q[2];
}

/* The equilibrium part comes from the wall node's own dynamics: the
 * Guo-VOF BGK equilibrium on fluid cells, zero on no-dynamics and
 * bounce-back cells (src/core/dynamics.hh:773-777,946-950), as
 * cell.computeEquilibrium gives it in the reference. */
void or_guo_wall_apply(int nx, int ny, int nz, const int periodic[3], double *f, const unsigned char *mask,
                       long n, const long *cell, const int *dir, const double *delta, const double *vel)
{
    for (long k = 0; k < n; ++k) {
        const long c = cell[k];
        const int x = (int)(c / ((long)ny * nz)), y = (int)((c / nz) % ny), z = (int)(c % nz);
        const int *d = C[dir[k]];
        const double *f1 = f + 19 * wrapped(nx, ny, nz, periodic, x + d[0], y + d[1], z + d[2]);
        const double *f2 = f + 19 * wrapped(nx, ny, nz, periodic, x + 2 * d[0], y + 2 * d[1], z + 2 * d[2]);
        double j1[3], j2[3];
        const double rb1 = rho_bar(f1), rb2 = rho_bar(f2);
        get_j(f1, j1);
        get_j(f2, j2);
        const double j1sqr = j1[0] * j1[0] + j1[1] * j1[1] + j1[2] * j1[2];
        const double j2sqr = j2[0] * j2[0] + j2[1] * j2[1] + j2[2] * j2[2];
        double fneq1[19], fneq2[19];
        for (int i = 0; i < 19; ++i) {
            fneq1[i] = f1[i] - feq_i(i, rb1, j1, j1sqr);
            fneq2[i] = f2[i] - feq_i(i, rb2, j2, j2sqr);
        }
        const double rbw = rb1;
        const double *u = vel + 3 * k;
        const double wj[3] = {(rbw + 1.0) * u[0], (rbw + 1.0) * u[1], (rbw + 1.0) * u[2]};
        const double dl = delta[k];
        double jw[3], fneqw[19];
        if (dl < 0.75) {
            const double s = (1. - dl) / (1. + dl);
            for (int a = 0; a < 3; ++a) jw[a] = (wj[a] + (dl - 1.) * j1[a]) + s * (2. * wj[a] + (dl - 1.) * j2[a]);
            for (int i = 0; i < 19; ++i) fneqw[i] = dl * fneq1[i] + (1. - dl) * fneq2[i];
        } else {
            for (int a = 0; a < 3; ++a) jw[a] = 1. / dl * (wj[a] + (dl - 1.) * j1[a]);
            memcpy(fneqw, fneq1, sizeof(fneqw));
        }
        const double jsw = jw[0] * jw[0] + jw[1] * jw[1] + jw[2] * jw[2];
        double *fw = f + 19 * c;
        const int has_eq = !mask || mask[c] == 0 || mask[c] == 3;
        for (int i = 0; i < 19; ++i) fw[i] = (has_eq ? feq_i(i, rbw, jw, jsw) : 0.0) + fneqw[i];
    }
}

/* ------------------------------------------- near-wall immersed boundary */
/* blend weight of a vertex: distance to the boundary minus the kernel's
 * half support; `near` says whether it is a boundary node */
static double blend_of(int kernel, const or_boundary *b, const double *p, int *near)
{
    const double hs = or_dirac_half_support(kernel);
    *near = or_boundary_less_than(b, p, hs + 1.);
    return or_boundary_distance(b, p) - hs;
}

void or_interpolate_near(int kernel, int nx, int ny, int nz, const int periodic[3], const double *ufield,
                         long nv, const double *pos, const or_boundary *b, double *vel)
{
    for (long k = 0; k < nv; ++k) {
        const double *p = pos + 3 * k;
        int near = 0;
        const double w = b ? blend_of(kernel, b, p, &near) : 2.0;
        double ud[3], ut[3];
        if (!near || w >= 1) {
            or_interpolate(kernel, nx, ny, nz, periodic, ufield, 1, p, vel + 3 * k);
        } else if (w <= 0) {
            or_interpolate(OR_KERNEL_TOPHAT, nx, ny, nz, periodic, ufield, 1, p, vel + 3 * k);
        } else {
            or_interpolate(kernel, nx, ny, nz, periodic, ufield, 1, p, ud);
            or_interpolate(OR_KERNEL_TOPHAT, nx, ny, nz, periodic, ufield, 1, p, ut);
            for (int a = 0; a < 3; ++a) vel[3 * k + a] = (1. - w) * ut[a] + w * ud[a];
        }
    }
}

void or_spread_near(int kernel, int nx, int ny, int nz, const int periodic[3], double *ffield, long nv,
                    const double *pos, const double *fvert, const or_boundary *b)
{
    /* every bulk node first, in order; then the boundary nodes, in order */
    for (long k = 0; k < nv; ++k) {
        int near = 0;
        if (b) blend_of(kernel, b, pos + 3 * k, &near);
        if (!near) or_spread(kernel, nx, ny, nz, periodic, ffield, 1, pos + 3 * k, fvert + 3 * k);
    }
    if (!b) return;
    for (long k = 0; k < nv; ++k) {
        int near = 0;
        const double *p = pos + 3 * k, *F = fvert + 3 * k;
        const double w = blend_of(kernel, b, p, &near);
        if (!near) continue;
        if (w >= 1) {
            or_spread(kernel, nx, ny, nz, periodic, ffield, 1, p, F);
        } else if (w <= 0) {
            or_spread(OR_KERNEL_TOPHAT, nx, ny, nz, periodic, ffield, 1, p, F);
        } else {
            const double Fd[3] = {F[0] * w, F[1] * w, F[2] * w};
            const double Ft[3] = {F[0] * (1. - w), F[1] * (1. - w), F[2] * (1. - w)};
            or_spread(kernel, nx, ny, nz, periodic, ffield, 1, p, Fd);
            or_spread(OR_KERNEL_TOPHAT, nx, ny, nz, periodic, ffield, 1, p, Ft);
        }
    }
}

void or_move_reflect(long nv, double *pos, const double *vel, const or_boundary *b)
{
    for (long k = 0; k < nv; ++k) {
        double *p = pos + 3 * k;
        const double *v = vel + 3 * k;
        const double old[3] = {p[0], p[1], p[2]};
        for (int a = 0; a < 3; ++a) p[a] += v[a];
        double t;
        if (!b || !or_boundary_trace_ray(b, old, p, &t)) continue;
        for (int a = 0; a < 3; ++a) p[a] = old[a] + t * v[a];
        double n[3];
        or_boundary_normal(b, p, n);
        const double vn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
        for (int a = 0; a < 3; ++a) p[a] += (v[a] - 2. * vn * n[a]) * (1. - t);
    }
}
