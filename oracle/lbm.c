/*
 * oracle/lbm.c -- FP64 CPU restatement of the reference's D3Q19 step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).  Never linked into
 * the product library.
 *
 * Follows, operation by operation:
 *   D3Q19 vectors / weights      src/latticeBoltzmann/nearestNeighborLattices3D.hh:165-199
 *   fBar = f - t_i               src/latticeBoltzmann/roundOffPolicy.h:74-90
 *   rhoBar, j                    src/latticeBoltzmann/momentTemplates3D.h:375-399
 *   u = j/rho + F/2              src/basicDynamics/externalForceDynamics.hh:59-73
 *   omega(phase)                 src/fsi/VariableViscosityDynamics.hh:131-141
 *   BGK (ma2, fBar form)         src/latticeBoltzmann/dynamicsTemplates3D.h:759-838
 *   Guo forcing (VOF copy)       src/fsi/VariableViscosityDynamics.hh:47-129
 *   collide driver               src/fsi/VariableViscosityDynamics.hh:151-172
 *   bounce-back                  src/core/dynamics.hh:763-770
 *   streaming f_i(x+c_i) = f*_i  src/atomicBlock/blockLattice3D.hh:253-311,502-552
 *   periodic envelope            src/multiBlock/multiBlock3D.cpp:386
 *   velocity-wall density rule   src/boundaryCondition/boundaryDynamics.hh:631-690
 *   regularized FD wall          src/boundaryCondition/finiteDifferenceBoundaryProcessor3D.hh:68-115
 *   o1 / central FD stencils     src/finiteDifference/finiteDifference3D.h:67-140
 *   Pi -> fNeq                   src/latticeBoltzmann/offEquilibriumTemplates.h:91-115
 *
 * Domain faces on non-periodic axes reflect: a population that would leave
 * the box re-enters the same cell in the opposite direction.  Palabos leaves
 * an unfilled envelope there instead; both choices only touch populations
 * that wall cells (bounce-back or velocity-wall) discard or send back out, so
 * fluid cells are identical.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>

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

void or_d3q19(int c[19][3], double t[19])
{
    memcpy(c, C, sizeof(C));
    memcpy(t, TW, sizeof(TW));
}

static inline int opp(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }

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

static void bgk_ma2_collision_base(double *f, double rhoBar, const double j[3],
                                   double omega, double invRho)
{
    double one_m_omega = 1.0 - omega;
    double t0_omega = T0 * omega, t1_omega = T1 * omega, t4_omega = T4 * omega;
    double jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
    double kx = 3.0 * j[0], ky = 3.0 * j[1], kz = 3.0 * j[2];
    double kxSqr_ = invRho / 2.0 * kx * kx;
    double kySqr_ = invRho / 2.0 * ky * ky;
    double kzSqr_ = invRho / 2.0 * kz * kz;
    double kxky_ = invRho * kx * ky;
    double kxkz_ = invRho * kx * kz;
    double kykz_ = invRho * ky * kz;
    double C1 = rhoBar + invRho * 3.0 * jSqr;
    double C2, C3;

    C3 = -kxSqr_ - kySqr_ - kzSqr_;
    f[0] *= one_m_omega; f[0] += t0_omega * (C1 + C3);

#define PAIR(a, b, tw, c2, c3)                                   \
    C2 = (c2); C3 = (c3);                                        \
    f[a] *= one_m_omega; f[a] += tw * (C1 + C2 + C3);            \
    f[b] *= one_m_omega; f[b] += tw * (C1 - C2 + C3);
    PAIR(1, 10, t1_omega, -kx, -kySqr_ - kzSqr_)
    PAIR(2, 11, t1_omega, -ky, -kxSqr_ - kzSqr_)
    PAIR(3, 12, t1_omega, -kz, -kxSqr_ - kySqr_)
    PAIR(4, 13, t4_omega, -kx - ky, kxky_ - kzSqr_)
    PAIR(5, 14, t4_omega, -kx + ky, -kxky_ - kzSqr_)
    PAIR(6, 15, t4_omega, -kx - kz, kxkz_ - kySqr_)
    PAIR(7, 16, t4_omega, -kx + kz, -kxkz_ - kySqr_)
    PAIR(8, 17, t4_omega, -ky - kz, kykz_ - kxSqr_)
    PAIR(9, 18, t4_omega, -ky + kz, -kykz_ - kxSqr_)
#undef PAIR
}

static void add_guo_force(double *f, const double *F, const double u[3], double omega)
{
    const double mu = 1.0 * (1.0 - omega / 2.0);
    const double o6 = 1.0 / 6.0, o12 = 1.0 / 12.0;
    f[0] += -mu * (F[0] * u[0] + F[1] * u[1] + F[2] * u[2]);
    f[1] += o6 * mu * (F[0] * (-1.0 + 2 * u[0]) - F[1] * u[1] - F[2] * u[2]);
    f[2] += -o6 * mu * (F[0] * u[0] + F[1] * (1.0 - 2 * u[1]) + F[2] * u[2]);
    f[3] += -o6 * mu * (F[0] * u[0] + F[1] * u[1] + F[2] * (1.0 - 2 * u[2]));
    f[4] += o12 * mu * (F[0] * (-1.0 + 2 * u[0] + 3 * u[1]) + F[1] * (-1.0 + 2 * u[1] + 3 * u[0]) - F[2] * u[2]);
    f[5] += o12 * mu * (F[0] * (-1.0 + 2 * u[0] - 3 * u[1]) + F[1] * (1.0 + 2 * u[1] - 3 * u[0]) - F[2] * u[2]);
    f[6] += o12 * mu * (F[0] * (-1.0 + 2 * u[0] + 3 * u[2]) - F[1] * u[1] + F[2] * (-1.0 + 2 * u[2] + 3 * u[0]));
    f[7] += o12 * mu * (F[0] * (-1.0 + 2 * u[0] - 3 * u[2]) - F[1] * u[1] + F[2] * (1.0 + 2 * u[2] - 3 * u[0]));
    f[8] += -o12 * mu * (F[0] * u[0] + F[1] * (1.0 - 2 * u[1] - 3 * u[2]) + F[2] * (1.0 - 2 * u[2] - 3 * u[1]));
    f[9] += -o12 * mu * (F[0] * u[0] + F[1] * (1.0 - 2 * u[1] + 3 * u[2]) + F[2] * (-1.0 - 2 * u[2] + 3 * u[1]));
    f[10] += o6 * mu * (F[0] * (1.0 + 2 * u[0]) - F[1] * u[1] - F[2] * u[2]);
    f[11] += -o6 * mu * (F[0] * u[0] + F[1] * (-1.0 - 2 * u[1]) + F[2] * u[2]);
    f[12] += -o6 * mu * (F[0] * u[0] + F[1] * u[1] + F[2] * (-1.0 - 2 * u[2]));
    f[13] += o12 * mu * (F[0] * (1.0 + 2 * u[0] + 3 * u[1]) + F[1] * (1.0 + 2 * u[1] + 3 * u[0]) - F[2] * u[2]);
    f[14] += o12 * mu * (F[0] * (1.0 + 2 * u[0] - 3 * u[1]) + F[1] * (-1.0 + 2 * u[1] - 3 * u[0]) - F[2] * u[2]);
    f[15] += o12 * mu * (F[0] * (1.0 + 2 * u[0] + 3 * u[2]) - F[1] * u[1] + F[2] * (1.0 + 2 * u[2] + 3 * u[0]));
    f[16] += o12 * mu * (F[0] * (1.0 + 2 * u[0] - 3 * u[2]) - F[1] * u[1] + F[2] * (-1.0 + 2 * u[2] - 3 * u[0]));
    f[17] += -o12 * mu * (F[0] * u[0] + F[1] * (-1.0 - 2 * u[1] - 3 * u[2]) + F[2] * (-1.0 - 2 * u[2] - 3 * u[1]));
    f[18] += -o12 * mu * (F[0] * u[0] + F[1] * (-1.0 - 2 * u[1] + 3 * u[2]) + F[2] * (1.0 - 2 * u[2] + 3 * u[1]));
}

void or_collide_cell(double f[19], const double F[3], double phase,
                     double omega_out, double omega_in)
{
    double rhoBar = rho_bar(f);
    double j[3], u[3], jj[3];
    get_j(f, j);
    double invRho = 1.0 / (rhoBar + 1.0);
    for (int d = 0; d < 3; ++d) u[d] = j[d] * invRho + F[d] / 2.0;
    double rho = rhoBar + 1.0;
    for (int d = 0; d < 3; ++d) jj[d] = rho * u[d];
    double omega = omega_out * omega_in / (phase * omega_out + (1.0 - phase) * omega_in);
    bgk_ma2_collision_base(f, rhoBar, jj, omega, 1.0 / (rhoBar + 1.0));
    add_guo_force(f, F, u, omega);
}

static inline double feq_i(int i, double rhoBar, double invRho, const double j[3], double jSqr)
{
    double c_j = C[i][0] * j[0] + C[i][1] * j[1] + C[i][2] * j[2];
    return TW[i] * (rhoBar + 3. * c_j + invRho * (4.5 * c_j * c_j - 1.5 * jSqr));
}

/* Pi in Palabos symmetric order xx, xy, xz, yy, yz, zz. */
static inline double pi_to_fneq(int i, const double pi[6])
{
    double fNeq = 0.0;
    int iPi = 0;
    for (int a = 0; a < 3; ++a) {
        fNeq += pi[iPi] * (C[i][a] * C[i][a] - 1.0 / 3.0);
        ++iPi;
        for (int b = a + 1; b < 3; ++b) {
            fNeq += pi[iPi] * 2.0 * C[i][a] * C[i][b];
            ++iPi;
        }
    }
    fNeq *= TW[i] * 3.0 * 3.0 / 2.0;
    return fNeq;
}

static inline long idx(int nx, int ny, int nz, int x, int y, int z)
{
    (void)nx;
    return ((long)x * ny + y) * nz + z;
}

static inline int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

/* Velocity as Cell::computeVelocity reports it: stored wall velocity on
 * velocity-wall cells, zero on bounce-back and no-dynamics cells
 * (src/core/dynamics.hh:796-801,970-975). */
static void cell_velocity(const double *f, const double *F, unsigned char m,
                          const double *uw, double u[3])
{
    if (m == 3) {
        u[0] = uw[0]; u[1] = uw[1]; u[2] = uw[2];
        return;
    }
    if (m == 1 || m == 2) {
        u[0] = u[1] = u[2] = 0.0;
        return;
    }
    double rb = rho_bar(f), j[3];
    get_j(f, j);
    double invRho = 1.0 / (rb + 1.0);
    for (int d = 0; d < 3; ++d) u[d] = j[d] * invRho + F[d] / 2.0;
}

static double wall_rho_bar(const double *f, const double *F, int axis, int orient, const double *uw)
{
    double onWall = 0.0, normal = 0.0;
    for (int i = 0; i < 19; ++i) {
        if (C[i][axis] == 0) onWall += f[i];
    }
    for (int i = 0; i < 19; ++i) {
        if (C[i][axis] == orient) normal += f[i];
    }
    double velNormal = (double)orient * (uw[axis] - 0.5 * F[axis]);
    return (2.0 * normal + onWall - 1.0 * velNormal) / (1.0 + velNormal);
}

void or_velocity_walls(int nx, int ny, int nz, const int periodic[3],
                       double omega_out, double *f, const double *force,
                       const unsigned char *mask, int wall_axis,
                       const double *wall_u)
{
    const int n[3] = {nx, ny, nz};
    const long ncell = (long)nx * ny * nz;
    int any = 0;
    for (long c = 0; c < ncell && !any; ++c) any = mask && mask[c] == 3;
    if (!any) return;
    double *out = (double *)malloc(sizeof(double) * 19 * ncell);
    unsigned char *touched = (unsigned char *)calloc(ncell, 1);
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                long c = idx(nx, ny, nz, x, y, z);
                if (mask[c] != 3) continue;
                int p[3] = {x, y, z};
                int side = p[wall_axis] == 0 ? 0 : 1;
                int orient = side == 0 ? -1 : 1;
                const double *uw = wall_u + 3 * side;
                const double *fc = f + 19 * c;
                double rb = wall_rho_bar(fc, force + 3 * c, wall_axis, orient, uw);
                double rho = rb + 1.0;
                double du[3][3]; /* du[deriv][comp] */
                for (int d = 0; d < 3; ++d) {
                    if (d == wall_axis) {
                        int q[3] = {x, y, z};
                        q[d] -= orient;
                        long c1 = idx(nx, ny, nz, q[0], q[1], q[2]);
                        double u1[3];
                        cell_velocity(f + 19 * c1, force + 3 * c1, mask[c1],
                                      wall_u + 3 * (q[d] == 0 ? 0 : 1), u1);
                        for (int a = 0; a < 3; ++a) du[d][a] = -orient * (-uw[a] + u1[a]);
                    } else {
                        int qp[3] = {x, y, z}, qm[3] = {x, y, z};
                        qp[d] += 1; qm[d] -= 1;
                        if (periodic[d]) { qp[d] = wrap(qp[d], n[d]); qm[d] = wrap(qm[d], n[d]); }
                        else {
                            if (qp[d] >= n[d]) qp[d] = n[d] - 1;
                            if (qm[d] < 0) qm[d] = 0;
                        }
                        long cp = idx(nx, ny, nz, qp[0], qp[1], qp[2]);
                        long cm = idx(nx, ny, nz, qm[0], qm[1], qm[2]);
                        double up[3], um[3];
                        cell_velocity(f + 19 * cp, force + 3 * cp, mask[cp], uw, up);
                        cell_velocity(f + 19 * cm, force + 3 * cm, mask[cm], uw, um);
                        for (int a = 0; a < 3; ++a) du[d][a] = (up[a] - um[a]) / 2.0;
                    }
                }
                double sToPi = -rho / 3.0 / omega_out;
                double pi[6];
                pi[0] = 2.0 * du[0][0] * sToPi;
                pi[3] = 2.0 * du[1][1] * sToPi;
                pi[5] = 2.0 * du[2][2] * sToPi;
                pi[1] = (du[0][1] + du[1][0]) * sToPi;
                pi[2] = (du[0][2] + du[2][0]) * sToPi;
                pi[4] = (du[1][2] + du[2][1]) * sToPi;
                double j[3] = {rho * uw[0], rho * uw[1], rho * uw[2]};
                double jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
                double invRho = 1.0 / (rb + 1.0);
                for (int i = 0; i < 19; ++i)
                    out[19 * c + i] = feq_i(i, rb, invRho, j, jSqr) + pi_to_fneq(i, pi);
                touched[c] = 1;
            }
    for (long c = 0; c < ncell; ++c)
        if (touched[c]) memcpy(f + 19 * c, out + 19 * c, sizeof(double) * 19);
    free(out);
    free(touched);
}

void or_lbm_step(int nx, int ny, int nz, const int periodic[3],
                 double omega_out, double omega_in,
                 double *f, double *ftmp, const double *force,
                 const double *phase, const unsigned char *mask,
                 int wall_axis, const double *wall_u)
{
    const long ncell = (long)nx * ny * nz;
    const int n[3] = {nx, ny, nz};
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                long c = idx(nx, ny, nz, x, y, z);
                double g[19];
                memcpy(g, f + 19 * c, sizeof(g));
                unsigned char m = mask ? mask[c] : 0;
                if (m == 0 || m == 3) {
                    or_collide_cell(g, force + 3 * c, phase ? phase[c] : 0.0, omega_out, omega_in);
                } else if (m == 1) {
                    for (int i = 1; i <= 9; ++i) {
                        double tmp = g[i];
                        g[i] = g[i + 9];
                        g[i + 9] = tmp;
                    }
                }
                /* push streaming; reflect at non-periodic faces */
                for (int i = 0; i < 19; ++i) {
                    int q[3] = {x + C[i][0], y + C[i][1], z + C[i][2]};
                    int out = 0;
                    for (int d = 0; d < 3; ++d) {
                        if (q[d] < 0 || q[d] >= n[d]) {
                            if (periodic[d]) q[d] = wrap(q[d], n[d]);
                            else out = 1;
                        }
                    }
                    if (out) ftmp[19 * c + opp(i)] = g[i];
                    else ftmp[19 * idx(nx, ny, nz, q[0], q[1], q[2]) + i] = g[i];
                }
            }
    memcpy(f, ftmp, sizeof(double) * 19 * ncell);
    or_velocity_walls(nx, ny, nz, periodic, omega_out, f, force, mask, wall_axis, wall_u);
}

void or_lbm_moments(int nx, int ny, int nz, const double *f,
                    const double *force, const unsigned char *mask,
                    int wall_axis, const double *wall_u,
                    double *rho, double *u)
{
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                long c = idx(nx, ny, nz, x, y, z);
                unsigned char m = mask ? mask[c] : 0;
                if (m == 3) {
                    int p[3] = {x, y, z};
                    int side = p[wall_axis] == 0 ? 0 : 1;
                    const double *uw = wall_u + 3 * side;
                    rho[c] = wall_rho_bar(f + 19 * c, force + 3 * c, wall_axis, side ? 1 : -1, uw) + 1.0;
                    for (int d = 0; d < 3; ++d) u[3 * c + d] = uw[d];
                } else if (m == 1 || m == 2) {
                    /* default-constructed BounceBack / NoDynamics report rho = 0
                     * (src/core/dynamics.h:564,673, src/core/dynamics.hh:786-788) */
                    rho[c] = 0.0;
                    u[3 * c] = u[3 * c + 1] = u[3 * c + 2] = 0.0;
                } else {
                    rho[c] = rho_bar(f + 19 * c) + 1.0;
                    cell_velocity(f + 19 * c, force + 3 * c, 0, 0, u + 3 * c);
                }
            }
}

void or_lbm_init_equilibrium(long ncell, const double *rho, const double *u, double *f)
{
    for (long c = 0; c < ncell; ++c) {
        double r = rho[c];
        double j[3] = {r * u[3 * c], r * u[3 * c + 1], r * u[3 * c + 2]};
        double jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
        for (int i = 0; i < 19; ++i) f[19 * c + i] = feq_i(i, r - 1.0, 1.0 / r, j, jSqr);
    }
}

void or_lbm_conserved(long ncell, const double *f, const unsigned char *mask, double out[4])
{
    double m = 0, jx = 0, jy = 0, jz = 0;
    for (long c = 0; c < ncell; ++c) {
        if (mask && mask[c] == 1) continue;
        const double *g = f + 19 * c;
        double j[3];
        get_j(g, j);
        m += 1.0 + rho_bar(g);
        jx += j[0]; jy += j[1]; jz += j[2];
    }
    out[0] = m; out[1] = jx; out[2] = jy; out[3] = jz;
}
