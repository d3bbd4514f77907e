/*
 * CPU restatement (FP64) of the rigid bodies of the growth packing.
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).
 *
 *   or_shape_inertia   ParticleShape::compute_properties, the rest-shape
 *                      moment of inertia (src/fsi/ParticleShape.hh:249-360;
 *                      about the mesh origin, as there)
 *   or_rigid_update    RigidParticle3D::update (src/fsi/RigidParticle.hh:30-39)
 *                      with Quaternion::to_rot_matrix (src/fsi/Quaternion.h)
 *   or_rigid_move      RigidParticle3D::move_vertices (src/fsi/RigidParticle.hh:57-98)
 *                      then map_center_of_mass_to_periodic_grid
 *                      (src/fsi/ParticleBase.hh:110-136, src/fsi/Periodicity.h:83-87)
 *
 * Rigid state, 16 doubles: com (3), velocity (3), angular velocity (3, world
 * frame), orientation quaternion (4: w, x, y, z), scale, density, damping.
 * Expressions keep the reference's operand order.
 */
#include <math.h>
#include <float.h>

#include "oracle.h"

static void cross3(const double a[3], const double b[3], double r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

static void rot_matrix(const double q[4], double m[9])
{
    m[0] = 1 - 2 * (q[2] * q[2] + q[3] * q[3]);
    m[1] = 2 * (q[1] * q[2] - q[0] * q[3]);
    m[2] = 2 * (q[0] * q[2] + q[1] * q[3]);
    m[3] = 2 * (q[1] * q[2] + q[0] * q[3]);
    m[4] = 1 - 2 * (q[1] * q[1] + q[3] * q[3]);
    m[5] = 2 * (q[2] * q[3] - q[0] * q[1]);
    m[6] = 2 * (q[1] * q[3] - q[0] * q[2]);
    m[7] = 2 * (q[0] * q[1] + q[2] * q[3]);
    m[8] = 1 - 2 * (q[1] * q[1] + q[2] * q[2]);
}

static void apply_rotation(const double q[4], const double v[3], double r[3])
{
    r[0] = (1 - 2 * (q[2] * q[2] + q[3] * q[3])) * v[0] + 2 * (q[1] * q[2] - q[0] * q[3]) * v[1] +
           2 * (q[0] * q[2] + q[1] * q[3]) * v[2];
    r[1] = 2 * (q[1] * q[2] + q[0] * q[3]) * v[0] + (1 - 2 * (q[1] * q[1] + q[3] * q[3])) * v[1] +
           2 * (q[2] * q[3] - q[0] * q[1]) * v[2];
    r[2] = 2 * (q[1] * q[3] - q[0] * q[2]) * v[0] + 2 * (q[0] * q[1] + q[2] * q[3]) * v[1] +
           (1 - 2 * (q[1] * q[1] + q[2] * q[2])) * v[2];
}

static void apply_inv_rotation(const double q[4], const double v[3], double r[3])
{
    r[0] = (1 - 2 * (q[2] * q[2] + q[3] * q[3])) * v[0] + 2 * (q[1] * q[2] + q[0] * q[3]) * v[1] +
           2 * (q[1] * q[3] - q[0] * q[2]) * v[2];
    r[1] = 2 * (q[1] * q[2] - q[0] * q[3]) * v[0] + (1 - 2 * (q[1] * q[1] + q[3] * q[3])) * v[1] +
           2 * (q[0] * q[1] + q[2] * q[3]) * v[2];
    r[2] = 2 * (q[0] * q[2] + q[1] * q[3]) * v[0] + 2 * (q[2] * q[3] - q[0] * q[1]) * v[1] +
           (1 - 2 * (q[1] * q[1] + q[2] * q[2])) * v[2];
}

static void mat_vec(const double m[9], const double v[3], double r[3])
{
    r[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    r[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
    r[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

/* Matrix<T,3>::solve (src/fsi/Matrix3.h:129-144): the adjugate over the
 * determinant */
static void mat_solve(const double m[9], const double b[3], double r[3])
{
    const double det = m[0] * m[4] * m[8] - m[0] * m[5] * m[7] - m[1] * m[3] * m[8] + m[1] * m[5] * m[6] +
                       m[2] * m[3] * m[7] - m[2] * m[4] * m[6];
    r[0] = (b[2] * (m[1] * m[5] - m[2] * m[4]) + b[1] * (m[2] * m[7] - m[1] * m[8]) + b[0] * (m[4] * m[8] - m[5] * m[7])) / det;
    r[1] = (b[2] * (m[2] * m[3] - m[0] * m[5]) + b[1] * (m[0] * m[8] - m[2] * m[6]) + b[0] * (m[5] * m[6] - m[3] * m[8])) / det;
    r[2] = (b[2] * (m[0] * m[4] - m[1] * m[3]) + b[1] * (m[1] * m[6] - m[0] * m[7]) + b[0] * (m[3] * m[7] - m[4] * m[6])) / det;
}

void or_shape_inertia(long nv, long nt, const int *tri, const double *rest, double I[9])
{
    double in[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    (void)nv;
    for (long t = 0; t < nt; ++t) {
        const double *v0 = rest + 3 * tri[3 * t], *v1 = rest + 3 * tri[3 * t + 1], *v2 = rest + 3 * tri[3 * t + 2];
        const double a[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const double b[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        double N[3];
        cross3(a, b, N);
        double f1[3], f2[3], f3[3], g0[3], g1[3], g2[3];
        for (int d = 0; d < 3; ++d) {
            const double tmp0 = v0[d] + v1[d];
            f1[d] = tmp0 + v2[d];
            const double tmp1 = v0[d] * v0[d];
            const double tmp2 = tmp1 + v1[d] * tmp0;
            f2[d] = tmp2 + v2[d] * f1[d];
            f3[d] = v0[d] * tmp1 + v1[d] * tmp2 + v2[d] * f2[d];
            g0[d] = f2[d] + v0[d] * (f1[d] + v0[d]);
            g1[d] = f2[d] + v1[d] * (f1[d] + v1[d]);
            g2[d] = f2[d] + v2[d] * (f1[d] + v2[d]);
        }
        in[0] += N[0] * f1[0];
        in[1] += N[0] * f2[0];
        in[2] += N[1] * f2[1];
        in[3] += N[2] * f2[2];
        in[4] += N[0] * f3[0];
        in[5] += N[1] * f3[1];
        in[6] += N[2] * f3[2];
        in[7] += N[0] * (v0[1] * g0[0] + v1[1] * g1[0] + v2[1] * g2[0]);
        in[8] += N[1] * (v0[2] * g0[1] + v1[2] * g1[1] + v2[2] * g2[1]);
        in[9] += N[2] * (v0[0] * g0[2] + v1[0] * g1[2] + v2[0] * g2[2]);
    }
    for (int k = 4; k <= 6; ++k) in[k] *= (double)(1.0 / 60.0);
    for (int k = 7; k <= 9; ++k) in[k] *= (double)(1.0 / 120.0);
    I[0] = in[5] + in[6];
    I[1] = -in[7];
    I[2] = -in[9];
    I[3] = I[1];
    I[4] = in[4] + in[6];
    I[5] = -in[8];
    I[6] = I[2];
    I[7] = I[5];
    I[8] = in[4] + in[5];
}

void or_rigid_update(long nv, const double *rest, const double *state, double *pos)
{
    double R[9];
    rot_matrix(state + 9, R);
    const double s = state[13];
    for (long i = 0; i < nv; ++i) {
        const double v[3] = {s * rest[3 * i], s * rest[3 * i + 1], s * rest[3 * i + 2]};
        double r[3];
        mat_vec(R, v, r);
        for (int d = 0; d < 3; ++d) pos[3 * i + d] = state[d] + r[d];
    }
}

void or_rigid_move(long nv, const double *rest, const double I[9], double volume, const int periodic[3],
                   const int n[3], double *state, double *pos, const double *force)
{
    double *com = state, *vel = state + 3, *w = state + 6, *q = state + 9;
    const double scale = state[13], density = state[14], damping = state[15];
    double F[3] = {0, 0, 0}, tq[3] = {0, 0, 0};
    for (long i = 0; i < nv; ++i) {
        const double *f = force + 3 * i;
        const double r[3] = {pos[3 * i] - com[0], pos[3 * i + 1] - com[1], pos[3 * i + 2] - com[2]};
        double c[3];
        cross3(r, f, c);
        for (int d = 0; d < 3; ++d) {
            F[d] += f[d];
            tq[d] += c[d];
        }
    }
    double wb[3], Ib[9], Iw[3], tb[3], cw[3], rhs[3], sol[3], wd[3];
    apply_inv_rotation(q, w, wb);
    const double fac = density * pow(scale, 5);
    for (int k = 0; k < 9; ++k) Ib[k] = fac * I[k];
    const double mass = (scale * scale * scale) * volume * density;
    double vdot[3];
    for (int d = 0; d < 3; ++d) vdot[d] = F[d] / mass - damping * vel[d];
    apply_inv_rotation(q, tq, tb);
    mat_vec(Ib, wb, Iw);
    cross3(wb, Iw, cw);
    for (int d = 0; d < 3; ++d) rhs[d] = tb[d] - cw[d];
    mat_solve(Ib, rhs, sol);
    apply_rotation(q, sol, wd);
    for (int d = 0; d < 3; ++d) {
        wd[d] = wd[d] - damping * w[d];
        vel[d] += vdot[d];
        w[d] += wd[d];
        com[d] += vel[d];
    }
    const double om = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (om >= (double)10 * DBL_EPSILON) {
        const double ax[3] = {w[0] / om, w[1] / om, w[2] / om};
        const double sh = sin(om / 2);
        const double p[4] = {cos(om / 2), sh * ax[0], sh * ax[1], sh * ax[2]};
        const double o[4] = {q[0], q[1], q[2], q[3]};
        q[0] = p[0] * o[0] - (p[1] * o[1] + p[2] * o[2] + p[3] * o[3]);
        q[1] = p[0] * o[1] + o[0] * p[1] + p[2] * o[3] - p[3] * o[2];
        q[2] = p[0] * o[2] + o[0] * p[2] + p[3] * o[1] - p[1] * o[3];
        q[3] = p[0] * o[3] + o[0] * p[3] + p[1] * o[2] - p[2] * o[1];
    }
    or_rigid_update(nv, rest, state, pos);
    /* centre of mass onto the periodic grid [-1/2, n - 1/2), the vertices
     * shifted with it */
    for (int d = 0; d < 3; ++d) {
        if (!periodic[d]) continue;
        const double L = (double)n[d], x0 = -0.5, Linv = 1. / L;
        const double c = com[d] - L * floor((com[d] - x0) * Linv);
        const double delta = c - com[d];
        for (long i = 0; i < nv; ++i) pos[3 * i + d] += delta;
        com[d] = c;
    }
}
