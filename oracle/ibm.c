/*
 * oracle/ibm.c -- FP64 CPU restatement of the reference's immersed-boundary
 * coupling (fsi module).
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).  Never linked into
 * the product library.
 *
 * Follows:
 *   Dirac kernels (Roma / Peskin / top hat)   src/fsi/Dirac.h:23-101
 *   support box  x0 = ceil(x - h)             src/fsi/Dirac.h:131-139
 *   periodic index remap                      src/fsi/Periodicity.h:110-115
 *   interpolation (general interpolator)      src/fsi/ImmersedBoundaryDynamics_interpolation.hh:14-78
 *   spreading (general spreading)             src/fsi/ImmersedBoundaryDynamics_forces.hh:22-85
 *   capsule Skalak / NeoHookean forces        src/fsi/DeformableCapsule.hh:60-150
 *   RBC WLC-POW / area / volume / bending     src/fsi/RBCParticle.hh:27-44,140-247
 *   triangle helpers                          src/fsi/TriangleUtils.h:28-164
 *   links, adjacent pairs, rest angles        src/fsi/ParticleShape.hh:21-212
 *   area / volume / COM / bounding box        src/fsi/DeformableParticle3D.hh:54-82, src/fsi/ParticleBase.hh:46-69
 *   moment of inertia                         src/fsi/DeformableParticle3D.hh:173-272
 *   ray / triangle z-hits                     src/fsi/RayTracer.hh:19-39,80-108
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* ---------------------------------------------------------------- Dirac */
double or_dirac_1d(int kernel, double x)
{
    double dx = fabs(x);
    switch (kernel) {
    case 0:
        return (1.0 / 6.0) * (dx > 0.5 ? (5.0 - 3 * dx - sqrt(-3 * (1 - dx) * (1 - dx) + 1))
                                       : 2.0 * (1 + sqrt(1 - 3 * dx * dx)));
    case 1:
        return 0.125 * (dx <= 1 ? (3 - 2 * dx + sqrt(1 + 4 * dx * (1 - dx)))
                                : (5 - 2 * dx - sqrt(-7 + 4 * dx * (3 - dx))));
    default:
        return 1. - fabs(x);
    }
}

int or_dirac_support(int kernel) { return kernel == 0 ? 3 : (kernel == 1 ? 4 : 2); }
double or_dirac_half_support(int kernel) { return kernel == 0 ? 1.5 : (kernel == 1 ? 2.0 : 1.0); }
long or_support_base(int kernel, double x) { return (long)ceil(x - or_dirac_half_support(kernel)); }

static inline long remap_periodic(long i, long n)
{
    long v = ((i + 100 * n) % n);
    return v < 0 ? v + n : v;
}

void or_interpolate(int kernel, int nx, int ny, int nz, const int periodic[3],
                    const double *ufield, long nv, const double *pos, double *vel)
{
    const long n[3] = {nx, ny, nz};
    const int S = or_dirac_support(kernel);
    for (long k = 0; k < nv; ++k) {
        const double *p = pos + 3 * k;
        long b[3];
        for (int d = 0; d < 3; ++d) b[d] = or_support_base(kernel, p[d]);
        double v[3] = {0, 0, 0};
        for (long i = b[0]; i < b[0] + S; ++i) {
            long i2 = periodic[0] ? remap_periodic(i, n[0]) : i;
            if (i2 < 0 || i2 >= n[0]) continue;
            double wx = or_dirac_1d(kernel, (double)i - p[0]);
            for (long j = b[1]; j < b[1] + S; ++j) {
                long j2 = periodic[1] ? remap_periodic(j, n[1]) : j;
                if (j2 < 0 || j2 >= n[1]) continue;
                double wxy = wx * or_dirac_1d(kernel, (double)j - p[1]);
                for (long l = b[2]; l < b[2] + S; ++l) {
                    long l2 = periodic[2] ? remap_periodic(l, n[2]) : l;
                    if (l2 < 0 || l2 >= n[2]) continue;
                    double w = wxy * or_dirac_1d(kernel, (double)l - p[2]);
                    const double *u = ufield + 3 * ((i2 * n[1] + j2) * n[2] + l2);
                    v[0] += w * u[0];
                    v[1] += w * u[1];
                    v[2] += w * u[2];
                }
            }
        }
        vel[3 * k] = v[0];
        vel[3 * k + 1] = v[1];
        vel[3 * k + 2] = v[2];
    }
}

void or_spread(int kernel, int nx, int ny, int nz, const int periodic[3],
               double *ffield, long nv, const double *pos, const double *fvert)
{
    const long n[3] = {nx, ny, nz};
    const int S = or_dirac_support(kernel);
    for (long k = 0; k < nv; ++k) {
        const double *p = pos + 3 * k;
        const double *F = fvert + 3 * k;
        long b[3];
        for (int d = 0; d < 3; ++d) b[d] = or_support_base(kernel, p[d]);
        for (long i = b[0]; i < b[0] + S; ++i) {
            long i2 = periodic[0] ? remap_periodic(i, n[0]) : i;
            if (i2 < 0 || i2 >= n[0]) continue;
            double wx = or_dirac_1d(kernel, (double)i - p[0]);
            for (long j = b[1]; j < b[1] + S; ++j) {
                long j2 = periodic[1] ? remap_periodic(j, n[1]) : j;
                if (j2 < 0 || j2 >= n[1]) continue;
                double wxy = wx * or_dirac_1d(kernel, (double)j - p[1]);
                for (long l = b[2]; l < b[2] + S; ++l) {
                    long l2 = periodic[2] ? remap_periodic(l, n[2]) : l;
                    if (l2 < 0 || l2 >= n[2]) continue;
                    double w = wxy * or_dirac_1d(kernel, (double)l - p[2]);
                    double *f = ffield + 3 * ((i2 * n[1] + j2) * n[2] + l2);
                    f[0] += w * F[0];
                    f[1] += w * F[1];
                    f[2] += w * F[2];
                }
            }
        }
    }
}

/* -------------------------------------------------------- vector helpers */
static inline void sub3(const double *a, const double *b, double *r)
{
    r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
static inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double *a, const double *b, double *r)
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double norm3(const double *a) { return sqrt(dot3(a, a)); }

/* ----------------------------------------------------- capsule membrane */
void or_capsule_forces(long nv, long nt, const int *tri,
                       const double *rest_pos, const double *pos,
                       double G, double Cc, int shear_model, int bug_compatible,
                       double *force)
{
    (void)nv;
    for (long t = 0; t < nt; ++t) {
        const int ids[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
        const double *p0 = pos + 3 * ids[0], *p1 = pos + 3 * ids[1], *p2 = pos + 3 * ids[2];
        const double *r0 = rest_pos + 3 * ids[0], *r1 = rest_pos + 3 * ids[1], *r2 = rest_pos + 3 * ids[2];
        double e0[3], e1[3], e0_0[3], e1_0[3], cr[3];
        sub3(p1, p0, e0);
        sub3(p2, p0, e1);
        sub3(r1, r0, e0_0);
        sub3(r2, r0, e1_0);
        /* rest triangle area, as ParticleShape computes it */
        cross3(e0_0, e1_0, cr);
        const double area0 = 0.5 * norm3(cr);

        const double l = norm3(e0), lp = norm3(e1);
        const double l0 = norm3(e0_0), l0p = norm3(e1_0);
        const double e0e1 = dot3(e0, e1);
        cross3(e0, e1, cr);
        const double sin_phi = norm3(cr) / (l * lp);
        const double cos_phi = e0e1 / (l * lp);
        cross3(e0_0, e1_0, cr);
        const double sin_phi0 = norm3(cr) / (l0 * l0p);
        const double cos_phi0 = dot3(e0_0, e1_0) / (l0 * l0p);

        const double a = l / l0;
        const double b = (lp * cos_phi / l0p - l * cos_phi0 / l0) / sin_phi0;
        const double c = bug_compatible ? lp * sin_phi / (lp * sin_phi0)
                                        : lp * sin_phi / (l0p * sin_phi0);
        const double D11 = (a * a - 1.) / 2.;
        const double D22 = (b * b + c * c - 1) / 2.;
        const double D12 = a * b / 2.;
        const double I1 = 2. * (D11 + D22);
        const double I2 = I1 + 4 * (D11 * D22 - D12 * D12);

        double gl[3][3], glp[3][3], gc[3][3];
        for (int d = 0; d < 3; ++d) {
            gl[0][d] = e0[d] / (-l);
            gl[1][d] = e0[d] / l;
            gl[2][d] = 0.0;
            glp[0][d] = e1[d] / (-lp);
            glp[1][d] = 0.0;
            glp[2][d] = e1[d] / lp;
            gc[0][d] = (-1. / (l * lp) + e0e1 / (l * l * l * lp)) * e0[d] +
                       (-1. / (l * lp) + e0e1 / (l * lp * lp * lp)) * e1[d];
            gc[1][d] = (1. / (l * lp)) * e1[d] - (e0e1 / (l * l * l * lp)) * e0[d];
            gc[2][d] = (1. / (l * lp)) * e0[d] - (e0e1 / (l * lp * lp * lp)) * e1[d];
        }
        for (int i = 0; i < 3; ++i) {
            double *F = force + 3 * ids[i];
            for (int d = 0; d < 3; ++d) {
                const double gD11 = (l / (l0 * l0)) * gl[i][d];
                const double gD22 = (1. / (sin_phi0 * sin_phi0)) *
                    ((lp / (l0p * l0p) - l * cos_phi * cos_phi0 / (l0 * l0p)) * glp[i][d] +
                     (l * (cos_phi0 * cos_phi0) / (l0 * l0) - lp * cos_phi * cos_phi0 / (l0p * l0)) * gl[i][d] -
                     (l * lp * cos_phi0 / (l0 * l0p)) * gc[i][d]);
                const double gD12 = (1. / (2. * l0p * sin_phi0)) *
                    ((lp * cos_phi / l0p - 2. * l * cos_phi0 / l0) * gl[i][d] +
                     (l * cos_phi / l0p) * glp[i][d] +
                     (l * lp / l0p) * gc[i][d]);
                const double gI1 = 2. * (gD11 + gD22);
                const double gI2 = (2. + 4. * D22) * gD11 + (2. + 4. * D11) * gD22 - 8. * D12 * gD12;
                if (shear_model == 0)
                    F[d] += -G / 2. * area0 * ((I1 + 1) * gI1 + (Cc * I2 - 1) * gI2);
                else
                    F[d] += -G / 2. * area0 * (gI1 - (1. / ((1 + I2) * (1 + I2))) * gI2);
            }
        }
    }
}

/* -------------------------------------------------- rest-state topology */
typedef struct { int i0, i1, t1, t2; double len; } edge_rec;

static int link_cmp(const void *a, const void *b)
{
    const edge_rec *x = (const edge_rec *)a, *y = (const edge_rec *)b;
    if (x->i0 != y->i0) return x->i0 < y->i0 ? -1 : 1;
    if (x->i1 != y->i1) return x->i1 < y->i1 ? -1 : 1;
    return 0;
}

long or_build_links(long nv, long nt, const int *tri, const double *rest_pos,
                    int *link_i, double *link_len, int *link_t)
{
    (void)nv;
    /* Every edge is visited in triangle order; the first visit fixes t1,
     * the next visit by another triangle sets t2 (as std::set insertion). */
    edge_rec *all = (edge_rec *)malloc(sizeof(edge_rec) * 3 * nt);
    long n = 0;
    for (long t = 0; t < nt; ++t) {
        int ids[4] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2], tri[3 * t]};
        for (int e = 0; e < 3; ++e) {
            edge_rec L;
            L.i0 = ids[e] < ids[e + 1] ? ids[e] : ids[e + 1];
            L.i1 = ids[e] < ids[e + 1] ? ids[e + 1] : ids[e];
            long found = -1;
            for (long k = 0; k < n; ++k)
                if (all[k].i0 == L.i0 && all[k].i1 == L.i1) { found = k; break; }
            if (found < 0) {
                double d[3];
                sub3(rest_pos + 3 * L.i0, rest_pos + 3 * L.i1, d);
                L.len = norm3(d);
                L.t1 = (int)t;
                L.t2 = -1;
                all[n++] = L;
            } else if (all[found].t1 != t) {
                all[found].t2 = (int)t;
            }
        }
    }
    qsort(all, n, sizeof(edge_rec), link_cmp);
    for (long k = 0; k < n; ++k) {
        link_i[2 * k] = all[k].i0;
        link_i[2 * k + 1] = all[k].i1;
        link_len[k] = all[k].len;
        link_t[2 * k] = all[k].t1;
        link_t[2 * k + 1] = all[k].t2;
    }
    free(all);
    return n;
}

static void tri_normal(const double *v0, const double *v1, const double *v2, double *n)
{
    double a[3], b[3];
    sub3(v1, v0, a);
    sub3(v2, v0, b);
    cross3(a, b, n);
    double nn = norm3(n);
    n[0] /= nn; n[1] /= nn; n[2] /= nn;
}

static void cos_sin_pair(const double *v0, const double *v1, const double *v2, const double *v3,
                         double *cs, double *sn)
{
    double n1[3], n2[3], d[3], e[3], cr[3];
    tri_normal(v0, v1, v2, n1);
    tri_normal(v0, v2, v3, n2);
    *cs = dot3(n1, n2);
    sub3(n1, n2, d);
    sub3(v1, v3, e);
    cross3(n1, n2, cr);
    *sn = ((dot3(d, e) > 0) ? 1.0 : -1.0) * norm3(cr);
}

long or_build_pairs(long nv, long nt, const int *tri, const double *rest_pos,
                    long nlinks, const int *link_i, const int *link_t,
                    int *pair_i, double *pair_cs0)
{
    (void)nv; (void)nt;
    long np = 0;
    for (long k = 0; k < nlinks; ++k) {
        int ta = link_t[2 * k], tb = link_t[2 * k + 1];
        if (ta == -1 || tb == -1) continue;
        int i0 = link_i[2 * k], i1 = link_i[2 * k + 1];
        int t1[3] = {tri[3 * ta], tri[3 * ta + 1], tri[3 * ta + 2]};
        int t2[3] = {tri[3 * tb], tri[3 * tb + 1], tri[3 * tb + 2]};
        /* rotate (a,b,c) -> (c,a,b) until the shared edge is (0,2) / (0,1) */
        while (!((t1[0] == i0 && t1[2] == i1) || (t1[0] == i1 && t1[2] == i0))) {
            int tmp = t1[0]; t1[0] = t1[2]; t1[2] = t1[1]; t1[1] = tmp;
        }
        while (!((t2[0] == i0 && t2[1] == i1) || (t2[0] == i1 && t2[1] == i0))) {
            int tmp = t2[0]; t2[0] = t2[2]; t2[2] = t2[1]; t2[1] = tmp;
        }
        int q[4] = {t1[0], t1[1], t1[2], t2[2]};
        memcpy(pair_i + 4 * np, q, sizeof(q));
        cos_sin_pair(rest_pos + 3 * q[0], rest_pos + 3 * q[1], rest_pos + 3 * q[2],
                     rest_pos + 3 * q[3], pair_cs0 + 2 * np, pair_cs0 + 2 * np + 1);
        ++np;
    }
    return np;
}

/* ---------------------------------------------------------- RBC membrane */
static void tri_area_grad(const double *v0, const double *v1, const double *v2,
                          double *area, double *g0, double *g1, double *g2)
{
    const double v01v11 = v0[1] - v1[1];
    const double v02v22 = v0[2] - v2[2];
    const double v02v12 = v0[2] - v1[2];
    const double v01v21 = v0[1] - v2[1];
    const double v00v20 = v0[0] - v2[0];
    const double v00v10 = v0[0] - v1[0];
    const double a = v01v11 * v02v22 - v02v12 * v01v21;
    const double b = v02v12 * v00v20 - (v0[0] - v1[0]) * v02v22;
    const double c = v00v10 * v01v21 - v01v11 * v00v20;
    *area = 0.5 * sqrt(a * a + b * b + c * c);
    const double denom = 1. / (4. * *area);
    g0[0] = (b * (v2[2] - v1[2]) + c * (v1[1] - v2[1])) * denom;
    g0[1] = (a * (v1[2] - v2[2]) + c * (v2[0] - v1[0])) * denom;
    g0[2] = (a * (v2[1] - v1[1]) + b * (v1[0] - v2[0])) * denom;
    g1[0] = (b * v02v22 + c * (v2[1] - v0[1])) * denom;
    g1[1] = (a * (v2[2] - v0[2]) + c * v00v20) * denom;
    g1[2] = (a * v01v21 + b * (v2[0] - v0[0])) * denom;
    g2[0] = (b * (v1[2] - v0[2]) + c * v01v11) * denom;
    g2[1] = (a * (v02v12) + c * (v1[0] - v0[0])) * denom;
    g2[2] = (a * (v1[1] - v0[1]) + b * v00v10) * denom;
}

static void grad_angle_pair(const double *v0, const double *v1, const double *v2, const double *v3,
                            double *cs, double *sn, double g[4][3])
{
    double e0[3], e1[3], e2[3], n1[3], n2[3], n1u[3], n2u[3], e1u[3], d[3], e[3], cr[3];
    sub3(v1, v0, e0);
    sub3(v2, v0, e1);
    sub3(v3, v0, e2);
    const double e1n = norm3(e1);
    for (int k = 0; k < 3; ++k) e1u[k] = e1[k] / e1n;
    cross3(e0, e1, n1);
    cross3(e1, e2, n2);
    const double n1n = norm3(n1), n2n = norm3(n2);
    for (int k = 0; k < 3; ++k) { n1u[k] = n1[k] / n1n; n2u[k] = n2[k] / n2n; }
    *cs = dot3(n1u, n2u);
    sub3(n1, n2, d);
    sub3(v1, v3, e);
    cross3(n1u, n2u, cr);
    *sn = ((dot3(d, e) > 0) ? 1.0 : -1.0) * norm3(cr);
    double t1[3], t2[3];
    sub3(e1, e2, t1);
    sub3(e1, e0, t2);
    const double a0 = dot3(e1u, t1) / n2n, b0 = dot3(e1u, t2) / n1n;
    const double a2 = dot3(e1u, e2) / n2n, b2 = dot3(e1u, e0) / n1n;
    for (int k = 0; k < 3; ++k) {
        g[0][k] = a0 * n2u[k] + b0 * n1u[k];
        g[1][k] = (-e1n / n1n) * n1u[k];
        g[2][k] = a2 * n2u[k] + b2 * n1u[k];
        g[3][k] = (-e1n / n2n) * n2u[k];
    }
}

void or_rbc_forces(long nv, long nt, const int *tri, const double *tri_area0,
                   long nl, const int *link_i, const double *link_len,
                   long np, const int *pair_i, const double *pair_cs0,
                   double area0_total, const or_rbc_params *p,
                   double area, double volume, const double *pos,
                   double *force)
{
    (void)nv;
    const double m = 2;
    const double x0 = 1. / 2.2;
    const double ks_coeff1 = 0.25 * (6. * x0 + (3 * x0 * x0 - 2 * x0 * x0 * x0) / ((1 - x0) * (1 - x0)));
    const double ks_coeff2 = x0 / (2. * (1 - x0) * (1 - x0) * (1 - x0)) + 0.25 * (1. - 1. / ((1 - x0) * (1 - x0)));
    for (long k = 0; k < nl; ++k) {
        const double l0 = link_len[k];
        const double kp_coeff1 = -1. / pow(l0, m);
        const double kp_coeff2 = (m + 1) / pow(l0, m);
        const double G_lhs = 4. * l0 * p->shear_modulus / sqrt(3);
        const double ks = -kp_coeff1 * G_lhs / (ks_coeff1 * kp_coeff2 - ks_coeff2 * kp_coeff1);
        const double kp = ks_coeff1 * G_lhs / (ks_coeff1 * kp_coeff2 - ks_coeff2 * kp_coeff1);
        const int a = link_i[2 * k], b = link_i[2 * k + 1];
        double dv[3];
        sub3(pos + 3 * a, pos + 3 * b, dv);
        const double L = norm3(dv);
        const double x = L / (l0 * 2.2);
        const double df = -0.25 * ks * (6. * x + (3. * x * x - 2. * x * x * x) / ((1 - x) * (1 - x))) + kp / pow(L, m);
        for (int d = 0; d < 3; ++d) {
            force[3 * a + d] += df * dv[d] / L;
            force[3 * b + d] -= df * dv[d] / L;
        }
    }
    for (long t = 0; t < nt; ++t) {
        const int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
        const double *v0 = pos + 3 * i0, *v1 = pos + 3 * i1, *v2 = pos + 3 * i2;
        double A, g0[3], g1[3], g2[3];
        tri_area_grad(v0, v1, v2, &A, g0, g1, g2);
        double df_area = 0.;
        df_area += p->k_area_global * (area0_total - area) / area0_total;
        df_area += p->k_area_local * (tri_area0[t] - A) / tri_area0[t];
        for (int d = 0; d < 3; ++d) {
            force[3 * i0 + d] += df_area * g0[d];
            force[3 * i1 + d] += df_area * g1[d];
            force[3 * i2 + d] += df_area * g2[d];
        }
        cross3(v1, v2, g0);
        cross3(v2, v0, g1);
        cross3(v0, v1, g2);
        const double df_vol = p->k_volume * (p->vol_desired - volume) / p->vol_desired;
        for (int d = 0; d < 3; ++d) {
            force[3 * i0 + d] += df_vol * (g0[d] / 6.);
            force[3 * i1 + d] += df_vol * (g1[d] / 6.);
            force[3 * i2 + d] += df_vol * (g2[d] / 6.);
        }
    }
    for (long k = 0; k < np; ++k) {
        const int *q = pair_i + 4 * k;
        double cs, sn, g[4][3];
        grad_angle_pair(pos + 3 * q[0], pos + 3 * q[1], pos + 3 * q[2], pos + 3 * q[3], &cs, &sn, g);
        const double pref = p->k_bend * (-pair_cs0[2 * k] * sn + pair_cs0[2 * k + 1] * cs);
        for (int v = 0; v < 4; ++v)
            for (int d = 0; d < 3; ++d) force[3 * q[v] + d] += pref * g[v][d];
    }
}

/* ------------------------------------------------- particle geometry */
void or_particle_update(long nv, long nt, const int *tri, const double *pos,
                        double out[5], double bbox[6])
{
    for (int d = 0; d < 3; ++d) { bbox[2 * d] = DBL_MAX; bbox[2 * d + 1] = -DBL_MAX; }
    for (long v = 0; v < nv; ++v)
        for (int d = 0; d < 3; ++d) {
            if (pos[3 * v + d] < bbox[2 * d]) bbox[2 * d] = pos[3 * v + d];
            if (pos[3 * v + d] > bbox[2 * d + 1]) bbox[2 * d + 1] = pos[3 * v + d];
        }
    double area = 0, vol = 0, com[3] = {0, 0, 0};
    for (long t = 0; t < nt; ++t) {
        const double *v0 = pos + 3 * tri[3 * t], *v1 = pos + 3 * tri[3 * t + 1], *v2 = pos + 3 * tri[3 * t + 2];
        double a[3], b[3], n[3], cr[3];
        sub3(v1, v0, a);
        sub3(v2, v0, b);
        cross3(a, b, n);
        double dA = 0.5 * norm3(n);
        area += dA;
        cross3(v1, v2, cr);
        vol += dot3(v0, cr) / 6.;
        for (int d = 0; d < 3; ++d) com[d] += ((v0[d] + v1[d] + v2[d]) / 3.0) * dA;
    }
    out[0] = area;
    out[1] = vol;
    for (int d = 0; d < 3; ++d) out[2 + d] = com[d] / area;
}

void or_moment_of_inertia(long nv, long nt, const int *tri, const double *pos, double I[9])
{
    (void)nv;
    double in[10] = {0};
    for (long t = 0; t < nt; ++t) {
        const double *v0 = pos + 3 * tri[3 * t], *v1 = pos + 3 * tri[3 * t + 1], *v2 = pos + 3 * tri[3 * t + 2];
        double a[3], b[3], N[3];
        sub3(v1, v0, a);
        sub3(v2, v0, b);
        cross3(a, b, N);
        double f1[3], f2[3], f3[3], g0[3], g1[3], g2[3];
        for (int d = 0; d < 3; ++d) {
            double tmp0 = v0[d] + v1[d];
            f1[d] = tmp0 + v2[d];
            double tmp1 = v0[d] * v0[d];
            double tmp2 = tmp1 + v1[d] * tmp0;
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
    in[0] *= 1.0 / 6.0;
    for (int k = 1; k <= 3; ++k) in[k] *= 1.0 / 24.0;
    for (int k = 4; k <= 6; ++k) in[k] *= 1.0 / 60.0;
    for (int k = 7; k <= 9; ++k) in[k] *= 1.0 / 120.0;
    const double V = in[0];
    const double c[3] = {in[1] / V, in[2] / V, in[3] / V};
    I[0] = in[5] + in[6] - V * (c[1] * c[1] + c[2] * c[2]);
    I[1] = -in[7] + V * c[0] * c[1];
    I[2] = -in[9] + V * c[2] * c[0];
    I[3] = I[1];
    I[4] = in[4] + in[6] - V * (c[2] * c[2] + c[0] * c[0]);
    I[5] = -in[8] + V * c[1] * c[2];
    I[6] = I[2];
    I[7] = I[5];
    I[8] = in[4] + in[5] - V * (c[0] * c[0] + c[1] * c[1]);
}

/* ------------------------------------------------------- voxelization */
long or_ray_hits_z(long nv, long nt, const int *tri, const double *pos, long cap, double *hits)
{
    (void)nv;
    long n = 0;
    for (long t = 0; t < nt; ++t) {
        const double *v0 = pos + 3 * tri[3 * t], *v1 = pos + 3 * tri[3 * t + 1], *v2 = pos + 3 * tri[3 * t + 2];
        long x0 = (long)ceil(fmin(v0[0], fmin(v1[0], v2[0])));
        long x1 = (long)floor(fmax(v0[0], fmax(v1[0], v2[0])));
        long y0 = (long)ceil(fmin(v0[1], fmin(v1[1], v2[1])));
        long y1 = (long)floor(fmax(v0[1], fmax(v1[1], v2[1])));
        double e1[3], e2[3], q[3], s[3], r[3];
        const double dir[3] = {0, 0, 1};
        sub3(v1, v0, e1);
        sub3(v2, v0, e2);
        cross3(dir, e2, q);
        const double a = dot3(e1, q);
        for (long i = x0; i <= x1; ++i)
            for (long j = y0; j <= y1; ++j) {
                if (fabs(a) < DBL_EPSILON) continue;
                const double f = 1 / a;
                const double o[3] = {(double)i, (double)j, 0.};
                sub3(o, v0, s);
                const double u = f * dot3(s, q);
                if (u < 0.0) continue;
                cross3(s, e1, r);
                const double v = f * dot3(dir, r);
                if (v < 0.0 || u + v > 1.0) continue;
                if (n < cap) {
                    hits[3 * n] = (double)i;
                    hits[3 * n + 1] = (double)j;
                    hits[3 * n + 2] = f * dot3(e2, r);
                }
                ++n;
            }
    }
    return n;
}

/* Column fill of the phase field from z-ray hits, as apply_voxelization
 * does (src/fsi/ImmersedBoundaryDynamics_voxelization.hh:121-165): hits are
 * mapped into [-0.5, nz-0.5) on a periodic z axis (src/fsi/ZBuffer.h:39-49),
 * columns wrap on periodic x / y (src/fsi/ZBuffer.h:211-213), each column's
 * hits are sorted and the nodes between consecutive hits alternate
 * outside / inside, a later segment overwriting a shared end node. */
static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

void or_voxel_fill(int nx, int ny, int nz, const int periodic[3], long nh,
                   const double *hits, unsigned char *phase)
{
    long ncol = (long)nx * ny;
    long *cnt = (long *)calloc(ncol + 1, sizeof(long));
    double *zs = (double *)malloc(sizeof(double) * (nh > 0 ? nh : 1));
    long *col_of = (long *)malloc(sizeof(long) * (nh > 0 ? nh : 1));
    for (long k = 0; k < nh; ++k) {
        long i = (long)hits[3 * k], j = (long)hits[3 * k + 1];
        if (periodic[0]) i = ((i % nx) + nx) % nx;
        if (periodic[1]) j = ((j % ny) + ny) % ny;
        if (i < 0 || i >= nx || j < 0 || j >= ny) { col_of[k] = -1; continue; }
        col_of[k] = i * ny + j;
        cnt[col_of[k] + 1]++;
    }
    for (long c = 0; c < ncol; ++c) cnt[c + 1] += cnt[c];
    long *fill = (long *)calloc(ncol, sizeof(long));
    for (long k = 0; k < nh; ++k) {
        if (col_of[k] < 0) continue;
        double z = hits[3 * k + 2];
        if (periodic[2]) z = z - (double)nz * floor((z + 0.5) * (1.0 / nz));
        zs[cnt[col_of[k]] + fill[col_of[k]]++] = z;
    }
    for (long c = 0; c < ncol; ++c) {
        long n = cnt[c + 1] - cnt[c];
        double *z = zs + cnt[c];
        qsort(z, n, sizeof(double), cmp_double);
        unsigned char *ph = phase + c * nz;
        int inside = 0;
        double z_last = -10000;
        for (long k = 0; k < n; ++k) {
            long lo = (long)ceil(z_last), hi = (long)floor(z[k]);
            if (lo < 0) lo = 0;
            if (hi > nz - 1) hi = nz - 1;
            for (long q = lo; q <= hi; ++q) ph[q] = inside ? 1 : 0;
            z_last = z[k];
            inside = !inside;
        }
        if (ceil(z_last) < nz - 1) {
            long lo = (long)ceil(z_last);
            if (lo < 0) lo = 0;
            for (long q = lo; q <= nz - 1; ++q) ph[q] = inside ? 1 : 0;
        }
    }
    free(cnt); free(zs); free(col_of); free(fill);
}
