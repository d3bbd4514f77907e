/*
 * oracle/interactions.c -- FP64 CPU restatement of the reference's
 * particle-particle and particle-wall interaction forces.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).  Never linked into
 * the product library.
 *
 * Follows, operation by operation:
 *   MorsePotential / SpringPotential       src/fsi/Potentials.h:28-72
 *   WallInteraction::apply_force          src/fsi/ForceDecorator.hh:27-35
 *   collision domain and node insertion   src/fsi/ImmersedBoundaryDynamics_forces.hh:332-375
 *   Grid (cell size 1, one-cell envelope)  src/fsi/Grid.hh:30-97
 *   CollisionHandler pair enumeration     src/fsi/CollisionForces.h:84-161
 *   wall collision forces                 src/fsi/CollisionForces.h:167-187
 *   collide_all_in_cells / handle_collision src/fsi/CollisionForces.h:191-212
 *   minimal-image difference              src/fsi/Periodicity.h:27-35,91-99,268-275
 *   boundary / rectangle intersection     src/fsi/Boundary.h:227-242,316-320, src/fsi/geometry.h:20-28
 *
 * The collision handler is restated in its own scatter form: nodes are
 * binned into per-cell lists in insertion order and every pair is visited
 * in the reference's loop order, adding +df to the first node and -df to
 * the second, so the sums round exactly as the reference's do.
 */
#include "oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double sqr(double x) { return x * x; }

double or_potential_force(const or_potential *u, double dist)
{
    if (u->kind == OR_POTENTIAL_MORSE) {
        /* r_min = p[0], beta = p[1], strength = p[2] */
        if (dist < 0) dist = 0;
        const double two_beta_strength = 2 * u->p[1] * u->p[2];
        const double x = exp(u->p[1] * (u->p[0] - dist));
        return two_beta_strength * x * (x - 1);
    }
    /* spring: min_dist = p[0], strength = p[1] */
    return (u->p[0] - dist) * u->p[1];
}

double or_potential_cutoff(const or_potential *u) { return u->p[0]; }

void or_wall_interaction(const or_boundary *b, const or_potential *u, long nv, const double *pos, double *force)
{
    const double cut = or_potential_cutoff(u);
    for (long v = 0; v < nv; ++v) {
        const double *p = pos + 3 * v;
        const double d = or_boundary_distance(b, p);
        if (d <= cut) {
            const double F = or_potential_force(u, d);
            double n[3];
            or_boundary_normal(b, p, n);
            for (int k = 0; k < 3; ++k) force[3 * v + k] -= F * n[k];
        }
    }
}

/* Grid geometry of the collision handler over the lattice box
 * [0, n-1]^3: bulk cells 1..num, envelope cells 0 and num+1. */
void or_collision_grid(int nx, int ny, int nz, double cutoff, or_collision_grid_t *g)
{
    const int n[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d) {
        const double lo = 0 - 0.5, hi = (double)(n[d] - 1) + 0.5 - DBL_EPSILON;
        g->bb_lo[d] = lo;
        g->bb_hi[d] = hi;
        g->num[d] = (long)floor((hi - lo) / 1.0);
        g->cs[d] = (hi - lo) / (double)g->num[d];
        g->x0g[d] = lo - g->cs[d];
        g->lo[d] = lo - cutoff;
        g->hi[d] = hi + cutoff;
        g->L[d] = (double)n[d];
        g->Linv[d] = 1. / g->L[d];
    }
}

long or_collision_cell(const or_collision_grid_t *g, const double p[3])
{
    long ix[3];
    for (int d = 0; d < 3; ++d)
        if (!(p[d] > g->lo[d] && p[d] < g->hi[d])) return -1;
    for (int d = 0; d < 3; ++d) {
        ix[d] = (long)floor((p[d] - g->x0g[d]) / g->cs[d]);
        /* the reference indexes past its grid here (cut-offs above one cell
         * size); such nodes are left out */
        if (ix[d] < 0 || ix[d] > g->num[d] + 1) return -1;
    }
    return ix[2] + (g->num[2] + 2) * (ix[1] + (g->num[1] + 2) * ix[0]);
}

typedef struct {
    const or_collision_grid_t *g;
    const int *periodic;
    const or_potential *u;
    double cut2;
    const double *pos;
    const int *obj;
    double *force;
    const long *head, *list; /* per cell: list[head[c] .. head[c+1]) */
    long pairs;
} coll_ctx;

static double min_mod_diff(const coll_ctx *c, int d, double x, double y)
{
    const double v = x - y;
    if (!c->periodic[d]) return v;
    return v - floor(v * c->g->Linv[d] + 0.5) * c->g->L[d];
}

static void handle(coll_ctx *c, long a, long b)
{
    if (c->obj[a] == c->obj[b]) return;
    double dr[3];
    for (int d = 0; d < 3; ++d) dr[d] = min_mod_diff(c, d, c->pos[3 * b + d], c->pos[3 * a + d]);
    const double d2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    if (d2 <= c->cut2) {
        const double dist = sqrt(d2);
        const double s = -or_potential_force(c->u, dist) / dist;
        for (int d = 0; d < 3; ++d) {
            const double df = s * dr[d];
            c->force[3 * a + d] += df;
            c->force[3 * b + d] -= df;
        }
        c->pairs++;
    }
}

static long cell_at(const coll_ctx *c, long i, long j, long k)
{
    return k + (c->g->num[2] + 2) * (j + (c->g->num[1] + 2) * i);
}

static void all_in_cells(coll_ctx *c, long c1, long c2)
{
    for (long l = c->head[c1]; l < c->head[c1 + 1]; ++l)
        for (long m = c->head[c2]; m < c->head[c2 + 1]; ++m) handle(c, c->list[l], c->list[m]);
}

static int pipe_contains_margin(const or_boundary *b, double y, double z, double margin)
{
    const double q[3] = {0, y, z};
    return or_boundary_contains(b, q, margin);
}

static double clampd(double v, double lo, double hi)
{
    if (v >= hi) return hi;
    if (v <= lo) return lo;
    return v;
}

int or_boundary_intersects_rect(const or_boundary *b, const double r[6], double margin)
{
    if (b->kind == OR_BOUNDARY_PIPE) {
        const double y0 = b->p[0], z0 = b->p[1], R = b->p[2];
        const double tmp = sqr(y0 - clampd(y0, r[2], r[3])) + sqr(z0 - clampd(z0, r[4], r[5]));
        if (tmp <= sqr(R))
            return !(pipe_contains_margin(b, r[2], r[4], margin) && pipe_contains_margin(b, r[2], r[5], margin) &&
                     pipe_contains_margin(b, r[3], r[4], margin) && pipe_contains_margin(b, r[3], r[5], margin));
        return 0;
    }
    const double y0 = b->p[0], y1 = b->p[1];
    return (r[2] < (y0 + margin) && r[3] > (y0 + margin)) || (r[2] < (y1 - margin) && r[3] > (y1 - margin));
}

long or_collision_forces(int nx, int ny, int nz, const int periodic[3], const or_potential *u, long nv,
                         const double *pos, const int *obj, const or_boundary *wall, double *force)
{
    const double cut = or_potential_cutoff(u);
    or_collision_grid_t g;
    or_collision_grid(nx, ny, nz, cut, &g);
    const long ncell = (g.num[0] + 2) * (g.num[1] + 2) * (g.num[2] + 2);
    long *head = (long *)calloc((size_t)ncell + 1, sizeof(long));
    long *list = (long *)malloc(sizeof(long) * (size_t)(nv > 0 ? nv : 1));
    long *cell = (long *)malloc(sizeof(long) * (size_t)(nv > 0 ? nv : 1));
    /* counting sort: per-cell lists in insertion (vertex) order */
    for (long v = 0; v < nv; ++v) {
        cell[v] = or_collision_cell(&g, pos + 3 * v);
        if (cell[v] >= 0) head[cell[v] + 1]++;
    }
    for (long c = 0; c < ncell; ++c) head[c + 1] += head[c];
    {
        long *fill = (long *)malloc(sizeof(long) * (size_t)ncell);
        memcpy(fill, head, sizeof(long) * (size_t)ncell);
        for (long v = 0; v < nv; ++v)
            if (cell[v] >= 0) list[fill[cell[v]]++] = v;
        free(fill);
    }
    coll_ctx c = {&g, periodic, u, sqr(cut), pos, obj, force, head, list, 0};
    const long X0 = 1, X1 = g.num[0], Y0 = 1, Y1 = g.num[1], Z0 = 1, Z1 = g.num[2];
    static const int off[7][3] = {{0, 1, 0}, {0, 1, 1}, {0, 0, 1}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}};
    for (long i = X0; i <= X1; ++i)
        for (long j = Y0; j <= Y1; ++j)
            for (long k = Z0; k <= Z1; ++k) {
                const long ce = cell_at(&c, i, j, k);
                long nb[7];
                for (int o = 0; o < 7; ++o) nb[o] = cell_at(&c, i + off[o][0], j + off[o][1], k + off[o][2]);
                for (long l = head[ce]; l < head[ce + 1]; ++l) {
                    for (long m = l + 1; m < head[ce + 1]; ++m) handle(&c, list[l], list[m]);
                    for (int o = 0; o < 7; ++o)
                        for (long m = head[nb[o]]; m < head[nb[o] + 1]; ++m) handle(&c, list[l], list[m]);
                }
            }
    if (periodic[0])
        for (long j = Y0; j <= Y1; ++j)
            for (long k = Z0; k <= Z1; ++k) all_in_cells(&c, cell_at(&c, X0, j, k), cell_at(&c, X1, j, k));
    if (periodic[1])
        for (long i = X0; i <= X1; ++i)
            for (long k = Z0; k <= Z1; ++k) all_in_cells(&c, cell_at(&c, i, Y0, k), cell_at(&c, i, Y1, k));
    if (periodic[2])
        for (long i = X0; i <= X1; ++i)
            for (long j = Y0; j <= Y1; ++j) all_in_cells(&c, cell_at(&c, i, j, Z0), cell_at(&c, i, j, Z1));
    if (periodic[0] && periodic[1])
        for (long k = Z0; k <= Z1; ++k) all_in_cells(&c, cell_at(&c, X0, Y0, k), cell_at(&c, X1, Y1, k));
    if (periodic[0] && periodic[2])
        for (long j = Y0; j <= Y1; ++j) all_in_cells(&c, cell_at(&c, X0, j, Z0), cell_at(&c, X1, j, Z1));
    if (periodic[1] && periodic[2])
        for (long i = X0; i <= X1; ++i) all_in_cells(&c, cell_at(&c, i, Y0, Z0), cell_at(&c, i, Y1, Z1));
    if (periodic[0] && periodic[1] && periodic[2])
        all_in_cells(&c, cell_at(&c, X0, Y0, Z0), cell_at(&c, X1, Y1, Z1));

    if (wall) {
        for (long i = X0; i <= X1; ++i)
            for (long j = Y0; j <= Y1; ++j)
                for (long k = Z0; k <= Z1; ++k) {
                    const double x0 = g.bb_lo[0] + (double)(i - 1) * g.cs[0];
                    const double y0 = g.bb_lo[1] + (double)(j - 1) * g.cs[1];
                    const double z0 = g.bb_lo[2] + (double)(k - 1) * g.cs[2];
                    const double r[6] = {x0, x0 + g.cs[0], y0, y0 + g.cs[1], z0, z0 + g.cs[2]};
                    if (!or_boundary_intersects_rect(wall, r, cut)) continue;
                    const long ce = cell_at(&c, i, j, k);
                    for (long l = head[ce]; l < head[ce + 1]; ++l) {
                        const long v = list[l];
                        const double d = or_boundary_distance(wall, pos + 3 * v);
                        if (d <= cut) {
                            const double F = or_potential_force(u, d);
                            double n[3];
                            or_boundary_normal(wall, pos + 3 * v, n);
                            for (int q = 0; q < 3; ++q) force[3 * v + q] -= F * n[q];
                        }
                    }
                }
    }
    free(head);
    free(list);
    free(cell);
    return c.pairs;
}
