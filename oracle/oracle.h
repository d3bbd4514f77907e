/*
 * CPU restatement (FP64, plain C) of the reference algorithms on the
 * D3Q19 + immersed-boundary hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg, and
 * by nothing else.  The product path (biblbm_amd, libbiblbm.so) never calls
 * it.
 *
 * Layout conventions (shared with the product C-ABI so that arrays can be
 * compared directly):
 *   cell index        c = (x*ny + y)*nz + z          (z fastest, as Palabos)
 *   populations       f[c*19 + i], stored as fBar_i = f_i - t_i
 *   force             F[c*3 + d]
 *   lattice vectors   Palabos D3Q19 ordering: i and i+9 opposite
 *   mask              0 fluid (Guo-VOF BGK), 1 bounce-back, 2 no dynamics,
 *                     3 velocity-wall node (regularized FD boundary)
 */
#ifndef BIBLBM_ORACLE_H
#define BIBLBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- lattice ---------------------------------------------------------- */
void or_d3q19(int c[19][3], double t[19]);

void or_collide_cell(double f[19], const double F[3], double phase,
                     double omega_out, double omega_in);

/* One collide-and-stream step followed by the velocity-wall post-stream
 * pass.  `ftmp` is scratch of the same size as f.  wall_u[side*3+d] is the
 * velocity of the wall plane at index 0 (side 0) / n-1 (side 1) of
 * `wall_axis`; both are ignored when no cell has mask 3. */
void or_lbm_step(int nx, int ny, int nz, const int periodic[3],
                 double omega_out, double omega_in,
                 double *f, double *ftmp, const double *force,
                 const double *phase, const unsigned char *mask,
                 int wall_axis, const double *wall_u);

/* The regularized velocity-wall pass alone (what lattice.initialize() runs). */
void or_velocity_walls(int nx, int ny, int nz, const int periodic[3],
                       double omega_out, double *f, const double *force,
                       const unsigned char *mask, int wall_axis,
                       const double *wall_u);

/* rho and u = j/rho + F/2 per cell (velocity-wall cells report the wall
 * velocity and the boundary density rule). */
void or_lbm_moments(int nx, int ny, int nz, const double *f,
                    const double *force, const unsigned char *mask,
                    int wall_axis, const double *wall_u,
                    double *rho, double *u);

/* Equilibrium initialisation fBar_i = feq_i(rho, j = rho u). */
void or_lbm_init_equilibrium(long ncell, const double *rho, const double *u,
                             double *f);

/* Sums over all cells (FP64): mass = sum(1 + sum_i fBar_i), momentum j. */
void or_lbm_conserved(long ncell, const double *f, const unsigned char *mask,
                      double out[4]);

/* ---- immersed boundary ---------------------------------------------- */
/* kernel: 0 = Roma 3-point, 1 = Peskin 4-point, 2 = top hat 2-point */
double or_dirac_1d(int kernel, double r);
int    or_dirac_support(int kernel);
double or_dirac_half_support(int kernel);
/* base index of the support box: ceil(x - h) */
long   or_support_base(int kernel, double x);

/* Interpolation of a (nx,ny,nz,3) velocity field at nv points with the
 * given kernel; periodic axes wrap, other axes skip out-of-range nodes. */
void or_interpolate(int kernel, int nx, int ny, int nz, const int periodic[3],
                    const double *ufield, long nv, const double *pos,
                    double *vel);

/* Spreading: F(x) += sum_k phi(x - X_k) F_k on a (nx,ny,nz,3) field. */
void or_spread(int kernel, int nx, int ny, int nz, const int periodic[3],
               double *ffield, long nv, const double *pos,
               const double *fvert);

/* Membrane models.  tri[3*t], rest_pos[3*v], pos[3*v]; forces are ADDED to
 * force[3*v].  shear_model: 0 Skalak, 1 NeoHookean.  `bug_compatible`
 * reproduces the reference's c = l' sin(phi) / (l' sin(phi0)) term. */
void or_capsule_forces(long nv, long nt, const int *tri,
                       const double *rest_pos, const double *pos,
                       double G, double C, int shear_model, int bug_compatible,
                       double *force);

/* Rest-state topology (reference ParticleShape::compute_connectivity):
 * returns counts; arrays must be sized 3*nt for links / pairs. */
long or_build_links(long nv, long nt, const int *tri, const double *rest_pos,
                    int *link_i, double *link_len, int *link_t);
long or_build_pairs(long nv, long nt, const int *tri, const double *rest_pos,
                    long nlinks, const int *link_i, const int *link_t,
                    int *pair_i, double *pair_cs0);

typedef struct {
    double shear_modulus, L0;
    double k_area_global, k_area_local;
    double k_volume, vol_desired;
    double k_bend, theta0;
} or_rbc_params;

void or_rbc_forces(long nv, long nt, const int *tri, const double *tri_area0,
                   long nl, const int *link_i, const double *link_len,
                   long np, const int *pair_i, const double *pair_cs0,
                   double area0_total, const or_rbc_params *p,
                   double area, double volume, const double *pos,
                   double *force);

/* Particle geometry: area, signed volume, area-weighted centroid COM,
 * bounding box [x0,x1,y0,y1,z0,z1]. */
void or_particle_update(long nv, long nt, const int *tri, const double *pos,
                        double out_area_vol_com[5], double bbox[6]);

/* Moment of inertia of the enclosed volume about the COM (divergence
 * theorem integrals). */
void or_moment_of_inertia(long nv, long nt, const int *tri, const double *pos,
                          double I[9]);

/* Voxelization: z-hits of rays at integer (x,y) along +z. Returns hit
 * count (<= cap); hits[k] = {x, y, z}. */
long or_ray_hits_z(long nv, long nt, const int *tri, const double *pos,
                   long cap, double *hits);

/* Phase field (ncell bytes) from z-ray hits: column sort + parity fill. */
void or_voxel_fill(int nx, int ny, int nz, const int periodic[3], long nh,
                   const double *hits, unsigned char *phase);

/* ---- curved walls ------------------------------------------------- */
#define OR_KERNEL_TOPHAT 2
#define OR_BOUNDARY_PIPE 0   /* along x: p = {y0, z0, radius, length} */
#define OR_BOUNDARY_PLATES 1 /* y0 < y < y1: p = {y0, y1} */
typedef struct {
    int kind;
    double p[4];
} or_boundary;

int    or_boundary_contains(const or_boundary *b, const double p[3], double margin);
int    or_boundary_less_than(const or_boundary *b, const double p[3], double dist);
double or_boundary_distance(const or_boundary *b, const double p[3]);
int    or_boundary_trace_ray(const or_boundary *b, const double x0[3], const double x1[3], double *t);
void   or_boundary_normal(const or_boundary *b, const double p[3], double n[3]);

/* Guo wall nodes of an nx*ny*nz lattice at integer positions: cells outside
 * the boundary with a lattice link crossing it.  Returns the count (arrays
 * filled up to `cap`): cell index, lattice direction (1..18) towards the
 * fluid, delta = 1 - t, wall position. */
long or_guo_wall_nodes(int nx, int ny, int nz, const or_boundary *b, long cap, long *cell, int *dir,
                       double *delta, double *wall_pos);

/* The Guo extrapolation over the wall nodes, in node order, on post-stream
 * populations f (cell-major, fBar).  vel[3*k]: wall velocity of node k.
 * mask (may be NULL: all fluid) gives each wall node's dynamics. */
void or_guo_wall_apply(int nx, int ny, int nz, const int periodic[3], double *f, const unsigned char *mask,
                       long n, const long *cell, const int *dir, const double *delta, const double *vel);

/* Near-wall interpolation / spreading (b may be NULL: plain kernel) and
 * vertex advection with reflection off the boundary. */
void or_interpolate_near(int kernel, int nx, int ny, int nz, const int periodic[3], const double *ufield,
                         long nv, const double *pos, const or_boundary *b, double *vel);
void or_spread_near(int kernel, int nx, int ny, int nz, const int periodic[3], double *ffield, long nv,
                    const double *pos, const double *fvert, const or_boundary *b);
void or_move_reflect(long nv, double *pos, const double *vel, const or_boundary *b);

/* ---- interaction forces ------------------------------------------- */
#define OR_POTENTIAL_MORSE 0  /* p = {r_min, beta, strength} */
#define OR_POTENTIAL_SPRING 1 /* p = {min_dist, strength} */
typedef struct {
    int kind;
    double p[3];
} or_potential;

double or_potential_force(const or_potential *u, double dist);
double or_potential_cutoff(const or_potential *u);

/* Wall interaction decorator: force -= F(d) n for every vertex at
 * distance d <= cutoff from the boundary. */
void or_wall_interaction(const or_boundary *b, const or_potential *u, long nv, const double *pos, double *force);

/* Collision-handler grid over an nx*ny*nz lattice box: bounding box
 * [-0.5, n - 0.5 - eps], cells of size ~1 (num per axis) plus a one-cell
 * envelope, nodes taken if strictly inside the box enlarged by the
 * cut-off. */
typedef struct {
    double bb_lo[3], bb_hi[3], cs[3], x0g[3], lo[3], hi[3], L[3], Linv[3];
    long num[3];
} or_collision_grid_t;
void or_collision_grid(int nx, int ny, int nz, double cutoff, or_collision_grid_t *g);
/* cell index k + (num_z+2)*(j + (num_y+2)*i) of a node, or -1 if not taken */
long or_collision_cell(const or_collision_grid_t *g, const double p[3]);
int or_boundary_intersects_rect(const or_boundary *b, const double r[6], double margin);

/* Particle-particle interaction forces of nv nodes (obj: particle of each
 * node; nodes of one particle do not interact), ADDED to force, in the
 * reference's pair order; wall (may be NULL) adds the handler's wall
 * collision forces afterwards.  Returns the number of interacting pairs. */
long or_collision_forces(int nx, int ny, int nz, const int periodic[3], const or_potential *u, long nv,
                         const double *pos, const int *obj, const or_boundary *wall, double *force);

/* ---- rigid bodies (growth packing) ------------------------------------ */
/* Rest-shape moment of inertia about the mesh origin
 * (src/fsi/ParticleShape.hh:249-360), row-major. */
void or_shape_inertia(long nv, long nt, const int *tri, const double *rest, double I[9]);
/* state (16): com (3), velocity (3), angular velocity (3), quaternion
 * (w, x, y, z), scale, density, damping.  update: pos = com + R (scale v)
 * (src/fsi/RigidParticle.hh:30-39); move: RigidParticle3D::move_vertices
 * (src/fsi/RigidParticle.hh:57-98) with the vertex forces `force`, then the
 * centre of mass mapped onto the periodic grid with the vertices. */
void or_rigid_update(long nv, const double *rest, const double *state, double *pos);
void or_rigid_move(long nv, const double *rest, const double I[9], double volume, const int periodic[3],
                   const int n[3], double *state, double *pos, const double *force);

#ifdef __cplusplus
}
#endif
#endif
