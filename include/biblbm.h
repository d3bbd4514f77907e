/*
 * biblbm.h -- C-ABI of libbiblbm.so, the MI355X (gfx950) implementation of
 * the D3Q19 Guo-VOF collide-and-stream and the immersed-boundary coupling.
 *
 * Conventions
 *   - Every call returns an int status: BL_OK (0) or a BL_ERR_* code; the
 *     text of the last error on the calling thread is bl_last_error().
 *   - Handles are opaque.  Device memory is owned by the library; host
 *     arrays passed in are borrowed for the duration of the call only.
 *   - One HIP stream per lattice (default priority).  Calls are ordered on
 *     that stream; calls that return data to the host synchronize it.
 *   - Cell index c = (x*ny + y)*nz + z (z fastest).  Populations crossing the
 *     boundary are fBar_i = f_i - t_i in the Palabos D3Q19 ordering (i and
 *     i+9 opposite), 19 per cell, cell-major: buf[c*19 + i].
 *   - Cell mask: 0 fluid (Guo-VOF BGK), 1 bounce-back, 2 no dynamics,
 *     3 velocity-wall node (collides as fluid, then regularized after
 *     streaming).
 *
 * The reference interface each entry point replaces is cited next to it
 * (paths relative to the reference checkout).
 */
#ifndef BIBLBM_H
#define BIBLBM_H

#ifdef __cplusplus
extern "C" {
#endif

#define BL_OK 0
#define BL_ERR_ARG 1
#define BL_ERR_HIP 2
#define BL_ERR_STATE 3
#define BL_ERR_CAPACITY 4

/* storage of the populations */
#define BL_STORE_F64 0   /* FP64 fBar, FP64 collision                     */
#define BL_STORE_F32 1   /* FP32 fBar, FP32 collision                     */
#define BL_STORE_F16S 2  /* FP16 of fBar_i / t_i * 2^10, FP32 collision   */

/* immersed-boundary delta kernels */
#define BL_ORDER_ROWS 3  /* collide launch order: whole z-rows per XCD (bl_lattice_set_launch_parity) */

#define BL_KERNEL_ROMA 0    /* 3-point, src/fsi/Dirac.h:74-101   */
#define BL_KERNEL_PESKIN 1  /* 4-point, src/fsi/Dirac.h:23-44    */
#define BL_KERNEL_TOPHAT 2  /* 2-point, src/fsi/Dirac.h:51-66    */

/* membrane models */
#define BL_MODEL_CAPSULE_SKALAK 0      /* src/fsi/DeformableCapsule.hh:140-142 */
#define BL_MODEL_CAPSULE_NEOHOOKEAN 1  /* src/fsi/DeformableCapsule.hh:143-145 */
#define BL_MODEL_RBC 2                 /* src/fsi/RBCParticle.hh:140-247      */
/* Skalak in-plane law of the capsule (src/fsi/DeformableCapsule.hh:60-150)
 * with the RBC model's local + global area, volume and bending terms
 * (src/fsi/RBCParticle.hh:176-246) and without its WLC-POW links: the
 * "Skalak + bending + area + volume" membrane of the red-cell workloads.
 * params: {G, C, k_area_global, k_area_local, k_volume, vol_desired, k_bend} */
#define BL_MODEL_SKALAK_RBC 3
/* RigidParticle3D (src/fsi/RigidParticle.hh:13-98), added with bl_rigid_add */
#define BL_MODEL_RIGID 4

typedef struct bl_lattice_s *bl_lattice;
typedef struct bl_ibm_s *bl_ibm;
typedef struct bl_ensemble_s *bl_ensemble;
typedef struct bl_guo_wall_s *bl_guo_wall;
typedef struct bl_positions_s *bl_positions;

/* Analytic boundaries (src/fsi/Boundary.h): PipeBoundary(y0, z0, radius,
 * length) along x (:132-186), ParallelPlatesBoundary(y0, y1) normal to y
 * (:305-365).  Lattice coordinates: cell (x, y, z) sits at (x, y, z). */
#define BL_BOUNDARY_PIPE 0    /* p = {y0, z0, radius, length} */
#define BL_BOUNDARY_PLATES 1  /* p = {y0, y1, -, -}           */
typedef struct {
    int kind;
    double p[4];
} bl_boundary;

/* ------------------------------------------------------------ runtime */
/* Replaces plbInit (src/core/plbInit.cpp:38) for the device side. */
int bl_init(int device);
int bl_shutdown(void);
const char *bl_last_error(void);
/* name (NUL-terminated, <= cap bytes), compute units, HBM bytes */
int bl_device_info(char *name, int cap, int *compute_units, double *hbm_bytes);

/* ------------------------------------------------------------ lattice */
/* MultiBlockLattice3D(nx,ny,nz, dynamics) + periodicity().toggle
 * (src/multiBlock/multiBlockLattice3D.h, src/multiBlock/multiBlock3D.h:77). */
int bl_lattice_create(int nx, int ny, int nz, const int periodic[3], int storage,
                      bl_lattice *out);
int bl_lattice_destroy(bl_lattice lat);
/* GuoExternalForceVOFDynamics(omega_out, omega_in)
 * (src/fsi/VariableViscosityDynamics.h:48-54). */
int bl_lattice_set_dynamics(bl_lattice lat, double omega_out, double omega_in);
/* defineDynamics(... BounceBack / NoDynamics) and the velocity-boundary cell
 * set (src/dataProcessors/dataInitializerWrapper3D.h, src/core/dynamics.h:559,670,
 * src/boundaryCondition/boundaryCondition3D.hh:1053-1066).  NULL = all fluid. */
int bl_lattice_set_mask(bl_lattice lat, const unsigned char *mask);
/* No counterpart in the reference (its BounceBack cells all stream): on = 1
 * leaves inert cells alone in the collide-and-stream, bounce-back cells
 * whose 18 neighbours are all bounce-back (or past a non-periodic face),
 * for example the solid beyond a pipe wall.  What such a cell holds never
 * reaches a fluid cell (a bounce-back cell hands each neighbour back what
 * that neighbour sent), so every fluid, velocity-wall and no-dynamics cell,
 * every bounce-back cell next to one, the moments and the conservation sums
 * step bitwise as with on = 0; the inert cells keep stale populations.  The
 * set is rebuilt from each new mask; the odd step's z-segment kernel then
 * runs as the per-cell kernel.  Default 0. */
int bl_lattice_set_skip_inert(bl_lattice lat, int on);
/* n = inert cells skipped (0 with skipping off). */
int bl_lattice_inert_cells(bl_lattice lat, long long *n);
/* setBoundaryVelocity on the two planes normal to `axis` that carry mask 3
 * (index 0: u_lo, index n-1: u_hi). */
int bl_lattice_set_velocity_walls(bl_lattice lat, int axis, const double u_lo[3],
                                  const double u_hi[3]);
/* setExternalVector(lattice, domain, forceBeginsAt, g)
 * (src/dataProcessors/dataInitializerWrapper3D.h:301): the force every cell
 * carries outside IBM-touched regions. */
int bl_lattice_set_body_force(bl_lattice lat, const double g[3]);
/* Per-cell external force (ncell*3) and phase (ncell, 0 or 1): the
 * ForcedPhaseD3Q19 externals (src/fsi/VariableViscosityDynamics.h:25-42). */
int bl_lattice_set_force_field(bl_lattice lat, const double *force);
int bl_lattice_set_phase(bl_lattice lat, const unsigned char *phase);
/* initializeAtEquilibrium (src/core/cell.hh:124-132); NULL rho -> 1, NULL u -> 0 */
int bl_lattice_init_equilibrium(bl_lattice lat, const double *rho, const double *u);
/* Raw population transfer (the coprocessor layout of
 * src/coProcessors/coProcessor3D.h:84-92): buf[c*19+i] = fBar_i. */
int bl_lattice_put_populations(bl_lattice lat, const double *buf);
int bl_lattice_get_populations(bl_lattice lat, double *buf);
/* MultiBlock3D::initialize(): runs the post-stream boundary pass once
 * (src/multiBlock/multiBlock3D.cpp:330-332). */
int bl_lattice_initialize(bl_lattice lat);
/* MultiBlockLattice3D::collideAndStream(), n times
 * (src/multiBlock/multiBlockLattice3D.hh:307-347). */
int bl_collide_stream(bl_lattice lat, int n_steps);
/* computeDensity / computeVelocity (u = j/rho + F/2)
 * (src/dataProcessors/dataAnalysisFunctional3D.hh:772-784,
 *  src/basicDynamics/externalForceDynamics.hh:59-73).  Either may be NULL. */
int bl_lattice_get_moments(bl_lattice lat, double *rho, double *u);
int bl_lattice_get_force(bl_lattice lat, double *force);
/* FP64 device sums over non-bounce-back cells: mass, jx, jy, jz. */
int bl_lattice_conserved(bl_lattice lat, double out[4]);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  The sums of bl_lattice_conserved, bitwise equal,
 * written to `out` (4 doubles of device memory, or host memory the device
 * can write) in the order of the lattice stream, without host
 * synchronisation and without allocating anything: the per-block partial
 * sums live in a static device array of the library (8 MiB + 32 B per
 * device, held until the process ends), which limits the lattice to
 * 2^26 = 67 108 864 cells (BL_ERR_CAPACITY beyond).  Every call on one
 * device shares that array, so calls that overlap on one device (lattices
 * stepping on different streams) race. */
int bl_lattice_conserved_static(bl_lattice lat, double *out);
int bl_lattice_steps(bl_lattice lat, long long *steps);
/* Launch shape of the collide-stream kernel: threads per block (64, 128,
 * 192 or 256, or the negative of one: see bl_lattice_set_launch_parity)
 * and XCD-aware block remapping on/off, for both AA step
 * parities, without nontemporal hints.  No counterpart in the reference
 * (its block decomposition is fixed, src/multiBlock/defaultMultiBlockPolicy3D.h:119-129). */
int bl_lattice_set_launch(bl_lattice lat, int block, int xcd_remap);
/* The same for one parity (0: even steps, which address each cell's own
 * slots; 1: odd steps, which address the neighbours').  nontemporal: 0 no
 * hints, 1 nontemporal loads and stores of every population, 2 only of the
 * planes the odd step addresses z-aligned (its z +-1 accesses are
 * misaligned; on even steps 2 is 1).  xcd_remap: 0 keeps the
 * hardware's block order (round-robin over the 8 XCDs), 1 gives each XCD
 * one contiguous range of cells, a power of two k >= 2 deals chunks of k
 * consecutive blocks round-robin over the XCDs, BL_ORDER_ROWS (3) gives
 * each z-row of cells (split into ceil(nz / block) blocks) to one XCD, the
 * eight XCDs taking adjacent rows.  A negative block -B runs the odd step
 * over z-segments: one workgroup of B lanes per B consecutive cells of a
 * z-row, every population access z-aligned (the z +-1 shift goes through
 * LDS; nontemporal 2 is then 1).  It applies where z is periodic, nz is a
 * multiple of B and the lattice holds fewer than 2^32 population values;
 * elsewhere, and on even steps, -B acts as B.  It takes xcd_remap 0, 1 or
 * a power of two (dealing segments as blocks), not BL_ORDER_ROWS.  Results
 * do not change. */
int bl_lattice_set_launch_parity(bl_lattice lat, int parity, int block, int xcd_remap, int nontemporal);
/* out = {block, xcd_remap, nontemporal} of one parity. */
int bl_lattice_get_launch(bl_lattice lat, int parity, int out[3]);
/* HIP stream the lattice works on (hipStream_t), for event timing. */
int bl_lattice_stream(bl_lattice lat, void **stream);
/* Runs n collide-stream launches (no boundary pass) and returns their total
 * device time in milliseconds, measured with HIP events on the lattice
 * stream. */
int bl_lattice_time_collide(bl_lattice lat, int n_steps, double *ms);
/* Waits for all work on the device (the reference is synchronous; this has
 * no counterpart there). */
int bl_sync(void);
/* Plain device copy of n_bytes, repeated n_iter times: bandwidth reference
 * and counter calibration (no counterpart in the reference).  *ms = total. */
int bl_device_copy(long long n_bytes, int n_iter, double *ms);

/* ------------------------------------------------ immersed boundary */
/* ImmersedBoundaryDynamics3D(lattice, shape_library)
 * (src/fsi/ImmersedBoundaryDynamics.h:57-58); the delta kernel is the
 * reference's compile-time Dirac typedef (src/fsi/ImmersedBoundaryDynamics.h:55). */
int bl_ibm_create(bl_lattice lat, int kernel, bl_ibm *out);
int bl_ibm_destroy(bl_ibm g);
/* ParticleShapeLibrary::store_mesh (src/fsi/ParticleShapeFactory.hh:118-129):
 * rest geometry (links, adjacent pairs, areas, angles) is computed in FP64. */
int bl_shape_add(bl_ibm g, int nv, int nt, const double *rest_pos, const int *tri,
                 int *shape_id);
/* Shape rest quantities: area, volume, center (3), radius. */
int bl_shape_info(bl_ibm g, int shape_id, double out[6]);
/* add_particle (src/fsi/ImmersedBoundaryDynamics.hh:508-525).  params:
 *   capsule models: {G, C}
 *   RBC: {shear_modulus, L0, k_area_global, k_area_local, k_volume,
 *         vol_desired, k_bend, theta0}  (src/fsi/RBCParticle.h:22-58)
 * pos: nv*3 initial vertex positions.  bug_compatible (capsules) keeps the
 * reference's c = l' sin(phi) / (l' sin(phi0)) term
 * (src/fsi/DeformableCapsule.hh:96); 0 uses l0' in the denominator. */
int bl_particle_add(bl_ibm g, int shape_id, int model, const double *params,
                    int nparams, const double *pos, int bug_compatible, int *pid);
/* init() (src/fsi/ImmersedBoundaryDynamics.hh:148-152): uploads particles,
 * maps every centre of mass onto the periodic grid. */
int bl_ibm_init(bl_ibm g);
/* The four stages of ImmersedBoundaryWrapperFunctional3D::process
 * (src/fsi/DataProcessors.hh:94-108), preceded by the computeVelocity pass
 * and the setExternalVector(force := body force) reset of the progs. */
int bl_ibm_interpolate(bl_ibm g);
int bl_ibm_compute_forces(bl_ibm g);
int bl_ibm_spread(bl_ibm g);
int bl_ibm_move(bl_ibm g);
/* All of the above (and the phase voxelization when omega_in != omega_out). */
int bl_ibm_step(bl_ibm g);
/* n coupled time steps: IBM step + collide-and-stream each (the loop body of
 * progs/capsule_in_shear/capsule_in_shear.cpp:187-227).  use_graph captures
 * one step in a hipGraph and replays it. */
int bl_coupled_steps(bl_ibm g, int n_steps, int use_graph);
/* Captures the graphs bl_coupled_steps(g, n_steps, 1) replays from the
 * current state, without running a step, so that a timed run replays
 * only (no counterpart in the reference).  BL_ERR_STATE while the next step
 * must run eagerly (after a loaded force field or a body-force change). */
int bl_coupled_prepare(bl_ibm g, int n_steps);
int bl_ibm_count(bl_ibm g, int *particles, int *vertices);
/* Vertex arrays of one particle (pid) or of all (pid = -1), nv*3 doubles. */
int bl_particle_get(bl_ibm g, int pid, int what, double *out); /* what: 0 pos, 1 vel, 2 force */
int bl_particle_set_positions(bl_ibm g, int pid, const double *pos);
/* Vertex state of one particle (pid) or of all (pid = -1): what 0
 * positions (as bl_particle_set_positions), 1 velocities, 2 forces; nv*3
 * doubles.  The vertex arrays of ParticleBase3D (src/fsi/ParticleBase.h:30-36). */
int bl_particle_set_state(bl_ibm g, int pid, int what, const double *data);
/* The first nparams (<= 8) model parameters of particle pid, in the order
 * of bl_particle_add (e.g. vol_desired of an RBC, RBCParticle::params(),
 * src/fsi/RBCParticle.h:70-71). */
int bl_particle_set_params(bl_ibm g, int pid, const double *params, int nparams);
/* RBCParticle::relax_nodes(damping) (src/fsi/RBCParticle.hh:318-342) for
 * particle pid (models 2 and 3): node mass = area/3 of the incident
 * triangles, vel += F / M - damping vel, pos += vel, geometry update, new
 * membrane forces; the forces in place on entry are those of the current
 * positions (bl_particle_membrane_forces). */
int bl_rbc_relax_nodes(bl_ibm g, int pid, double damping);
/* compute_forces of every particle's membrane alone (no decorators, no
 * particle-particle forces) after a geometry update: the particles'
 * compute_forces() (src/fsi/RBCParticle.hh:140-247,
 * src/fsi/DeformableCapsule.hh:60-150). */
int bl_particle_membrane_forces(bl_ibm g);
/* Vertex forces (nv*3 over all particles, in particle order) that every
 * later coupled step adds to each vertex's force after the wall
 * decorators and before the membrane forces, until replaced; NULL removes
 * them.  The force decorators of a prog that have no device form
 * (ForceDecorator::apply_force, src/fsi/ForceDecorator.h:12-15) are
 * evaluated on the host and arrive here. */
int bl_ibm_set_vertex_forces(bl_ibm g, const double *force);
/* area, volume, centre of mass (3), inertia tensor (9) of the enclosed
 * volume (src/fsi/DeformableParticle3D.hh:54-82,173-272). */
int bl_particle_observables(bl_ibm g, int pid, double out[14]);
/* The phase field of the last voxelization (ncell bytes, 0/1). */
int bl_ibm_get_phase(bl_ibm g, unsigned char *phase);
/* Time of each IBM stage over the last bl_ibm_step, ms: stencil binning,
 * membrane forces, velocity staging + spreading + interpolation with
 * advection, particle geometry + centre-of-mass remap, voxelization. */
int bl_ibm_stage_times(bl_ibm g, double out[5]);
/* Tiles touched by the last binning and (tile, vertex) list entries
 * (diagnostics for the traffic model; no counterpart in the reference). */
int bl_ibm_tile_stats(bl_ibm g, long long out[2]);
/* Launch structure of bl_coupled_steps (no counterpart in the reference,
 * whose order it keeps): mode 1 carries the velocity staging (the progs'
 * computeVelocity pass, restricted to the touched tiles) in the
 * collide-and-stream launch, which then runs before the interpolation and
 * advection of the same step; 0 stages in a pass of its own before the
 * spreading; -1 (default) picks 1 for 65 536 vertices or more.  Without a
 * phase field only (omega_in == omega_out); otherwise the separate pass is
 * used.  The staged values are the same either way. */
int bl_ibm_set_fused_staging(bl_ibm g, int mode);
/* Stream layout of bl_coupled_steps without fused staging (no counterpart
 * in the reference, whose order it keeps): mode 1 runs a step on the
 * lattice stream alone, pairing stages that do not read each other's
 * output in one launch (binning | element forces, force gather | velocity
 * staging, spreading | interpolation); 0 forks the membrane forces and the
 * spreading onto a second stream; -1 (default) picks 1 below 65 536
 * vertices.  The results are the same either way. */
int bl_ibm_set_layered(bl_ibm g, int mode);
/* Launch parameters of the immersed-boundary kernels (no counterpart in
 * the reference, whose loops are serial), indexed by BL_IBM_*:
 *   BL_IBM_BIN_BLOCK     threads per binning workgroup (64, 128, 256;
 *                        also the paired binning | element-force launch)
 *   BL_IBM_ELEM_BLOCK    threads per element-force workgroup (64..256)
 *   BL_IBM_GATHER_LANES  lanes per vertex of the force gather (2, 4, 8)
 *   BL_IBM_INTERP_BLOCK  threads per interpolation workgroup (64..256)
 *   BL_IBM_SPREAD_SMALL  small-tile spreading waves per 4 vertices (1..8)
 *   BL_IBM_SPREAD_BIG    vertices per big-tile spreading workgroup (16..128)
 *   BL_IBM_XCD           workgroup order over the XCDs of the vertex,
 *                        element and small-tile launches: chunks of 2^k
 *                        consecutive workgroups on one XCD (k = 0, 2, 4, 6;
 *                        0 the hardware's order)
 * Defaults {256, 256, 4, 256, 4, 32, 0}.  Only the gather lane count changes
 * an order of summation (its butterfly): results agree to rounding across
 * lane counts; every other parameter leaves them bitwise the same.
 * bl_ibm_set_launch refuses values outside those sets (BL_ERR_ARG). */
#define BL_IBM_BIN_BLOCK 0
#define BL_IBM_ELEM_BLOCK 1
#define BL_IBM_GATHER_LANES 2
#define BL_IBM_INTERP_BLOCK 3
#define BL_IBM_SPREAD_SMALL 4
#define BL_IBM_SPREAD_BIG 5
#define BL_IBM_XCD 6
#define BL_IBM_NLAUNCH 7
int bl_ibm_set_launch(bl_ibm g, const int params[BL_IBM_NLAUNCH]);
int bl_ibm_get_launch(bl_ibm g, int params[BL_IBM_NLAUNCH]);

/* ------------------------------------------------ initial conditions */
/* shrink_rbc_volume (progs/common/calibrate_rbc_shape.h:13-72) on the
 * device, for particle pid or (pid = -1) every particle with a volume
 * constraint (models 2 and 3) at once: vol_desired ramps from the rest
 * shape's volume to vol_final over n_it steps of RBCParticle::relax_nodes
 * (src/fsi/RBCParticle.hh:318-342: node mass = area/3 of the incident
 * triangles, vel += F/M - damping vel, pos += vel, membrane forces), 2 n_it
 * + 1 steps in all; a particle whose area turns NaN starts again from its
 * starting state with twice the damping (an error once the damping would
 * exceed 2).  The particles keep the calibrated positions, velocities,
 * forces and vol_desired; damping_out (np entries, may be null) receives
 * the damping each particle finished with.  Needs bl_ibm_init; the lattice
 * is not touched. */
int bl_rbc_shrink_volume(bl_ibm g, int pid, double vol_final, int n_it, double damping, double *damping_out);
/* ParticlePositionInitializer (src/fsi/ParticlePositionInitializer.h:19-37,
 * .hh:18-86), host only: spheres placed one after another in the box
 * (x0, x1, y0, y1, z0, z1), inside the boundary (may be null) and apart from
 * every sphere placed before, drawn with the C library's rand() as the
 * reference draws them (seed it with srand for a given sequence).
 * generate fails with BL_ERR_ARG when 2 radius exceeds 2 max_radius, with
 * BL_ERR_CAPACITY when a point finds no room in 5e6 draws (the points
 * placed so far stay); get returns the points of one tag in placement
 * order (out may be null to ask for the count). */
int bl_positions_create(const int box[6], const bl_boundary *boundary, double max_radius, bl_positions *out);
int bl_positions_generate(bl_positions h, long long n, double radius, char tag);
int bl_positions_get(bl_positions h, char tag, double *out, long long cap, long long *n);
int bl_positions_destroy(bl_positions h);

/* ---------------------------------------------------- on-disk formats */
/* ParticleShape tag (src/fsi/ParticleShape.h:64,118): the shape descriptor
 * the lightweight output names. */
int bl_shape_set_tag(bl_ibm g, int shape_id, const char *tag);
/* ParticleBase3D::pack of one particle (src/fsi/ParticleBase.hh:218-224,
 * src/fsi/DeformableParticle3D.hh:277-287, src/fsi/DeformableCapsule.hh:
 * 217-223, src/fsi/RBCParticle.hh:346-350): type id (capsule 0, RBC 1, the
 * reference build's registration order), shape id, particle id, should-
 * voxelize, pos/vel/force per vertex, then the model's parameters.  A model
 * the reference types cannot hold (3, or a capsule without the c-term quirk)
 * gets type id 100 and model, flag and 8 parameters.  buf may be null to ask
 * for the length. */
int bl_particle_pack(bl_ibm g, int pid, char *buf, long long cap, long long *len);
/* Shape, membrane model and the reference's particle id (get_id,
 * src/fsi/ParticleBase.h) of particle index pid: the index for particles
 * added with bl_particle_add, the stored id for particles loaded from a
 * checkpoint (kept in increasing id order). */
int bl_particle_describe(bl_ibm g, int pid, int *shape_id, int *model, long long *id);
/* ImmersedBoundaryDynamics3D::save_checkpoint / load_checkpoint
 * (src/fsi/ImmersedBoundaryDynamics_io.hh:17-158) for one rank: header
 * <folder>/<name>.plb (particle count, 1, "<name>/0.plb") and the rank file
 * <folder>/<name>/0.plb (plint count, packs); the folder <folder>/<name>
 * must exist.  Loading replaces the particles (before bl_ibm_init; the
 * shapes must have been added in the same order) and keeps their vertex
 * velocities and forces for the upload at init. */
int bl_ibm_save_checkpoint(bl_ibm g, const char *folder, const char *name);
int bl_ibm_load_checkpoint(bl_ibm g, const char *folder, const char *name);
/* ParticleBase3D::write_vtk (src/fsi/ParticleBase.hh:270-398): VTU with
 * appended raw Float32 points and force_tot, Int32 cells. */
int bl_particle_write_vtk(bl_ibm g, int pid, const char *path);
/* ParticleBase3D::write_gmsh (src/fsi/ParticleBase.hh:400-437). */
int bl_particle_write_gmsh(bl_ibm g, int pid, const char *path);
/* write_lightweight_particle_data (src/fsi/ImmersedBoundaryDynamics_io.hh:
 * 178-200, src/fsi/DeformableParticle3D.hh:318-333): "iteration:<it>", then
 * per particle tag:id:com(3):gyration(xx,xy,xz,yy,yz,zz):volume. */
int bl_ibm_write_lightweight(bl_ibm g, const char *path, long long iteration);
/* Lattice state for a restart in this library's own layout (the
 * reference's parallelIO::save / load of the lattice,
 * progs/rbc_in_pipe/rbc_in_pipe.cpp:321-322,693, writes Palabos' format,
 * which is not restated): sizes, storage, step count, the populations as
 * stored and the external force of every cell.  Loading needs a lattice of
 * the same size and storage; the continuation is bitwise that of the saved
 * run. */
int bl_lattice_save_checkpoint(bl_lattice h, const char *path);
int bl_lattice_load_checkpoint(bl_lattice h, const char *path);

/* ---------------------------------------------------- curved walls */
/* GuoRigidWallBoundary(boundary, lattice) (src/fsi/GuoBoundaryCondition.hh:
 * 17-22, instantiator :62-135): the wall nodes are the cells outside the
 * boundary with a lattice link that crosses it; each keeps the link most
 * opposed to the wall normal (direction index 1..18 towards the fluid) and
 * delta = 1 - t of its crossing.  Fails (BL_ERR_ARG) if a node's two
 * neighbours along that link leave the lattice across a non-periodic face.
 * Nodes whose neighbours are wall nodes themselves (lattice nodes on the
 * boundary surface) keep the reference's one-after-the-other order. */
int bl_guo_wall_create(bl_lattice lat, const bl_boundary *b, bl_guo_wall *out);
int bl_guo_wall_destroy(bl_guo_wall w);
int bl_guo_wall_count(bl_guo_wall w, long long *n);
/* The node list in instantiation order (x, y, z loops); any pointer may be
 * NULL.  cell: (x*ny + y)*nz + z; wall_pos: 3 per node. */
int bl_guo_wall_nodes(bl_guo_wall w, long long *cell, int *dir, double *delta, double *wall_pos);
/* set_wall_velocity(Array) / set_wall_velocity(function) (:44-59): one
 * velocity for every node, or 3 per node (e.g. evaluated at wall_pos). */
int bl_guo_wall_set_velocity(bl_guo_wall w, const double vel[3]);
int bl_guo_wall_set_node_velocities(bl_guo_wall w, const double *vel);
/* insert() (:24-28): the extrapolation (GuoRigidWallBoundaryFunctional,
 * :138-199, second-order branch) runs after every collide-and-stream of the
 * lattice, and once in bl_lattice_initialize, as an integrated processor
 * does. */
int bl_guo_wall_insert(bl_guo_wall w);
/* defineDynamics(..., BoundaryOutside(boundary, margin), value)
 * (src/fsi/Boundary.h:114-128): mask[c] = value for every cell with
 * !contains(c, -margin).  The reference passes sqrt(2) to an integer
 * margin, i.e. 1.  Host only. */
int bl_boundary_outside_mask(const bl_boundary *b, int nx, int ny, int nz, double margin,
                             unsigned char *mask, unsigned char value);
/* ImmersedBoundaryDynamics3D::set_boundary (src/fsi/ImmersedBoundaryDynamics.h:93):
 * vertices nearer the boundary than the kernel's half support + 1 blend the
 * kernel with the top hat in interpolation and spreading
 * (src/fsi/ImmersedBoundaryDynamics_interpolation.hh:182-230,
 * src/fsi/ImmersedBoundaryDynamics_forces.hh:193-237) and are spread after
 * every other vertex; advected vertices reflect off the boundary
 * (src/fsi/DeformableParticle3D.hh:26-50).  NULL: no boundary. */
int bl_ibm_set_boundary(bl_ibm g, const bl_boundary *b);

/* ------------------------------------------------ interaction forces */
/* Interaction potentials (src/fsi/Potentials.h): MorsePotential(r_min,
 * beta, strength), force 2 beta strength x (x - 1) with
 * x = exp(beta (r_min - max(d, 0))), cut-off r_min (:28-56);
 * SpringPotential(min_dist, strength), force (min_dist - d) strength,
 * cut-off min_dist (:58-72). */
#define BL_POTENTIAL_MORSE 0  /* p = {r_min, beta, strength} */
#define BL_POTENTIAL_SPRING 1 /* p = {min_dist, strength, -}  */
typedef struct {
    int kind;
    double p[3];
} bl_potential;
/* The potential's force at distance d and its cut-off (host only). */
int bl_potential_eval(const bl_potential *u, double d, double *force, double *cutoff);
/* add_force_decorator(new WallInteraction(boundary, potential))
 * (src/fsi/ImmersedBoundaryDynamics.h:78, src/fsi/ForceDecorator.hh:27-35):
 * before the membrane forces of every step, each vertex at distance
 * d <= cut-off from the boundary gets force -= F(d) n.  Decorators apply
 * in the order added; at most 4. */
int bl_ibm_add_wall_interaction(bl_ibm g, const bl_boundary *b, const bl_potential *u);
/* add_pp_force(new PPPotentialForce(potential))
 * (src/fsi/ImmersedBoundaryDynamics.h:79, src/fsi/ImmersedBoundaryDynamics_forces.hh:
 * 287-302,332-375): after the membrane forces, vertices of different
 * particles interact through the collision handler
 * (src/fsi/CollisionForces.h:84-161) over the lattice box: unit cells,
 * pairs within the cell and with its 7 forward neighbours, and the
 * handler's face / edge / corner pairs across periodic axes.  Vertices
 * outside the box enlarged by the cut-off do not take part.  The cut-off
 * must not exceed the kernel's half support - 0.5 (BL_ERR_ARG, where the
 * reference exits).  At most 4. */
int bl_ibm_add_pp_force(bl_ibm g, const bl_potential *u);
/* compute_collision_forces(potential, boundary)
 * (src/fsi/ImmersedBoundaryDynamics_forces.hh:505-521): vertex forces :=
 * the collision handler's pair forces alone, plus, with a boundary (NULL:
 * none), its wall forces (src/fsi/CollisionForces.h:167-187) on vertices
 * in unit cells that the boundary, widened by the cut-off, crosses. */
int bl_ibm_compute_collision_forces(bl_ibm g, const bl_potential *u, const bl_boundary *b);

/* ------------------------------------------ rigid particles, packing */
/* A RigidParticle3D (src/fsi/RigidParticle.hh:13-39) of shape shape_id:
 * state = {com x, y, z, quaternion w, x, y, z, scale, density, damping,
 * relaxation factor}; velocity and angular velocity start at zero, the
 * vertices at com + R(q) (scale v_rest).  The rest-shape inertia is
 * ParticleShape's (src/fsi/ParticleShape.hh:249-360, about the mesh
 * origin).  Like bl_particle_add, before bl_ibm_init, whose periodic mapping
 * moves the centre of mass (and the vertices) onto the grid.  A boundary
 * holding rigid particles refuses the fluid-coupling stages (bl_ibm_step,
 * bl_coupled_steps, ...; BL_ERR_STATE): RigidParticle3D::compute_forces is
 * not implemented. */
int bl_rigid_add(bl_ibm g, int shape_id, const double state[11], int *pid);
/* move_vertices() of every rigid particle (src/fsi/ImmersedBoundaryDynamics.hh:
 * 454-493 with RigidParticle3D::move_vertices, src/fsi/RigidParticle.hh:57-98)
 * under the current vertex forces: force and torque summed over the
 * vertices in order, velocity, angular velocity, centre of mass and
 * orientation advanced, vertices regenerated, centre of mass mapped onto the
 * periodic grid with the vertices. */
int bl_rigid_move(bl_ibm g);
/* p->scale() = scale; p->update() for rigid particle pid (-1: all). */
int bl_rigid_set_scale(bl_ibm g, int pid, double scale);
/* out = {com (3), velocity (3), angular velocity (3), quaternion (4),
 * scale, density, damping, relaxation factor} of rigid particle pid. */
int bl_rigid_get(bl_ibm g, int pid, double out[17]);
/* The growth loop of create_particles (progs/common/create_particles.h:
 * 129-155) on the device: for it = 0..n_iter, every `every` iterations the
 * scale of all rigid particles becomes initial_scale + (1 - initial_scale)
 * it / n_iter (vertices regenerated); then bl_ibm_compute_collision_forces(u,
 * b) and bl_rigid_move.  The boundary's particles must all be rigid for the
 * packing to mean the reference's. */
int bl_rigid_grow(bl_ibm g, int n_iter, double initial_scale, int every, const bl_potential *u,
                  const bl_boundary *b);

/* ----------------------------------------------------------- ensembles */
/* No counterpart in the reference, which runs independent boxes as separate
 * programs (one rbc_in_shear run per box).  Groups n_boxes lattices (each
 * with its immersed boundary, or NULL / ibms == NULL for lattice-only boxes)
 * created and initialised through the entry points above; flags must be 0.
 * The handles stay owned by the caller and must outlive the ensemble. */
int bl_ensemble_create(int n_boxes, const bl_lattice *lattices, const bl_ibm *ibms, int flags,
                       bl_ensemble *out);
int bl_ensemble_destroy(bl_ensemble e);
/* n coupled steps of every box.  use_graph: two steps of all boxes are
 * captured once in one hipGraph (each box a parallel branch) and replayed. */
int bl_ensemble_steps(bl_ensemble e, int n_steps, int use_graph);
/* How bl_ensemble_steps runs graphs (no counterpart in the reference):
 * mode 1 batched, every stage of a step one launch over all boxes (each
 * box's workgroups with its own arguments), an error if the boxes do not
 * allow it; 0 one graph branch per box; -1 (default) batched when the
 * boxes allow it (same storage and AA parity, paired-stage immersed
 * boundary steps, no Guo walls, phase fields or particle-particle forces),
 * else branches.  Both give the values of the boxes stepped alone. */
int bl_ensemble_set_batched(bl_ensemble e, int mode);
/* 1 if the ensemble's captured steps are the batched ones. */
int bl_ensemble_is_batched(bl_ensemble e, int *batched);
/* The project's tuner for batched ensembles (no counterpart in the
 * reference): times n_steps (even) graph-replayed steps for each workgroup
 * order of the batched collide-and-stream launch (the modes of
 * bl_lattice_set_launch_parity except BL_ORDER_ROWS), per AA parity in
 * turn, then for each nontemporal level (0, 1, 2) of that launch with the
 * orders kept, and keeps the fastest.  The trials are real steps (the
 * report says how many); every order and level gives the same values. */
int bl_ensemble_tune(bl_ensemble e, int n_steps, char *report, int cap);
/* Sets the batched collide's workgroup order (0, 1 or a power of two, as
 * bl_lattice_set_launch_parity) and nontemporal level (0, 1, 2; on even
 * steps, which move no population along z, 2 acts as 1) for steps of AA
 * parity `parity` without tuning (no counterpart in the reference);
 * the next steps capture with it.  Every setting gives the same values. */
int bl_ensemble_set_launch(bl_ensemble e, int parity, int order, int nontemporal);
/* The same, returning the device time of the n steps in milliseconds. */
int bl_ensemble_time_steps(bl_ensemble e, int n_steps, int use_graph, double *ms);

/* ------------------------------------------------------------- tuning */
/* No counterpart in the reference (its block decomposition is fixed,
 * src/multiBlock/defaultMultiBlockPolicy3D.h:119-129).  Times the
 * collide-and-stream kernel of this lattice for every launch shape per AA
 * step parity (block 64..256, the block orders and nontemporal levels of
 * bl_lattice_set_launch_parity: 96 odd candidates, and on odd steps the
 * z-segment kernel for every segment length the lattice takes, in hardware
 * order and chunks of 4, 16 and 64, with and without nontemporal hints),
 * times the three fastest of each parity again (8 launch pairs each, the
 * other parity at its fastest), keeps the fastest of those, restores the
 * populations and step count the trials advanced, and records the result in
 * the process's tuning table (a lattice of the same size, storage and
 * device created later starts with it).  report (cap bytes, may be NULL)
 * receives a JSON summary. */
int bl_tune_collide(bl_lattice lat, char *report, int cap);
/* Times n_steps (even) coupled steps of each launch structure
 * (bl_ibm_set_fused_staging 0/1 without a phase field, bl_ibm_set_layered
 * 0/1), keeps the fastest and records it in the tuning table (key: device,
 * lattice size, storage, vertex count).  Every layout gives the same
 * results; the trials are real steps: the simulation advances by
 * 2 n_steps per layout tried (the report says how many). */
int bl_tune_coupled(bl_ibm g, int n_steps, char *report, int cap);
/* Times n_steps (even) coupled steps for the immersed-boundary launch
 * parameters of bl_ibm_set_launch, one parameter at a time over its
 * allowed values with the best of the others kept (coordinate descent,
 * two passes, each value other than the parameter's value at the start of
 * its sweep tried once: always 35 runs of 2 n_steps), keeps the fastest and
 * records it in the tuning table entry of bl_tune_coupled's key.  The
 * trials are real steps (the report says how many); with the gather lane
 * count the trajectory changes at rounding. */
int bl_tune_ibm(bl_ibm g, int n_steps, char *report, int cap);
/* Applies the tuning table's launch structure and immersed-boundary launch
 * parameters for this boundary, if any. */
int bl_tune_apply_coupled(bl_ibm g);
/* The tuning table to / from a file, one JSON object per line; loading
 * merges (an entry of the same key replaces the one held) and refuses
 * unreadable lines. */
int bl_tune_cache_save(const char *path);
int bl_tune_cache_load(const char *path);

#ifdef __cplusplus
}
#endif
#endif
