// Device-resident D3Q19 lattice: structure-of-arrays populations in an
// in-place AA pattern, cell mask, tile-gated external force, phase.
#pragma once

#include <vector>

#include "common.hpp"

struct bl_lattice_s {
    int nx = 0, ny = 0, nz = 0;
    long ncell = 0;
    long stride = 0;  // elements between population planes (multiple of 64)
    int periodic[3] = {1, 1, 1};
    int storage = BL_STORE_F64;
    size_t pop_bytes = 0;  // bytes per stored population value
    void* pops = nullptr;  // Q * stride values of the storage type

    unsigned char* mask = nullptr;   // ncell, or null (all fluid)
    bool has_vwall = false;
    // Inert cells (bl_lattice_set_skip_inert): bounce-back cells whose 18
    // neighbours are all bounce-back (or past a non-periodic face).  With
    // skip_inert set, one bit per cell (bit c & 63 of word c >> 6), built
    // from the mask, and the collide-and-stream leaves those cells alone.
    bool skip_inert = false;
    unsigned long long* inert = nullptr;
    long n_inert = 0;
    int wall_axis = 1;
    double wall_u[6] = {0, 0, 0, 0, 0, 0};

    double omega_out = 1.0, omega_in = 1.0;
    double omega_c0 = 1.0, omega_c1 = 1.0;  // omega for phase 0 / 1
    unsigned char* phase = nullptr;         // ncell, or null (phase 0): own_phase or an IBM's voxelization
    unsigned char* own_phase = nullptr;     // the phase set through the lattice (owned here)

    // external force: body force g everywhere except in tiles whose flag is
    // set, where force[d*stride + c] holds the full value.
    double g[3] = {0, 0, 0};
    // body force of the last collision: the value the untouched cells held
    // when the progs' computeVelocity pass of the next step reads the field
    // (the force reset to g comes after it); g itself until the first
    // collision
    double g_last[3] = {0, 0, 0};
    bool collided = false;
    void* force = nullptr;  // 3 * stride of double (F64) or float
    size_t force_bytes = 0;
    unsigned char* tile_flags = nullptr;
    int ntx = 0, nty = 0, ntz = 0;
    long ntiles = 0;
    bool force_field = false;  // tiles may carry a field
    bool flags_full = false;   // every tile flag was set (a full field was loaded); cleared by its consumer

    // Velocity staging carried by the next collide-and-stream (set by an
    // immersed boundary around one launch): on the cells of the tiles flagged
    // in tile_flags, u = j/rho + F/2 of the pre-collision populations with
    // the force of emit_force / emit_flags (the last collision's field) goes
    // to emit_u, one aligned 4-vector of the compute type per cell.
    void* emit_u = nullptr;
    const void* emit_force = nullptr;
    const unsigned char* emit_flags = nullptr;

    long long steps = 0;  // completed collide-and-stream steps (AA parity)
    int device = 0;       // the device current at creation (holds every buffer)
    // Bumped by every change that a captured step bakes into its kernel
    // arguments or launch choices (omega, body force, wall velocities, mask,
    // phase and force-field buffers, launch shape, inserted walls); captured
    // graphs keyed on an older value are captured again.
    unsigned long long gen = 0;
    hipStream_t stream = nullptr;

    // launch shape of the collide-and-stream kernel per step parity (index 0
    // even, 1 odd): threads per block, XCD-aware block order, nontemporal
    // population loads and stores (even step only; set by the tuner or a
    // cached tuning result)
    int block[2] = {256, 256};
    int xcd_remap[2] = {1, 1};
    int nontemporal[2] = {0, 0};

    // scratch
    double* d_stage = nullptr;  // ncell * 19 doubles, allocated on demand
    double* d_partials = nullptr;
    double* d_wall = nullptr;   // velocities read by the velocity-wall finite differences
    long nwall = 0;
    bl_ibm ibm = nullptr;
    std::vector<bl_guo_wall> guo;  // inserted Guo curved walls (not owned)
};

namespace bl {

// n / d for 0 <= n < 2^31 without a hardware divide: q = (t + ((n - t) >> 1))
// >> (l - 1) with t = mulhi(m, n), l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1 (Granlund & Montgomery, PLDI 1994, sec. 4).
struct FastDiv {
    unsigned m = 0;
    int l = 0;
};
FastDiv make_fastdiv(unsigned d);

struct LatView {
    int nx, ny, nz;
    FastDiv dy, dz;  // division by ny and by nz
    unsigned ncell;
    long stride;
    int per[3];
    void* pops;
    const unsigned char* mask;
    const unsigned long long* inert;  // inert-cell bits, or null (every cell steps)
    const unsigned char* phase;
    const unsigned char* flags;
    const void* force;
    double g[3];
    double g_last[3];  // body force of the last collision (read by the velocity staging)
    double om0, om1;
    double om_base;  // omega_out, as Dynamics::getOmega() reports it
    int ntx, nty, ntz;
    int wall_axis;
    double wall_u[6];
    int xcd_remap;
    int row_segs;  // blocks per z-row under BL_ORDER_ROWS
    void* emit_u;  // velocity staging carried by the collision (see bl_lattice_s)
    const void* emit_force;
    const unsigned char* emit_flags;
};

LatView view_of(const bl_lattice_s* L);

// One collide-and-stream launch of a lattice, as collide_stream_launch
// would issue it (collide_plan reads the lattice's state, collide_commit
// advances it: step count, last body force).
struct CollidePlan {
    LatView v;
    bool odd = false, mask = false, phase = false;
    int fm = 0, nt = 0, blk = 256, storage = 0;
    unsigned nblk = 0;
};
CollidePlan collide_plan(const bl_lattice_s* L);
// Whether the lattice meets the conditions of the z-segment kernel with
// segments of B lanes (a launch block of -B).
bool zseg_ok(const bl_lattice_s* L, int B);
void collide_commit(bl_lattice_s* L);
// Batches of boxes (ensembles): one launch over the boxes' views, box k
// taking workgroups [base[k], base[k + 1]) (device arrays); the plans must
// agree on everything collide_batchable checks.
bool collide_batchable(const CollidePlan& a, const CollidePlan& b);
// `order`: how the batch's workgroups are dealt over the XCDs (the modes
// of bl_lattice_set_launch_parity except BL_ORDER_ROWS); `emit`: the views
// carry velocity staging (force-field plans only); `nt`: the nontemporal
// level (0, 1, 2) of bl_lattice_set_launch_parity
int collide_batch_launch(const CollidePlan& p0, bool emit, const LatView* views, const int* base, int nbox,
                         unsigned nblk, int order, int nt, hipStream_t s);
// velocity-wall cells of one box: its view, wall-cell count and the buffer
// of the velocities its finite differences read
struct VwallPlan {
    LatView v;
    long nwall;
    double* buf;
};
// base / nblk: workgroups over the wall cells, base_u / nblk_u over twice
// as many (the neighbour-velocity pass)
int vwall_batch_launch(int storage, bool swapped, const VwallPlan* tab, const int* base, const int* base_u, int nbox,
                       unsigned nblk, unsigned nblk_u, hipStream_t s);
int ensure_force_field(bl_lattice_s* L);
int collide_stream_launch(bl_lattice_s* L);
int velocity_walls_launch(bl_lattice_s* L);  // post-stream passes: velocity walls, then Guo walls
int guo_walls_launch(bl_lattice_s* L);
// Launch shapes of an earlier tuning of a lattice of this size on this
// device (tune.cpp), if the tuning table holds one.
void tune_cache_apply(bl_lattice_s* L);
void guo_walls_detach(bl_lattice_s* L);

}  // namespace bl
