// Immersed-boundary state: shapes (rest geometry), particles, vertex arrays,
// tile binning of the delta-kernel stencils, voxelization buffers.
#pragma once

#include <vector>

#include "lattice.hpp"
#include "mesh.hpp"

namespace bl {
struct PPRec;
}

struct bl_ibm_s {
    bl_lattice lat = nullptr;
    int kernel = BL_KERNEL_ROMA;
    int S = 3;        // stencil width
    double h = 1.5;   // half support

    std::vector<bl::ShapeHost> shapes;
    struct PartHost {
        int shape = 0, model = 0, bug = 1;
        long long id = 0;  // the reference's particle id (the index unless loaded from a checkpoint)
        double params[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        std::vector<double> pos;
        std::vector<double> vel, force;  // from a checkpoint (else zero at init)
        double rigid[32] = {};  // BL_MODEL_RIGID: state, then the rest-shape inertia and volume (rigid.hip)
    };
    std::vector<PartHost> parts;
    bool inited = false;

    int np = 0;
    long NV = 0, NT = 0, NL = 0, NP = 0;  // totals over particles
    int nshape = 0;

    // ---- shape data (device, concatenated over shapes) ----
    int* s_voff = nullptr;  // nshape+1 (rest vertices, CSR heads)
    int* s_toff = nullptr;  // nshape+1
    int* s_loff = nullptr;
    int* s_poff = nullptr;
    int* s_vtoff = nullptr;  // nshape+1 offsets into CSR entry arrays
    int* s_vloff = nullptr;
    int* s_vpoff = nullptr;
    int* s_tri = nullptr;        // 3 per tri (local ids)
    double* s_tri_area0 = nullptr;
    double* s_tri_geo0 = nullptr;  // 4 per tri
    double* s_rest = nullptr;      // 3 per vertex
    int* s_link = nullptr;         // 2 per link
    double* s_link_len = nullptr;
    int* s_pair = nullptr;         // 4 per pair
    double* s_pair_cs0 = nullptr;  // 2 per pair
    int* s_vt_head = nullptr;      // per shape vertex: nv+1 heads (relative)
    int* s_vt = nullptr;
    int* s_vl_head = nullptr;
    int* s_vl = nullptr;
    int* s_vp_head = nullptr;
    int* s_vp = nullptr;
    double* s_area0 = nullptr;  // per shape total rest area

    // ---- particles ----
    int* p_shape = nullptr;
    int* p_model = nullptr;
    int* p_bug = nullptr;
    double* p_params = nullptr;  // 8 per particle
    long* p_voff = nullptr;      // np+1
    long* p_toff = nullptr;
    long* p_loff = nullptr;
    long* p_poff = nullptr;
    double* p_geom = nullptr;    // 11 per particle: area, volume, com(3), bbox(6)
    double* p_delta = nullptr;   // 3 per particle (COM remap shift)
    int* v_part = nullptr;       // particle of each vertex, triangle, pair, link
    int* v_inc = nullptr;        // per vertex: incidence list starts and counts, particle (8 ints)
    int* t_part = nullptr;
    int* r_part = nullptr;
    int* l_part = nullptr;

    // ---- vertices ----
    // Vertex state in `vreal`-byte reals: 8 (double) with FP64 populations
    // or rigid particles, 4 (float) with FP32 and FP16-shifted populations
    // (fixed at bl_ibm_init).  Host copies and the C-ABI stay FP64.
    int vreal = 8;
    void* pos = nullptr;
    void* vel = nullptr;
    void* force = nullptr;
    void* ec = nullptr;  // per-element force contributions: 9 per triangle, 12 per pair, 3 per link
    long ec_pair = 0, ec_link = 0;  // offsets of the pair and link sections in ec
    void* ut = nullptr;    // staged velocity, (u, 0) per cell (FP64 for FP64 storage, else FP32)
    bool binned = false;   // tile lists match the current positions
    bool bins_dirty = false;  // tile counters hold a binning no spreading has consumed
    bool prepped = false;     // the next binning's list reset is already enqueued

    // ---- near-wall blending and reflection (a boundary is set) ----
    bool has_boundary = false;
    bl_boundary boundary{};
    unsigned char* v_near = nullptr;  // per vertex: nearer than half support + 1
    void* v_blend = nullptr;          // per vertex: distance - half support (vreal)
    int* v_base_t = nullptr;          // per vertex: top-hat stencil base (3)
    void* v_w_t = nullptr;            // per vertex: top-hat weights (2 per axis, vreal)

    // ---- rigid particles (BL_MODEL_RIGID, rigid.hip) ----
    int n_rigid = 0;
    double* r_state = nullptr;  // 32 per particle (rigid ones only filled)

    // ---- interaction forces ----
    struct WallDeco {
        bl_boundary b;
        bl_potential u;
    };
    std::vector<WallDeco> wall_deco;  // wall-interaction decorators, in order
    std::vector<bl_potential> pp;     // particle-particle potentials, in order
    int* pp_head = nullptr;   // per collision cell: the last vertex pushed onto its list, -1 if empty
    int* pp_first = nullptr;  // per collision cell: its smallest vertex id, -1 if empty
    int* pp_next = nullptr;   // per vertex: next vertex of its cell's list
    int* pp_cell = nullptr;   // per vertex: collision cell, -1 if the vertex does not take part
    bl::PPRec* pp_rec = nullptr;  // per vertex: position, owner, successor in its cell (ascending id)
    void* pp_force = nullptr;     // per vertex: the step's particle-particle forces (3, vreal)
    void* ext_force = nullptr;    // per vertex: forces given by the host (bl_ibm_set_vertex_forces), 3 (vreal)
    bool has_ext = false;
    long pp_ncell = 0;

    // ---- stencil / tiles ----
    int* v_base = nullptr;    // 3 per vertex
    void* v_w = nullptr;      // 12 per vertex (3 axes x up to 4 weights, vreal)
    int* tile_count = nullptr;  // entries per tile (cleared by the spreading that consumes them)
    int* tile_list = nullptr;   // per tile a fixed-capacity list of the vertices whose stencil touches it
    int* active = nullptr;      // two lists of touched tiles, one per force-flag buffer, then the two
                                // lists of their tiles with more than 64 entries (ntiles each)
    int* n_active = nullptr;    // [0], [1] lengths of the two lists, [2] overflow flags, [4] arrival counter
                                // of the list reset, [5], [6] (tile, vertex) pairs of the two lists,
                                // [8], [9] lengths of the two big-tile lists
    int alt = 0;                // list and flag buffer the next binning writes (flags_alt)
    int last_list = 0;          // list of the last binning
    bool alt_full = false;      // flags_alt has every tile set: the next binning clears all of it
    unsigned char* flags_alt = nullptr;  // next step's tile flags (swapped with the lattice's)
    void* force_alt = nullptr;           // next step's force field (swapped with the lattice's)
    int* ch_part = nullptr;              // per geometry chunk: particle
    int* ch_idx = nullptr;               //                     chunk index within the particle
    int* p_choff = nullptr;              // np+1 chunk offsets
    int nchunk = 0;
    double* ch_partial = nullptr;        // 11 per chunk
    int* ch_tab = nullptr;               // per chunk: first triangle, triangles, bbox vertex range (4)
    int* t_vert = nullptr;               // per triangle: its three global vertex indices

    // ---- voxelization ----
    int* col_count = nullptr;
    int* col_off = nullptr;
    int* col_fill = nullptr;
    double* col_z = nullptr;
    long col_cap = 0;
    int* n_hits = nullptr;
    unsigned char* phase = nullptr;

    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double stage_ms[5] = {0, 0, 0, 0, 0};
    bool timing = false;
    // launch parameters of the immersed-boundary kernels (bl_ibm_set_launch,
    // chosen by bl_tune_ibm), indexed by BL_IBM_*: threads per workgroup of
    // the binning, element forces and interpolation, lanes per vertex of the
    // force gather, small-tile spreading waves per 4 vertices, vertices per
    // big-tile spreading workgroup, workgroup order over the XCDs
    int lp[BL_IBM_NLAUNCH] = {256, 256, 4, 256, 4, 32, 0};
    int fused_staging = -1;  // velocity staging carried by the collision: -1 auto (large vertex counts), 0 off, 1 on
    int layered = -1;        // stages paired on one stream: -1 auto (small vertex counts), 0 off (forked), 1 on

    hipStream_t aux[2] = {nullptr, nullptr};  // forked stages of a step (default priority)
    hipEvent_t fev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

    // captured runs of coupled steps: [0] two steps, [1] kLongRun steps,
    // each with the step parity / buffer state it was captured in
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    long long graph_parity[2] = {-1, -1};
    unsigned long long graph_gen[2][2] = {{0, 0}, {0, 0}};  // lattice and boundary generations of each capture
    // Bumped whenever captured steps stop matching the boundary's passes
    // (interactions, walls, launch structure, a loaded checkpoint).
    unsigned long long gen = 0;
};

namespace bl {
// One coupled step's launches on the lattice stream, without host
// synchronisation: IBM stages, collide-and-stream, velocity-wall pass.
// `forked` lets independent IBM stages run on the boundary's own extra
// streams; otherwise everything stays on the lattice stream.
int ibm_step_launch(bl_ibm g, bool forked);
// Ensembles of boxes whose coupled steps take the paired-stage layout:
// two coupled steps of every box as one launch per stage over all boxes
// (ensemble_batch_launch enqueues them on s; the boxes' host state is as
// after ensemble_batch_build, which plans them).  *out stays null when the
// boxes cannot be batched (differing storage, parity, stage layout, ...).
struct EnsembleBatch;
// order, nt: the batched collide's workgroup order and nontemporal level
// per AA parity
int ensemble_batch_build(const std::vector<bl_lattice>& lats, const std::vector<bl_ibm>& ibms, const int order[2],
                         const int nt[2], EnsembleBatch** out);
int ensemble_batch_launch(const EnsembleBatch* b, hipStream_t s);
void ensemble_batch_free(EnsembleBatch* b);
// Reads the overflow flags of the last steps (synchronises the stream).
int ibm_check_overflow(bl_ibm g);
// The lattice is going away: forget it (the boundary's buffers stay valid
// until bl_ibm_destroy).
void ibm_detach(bl_ibm g);
void ibm_invalidate_graphs(bl_ibm g);  // captured steps no longer match the lattice's passes
// BL_OK when the boundary is initialised and every periodic extent of its
// lattice holds one stencil width (the kernel stages need it).
int need_stencil(bl_ibm g);
// Rigid particles: state upload at bl_ibm_init (rigid.hip); BL_ERR_STATE
// from the fluid-coupling stages while the boundary holds rigid particles.
int rigid_init(bl_ibm g);
// FP64 vertex data <-> the boundary's vertex reals (n values, stream order
// on s; the host side is synchronised before return).
int vreal_upload(bl_ibm g, void* dst, long off, const double* src, long n, hipStream_t s);
int vreal_download(bl_ibm g, double* dst, const void* src, long off, long n, hipStream_t s);
int rigid_refuse_coupling(bl_ibm g);
// A full force field was loaded into the lattice (every tile flagged):
// steps must run eagerly until the flags are consistent again.
bool ibm_needs_settle(bl_ibm g);

// Interaction forces (interactions.hip).
struct DecoSet;
DecoSet deco_set(bl_ibm g);                       // the wall-interaction decorators
int interactions_alloc(bl_ibm g);                 // collision grid for the current vertices
void interactions_free(bl_ibm g);
int interactions_launch(bl_ibm g, hipStream_t s);  // particle-particle forces into pp_force
}  // namespace bl
