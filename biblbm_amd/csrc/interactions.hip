// Interaction forces of the immersed boundary: wall-interaction decorators
// (applied inside the membrane-force gather) and particle-particle forces
// through the reference's collision handler.
//
// The collision handler (src/fsi/CollisionForces.h:84-212) bins the nodes
// into unit cells over the lattice box and visits, for every bulk cell, the
// pairs inside the cell and with its 7 forward neighbours, then the
// face / edge / corner cell pairs across periodic axes; each visited pair of
// different particles within the cut-off adds df = -F(d)/d dr to its first
// node and -df to its second.  Here every vertex gathers its own share of
// exactly those pairs, so no atomics touch the forces:
//   k_pp_bin    cell of each vertex; vertices pushed onto per-cell lists
//   k_pp_pack   one vertex per occupied cell copies its list into a run of
//               slots, sorted by vertex id (the reference's insertion order)
//   k_pp_force  per vertex: the cells it pairs with under the handler's
//               rules, each with the role (first / second node) the vertex
//               has in those pairs; the pair force is evaluated with the
//               reference's operand order for that role
//   k_pp_clear  empties the touched cells for the next evaluation
// A vertex's sum runs over its partner cells in a fixed rule order and over
// each cell's slots in vertex order, so results do not depend on
// scheduling.
#include <algorithm>
#include <cmath>
#include <string>

#include "ibm.hpp"
#include "interactions.hpp"

namespace bl {

namespace {

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

__global__ void __launch_bounds__(256) k_pp_bin(long NV, const double* __restrict__ pos, CollGrid g,
                                                int* __restrict__ head, int* __restrict__ next,
                                                int* __restrict__ cell)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NV) return;
    const int c = coll_cell(g, pos + 3 * v);
    cell[v] = c;
    if (c >= 0) next[v] = atomicExch(head + c, (int)v);
}

__global__ void __launch_bounds__(256) k_pp_pack(long NV, const int* __restrict__ cell,
                                                 const int* __restrict__ head, const int* __restrict__ next,
                                                 int* __restrict__ ctr, int* __restrict__ slot,
                                                 int2* __restrict__ info)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NV) return;
    const int c = cell[v];
    if (c < 0 || head[c] != (int)v) return;  // one vertex per occupied cell packs it
    int n = 0;
    for (int w = (int)v; w >= 0; w = next[w]) ++n;
    const int base = atomicAdd(ctr, n);
    int k = 0;
    for (int w = (int)v; w >= 0; w = next[w]) {
        // insertion into the sorted run slot[base .. base+k)
        int j = k;
        while (j > 0 && slot[base + j - 1] > w) {
            slot[base + j] = slot[base + j - 1];
            --j;
        }
        slot[base + j] = w;
        ++k;
    }
    info[c] = make_int2(base, n);
}

// minimal-image difference x - y along axis d (src/fsi/Periodicity.h:27-35,91-99)
__device__ inline double mm_diff(const CollGrid& g, int d, double x, double y)
{
#pragma clang fp contract(off)
    const double v = x - y;
    if (!g.per[d]) return v;
    return v - floor(v * g.Linv[d] + 0.5) * g.L[d];
}

struct PPAcc {
    double f[3];
};

// The pairs of vertex v with the vertices of cell pc.  role 0: v is the
// first node of every pair; 1: the second; 2 (its own cell): first if it
// precedes the other vertex in insertion order, else second.
__device__ inline void pp_cell(const CollGrid& g, const bl_potential& u, const double* __restrict__ pos,
                               const int* __restrict__ vpart, const int2* __restrict__ info,
                               const int* __restrict__ slot, int v, int ov, const double pv[3], int pc, int role,
                               PPAcc& a)
{
#pragma clang fp contract(off)
    const int2 in = info[pc];
    for (int s = in.x; s < in.x + in.y; ++s) {
        const int w = slot[s];
        if (vpart[w] == ov) continue;
        const bool first = role == 0 || (role == 2 && v < w);
        const double pw[3] = {pos[3 * (long)w], pos[3 * (long)w + 1], pos[3 * (long)w + 2]};
        double dr[3];
        for (int d = 0; d < 3; ++d) dr[d] = first ? mm_diff(g, d, pw[d], pv[d]) : mm_diff(g, d, pv[d], pw[d]);
        const double d2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        if (d2 <= g.cut2) {
            const double dist = sqrt(d2);
            const double sc = -pot_force(u, dist) / dist;
            for (int d = 0; d < 3; ++d) {
                const double df = sc * dr[d];
                a.f[d] = first ? a.f[d] + df : a.f[d] - df;
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_pp_force(long NV, CollGrid g, bl_potential u, int has_wall, bl_boundary wall,
                                                  const double* __restrict__ pos, const int* __restrict__ vpart,
                                                  const int* __restrict__ cell, const int2* __restrict__ info,
                                                  const int* __restrict__ slot, double* __restrict__ force)
{
#pragma clang fp contract(off)
    const long vl = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (vl >= NV) return;
    const int c = cell[vl];
    if (c < 0) return;
    const int v = (int)vl;
    const int NZ = g.num[2] + 2, NY = g.num[1] + 2;
    const int i = c / (NY * NZ), j = (c / NZ) % NY, k = c % NZ;
    const int X1 = g.num[0], Y1 = g.num[1], Z1 = g.num[2];
    auto at = [&](int a, int b, int e) { return e + NZ * (b + NY * a); };
    auto bulk = [&](int a, int b, int e) { return a >= 1 && a <= X1 && b >= 1 && b <= Y1 && e >= 1 && e <= Z1; };
    const double pv[3] = {pos[3 * vl], pos[3 * vl + 1], pos[3 * vl + 2]};
    const int ov = vpart[v];
    PPAcc a{{force[3 * vl], force[3 * vl + 1], force[3 * vl + 2]}};
    // forward neighbours of the bulk loop (src/fsi/CollisionForces.h:96-104)
    constexpr int off[7][3] = {{0, 1, 0}, {0, 1, 1}, {0, 0, 1}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}};
    if (bulk(i, j, k)) {
        pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, c, 2, a);
        for (int o = 0; o < 7; ++o)
            pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(i + off[o][0], j + off[o][1], k + off[o][2]), 0, a);
    }
    for (int o = 0; o < 7; ++o) {
        const int a0 = i - off[o][0], b0 = j - off[o][1], e0 = k - off[o][2];
        if (bulk(a0, b0, e0)) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(a0, b0, e0), 1, a);
    }
    // cell pairs across periodic axes (:119-160): first cell at the low
    // bulk index of the axes involved, second at the high one
    const bool bi = i >= 1 && i <= X1, bj = j >= 1 && j <= Y1, bk = k >= 1 && k <= Z1;
    if (g.per[0] && bj && bk) {
        if (i == 1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(X1, j, k), 0, a);
        if (i == X1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(1, j, k), 1, a);
    }
    if (g.per[1] && bi && bk) {
        if (j == 1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(i, Y1, k), 0, a);
        if (j == Y1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(i, 1, k), 1, a);
    }
    if (g.per[2] && bi && bj) {
        if (k == 1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(i, j, Z1), 0, a);
        if (k == Z1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(i, j, 1), 1, a);
    }
    if (g.per[0] && g.per[1] && bk) {
        if (i == 1 && j == 1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(X1, Y1, k), 0, a);
        if (i == X1 && j == Y1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(1, 1, k), 1, a);
    }
    if (g.per[0] && g.per[2] && bj) {
        if (i == 1 && k == 1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(X1, j, Z1), 0, a);
        if (i == X1 && k == Z1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(1, j, 1), 1, a);
    }
    if (g.per[1] && g.per[2] && bi) {
        if (j == 1 && k == 1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(i, Y1, Z1), 0, a);
        if (j == Y1 && k == Z1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(i, 1, 1), 1, a);
    }
    if (g.per[0] && g.per[1] && g.per[2]) {
        if (i == 1 && j == 1 && k == 1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(X1, Y1, Z1), 0, a);
        if (i == X1 && j == Y1 && k == Z1) pp_cell(g, u, pos, vpart, info, slot, v, ov, pv, at(1, 1, 1), 1, a);
    }
    // the handler's wall forces (:167-187), after every pair
    if (has_wall && bulk(i, j, k)) {
        const double x0 = g.bb_lo[0] + (double)(i - 1) * g.cs[0];
        const double y0 = g.bb_lo[1] + (double)(j - 1) * g.cs[1];
        const double z0 = g.bb_lo[2] + (double)(k - 1) * g.cs[2];
        const double r[6] = {x0, x0 + g.cs[0], y0, y0 + g.cs[1], z0, z0 + g.cs[2]};
        if (bnd_intersects_rect(wall, r, g.cut)) {
            const double d = bnd_distance(wall, pv);
            if (d <= g.cut) {
                const double F = pot_force(u, d);
                double n[3];
                bnd_normal(wall, pv, n);
                for (int q = 0; q < 3; ++q) a.f[q] -= F * n[q];
            }
        }
    }
    force[3 * vl] = a.f[0];
    force[3 * vl + 1] = a.f[1];
    force[3 * vl + 2] = a.f[2];
}

__global__ void __launch_bounds__(256) k_pp_clear(long NV, const int* __restrict__ cell, int* __restrict__ head,
                                                  int2* __restrict__ info, int* __restrict__ ctr)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NV) return;
    if (v == 0) *ctr = 0;
    const int c = cell[v];
    if (c >= 0) {
        head[c] = -1;
        info[c] = make_int2(0, 0);
    }
}

__global__ void k_pp_init(long n, int* __restrict__ head, int2* __restrict__ info)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) {
        head[c] = -1;
        info[c] = make_int2(0, 0);
    }
}

__global__ void k_zero_f64(long n, double* __restrict__ p)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

CollGrid grid_for(bl_ibm g, const bl_potential& u)
{
    bl_lattice L = g->lat;
    return coll_grid(L->nx, L->ny, L->nz, L->periodic, pot_cutoff(u));
}

int check_potential(bl_ibm g, const bl_potential* u, const char* who, bool pp)
{
    if (!u) return fail(BL_ERR_ARG, std::string(who) + ": null potential");
    if (u->kind != BL_POTENTIAL_MORSE && u->kind != BL_POTENTIAL_SPRING)
        return fail(BL_ERR_ARG, std::string(who) + ": unknown potential kind");
    if (!(pot_cutoff(*u) > 0)) return fail(BL_ERR_ARG, std::string(who) + ": the cut-off distance must be positive");
    if (pp && pot_cutoff(*u) > g->h - 0.5)
        return fail(BL_ERR_ARG, std::string(who) + ": the interaction distance " + std::to_string(pot_cutoff(*u)) +
                                    " exceeds the kernel's half support - 0.5 (" + std::to_string(g->h - 0.5) + ")");
    return BL_OK;
}

// One collision-handler pass: pair forces (and the handler's wall forces)
// added to the vertex forces.
int pp_pass(bl_ibm g, const bl_potential& u, const bl_boundary* wall, hipStream_t s)
{
    const long NV = g->NV;
    if (NV == 0) return BL_OK;
    const CollGrid cg = grid_for(g, u);
    const bl_boundary w = wall ? *wall : bl_boundary{};
    hipLaunchKernelGGL(k_pp_bin, dim3(nblk(NV)), dim3(256), 0, s, NV, (const double*)g->pos, cg, g->pp_head,
                       g->pp_next, g->pp_cell);
    hipLaunchKernelGGL(k_pp_pack, dim3(nblk(NV)), dim3(256), 0, s, NV, (const int*)g->pp_cell,
                       (const int*)g->pp_head, (const int*)g->pp_next, g->pp_ctr, g->pp_slot, g->pp_info);
    hipLaunchKernelGGL(k_pp_force, dim3(nblk(NV)), dim3(256), 0, s, NV, cg, u, wall ? 1 : 0, w,
                       (const double*)g->pos, (const int*)g->v_part, (const int*)g->pp_cell,
                       (const int2*)g->pp_info, (const int*)g->pp_slot, g->force);
    hipLaunchKernelGGL(k_pp_clear, dim3(nblk(NV)), dim3(256), 0, s, NV, (const int*)g->pp_cell, g->pp_head,
                       g->pp_info, g->pp_ctr);
    BL_CHECK_LAUNCH();
    return BL_OK;
}

}  // namespace

DecoSet deco_set(bl_ibm g)
{
    DecoSet d{};
    d.n = (int)g->wall_deco.size();
    for (int k = 0; k < d.n; ++k) {
        d.b[k] = g->wall_deco[k].b;
        d.u[k] = g->wall_deco[k].u;
    }
    return d;
}

// Cell grid and per-vertex lists of the collision handler, sized for the
// current vertex count and lattice (synchronous; never under capture).
int interactions_alloc(bl_ibm g)
{
    if (!g->inited || !g->lat) return BL_OK;
    const bl_potential any{BL_POTENTIAL_SPRING, {0.5, 0, 0}};
    const long ncell = coll_ncell(grid_for(g, any));
    const size_t nv = (size_t)std::max<long>(g->NV, 1);
    if (g->pp_head && g->pp_ncell == ncell) return BL_OK;
    void* old[] = {g->pp_head, g->pp_info, g->pp_next, g->pp_cell, g->pp_slot, g->pp_ctr};
    for (void* p : old) hipFree(p);
    g->pp_head = g->pp_next = g->pp_cell = g->pp_slot = g->pp_ctr = nullptr;
    g->pp_info = nullptr;
    hipError_t e = hipMalloc(&g->pp_head, sizeof(int) * ncell);
    if (e == hipSuccess) e = hipMalloc(&g->pp_info, sizeof(int2) * ncell);
    if (e == hipSuccess) e = hipMalloc(&g->pp_next, sizeof(int) * nv);
    if (e == hipSuccess) e = hipMalloc(&g->pp_cell, sizeof(int) * nv);
    if (e == hipSuccess) e = hipMalloc(&g->pp_slot, sizeof(int) * nv);
    if (e == hipSuccess) e = hipMalloc(&g->pp_ctr, sizeof(int));
    if (e != hipSuccess) return fail(BL_ERR_HIP, std::string("collision grid: ") + hipGetErrorString(e));
    g->pp_ncell = ncell;
    hipStream_t s = g->lat->stream;
    hipLaunchKernelGGL(k_pp_init, dim3(nblk(ncell)), dim3(256), 0, s, ncell, g->pp_head, g->pp_info);
    BL_CHECK_LAUNCH();
    BL_HIP(hipMemsetAsync(g->pp_ctr, 0, sizeof(int), s));
    BL_HIP(hipStreamSynchronize(s));
    return BL_OK;
}

void interactions_free(bl_ibm g)
{
    void* p[] = {g->pp_head, g->pp_info, g->pp_next, g->pp_cell, g->pp_slot, g->pp_ctr};
    for (void* q : p) hipFree(q);
    g->pp_head = g->pp_next = g->pp_cell = g->pp_slot = g->pp_ctr = nullptr;
    g->pp_info = nullptr;
    g->pp_ncell = 0;
}

// The particle-particle forces of a step, after the membrane forces
// (src/fsi/ImmersedBoundaryDynamics_forces.hh:299-301).
int interactions_launch(bl_ibm g, hipStream_t s)
{
    int rc;
    for (const bl_potential& u : g->pp)
        if ((rc = pp_pass(g, u, nullptr, s))) return rc;
    return BL_OK;
}

}  // namespace bl

using namespace bl;

extern "C" {

int bl_potential_eval(const bl_potential* u, double d, double* force, double* cutoff)
{
    if (!u) return fail(BL_ERR_ARG, "bl_potential_eval: null potential");
    if (u->kind != BL_POTENTIAL_MORSE && u->kind != BL_POTENTIAL_SPRING)
        return fail(BL_ERR_ARG, "bl_potential_eval: unknown potential kind");
    if (force) *force = pot_force(*u, d);
    if (cutoff) *cutoff = pot_cutoff(*u);
    return BL_OK;
}

int bl_ibm_add_wall_interaction(bl_ibm g, const bl_boundary* b, const bl_potential* u)
{
    if (!g || !b) return fail(BL_ERR_ARG, "bl_ibm_add_wall_interaction: null argument");
    if (b->kind != BL_BOUNDARY_PIPE && b->kind != BL_BOUNDARY_PLATES)
        return fail(BL_ERR_ARG, "bl_ibm_add_wall_interaction: unknown boundary kind");
    int rc = check_potential(g, u, "bl_ibm_add_wall_interaction", false);
    if (rc) return rc;
    if ((int)g->wall_deco.size() >= kMaxWallDeco)
        return fail(BL_ERR_CAPACITY, "bl_ibm_add_wall_interaction: at most " + std::to_string(kMaxWallDeco) +
                                         " wall interactions");
    if (g->lat) BL_HIP(hipStreamSynchronize(g->lat->stream));
    g->wall_deco.push_back({*b, *u});
    ibm_invalidate_graphs(g);
    return BL_OK;
}

int bl_ibm_add_pp_force(bl_ibm g, const bl_potential* u)
{
    if (!g) return fail(BL_ERR_ARG, "bl_ibm_add_pp_force: null ibm handle");
    int rc = check_potential(g, u, "bl_ibm_add_pp_force", true);
    if (rc) return rc;
    if ((int)g->pp.size() >= kMaxPPForce)
        return fail(BL_ERR_CAPACITY, "bl_ibm_add_pp_force: at most " + std::to_string(kMaxPPForce) +
                                         " particle-particle forces");
    if (g->lat) BL_HIP(hipStreamSynchronize(g->lat->stream));
    g->pp.push_back(*u);
    ibm_invalidate_graphs(g);
    return interactions_alloc(g);
}

int bl_ibm_compute_collision_forces(bl_ibm g, const bl_potential* u, const bl_boundary* b)
{
    if (!g) return fail(BL_ERR_ARG, "null ibm handle");
    if (!g->inited) return fail(BL_ERR_STATE, "immersed boundary not initialised (call bl_ibm_init)");
    if (!g->lat) return fail(BL_ERR_STATE, "the lattice of this immersed boundary was destroyed");
    int rc = check_potential(g, u, "bl_ibm_compute_collision_forces", true);
    if (rc) return rc;
    if (b && b->kind != BL_BOUNDARY_PIPE && b->kind != BL_BOUNDARY_PLATES)
        return fail(BL_ERR_ARG, "bl_ibm_compute_collision_forces: unknown boundary kind");
    if ((rc = interactions_alloc(g))) return rc;
    hipStream_t s = g->lat->stream;
    if (g->NV) hipLaunchKernelGGL(k_zero_f64, dim3(nblk(3 * g->NV)), dim3(256), 0, s, 3 * g->NV, g->force);
    BL_CHECK_LAUNCH();
    return pp_pass(g, *u, b, s);
}

}  // extern "C"
