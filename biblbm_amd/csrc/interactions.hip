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
//   k_pp_link   every vertex of a cell finds its successor in ascending id
//               order (the reference's insertion order) and writes its record
//               (position, owner, successor); the cell's smallest id becomes
//               the cell's first vertex
//   k_pp_force  per vertex, four lanes: the cells it pairs with under the
//               handler's rules, each with the role (first / second node) the
//               vertex has in those pairs; the pair force is evaluated with
//               the reference's operand order for that role
//   k_pp_clear  empties the touched cells for the next evaluation
// A vertex's sum runs over its partner cells in a fixed rule order (dealt
// round-robin to four lanes, added by a fixed butterfly) and over each
// cell's vertices in id order, so results do not depend on scheduling.  No
// counter is shared between cells: same-address atomics serialise across
// the chip's dies.
#include <algorithm>
#include <cmath>
#include <string>

#include "ibm.hpp"
#include "interactions.hpp"

namespace bl {

namespace {

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Vertex positions and forces are in the boundary's vertex reals R; the
// handler's geometry and potentials are evaluated in FP64.
template <class R>
__global__ void __launch_bounds__(256) k_pp_bin(long NV, const R* __restrict__ pos, CollGrid g,
                                                int* __restrict__ head, int* __restrict__ next,
                                                int* __restrict__ cell)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NV) return;
    const double p[3] = {(double)pos[3 * v], (double)pos[3 * v + 1], (double)pos[3 * v + 2]};
    const int c = coll_cell(g, p);
    cell[v] = c;
    if (c >= 0) next[v] = atomicExch(head + c, (int)v);
}

// Every listed vertex: its successor in its cell (the smallest larger id)
// and, for the smallest id, the cell's first vertex.
template <class R>
__global__ void __launch_bounds__(256) k_pp_link(long NV, const int* __restrict__ cell,
                                                 const int* __restrict__ head, const int* __restrict__ next,
                                                 const R* __restrict__ pos, const int* __restrict__ vpart,
                                                 PPRec* __restrict__ rec, int* __restrict__ first)
{
    const long vl = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (vl >= NV) return;
    const int c = cell[vl];
    if (c < 0) return;
    const int v = (int)vl;
    int succ = -1;
    bool smallest = true;
    for (int w = head[c]; w >= 0; w = next[w]) {
        if (w > v && (succ < 0 || w < succ)) succ = w;
        if (w < v) smallest = false;
    }
    PPRec r;
    r.x = (double)pos[3 * vl];
    r.y = (double)pos[3 * vl + 1];
    r.z = (double)pos[3 * vl + 2];
    r.obj = vpart[v];
    r.succ = succ;
    rec[v] = r;
    if (smallest) first[c] = v;
}

// minimal-image difference x - y along axis d (src/fsi/Periodicity.h:27-35,91-99)
__device__ inline double mm_diff(const CollGrid& g, int d, double x, double y)
{
#pragma clang fp contract(off)
    const double v = x - y;
    if (!g.per[d]) return v;
    return v - floor(v * g.Linv[d] + 0.5) * g.L[d];
}

// The partner cells of a vertex in cell (i, j, k), as 22 entries in the
// handler's order: the cell itself and its 7 forward neighbours when it is a
// bulk cell (src/fsi/CollisionForces.h:89-117), the 7 cells it is a forward
// neighbour of, then the face / edge / corner pairs across periodic axes
// (:119-160: first cell at the low bulk index of the axes involved, second
// at the high one).  Returns the partner cell or -1, and the role: 0 the
// vertex is the first node of every pair, 1 the second, 2 (its own cell)
// first when it precedes the other vertex.
constexpr int kPartners = 22;

__device__ inline int partner(const CollGrid& g, int t, int i, int j, int k, int& role)
{
    const int NZ = g.num[2] + 2, NY = g.num[1] + 2;
    const int X1 = g.num[0], Y1 = g.num[1], Z1 = g.num[2];
    auto bulk = [&](int a, int b, int e) { return a >= 1 && a <= X1 && b >= 1 && b <= Y1 && e >= 1 && e <= Z1; };
    auto at = [&](int a, int b, int e) { return e + NZ * (b + NY * a); };
    // forward offsets, bits (dx << 2) | (dy << 1) | dz, in the handler's order
    // (0,1,0) (0,1,1) (0,0,1) (1,0,0) (1,1,0) (1,1,1) (1,0,1)
    constexpr unsigned kFwd = 2u | (3u << 3) | (1u << 6) | (4u << 9) | (6u << 12) | (7u << 15) | (5u << 18);
    if (t == 0) {
        role = 2;
        return bulk(i, j, k) ? at(i, j, k) : -1;
    }
    if (t <= 14) {
        const unsigned o = (kFwd >> (3 * ((t - 1) % 7))) & 7u;
        const int dx = (o >> 2) & 1, dy = (o >> 1) & 1, dz = o & 1;
        if (t <= 7) {
            role = 0;
            return bulk(i, j, k) ? at(i + dx, j + dy, k + dz) : -1;
        }
        role = 1;
        return bulk(i - dx, j - dy, k - dz) ? at(i - dx, j - dy, k - dz) : -1;
    }
    // periodic rules over the axis sets x, y, z, xy, xz, yz, xyz
    constexpr unsigned char kAxes[7] = {1, 2, 4, 3, 5, 6, 7};
    const unsigned m = kAxes[t - 15];
    const int ix[3] = {i, j, k}, hi[3] = {X1, Y1, Z1};
    bool low = true, high = true;
    for (int d = 0; d < 3; ++d) {
        const bool in_bulk = ix[d] >= 1 && ix[d] <= hi[d];
        if (m & (1u << d)) {
            if (!g.per[d]) return -1;
            low = low && ix[d] == 1;
            high = high && ix[d] == hi[d];
        } else if (!in_bulk) {
            return -1;
        }
    }
    int q[3] = {i, j, k};
    if (low) {
        role = 0;
        for (int d = 0; d < 3; ++d)
            if (m & (1u << d)) q[d] = hi[d];
        return at(q[0], q[1], q[2]);
    }
    if (high) {
        role = 1;
        for (int d = 0; d < 3; ++d)
            if (m & (1u << d)) q[d] = 1;
        return at(q[0], q[1], q[2]);
    }
    return -1;
}

// Four lanes per vertex: lane l takes partner entries l, l+4, ...; a fixed
// butterfly adds the lanes' sums to the vertex force.
constexpr int kPPLanes = 4;

// out[v] = in[v] (0 without `in`) + the vertex's pair forces, for every
// vertex (those outside the handler get in[v]).
template <class R>
__global__ void __launch_bounds__(256) k_pp_force(long NV, CollGrid g, bl_potential u, int has_wall, bl_boundary wall,
                                                  const int* __restrict__ cell, const int* __restrict__ first,
                                                  const PPRec* __restrict__ rec, const R* in, R* out)
{
#pragma clang fp contract(off)
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long vl = gid / kPPLanes;
    const int lane = (int)(gid % kPPLanes);
    const int c = vl < NV ? cell[vl] : -1;
    double f[3] = {0, 0, 0};
    if (in && lane == 0 && vl < NV) {
        f[0] = (double)in[3 * vl];
        f[1] = (double)in[3 * vl + 1];
        f[2] = (double)in[3 * vl + 2];
    }
    double pv[3] = {0, 0, 0};
    const int NZ = g.num[2] + 2, NY = g.num[1] + 2;
    int i = 0, j = 0, k = 0;
    if (c >= 0) {
        const int v = (int)vl;
        i = c / (NY * NZ);
        j = (c / NZ) % NY;
        k = c % NZ;
        const PPRec me = rec[v];
        pv[0] = me.x;
        pv[1] = me.y;
        pv[2] = me.z;
#pragma unroll 1
        for (int t = lane; t < kPartners; t += kPPLanes) {
            int role = 0;
            const int pc = partner(g, t, i, j, k, role);
            if (pc < 0) continue;
            for (int w = first[pc]; w >= 0;) {
                const PPRec r = rec[w];
                if (r.obj != me.obj) {
                    const bool fst = role == 0 || (role == 2 && v < w);
                    const double pw[3] = {r.x, r.y, r.z};
                    double dr[3];
                    for (int d = 0; d < 3; ++d)
                        dr[d] = fst ? mm_diff(g, d, pw[d], pv[d]) : mm_diff(g, d, pv[d], pw[d]);
                    const double d2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                    if (d2 <= g.cut2) {
                        const double dist = sqrt(d2);
                        const double sc = -pot_force(u, dist) / dist;
                        for (int d = 0; d < 3; ++d) {
                            const double df = sc * dr[d];
                            f[d] = fst ? f[d] + df : f[d] - df;
                        }
                    }
                }
                w = r.succ;
            }
        }
    }
#pragma unroll
    for (int o = 1; o < kPPLanes; o <<= 1) {
        f[0] += __shfl_xor(f[0], o);
        f[1] += __shfl_xor(f[1], o);
        f[2] += __shfl_xor(f[2], o);
    }
    if (vl >= NV || lane != 0) return;
    // the handler's wall forces (:167-187), after every pair
    const int X1 = g.num[0], Y1 = g.num[1], Z1 = g.num[2];
    if (c >= 0 && has_wall && i >= 1 && i <= X1 && j >= 1 && j <= Y1 && k >= 1 && k <= Z1) {
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
                for (int q = 0; q < 3; ++q) f[q] -= F * n[q];
            }
        }
    }
    out[3 * vl] = (R)f[0];
    out[3 * vl + 1] = (R)f[1];
    out[3 * vl + 2] = (R)f[2];
}

__global__ void __launch_bounds__(256) k_pp_clear(long NV, const int* __restrict__ cell, int* __restrict__ head,
                                                  int* __restrict__ first)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NV) return;
    const int c = cell[v];
    if (c >= 0) {
        head[c] = -1;
        first[c] = -1;
    }
}

__global__ void k_pp_init(long n, int* __restrict__ head, int* __restrict__ first)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) {
        head[c] = -1;
        first[c] = -1;
    }
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

// One collision-handler pass: out := in (0 without `in`) + pair forces (+
// the handler's wall forces) of every vertex.
template <class R>
void pp_launch(bl_ibm g, const CollGrid& cg, const bl_potential& u, const bl_boundary* wall, const void* in,
               void* out, hipStream_t s)
{
    const long NV = g->NV;
    const bl_boundary w = wall ? *wall : bl_boundary{};
    hipLaunchKernelGGL(k_pp_bin<R>, dim3(nblk(NV)), dim3(256), 0, s, NV, (const R*)g->pos, cg, g->pp_head,
                       g->pp_next, g->pp_cell);
    hipLaunchKernelGGL(k_pp_link<R>, dim3(nblk(NV)), dim3(256), 0, s, NV, (const int*)g->pp_cell,
                       (const int*)g->pp_head, (const int*)g->pp_next, (const R*)g->pos, (const int*)g->v_part,
                       g->pp_rec, g->pp_first);
    hipLaunchKernelGGL(k_pp_force<R>, dim3(nblk(kPPLanes * NV)), dim3(256), 0, s, NV, cg, u, wall ? 1 : 0, w,
                       (const int*)g->pp_cell, (const int*)g->pp_first, (const PPRec*)g->pp_rec, (const R*)in,
                       (R*)out);
}

int pp_pass(bl_ibm g, const bl_potential& u, const bl_boundary* wall, const void* in, void* out, hipStream_t s)
{
    const long NV = g->NV;
    if (NV == 0) return BL_OK;
    const CollGrid cg = grid_for(g, u);
    if (g->vreal == 4) pp_launch<float>(g, cg, u, wall, in, out, s);
    else pp_launch<double>(g, cg, u, wall, in, out, s);
    hipLaunchKernelGGL(k_pp_clear, dim3(nblk(NV)), dim3(256), 0, s, NV, (const int*)g->pp_cell, g->pp_head,
                       g->pp_first);
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
    interactions_free(g);
    hipError_t e = hipMalloc(&g->pp_head, sizeof(int) * ncell);
    if (e == hipSuccess) e = hipMalloc(&g->pp_first, sizeof(int) * ncell);
    if (e == hipSuccess) e = hipMalloc(&g->pp_next, sizeof(int) * nv);
    if (e == hipSuccess) e = hipMalloc(&g->pp_cell, sizeof(int) * nv);
    if (e == hipSuccess) e = hipMalloc(&g->pp_rec, sizeof(PPRec) * nv);
    if (e == hipSuccess) e = hipMalloc(&g->pp_force, (size_t)g->vreal * 3 * nv);
    if (e != hipSuccess) return fail(BL_ERR_HIP, std::string("collision grid: ") + hipGetErrorString(e));
    g->pp_ncell = ncell;
    hipStream_t s = g->lat->stream;
    hipLaunchKernelGGL(k_pp_init, dim3(nblk(ncell)), dim3(256), 0, s, ncell, g->pp_head, g->pp_first);
    BL_CHECK_LAUNCH();
    BL_HIP(hipStreamSynchronize(s));
    return BL_OK;
}

void interactions_free(bl_ibm g)
{
    void* p[] = {g->pp_head, g->pp_first, g->pp_next, g->pp_cell, g->pp_rec, g->pp_force};
    for (void* q : p) (void)hipFree(q);
    g->pp_head = g->pp_first = g->pp_next = g->pp_cell = nullptr;
    g->pp_rec = nullptr;
    g->pp_force = nullptr;
    g->pp_ncell = 0;
}

// The particle-particle forces of a step into their own per-vertex buffer;
// the membrane-force gather adds it last, as the reference adds them after
// the membrane forces (src/fsi/ImmersedBoundaryDynamics_forces.hh:299-301).
int interactions_launch(bl_ibm g, hipStream_t s)
{
    int rc;
    for (size_t k = 0; k < g->pp.size(); ++k)
        if ((rc = pp_pass(g, g->pp[k], nullptr, k ? g->pp_force : nullptr, g->pp_force, s))) return rc;
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
    // vertex forces := the handler's forces alone (set_forces_to_zero first)
    return pp_pass(g, *u, b, nullptr, g->force, g->lat->stream);
}

}  // extern "C"
