// Immersed-boundary coupling on gfx950.
//
// Per coupled step (the IBM functional of the reference,
// src/fsi/DataProcessors.hh:94-108, preceded by the progs' computeVelocity
// and force reset):
//   bin      every vertex's delta-kernel stencil into 4x4x4 lattice tiles
//            (counts, exclusive scan, list fill; atomics aggregated per
//            workgroup in LDS);
//   forces   per element (triangle / link / pair) contributions, then a
//            per-vertex gather over the CSR incidence lists (K5, no atomics);
//   stage    u = j/rho + F/2 on the cells of the touched tiles (from the
//            post-stream populations and the force of the last collision);
//   spread   one workgroup per touched tile: its vertex list is sorted in
//            LDS, every cell sums g + sum_k phi F_k in vertex order (K6,
//            deterministic, no float atomics);
//   interp   U_k = sum phi phi phi u, one thread per vertex in the
//            reference's node order (K4), fused with the advection x += U;
//   move     per-particle area / volume / COM / bbox in FP64 with a
//            fixed reduction order, COM mapped onto the periodic grid (K7);
//   voxelize z-ray hits per column, sorted, parity fill of the phase (K8),
//            only when the inner and outer viscosities differ.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <string>
#include <type_traits>
#include <vector>

#include "ibm.hpp"
#include "interactions.hpp"
#include "lattice_dev.hpp"
#include "walls.hpp"

namespace bl {

namespace {

constexpr int kTileCap = 512;      // list entries per 4x4x4 tile (stencils touching it)
constexpr int kSmallTile = 64;     // tiles with more entries are spread by a whole workgroup
constexpr int kMaxList = kTileCap; // entries the spreading kernel sorts in LDS

// ------------------------------------------------------------- helpers
// Vertex data (positions, velocities, forces, element forces, stencil
// weights, spreading sums) is held and computed in the real type R: double
// with FP64 populations, float with FP32 and FP16-shifted ones.  Rest
// geometry, per-particle area / volume / centroid reductions and the
// conservation sums stay FP64 in both.

// Delta kernels (src/fsi/Dirac.h:23-66,74-101), 1-D factor.
template <class R>
__device__ __forceinline__ R dirac(int kernel, R x)
{
    const R dx = fabs(x);
    if (kernel == BL_KERNEL_ROMA)
        return (R(1.0) / R(6.0)) * (dx > R(0.5) ? (R(5.0) - R(3) * dx - sqrt(-R(3) * (R(1) - dx) * (R(1) - dx) + R(1)))
                                                : R(2.0) * (R(1) + sqrt(R(1) - R(3) * dx * dx)));
    if (kernel == BL_KERNEL_PESKIN)
        return R(0.125) * (dx <= R(1) ? (R(3) - R(2) * dx + sqrt(R(1) + R(4) * dx * (R(1) - dx)))
                                      : (R(5) - R(2) * dx - sqrt(-R(7) + R(4) * dx * (R(3) - dx))));
    return R(1.) - fabs(x);
}

__device__ __forceinline__ int wrap_node(int i, int n, int per)
{
    if (per) {
        int r = i % n;
        return r < 0 ? r + n : r;
    }
    return (i < 0 || i >= n) ? -1 : i;
}

template <class R>
struct V3 {
    R x, y, z;
};
using D3 = V3<double>;
template <class R>
__device__ __forceinline__ V3<R> ld3(const R* p) { return {p[0], p[1], p[2]}; }
// a vertex in FP64 (geometry reductions, ray tracing)
template <class R>
__device__ __forceinline__ D3 ld3d(const R* p) { return {(double)p[0], (double)p[1], (double)p[2]}; }
template <class R>
__device__ __forceinline__ V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R>
__device__ __forceinline__ V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R>
__device__ __forceinline__ V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R>
__device__ __forceinline__ V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R>
__device__ __forceinline__ R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R>
__device__ __forceinline__ V3<R> cross(V3<R> a, V3<R> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R>
__device__ __forceinline__ R norm(V3<R> a) { return sqrt(dot(a, a)); }
template <class R>
struct Vec2;
template <>
struct Vec2<double> {
    using T = double2;
};
template <>
struct Vec2<float> {
    using T = float2;
};
template <class R>
__device__ __forceinline__ void st3(R* p, V3<R> a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }

struct Grid {
    int n[3];
    int per[3];
    int nt[3];
};

Grid grid_of(const bl_lattice_s* L)
{
    Grid g;
    g.n[0] = L->nx; g.n[1] = L->ny; g.n[2] = L->nz;
    for (int d = 0; d < 3; ++d) g.per[d] = L->periodic[d];
    g.nt[0] = L->ntx; g.nt[1] = L->nty; g.nt[2] = L->ntz;
    return g;
}

// Workgroup order of the immersed-boundary launches over the XCDs: chunks
// of 2^lc consecutive workgroups on one XCD, so that neighbouring vertices
// (which share stencil nodes, element records and tiles) meet in one L2;
// lc = 0 is the hardware's order.  Every order gives the same values.
__device__ __forceinline__ long ibm_block(long bx, long nbx, int lc)
{
    return lc > 0 ? (long)xcd_chunk_block((unsigned)bx, (unsigned)nbx, (unsigned)lc) : bx;
}

// ------------------------------------------------------------- binning
// Stencil bases are stored wrapped into [0, n) on periodic axes (raw on the
// others), so node k of a stencil is base + k minus n at most once.
__device__ __forceinline__ int node_at(int bw, int k, int n, int per)
{
    int v = bw + k;
    if (per) return v >= n ? v - n : v;
    return (v < 0 || v >= n) ? -1 : v;
}

// Distinct tiles touched along one axis by the S (<= 4) nodes of a stencil:
// at most 3 (two for the run up to the face, one more after a periodic
// wrap), kept in registers.
struct AxisTiles {
    int t0 = -1, t1 = -1, t2 = -1, m = 0;
    __device__ __forceinline__ int get(int i) const { return i == 0 ? t0 : (i == 1 ? t1 : t2); }
};

__device__ __forceinline__ AxisTiles axis_tiles(int bw, int S, int n, int per)
{
    AxisTiles r;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int node = k < S ? node_at(bw, k, n, per) : -1;
        const int t = node >> 2;
        const bool fresh = node >= 0 && t != r.t0 && t != r.t1 && t != r.t2;
        if (fresh) {
            if (r.m == 0) r.t0 = t;
            else if (r.m == 1) r.t1 = t;
            else r.t2 = t;
            ++r.m;
        }
    }
    return r;
}

// Tiles touched by a vertex's stencil: the product of the per-axis sets.
struct StencilTiles {
    AxisTiles ax[3];
    __device__ __forceinline__ void init(const Grid& gd, int S, const int* b)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a) ax[a] = axis_tiles(b[a], S, gd.n[a], gd.per[a]);
    }
    // tile (i, j, k) of the product, or -1 past the per-axis counts
    __device__ __forceinline__ int at(const Grid& gd, int i, int j, int k) const
    {
        if (i >= ax[0].m || j >= ax[1].m || k >= ax[2].m) return -1;
        return (ax[0].get(i) * gd.nt[1] + ax[1].get(j)) * gd.nt[2] + ax[2].get(k);
    }
};

// Neighbouring vertices of a membrane touch the same few tiles, so a
// workgroup first aggregates its (tile, vertex) pairs in an LDS hash table
// (open addressing, linear probing) and then issues one global atomic per
// distinct tile.  A pair whose tile finds the table full goes straight to
// the global counter; counts stay exact either way.
constexpr int kHashBits = 9;  // 512 slots: a workgroup's 256 vertices touch a few dozen tiles
constexpr int kHash = 1 << kHashBits;

__device__ __forceinline__ int hash_insert(int* keys, int t)
{
    unsigned h = ((unsigned)t * 2654435761u) >> (32 - kHashBits);
    for (int probe = 0; probe < kHash; ++probe) {
        const int old = atomicCAS(&keys[h], -1, t);
        if (old == -1 || old == t) return (int)h;
        h = (h + 1) & (kHash - 1);
    }
    return -1;
}

// slot of a tile inserted earlier in the same workgroup, -1 if its insert
// found the table full (then it was never inserted)
__device__ __forceinline__ int hash_find(const int* keys, int t)
{
    unsigned h = ((unsigned)t * 2654435761u) >> (32 - kHashBits);
    for (int probe = 0; probe < kHash; ++probe) {
        const int k = keys[h];
        if (k == t) return (int)h;
        if (k == -1) return -1;
        h = (h + 1) & (kHash - 1);
    }
    return -1;
}

__device__ __forceinline__ void hash_clear(int* keys, int* a, int* b)
{
    for (int i = threadIdx.x; i < kHash; i += blockDim.x) {
        keys[i] = -1;
        a[i] = 0;
        if (b) b[i] = 0;
    }
}

// Readies list `act` (length *na) / flag buffer `flags` for a binning:
// clears the flags of the tiles the list held (two steps back), then the
// last workgroup to finish empties the list and its pair count.  The first
// kResetBlocks workgroups of the grid that carries it take part (all their
// threads); they have read the length before they arrive, so no fence is
// needed, and a bounded number of them keeps the arrival atomic cheap.
constexpr int kResetBlocks = 256;
struct ListReset {
    const int* act = nullptr;  // null: nothing to reset
    int* na = nullptr;
    unsigned char* flags = nullptr;
    int* pairs = nullptr;
    int* arrive = nullptr;
    int* nbig = nullptr;
};

__device__ __forceinline__ void list_reset(const ListReset& lr, int bx, int nbx)
{
    const int nb = min(nbx, kResetBlocks);
    if (bx >= nb) return;
    __shared__ int n;
    if (threadIdx.x == 0) n = *lr.na;
    __syncthreads();
    for (int i = bx * blockDim.x + threadIdx.x; i < n; i += nb * blockDim.x) lr.flags[lr.act[i]] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        if (atomicAdd(lr.arrive, 1) == nb - 1) {
            *lr.na = 0;
            *lr.pairs = 0;
            *lr.nbig = 0;
            *lr.arrive = 0;
        }
    }
}

__global__ void __launch_bounds__(256) k_bin_prep(ListReset lr) { list_reset(lr, blockIdx.x, gridDim.x); }

// Near-wall vertices (a boundary is set and the vertex is nearer to it than
// the kernel's half support + 1, src/fsi/ImmersedBoundaryDynamics.hh:
// 346-357) spread after every other vertex, the kernel part (force scaled by
// the blend weight w) before the top-hat part (force scaled by 1 - w), as
// src/fsi/ImmersedBoundaryDynamics_forces.hh:193-237,259-277 orders them.
// List entries carry that order in their value: v for a bulk vertex,
// kNearKey | 2v for the kernel part and kNearKey | 2v + 1 for the top-hat
// part of a near-wall vertex, so sorting a tile's entries gives the
// reference's summation order.
constexpr int kNearKey = 1 << 30;

template <class R>
struct NearView {
    bl_boundary b;
    int on = 0;                      // a boundary is set
    unsigned char* near = nullptr;   // per vertex: near-wall
    R* blend = nullptr;              // per vertex: distance - half support
    int* base_t = nullptr;           // per vertex: top-hat stencil base (3)
    R* W_t = nullptr;                // per vertex: top-hat weights (2 per axis)
};

// The first pair that reaches a tile (outside a workgroup's table) lists it
// as touched and flags it as carrying spread force.
__device__ __forceinline__ void touch_tile(int t, int* act, int* na, unsigned char* flags)
{
    act[atomicAdd(na, 1)] = t;
    flags[t] = 1;
}

__device__ __forceinline__ void count_tiles(const StencilTiles& st, const Grid& gd, int* keys, int* cnt)
{
#pragma unroll 1
    for (int i = 0; i < st.ax[0].m; ++i)
#pragma unroll 1
        for (int j = 0; j < st.ax[1].m; ++j)
#pragma unroll 1
            for (int k = 0; k < st.ax[2].m; ++k) {
                const int sl = hash_insert(keys, st.at(gd, i, j, k));
                if (sl >= 0) atomicAdd(&cnt[sl], 1);
            }
}

__device__ __forceinline__ void fill_tiles(const StencilTiles& st, const Grid& gd, int entry, const int* keys,
                                           const int* cnt, int* rank, int* count, int* list, int cap, int* act,
                                           int* na, unsigned char* flags, int* status, int* pairs, int* big,
                                           int* nbig)
{
#pragma unroll 1
    for (int i = 0; i < st.ax[0].m; ++i)
#pragma unroll 1
        for (int j = 0; j < st.ax[1].m; ++j)
#pragma unroll 1
            for (int k = 0; k < st.ax[2].m; ++k) {
                const int tt = st.at(gd, i, j, k);
                const int sl = hash_find(keys, tt);
                int slot;
                if (sl >= 0) {
                    slot = cnt[sl] + atomicAdd(&rank[sl], 1);
                } else {  // the workgroup's table was full for this tile
                    slot = atomicAdd(&count[tt], 1);
                    if (slot == 0) touch_tile(tt, act, na, flags);
                    if (slot == kSmallTile) big[atomicAdd(nbig, 1)] = tt;
                    atomicAdd(pairs, 1);
                    if (slot >= cap) atomicOr(&status[2], 1);
                }
                if (slot < cap) list[(long)tt * cap + slot] = entry;
            }
}

// Support-box base (wrapped on periodic axes) and the S weights of one axis.
template <class R>
__device__ __forceinline__ int axis_stencil(int kernel, int S, R h, R x, int n, int per, R* w)
{
    const int raw = (int)ceil(x - h);
#pragma unroll 1
    for (int k = 0; k < 4; ++k)
        if (k < S) w[k] = dirac<R>(kernel, (R)(raw + k) - x);
    int bw = raw;
    if (per) {
        bw = raw % n;
        if (bw < 0) bw += n;
    }
    return bw;
}

// Binning in one pass: stencil base and 1-D weights per vertex, the tiles
// its stencil touches, and every (tile, vertex) pair written into the
// tile's fixed-capacity list (entry values as described at kNearKey; a
// near-wall vertex also gets its blend weight, top-hat base and weights and
// the tiles of its top-hat stencil).  Per workgroup the pairs are counted
// per distinct tile in LDS, one global atomic per tile reserves the
// workgroup's run, and the slots inside the run come from LDS.  The order
// within a list is whatever the atomics give; spreading sorts.  The add that
// takes a tile past kSmallTile entries appends it to the big-tile list.
template <class R>
struct BinArgs {
    Grid gd;
    int kernel, S;
    double h;
    long NV;
    const R* pos;
    int* base;
    R* W;
    int *count, *list;
    int cap;
    int *act, *na;
    unsigned char* flags;
    int *status, *pairs, *big, *nbig;
    NearView<R> nw;
    int xcd;  // workgroup order (ibm_block)
};

template <class R>
__device__ __forceinline__ void bin_body(long bx, const BinArgs<R>& A)
{
    const Grid& gd = A.gd;
    const int kernel = A.kernel, S = A.S, cap = A.cap;
    const R h = (R)A.h;
    const long NV = A.NV;
    const R* __restrict__ pos = A.pos;
    int* __restrict__ base = A.base;
    R* __restrict__ W = A.W;
    int* __restrict__ count = A.count;
    int* __restrict__ list = A.list;
    int* __restrict__ act = A.act;
    int* __restrict__ na = A.na;
    unsigned char* __restrict__ flags = A.flags;
    int* __restrict__ status = A.status;
    int* __restrict__ pairs = A.pairs;
    int* __restrict__ big = A.big;
    int* __restrict__ nbig = A.nbig;
    const NearView<R>& nw = A.nw;
    __shared__ int keys[kHash];
    __shared__ int cnt[kHash];  // pairs per slot, then the slot's first index in the tile's list
    __shared__ int rank[kHash];
    __shared__ int fresh[kHash];  // tiles this workgroup reached first
    __shared__ int occ[kHash];    // occupied slots, compacted
    __shared__ int n_fresh, n_pairs, fresh_base, n_occ;
    hash_clear(keys, cnt, rank);
    if (threadIdx.x == 0) {
        n_fresh = 0;
        n_pairs = 0;
        n_occ = 0;
    }
    __syncthreads();
    const long v = bx * blockDim.x + threadIdx.x;
    StencilTiles st, stt;
    int e0 = -1, e1 = -1;  // list entries of the kernel part and the top-hat part (-1: none)
    // a vertex whose position is not finite (a membrane that blew up) or
    // absurdly far off touches no tile: zero weights at base 0, reported
    bool off = false;
    if (v < NV)
        for (int a = 0; a < 3; ++a) off = off || !(fabs(pos[3 * v + a]) < R(1e9));
    if (v < NV && off) {
        for (int a = 0; a < 3; ++a) {
            base[3 * v + a] = 0;
            for (int k = 0; k < 4; ++k) W[12 * v + 4 * a + k] = R(0.0);
        }
        if (nw.on) {
            nw.near[v] = 0;
            nw.blend[v] = R(1.0);
        }
        atomicOr(&status[2], 4);
    }
    if (v < NV && !off) {
        int b[3];
        for (int a = 0; a < 3; ++a) {
            R w[4] = {R(0.0), R(0.0), R(0.0), R(0.0)};
            b[a] = axis_stencil<R>(kernel, S, h, pos[3 * v + a], gd.n[a], gd.per[a], w);
            base[3 * v + a] = b[a];
            for (int k = 0; k < 4; ++k) W[12 * v + 4 * a + k] = w[k];
        }
        st.init(gd, S, b);
        e0 = (int)v;
        if (nw.on) {
            // the wall distance in FP64 (the boundary's own arithmetic)
            const double p[3] = {(double)pos[3 * v], (double)pos[3 * v + 1], (double)pos[3 * v + 2]};
            const bool nr = bnd_less_than(nw.b, p, A.h + 1.);
            const R wb = (R)(bnd_distance(nw.b, p) - A.h);
            nw.near[v] = nr;
            nw.blend[v] = wb;
            if (nr) {
                e0 = wb > R(0) ? (kNearKey | (int)(2 * v)) : -1;
                if (wb < R(1)) {
                    int bt[3];
                    for (int a = 0; a < 3; ++a) {
                        R w[4];
                        bt[a] = axis_stencil<R>(BL_KERNEL_TOPHAT, 2, R(1.0), pos[3 * v + a], gd.n[a], gd.per[a], w);
                        nw.base_t[3 * v + a] = bt[a];
                        nw.W_t[6 * v + 2 * a] = w[0];
                        nw.W_t[6 * v + 2 * a + 1] = w[1];
                    }
                    stt.init(gd, 2, bt);
                    e1 = kNearKey | (int)(2 * v + 1);
                }
            }
        }
        // runtime loops: the kernel runs once per launch on few workgroups,
        // so compact code (instruction fetch) beats unrolling
        if (e0 >= 0) count_tiles(st, gd, keys, cnt);
        if (e1 >= 0) count_tiles(stt, gd, keys, cnt);
    }
    __syncthreads();
    // one global atomic per distinct tile; the occupied slots are compacted
    // first so that each lane waits for one atomic, not one per slot it scans
    for (int i = threadIdx.x; i < kHash; i += blockDim.x)
        if (cnt[i] > 0) occ[atomicAdd(&n_occ, 1)] = i;
    __syncthreads();
    for (int k = threadIdx.x; k < n_occ; k += blockDim.x) {
        const int i = occ[k];
        const int c = cnt[i];
        const int t = keys[i];
        const int old = atomicAdd(&count[t], c);
        cnt[i] = old;
        if (old == 0) fresh[atomicAdd(&n_fresh, 1)] = t;
        if (old <= kSmallTile && old + c > kSmallTile) big[atomicAdd(nbig, 1)] = t;
        atomicAdd(&n_pairs, c);
        if (old + c > cap) atomicOr(&status[2], 1);
    }
    __syncthreads();
    // the touched-list reservation is issued before the list fill and its
    // result read after it, so the two round trips overlap
    int fb = 0;
    if (threadIdx.x == 0) {
        if (n_fresh) fb = atomicAdd(na, n_fresh);
        if (n_pairs) atomicAdd(pairs, n_pairs);
    }
    if (e0 >= 0) fill_tiles(st, gd, e0, keys, cnt, rank, count, list, cap, act, na, flags, status, pairs, big, nbig);
    if (e1 >= 0) fill_tiles(stt, gd, e1, keys, cnt, rank, count, list, cap, act, na, flags, status, pairs, big, nbig);
    if (threadIdx.x == 0) fresh_base = fb;
    __syncthreads();
    for (int i = threadIdx.x; i < n_fresh; i += blockDim.x) {
        act[fresh_base + i] = fresh[i];
        flags[fresh[i]] = 1;
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_bin(BinArgs<R> a)
{
    bin_body<R>(ibm_block(blockIdx.x, gridDim.x, a.xcd), a);
}

__global__ void k_clear_bins(int* count, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) count[i] = 0;
}

__global__ void k_zero_u8(unsigned char* p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_fill_i32(int* p, long n, int v)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}


// ------------------------------------------------------------ membrane
template <class R>
struct ForceView {
    int np;
    const long *p_voff, *p_toff, *p_loff, *p_poff;
    const int *p_shape, *p_model, *p_bug;
    const double* p_params;
    const double* p_geom;
    const int *s_voff, *s_toff, *s_loff, *s_poff, *s_vtoff, *s_vloff, *s_vpoff;
    const int* s_tri;
    const double *s_tri_area0, *s_tri_geo0;
    const int* s_link;
    const double* s_link_len;
    const int* s_pair;
    const double* s_pair_cs0;
    const int *s_vt_head, *s_vt, *s_vl_head, *s_vl, *s_vp_head, *s_vp;
    const double* s_area0;
    const R* pos;
    const int *v_part, *t_part, *r_part, *l_part;  // particle of each vertex / triangle / pair / link
    const int4* v_inc;  // per vertex: {link list start, links, triangle list start, triangles},
                        // {pair list start, pairs, particle, 0}; counts of unused kinds are 0
};

// Skalak / NeoHookean in-plane forces of one triangle
// (src/fsi/DeformableCapsule.hh:60-150), with the reference's
// c = l' sin(phi) / (l' sin(phi0)) when `bug` is set.
template <class R>
__device__ void capsule_triangle(V3<R> p0, V3<R> p1, V3<R> p2, const double* geo0, double area0_, double G_,
                                 double Cc_, bool skalak, bool bug, int only, V3<R> out[3])
{
    using V = V3<R>;
    const R area0 = (R)area0_, G = (R)G_, Cc = (R)Cc_;
    const V e0 = p1 - p0, e1 = p2 - p0;
    const R l = norm(e0), lp = norm(e1);
    const R l0 = (R)geo0[0], l0p = (R)geo0[1], cos_phi0 = (R)geo0[2], sin_phi0 = (R)geo0[3];
    const R e0e1 = dot(e0, e1);
    const R sin_phi = norm(cross(e0, e1)) / (l * lp);
    const R cos_phi = e0e1 / (l * lp);
    const R a = l / l0;
    const R b = (lp * cos_phi / l0p - l * cos_phi0 / l0) / sin_phi0;
    const R c = bug ? lp * sin_phi / (lp * sin_phi0) : lp * sin_phi / (l0p * sin_phi0);
    const R D11 = (a * a - R(1.)) / R(2.);
    const R D22 = (b * b + c * c - R(1)) / R(2.);
    const R D12 = a * b / R(2.);
    const R I1 = R(2.) * (D11 + D22);
    const R I2 = I1 + R(4) * (D11 * D22 - D12 * D12);
    V gl[3], glp[3], gc[3];
    gl[0] = e0 / (-l);
    gl[1] = e0 / l;
    gl[2] = {R(0), R(0), R(0)};
    glp[0] = e1 / (-lp);
    glp[1] = {R(0), R(0), R(0)};
    glp[2] = e1 / lp;
    gc[0] = (-R(1.) / (l * lp) + e0e1 / (l * l * l * lp)) * e0 + (-R(1.) / (l * lp) + e0e1 / (l * lp * lp * lp)) * e1;
    gc[1] = (R(1.) / (l * lp)) * e1 - (e0e1 / (l * l * l * lp)) * e0;
    gc[2] = (R(1.) / (l * lp)) * e0 - (e0e1 / (l * lp * lp * lp)) * e1;
    const R f1 = l / (l0 * l0);
    const R f2a = lp / (l0p * l0p) - l * cos_phi * cos_phi0 / (l0 * l0p);
    const R f2b = l * (cos_phi0 * cos_phi0) / (l0 * l0) - lp * cos_phi * cos_phi0 / (l0p * l0);
    const R f2c = l * lp * cos_phi0 / (l0 * l0p);
    const R f2 = R(1.) / (sin_phi0 * sin_phi0);
    const R f3 = R(1.) / (R(2.) * l0p * sin_phi0);
    const R f3a = lp * cos_phi / l0p - R(2.) * l * cos_phi0 / l0;
    const R f3b = l * cos_phi / l0p;
    const R f3c = l * lp / l0p;
    for (int i = 0; i < 3; ++i) {
        if (only >= 0 && i != only) continue;
        const V gD11 = f1 * gl[i];
        const V gD22 = f2 * (f2a * glp[i] + f2b * gl[i] - f2c * gc[i]);
        const V gD12 = f3 * (f3a * gl[i] + f3b * glp[i] + f3c * gc[i]);
        const V gI1 = R(2.) * (gD11 + gD22);
        const V gI2 = (R(2.) + R(4.) * D22) * gD11 + (R(2.) + R(4.) * D11) * gD22 - (R(8.) * D12) * gD12;
        if (skalak)
            out[i] = (-G / R(2.) * area0) * ((I1 + R(1)) * gI1 + (Cc * I2 - R(1)) * gI2);
        else
            out[i] = (-G / R(2.) * area0) * (gI1 - (R(1.) / ((R(1) + I2) * (R(1) + I2))) * gI2);
    }
}

// Area (local + global) and volume constraint forces of one triangle
// (src/fsi/RBCParticle.hh:176-209, src/fsi/TriangleUtils.h:60-129).
template <class R>
__device__ void area_volume_triangle(V3<R> v0, V3<R> v1, V3<R> v2, double area0_, double A0tot_, double A_, double V_,
                                     double k_ag_, double k_al_, double k_v_, double V0_, V3<R> out[3])
{
    const R area0 = (R)area0_, A0tot = (R)A0tot_, A = (R)A_, V = (R)V_, k_ag = (R)k_ag_, k_al = (R)k_al_,
            k_v = (R)k_v_, V0 = (R)V0_;
    const R v01v11 = v0.y - v1.y, v02v22 = v0.z - v2.z, v02v12 = v0.z - v1.z;
    const R v01v21 = v0.y - v2.y, v00v20 = v0.x - v2.x, v00v10 = v0.x - v1.x;
    const R a = v01v11 * v02v22 - v02v12 * v01v21;
    const R b = v02v12 * v00v20 - (v0.x - v1.x) * v02v22;
    const R c = v00v10 * v01v21 - v01v11 * v00v20;
    const R la = R(0.5) * sqrt(a * a + b * b + c * c);
    const R denom = R(1.) / (R(4.) * la);
    V3<R> g0, g1, g2;
    g0.x = (b * (v2.z - v1.z) + c * (v1.y - v2.y)) * denom;
    g0.y = (a * (v1.z - v2.z) + c * (v2.x - v1.x)) * denom;
    g0.z = (a * (v2.y - v1.y) + b * (v1.x - v2.x)) * denom;
    g1.x = (b * v02v22 + c * (v2.y - v0.y)) * denom;
    g1.y = (a * (v2.z - v0.z) + c * v00v20) * denom;
    g1.z = (a * v01v21 + b * (v2.x - v0.x)) * denom;
    g2.x = (b * (v1.z - v0.z) + c * v01v11) * denom;
    g2.y = (a * (v02v12) + c * (v1.x - v0.x)) * denom;
    g2.z = (a * (v1.y - v0.y) + b * v00v10) * denom;
    R df_area = R(0.);
    df_area += k_ag * (A0tot - A) / A0tot;
    df_area += k_al * (area0 - la) / area0;
    const R df_vol = k_v * (V0 - V) / V0;
    out[0] = out[0] + df_area * g0;
    out[1] = out[1] + df_area * g1;
    out[2] = out[2] + df_area * g2;
    out[0] = out[0] + df_vol * (cross(v1, v2) / R(6.));
    out[1] = out[1] + df_vol * (cross(v2, v0) / R(6.));
    out[2] = out[2] + df_vol * (cross(v0, v1) / R(6.));
}

// Force of one link on its first vertex (the second gets the negative):
// worm-like chain + power-law repulsion (src/fsi/RBCParticle.hh:149-171).
// The link's coefficients depend on the rest length and the modulus only
// and are formed in FP64.
template <class R>
__device__ V3<R> link_force(const ForceView<R>& f, int s, long vo, const double* P, long link)
{
    const double m = 2;
    const double x0 = 1. / 2.2;
    const double ks_coeff1 = 0.25 * (6. * x0 + (3 * x0 * x0 - 2 * x0 * x0 * x0) / ((1 - x0) * (1 - x0)));
    const double ks_coeff2 = x0 / (2. * (1 - x0) * (1 - x0) * (1 - x0)) + 0.25 * (1. - 1. / ((1 - x0) * (1 - x0)));
    const long sl = f.s_loff[s] + link;
    const double l0d = f.s_link_len[sl];
    const double kp_coeff1 = -1. / pow(l0d, m);
    const double kp_coeff2 = (m + 1) / pow(l0d, m);
    const double G_lhs = 4. * l0d * P[0] / sqrt(3.0);
    const R ks = (R)(-kp_coeff1 * G_lhs / (ks_coeff1 * kp_coeff2 - ks_coeff2 * kp_coeff1));
    const R kp = (R)(ks_coeff1 * G_lhs / (ks_coeff1 * kp_coeff2 - ks_coeff2 * kp_coeff1));
    const R l0 = (R)l0d;
    const int a = f.s_link[2 * sl], b = f.s_link[2 * sl + 1];
    const V3<R> dv = ld3(f.pos + 3 * (vo + a)) - ld3(f.pos + 3 * (vo + b));
    const R L = norm(dv);
    const R x = L / (l0 * R(2.2));
    const R df = -R(0.25) * ks * (R(6.) * x + (R(3.) * x * x - R(2.) * x * x * x) / ((R(1) - x) * (R(1) - x))) +
                 kp / pow(L, (R)m);
    return (df / L) * dv;
}

// Forces of one triangle on its three vertices.
template <class R>
__device__ void tri_forces(const ForceView<R>& f, int p, int s, int model, long vo, const double* P, long tri,
                           V3<R> out[3])
{
    const long st = f.s_toff[s] + tri;
    const int i0 = f.s_tri[3 * st], i1 = f.s_tri[3 * st + 1], i2 = f.s_tri[3 * st + 2];
    const V3<R> v0 = ld3(f.pos + 3 * (vo + i0)), v1 = ld3(f.pos + 3 * (vo + i1)), v2 = ld3(f.pos + 3 * (vo + i2));
    out[0] = out[1] = out[2] = V3<R>{R(0), R(0), R(0)};
    const double area0 = f.s_tri_area0[st];
    if (model == BL_MODEL_CAPSULE_SKALAK || model == BL_MODEL_CAPSULE_NEOHOOKEAN || model == 3)
        capsule_triangle<R>(v0, v1, v2, f.s_tri_geo0 + 4 * st, area0, P[0], P[1],
                            model != BL_MODEL_CAPSULE_NEOHOOKEAN, f.p_bug[p] != 0, -1, out);
    if (model == BL_MODEL_RBC || model == 3)
        area_volume_triangle<R>(v0, v1, v2, area0, f.s_area0[s], f.p_geom[11 * p], f.p_geom[11 * p + 1], P[2], P[3],
                                P[4], P[5], out);
}

// Bending forces of one dihedral pair on its four vertices.
template <class R>
__device__ void pair_forces(const ForceView<R>& f, int s, long vo, const double* P, long pair, V3<R> out[4])
{
    using V = V3<R>;
    const long sp = f.s_poff[s] + pair;
    const int* q = f.s_pair + 4 * sp;
    const V v0 = ld3(f.pos + 3 * (vo + q[0])), v1 = ld3(f.pos + 3 * (vo + q[1]));
    const V v2 = ld3(f.pos + 3 * (vo + q[2])), v3 = ld3(f.pos + 3 * (vo + q[3]));
    const V e0 = v1 - v0, e1 = v2 - v0, e2 = v3 - v0;
    const R e1n = norm(e1);
    const V e1u = e1 / e1n;
    const V n1 = cross(e0, e1), n2 = cross(e1, e2);
    const R n1n = norm(n1), n2n = norm(n2);
    const V n1u = n1 / n1n, n2u = n2 / n2n;
    const R cs = dot(n1u, n2u);
    const R sn = ((dot(n1 - n2, v1 - v3) > R(0)) ? R(1.0) : -R(1.0)) * norm(cross(n1u, n2u));
    const R pref = (R)P[6] * (-(R)f.s_pair_cs0[2 * sp] * sn + (R)f.s_pair_cs0[2 * sp + 1] * cs);
    out[0] = pref * ((dot(e1u, e1 - e2) / n2n) * n2u + (dot(e1u, e1 - e0) / n1n) * n1u);
    out[1] = pref * ((-e1n / n1n) * n1u);
    out[2] = pref * ((dot(e1u, e2) / n2n) * n2u + (dot(e1u, e0) / n1n) * n1u);
    out[3] = pref * ((-e1n / n2n) * n2u);
}

// One thread per element, over the concatenation [triangles | pairs | links]
// of all particles: the element's forces on its vertices go to its own slots
// of `ec` (9 values per triangle, 12 per pair, 3 per link).  Elements that
// the particle's model does not use are skipped.
template <class R>
struct ElemArgs {
    ForceView<R> f;
    long NT, NP, NL, ec_pair, ec_link;
    R* ec;
    int xcd;  // workgroup order (ibm_block)
};

template <class R>
__device__ __forceinline__ void elem_body(long bx, const ElemArgs<R>& A)
{
    const ForceView<R>& f = A.f;
    const long NT = A.NT, NP = A.NP, NL = A.NL, ec_pair = A.ec_pair, ec_link = A.ec_link;
    R* __restrict__ ec = A.ec;
    const long e = bx * blockDim.x + threadIdx.x;
    if (e < NT) {
        const int p = f.t_part[e];
        V3<R> out[3];
        tri_forces<R>(f, p, f.p_shape[p], f.p_model[p], f.p_voff[p], f.p_params + 8 * p, e - f.p_toff[p], out);
        for (int k = 0; k < 3; ++k) st3(ec + 9 * e + 3 * k, out[k]);
    } else if (e < NT + NP) {
        const long g = e - NT;
        const int p = f.r_part[g];
        const int model = f.p_model[p];
        if (model != BL_MODEL_RBC && model != 3) return;
        V3<R> out[4];
        pair_forces<R>(f, f.p_shape[p], f.p_voff[p], f.p_params + 8 * p, g - f.p_poff[p], out);
        for (int k = 0; k < 4; ++k) st3(ec + ec_pair + 12 * g + 3 * k, out[k]);
    } else if (e < NT + NP + NL) {
        const long g = e - NT - NP;
        const int p = f.l_part[g];
        if (f.p_model[p] != BL_MODEL_RBC) return;
        st3(ec + ec_link + 3 * g, link_force<R>(f, f.p_shape[p], f.p_voff[p], f.p_params + 8 * p, g - f.p_loff[p]));
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_elem_forces(ElemArgs<R> a)
{
    elem_body<R>(ibm_block(blockIdx.x, gridDim.x, a.xcd), a);
}

// Force on every vertex: the contributions of its incident elements --
// links (src/fsi/RBCParticle.hh:149-171), triangles
// (src/fsi/DeformableCapsule.hh:60-150, src/fsi/RBCParticle.hh:176-209) and
// pairs (src/fsi/RBCParticle.hh:214-246) -- from the per-element buffer.
// The incidence lists hold element * 4 + the vertex's slot in the element.
// Four lanes per vertex: lane l sums entries l, l+4, ... of the vertex's
// list [links | triangles | pairs] in order, then a fixed butterfly adds the
// lanes, so the result does not depend on scheduling.  No atomics.
// Lane 0 starts from the wall-interaction decorators' forces, which the
// reference applies before the membrane forces, and the sum ends with the
// particle-particle forces (`fpp`), which it adds after them
// (src/fsi/ImmersedBoundaryDynamics_forces.hh:293-301).
// GL (2, 4 or 8) lanes per vertex, a launch parameter (bl_ibm_set_launch):
// the lane count fixes the butterfly, so results agree to rounding across
// lane counts and are bitwise repeatable for each.
constexpr int kGatherLanes = 4;  // the default

template <class R>
struct GatherArgs {
    ForceView<R> f;
    long NV, ec_pair, ec_link;
    const R* ec;
    R* force;
    DecoSet deco;
    const R* fpp;
    const R* ext;  // host-given vertex forces, added after the decorators (or null)
    int xcd;       // workgroup order (ibm_block)
};

template <class R, int GL>
__device__ __forceinline__ void gather_body(long bx, const GatherArgs<R>& A)
{
    const ForceView<R>& f = A.f;
    const long NV = A.NV, ec_pair = A.ec_pair, ec_link = A.ec_link;
    const R* __restrict__ ec = A.ec;
    R* __restrict__ force = A.force;
    const DecoSet& deco = A.deco;
    const R* __restrict__ fpp = A.fpp;
    const long gid = bx * blockDim.x + threadIdx.x;
    const long v = gid / GL;
    const int lane = (int)(gid % GL);
    R c0 = R(0), c1 = R(0), c2 = R(0);
    if (deco.n && lane == 0 && v < NV) {
        // the decorators' potentials and wall geometry in FP64
        const double p[3] = {(double)f.pos[3 * v], (double)f.pos[3 * v + 1], (double)f.pos[3 * v + 2]};
        double w[3] = {0, 0, 0};
        deco_apply(deco, p, w);
        c0 = (R)w[0];
        c1 = (R)w[1];
        c2 = (R)w[2];
    }
    if (A.ext && lane == 0 && v < NV) {
        c0 += A.ext[3 * v];
        c1 += A.ext[3 * v + 1];
        c2 += A.ext[3 * v + 2];
    }
    if (v < NV) {
        // the vertex's incidence lists in one load (built at init), so the
        // element forces are two dependent loads away
        const int4 ia = f.v_inc[2 * v], ib = f.v_inc[2 * v + 1];
        const int p = ib.z;
        const int nl = ia.y, nt = ia.w, npr = ib.y;
        const int* Ll = f.s_vl + ia.x;
        const int* Lt = f.s_vt + ia.z;
        const int* Lp = f.s_vp + ib.x;
        const R* El = ec + ec_link + 3 * f.p_loff[p];
        const R* Et = ec + 9 * f.p_toff[p];
        const R* Ep = ec + ec_pair + 12 * f.p_poff[p];
        for (int e = lane; e < nl + nt + npr; e += GL) {
            if (e < nl) {
                const int code = Ll[e];
                const R* x = El + 3 * (code >> 2);
                const R sg = (code & 3) == 0 ? R(1.0) : -R(1.0);
                c0 += sg * x[0];
                c1 += sg * x[1];
                c2 += sg * x[2];
            } else if (e < nl + nt) {
                const int code = Lt[e - nl];
                const R* x = Et + 9 * (code >> 2) + 3 * (code & 3);
                c0 += x[0];
                c1 += x[1];
                c2 += x[2];
            } else {
                const int code = Lp[e - nl - nt];
                const R* x = Ep + 12 * (code >> 2) + 3 * (code & 3);
                c0 += x[0];
                c1 += x[1];
                c2 += x[2];
            }
        }
    }
#pragma unroll
    for (int o = 1; o < GL; o <<= 1) {
        c0 += __shfl_xor(c0, o);
        c1 += __shfl_xor(c1, o);
        c2 += __shfl_xor(c2, o);
    }
    if (lane == 0 && v < NV) {
        if (fpp) {
            c0 += fpp[3 * v];
            c1 += fpp[3 * v + 1];
            c2 += fpp[3 * v + 2];
        }
        force[3 * v] = c0;
        force[3 * v + 1] = c1;
        force[3 * v + 2] = c2;
    }
}

template <class R, int GL>
__global__ void __launch_bounds__(256) k_vertex_gather(GatherArgs<R> a)
{
    gather_body<R, GL>(ibm_block(blockIdx.x, gridDim.x, a.xcd), a);
}

// ------------------------------------------------- velocity staging
// The progs' computeVelocity pass (staged_velocity), restricted to the
// cells of the tiles this step's stencils touch (`touched`, the step's tile
// flags).  `lv` carries the force field and flags of the last collision.
// One thread per cell in z-fastest order, so population reads coalesce as
// in the collide kernel and the staged vectors are written contiguously.
// (With the staging carried by the collision, k_collide_stream writes the
// same values from the populations it loads.)
struct StageArgs {
    LatView lv;
    const unsigned char* touched;
    void* ut;  // Vel4 of the storage's compute type
};

template <class S, bool SWAPPED>
__device__ __forceinline__ void stage_body(unsigned bx, const StageArgs& A)
{
    using C = typename Pop<S>::C;
    using FT = typename Pop<S>::F;
    const LatView& lv = A.lv;
    const unsigned char* __restrict__ touched = A.touched;
    Vel4<C>* __restrict__ ut = (Vel4<C>*)A.ut;
    const unsigned c = bx * blockDim.x + threadIdx.x;
    if (c >= lv.ncell) return;
    int x, y, z;
    decode(c, lv, x, y, z);
    const unsigned t = ((unsigned)(x >> 2) * lv.nty + (y >> 2)) * lv.ntz + (z >> 2);
    if (!touched[t]) return;
    const unsigned char m = lv.mask ? lv.mask[c] : (unsigned char)M_FLUID;
    // the moments as the collision forms them (working form of the storage),
    // so that a staging carried by the collision gives the same values
    C rb = C(0), j[3] = {C(0), C(0), C(0)};
    if (m == M_FLUID) {
        Nbr nb;
        if (SWAPPED) nb.init(lv, x, y, z);
        C f[Q];
        load_post_w<S>(lv, nb, c, SWAPPED, f);
        wmoments<S>(f, rb, j);
    }
    const C invRho = C(1) / (rb + C(1));
    ut[c] = staged_velocity_m<FT, C>(lv, lv.flags, lv.force, m, c, x, y, z, j, invRho);
}

template <class S, bool SWAPPED>
__global__ void __launch_bounds__(256) k_stage_u(StageArgs a) { stage_body<S, SWAPPED>(blockIdx.x, a); }

// ------------------------------------------------------- interpolation
// U_k = sum over the stencil nodes of phi(dx) phi(dy) phi(dz) u
// (src/fsi/ImmersedBoundaryDynamics_interpolation.hh:14-78): a group of
// lanes per vertex, lane (j, k) walks the x-line of its (y, z) node pair
// (nodes outside a non-periodic face skipped, weight (phi_x phi_y) phi_z as
// the reference forms it), then the lanes are added by a fixed butterfly,
// all in FP64.  The lanes of one x-step read consecutive cells of a few
// z-rows.  With MOVE the vertex is then advanced, x += U (explicit Euler,
// src/fsi/DeformableParticle3D.hh:26-36).
// Explicit Euler step that reflects off the boundary
// (DeformableParticle3D::move_vertices, src/fsi/DeformableParticle3D.hh:26-50),
// in FP64 (the boundary's own arithmetic) for either vertex type.
template <class R>
__device__ __forceinline__ void move_reflect(const bl_boundary& b, R* pr, const R* ur)
{
#pragma clang fp contract(off)
    const double u[3] = {(double)ur[0], (double)ur[1], (double)ur[2]};
    const double old[3] = {(double)pr[0], (double)pr[1], (double)pr[2]};
    const double nw[3] = {old[0] + u[0], old[1] + u[1], old[2] + u[2]};
    double t;
    if (!bnd_trace_ray(b, old, nw, t)) {
        pr[0] = (R)nw[0];
        pr[1] = (R)nw[1];
        pr[2] = (R)nw[2];
        return;
    }
    double q[3] = {old[0] + t * u[0], old[1] + t * u[1], old[2] + t * u[2]};
    double n[3];
    bnd_normal(b, q, n);
    const double vn = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    for (int a = 0; a < 3; ++a) pr[a] = (R)(q[a] + (u[a] - 2. * vn * n[a]) * (1. - t));
}

// A near-wall vertex (blend weight w < 1) also interpolates with the top
// hat, lanes (j, k) < 2 over its 2x2x2 stencil, and takes (1 - w) u_tophat
// + w u_kernel, or u_tophat alone for w <= 0
// (src/fsi/ImmersedBoundaryDynamics_interpolation.hh:182-230).
template <class R>
struct InterpArgs {
    Grid gd;
    long NV;
    const int* base;
    const R* W;
    const void* ut;  // Vel4<C>
    R *vel, *pos;
    ListReset lr;
    NearView<R> nw;
    int xcd;  // workgroup order of the vertices (ibm_block; the list reset keeps the hardware's)
};

template <class C, int SW, bool MOVE, class R>
__device__ __forceinline__ void interp_body(long bx, int nbx, const InterpArgs<R>& A)
{
    const Grid& gd = A.gd;
    const long NV = A.NV;
    const int* __restrict__ base = A.base;
    const R* __restrict__ W = A.W;
    const Vel4<C>* __restrict__ ut = (const Vel4<C>*)A.ut;
    R* __restrict__ vel = A.vel;
    R* __restrict__ pos = A.pos;
    const NearView<R>& nw = A.nw;
    if (A.lr.act) list_reset(A.lr, (int)bx, nbx);  // the next binning's list, carried by this launch
    constexpr int LA = SW == 2 ? 2 : 4;
    constexpr int LPV = LA * LA;
    const long gid = ibm_block(bx, nbx, A.xcd) * blockDim.x + threadIdx.x;
    const long v = gid / LPV;
    const int r = (int)(gid % LPV);
    const int j = r / LA, k = r % LA;
    R a0 = R(0), a1 = R(0), a2 = R(0);
    if (v < NV && j < SW && k < SW) {
        const int yj = node_at(base[3 * v + 1], j, gd.n[1], gd.per[1]);
        const int zk = node_at(base[3 * v + 2], k, gd.n[2], gd.per[2]);
        if (yj >= 0 && zk >= 0) {
            const R* Wv = W + 12 * v;
            const R wy = Wv[4 + j], wz = Wv[8 + k];
            const int bx = base[3 * v];
#pragma unroll
            for (int i = 0; i < SW; ++i) {
                const int xi = node_at(bx, i, gd.n[0], gd.per[0]);
                if (xi < 0) continue;
                const R w = (Wv[i] * wy) * wz;
                const Vel4<C> u = ut[((size_t)xi * gd.n[1] + yj) * gd.n[2] + zk];
                a0 += w * (R)u.x;
                a1 += w * (R)u.y;
                a2 += w * (R)u.z;
            }
        }
    }
#pragma unroll
    for (int o = 1; o < LPV; o <<= 1) {
        a0 += __shfl_xor(a0, o);
        a1 += __shfl_xor(a1, o);
        a2 += __shfl_xor(a2, o);
    }
    if (nw.on && v < NV && nw.near[v] && nw.blend[v] < R(1)) {  // the same for all lanes of the vertex
        const R wb = nw.blend[v];
        R t0 = R(0), t1 = R(0), t2 = R(0);
        if (j < 2 && k < 2) {
            const int* bt = nw.base_t + 3 * v;
            const R* Wt = nw.W_t + 6 * v;
            const int yj = node_at(bt[1], j, gd.n[1], gd.per[1]);
            const int zk = node_at(bt[2], k, gd.n[2], gd.per[2]);
            if (yj >= 0 && zk >= 0) {
                const R wy = Wt[2 + j], wz = Wt[4 + k];
                for (int i = 0; i < 2; ++i) {
                    const int xi = node_at(bt[0], i, gd.n[0], gd.per[0]);
                    if (xi < 0) continue;
                    const R w = (Wt[i] * wy) * wz;
                    const Vel4<C> u = ut[((size_t)xi * gd.n[1] + yj) * gd.n[2] + zk];
                    t0 += w * (R)u.x;
                    t1 += w * (R)u.y;
                    t2 += w * (R)u.z;
                }
            }
        }
#pragma unroll
        for (int o = 1; o < LPV; o <<= 1) {
            t0 += __shfl_xor(t0, o);
            t1 += __shfl_xor(t1, o);
            t2 += __shfl_xor(t2, o);
        }
        if (wb <= R(0)) {
            a0 = t0;
            a1 = t1;
            a2 = t2;
        } else {
            a0 = (R(1.) - wb) * t0 + wb * a0;
            a1 = (R(1.) - wb) * t1 + wb * a1;
            a2 = (R(1.) - wb) * t2 + wb * a2;
        }
    }
    if (r == 0 && v < NV) {
        vel[3 * v] = a0;
        vel[3 * v + 1] = a1;
        vel[3 * v + 2] = a2;
        if (MOVE) {
            if (nw.on) {
                const R u[3] = {a0, a1, a2};
                move_reflect<R>(nw.b, pos + 3 * v, u);
            } else {
                pos[3 * v] += a0;
                pos[3 * v + 1] += a1;
                pos[3 * v + 2] += a2;
            }
        }
    }
}

template <class C, int SW, bool MOVE, class R>
__global__ void __launch_bounds__(256) k_interp(InterpArgs<R> a)
{
    interp_body<C, SW, MOVE, R>(blockIdx.x, gridDim.x, a);
}

// ----------------------------------------------------------- spreading
// Every cell of a touched 4x4x4 tile sums g + phi phi phi F_k over the
// tile's vertices in increasing vertex order, the order of the reference's
// vertex loop (src/fsi/ImmersedBoundaryDynamics_forces.hh:53-85), and writes
// its force; one lane per cell.  The tile's list is sorted first.  Then, 64
// entries at a time, lane e loads entry e and forms its weights on the
// tile's own 4 cells per axis (zero outside the stencil) and its force in
// registers (all zero past the list's end); two halves of 32 go through LDS
// in turn, and the cells walk each half in order, four entries per step,
// with broadcast LDS reads.
// One launch (k_spread) holds two kinds of workgroups:
//   small: tiles with at most kSmallTile entries (the touched-tile list,
//          skipping the others), one wave per tile, four independent waves
//          per workgroup, keys sorted in registers, one running sum from g
//          in vertex order;
//   big:   the big-tile list, one workgroup of four waves per tile; each
//          wave sorts a run of 64 keys in registers and the runs are merged
//          by rank in LDS (a bitonic sort in LDS past 256 entries); wave w
//          sums the w-th contiguous quarter of the sorted list and the four
//          partial sums are added to g in wave order.
// Each clears the tile's count once done, so the next binning starts from
// zero without a clearing pass.
constexpr int kChunk = 32;
constexpr int kSpreadWaves = 4;

// Orders the LDS accesses of one wave (a wave's LDS operations complete in
// issue order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Weights of list entry `e` on the tile's cells (rows 0-11: axis-major, 4
// cells per axis) and its force (rows 12-14), into registers.  The vertex's
// weights are loaded whole (not indexed by the stencil offset) so that no
// load waits for the base.  A near-wall vertex's kernel part carries the
// force times w (w < 1), its top-hat part the force times 1 - w (w > 0),
// the top hat's two weights per axis.
template <class R>
__device__ __forceinline__ void entry_weights(const Grid& gd, int Sw, const int org[3], int e,
                                              const int* __restrict__ base, const R* __restrict__ W,
                                              const R* __restrict__ vforce, const NearView<R>& nw, R ev[15])
{
    using R2 = typename Vec2<R>::T;
    const bool nr = (e & kNearKey) != 0;
    const long v = nr ? (long)((e & ~kNearKey) >> 1) : (long)e;
    const bool top = nr && (e & 1);
    R w[12];
    const int* bv;
    int sw;
    if (top) {
        const R2* W2 = (const R2*)(nw.W_t + 6 * v);
        for (int k = 0; k < 3; ++k) {
            const R2 p = W2[k];
            w[4 * k] = p.x;
            w[4 * k + 1] = p.y;
            w[4 * k + 2] = R(0.0);
            w[4 * k + 3] = R(0.0);
        }
        bv = nw.base_t + 3 * v;
        sw = 2;
    } else {
        const R2* W2 = (const R2*)(W + 12 * v);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const R2 p = W2[k];
            w[2 * k] = p.x;
            w[2 * k + 1] = p.y;
        }
        bv = base + 3 * v;
        sw = Sw;
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int b = bv[ax];
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            int d = org[ax] + l - b;
            if (gd.per[ax] && d < 0) d += gd.n[ax];
            const R* wa = w + 4 * ax;
            const R x = d == 0 ? wa[0] : (d == 1 ? wa[1] : (d == 2 ? wa[2] : wa[3]));
            ev[4 * ax + l] = (unsigned)d < (unsigned)sw ? x : R(0.0);
        }
    }
    R F0 = vforce[3 * v], F1 = vforce[3 * v + 1], F2 = vforce[3 * v + 2];
    if (nr) {
        const R wb = nw.blend[v];
        const bool scale = top ? wb > R(0) : wb < R(1);
        const R s = top ? R(1.) - wb : wb;
        if (scale) {
            F0 *= s;
            F1 *= s;
            F2 *= s;
        }
    }
    ev[12] = F0;
    ev[13] = F1;
    ev[14] = F2;
}

// Adds entries [0, cnt) of a staged half (cnt <= kChunk; the rows are zero
// from cnt up to the next multiple of 4, so a padded entry adds nothing)
// in entry order; the LDS reads of four entries are issued together.
template <class R>
__device__ __forceinline__ void chunk_sum(const R (*ew)[kChunk + 1], int cnt, int lx, int ly, int lz, R& a0, R& a1,
                                          R& a2)
{
    for (int e = 0; e < cnt; e += 4) {
        R w[4], f0[4], f1[4], f2[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            w[k] = (ew[lx][e + k] * ew[4 + ly][e + k]) * ew[8 + lz][e + k];
            f0[k] = ew[12][e + k];
            f1[k] = ew[13][e + k];
            f2[k] = ew[14][e + k];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            a0 += w[k] * f0[k];
            a1 += w[k] * f1[k];
            a2 += w[k] * f2[k];
        }
    }
}

// Sums the sorted entries [lo, hi) of a tile into a0..a2, 64 at a time;
// key(e) gives the e-th sorted entry.
template <class R, class KeyOf>
__device__ __forceinline__ void sum_entries(const Grid& gd, int Sw, const int org[3], int lo, int hi, KeyOf key,
                                            const int* __restrict__ base, const R* __restrict__ W,
                                            const R* __restrict__ vforce, const NearView<R>& nw,
                                            R (*ew)[kChunk + 1], int lane, int lx, int ly, int lz, R& a0, R& a1,
                                            R& a2)
{
    for (int c0 = lo; c0 < hi; c0 += 64) {
        const int e = c0 + lane;
        R ev[15];
#pragma unroll
        for (int r = 0; r < 15; ++r) ev[r] = R(0.0);
        if (e < hi) entry_weights(gd, Sw, org, key(e), base, W, vforce, nw, ev);
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int cnt = min(kChunk, hi - c0 - half * kChunk);
            if (cnt <= 0) break;
            if ((lane >> 5) == half)
#pragma unroll
                for (int r = 0; r < 15; ++r) ew[r][lane & (kChunk - 1)] = ev[r];
            wave_sync();
            chunk_sum(ew, cnt, lx, ly, lz, a0, a1, a2);
            wave_sync();
        }
    }
}

__device__ __forceinline__ void tile_origin(const Grid& gd, int t, int org[3])
{
    const int tz = t % gd.nt[2], rr = t / gd.nt[2], ty = rr % gd.nt[1], tx = rr / gd.nt[1];
    org[0] = tx * kTile;
    org[1] = ty * kTile;
    org[2] = tz * kTile;
}

template <class FT, class R>
__device__ __forceinline__ void store_cell(const LatView& lv, const Grid& gd, const int org[3], int lx, int ly,
                                           int lz, R a0, R a1, R a2, FT* __restrict__ Fout)
{
    const int x = org[0] + lx, y = org[1] + ly, z = org[2] + lz;
    if (x < gd.n[0] && y < gd.n[1] && z < gd.n[2]) {
        const long c = ((long)x * gd.n[1] + y) * gd.n[2] + z;
        Fout[c] = (FT)a0;
        Fout[lv.stride + c] = (FT)a1;
        Fout[2 * lv.stride + c] = (FT)a2;
    }
}

// ascending bitonic sort of one key per lane across the wave
__device__ __forceinline__ int wave_sort(int key, int lane)
{
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int other = __shfl_xor(key, stride);
            const bool up = (lane & size) == 0;
            const bool lower = (lane & stride) == 0;
            key = (lower == up) ? min(key, other) : max(key, other);
        }
    return key;
}

// number of entries of the ascending run r[0, 64) that are below `key`
__device__ __forceinline__ int rank_in_run(const int* r, int key)
{
    int lo = 0;  // count among r[0, 63), then r[63] when all of those are below
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
        if (r[lo + step - 1] < key) lo += step;
    return lo + (r[lo] < key ? 1 : 0);
}

template <class FT, class R>
__device__ __forceinline__ void spread_small(int blk, int nb, R (*ew_all)[15][kChunk + 1], const LatView& lv,
                                             const Grid& gd, int Sw, const int* __restrict__ active,
                                             const int* __restrict__ nactive, int* __restrict__ count, int cap,
                                             const int* __restrict__ list, const int* __restrict__ base,
                                             const R* __restrict__ W, const R* __restrict__ vforce,
                                             const NearView<R>& nw, FT* __restrict__ Fout)
{
    const int wv = threadIdx.x >> 6;
    R(*ew)[kChunk + 1] = ew_all[wv];
    const int na = *nactive;
    const int lane = threadIdx.x & 63;
    const int lx = lane >> 4, ly = (lane >> 2) & 3, lz = lane & 3;
    const int stride = nb * kSpreadWaves;
    int a = blk * kSpreadWaves + wv;
    int t = a < na ? active[a] : 0;
    while (a < na) {
        // the list's first slots are loaded with the count (not after it),
        // and the next tile's id while this one is summed
        const int raw = list[(long)t * cap + lane];
        const int n = count[t];
        const int a_next = a + stride;
        const int t_next = a_next < na ? active[a_next] : 0;
        // a big tile is its workgroup's; a count already cleared (0) is one
        // that workgroup has finished (a listed tile has at least one entry)
        if (n > 0 && n <= kSmallTile) {
            const int key = wave_sort(lane < n ? raw : 0x7fffffff, lane);
            int org[3];
            tile_origin(gd, t, org);
            R acc0 = (R)lv.g[0], acc1 = (R)lv.g[1], acc2 = (R)lv.g[2];
            sum_entries<R>(gd, Sw, org, 0, n, [&](int) { return key; }, base, W, vforce, nw, ew, lane, lx, ly, lz,
                           acc0, acc1, acc2);
            store_cell(lv, gd, org, lx, ly, lz, acc0, acc1, acc2, Fout);
            if (lane == 0) count[t] = 0;
        }
        a = a_next;
        t = t_next;
    }
}

template <class FT, class R>
__device__ __forceinline__ void spread_big(int blk, int nb, int* keys, int* runs, R (*ew_all)[15][kChunk + 1],
                                           R (*part)[3][64], const LatView& lv, const Grid& gd, int Sw,
                                           const int* __restrict__ big, const int* __restrict__ nbig,
                                           int* __restrict__ count, int cap, const int* __restrict__ list,
                                           const int* __restrict__ base, const R* __restrict__ W,
                                           const R* __restrict__ vforce, const NearView<R>& nw,
                                           FT* __restrict__ Fout)
{
    const int wv = threadIdx.x >> 6;
    R(*ew)[kChunk + 1] = ew_all[wv];
    const int nt = *nbig;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int lx = lane >> 4, ly = (lane >> 2) & 3, lz = lane & 3;
    for (int a = blk; a < nt; a += nb) {
        const int t = big[a];
        const int n = min(count[t], cap);  // a longer list was cut at binning and reported there
        const long o = (long)t * cap;
        if (n <= 64 * kSpreadWaves) {
            // four register-sorted runs, merged by rank (keys are distinct)
            const int e = wv * 64 + lane;
            const int key = wave_sort(e < n ? list[o + e] : 0x7fffffff, lane);
            runs[e] = key;
            __syncthreads();
            if (e < n) {
                int r = lane;
#pragma unroll
                for (int w2 = 0; w2 < kSpreadWaves; ++w2)
                    if (w2 != wv) r += rank_in_run(runs + 64 * w2, key);
                keys[r] = key;
            }
        } else {
            int n2 = 1;
            while (n2 < n) n2 <<= 1;
            for (int k = tid; k < n2; k += blockDim.x) keys[k] = k < n ? list[o + k] : 0x7fffffff;
            __syncthreads();
            for (int size = 2; size <= n2; size <<= 1)
                for (int stride2 = size >> 1; stride2 > 0; stride2 >>= 1) {
                    for (int k = tid; k < n2; k += blockDim.x) {
                        const int partner = k ^ stride2;
                        if (partner > k) {
                            const bool up = (k & size) == 0;
                            const int a0 = keys[k], a1 = keys[partner];
                            if ((a0 > a1) == up) {
                                keys[k] = a1;
                                keys[partner] = a0;
                            }
                        }
                    }
                    __syncthreads();
                }
        }
        __syncthreads();
        int org[3];
        tile_origin(gd, t, org);
        // wave w sums the contiguous range [lo, hi) of the sorted entries
        const int q = (n + kSpreadWaves - 1) / kSpreadWaves;
        const int lo = min(n, wv * q), hi = min(n, lo + q);
        R acc0 = R(0), acc1 = R(0), acc2 = R(0);
        sum_entries<R>(gd, Sw, org, lo, hi, [&](int k) { return keys[k]; }, base, W, vforce, nw, ew, lane, lx, ly,
                       lz, acc0, acc1, acc2);
        part[wv][0][lane] = acc0;
        part[wv][1][lane] = acc1;
        part[wv][2][lane] = acc2;
        __syncthreads();
        if (wv == 0) {
            R s0 = (R)lv.g[0], s1 = (R)lv.g[1], s2 = (R)lv.g[2];
            for (int w = 0; w < kSpreadWaves; ++w) {
                s0 += part[w][0][lane];
                s1 += part[w][1][lane];
                s2 += part[w][2][lane];
            }
            store_cell(lv, gd, org, lx, ly, lz, s0, s1, s2, Fout);
            if (lane == 0) count[t] = 0;
        }
        __syncthreads();  // keys and partial sums are rewritten for the next tile
    }
}

// Both kinds in one launch: workgroups [0, n_small) take the small tiles,
// the others the big ones.
template <class R>
struct SpreadArgs {
    LatView lv;
    Grid gd;
    int Sw, n_small;
    const int *active, *nactive, *big, *nbig;
    int* count;
    int cap;
    const int *list, *base;
    const R *W, *vforce;
    NearView<R> nw;
    void* Fout;  // force field in the storage's force type
    int xcd;     // order of the small-tile workgroups over the XCDs (ibm_block)
};

template <class S, class R>
__device__ __forceinline__ void spread_body(int bx, int nbx, const SpreadArgs<R>& A)
{
    using FT = typename Pop<S>::F;
    __shared__ int keys[kMaxList];
    __shared__ int runs[64 * kSpreadWaves];
    __shared__ R ew_all[kSpreadWaves][15][kChunk + 1];
    __shared__ R part[kSpreadWaves][3][64];
    if (bx < A.n_small)
        spread_small<FT, R>((int)ibm_block(bx, A.n_small, A.xcd), A.n_small, ew_all, A.lv, A.gd, A.Sw, A.active,
                            A.nactive, A.count, A.cap, A.list, A.base, A.W, A.vforce, A.nw, (FT*)A.Fout);
    else
        spread_big<FT, R>(bx - A.n_small, nbx - A.n_small, keys, runs, ew_all, part, A.lv, A.gd, A.Sw, A.big, A.nbig,
                          A.count, A.cap, A.list, A.base, A.W, A.vforce, A.nw, (FT*)A.Fout);
}

template <class S, class R>
__global__ void __launch_bounds__(64 * kSpreadWaves) k_spread(SpreadArgs<R> a)
{
    spread_body<S, R>(blockIdx.x, gridDim.x, a);
}

// ------------------------------------------- stages paired in one launch
// Two stages of a coupled step that do not read each other's output run as
// one launch: workgroups [0, n1) take the first, the others the second
// (the choice is uniform over a workgroup, so its barriers stay whole).
// Binning | element forces: both read the positions only.
template <class R>
__global__ void __launch_bounds__(256) k_bin_elem(BinArgs<R> b, int n1, ElemArgs<R> e)
{
    const int bx = blockIdx.x;
    if (bx < n1) bin_body<R>(ibm_block(bx, n1, b.xcd), b);
    else elem_body<R>(ibm_block(bx - n1, (int)gridDim.x - n1, e.xcd), e);
}

// Vertex force gather (element forces) | velocity staging (the binning's
// tiles, the last collision's field).
template <class S, bool SWAPPED, class R, int GL>
__global__ void __launch_bounds__(256) k_gather_stage(GatherArgs<R> ga, int n1, StageArgs sa)
{
    const int bx = blockIdx.x;
    if (bx < n1) gather_body<R, GL>(ibm_block(bx, n1, ga.xcd), ga);
    else stage_body<S, SWAPPED>(bx - n1, sa);
}

// Spreading (lists, weights and vertex forces) | interpolation + advection
// (weights and staged velocity; it moves the vertices, which the spreading
// does not read).
template <class S, int SW, class R>
__global__ void __launch_bounds__(256) k_spread_interp(SpreadArgs<R> sp, int n1, InterpArgs<R> ia)
{
    const int bx = blockIdx.x;
    if (bx < n1) spread_body<S, R>(bx, n1, sp);
    else interp_body<typename Pop<S>::C, SW, true, R>(bx - n1, (int)gridDim.x - n1, ia);
}

// ------------------------------------------------------- calibration
// shrink_rbc_volume (progs/common/calibrate_rbc_shape.h:13-72) with
// RBCParticle::relax_nodes (src/fsi/RBCParticle.hh:318-342) as its step,
// for every selected particle at once.  Node mass: area / 3 of each
// incident triangle (tri::triangle_area, src/fsi/TriangleUtils.h:52-55),
// summed in triangle order (the vertex's incidence list is in element
// order).  The step counter lives on the device so that the launches of
// any step are the same.
template <class R>
__global__ void __launch_bounds__(256) k_calib_mass(ForceView<R> f, long NV, const int* __restrict__ active,
                                                    double* __restrict__ mass, int* __restrict__ step)
{
#pragma clang fp contract(off)
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v == 0) *step += 1;
    if (v >= NV) return;
    const int p = f.v_part[v];
    if (!active[p]) return;
    const int s = f.p_shape[p];
    const long vo = f.p_voff[p];
    const long sv = f.s_voff[s] + (v - vo);
    const int* ht = f.s_vt_head + sv + s;
    const int* Lt = f.s_vt + f.s_vtoff[s] + ht[0];
    const int nt = ht[1] - ht[0];
    double M = 0.0;
    for (int k = 0; k < nt; ++k) {
        const long st = f.s_toff[s] + (Lt[k] >> 2);
        const D3 v0 = ld3d(f.pos + 3 * (vo + f.s_tri[3 * st])), v1 = ld3d(f.pos + 3 * (vo + f.s_tri[3 * st + 1])),
                 v2 = ld3d(f.pos + 3 * (vo + f.s_tri[3 * st + 2]));
        M += (0.5 * norm(cross(v1 - v0, v2 - v0))) / 3.;
    }
    mass[v] = M;
}

// One relax step of the selected particles: vel += F / M - damping vel,
// pos += vel; the first vertex of each sets vol_desired of this step (the
// ramp from vol_start to vol_final over the first n_it steps).  The step's
// arithmetic is FP64; the vertex state keeps the boundary's type.
template <class R>
__global__ void __launch_bounds__(256) k_calib_relax(long NV, const int* __restrict__ v_part,
                                                     const long* __restrict__ p_voff, const int* __restrict__ active,
                                                     const double* __restrict__ damping,
                                                     const double* __restrict__ vol_start,
                                                     const double* __restrict__ vol_final, int n_it,
                                                     const int* __restrict__ step, const double* __restrict__ mass,
                                                     const R* __restrict__ force, R* __restrict__ vel,
                                                     R* __restrict__ pos, double* __restrict__ params)
{
#pragma clang fp contract(off)
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NV) return;
    const int p = v_part[v];
    if (!active[p]) return;
    const int it = *step - 1;
    if (v == p_voff[p] && it <= n_it)
        params[8 * p + 5] = vol_start[p] + ((double)it / (double)n_it) * (vol_final[p] - vol_start[p]);
    const double M = mass[v], d = damping[p];
    for (int k = 0; k < 3; ++k) {
        const double u0 = (double)vel[3 * v + k];
        const double u = u0 + ((double)force[3 * v + k] / M - d * u0);
        vel[3 * v + k] = (R)u;
        pos[3 * v + k] = (R)((double)pos[3 * v + k] + u);
    }
}

// --------------------------------------------------------------- motion
template <class R>
__global__ void k_move(long NV, R* __restrict__ pos, const R* __restrict__ vel)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 3 * NV) return;
    pos[v] += vel[v];
}

// Per-particle area, signed volume, area-weighted centroid
// (src/fsi/DeformableParticle3D.hh:54-82) and bounding box
// (src/fsi/ParticleBase.hh:51-69) in FP64: each particle's triangles are cut
// into chunks of 256 (one workgroup each, fixed-order tree in LDS), the
// chunk partials are then summed in chunk order, so the result does not
// depend on scheduling.
// Chunk c (its row of `chunks`: first triangle, triangles, vertex range
// of its bounding-box share, all global indices) -> 11 partials.
template <class R>
struct GeomArgs {
    ForceView<R> f;
    const int4* chunks;
    const int* t_vert;
    double* partial;
    int nchunk, np;
    Grid gd;
    const int* p_choff;
    const long* p_voff;
    double *geom, *delta;
    R* pos;
    int remap;
    const int* p_model;
};

template <class R>
__device__ __forceinline__ void geom_partial_body(int c, const GeomArgs<R>& A)
{
    const ForceView<R>& f = A.f;
    const int4* __restrict__ chunks = A.chunks;
    const int* __restrict__ t_vert = A.t_vert;
    double* __restrict__ partial = A.partial;
    __shared__ double sh[11][256];
    const int4 ch = chunks[c];
    double acc[5] = {0, 0, 0, 0, 0};
    if ((int)threadIdx.x < ch.y) {
        const int* tv = t_vert + 3 * ((long)ch.x + threadIdx.x);
        const D3 v0 = ld3d(f.pos + 3 * (long)tv[0]), v1 = ld3d(f.pos + 3 * (long)tv[1]),
                 v2 = ld3d(f.pos + 3 * (long)tv[2]);
        const double dA = 0.5 * norm(cross(v1 - v0, v2 - v0));
        acc[0] = dA;
        acc[1] = dot(v0, cross(v1, v2)) / 6.;
        acc[2] = ((v0.x + v1.x + v2.x) / 3.0) * dA;
        acc[3] = ((v0.y + v1.y + v2.y) / 3.0) * dA;
        acc[4] = ((v0.z + v1.z + v2.z) / 3.0) * dA;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (long v = ch.z + threadIdx.x; v < ch.w; v += blockDim.x)
        for (int d = 0; d < 3; ++d) {
            const double x = (double)f.pos[3 * v + d];
            lo[d] = fmin(lo[d], x);
            hi[d] = fmax(hi[d], x);
        }
    for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] = acc[q];
    for (int d = 0; d < 3; ++d) {
        sh[5 + 2 * d][threadIdx.x] = lo[d];
        sh[6 + 2 * d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
            for (int d = 0; d < 3; ++d) {
                sh[5 + 2 * d][threadIdx.x] = fmin(sh[5 + 2 * d][threadIdx.x], sh[5 + 2 * d][threadIdx.x + w]);
                sh[6 + 2 * d][threadIdx.x] = fmax(sh[6 + 2 * d][threadIdx.x], sh[6 + 2 * d][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 11) partial[11 * (long)c + threadIdx.x] = sh[threadIdx.x][0];
}

template <class R>
__global__ void __launch_bounds__(256) k_geom_partial(GeomArgs<R> a) { geom_partial_body<R>(blockIdx.x, a); }

// Chunk partials -> geometry; the centre of mass is then mapped onto the
// periodic grid (src/fsi/ParticleBase.hh:113-136, src/fsi/Periodicity.h:83-87)
// and the whole particle shifted by the same amount.
template <class R>
__device__ __forceinline__ void geom_final_body(int p, const GeomArgs<R>& A)
{
    const Grid& gd = A.gd;
    const int* __restrict__ p_choff = A.p_choff;
    const double* __restrict__ partial = A.partial;
    const long* __restrict__ p_voff = A.p_voff;
    double* __restrict__ geom = A.geom;
    double* __restrict__ delta = A.delta;
    R* __restrict__ pos = A.pos;
    const int remap = A.remap;
    const int* __restrict__ p_model = A.p_model;
    __shared__ double sft[3];
    __shared__ double part[64][11];  // a batch of chunk partials, loaded by all lanes at once
    const int c0 = p_choff[p], c1 = p_choff[p + 1];
    double acc[5] = {0, 0, 0, 0, 0};
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int b = c0; b < c1; b += 64) {
        const int nb = min(64, c1 - b);
        if ((int)threadIdx.x < nb)
            for (int q = 0; q < 11; ++q) part[threadIdx.x][q] = partial[11 * (long)(b + threadIdx.x) + q];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < nb; ++k) {  // chunk order
                const double* P = part[k];
                for (int q = 0; q < 5; ++q) acc[q] += P[q];
                for (int d = 0; d < 3; ++d) {
                    lo[d] = fmin(lo[d], P[5 + 2 * d]);
                    hi[d] = fmax(hi[d], P[6 + 2 * d]);
                }
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* G = geom + 11 * p;
        const double area = acc[0];
        double com[3] = {acc[2] / area, acc[3] / area, acc[4] / area};
        for (int d = 0; d < 3; ++d) {
            sft[d] = 0.0;
            // a rigid particle maps its own centre of mass (rigid.hip)
            if (!remap || !gd.per[d] || p_model[p] == BL_MODEL_RIGID) continue;
            const double L = (double)gd.n[d];
            const double x0 = -0.5;
            const double r = com[d] - L * floor((com[d] - x0) * (1. / L));
            sft[d] = r - com[d];
            com[d] = r;
        }
        G[0] = area;
        G[1] = acc[1];
        for (int d = 0; d < 3; ++d) {
            G[2 + d] = com[d];
            G[5 + 2 * d] = lo[d] + sft[d];
            G[6 + 2 * d] = hi[d] + sft[d];
            delta[3 * p + d] = sft[d];
        }
    }
    __syncthreads();
    const double s0 = sft[0], s1 = sft[1], s2 = sft[2];
    if (s0 == 0.0 && s1 == 0.0 && s2 == 0.0) return;
    for (long v = p_voff[p] + threadIdx.x; v < p_voff[p + 1]; v += blockDim.x) {
        pos[3 * v] = (R)((double)pos[3 * v] + s0);
        pos[3 * v + 1] = (R)((double)pos[3 * v + 1] + s1);
        pos[3 * v + 2] = (R)((double)pos[3 * v + 2] + s2);
    }
}

template <class R>
__global__ void __launch_bounds__(64) k_geom_final(GeomArgs<R> a) { geom_final_body<R>(blockIdx.x, a); }

// --------------------------------------------------------- voxelization
// z-rays through integer (x, y) columns (RayTracer::find_intersections_z,
// src/fsi/RayTracer.hh:19-39,80-108).  Pass 0 counts hits per column,
// pass 1 writes them (z mapped into [-0.5, nz - 0.5) on a periodic z axis,
// as ZBuffer::put does, src/fsi/ZBuffer.h:39-49).
template <class R>
__global__ void k_ray_hits(ForceView<R> f, Grid gd, long NT, int pass, int* col_count, const int* col_off,
                           int* col_fill, double* col_z, long cap, int* status)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= NT) return;
    const int p = f.t_part[g];
    const int s = f.p_shape[p];
    const long vo = f.p_voff[p];
    const int* tri = f.s_tri + 3 * (f.s_toff[s] + (g - f.p_toff[p]));
    const D3 a = ld3d(f.pos + 3 * (vo + tri[0])), b = ld3d(f.pos + 3 * (vo + tri[1])),
             c = ld3d(f.pos + 3 * (vo + tri[2]));
    // a triangle with a non-finite corner, or wider than the lattice, is
    // not traced (its membrane diverged; the binning reports it)
    const double ext = fmax(fmax(fabs(a.x), fabs(b.x)), fmax(fmax(fabs(c.x), fabs(a.y)), fmax(fabs(b.y), fabs(c.y))));
    if (!(ext < 1e9)) return;
    const long x0 = (long)ceil(fmin(a.x, fmin(b.x, c.x))), x1 = (long)floor(fmax(a.x, fmax(b.x, c.x)));
    const long y0 = (long)ceil(fmin(a.y, fmin(b.y, c.y))), y1 = (long)floor(fmax(a.y, fmax(b.y, c.y)));
    if (x1 - x0 > (long)gd.n[0] + 4 || y1 - y0 > (long)gd.n[1] + 4) return;
    const D3 e1 = b - a, e2 = c - a;
    const D3 dir = {0, 0, 1};
    const D3 q = cross(dir, e2);
    const double det = dot(e1, q);
    if (fabs(det) < 2.220446049250313e-16) return;
    const double fi = 1 / det;
    for (long i = x0; i <= x1; ++i)
        for (long j = y0; j <= y1; ++j) {
            const D3 o = {(double)i, (double)j, 0.};
            const D3 sv = o - a;
            const double u = fi * dot(sv, q);
            if (u < 0.0) continue;
            const D3 r = cross(sv, e1);
            const double v = fi * dot(dir, r);
            if (v < 0.0 || u + v > 1.0) continue;
            const int cx_ = wrap_node((int)i, gd.n[0], gd.per[0]);
            const int cy_ = wrap_node((int)j, gd.n[1], gd.per[1]);
            if (cx_ < 0 || cy_ < 0) continue;
            const long col = (long)cx_ * gd.n[1] + cy_;
            if (pass == 0) {
                atomicAdd(&col_count[col], 1);
            } else {
                double z = fi * dot(e2, r);
                if (gd.per[2]) {
                    const double L = (double)gd.n[2];
                    z = z - L * floor((z + 0.5) * (1. / L));
                }
                const long slot = (long)col_off[col] + atomicAdd(&col_fill[col], 1);
                if (slot < cap) col_z[slot] = z;
                else atomicOr(&status[2], 2);
            }
        }
}

// Per column: sort the hits and fill the phase as apply_voxelization does
// (src/fsi/ImmersedBoundaryDynamics_voxelization.hh:121-165): alternate
// outside/inside between consecutive hits, later segments overwrite the
// shared end node.
__global__ void k_fill_columns(Grid gd, const int* col_count, const int* col_off, double* col_z, long cap,
                               unsigned char* phase)
{
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (long)gd.n[0] * gd.n[1]) return;
    const int nz = gd.n[2];
    unsigned char* ph = phase + col * nz;
    int n = col_count[col];
    const long o = col_off[col];
    if (o + n > cap) n = (int)(cap > o ? cap - o : 0);
    double* zs = col_z + o;
    for (int i = 1; i < n; ++i) {  // insertion sort (few hits per column)
        const double key = zs[i];
        int j = i - 1;
        while (j >= 0 && zs[j] > key) {
            zs[j + 1] = zs[j];
            --j;
        }
        zs[j + 1] = key;
    }
    bool inside = false;
    double z_last = -10000;
    for (int k = 0; k < n; ++k) {
        const long zlo = max((long)ceil(z_last), 0L), zhi = min((long)floor(zs[k]), (long)nz - 1);
        for (long z = zlo; z <= zhi; ++z) ph[z] = inside ? 1 : 0;
        z_last = zs[k];
        inside = !inside;
    }
    if (ceil(z_last) < nz - 1) {
        const long zlo = max((long)ceil(z_last), 0L);
        for (long z = zlo; z <= nz - 1; ++z) ph[z] = inside ? 1 : 0;
    }
}

__global__ void k_scan_cols(const int* in, long n, int* off, int* total)
{
    // single-workgroup exclusive scan (columns are few: nx*ny)
    __shared__ int sh[256];
    int carry = 0;
    for (long c0 = 0; c0 < n; c0 += 256) {
        const long i = c0 + threadIdx.x;
        const int v = i < n ? in[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int t = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) off[i] = carry + sh[threadIdx.x] - v;
        const int tot = sh[255];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

}  // namespace

// ======================================================== host helpers
namespace {

template <class T>
int upload(T*& dst, const std::vector<T>& src)
{
    hipFree(dst);
    dst = nullptr;
    const size_t n = std::max<size_t>(src.size(), 1);
    BL_HIP(hipMalloc(&dst, n * sizeof(T)));
    if (!src.empty()) BL_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return BL_OK;
}

template <class T>
int alloc(T*& dst, size_t n)
{
    hipFree(dst);
    dst = nullptr;
    BL_HIP(hipMalloc(&dst, std::max<size_t>(n, 1) * sizeof(T)));
    return BL_OK;
}

int alloc_bytes(void*& dst, size_t n)
{
    hipFree(dst);
    dst = nullptr;
    BL_HIP(hipMalloc(&dst, std::max<size_t>(n, 1)));
    return BL_OK;
}

// Runs fn(R{}) with R the boundary's vertex real type.
template <class Fn>
auto with_vreal(bl_ibm g, Fn&& fn)
{
    if (g->vreal == 4) return fn(float{});
    return fn(double{});
}

// The vertex real type that goes with a population storage type.
template <class S>
struct VRof {
    using T = float;
};
template <>
struct VRof<double> {
    using T = double;
};

// The kernels templated on the storage use its vertex type; a boundary
// whose vertex type differs (rigid particles on a reduced-precision
// lattice) never reaches them (rigid_refuse_coupling).
int vreal_matches_storage(bl_ibm g)
{
    const int want = g->lat->storage == BL_STORE_F64 ? 8 : 4;
    if (g->vreal != want)
        return fail(BL_ERR_STATE, "immersed boundary: vertex precision does not match the lattice storage");
    return BL_OK;
}

template <class R>
ForceView<R> force_view(bl_ibm g)
{
    ForceView<R> f;
    f.np = g->np;
    f.p_voff = g->p_voff; f.p_toff = g->p_toff; f.p_loff = g->p_loff; f.p_poff = g->p_poff;
    f.p_shape = g->p_shape; f.p_model = g->p_model; f.p_bug = g->p_bug;
    f.p_params = g->p_params; f.p_geom = g->p_geom;
    f.s_voff = g->s_voff; f.s_toff = g->s_toff; f.s_loff = g->s_loff; f.s_poff = g->s_poff;
    f.s_vtoff = g->s_vtoff; f.s_vloff = g->s_vloff; f.s_vpoff = g->s_vpoff;
    f.s_tri = g->s_tri; f.s_tri_area0 = g->s_tri_area0; f.s_tri_geo0 = g->s_tri_geo0;
    f.s_link = g->s_link; f.s_link_len = g->s_link_len;
    f.s_pair = g->s_pair; f.s_pair_cs0 = g->s_pair_cs0;
    f.s_vt_head = g->s_vt_head; f.s_vt = g->s_vt;
    f.s_vl_head = g->s_vl_head; f.s_vl = g->s_vl;
    f.s_vp_head = g->s_vp_head; f.s_vp = g->s_vp;
    f.s_area0 = g->s_area0;
    f.pos = (const R*)g->pos;
    f.v_part = g->v_part; f.t_part = g->t_part; f.r_part = g->r_part; f.l_part = g->l_part;
    f.v_inc = (const int4*)g->v_inc;
    return f;
}

template <class R>
NearView<R> near_view(bl_ibm g)
{
    NearView<R> nw;
    nw.b = g->boundary;
    nw.on = g->has_boundary && g->v_near ? 1 : 0;
    nw.near = g->v_near;
    nw.blend = (R*)g->v_blend;
    nw.base_t = g->v_base_t;
    nw.W_t = (R*)g->v_w_t;
    return nw;
}

// Per-vertex near-wall buffers, once both a boundary and the vertices exist.
int alloc_near(bl_ibm g)
{
    if (!g->has_boundary || !g->inited || g->v_near) return BL_OK;
    const size_t nv = (size_t)std::max<long>(g->NV, 1);
    hipError_t e = hipMalloc(&g->v_near, nv);
    if (e == hipSuccess) e = hipMalloc(&g->v_blend, (size_t)g->vreal * nv);
    if (e == hipSuccess) e = hipMalloc(&g->v_base_t, sizeof(int) * 3 * nv);
    if (e == hipSuccess) e = hipMalloc(&g->v_w_t, (size_t)g->vreal * 6 * nv);
    if (e != hipSuccess) return fail(BL_ERR_HIP, std::string("near-wall buffers: ") + hipGetErrorString(e));
    return BL_OK;
}

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

int need_init(bl_ibm g)
{
    if (!g) return fail(BL_ERR_ARG, "null ibm handle");
    if (!g->inited) return fail(BL_ERR_STATE, "immersed boundary not initialised (call bl_ibm_init)");
    if (!g->lat) return fail(BL_ERR_STATE, "the lattice of this immersed boundary was destroyed");
    return BL_OK;
}

void mark(bl_ibm g, int k)
{
    if (g->timing) hipEventRecord(g->ev[k], g->lat->stream);
}

template <class R>
GeomArgs<R> geom_args(bl_ibm g, int remap)
{
    GeomArgs<R> a;
    a.f = force_view<R>(g);
    a.chunks = (const int4*)g->ch_tab;
    a.t_vert = g->t_vert;
    a.partial = g->ch_partial;
    a.nchunk = g->nchunk;
    a.np = g->np;
    a.gd = grid_of(g->lat);
    a.p_choff = g->p_choff;
    a.p_voff = g->p_voff;
    a.geom = g->p_geom;
    a.delta = g->p_delta;
    a.pos = (R*)g->pos;
    a.remap = remap;
    a.p_model = g->p_model;
    return a;
}

int update_particles(bl_ibm g, int remap, hipStream_t s)
{
    if (g->np == 0) return BL_OK;
    with_vreal(g, [&](auto tag) {
        using R = decltype(tag);
        const GeomArgs<R> a = geom_args<R>(g, remap);
        hipLaunchKernelGGL(k_geom_partial<R>, dim3(g->nchunk), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_geom_final<R>, dim3(g->np), dim3(64), 0, s, a);
        return 0;
    });
    BL_CHECK_LAUNCH();
    return BL_OK;
}

int update_particles(bl_ibm g, int remap) { return update_particles(g, remap, g->lat->stream); }

// Stencil binning into the tile lists, the touched-tile list and the force
// flags of this step (buffer g->alt), on stream s.  The spreading kernels
// clear the counts of the tiles they consume; a binning that no spreading
// consumed leaves them set, and the next one clears first.  `done`
// (optional) is recorded once the lists are complete.
//
// Before it, list g->alt and flag buffer flags_alt are readied: counters a
// binning left unconsumed are cleared, the flags of the tiles the list held
// (two steps back) are cleared and the list emptied.  That may run as soon
// as the last staging has read those flags, carried by the interpolation
// launch of the step before (prep_rare + list_reset_of), or here.
ListReset list_reset_of(bl_ibm g)
{
    bl_lattice L = g->lat;
    ListReset lr;
    lr.act = g->active + (long)g->alt * L->ntiles;
    lr.na = g->n_active + g->alt;
    lr.flags = g->flags_alt;
    lr.pairs = g->n_active + 5 + g->alt;
    lr.arrive = g->n_active + 4;
    lr.nbig = g->n_active + 8 + g->alt;
    return lr;
}

void prep_rare(bl_ibm g, hipStream_t s)
{
    bl_lattice L = g->lat;
    if (g->bins_dirty) {
        hipLaunchKernelGGL(k_clear_bins, dim3(nblk(L->ntiles)), dim3(256), 0, s, g->tile_count, L->ntiles);
        g->bins_dirty = false;
    }
    if (g->alt_full) {
        hipLaunchKernelGGL(k_zero_u8, dim3(nblk(L->ntiles)), dim3(256), 0, s, g->flags_alt, L->ntiles);
        g->alt_full = false;
    }
}

int prep_bin(bl_ibm g, hipStream_t s)
{
    prep_rare(g, s);
    hipLaunchKernelGGL(k_bin_prep, dim3((unsigned)std::min<long>(nblk(g->lat->ntiles), kResetBlocks)), dim3(256), 0, s,
                       list_reset_of(g));
    BL_CHECK_LAUNCH();
    g->prepped = true;
    return BL_OK;
}

template <class R>
BinArgs<R> bin_args(bl_ibm g)
{
    bl_lattice L = g->lat;
    BinArgs<R> a;
    a.gd = grid_of(L);
    a.kernel = g->kernel;
    a.S = g->S;
    a.h = g->h;
    a.NV = g->NV;
    a.pos = (const R*)g->pos;
    a.base = g->v_base;
    a.W = (R*)g->v_w;
    a.count = g->tile_count;
    a.list = g->tile_list;
    a.cap = kTileCap;
    a.act = g->active + (long)g->alt * L->ntiles;
    a.na = g->n_active + g->alt;
    a.flags = g->flags_alt;
    a.status = g->n_active;
    a.pairs = g->n_active + 5 + g->alt;
    a.big = g->active + (2L + g->alt) * L->ntiles;
    a.nbig = g->n_active + 8 + g->alt;
    a.nw = near_view<R>(g);
    a.xcd = g->lp[BL_IBM_XCD];
    return a;
}

// host state after a binning has been enqueued
void binned(bl_ibm g)
{
    g->binned = true;
    g->bins_dirty = true;
    g->last_list = g->alt;
}

int do_bin(bl_ibm g, hipStream_t s, hipEvent_t done = nullptr)
{
    int rc;
    if (!g->prepped && (rc = prep_bin(g, s))) return rc;
    g->prepped = false;
    if (g->NV)
        with_vreal(g, [&](auto tag) {
            using R = decltype(tag);
            const int bb = g->lp[BL_IBM_BIN_BLOCK];
            hipLaunchKernelGGL(k_bin<R>, dim3(nblk(g->NV, bb)), dim3(bb), 0, s, bin_args<R>(g));
            return 0;
        });
    if (done) BL_HIP(hipEventRecord(done, s));
    BL_CHECK_LAUNCH();
    binned(g);
    return BL_OK;
}

LatView ibm_view(bl_ibm g)
{
    LatView lv = view_of(g->lat);
    if (!g->lat->force_field) lv.flags = nullptr;
    return lv;
}

// Velocity staging of the step whose binning is the last one.  Before that
// binning's spreading (`spread` false) the last collision's force field and
// flags are the lattice's, and the binning's flags are in flags_alt; after
// it (`spread` true) the two have been swapped.
StageArgs stage_args(bl_ibm g, bool spread)
{
    bl_lattice L = g->lat;
    StageArgs a;
    a.lv = ibm_view(g);
    a.touched = g->flags_alt;
    a.ut = g->ut;
    if (spread) {
        a.lv.force = g->force_alt;
        if (a.lv.flags) a.lv.flags = g->flags_alt;
        a.touched = L->tile_flags;
    }
    return a;
}

// u on the cells of the touched tiles, one thread per cell
unsigned stage_blocks(bl_ibm g) { return (unsigned)((g->lat->ncell + 255) / 256); }

template <class S>
void launch_stage_s(const StageArgs& a, unsigned nblocks, bool swapped, hipStream_t s)
{
    if (swapped)
        hipLaunchKernelGGL((k_stage_u<S, true>), dim3(nblocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_stage_u<S, false>), dim3(nblocks), dim3(256), 0, s, a);
}

int launch_stage(bl_ibm g, hipStream_t s, bool spread)
{
    bl_lattice L = g->lat;
    if (g->NV == 0) return BL_OK;
    const StageArgs a = stage_args(g, spread);
    const bool swapped = (L->steps & 1) != 0;
    if (L->storage == BL_STORE_F64) launch_stage_s<double>(a, stage_blocks(g), swapped, s);
    else if (L->storage == BL_STORE_F32) launch_stage_s<float>(a, stage_blocks(g), swapped, s);
    else launch_stage_s<__half>(a, stage_blocks(g), swapped, s);
    BL_CHECK_LAUNCH();
    return BL_OK;
}

template <class R>
InterpArgs<R> interp_args(bl_ibm g, const ListReset& lr)
{
    InterpArgs<R> a;
    a.gd = grid_of(g->lat);
    a.NV = g->NV;
    a.base = g->v_base;
    a.W = (const R*)g->v_w;
    a.ut = g->ut;
    a.vel = (R*)g->vel;
    a.pos = (R*)g->pos;
    a.lr = lr;
    a.nw = near_view<R>(g);
    a.xcd = g->lp[BL_IBM_XCD];
    return a;
}

unsigned interp_blocks(bl_ibm g, int blk = 256) { return (unsigned)nblk((g->S == 2 ? 4 : 16) * g->NV, blk); }

template <class C, bool MOVE>
void launch_interp_c(bl_ibm g, hipStream_t s, const ListReset& lr)
{
    using R = C;  // the staged velocity's type is the vertex type
    const InterpArgs<R> a = interp_args<R>(g, lr);
    const int ib = g->lp[BL_IBM_INTERP_BLOCK];
    const dim3 grid(interp_blocks(g, ib));
    if (g->S == 4) hipLaunchKernelGGL((k_interp<C, 4, MOVE, R>), grid, dim3(ib), 0, s, a);
    else if (g->S == 3) hipLaunchKernelGGL((k_interp<C, 3, MOVE, R>), grid, dim3(ib), 0, s, a);
    else hipLaunchKernelGGL((k_interp<C, 2, MOVE, R>), grid, dim3(ib), 0, s, a);
}

// The list reset carried by a launch that follows this step's spreading:
// the flags and list of two steps back are free once the staging has read
// them.
ListReset reset_next_of(bl_ibm g, hipStream_t s)
{
    prep_rare(g, s);
    g->prepped = true;
    return list_reset_of(g);
}

// vertex velocities from the staged u; with `move` the vertices advance.
// `reset_next` (after this step's spreading, once its staging is done):
// the launch also readies the list of the next binning.
int launch_interp(bl_ibm g, bool move, hipStream_t s, bool reset_next = false)
{
    if (g->NV == 0) return BL_OK;
    int rc = vreal_matches_storage(g);
    if (rc) return rc;
    ListReset lr;
    if (reset_next) lr = reset_next_of(g, s);
    const bool f64 = g->lat->storage == BL_STORE_F64;
    if (f64 && move) launch_interp_c<double, true>(g, s, lr);
    else if (f64) launch_interp_c<double, false>(g, s, lr);
    else if (move) launch_interp_c<float, true>(g, s, lr);
    else launch_interp_c<float, false>(g, s, lr);
    BL_CHECK_LAUNCH();
    return BL_OK;
}

// Spreading of the last binning's lists into the boundary's second force
// buffer; `blocks` receives the launch's workgroup count.
// `share`: boxes whose spreading shares one launch (a batched ensemble);
// the grid's lower bounds are divided among them.
template <class R>
SpreadArgs<R> spread_args(bl_ibm g, unsigned& blocks, int share = 1)
{
    bl_lattice L = g->lat;
    SpreadArgs<R> a;
    a.lv = ibm_view(g);
    a.gd = grid_of(L);
    a.Sw = g->S;
    a.active = g->active + (long)g->alt * L->ntiles;
    a.nactive = g->n_active + g->alt;
    a.big = g->active + (2L + g->alt) * L->ntiles;
    a.nbig = g->n_active + 8 + g->alt;
    a.count = g->tile_count;
    a.cap = kTileCap;
    a.list = g->tile_list;
    a.base = g->v_base;
    a.W = (const R*)g->v_w;
    a.vforce = (const R*)g->force;
    a.nw = near_view<R>(g);
    a.Fout = g->force_alt;
    a.xcd = g->lp[BL_IBM_XCD];
    // about one small-tile wave per touched tile and one workgroup per big
    // tile: the touched tiles are at most every tile and, for membranes,
    // fewer than the vertices (by default a wave per vertex); big tiles
    // fewer than a 32nd of the vertices (by default a workgroup per 32
    // vertices).  Both are launch parameters: the grid does not change any
    // sum.
    const long sw = (long)g->NV * g->lp[BL_IBM_SPREAD_SMALL] / 4;
    const long min_small = std::max<long>(64, 1024 / share), min_big = std::max<long>(4, 256 / share);
    const long small_waves = std::max<long>(64, std::min<long>(L->ntiles, std::max<long>(sw, min_small)));
    const int gs = (int)std::min<long>((small_waves + kSpreadWaves - 1) / kSpreadWaves, 4096);
    const int gb = (int)std::max<long>(
        1, std::min<long>(L->ntiles, std::max<long>(g->NV / g->lp[BL_IBM_SPREAD_BIG], min_big)));
    a.n_small = gs;
    blocks = (unsigned)(gs + gb);
    return a;
}

// Host state once a spreading is enqueued: its force field and flags trade
// places with the lattice's, so that the last collision's field stays
// readable (velocity staging) while the next one is formed.
void spread_swap(bl_ibm g)
{
    bl_lattice L = g->lat;
    if (L->flags_full) {  // the buffer leaving the lattice has every tile set
        g->alt_full = true;
        L->flags_full = false;
    }
    std::swap(L->tile_flags, g->flags_alt);  // this step's tiles carry the force now
    std::swap(L->force, g->force_alt);
    g->alt ^= 1;
    g->bins_dirty = false;
}

int launch_spread(bl_ibm g, hipStream_t s)
{
    bl_lattice L = g->lat;
    if (g->NV == 0) return BL_OK;
    int rc = vreal_matches_storage(g);
    if (rc) return rc;
    unsigned nb;
    if (L->storage == BL_STORE_F64) {
        const SpreadArgs<double> a = spread_args<double>(g, nb);
        hipLaunchKernelGGL((k_spread<double, double>), dim3(nb), dim3(256), 0, s, a);
    } else {
        const SpreadArgs<float> a = spread_args<float>(g, nb);
        if (L->storage == BL_STORE_F32) hipLaunchKernelGGL((k_spread<float, float>), dim3(nb), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_spread<__half, float>), dim3(nb), dim3(256), 0, s, a);
    }
    BL_CHECK_LAUNCH();
    spread_swap(g);
    return BL_OK;
}

int do_interpolate(bl_ibm g)
{
    int rc;
    if ((rc = do_bin(g, g->lat->stream))) return rc;
    if ((rc = launch_stage(g, g->lat->stream, false))) return rc;
    return launch_interp(g, false, g->lat->stream);
}

// Vertex forces on stream s: particle-particle forces into their own
// buffer, per-element membrane forces, then the per-vertex sum.
template <class R>
ElemArgs<R> elem_args(bl_ibm g)
{
    ElemArgs<R> a;
    a.f = force_view<R>(g);
    a.NT = g->NT;
    a.NP = g->NP;
    a.NL = g->NL;
    a.ec_pair = g->ec_pair;
    a.ec_link = g->ec_link;
    a.ec = (R*)g->ec;
    a.xcd = g->lp[BL_IBM_XCD];
    return a;
}

// `walls`: with the wall-interaction decorators and the particle-particle
// forces of this step (when there are any)
template <class R>
GatherArgs<R> gather_args(bl_ibm g, bool walls)
{
    GatherArgs<R> a;
    a.f = force_view<R>(g);
    a.NV = g->NV;
    a.ec_pair = g->ec_pair;
    a.ec_link = g->ec_link;
    a.ec = (const R*)g->ec;
    a.force = (R*)g->force;
    if (walls) a.deco = deco_set(g);
    else a.deco.n = 0;
    a.fpp = walls && !g->pp.empty() ? (const R*)g->pp_force : nullptr;
    a.ext = walls && g->has_ext ? (const R*)g->ext_force : nullptr;
    a.xcd = g->lp[BL_IBM_XCD];
    return a;
}

// per-element membrane forces, then the per-vertex sum (`walls`: with the
// decorators, particle-particle and host-given forces)
template <class R, int GL>
void launch_gather(bl_ibm g, hipStream_t s, bool walls)
{
    hipLaunchKernelGGL((k_vertex_gather<R, GL>), dim3(nblk(GL * g->NV)), dim3(256), 0, s, gather_args<R>(g, walls));
}

void launch_membrane(bl_ibm g, hipStream_t s, bool walls)
{
    const int eb = g->lp[BL_IBM_ELEM_BLOCK];
    with_vreal(g, [&](auto tag) {
        using R = decltype(tag);
        hipLaunchKernelGGL(k_elem_forces<R>, dim3(nblk(g->NT + g->NP + g->NL, eb)), dim3(eb), 0, s, elem_args<R>(g));
        const int gl = g->lp[BL_IBM_GATHER_LANES];
        if (gl == 2) launch_gather<R, 2>(g, s, walls);
        else if (gl == 8) launch_gather<R, 8>(g, s, walls);
        else launch_gather<R, 4>(g, s, walls);
        return 0;
    });
}

int do_forces(bl_ibm g, hipStream_t s)
{
    int rc;
    const bool pp = !g->pp.empty() && g->NV;
    if (pp && (rc = interactions_launch(g, s))) return rc;
    if (g->NV) launch_membrane(g, s, true);
    BL_CHECK_LAUNCH();
    return BL_OK;
}

int do_spread(bl_ibm g)
{
    int rc;
    if (!g->binned && (rc = do_bin(g, g->lat->stream))) return rc;
    return launch_spread(g, g->lat->stream);
}

int do_move(bl_ibm g)
{
    hipStream_t s = g->lat->stream;
    if (g->NV)
        with_vreal(g, [&](auto tag) {
            using R = decltype(tag);
            hipLaunchKernelGGL(k_move<R>, dim3(nblk(3 * g->NV)), dim3(256), 0, s, g->NV, (R*)g->pos,
                               (const R*)g->vel);
            return 0;
        });
    BL_CHECK_LAUNCH();
    g->binned = false;
    return update_particles(g, 1);
}

int do_voxelize(bl_ibm g, hipStream_t s)
{
    bl_lattice L = g->lat;
    if (L->omega_c0 == L->omega_c1 || g->NT == 0) return BL_OK;
    const Grid gd = grid_of(L);
    const long ncol = (long)L->nx * L->ny;
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(ncol)), dim3(256), 0, s, g->col_count, ncol, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(ncol)), dim3(256), 0, s, g->col_fill, ncol, 0);
    with_vreal(g, [&](auto tag) {
        using R = decltype(tag);
        const ForceView<R> f = force_view<R>(g);
        hipLaunchKernelGGL(k_ray_hits<R>, dim3(nblk(g->NT)), dim3(256), 0, s, f, gd, g->NT, 0, g->col_count,
                           (const int*)g->col_off, g->col_fill, g->col_z, g->col_cap, g->n_active);
        hipLaunchKernelGGL(k_scan_cols, dim3(1), dim3(256), 0, s, (const int*)g->col_count, ncol, g->col_off,
                           g->n_hits);
        hipLaunchKernelGGL(k_ray_hits<R>, dim3(nblk(g->NT)), dim3(256), 0, s, f, gd, g->NT, 1, g->col_count,
                           (const int*)g->col_off, g->col_fill, g->col_z, g->col_cap, g->n_active);
        return 0;
    });
    hipLaunchKernelGGL(k_fill_columns, dim3(nblk(ncol)), dim3(256), 0, s, gd, (const int*)g->col_count,
                       (const int*)g->col_off, g->col_z, g->col_cap, g->phase);
    BL_CHECK_LAUNCH();
    if (L->phase != g->phase) {
        L->phase = g->phase;
        ++L->gen;
    }
    return BL_OK;
}

// Overflow flags accumulate on the device over any number of steps; this
// reads and clears them.
int check_overflow(bl_ibm g)
{
    int st[3] = {0, 0, 0};
    BL_HIP(hipMemcpyAsync(st, g->n_active, sizeof(st), hipMemcpyDeviceToHost, g->lat->stream));
    BL_HIP(hipStreamSynchronize(g->lat->stream));
    if (st[2]) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, g->lat->stream, g->n_active + 2, 1L, 0);
        BL_CHECK_LAUNCH();
    }
    if (st[2] & 1)
        return fail(BL_ERR_CAPACITY, "immersed boundary: more than " + std::to_string(kTileCap) +
                                         " stencils touch one 4x4x4 tile");
    if (st[2] & 2) return fail(BL_ERR_CAPACITY, "voxelization: hit buffer capacity exceeded");
    if (st[2] & 4)
        return fail(BL_ERR_STATE, "immersed boundary: a vertex position is not finite (the membrane diverged)");
    return BL_OK;
}

// One IBM step, in one stream (with the stage timing events) ...
int step_serial(bl_ibm g)
{
    int rc;
    hipStream_t s = g->lat->stream;
    mark(g, 0);
    if ((rc = do_bin(g, s))) return rc;
    mark(g, 1);
    if ((rc = do_forces(g, s))) return rc;
    mark(g, 2);
    if ((rc = launch_stage(g, s, false))) return rc;
    if ((rc = launch_spread(g, s))) return rc;
    if ((rc = launch_interp(g, true, s, true))) return rc;
    mark(g, 3);
    g->binned = false;
    if ((rc = update_particles(g, 1, s))) return rc;
    mark(g, 4);
    if ((rc = do_voxelize(g, s))) return rc;
    mark(g, 5);
    return BL_OK;
}

// ... or forked over the lattice stream M and the boundary's stream A, so
// that independent stages overlap:
//   A: particle-particle forces (when there are any) and membrane forces
//      (read the positions before the move); after the binning, spreading
//      into the second force buffer (it needs the lists and the forces,
//      not the staging)
//   M: binning, velocity staging (the last collision's force field and the
//      new tile flags), then, once the forces have read the positions,
//      interpolation + advection (the launch also readies the next
//      binning's list: the flags and list two steps back are free once
//      staged), geometry, voxelization; then joins A (the collision reads
//      the spread force field).
// The chain that ends at the collision stays on one queue; only the start
// of A, the binning's lists and the spread field cross between queues.
int step_forked(bl_ibm g)
{
    int rc;
    hipStream_t M = g->lat->stream, A = g->aux[0];
    hipEvent_t* e = g->fev;
    BL_HIP(hipEventRecord(e[0], M));
    BL_HIP(hipStreamWaitEvent(A, e[0], 0));
    if ((rc = do_forces(g, A))) return rc;
    BL_HIP(hipEventRecord(e[1], A));
    if ((rc = do_bin(g, M, e[2]))) return rc;
    if ((rc = launch_stage(g, M, false))) return rc;
    BL_HIP(hipStreamWaitEvent(A, e[2], 0));
    if ((rc = launch_spread(g, A))) return rc;
    BL_HIP(hipEventRecord(e[4], A));
    BL_HIP(hipStreamWaitEvent(M, e[1], 0));
    if ((rc = launch_interp(g, true, M, true))) return rc;
    g->binned = false;
    if ((rc = update_particles(g, 1, M))) return rc;
    if ((rc = do_voxelize(g, M))) return rc;
    BL_HIP(hipStreamWaitEvent(M, e[4], 0));
    return BL_OK;
}

// Whether coupled steps carry the velocity staging in the collision
// (step_fused).  Only without a phase field: the collision of a step reads
// the phase voxelized from the moved vertices, which need the staged
// velocity first.  By default for 65 536 vertices or more, where the
// staging pass reads populations over most of the lattice.
bool use_fused(bl_ibm g)
{
    bl_lattice L = g->lat;
    if (g->fused_staging == 0 || g->NV == 0 || L->omega_c0 != L->omega_c1) return false;
    return g->fused_staging == 1 || g->NV >= 65536;
}

// One coupled step with the velocity staging carried by the collision.
// The reference's order (computeVelocity, interpolate, spread, move,
// collide) has the collision depend only on the spread force, and the
// interpolation only on the staged velocity, so the step runs as
//   A: particle-particle and membrane forces (positions before the move)
//   M: binning; then, after A, spreading; collide-and-stream, which also
//      stages u (pre-collision populations, the last collision's force
//      field) on the touched tiles; post-stream walls; interpolation +
//      advection (readying the next binning's list); geometry.
// The staged values are those of k_stage_u; one pass over the touched
// populations is saved.
int step_fused(bl_ibm g)
{
    int rc;
    bl_lattice L = g->lat;
    hipStream_t M = L->stream, A = g->aux[0];
    hipEvent_t* e = g->fev;
    BL_HIP(hipEventRecord(e[0], M));
    BL_HIP(hipStreamWaitEvent(A, e[0], 0));
    if ((rc = do_forces(g, A))) return rc;
    BL_HIP(hipEventRecord(e[1], A));
    if ((rc = do_bin(g, M))) return rc;
    BL_HIP(hipStreamWaitEvent(M, e[1], 0));
    if ((rc = launch_spread(g, M))) return rc;
    // after the swap: the lattice holds this step's field and flags (the
    // touched tiles), the boundary's spare buffers the last collision's
    L->emit_u = g->ut;
    L->emit_force = g->force_alt;
    L->emit_flags = g->flags_alt;
    rc = collide_stream_launch(L);
    L->emit_u = nullptr;
    L->emit_force = nullptr;
    L->emit_flags = nullptr;
    if (rc) return rc;
    if ((rc = velocity_walls_launch(L))) return rc;
    if ((rc = launch_interp(g, true, M, true))) return rc;
    g->binned = false;
    return update_particles(g, 1, M);
}

// One coupled step on the lattice stream alone, as four launches before
// the collision, each pairing two stages that do not read each other's
// output:
//   binning | membrane element forces (both read the positions only)
//   vertex force gather | velocity staging (the binning's tiles, the last
//     collision's field and flags)
//   spreading | interpolation + advection (the launch also readies the
//     next binning's list)
//   geometry (then voxelization, with a phase field)
// The reference's order and every value are those of the other layouts;
// no stage waits on another queue.
// force gather (GL lanes per vertex) | velocity staging, one launch
template <class S, class R, int GL>
void launch_gather_stage(bl_ibm g, hipStream_t s)
{
    const int ng = (int)nblk(GL * g->NV);
    const StageArgs sa = stage_args(g, false);
    const dim3 g2(ng + stage_blocks(g));
    if (g->lat->steps & 1)
        hipLaunchKernelGGL((k_gather_stage<S, true, R, GL>), g2, dim3(256), 0, s, gather_args<R>(g, true), ng, sa);
    else
        hipLaunchKernelGGL((k_gather_stage<S, false, R, GL>), g2, dim3(256), 0, s, gather_args<R>(g, true), ng, sa);
}

template <class S>
int step_layered_s(bl_ibm g)
{
    using R = typename VRof<S>::T;
    int rc;
    bl_lattice L = g->lat;
    if ((rc = vreal_matches_storage(g))) return rc;
    hipStream_t s = L->stream;
    if (!g->pp.empty() && (rc = interactions_launch(g, s))) return rc;
    if (!g->prepped && (rc = prep_bin(g, s))) return rc;
    g->prepped = false;
    const int bb = g->lp[BL_IBM_BIN_BLOCK];  // both halves of the paired launch
    const int nb = (int)nblk(g->NV, bb), ne = (int)nblk(g->NT + g->NP + g->NL, bb);
    hipLaunchKernelGGL(k_bin_elem<R>, dim3(nb + ne), dim3(bb), 0, s, bin_args<R>(g), nb, elem_args<R>(g));
    binned(g);
    const int gl = g->lp[BL_IBM_GATHER_LANES];
    if (gl == 2) launch_gather_stage<S, R, 2>(g, s);
    else if (gl == 8) launch_gather_stage<S, R, 8>(g, s);
    else launch_gather_stage<S, R, 4>(g, s);
    unsigned nsp;
    const SpreadArgs<R> sp = spread_args<R>(g, nsp);
    spread_swap(g);
    const InterpArgs<R> ia = interp_args<R>(g, reset_next_of(g, s));
    const dim3 g3(nsp + interp_blocks(g));
    if (g->S == 4) hipLaunchKernelGGL((k_spread_interp<S, 4, R>), g3, dim3(256), 0, s, sp, (int)nsp, ia);
    else if (g->S == 3) hipLaunchKernelGGL((k_spread_interp<S, 3, R>), g3, dim3(256), 0, s, sp, (int)nsp, ia);
    else hipLaunchKernelGGL((k_spread_interp<S, 2, R>), g3, dim3(256), 0, s, sp, (int)nsp, ia);
    BL_CHECK_LAUNCH();
    g->binned = false;
    if ((rc = update_particles(g, 1, s))) return rc;
    return do_voxelize(g, s);
}

int step_layered(bl_ibm g)
{
    const int st = g->lat->storage;
    if (st == BL_STORE_F64) return step_layered_s<double>(g);
    if (st == BL_STORE_F32) return step_layered_s<float>(g);
    return step_layered_s<__half>(g);
}

// Whether coupled steps without fused staging run as step_layered (one
// stream) rather than step_forked (two): by default below 65 536 vertices,
// where the stages are a few workgroups each and a wait across queues
// costs more than the overlap gains.
bool use_layered(bl_ibm g)
{
    if (g->NV == 0 || g->layered == 0) return false;
    return g->layered == 1 || g->NV < 65536;
}

// n coupled steps enqueued on the lattice stream
int steps_forked(bl_ibm g, int n)
{
    int rc;
    for (int k = 0; k < n; ++k) {
        if (use_fused(g)) {
            if ((rc = step_fused(g))) return rc;
            continue;
        }
        if ((rc = use_layered(g) ? step_layered(g) : step_forked(g))) return rc;
        if ((rc = collide_stream_launch(g->lat))) return rc;
        if ((rc = velocity_walls_launch(g->lat))) return rc;
    }
    return BL_OK;
}

// ------------------------------------------------ batches of boxes
// An ensemble of independent boxes whose steps all take the paired-stage
// layout (step_layered) runs each stage as one launch over every box:
// LayeredEntry holds one box's arguments of one step, as step_layered
// forms them; a batched kernel finds its box with batch_box over the
// launch's per-box workgroup offsets and runs the box-local workgroup with
// that box's arguments.  plan_layered advances the boundary's host state
// exactly as step_layered does, without launching.
template <class R>
struct LayeredEntry {
    ListReset reset;
    int nr;  // workgroups of the binning's list reset (0: the list is ready)
    BinArgs<R> bin;
    ElemArgs<R> elem;
    int nb, ne;
    GatherArgs<R> gather;
    StageArgs stage;
    int ng, nst;
    SpreadArgs<R> spread;
    InterpArgs<R> interp;
    int nsp, nip;
    GeomArgs<R> geom;
};

// Whether a boundary's coupled step can join a batch: the paired-stage
// layout, no particle-particle forces or phase field, no pending clear of
// its tile buffers, and the vertex type of its storage.
bool layered_batchable(bl_ibm g)
{
    const bl_lattice L = g->lat;
    return L && g->inited && g->NV > 0 && use_layered(g) && !use_fused(g) && g->pp.empty() &&
           L->omega_c0 == L->omega_c1 && !g->bins_dirty && !g->alt_full && !ibm_needs_settle(g) && !g->timing &&
           g->n_rigid == 0 && g->vreal == (L->storage == BL_STORE_F64 ? 8 : 4);
}

// One step's entry; BL_ERR_STATE when a clear of the tile buffers would be
// needed in the middle of the step (the batch is then not used).
template <class R>
int plan_layered(bl_ibm g, LayeredEntry<R>& E, int share)
{
    bl_lattice L = g->lat;
    E.nr = 0;
    if (!g->prepped) {
        if (g->bins_dirty || g->alt_full) return BL_ERR_STATE;
        E.reset = list_reset_of(g);
        E.nr = (int)std::min<long>(nblk(L->ntiles), kResetBlocks);
        g->prepped = true;
    }
    g->prepped = false;
    E.bin = bin_args<R>(g);
    E.elem = elem_args<R>(g);
    E.nb = (int)nblk(g->NV);
    E.ne = (int)nblk(g->NT + g->NP + g->NL);
    binned(g);
    E.gather = gather_args<R>(g, true);
    E.stage = stage_args(g, false);
    E.ng = (int)nblk((long)g->lp[BL_IBM_GATHER_LANES] * g->NV);
    E.nst = (int)stage_blocks(g);
    unsigned nsp;
    E.spread = spread_args<R>(g, nsp, share);
    E.nsp = (int)nsp;
    spread_swap(g);
    if (g->bins_dirty || g->alt_full) return BL_ERR_STATE;
    g->prepped = true;
    E.interp = interp_args<R>(g, list_reset_of(g));
    E.nip = (int)interp_blocks(g);
    g->binned = false;
    E.geom = geom_args<R>(g, 1);
    return BL_OK;
}

template <class R>
__global__ void __launch_bounds__(256) k_reset_batch(const LayeredEntry<R>* __restrict__ tab,
                                                     const int* __restrict__ base, int nbox)
{
    const int box = batch_box(base, nbox, (int)blockIdx.x);
    const LayeredEntry<R>& E = tab[box];
    list_reset(E.reset, (int)blockIdx.x - base[box], E.nr);
}

template <class R>
__global__ void __launch_bounds__(256) k_bin_elem_batch(const LayeredEntry<R>* __restrict__ tab,
                                                        const int* __restrict__ base, int nbox)
{
    const int box = batch_box(base, nbox, (int)blockIdx.x);
    const LayeredEntry<R>& E = tab[box];
    const int bx = (int)blockIdx.x - base[box];
    if (bx < E.nb) bin_body<R>(ibm_block(bx, E.nb, E.bin.xcd), E.bin);
    else elem_body<R>(ibm_block(bx - E.nb, E.ne, E.elem.xcd), E.elem);
}

template <class S, bool SWAPPED, class R, int GL>
__global__ void __launch_bounds__(256) k_gather_stage_batch(const LayeredEntry<R>* __restrict__ tab,
                                                            const int* __restrict__ base, int nbox)
{
    const int box = batch_box(base, nbox, (int)blockIdx.x);
    const LayeredEntry<R>& E = tab[box];
    const int bx = (int)blockIdx.x - base[box];
    if (bx < E.ng) gather_body<R, GL>(ibm_block(bx, E.ng, E.gather.xcd), E.gather);
    else stage_body<S, SWAPPED>(bx - E.ng, E.stage);
}

template <class S, int SW, class R>
__global__ void __launch_bounds__(256) k_spread_interp_batch(const LayeredEntry<R>* __restrict__ tab,
                                                             const int* __restrict__ base, int nbox)
{
    const int box = batch_box(base, nbox, (int)blockIdx.x);
    const LayeredEntry<R>& E = tab[box];
    const int bx = (int)blockIdx.x - base[box];
    if (bx < E.nsp) spread_body<S, R>(bx, E.nsp, E.spread);
    else interp_body<typename Pop<S>::C, SW, true, R>(bx - E.nsp, E.nip, E.interp);
}

// The interpolation with advection alone (batches whose velocity staging
// rides on the collision: it runs after the collide-and-stream).
template <class S, int SW, class R>
__global__ void __launch_bounds__(256) k_interp_batch(const LayeredEntry<R>* __restrict__ tab,
                                                      const int* __restrict__ base, int nbox)
{
    const int box = batch_box(base, nbox, (int)blockIdx.x);
    const LayeredEntry<R>& E = tab[box];
    interp_body<typename Pop<S>::C, SW, true, R>((int)blockIdx.x - base[box], E.nip, E.interp);
}

template <class R>
__global__ void __launch_bounds__(256) k_geom_partial_batch(const LayeredEntry<R>* __restrict__ tab,
                                                            const int* __restrict__ base, int nbox)
{
    const int box = batch_box(base, nbox, (int)blockIdx.x);
    geom_partial_body<R>((int)blockIdx.x - base[box], tab[box].geom);
}

template <class R>
__global__ void __launch_bounds__(64) k_geom_final_batch(const LayeredEntry<R>* __restrict__ tab,
                                                         const int* __restrict__ base, int nbox)
{
    const int box = batch_box(base, nbox, (int)blockIdx.x);
    geom_final_body<R>((int)blockIdx.x - base[box], tab[box].geom);
}

}  // namespace

// Two coupled steps of every box of an ensemble as batched launches: the
// entries and workgroup offsets of every launch in one device buffer.
struct EnsembleBatch {
    int storage = 0, sw = 4, gl = 4;
    int order[2] = {0, 0};  // collide workgroup order per AA parity (collide_block modes)
    int nt[2] = {0, 0};     // and nontemporal level
    bool any_ibm = false, any_vwall = false;
    // velocity staging carried by the batched collision (every box has an
    // immersed boundary): no staging pass, the interpolation after the
    // collide-and-stream and the walls, as a single box's fused step
    bool fused = false;
    void* dev = nullptr;  // entries and offsets of both steps
    struct Launch {
        size_t tab = 0, base = 0;  // byte offsets into dev
        int nbox = 0;
        unsigned nblk = 0;
    };
    struct Step {
        bool swapped = false;  // the IBM stages' AA parity (before the collision)
        Launch reset, bin_elem, gather_stage, spread_interp, interp, geom_partial, geom_final, collide, vwall, vwall_u;
        CollidePlan c0;        // template choices of the collide launch
        bool vw_swapped = false;
    } step[2];
};

namespace {

// Appends n objects to a host byte image (8-byte aligned), returning their offset.
template <class T>
size_t put(std::vector<char>& img, const T* p, size_t n)
{
    size_t off = (img.size() + 15) & ~(size_t)15;
    img.resize(off + sizeof(T) * n);
    if (n) memcpy(img.data() + off, p, sizeof(T) * n);
    return off;
}

// the workgroup offsets of a launch over boxes with the given counts
EnsembleBatch::Launch put_launch(std::vector<char>& img, size_t tab, const std::vector<int>& counts)
{
    EnsembleBatch::Launch l;
    std::vector<int> base(counts.size() + 1, 0);
    for (size_t k = 0; k < counts.size(); ++k) base[k + 1] = base[k] + counts[k];
    l.tab = tab;
    l.base = put(img, base.data(), base.size());
    l.nbox = (int)counts.size();
    l.nblk = (unsigned)base.back();
    return l;
}

template <class R>
int batch_plan_ibm(const std::vector<bl_ibm>& gs, std::vector<char>& img, EnsembleBatch::Step& st, bool fused,
                   std::vector<LatView>* emit)
{
    std::vector<LayeredEntry<R>> E(gs.size());
    for (size_t k = 0; k < gs.size(); ++k) {
        const int rc = plan_layered<R>(gs[k], E[k], (int)gs.size());
        if (rc) return rc;
        if (fused) {
            // after the step's spread swap: the boundary's spare field and
            // flags are the last collision's (what the staging reads)
            LatView& v = (*emit)[k];
            v.emit_u = gs[k]->ut;
            v.emit_force = gs[k]->force_alt;
            v.emit_flags = gs[k]->flags_alt;
        }
    }
    const size_t tab = put(img, E.data(), E.size());
    std::vector<int> nr, nbe, ngs, nsi, nip, nch, npa;
    for (size_t k = 0; k < gs.size(); ++k) {
        nr.push_back(E[k].nr);
        nbe.push_back(E[k].nb + E[k].ne);
        // fused: the gather alone and the spreading alone (every workgroup
        // index stays below the first stage's count)
        ngs.push_back(E[k].ng + (fused ? 0 : E[k].nst));
        nsi.push_back(E[k].nsp + (fused ? 0 : E[k].nip));
        nip.push_back(fused ? E[k].nip : 0);
        nch.push_back(gs[k]->nchunk);
        npa.push_back(gs[k]->np);
    }
    st.reset = put_launch(img, tab, nr);
    st.bin_elem = put_launch(img, tab, nbe);
    st.gather_stage = put_launch(img, tab, ngs);
    st.spread_interp = put_launch(img, tab, nsi);
    st.interp = put_launch(img, tab, nip);
    st.geom_partial = put_launch(img, tab, nch);
    st.geom_final = put_launch(img, tab, npa);
    return BL_OK;
}

// The immersed-boundary launches of one batched step: before the collision
// (after = false) the binning with element forces, the gather (with the
// staging unless fused) and the spreading (with the interpolation unless
// fused), then the geometry unless fused; after the collision and the
// walls (after = true, fused only) the interpolation and the geometry.
template <class S>
void batch_launch_ibm(const EnsembleBatch& B, const EnsembleBatch::Step& st, bool after, hipStream_t s)
{
    using R = typename VRof<S>::T;
    const char* d = (const char*)B.dev;
    auto tab = [&](const EnsembleBatch::Launch& l) { return (const LayeredEntry<R>*)(d + l.tab); };
    auto base = [&](const EnsembleBatch::Launch& l) { return (const int*)(d + l.base); };
    auto go = [&](auto kern, const EnsembleBatch::Launch& l, int blk) {
        if (l.nblk) hipLaunchKernelGGL(kern, dim3(l.nblk), dim3(blk), 0, s, tab(l), base(l), l.nbox);
    };
    if (!after) {
        go(k_reset_batch<R>, st.reset, 256);
        go(k_bin_elem_batch<R>, st.bin_elem, 256);
        auto gs = [&](auto sw_tag) {
            constexpr bool SWP = decltype(sw_tag)::value;
            if (B.gl == 2) go(k_gather_stage_batch<S, SWP, R, 2>, st.gather_stage, 256);
            else if (B.gl == 8) go(k_gather_stage_batch<S, SWP, R, 8>, st.gather_stage, 256);
            else go(k_gather_stage_batch<S, SWP, R, 4>, st.gather_stage, 256);
        };
        if (st.swapped) gs(std::true_type{});
        else gs(std::false_type{});
        if (B.sw == 4) go(k_spread_interp_batch<S, 4, R>, st.spread_interp, 256);
        else if (B.sw == 3) go(k_spread_interp_batch<S, 3, R>, st.spread_interp, 256);
        else go(k_spread_interp_batch<S, 2, R>, st.spread_interp, 256);
        if (B.fused) return;
    } else {
        if (B.sw == 4) go(k_interp_batch<S, 4, R>, st.interp, 256);
        else if (B.sw == 3) go(k_interp_batch<S, 3, R>, st.interp, 256);
        else go(k_interp_batch<S, 2, R>, st.interp, 256);
    }
    go(k_geom_partial_batch<R>, st.geom_partial, 256);
    go(k_geom_final_batch<R>, st.geom_final, 64);
}

}  // namespace

int ensemble_batch_build(const std::vector<bl_lattice>& lats, const std::vector<bl_ibm>& ibms, const int order[2],
                         const int nt[2], EnsembleBatch** out)
{
    *out = nullptr;
    const size_t n = lats.size();
    if (n < 2) return BL_OK;
    std::vector<bl_ibm> gs;
    int sw = 0, gl = 0;
    for (size_t k = 0; k < n; ++k) {
        const bl_lattice L = lats[k];
        if (!L->guo.empty() || L->storage != lats[0]->storage || (L->steps & 1) != (lats[0]->steps & 1)) return BL_OK;
        if (!collide_batchable(collide_plan(L), collide_plan(lats[0]))) return BL_OK;
        const bl_ibm g = ibms[k];
        if (!g) continue;
        if (!layered_batchable(g)) return BL_OK;
        if ((sw && g->S != sw) || (gl && g->lp[BL_IBM_GATHER_LANES] != gl)) return BL_OK;
        sw = g->S;
        gl = g->lp[BL_IBM_GATHER_LANES];
        gs.push_back(g);
    }
    // plan both steps; the host state of every box advances two steps and
    // its step count is put back afterwards (nothing has run)
    std::vector<long long> s0(n);
    for (size_t k = 0; k < n; ++k) s0[k] = lats[k]->steps;
    EnsembleBatch* B = new EnsembleBatch();
    B->storage = lats[0]->storage;
    B->sw = sw ? sw : 4;
    B->gl = gl ? gl : 4;
    B->any_ibm = !gs.empty();
    B->fused = gs.size() == n && collide_plan(lats[0]).fm == 2;
    B->order[0] = order[0];
    B->order[1] = order[1];
    B->nt[0] = nt[0];
    B->nt[1] = nt[1];
    std::vector<char> img;
    int rc = BL_OK;
    for (int k = 0; k < 2 && rc == BL_OK; ++k) {
        EnsembleBatch::Step& st = B->step[k];
        st.swapped = (lats[0]->steps & 1) != 0;
        // the collide views after the boundaries' plans (their spread swap
        // hands each lattice this step's force field), with the staging
        // pointers the plans recorded when fused
        std::vector<LatView> emit(n);
        if (!gs.empty()) {
            if (B->storage == BL_STORE_F64) rc = batch_plan_ibm<double>(gs, img, st, B->fused, &emit);
            else rc = batch_plan_ibm<float>(gs, img, st, B->fused, &emit);
            if (rc) break;
        }
        std::vector<LatView> views;
        std::vector<int> cnt;
        st.c0 = collide_plan(lats[0]);
        for (size_t k = 0; k < n; ++k) {
            const bl_lattice L = lats[k];
            LatView v = collide_plan(L).v;
            if (B->fused) {
                v.emit_u = emit[k].emit_u;
                v.emit_force = emit[k].emit_force;
                v.emit_flags = emit[k].emit_flags;
            }
            views.push_back(v);
            cnt.push_back((int)((L->ncell + 255) / 256));
            collide_commit(L);
        }
        st.collide = put_launch(img, put(img, views.data(), views.size()), cnt);
        std::vector<VwallPlan> vw;
        std::vector<int> vcnt, ucnt;
        st.vw_swapped = (lats[0]->steps & 1) != 0;
        for (bl_lattice L : lats)
            if (L->has_vwall && L->mask) {
                vw.push_back(VwallPlan{view_of(L), L->nwall, L->d_wall});
                vcnt.push_back((int)((L->nwall + 255) / 256));
                ucnt.push_back((int)((2 * L->nwall + 255) / 256));
            }
        B->any_vwall = !vw.empty();
        const size_t vtab = put(img, vw.data(), vw.size());
        st.vwall = put_launch(img, vtab, vcnt);
        st.vwall_u = put_launch(img, vtab, ucnt);
    }
    for (size_t k = 0; k < n; ++k) lats[k]->steps = s0[k];
    if (rc) {  // a clear was due mid-step: the ensemble uses per-box branches
        delete B;
        return rc == BL_ERR_STATE ? BL_OK : rc;
    }
    if (hipMalloc(&B->dev, img.size()) != hipSuccess ||
        hipMemcpy(B->dev, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(B->dev);
        delete B;
        return fail(BL_ERR_HIP, "ensemble batch: device tables");
    }
    *out = B;
    return BL_OK;
}

int ensemble_batch_launch(const EnsembleBatch* B, hipStream_t s)
{
    const char* d = (const char*)B->dev;
    for (int k = 0; k < 2; ++k) {
        const EnsembleBatch::Step& st = B->step[k];
        auto ibm = [&](bool after) {
            if (B->storage == BL_STORE_F64) batch_launch_ibm<double>(*B, st, after, s);
            else if (B->storage == BL_STORE_F32) batch_launch_ibm<float>(*B, st, after, s);
            else batch_launch_ibm<__half>(*B, st, after, s);
        };
        if (B->any_ibm) {
            ibm(false);
            BL_CHECK_LAUNCH();
        }
        int rc = collide_batch_launch(st.c0, B->fused, (const LatView*)(d + st.collide.tab),
                                      (const int*)(d + st.collide.base), st.collide.nbox, st.collide.nblk,
                                      B->order[st.c0.odd ? 1 : 0], B->nt[st.c0.odd ? 1 : 0], s);
        if (rc) return rc;
        if (st.vwall.nblk &&
            (rc = vwall_batch_launch(B->storage, st.vw_swapped, (const VwallPlan*)(d + st.vwall.tab),
                                     (const int*)(d + st.vwall.base), (const int*)(d + st.vwall_u.base),
                                     st.vwall.nbox, st.vwall.nblk, st.vwall_u.nblk, s)))
            return rc;
        if (B->fused) {
            ibm(true);
            BL_CHECK_LAUNCH();
        }
    }
    return BL_OK;
}

void ensemble_batch_free(EnsembleBatch* B)
{
    if (!B) return;
    hipFree(B->dev);
    delete B;
}

namespace {

}  // namespace

int ibm_step_launch(bl_ibm g, bool forked)
{
    int rc;
    if (forked) return steps_forked(g, 1);
    // on the lattice stream alone: paired stages unless stage times are taken
    if ((rc = !g->timing && use_layered(g) ? step_layered(g) : step_serial(g))) return rc;
    if ((rc = collide_stream_launch(g->lat))) return rc;
    return velocity_walls_launch(g->lat);
}

int ibm_check_overflow(bl_ibm g) { return check_overflow(g); }

void ibm_detach(bl_ibm g)
{
    for (auto& st : g->aux)
        if (st) hipStreamSynchronize(st);
    g->lat = nullptr;
}

int vreal_upload(bl_ibm g, void* dst, long off, const double* src, long n, hipStream_t s)
{
    if (n <= 0) return BL_OK;
    if (g->vreal == 8) {
        BL_HIP(hipMemcpyAsync((double*)dst + off, src, sizeof(double) * n, hipMemcpyHostToDevice, s));
        BL_HIP(hipStreamSynchronize(s));
        return BL_OK;
    }
    std::vector<float> tmp(src, src + n);
    BL_HIP(hipMemcpyAsync((float*)dst + off, tmp.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    BL_HIP(hipStreamSynchronize(s));
    return BL_OK;
}

int vreal_download(bl_ibm g, double* dst, const void* src, long off, long n, hipStream_t s)
{
    if (n <= 0) return BL_OK;
    if (g->vreal == 8) {
        BL_HIP(hipMemcpyAsync(dst, (const double*)src + off, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        BL_HIP(hipStreamSynchronize(s));
        return BL_OK;
    }
    std::vector<float> tmp(n);
    BL_HIP(hipMemcpyAsync(tmp.data(), (const float*)src + off, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    BL_HIP(hipStreamSynchronize(s));
    for (long i = 0; i < n; ++i) dst[i] = (double)tmp[i];
    return BL_OK;
}

void ibm_invalidate_graphs(bl_ibm g)
{
    g->graph_parity[0] = g->graph_parity[1] = -1;
    ++g->gen;
}

// Steps run eagerly while a loaded full force field is consumed, and for the
// first step after the body force changed (its velocity staging reads the
// previous body force; a captured step would bake that value in).
bool ibm_needs_settle(bl_ibm g)
{
    const bl_lattice L = g->lat;
    if (!L) return false;
    const bool g_changed = L->g_last[0] != L->g[0] || L->g_last[1] != L->g[1] || L->g_last[2] != L->g[2];
    return L->flags_full || g->alt_full || g_changed;
}

// The stages that place delta-kernel stencils on the lattice (binning,
// interpolation, spreading) need every periodic extent to hold one stencil
// width, so that a stencil never covers a cell twice; the membrane and
// collision-handler forces do not.
int need_stencil(bl_ibm g)
{
    int rc = need_init(g);
    if (rc) return rc;
    if ((rc = rigid_refuse_coupling(g))) return rc;
    const bl_lattice L = g->lat;
    const int n3[3] = {L->nx, L->ny, L->nz};
    for (int d = 0; d < 3; ++d)
        if (L->periodic[d] && n3[d] < 4)
            return fail(BL_ERR_ARG, "immersed boundary: the kernel stages need periodic extents of at least 4 "
                                    "(one stencil width)");
    return BL_OK;
}

}  // namespace bl

using namespace bl;

extern "C" {

int bl_ibm_create(bl_lattice lat, int kernel, bl_ibm* out)
{
    if (!lat || !out) return fail(BL_ERR_ARG, "bl_ibm_create: null argument");
    if (kernel < BL_KERNEL_ROMA || kernel > BL_KERNEL_TOPHAT) return fail(BL_ERR_ARG, "bl_ibm_create: unknown kernel");
    if (lat->ibm) return fail(BL_ERR_STATE, "bl_ibm_create: lattice already has an immersed boundary");
    bl_ibm g = new bl_ibm_s();
    g->lat = lat;
    g->kernel = kernel;
    g->S = kernel == BL_KERNEL_ROMA ? 3 : (kernel == BL_KERNEL_PESKIN ? 4 : 2);
    g->h = kernel == BL_KERNEL_ROMA ? 1.5 : (kernel == BL_KERNEL_PESKIN ? 2.0 : 1.0);
    lat->ibm = g;
    *out = g;
    return BL_OK;
}

int bl_ibm_destroy(bl_ibm g)
{
    if (!g) return BL_OK;
    if (g->lat) {
        hipStreamSynchronize(g->lat->stream);
        if (g->lat->phase == g->phase) g->lat->phase = g->lat->own_phase;
        g->lat->ibm = nullptr;
        ++g->lat->gen;
    }
    void* ptrs[] = {g->s_voff, g->s_toff, g->s_loff, g->s_poff, g->s_vtoff, g->s_vloff, g->s_vpoff, g->s_tri,
                    g->s_tri_area0, g->s_tri_geo0, g->s_rest, g->s_link, g->s_link_len, g->s_pair, g->s_pair_cs0,
                    g->s_vt_head, g->s_vt, g->s_vl_head, g->s_vl, g->s_vp_head, g->s_vp, g->s_area0, g->p_shape,
                    g->p_model, g->p_bug, g->p_params, g->p_voff, g->p_toff, g->p_loff, g->p_poff, g->p_geom,
                    g->p_delta, g->pos, g->vel, g->force, g->ec, g->ut, g->v_base, g->v_w,
                    g->tile_count, g->tile_list, g->active, g->n_active,
                    g->col_count, g->col_off, g->col_fill, g->col_z, g->n_hits, g->phase,
                    g->flags_alt, g->force_alt, g->ch_part, g->ch_idx, g->p_choff, g->ch_partial, g->ch_tab, g->t_vert, g->v_inc, g->v_near, g->v_blend, g->v_base_t, g->v_w_t, g->v_part, g->t_part, g->r_part,
                    g->l_part, g->ext_force, g->r_state};
    for (void* p : ptrs) hipFree(p);
    interactions_free(g);
    for (auto& e : g->ev)
        if (e) hipEventDestroy(e);
    for (auto& e : g->fev)
        if (e) hipEventDestroy(e);
    for (auto& st : g->aux)
        if (st) {
            hipStreamSynchronize(st);
            hipStreamDestroy(st);
        }
    for (auto& ge : g->graph)
        if (ge) hipGraphExecDestroy(ge);
    delete g;
    return BL_OK;
}

int bl_shape_add(bl_ibm g, int nv, int nt, const double* rest_pos, const int* tri, int* shape_id)
{
    if (!g || !rest_pos || !tri || nv < 3 || nt < 1) return fail(BL_ERR_ARG, "bl_shape_add: bad argument");
    if (g->inited) return fail(BL_ERR_STATE, "bl_shape_add: shapes must be added before bl_ibm_init");
    ShapeHost sh;
    std::string err;
    if (!build_shape(nv, nt, rest_pos, tri, sh, &err)) return fail(BL_ERR_ARG, "bl_shape_add: " + err);
    g->shapes.push_back(std::move(sh));
    if (shape_id) *shape_id = (int)g->shapes.size() - 1;
    return BL_OK;
}

int bl_shape_info(bl_ibm g, int shape_id, double out[6])
{
    if (!g || !out || shape_id < 0 || shape_id >= (int)g->shapes.size()) return fail(BL_ERR_ARG, "bl_shape_info: bad argument");
    const ShapeHost& s = g->shapes[shape_id];
    out[0] = s.area;
    out[1] = s.volume;
    for (int d = 0; d < 3; ++d) out[2 + d] = s.center[d];
    out[5] = s.radius;
    return BL_OK;
}

int bl_particle_add(bl_ibm g, int shape_id, int model, const double* params, int nparams, const double* pos,
                    int bug_compatible, int* pid)
{
    if (!g || !params || !pos) return fail(BL_ERR_ARG, "bl_particle_add: null argument");
    if (g->inited) return fail(BL_ERR_STATE, "bl_particle_add: particles must be added before bl_ibm_init");
    if (shape_id < 0 || shape_id >= (int)g->shapes.size()) return fail(BL_ERR_ARG, "bl_particle_add: unknown shape");
    if (model < 0 || model > 3) return fail(BL_ERR_ARG, "bl_particle_add: unknown model");
    const int need = model <= 1 ? 2 : 7;
    if (nparams < need || nparams > 8) return fail(BL_ERR_ARG, "bl_particle_add: wrong number of parameters");
    bl_ibm_s::PartHost p;
    p.shape = shape_id;
    p.model = model;
    p.bug = bug_compatible ? 1 : 0;
    for (int k = 0; k < nparams; ++k) p.params[k] = params[k];
    if ((model == BL_MODEL_RBC || model == 3) && !(p.params[5] > 0))
        return fail(BL_ERR_ARG, "bl_particle_add: vol_desired must be positive");
    const int nv = g->shapes[shape_id].nv;
    p.pos.assign(pos, pos + 3 * (size_t)nv);
    p.id = (long long)g->parts.size();
    g->parts.push_back(std::move(p));
    if (pid) *pid = (int)g->parts.size() - 1;
    return BL_OK;
}

int bl_ibm_init(bl_ibm g)
{
    if (!g) return fail(BL_ERR_ARG, "null ibm handle");
    if (g->inited) return fail(BL_ERR_STATE, "bl_ibm_init: already initialised");
    bl_lattice L = g->lat;
    int rc = ensure_force_field(L);
    if (rc) return rc;
    // shapes
    std::vector<int> s_voff{0}, s_toff{0}, s_loff{0}, s_poff{0}, s_vtoff{0}, s_vloff{0}, s_vpoff{0};
    std::vector<int> s_tri, s_link, s_pair, vt_head, vt, vl_head, vl, vp_head, vp;
    std::vector<double> s_area0v, s_geo, s_rest, s_len, s_cs;
    std::vector<double> s_tarea;
    for (const ShapeHost& s : g->shapes) {
        s_voff.push_back(s_voff.back() + s.nv);
        s_toff.push_back(s_toff.back() + s.nt);
        s_loff.push_back(s_loff.back() + s.nl);
        s_poff.push_back(s_poff.back() + s.np);
        s_vtoff.push_back(s_vtoff.back() + (int)s.v_tri.size());
        s_vloff.push_back(s_vloff.back() + (int)s.v_link.size());
        s_vpoff.push_back(s_vpoff.back() + (int)s.v_pair.size());
        s_tri.insert(s_tri.end(), s.tri.begin(), s.tri.end());
        s_tarea.insert(s_tarea.end(), s.tri_area0.begin(), s.tri_area0.end());
        s_geo.insert(s_geo.end(), s.tri_geo0.begin(), s.tri_geo0.end());
        s_rest.insert(s_rest.end(), s.rest.begin(), s.rest.end());
        s_link.insert(s_link.end(), s.link.begin(), s.link.end());
        s_len.insert(s_len.end(), s.link_len.begin(), s.link_len.end());
        s_pair.insert(s_pair.end(), s.pair.begin(), s.pair.end());
        s_cs.insert(s_cs.end(), s.pair_cs0.begin(), s.pair_cs0.end());
        vt_head.insert(vt_head.end(), s.v_tri_off.begin(), s.v_tri_off.end());
        vt.insert(vt.end(), s.v_tri.begin(), s.v_tri.end());
        vl_head.insert(vl_head.end(), s.v_link_off.begin(), s.v_link_off.end());
        vl.insert(vl.end(), s.v_link.begin(), s.v_link.end());
        vp_head.insert(vp_head.end(), s.v_pair_off.begin(), s.v_pair_off.end());
        vp.insert(vp.end(), s.v_pair.begin(), s.v_pair.end());
        s_area0v.push_back(s.area);
    }
    g->nshape = (int)g->shapes.size();
    if ((rc = upload(g->s_voff, s_voff)) || (rc = upload(g->s_toff, s_toff)) || (rc = upload(g->s_loff, s_loff)) ||
        (rc = upload(g->s_poff, s_poff)) || (rc = upload(g->s_vtoff, s_vtoff)) || (rc = upload(g->s_vloff, s_vloff)) ||
        (rc = upload(g->s_vpoff, s_vpoff)) || (rc = upload(g->s_tri, s_tri)) || (rc = upload(g->s_tri_area0, s_tarea)) ||
        (rc = upload(g->s_tri_geo0, s_geo)) || (rc = upload(g->s_rest, s_rest)) || (rc = upload(g->s_link, s_link)) ||
        (rc = upload(g->s_link_len, s_len)) || (rc = upload(g->s_pair, s_pair)) || (rc = upload(g->s_pair_cs0, s_cs)) ||
        (rc = upload(g->s_vt_head, vt_head)) || (rc = upload(g->s_vt, vt)) || (rc = upload(g->s_vl_head, vl_head)) ||
        (rc = upload(g->s_vl, vl)) || (rc = upload(g->s_vp_head, vp_head)) || (rc = upload(g->s_vp, vp)) ||
        (rc = upload(g->s_area0, s_area0v)))
        return rc;
    // particles
    g->np = (int)g->parts.size();
    std::vector<long> voff{0}, toff{0}, loff{0}, poff{0};
    std::vector<int> pshape, pmodel, pbug;
    std::vector<double> params, pos;
    for (const auto& p : g->parts) {
        const ShapeHost& s = g->shapes[p.shape];
        voff.push_back(voff.back() + s.nv);
        toff.push_back(toff.back() + s.nt);
        loff.push_back(loff.back() + s.nl);
        poff.push_back(poff.back() + s.np);
        pshape.push_back(p.shape);
        pmodel.push_back(p.model);
        pbug.push_back(p.bug);
        params.insert(params.end(), p.params, p.params + 8);
        pos.insert(pos.end(), p.pos.begin(), p.pos.end());
    }
    g->NV = voff.back();
    if (g->NV >= (1L << 29)) return fail(BL_ERR_CAPACITY, "bl_ibm_init: at most 2^29 vertices");
    g->NT = toff.back();
    g->NL = loff.back();
    g->NP = poff.back();
    {
        std::vector<int> vp, tp, rp, lp;
        vp.reserve(g->NV);
        tp.reserve(g->NT);
        rp.reserve(g->NP);
        lp.reserve(g->NL);
        for (int p = 0; p < g->np; ++p) {
            vp.insert(vp.end(), voff[p + 1] - voff[p], p);
            tp.insert(tp.end(), toff[p + 1] - toff[p], p);
            rp.insert(rp.end(), poff[p + 1] - poff[p], p);
            lp.insert(lp.end(), loff[p + 1] - loff[p], p);
        }
        // per vertex: where its link / triangle / pair incidence lists start
        // in the concatenated lists and how many entries its particle's
        // model uses (links for RBCs, pairs for RBCs and model 3), and the
        // particle
        std::vector<int> inc;
        inc.reserve(8 * (size_t)g->NV);
        for (int p = 0; p < g->np; ++p) {
            const int sh = g->parts[p].shape, model = g->parts[p].model;
            const ShapeHost& S = g->shapes[sh];
            const bool links = model == BL_MODEL_RBC, pairs = model == BL_MODEL_RBC || model == 3;
            for (int i = 0; i < S.nv; ++i) {
                inc.push_back(s_vloff[sh] + S.v_link_off[i]);
                inc.push_back(links ? S.v_link_off[i + 1] - S.v_link_off[i] : 0);
                inc.push_back(s_vtoff[sh] + S.v_tri_off[i]);
                inc.push_back(S.v_tri_off[i + 1] - S.v_tri_off[i]);
                inc.push_back(s_vpoff[sh] + S.v_pair_off[i]);
                inc.push_back(pairs ? S.v_pair_off[i + 1] - S.v_pair_off[i] : 0);
                inc.push_back(p);
                inc.push_back(0);
            }
        }
        if ((rc = upload(g->v_part, vp)) || (rc = upload(g->t_part, tp)) || (rc = upload(g->r_part, rp)) ||
            (rc = upload(g->l_part, lp)) || (rc = upload(g->v_inc, inc)))
            return rc;
    }
    // vertex reals: FP32 with reduced-precision populations; rigid particles
    // (their own FP64 dynamics, no fluid coupling) keep FP64
    bool any_rigid = false;
    for (const auto& p : g->parts) any_rigid = any_rigid || p.model == BL_MODEL_RIGID;
    g->vreal = (L->storage == BL_STORE_F64 || any_rigid) ? 8 : 4;
    const size_t vr = (size_t)g->vreal;
    const size_t nec = 9 * (size_t)g->NT + 12 * (size_t)g->NP + 3 * (size_t)g->NL;
    if ((rc = upload(g->p_voff, voff)) || (rc = upload(g->p_toff, toff)) || (rc = upload(g->p_loff, loff)) ||
        (rc = upload(g->p_poff, poff)) || (rc = upload(g->p_shape, pshape)) || (rc = upload(g->p_model, pmodel)) ||
        (rc = upload(g->p_bug, pbug)) || (rc = upload(g->p_params, params)) ||
        (rc = alloc_bytes(g->pos, vr * 3 * g->NV)) || (rc = alloc_bytes(g->vel, vr * 3 * g->NV)) ||
        (rc = alloc_bytes(g->force, vr * 3 * g->NV)) || (rc = alloc(g->p_geom, 11 * (size_t)g->np)) ||
        (rc = alloc(g->p_delta, 3 * (size_t)g->np)) || (rc = alloc_bytes(g->ec, vr * nec)) ||
        (rc = alloc(g->v_base, 3 * g->NV)) || (rc = alloc_bytes(g->v_w, vr * 12 * g->NV)))
        return rc;
    if ((rc = vreal_upload(g, g->pos, 0, pos.data(), (long)pos.size(), L->stream))) return rc;
    g->ec_pair = 9 * g->NT;
    g->ec_link = g->ec_pair + 12 * g->NP;
    {
        const size_t csz = L->storage == BL_STORE_F64 ? sizeof(Vel4<double>) : sizeof(Vel4<float>);
        g->ut = nullptr;
        BL_HIP(hipMalloc(&g->ut, csz * (size_t)L->ncell));
        BL_HIP(hipMemset(g->ut, 0, csz * (size_t)L->ncell));
    }
    BL_HIP(hipMemset(g->vel, 0, vr * 3 * std::max<long>(g->NV, 1)));
    BL_HIP(hipMemset(g->force, 0, vr * 3 * std::max<long>(g->NV, 1)));
    {  // vertex velocities and forces of particles loaded from a checkpoint
        long v0 = 0;
        for (const auto& p : g->parts) {
            const long n = (long)p.pos.size();
            if ((long)p.vel.size() == n && (rc = vreal_upload(g, g->vel, 3 * v0, p.vel.data(), n, L->stream)))
                return rc;
            if ((long)p.force.size() == n && (rc = vreal_upload(g, g->force, 3 * v0, p.force.data(), n, L->stream)))
                return rc;
            v0 += n / 3;
        }
    }
    // tiles
    if ((rc = alloc(g->tile_count, L->ntiles)) || (rc = alloc(g->tile_list, (size_t)L->ntiles * kTileCap)) ||
        (rc = alloc(g->active, 4 * (size_t)L->ntiles)) || (rc = alloc(g->n_active, 16)))
        return rc;
    BL_HIP(hipMemset(g->n_active, 0, 16 * sizeof(int)));
    if ((rc = alloc(g->flags_alt, L->ntiles))) return rc;
    BL_HIP(hipMemset(g->flags_alt, 0, L->ntiles));
    BL_HIP(hipMalloc(&g->force_alt, 3 * L->stride * L->force_bytes));
    BL_HIP(hipMemset(g->force_alt, 0, 3 * L->stride * L->force_bytes));
    {
        // chunks of 256 triangles per particle; chunk k also takes the
        // bounding box of vertices [nv k / nch, nv (k + 1) / nch)
        std::vector<int> chp, chi, choff{0}, tab, tvert;
        tvert.reserve(3 * (size_t)g->NT);
        for (int p = 0; p < g->np; ++p) {
            const ShapeHost& S = g->shapes[g->parts[p].shape];
            const long nt = toff[p + 1] - toff[p], nv = voff[p + 1] - voff[p];
            const int nch = (int)std::max<long>(1, (nt + 255) / 256);
            for (int k = 0; k < nch; ++k) {
                chp.push_back(p);
                chi.push_back(k);
                tab.push_back((int)(toff[p] + 256L * k));
                tab.push_back((int)std::max<long>(0, std::min<long>(256, nt - 256L * k)));
                tab.push_back((int)(voff[p] + nv * k / nch));
                tab.push_back((int)(voff[p] + nv * (k + 1) / nch));
            }
            choff.push_back(choff.back() + nch);
            for (long t = 0; t < 3 * nt; ++t) tvert.push_back((int)(voff[p] + S.tri[t]));
        }
        g->nchunk = (int)chp.size();
        if ((rc = upload(g->ch_part, chp)) || (rc = upload(g->ch_idx, chi)) || (rc = upload(g->p_choff, choff)) ||
            (rc = upload(g->ch_tab, tab)) || (rc = upload(g->t_vert, tvert)) ||
            (rc = alloc(g->ch_partial, 11 * (size_t)std::max(g->nchunk, 1))))
            return rc;
    }
    BL_HIP(hipMemset(g->tile_count, 0, L->ntiles * sizeof(int)));
    // voxelization
    const long ncol = (long)L->nx * L->ny;
    g->col_cap = std::max<long>(64 * g->NT, 1024);
    if ((rc = alloc(g->col_count, ncol)) || (rc = alloc(g->col_off, ncol)) || (rc = alloc(g->col_fill, ncol)) ||
        (rc = alloc(g->col_z, g->col_cap)) || (rc = alloc(g->n_hits, 1)) || (rc = alloc(g->phase, L->ncell)))
        return rc;
    BL_HIP(hipMemset(g->phase, 0, L->ncell));
    for (auto& e : g->ev) BL_HIP(hipEventCreate(&e));
    for (auto& st : g->aux) BL_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (auto& e : g->fev) BL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    g->inited = true;
    if ((rc = alloc_near(g))) return rc;
    if (!g->pp.empty() && (rc = interactions_alloc(g))) return rc;
    // update() of every particle and the centre-of-mass mapping of add_particle
    // (rigid particles map their own centre, rigid_init)
    rc = update_particles(g, 1);
    if (rc) return rc;
    if ((rc = rigid_init(g))) return rc;
    BL_HIP(hipStreamSynchronize(L->stream));
    return BL_OK;
}

int bl_ibm_set_boundary(bl_ibm g, const bl_boundary* b)
{
    if (!g) return fail(BL_ERR_ARG, "null ibm handle");
    if (b && b->kind != BL_BOUNDARY_PIPE && b->kind != BL_BOUNDARY_PLATES)
        return fail(BL_ERR_ARG, "bl_ibm_set_boundary: unknown boundary kind");
    if (g->lat) BL_HIP(hipStreamSynchronize(g->lat->stream));
    g->has_boundary = b != nullptr;
    if (b) g->boundary = *b;
    ibm_invalidate_graphs(g);
    return alloc_near(g);
}

int bl_ibm_interpolate(bl_ibm g)
{
    int rc = need_stencil(g);
    return rc ? rc : do_interpolate(g);
}

int bl_ibm_compute_forces(bl_ibm g)
{
    int rc = need_init(g);
    if (!rc) rc = rigid_refuse_coupling(g);
    return rc ? rc : do_forces(g, g->lat->stream);
}

int bl_ibm_spread(bl_ibm g)
{
    int rc = need_stencil(g);
    if (rc) return rc;
    if ((rc = do_spread(g))) return rc;
    return check_overflow(g);
}

int bl_ibm_move(bl_ibm g)
{
    int rc = need_init(g);
    if (!rc) rc = rigid_refuse_coupling(g);
    if (rc) return rc;
    if ((rc = do_move(g))) return rc;
    return do_voxelize(g, g->lat->stream);
}

int bl_ibm_step(bl_ibm g)
{
    int rc = need_stencil(g);
    if (rc) return rc;
    g->timing = true;
    rc = step_serial(g);
    g->timing = false;
    if (rc) return rc;
    BL_HIP(hipEventSynchronize(g->ev[5]));
    for (int k = 0; k < 5; ++k) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, g->ev[k], g->ev[k + 1]);
        g->stage_ms[k] = ms;
    }
    return check_overflow(g);
}

namespace {
constexpr int kLongRun = 32;  // steps per replay of the long captured run

// Captures `n` (even) coupled steps into graph slot `slot` unless the slot
// holds a capture made in the same step parity / buffer / list-reset state
// and the same lattice and boundary generations (no setter ran since).
int ensure_graph(bl_ibm g, int slot, int n)
{
    bl_lattice L = g->lat;
    const long long par = (L->steps & 1) | ((long long)g->alt << 1) | ((long long)g->prepped << 2);
    if (g->graph[slot] && g->graph_parity[slot] == par && g->graph_gen[slot][0] == L->gen &&
        g->graph_gen[slot][1] == g->gen)
        return BL_OK;
    if (g->graph[slot]) {
        hipGraphExecDestroy(g->graph[slot]);
        g->graph[slot] = nullptr;
    }
    hipGraph_t graph;
    BL_HIP(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
    const long long s0 = L->steps;
    int rc = steps_forked(g, n);
    hipError_t e = hipStreamEndCapture(L->stream, &graph);
    L->steps = s0;  // capture did not execute anything
    if (rc) return rc;
    if (e != hipSuccess) return fail(BL_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&g->graph[slot], graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(BL_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    g->graph_parity[slot] = par;
    // after the capture: a capture that set state for the first time (the
    // phase field of a voxelization) keys on the state it captured
    g->graph_gen[slot][0] = L->gen;
    g->graph_gen[slot][1] = g->gen;
    return BL_OK;
}

int replay(bl_ibm g, int slot, int n, int reps)
{
    for (int r = 0; r < reps; ++r) {
        BL_HIP(hipGraphLaunch(g->graph[slot], g->lat->stream));
        g->lat->steps += n;
    }
    return BL_OK;
}
}  // namespace

int bl_coupled_steps(bl_ibm g, int n_steps, int use_graph)
{
    int rc = need_stencil(g);
    if (rc) return rc;
    bl_lattice L = g->lat;
    if (n_steps < 0) return fail(BL_ERR_ARG, "bl_coupled_steps: negative step count");
    int done = 0;
    // a force field loaded into the lattice is consumed over the next two
    // steps; those run eagerly, so no captured step depends on it
    for (; done < n_steps && ibm_needs_settle(g); ++done)
        if ((rc = steps_forked(g, 1))) return rc;
    if (use_graph) {
        // Runs of kLongRun steps, then pairs (the two AA parities), each
        // captured once per step parity and replayed.
        const int left = n_steps - done;
        if (left >= kLongRun) {
            if ((rc = ensure_graph(g, 1, kLongRun))) return rc;
            if ((rc = replay(g, 1, kLongRun, left / kLongRun))) return rc;
            done += left / kLongRun * kLongRun;
        }
        if (n_steps - done >= 2) {
            if ((rc = ensure_graph(g, 0, 2))) return rc;
            const int reps = (n_steps - done) / 2;
            if ((rc = replay(g, 0, 2, reps))) return rc;
            done += 2 * reps;
        }
    }
    if (done < n_steps && (rc = steps_forked(g, n_steps - done))) return rc;
    return check_overflow(g);
}

int bl_coupled_prepare(bl_ibm g, int n_steps)
{
    int rc = need_stencil(g);
    if (rc) return rc;
    if (n_steps < 0) return fail(BL_ERR_ARG, "bl_coupled_prepare: negative step count");
    if (ibm_needs_settle(g))
        return fail(BL_ERR_STATE, "bl_coupled_prepare: the next step runs eagerly (a loaded force field or a changed "
                                  "body force); run it first");
    // the graphs bl_coupled_steps(g, n_steps, 1) replays from this state: runs
    // of kLongRun steps keep the step parity, so both are captured for it
    if (n_steps >= kLongRun && (rc = ensure_graph(g, 1, kLongRun))) return rc;
    if (n_steps % kLongRun >= 2 && (rc = ensure_graph(g, 0, 2))) return rc;
    return BL_OK;
}

int bl_ibm_count(bl_ibm g, int* particles, int* vertices)
{
    if (!g) return fail(BL_ERR_ARG, "null ibm handle");
    if (particles) *particles = (int)g->parts.size();
    long nv = 0;
    for (const auto& p : g->parts) nv += g->shapes[p.shape].nv;
    if (vertices) *vertices = (int)nv;
    return BL_OK;
}

int bl_particle_get(bl_ibm g, int pid, int what, double* out)
{
    int rc = need_init(g);
    if (rc) return rc;
    if (!out || what < 0 || what > 2 || pid < -1 || pid >= g->np) return fail(BL_ERR_ARG, "bl_particle_get: bad argument");
    std::vector<long> voff{0};
    for (const auto& p : g->parts) voff.push_back(voff.back() + g->shapes[p.shape].nv);
    const long v0 = pid < 0 ? 0 : voff[pid], v1 = pid < 0 ? voff.back() : voff[pid + 1];
    const void* src = what == 0 ? g->pos : (what == 1 ? g->vel : g->force);
    return vreal_download(g, out, src, 3 * v0, 3 * (v1 - v0), g->lat->stream);
}

int bl_particle_set_positions(bl_ibm g, int pid, const double* pos)
{
    int rc = need_init(g);
    if (rc) return rc;
    if (!pos || pid < -1 || pid >= g->np) return fail(BL_ERR_ARG, "bl_particle_set_positions: bad argument");
    std::vector<long> voff{0};
    for (const auto& p : g->parts) voff.push_back(voff.back() + g->shapes[p.shape].nv);
    const long v0 = pid < 0 ? 0 : voff[pid], v1 = pid < 0 ? voff.back() : voff[pid + 1];
    if ((rc = vreal_upload(g, g->pos, 3 * v0, pos, 3 * (v1 - v0), g->lat->stream))) return rc;
    rc = update_particles(g, 0);
    if (rc) return rc;
    BL_HIP(hipStreamSynchronize(g->lat->stream));
    return BL_OK;
}

}  // extern "C"

namespace {
// Membrane forces alone (RBCParticle::compute_forces): no decorators, no
// particle-particle forces.
int membrane_forces(bl_ibm g, hipStream_t s)
{
    launch_membrane(g, s, false);
    BL_CHECK_LAUNCH();
    return BL_OK;
}

// One calibration step (node masses, then the relax update) of the
// selected particles.
int calib_step(bl_ibm g, hipStream_t s, const int* active, const double* damp, const double* v0, const double* v1,
               int n_it, double* mass, int* step)
{
    with_vreal(g, [&](auto tag) {
        using R = decltype(tag);
        hipLaunchKernelGGL(k_calib_mass<R>, dim3(nblk(g->NV)), dim3(256), 0, s, force_view<R>(g), g->NV, active, mass,
                           step);
        hipLaunchKernelGGL(k_calib_relax<R>, dim3(nblk(g->NV)), dim3(256), 0, s, g->NV, (const int*)g->v_part,
                           (const long*)g->p_voff, active, damp, v0, v1, n_it, (const int*)step, (const double*)mass,
                           (const R*)g->force, (R*)g->vel, (R*)g->pos, g->p_params);
        return 0;
    });
    BL_CHECK_LAUNCH();
    return BL_OK;
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { hipFree(p); }
    int alloc(size_t n)
    {
        BL_HIP(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)));
        return BL_OK;
    }
};
}  // namespace

extern "C" {

int bl_rbc_shrink_volume(bl_ibm g, int pid, double vol_final, int n_it, double damping, double* damping_out)
{
    int rc = need_init(g);
    if (rc) return rc;
    if (pid < -1 || pid >= g->np) return fail(BL_ERR_ARG, "bl_rbc_shrink_volume: bad particle id");
    if (n_it < 1 || !(damping > 0) || !(vol_final > 0))
        return fail(BL_ERR_ARG, "bl_rbc_shrink_volume: need n_it >= 1, damping > 0, vol_final > 0");
    const int np = g->np;
    std::vector<int> todo(np, 0);
    std::vector<double> damp(np, damping), v0(np, 0.0), v1(np, vol_final);
    std::vector<long> voff(np + 1, 0);
    for (int p = 0; p < np; ++p) {
        const auto& P = g->parts[p];
        voff[p + 1] = voff[p] + g->shapes[P.shape].nv;
        v0[p] = g->shapes[P.shape].volume;  // the rest shape's volume
        const bool rbc = P.model == BL_MODEL_RBC || P.model == 3;
        if (pid == p && !rbc) return fail(BL_ERR_ARG, "bl_rbc_shrink_volume: the particle has no volume constraint");
        todo[p] = (pid == -1 || pid == p) && rbc;
    }
    hipStream_t s = g->lat->stream;
    const size_t nv3 = 3 * (size_t)g->NV;
    const size_t vr = (size_t)g->vreal;
    DevBuf<int> d_active, d_step;
    DevBuf<double> d_damp, d_v0, d_v1, d_mass, prm0;
    DevBuf<char> pos0, vel0, frc0;  // the vertex state, in the boundary's reals
    if ((rc = d_active.alloc(np)) || (rc = d_step.alloc(1)) || (rc = d_damp.alloc(np)) || (rc = d_v0.alloc(np)) ||
        (rc = d_v1.alloc(np)) || (rc = d_mass.alloc(g->NV)) || (rc = pos0.alloc(vr * nv3)) ||
        (rc = vel0.alloc(vr * nv3)) || (rc = frc0.alloc(vr * nv3)) || (rc = prm0.alloc(8 * (size_t)np)))
        return rc;
    BL_HIP(hipMemcpyAsync(d_v0.p, v0.data(), sizeof(double) * np, hipMemcpyHostToDevice, s));
    BL_HIP(hipMemcpyAsync(d_v1.p, v1.data(), sizeof(double) * np, hipMemcpyHostToDevice, s));
    // the starting state: geometry and forces of the current positions with
    // the current parameters (rbc.compute_forces() before the loop)
    if ((rc = update_particles(g, 0, s)) || (rc = membrane_forces(g, s))) return rc;
    BL_HIP(hipMemcpyAsync(pos0.p, g->pos, vr * nv3, hipMemcpyDeviceToDevice, s));
    BL_HIP(hipMemcpyAsync(vel0.p, g->vel, vr * nv3, hipMemcpyDeviceToDevice, s));
    BL_HIP(hipMemcpyAsync(frc0.p, g->force, vr * nv3, hipMemcpyDeviceToDevice, s));
    BL_HIP(hipMemcpyAsync(prm0.p, g->p_params, sizeof(double) * 8 * np, hipMemcpyDeviceToDevice, s));
    for (;;) {
        bool any = false;
        for (int p = 0; p < np; ++p) any = any || todo[p];
        if (!any) break;
        BL_HIP(hipMemcpyAsync(d_active.p, todo.data(), sizeof(int) * np, hipMemcpyHostToDevice, s));
        BL_HIP(hipMemcpyAsync(d_damp.p, damp.data(), sizeof(double) * np, hipMemcpyHostToDevice, s));
        BL_HIP(hipMemsetAsync(d_step.p, 0, sizeof(int), s));
        for (int it = 0; it <= 2 * n_it; ++it) {
            if ((rc = calib_step(g, s, d_active.p, d_damp.p, d_v0.p, d_v1.p, n_it, d_mass.p, d_step.p))) return rc;
            if ((rc = update_particles(g, 0, s)) || (rc = membrane_forces(g, s))) return rc;
        }
        // a particle whose area went NaN starts over with twice the damping
        std::vector<double> geom(11 * (size_t)np);
        BL_HIP(hipMemcpyAsync(geom.data(), g->p_geom, sizeof(double) * geom.size(), hipMemcpyDeviceToHost, s));
        BL_HIP(hipStreamSynchronize(s));
        for (int p = 0; p < np; ++p) {
            if (!todo[p]) continue;
            const double area = geom[11 * p];
            if (area == area) {
                todo[p] = 0;
                continue;
            }
            damp[p] *= 2.0;
            const size_t a = vr * 3 * (size_t)voff[p], n = vr * 3 * (size_t)(voff[p + 1] - voff[p]);
            BL_HIP(hipMemcpyAsync((char*)g->pos + a, pos0.p + a, n, hipMemcpyDeviceToDevice, s));
            BL_HIP(hipMemcpyAsync((char*)g->vel + a, vel0.p + a, n, hipMemcpyDeviceToDevice, s));
            BL_HIP(hipMemcpyAsync((char*)g->force + a, frc0.p + a, n, hipMemcpyDeviceToDevice, s));
            BL_HIP(hipMemcpyAsync(g->p_params + 8 * p, prm0.p + 8 * p, sizeof(double) * 8, hipMemcpyDeviceToDevice, s));
            if (damp[p] > 2.0) {  // left in its starting state
                if ((rc = update_particles(g, 0, s))) return rc;
                BL_HIP(hipStreamSynchronize(s));
                return fail(BL_ERR_STATE, "bl_rbc_shrink_volume: particle " + std::to_string(p) +
                                              " diverges for every damping up to 2");
            }
        }
        // restored particles need their geometry back before the next attempt
        if ((rc = update_particles(g, 0, s))) return rc;
    }
    // the calibrated vol_desired stays with the particles (host copy too)
    std::vector<double> prm(8 * (size_t)np);
    BL_HIP(hipMemcpyAsync(prm.data(), g->p_params, sizeof(double) * prm.size(), hipMemcpyDeviceToHost, s));
    BL_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < np; ++p) {
        for (int k = 0; k < 8; ++k) g->parts[p].params[k] = prm[8 * p + k];
        if (damping_out) damping_out[p] = damp[p];
    }
    g->binned = false;
    ibm_invalidate_graphs(g);
    return BL_OK;
}

int bl_particle_set_state(bl_ibm g, int pid, int what, const double* data)
{
    int rc = need_init(g);
    if (rc) return rc;
    if (!data || what < 0 || what > 2 || pid < -1 || pid >= g->np)
        return fail(BL_ERR_ARG, "bl_particle_set_state: bad argument");
    if (what == 0) return bl_particle_set_positions(g, pid, data);
    long v0 = 0, v1 = 0;
    for (int p = 0; p < g->np; ++p) {
        const long n = g->shapes[g->parts[p].shape].nv;
        if (p < pid) v0 += n;
        if (pid < 0 || p <= pid) v1 += n;
    }
    if (pid < 0) v0 = 0;
    void* dst = what == 1 ? g->vel : g->force;
    return vreal_upload(g, dst, 3 * v0, data, 3 * (v1 - v0), g->lat->stream);
}

int bl_particle_set_params(bl_ibm g, int pid, const double* params, int nparams)
{
    int rc = need_init(g);
    if (rc) return rc;
    if (!params || pid < 0 || pid >= g->np || nparams < 1 || nparams > 8)
        return fail(BL_ERR_ARG, "bl_particle_set_params: bad argument");
    double* h = g->parts[pid].params;
    for (int k = 0; k < nparams; ++k) h[k] = params[k];
    BL_HIP(hipMemcpyAsync(g->p_params + 8 * (size_t)pid, h, sizeof(double) * 8, hipMemcpyHostToDevice, g->lat->stream));
    BL_HIP(hipStreamSynchronize(g->lat->stream));
    ibm_invalidate_graphs(g);
    return BL_OK;
}

int bl_rbc_relax_nodes(bl_ibm g, int pid, double damping)
{
    int rc = need_init(g);
    if (rc) return rc;
    if (pid < 0 || pid >= g->np || !(damping >= 0)) return fail(BL_ERR_ARG, "bl_rbc_relax_nodes: bad argument");
    const int model = g->parts[pid].model;
    if (model != BL_MODEL_RBC && model != 3) return fail(BL_ERR_ARG, "bl_rbc_relax_nodes: not a red-cell model");
    hipStream_t s = g->lat->stream;
    const int np = g->np;
    std::vector<int> act(np, 0);
    act[pid] = 1;
    std::vector<double> damp(np, damping);
    DevBuf<int> d_active, d_step;
    DevBuf<double> d_damp, d_mass, d_v;
    if ((rc = d_active.alloc(np)) || (rc = d_step.alloc(1)) || (rc = d_damp.alloc(np)) || (rc = d_mass.alloc(g->NV)) ||
        (rc = d_v.alloc(np)))
        return rc;
    BL_HIP(hipMemcpyAsync(d_active.p, act.data(), sizeof(int) * np, hipMemcpyHostToDevice, s));
    BL_HIP(hipMemcpyAsync(d_damp.p, damp.data(), sizeof(double) * np, hipMemcpyHostToDevice, s));
    BL_HIP(hipMemsetAsync(d_step.p, 0, sizeof(int), s));
    // the vertex forces in place are those of the current positions
    // (compute_forces before the first step); n_it = -1: vol_desired stays
    // as the parameters hold it
    if ((rc = calib_step(g, s, d_active.p, d_damp.p, d_v.p, d_v.p, -1, d_mass.p, d_step.p))) return rc;
    if ((rc = update_particles(g, 0, s)) || (rc = membrane_forces(g, s))) return rc;
    BL_HIP(hipStreamSynchronize(s));
    g->binned = false;
    return BL_OK;
}

int bl_ibm_set_vertex_forces(bl_ibm g, const double* force)
{
    int rc = need_init(g);
    if (rc) return rc;
    hipStream_t s = g->lat->stream;
    if (!force) {
        if (g->has_ext) ibm_invalidate_graphs(g);
        g->has_ext = false;
        return BL_OK;
    }
    if (!g->ext_force) BL_HIP(hipMalloc(&g->ext_force, (size_t)g->vreal * 3 * std::max<long>(g->NV, 1)));
    if ((rc = vreal_upload(g, g->ext_force, 0, force, 3 * g->NV, s))) return rc;
    if (!g->has_ext) ibm_invalidate_graphs(g);
    g->has_ext = true;
    return BL_OK;
}

int bl_particle_membrane_forces(bl_ibm g)
{
    int rc = need_init(g);
    if (rc) return rc;
    hipStream_t s = g->lat->stream;
    if ((rc = update_particles(g, 0, s)) || (rc = membrane_forces(g, s))) return rc;
    BL_HIP(hipStreamSynchronize(s));
    return BL_OK;
}

int bl_particle_observables(bl_ibm g, int pid, double out[14])
{
    int rc = need_init(g);
    if (rc) return rc;
    if (!out || pid < 0 || pid >= g->np) return fail(BL_ERR_ARG, "bl_particle_observables: bad argument");
    double geom[11];
    BL_HIP(hipMemcpyAsync(geom, g->p_geom + 11 * pid, sizeof(geom), hipMemcpyDeviceToHost, g->lat->stream));
    const int nv = g->shapes[g->parts[pid].shape].nv;
    std::vector<double> pos(3 * (size_t)nv);
    rc = bl_particle_get(g, pid, 0, pos.data());
    if (rc) return rc;
    // Moment of inertia of the enclosed volume about its centroid
    // (src/fsi/DeformableParticle3D.hh:173-272), FP64 on the host.
    const ShapeHost& s = g->shapes[g->parts[pid].shape];
    double in[10] = {0};
    for (int t = 0; t < s.nt; ++t) {
        const double* v[3] = {&pos[3 * s.tri[3 * t]], &pos[3 * s.tri[3 * t + 1]], &pos[3 * s.tri[3 * t + 2]]};
        const double a[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
        const double b[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
        const double N[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        double f1[3], f2[3], f3[3], g0[3], g1[3], g2[3];
        for (int d = 0; d < 3; ++d) {
            const double tmp0 = v[0][d] + v[1][d];
            f1[d] = tmp0 + v[2][d];
            const double tmp1 = v[0][d] * v[0][d];
            const double tmp2 = tmp1 + v[1][d] * tmp0;
            f2[d] = tmp2 + v[2][d] * f1[d];
            f3[d] = v[0][d] * tmp1 + v[1][d] * tmp2 + v[2][d] * f2[d];
            g0[d] = f2[d] + v[0][d] * (f1[d] + v[0][d]);
            g1[d] = f2[d] + v[1][d] * (f1[d] + v[1][d]);
            g2[d] = f2[d] + v[2][d] * (f1[d] + v[2][d]);
        }
        in[0] += N[0] * f1[0];
        in[1] += N[0] * f2[0];
        in[2] += N[1] * f2[1];
        in[3] += N[2] * f2[2];
        in[4] += N[0] * f3[0];
        in[5] += N[1] * f3[1];
        in[6] += N[2] * f3[2];
        in[7] += N[0] * (v[0][1] * g0[0] + v[1][1] * g1[0] + v[2][1] * g2[0]);
        in[8] += N[1] * (v[0][2] * g0[1] + v[1][2] * g1[1] + v[2][2] * g2[1]);
        in[9] += N[2] * (v[0][0] * g0[2] + v[1][0] * g1[2] + v[2][0] * g2[2]);
    }
    in[0] *= 1.0 / 6.0;
    for (int k = 1; k <= 3; ++k) in[k] *= 1.0 / 24.0;
    for (int k = 4; k <= 6; ++k) in[k] *= 1.0 / 60.0;
    for (int k = 7; k <= 9; ++k) in[k] *= 1.0 / 120.0;
    const double V = in[0];
    const double c[3] = {in[1] / V, in[2] / V, in[3] / V};
    double* I = out + 5;
    I[0] = in[5] + in[6] - V * (c[1] * c[1] + c[2] * c[2]);
    I[1] = -in[7] + V * c[0] * c[1];
    I[2] = -in[9] + V * c[2] * c[0];
    I[3] = I[1];
    I[4] = in[4] + in[6] - V * (c[2] * c[2] + c[0] * c[0]);
    I[5] = -in[8] + V * c[1] * c[2];
    I[6] = I[2];
    I[7] = I[5];
    I[8] = in[4] + in[5] - V * (c[0] * c[0] + c[1] * c[1]);
    for (int k = 0; k < 5; ++k) out[k] = geom[k];
    return BL_OK;
}

int bl_ibm_get_phase(bl_ibm g, unsigned char* phase)
{
    int rc = need_init(g);
    if (rc) return rc;
    if (!phase) return fail(BL_ERR_ARG, "bl_ibm_get_phase: null argument");
    BL_HIP(hipMemcpyAsync(phase, g->phase, g->lat->ncell, hipMemcpyDeviceToHost, g->lat->stream));
    BL_HIP(hipStreamSynchronize(g->lat->stream));
    return BL_OK;
}

int bl_ibm_tile_stats(bl_ibm g, long long out[2])
{
    int rc = need_init(g);
    if (rc) return rc;
    if (!out) return fail(BL_ERR_ARG, "bl_ibm_tile_stats: null argument");
    int st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    BL_HIP(hipMemcpyAsync(st, g->n_active, sizeof(st), hipMemcpyDeviceToHost, g->lat->stream));
    BL_HIP(hipStreamSynchronize(g->lat->stream));
    out[0] = st[g->last_list];
    out[1] = st[5 + g->last_list];
    return BL_OK;
}

int bl_ibm_set_fused_staging(bl_ibm g, int mode)
{
    if (!g) return fail(BL_ERR_ARG, "null ibm handle");
    if (mode < -1 || mode > 1) return fail(BL_ERR_ARG, "bl_ibm_set_fused_staging: mode must be -1, 0 or 1");
    if (mode != g->fused_staging) ibm_invalidate_graphs(g);
    g->fused_staging = mode;
    return BL_OK;
}

int bl_ibm_set_layered(bl_ibm g, int mode)
{
    if (!g) return fail(BL_ERR_ARG, "bl_ibm_set_layered: null handle");
    if (mode < -1 || mode > 1) return fail(BL_ERR_ARG, "bl_ibm_set_layered: mode must be -1, 0 or 1");
    if (mode != g->layered) ibm_invalidate_graphs(g);
    g->layered = mode;
    return BL_OK;
}

int bl_ibm_set_launch(bl_ibm g, const int params[BL_IBM_NLAUNCH])
{
    if (!g || !params) return fail(BL_ERR_ARG, "bl_ibm_set_launch: null argument");
    auto in = [](int v, std::initializer_list<int> ok) { return std::find(ok.begin(), ok.end(), v) != ok.end(); };
    const int* p = params;
    if (!in(p[BL_IBM_BIN_BLOCK], {64, 128, 256}) || !in(p[BL_IBM_ELEM_BLOCK], {64, 128, 256}) ||
        !in(p[BL_IBM_GATHER_LANES], {2, 4, 8}) || !in(p[BL_IBM_INTERP_BLOCK], {64, 128, 256}) ||
        !in(p[BL_IBM_SPREAD_SMALL], {1, 2, 4, 8}) || !in(p[BL_IBM_SPREAD_BIG], {16, 32, 64, 128}) ||
        !in(p[BL_IBM_XCD], {0, 2, 4, 6}))
        return fail(BL_ERR_ARG, "bl_ibm_set_launch: a launch parameter is outside its allowed values");
    bool same = true;
    for (int k = 0; k < BL_IBM_NLAUNCH; ++k) same = same && g->lp[k] == p[k];
    if (same) return BL_OK;
    if (g->lat) BL_HIP(hipStreamSynchronize(g->lat->stream));
    for (int k = 0; k < BL_IBM_NLAUNCH; ++k) g->lp[k] = p[k];
    ibm_invalidate_graphs(g);
    return BL_OK;
}

int bl_ibm_get_launch(bl_ibm g, int params[BL_IBM_NLAUNCH])
{
    if (!g || !params) return fail(BL_ERR_ARG, "bl_ibm_get_launch: null argument");
    for (int k = 0; k < BL_IBM_NLAUNCH; ++k) params[k] = g->lp[k];
    return BL_OK;
}

int bl_ibm_stage_times(bl_ibm g, double out[5])
{
    if (!g || !out) return fail(BL_ERR_ARG, "bl_ibm_stage_times: null argument");
    for (int k = 0; k < 5; ++k) out[k] = g->stage_ms[k];
    return BL_OK;
}

}  // extern "C"
