// D3Q19 Guo-VOF collide-and-stream for gfx950.
//
// One fused pass per time step over structure-of-arrays populations, in place
// with the AA access pattern:
//   even step: read f_i(x) from plane i at x, collide, write f*_i(x) to plane
//              opp(i) at x (purely local);
//   odd step:  read f_i(x) from plane opp(i) at x - c_i, collide, write
//              f*_i(x) to plane i at x + c_i.
// Every slot is read and written by the same thread, so no step needs a
// second lattice.  Non-periodic faces reflect (a population leaving the box
// re-enters the same cell in the opposite direction); wall cells (bounce-back,
// velocity wall) sit inside the box.
//
// The collision restates GuoExternalForceVOFDynamics::collide
// (src/fsi/VariableViscosityDynamics.hh:151-172): fBar moments
// (src/latticeBoltzmann/momentTemplates3D.h:375-399), u = j/rho + F/2
// (src/basicDynamics/externalForceDynamics.hh:59-73), phase-dependent omega
// (src/fsi/VariableViscosityDynamics.hh:131-141), BGK in the fBar form
// (src/latticeBoltzmann/dynamicsTemplates3D.h:759-838) and the Guo source
// (src/fsi/VariableViscosityDynamics.hh:47-129).
#include "ibm.hpp"
#include "lattice_dev.hpp"

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

namespace bl {

FastDiv make_fastdiv(unsigned d)
{
    FastDiv f;
    while ((1ull << f.l) < d) ++f.l;
    f.m = (unsigned)((((unsigned long long)1 << 32) * ((1ull << f.l) - d)) / d + 1);
    return f;
}

LatView view_of(const bl_lattice_s* L)
{
    LatView v;
    v.nx = L->nx; v.ny = L->ny; v.nz = L->nz;
    v.dy = make_fastdiv((unsigned)L->ny);
    v.dz = make_fastdiv((unsigned)L->nz);
    v.ncell = (unsigned)L->ncell;
    v.stride = L->stride;
    for (int d = 0; d < 3; ++d) v.per[d] = L->periodic[d];
    v.pops = L->pops;
    v.mask = L->mask;
    v.inert = L->skip_inert ? L->inert : nullptr;
    v.phase = L->phase;
    v.flags = L->tile_flags;
    v.force = L->force;
    for (int d = 0; d < 3; ++d) v.g[d] = L->g[d];
    for (int d = 0; d < 3; ++d) v.g_last[d] = L->g_last[d];
    v.om0 = L->omega_c0;
    v.om1 = L->omega_c1;
    v.om_base = L->omega_out;
    v.ntx = L->ntx; v.nty = L->nty; v.ntz = L->ntz;
    v.wall_axis = L->wall_axis;
    for (int k = 0; k < 6; ++k) v.wall_u[k] = L->wall_u[k];
    v.xcd_remap = L->xcd_remap[0];
    v.row_segs = 1;
    v.emit_u = L->emit_u;
    v.emit_force = L->emit_force;
    v.emit_flags = L->emit_flags;
    return v;
}

// ------------------------------------------------------ collide & stream
// FMODE: 0 no force, 1 uniform body force, 2 body force + tile-gated field
// ODD steps address the neighbours' slots: element offsets from the start
// of the population buffer, 32-bit (`unsigned`) unless the buffer holds
// 2^32 values or more (`size_t`).  A slot past a non-periodic face is
// replaced by the cell's own slot of the opposite population (reflection),
// chosen per lane without branching.  Waves whose lanes all sit one cell or
// more inside the x and y faces (nearly all of them) take the x/y part of
// every neighbour offset as a wave-uniform value; only z is per lane.
template <class K>
__device__ __forceinline__ K odd_slot(const NbrOff& nb, bool inner, int sx, int sy, unsigned c, long st,
                                      int dx, int dy, int dz, int p_valid, int p_face)
{
    bool valid;
    const int d = inner ? nb.off_inner(dx, dy, dz, sx, sy, valid) : nb.off(dx, dy, dz, valid);
    return valid ? (K)p_valid * (K)st + (K)(c + d) : (K)p_face * (K)st + (K)c;
}

// What one cell's update reads besides its populations: the mask code, the
// collision force and relaxation rate, and for the velocity staging (EMIT)
// whether the cell's tile is flagged and the last collision's force at the
// cell.  Gathered in one place so that the z-row kernel can issue these
// loads before its first barrier.
template <class C>
struct CellIn {
    unsigned char m;
    bool emit;
    C F[3], omega, Fe[3];
};

// The tile's flags (the force field's, v.flags, and the last collision's,
// v.emit_flags) load first, since the force loads wait for them.
struct TileFlags {
    bool cur, emit;
};

template <int FMODE, bool EMIT>
__device__ __forceinline__ TileFlags cs_flags(const LatView& v, int x, int y, int z)
{
    TileFlags tf{false, false};
    if (FMODE == 2 || EMIT) {
        const long t = ((long)(x >> 2) * v.nty + (y >> 2)) * v.ntz + (z >> 2);
        if (FMODE == 2) tf.cur = v.flags[t] != 0;
        if (EMIT && v.emit_flags) tf.emit = v.emit_flags[t] != 0;
    }
    return tf;
}

template <class S, bool MASK, int FMODE, bool PHASE, bool EMIT>
__device__ __forceinline__ CellIn<typename Pop<S>::C> cs_inputs(const LatView& v, unsigned c, const TileFlags& tf)
{
    using C = typename Pop<S>::C;
    using FT = typename Pop<S>::F;
    CellIn<C> in;
    in.m = MASK ? v.mask[c] : (unsigned char)M_FLUID;
    in.F[0] = in.F[1] = in.F[2] = C(0);
    if (FMODE == 1) { in.F[0] = (C)v.g[0]; in.F[1] = (C)v.g[1]; in.F[2] = (C)v.g[2]; }
    if (FMODE == 2) cell_force_fl<FT, C>(v, v.g, tf.cur, v.force, c, in.F);
    in.omega = PHASE ? (v.phase[c] ? (C)v.om1 : (C)v.om0) : (C)v.om0;
    in.emit = EMIT && tf.cur;
    in.Fe[0] = in.Fe[1] = in.Fe[2] = C(0);
    if (EMIT && in.emit && in.m == M_FLUID) cell_force_fl<FT, C>(v, v.g_last, tf.emit, v.emit_force, c, in.Fe);
    return in;
}

// The update of one cell from its loaded populations f (in the collide
// order) and its other inputs: moments, the velocity staging (EMIT), then
// the collision of fluid and velocity-wall cells or the swap of a
// bounce-back cell.
template <class S, bool FORCE, bool EMIT>
__device__ __forceinline__ void cs_update(const LatView& v, unsigned c, int x, int y, int z, typename Pop<S>::C f[Q],
                                          const CellIn<typename Pop<S>::C>& in)
{
    using C = typename Pop<S>::C;
    // the moments of the loaded populations, shared by the staging (EMIT)
    // and the collision: the same expressions either would form
    C rhoBar, jm[3];
    wmoments<S>(f, rhoBar, jm);
    const C invRho = C(1) / (rhoBar + C(1));
    if (EMIT && in.emit) ((Vel4<C>*)v.emit_u)[c] = staged_velocity_f<C>(v, in.m, x, y, z, jm, invRho, in.Fe);
    if (in.m == M_FLUID || in.m == M_VWALL) {
        wcollide<S, FORCE>(f, in.F, in.omega, rhoBar, jm, invRho);
    } else if (in.m == M_BB) {
#pragma unroll
        for (int i = 1; i <= 9; ++i) {
            const C t = f[i];
            f[i] = f[i + 9];
            f[i + 9] = t;
        }
    }
}

// One cell of the collide-and-stream pass.  EMIT: the velocity staging of an
// immersed-boundary step rides on the launch (bl_lattice_s::emit_u): cells
// of flagged tiles also write the computeVelocity value of the populations
// they loaded, with the last collision's force field.  NT: nontemporal hints
// on the population loads and stores, 1 on all of them, 2 on the planes the
// odd step addresses z-aligned only (on even steps 2 is 1).
template <class S, bool ODD, bool MASK, int FMODE, bool PHASE, class K, bool EMIT, int NT>
__device__ __forceinline__ void cs_cell(const LatView& v, unsigned c, int x, int y, int z)
{
    using P = Pop<S>;
    using C = typename P::C;
    S* __restrict__ A = (S*)v.pops;
    const long st = v.stride;
    // an inert bounce-back cell: what it holds never reaches a fluid cell
    // (a bounce-back cell hands each neighbour back what that neighbour
    // sent, and every neighbour of an inert cell is bounce-back); only its
    // staged velocity, zero on a bounce-back cell, is still written
    if (MASK && v.inert && ((v.inert[c >> 6] >> (c & 63u)) & 1ull)) {
        if (EMIT && cs_flags<FMODE, EMIT>(v, x, y, z).cur) ((Vel4<C>*)v.emit_u)[c] = Vel4<C>{C(0), C(0), C(0), C(0)};
        return;
    }

    NbrOff nb;
    bool inner = false;
    const int sx = v.ny * v.nz, sy = v.nz;
    if (ODD) {
        nb.init(v, x, y, z);
        inner = __all(x > 0 && x < v.nx - 1 && y > 0 && y < v.ny - 1) != 0;
    }

    // the tile flags first: the force loads wait for them, and so they
    // arrive while the populations load
    const TileFlags tf = cs_flags<FMODE, EMIT>(v, x, y, z);
    C f[Q];
    // inner waves of a lattice periodic along z: every neighbour exists (x
    // and y one cell or more from their faces, z wraps), so no address needs
    // the reflection's select (a uniform branch; the other waves take the
    // general form)
    const bool all_valid = ODD && inner && v.per[2];
    if (ODD && all_valid) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            bool valid;
            const K k = (K)opp(i) * (K)st + (K)(c + nb.off_inner(-cx(i), -cy(i), -cz(i), sx, sy, valid));
            f[i] = (NT == 1 || (NT == 2 && cz(i) == 0)) ? P::template ldw<K, true>(A, k, i) : P::template ldw<K, false>(A, k, i);
        }
    } else if (ODD) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const K k = odd_slot<K>(nb, inner, sx, sy, c, st, -cx(i), -cy(i), -cz(i), opp(i), i);
            f[i] = (NT == 1 || (NT == 2 && cz(i) == 0)) ? P::template ldw<K, true>(A, k, i) : P::template ldw<K, false>(A, k, i);
        }
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = P::template ldw<unsigned, NT != 0>(A + i * st, c, i);
    }

    cs_update<S, FMODE != 0, EMIT>(v, c, x, y, z, f, cs_inputs<S, MASK, FMODE, PHASE, EMIT>(v, c, tf));

    if (ODD && all_valid) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            bool valid;
            const K k = (K)i * (K)st + (K)(c + nb.off_inner(cx(i), cy(i), cz(i), sx, sy, valid));
            if (NT == 1 || (NT == 2 && cz(i) == 0)) P::template stw<K, true>(A, k, i, f[i]);
            else P::template stw<K, false>(A, k, i, f[i]);
        }
    } else if (ODD) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const K k = odd_slot<K>(nb, inner, sx, sy, c, st, cx(i), cy(i), cz(i), i, opp(i));
            if (NT == 1 || (NT == 2 && cz(i) == 0)) P::template stw<K, true>(A, k, i, f[i]);
            else P::template stw<K, false>(A, k, i, f[i]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) P::template stw<unsigned, NT != 0>(A + opp(i) * st, c, i, f[i]);
    }
}

// Slot of population i in the LDS rows of the z-segment kernel: the ten
// populations with c_z != 0, in index order (-1 for the others).
__host__ __device__ constexpr int zslot(int i)
{
    int k = 0;
    for (int j = 0; j < i; ++j) k += cz(j) != 0;
    return cz(i) != 0 ? k : -1;
}
static_assert(zslot(3) == 0 && zslot(7) == 2 && zslot(18) == 9, "ten populations move along z");

// The odd step over z-segments (launch block -B: one workgroup of B lanes
// per segment of B cells of a z-row, nz a multiple of B, z periodic): lane
// l reads and writes slot zb + l of every row it addresses, so the
// population accesses are z-aligned.  The z +-1 shift of the ten
// populations with c_z != 0 goes through LDS: after the loads lane l takes
// the value that lane l - c_z loaded (the first and last lanes also load
// the one cell before and after the segment), before the stores it hands
// its value to lane l + c_z, and the value leaving the segment is stored
// by its own lane.  A row past a non-periodic x or y face reflects, as in
// cs_cell: its populations come from, and go to, the cell's own slots of
// the opposite plane, with no shift.  Row offsets are uniform over the
// workgroup, so addresses are a scalar base plus the lane's z.
// v.row_segs holds nz / B; v.xcd_remap deals the segments over the XCDs
// as collide_block deals blocks.
template <class S, bool MASK, int FMODE, bool PHASE, bool EMIT, bool NT>
__global__ void __launch_bounds__(256) k_collide_zseg(LatView v)
{
    using P = Pop<S>;
    using C = typename P::C;
    extern __shared__ double lds_rows[];
    const int B = blockDim.x, W = B + 2;
    C* rot = (C*)lds_rows;  // [10][B + 2]: slot l + 1 the segment's cell l, 0 and B + 1 its neighbours
    const int nz = v.nz;
    const unsigned nseg = (unsigned)v.row_segs;
    const unsigned bid = collide_block(blockIdx.x, gridDim.x, v.xcd_remap);
    const unsigned row = bid / nseg;
    const int zb = (int)(bid - row * nseg) * B;
    const int l = threadIdx.x, z = zb + l;
    const unsigned q = fdiv(row, v.dy);
    const int x = (int)q, y = (int)(row - q * (unsigned)v.ny);
    const unsigned c = row * (unsigned)nz + (unsigned)z;
    const unsigned own = row * (unsigned)nz;
    S* __restrict__ A = (S*)v.pops;
    const unsigned st = (unsigned)v.stride;  // Q * stride < 2^32 (host checks)
    // start of row (x + a - 1, y + b - 1), or -1 past a non-periodic face
    int R[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int xs = wrapc(x + a - 1, v.nx, v.per[0]), ys = wrapc(y + b - 1, v.ny, v.per[1]);
            R[a][b] = (xs | ys) < 0 ? -1 : (xs * v.ny + ys) * nz;
        }
    const unsigned zlo = zb == 0 ? nz - 1 : zb - 1;    // the cell before the segment
    const unsigned zhi = zb + B == nz ? 0 : zb + B;    // and after it

    // the tile flags first: the force loads wait for them, and so they
    // arrive while the populations load
    const TileFlags tf = cs_flags<FMODE, EMIT>(v, x, y, z);
    C f[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int rb = R[1 - cx(i)][1 - cy(i)];
        const unsigned base = rb >= 0 ? (unsigned)opp(i) * st + (unsigned)rb : (unsigned)i * st + own;
        f[i] = P::template ldw<unsigned, NT>(A + base, (unsigned)z, i);
    }
    // the cell's other inputs load beside the populations, not after the
    // barrier below, which no load crosses
    const CellIn<C> in = cs_inputs<S, MASK, FMODE, PHASE, EMIT>(v, c, tf);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (cz(i) == 0) continue;
        const int rb = R[1 - cx(i)][1 - cy(i)];
        const unsigned base = rb >= 0 ? (unsigned)opp(i) * st + (unsigned)rb : (unsigned)i * st + own;
        C* r = rot + zslot(i) * W;
        r[l + 1] = f[i];
        if (cz(i) > 0 && l == 0) r[0] = P::template ldw<unsigned, NT>(A + base, zlo, i);
        if (cz(i) < 0 && l == B - 1) r[B + 1] = P::template ldw<unsigned, NT>(A + base, zhi, i);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (cz(i) == 0) continue;
        const C w = rot[zslot(i) * W + l + 1 - cz(i)];
        f[i] = R[1 - cx(i)][1 - cy(i)] >= 0 ? w : f[i];
    }

    cs_update<S, FMODE != 0, EMIT>(v, c, x, y, z, f, in);

    __syncthreads();  // every lane has read its shifted values
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int rb = R[1 + cx(i)][1 + cy(i)];
        const unsigned base = rb >= 0 ? (unsigned)i * st + (unsigned)rb : (unsigned)opp(i) * st + own;
        if (cz(i) == 0) {
            P::template stw<unsigned, NT>(A + base, (unsigned)z, i, f[i]);
            continue;
        }
        rot[zslot(i) * W + l + 1] = f[i];
        // the value leaving the segment
        if (rb >= 0 && cz(i) > 0 && l == B - 1) P::template stw<unsigned, NT>(A + base, zhi, i, f[i]);
        if (rb >= 0 && cz(i) < 0 && l == 0) P::template stw<unsigned, NT>(A + base, zlo, i, f[i]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (cz(i) == 0) continue;
        const int rb = R[1 + cx(i)][1 + cy(i)];
        const unsigned base = rb >= 0 ? (unsigned)i * st + (unsigned)rb : (unsigned)opp(i) * st + own;
        const int src = l - cz(i);  // the lane whose value lands on slot z
        const C w = rot[zslot(i) * W + src + 1];
        if (rb < 0 || (src >= 0 && src < B)) P::template stw<unsigned, NT>(A + base, (unsigned)z, i, rb >= 0 ? w : f[i]);
    }
}

template <class S, bool ODD, bool MASK, int FMODE, bool PHASE, class K, bool EMIT = false, int NT = 0>
__global__ void __launch_bounds__(256) k_collide_stream(LatView v)
{
    unsigned c;
    int x, y, z;
    if (v.xcd_remap == BL_ORDER_ROWS) {
        if (!row_cell(v, c, x, y, z)) return;
    } else {
        c = collide_block(blockIdx.x, gridDim.x, v.xcd_remap) * blockDim.x + threadIdx.x;
        if (c >= v.ncell) return;
        decode(c, v, x, y, z);
    }
    cs_cell<S, ODD, MASK, FMODE, PHASE, K, EMIT, NT>(v, c, x, y, z);
}

template <class S, bool ODD, bool MASK, int FMODE, bool PHASE, class K, bool EMIT, int NT>
void launch_one(const LatView& v, unsigned nblk, int blk, hipStream_t s)
{
    if constexpr (ODD && std::is_same<K, unsigned>::value) {
        if (blk < 0) {  // collide_plan checked zseg_ok: nblk segments of -blk lanes
            const size_t lds = 10 * (size_t)(2 - blk) * sizeof(typename Pop<S>::C);
            hipLaunchKernelGGL((k_collide_zseg<S, MASK, FMODE, PHASE, EMIT, NT != 0>), dim3(nblk), dim3(-blk), lds, s, v);
            return;
        }
    }
    hipLaunchKernelGGL((k_collide_stream<S, ODD, MASK, FMODE, PHASE, K, EMIT, NT>), dim3(nblk), dim3(blk), 0, s, v);
}

template <class S, bool ODD, bool MASK, int FMODE, int NT>
void launch_cs4(const LatView& v, bool phase, unsigned nblk, int blk, hipStream_t s)
{
    const bool big = ODD && (size_t)Q * (size_t)v.stride > 0xffffffffull;
    if constexpr (FMODE == 2) {
        if (v.emit_u) {  // the launcher passes emit_u only without a phase field
            if (big) launch_one<S, ODD, MASK, 2, false, size_t, true, NT>(v, nblk, blk, s);
            else launch_one<S, ODD, MASK, 2, false, unsigned, true, NT>(v, nblk, blk, s);
            return;
        }
    }
    if (big) {
        if (phase) launch_one<S, ODD, MASK, FMODE, true, size_t, false, NT>(v, nblk, blk, s);
        else launch_one<S, ODD, MASK, FMODE, false, size_t, false, NT>(v, nblk, blk, s);
    } else {
        if (phase) launch_one<S, ODD, MASK, FMODE, true, unsigned, false, NT>(v, nblk, blk, s);
        else launch_one<S, ODD, MASK, FMODE, false, unsigned, false, NT>(v, nblk, blk, s);
    }
}

template <class S, bool ODD, bool MASK, int FMODE>
void launch_cs3(const LatView& v, bool phase, unsigned nblk, int blk, hipStream_t s, int nt)
{
    if (nt == 0) launch_cs4<S, ODD, MASK, FMODE, 0>(v, phase, nblk, blk, s);
    else if (nt == 1 || !ODD) launch_cs4<S, ODD, MASK, FMODE, 1>(v, phase, nblk, blk, s);
    else launch_cs4<S, ODD, MASK, FMODE, 2>(v, phase, nblk, blk, s);
}

template <class S, bool ODD, bool MASK>
void launch_cs2(const LatView& v, int fmode, bool phase, unsigned nblk, int blk, hipStream_t s, int nt)
{
    if (fmode == 0) launch_cs3<S, ODD, MASK, 0>(v, phase, nblk, blk, s, nt);
    else if (fmode == 1) launch_cs3<S, ODD, MASK, 1>(v, phase, nblk, blk, s, nt);
    else launch_cs3<S, ODD, MASK, 2>(v, phase, nblk, blk, s, nt);
}

template <class S>
void launch_cs(const LatView& v, bool odd, bool mask, int fmode, bool phase, unsigned nblk, int blk,
               hipStream_t s, int nt)
{
    if (odd) {
        if (mask) launch_cs2<S, true, true>(v, fmode, phase, nblk, blk, s, nt);
        else launch_cs2<S, true, false>(v, fmode, phase, nblk, blk, s, nt);
    } else {
        if (mask) launch_cs2<S, false, true>(v, fmode, phase, nblk, blk, s, nt);
        else launch_cs2<S, false, false>(v, fmode, phase, nblk, blk, s, nt);
    }
}

// The z-segment kernel's conditions for segments of B lanes: z periodic,
// nz a multiple of B, 32-bit slot offsets.
bool zseg_ok(const bl_lattice_s* L, int B)
{
    return L->periodic[2] && B >= 64 && B <= 256 && B % 64 == 0 && L->nz % B == 0 &&
           (size_t)Q * (size_t)L->stride <= 0xffffffffull;
}

static int force_mode(const bl_lattice_s* L)
{
    if (L->tile_flags && L->force_field) return 2;
    if (L->g[0] != 0.0 || L->g[1] != 0.0 || L->g[2] != 0.0) return 1;
    return 0;
}

CollidePlan collide_plan(const bl_lattice_s* L)
{
    CollidePlan p;
    LatView& v = p.v;
    v = view_of(L);
    p.odd = (L->steps & 1) != 0;
    p.fm = force_mode(L);
    if (p.fm != 2) v.flags = nullptr;
    p.phase = L->phase != nullptr && L->omega_c0 != L->omega_c1;
    if (p.fm != 2 || p.phase) v.emit_u = nullptr;
    p.mask = L->mask != nullptr;
    p.storage = L->storage;
    const int par = p.odd ? 1 : 0;
    p.blk = L->block[par];
    v.xcd_remap = L->xcd_remap[par];
    p.nt = L->nontemporal[par];
    if (p.blk < 0) {
        // the z-segment kernel steps every cell (its workgroups exchange
        // values through LDS), so it does not run with inert-cell skipping
        if (p.odd && zseg_ok(L, -p.blk) && !v.inert) {
            v.row_segs = L->nz / -p.blk;
            p.nblk = (unsigned)((long)L->nx * L->ny * v.row_segs);
            return p;
        }
        p.blk = -p.blk;  // even steps, and lattices the z-segment kernel cannot take
    }
    p.nblk = (unsigned)((L->ncell + p.blk - 1) / p.blk);
    if (v.xcd_remap == BL_ORDER_ROWS) {
        v.row_segs = (L->nz + p.blk - 1) / p.blk;
        p.nblk = (unsigned)(((long)L->nx * L->ny + 7) / 8 * 8 * v.row_segs);
    }
    return p;
}

void collide_commit(bl_lattice_s* L)
{
    L->steps += 1;
    for (int d = 0; d < 3; ++d) L->g_last[d] = L->g[d];
    L->collided = true;
}

int collide_stream_launch(bl_lattice_s* L)
{
    const CollidePlan p = collide_plan(L);
    if (p.storage == BL_STORE_F64) launch_cs<double>(p.v, p.odd, p.mask, p.fm, p.phase, p.nblk, p.blk, L->stream, p.nt);
    else if (p.storage == BL_STORE_F32) launch_cs<float>(p.v, p.odd, p.mask, p.fm, p.phase, p.nblk, p.blk, L->stream, p.nt);
    else launch_cs<__half>(p.v, p.odd, p.mask, p.fm, p.phase, p.nblk, p.blk, L->stream, p.nt);
    BL_CHECK_LAUNCH();
    collide_commit(L);
    return BL_OK;
}

// ------------------------------------------- batches of independent boxes
// One collide-and-stream launch over the lattices of an ensemble: workgroup
// b belongs to box k with base[k] <= b < base[k + 1] (a binary search over
// the box offsets, uniform over the workgroup) and runs local block
// b - base[k] of that box with its own view.  Every box shares the template
// choices (storage, parity, mask, force mode); the launch shape is 256
// threads, in the order `order` deals them over the XCDs (any order gives
// the same values), with the nontemporal level NT of cs_cell.
// EMIT: every box's view carries its velocity staging (emit_u and the last
// collision's field), as a single lattice's step with fused staging.
template <class S, bool ODD, bool MASK, int FMODE, bool EMIT = false, int NT = 0>
__global__ void __launch_bounds__(256) k_collide_batch(const LatView* __restrict__ views, const int* __restrict__ base,
                                                       int nbox, int order)
{
    // `order` deals the batch's workgroups over the XCDs (collide_block)
    const int b = (int)collide_block(blockIdx.x, gridDim.x, order);
    const int box = batch_box(base, nbox, b);
    const LatView v = views[box];
    const unsigned c = (unsigned)(b - base[box]) * blockDim.x + threadIdx.x;
    if (c >= v.ncell) return;
    int x, y, z;
    decode(c, v, x, y, z);
    cs_cell<S, ODD, MASK, FMODE, false, unsigned, EMIT, NT>(v, c, x, y, z);
}

template <class S, bool ODD, bool MASK, int FM, bool EMIT>
void launch_collide_batch3(int nt, const LatView* views, const int* base, int nbox, unsigned nblk, int order,
                           hipStream_t s)
{
    if (nt == 1)
        hipLaunchKernelGGL((k_collide_batch<S, ODD, MASK, FM, EMIT, 1>), dim3(nblk), dim3(256), 0, s, views, base,
                           nbox, order);
    else if (nt == 2)
        hipLaunchKernelGGL((k_collide_batch<S, ODD, MASK, FM, EMIT, 2>), dim3(nblk), dim3(256), 0, s, views, base,
                           nbox, order);
    else
        hipLaunchKernelGGL((k_collide_batch<S, ODD, MASK, FM, EMIT, 0>), dim3(nblk), dim3(256), 0, s, views, base,
                           nbox, order);
}

template <class S, bool ODD, bool MASK>
void launch_collide_batch2(int fm, bool emit, const LatView* views, const int* base, int nbox, unsigned nblk, int order,
                           int nt, hipStream_t s)
{
    if (fm == 0) launch_collide_batch3<S, ODD, MASK, 0, false>(nt, views, base, nbox, nblk, order, s);
    else if (fm == 1) launch_collide_batch3<S, ODD, MASK, 1, false>(nt, views, base, nbox, nblk, order, s);
    else if (emit) launch_collide_batch3<S, ODD, MASK, 2, true>(nt, views, base, nbox, nblk, order, s);
    else launch_collide_batch3<S, ODD, MASK, 2, false>(nt, views, base, nbox, nblk, order, s);
}

template <class S>
void launch_collide_batch1(bool odd, bool mask, int fm, bool emit, const LatView* views, const int* base, int nbox,
                           unsigned nblk, int order, int nt, hipStream_t s)
{
    if (odd) {
        if (mask) launch_collide_batch2<S, true, true>(fm, emit, views, base, nbox, nblk, order, nt, s);
        else launch_collide_batch2<S, true, false>(fm, emit, views, base, nbox, nblk, order, nt, s);
    } else {
        if (mask) launch_collide_batch2<S, false, true>(fm, emit, views, base, nbox, nblk, order, nt, s);
        else launch_collide_batch2<S, false, false>(fm, emit, views, base, nbox, nblk, order, nt, s);
    }
}

bool collide_batchable(const CollidePlan& a, const CollidePlan& b)
{
    return a.storage == b.storage && a.odd == b.odd && a.mask == b.mask && a.fm == b.fm && !a.phase && !b.phase &&
           !a.v.emit_u && !b.v.emit_u && (size_t)Q * a.v.stride <= 0xffffffffull &&
           (size_t)Q * b.v.stride <= 0xffffffffull;
}

int collide_batch_launch(const CollidePlan& p0, bool emit, const LatView* views, const int* base, int nbox,
                         unsigned nblk, int order, int nt, hipStream_t s)
{
    if (emit && p0.fm != 2) return fail(BL_ERR_STATE, "batched collide: velocity staging needs the force field");
    if (nt < 0 || nt > 2) return fail(BL_ERR_ARG, "batched collide: nontemporal level must be 0, 1 or 2");
    if (p0.storage == BL_STORE_F64)
        launch_collide_batch1<double>(p0.odd, p0.mask, p0.fm, emit, views, base, nbox, nblk, order, nt, s);
    else if (p0.storage == BL_STORE_F32)
        launch_collide_batch1<float>(p0.odd, p0.mask, p0.fm, emit, views, base, nbox, nblk, order, nt, s);
    else launch_collide_batch1<__half>(p0.odd, p0.mask, p0.fm, emit, views, base, nbox, nblk, order, nt, s);
    BL_CHECK_LAUNCH();
    return BL_OK;
}

// ------------------------------------------------ velocity-wall boundary
// Regularized finite-difference velocity boundary applied after streaming
// (PlaneFdBoundaryFunctional3D::processCell,
// src/boundaryCondition/finiteDifferenceBoundaryProcessor3D.hh:68-115) with
// the density rule of VelocityDirichletBoundaryDynamics::computeRhoBar
// (src/boundaryCondition/boundaryDynamics.hh:631-660).  Two launches: the
// first computes, from the post-stream state, the velocities that the
// finite differences read (wall cells and the interior cells next to
// them); the second rewrites every wall cell from its own post-stream
// populations and those velocities.  A cell's population slots are its
// own in either AA layout, so no wall cell reads a value another replaced.
template <class S>
__device__ void cell_velocity_post(const LatView& v, long c, int x, int y, int z, bool swapped, double u[3])
{
    using C = typename Pop<S>::C;
    using FT = typename Pop<S>::F;
    const unsigned char m = v.mask ? v.mask[c] : (unsigned char)M_FLUID;
    if (m == M_VWALL) {
        const int p[3] = {x, y, z};
        const int side = p[v.wall_axis] == 0 ? 0 : 1;
        for (int d = 0; d < 3; ++d) u[d] = v.wall_u[3 * side + d];
        return;
    }
    if (m == M_BB || m == M_NONE) {
        u[0] = u[1] = u[2] = 0.0;
        return;
    }
    Nbr n;
    n.init(v, x, y, z);
    C f[Q];
    load_post<S>(v, n, c, swapped, f);
    C F[3];
    cell_force<FT, C>(v, c, x, y, z, F);
    const C rb = sum_rho(f);
    C j[3];
    sum_j(f, j);
    const C invRho = C(1) / (rb + C(1));
    for (int d = 0; d < 3; ++d) u[d] = (double)(j[d] * invRho + F[d] / C(2));
}

__device__ __forceinline__ void wall_cell_coords(const LatView& v, long w, int& x, int& y, int& z)
{
    const int a = v.wall_axis;
    const int n[3] = {v.nx, v.ny, v.nz};
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    const long per_plane = (long)n[t1] * n[t2];
    const int side = (int)(w / per_plane);
    const long r = w % per_plane;
    int p[3];
    p[a] = side == 0 ? 0 : n[a] - 1;
    p[t1] = (int)(r / n[t2]);
    p[t2] = (int)(r % n[t2]);
    x = p[0]; y = p[1]; z = p[2];
}

// Wall-plane index of a cell on plane `side` of the wall axis.
__device__ __forceinline__ long wall_index(const LatView& v, int side, const int p[3])
{
    const int a = v.wall_axis;
    const int n[3] = {v.nx, v.ny, v.nz};
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    return (long)side * n[t1] * n[t2] + (long)p[t1] * n[t2] + p[t2];
}

// The velocities the finite differences of the wall cells read, computed
// once per cell: cells w < nwall are the wall-plane cells themselves
// (tangential neighbours), cells nwall + w the first interior cell next to
// wall cell w (normal neighbour).  Three per cell in `out`.
template <class S>
__device__ __forceinline__ void vwall_u_cell(const LatView& v, long nwall, bool swapped, double* out, long w)
{
    if (w >= 2 * nwall) return;
    const bool inner = w >= nwall;
    int x, y, z;
    wall_cell_coords(v, inner ? w - nwall : w, x, y, z);
    if (inner) {
        int q[3] = {x, y, z};
        const int a = v.wall_axis;
        q[a] += q[a] == 0 ? 1 : -1;
        x = q[0];
        y = q[1];
        z = q[2];
    }
    double u[3];
    cell_velocity_post<S>(v, ((long)x * v.ny + y) * v.nz + z, x, y, z, swapped, u);
    double* o = out + 3 * w;
    o[0] = u[0];
    o[1] = u[1];
    o[2] = u[2];
}

template <class S>
__global__ void k_vwall_u(LatView v, long nwall, bool swapped, double* out)
{
    vwall_u_cell<S>(v, nwall, swapped, out, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <class S>
__device__ __forceinline__ void vwall_compute_cell(const LatView& v, long nwall, bool swapped, const double* un,
                                                   long w)
{
    using C = typename Pop<S>::C;
    using FT = typename Pop<S>::F;
    if (w >= nwall) return;
    int x, y, z;
    wall_cell_coords(v, w, x, y, z);
    const long c = ((long)x * v.ny + y) * v.nz + z;
    if (!v.mask || v.mask[c] != M_VWALL) return;
    const int a = v.wall_axis;
    const int p[3] = {x, y, z};
    const int side = p[a] == 0 ? 0 : 1;
    const int orient = side == 0 ? -1 : 1;
    const double* uw = v.wall_u + 3 * side;

    Nbr n;
    n.init(v, x, y, z);
    C fc[Q];
    load_post<S>(v, n, c, swapped, fc);
    C Fc[3];
    cell_force<FT, C>(v, c, x, y, z, Fc);
    double onWall = 0.0, normal = 0.0;
    #pragma unroll
    for (int i = 0; i < Q; ++i)
        if (ci(i, a) == 0) onWall += (double)fc[i];
    #pragma unroll
    for (int i = 0; i < Q; ++i)
        if (ci(i, a) == orient) normal += (double)fc[i];
    const double velNormal = (double)orient * (uw[a] - 0.5 * (double)Fc[a]);
    const double rb = (2.0 * normal + onWall - 1.0 * velNormal) / (1.0 + velNormal);
    const double rho = rb + 1.0;

    double du[3][3];
    const int n3[3] = {v.nx, v.ny, v.nz};
    for (int d = 0; d < 3; ++d) {
        if (d == a) {
            const double* u1 = un + 3 * (nwall + w);
            for (int k = 0; k < 3; ++k) du[d][k] = -orient * (-uw[k] + u1[k]);
        } else {
            int qp[3] = {x, y, z}, qm[3] = {x, y, z};
            qp[d] += 1;
            qm[d] -= 1;
            if (v.per[d]) {
                if (qp[d] >= n3[d]) qp[d] -= n3[d];
                if (qm[d] < 0) qm[d] += n3[d];
            } else {
                if (qp[d] >= n3[d]) qp[d] = n3[d] - 1;
                if (qm[d] < 0) qm[d] = 0;
            }
            const double* up = un + 3 * wall_index(v, side, qp);
            const double* um = un + 3 * wall_index(v, side, qm);
            for (int k = 0; k < 3; ++k) du[d][k] = (up[k] - um[k]) / 2.0;
        }
    }
    const double sToPi = -rho / 3.0 / v.om_base;
    double pi[6];
    pi[0] = 2.0 * du[0][0] * sToPi;
    pi[3] = 2.0 * du[1][1] * sToPi;
    pi[5] = 2.0 * du[2][2] * sToPi;
    pi[1] = (du[0][1] + du[1][0]) * sToPi;
    pi[2] = (du[0][2] + du[2][0]) * sToPi;
    pi[4] = (du[1][2] + du[2][1]) * sToPi;
    const double jv[3] = {rho * uw[0], rho * uw[1], rho * uw[2]};
    const double jSqr = jv[0] * jv[0] + jv[1] * jv[1] + jv[2] * jv[2];
    const double invRho = 1.0 / (rb + 1.0);
    #pragma unroll
    for (int i = 0; i < Q; ++i) {
        const double c_j = cx(i) * jv[0] + cy(i) * jv[1] + cz(i) * jv[2];
        const double feq = tw(i) * (rb + 3. * c_j + invRho * (4.5 * c_j * c_j - 1.5 * jSqr));
        // Pi -> fNeq (src/latticeBoltzmann/offEquilibriumTemplates.h:91-115)
        double fneq = 0.0;
        int iPi = 0;
        for (int al = 0; al < 3; ++al) {
            fneq += pi[iPi] * (ci(i, al) * ci(i, al) - 1.0 / 3.0);
            ++iPi;
            for (int be = al + 1; be < 3; ++be) {
                fneq += pi[iPi] * 2.0 * ci(i, al) * ci(i, be);
                ++iPi;
            }
        }
        fneq *= tw(i) * 3.0 * 3.0 / 2.0;
        fc[i] = (C)(feq + fneq);
    }
    store_post<S>(v, n, c, swapped, fc);
}

template <class S>
__global__ void k_vwall_compute(LatView v, long nwall, bool swapped, const double* un)
{
    vwall_compute_cell<S>(v, nwall, swapped, un, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// The velocity-wall passes of a batch of boxes (each with velocity walls
// and no Guo wall): the same two launches, each over every box's wall
// cells.
// PHASE 0 neighbour velocities (base counts the workgroups over 2 nwall
// cells), 1 the wall cells.
template <class S, int PHASE>
__global__ void k_vwall_batch(const VwallPlan* __restrict__ tab, const int* __restrict__ base, int nbox, bool swapped)
{
    const int box = batch_box(base, nbox, (int)blockIdx.x);
    // the box's plan is read in place (a copy of it, which the wall
    // velocity's indexed read would put in scratch memory, is not made)
    const VwallPlan& w = tab[box];
    const long i = (long)((int)blockIdx.x - base[box]) * blockDim.x + threadIdx.x;
    if (PHASE == 0) vwall_u_cell<S>(w.v, w.nwall, swapped, w.buf, i);
    else vwall_compute_cell<S>(w.v, w.nwall, swapped, w.buf, i);
}

template <class S>
void vwall_batch_s(bool swapped, const VwallPlan* tab, const int* base, const int* base_u, int nbox, unsigned nblk,
                   unsigned nblk_u, hipStream_t s)
{
    hipLaunchKernelGGL((k_vwall_batch<S, 0>), dim3(nblk_u), dim3(256), 0, s, tab, base_u, nbox, swapped);
    hipLaunchKernelGGL((k_vwall_batch<S, 1>), dim3(nblk), dim3(256), 0, s, tab, base, nbox, swapped);
}

int vwall_batch_launch(int storage, bool swapped, const VwallPlan* tab, const int* base, const int* base_u, int nbox,
                       unsigned nblk, unsigned nblk_u, hipStream_t s)
{
    if (storage == BL_STORE_F64) vwall_batch_s<double>(swapped, tab, base, base_u, nbox, nblk, nblk_u, s);
    else if (storage == BL_STORE_F32) vwall_batch_s<float>(swapped, tab, base, base_u, nbox, nblk, nblk_u, s);
    else vwall_batch_s<__half>(swapped, tab, base, base_u, nbox, nblk, nblk_u, s);
    BL_CHECK_LAUNCH();
    return BL_OK;
}

int velocity_walls_launch(bl_lattice_s* L)
{
    if (!L->has_vwall || !L->mask) return guo_walls_launch(L);
    const LatView v = view_of(L);
    const bool swapped = (L->steps & 1) != 0;
    const unsigned nb = (unsigned)((L->nwall + 255) / 256), nbu = (unsigned)((2 * L->nwall + 255) / 256);
    if (L->storage == BL_STORE_F64) {
        hipLaunchKernelGGL(k_vwall_u<double>, dim3(nbu), dim3(256), 0, L->stream, v, L->nwall, swapped, L->d_wall);
        hipLaunchKernelGGL(k_vwall_compute<double>, dim3(nb), dim3(256), 0, L->stream, v, L->nwall, swapped,
                           (const double*)L->d_wall);
    } else if (L->storage == BL_STORE_F32) {
        hipLaunchKernelGGL(k_vwall_u<float>, dim3(nbu), dim3(256), 0, L->stream, v, L->nwall, swapped, L->d_wall);
        hipLaunchKernelGGL(k_vwall_compute<float>, dim3(nb), dim3(256), 0, L->stream, v, L->nwall, swapped,
                           (const double*)L->d_wall);
    } else {
        hipLaunchKernelGGL(k_vwall_u<__half>, dim3(nbu), dim3(256), 0, L->stream, v, L->nwall, swapped, L->d_wall);
        hipLaunchKernelGGL(k_vwall_compute<__half>, dim3(nb), dim3(256), 0, L->stream, v, L->nwall, swapped,
                           (const double*)L->d_wall);
    }
    BL_CHECK_LAUNCH();
    return guo_walls_launch(L);
}

// ----------------------------------------------- transfers and moments
// buf: ncell*19 doubles in cell-major order (host layout).
template <class S>
__global__ void k_put(LatView v, bool swapped, const double* buf)
{
    using C = typename Pop<S>::C;
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= v.ncell) return;
    int x, y, z;
    decode(c, v.ny, v.nz, x, y, z);
    Nbr n;
    n.init(v, x, y, z);
    C f[Q];
    for (int i = 0; i < Q; ++i) f[i] = (C)buf[(long)c * Q + i];
    store_post<S>(v, n, c, swapped, f);
}

template <class S>
__global__ void k_get(LatView v, bool swapped, double* buf)
{
    using C = typename Pop<S>::C;
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= v.ncell) return;
    int x, y, z;
    decode(c, v.ny, v.nz, x, y, z);
    Nbr n;
    n.init(v, x, y, z);
    C f[Q];
    load_post<S>(v, n, c, swapped, f);
    for (int i = 0; i < Q; ++i) buf[(long)c * Q + i] = (double)f[i];
}

// rho, u as computeDensity / computeVelocity report them.
template <class S>
__global__ void k_moments(LatView v, bool swapped, double* rho, double* u)
{
    using C = typename Pop<S>::C;
    using FT = typename Pop<S>::F;
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= v.ncell) return;
    int x, y, z;
    decode(c, v.ny, v.nz, x, y, z);
    Nbr n;
    n.init(v, x, y, z);
    C f[Q];
    load_post<S>(v, n, c, swapped, f);
    C F[3];
    cell_force<FT, C>(v, c, x, y, z, F);
    const unsigned char m = v.mask ? v.mask[c] : (unsigned char)M_FLUID;
    double r, uu[3];
    if (m == M_VWALL) {
        const int p[3] = {x, y, z};
        const int a = v.wall_axis;
        const int side = p[a] == 0 ? 0 : 1;
        const int orient = side == 0 ? -1 : 1;
        const double* uw = v.wall_u + 3 * side;
        double onWall = 0.0, normal = 0.0;
        for (int i = 0; i < Q; ++i)
            if (ci(i, a) == 0) onWall += (double)f[i];
        for (int i = 0; i < Q; ++i)
            if (ci(i, a) == orient) normal += (double)f[i];
        const double velNormal = (double)orient * (uw[a] - 0.5 * (double)F[a]);
        r = (2.0 * normal + onWall - 1.0 * velNormal) / (1.0 + velNormal) + 1.0;
        for (int d = 0; d < 3; ++d) uu[d] = uw[d];
    } else if (m == M_BB || m == M_NONE) {
        // default-constructed BounceBack / NoDynamics: rho = 0, u = 0
        r = 0.0;
        uu[0] = uu[1] = uu[2] = 0.0;
    } else {
        const C rb = sum_rho(f);
        C j[3];
        sum_j(f, j);
        const C invRho = C(1) / (rb + C(1));
        r = (double)(rb + C(1));
        for (int d = 0; d < 3; ++d) uu[d] = (double)(j[d] * invRho + F[d] / C(2));
    }
    if (rho) rho[c] = r;
    if (u) for (int d = 0; d < 3; ++d) u[3 * (long)c + d] = uu[d];
}

template <class S>
__global__ void k_get_force(LatView v, double* out)
{
    using C = typename Pop<S>::C;
    using FT = typename Pop<S>::F;
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= v.ncell) return;
    int x, y, z;
    decode(c, v.ny, v.nz, x, y, z);
    C F[3];
    cell_force<FT, C>(v, c, x, y, z, F);
    for (int d = 0; d < 3; ++d) out[3 * (long)c + d] = (double)F[d];
}

// Equilibrium initialisation (iniCellAtEquilibrium, src/core/cell.hh:124-132).
template <class S>
__global__ void k_init_eq(LatView v, bool swapped, const double* rho, const double* u)
{
    using C = typename Pop<S>::C;
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= v.ncell) return;
    int x, y, z;
    decode(c, v.ny, v.nz, x, y, z);
    const double r = rho ? rho[c] : 1.0;
    double j[3] = {0, 0, 0};
    if (u) for (int d = 0; d < 3; ++d) j[d] = r * u[3 * (long)c + d];
    const double jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
    const double rb = r - 1.0, invRho = 1.0 / r;
    C f[Q];
    for (int i = 0; i < Q; ++i) {
        const double c_j = cx(i) * j[0] + cy(i) * j[1] + cz(i) * j[2];
        f[i] = (C)(tw(i) * (rb + 3. * c_j + invRho * (4.5 * c_j * c_j - 1.5 * jSqr)));
    }
    Nbr n;
    n.init(v, x, y, z);
    store_post<S>(v, n, c, swapped, f);
}

// FP64 sums of mass and momentum over non-bounce-back cells: one partial
// per block (fixed-order tree in LDS), then one block sums the partials in
// order, so the result does not depend on scheduling.
template <class S>
__global__ void __launch_bounds__(256) k_conserved_partial(LatView v, bool swapped, double* partial)
{
    using C = typename Pop<S>::C;
    __shared__ double sh[4][256];
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    if (c < v.ncell && !(v.mask && v.mask[c] == M_BB)) {
        int x, y, z;
        decode(c, v.ny, v.nz, x, y, z);
        Nbr n;
        n.init(v, x, y, z);
        C f[Q];
        load_post<S>(v, n, c, swapped, f);
        double fd[Q];
        for (int i = 0; i < Q; ++i) fd[i] = (double)f[i];
        double j[3];
        sum_j(fd, j);
        acc[0] = 1.0 + sum_rho(fd);
        acc[1] = j[0]; acc[2] = j[1]; acc[3] = j[2];
    }
    for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partial[4 * (long)blockIdx.x + k] = sh[k][0];
}

__global__ void __launch_bounds__(256) k_sum_partials(const double* partial, long n, double* out)
{
    __shared__ double sh[4][256];
    double acc[4] = {0, 0, 0, 0};
    for (long b = threadIdx.x; b < n; b += blockDim.x)
        for (int k = 0; k < 4; ++k) acc[k] += partial[4 * b + k];
    for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) out[k] = sh[k][0];
}

}  // namespace bl

// ---------------------------------------- allocation-free conservation sums
// bl_lattice_conserved_static keeps the per-block partial sums of
// k_conserved_partial in this file-scope array instead of the lattice's own
// buffer: 4 doubles per 256-cell block for lattices of up to kStaticSumCells
// cells, then the 4 results (8 MiB + 32 B; every device the library runs on
// holds its own copy until the process ends).  Host code never passes the
// array's name to a kernel: the device address of each device's copy comes
// from hipGetSymbolAddress (static_scratch_init) and is kept per device.
constexpr long kStaticSumCells = 1L << 26;
constexpr long kStaticSumBlocks = kStaticSumCells / 256;
static __device__ __align__(256) double scratch[4 * kStaticSumBlocks + 4];

namespace bl {

constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` per device (null: not looked up yet)

// Looks up the current device's copy of `scratch` (once per device); the
// caller makes `device` current first.
int static_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(BL_ERR_ARG, "static scratch: device index beyond 63");
    if (g_scratch_dev[device]) return BL_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(BL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = (double*)p;
    return BL_OK;
}

// ------------------------------------------------------------ fills
// Streaming copy (bandwidth reference): grid-stride over 16-byte elements.
__global__ void __launch_bounds__(256) k_copy16(const int4* __restrict__ a, int4* __restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

__global__ void k_fill_u8(unsigned char* p, long n, unsigned char val)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}

template <class FT>
__global__ void k_put_force(FT* F, long stride, unsigned ncell, const double* buf)
{
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    for (int d = 0; d < 3; ++d) F[d * stride + c] = (FT)buf[3 * (long)c + d];
}

int ensure_force_field(bl_lattice_s* L)
{
    if (L->force) return BL_OK;
    L->force_bytes = L->storage == BL_STORE_F64 ? sizeof(double) : sizeof(float);
    BL_HIP(hipMalloc(&L->force, 3 * L->stride * L->force_bytes));
    BL_HIP(hipMemsetAsync(L->force, 0, 3 * L->stride * L->force_bytes, L->stream));
    BL_HIP(hipMalloc(&L->tile_flags, L->ntiles));
    const unsigned nb = (unsigned)((L->ntiles + 255) / 256);
    hipLaunchKernelGGL(k_fill_u8, dim3(nb), dim3(256), 0, L->stream, L->tile_flags, L->ntiles, (unsigned char)0);
    BL_CHECK_LAUNCH();
    L->force_field = true;
    ++L->gen;
    return BL_OK;
}

// Inert cells: bounce-back cells whose 18 neighbours are all bounce-back or
// past a non-periodic face; one thread per word of 64 cells.
__global__ void k_mark_inert(LatView v, unsigned long long* bits)
{
    const unsigned w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (v.ncell + 63u) / 64u) return;
    unsigned long long b = 0;
    for (unsigned l = 0; l < 64u; ++l) {
        const unsigned c = w * 64u + l;
        if (c >= v.ncell) break;
        if (v.mask[c] != M_BB) continue;
        int x, y, z;
        decode(c, v, x, y, z);
        Nbr n;
        n.init(v, x, y, z);
        bool inert = true;
        for (int i = 1; i < Q; ++i) {
            const int d = n.at(v, cx(i), cy(i), cz(i));
            if (d >= 0 && v.mask[d] != M_BB) inert = false;
        }
        if (inert) b |= 1ull << l;
    }
    bits[w] = b;
}

// (Re)builds the inert-cell bits from the mask when skipping is on (none
// without a mask), frees them otherwise.
int mark_inert(bl_lattice_s* L)
{
    if (!L->skip_inert || !L->mask) {
        hipFree(L->inert);
        L->inert = nullptr;
        L->n_inert = 0;
        return BL_OK;
    }
    const long nw = (L->ncell + 63) / 64;
    if (!L->inert) BL_HIP(hipMalloc(&L->inert, sizeof(unsigned long long) * nw));
    LatView v = view_of(L);
    hipLaunchKernelGGL(k_mark_inert, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, L->stream, v, L->inert);
    BL_CHECK_LAUNCH();
    std::vector<unsigned long long> h(nw);
    BL_HIP(hipMemcpyAsync(h.data(), L->inert, sizeof(unsigned long long) * nw, hipMemcpyDeviceToHost, L->stream));
    BL_HIP(hipStreamSynchronize(L->stream));
    long n = 0;
    for (unsigned long long b : h) n += __builtin_popcountll(b);
    L->n_inert = n;
    return BL_OK;
}

}  // namespace bl

// =================================================================== C-ABI
using namespace bl;

namespace {

bool g_inited = false;
int g_device = 0;

int stage_alloc(bl_lattice L)
{
    if (!L->d_stage) BL_HIP(hipMalloc(&L->d_stage, sizeof(double) * Q * L->ncell));
    return BL_OK;
}

#define BL_DISPATCH(L, KERNEL, NB, ...)                                                            \
    do {                                                                                           \
        if ((L)->storage == BL_STORE_F64)                                                          \
            hipLaunchKernelGGL(KERNEL<double>, dim3(NB), dim3(256), 0, (L)->stream, __VA_ARGS__);  \
        else if ((L)->storage == BL_STORE_F32)                                                     \
            hipLaunchKernelGGL(KERNEL<float>, dim3(NB), dim3(256), 0, (L)->stream, __VA_ARGS__);   \
        else                                                                                       \
            hipLaunchKernelGGL(KERNEL<__half>, dim3(NB), dim3(256), 0, (L)->stream, __VA_ARGS__);  \
    } while (0)

}  // namespace

extern "C" {

int bl_init(int device)
{
    int n = 0;
    BL_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(BL_ERR_ARG, "bl_init: no such device");
    BL_HIP(hipSetDevice(device));
    int rc = static_scratch_init(device);
    if (rc) return rc;
    g_device = device;
    g_inited = true;
    return BL_OK;
}

int bl_shutdown(void)
{
    if (g_inited) BL_HIP(hipDeviceSynchronize());
    g_inited = false;
    return BL_OK;
}

int bl_device_info(char* name, int cap, int* cus, double* hbm)
{
    hipDeviceProp_t p;
    BL_HIP(hipGetDeviceProperties(&p, g_device));
    if (name && cap > 0) {
        std::strncpy(name, p.gcnArchName, cap - 1);
        name[cap - 1] = 0;
    }
    if (cus) *cus = p.multiProcessorCount;
    if (hbm) *hbm = (double)p.totalGlobalMem;
    return BL_OK;
}

int bl_lattice_create(int nx, int ny, int nz, const int periodic[3], int storage, bl_lattice* out)
{
    if (!out) return fail(BL_ERR_ARG, "bl_lattice_create: null output handle");
    if (nx < 3 || ny < 3 || nz < 3) return fail(BL_ERR_ARG, "bl_lattice_create: each extent must be >= 3");
    const long ncell = (long)nx * ny * nz;
    if (ncell >= (1L << 31)) return fail(BL_ERR_ARG, "bl_lattice_create: more than 2^31 cells");
    if (storage < BL_STORE_F64 || storage > BL_STORE_F16S) return fail(BL_ERR_ARG, "bl_lattice_create: unknown storage");
    if (!g_inited) {
        int rc = bl_init(0);
        if (rc) return rc;
    }
    int dev = 0;
    BL_HIP(hipGetDevice(&dev));
    bl_lattice L = new bl_lattice_s();
    L->device = dev;
    L->nx = nx; L->ny = ny; L->nz = nz;
    L->ncell = ncell;
    L->stride = (ncell + 127) / 128 * 128;
    for (int d = 0; d < 3; ++d) L->periodic[d] = periodic ? (periodic[d] != 0) : 1;
    L->storage = storage;
    L->pop_bytes = storage == BL_STORE_F64 ? 8 : (storage == BL_STORE_F32 ? 4 : 2);
    L->ntx = (nx + kTile - 1) / kTile;
    L->nty = (ny + kTile - 1) / kTile;
    L->ntz = (nz + kTile - 1) / kTile;
    L->ntiles = (long)L->ntx * L->nty * L->ntz;
    hipError_t e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&L->pops, Q * L->stride * L->pop_bytes);
    if (e == hipSuccess) e = hipMalloc(&L->d_partials, sizeof(double) * 4 * ((ncell + 255) / 256 + 1));
    if (e != hipSuccess) {
        delete L;
        return fail(BL_ERR_HIP, std::string("bl_lattice_create: ") + hipGetErrorString(e));
    }
    tune_cache_apply(L);
    *out = L;
    return bl_lattice_init_equilibrium(L, nullptr, nullptr);
}

int bl_lattice_destroy(bl_lattice L)
{
    if (!L) return BL_OK;
    if (L->stream) hipStreamSynchronize(L->stream);
    if (L->ibm) bl::ibm_detach(L->ibm);  // the boundary keeps its own buffers; it cannot step any more
    bl::guo_walls_detach(L);
    hipFree(L->pops);
    hipFree(L->mask);
    hipFree(L->inert);
    hipFree(L->own_phase);
    hipFree(L->force);
    hipFree(L->tile_flags);
    hipFree(L->d_stage);
    hipFree(L->d_partials);
    hipFree(L->d_wall);
    if (L->stream) hipStreamDestroy(L->stream);
    delete L;
    return BL_OK;
}

int bl_lattice_set_dynamics(bl_lattice L, double omega_out, double omega_in)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    if (!(omega_out > 0.0 && omega_out < 2.0 && omega_in > 0.0 && omega_in < 2.0))
        return fail(BL_ERR_ARG, "bl_lattice_set_dynamics: omega must lie in (0, 2)");
    L->omega_out = omega_out;
    L->omega_in = omega_in;
    // omega(phase) = w0 w1 / (c w0 + (1-c) w1) at c = 0 and c = 1, evaluated
    // exactly as the reference does (src/fsi/VariableViscosityDynamics.hh:131-141).
    const double c0 = 0.0, c1 = 1.0;
    L->omega_c0 = omega_out * omega_in / (c0 * omega_out + (1. - c0) * omega_in);
    L->omega_c1 = omega_out * omega_in / (c1 * omega_out + (1. - c1) * omega_in);
    ++L->gen;
    return BL_OK;
}

static int set_mask_cells(bl_lattice L, const unsigned char* mask);

int bl_lattice_set_mask(bl_lattice L, const unsigned char* mask)
{
    int rc = set_mask_cells(L, mask);
    return rc ? rc : bl::mark_inert(L);
}

int bl_lattice_set_skip_inert(bl_lattice L, int on)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    if (on != 0 && on != 1) return fail(BL_ERR_ARG, "bl_lattice_set_skip_inert: on must be 0 or 1");
    BL_HIP(hipStreamSynchronize(L->stream));
    L->skip_inert = on != 0;
    ++L->gen;
    return bl::mark_inert(L);
}

int bl_lattice_inert_cells(bl_lattice L, long long* n)
{
    if (!L || !n) return fail(BL_ERR_ARG, "bl_lattice_inert_cells: bad argument");
    *n = L->skip_inert ? (long long)L->n_inert : 0;
    return BL_OK;
}

static int set_mask_cells(bl_lattice L, const unsigned char* mask)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    BL_HIP(hipStreamSynchronize(L->stream));
    ++L->gen;
    if (!mask) {
        hipFree(L->mask);
        L->mask = nullptr;
        L->has_vwall = false;
        return BL_OK;
    }
    bool vw = false;
    for (long c = 0; c < L->ncell; ++c) {
        if (mask[c] > M_VWALL) return fail(BL_ERR_ARG, "bl_lattice_set_mask: unknown mask code");
        vw = vw || mask[c] == M_VWALL;
    }
    if (!L->mask) BL_HIP(hipMalloc(&L->mask, L->ncell));
    BL_HIP(hipMemcpy(L->mask, mask, L->ncell, hipMemcpyHostToDevice));
    L->has_vwall = vw;
    if (vw) {
        // velocity-wall cells must lie on the two planes of the wall axis
        const int n[3] = {L->nx, L->ny, L->nz};
        const int a = L->wall_axis;
        for (int x = 0; x < L->nx; ++x)
            for (int y = 0; y < L->ny; ++y)
                for (int z = 0; z < L->nz; ++z) {
                    const int p[3] = {x, y, z};
                    if (mask[((long)x * L->ny + y) * L->nz + z] == M_VWALL && p[a] != 0 && p[a] != n[a] - 1)
                        return fail(BL_ERR_ARG, "bl_lattice_set_mask: velocity-wall cells must lie on the planes "
                                                "0 and n-1 of the wall axis (set it first)");
                }
        const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
        const long nwall = 2L * n[t1] * n[t2];
        if (nwall > L->nwall) {
            hipFree(L->d_wall);
            // three neighbour velocities per wall cell and per interior
            // cell next to it
            BL_HIP(hipMalloc(&L->d_wall, sizeof(double) * 6 * nwall));
        }
        L->nwall = nwall;
    }
    return BL_OK;
}

int bl_lattice_set_velocity_walls(bl_lattice L, int axis, const double u_lo[3], const double u_hi[3])
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    if (axis < 0 || axis > 2) return fail(BL_ERR_ARG, "bl_lattice_set_velocity_walls: axis must be 0, 1 or 2");
    if (L->periodic[axis]) return fail(BL_ERR_ARG, "bl_lattice_set_velocity_walls: wall axis must be non-periodic");
    L->wall_axis = axis;
    for (int d = 0; d < 3; ++d) {
        L->wall_u[d] = u_lo ? u_lo[d] : 0.0;
        L->wall_u[3 + d] = u_hi ? u_hi[d] : 0.0;
    }
    ++L->gen;
    return BL_OK;
}

int bl_lattice_set_body_force(bl_lattice L, const double g[3])
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    for (int d = 0; d < 3; ++d) L->g[d] = g ? g[d] : 0.0;
    if (!L->collided)
        for (int d = 0; d < 3; ++d) L->g_last[d] = L->g[d];
    ++L->gen;
    return BL_OK;
}

int bl_lattice_set_force_field(bl_lattice L, const double* force)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    if (!force) return fail(BL_ERR_ARG, "bl_lattice_set_force_field: null array");
    int rc = ensure_force_field(L);
    if (rc) return rc;
    double* d = nullptr;
    BL_HIP(hipMalloc(&d, sizeof(double) * 3 * L->ncell));
    BL_HIP(hipMemcpy(d, force, sizeof(double) * 3 * L->ncell, hipMemcpyHostToDevice));
    const unsigned nb = (unsigned)((L->ncell + 255) / 256);
    if (L->storage == BL_STORE_F64)
        hipLaunchKernelGGL(k_put_force<double>, dim3(nb), dim3(256), 0, L->stream, (double*)L->force, L->stride, (unsigned)L->ncell, (const double*)d);
    else
        hipLaunchKernelGGL(k_put_force<float>, dim3(nb), dim3(256), 0, L->stream, (float*)L->force, L->stride, (unsigned)L->ncell, (const double*)d);
    const unsigned nt = (unsigned)((L->ntiles + 255) / 256);
    hipLaunchKernelGGL(k_fill_u8, dim3(nt), dim3(256), 0, L->stream, L->tile_flags, L->ntiles, (unsigned char)1);
    BL_CHECK_LAUNCH();
    L->flags_full = true;
    ++L->gen;
    BL_HIP(hipStreamSynchronize(L->stream));
    hipFree(d);
    return BL_OK;
}

int bl_lattice_set_phase(bl_lattice L, const unsigned char* phase)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    BL_HIP(hipStreamSynchronize(L->stream));
    ++L->gen;
    if (!phase) {
        hipFree(L->own_phase);
        L->own_phase = nullptr;
        L->phase = nullptr;
        return BL_OK;
    }
    for (long c = 0; c < L->ncell; ++c)
        if (phase[c] > 1) return fail(BL_ERR_ARG, "bl_lattice_set_phase: phase must be 0 or 1");
    if (!L->own_phase) BL_HIP(hipMalloc(&L->own_phase, L->ncell));
    BL_HIP(hipMemcpy(L->own_phase, phase, L->ncell, hipMemcpyHostToDevice));
    L->phase = L->own_phase;
    return BL_OK;
}

int bl_lattice_init_equilibrium(bl_lattice L, const double* rho, const double* u)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    double *dr = nullptr, *du = nullptr;
    if (rho) {
        BL_HIP(hipMalloc(&dr, sizeof(double) * L->ncell));
        BL_HIP(hipMemcpy(dr, rho, sizeof(double) * L->ncell, hipMemcpyHostToDevice));
    }
    if (u) {
        BL_HIP(hipMalloc(&du, sizeof(double) * 3 * L->ncell));
        BL_HIP(hipMemcpy(du, u, sizeof(double) * 3 * L->ncell, hipMemcpyHostToDevice));
    }
    const LatView v = view_of(L);
    const bool swapped = (L->steps & 1) != 0;
    const unsigned nb = (unsigned)((L->ncell + 255) / 256);
    BL_DISPATCH(L, k_init_eq, nb, v, swapped, (const double*)dr, (const double*)du);
    BL_CHECK_LAUNCH();
    BL_HIP(hipStreamSynchronize(L->stream));
    hipFree(dr);
    hipFree(du);
    return BL_OK;
}

int bl_lattice_put_populations(bl_lattice L, const double* buf)
{
    if (!L || !buf) return fail(BL_ERR_ARG, "bl_lattice_put_populations: null argument");
    int rc = stage_alloc(L);
    if (rc) return rc;
    BL_HIP(hipMemcpyAsync(L->d_stage, buf, sizeof(double) * Q * L->ncell, hipMemcpyHostToDevice, L->stream));
    const LatView v = view_of(L);
    const unsigned nb = (unsigned)((L->ncell + 255) / 256);
    BL_DISPATCH(L, k_put, nb, v, (L->steps & 1) != 0, (const double*)L->d_stage);
    BL_CHECK_LAUNCH();
    BL_HIP(hipStreamSynchronize(L->stream));
    return BL_OK;
}

int bl_lattice_get_populations(bl_lattice L, double* buf)
{
    if (!L || !buf) return fail(BL_ERR_ARG, "bl_lattice_get_populations: null argument");
    int rc = stage_alloc(L);
    if (rc) return rc;
    const LatView v = view_of(L);
    const unsigned nb = (unsigned)((L->ncell + 255) / 256);
    BL_DISPATCH(L, k_get, nb, v, (L->steps & 1) != 0, L->d_stage);
    BL_CHECK_LAUNCH();
    BL_HIP(hipMemcpyAsync(buf, L->d_stage, sizeof(double) * Q * L->ncell, hipMemcpyDeviceToHost, L->stream));
    BL_HIP(hipStreamSynchronize(L->stream));
    return BL_OK;
}

int bl_lattice_initialize(bl_lattice L)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    int rc = velocity_walls_launch(L);
    if (rc) return rc;
    BL_HIP(hipStreamSynchronize(L->stream));
    return BL_OK;
}

int bl_collide_stream(bl_lattice L, int n_steps)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    if (n_steps < 0) return fail(BL_ERR_ARG, "bl_collide_stream: negative step count");
    for (int s = 0; s < n_steps; ++s) {
        int rc = collide_stream_launch(L);
        if (rc) return rc;
        rc = velocity_walls_launch(L);
        if (rc) return rc;
    }
    return BL_OK;
}

int bl_lattice_get_moments(bl_lattice L, double* rho, double* u)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    double *dr = nullptr, *du = nullptr;
    BL_HIP(hipMalloc(&dr, sizeof(double) * L->ncell));
    BL_HIP(hipMalloc(&du, sizeof(double) * 3 * L->ncell));
    LatView v = view_of(L);
    if (!L->force_field) v.flags = nullptr;
    const unsigned nb = (unsigned)((L->ncell + 255) / 256);
    BL_DISPATCH(L, k_moments, nb, v, (L->steps & 1) != 0, dr, du);
    BL_CHECK_LAUNCH();
    if (rho) BL_HIP(hipMemcpyAsync(rho, dr, sizeof(double) * L->ncell, hipMemcpyDeviceToHost, L->stream));
    if (u) BL_HIP(hipMemcpyAsync(u, du, sizeof(double) * 3 * L->ncell, hipMemcpyDeviceToHost, L->stream));
    BL_HIP(hipStreamSynchronize(L->stream));
    hipFree(dr);
    hipFree(du);
    return BL_OK;
}

int bl_lattice_get_force(bl_lattice L, double* force)
{
    if (!L || !force) return fail(BL_ERR_ARG, "bl_lattice_get_force: null argument");
    double* d = nullptr;
    BL_HIP(hipMalloc(&d, sizeof(double) * 3 * L->ncell));
    LatView v = view_of(L);
    if (!L->force_field) v.flags = nullptr;
    const unsigned nb = (unsigned)((L->ncell + 255) / 256);
    BL_DISPATCH(L, k_get_force, nb, v, d);
    BL_CHECK_LAUNCH();
    BL_HIP(hipMemcpyAsync(force, d, sizeof(double) * 3 * L->ncell, hipMemcpyDeviceToHost, L->stream));
    BL_HIP(hipStreamSynchronize(L->stream));
    hipFree(d);
    return BL_OK;
}

int bl_lattice_conserved(bl_lattice L, double out[4])
{
    if (!L || !out) return fail(BL_ERR_ARG, "bl_lattice_conserved: null argument");
    const LatView v = view_of(L);
    const long nb = (L->ncell + 255) / 256;
    if (L->storage == BL_STORE_F64)
        hipLaunchKernelGGL(k_conserved_partial<double>, dim3(nb), dim3(256), 0, L->stream, v, (L->steps & 1) != 0, L->d_partials);
    else if (L->storage == BL_STORE_F32)
        hipLaunchKernelGGL(k_conserved_partial<float>, dim3(nb), dim3(256), 0, L->stream, v, (L->steps & 1) != 0, L->d_partials);
    else
        hipLaunchKernelGGL(k_conserved_partial<__half>, dim3(nb), dim3(256), 0, L->stream, v, (L->steps & 1) != 0, L->d_partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, L->stream, (const double*)L->d_partials, nb,
                       L->d_partials + 4 * nb);
    BL_CHECK_LAUNCH();
    BL_HIP(hipMemcpyAsync(out, L->d_partials + 4 * nb, sizeof(double) * 4, hipMemcpyDeviceToHost, L->stream));
    BL_HIP(hipStreamSynchronize(L->stream));
    return BL_OK;
}

int bl_lattice_conserved_static(bl_lattice L, double* out)
{
    if (!L || !out) return fail(BL_ERR_ARG, "bl_lattice_conserved_static: null argument");
    if (L->ncell > kStaticSumCells)
        return fail(BL_ERR_CAPACITY, "bl_lattice_conserved_static: at most 67108864 cells (the static scratch "
                                     "holds 4 partial sums per 256 cells)");
    if (L->device < 0 || L->device >= kMaxDevices || !g_scratch_dev[L->device]) {
        // first call on a device that bl_init did not set up: look its copy
        // up with that device current
        int cur = -1;
        BL_HIP(hipGetDevice(&cur));
        if (cur != L->device)
            return fail(BL_ERR_STATE, "bl_lattice_conserved_static: make the lattice's device current (bl_init) first");
        int rc = static_scratch_init(L->device);
        if (rc) return rc;
    }
    double* scr = g_scratch_dev[L->device];
    const LatView v = view_of(L);
    const long nb = (L->ncell + 255) / 256;
    const bool swapped = (L->steps & 1) != 0;
    if (L->storage == BL_STORE_F64)
        hipLaunchKernelGGL(k_conserved_partial<double>, dim3(nb), dim3(256), 0, L->stream, v, swapped, scr);
    else if (L->storage == BL_STORE_F32)
        hipLaunchKernelGGL(k_conserved_partial<float>, dim3(nb), dim3(256), 0, L->stream, v, swapped, scr);
    else
        hipLaunchKernelGGL(k_conserved_partial<__half>, dim3(nb), dim3(256), 0, L->stream, v, swapped, scr);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, L->stream, (const double*)scr, nb, out);
    BL_CHECK_LAUNCH();
    return BL_OK;
}

int bl_lattice_steps(bl_lattice L, long long* steps)
{
    if (!L || !steps) return fail(BL_ERR_ARG, "bl_lattice_steps: null argument");
    *steps = L->steps;
    return BL_OK;
}

int bl_lattice_set_launch(bl_lattice L, int block, int xcd_remap)
{
    int rc = bl_lattice_set_launch_parity(L, 0, block, xcd_remap ? 1 : 0, 0);
    return rc ? rc : bl_lattice_set_launch_parity(L, 1, block, xcd_remap ? 1 : 0, 0);
}

int bl_lattice_set_launch_parity(bl_lattice L, int parity, int block, int xcd_remap, int nontemporal)
{
    if (!L) return fail(BL_ERR_ARG, "null lattice");
    if (parity != 0 && parity != 1) return fail(BL_ERR_ARG, "bl_lattice_set_launch_parity: parity must be 0 or 1");
    // the collide-stream kernels are compiled for at most 256 threads per
    // block; a negative block asks for the odd step's z-segment kernel
    const int lanes = block < 0 ? -block : block;
    if (lanes < 64 || lanes > 256 || lanes % 64)
        return fail(BL_ERR_ARG, "bl_lattice_set_launch: block must be +-64, +-128, +-192 or +-256");
    if (block < 0 && xcd_remap == BL_ORDER_ROWS)
        return fail(BL_ERR_ARG, "bl_lattice_set_launch: z-segment blocks take xcd_remap 0, 1 or a power of two");
    if (xcd_remap < 0 || xcd_remap > (1 << 20) ||
        (xcd_remap > 1 && xcd_remap != BL_ORDER_ROWS && (xcd_remap & (xcd_remap - 1))))
        return fail(BL_ERR_ARG,
                    "bl_lattice_set_launch: xcd_remap must be 0, 1, BL_ORDER_ROWS (3) or a power of two up to 2^20");
    if (nontemporal < 0 || nontemporal > 2)
        return fail(BL_ERR_ARG, "bl_lattice_set_launch_parity: nontemporal must be 0, 1 or 2");
    L->block[parity] = block;
    L->xcd_remap[parity] = xcd_remap;
    L->nontemporal[parity] = parity == 0 && nontemporal ? 1 : nontemporal;
    ++L->gen;
    return BL_OK;
}

int bl_lattice_get_launch(bl_lattice L, int parity, int out[3])
{
    if (!L || !out || (parity != 0 && parity != 1)) return fail(BL_ERR_ARG, "bl_lattice_get_launch: bad argument");
    out[0] = L->block[parity];
    out[1] = L->xcd_remap[parity];
    out[2] = L->nontemporal[parity];
    return BL_OK;
}

int bl_lattice_stream(bl_lattice L, void** stream)
{
    if (!L || !stream) return fail(BL_ERR_ARG, "bl_lattice_stream: null argument");
    *stream = (void*)L->stream;
    return BL_OK;
}

int bl_lattice_time_collide(bl_lattice L, int n_steps, double* ms)
{
    if (!L || !ms || n_steps <= 0) return fail(BL_ERR_ARG, "bl_lattice_time_collide: bad argument");
    hipEvent_t a, b;
    BL_HIP(hipEventCreate(&a));
    BL_HIP(hipEventCreate(&b));
    BL_HIP(hipEventRecord(a, L->stream));
    for (int s = 0; s < n_steps; ++s) {
        int rc = collide_stream_launch(L);
        if (rc) return rc;
    }
    BL_HIP(hipEventRecord(b, L->stream));
    BL_HIP(hipEventSynchronize(b));
    float t = 0.f;
    BL_HIP(hipEventElapsedTime(&t, a, b));
    *ms = (double)t;
    hipEventDestroy(a);
    hipEventDestroy(b);
    return BL_OK;
}

int bl_sync(void)
{
    BL_HIP(hipDeviceSynchronize());
    return BL_OK;
}

int bl_device_copy(long long n_bytes, int n_iter, double* ms)
{
    if (n_bytes <= 0 || n_iter <= 0 || !ms) return fail(BL_ERR_ARG, "bl_device_copy: bad argument");
    const long n = (long)((n_bytes + 15) / 16);
    int4 *a = nullptr, *b = nullptr;
    BL_HIP(hipMalloc(&a, n * sizeof(int4)));
    BL_HIP(hipMalloc(&b, n * sizeof(int4)));
    BL_HIP(hipMemset(a, 1, n * sizeof(int4)));
    hipEvent_t e0, e1;
    BL_HIP(hipEventCreate(&e0));
    BL_HIP(hipEventCreate(&e1));
    const unsigned nb = (unsigned)std::min<long>((n + 255) / 256, 1L << 20);
    hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, nullptr, (const int4*)a, b, n);
    BL_HIP(hipEventRecord(e0, nullptr));
    for (int k = 0; k < n_iter; ++k)
        hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, nullptr, (const int4*)a, b, n);
    BL_HIP(hipEventRecord(e1, nullptr));
    BL_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    BL_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = (double)t;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return BL_OK;
}

}  // extern "C"
