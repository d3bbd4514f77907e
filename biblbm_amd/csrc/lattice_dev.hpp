// Device-side helpers shared by the lattice and immersed-boundary kernels:
// storage conversion, AA-pattern slot mapping, D3Q19 moments and the
// Guo-VOF BGK collision.
#pragma once

#include "lattice.hpp"

namespace bl {

// ------------------------------------------------------------ storage I/O
// Nontemporal loads and stores (streaming hint: the line is not expected to
// be reused); __half goes through its 16-bit pattern.
template <class T>
__device__ __forceinline__ T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }
template <class T>
__device__ __forceinline__ void st_nt(T* p, T v) { __builtin_nontemporal_store(v, p); }
template <>
__device__ __forceinline__ __half ld_nt(const __half* p)
{
    return __ushort_as_half(__builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p)));
}
template <>
__device__ __forceinline__ void st_nt(__half* p, __half v)
{
    __builtin_nontemporal_store(__half_as_ushort(v), reinterpret_cast<unsigned short*>(p));
}

template <class S>
struct Pop;
// Each storage also loads and stores its populations in the collision's
// working form (ldw / stw): fBar_i itself for FP64 and FP32, and for the
// FP16-shifted storage the stored value g_i = fBar_i / t_i * 2^10 as a
// float (kRaw): the weights t_i and the scale then fold into the moments
// and the collision's constants (wmoments, wcollide) instead of costing
// two multiplies per population.
template <>
struct Pop<double> {
    using C = double;
    using F = double;
    static constexpr bool kRaw = false;
    template <class K, bool NT = false>
    __device__ static C ld(const double* A, K k, int) { return NT ? ld_nt(A + k) : A[k]; }
    __device__ static double enc(C v, int) { return v; }
    __device__ static C dec(double s, int) { return s; }
    template <class K, bool NT = false>
    __device__ static void st(double* A, K k, int, C v)
    {
        if (NT) st_nt(A + k, v);
        else A[k] = v;
    }
    template <class K, bool NT = false>
    __device__ static C ldw(const double* A, K k, int i) { return ld<K, NT>(A, k, i); }
    template <class K, bool NT = false>
    __device__ static void stw(double* A, K k, int i, C v) { st<K, NT>(A, k, i, v); }
};
template <>
struct Pop<float> {
    using C = float;
    using F = float;
    static constexpr bool kRaw = false;
    template <class K, bool NT = false>
    __device__ static C ld(const float* A, K k, int) { return NT ? ld_nt(A + k) : A[k]; }
    __device__ static float enc(C v, int) { return v; }
    __device__ static C dec(float s, int) { return s; }
    template <class K, bool NT = false>
    __device__ static void st(float* A, K k, int, C v)
    {
        if (NT) st_nt(A + k, v);
        else A[k] = v;
    }
    template <class K, bool NT = false>
    __device__ static C ldw(const float* A, K k, int i) { return ld<K, NT>(A, k, i); }
    template <class K, bool NT = false>
    __device__ static void stw(float* A, K k, int i, C v) { st<K, NT>(A, k, i, v); }
};
template <>
struct Pop<__half> {
    using C = float;
    using F = float;
    static constexpr bool kRaw = true;
    template <class K, bool NT = false>
    __device__ static C ld(const __half* A, K k, int i)
    {
        return __half2float(NT ? ld_nt(A + k) : A[k]) * (float)(tw(i) / kF16Scale);
    }
    __device__ static __half enc(C v, int i) { return __float2half_rn(v * (float)(kF16Scale / tw(i))); }
    __device__ static C dec(__half s, int i) { return __half2float(s) * (float)(tw(i) / kF16Scale); }
    template <class K, bool NT = false>
    __device__ static void st(__half* A, K k, int i, C v)
    {
        const __half h = enc(v, i);
        if (NT) st_nt(A + k, h);
        else A[k] = h;
    }
    template <class K, bool NT = false>
    __device__ static C ldw(const __half* A, K k, int)
    {
        return __half2float(NT ? ld_nt(A + k) : A[k]);
    }
    template <class K, bool NT = false>
    __device__ static void stw(__half* A, K k, int, C v)
    {
        const __half h = __float2half_rn(v);
        if (NT) st_nt(A + k, h);
        else A[k] = h;
    }
};

// --------------------------------------------------------- index helpers
__device__ __forceinline__ int wrapc(int v, int n, int per)
{
    if (v < 0) return per ? v + n : -1;
    if (v >= n) return per ? v - n : -1;
    return v;
}

// Workgroups are dealt round-robin over the 8 XCDs; remap so that each XCD
// walks one contiguous range of cells and neighbouring blocks share its L2.
__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nb)
{
    const unsigned per = nb >> 3, rem = nb & 7u;
    const unsigned xcd = bid & 7u, k = bid >> 3;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

// Between the two: chunks of 2^lc consecutive blocks dealt round-robin over
// the XCDs, so each XCD's L2 sees runs of 2^lc blocks (neighbouring blocks
// share partial cache lines there) while all XCDs stream from nearby
// addresses.  Blocks past the last whole round keep their own index.
__device__ __forceinline__ unsigned xcd_chunk_block(unsigned bid, unsigned nb, unsigned lc)
{
    const unsigned span = 8u << lc;
    const unsigned full = nb / span * span;
    if (bid >= full) return bid;
    const unsigned x = bid & 7u, k = bid >> 3;
    return ((((k >> lc) << 3) + x) << lc) + (k & ((1u << lc) - 1u));
}

// The block order of the collide launch: mode 0 the hardware's, 1 one
// contiguous range per XCD, 2^k (k >= 1) chunks of 2^k blocks.
__device__ __forceinline__ unsigned collide_block(unsigned bid, unsigned nb, int mode)
{
    if (mode == 0) return bid;
    if (mode == 1) return xcd_block(bid, nb);
    return xcd_chunk_block(bid, nb, (unsigned)(31 - __builtin_clz((unsigned)mode)));
}

// Batched launches over boxes: the box of workgroup b, the k with
// base[k] <= b < base[k + 1] (base ascending, nbox boxes; uniform over the
// workgroup, so the search runs on scalar registers).
__device__ __forceinline__ int batch_box(const int* __restrict__ base, int nbox, int b)
{
    int lo = 0, hi = nbox;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (base[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct Nbr {
    int X[3], Y[3], Z[3];
    __device__ __forceinline__ void init(const LatView& v, int x, int y, int z)
    {
        X[0] = wrapc(x - 1, v.nx, v.per[0]); X[1] = x; X[2] = wrapc(x + 1, v.nx, v.per[0]);
        Y[0] = wrapc(y - 1, v.ny, v.per[1]); Y[1] = y; Y[2] = wrapc(y + 1, v.ny, v.per[1]);
        Z[0] = wrapc(z - 1, v.nz, v.per[2]); Z[1] = z; Z[2] = wrapc(z + 1, v.nz, v.per[2]);
    }
    // cell index of x + (dx,dy,dz), or -1 outside a non-periodic face
    __device__ __forceinline__ int at(const LatView& v, int dx, int dy, int dz) const
    {
        const int a = X[dx + 1], b = Y[dy + 1], c = Z[dz + 1];
        if ((a | b | c) < 0) return -1;
        return (a * v.ny + b) * v.nz + c;
    }
};

// Neighbour addressing of the AA odd step without branches: per axis the
// element offset of the -1 / +1 neighbour (wrapped across a periodic face)
// and whether that neighbour exists (not past a non-periodic face).
struct NbrOff {
    int o[3][3];     // [axis][d + 1]
    unsigned ok[3];  // bit d + 1: the neighbour at d exists
    __device__ __forceinline__ void axis(int a, int p, int n, int per, int s)
    {
        const bool lo = p > 0, hi = p < n - 1;
        o[a][0] = lo ? -s : (n - 1) * s;
        o[a][1] = 0;
        o[a][2] = hi ? s : -(n - 1) * s;
        ok[a] = 2u | ((lo || per) ? 1u : 0u) | ((hi || per) ? 4u : 0u);
    }
    __device__ __forceinline__ void init(const LatView& v, int x, int y, int z)
    {
        axis(0, x, v.nx, v.per[0], v.ny * v.nz);
        axis(1, y, v.ny, v.per[1], v.nz);
        axis(2, z, v.nz, v.per[2], 1);
    }
    // element offset of the cell x + (dx, dy, dz); `valid` false past a face
    __device__ __forceinline__ int off(int dx, int dy, int dz, bool& valid) const
    {
        valid = ((ok[0] >> (dx + 1)) & (ok[1] >> (dy + 1)) & (ok[2] >> (dz + 1)) & 1u) != 0;
        return o[0][dx + 1] + o[1][dy + 1] + o[2][dz + 1];
    }
    // the same where x and y are at least one cell away from their faces in
    // every lane of the wave: their part of the offset, dx*sx + dy*sy, is
    // then the same in every lane (sx, sy wave-uniform)
    __device__ __forceinline__ int off_inner(int dx, int dy, int dz, int sx, int sy, bool& valid) const
    {
        valid = ((ok[2] >> (dz + 1)) & 1u) != 0;
        return dx * sx + dy * sy + o[2][dz + 1];
    }
};

__device__ __forceinline__ unsigned fdiv(unsigned n, FastDiv f)
{
    if (f.l == 0) return n;
    const unsigned t = __umulhi(f.m, n);
    return (t + ((n - t) >> 1)) >> (f.l - 1);
}

// BL_ORDER_ROWS: hardware block b runs on XCD b % 8; its k = b / 8-th block
// there covers segment k % s of z-row 8 (k / s) + b % 8, s = row_segs blocks
// per row.  Every cache line that a row's misaligned z +-1 accesses share
// (with its own wrap-around) then stays in one XCD's L2, while the eight
// XCDs work on adjacent rows.  False for a thread past the last row or
// past the row's end.
__device__ __forceinline__ bool row_cell(const LatView& v, unsigned& c, int& x, int& y, int& z)
{
    const unsigned b = blockIdx.x, k = b >> 3;
    const unsigned s = (unsigned)v.row_segs;
    const unsigned kr = k / s;
    const unsigned row = (kr << 3) + (b & 7u);
    const unsigned zz = (k - kr * s) * blockDim.x + threadIdx.x;
    if (row >= (unsigned)v.nx * (unsigned)v.ny || zz >= (unsigned)v.nz) return false;
    const unsigned q = fdiv(row, v.dy);
    x = (int)q;
    y = (int)(row - q * (unsigned)v.ny);
    z = (int)zz;
    c = row * (unsigned)v.nz + zz;
    return true;
}

__device__ __forceinline__ void decode(unsigned c, const LatView& v, int& x, int& y, int& z)
{
    const unsigned r = fdiv(c, v.dz);
    z = (int)(c - r * (unsigned)v.nz);
    const unsigned q = fdiv(r, v.dy);
    y = (int)(r - q * (unsigned)v.ny);
    x = (int)q;
}

__device__ __forceinline__ void decode(unsigned c, int ny, int nz, int& x, int& y, int& z)
{
    z = (int)(c % (unsigned)nz);
    const unsigned r = c / (unsigned)nz;
    y = (int)(r % (unsigned)ny);
    x = (int)(r / (unsigned)ny);
}

// Post-stream population f_i(x) after `swapped` (odd number of completed
// steps) or natural layout: returns (plane, cell).
__device__ __forceinline__ long post_slot(const LatView& v, const Nbr& n, long c, int i, bool swapped,
                                         int& plane)
{
    if (!swapped) { plane = i; return c; }
    const long s = n.at(v, -cx(i), -cy(i), -cz(i));
    if (s < 0) { plane = i; return c; }
    plane = opp(i);
    return s;
}

template <class S>
__device__ __forceinline__ void load_post(const LatView& v, const Nbr& n, long c, bool swapped,
                                          typename Pop<S>::C f[Q])
{
    const S* A = (const S*)v.pops;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        int pl;
        const long k = post_slot(v, n, c, i, swapped, pl);
        f[i] = Pop<S>::ld(A + pl * v.stride, (unsigned)k, i);
    }
}

// The same in the collision's working form (Pop<S>::ldw).
template <class S>
__device__ __forceinline__ void load_post_w(const LatView& v, const Nbr& n, long c, bool swapped,
                                            typename Pop<S>::C f[Q])
{
    const S* A = (const S*)v.pops;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        int pl;
        const long k = post_slot(v, n, c, i, swapped, pl);
        f[i] = Pop<S>::template ldw<unsigned>(A + pl * v.stride, (unsigned)k, i);
    }
}

template <class S>
__device__ __forceinline__ void store_post(const LatView& v, const Nbr& n, long c, bool swapped,
                                           const typename Pop<S>::C f[Q])
{
    S* A = (S*)v.pops;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        int pl;
        const long k = post_slot(v, n, c, i, swapped, pl);
        Pop<S>::st(A + pl * v.stride, (unsigned)k, i, f[i]);
    }
}

// External force of cell c in the field (flags, force): the field's value
// in a flagged tile, the body force gb elsewhere; cell_force_fl with the
// tile's flag already loaded.
template <class FT, class C>
__device__ __forceinline__ void cell_force_fl(const LatView& v, const double* gb, bool flagged, const void* force,
                                              long c, C F[3])
{
    F[0] = (C)gb[0]; F[1] = (C)gb[1]; F[2] = (C)gb[2];
    if (flagged) {
        const FT* Fv = (const FT*)force;
        F[0] = (C)Fv[c];
        F[1] = (C)Fv[v.stride + c];
        F[2] = (C)Fv[2 * v.stride + c];
    }
}

template <class FT, class C>
__device__ __forceinline__ void cell_force_in(const LatView& v, const double* gb, const unsigned char* flags,
                                              const void* force, long c, int x, int y, int z, C F[3])
{
    bool fl = false;
    if (flags) {
        const long t = ((long)(x >> 2) * v.nty + (y >> 2)) * v.ntz + (z >> 2);
        fl = flags[t] != 0;
    }
    cell_force_fl<FT, C>(v, gb, fl, force, c, F);
}

// External force of cell c.
template <class FT, class C>
__device__ __forceinline__ void cell_force(const LatView& v, long c, int x, int y, int z, C F[3])
{
    cell_force_in<FT, C>(v, v.g, v.flags, v.force, c, x, y, z, F);
}

// ------------------------------------------------------------- collision
template <class C>
__device__ __forceinline__ C sum_rho(const C f[Q])
{
    return f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] +
           f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
}

template <class C>
__device__ __forceinline__ void sum_j(const C f[Q], C j[3])
{
    const C xm = f[1] + f[4] + f[5] + f[6] + f[7];
    const C xp = f[10] + f[13] + f[14] + f[15] + f[16];
    const C ym = f[2] + f[4] + f[8] + f[9] + f[14];
    const C yp = f[5] + f[11] + f[13] + f[17] + f[18];
    const C zm = f[3] + f[6] + f[8] + f[16] + f[18];
    const C zp = f[7] + f[9] + f[12] + f[15] + f[17];
    j[0] = xp - xm;
    j[1] = yp - ym;
    j[2] = zp - zm;
}

// The collision with the moments of f already formed (rhoBar = sum_rho(f),
// j = sum_j(f), invRho = 1 / (rhoBar + 1)): a caller that also needs them
// (the velocity staging carried by the collision) forms them once.
template <class C, bool FORCE>
__device__ __forceinline__ void collide_m(C f[Q], const C F[3], C omega, C rhoBar, const C jm[3], C invRho)
{
    const C j[3] = {jm[0], jm[1], jm[2]};
    C u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = FORCE ? j[d] * invRho + F[d] / C(2) : j[d] * invRho;
    const C rho = rhoBar + C(1);
    const C jx = rho * u[0], jy = rho * u[1], jz = rho * u[2];

    const C one_m_omega = C(1) - omega;
    const C t0_omega = C(1.0 / 3.0) * omega;
    const C t1_omega = C(1.0 / 18.0) * omega;
    const C t4_omega = C(1.0 / 36.0) * omega;
    const C jSqr = jx * jx + jy * jy + jz * jz;
    const C kx = C(3) * jx, ky = C(3) * jy, kz = C(3) * jz;
    const C kxSqr_ = invRho / C(2) * kx * kx;
    const C kySqr_ = invRho / C(2) * ky * ky;
    const C kzSqr_ = invRho / C(2) * kz * kz;
    const C kxky_ = invRho * kx * ky;
    const C kxkz_ = invRho * kx * kz;
    const C kykz_ = invRho * ky * kz;
    const C C1 = rhoBar + invRho * C(3) * jSqr;
    C C2, C3;
    C3 = -kxSqr_ - kySqr_ - kzSqr_;
    f[0] = f[0] * one_m_omega + t0_omega * (C1 + C3);
#define BL_PAIR(a, b, tw_, c2, c3)                       \
    C2 = (c2);                                           \
    C3 = (c3);                                           \
    f[a] = f[a] * one_m_omega + tw_ * (C1 + C2 + C3);    \
    f[b] = f[b] * one_m_omega + tw_ * (C1 - C2 + C3);
    BL_PAIR(1, 10, t1_omega, -kx, -kySqr_ - kzSqr_)
    BL_PAIR(2, 11, t1_omega, -ky, -kxSqr_ - kzSqr_)
    BL_PAIR(3, 12, t1_omega, -kz, -kxSqr_ - kySqr_)
    BL_PAIR(4, 13, t4_omega, -kx - ky, kxky_ - kzSqr_)
    BL_PAIR(5, 14, t4_omega, -kx + ky, -kxky_ - kzSqr_)
    BL_PAIR(6, 15, t4_omega, -kx - kz, kxkz_ - kySqr_)
    BL_PAIR(7, 16, t4_omega, -kx + kz, -kxkz_ - kySqr_)
    BL_PAIR(8, 17, t4_omega, -ky - kz, kykz_ - kxSqr_)
    BL_PAIR(9, 18, t4_omega, -ky + kz, -kykz_ - kxSqr_)
#undef BL_PAIR

    if (FORCE) {
        const C mu = C(1) - omega / C(2);
        const C o6 = C(1.0 / 6.0), o12 = C(1.0 / 12.0);
        const C Fx = F[0], Fy = F[1], Fz = F[2];
        const C ux = u[0], uy = u[1], uz = u[2];
        f[0] += -mu * (Fx * ux + Fy * uy + Fz * uz);
        f[1] += o6 * mu * (Fx * (C(-1) + 2 * ux) - Fy * uy - Fz * uz);
        f[2] += -o6 * mu * (Fx * ux + Fy * (C(1) - 2 * uy) + Fz * uz);
        f[3] += -o6 * mu * (Fx * ux + Fy * uy + Fz * (C(1) - 2 * uz));
        f[4] += o12 * mu * (Fx * (C(-1) + 2 * ux + 3 * uy) + Fy * (C(-1) + 2 * uy + 3 * ux) - Fz * uz);
        f[5] += o12 * mu * (Fx * (C(-1) + 2 * ux - 3 * uy) + Fy * (C(1) + 2 * uy - 3 * ux) - Fz * uz);
        f[6] += o12 * mu * (Fx * (C(-1) + 2 * ux + 3 * uz) - Fy * uy + Fz * (C(-1) + 2 * uz + 3 * ux));
        f[7] += o12 * mu * (Fx * (C(-1) + 2 * ux - 3 * uz) - Fy * uy + Fz * (C(1) + 2 * uz - 3 * ux));
        f[8] += -o12 * mu * (Fx * ux + Fy * (C(1) - 2 * uy - 3 * uz) + Fz * (C(1) - 2 * uz - 3 * uy));
        f[9] += -o12 * mu * (Fx * ux + Fy * (C(1) - 2 * uy + 3 * uz) + Fz * (C(-1) - 2 * uz + 3 * uy));
        f[10] += o6 * mu * (Fx * (C(1) + 2 * ux) - Fy * uy - Fz * uz);
        f[11] += -o6 * mu * (Fx * ux + Fy * (C(-1) - 2 * uy) + Fz * uz);
        f[12] += -o6 * mu * (Fx * ux + Fy * uy + Fz * (C(-1) - 2 * uz));
        f[13] += o12 * mu * (Fx * (C(1) + 2 * ux + 3 * uy) + Fy * (C(1) + 2 * uy + 3 * ux) - Fz * uz);
        f[14] += o12 * mu * (Fx * (C(1) + 2 * ux - 3 * uy) + Fy * (C(-1) + 2 * uy - 3 * ux) - Fz * uz);
        f[15] += o12 * mu * (Fx * (C(1) + 2 * ux + 3 * uz) - Fy * uy + Fz * (C(1) + 2 * uz + 3 * ux));
        f[16] += o12 * mu * (Fx * (C(1) + 2 * ux - 3 * uz) - Fy * uy + Fz * (C(-1) + 2 * uz - 3 * ux));
        f[17] += -o12 * mu * (Fx * ux + Fy * (C(-1) - 2 * uy - 3 * uz) + Fz * (C(-1) - 2 * uz - 3 * uy));
        f[18] += -o12 * mu * (Fx * ux + Fy * (C(-1) - 2 * uy + 3 * uz) + Fz * (C(1) - 2 * uz + 3 * uy));
    }
}

// The collision in the FP16-shifted storage's working form g_i = fBar_i /
// t_i * 2^10 (Pop<__half>::ldw): the same BGK and Guo source as collide_m,
// with every term's weight t_i cancelled against the 1 / t_i of the form
// (the equilibrium's t_i omega becomes 2^10 omega; the Guo source's 1, 1/6
// and 1/12, which are 3 t_i, become 3 * 2^10 (1 - omega/2)).
template <bool FORCE>
__device__ __forceinline__ void collide_raw(float g[Q], const float F[3], float omega, float rhoBar,
                                            const float jm[3], float invRho)
{
    using C = float;
    const C s = C(kF16Scale);
    const C j[3] = {jm[0], jm[1], jm[2]};
    C u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = FORCE ? j[d] * invRho + F[d] / C(2) : j[d] * invRho;
    const C rho = rhoBar + C(1);
    const C jx = rho * u[0], jy = rho * u[1], jz = rho * u[2];

    const C one_m_omega = C(1) - omega;
    const C ws = s * omega;
    const C jSqr = jx * jx + jy * jy + jz * jz;
    const C kx = C(3) * jx, ky = C(3) * jy, kz = C(3) * jz;
    const C kxSqr_ = invRho / C(2) * kx * kx;
    const C kySqr_ = invRho / C(2) * ky * ky;
    const C kzSqr_ = invRho / C(2) * kz * kz;
    const C kxky_ = invRho * kx * ky;
    const C kxkz_ = invRho * kx * kz;
    const C kykz_ = invRho * ky * kz;
    const C C1 = rhoBar + invRho * C(3) * jSqr;
    C C2, C3;
    C3 = -kxSqr_ - kySqr_ - kzSqr_;
    g[0] = g[0] * one_m_omega + ws * (C1 + C3);
#define BL_PAIR(a, b, c2, c3)                          \
    C2 = (c2);                                         \
    C3 = (c3);                                         \
    g[a] = g[a] * one_m_omega + ws * (C1 + C2 + C3);   \
    g[b] = g[b] * one_m_omega + ws * (C1 - C2 + C3);
    BL_PAIR(1, 10, -kx, -kySqr_ - kzSqr_)
    BL_PAIR(2, 11, -ky, -kxSqr_ - kzSqr_)
    BL_PAIR(3, 12, -kz, -kxSqr_ - kySqr_)
    BL_PAIR(4, 13, -kx - ky, kxky_ - kzSqr_)
    BL_PAIR(5, 14, -kx + ky, -kxky_ - kzSqr_)
    BL_PAIR(6, 15, -kx - kz, kxkz_ - kySqr_)
    BL_PAIR(7, 16, -kx + kz, -kxkz_ - kySqr_)
    BL_PAIR(8, 17, -ky - kz, kykz_ - kxSqr_)
    BL_PAIR(9, 18, -ky + kz, -kykz_ - kxSqr_)
#undef BL_PAIR

    if (FORCE) {
        const C K = C(3) * s * (C(1) - omega / C(2));
        const C Fx = F[0], Fy = F[1], Fz = F[2];
        const C ux = u[0], uy = u[1], uz = u[2];
        g[0] += -K * (Fx * ux + Fy * uy + Fz * uz);
        g[1] += K * (Fx * (C(-1) + 2 * ux) - Fy * uy - Fz * uz);
        g[2] += -K * (Fx * ux + Fy * (C(1) - 2 * uy) + Fz * uz);
        g[3] += -K * (Fx * ux + Fy * uy + Fz * (C(1) - 2 * uz));
        g[4] += K * (Fx * (C(-1) + 2 * ux + 3 * uy) + Fy * (C(-1) + 2 * uy + 3 * ux) - Fz * uz);
        g[5] += K * (Fx * (C(-1) + 2 * ux - 3 * uy) + Fy * (C(1) + 2 * uy - 3 * ux) - Fz * uz);
        g[6] += K * (Fx * (C(-1) + 2 * ux + 3 * uz) - Fy * uy + Fz * (C(-1) + 2 * uz + 3 * ux));
        g[7] += K * (Fx * (C(-1) + 2 * ux - 3 * uz) - Fy * uy + Fz * (C(1) + 2 * uz - 3 * ux));
        g[8] += -K * (Fx * ux + Fy * (C(1) - 2 * uy - 3 * uz) + Fz * (C(1) - 2 * uz - 3 * uy));
        g[9] += -K * (Fx * ux + Fy * (C(1) - 2 * uy + 3 * uz) + Fz * (C(-1) - 2 * uz + 3 * uy));
        g[10] += K * (Fx * (C(1) + 2 * ux) - Fy * uy - Fz * uz);
        g[11] += -K * (Fx * ux + Fy * (C(-1) - 2 * uy) + Fz * uz);
        g[12] += -K * (Fx * ux + Fy * uy + Fz * (C(-1) - 2 * uz));
        g[13] += K * (Fx * (C(1) + 2 * ux + 3 * uy) + Fy * (C(1) + 2 * uy + 3 * ux) - Fz * uz);
        g[14] += K * (Fx * (C(1) + 2 * ux - 3 * uy) + Fy * (C(-1) + 2 * uy - 3 * ux) - Fz * uz);
        g[15] += K * (Fx * (C(1) + 2 * ux + 3 * uz) - Fy * uy + Fz * (C(1) + 2 * uz + 3 * ux));
        g[16] += K * (Fx * (C(1) + 2 * ux - 3 * uz) - Fy * uy + Fz * (C(-1) + 2 * uz - 3 * ux));
        g[17] += -K * (Fx * ux + Fy * (C(-1) - 2 * uy - 3 * uz) + Fz * (C(-1) - 2 * uz - 3 * uy));
        g[18] += -K * (Fx * ux + Fy * (C(-1) - 2 * uy + 3 * uz) + Fz * (C(1) - 2 * uz + 3 * uy));
    }
}

// The moments rhoBar = sum fBar_i and j = sum c_i fBar_i of populations
// in a storage's working form: plain sums, or for the FP16-shifted form
// the axis and diagonal sums weighted by t_i / 2^10.
template <class S>
__device__ __forceinline__ void wmoments(const typename Pop<S>::C f[Q], typename Pop<S>::C& rhoBar,
                                         typename Pop<S>::C j[3])
{
    using C = typename Pop<S>::C;
    if constexpr (Pop<S>::kRaw) {
        const C inv = C(1.0 / kF16Scale);
        const C t0 = C(1.0 / 3.0), t1 = C(1.0 / 18.0), t4 = C(1.0 / 36.0);
        const C ax = f[1] + f[2] + f[3] + f[10] + f[11] + f[12];
        const C dg = f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
        rhoBar = (t0 * f[0] + t1 * ax + t4 * dg) * inv;
        j[0] = (t1 * (f[10] - f[1]) + t4 * ((f[13] + f[14] + f[15] + f[16]) - (f[4] + f[5] + f[6] + f[7]))) * inv;
        j[1] = (t1 * (f[11] - f[2]) + t4 * ((f[5] + f[13] + f[17] + f[18]) - (f[4] + f[8] + f[9] + f[14]))) * inv;
        j[2] = (t1 * (f[12] - f[3]) + t4 * ((f[7] + f[9] + f[15] + f[17]) - (f[6] + f[8] + f[16] + f[18]))) * inv;
    } else {
        rhoBar = sum_rho(f);
        sum_j(f, j);
    }
}

// The collision of populations in a storage's working form.
template <class S, bool FORCE>
__device__ __forceinline__ void wcollide(typename Pop<S>::C f[Q], const typename Pop<S>::C F[3],
                                         typename Pop<S>::C omega, typename Pop<S>::C rhoBar,
                                         const typename Pop<S>::C j[3], typename Pop<S>::C invRho)
{
    if constexpr (Pop<S>::kRaw) collide_raw<FORCE>(f, F, omega, rhoBar, j, invRho);
    else collide_m<typename Pop<S>::C, FORCE>(f, F, omega, rhoBar, j, invRho);
}

// ------------------------------------------------------ velocity staging
// Staged velocity: one aligned 4-vector per cell (u, 0) so that a z-row of
// a stencil is one contiguous run of memory.
template <class C>
struct alignas(4 * sizeof(C)) Vel4 {
    C x, y, z, w;
};

// The progs' computeVelocity value of one cell (u = j/rho + F/2 from the
// post-stream populations' moments and the last collision's force,
// src/dataProcessors/dataAnalysisFunctional3D.hh:772-784; zero on
// bounce-back / no-dynamics cells, the wall velocity on velocity-wall
// cells), from the cell's moments j and invRho = 1 / (rhoBar + 1) as
// wmoments forms them (the collision that carries the staging forms them
// anyway): staged_velocity_f with the force at the cell, staged_velocity_m
// reading it from (flags, force), the last collision's force field, with
// the last collision's body force in the tiles it does not flag.
template <class C>
__device__ __forceinline__ Vel4<C> staged_velocity_f(const LatView& v, unsigned char m, int x, int y, int z,
                                                     const C j[3], C invRho, const C Fc[3])
{
    C u[3] = {C(0), C(0), C(0)};
    if (m == M_VWALL) {
        const int pp[3] = {x, y, z};
        const int side = pp[v.wall_axis] == 0 ? 0 : 1;
        for (int d = 0; d < 3; ++d) u[d] = (C)v.wall_u[3 * side + d];
    } else if (m == M_FLUID) {
        for (int d = 0; d < 3; ++d) u[d] = j[d] * invRho + Fc[d] / C(2);
    }
    return Vel4<C>{u[0], u[1], u[2], C(0)};
}

template <class FT, class C>
__device__ __forceinline__ Vel4<C> staged_velocity_m(const LatView& v, const unsigned char* flags, const void* force,
                                                     unsigned char m, long c, int x, int y, int z, const C j[3],
                                                     C invRho)
{
    C Fc[3] = {C(0), C(0), C(0)};
    if (m == M_FLUID) cell_force_in<FT, C>(v, v.g_last, flags, force, c, x, y, z, Fc);
    return staged_velocity_f<C>(v, m, x, y, z, j, invRho, Fc);
}

}  // namespace bl
