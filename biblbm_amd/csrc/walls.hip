// Curved walls: the Guo-Zheng-Shi extrapolation wall of the fsi module
// (src/fsi/GuoBoundaryCondition.hh) on the device lattice, and the boundary
// mask helper.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <unordered_map>
#include <vector>

#include "ibm.hpp"
#include "lattice.hpp"
#include "lattice_dev.hpp"
#include "walls.hpp"

struct bl_guo_wall_s {
    bl_lattice lat = nullptr;
    bl_boundary b{};
    long n = 0;
    std::vector<long long> cell;
    std::vector<int> dir;
    std::vector<double> delta, wall_pos, vel;
    double* d_vel = nullptr;  // 3 per node, in node order
    // The reference runs the nodes one after the other in place, so a node
    // whose link reads another wall node (lattice nodes on the boundary
    // surface) sees it updated if it comes earlier in node order and not
    // yet updated if it comes later.  Here the later ones are read from a
    // snapshot taken before the pass, and the nodes run in levels: a node's
    // level is one more than the highest level among the earlier nodes it
    // reads, and each level is one parallel launch.
    long n_snap = 0;
    std::vector<long> level_off;    // entries of level l: [level_off[l], level_off[l+1])
    int* d_order = nullptr;         // node indices by level, ascending within a level
    // per entry (level order): the node's cell, its link direction and, per
    // neighbour along the link, -1 (the lattice) or a snapshot slot; and the
    // node's link fraction
    int4* d_ent = nullptr;
    double* d_tdelta = nullptr;
    unsigned* d_snap_cell = nullptr;
    double* d_snap = nullptr;       // 19 per snapshot slot
    bool inserted = false;
};

namespace bl {
namespace {

// GuoRigidWallBoundaryFunctional::process (src/fsi/GuoBoundaryCondition.hh:
// 138-199, depth 2): moments and non-equilibrium parts of the two cells
// along the node's link, extrapolated wall momentum and non-equilibrium
// part, and the node's post-stream populations rewritten as
// feq(rhoBar_1, j_w) + fNeq_w.  The equilibrium comes from the node's own
// dynamics: zero on no-dynamics and bounce-back cells
// (src/core/dynamics.hh:773-777,946-950).  One thread per node, in the
// levels and with the snapshot described at bl_guo_wall_s, which give the
// result of the reference's node-by-node pass.  Arithmetic in FP64
// whatever the storage.
// invRho = 1 / (rhoBar + 1), formed once per cell by the caller
__device__ __forceinline__ double feq_d(int i, double rhoBar, double invRho, const double* j, double jSqr)
{
    const double c_j = cx(i) * j[0] + cy(i) * j[1] + cz(i) * j[2];
    return tw(i) * (rhoBar + 3. * c_j + invRho * (4.5 * c_j * c_j - 1.5 * jSqr));
}

__device__ __forceinline__ int wrap_axis(int q, int n, int per) { return per ? (q < 0 ? q + n : (q >= n ? q - n : q)) : q; }

__device__ __forceinline__ long nbr_cell(const LatView& v, int x, int y, int z, int i, int m, int& qx, int& qy,
                                         int& qz)
{
    qx = wrap_axis(x + m * cx(i), v.nx, v.per[0]);
    qy = wrap_axis(y + m * cy(i), v.ny, v.per[1]);
    qz = wrap_axis(z + m * cz(i), v.nz, v.per[2]);
    return ((long)qx * v.ny + qy) * v.nz + qz;
}

template <class S>
__device__ __forceinline__ void load_cell(const LatView& v, long q, int qx, int qy, int qz, bool swapped, double* f)
{
    using C = typename Pop<S>::C;
    Nbr nb;
    nb.init(v, qx, qy, qz);
    C g[Q];
    load_post<S>(v, nb, q, swapped, g);
    for (int p = 0; p < Q; ++p) f[p] = (double)g[p];
}

// The extrapolation for the node at cell c (mask code m, link fraction dl,
// wall velocity u) from its two neighbours' populations f1, f2.
template <class S>
__device__ void guo_node(const LatView& v, unsigned c, unsigned char m, double dl, const double* u, int x, int y,
                         int z, const double* f1, const double* f2, bool swapped)
{
    using C = typename Pop<S>::C;
    double j1[3], j2[3];
    const double rb1 = sum_rho(f1), rb2 = sum_rho(f2);
    sum_j(f1, j1);
    sum_j(f2, j2);
    const double j1sqr = j1[0] * j1[0] + j1[1] * j1[1] + j1[2] * j1[2];
    const double j2sqr = j2[0] * j2[0] + j2[1] * j2[1] + j2[2] * j2[2];
    const double wj[3] = {(rb1 + 1.0) * u[0], (rb1 + 1.0) * u[1], (rb1 + 1.0) * u[2]};
    double jw[3];
    if (dl < 0.75) {
        const double s = (1. - dl) / (1. + dl);
        for (int a = 0; a < 3; ++a) jw[a] = (wj[a] + (dl - 1.) * j1[a]) + s * (2. * wj[a] + (dl - 1.) * j2[a]);
    } else {
        for (int a = 0; a < 3; ++a) jw[a] = 1. / dl * (wj[a] + (dl - 1.) * j1[a]);
    }
    const double jsw = jw[0] * jw[0] + jw[1] * jw[1] + jw[2] * jw[2];
    const bool has_eq = m == M_FLUID || m == M_VWALL;
    const double inv1 = 1.0 / (rb1 + 1.0), inv2 = 1.0 / (rb2 + 1.0);
    C fw[Q];
    for (int p = 0; p < Q; ++p) {
        const double fneq1 = f1[p] - feq_d(p, rb1, inv1, j1, j1sqr);
        const double fneqw = dl < 0.75 ? dl * fneq1 + (1. - dl) * (f2[p] - feq_d(p, rb2, inv2, j2, j2sqr)) : fneq1;
        fw[p] = (C)((has_eq ? feq_d(p, rb1, inv1, jw, jsw) : 0.0) + fneqw);
    }
    Nbr nb;
    nb.init(v, x, y, z);
    store_post<S>(v, nb, c, swapped, fw);
}

// One level of nodes, one thread each.  Every input loads with no branch
// in front of it (the entry and node data, then both neighbours' 38
// populations from the lattice at once); a neighbour that is a wall node of
// a later entry then takes its populations from the snapshot instead (the
// lattice value read for it is dropped).
template <class S>
__global__ void __launch_bounds__(64) k_guo_wall(LatView v, const int4* __restrict__ ent,
                                                 const int* __restrict__ order, const double* __restrict__ tdelta,
                                                 const double* __restrict__ vel, long n,
                                                 const double* __restrict__ snap, bool swapped)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int4 e = ent[t];
    const long k = order[t];
    const double dl = tdelta[t];
    const unsigned c = (unsigned)e.x;
    int x, y, z;
    decode(c, v.ny, v.nz, x, y, z);
    const unsigned char m = v.mask ? v.mask[c] : (unsigned char)M_FLUID;
    double f[2][Q];
#pragma unroll
    for (int mm = 0; mm < 2; ++mm) {
        int qx, qy, qz;
        const long q = nbr_cell(v, x, y, z, e.y, mm + 1, qx, qy, qz);
        load_cell<S>(v, q, qx, qy, qz, swapped, f[mm]);
    }
    if (e.z >= 0)
        for (int p = 0; p < Q; ++p) f[0][p] = snap[(long)Q * e.z + p];
    if (e.w >= 0)
        for (int p = 0; p < Q; ++p) f[1][p] = snap[(long)Q * e.w + p];
    guo_node<S>(v, c, m, dl, vel + 3 * k, x, y, z, f[0], f[1], swapped);
}

// populations of the wall nodes that earlier nodes read, taken before the
// pass changes them
template <class S>
__global__ void __launch_bounds__(64) k_guo_snapshot(LatView v, long n, const unsigned* __restrict__ cell,
                                                     double* __restrict__ snap, bool swapped)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int x, y, z;
    decode(cell[s], v.ny, v.nz, x, y, z);
    load_cell<S>(v, cell[s], x, y, z, swapped, snap + (long)Q * s);
}

template <class S>
void launch_guo(bl_guo_wall w, const LatView& v, bool swapped, hipStream_t s)
{
    if (w->n_snap)
        hipLaunchKernelGGL(k_guo_snapshot<S>, dim3((unsigned)((w->n_snap + 63) / 64)), dim3(64), 0, s, v, w->n_snap,
                           (const unsigned*)w->d_snap_cell, w->d_snap, swapped);
    for (size_t l = 0; l + 1 < w->level_off.size(); ++l) {
        const long a = w->level_off[l], n = w->level_off[l + 1] - a;
        hipLaunchKernelGGL(k_guo_wall<S>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, v,
                           (const int4*)w->d_ent + a, (const int*)w->d_order + a, (const double*)w->d_tdelta + a,
                           (const double*)w->d_vel, n, (const double*)w->d_snap, swapped);
    }
}

int upload_vel(bl_guo_wall w)
{
    if (w->n == 0) return BL_OK;
    BL_HIP(hipMemcpyAsync(w->d_vel, w->vel.data(), sizeof(double) * 3 * w->n, hipMemcpyHostToDevice, w->lat->stream));
    BL_HIP(hipStreamSynchronize(w->lat->stream));
    return BL_OK;
}

int need_wall(bl_guo_wall w)
{
    if (!w) return fail(BL_ERR_ARG, "null Guo wall handle");
    if (!w->lat) return fail(BL_ERR_STATE, "the lattice of this Guo wall was destroyed");
    return BL_OK;
}

}  // namespace

int guo_walls_launch(bl_lattice_s* L)
{
    if (L->guo.empty()) return BL_OK;
    const LatView v = view_of(L);
    const bool swapped = (L->steps & 1) != 0;
    for (bl_guo_wall w : L->guo) {
        if (w->n == 0) continue;
        if (L->storage == BL_STORE_F64) launch_guo<double>(w, v, swapped, L->stream);
        else if (L->storage == BL_STORE_F32) launch_guo<float>(w, v, swapped, L->stream);
        else launch_guo<__half>(w, v, swapped, L->stream);
    }
    BL_CHECK_LAUNCH();
    return BL_OK;
}

void guo_walls_detach(bl_lattice_s* L)
{
    for (bl_guo_wall w : L->guo) {
        w->lat = nullptr;
        w->inserted = false;
    }
    L->guo.clear();
}

}  // namespace bl

using namespace bl;

extern "C" {

int bl_guo_wall_create(bl_lattice lat, const bl_boundary* b, bl_guo_wall* out)
{
    if (!lat || !b || !out) return fail(BL_ERR_ARG, "bl_guo_wall_create: null argument");
    if (b->kind != BL_BOUNDARY_PIPE && b->kind != BL_BOUNDARY_PLATES)
        return fail(BL_ERR_ARG, "bl_guo_wall_create: unknown boundary kind");
    const int n3[3] = {lat->nx, lat->ny, lat->nz};
    bl_guo_wall w = new bl_guo_wall_s();
    w->lat = lat;
    w->b = *b;
    // GuoRigidWallBoundaryInstantiator::process (src/fsi/GuoBoundaryCondition.hh:62-135)
    for (int x = 0; x < n3[0]; ++x)
        for (int y = 0; y < n3[1]; ++y)
            for (int z = 0; z < n3[2]; ++z) {
                const double pos[3] = {(double)x, (double)y, (double)z};
                if (bnd_contains(*b, pos, 0)) continue;
                double cos_min = 2, delta_min = -1, wmin[3] = {0, 0, 0};
                int dmin = 0;
                bool success = false;
                for (int i = 1; i < Q; ++i) {
                    const double tip[3] = {pos[0] + cx(i), pos[1] + cy(i), pos[2] + cz(i)};
                    double t = -1;
                    if (!bnd_trace_ray(*b, pos, tip, t)) continue;
                    const double nrm = std::sqrt((double)(cx(i) * cx(i) + cy(i) * cy(i) + cz(i) * cz(i)));
                    const double d2[3] = {cx(i) / nrm, cy(i) / nrm, cz(i) / nrm};
                    // the reference steps along the normalised direction
                    const double wp[3] = {pos[0] + d2[0] * t, pos[1] + d2[1] * t, pos[2] + d2[2] * t};
                    double nv[3];
                    bnd_normal(*b, wp, nv);
                    const double ca = nv[0] * d2[0] + nv[1] * d2[1] + nv[2] * d2[2];
                    if (ca < cos_min) {
                        dmin = i;
                        cos_min = ca;
                        delta_min = 1 - t;
                        std::copy(wp, wp + 3, wmin);
                    }
                    success = true;
                }
                if (!success) continue;
                w->cell.push_back(((long long)x * n3[1] + y) * n3[2] + z);
                w->dir.push_back(dmin);
                w->delta.push_back(delta_min);
                w->wall_pos.insert(w->wall_pos.end(), wmin, wmin + 3);
            }
    w->n = (long)w->cell.size();
    w->vel.assign(3 * w->n, 0.0);
    // the two neighbours along each node's link must lie in the lattice;
    // reads of wall nodes decide the snapshot and the levels
    std::unordered_map<long long, int> node_of;
    for (long k = 0; k < w->n; ++k) node_of[w->cell[k]] = (int)k;
    std::vector<int> level(w->n, 0), src(2 * w->n, -1), snap_node;
    std::unordered_map<int, int> snap_slot;
    for (long k = 0; k < w->n; ++k) {
        const long long c = w->cell[k];
        const int p[3] = {(int)(c / ((long long)n3[1] * n3[2])), (int)((c / n3[2]) % n3[1]), (int)(c % n3[2])};
        const int d[3] = {cx(w->dir[k]), cy(w->dir[k]), cz(w->dir[k])};
        for (int m = 1; m <= 2; ++m) {
            int q[3];
            bool inside = true;
            for (int a = 0; a < 3; ++a) {
                q[a] = p[a] + m * d[a];
                if (q[a] < 0 || q[a] >= n3[a]) {
                    if (lat->periodic[a]) q[a] = (q[a] + n3[a]) % n3[a];
                    else inside = false;
                }
            }
            if (!inside) {
                delete w;
                return fail(BL_ERR_ARG, "bl_guo_wall_create: wall node (" + std::to_string(p[0]) + ", " +
                                            std::to_string(p[1]) + ", " + std::to_string(p[2]) +
                                            ") has a neighbour along its link outside the lattice");
            }
            const auto it = node_of.find(((long long)q[0] * n3[1] + q[1]) * n3[2] + q[2]);
            if (it == node_of.end()) continue;
            const int j = it->second;
            if (j < (int)k) {  // already updated when k runs
                level[k] = std::max(level[k], level[j] + 1);
            } else {           // not yet: its value from before the pass
                auto sl = snap_slot.find(j);
                if (sl == snap_slot.end()) {
                    sl = snap_slot.emplace(j, (int)snap_node.size()).first;
                    snap_node.push_back(j);
                }
                src[2 * k + m - 1] = sl->second;
            }
        }
    }
    const int nlev = w->n ? *std::max_element(level.begin(), level.end()) + 1 : 0;
    std::vector<int> order;
    std::vector<int4> ent;
    std::vector<double> tdelta;
    w->level_off.assign(1, 0);
    for (int l = 0; l < nlev; ++l) {
        for (long k = 0; k < w->n; ++k)
            if (level[k] == l) {
                order.push_back((int)k);
                ent.push_back(make_int4((int)w->cell[k], w->dir[k], src[2 * k], src[2 * k + 1]));
                tdelta.push_back(w->delta[k]);
            }
        w->level_off.push_back((long)order.size());
    }
    w->n_snap = (long)snap_node.size();
    if (w->n) {
        std::vector<unsigned> snap_cell;
        for (int j : snap_node) snap_cell.push_back((unsigned)w->cell[j]);
        hipError_t e = hipMalloc(&w->d_vel, sizeof(double) * 3 * w->n);
        if (e == hipSuccess) e = hipMalloc(&w->d_order, sizeof(int) * w->n);
        if (e == hipSuccess) e = hipMalloc(&w->d_ent, sizeof(int4) * w->n);
        if (e == hipSuccess) e = hipMalloc(&w->d_tdelta, sizeof(double) * w->n);
        if (e == hipSuccess) e = hipMalloc(&w->d_snap_cell, sizeof(unsigned) * std::max<size_t>(snap_cell.size(), 1));
        if (e == hipSuccess) e = hipMalloc(&w->d_snap, sizeof(double) * Q * std::max<size_t>(snap_cell.size(), 1));
        if (e == hipSuccess) e = hipMemset(w->d_vel, 0, sizeof(double) * 3 * w->n);
        if (e == hipSuccess) e = hipMemcpy(w->d_order, order.data(), sizeof(int) * w->n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(w->d_ent, ent.data(), sizeof(int4) * w->n, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(w->d_tdelta, tdelta.data(), sizeof(double) * w->n, hipMemcpyHostToDevice);
        if (e == hipSuccess && !snap_cell.empty())
            e = hipMemcpy(w->d_snap_cell, snap_cell.data(), sizeof(unsigned) * snap_cell.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            bl_guo_wall_destroy(w);
            return fail(BL_ERR_HIP, std::string("bl_guo_wall_create: ") + hipGetErrorString(e));
        }
    }
    *out = w;
    return BL_OK;
}

int bl_guo_wall_destroy(bl_guo_wall w)
{
    if (!w) return BL_OK;
    if (w->lat) {
        hipStreamSynchronize(w->lat->stream);
        auto& v = w->lat->guo;
        v.erase(std::remove(v.begin(), v.end(), w), v.end());
        ++w->lat->gen;
        if (w->lat->ibm) ibm_invalidate_graphs(w->lat->ibm);
    }
    hipFree(w->d_vel);
    hipFree(w->d_order);
    hipFree(w->d_ent);
    hipFree(w->d_tdelta);
    hipFree(w->d_snap_cell);
    hipFree(w->d_snap);
    delete w;
    return BL_OK;
}

int bl_guo_wall_count(bl_guo_wall w, long long* n)
{
    if (!w || !n) return fail(BL_ERR_ARG, "bl_guo_wall_count: null argument");
    *n = w->n;
    return BL_OK;
}

int bl_guo_wall_nodes(bl_guo_wall w, long long* cell, int* dir, double* delta, double* wall_pos)
{
    if (!w) return fail(BL_ERR_ARG, "null Guo wall handle");
    if (cell) std::copy(w->cell.begin(), w->cell.end(), cell);
    if (dir) std::copy(w->dir.begin(), w->dir.end(), dir);
    if (delta) std::copy(w->delta.begin(), w->delta.end(), delta);
    if (wall_pos) std::copy(w->wall_pos.begin(), w->wall_pos.end(), wall_pos);
    return BL_OK;
}

int bl_guo_wall_set_velocity(bl_guo_wall w, const double vel[3])
{
    int rc = need_wall(w);
    if (rc) return rc;
    if (!vel) return fail(BL_ERR_ARG, "bl_guo_wall_set_velocity: null velocity");
    for (long k = 0; k < w->n; ++k)
        for (int a = 0; a < 3; ++a) w->vel[3 * k + a] = vel[a];
    return upload_vel(w);
}

int bl_guo_wall_set_node_velocities(bl_guo_wall w, const double* vel)
{
    int rc = need_wall(w);
    if (rc) return rc;
    if (!vel && w->n) return fail(BL_ERR_ARG, "bl_guo_wall_set_node_velocities: null array");
    std::copy(vel, vel + 3 * w->n, w->vel.begin());
    return upload_vel(w);
}

int bl_guo_wall_insert(bl_guo_wall w)
{
    int rc = need_wall(w);
    if (rc) return rc;
    if (w->inserted) return fail(BL_ERR_STATE, "bl_guo_wall_insert: already inserted");
    w->lat->guo.push_back(w);
    w->inserted = true;
    ++w->lat->gen;
    if (w->lat->ibm) ibm_invalidate_graphs(w->lat->ibm);
    return BL_OK;
}

int bl_boundary_outside_mask(const bl_boundary* b, int nx, int ny, int nz, double margin, unsigned char* mask,
                             unsigned char value)
{
    if (!b || !mask || nx <= 0 || ny <= 0 || nz <= 0) return fail(BL_ERR_ARG, "bl_boundary_outside_mask: bad argument");
    if (b->kind != BL_BOUNDARY_PIPE && b->kind != BL_BOUNDARY_PLATES)
        return fail(BL_ERR_ARG, "bl_boundary_outside_mask: unknown boundary kind");
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const double p[3] = {(double)x, (double)y, (double)z};
                if (!bnd_contains(*b, p, -margin)) mask[((long)x * ny + y) * nz + z] = value;
            }
    return BL_OK;
}

}  // extern "C"
