// Rigid particles (BL_MODEL_RIGID) and the rigid-growth packing.
//
// RigidParticle3D (src/fsi/RigidParticle.hh:13-98): vertex positions are
// com + R(q) (scale v_rest); a move sums the vertex forces and torques,
// integrates velocity, angular velocity (Euler's equations with the
// rest-shape inertia scaled by density scale^5), centre of mass and
// orientation (unit-norm exponential map), regenerates the vertices and maps
// the centre of mass onto the periodic grid (src/fsi/ParticleBase.hh:110-136)
// with the vertices.  The growth loop of progs/common/create_particles.h:
// 129-155 alternates the collision handler's forces (bl_ibm_compute_collision_forces)
// with these moves while the scale grows.
//
// State per particle (r_state, 32 doubles): com (3), velocity (3), angular
// velocity (3, world frame), quaternion (w, x, y, z), scale, density,
// damping, relaxation factor, rest-shape inertia (9, row-major, about the
// mesh origin), rest volume, com before the periodic mapping (3).
// Expressions keep the reference's operand order; contraction of a*b+c is
// off here so that the arithmetic is the reference's operation by operation.
#include <cmath>
#include <cfloat>

#include <string>

#include "ibm.hpp"

#pragma clang fp contract(off)

namespace bl {
namespace {

struct Grid3 {
    int n[3];
    int per[3];
};

constexpr int kRS = 32;
constexpr int kCom = 0, kVel = 3, kW = 6, kQ = 9, kScale = 13, kDens = 14, kDamp = 15, kRelax = 16, kI = 17,
              kVol = 26, kComPre = 27;

__host__ __device__ inline void cross3(const double a[3], const double b[3], double r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

// Quaternion::to_rot_matrix, apply_rotation, apply_inv_rotation (src/fsi/Quaternion.h)
__host__ __device__ inline void rot_matrix(const double* q, double m[9])
{
    m[0] = 1 - 2 * (q[2] * q[2] + q[3] * q[3]);
    m[1] = 2 * (q[1] * q[2] - q[0] * q[3]);
    m[2] = 2 * (q[0] * q[2] + q[1] * q[3]);
    m[3] = 2 * (q[1] * q[2] + q[0] * q[3]);
    m[4] = 1 - 2 * (q[1] * q[1] + q[3] * q[3]);
    m[5] = 2 * (q[2] * q[3] - q[0] * q[1]);
    m[6] = 2 * (q[1] * q[3] - q[0] * q[2]);
    m[7] = 2 * (q[0] * q[1] + q[2] * q[3]);
    m[8] = 1 - 2 * (q[1] * q[1] + q[2] * q[2]);
}

__device__ inline void apply_rotation(const double* q, const double v[3], double r[3])
{
    r[0] = (1 - 2 * (q[2] * q[2] + q[3] * q[3])) * v[0] + 2 * (q[1] * q[2] - q[0] * q[3]) * v[1] +
           2 * (q[0] * q[2] + q[1] * q[3]) * v[2];
    r[1] = 2 * (q[1] * q[2] + q[0] * q[3]) * v[0] + (1 - 2 * (q[1] * q[1] + q[3] * q[3])) * v[1] +
           2 * (q[2] * q[3] - q[0] * q[1]) * v[2];
    r[2] = 2 * (q[1] * q[3] - q[0] * q[2]) * v[0] + 2 * (q[0] * q[1] + q[2] * q[3]) * v[1] +
           (1 - 2 * (q[1] * q[1] + q[2] * q[2])) * v[2];
}

__device__ inline void apply_inv_rotation(const double* q, const double v[3], double r[3])
{
    r[0] = (1 - 2 * (q[2] * q[2] + q[3] * q[3])) * v[0] + 2 * (q[1] * q[2] + q[0] * q[3]) * v[1] +
           2 * (q[1] * q[3] - q[0] * q[2]) * v[2];
    r[1] = 2 * (q[1] * q[2] - q[0] * q[3]) * v[0] + (1 - 2 * (q[1] * q[1] + q[3] * q[3])) * v[1] +
           2 * (q[0] * q[1] + q[2] * q[3]) * v[2];
    r[2] = 2 * (q[0] * q[2] + q[1] * q[3]) * v[0] + 2 * (q[2] * q[3] - q[0] * q[1]) * v[1] +
           (1 - 2 * (q[1] * q[1] + q[2] * q[2])) * v[2];
}

__host__ __device__ inline void mat_vec(const double m[9], const double v[3], double r[3])
{
    r[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    r[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
    r[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

// Matrix<T,3>::solve (src/fsi/Matrix3.h:129-144): adjugate over determinant
__device__ inline void mat_solve(const double m[9], const double b[3], double r[3])
{
    const double det = m[0] * m[4] * m[8] - m[0] * m[5] * m[7] - m[1] * m[3] * m[8] + m[1] * m[5] * m[6] +
                       m[2] * m[3] * m[7] - m[2] * m[4] * m[6];
    r[0] = (b[2] * (m[1] * m[5] - m[2] * m[4]) + b[1] * (m[2] * m[7] - m[1] * m[8]) + b[0] * (m[4] * m[8] - m[5] * m[7])) / det;
    r[1] = (b[2] * (m[2] * m[3] - m[0] * m[5]) + b[1] * (m[0] * m[8] - m[2] * m[6]) + b[0] * (m[5] * m[6] - m[3] * m[8])) / det;
    r[2] = (b[2] * (m[0] * m[4] - m[1] * m[3]) + b[1] * (m[1] * m[6] - m[0] * m[7]) + b[0] * (m[3] * m[7] - m[4] * m[6])) / det;
}

// One thread per particle: RigidParticle3D::move_vertices up to the vertex
// update -- force and torque summed over the vertices in order, the
// accelerations, velocity, angular velocity, centre of mass, orientation --
// and the periodic mapping of the centre of mass (the pre-mapping centre is
// kept for the vertex update, which precedes the shift in the reference).
__global__ void __launch_bounds__(64) k_rigid_dynamics(int np, const int* __restrict__ p_model,
                                                       const long* __restrict__ p_voff, const double* __restrict__ pos,
                                                       const double* __restrict__ force, double* __restrict__ rs,
                                                       Grid3 gd)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || p_model[p] != BL_MODEL_RIGID) return;
    double* S = rs + (long)kRS * p;
    double com[3] = {S[kCom], S[kCom + 1], S[kCom + 2]};
    double F[3] = {0, 0, 0}, tq[3] = {0, 0, 0};
    for (long v = p_voff[p]; v < p_voff[p + 1]; ++v) {
        const double f[3] = {force[3 * v], force[3 * v + 1], force[3 * v + 2]};
        const double r[3] = {pos[3 * v] - com[0], pos[3 * v + 1] - com[1], pos[3 * v + 2] - com[2]};
        double c[3];
        cross3(r, f, c);
        for (int d = 0; d < 3; ++d) {
            F[d] += f[d];
            tq[d] += c[d];
        }
    }
    const double* q = S + kQ;
    const double scale = S[kScale], density = S[kDens], damping = S[kDamp];
    double w[3] = {S[kW], S[kW + 1], S[kW + 2]}, vel[3] = {S[kVel], S[kVel + 1], S[kVel + 2]};
    double wb[3], Ib[9], Iw[3], tb[3], cw[3], rhs[3], sol[3], wd[3], vdot[3];
    apply_inv_rotation(q, w, wb);
    const double fac = density * pow(scale, 5.0);
    for (int k = 0; k < 9; ++k) Ib[k] = fac * S[kI + k];
    const double mass = (scale * scale * scale) * S[kVol] * density;
    for (int d = 0; d < 3; ++d) vdot[d] = F[d] / mass - damping * vel[d];
    apply_inv_rotation(q, tq, tb);
    mat_vec(Ib, wb, Iw);
    cross3(wb, Iw, cw);
    for (int d = 0; d < 3; ++d) rhs[d] = tb[d] - cw[d];
    mat_solve(Ib, rhs, sol);
    apply_rotation(q, sol, wd);
    for (int d = 0; d < 3; ++d) {
        wd[d] = wd[d] - damping * w[d];
        vel[d] += vdot[d];
        w[d] += wd[d];
        com[d] += vel[d];
    }
    double qn[4] = {q[0], q[1], q[2], q[3]};
    const double om = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (om >= (double)10 * DBL_EPSILON) {
        const double ax[3] = {w[0] / om, w[1] / om, w[2] / om};
        const double sh = sin(om / 2);
        const double pq[4] = {cos(om / 2), sh * ax[0], sh * ax[1], sh * ax[2]};
        qn[0] = pq[0] * q[0] - (pq[1] * q[1] + pq[2] * q[2] + pq[3] * q[3]);
        qn[1] = pq[0] * q[1] + q[0] * pq[1] + pq[2] * q[3] - pq[3] * q[2];
        qn[2] = pq[0] * q[2] + q[0] * pq[2] + pq[3] * q[1] - pq[1] * q[3];
        qn[3] = pq[0] * q[3] + q[0] * pq[3] + pq[1] * q[2] - pq[2] * q[1];
    }
    for (int d = 0; d < 3; ++d) {
        S[kVel + d] = vel[d];
        S[kW + d] = w[d];
        S[kComPre + d] = com[d];
        double c = com[d];
        if (gd.per[d]) {
            const double L = (double)gd.n[d], x0 = -0.5, Linv = 1. / L;
            c = com[d] - L * floor((com[d] - x0) * Linv);
        }
        S[kCom + d] = c;
    }
    for (int k = 0; k < 4; ++k) S[kQ + k] = qn[k];
}

// One workgroup per particle: the vertices regenerated as
// com_pre + R (scale v_rest) (update()), then, with `shift`, moved by the
// periodic mapping's com - com_pre (set_center_of_mass).  With regen = 0
// only the shift (the mapping of a freshly added particle).
__global__ void __launch_bounds__(256) k_rigid_place(const int* __restrict__ p_model, const int* __restrict__ p_shape,
                                                    const long* __restrict__ p_voff, const int* __restrict__ s_voff,
                                                    const double* __restrict__ s_rest, const double* __restrict__ rs,
                                                    double* __restrict__ pos, int regen, int shift, int pid)
{
    const int p = pid >= 0 ? pid : (int)blockIdx.x;
    if (p_model[p] != BL_MODEL_RIGID) return;
    const double* S = rs + (long)kRS * p;
    double R[9];
    rot_matrix(S + kQ, R);
    const double s = S[kScale];
    const double* rest = s_rest + 3 * (long)s_voff[p_shape[p]];
    const long v0 = p_voff[p], nv = p_voff[p + 1] - v0;
    for (long i = threadIdx.x; i < nv; i += blockDim.x) {
        double* x = pos + 3 * (v0 + i);
        if (regen) {
            const double v[3] = {s * rest[3 * i], s * rest[3 * i + 1], s * rest[3 * i + 2]};
            double r[3];
            mat_vec(R, v, r);
            for (int d = 0; d < 3; ++d) x[d] = S[kComPre + d] + r[d];
        }
        if (shift)
            for (int d = 0; d < 3; ++d) x[d] += S[kCom + d] - S[kComPre + d];
    }
}

__global__ void k_rigid_scale(int np, const int* __restrict__ p_model, double* __restrict__ rs, int pid, double scale)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || p_model[p] != BL_MODEL_RIGID || (pid >= 0 && p != pid)) return;
    rs[(long)kRS * p + kScale] = scale;
    for (int d = 0; d < 3; ++d) rs[(long)kRS * p + kComPre + d] = rs[(long)kRS * p + kCom + d];
}

// The periodic mapping of init() (src/fsi/ImmersedBoundaryDynamics.hh:148-152)
// for freshly uploaded rigid particles: com_pre := com, com := mapped.
__global__ void k_rigid_map(int np, const int* __restrict__ p_model, double* __restrict__ rs, Grid3 gd)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || p_model[p] != BL_MODEL_RIGID) return;
    double* S = rs + (long)kRS * p;
    for (int d = 0; d < 3; ++d) {
        const double com = S[kCom + d];
        S[kComPre + d] = com;
        if (gd.per[d]) {
            const double L = (double)gd.n[d], x0 = -0.5, Linv = 1. / L;
            S[kCom + d] = com - L * floor((com - x0) * Linv);
        }
    }
}

Grid3 grid3_of(const bl_lattice_s* L)
{
    Grid3 g;
    g.n[0] = L->nx; g.n[1] = L->ny; g.n[2] = L->nz;
    for (int d = 0; d < 3; ++d) g.per[d] = L->periodic[d];
    return g;
}

// ParticleShape::compute_properties (src/fsi/ParticleShape.hh:249-360): the
// rest-shape moment of inertia about the mesh origin, in triangle order.
void shape_inertia(const ShapeHost& s, double I[9])
{
    double in[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < s.nt; ++t) {
        const double* v0 = &s.rest[3 * s.tri[3 * t]];
        const double* v1 = &s.rest[3 * s.tri[3 * t + 1]];
        const double* v2 = &s.rest[3 * s.tri[3 * t + 2]];
        const double a[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const double b[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        double N[3];
        cross3(a, b, N);
        double f1[3], f2[3], f3[3], g0[3], g1[3], g2[3];
        for (int d = 0; d < 3; ++d) {
            const double tmp0 = v0[d] + v1[d];
            f1[d] = tmp0 + v2[d];
            const double tmp1 = v0[d] * v0[d];
            const double tmp2 = tmp1 + v1[d] * tmp0;
            f2[d] = tmp2 + v2[d] * f1[d];
            f3[d] = v0[d] * tmp1 + v1[d] * tmp2 + v2[d] * f2[d];
            g0[d] = f2[d] + v0[d] * (f1[d] + v0[d]);
            g1[d] = f2[d] + v1[d] * (f1[d] + v1[d]);
            g2[d] = f2[d] + v2[d] * (f1[d] + v2[d]);
        }
        in[4] += N[0] * f3[0];
        in[5] += N[1] * f3[1];
        in[6] += N[2] * f3[2];
        in[7] += N[0] * (v0[1] * g0[0] + v1[1] * g1[0] + v2[1] * g2[0]);
        in[8] += N[1] * (v0[2] * g0[1] + v1[2] * g1[1] + v2[2] * g2[1]);
        in[9] += N[2] * (v0[0] * g0[2] + v1[0] * g1[2] + v2[0] * g2[2]);
    }
    for (int k = 4; k <= 6; ++k) in[k] *= (double)(1.0 / 60.0);
    for (int k = 7; k <= 9; ++k) in[k] *= (double)(1.0 / 120.0);
    I[0] = in[5] + in[6];
    I[1] = -in[7];
    I[2] = -in[9];
    I[3] = I[1];
    I[4] = in[4] + in[6];
    I[5] = -in[8];
    I[6] = I[2];
    I[7] = I[5];
    I[8] = in[4] + in[5];
}

int need_rigid(bl_ibm g, int pid, const char* what)
{
    if (!g) return fail(BL_ERR_ARG, "null ibm handle");
    if (!g->inited) return fail(BL_ERR_STATE, "immersed boundary not initialised (call bl_ibm_init)");
    if (!g->lat) return fail(BL_ERR_STATE, "the lattice of this immersed boundary was destroyed");
    if (g->n_rigid == 0) return fail(BL_ERR_STATE, std::string(what) + ": no rigid particles");
    if (pid < -1 || pid >= g->np || (pid >= 0 && g->parts[pid].model != BL_MODEL_RIGID))
        return fail(BL_ERR_ARG, std::string(what) + ": not a rigid particle");
    return BL_OK;
}

int rigid_move_launch(bl_ibm g, hipStream_t s)
{
    const Grid3 gd = grid3_of(g->lat);
    hipLaunchKernelGGL(k_rigid_dynamics, dim3((g->np + 63) / 64), dim3(64), 0, s, g->np, (const int*)g->p_model,
                       (const long*)g->p_voff, (const double*)g->pos, (const double*)g->force, g->r_state, gd);
    hipLaunchKernelGGL(k_rigid_place, dim3(g->np), dim3(256), 0, s, (const int*)g->p_model, (const int*)g->p_shape,
                       (const long*)g->p_voff, (const int*)g->s_voff, (const double*)g->s_rest,
                       (const double*)g->r_state, (double*)g->pos, 1, 1, -1);
    BL_CHECK_LAUNCH();
    g->binned = false;
    return BL_OK;
}

int rigid_scale_launch(bl_ibm g, int pid, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_rigid_scale, dim3((g->np + 63) / 64), dim3(64), 0, s, g->np, (const int*)g->p_model,
                       g->r_state, pid, scale);
    hipLaunchKernelGGL(k_rigid_place, dim3(pid >= 0 ? 1 : g->np), dim3(256), 0, s, (const int*)g->p_model,
                       (const int*)g->p_shape, (const long*)g->p_voff, (const int*)g->s_voff,
                       (const double*)g->s_rest, (const double*)g->r_state, (double*)g->pos, 1, 0, pid);
    BL_CHECK_LAUNCH();
    g->binned = false;
    return BL_OK;
}

}  // namespace

int rigid_init(bl_ibm g)
{
    g->n_rigid = 0;
    for (const auto& p : g->parts) g->n_rigid += p.model == BL_MODEL_RIGID ? 1 : 0;
    if (g->n_rigid == 0) return BL_OK;
    std::vector<double> rs((size_t)kRS * g->np, 0.0);
    for (int p = 0; p < g->np; ++p)
        if (g->parts[p].model == BL_MODEL_RIGID)
            for (int k = 0; k < kRS; ++k) rs[(size_t)kRS * p + k] = g->parts[p].rigid[k];
    BL_HIP(hipMalloc(&g->r_state, sizeof(double) * rs.size()));
    BL_HIP(hipMemcpy(g->r_state, rs.data(), sizeof(double) * rs.size(), hipMemcpyHostToDevice));
    // init(): the centre of mass onto the periodic grid, the vertices with it
    hipStream_t s = g->lat->stream;
    hipLaunchKernelGGL(k_rigid_map, dim3((g->np + 63) / 64), dim3(64), 0, s, g->np, (const int*)g->p_model, g->r_state,
                       grid3_of(g->lat));
    hipLaunchKernelGGL(k_rigid_place, dim3(g->np), dim3(256), 0, s, (const int*)g->p_model, (const int*)g->p_shape,
                       (const long*)g->p_voff, (const int*)g->s_voff, (const double*)g->s_rest,
                       (const double*)g->r_state, (double*)g->pos, 0, 1, -1);
    BL_CHECK_LAUNCH();
    return BL_OK;
}

int rigid_refuse_coupling(bl_ibm g)
{
    if (g->n_rigid == 0) return BL_OK;
    return fail(BL_ERR_STATE, "rigid particles move only through bl_rigid_move / bl_rigid_grow: their fluid coupling "
                              "(RigidParticle3D::compute_forces) is not implemented");
}

}  // namespace bl

using namespace bl;

extern "C" {

int bl_rigid_add(bl_ibm g, int shape_id, const double state[11], int* pid)
{
    if (!g || !state) return fail(BL_ERR_ARG, "bl_rigid_add: null argument");
    if (g->inited) return fail(BL_ERR_STATE, "bl_rigid_add: particles must be added before bl_ibm_init");
    if (shape_id < 0 || shape_id >= (int)g->shapes.size()) return fail(BL_ERR_ARG, "bl_rigid_add: unknown shape");
    if (!(state[7] > 0) || !(state[8] > 0) || !(state[9] >= 0))
        return fail(BL_ERR_ARG, "bl_rigid_add: scale and density must be positive, damping not negative");
    const ShapeHost& S = g->shapes[shape_id];
    if (!(S.volume > 0)) return fail(BL_ERR_ARG, "bl_rigid_add: the shape encloses no volume");
    bl_ibm_s::PartHost p;
    p.shape = shape_id;
    p.model = BL_MODEL_RIGID;
    double* R = p.rigid;
    for (int d = 0; d < 3; ++d) R[kCom + d] = state[d];
    for (int k = 0; k < 4; ++k) R[kQ + k] = state[3 + k];
    R[kScale] = state[7];
    R[kDens] = state[8];
    R[kDamp] = state[9];
    R[kRelax] = state[10];
    shape_inertia(S, R + kI);
    R[kVol] = S.volume;
    for (int d = 0; d < 3; ++d) R[kComPre + d] = state[d];
    // update(): com + R(q) (scale v_rest)
    double M[9];
    rot_matrix(R + kQ, M);
    p.pos.resize(3 * (size_t)S.nv);
    for (int i = 0; i < S.nv; ++i) {
        const double v[3] = {R[kScale] * S.rest[3 * i], R[kScale] * S.rest[3 * i + 1], R[kScale] * S.rest[3 * i + 2]};
        double r[3];
        mat_vec(M, v, r);
        for (int d = 0; d < 3; ++d) p.pos[3 * i + d] = R[kCom + d] + r[d];
    }
    p.id = (long long)g->parts.size();
    g->parts.push_back(std::move(p));
    if (pid) *pid = (int)g->parts.size() - 1;
    return BL_OK;
}

int bl_rigid_move(bl_ibm g)
{
    int rc = need_rigid(g, -1, "bl_rigid_move");
    if (rc) return rc;
    if ((rc = rigid_move_launch(g, g->lat->stream))) return rc;
    BL_HIP(hipStreamSynchronize(g->lat->stream));
    return BL_OK;
}

int bl_rigid_set_scale(bl_ibm g, int pid, double scale)
{
    int rc = need_rigid(g, pid, "bl_rigid_set_scale");
    if (rc) return rc;
    if (!(scale > 0)) return fail(BL_ERR_ARG, "bl_rigid_set_scale: scale must be positive");
    if ((rc = rigid_scale_launch(g, pid, scale, g->lat->stream))) return rc;
    BL_HIP(hipStreamSynchronize(g->lat->stream));
    return BL_OK;
}

int bl_rigid_get(bl_ibm g, int pid, double out[17])
{
    int rc = need_rigid(g, pid, "bl_rigid_get");
    if (rc) return rc;
    if (pid < 0 || !out) return fail(BL_ERR_ARG, "bl_rigid_get: one particle and an output array");
    BL_HIP(hipStreamSynchronize(g->lat->stream));
    BL_HIP(hipMemcpy(out, g->r_state + (long)kRS * pid, sizeof(double) * 17, hipMemcpyDeviceToHost));
    return BL_OK;
}

int bl_rigid_grow(bl_ibm g, int n_iter, double initial_scale, int every, const bl_potential* u, const bl_boundary* b)
{
    int rc = need_rigid(g, -1, "bl_rigid_grow");
    if (rc) return rc;
    if (n_iter < 1 || every < 1 || !(initial_scale > 0) || !u)
        return fail(BL_ERR_ARG, "bl_rigid_grow: n_iter and every must be positive, initial_scale positive, a potential");
    hipStream_t s = g->lat->stream;
    for (int it = 0; it <= n_iter; ++it) {
        if (it % every == 0) {
            const double fraction = (double)it / (double)n_iter;
            const double scale = initial_scale + (1. - initial_scale) * fraction;
            if ((rc = rigid_scale_launch(g, -1, scale, s))) return rc;
        }
        if ((rc = bl_ibm_compute_collision_forces(g, u, b))) return rc;
        if ((rc = rigid_move_launch(g, s))) return rc;
    }
    BL_HIP(hipStreamSynchronize(s));
    return BL_OK;
}

}  // extern "C"
