/*
 * biblbm_calibrate.h -- initial-condition helpers of the progs
 * (progs/common/calibrate_rbc_shape.h) as this project runs them: the RBC
 * parameter recipe and the volume-shrink calibration of every step on the
 * device at once (bl_rbc_shrink_volume).  Progs that include the
 * reference's own calibrate_rbc_shape.h get its loop over
 * RBCParticle::relax_nodes instead, which runs one device step per call.
 */
#ifndef BIBLBM_CALIBRATE_H
#define BIBLBM_CALIBRATE_H
#include "biblbm_plb.h"

namespace plb {
namespace fsi {

// create_rbc_params (progs/common/calibrate_rbc_shape.h:75-112): L0 the mean
// link length (distinct edges in (i0, i1) order), vol_desired the rest
// volume, the area stiffness beyond K() = 2 G split 0.8 / 0.2 between
// global and local, k_volume = 1000 G / radius, theta0 of a regular
// triangulation with the shape's vertex count.
template <class T>
RBCParameters<T> create_rbc_params(const ParticleShape<T>* shape, T shear_modulus, T K_area, T K_bend)
{
    std::vector<std::pair<plint, plint>> edges;
    for (const auto& t : shape->triangles)
        for (int k = 0; k < 3; ++k) {
            const plint a = t[k], b = t[(k + 1) % 3];
            edges.push_back(std::make_pair(std::min(a, b), std::max(a, b)));
        }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    T lsum = 0;
    for (const auto& e : edges) {
        const Array<T, 3>&a = shape->vertices[e.first], &b = shape->vertices[e.second];
        lsum += std::sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]) + (b[2] - a[2]) * (b[2] - a[2]));
    }
    T vol = 0;
    for (const auto& t : shape->triangles) {
        const Array<T, 3>&a = shape->vertices[t[0]], &b = shape->vertices[t[1]], &c = shape->vertices[t[2]];
        vol += (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                a[2] * (b[0] * c[1] - b[1] * c[0])) / 6;
    }
    const T nv = (T)shape->vertices.size(), s3 = std::sqrt((T)3);
    RBCParameters<T> p;
    p.shear_modulus = shear_modulus;
    p.L0 = lsum / (T)edges.size();
    p.vol_desired = vol;
    const T delta = K_area - 2 * shear_modulus;
    p.k_area_global = delta * 0.8;
    p.k_area_local = delta * 0.2;
    p.k_volume = 1000. * shear_modulus / shape->get_radius();
    p.k_bend = K_bend;
    p.theta0 = std::acos((s3 * (nv - 2.) - 5 * M_PI) / (s3 * (nv - 2.) - 3. * M_PI));
    return p;
}

// shrink_rbc_volume(rbc, vol_final, N_it, m, damping)
// (progs/common/calibrate_rbc_shape.h:13-72) on the device: the particle's
// vertices relax while vol_desired ramps to vol_final; on return the
// particle holds the calibrated vertices and vol_desired.  A scratch
// lattice and immersed boundary carry the device state.
template <class T>
void shrink_rbc_volume(RBCParticle<T>& rbc, T vol_final, int N_it, T m = 1, T damping = 1.e-2)
{
    (void)m;
    bl_lattice L = nullptr;
    bl_ibm g = nullptr;
    const int per[3] = {0, 0, 0};
    auto cleanup = [&]() {
        if (g) bl_ibm_destroy(g);
        if (L) bl_lattice_destroy(L);
    };
    try {
        blCheck(bl_lattice_create(4, 4, 4, per, BL_STORE_F64, &L), "shrink_rbc_volume");
        blCheck(bl_ibm_create(L, BL_KERNEL_PESKIN, &g), "shrink_rbc_volume");
        const ParticleShape<T>* s = rbc.get_shape();
        std::vector<double> v, x;
        std::vector<int> t;
        for (const auto& a : s->vertices) v.insert(v.end(), {a[0], a[1], a[2]});
        for (const auto& a : s->triangles) t.insert(t.end(), {(int)a[0], (int)a[1], (int)a[2]});
        for (auto it = rbc.begin(); it != rbc.end(); ++it) x.insert(x.end(), {it->pos[0], it->pos[1], it->pos[2]});
        int sid = -1, pid = -1;
        blCheck(bl_shape_add(g, (int)s->vertices.size(), (int)s->triangles.size(), v.data(), t.data(), &sid),
                "shrink_rbc_volume");
        const std::vector<double> prm = rbc.device_params();
        blCheck(bl_particle_add(g, sid, BL_MODEL_RBC, prm.data(), (int)prm.size(), x.data(), 1, &pid),
                "shrink_rbc_volume");
        blCheck(bl_ibm_init(g), "shrink_rbc_volume");
        blCheck(bl_rbc_shrink_volume(g, pid, vol_final, N_it, damping, nullptr), "shrink_rbc_volume");
        blCheck(bl_particle_get(g, pid, 0, x.data()), "shrink_rbc_volume");
        for (plint k = 0; k < rbc.count_nodes(); ++k) rbc.get_node(k).pos = Array<T, 3>(x[3 * k], x[3 * k + 1], x[3 * k + 2]);
        rbc.params().vol_desired = vol_final;
        rbc.update();
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

}  // namespace fsi
}  // namespace plb

#endif
