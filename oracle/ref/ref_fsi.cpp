// Golden vectors for the fsi membrane, Dirac and voxelization code, produced
// by the reference classes themselves (the parts of src/fsi that build
// without MPI):
//   ParticleShape rest geometry / connectivity   src/fsi/ParticleShape.hh:21-212
//   gmsh reader                                   src/fsi/ParticleShapeFactory.hh:32-83
//   DeformableCapsuleParticle::compute_forces     src/fsi/DeformableCapsule.hh:60-150
//   RBCParticle::compute_forces                   src/fsi/RBCParticle.hh:140-247
//   DeformableParticle3D::update / inertia        src/fsi/DeformableParticle3D.hh:54-82,173-272
//   RayTracer::find_intersections_z               src/fsi/RayTracer.hh:19-39
//   Roma / Peskin / TopHat Dirac + support box    src/fsi/Dirac.h:23-139
//
// The Skalak shear law is selected through the particle's own serialization
// (pack, overwrite the 4-byte shear-model field, re-create through the
// particle factory), since the class exposes no setter.
//
// Usage: ref_fsi [reference_root]   (writes ./fsi/...)
#include "palabos3D.h"
#include "palabos3D.hh"
#include <cstring>
#include <cstdio>
#include "fsi/ParticleBase.h"
#include "fsi/ParticleShape.h"
#include "fsi/ParticleShapeFactory.h"
#include "fsi/DeformableParticle3D.h"
#include "fsi/RBCParticle.h"
#include "fsi/DeformableCapsule.h"
#include "fsi/ParticleFactory.h"
#include "fsi/Transform.h"
#include "fsi/Dirac.h"
#include "fsi/ParticleBase.hh"
#include "fsi/RBCParticle.hh"
#include "fsi/ParticleShape.hh"
#include "fsi/ParticleShapeFactory.hh"
#include "fsi/DeformableParticle3D.hh"
#include "fsi/Transform.hh"
#include "fsi/ParticleFactory.hh"
#include "fsi/DeformableCapsule.hh"
#include "fsi/RayTracer.hh"
#include "common.h"

using namespace plb;
using namespace plb::fsi;
typedef double T;

namespace {

struct HitRecorder {
    std::vector<T> hits;
    void put(plint i, plint j, T z)
    {
        hits.push_back((T)i);
        hits.push_back((T)j);
        hits.push_back(z);
    }
};

template <class P>
void set_positions(P& p, const std::vector<T>& pos)
{
    for (plint i = 0; i < p.count_nodes(); ++i)
        for (int d = 0; d < 3; ++d) p.get_node(i).pos[d] = pos[3 * i + d];
    p.update();
}

template <class P>
std::vector<T> forces_of(P& p)
{
    p.reset_forces();
    p.compute_forces();
    std::vector<T> f(3 * p.count_nodes());
    for (plint i = 0; i < p.count_nodes(); ++i)
        for (int d = 0; d < 3; ++d) f[3 * i + d] = p.get_node(i).force[d];
    return f;
}

void mesh_case(const std::string& ref, const std::string& mesh, const std::string& name,
               ParticleShapeLibrary<T>& lib, unsigned long long seed)
{
    const std::string dir = "fsi/" + name;
    refio::ensure_dir(dir);
    lib.read_and_store_mesh(ref + "/progs/" + mesh, name);
    ParticleShape<T>* shape = lib.get_by_tag(name);
    const long nv = shape->count_vertices(), nt = shape->count_triangles();

    std::vector<T> rest(3 * nv);
    for (long i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) rest[3 * i + d] = shape->get_vertex(i)[d];
    std::vector<int> tri(3 * nt);
    std::vector<T> tri_area(nt);
    for (long t = 0; t < nt; ++t) {
        const Triangle<T>& tr = shape->get_triangle(t);
        tri[3 * t] = tr.i0; tri[3 * t + 1] = tr.i1; tri[3 * t + 2] = tr.i2;
        tri_area[t] = tr.area;
    }
    std::vector<int> link_i, link_t;
    std::vector<T> link_len;
    for (typename ParticleShape<T>::link_const_iterator it = shape->links_begin(); it != shape->links_end(); ++it) {
        link_i.push_back(it->i0); link_i.push_back(it->i1);
        link_t.push_back(it->t1); link_t.push_back(it->t2);
        link_len.push_back(it->length);
    }
    std::vector<int> pair_i;
    std::vector<T> pair_cs;
    for (typename ParticleShape<T>::triangle_pair_const_iterator it = shape->triangle_pairs_begin();
         it != shape->triangle_pairs_end(); ++it) {
        pair_i.push_back(it->i0); pair_i.push_back(it->i1); pair_i.push_back(it->i2); pair_i.push_back(it->i3);
        pair_cs.push_back(it->cos_theta0); pair_cs.push_back(it->sin_theta0);
    }
    const long nl = link_len.size(), np = pair_cs.size() / 2;
    Array<T, 3> c = shape->get_center();
    T scal[6] = {shape->get_area(), shape->get_volume(), c[0], c[1], c[2], shape->get_radius()};
    refio::write_npy(dir + "/rest_pos.npy", rest, {nv, 3L});
    refio::write_npy(dir + "/tri.npy", tri, {nt, 3L});
    refio::write_npy(dir + "/tri_area.npy", tri_area, {nt});
    refio::write_npy(dir + "/link_i.npy", link_i, {nl, 2L});
    refio::write_npy(dir + "/link_t.npy", link_t, {nl, 2L});
    refio::write_npy(dir + "/link_len.npy", link_len, {nl});
    refio::write_npy(dir + "/pair_i.npy", pair_i, {np, 4L});
    refio::write_npy(dir + "/pair_cs0.npy", pair_cs, {np, 2L});
    refio::write_npy(dir + "/shape_scalars.npy", scal, {6L});

    // Seeded deformation: affine stretch + shear about the centre, per-vertex
    // noise, then a translation into a lattice-sized box.
    refio::Rng rng(seed);
    const T A[3][3] = {{1.12, 0.05, 0.0}, {0.0, 0.93, 0.02}, {0.01, 0.0, 0.97}};
    const T shift[3] = {20.3, 17.1, 25.6};
    const T lm = shape->get_triangle(0).l_mean;
    std::vector<T> pos(3 * nv);
    for (long i = 0; i < nv; ++i) {
        T r[3];
        for (int d = 0; d < 3; ++d) r[d] = rest[3 * i + d] - c[d];
        for (int d = 0; d < 3; ++d)
            pos[3 * i + d] = c[d] + A[d][0] * r[0] + A[d][1] * r[1] + A[d][2] * r[2] +
                             0.03 * lm * rng.uniform(-1., 1.) + shift[d];
    }
    refio::write_npy(dir + "/pos.npy", pos, {nv, 3L});

    // Capsule, NeoHookean (the class default).
    CapsuleParameters<T> cp;
    cp.G = 0.02; cp.C = 10.; cp.nu_s = 1.; cp.kb = 0.;
    DeformableCapsuleParticle<T> neo(shape, cp);
    set_positions(neo, pos);
    refio::write_npy(dir + "/force_neohookean.npy", forces_of(neo), {nv, 3L});

    // Capsule, Skalak (via serialization: shear_model_ follows the header
    // (3 x plint + bool) and the vertex block (pos, vel, force per vertex)).
    std::vector<char> buf;
    neo.pack(buf);
    const size_t off = 3 * sizeof(plint) + sizeof(bool) + nv * 9 * sizeof(T);
    if (buf.size() != off + sizeof(int) + 4 * sizeof(T)) {
        std::fprintf(stderr, "unexpected capsule pack size %zu\n", buf.size());
        std::exit(1);
    }
    int skalak = 0;  // enum ShearModel {Skalak, NeoHookean}
    std::memcpy(&buf[off], &skalak, sizeof(int));
    char* it = &buf[0];
    ParticleBase3D<T>* sk = particleFactory<T>().create(it, lib);
    DeformableCapsuleParticle<T>* skp = dynamic_cast<DeformableCapsuleParticle<T>*>(sk);
    skp->update();
    refio::write_npy(dir + "/force_skalak.npy", forces_of(*skp), {nv, 3L});
    delete sk;

    // RBC (WLC-POW links, local + global area, volume, bending).
    RBCParameters<T> rp;
    rp.shear_modulus = 0.015; rp.L0 = lm; rp.k_area_global = 0.6; rp.k_area_local = 0.05;
    rp.k_volume = 0.8; rp.vol_desired = 0.9 * shape->get_volume(); rp.k_bend = 0.012; rp.theta0 = 0.;
    RBCParticle<T> rbc(shape, rp);
    set_positions(rbc, pos);
    refio::write_npy(dir + "/force_rbc.npy", forces_of(rbc), {nv, 3L});
    T rbcp[8] = {rp.shear_modulus, rp.L0, rp.k_area_global, rp.k_area_local,
                 rp.k_volume, rp.vol_desired, rp.k_bend, rp.theta0};
    refio::write_npy(dir + "/rbc_params.npy", rbcp, {8L});

    // Geometry of the deformed particle.
    geo::Rect<T> bb = rbc.bounding_box();
    T geom[11] = {rbc.area(), rbc.volume(), rbc.center_of_mass()[0], rbc.center_of_mass()[1],
                  rbc.center_of_mass()[2], bb.x0, bb.x1, bb.y0, bb.y1, bb.z0, bb.z1};
    refio::write_npy(dir + "/geometry.npy", geom, {11L});
    Matrix<T, 3> I;
    rbc.compute_moment_of_intertia(I);
    T Iv[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Iv[3 * a + b] = I(a, b);
    refio::write_npy(dir + "/inertia.npy", Iv, {3L, 3L});

    // z-ray hits (before periodic remapping).
    HitRecorder rec;
    rbc.voxelizer().find_intersections_z(rec);
    refio::write_npy(dir + "/ray_hits_z.npy", rec.hits, {(long)rec.hits.size() / 3, 3L});
    std::printf("  %s: %ld vertices, %ld triangles, %ld links, %ld pairs, %ld hits\n",
                name.c_str(), nv, nt, nl, np, (long)rec.hits.size() / 3);
}

void dirac_case()
{
    refio::ensure_dir("fsi/dirac");
    refio::Rng rng(99);
    const long n = 2001;
    std::vector<T> x(n), roma(n), peskin(n), tophat(n);
    for (long k = 0; k < n; ++k) {
        x[k] = -2.0 + 4.0 * k / (n - 1);
        roma[k] = std::fabs(x[k]) <= 1.5 ? RomaDirac<T, 3>::eval(x[k]) : 0.;
        peskin[k] = PeskinDirac<T, 3>::eval(x[k]);
        tophat[k] = std::fabs(x[k]) <= 1.0 ? TopHatDirac<T, 3>::eval(x[k]) : 0.;
    }
    refio::write_npy("fsi/dirac/x.npy", x, {n});
    refio::write_npy("fsi/dirac/roma.npy", roma, {n});
    refio::write_npy("fsi/dirac/peskin.npy", peskin, {n});
    refio::write_npy("fsi/dirac/tophat.npy", tophat, {n});

    // Support boxes for random points, including exact integers and
    // half-integers (the ceil() edge cases).
    const long m = 3000;
    std::vector<T> p(3 * m);
    std::vector<long long> boxes(3 * 3 * m);
    for (long k = 0; k < m; ++k) {
        for (int d = 0; d < 3; ++d) {
            T v = rng.uniform(-3., 40.);
            if (k % 3 == 1) v = std::floor(v);
            if (k % 3 == 2) v = std::floor(v) + 0.5;
            p[3 * k + d] = v;
        }
        Array<T, 3> pos(p[3 * k], p[3 * k + 1], p[3 * k + 2]);
        Box3D b;
        get_dirac_compact_support_box<T, RomaDirac<T, 3> >(pos, b);
        boxes[9 * k + 0] = b.x0; boxes[9 * k + 1] = b.y0; boxes[9 * k + 2] = b.z0;
        get_dirac_compact_support_box<T, PeskinDirac<T, 3> >(pos, b);
        boxes[9 * k + 3] = b.x0; boxes[9 * k + 4] = b.y0; boxes[9 * k + 5] = b.z0;
        get_dirac_compact_support_box<T, TopHatDirac<T, 3> >(pos, b);
        boxes[9 * k + 6] = b.x0; boxes[9 * k + 7] = b.y0; boxes[9 * k + 8] = b.z0;
    }
    refio::write_npy("fsi/dirac/points.npy", p, {m, 3L});
    refio::write_npy("fsi/dirac/box_base.npy", boxes, {m, 3L, 3L});
}

}  // namespace

int main(int argc, char** argv)
{
    plbInit(&argc, &argv);
    const std::string ref = argc > 1 ? argv[1] : "/root/reference";
    refio::ensure_dir("fsi");
    ParticleShapeLibrary<T> lib;
    mesh_case(ref, "rbc_in_pipe/sphere_TRImesh_626n_1248el.msh", "sphere626", lib, 21);
    mesh_case(ref, "capsule_in_shear/ellipsoid_12_3_3_TRImesh_414n_824el.msh", "ellipsoid414", lib, 22);
    dirac_case();
    std::printf("ref_fsi: done\n");
    return 0;
}
