// Golden files for the fsi on-disk formats, written by the reference's own
// classes (the parts of src/fsi that build without MPI):
//   particle pack layouts        src/fsi/ParticleBase.hh:218-224,
//                                src/fsi/DeformableParticle3D.hh:277-287,
//                                src/fsi/DeformableCapsule.hh:217-223,
//                                src/fsi/RBCParticle.hh:95-105,346-350
//   VTU particle mesh            src/fsi/ParticleBase.hh:270-398
//   gmsh particle mesh           src/fsi/ParticleBase.hh:400-437
//   lightweight text             src/fsi/DeformableParticle3D.hh:318-333,
//                                src/fsi/ImmersedBoundaryDynamics_io.hh:178-200
//   checkpoint header names      src/fsi/ImmersedBoundaryDynamics_io.hh:17-73
//                                (FileName composition; the writer itself is
//                                in the MPI-only ImmersedBoundaryDynamics3D)
// The particles sit on the 626- and 414-vertex meshes at the deformed
// positions of the fsi goldens (tests/golden/fsi/<mesh>/pos.npy, re-created
// here from the same seed), with seeded vertex velocities and the forces of
// their own compute_forces.
//
// Usage: ref_io [reference_root]   (writes ./io/...)
#include "palabos3D.h"
#include "palabos3D.hh"
#include <cstring>
#include <cstdio>
#include <fstream>
#include "fsi/ParticleBase.h"
#include "fsi/ParticleShape.h"
#include "fsi/ParticleShapeFactory.h"
#include "fsi/DeformableParticle3D.h"
#include "fsi/RBCParticle.h"
#include "fsi/DeformableCapsule.h"
#include "fsi/ParticleFactory.h"
#include "fsi/ParticleBase.hh"
#include "fsi/RBCParticle.hh"
#include "fsi/ParticleShape.hh"
#include "fsi/ParticleShapeFactory.hh"
#include "fsi/DeformableParticle3D.hh"
#include "fsi/ParticleFactory.hh"
#include "fsi/DeformableCapsule.hh"
#include "common.h"

using namespace plb;
using namespace plb::fsi;
typedef double T;

namespace {

void write_bytes(const std::string& path, const std::vector<char>& b)
{
    std::ofstream out(path.c_str(), std::ios::binary);
    out.write(b.empty() ? (const char*)0 : &b[0], b.size());
}

template <class P>
void set_state(P& p, const std::vector<T>& pos, const std::vector<T>& vel)
{
    for (plint i = 0; i < p.count_nodes(); ++i)
        for (int d = 0; d < 3; ++d) {
            p.get_node(i).pos[d] = pos[3 * i + d];
            p.get_node(i).vel[d] = vel[3 * i + d];
        }
    p.update();
    p.reset_forces();
    p.compute_forces();
}

void mesh_case(const std::string& ref, const std::string& mesh, const std::string& name,
               ParticleShapeLibrary<T>& lib, unsigned long long seed)
{
    const std::string dir = "io/" + name;
    refio::ensure_dir(dir);
    lib.read_and_store_mesh(ref + "/progs/" + mesh, name);
    ParticleShape<T>* shape = lib.get_by_tag(name);
    const long nv = shape->count_vertices();

    // the deformed positions of ref_fsi (same seed, same recipe)
    refio::Rng rng(seed);
    const T A[3][3] = {{1.12, 0.05, 0.0}, {0.0, 0.93, 0.02}, {0.01, 0.0, 0.97}};
    const T shift[3] = {20.3, 17.1, 25.6};
    const T lm = shape->get_triangle(0).l_mean;
    Array<T, 3> c = shape->get_center();
    std::vector<T> pos(3 * nv), vel(3 * nv);
    for (long i = 0; i < nv; ++i) {
        T r[3];
        for (int d = 0; d < 3; ++d) r[d] = shape->get_vertex(i)[d] - c[d];
        for (int d = 0; d < 3; ++d)
            pos[3 * i + d] = c[d] + A[d][0] * r[0] + A[d][1] * r[1] + A[d][2] * r[2] +
                             0.03 * lm * rng.uniform(-1., 1.) + shift[d];
    }
    refio::Rng rv(seed + 100);
    for (long k = 0; k < 3 * nv; ++k) vel[k] = 1e-3 * rv.uniform(-1., 1.);
    refio::write_npy(dir + "/pos.npy", pos, {nv, 3L});
    refio::write_npy(dir + "/vel.npy", vel, {nv, 3L});

    // NeoHookean capsule (the class default shear model), id 3
    CapsuleParameters<T> cp;
    cp.G = 0.02; cp.C = 10.; cp.nu_s = 0.5; cp.kb = 0.001;
    DeformableCapsuleParticle<T> neo(shape, cp);
    neo.set_id(3);
    set_state(neo, pos, vel);
    std::vector<char> buf;
    neo.pack(buf);
    write_bytes(dir + "/pack_neohookean.bin", buf);

    // Skalak capsule: the same particle with the shear-model field (after the
    // header and the vertex block) switched, re-created through the factory
    std::vector<char> sk_buf = buf;
    const size_t off = 3 * sizeof(plint) + sizeof(bool) + nv * 9 * sizeof(T);
    int skalak = 0;  // enum ShearModel {Skalak, NeoHookean}
    std::memcpy(&sk_buf[off], &skalak, sizeof(int));
    char* it = &sk_buf[0];
    ParticleBase3D<T>* sk = particleFactory<T>().create(it, lib);
    sk->set_id(4);
    DeformableCapsuleParticle<T>* skp = dynamic_cast<DeformableCapsuleParticle<T>*>(sk);
    set_state(*skp, pos, vel);
    buf.clear();
    skp->pack(buf);
    write_bytes(dir + "/pack_skalak.bin", buf);
    delete sk;

    // RBC, id 5
    RBCParameters<T> rp;
    rp.shear_modulus = 0.015; rp.L0 = lm; rp.k_area_global = 0.6; rp.k_area_local = 0.05;
    rp.k_volume = 0.8; rp.vol_desired = 0.9 * shape->get_volume(); rp.k_bend = 0.012; rp.theta0 = 0.;
    RBCParticle<T> rbc(shape, rp);
    rbc.set_id(5);
    set_state(rbc, pos, vel);
    buf.clear();
    rbc.pack(buf);
    write_bytes(dir + "/pack_rbc.bin", buf);
    T rbcp[8] = {rp.shear_modulus, rp.L0, rp.k_area_global, rp.k_area_local,
                 rp.k_volume, rp.vol_desired, rp.k_bend, rp.theta0};
    refio::write_npy(dir + "/rbc_params.npy", rbcp, {8L});

    // mesh writers and the lightweight line of the RBC
    rbc.write_vtk(FileName(dir + "/rbc.vtu"));
    rbc.write_gmsh(FileName(dir + "/rbc.msh"));
    {
        std::ofstream out((dir + "/lightweight.txt").c_str());
        out << "iteration:" << 120 << std::endl;
        neo.write_lightweight(out);
        out << std::endl;
        rbc.write_lightweight(out);
        out << std::endl;
    }
    long long ids[2] = {(long long)DeformableCapsuleParticle<T>::create(shape)->get_type_id(),
                        (long long)RBCParticle<T>::create(shape)->get_type_id()};
    refio::write_npy(dir + "/type_ids.npy", ids, {2L});
    std::printf("  %s: %ld vertices, capsule type %lld, rbc type %lld, shape id %ld\n", name.c_str(), nv, ids[0],
                ids[1], (long)shape->get_id());
}

// Names of the checkpoint header and of the per-rank file, composed with
// FileName exactly as save_checkpoint composes them
// (src/fsi/ImmersedBoundaryDynamics_io.hh:21-37).
void header_case()
{
    refio::ensure_dir("io/checkpoint");
    FileName header_file("checkpoint_fsi");
    std::string base_folder = "out/";
    header_file.setPath(base_folder);
    header_file.setExt("plb");
    FileName proc_file;
    proc_file.setPath(header_file.getName() + "/");
    proc_file.setExt("plb");
    FileName fn = proc_file;
    fn.setName(util::val2str(0));
    std::ofstream out("io/checkpoint/names.txt");
    out << header_file.get() << std::endl;
    out << fn.get() << std::endl;
    FileName local = fn;
    local.setPath(base_folder + local.getPath());
    out << local.get() << std::endl;
}

}  // namespace

int main(int argc, char** argv)
{
    plbInit(&argc, &argv);
    const std::string ref = argc > 1 ? argv[1] : "/root/reference";
    refio::ensure_dir("io");
    ParticleShapeLibrary<T> lib;
    mesh_case(ref, "rbc_in_pipe/sphere_TRImesh_626n_1248el.msh", "sphere626", lib, 21);
    mesh_case(ref, "capsule_in_shear/ellipsoid_12_3_3_TRImesh_414n_824el.msh", "ellipsoid414", lib, 22);
    header_case();
    std::printf("ref_io: done\n");
    return 0;
}
