// Golden files for the initial-condition pipeline, written by the
// reference's own code (header-only parts that build without MPI):
//   create_rbc_params          progs/common/calibrate_rbc_shape.h:75-112
//   shrink_rbc_volume          progs/common/calibrate_rbc_shape.h:13-72
//   RBCParticle::relax_nodes   src/fsi/RBCParticle.hh:318-342
//   ParticlePositionInitializer src/fsi/ParticlePositionInitializer.hh:18-86
//
// Usage: ref_init [reference_root]   (writes ./init/...)
#include "palabos3D.h"
#include "palabos3D.hh"
#include <cstdio>
#include <cstdlib>
#include "fsi/ParticleBase.h"
#include "fsi/ParticleShape.h"
#include "fsi/ParticleShapeFactory.h"
#include "fsi/DeformableParticle3D.h"
#include "fsi/RBCParticle.h"
#include "fsi/Boundary.h"
#include "fsi/ParticlePositionInitializer.h"
#include "fsi/ParticleBase.hh"
#include "fsi/RBCParticle.hh"
#include "fsi/ParticleShape.hh"
#include "fsi/ParticleShapeFactory.hh"
#include "fsi/DeformableParticle3D.hh"
#include "fsi/ParticlePositionInitializer.hh"
#include "fsi/Grid.hh"
#include "fsi/ParticleFactory.h"
#include "fsi/ParticleFactory.hh"
#include "../progs/common/calibrate_rbc_shape.h"
#include "common.h"

using namespace plb;
using namespace plb::fsi;
typedef double T;

namespace {

void write_state(const std::string& dir, const RBCParticle<T>& rbc)
{
    const long nv = rbc.count_nodes();
    std::vector<T> pos(3 * nv), vel(3 * nv), frc(3 * nv);
    for (long i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) {
            pos[3 * i + d] = rbc.get_node(i).pos[d];
            vel[3 * i + d] = rbc.get_node(i).vel[d];
            frc[3 * i + d] = rbc.get_node(i).force[d];
        }
    refio::write_npy(dir + "/pos.npy", pos, {nv, 3L});
    refio::write_npy(dir + "/vel.npy", vel, {nv, 3L});
    refio::write_npy(dir + "/force.npy", frc, {nv, 3L});
    T g[2] = {rbc.area(), rbc.volume()};
    refio::write_npy(dir + "/area_volume.npy", g, {2L});
}

// One calibration: create_rbc_params(shape, G, K_area, K_bend), then
// shrink_rbc_volume(rbc, ratio * V0, n_it, 1, damping).
void calibration_case(const std::string& name, ParticleShapeLibrary<T>& lib, const std::string& tag, T G,
                      T K_area, T K_bend, T ratio, int n_it, T damping)
{
    const std::string dir = "init/" + name;
    refio::ensure_dir(dir);
    ParticleShape<T>* shape = lib.get_by_tag(tag);
    RBCParameters<T> rp = create_rbc_params<T>(shape, G, K_area, K_bend);
    T prm[8] = {rp.shear_modulus, rp.L0, rp.k_area_global, rp.k_area_local,
                rp.k_volume, rp.vol_desired, rp.k_bend, rp.theta0};
    refio::write_npy(dir + "/rbc_params.npy", prm, {8L});
    T args[5] = {ratio * shape->get_volume(), (T)n_it, damping, G, K_area};
    refio::write_npy(dir + "/args.npy", args, {5L});
    RBCParticle<T> rbc(shape, rp);
    shrink_rbc_volume(rbc, ratio * shape->get_volume(), n_it, (T)1, damping);
    write_state(dir, rbc);
}

void placement_case()
{
    refio::ensure_dir("init/placement");
    std::srand(5);
    const Box3D bb(0, 63, 0, 31, 0, 31);
    PipeBoundary<T> pipe(15.5, 15.5, 14.0, 64.0);
    ParticlePositionInitializer<T> init(bb, &pipe, 3.2);
    const bool okR = init.generate_points(40, 3.2, 'R');
    const bool okP = init.generate_points(30, 1.7, 'P');
    std::vector<Array<T, 3> > r, p;
    init.get_points('R', r);
    init.get_points('P', p);
    std::vector<T> out;
    for (size_t k = 0; k < r.size(); ++k) out.insert(out.end(), {r[k][0], r[k][1], r[k][2]});
    refio::write_npy("init/placement/R.npy", out, {(long)r.size(), 3L});
    out.clear();
    for (size_t k = 0; k < p.size(); ++k) out.insert(out.end(), {p[k][0], p[k][1], p[k][2]});
    refio::write_npy("init/placement/P.npy", out, {(long)p.size(), 3L});
    // the next draw of the generator after the placement
    long long next[1] = {(long long)std::rand()};
    refio::write_npy("init/placement/next_rand.npy", next, {1L});
    std::printf("  placement: %zu + %zu points (%d %d)\n", r.size(), p.size(), (int)okR, (int)okP);
}

}  // namespace

int main(int argc, char** argv)
{
    plbInit(&argc, &argv);
    const std::string ref = argc > 1 ? argv[1] : "/root/reference";
    refio::ensure_dir("init");
    ParticleShapeLibrary<T> lib;
    lib.read_and_store_mesh(ref + "/progs/rbc_in_pipe/sphere_TRImesh_626n_1248el.msh", "sphere626");
    calibration_case("shrink626", lib, "sphere626", 0.02, 0.4, 0.002, 0.8, 40, 1e-2);
    placement_case();
    std::printf("ref_init: done\n");
    return 0;
}
