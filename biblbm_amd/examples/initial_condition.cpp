// An initial red-cell distribution, written against the C++ layer of
// include/biblbm_plb.h in the structure of the reference prog
// progs/generate_initial_distribution/initial_condition.cpp: read the mesh,
// create_rbc_params, shrink_rbc_volume (on the device), then
// ParticlePositionInitializer places spheres enclosing the calibrated cell
// in a pipe, and the cells go into an immersed boundary whose checkpoint
// (save_checkpoint("checkpoint_fsi"), in the current folder) a flow prog
// can start from.  The reference prog's rigid-growth packing that follows
// is not part of this build.
//
// usage: initial_condition mesh.msh G K_area K_bend ratio n_it damping seed n_cells out_prefix
//   G, K_area, K_bend  create_rbc_params(shape, G, K_area, K_bend)
//   ratio, n_it, damping  shrink_rbc_volume(rbc, ratio * V0, n_it, 1, damping)
//   seed               srand(seed) before the placement
//   n_cells            cells placed in the pipe y0 = z0 = 15.5, radius 14,
//                      in the box (0, 63, 0, 31, 0, 31), sphere radius 3.2
//                      (then 30 platelet spheres of radius 1.7, tag 'P')
// Writes <out_prefix>_shape.txt (calibrated vertices), <out_prefix>_R.txt and
// <out_prefix>_P.txt (sphere centres).
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "biblbm_calibrate.h"

using namespace plb;
using namespace plb::fsi;
typedef double T;
#define DESCRIPTOR descriptors::ForcedPhaseD3Q19Descriptor

static void write_points(const std::string& file, const std::vector<Array<T, 3>>& pts)
{
    std::ofstream out(file.c_str());
    out.precision(17);
    for (const auto& p : pts) out << p[0] << " " << p[1] << " " << p[2] << "\n";
}

int main(int argc, char** argv)
{
    if (argc < 11) {
        std::cerr << "usage: initial_condition mesh.msh G K_area K_bend ratio n_it damping seed n_cells out_prefix\n";
        return 2;
    }
    try {
        const std::string mesh = argv[1];
        const T G = std::atof(argv[2]), K_area = std::atof(argv[3]), K_bend = std::atof(argv[4]);
        const T ratio = std::atof(argv[5]);
        const int n_it = std::atoi(argv[6]);
        const T damping = std::atof(argv[7]);
        const unsigned seed = (unsigned)std::atoi(argv[8]);
        const plint n_cells = std::atol(argv[9]);
        const std::string out = argv[10];
        blCheck(bl_init(0), "bl_init");

        ParticleShapeLibrary<T> shape_library;
        shape_library.read_and_store_mesh(mesh, "RBC");
        const ParticleShape<T>* rbc_shape = shape_library.get_by_tag("RBC");

        // Create RBC and calibrate shape
        RBCParticle<T> rbc(rbc_shape, create_rbc_params(rbc_shape, G, K_area, K_bend));
        T V0 = rbc.params_ref().vol_desired;
        shrink_rbc_volume(rbc, ratio * V0, n_it, (T)1, damping);
        write_points(out + "_shape.txt", rbc.positions());

        // Generate initial distribution
        PipeBoundary<T> boundary(15.5, 15.5, 14.0, 64.0);
        std::srand(seed);
        ParticlePositionInitializer<T> position_initializer(Box3D(0, 63, 0, 31, 0, 31), &boundary, 3.2);
        const bool okR = position_initializer.generate_points(n_cells, 3.2, 'R');
        const bool okP = position_initializer.generate_points(30, 1.7, 'P');
        std::vector<Array<T, 3>> R, P;
        position_initializer.get_points('R', R);
        position_initializer.get_points('P', P);
        write_points(out + "_R.txt", R);
        write_points(out + "_P.txt", P);

        // Calibrated cells at the sphere centres, into an immersed boundary
        // whose checkpoint seeds a flow run
        MultiBlockLattice3D<T, DESCRIPTOR> lattice(64, 32, 32, new GuoExternalForceVOFDynamics<T, DESCRIPTOR>(1.0, 1.0));
        lattice.periodicity().toggle(0, true);
        ImmersedBoundaryDynamics3D<T, DESCRIPTOR> fsi(lattice, shape_library, BL_KERNEL_PESKIN);
        std::vector<std::unique_ptr<RBCParticle<T>>> cells;
        for (const auto& c : R) {
            cells.emplace_back(new RBCParticle<T>(rbc));
            cells.back()->set_center_of_mass(c);
            fsi.add_particle(cells.back().get());
        }
        fsi.init();
        fsi.save_checkpoint("checkpoint_fsi");
        std::cout << "initial_condition: calibrated volume ratio " << ratio << ", placed " << R.size() << " cells ("
                  << (okR ? "all" : "not all") << ") and " << P.size() << " platelets (" << (okP ? "all" : "not all")
                  << "), " << fsi.count_particles() << " particles in the checkpoint" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "initial_condition: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
