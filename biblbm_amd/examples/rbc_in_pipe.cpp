// A red blood cell in pressure-driven pipe flow, written against the C++
// layer of include/biblbm_plb.h in the structure of the reference prog
// progs/rbc_in_pipe/rbc_in_pipe.cpp (setup_flow, :286-330, and the main
// loop, :748-790): a lattice sized by the pipe's bounding box, periodic
// along the pipe, the Guo curved wall on the pipe surface (or bounce-back
// on the cells outside it), no dynamics beyond the boundary, a body force,
// one RBC read from a gmsh mesh, and the loop computeVelocity ->
// setExternalVector -> applyProcessingFunctional(wrap_ibm_dynamics3D(fsi))
// -> collideAndStream.
//
// usage: rbc_in_pipe mesh.msh R L steps out_prefix [second_order] [ibm_boundary] [storage] [interactions]
//                    [checkpoint_interval] [start_from_checkpoint]
//   R, L          pipe radius and length (lattice units)
//   second_order  1 (default): Guo curved wall; 0: bounce-back surface
//   ibm_boundary  1: the immersed boundary blends and reflects at the pipe
//                 wall (set_boundary, as the rbc_in_pipe_distribution prog
//                 does); 0 (default): not, as rbc_in_pipe itself
//   storage       f64 (default), f32 or f16s
//   interactions  1: the prog's Morse wall interaction (r_min 2, beta 2,
//                 strength 1e-4) and Morse particle-particle force (r_min 1,
//                 beta 1, strength 1e-3), the values of its input.xml
//                 (read_parameters, :150-178, and main, :735-741); 0 (default): none
//   checkpoint_interval   > 0: at every multiple of it, write_checkpoint
//                 (:686-696: parallelIO::save of the lattice to
//                 checkpoint_lbm, fsi.save_checkpoint("checkpoint_fsi"),
//                 it.write_checkpoint("checkpoint_time")) and the lightweight
//                 particle data, in the current folder; 0 (default): none
//   start_from_checkpoint 1: resume from those files (setup_flow :320-322,
//                 create_particles :348-351, create_time :677-684)
// Writes <out_prefix>_final.txt (vertex positions) and
// <out_prefix>_profile.txt (u_x across the pipe at mid-length).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "biblbm_plb.h"

using namespace plb;
using namespace plb::fsi;
typedef double T;
#define DESCRIPTOR descriptors::ForcedPhaseD3Q19Descriptor

// the prog's BoundarySurface: every cell outside the boundary
class BoundarySurface {
public:
    explicit BoundarySurface(const Boundary<T>& b) : b_(b) {}
    bool operator()(plint x, plint y, plint z) const { return !b_.contains(Array<T, 3>(x, y, z)); }

private:
    const Boundary<T>& b_;
};

int main(int argc, char** argv)
{
    if (argc < 6) {
        std::cerr << "usage: rbc_in_pipe mesh.msh R L steps out_prefix [second_order] [ibm_boundary] [storage] "
                     "[interactions]\n";
        return 2;
    }
    try {
        const std::string mesh = argv[1];
        const T R = std::atof(argv[2]);
        const T length = std::atof(argv[3]);
        const plint Nt = std::atol(argv[4]);
        const std::string out = argv[5];
        const bool second_order = argc > 6 ? std::atoi(argv[6]) != 0 : true;
        const bool ibm_boundary = argc > 7 ? std::atoi(argv[7]) != 0 : false;
        const std::string storage = argc > 8 ? argv[8] : "f64";
        const bool interactions = argc > 9 ? std::atoi(argv[9]) != 0 : false;
        const plint checkpoint_interval = argc > 10 ? std::atol(argv[10]) : 0;
        const bool load_checkpoint = argc > 11 ? std::atoi(argv[11]) != 0 : false;

        blCheck(bl_init(0), "bl_init");
        const T tau = 1.0, u_max = 0.02;
        const T omega = 1 / tau;
        const T nu = (tau - 0.5) / 3;
        const plint margin = 2;
        PipeBoundary<T> pipe_boundary(R + margin, R + margin, R, length);
        const Box3D box = pipe_boundary.get_bounding_box(margin);
        // Poiseuille: u_max = g R^2 / (4 nu)
        const Array<T, 3> external_force(4 * nu * u_max / (R * R), 0, 0);

        ParticleShapeLibrary<T> shape_library;
        shape_library.read_and_store_mesh(mesh, "rbc");
        RBCParameters<T> p;
        p.shear_modulus = 0.0015;
        p.L0 = 1.0;
        p.k_area_global = 0.5;
        p.k_area_local = 0.005;
        p.k_volume = 0.5;
        p.k_bend = 0.0005;
        p.theta0 = 0.0;
        const ParticleShape<T>* shape = shape_library.get_by_tag("rbc");
        const T vol = [&] {
            T V = 0;
            for (const auto& t : shape->triangles) {
                const Array<T, 3>&a = shape->vertices[t[0]], &b = shape->vertices[t[1]], &c = shape->vertices[t[2]];
                V += (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                      a[2] * (b[0] * c[1] - b[1] * c[0])) / 6;
            }
            return V;
        }();
        p.vol_desired = vol;
        RBCParticle<T> rbc(shape, p);
        rbc.set_center_of_mass(Array<T, 3>(length / 2, R + margin, R + margin));

        MultiBlockLattice3D<T, DESCRIPTOR> lattice(box.x1 - box.x0 + 1, box.y1 - box.y0 + 1, box.z1 - box.z0 + 1,
                                                   new GuoExternalForceVOFDynamics<T, DESCRIPTOR>(omega, omega));
        if (storage == "f32") lattice.setStorage(BL_STORE_F32);
        if (storage == "f16s") lattice.setStorage(BL_STORE_F16S);
        const Box3D domain = lattice.getBoundingBox();
        lattice.periodicity().toggle(0, true);
        lattice.periodicity().toggle(1, false);
        lattice.periodicity().toggle(2, false);

        std::unique_ptr<GuoRigidWallBoundary<T, DESCRIPTOR>> boundary_condition;
        if (second_order) {
            boundary_condition.reset(new GuoRigidWallBoundary<T, DESCRIPTOR>(pipe_boundary, lattice));
            boundary_condition->insert();
        } else {
            defineDynamics<T, DESCRIPTOR>(lattice, domain, new BoundarySurface(pipe_boundary),
                                          new BounceBack<T, DESCRIPTOR>());
        }
        defineDynamics<T, DESCRIPTOR>(lattice, domain, new BoundaryOutside<T>(pipe_boundary, std::sqrt(2)),
                                      new NoDynamics<T, DESCRIPTOR>());
        setExternalVector(lattice, domain, DESCRIPTOR<T>::ExternalField::forceBeginsAt, external_force);
        initializeAtEquilibrium(lattice, domain, [](plint, plint, plint, T& rho, Array<T, 3>& u) {
            rho = 1;
            u = Array<T, 3>(0, 0, 0);
        });
        lattice.initialize();
        if (load_checkpoint) parallelIO::load("checkpoint_lbm", lattice, false);

        ImmersedBoundaryDynamics3D<T, DESCRIPTOR> fsi(lattice, shape_library, BL_KERNEL_PESKIN);
        if (ibm_boundary) fsi.set_boundary(&pipe_boundary);
        WallInteraction<T, MorsePotential<T>> wall_interaction(pipe_boundary, MorsePotential<T>(2., 2., 1e-4));
        PPPotentialForce<T, MorsePotential<T>> pp_interaction(MorsePotential<T>(1., 1., 1e-3));
        if (interactions) {
            fsi.add_force_decorator(&wall_interaction);
            fsi.add_pp_force(&pp_interaction);
        }
        if (load_checkpoint) fsi.load_checkpoint("checkpoint_fsi");
        else fsi.add_particle(&rbc);
        fsi.init();
        Time it(0);
        if (load_checkpoint) it.read_checkpoint("checkpoint_time");
        std::unique_ptr<MultiTensorField3D<T, 3>> velocity = generateMultiTensorField<T, 3>(domain);

        std::cout << "rbc_in_pipe: " << lattice.getNx() << "x" << lattice.getNy() << "x" << lattice.getNz()
                  << ", R " << R << ", " << Nt << " steps, " << (second_order ? "Guo curved wall" : "bounce-back")
                  << (boundary_condition ? ", " + std::to_string(std::distance(boundary_condition->begin(),
                                                                               boundary_condition->end())) +
                                               " wall nodes"
                                         : std::string())
                  << ", storage " << storage << std::endl;
        for (; it < Nt; ++it) {
            if (checkpoint_interval > 0 && it.is_multiple_of(checkpoint_interval)) {
                parallelIO::save(lattice, "checkpoint_lbm", false);
                fsi.save_checkpoint("checkpoint_fsi");
                it.write_checkpoint("checkpoint_time");
                fsi.write_lightweight_particle_data(it.to_plint());
            }
            setExternalVector(lattice, domain, DESCRIPTOR<T>::ExternalField::forceBeginsAt, external_force);
            applyProcessingFunctional(wrap_ibm_dynamics3D(fsi), domain, lattice, *velocity);
            lattice.collideAndStream();
        }
        std::vector<Array<T, 3>> pos;
        fsi.get_particle(0, pos);
        std::ofstream fpos((out + "_final.txt").c_str());
        fpos.precision(17);
        for (const auto& x : pos) fpos << x[0] << " " << x[1] << " " << x[2] << "\n";
        computeVelocity(lattice, *velocity, domain);
        std::ofstream fprof((out + "_profile.txt").c_str());
        fprof.precision(17);
        const plint xm = lattice.getNx() / 2, zc = (plint)(R + margin);
        for (plint y = 0; y < lattice.getNy(); ++y) fprof << y << " " << velocity->get(xm, y, zc)[0] << "\n";
        std::cout << "done: total particle volume " << fsi.compute_total_volume() << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "rbc_in_pipe: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
