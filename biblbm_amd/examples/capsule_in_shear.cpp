// A capsule in simple shear flow, written against the C++ layer of
// include/biblbm_plb.h in the structure of the reference prog
// progs/capsule_in_shear/capsule_in_shear.cpp: a lattice with the Guo-VOF
// dynamics, periodic in x and z, interp velocity boundaries on the y faces
// moving at -/+ U, a linear shear start, one Skalak capsule read from a gmsh
// mesh, and the coupled loop computeVelocity -> setExternalVector ->
// applyProcessingFunctional(wrap_ibm_dynamics3D(fsi)) -> collideAndStream.
//
// usage: capsule_in_shear mesh.msh N steps Ca radius out_prefix [storage]
//   N         lattice side (N^3 cells)
//   steps     coupled steps
//   Ca        capillary number G = mu * gamma * radius / Ca
//   storage   f64 (default), f32 or f16s
// Writes <out_prefix>_deformation.txt (step, volume, Taylor deformation
// index every 10 steps) and <out_prefix>_final.txt (vertex positions).
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

// Taylor deformation index from the inertia tensor of the enclosed volume:
// the ellipsoid with that tensor has semi-axes a_k^2 ~ (I1 + I2 + I3)/2 - I_k.
static T taylor_deformation(const T I[9])
{
    // eigenvalues of a symmetric 3x3 matrix (closed form)
    const T p1 = I[1] * I[1] + I[2] * I[2] + I[5] * I[5];
    const T q = (I[0] + I[4] + I[8]) / 3;
    const T p2 = (I[0] - q) * (I[0] - q) + (I[4] - q) * (I[4] - q) + (I[8] - q) * (I[8] - q) + 2 * p1;
    const T p = std::sqrt(p2 / 6);
    T e[3] = {q, q, q};
    if (p > 0) {
        const T B[9] = {(I[0] - q) / p, I[1] / p, I[2] / p, I[3] / p, (I[4] - q) / p, I[5] / p,
                        I[6] / p, I[7] / p, (I[8] - q) / p};
        const T detB = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                       B[2] * (B[3] * B[7] - B[4] * B[6]);
        const T r = std::max(T(-1), std::min(T(1), detB / 2));
        const T phi = std::acos(r) / 3;
        e[0] = q + 2 * p * std::cos(phi);
        e[2] = q + 2 * p * std::cos(phi + 2 * M_PI / 3);
        e[1] = 3 * q - e[0] - e[2];
    }
    const T s = (e[0] + e[1] + e[2]) / 2;
    T amax = 0, amin = 1e300;
    for (int k = 0; k < 3; ++k) {
        const T a = std::sqrt(std::max(T(0), s - e[k]));
        amax = std::max(amax, a);
        amin = std::min(amin, a);
    }
    return (amax - amin) / (amax + amin);
}

int main(int argc, char** argv)
{
    if (argc < 7) {
        std::cerr << "usage: capsule_in_shear mesh.msh N steps Ca radius out_prefix [f64|f32|f16s]\n";
        return 2;
    }
    try {
        const std::string mesh = argv[1];
        const plint N = std::atol(argv[2]);
        const plint Nt = std::atol(argv[3]);
        const T Ca = std::atof(argv[4]);
        const T radius = std::atof(argv[5]);
        const std::string out = argv[6];
        const std::string storage = argc > 7 ? argv[7] : "f64";

        blCheck(bl_init(0), "bl_init");
        const T tau = 1.0, U = 0.01;
        const T omega = 1 / tau;
        const T nu_lb = (tau - 0.5) / 3;
        const T gamma_lb = 2 * U / (N - 1);

        ParticleShapeLibrary<T> shape_library;
        shape_library.read_and_store_mesh(mesh, "capsule");
        CapsuleParameters<T> params;
        params.G = nu_lb * gamma_lb * radius / Ca;
        params.C = 10;
        DeformableCapsuleParticle<T> capsule(shape_library.get_by_tag("capsule"), params);
        capsule.set_shear_model(DeformableCapsuleParticle<T>::Skalak);
        capsule.set_center_of_mass(Array<T, 3>(N / 2.0, (N - 1) / 2.0, N / 2.0));

        MultiBlockLattice3D<T, DESCRIPTOR> lattice(N, N, N, new GuoExternalForceVOFDynamics<T, DESCRIPTOR>(omega, omega));
        if (storage == "f32") lattice.setStorage(BL_STORE_F32);
        if (storage == "f16s") lattice.setStorage(BL_STORE_F16S);
        const Box3D domain = lattice.getBoundingBox();
        lattice.periodicity().toggle(0, true);
        lattice.periodicity().toggle(1, false);
        lattice.periodicity().toggle(2, true);

        // shear flow setup (progs/common/shear_flow_setup.h): velocity
        // boundaries on the y faces, linear profile between them
        std::unique_ptr<OnLatticeBoundaryCondition3D<T, DESCRIPTOR>> bc(createInterpBoundaryCondition3D<T, DESCRIPTOR>());
        const Box3D bottom(0, N - 1, 0, 0, 0, N - 1), top(0, N - 1, N - 1, N - 1, 0, N - 1);
        bc->addVelocityBoundary1N(bottom, lattice);
        bc->addVelocityBoundary1P(top, lattice);
        setBoundaryVelocity(lattice, bottom, Array<T, 3>(-U, 0, 0));
        setBoundaryVelocity(lattice, top, Array<T, 3>(U, 0, 0));
        initializeAtEquilibrium(lattice, domain, [&](plint, plint iY, plint, T& rho, Array<T, 3>& u) {
            rho = 1;
            u = Array<T, 3>(U * (2.0 * iY / (N - 1) - 1), 0, 0);
        });
        lattice.initialize();

        ImmersedBoundaryDynamics3D<T, DESCRIPTOR> fsi(lattice, shape_library, BL_KERNEL_PESKIN);
        fsi.add_particle(&capsule);
        fsi.init();
        std::unique_ptr<MultiTensorField3D<T, 3>> velocity = generateMultiTensorField<T, 3>(domain);

        std::ofstream fdef((out + "_deformation.txt").c_str());
        fdef.precision(17);
        std::cout << "capsule_in_shear: " << N << "^3, " << Nt << " steps, Ca " << Ca << ", G " << params.G
                  << ", storage " << storage << std::endl;
        for (plint it = 0; it <= Nt; ++it) {
            if (it % 10 == 0) {
                T area, vol, I[9];
                Array<T, 3> com;
                fsi.observables(0, area, vol, com, I);
                fdef << it << " " << vol << " " << taylor_deformation(I) << "\n";
            }
            if (it == Nt) break;
            setExternalVector(lattice, domain, DESCRIPTOR<T>::ExternalField::forceBeginsAt, Array<T, 3>(0, 0, 0));
            applyProcessingFunctional(wrap_ibm_dynamics3D(fsi), domain, lattice, *velocity);
            lattice.collideAndStream();
        }
        std::vector<Array<T, 3>> pos;
        fsi.get_particle(0, pos);
        std::ofstream fpos((out + "_final.txt").c_str());
        fpos.precision(17);
        for (const auto& x : pos) fpos << x[0] << " " << x[1] << " " << x[2] << "\n";
        computeVelocity(lattice, *velocity, domain);
        const Array<T, 3> uc = velocity->get(N / 2, (N - 1) / 2, N / 2);
        std::cout << "done: centre velocity " << uc[0] << " " << uc[1] << " " << uc[2]
                  << ", total particle volume " << fsi.compute_total_volume() << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "capsule_in_shear: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
