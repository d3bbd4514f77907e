// Times the reference's own lattice passes of one coupled step, compiled
// from the reference sources: the loop of progs/capsule_in_shear/
// capsule_in_shear.cpp:187-227 without the immersed-boundary functional
// (ImmersedBoundaryDynamics3D needs MPI and is not built here), i.e.
//   computeVelocity(lattice, velocity, domain)        (:191)
//   setExternalVector(lattice, domain, force, 0)      (:205)
//   lattice.collideAndStream()                        (:223)
// with GuoExternalForceVOFDynamics (src/fsi/VariableViscosityDynamics.hh:
// 151-172) on an n^3 periodic box started from the Taylor-Green field
// (u0, k = 2 pi / n, density from its pressure), tau given.
//
//   ref_time n steps tau u0
//
// Prints one JSON line: cells, steps, seconds of the timed steps, MLUPS,
// seconds of the set-up.  Serial build: one core.
#include "palabos3D.h"
#include "palabos3D.hh"
#include "fsi/VariableViscosityDynamics.h"
#include "fsi/VariableViscosityDynamics.hh"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

using namespace plb;
typedef double T;
#define DESCRIPTOR descriptors::ForcedPhaseD3Q19Descriptor

int main(int argc, char** argv)
{
    plbInit(&argc, &argv);
    if (argc < 5) {
        std::fprintf(stderr, "usage: ref_time n steps tau u0\n");
        return 2;
    }
    const plint n = std::atol(argv[1]);
    const int steps = std::atoi(argv[2]);
    const T tau = std::atof(argv[3]), u0 = std::atof(argv[4]);
    if (n < 4 || steps < 1 || !(tau > 0.5)) {
        std::fprintf(stderr, "ref_time: need n >= 4, steps >= 1, tau > 0.5\n");
        return 2;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const T omega = 1. / tau;
    MultiBlockLattice3D<T, DESCRIPTOR> lattice(n, n, n, new GuoExternalForceVOFDynamics<T, DESCRIPTOR>(omega, omega));
    lattice.toggleInternalStatistics(false);
    lattice.periodicity().toggleAll(true);
    const T k = 2. * M_PI / (T)n;
    for (plint x = 0; x < n; ++x)
        for (plint y = 0; y < n; ++y)
            for (plint z = 0; z < n; ++z) {
                Cell<T, DESCRIPTOR>& cell = lattice.get(x, y, z);
                const T rho = 1. + 3. * u0 * u0 / 16. * (std::cos(2 * k * x) + std::cos(2 * k * y)) *
                                       (std::cos(2 * k * z) + 2.);
                const Array<T, 3> u(u0 * std::sin(k * x) * std::cos(k * y) * std::cos(k * z),
                                    -u0 * std::cos(k * x) * std::sin(k * y) * std::cos(k * z), 0.);
                const Array<T, 3> j(rho * u[0], rho * u[1], rho * u[2]);
                const T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
                for (int i = 0; i < 19; ++i)
                    cell[i] = dynamicsTemplates<T, DESCRIPTOR>::bgk_ma2_equilibrium(i, rho - 1., 1. / rho, j, jSqr);
                T* ext = cell.getExternal(0);
                for (int d = 0; d < 4; ++d) ext[d] = 0.;
            }
    lattice.initialize();
    const Box3D domain = lattice.getBoundingBox();
    MultiTensorField3D<T, 3> velocity(n, n, n);
    const auto t1 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) {
        computeVelocity(lattice, velocity, domain);
        setExternalVector(lattice, domain, DESCRIPTOR<T>::ExternalField::forceBeginsAt, Array<T, 3>(0., 0., 0.));
        lattice.collideAndStream();
    }
    const auto t2 = std::chrono::steady_clock::now();
    const double setup = std::chrono::duration<double>(t1 - t0).count();
    const double sec = std::chrono::duration<double>(t2 - t1).count();
    const double cells = (double)n * (double)n * (double)n;
    std::printf("{\"cells\": %.0f, \"steps\": %d, \"seconds\": %.6f, \"mlups\": %.4f, \"setup_seconds\": %.3f, "
                "\"velocity_x_probe\": %.17g}\n",
                cells, steps, sec, cells * steps / sec / 1e6, setup, velocity.get(n / 8, n / 4, 0)[0]);
    return 0;
}
