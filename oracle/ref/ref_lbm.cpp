// Golden vectors for the D3Q19 Guo-VOF collide-and-stream path, produced by
// the reference itself: Palabos MultiBlockLattice3D::collideAndStream
// (src/multiBlock/multiBlockLattice3D.hh:307-347) driving
// GuoExternalForceVOFDynamics (src/fsi/VariableViscosityDynamics.hh:151-172),
// BounceBack (src/core/dynamics.hh:763-770) and the interpolation velocity
// boundary (src/boundaryCondition/boundaryCondition3D.hh:1184-1188,
// src/boundaryCondition/finiteDifferenceBoundaryProcessor3D.hh:68-115).
//
// Each case writes its initial state (after lattice.initialize()), the
// per-cell external force and phase, and the populations after N steps, in the
// Palabos cell layout: index ((x*ny + y)*nz + z)*19 + i, fBar = f - t_i.
#include "palabos3D.h"
#include "palabos3D.hh"
#include "fsi/VariableViscosityDynamics.h"
#include "fsi/VariableViscosityDynamics.hh"
#include "common.h"

using namespace plb;
typedef double T;
#define DESCRIPTOR descriptors::ForcedPhaseD3Q19Descriptor

namespace {

struct Case {
    std::string name;
    plint nx, ny, nz;
    bool periodic[3];
    T tau, lambda;
    int steps;
    T force_amp, body_fx;
    bool random_phase;
    int walls;          // 0 = none, 1 = bounce-back planes y=0,ny-1, 2 = interp velocity planes
    T wall_u;           // interp planes: u_x = -wall_u at y=0, +wall_u at y=ny-1
    unsigned long long seed;
};

void dump_lattice(MultiBlockLattice3D<T, DESCRIPTOR>& lat, const Case& c, const std::string& tag)
{
    std::vector<T> f(c.nx * c.ny * c.nz * 19);
    for (plint x = 0; x < c.nx; ++x)
        for (plint y = 0; y < c.ny; ++y)
            for (plint z = 0; z < c.nz; ++z) {
                Cell<T, DESCRIPTOR>& cell = lat.get(x, y, z);
                for (int i = 0; i < 19; ++i) f[((x * c.ny + y) * c.nz + z) * 19 + i] = cell[i];
            }
    refio::write_npy(c.name + "/" + tag + ".npy", f, {(long)c.nx, (long)c.ny, (long)c.nz, 19L});
}

void run_case(const Case& c)
{
    refio::ensure_dir(c.name);
    refio::Rng rng(c.seed);
    const T omega = 1. / c.tau;
    const T omega_in = 1. / (0.5 + c.lambda * (c.tau - 0.5));
    MultiBlockLattice3D<T, DESCRIPTOR> lat(c.nx, c.ny, c.nz,
        new GuoExternalForceVOFDynamics<T, DESCRIPTOR>(omega, omega_in));
    lat.toggleInternalStatistics(false);
    for (int d = 0; d < 3; ++d) lat.periodicity().toggle(d, c.periodic[d]);

    OnLatticeBoundaryCondition3D<T, DESCRIPTOR>* bc = 0;
    if (c.walls == 1) {
        defineDynamics(lat, Box3D(0, c.nx - 1, 0, 0, 0, c.nz - 1), new BounceBack<T, DESCRIPTOR>());
        defineDynamics(lat, Box3D(0, c.nx - 1, c.ny - 1, c.ny - 1, 0, c.nz - 1), new BounceBack<T, DESCRIPTOR>());
    } else if (c.walls == 2) {
        bc = createInterpBoundaryCondition3D<T, DESCRIPTOR>();
        bc->addVelocityBoundary1N(Box3D(0, c.nx - 1, 0, 0, 0, c.nz - 1), lat);
        bc->addVelocityBoundary1P(Box3D(0, c.nx - 1, c.ny - 1, c.ny - 1, 0, c.nz - 1), lat);
        setBoundaryVelocity(lat, Box3D(0, c.nx - 1, 0, 0, 0, c.nz - 1), Array<T, 3>(-c.wall_u, 0., 0.));
        setBoundaryVelocity(lat, Box3D(0, c.nx - 1, c.ny - 1, c.ny - 1, 0, c.nz - 1), Array<T, 3>(c.wall_u, 0., 0.));
    }

    // Seeded initial state: equilibrium at random (rho, u) plus a small
    // non-equilibrium perturbation; random force and phase fields.
    std::vector<T> force(c.nx * c.ny * c.nz * 3), phase(c.nx * c.ny * c.nz);
    for (plint x = 0; x < c.nx; ++x)
        for (plint y = 0; y < c.ny; ++y)
            for (plint z = 0; z < c.nz; ++z) {
                const long id = (x * c.ny + y) * c.nz + z;
                Cell<T, DESCRIPTOR>& cell = lat.get(x, y, z);
                T rho = rng.uniform(0.98, 1.02);
                Array<T, 3> u(rng.uniform(-0.04, 0.04), rng.uniform(-0.04, 0.04), rng.uniform(-0.04, 0.04));
                Array<T, 3> j(rho * u[0], rho * u[1], rho * u[2]);
                T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
                for (int i = 0; i < 19; ++i) {
                    T feq = dynamicsTemplates<T, DESCRIPTOR>::bgk_ma2_equilibrium(i, rho - 1., 1. / rho, j, jSqr);
                    cell[i] = feq + 1e-3 * DESCRIPTOR<T>::t[i] * rng.uniform(-1., 1.);
                }
                T* ext = cell.getExternal(0);
                for (int d = 0; d < 3; ++d) {
                    ext[d] = c.force_amp * rng.uniform(-1., 1.) + (d == 0 ? c.body_fx : 0.);
                    force[id * 3 + d] = ext[d];
                }
                ext[3] = c.random_phase ? (rng.uniform() < 0.5 ? 1. : 0.) : 0.;
                phase[id] = ext[3];
            }
    lat.duplicateOverlaps(modif::allVariables);
    lat.initialize();

    dump_lattice(lat, c, "f0");
    refio::write_npy(c.name + "/force.npy", force, {(long)c.nx, (long)c.ny, (long)c.nz, 3L});
    refio::write_npy(c.name + "/phase.npy", phase, {(long)c.nx, (long)c.ny, (long)c.nz});
    T params[6] = {omega, omega_in, (T)c.steps, c.wall_u, (T)c.walls, 0.};
    refio::write_npy(c.name + "/params.npy", params, {6L});
    int per[3] = {c.periodic[0], c.periodic[1], c.periodic[2]};
    refio::write_npy(c.name + "/periodic.npy", per, {3L});

    for (int s = 0; s < c.steps; ++s) lat.collideAndStream();
    dump_lattice(lat, c, "fN");

    // Macroscopic velocity as computeVelocity() reports it (u = j/rho + F/2,
    // stored wall velocity on interpolation-boundary cells).
    std::auto_ptr<MultiTensorField3D<T, 3> > vel = computeVelocity(lat);
    std::auto_ptr<MultiScalarField3D<T> > rho = computeDensity(lat);
    std::vector<T> u(c.nx * c.ny * c.nz * 3), r(c.nx * c.ny * c.nz);
    for (plint x = 0; x < c.nx; ++x)
        for (plint y = 0; y < c.ny; ++y)
            for (plint z = 0; z < c.nz; ++z) {
                const long id = (x * c.ny + y) * c.nz + z;
                for (int d = 0; d < 3; ++d) u[id * 3 + d] = vel->get(x, y, z)[d];
                r[id] = rho->get(x, y, z);
            }
    refio::write_npy(c.name + "/uN.npy", u, {(long)c.nx, (long)c.ny, (long)c.nz, 3L});
    refio::write_npy(c.name + "/rhoN.npy", r, {(long)c.nx, (long)c.ny, (long)c.nz});
    delete bc;
}

}  // namespace

int main(int argc, char** argv)
{
    plbInit(&argc, &argv);
    refio::ensure_dir("lbm");
    Case a = {"lbm/periodic_vof", 12, 10, 14, {true, true, true}, 0.8, 5.0, 10, 1e-3, 0., true, 0, 0., 11};
    Case b = {"lbm/shear_interp", 12, 11, 10, {true, false, true}, 0.9, 1.0, 20, 1e-4, 0., false, 2, 0.02, 12};
    Case c = {"lbm/bb_channel", 10, 12, 8, {true, false, true}, 1.0, 1.0, 30, 0., 1e-5, false, 1, 0., 13};
    run_case(a);
    run_case(b);
    run_case(c);
    std::printf("ref_lbm: wrote 3 cases\n");
    return 0;
}
