// Golden vectors for the curved-wall path, produced by the reference itself:
//   GuoRigidWallBoundary instantiation     src/fsi/GuoBoundaryCondition.hh:62-135
//   GuoRigidWallBoundaryFunctional         src/fsi/GuoBoundaryCondition.hh:138-199
//   PipeBoundary / ParallelPlatesBoundary  src/fsi/Boundary.h:132-365
//   BoundaryOutside -> NoDynamics          src/fsi/Boundary.h:114-128
// set up as progs/rbc_in_pipe/rbc_in_pipe.cpp:286-330 does it (Guo-VOF
// dynamics everywhere, the wall functional inserted at level 1, cells beyond
// the boundary margin switched to NoDynamics), then driven by Palabos'
// collideAndStream (src/multiBlock/multiBlockLattice3D.hh:307-347).
//
// Each case writes the wall-node list (cell, direction, delta, wall position,
// wall velocity), the dynamics mask (0 Guo-VOF, 2 no dynamics), the external
// force, the populations after lattice.initialize() (which runs the wall
// functional once), after one step and after N steps, in the Palabos cell
// layout
// ((x*ny + y)*nz + z)*19 + i, fBar = f - t_i.
#include "palabos3D.h"
#include "palabos3D.hh"
#include "fsi/VariableViscosityDynamics.h"
#include "fsi/VariableViscosityDynamics.hh"
#include "fsi/Boundary.h"
#include "fsi/GuoBoundaryCondition.h"
#include "fsi/GuoBoundaryCondition.hh"
#include "common.h"
#include <cmath>

using namespace plb;
using namespace plb::fsi;
typedef double T;
#define DESCRIPTOR descriptors::ForcedPhaseD3Q19Descriptor

namespace {

struct PlateVelocity {
    T mid, u;
    Array<T, 3> operator()(const Array<T, 3>& wall_pos) const
    {
        return Array<T, 3>(wall_pos[1] < mid ? -u : u, 0., 0.);
    }
};

struct Case {
    std::string name;
    plint nx, ny, nz;
    bool periodic[3];
    int kind;          // 0 pipe, 1 plates
    T p[4];            // pipe: y0, z0, radius, length; plates: y0, y1
    T tau, body_fx, wall_u;
    int steps;
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
    MultiBlockLattice3D<T, DESCRIPTOR> lat(c.nx, c.ny, c.nz, new GuoExternalForceVOFDynamics<T, DESCRIPTOR>(omega, omega));
    lat.toggleInternalStatistics(false);
    for (int d = 0; d < 3; ++d) lat.periodicity().toggle(d, c.periodic[d]);

    PipeBoundary<T> pipe(c.p[0], c.p[1], c.p[2], c.p[3]);
    ParallelPlatesBoundary<T> plates(c.p[0], c.p[1]);
    const Boundary<T>& boundary = c.kind == 0 ? static_cast<const Boundary<T>&>(pipe)
                                              : static_cast<const Boundary<T>&>(plates);

    GuoRigidWallBoundary<T, DESCRIPTOR> wall(boundary, lat);
    if (c.kind == 0) wall.set_wall_velocity(Array<T, 3>(c.wall_u, 0., 0.));
    else wall.set_wall_velocity(PlateVelocity{(T)0.5 * (c.p[0] + c.p[1]), c.wall_u});
    wall.insert();
    BoundaryOutside<T> outside(boundary, std::sqrt(2));
    defineDynamics<T, DESCRIPTOR>(lat, lat.getBoundingBox(), new BoundaryOutside<T>(boundary, std::sqrt(2)),
                                  new NoDynamics<T, DESCRIPTOR>());

    std::vector<T> force(c.nx * c.ny * c.nz * 3);
    std::vector<unsigned char> mask(c.nx * c.ny * c.nz);
    for (plint x = 0; x < c.nx; ++x)
        for (plint y = 0; y < c.ny; ++y)
            for (plint z = 0; z < c.nz; ++z) {
                const long id = (x * c.ny + y) * c.nz + z;
                Cell<T, DESCRIPTOR>& cell = lat.get(x, y, z);
                T rho = rng.uniform(0.99, 1.01);
                Array<T, 3> u(rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02));
                Array<T, 3> j(rho * u[0], rho * u[1], rho * u[2]);
                T jSqr = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
                for (int i = 0; i < 19; ++i) {
                    T feq = dynamicsTemplates<T, DESCRIPTOR>::bgk_ma2_equilibrium(i, rho - 1., 1. / rho, j, jSqr);
                    cell[i] = feq + 1e-3 * DESCRIPTOR<T>::t[i] * rng.uniform(-1., 1.);
                }
                T* ext = cell.getExternal(0);
                ext[0] = c.body_fx;
                ext[1] = 0.;
                ext[2] = 0.;
                ext[3] = 0.;
                force[id * 3] = c.body_fx;
                mask[id] = outside(x, y, z) ? 2 : 0;
            }
    lat.duplicateOverlaps(modif::allVariables);
    lat.initialize();

    std::vector<long long> cell_ids;
    std::vector<int> dirs;
    std::vector<T> deltas, wall_pos, wall_vel;
    for (GuoRigidWallBoundary<T, DESCRIPTOR>::const_iterator it = wall.begin(); it != wall.end(); ++it) {
        cell_ids.push_back(((long long)it->location.x * c.ny + it->location.y) * c.nz + it->location.z);
        int dir = -1;
        for (int i = 1; i < 19; ++i)
            if (DESCRIPTOR<T>::c[i][0] == it->direction.x && DESCRIPTOR<T>::c[i][1] == it->direction.y &&
                DESCRIPTOR<T>::c[i][2] == it->direction.z)
                dir = i;
        dirs.push_back(dir);
        deltas.push_back(it->delta);
        for (int d = 0; d < 3; ++d) {
            wall_pos.push_back(it->wall_position[d]);
            wall_vel.push_back(it->velocity[d]);
        }
    }
    const long nw = (long)cell_ids.size();
    refio::write_npy(c.name + "/wall_cell.npy", cell_ids, {nw});
    refio::write_npy(c.name + "/wall_dir.npy", dirs, {nw});
    refio::write_npy(c.name + "/wall_delta.npy", deltas, {nw});
    refio::write_npy(c.name + "/wall_pos.npy", wall_pos, {nw, 3L});
    refio::write_npy(c.name + "/wall_vel.npy", wall_vel, {nw, 3L});
    refio::write_npy(c.name + "/mask.npy", mask, {(long)c.nx, (long)c.ny, (long)c.nz});
    refio::write_npy(c.name + "/force.npy", force, {(long)c.nx, (long)c.ny, (long)c.nz, 3L});
    T params[8] = {omega, (T)c.steps, (T)c.kind, c.p[0], c.p[1], c.p[2], c.p[3], c.wall_u};
    refio::write_npy(c.name + "/params.npy", params, {8L});
    int per[3] = {c.periodic[0], c.periodic[1], c.periodic[2]};
    refio::write_npy(c.name + "/periodic.npy", per, {3L});

    dump_lattice(lat, c, "f0");
    lat.collideAndStream();
    dump_lattice(lat, c, "f1");
    for (int s = 1; s < c.steps; ++s) lat.collideAndStream();
    dump_lattice(lat, c, "fN");
}

}  // namespace

int main(int argc, char** argv)
{
    plbInit(&argc, &argv);
    refio::ensure_dir("guo");
    Case a = {"guo/pipe_static", 6, 22, 21, {true, false, false}, 0, {10.3, 10.1, 8.6, 6.0}, 0.9, 2e-5, 0.0, 12, 21};
    Case b = {"guo/pipe_moving", 5, 19, 20, {true, false, false}, 0, {9.2, 9.6, 7.3, 5.0}, 0.7, 0.0, 0.02, 10, 22};
    Case c = {"guo/plates_couette", 6, 15, 7, {true, false, true}, 1, {2.35, 11.6, 0., 0.}, 1.0, 1e-5, 0.015, 12, 23};
    run_case(a);
    run_case(b);
    run_case(c);
    std::printf("ref_guo: wrote 3 cases\n");
    return 0;
}
