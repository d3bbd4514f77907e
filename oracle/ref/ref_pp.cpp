// Golden vectors for the interaction forces, produced by the reference itself:
//   MorsePotential / SpringPotential     src/fsi/Potentials.h:28-72
//   WallInteraction::apply_force         src/fsi/ForceDecorator.hh:27-35
//   CollisionHandler                     src/fsi/CollisionForces.h:57-222
//   Grid                                 src/fsi/Grid.hh:30-97
//   Arithmetic3D (minimal image)         src/fsi/Periodicity.h:128-280
// The handler is set up as ImmersedBoundaryDynamics3D's
// compute_particle_particle_interaction_impl sets it up in a serial run
// (src/fsi/ImmersedBoundaryDynamics_forces.hh:336-375: bounding box of the
// lattice domain +-0.5, nodes taken when strictly inside that box enlarged
// by the cut-off); ImmersedBoundaryDynamics3D itself needs MPI and is not
// built here.
//
// pp/<case>/: pos (n,3), obj (n), params (kind, p0, p1, p2, nx, ny, nz,
//   px, py, pz, wall kind, wall p0..p3), force (n,3) after
//   compute_collision_forces (+ compute_wall_collision_forces when a wall
//   is given).
// pp/wall_<case>/: a capsule's vertex positions near a boundary and the
//   forces WallInteraction::apply_force leaves on them (starting from zero).
#include "palabos3D.h"
#include "palabos3D.hh"
#include <cstdio>
#include "fsi/ParticleBase.h"
#include "fsi/ParticleShape.h"
#include "fsi/ParticleShapeFactory.h"
#include "fsi/DeformableParticle3D.h"
#include "fsi/DeformableCapsule.h"
#include "fsi/Boundary.h"
#include "fsi/geometry.h"
#include "fsi/Periodicity.h"
#include "fsi/Grid.h"
#include "fsi/Grid.hh"
#include "fsi/CollisionForces.h"
#include "fsi/Potentials.h"
#include "fsi/ForceDecorator.h"
#include "fsi/ForceDecorator.hh"
#include "fsi/ParticleBase.hh"
#include "fsi/ParticleShape.hh"
#include "fsi/ParticleShapeFactory.hh"
#include "fsi/DeformableParticle3D.hh"
#include "fsi/ParticleFactory.h"
#include "fsi/ParticleFactory.hh"
#include "fsi/Transform.h"
#include "fsi/Transform.hh"
#include "fsi/DeformableCapsule.hh"
#include "fsi/RayTracer.hh"
#include "common.h"

using namespace plb;
using namespace plb::fsi;
typedef double T;

namespace {

struct PPCase {
    std::string name;
    plint nx, ny, nz;
    bool per[3];
    int kind;        // 0 morse, 1 spring
    T p[3];
    int wall;        // -1 none, 0 pipe, 1 plates
    T w[4];
    int nobj, per_obj;
    T spread;        // cluster radius
    unsigned long long seed;
};

// The nodes: positions, the force slots the handler adds into, owners.
struct Nodes {
    std::vector<Array<T, 3> > pos, force;
    std::vector<plint> obj;
};

template <class Interaction, bool PX, bool PY, bool PZ>
long run_handler(const PPCase& c, const Interaction& interaction, Nodes& nd, const Boundary<T>* wall)
{
    const Box3D domain(0, c.nx - 1, 0, c.ny - 1, 0, c.nz - 1);
    Arithmetic3D<T, PX, PY, PZ> arithmetic(domain);
    geo::Rect<T> bb(domain.x0 - 0.5, domain.x1 + 0.5 - std::numeric_limits<T>::epsilon(), domain.y0 - 0.5,
                    domain.y1 + 0.5 - std::numeric_limits<T>::epsilon(), domain.z0 - 0.5,
                    domain.z1 + 0.5 - std::numeric_limits<T>::epsilon());
    geo::Rect<T> coll_domain = bb.enlarge(interaction.get_cutoff_distance());
    collision::CollisionHandler<T, Arithmetic3D<T, PX, PY, PZ>, Interaction> handler(bb, arithmetic, interaction);
    for (size_t i = 0; i < nd.pos.size(); ++i)
        if (coll_domain.contains(nd.pos[i])) handler.add_node(nd.pos[i], &nd.force[i], nd.obj[i]);
    handler.compute_collision_forces();
    if (wall) handler.compute_wall_collision_forces(*wall);
    return 0;
}

template <class Interaction>
long dispatch(const PPCase& c, const Interaction& it, Nodes& nd, const Boundary<T>* wall)
{
    const int code = (c.per[0] ? 4 : 0) | (c.per[1] ? 2 : 0) | (c.per[2] ? 1 : 0);
    switch (code) {
    case 0: return run_handler<Interaction, false, false, false>(c, it, nd, wall);
    case 1: return run_handler<Interaction, false, false, true>(c, it, nd, wall);
    case 2: return run_handler<Interaction, false, true, false>(c, it, nd, wall);
    case 3: return run_handler<Interaction, false, true, true>(c, it, nd, wall);
    case 4: return run_handler<Interaction, true, false, false>(c, it, nd, wall);
    case 5: return run_handler<Interaction, true, false, true>(c, it, nd, wall);
    case 6: return run_handler<Interaction, true, true, false>(c, it, nd, wall);
    default: return run_handler<Interaction, true, true, true>(c, it, nd, wall);
    }
}

void pp_case(const PPCase& c)
{
    const std::string dir = "pp/" + c.name;
    refio::ensure_dir(dir);
    refio::Rng rng(c.seed);
    Nodes nd;
    const T ext[3] = {(T)c.nx, (T)c.ny, (T)c.nz};
    for (int o = 0; o < c.nobj; ++o) {
        // cluster centres anywhere in [-1, n] so that faces, edges, corners
        // and the envelope all see nodes
        T ctr[3];
        for (int d = 0; d < 3; ++d) ctr[d] = rng.uniform(-1.2, ext[d] + 0.2);
        for (int k = 0; k < c.per_obj; ++k) {
            Array<T, 3> x;
            for (int d = 0; d < 3; ++d) x[d] = ctr[d] + rng.uniform(-c.spread, c.spread);
            nd.pos.push_back(x);
            nd.force.push_back(Array<T, 3>(0., 0., 0.));
            nd.obj.push_back(o);
        }
    }
    // a few exact lattice-aligned and half-integer positions (cell edges)
    for (int k = 0; k < 6 && !nd.pos.empty(); ++k) {
        Array<T, 3> x = nd.pos[k * 7 % nd.pos.size()];
        x[k % 3] = std::floor(x[k % 3]) + ((k & 1) ? 0.5 : 0.0);
        nd.pos.push_back(x);
        nd.force.push_back(Array<T, 3>(0., 0., 0.));
        nd.obj.push_back(c.nobj + k);
    }
    PipeBoundary<T> pipe(c.w[0], c.w[1], c.w[2], c.w[3]);
    ParallelPlatesBoundary<T> plates(c.w[0], c.w[1]);
    const Boundary<T>* wall = c.wall == 0 ? (const Boundary<T>*)&pipe : c.wall == 1 ? (const Boundary<T>*)&plates : 0;
    if (c.kind == 0) {
        MorsePotential<T> morse(c.p[0], c.p[1], c.p[2]);
        dispatch(c, morse, nd, wall);
    } else {
        SpringPotential<T> spring(c.p[0], c.p[1]);
        dispatch(c, spring, nd, wall);
    }
    const long n = nd.pos.size();
    std::vector<T> pos(3 * n), force(3 * n);
    std::vector<int> obj(n);
    for (long i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            pos[3 * i + d] = nd.pos[i][d];
            force[3 * i + d] = nd.force[i][d];
        }
        obj[i] = (int)nd.obj[i];
    }
    const T params[14] = {(T)c.kind, c.p[0], c.p[1], c.p[2], (T)c.nx, (T)c.ny, (T)c.nz,
                          (T)c.per[0], (T)c.per[1], (T)c.per[2], (T)c.wall, c.w[0], c.w[1], c.w[2]};
    const T wlen = c.w[3];
    std::vector<T> pv(params, params + 14);
    pv.push_back(wlen);
    refio::write_npy(dir + "/pos.npy", pos, {n, 3L});
    refio::write_npy(dir + "/obj.npy", obj, {n});
    refio::write_npy(dir + "/params.npy", pv, {(long)pv.size()});
    refio::write_npy(dir + "/force.npy", force, {n, 3L});
}

void wall_case(const std::string& ref, ParticleShapeLibrary<T>& lib, const std::string& name, int wall_kind,
               const T w[4], int kind, const T p[3], const T centre[3], T radius, unsigned long long seed)
{
    const std::string dir = "pp/wall_" + name;
    refio::ensure_dir(dir);
    ParticleShape<T>* shape = lib.get_by_tag("s626");
    CapsuleParameters<T> cp;
    cp.G = 0.02; cp.C = 10.; cp.nu_s = 1.; cp.kb = 0.;
    DeformableCapsuleParticle<T> cap(shape, cp);
    const long nv = cap.count_nodes();
    const Array<T, 3> c0 = shape->get_center();
    const T r0 = shape->get_radius();
    refio::Rng rng(seed);
    std::vector<T> pos(3 * nv);
    for (long i = 0; i < nv; ++i) {
        for (int d = 0; d < 3; ++d) {
            const T x = centre[d] + (shape->get_vertex(i)[d] - c0[d]) * (radius / r0) + 0.05 * rng.uniform(-1., 1.);
            cap.get_node(i).pos[d] = x;
            pos[3 * i + d] = x;
        }
    }
    cap.update();
    cap.reset_forces();
    PipeBoundary<T> pipe(w[0], w[1], w[2], w[3]);
    ParallelPlatesBoundary<T> plates(w[0], w[1]);
    const Boundary<T>& b = wall_kind == 0 ? (const Boundary<T>&)pipe : (const Boundary<T>&)plates;
    if (kind == 0) {
        WallInteraction<T, MorsePotential<T> > deco(b, MorsePotential<T>(p[0], p[1], p[2]));
        deco.apply_force(&cap);
    } else {
        WallInteraction<T, SpringPotential<T> > deco(b, SpringPotential<T>(p[0], p[1]));
        deco.apply_force(&cap);
    }
    std::vector<T> force(3 * nv);
    for (long i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) force[3 * i + d] = cap.get_node(i).force[d];
    const T params[9] = {(T)kind, p[0], p[1], p[2], (T)wall_kind, w[0], w[1], w[2], w[3]};
    refio::write_npy(dir + "/pos.npy", pos, {nv, 3L});
    refio::write_npy(dir + "/params.npy", params, {9L});
    refio::write_npy(dir + "/force.npy", force, {nv, 3L});
}

}  // namespace

int main(int argc, char** argv)
{
    plbInit(&argc, &argv);
    const std::string ref = argc > 1 ? argv[1] : "/root/reference";
    refio::ensure_dir("pp");
    //        name              nx  ny  nz  periodic       kind params          wall  wall params         objs per spread seed
    const PPCase cases[] = {
        {"morse_closed",       12, 10,  9, {false, false, false}, 0, {1.0, 1.0, 1e-3}, -1, {0, 0, 0, 0},        14, 20, 1.3, 41},
        {"morse_periodic_x",   12, 10,  9, {true, false, false},  0, {1.0, 1.0, 1e-3}, -1, {0, 0, 0, 0},        14, 20, 1.3, 42},
        {"spring_periodic_xz", 11,  9, 10, {true, false, true},   1, {1.0, 5e-2, 0},   -1, {0, 0, 0, 0},        14, 20, 1.3, 43},
        {"morse_periodic_xyz", 10,  9,  8, {true, true, true},    0, {0.8, 2.0, 1e-2}, -1, {0, 0, 0, 0},        16, 18, 1.2, 44},
        {"spring_periodic_yz",  9, 10, 11, {false, true, true},   1, {0.6, 0.5, 0},    -1, {0, 0, 0, 0},        16, 18, 1.2, 45},
        {"spring_thin_xyz",     3,  4,  3, {true, true, true},    1, {1.0, 5e-2, 0},   -1, {0, 0, 0, 0},         6, 10, 0.9, 46},
        {"morse_wall_plates",  12, 10,  9, {true, false, true},   0, {1.0, 1.0, 1e-3},  1, {0.5, 8.5, 0, 0},    14, 20, 1.3, 47},
        {"spring_wall_pipe",   12, 13, 13, {true, false, false},  1, {1.0, 5e-2, 0},    0, {6.0, 6.0, 5.2, 12}, 14, 20, 1.3, 48},
    };
    for (size_t k = 0; k < sizeof(cases) / sizeof(cases[0]); ++k) pp_case(cases[k]);

    ParticleShapeLibrary<T> lib;
    lib.read_and_store_mesh(ref + "/progs/rbc_in_pipe/sphere_TRImesh_626n_1248el.msh", "s626");
    {
        const T w[4] = {0.0, 20.0, 0, 0}, p[3] = {2.0, 2.0, 1e-4}, ctr[3] = {15.0, 4.6, 15.0};
        wall_case(ref, lib, "plates_morse", 1, w, 0, p, ctr, 4.0, 51);
    }
    {
        const T w[4] = {16.0, 16.0, 14.0, 40.0}, p[3] = {1.5, 0.05, 0}, ctr[3] = {20.0, 24.9, 17.0};
        wall_case(ref, lib, "pipe_spring", 0, w, 1, p, ctr, 4.0, 52);
    }
    {
        const T w[4] = {16.0, 16.0, 14.0, 40.0}, p[3] = {2.0, 2.0, 1e-4}, ctr[3] = {20.0, 9.0, 20.5};
        wall_case(ref, lib, "pipe_morse", 0, w, 0, p, ctr, 4.5, 53);
    }
    std::printf("ref_pp: done\n");
    return 0;
}
