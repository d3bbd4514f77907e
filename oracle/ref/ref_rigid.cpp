// Golden vectors for rigid-growth packing, produced by the reference's own
// classes: the growth loop of progs/common/create_particles.h:129-155 with
//   RigidParticle3D (update, move_vertices, get_mass)  src/fsi/RigidParticle.hh:13-98
//   ParticleShape moment of inertia                    src/fsi/ParticleShape.hh:249-360
//   Quaternion                                         src/fsi/Quaternion.h
//   map_center_of_mass_to_periodic_grid                src/fsi/ParticleBase.hh:110-136
//   CollisionHandler + wall forces, SpringPotential    src/fsi/CollisionForces.h, src/fsi/Potentials.h
// ImmersedBoundaryDynamics3D needs MPI and is not built here, so its
// compute_collision_forces (src/fsi/ImmersedBoundaryDynamics_forces.hh:505-521,
// 332-375: forces reset, nodes strictly inside the lattice box +-0.5
// enlarged by the cut-off, pair forces, wall forces) and move_vertices
// (src/fsi/ImmersedBoundaryDynamics.hh:454-493: every particle in id order
// moves, then its centre of mass is mapped onto the periodic grid) are
// restated in the loop below with those same calls.  Positions and
// orientations are drawn from the fixture's seeded generator (the prog's
// std::rand draws and ParticlePositionInitializer are pinned elsewhere).
//
// rigid/<case>/: rest (nv,3), tri (nt,3), params, com0 (n,3), quat0 (n,4),
//   and after iteration k of the loop (k = 10 and the last): com_k, vel_k,
//   angvel_k, quat_k (n,4), scale_k, pos_k (n*nv,3).
#include "palabos3D.h"
#include "palabos3D.hh"
#include <cstdio>
#include <cmath>
#include "fsi/ParticleBase.h"
#include "fsi/ParticleShape.h"
#include "fsi/ParticleShapeFactory.h"
#include "fsi/RigidParticle.h"
#include "fsi/Boundary.h"
#include "fsi/geometry.h"
#include "fsi/Periodicity.h"
#include "fsi/Grid.h"
#include "fsi/Grid.hh"
#include "fsi/CollisionForces.h"
#include "fsi/Potentials.h"
#include "fsi/ParticleBase.hh"
#include "fsi/ParticleShape.hh"
#include "fsi/ParticleShapeFactory.hh"
#include "fsi/ParticleFactory.h"
#include "fsi/ParticleFactory.hh"
#include "fsi/Transform.h"
#include "fsi/Transform.hh"
#include "fsi/RigidParticle.hh"
#include "fsi/RayTracer.hh"
#include "common.h"

using namespace plb;
using namespace plb::fsi;
typedef double T;

namespace {

typedef Arithmetic3D<T, true, false, false> Arith;

struct State {
    std::vector<T> com, vel, angvel, quat, scale, pos;
};

State snapshot(const std::vector<RigidParticle3D<T>*>& ps)
{
    State s;
    for (size_t p = 0; p < ps.size(); ++p) {
        RigidParticle3D<T>* r = ps[p];
        for (int d = 0; d < 3; ++d) {
            s.com.push_back(r->center_of_mass()[d]);
            s.vel.push_back(r->velocity()[d]);
            s.angvel.push_back(r->angular_velocity()[d]);
        }
        for (int d = 0; d < 4; ++d) s.quat.push_back(r->orientation()[d]);
        s.scale.push_back(r->scale());
        for (plint i = 0; i < r->count_nodes(); ++i)
            for (int d = 0; d < 3; ++d) s.pos.push_back(r->get_node(i).pos[d]);
    }
    return s;
}

void write_state(const std::string& dir, const std::string& tag, const State& s, long n, long nv)
{
    refio::write_npy(dir + "/com_" + tag + ".npy", s.com, {n, 3L});
    refio::write_npy(dir + "/vel_" + tag + ".npy", s.vel, {n, 3L});
    refio::write_npy(dir + "/angvel_" + tag + ".npy", s.angvel, {n, 3L});
    refio::write_npy(dir + "/quat_" + tag + ".npy", s.quat, {n, 4L});
    refio::write_npy(dir + "/scale_" + tag + ".npy", s.scale, {n});
    refio::write_npy(dir + "/pos_" + tag + ".npy", s.pos, {n * nv, 3L});
}

void rigid_case(ParticleShapeLibrary<T>& lib, const std::string& name, int n, plint nit, unsigned long long seed)
{
    const std::string dir = "rigid/" + name;
    refio::ensure_dir("rigid");
    refio::ensure_dir(dir);
    const plint nx = 40, ny = 28, nz = 28;
    const Box3D domain(0, nx - 1, 0, ny - 1, 0, nz - 1);
    const T pipe_y = 13.5, pipe_z = 13.5, pipe_r = 13.0;
    const T initial_scale = 0.2, density = 1.0, damping = 0.1;
    const T spring_k = 1.0, spring_r = 5e-2;
    const int every = 10;
    ParticleShape<T>* shape = lib.get_by_tag("s314");
    const long nv = shape->count_vertices();

    PipeBoundary<T> boundary2(pipe_y, pipe_z, pipe_r - 1.5, (T)(nx - 1));
    const T r_init = initial_scale * shape->get_radius() + 1.;

    // positions: inside the shrunk pipe by the initial radius, apart by two
    // initial radii; orientations as the prog draws them
    refio::Rng rng(seed);
    std::vector<Array<T, 3> > cpos;
    std::vector<Quaternion<T> > q0;
    while ((int)cpos.size() < n) {
        Array<T, 3> x(rng.uniform(0., (T)nx - 1.), rng.uniform(0., (T)ny - 1.), rng.uniform(0., (T)nz - 1.));
        if (!boundary2.contains(x, r_init)) continue;
        bool ok = true;
        for (size_t k = 0; k < cpos.size() && ok; ++k) ok = norm(cpos[k] - x) > 2 * r_init;
        if (ok) cpos.push_back(x);
    }
    for (int k = 0; k < n; ++k) {
        Array<T, 3> dir(rng.uniform(0., 1.), rng.uniform(0., 1.), rng.uniform(0., 1.));
        const T dn = norm(dir);
        const T theta = 2. * M_PI * rng.uniform(0., 1.);
        q0.push_back(dn > std::numeric_limits<T>::epsilon() ? Quaternion<T>(theta, dir / dn) : Quaternion<T>());
    }

    std::vector<RigidParticle3D<T>*> ps;
    for (int k = 0; k < n; ++k) {
        RigidParticle3D<T>* p = new RigidParticle3D<T>(shape);
        p->scale() = initial_scale;
        p->density() = density;
        p->damping() = damping;
        p->update();
        p->set_center_of_mass(cpos[k]);
        p->orientation() = q0[k];
        p->update();
        p->set_id(k);
        ps.push_back(p);
    }
    std::vector<T> com0, quat0;
    for (int k = 0; k < n; ++k) {
        for (int d = 0; d < 3; ++d) com0.push_back(cpos[k][d]);
        for (int d = 0; d < 4; ++d) quat0.push_back(q0[k][d]);
    }

    Arith arithmetic(domain);
    geo::Rect<T> bb(domain.x0 - 0.5, domain.x1 + 0.5 - std::numeric_limits<T>::epsilon(), domain.y0 - 0.5,
                    domain.y1 + 0.5 - std::numeric_limits<T>::epsilon(), domain.z0 - 0.5,
                    domain.z1 + 0.5 - std::numeric_limits<T>::epsilon());
    SpringPotential<T> potential(spring_k, spring_r);
    geo::Rect<T> coll_domain = bb.enlarge(potential.get_cutoff_distance());
    for (plint it = 0; it <= nit; ++it) {
        if (it % every == 0) {
            const T fraction = (T)it / (T)nit;
            const T scale = initial_scale + (1. - initial_scale) * fraction;
            for (int k = 0; k < n; ++k) {
                ps[k]->scale() = scale;
                ps[k]->update();
            }
        }
        // compute_collision_forces(potential, &boundary2)
        for (int k = 0; k < n; ++k) ps[k]->reset_forces();
        collision::CollisionHandler<T, Arith, SpringPotential<T> > handler(bb, arithmetic, potential);
        for (int k = 0; k < n; ++k)
            for (plint i = 0; i < ps[k]->count_nodes(); ++i)
                if (coll_domain.contains(ps[k]->get_node(i).pos))
                    handler.add_node(ps[k]->get_node(i).pos, &(ps[k]->get_node(i).force), ps[k]->get_id());
        handler.compute_collision_forces();
        handler.compute_wall_collision_forces(boundary2);
        // move_vertices()
        for (int k = 0; k < n; ++k) {
            ps[k]->move_vertices(0);
            ps[k]->map_center_of_mass_to_periodic_grid(arithmetic);
        }
        if (it == 10) write_state(dir, "10", snapshot(ps), n, nv);
    }
    write_state(dir, "end", snapshot(ps), n, nv);

    std::vector<T> rest(3 * nv);
    for (long i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) rest[3 * i + d] = shape->get_vertex(i)[d];
    std::vector<int> tri;
    for (typename ParticleShape<T>::triangle_const_iterator t = shape->triangles_begin(); t != shape->triangles_end(); ++t) {
        tri.push_back((int)t->i0);
        tri.push_back((int)t->i1);
        tri.push_back((int)t->i2);
    }
    const Matrix<T, 3> I = shape->get_moment_of_inertia();
    std::vector<T> inertia(I.data(), I.data() + 9);
    const T params[] = {(T)nx, (T)ny, (T)nz, 1, 0, 0, pipe_y, pipe_z, pipe_r - 1.5, (T)(nx - 1), spring_k, spring_r,
                        initial_scale, (T)nit, (T)every, density, damping, (T)n, shape->get_volume()};
    refio::write_npy(dir + "/rest.npy", rest, {nv, 3L});
    refio::write_npy(dir + "/tri.npy", tri, {(long)tri.size() / 3, 3L});
    refio::write_npy(dir + "/inertia.npy", inertia, {3L, 3L});
    refio::write_npy(dir + "/params.npy", params, {(long)(sizeof(params) / sizeof(params[0]))});
    refio::write_npy(dir + "/com0.npy", com0, {(long)n, 3L});
    refio::write_npy(dir + "/quat0.npy", quat0, {(long)n, 4L});
    for (int k = 0; k < n; ++k) delete ps[k];
}

}  // namespace

int main(int argc, char** argv)
{
    plbInit(&argc, &argv);
    const std::string ref = argc > 1 ? argv[1] : "/root/reference";
    ParticleShapeLibrary<T> lib;
    lib.read_and_store_mesh(ref + "/progs/capsule_in_shear/sphere_TRImesh_314n_624el.msh", "s314");
    //          name        particles iterations seed
    rigid_case(lib, "pipe14", 14, 400, 71);
    return 0;
}
