/*
 * biblbm_plb.h -- C++ host layer over the C-ABI (biblbm.h) with the names,
 * argument meaning and call order the reference's progs use on the coupled
 * lattice / immersed-boundary path, so that a prog's setup and time loop
 * read as they do against Palabos + fsi:
 *
 *   MultiBlockLattice3D<T, ForcedPhaseD3Q19Descriptor>(nx, ny, nz,
 *       new GuoExternalForceVOFDynamics<T, D>(omega, omega_in))
 *                                     src/multiBlock/multiBlockLattice3D.h,
 *                                     src/fsi/VariableViscosityDynamics.h:48-54
 *   periodicity().toggle, defineDynamics(BounceBack / NoDynamics)
 *                                     src/multiBlock/multiBlock3D.h:77,
 *                                     src/core/dynamics.h:559,670
 *   createInterpBoundaryCondition3D, addVelocityBoundary{0,1,2}{N,P},
 *   setBoundaryVelocity               src/boundaryCondition/boundaryCondition3D.hh:1053-1066
 *   initializeAtEquilibrium, setExternalVector, computeVelocity, computeDensity
 *                                     src/dataProcessors/dataInitializerWrapper3D.h,
 *                                     src/dataProcessors/dataAnalysisWrapper3D.hh:1660
 *   lattice.initialize(), lattice.collideAndStream()
 *                                     src/multiBlock/multiBlockLattice3D.hh:307-347
 *   fsi::ParticleShapeLibrary, DeformableCapsuleParticle, RBCParticle
 *                                     src/fsi/ParticleShapeFactory.hh:32-129,
 *                                     src/fsi/DeformableCapsule.h:22-64, src/fsi/RBCParticle.h:22-58
 *   fsi::PipeBoundary, fsi::ParallelPlatesBoundary, fsi::BoundaryOutside,
 *   fsi::GuoRigidWallBoundary (insert, set_wall_velocity)
 *                                     src/fsi/Boundary.h:114-365,
 *                                     src/fsi/GuoBoundaryCondition.h:19-52
 *   fsi::MorsePotential, fsi::SpringPotential, fsi::PPPotentialForce,
 *   fsi::WallInteraction          src/fsi/Potentials.h:28-72,
 *                                     src/fsi/ParticleParticleInteraction.h:17-40,
 *                                     src/fsi/ForceDecorator.hh:18-35
 *   fsi::ImmersedBoundaryDynamics3D: add_particle, init, get_particle,
 *   set_boundary, add_force_decorator, add_pp_force,
 *   compute_collision_forces, compute_total_volume, write_particles_as_vtk,
 *   write_lightweight_particle_data, save_checkpoint / load_checkpoint,
 *   fsi::Time and parallelIO::save / load of the lattice
 *                                     src/fsi/ImmersedBoundaryDynamics_io.hh:17-200,
 *                                     src/fsi/Time.h:15-48;
 *   create_rbc_params, shrink_rbc_volume (on the device),
 *   fsi::ParticlePositionInitializer
 *                                     progs/common/calibrate_rbc_shape.h:13-112,
 *                                     src/fsi/ParticlePositionInitializer.hh:18-86; and the
 *   per-step functional
 *   applyProcessingFunctional(wrap_ibm_dynamics3D(ibm), domain, lattice, u)
 *                                     src/fsi/ImmersedBoundaryDynamics.h:57-114,
 *                                     src/fsi/DataProcessors.hh:94-131
 *
 * State lives on the device; the objects here hold configuration until the
 * lattice is first used and then forward to the C-ABI.  Errors throw
 * std::runtime_error with the library's message (the reference prints and
 * exits).  Differences from the reference's objects:
 *   - one velocity-boundary axis per lattice (both of its faces);
 *   - setExternalVector only on the whole domain (a body force); the
 *     immersed boundary owns every other force;
 *   - computeVelocity / computeDensity copy the whole field to the host,
 *     so a loop should call them only on the steps it reads them;
 *   - the immersed-boundary functional runs when the next collideAndStream
 *     (or any read of particle state) needs it, fused with the collision;
 *   - population storage can be switched to FP32 or FP16-shifted before first
 *     use (setStorage), which the reference has no notion of.
 */
#ifndef BIBLBM_PLB_H
#define BIBLBM_PLB_H

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "biblbm.h"

namespace plb {

typedef long plint;
typedef unsigned long pluint;

inline void blCheck(int rc, const char* what)
{
    if (rc != BL_OK) throw std::runtime_error(std::string(what) + ": " + bl_last_error());
}

// io::mkdir (one level; an existing folder is fine)
inline void makeDir(const std::string& path) { ::mkdir(path.c_str(), 0755); }

template <typename T, int n>
struct Array {
    T d[n];
    Array() : d() {}
    Array(T a, T b, T c) : d{a, b, c} {}
    T& operator[](int i) { return d[i]; }
    const T& operator[](int i) const { return d[i]; }
};

struct Box3D {
    plint x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
    Box3D() {}
    Box3D(plint x0_, plint x1_, plint y0_, plint y1_, plint z0_, plint z1_)
        : x0(x0_), x1(x1_), y0(y0_), y1(y1_), z0(z0_), z1(z1_)
    {
    }
    plint getNx() const { return x1 - x0 + 1; }
    plint getNy() const { return y1 - y0 + 1; }
    plint getNz() const { return z1 - z0 + 1; }
};

namespace util {
template <typename T>
T sqr(T x) { return x * x; }
}  // namespace util

namespace descriptors {
// D3Q19 with external force (3) and phase (1), ForcedPhaseD3Q19Descriptor
// (src/fsi/VariableViscosityDynamics.h:25-42).
template <typename T>
struct ForcedPhaseD3Q19Descriptor {
    struct ExternalField {
        enum { forceBeginsAt = 0, phaseBeginsAt = 3 };
    };
};
}  // namespace descriptors

// ------------------------------------------------------------ dynamics
template <typename T, template <typename U> class Descriptor>
struct Dynamics {
    virtual ~Dynamics() {}
    virtual unsigned char maskCode() const = 0;
};

template <typename T, template <typename U> class Descriptor>
struct GuoExternalForceVOFDynamics : Dynamics<T, Descriptor> {
    T omega, omega_in;
    explicit GuoExternalForceVOFDynamics(T omega_, T omega_in_) : omega(omega_), omega_in(omega_in_) {}
    explicit GuoExternalForceVOFDynamics(T omega_) : omega(omega_), omega_in(omega_) {}
    unsigned char maskCode() const override { return 0; }
};

template <typename T, template <typename U> class Descriptor>
struct BounceBack : Dynamics<T, Descriptor> {
    explicit BounceBack(T = T(1)) {}
    unsigned char maskCode() const override { return 1; }
};

template <typename T, template <typename U> class Descriptor>
struct NoDynamics : Dynamics<T, Descriptor> {
    unsigned char maskCode() const override { return 2; }
};

template <typename T, template <typename U> class Descriptor>
class MultiBlockLattice3D;

// host-side fields filled from the lattice
template <typename T>
class MultiScalarField3D {
public:
    MultiScalarField3D(plint nx, plint ny, plint nz) : nx_(nx), ny_(ny), nz_(nz), v_(nx * ny * nz) {}
    T& get(plint x, plint y, plint z) { return v_[(x * ny_ + y) * nz_ + z]; }
    std::vector<T>& data() { return v_; }
    void fill(bl_lattice l) { blCheck(bl_lattice_get_moments(l, v_.data(), nullptr), "computeDensity"); }

private:
    plint nx_, ny_, nz_;
    std::vector<T> v_;
};

template <typename T, int nDim>
class MultiTensorField3D {
public:
    MultiTensorField3D(plint nx, plint ny, plint nz) : nx_(nx), ny_(ny), nz_(nz), v_(nx * ny * nz * nDim) {}
    Array<T, nDim>& get(plint x, plint y, plint z)
    {
        return *reinterpret_cast<Array<T, nDim>*>(&v_[nDim * ((x * ny_ + y) * nz_ + z)]);
    }
    std::vector<T>& data() { return v_; }
    void fill(bl_lattice l) { blCheck(bl_lattice_get_moments(l, nullptr, v_.data()), "computeVelocity"); }

private:
    plint nx_, ny_, nz_;
    std::vector<T> v_;
};

template <typename T>
std::unique_ptr<MultiScalarField3D<T>> generateMultiScalarField(const Box3D& b)
{
    return std::unique_ptr<MultiScalarField3D<T>>(new MultiScalarField3D<T>(b.getNx(), b.getNy(), b.getNz()));
}

template <typename T, int nDim>
std::unique_ptr<MultiTensorField3D<T, nDim>> generateMultiTensorField(const Box3D& b)
{
    return std::unique_ptr<MultiTensorField3D<T, nDim>>(
        new MultiTensorField3D<T, nDim>(b.getNx(), b.getNy(), b.getNz()));
}

namespace fsi {
template <typename T, template <typename U> class Descriptor, class Periodicity>
class ImmersedBoundaryDynamics3D;
}

// ------------------------------------------------------------- lattice
template <typename T, template <typename U> class Descriptor>
class MultiBlockLattice3D {
public:
    class Periodicity3D {
    public:
        explicit Periodicity3D(MultiBlockLattice3D& l) : l_(l) {}
        void toggle(int d, bool p) { l_.setPeriodic(d, p); }
        void toggleAll(bool p) { for (int d = 0; d < 3; ++d) l_.setPeriodic(d, p); }

    private:
        MultiBlockLattice3D& l_;
    };

    MultiBlockLattice3D(plint nx, plint ny, plint nz, Dynamics<T, Descriptor>* background)
        : nx_(nx), ny_(ny), nz_(nz), mask_(nx * ny * nz, background->maskCode())
    {
        auto* guo = dynamic_cast<GuoExternalForceVOFDynamics<T, Descriptor>*>(background);
        if (guo) {
            omega_ = guo->omega;
            omega_in_ = guo->omega_in;
        }
        delete background;  // the lattice takes ownership, as in the reference
    }
    ~MultiBlockLattice3D()
    {
        if (h_) bl_lattice_destroy(h_);
    }
    MultiBlockLattice3D(const MultiBlockLattice3D&) = delete;
    MultiBlockLattice3D& operator=(const MultiBlockLattice3D&) = delete;

    Box3D getBoundingBox() const { return Box3D(0, nx_ - 1, 0, ny_ - 1, 0, nz_ - 1); }
    plint getNx() const { return nx_; }
    plint getNy() const { return ny_; }
    plint getNz() const { return nz_; }
    Periodicity3D periodicity() { return Periodicity3D(*this); }
    void toggleInternalStatistics(bool) {}
    void setStorage(int storage)
    {
        if (h_) throw std::runtime_error("setStorage: the lattice is already in use");
        storage_ = storage;
    }

    void initialize() { blCheck(bl_lattice_initialize(handle()), "initialize"); }
    void collideAndStream()
    {
        if (pending_ibm_) {
            bl_ibm g = pending_ibm_;
            pending_ibm_ = nullptr;
            blCheck(bl_coupled_steps(g, 1, 0), "collideAndStream (with the immersed boundary)");
        } else {
            blCheck(bl_collide_stream(handle(), 1), "collideAndStream");
        }
    }

    // the device lattice, created with the configuration so far on first use
    bl_lattice handle()
    {
        if (h_) return h_;
        const int per[3] = {periodic_[0], periodic_[1], periodic_[2]};
        blCheck(bl_lattice_create((int)nx_, (int)ny_, (int)nz_, per, storage_, &h_), "MultiBlockLattice3D");
        blCheck(bl_lattice_set_dynamics(h_, omega_, omega_in_), "GuoExternalForceVOFDynamics");
        pushMask();
        blCheck(bl_lattice_set_body_force(h_, g_.d), "setExternalVector");
        if (!rho_.empty()) blCheck(bl_lattice_init_equilibrium(h_, rho_.data(), u_.data()), "initializeAtEquilibrium");
        rho_.clear();
        u_.clear();
        return h_;
    }

    // ---- used by the free functions below ----
    void setPeriodic(int d, bool p)
    {
        if (h_) throw std::runtime_error("periodicity: the lattice is already in use");
        periodic_[d] = p ? 1 : 0;
    }
    void setMask(const Box3D& b, unsigned char code)
    {
        for (plint x = std::max<plint>(b.x0, 0); x <= std::min(b.x1, nx_ - 1); ++x)
            for (plint y = std::max<plint>(b.y0, 0); y <= std::min(b.y1, ny_ - 1); ++y)
                for (plint z = std::max<plint>(b.z0, 0); z <= std::min(b.z1, nz_ - 1); ++z)
                    mask_[(x * ny_ + y) * nz_ + z] = code;
        if (h_) pushMask();
    }
    template <class Pred>
    void setMaskWhere(const Box3D& b, Pred&& pred, unsigned char code)
    {
        for (plint x = std::max<plint>(b.x0, 0); x <= std::min(b.x1, nx_ - 1); ++x)
            for (plint y = std::max<plint>(b.y0, 0); y <= std::min(b.y1, ny_ - 1); ++y)
                for (plint z = std::max<plint>(b.z0, 0); z <= std::min(b.z1, nz_ - 1); ++z)
                    if (pred(x, y, z)) mask_[(x * ny_ + y) * nz_ + z] = code;
        if (h_) pushMask();
    }
    void setWall(int axis, bool hi, const Array<T, 3>& u)
    {
        if (wall_axis_ >= 0 && wall_axis_ != axis)
            throw std::runtime_error("velocity boundaries on more than one axis are not supported");
        wall_axis_ = axis;
        for (int d = 0; d < 3; ++d) wall_u_[hi ? 3 + d : d] = u[d];
        if (h_) pushMask();
    }
    void setBodyForce(const Array<T, 3>& g)
    {
        g_ = g;
        if (h_) blCheck(bl_lattice_set_body_force(h_, g_.d), "setExternalVector");
    }
    template <class F>
    void setEquilibrium(const Box3D& b, F&& rho_u)
    {
        if (rho_.empty()) {
            rho_.assign(nx_ * ny_ * nz_, T(1));
            u_.assign(3 * nx_ * ny_ * nz_, T(0));
            if (h_) {  // start from the current state of the cells outside the box
                blCheck(bl_lattice_get_moments(h_, rho_.data(), u_.data()), "initializeAtEquilibrium");
                for (T& r : rho_)
                    if (r == T(0)) r = T(1);
            }
        }
        for (plint x = b.x0; x <= b.x1; ++x)
            for (plint y = b.y0; y <= b.y1; ++y)
                for (plint z = b.z0; z <= b.z1; ++z) {
                    const plint c = (x * ny_ + y) * nz_ + z;
                    Array<T, 3> u;
                    rho_u(x, y, z, rho_[c], u);
                    for (int d = 0; d < 3; ++d) u_[3 * c + d] = u[d];
                }
        if (h_) {
            blCheck(bl_lattice_init_equilibrium(h_, rho_.data(), u_.data()), "initializeAtEquilibrium");
            rho_.clear();
            u_.clear();
        }
    }
    void setPendingBoundaryStep(bl_ibm g) { pending_ibm_ = g; }
    void flushPendingBoundaryStep()
    {
        if (pending_ibm_) {
            bl_ibm g = pending_ibm_;
            pending_ibm_ = nullptr;
            blCheck(bl_ibm_step(g), "immersed boundary step");
        }
    }

private:
    void pushMask()
    {
        bool any = false;
        for (unsigned char m : mask_) any = any || m != 0;
        blCheck(bl_lattice_set_mask(h_, any ? mask_.data() : nullptr), "defineDynamics");
        if (wall_axis_ >= 0)
            blCheck(bl_lattice_set_velocity_walls(h_, wall_axis_, wall_u_, wall_u_ + 3), "setBoundaryVelocity");
    }

    plint nx_, ny_, nz_;
    int periodic_[3] = {1, 1, 1};
    int storage_ = BL_STORE_F64;
    T omega_ = 1, omega_in_ = 1;
    std::vector<unsigned char> mask_;
    int wall_axis_ = -1;
    double wall_u_[6] = {0, 0, 0, 0, 0, 0};
    Array<T, 3> g_;
    std::vector<T> rho_, u_;
    bl_lattice h_ = nullptr;
    bl_ibm pending_ibm_ = nullptr;
};

template <typename T, template <typename U> class Descriptor>
void defineDynamics(MultiBlockLattice3D<T, Descriptor>& lattice, const Box3D& box, Dynamics<T, Descriptor>* dyn)
{
    lattice.setMask(box, dyn->maskCode());
    delete dyn;
}

// defineDynamics over the cells of the box where a domain functional holds
// (src/dataProcessors/dataInitializerWrapper3D.hh); takes ownership of both.
template <typename T, template <typename U> class Descriptor, class DomainFunctional>
void defineDynamics(MultiBlockLattice3D<T, Descriptor>& lattice, const Box3D& box, DomainFunctional* domain,
                    Dynamics<T, Descriptor>* dyn)
{
    std::unique_ptr<DomainFunctional> own(domain);
    lattice.setMaskWhere(box, [&](plint x, plint y, plint z) { return (*own)(x, y, z); }, dyn->maskCode());
    delete dyn;
}

template <typename T, template <typename U> class Descriptor>
void initializeAtEquilibrium(MultiBlockLattice3D<T, Descriptor>& lattice, const Box3D& box, T rho,
                             const Array<T, 3>& u)
{
    lattice.setEquilibrium(box, [&](plint, plint, plint, T& r, Array<T, 3>& v) {
        r = rho;
        v = u;
    });
}

// initializeAtEquilibrium with a functor f(x, y, z, rho, u)
template <typename T, template <typename U> class Descriptor, class F>
void initializeAtEquilibrium(MultiBlockLattice3D<T, Descriptor>& lattice, const Box3D& box, F f)
{
    lattice.setEquilibrium(box, f);
}

template <typename T, template <typename U> class Descriptor>
void setExternalVector(MultiBlockLattice3D<T, Descriptor>& lattice, const Box3D& box, int forceBeginsAt,
                       const Array<T, 3>& g)
{
    const Box3D all = lattice.getBoundingBox();
    if (forceBeginsAt != Descriptor<T>::ExternalField::forceBeginsAt || box.x0 > all.x0 || box.x1 < all.x1 ||
        box.y0 > all.y0 || box.y1 < all.y1 || box.z0 > all.z0 || box.z1 < all.z1)
        throw std::runtime_error("setExternalVector: only the force on the whole domain (a body force)");
    lattice.setBodyForce(g);
}

template <typename T, template <typename U> class Descriptor>
void computeVelocity(MultiBlockLattice3D<T, Descriptor>& lattice, MultiTensorField3D<T, 3>& u, const Box3D&)
{
    lattice.flushPendingBoundaryStep();
    u.fill(lattice.handle());
}

template <typename T, template <typename U> class Descriptor>
void computeDensity(MultiBlockLattice3D<T, Descriptor>& lattice, MultiScalarField3D<T>& rho, const Box3D&)
{
    lattice.flushPendingBoundaryStep();
    rho.fill(lattice.handle());
}

// Regularized finite-difference velocity boundary (the interp boundary
// condition of the progs): cells of the box become velocity-wall nodes.
template <typename T, template <typename U> class Descriptor>
class OnLatticeBoundaryCondition3D {
public:
    void addVelocityBoundary0N(const Box3D& b, MultiBlockLattice3D<T, Descriptor>& l) { add(b, l, 0, false); }
    void addVelocityBoundary0P(const Box3D& b, MultiBlockLattice3D<T, Descriptor>& l) { add(b, l, 0, true); }
    void addVelocityBoundary1N(const Box3D& b, MultiBlockLattice3D<T, Descriptor>& l) { add(b, l, 1, false); }
    void addVelocityBoundary1P(const Box3D& b, MultiBlockLattice3D<T, Descriptor>& l) { add(b, l, 1, true); }
    void addVelocityBoundary2N(const Box3D& b, MultiBlockLattice3D<T, Descriptor>& l) { add(b, l, 2, false); }
    void addVelocityBoundary2P(const Box3D& b, MultiBlockLattice3D<T, Descriptor>& l) { add(b, l, 2, true); }

private:
    void add(const Box3D& b, MultiBlockLattice3D<T, Descriptor>& l, int axis, bool hi)
    {
        l.setMask(b, 3);
        l.setWall(axis, hi, Array<T, 3>());
    }
};

template <typename T, template <typename U> class Descriptor>
OnLatticeBoundaryCondition3D<T, Descriptor>* createInterpBoundaryCondition3D()
{
    return new OnLatticeBoundaryCondition3D<T, Descriptor>();
}

// the wall velocity of the face the box lies on (index 0 or n-1 of the
// velocity-boundary axis)
template <typename T, template <typename U> class Descriptor>
void setBoundaryVelocity(MultiBlockLattice3D<T, Descriptor>& lattice, const Box3D& b, const Array<T, 3>& u)
{
    const plint lo[3] = {b.x0, b.y0, b.z0}, hi[3] = {b.x1, b.y1, b.z1};
    const plint n[3] = {lattice.getNx(), lattice.getNy(), lattice.getNz()};
    for (int a = 0; a < 3; ++a)
        if (lo[a] == hi[a] && (lo[a] == 0 || lo[a] == n[a] - 1)) {
            lattice.setWall(a, lo[a] == n[a] - 1, u);
            return;
        }
    throw std::runtime_error("setBoundaryVelocity: the box must be one face plane of the domain");
}

// ================================================================== fsi
namespace fsi {

// Analytic boundaries (src/fsi/Boundary.h): the same expressions as the
// device code uses; desc() is what crosses the C-ABI.
template <class T>
class Boundary {
public:
    virtual ~Boundary() {}
    virtual bool contains(const Array<T, 3>& p, T margin = 0) const = 0;
    virtual Box3D get_bounding_box(plint margin) const = 0;
    virtual bl_boundary desc() const = 0;
};

template <class T>
class PipeBoundary : public Boundary<T> {
public:
    PipeBoundary() {}
    PipeBoundary(T y0_, T z0_, T radius_, T length_) : y0(y0_), z0(z0_), radius(radius_), length(length_) {}
    Box3D get_bounding_box(plint margin) const override
    {
        return Box3D((plint)std::floor(x0), (plint)std::ceil(x0 + std::floor(length)),
                     (plint)std::ceil(y0 - radius) - margin, (plint)std::floor(y0 + radius) + margin,
                     (plint)std::ceil(z0 - radius) - margin, (plint)std::floor(z0 + radius) + margin);
    }
    bool contains(const Array<T, 3>& p, T margin = 0) const override
    {
        return ((p[1] - y0) * (p[1] - y0) + (p[2] - z0) * (p[2] - z0)) < (radius - margin) * (radius - margin);
    }
    bl_boundary desc() const override { return bl_boundary{BL_BOUNDARY_PIPE, {y0, z0, radius, length}}; }
    T x0 = 0, y0 = 0, z0 = 0, radius = 0, length = 0;
};

template <class T>
class ParallelPlatesBoundary : public Boundary<T> {
public:
    ParallelPlatesBoundary(T y0_, T y1_) : y0(y0_), y1(y1_) {}
    bool contains(const Array<T, 3>& p, T margin = 0) const override { return p[1] > y0 + margin && p[1] < y1 - margin; }
    Box3D get_bounding_box(plint) const override { return Box3D(0, 0, (plint)y0, (plint)y1, 0, 0); }
    bl_boundary desc() const override { return bl_boundary{BL_BOUNDARY_PLATES, {y0, y1, 0, 0}}; }

private:
    T y0, y1;
};

// Cells farther outside than `margin` (src/fsi/Boundary.h:114-128).  The
// margin is an integer, as in the reference (whose progs pass sqrt(2): 1).
template <class T>
class BoundaryOutside {
public:
    BoundaryOutside(const Boundary<T>& b, plint margin) : b_(b), margin_((T)margin) {}
    bool operator()(plint x, plint y, plint z) const { return !b_.contains(Array<T, 3>(x, y, z), -margin_); }

private:
    const Boundary<T>& b_;
    T margin_;
};

// Interaction potentials (src/fsi/Potentials.h:28-72): the force at a
// distance and the cut-off; desc() is what crosses the C-ABI.
// ParticlePositionInitializer (src/fsi/ParticlePositionInitializer.h:19-37,
// .hh:18-86): spheres drawn with the C library's rand() inside the box and
// the boundary, apart from every sphere placed before.
template <class T>
class ParticlePositionInitializer {
public:
    ParticlePositionInitializer(const Box3D& box, const Boundary<T>* boundary, T max_radius)
    {
        const int b[6] = {(int)box.x0, (int)box.x1, (int)box.y0, (int)box.y1, (int)box.z0, (int)box.z1};
        if (boundary) {
            const bl_boundary d = boundary->desc();
            blCheck(bl_positions_create(b, &d, max_radius, &h_), "ParticlePositionInitializer");
        } else {
            blCheck(bl_positions_create(b, nullptr, max_radius, &h_), "ParticlePositionInitializer");
        }
    }
    ~ParticlePositionInitializer() { bl_positions_destroy(h_); }
    ParticlePositionInitializer(const ParticlePositionInitializer&) = delete;
    ParticlePositionInitializer& operator=(const ParticlePositionInitializer&) = delete;
    bool generate_points(pluint N, T radius, char tag)
    {
        return bl_positions_generate(h_, (long long)N, radius, tag) == BL_OK;
    }
    void get_points(char tag, std::vector<Array<T, 3>>& points)
    {
        long long n = 0;
        blCheck(bl_positions_get(h_, tag, nullptr, 0, &n), "get_points");
        std::vector<double> buf(3 * (size_t)n);
        blCheck(bl_positions_get(h_, tag, buf.data(), n, &n), "get_points");
        for (long long k = 0; k < n; ++k) points.push_back(Array<T, 3>(buf[3 * k], buf[3 * k + 1], buf[3 * k + 2]));
    }

private:
    bl_positions h_ = nullptr;
};

template <class T>
class MorsePotential {
public:
    MorsePotential(T r_min_, T beta_, T strength_) : r_min(r_min_), beta(beta_), strength(strength_) {}
    T operator()(T dist) const
    {
        double f = 0;
        const bl_potential u = desc();
        blCheck(bl_potential_eval(&u, dist, &f, nullptr), "MorsePotential");
        return (T)f;
    }
    T get_cutoff_distance() const { return r_min; }
    T get_repulsion_distance() const { return r_min; }
    bl_potential desc() const { return bl_potential{BL_POTENTIAL_MORSE, {r_min, beta, strength}}; }

private:
    T r_min, beta, strength;
};

template <class T>
class SpringPotential {
public:
    SpringPotential(T min_dist_, T strength_) : min_dist(min_dist_), strength(strength_) {}
    T operator()(T dist) const { return (min_dist - dist) * strength; }
    T get_cutoff_distance() const { return min_dist; }
    T get_repulsion_distance() const { return min_dist; }
    bl_potential desc() const { return bl_potential{BL_POTENTIAL_SPRING, {min_dist, strength, 0}}; }

private:
    T min_dist, strength;
};

// Particle-particle forces (src/fsi/ParticleParticleInteraction.h:17-40).
template <class T>
struct PPInteractionForce {
    virtual ~PPInteractionForce() {}
    T operator()(T dist) const { return eval(dist); }
    virtual T eval(T dist) const = 0;
    virtual T get_cutoff_distance() const = 0;
    virtual bl_potential desc() const = 0;
};

template <class T, class Interaction>
class PPPotentialForce : public PPInteractionForce<T> {
public:
    explicit PPPotentialForce(const Interaction& potential) : potential_(potential) {}
    T eval(T dist) const override { return potential_(dist); }
    T get_cutoff_distance() const override { return potential_.get_cutoff_distance(); }
    bl_potential desc() const override { return potential_.desc(); }

private:
    Interaction potential_;
};

// Force decorators (src/fsi/ForceDecorator.h:12-27): applied to every
// particle before its membrane forces.  The device runs the wall
// interaction; a decorator without a device form is refused when added.
template <class T>
struct ForceDecorator {
    virtual ~ForceDecorator() {}
    virtual bool device_form(bl_boundary& b, bl_potential& u) const
    {
        (void)b;
        (void)u;
        return false;
    }
};

// WallInteraction(boundary, interaction) (src/fsi/ForceDecorator.hh:18-35):
// force -= F(d) n on vertices within the cut-off of the boundary; the
// boundary must outlive the decorator.
template <class T, class Interaction>
class WallInteraction : public ForceDecorator<T> {
public:
    WallInteraction(const Boundary<T>& boundary, const Interaction& interaction)
        : boundary_(&boundary), interaction_(interaction)
    {
    }
    bool device_form(bl_boundary& b, bl_potential& u) const override
    {
        b = boundary_->desc();
        u = interaction_.desc();
        return true;
    }

private:
    const Boundary<T>* boundary_;
    Interaction interaction_;
};

// GuoRigidWallBoundary(boundary, lattice) (src/fsi/GuoBoundaryCondition.h:
// 19-52): the wall nodes are found when it is constructed (the lattice is
// created then); insert() makes the extrapolation run after every
// collideAndStream and in initialize().
template <typename T, template <typename U> class Descriptor>
class GuoRigidWallBoundary {
public:
    struct NodeInfo {
        Array<plint, 3> location;
        T delta;
        Array<plint, 3> direction;
        Array<T, 3> wall_position;
        Array<T, 3> velocity;
    };
    typedef typename std::vector<NodeInfo>::const_iterator const_iterator;

    GuoRigidWallBoundary(const Boundary<T>& boundary, MultiBlockLattice3D<T, Descriptor>& lattice)
        : boundary_(boundary), lattice_(lattice)
    {
        const bl_boundary b = boundary.desc();
        blCheck(bl_guo_wall_create(lattice.handle(), &b, &w_), "GuoRigidWallBoundary");
        long long n = 0;
        blCheck(bl_guo_wall_count(w_, &n), "GuoRigidWallBoundary");
        std::vector<long long> cell(n);
        std::vector<int> dir(n);
        std::vector<double> delta(n), wp(3 * n);
        blCheck(bl_guo_wall_nodes(w_, cell.data(), dir.data(), delta.data(), wp.data()), "GuoRigidWallBoundary");
        static const int C[19][3] = {{0, 0, 0},  {-1, 0, 0}, {0, -1, 0},  {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
                                     {-1, 0, 1}, {0, -1, -1}, {0, -1, 1}, {1, 0, 0},  {0, 1, 0},  {0, 0, 1},   {1, 1, 0},
                                     {1, -1, 0}, {1, 0, 1},  {1, 0, -1}, {0, 1, 1},   {0, 1, -1}};
        const plint ny = lattice.getNy(), nz = lattice.getNz();
        nodes_.resize(n);
        for (long long k = 0; k < n; ++k) {
            NodeInfo& i = nodes_[k];
            i.location = Array<plint, 3>(cell[k] / (ny * nz), (cell[k] / nz) % ny, cell[k] % nz);
            i.direction = Array<plint, 3>(C[dir[k]][0], C[dir[k]][1], C[dir[k]][2]);
            i.delta = delta[k];
            i.wall_position = Array<T, 3>(wp[3 * k], wp[3 * k + 1], wp[3 * k + 2]);
        }
    }
    ~GuoRigidWallBoundary()
    {
        if (w_) bl_guo_wall_destroy(w_);
    }
    GuoRigidWallBoundary(const GuoRigidWallBoundary&) = delete;
    GuoRigidWallBoundary& operator=(const GuoRigidWallBoundary&) = delete;

    const Boundary<T>& boundary() const { return boundary_; }
    void insert() { blCheck(bl_guo_wall_insert(w_), "GuoRigidWallBoundary::insert"); }
    void set_wall_velocity(const Array<T, 3>& velocity)
    {
        const double v[3] = {velocity[0], velocity[1], velocity[2]};
        blCheck(bl_guo_wall_set_velocity(w_, v), "set_wall_velocity");
        for (NodeInfo& i : nodes_) i.velocity = velocity;
    }
    template <class VelocityFunction>
    void set_wall_velocity(VelocityFunction velocity_function)
    {
        std::vector<double> v;
        for (NodeInfo& i : nodes_) {
            i.velocity = velocity_function(i.wall_position);
            v.insert(v.end(), {i.velocity[0], i.velocity[1], i.velocity[2]});
        }
        blCheck(bl_guo_wall_set_node_velocities(w_, v.data()), "set_wall_velocity");
    }
    const_iterator begin() const { return nodes_.begin(); }
    const_iterator end() const { return nodes_.end(); }

private:
    const Boundary<T>& boundary_;
    MultiBlockLattice3D<T, Descriptor>& lattice_;
    std::vector<NodeInfo> nodes_;
    bl_guo_wall w_ = nullptr;
};

template <class T>
class ParticleShape {
public:
    std::vector<Array<T, 3>> vertices;
    std::vector<Array<plint, 3>> triangles;
    std::string tag;
    plint count_nodes() const { return (plint)vertices.size(); }
    // largest vertex distance from the origin (src/fsi/ParticleShape.hh:361-392)
    T get_radius() const
    {
        T r = 0;
        for (const auto& v : vertices) r = std::max(r, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        return r;
    }
};

template <class T>
class ParticleShapeLibrary {
public:
    const ParticleShape<T>* store_mesh(const std::vector<Array<T, 3>>& v, const std::vector<Array<plint, 3>>& t,
                                       const std::string& tag)
    {
        std::unique_ptr<ParticleShape<T>> s(new ParticleShape<T>());
        s->vertices = v;
        s->triangles = t;
        s->tag = tag;
        if (!shapes_.count(tag)) order_.push_back(tag);
        shapes_[tag] = std::move(s);
        return shapes_[tag].get();
    }
    // gmsh ASCII 2.x: $Nodes and the 3-node triangles (type 2) of $Elements
    // (src/fsi/ParticleShapeFactory.hh:32-83)
    const ParticleShape<T>* read_and_store_mesh(const std::string& file, const std::string& tag)
    {
        std::ifstream in(file.c_str());
        if (!in) throw std::runtime_error("read_and_store_mesh: cannot open " + file);
        std::string line;
        std::vector<Array<T, 3>> v;
        std::vector<Array<plint, 3>> t;
        std::map<long, plint> id;
        while (std::getline(in, line)) {
            if (line == "$Nodes") {
                long n;
                in >> n;
                for (long k = 0; k < n; ++k) {
                    long i;
                    T x, y, z;
                    in >> i >> x >> y >> z;
                    id[i] = (plint)v.size();
                    v.push_back(Array<T, 3>(x, y, z));
                }
            } else if (line == "$Elements") {
                long n;
                in >> n;
                std::getline(in, line);
                for (long k = 0; k < n; ++k) {
                    std::getline(in, line);
                    std::istringstream ls(line);
                    long num, type, ntags;
                    ls >> num >> type >> ntags;
                    for (long q = 0; q < ntags; ++q) {
                        long skip;
                        ls >> skip;
                    }
                    if (type != 2) continue;
                    long a, b, c;
                    ls >> a >> b >> c;
                    t.push_back(Array<plint, 3>(id.at(a), id.at(b), id.at(c)));
                }
            }
        }
        if (v.empty() || t.empty()) throw std::runtime_error("read_and_store_mesh: no triangle mesh in " + file);
        return store_mesh(v, t, tag);
    }
    const ParticleShape<T>* get_by_tag(const std::string& tag) const { return shapes_.at(tag).get(); }
    // shape ids are the storage order (the reference library's get_id)
    plint count() const { return (plint)order_.size(); }
    const ParticleShape<T>* get_by_id(plint id) const { return shapes_.at(order_.at((size_t)id)).get(); }

private:
    std::map<std::string, std::unique_ptr<ParticleShape<T>>> shapes_;
    std::vector<std::string> order_;
};

template <class T>
struct CapsuleParameters {
    T G = 0, C = 0, nu_s = 0, kb = 0;
};

template <class T>
struct RBCParameters {
    T shear_modulus = 0, L0 = 0, k_area_global = 0, k_area_local = 0, k_volume = 0, vol_desired = 0, k_bend = 0,
      theta0 = 0;
    T G() const { return shear_modulus; }
};

template <class T>
class ParticleBase3D {
public:
    explicit ParticleBase3D(const ParticleShape<T>* shape) : shape_(shape), pos_(shape->vertices) {}
    virtual ~ParticleBase3D() {}
    const ParticleShape<T>* get_shape() const { return shape_; }
    plint count_nodes() const { return (plint)pos_.size(); }
    const Array<T, 3>& get_vertex(plint i) const { return pos_[i]; }
    std::vector<Array<T, 3>>& positions() { return pos_; }
    const std::vector<Array<T, 3>>& positions() const { return pos_; }
    // area-weighted centroid of the surface (src/fsi/DeformableParticle3D.hh:54-82)
    Array<T, 3> center_of_mass() const
    {
        T A = 0;
        Array<T, 3> c;
        for (const auto& tr : shape_->triangles) {
            const Array<T, 3>&a = pos_[tr[0]], &b = pos_[tr[1]], &d = pos_[tr[2]];
            const T e0[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const T e1[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
            const T n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
            const T ar = T(0.5) * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            A += ar;
            for (int k = 0; k < 3; ++k) c[k] += ar * (a[k] + b[k] + d[k]) / T(3);
        }
        for (int k = 0; k < 3; ++k) c[k] /= A;
        return c;
    }
    // src/fsi/ParticleBase.hh:121-140
    void set_center_of_mass(const Array<T, 3>& target)
    {
        const Array<T, 3> c = center_of_mass();
        for (auto& p : pos_)
            for (int k = 0; k < 3; ++k) p[k] += target[k] - c[k];
    }
    // rotation by `angle` about `axis` (right-handed) around the centre of mass
    void rotate(const Array<T, 3>& axis, T angle)
    {
        const T nrm = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
        const T k[3] = {axis[0] / nrm, axis[1] / nrm, axis[2] / nrm};
        const T cs = std::cos(angle), sn = std::sin(angle);
        const Array<T, 3> c = center_of_mass();
        for (auto& p : pos_) {
            const T v[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
            const T kv = k[0] * v[0] + k[1] * v[1] + k[2] * v[2];
            const T kx[3] = {k[1] * v[2] - k[2] * v[1], k[2] * v[0] - k[0] * v[2], k[0] * v[1] - k[1] * v[0]};
            for (int d = 0; d < 3; ++d) p[d] = c[d] + v[d] * cs + kx[d] * sn + k[d] * kv * (1 - cs);
        }
    }
    // aligns the gyration tensor's major axis with `pref`
    // (src/fsi/DeformableParticle3D.hh:100-170)
    void set_major_axis_orientation(const Array<T, 3>& pref)
    {
        const Array<T, 3> c = center_of_mass();
        T A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (const auto& p : pos_)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) A[i][j] += (p[i] - c[i]) * (p[j] - c[j]) / (T)pos_.size();
        Array<T, 3> dir = major_axis(A);
        Array<T, 3> ax(dir[1] * pref[2] - dir[2] * pref[1], dir[2] * pref[0] - dir[0] * pref[2],
                       dir[0] * pref[1] - dir[1] * pref[0]);
        const T na = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        const T np = std::sqrt(pref[0] * pref[0] + pref[1] * pref[1] + pref[2] * pref[2]);
        if (na == T(0)) return;
        T angle = std::asin(na / np);
        if (dir[0] * pref[0] + dir[1] * pref[1] + dir[2] * pref[2] < 0)
            for (int k = 0; k < 3; ++k) ax[k] = -ax[k];
        rotate(ax, angle);
    }
    virtual int model() const = 0;
    virtual std::vector<double> params() const = 0;
    virtual int bug_compatible() const { return 1; }

private:
    // eigenvector of the largest |eigenvalue| of a symmetric 3x3 matrix (cyclic Jacobi)
    static Array<T, 3> major_axis(T A[3][3])
    {
        T V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int sweep = 0; sweep < 50; ++sweep) {
            const T off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
            if (off < T(1e-30)) break;
            for (int p = 0; p < 2; ++p)
                for (int q = p + 1; q < 3; ++q) {
                    if (A[p][q] == T(0)) continue;
                    const T th = (A[q][q] - A[p][p]) / (2 * A[p][q]);
                    const T t = (th >= 0 ? 1 : -1) / (std::abs(th) + std::sqrt(th * th + 1));
                    const T cs = 1 / std::sqrt(t * t + 1), sn = t * cs;
                    for (int k = 0; k < 3; ++k) {
                        const T akp = A[k][p], akq = A[k][q];
                        A[k][p] = cs * akp - sn * akq;
                        A[k][q] = sn * akp + cs * akq;
                    }
                    for (int k = 0; k < 3; ++k) {
                        const T apk = A[p][k], aqk = A[q][k];
                        A[p][k] = cs * apk - sn * aqk;
                        A[q][k] = sn * apk + cs * aqk;
                    }
                    for (int k = 0; k < 3; ++k) {
                        const T vkp = V[k][p], vkq = V[k][q];
                        V[k][p] = cs * vkp - sn * vkq;
                        V[k][q] = sn * vkp + cs * vkq;
                    }
                }
        }
        int k = 0;
        for (int i = 1; i < 3; ++i)
            if (std::abs(A[i][i]) > std::abs(A[k][k])) k = i;
        const T n = std::sqrt(V[0][k] * V[0][k] + V[1][k] * V[1][k] + V[2][k] * V[2][k]);
        return Array<T, 3>(V[0][k] / n, V[1][k] / n, V[2][k] / n);
    }

protected:
    const ParticleShape<T>* shape_;
    std::vector<Array<T, 3>> pos_;
};

// DeformableCapsuleParticle (src/fsi/DeformableCapsule.h:33-64): NeoHookean
// unless the Skalak law is selected (the reference picks it through its
// serialized state; here set_shear_model).
template <class T>
class DeformableCapsuleParticle : public ParticleBase3D<T> {
public:
    enum ShearModel { Skalak, NeoHookean };
    DeformableCapsuleParticle(const ParticleShape<T>* s, const CapsuleParameters<T>& p)
        : ParticleBase3D<T>(s), p_(p)
    {
    }
    void set_shear_model(ShearModel m) { model_ = m; }
    int model() const override { return model_ == Skalak ? BL_MODEL_CAPSULE_SKALAK : BL_MODEL_CAPSULE_NEOHOOKEAN; }
    std::vector<double> params() const override { return {p_.G, p_.C}; }
    const CapsuleParameters<T>& get_params() const { return p_; }

private:
    CapsuleParameters<T> p_;
    ShearModel model_ = NeoHookean;
};

template <class T>
class RBCParticle : public ParticleBase3D<T> {
public:
    RBCParticle(const ParticleShape<T>* s, const RBCParameters<T>& p) : ParticleBase3D<T>(s), p_(p) {}
    int model() const override { return BL_MODEL_RBC; }
    std::vector<double> params() const override
    {
        return {p_.shear_modulus, p_.L0, p_.k_area_global, p_.k_area_local, p_.k_volume, p_.vol_desired, p_.k_bend,
                p_.theta0};
    }
    RBCParameters<T>& params_ref() { return p_; }
    const RBCParameters<T>& params_ref() const { return p_; }

private:
    RBCParameters<T> p_;
};

// create_rbc_params (progs/common/calibrate_rbc_shape.h:75-112): L0 the mean
// link length (distinct edges in (i0, i1) order), vol_desired the rest
// volume, the area stiffness beyond K() = 2 G split 0.8 / 0.2 between
// global and local, k_volume = 1000 G / radius, theta0 of a regular
// triangulation with the shape's vertex count.
template <class T>
RBCParameters<T> create_rbc_params(const ParticleShape<T>* shape, T shear_modulus, T K_area, T K_bend)
{
    std::vector<std::pair<plint, plint>> edges;
    for (const auto& t : shape->triangles)
        for (int k = 0; k < 3; ++k) {
            const plint a = t[k], b = t[(k + 1) % 3];
            edges.push_back(std::make_pair(std::min(a, b), std::max(a, b)));
        }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    T lsum = 0;
    for (const auto& e : edges) {
        const Array<T, 3>&a = shape->vertices[e.first], &b = shape->vertices[e.second];
        lsum += std::sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]) + (b[2] - a[2]) * (b[2] - a[2]));
    }
    T vol = 0;
    for (const auto& t : shape->triangles) {
        const Array<T, 3>&a = shape->vertices[t[0]], &b = shape->vertices[t[1]], &c = shape->vertices[t[2]];
        vol += (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                a[2] * (b[0] * c[1] - b[1] * c[0])) / 6;
    }
    const T nv = (T)shape->vertices.size(), s3 = std::sqrt((T)3);
    RBCParameters<T> p;
    p.shear_modulus = shear_modulus;
    p.L0 = lsum / (T)edges.size();
    p.vol_desired = vol;
    const T delta = K_area - 2 * shear_modulus;
    p.k_area_global = delta * 0.8;
    p.k_area_local = delta * 0.2;
    p.k_volume = 1000. * shear_modulus / shape->get_radius();
    p.k_bend = K_bend;
    p.theta0 = std::acos((s3 * (nv - 2.) - 5 * M_PI) / (s3 * (nv - 2.) - 3. * M_PI));
    return p;
}

// shrink_rbc_volume(rbc, vol_final, N_it, m, damping)
// (progs/common/calibrate_rbc_shape.h:13-72) on the device: the particle's
// vertices relax while vol_desired ramps to vol_final; on return the
// particle holds the calibrated vertices and vol_desired.  A scratch
// lattice and immersed boundary carry the device state.
template <class T>
void shrink_rbc_volume(RBCParticle<T>& rbc, T vol_final, int N_it, T m = 1, T damping = 1.e-2)
{
    (void)m;
    bl_lattice L = nullptr;
    bl_ibm g = nullptr;
    const int per[3] = {0, 0, 0};
    auto cleanup = [&]() {
        if (g) bl_ibm_destroy(g);
        if (L) bl_lattice_destroy(L);
    };
    try {
        blCheck(bl_lattice_create(4, 4, 4, per, BL_STORE_F64, &L), "shrink_rbc_volume");
        blCheck(bl_ibm_create(L, BL_KERNEL_PESKIN, &g), "shrink_rbc_volume");
        const ParticleShape<T>* s = rbc.get_shape();
        std::vector<double> v, x;
        std::vector<int> t;
        for (const auto& a : s->vertices) v.insert(v.end(), {a[0], a[1], a[2]});
        for (const auto& a : s->triangles) t.insert(t.end(), {(int)a[0], (int)a[1], (int)a[2]});
        for (const auto& a : rbc.positions()) x.insert(x.end(), {a[0], a[1], a[2]});
        int sid = -1, pid = -1;
        blCheck(bl_shape_add(g, (int)s->vertices.size(), (int)s->triangles.size(), v.data(), t.data(), &sid),
                "shrink_rbc_volume");
        const std::vector<double> prm = rbc.params();
        blCheck(bl_particle_add(g, sid, BL_MODEL_RBC, prm.data(), (int)prm.size(), x.data(), 1, &pid),
                "shrink_rbc_volume");
        blCheck(bl_ibm_init(g), "shrink_rbc_volume");
        blCheck(bl_rbc_shrink_volume(g, pid, vol_final, N_it, damping, nullptr), "shrink_rbc_volume");
        blCheck(bl_particle_get(g, pid, 0, x.data()), "shrink_rbc_volume");
        for (size_t k = 0; k < rbc.positions().size(); ++k)
            rbc.positions()[k] = Array<T, 3>(x[3 * k], x[3 * k + 1], x[3 * k + 2]);
        rbc.params_ref().vol_desired = vol_final;
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

template <typename T, template <typename U> class Descriptor, class Periodicity = void>
class ImmersedBoundaryDynamics3D {
public:
    ImmersedBoundaryDynamics3D(MultiBlockLattice3D<T, Descriptor>& lattice, const ParticleShapeLibrary<T>& lib,
                               int kernel = BL_KERNEL_ROMA)
        : lattice_(lattice), lib_(lib), kernel_(kernel)
    {
    }
    ~ImmersedBoundaryDynamics3D()
    {
        if (g_) bl_ibm_destroy(g_);
    }
    ImmersedBoundaryDynamics3D(const ImmersedBoundaryDynamics3D&) = delete;
    ImmersedBoundaryDynamics3D& operator=(const ImmersedBoundaryDynamics3D&) = delete;

    // the particle is copied (its shape, model, parameters and positions)
    plint add_particle(const ParticleBase3D<T>* p)
    {
        if (g_) throw std::runtime_error("add_particle: particles must be added before init()");
        parts_.push_back(p);
        return (plint)parts_.size() - 1;
    }
    // load_checkpoint (src/fsi/ImmersedBoundaryDynamics_io.hh:75-158): the
    // particles of <name>.plb (and its rank file) replace the added ones at
    // init(); shape ids refer to the library's storage order
    void load_checkpoint(const std::string& name)
    {
        if (g_) throw std::runtime_error("load_checkpoint: call before init()");
        checkpoint_ = name;
    }
    // save_checkpoint (src/fsi/ImmersedBoundaryDynamics_io.hh:17-73):
    // <name>.plb and <name>/0.plb
    void save_checkpoint(const std::string& name)
    {
        if (!g_) throw std::runtime_error("save_checkpoint: call init() first");
        lattice_.flushPendingBoundaryStep();
        makeDir(name);
        blCheck(bl_ibm_save_checkpoint(g_, ".", name.c_str()), "save_checkpoint");
    }
    void init()
    {
        blCheck(bl_ibm_create(lattice_.handle(), kernel_, &g_), "ImmersedBoundaryDynamics3D");
        std::map<const ParticleShape<T>*, int> sid;
        auto add_shape = [&](const ParticleShape<T>* s) {
            if (sid.count(s)) return;
            std::vector<double> v;
            std::vector<int> t;
            for (const auto& x : s->vertices) v.insert(v.end(), {x[0], x[1], x[2]});
            for (const auto& x : s->triangles) t.insert(t.end(), {(int)x[0], (int)x[1], (int)x[2]});
            int id = -1;
            blCheck(bl_shape_add(g_, (int)s->vertices.size(), (int)s->triangles.size(), v.data(), t.data(), &id),
                    "store_mesh");
            blCheck(bl_shape_set_tag(g_, id, s->tag.c_str()), "store_mesh");
            sid[s] = id;
            shape_nv_.push_back((plint)s->vertices.size());
        };
        for (plint k = 0; k < lib_.count(); ++k) add_shape(lib_.get_by_id(k));
        if (!checkpoint_.empty()) {
            blCheck(bl_ibm_load_checkpoint(g_, ".", checkpoint_.c_str()), "load_checkpoint");
            int n = 0;
            blCheck(bl_ibm_count(g_, &n, nullptr), "load_checkpoint");
            for (int p = 0; p < n; ++p) {
                int s = -1, model = -1;
                long long id = -1;
                blCheck(bl_particle_describe(g_, p, &s, &model, &id), "load_checkpoint");
                nv_.push_back(shape_nv_.at((size_t)s));
            }
            parts_.clear();
        }
        for (const ParticleBase3D<T>* p : parts_) {
            const ParticleShape<T>* s = p->get_shape();
            add_shape(s);
            nv_.push_back(p->count_nodes());
            std::vector<double> pos;
            for (const auto& x : p->positions()) pos.insert(pos.end(), {x[0], x[1], x[2]});
            const std::vector<double> prm = p->params();
            int pid = -1;
            blCheck(bl_particle_add(g_, sid[s], p->model(), prm.data(), (int)prm.size(), pos.data(),
                                    p->bug_compatible(), &pid),
                    "add_particle");
        }
        blCheck(bl_ibm_init(g_), "init");
        if (boundary_) set_boundary(boundary_);
        for (const ForceDecorator<T>* d : decorators_) push_decorator(d);
        for (const PPInteractionForce<T>* f : pp_forces_) push_pp_force(f);
    }
    bl_ibm handle() const { return g_; }

    // add_force_decorator / add_pp_force (src/fsi/ImmersedBoundaryDynamics.h:
    // 78-79): applied in every step's force computation, in the order added;
    // the objects must outlive this one
    void add_force_decorator(ForceDecorator<T>* d)
    {
        decorators_.push_back(d);
        if (g_) push_decorator(d);
    }
    void add_pp_force(PPInteractionForce<T>* f)
    {
        pp_forces_.push_back(f);
        if (g_) push_pp_force(f);
    }
    // compute_collision_forces(interaction, boundary)
    // (src/fsi/ImmersedBoundaryDynamics_forces.hh:505-521): vertex forces :=
    // the collision handler's forces alone (read them with get_forces)
    template <class InteractionFunctional>
    void compute_collision_forces(const InteractionFunctional& interaction, const Boundary<T>* boundary = 0)
    {
        if (!g_) throw std::runtime_error("compute_collision_forces: call init() first");
        lattice_.flushPendingBoundaryStep();
        const bl_potential u = interaction.desc();
        if (boundary) {
            const bl_boundary b = boundary->desc();
            blCheck(bl_ibm_compute_collision_forces(g_, &u, &b), "compute_collision_forces");
        } else {
            blCheck(bl_ibm_compute_collision_forces(g_, &u, nullptr), "compute_collision_forces");
        }
    }
    // current vertex forces of particle i into a host copy
    bool get_forces(plint i, std::vector<Array<T, 3>>& f)
    {
        if (i < 0 || i >= (plint)nv_.size()) return false;
        lattice_.flushPendingBoundaryStep();
        std::vector<double> buf(3 * nv_[i]);
        blCheck(bl_particle_get(g_, (int)i, 2, buf.data()), "get_forces");
        f.resize(nv_[i]);
        for (size_t k = 0; k < f.size(); ++k) f[k] = Array<T, 3>(buf[3 * k], buf[3 * k + 1], buf[3 * k + 2]);
        return true;
    }

    // set_boundary (src/fsi/ImmersedBoundaryDynamics.h:93): near-wall kernel
    // blending and vertex reflection; the boundary must outlive this object
    void set_boundary(const Boundary<T>* b)
    {
        boundary_ = b;
        if (!g_) return;
        lattice_.flushPendingBoundaryStep();
        if (b) {
            const bl_boundary d = b->desc();
            blCheck(bl_ibm_set_boundary(g_, &d), "set_boundary");
        } else {
            blCheck(bl_ibm_set_boundary(g_, nullptr), "set_boundary");
        }
    }

    // current vertices of particle i into a host copy (src/fsi/ImmersedBoundaryDynamics.h:95)
    bool get_particle(plint i, std::vector<Array<T, 3>>& pos)
    {
        if (i < 0 || i >= (plint)nv_.size()) return false;
        lattice_.flushPendingBoundaryStep();
        std::vector<double> buf(3 * nv_[i]);
        blCheck(bl_particle_get(g_, (int)i, 0, buf.data()), "get_particle");
        pos.resize(nv_[i]);
        for (size_t k = 0; k < pos.size(); ++k) pos[k] = Array<T, 3>(buf[3 * k], buf[3 * k + 1], buf[3 * k + 2]);
        return true;
    }
    // area, volume, centre of mass, inertia tensor of the enclosed volume
    // (src/fsi/DeformableParticle3D.hh:54-82,173-272)
    void observables(plint i, T& area, T& volume, Array<T, 3>& com, T inertia[9])
    {
        lattice_.flushPendingBoundaryStep();
        double o[14];
        blCheck(bl_particle_observables(g_, (int)i, o), "observables");
        area = o[0];
        volume = o[1];
        com = Array<T, 3>(o[2], o[3], o[4]);
        for (int k = 0; k < 9; ++k) inertia[k] = o[5 + k];
    }
    T compute_total_volume()
    {
        T V = 0, a;
        Array<T, 3> c;
        T I[9];
        for (plint i = 0; i < (plint)nv_.size(); ++i) {
            T v;
            observables(i, a, v, c, I);
            V += v;
        }
        return V;
    }
    plint count_particles() const { return (plint)nv_.size(); }
    // write_particles_as_vtk (src/fsi/ImmersedBoundaryDynamics_io.hh:160-176):
    // obj<id>-<it>.vtu per particle in the folder `dir` (the reference's
    // output directory)
    void write_particles_as_vtk(plint it, const std::string& dir = "")
    {
        lattice_.flushPendingBoundaryStep();
        for (plint i = 0; i < (plint)nv_.size(); ++i) {
            int s = -1, model = -1;
            long long id = -1;
            blCheck(bl_particle_describe(g_, (int)i, &s, &model, &id), "write_particles_as_vtk");
            std::ostringstream name;
            name << dir << "obj" << id << "-" << it << ".vtu";
            blCheck(bl_particle_write_vtk(g_, (int)i, name.str().c_str()), "write_particles_as_vtk");
        }
    }
    // write_lightweight_particle_data (src/fsi/ImmersedBoundaryDynamics_io.hh:
    // 178-200): objsP0-<it>.txt in `dir`
    void write_lightweight_particle_data(plint it, const std::string& dir = "")
    {
        lattice_.flushPendingBoundaryStep();
        std::ostringstream name;
        name << dir << "objsP0-" << it << ".txt";
        blCheck(bl_ibm_write_lightweight(g_, name.str().c_str(), (long long)it), "write_lightweight_particle_data");
    }
    void schedule_step() { lattice_.setPendingBoundaryStep(g_); }

private:
    MultiBlockLattice3D<T, Descriptor>& lattice_;
    const ParticleShapeLibrary<T>& lib_;
    int kernel_;
    std::string checkpoint_;
    std::vector<plint> shape_nv_, nv_;  // vertices per shape id / per particle
    void push_decorator(const ForceDecorator<T>* d)
    {
        bl_boundary b;
        bl_potential u;
        if (!d->device_form(b, u))
            throw std::runtime_error("add_force_decorator: only wall-interaction decorators run on the device");
        lattice_.flushPendingBoundaryStep();
        blCheck(bl_ibm_add_wall_interaction(g_, &b, &u), "add_force_decorator");
    }
    void push_pp_force(const PPInteractionForce<T>* f)
    {
        const bl_potential u = f->desc();
        lattice_.flushPendingBoundaryStep();
        blCheck(bl_ibm_add_pp_force(g_, &u), "add_pp_force");
    }

    std::vector<const ParticleBase3D<T>*> parts_;
    std::vector<const ForceDecorator<T>*> decorators_;
    std::vector<const PPInteractionForce<T>*> pp_forces_;
    const Boundary<T>* boundary_ = nullptr;
    bl_ibm g_ = nullptr;
};

// wrap_ibm_dynamics3D (src/fsi/DataProcessors.hh:129-131): the per-step
// functional; applyProcessingFunctional takes ownership and runs it
template <typename T, template <typename U> class Descriptor, class Periodicity>
struct ImmersedBoundaryWrapperFunctional3D {
    ImmersedBoundaryDynamics3D<T, Descriptor, Periodicity>& ibm;
};

template <typename T, template <typename U> class Descriptor, class Periodicity>
ImmersedBoundaryWrapperFunctional3D<T, Descriptor, Periodicity>* wrap_ibm_dynamics3D(
    ImmersedBoundaryDynamics3D<T, Descriptor, Periodicity>& ibm)
{
    return new ImmersedBoundaryWrapperFunctional3D<T, Descriptor, Periodicity>{ibm};
}

}  // namespace fsi

// The immersed-boundary step is scheduled and runs fused with the next
// collideAndStream (or when particle state is read first).
template <typename T, template <typename U> class Descriptor, class Periodicity, class Field>
void applyProcessingFunctional(fsi::ImmersedBoundaryWrapperFunctional3D<T, Descriptor, Periodicity>* f, const Box3D&,
                               MultiBlockLattice3D<T, Descriptor>&, Field&)
{
    std::unique_ptr<fsi::ImmersedBoundaryWrapperFunctional3D<T, Descriptor, Periodicity>> own(f);
    own->ibm.schedule_step();
}

// parallelIO::save / load of the lattice (progs/rbc_in_pipe/rbc_in_pipe.cpp:
// 321-322,693): <name>.blc in this library's checkpoint layout (populations
// as stored, step count, force field), not Palabos' serializer format.
namespace parallelIO {
template <typename T, template <typename U> class Descriptor>
void save(MultiBlockLattice3D<T, Descriptor>& lattice, const std::string& name, bool = false)
{
    lattice.flushPendingBoundaryStep();
    blCheck(bl_lattice_save_checkpoint(lattice.handle(), (name + ".blc").c_str()), "parallelIO::save");
}
template <typename T, template <typename U> class Descriptor>
void load(const std::string& name, MultiBlockLattice3D<T, Descriptor>& lattice, bool = false)
{
    blCheck(bl_lattice_load_checkpoint(lattice.handle(), (name + ".blc").c_str()), "parallelIO::load");
}
}  // namespace parallelIO

namespace fsi {
// Iteration counter with its checkpoint (src/fsi/Time.h:15-48).
class Time {
public:
    Time() : it_(0) {}
    Time(plint it) : it_(it) {}
    void write_checkpoint(const std::string& file) const
    {
        std::ofstream out(file.c_str());
        out << it_;
    }
    void read_checkpoint(const std::string& file)
    {
        std::ifstream in(file.c_str());
        if (!(in >> it_)) throw std::runtime_error("Time::read_checkpoint: cannot read " + file);
    }
    bool is_multiple_of(plint n) const { return (it_ % n) == 0; }
    bool operator<(plint it) const { return it_ < it; }
    bool operator<=(plint it) const { return it_ <= it; }
    bool operator==(plint it) const { return it_ == it; }
    bool operator!=(plint it) const { return it_ != it; }
    Time& operator++()
    {
        ++it_;
        return *this;
    }
    plint to_plint() const { return it_; }

private:
    plint it_;
};
}  // namespace fsi

}  // namespace plb

#endif
