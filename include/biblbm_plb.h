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
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
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
    Array(T a, T b, T c, T e) : d{a, b, c, e} {}
    T& operator[](int i) { return d[i]; }
    const T& operator[](int i) const { return d[i]; }
    void resetToZero()
    {
        for (int k = 0; k < n; ++k) d[k] = T();
    }
    Array& operator+=(const Array& o)
    {
        for (int k = 0; k < n; ++k) d[k] += o.d[k];
        return *this;
    }
    Array& operator-=(const Array& o)
    {
        for (int k = 0; k < n; ++k) d[k] -= o.d[k];
        return *this;
    }
    Array& operator*=(T s)
    {
        for (int k = 0; k < n; ++k) d[k] *= s;
        return *this;
    }
    Array& operator/=(T s)
    {
        for (int k = 0; k < n; ++k) d[k] /= s;
        return *this;
    }
};

// Array arithmetic and the vector helpers the progs use (Palabos'
// src/core/array.h, src/core/util.h)
template <typename T, int n>
Array<T, n> operator+(Array<T, n> a, const Array<T, n>& b) { return a += b; }
template <typename T, int n>
Array<T, n> operator-(Array<T, n> a, const Array<T, n>& b) { return a -= b; }
template <typename T, int n>
Array<T, n> operator-(Array<T, n> a)
{
    for (int k = 0; k < n; ++k) a[k] = -a[k];
    return a;
}
template <typename T, int n>
Array<T, n> operator*(Array<T, n> a, T s) { return a *= s; }
template <typename T, int n>
Array<T, n> operator*(T s, Array<T, n> a) { return a *= s; }
template <typename T, int n>
Array<T, n> operator/(Array<T, n> a, T s) { return a /= s; }
template <typename T, int n>
T dot(const Array<T, n>& a, const Array<T, n>& b)
{
    T r = T();
    for (int k = 0; k < n; ++k) r += a[k] * b[k];
    return r;
}
template <typename T, int n>
T norm(const Array<T, n>& a) { return std::sqrt(dot(a, a)); }
template <typename T>
Array<T, 3> crossProduct(const Array<T, 3>& a, const Array<T, 3>& b)
{
    return Array<T, 3>(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}

struct Dot3D {
    plint x = 0, y = 0, z = 0;
    Dot3D() {}
    Dot3D(plint x_, plint y_, plint z_) : x(x_), y(y_), z(z_) {}
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
    plint nCells() const { return getNx() * getNy() * getNz(); }
    Box3D shift(plint dx, plint dy, plint dz) const { return Box3D(x0 + dx, x1 + dx, y0 + dy, y1 + dy, z0 + dz, z1 + dz); }
};

// ------------------------------------------------ runtime (single process)
// Errors reading arguments or configuration (src/core/runTimeDiagnostics.h:92;
// the progs catch it around global::argv and XMLreader reads).
class PlbIOException : public std::exception {
public:
    explicit PlbIOException(const std::string& m) : m_(m) {}
    const char* what() const noexcept override { return m_.c_str(); }

private:
    std::string m_;
};

// pcout / pcerr: the rank-0 streams of the reference; one process here.
static std::ostream& pcout = std::cout;
static std::ostream& pcerr = std::cerr;

namespace global {
struct ProgramArgs {
    std::vector<std::string> argv;
};
inline ProgramArgs& program_args()
{
    static ProgramArgs a;
    return a;
}
inline int argc() { return (int)program_args().argv.size(); }

// global::argv(i).read(value): the i-th command-line word parsed as the
// value's type; PlbIOException when it is missing or does not parse.
class ArgvReader {
public:
    ArgvReader(int i, bool ok, const std::string& s) : i_(i), ok_(ok), s_(s) {}
    template <typename V>
    void read(V& v) const
    {
        if (!ok_) throw PlbIOException("Index " + std::to_string(i_) + " of command-line arguments out of range");
        std::istringstream is(s_);
        is >> v;
        if (is.fail()) throw PlbIOException("Cannot read command-line argument " + std::to_string(i_) + ": " + s_);
    }
    void read(std::string& v) const
    {
        if (!ok_) throw PlbIOException("Index " + std::to_string(i_) + " of command-line arguments out of range");
        v = s_;
    }

private:
    int i_;
    bool ok_;
    std::string s_;
};
inline ArgvReader argv(int i)
{
    const auto& a = program_args().argv;
    return ArgvReader(i, i >= 0 && i < (int)a.size(), i >= 0 && i < (int)a.size() ? a[i] : std::string());
}

class Directories {
public:
    void setOutputDir(const std::string& d) { out_ = d; }
    void setInputDir(const std::string& d) { in_ = d; }
    void setLogOutDir(const std::string& d) { log_ = d; }
    const std::string& getOutputDir() const { return out_; }
    const std::string& getInputDir() const { return in_; }
    const std::string& getLogOutDir() const { return log_; }

private:
    std::string out_ = "./", in_ = "./", log_ = "./";
};
inline Directories& directories()
{
    static Directories d;
    return d;
}

class IOpolicyClass {
public:
    void activateParallelIO(bool v) { parallel_ = v; }
    bool useParallelIO() const { return parallel_; }

private:
    bool parallel_ = false;
};
inline IOpolicyClass& IOpolicy()
{
    static IOpolicyClass p;
    return p;
}

// The reference's MPI manager for one process: rank 0 of 1, broadcasts and
// barriers have nothing to do.
class MpiManager {
public:
    int getRank() const { return 0; }
    int getSize() const { return 1; }
    bool isMainProcessor() const { return true; }
    template <typename V>
    void bCast(V*, int, int = 0) const
    {
    }
    void barrier() const {}
};
inline MpiManager& mpi()
{
    static MpiManager m;
    return m;
}
}  // namespace global

// plbInit (src/core/plbInit.cpp:38): keeps the command line and makes the
// first device current.
inline void plbInit(int* argc, char*** argv, bool = false)
{
    auto& a = global::program_args().argv;
    a.clear();
    if (argc && argv)
        for (int i = 0; i < *argc; ++i) a.push_back((*argv)[i]);
    blCheck(bl_init(0), "plbInit");
}

namespace io {
inline void mkdir(const char* path) { ::mkdir(path, 0755); }
}  // namespace io

// createFileName(name, number, width): name followed by the zero-padded
// number (src/io/imageWriter.h:131)
inline std::string createFileName(const std::string& name, plint number, int width)
{
    std::ostringstream s;
    s << name;
    s.width(width);
    s.fill('0');
    s << number;
    return s.str();
}

// Processing-functional vocabulary of the reference
// (src/atomicBlock/dataProcessingFunctional3D.h): progs derive their own
// functionals from these bases; applyProcessingFunctional runs them on the
// host copy of a field.
namespace BlockDomain {
enum DomainT { bulk, envelope, bulkAndEnvelope };
}
namespace modif {
enum ModifT { nothing, staticVariables, dynamicVariables, allVariables, dataStructure, undefined };
}

// A host field block (the TensorField3D / ScalarField3D a functional
// receives): cell (x, y, z) of the box at getLocation().
template <typename T, int nDim>
class TensorField3D {
public:
    TensorField3D(plint nx, plint ny, plint nz, Array<T, nDim>* data) : nx_(nx), ny_(ny), nz_(nz), d_(data) {}
    Array<T, nDim>& get(plint x, plint y, plint z) { return d_[(x * ny_ + y) * nz_ + z]; }
    const Array<T, nDim>& get(plint x, plint y, plint z) const { return d_[(x * ny_ + y) * nz_ + z]; }
    Dot3D getLocation() const { return Dot3D(0, 0, 0); }
    Box3D getBoundingBox() const { return Box3D(0, nx_ - 1, 0, ny_ - 1, 0, nz_ - 1); }

private:
    plint nx_, ny_, nz_;
    Array<T, nDim>* d_;
};

template <typename T>
class ScalarField3D {
public:
    ScalarField3D(plint nx, plint ny, plint nz, T* data) : nx_(nx), ny_(ny), nz_(nz), d_(data) {}
    T& get(plint x, plint y, plint z) { return d_[(x * ny_ + y) * nz_ + z]; }
    const T& get(plint x, plint y, plint z) const { return d_[(x * ny_ + y) * nz_ + z]; }
    Dot3D getLocation() const { return Dot3D(0, 0, 0); }
    Box3D getBoundingBox() const { return Box3D(0, nx_ - 1, 0, ny_ - 1, 0, nz_ - 1); }

private:
    plint nx_, ny_, nz_;
    T* d_;
};

template <typename T, int nDim>
struct BoxProcessingFunctional3D_T {
    virtual ~BoxProcessingFunctional3D_T() {}
    virtual void process(Box3D domain, TensorField3D<T, nDim>& field) = 0;
    virtual BoxProcessingFunctional3D_T* clone() const = 0;
    virtual BlockDomain::DomainT appliesTo() const { return BlockDomain::bulk; }
    virtual void getModificationPattern(std::vector<bool>& isWritten) const { isWritten.assign(1, true); }
    virtual void getTypeOfModification(std::vector<modif::ModifT>& m) const { m.assign(1, modif::staticVariables); }
};

template <typename T>
struct BoxProcessingFunctional3D_S {
    virtual ~BoxProcessingFunctional3D_S() {}
    virtual void process(Box3D domain, ScalarField3D<T>& field) = 0;
    virtual BoxProcessingFunctional3D_S* clone() const = 0;
    virtual BlockDomain::DomainT appliesTo() const { return BlockDomain::bulk; }
    virtual void getModificationPattern(std::vector<bool>& isWritten) const { isWritten.assign(1, true); }
    virtual void getTypeOfModification(std::vector<modif::ModifT>& m) const { m.assign(1, modif::staticVariables); }
};

// A predicate over lattice cells (src/dataProcessors/dataInitializerFunctional3D.h).
struct DomainFunctional3D {
    virtual ~DomainFunctional3D() {}
    virtual bool operator()(plint iX, plint iY, plint iZ) const = 0;
    virtual DomainFunctional3D* clone() const = 0;
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
    Box3D getBoundingBox() const { return Box3D(0, nx_ - 1, 0, ny_ - 1, 0, nz_ - 1); }
    ScalarField3D<T> view() { return ScalarField3D<T>(nx_, ny_, nz_, v_.data()); }
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
    Box3D getBoundingBox() const { return Box3D(0, nx_ - 1, 0, ny_ - 1, 0, nz_ - 1); }
    TensorField3D<T, nDim> view()
    {
        return TensorField3D<T, nDim>(nx_, ny_, nz_, reinterpret_cast<Array<T, nDim>*>(v_.data()));
    }
    void fill(bl_lattice l) { blCheck(bl_lattice_get_moments(l, nullptr, v_.data()), "computeVelocity"); }
    void fill_force(bl_lattice l) { blCheck(bl_lattice_get_force(l, v_.data()), "ForceFieldFunctional"); }

private:
    plint nx_, ny_, nz_;
    std::vector<T> v_;
};

// What generateMulti*Field / generateMultiBlockLattice return: the
// std::auto_ptr the reference's progs hold (a std::unique_ptr takes it over
// as well).
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wdeprecated-declarations"
template <class F>
using OwnedPtr = std::auto_ptr<F>;
#pragma GCC diagnostic pop

// generateMultiScalarField / generateMultiTensorField
// (src/multiBlock/multiDataField3D.h): a host field over the box; the
// envelope width has no meaning on one device.
template <typename T>
OwnedPtr<MultiScalarField3D<T>> generateMultiScalarField(const Box3D& b, plint = 0)
{
    return OwnedPtr<MultiScalarField3D<T>>(new MultiScalarField3D<T>(b.getNx(), b.getNy(), b.getNz()));
}

template <typename T, int nDim>
OwnedPtr<MultiTensorField3D<T, nDim>> generateMultiTensorField(const Box3D& b, plint = 0)
{
    return OwnedPtr<MultiTensorField3D<T, nDim>>(new MultiTensorField3D<T, nDim>(b.getNx(), b.getNy(), b.getNz()));
}

// applyProcessingFunctional of a prog's own functional on a host field
template <typename T, int nDim>
void applyProcessingFunctional(BoxProcessingFunctional3D_T<T, nDim>* f, const Box3D& domain,
                               MultiTensorField3D<T, nDim>& field)
{
    std::unique_ptr<BoxProcessingFunctional3D_T<T, nDim>> own(f);
    TensorField3D<T, nDim> v = field.view();
    own->process(domain, v);
}

template <typename T>
void applyProcessingFunctional(BoxProcessingFunctional3D_S<T>* f, const Box3D& domain, MultiScalarField3D<T>& field)
{
    std::unique_ptr<BoxProcessingFunctional3D_S<T>> own(f);
    ScalarField3D<T> v = field.view();
    own->process(domain, v);
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
            if (pre_step_) pre_step_();
            blCheck(bl_coupled_steps(g, 1, 0), "collideAndStream (with the immersed boundary)");
        } else {
            blCheck(bl_collide_stream(handle(), 1), "collideAndStream");
        }
    }

    // the device lattice, created with the configuration so far on first use
    bl_lattice handle()
    {
        if (h_) return h_;
        int per[3] = {periodic_[0], periodic_[1], periodic_[2]};
        // velocity boundaries on the two faces of an axis bound that axis
        // (the progs' shear setup leaves y at Palabos's default,
        // non-periodic, and marks its faces as velocity boundaries)
        if (wall_axis_ >= 0) per[wall_axis_] = 0;
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
    int wallAxis() const { return wall_axis_; }
    unsigned char maskAt(plint x, plint y, plint z) const { return mask_[(x * ny_ + y) * nz_ + z]; }
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
    // the immersed boundary of this lattice (its phase field)
    void setBoundaryHandle(bl_ibm g) { ibm_ = g; }
    // runs before each scheduled immersed-boundary step (host decorators)
    void setPreStepHook(std::function<void()> f) { pre_step_ = std::move(f); }
    bl_ibm boundaryHandle() const { return ibm_; }
    void flushPendingBoundaryStep()
    {
        if (pending_ibm_) {
            bl_ibm g = pending_ibm_;
            pending_ibm_ = nullptr;
            if (pre_step_) pre_step_();
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
    bl_ibm ibm_ = nullptr;
    std::function<void()> pre_step_;
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

// setBoundaryVelocity with a velocity functor f(iX, iY, iZ, u)
// (src/boundaryCondition/boundaryInstantiator3D.h): evaluated on the
// velocity-boundary cells of the box; the device wall takes one velocity per
// face plane, so the functor must give one value on each.
template <typename T, template <typename U> class Descriptor, class VelocityFunction>
void setBoundaryVelocity(MultiBlockLattice3D<T, Descriptor>& lattice, const Box3D& box, VelocityFunction f)
{
    const int a = lattice.wallAxis();
    if (a < 0) throw std::runtime_error("setBoundaryVelocity: no velocity boundary was added");
    const plint n[3] = {lattice.getNx(), lattice.getNy(), lattice.getNz()};
    for (int side = 0; side < 2; ++side) {
        const plint plane = side ? n[a] - 1 : 0;
        const plint lo[3] = {box.x0, box.y0, box.z0}, hi[3] = {box.x1, box.y1, box.z1};
        if (plane < lo[a] || plane > hi[a]) continue;
        bool have = false;
        Array<T, 3> u0;
        for (plint x = std::max<plint>(box.x0, 0); x <= std::min(box.x1, n[0] - 1); ++x)
            for (plint y = std::max<plint>(box.y0, 0); y <= std::min(box.y1, n[1] - 1); ++y)
                for (plint z = std::max<plint>(box.z0, 0); z <= std::min(box.z1, n[2] - 1); ++z) {
                    const plint p[3] = {x, y, z};
                    if (p[a] != plane || lattice.maskAt(x, y, z) != 3) continue;
                    Array<T, 3> u;
                    f(x, y, z, u);
                    if (!have) {
                        u0 = u;
                        have = true;
                    } else if (u[0] != u0[0] || u[1] != u0[1] || u[2] != u0[2]) {
                        throw std::runtime_error("setBoundaryVelocity: the device wall takes one velocity per face plane");
                    }
                }
        if (have) lattice.setWall(a, side == 1, u0);
    }
}

// generateMultiBlockLattice(domain, dynamics, envelopeWidth)
// (src/multiBlock/multiBlockGenerator3D.h): a lattice over the box; the
// device lattice's cells are indexed from the box's lower corner, which
// must be the origin.
template <typename T, template <typename U> class Descriptor>
OwnedPtr<MultiBlockLattice3D<T, Descriptor>> generateMultiBlockLattice(const Box3D& domain,
                                                                        Dynamics<T, Descriptor>* background,
                                                                        plint = 0)
{
    if (domain.x0 != 0 || domain.y0 != 0 || domain.z0 != 0) {
        delete background;
        throw std::runtime_error("generateMultiBlockLattice: the device lattice needs a bounding box that starts at "
                                 "(0, 0, 0)");
    }
    return OwnedPtr<MultiBlockLattice3D<T, Descriptor>>(
        new MultiBlockLattice3D<T, Descriptor>(domain.getNx(), domain.getNy(), domain.getNz(), background));
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
    void writeVTK(const std::string& file) const;
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

template <class T>
class ParticleBase3D;

// Force decorators (src/fsi/ForceDecorator.h:12-27): applied to every
// particle before its membrane forces.  The device runs the wall
// interaction; a prog's own decorator (apply_force on the particle) runs on
// the host before each coupled step, its vertex forces handed to the device
// (bl_ibm_set_vertex_forces), after the device decorators.
template <class T>
struct ForceDecorator {
    virtual ~ForceDecorator() {}
    virtual void apply_force(ParticleBase3D<T>*) {}
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

// Periodicity of the immersed boundary's domain, a compile-time choice in
// the reference (src/fsi/Periodicity.h): the device takes it from the
// lattice, so this only answers get_x / get_y / get_z.
template <class T, bool px, bool py, bool pz>
struct Periodicity3D {
    static bool get_x() { return px; }
    static bool get_y() { return py; }
    static bool get_z() { return pz; }
};

namespace geo {
template <class T>
struct Rect {
    T x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
};
}  // namespace geo

// Small dense matrix, row-major (src/fsi/Matrix3.h)
template <class T, int N>
class Matrix {
public:
    Matrix() { reset_to_zero(); }
    T& operator()(int i, int j) { return m_[i * N + j]; }
    const T& operator()(int i, int j) const { return m_[i * N + j]; }
    T& operator[](int k) { return m_[k]; }
    const T& operator[](int k) const { return m_[k]; }
    void reset_to_zero()
    {
        for (int k = 0; k < N * N; ++k) m_[k] = T();
    }
    void resetToZero() { reset_to_zero(); }

private:
    T m_[N * N];
};

// Unit quaternion of a rotation (src/fsi/Quaternion.h): Quaternion(angle,
// axis) turns by angle about the unit axis; products compose rotations
// (q1 * q2 applies q2 first).
template <class T>
class Quaternion {
public:
    Quaternion() : q0(1), q1(0), q2(0), q3(0) {}
    Quaternion(T angle, const Array<T, 3>& axis)
        : q0(std::cos(angle / 2)), q1(axis[0] * std::sin(angle / 2)), q2(axis[1] * std::sin(angle / 2)),
          q3(axis[2] * std::sin(angle / 2))
    {
    }
    Quaternion(T a, T b, T c, T d) : q0(a), q1(b), q2(c), q3(d) {}
    void set_to_unity() { q0 = 1, q1 = q2 = q3 = 0; }
    Quaternion operator*(const Quaternion& o) const
    {
        return Quaternion(q0 * o.q0 - q1 * o.q1 - q2 * o.q2 - q3 * o.q3, q0 * o.q1 + q1 * o.q0 + q2 * o.q3 - q3 * o.q2,
                          q0 * o.q2 - q1 * o.q3 + q2 * o.q0 + q3 * o.q1, q0 * o.q3 + q1 * o.q2 - q2 * o.q1 + q3 * o.q0);
    }
    T norm() const { return std::sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3); }
    void to_rot_matrix(Matrix<T, 3>& R) const
    {
        R(0, 0) = 1 - 2 * (q2 * q2 + q3 * q3);
        R(0, 1) = 2 * (q1 * q2 - q0 * q3);
        R(0, 2) = 2 * (q1 * q3 + q0 * q2);
        R(1, 0) = 2 * (q1 * q2 + q0 * q3);
        R(1, 1) = 1 - 2 * (q1 * q1 + q3 * q3);
        R(1, 2) = 2 * (q2 * q3 - q0 * q1);
        R(2, 0) = 2 * (q1 * q3 - q0 * q2);
        R(2, 1) = 2 * (q2 * q3 + q0 * q1);
        R(2, 2) = 1 - 2 * (q1 * q1 + q2 * q2);
    }
    Array<T, 3> apply_rotation(const Array<T, 3>& v) const
    {
        Matrix<T, 3> R;
        to_rot_matrix(R);
        return Array<T, 3>(R(0, 0) * v[0] + R(0, 1) * v[1] + R(0, 2) * v[2], R(1, 0) * v[0] + R(1, 1) * v[1] + R(1, 2) * v[2],
                           R(2, 0) * v[0] + R(2, 1) * v[1] + R(2, 2) * v[2]);
    }
    Array<T, 3> apply_inv_rotation(const Array<T, 3>& v) const { return Quaternion(q0, -q1, -q2, -q3).apply_rotation(v); }
    T q0, q1, q2, q3;
};

template <class T>
std::ostream& operator<<(std::ostream& o, const Quaternion<T>& q)
{
    return o << q.q0 << " " << q.q1 << " " << q.q2 << " " << q.q3;
}

// Translations and rotations applied in the order added (src/fsi/Transform.h)
template <class T>
class Transform {
public:
    Transform& translate(const Array<T, 3>& t)
    {
        ops_.push_back(Op{false, t, Quaternion<T>()});
        return *this;
    }
    Transform& rotate(const Quaternion<T>& q)
    {
        ops_.push_back(Op{true, Array<T, 3>(), q});
        return *this;
    }
    Array<T, 3> apply(Array<T, 3> p) const
    {
        for (const Op& o : ops_) p = o.rot ? o.q.apply_rotation(p) : p + o.t;
        return p;
    }

private:
    struct Op {
        bool rot;
        Array<T, 3> t;
        Quaternion<T> q;
    };
    std::vector<Op> ops_;
};

// One vertex of a particle (src/fsi/ParticleBase.h:30-36)
template <class T>
struct Vertex {
    Array<T, 3> pos;
    Array<T, 3> vel;
    Array<T, 3> force;
};

// Area, signed volume and area-weighted centroid of a closed triangle mesh
// (the update() of src/fsi/DeformableParticle3D.hh:54-82), for the host side
// of particle set-up (placing and orienting before the device takes over).
template <class T>
void mesh_geometry(const std::vector<Array<T, 3>>& p, const std::vector<Array<plint, 3>>& tri, T& area, T& volume,
                   Array<T, 3>& com)
{
    area = volume = 0;
    com.resetToZero();
    for (const auto& t : tri) {
        const Array<T, 3>&a = p[t[0]], &b = p[t[1]], &c = p[t[2]];
        const Array<T, 3> n = crossProduct(b - a, c - a);
        const T dA = norm(n) / 2;
        area += dA;
        volume += dot(a, crossProduct(b, c)) / 6;
        com += (a + b + c) * (dA / 3);
    }
    com /= area;
}

template <class T>
class ParticleShapeLibrary;

template <class T>
class ParticleShape {
public:
    // a mesh edge (i0 < i1) with its rest length (src/fsi/ParticleShape.h)
    struct Link {
        plint i0, i1;
        T length;
    };
    typedef typename std::vector<Link>::const_iterator link_const_iterator;

    std::vector<Array<T, 3>> vertices;
    std::vector<Array<plint, 3>> triangles;
    std::string tag;
    plint id = 0;

    plint count_nodes() const { return (plint)vertices.size(); }
    plint count_vertices() const { return (plint)vertices.size(); }
    plint count_triangles() const { return (plint)triangles.size(); }
    const Array<T, 3>& get_vertex(plint i) const { return vertices[i]; }
    plint get_id() const { return id; }
    const std::string& get_tag() const { return tag; }
    // largest vertex distance from the origin (src/fsi/ParticleShape.hh:361-392)
    T get_radius() const
    {
        T r = 0;
        for (const auto& v : vertices) r = std::max(r, norm(v));
        return r;
    }
    T get_area() const
    {
        T a, v;
        Array<T, 3> c;
        mesh_geometry(vertices, triangles, a, v, c);
        return a;
    }
    T get_volume() const
    {
        T a, v;
        Array<T, 3> c;
        mesh_geometry(vertices, triangles, a, v, c);
        return v;
    }
    Array<T, 3> get_center() const
    {
        T a, v;
        Array<T, 3> c;
        mesh_geometry(vertices, triangles, a, v, c);
        return c;
    }
    // distinct edges in (i0, i1) order, i0 < i1 (the link set of
    // compute_connectivity, src/fsi/ParticleShape.hh:90-212)
    link_const_iterator links_begin() const
    {
        ensure_links();
        return links_.begin();
    }
    link_const_iterator links_end() const
    {
        ensure_links();
        return links_.end();
    }
    plint count_links() const
    {
        ensure_links();
        return (plint)links_.size();
    }

private:
    void ensure_links() const
    {
        if (have_links_) return;
        std::vector<std::pair<plint, plint>> e;
        for (const auto& t : triangles)
            for (int k = 0; k < 3; ++k) {
                const plint a = t[k], b = t[(k + 1) % 3];
                e.push_back(std::make_pair(std::min(a, b), std::max(a, b)));
            }
        std::sort(e.begin(), e.end());
        e.erase(std::unique(e.begin(), e.end()), e.end());
        links_.clear();
        for (const auto& x : e) links_.push_back(Link{x.first, x.second, norm(vertices[x.second] - vertices[x.first])});
        have_links_ = true;
    }
    mutable std::vector<Link> links_;
    mutable bool have_links_ = false;
};

template <class T>
class ParticleShapeLibrary {
public:
    ParticleShape<T>* store_mesh(const std::vector<Array<T, 3>>& v, const std::vector<Array<plint, 3>>& t,
                                 const std::string& tag)
    {
        std::unique_ptr<ParticleShape<T>> s(new ParticleShape<T>());
        s->vertices = v;
        s->triangles = t;
        s->tag = tag;
        if (!shapes_.count(tag)) order_.push_back(tag);
        s->id = (plint)(std::find(order_.begin(), order_.end(), tag) - order_.begin());
        shapes_[tag] = std::move(s);
        return shapes_[tag].get();
    }
    // gmsh ASCII 2.x: $Nodes and the 3-node triangles (type 2) of $Elements
    // (src/fsi/ParticleShapeFactory.hh:32-83)
    ParticleShape<T>* read_and_store_mesh(const std::string& file, const std::string& tag)
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
    ParticleShape<T>* get_by_tag(const std::string& tag) { return shapes_.at(tag).get(); }
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

// RBCParameters (src/fsi/RBCParticle.h:22-58): the model's parameters; K()
// is the area modulus of the in-plane links alone, 2 G
// (src/fsi/RBCParticle.hh:76-79), Young's modulus and Poisson ratio follow
// from G and K + the added area stiffness.
template <class T>
struct RBCParameters {
    T shear_modulus = 0, L0 = 0, k_area_global = 0, k_area_local = 0, k_volume = 0, vol_desired = 0, k_bend = 0,
      theta0 = 0;
    T G() const { return shear_modulus; }
    T K() const { return 2 * shear_modulus; }
    T youngs_modulus() const
    {
        const T Ka = K() + k_area_global + k_area_local;
        return 4 * Ka * G() / (Ka + G());
    }
    T poisson_ratio() const
    {
        const T Ka = K() + k_area_global + k_area_local;
        return (Ka - G()) / (Ka + G());
    }
    T cos_theta0() const { return std::cos(theta0); }
    T sin_theta0() const { return std::sin(theta0); }
};

// SemiRigidParticleParams (src/fsi/SemiRigidParticle.h:16-34)
template <class T>
struct SemiRigidParticleParams {
    T shear_modulus() const { return std::sqrt((T)3) * k_in_plane / (4 * l0); }
    T l0 = 0, k_in_plane = 0, k_out_of_plane = 0, k_bend = 0;
};

// One particle on the device outside any immersed boundary (a 4^3 lattice
// it never touches): the RBC calibration steps, membrane forces and
// volume observables of a particle being set up run here.
class ParticleDevice {
public:
    template <class T>
    ParticleDevice(const ParticleShape<T>* s, int model, const std::vector<double>& prm, int bug,
                   const std::vector<Vertex<T>>& v)
    {
        const int per[3] = {0, 0, 0};
        blCheck(bl_lattice_create(4, 4, 4, per, BL_STORE_F64, &L_), "particle device");
        blCheck(bl_ibm_create(L_, BL_KERNEL_PESKIN, &g_), "particle device");
        std::vector<double> rest, pos;
        std::vector<int> tri;
        for (const auto& x : s->vertices) rest.insert(rest.end(), {x[0], x[1], x[2]});
        for (const auto& x : s->triangles) tri.insert(tri.end(), {(int)x[0], (int)x[1], (int)x[2]});
        for (const auto& x : v) pos.insert(pos.end(), {x.pos[0], x.pos[1], x.pos[2]});
        int sid = -1, pid = -1;
        blCheck(bl_shape_add(g_, (int)s->vertices.size(), (int)s->triangles.size(), rest.data(), tri.data(), &sid),
                "particle device");
        blCheck(bl_particle_add(g_, sid, model, prm.data(), (int)prm.size(), pos.data(), bug, &pid), "particle device");
        blCheck(bl_ibm_init(g_), "particle device");
        nv_ = v.size();
    }
    ~ParticleDevice()
    {
        if (g_) bl_ibm_destroy(g_);
        if (L_) bl_lattice_destroy(L_);
    }
    ParticleDevice(const ParticleDevice&) = delete;
    ParticleDevice& operator=(const ParticleDevice&) = delete;
    template <class T>
    void upload(const std::vector<Vertex<T>>& v, const std::vector<double>& prm)
    {
        std::vector<double> b(3 * v.size());
        for (int what = 0; what < 3; ++what) {
            for (size_t k = 0; k < v.size(); ++k) {
                const Array<T, 3>& a = what == 0 ? v[k].pos : (what == 1 ? v[k].vel : v[k].force);
                for (int d = 0; d < 3; ++d) b[3 * k + d] = a[d];
            }
            blCheck(bl_particle_set_state(g_, 0, what, b.data()), "particle device");
        }
        blCheck(bl_particle_set_params(g_, 0, prm.data(), (int)std::min<size_t>(prm.size(), 8)), "particle device");
    }
    template <class T>
    void download(std::vector<Vertex<T>>& v)
    {
        std::vector<double> b(3 * v.size());
        for (int what = 0; what < 3; ++what) {
            blCheck(bl_particle_get(g_, 0, what, b.data()), "particle device");
            for (size_t k = 0; k < v.size(); ++k) {
                Array<T, 3>& a = what == 0 ? v[k].pos : (what == 1 ? v[k].vel : v[k].force);
                for (int d = 0; d < 3; ++d) a[d] = b[3 * k + d];
            }
        }
    }
    bl_ibm handle() const { return g_; }

private:
    bl_lattice L_ = nullptr;
    bl_ibm g_ = nullptr;
    size_t nv_ = 0;
};

template <class T>
class ParticleBase3D {
public:
    typedef typename std::vector<Vertex<T>>::iterator vertex_iterator;
    typedef typename std::vector<Vertex<T>>::const_iterator vertex_const_iterator;

    explicit ParticleBase3D(const ParticleShape<T>* shape) : shape_(shape)
    {
        vertices_.resize(shape->vertices.size());
        for (size_t k = 0; k < vertices_.size(); ++k) vertices_[k].pos = shape->vertices[k];
        update();
    }
    virtual ~ParticleBase3D() {}
    virtual ParticleBase3D* clone() const = 0;
    virtual plint get_type_id() const = 0;

    const ParticleShape<T>* shape() const { return shape_; }
    const ParticleShape<T>* get_shape() const { return shape_; }
    plint get_shape_id() const { return shape_->get_id(); }
    plint count_nodes() const { return (plint)vertices_.size(); }
    Vertex<T>& get_node(plint i) { return vertices_[i]; }
    const Vertex<T>& get_node(plint i) const { return vertices_[i]; }
    vertex_iterator begin() { return vertices_.begin(); }
    vertex_iterator end() { return vertices_.end(); }
    vertex_const_iterator begin() const { return vertices_.begin(); }
    vertex_const_iterator end() const { return vertices_.end(); }
    const Array<T, 3>& get_vertex(plint i) const { return vertices_[i].pos; }
    std::vector<Array<T, 3>> positions() const
    {
        std::vector<Array<T, 3>> p;
        for (const auto& v : vertices_) p.push_back(v.pos);
        return p;
    }

    plint get_id() const { return id_; }
    void set_id(plint id) { id_ = id; }
    const Array<T, 3>& center_of_mass() const { return com_; }
    const T& volume() const { return volume_; }
    const T& area() const { return area_; }
    geo::Rect<T> bounding_box() const { return bbox_; }

    // geometry of the current vertices (src/fsi/ParticleBase.hh:51-69,
    // src/fsi/DeformableParticle3D.hh:54-82)
    virtual void update()
    {
        std::vector<Array<T, 3>> p = positions();
        mesh_geometry(p, shape_->triangles, area_, volume_, com_);
        bbox_.x0 = bbox_.y0 = bbox_.z0 = std::numeric_limits<T>::max();
        bbox_.x1 = bbox_.y1 = bbox_.z1 = -std::numeric_limits<T>::max();
        for (const auto& q : p) {
            bbox_.x0 = std::min(bbox_.x0, q[0]), bbox_.x1 = std::max(bbox_.x1, q[0]);
            bbox_.y0 = std::min(bbox_.y0, q[1]), bbox_.y1 = std::max(bbox_.y1, q[1]);
            bbox_.z0 = std::min(bbox_.z0, q[2]), bbox_.z1 = std::max(bbox_.z1, q[2]);
        }
    }
    // src/fsi/ParticleBase.hh:121-140
    virtual void set_center_of_mass(const Array<T, 3>& target)
    {
        const Array<T, 3> shift = target - com_;
        for (auto& v : vertices_) v.pos += shift;
        update();
    }
    virtual void reset_forces()
    {
        for (auto& v : vertices_) v.force.resetToZero();
    }
    // the current vertices as a new shape of the library (src/fsi/ParticleBase.h:113)
    void store_shape(ParticleShapeLibrary<T>& lib, const std::string& tag) const
    {
        lib.store_mesh(positions(), shape_->triangles, tag);
    }

    // the device form of the model: BL_MODEL_* and its parameters
    virtual int model() const = 0;
    virtual std::vector<double> device_params() const = 0;
    virtual int bug_compatible() const { return 1; }
    // set by an immersed boundary that holds this object as its copy of a
    // particle: observables then come from that boundary's device state
    void attach(bl_ibm g, int index)
    {
        held_ = g;
        held_index_ = index;
    }
    bl_ibm held_handle() const { return held_; }
    int held_index() const { return held_index_; }

protected:
    // the particle's own device context (set-up steps outside any boundary)
    ParticleDevice& device()
    {
        if (!dev_) dev_ = std::make_shared<ParticleDevice>(shape_, model(), device_params(), bug_compatible(), vertices_);
        return *dev_;
    }
    // particle index and handle to read observables from
    bl_ibm observable_source(int& index)
    {
        if (held_) {
            index = held_index_;
            return held_;
        }
        device().upload(vertices_, device_params());
        index = 0;
        return dev_->handle();
    }
    std::vector<Vertex<T>> vertices_;
    T area_ = 0, volume_ = 0;
    Array<T, 3> com_;
    geo::Rect<T> bbox_;

private:
    const ParticleShape<T>* shape_;
    plint id_ = 0;
    bl_ibm held_ = nullptr;
    int held_index_ = -1;
    std::shared_ptr<ParticleDevice> dev_;
};

// eigenvalues / eigenvectors (columns of Q) of a symmetric 3x3 matrix by
// cyclic Jacobi rotations (the role of linalg::diagonalize, src/fsi/linalg.h:22-140)
template <class T>
void diagonalize(Matrix<T, 3> A, Matrix<T, 3>& Q, Array<T, 3>& eig)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Q(i, j) = i == j ? 1 : 0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        const T off = A(0, 1) * A(0, 1) + A(0, 2) * A(0, 2) + A(1, 2) * A(1, 2);
        if (off < T(1e-30)) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A(p, q) == T(0)) continue;
                const T th = (A(q, q) - A(p, p)) / (2 * A(p, q));
                const T t = (th >= 0 ? 1 : -1) / (std::abs(th) + std::sqrt(th * th + 1));
                const T cs = 1 / std::sqrt(t * t + 1), sn = t * cs;
                for (int k = 0; k < 3; ++k) {
                    const T akp = A(k, p), akq = A(k, q);
                    A(k, p) = cs * akp - sn * akq;
                    A(k, q) = sn * akp + cs * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const T apk = A(p, k), aqk = A(q, k);
                    A(p, k) = cs * apk - sn * aqk;
                    A(q, k) = sn * apk + cs * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const T vkp = Q(k, p), vkq = Q(k, q);
                    Q(k, p) = cs * vkp - sn * vkq;
                    Q(k, q) = sn * vkp + cs * vkq;
                }
            }
    }
    eig = Array<T, 3>(A(0, 0), A(1, 1), A(2, 2));
}

// DeformableParticle3D (src/fsi/DeformableParticle3D.h:20-50): orientation
// helpers and the inertia of the enclosed volume.
template <class T>
class DeformableParticle3D : public ParticleBase3D<T> {
public:
    explicit DeformableParticle3D(const ParticleShape<T>* s) : ParticleBase3D<T>(s) {}
    void transform_vertices(const Transform<T>& tr)
    {
        for (auto& v : this->vertices_) v.pos = tr.apply(v.pos);
        this->update();
    }
    // tensor of gyration of the vertices about the centre of mass
    // (src/fsi/DeformableParticle3D.hh:111-125)
    void compute_tensor_of_gyration(Matrix<T, 3>& A) const
    {
        A.reset_to_zero();
        const Array<T, 3>& c = this->center_of_mass();
        for (const auto& v : this->vertices_)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) A(i, j) += (v.pos[i] - c[i]) * (v.pos[j] - c[j]);
        for (int k = 0; k < 9; ++k) A[k] /= (T)this->count_nodes();
    }
    void compute_orientation(Matrix<T, 3>& Q, Array<T, 3>& eig) const
    {
        Matrix<T, 3> A;
        compute_tensor_of_gyration(A);
        diagonalize(A, Q, eig);
    }
    // turns the axis of the smallest / largest gyration eigenvalue onto pref
    void set_minor_axis_orientation(const Array<T, 3>& pref) { align_axis(pref, true); }
    void set_major_axis_orientation(const Array<T, 3>& pref) { align_axis(pref, false); }
    // inertia tensor of the enclosed volume (unit density), from the device
    // (src/fsi/DeformableParticle3D.hh:173-272)
    void compute_moment_of_intertia(Matrix<T, 3>& I)
    {
        int index = 0;
        const bl_ibm g = this->observable_source(index);
        double o[14];
        blCheck(bl_particle_observables(g, index, o), "compute_moment_of_intertia");
        for (int k = 0; k < 9; ++k) I[k] = o[5 + k];
    }

private:
    void align_axis(const Array<T, 3>& pref, bool minor)
    {
        Matrix<T, 3> Q;
        Array<T, 3> e;
        compute_orientation(Q, e);
        const Array<T, 3> a(std::abs(e[0]), std::abs(e[1]), std::abs(e[2]));
        const int k = minor ? ((a[0] < a[1] && a[0] < a[2]) ? 0 : (a[1] < a[2] ? 1 : 2))
                            : ((a[0] > a[1] && a[0] > a[2]) ? 0 : (a[1] > a[2] ? 1 : 2));
        Array<T, 3> dir(Q(0, k), Q(1, k), Q(2, k));
        dir /= norm(dir);
        Array<T, 3> axis = crossProduct(dir, pref);
        const T na = norm(axis);
        if (na == T(0)) return;
        const T angle = std::asin(std::min(T(1), na / (norm(dir) * norm(pref))));
        axis /= na;
        if (dot(pref, dir) < 0) axis = -axis;
        const Array<T, 3> c = this->center_of_mass();
        Transform<T> tr;
        tr.translate(-c).rotate(Quaternion<T>(angle, axis)).translate(c);
        transform_vertices(tr);
    }
};

// DeformableCapsuleParticle (src/fsi/DeformableCapsule.h:33-64): NeoHookean
// unless the Skalak law is selected (the reference picks it through its
// serialized state; here set_shear_model).
template <class T>
class DeformableCapsuleParticle : public DeformableParticle3D<T> {
public:
    enum ShearModel { Skalak, NeoHookean };
    DeformableCapsuleParticle(const ParticleShape<T>* s, const CapsuleParameters<T>& p = CapsuleParameters<T>())
        : DeformableParticle3D<T>(s), p_(p)
    {
    }
    DeformableCapsuleParticle* clone() const override { return new DeformableCapsuleParticle(*this); }
    plint get_type_id() const override { return 0; }
    void set_shear_model(ShearModel m) { model_ = m; }
    int model() const override { return model_ == Skalak ? BL_MODEL_CAPSULE_SKALAK : BL_MODEL_CAPSULE_NEOHOOKEAN; }
    std::vector<double> device_params() const override { return {p_.G, p_.C}; }
    CapsuleParameters<T>& params() { return p_; }
    const CapsuleParameters<T>& params() const { return p_; }
    const CapsuleParameters<T>& get_params() const { return p_; }

private:
    CapsuleParameters<T> p_;
    ShearModel model_ = NeoHookean;
};

// RBCParticle (src/fsi/RBCParticle.h:60-90): compute_forces and
// relax_nodes run on the particle's device context.
template <class T>
class RBCParticle : public DeformableParticle3D<T> {
public:
    typedef typename DeformableParticle3D<T>::vertex_const_iterator vertex_const_iterator;
    typedef typename DeformableParticle3D<T>::vertex_iterator vertex_iterator;
    explicit RBCParticle(const ParticleShape<T>* s) : DeformableParticle3D<T>(s) {}
    RBCParticle(const ParticleShape<T>* s, const RBCParameters<T>& p) : DeformableParticle3D<T>(s), p_(p) {}
    RBCParticle* clone() const override { return new RBCParticle(*this); }
    plint get_type_id() const override { return 1; }
    int model() const override { return BL_MODEL_RBC; }
    std::vector<double> device_params() const override
    {
        return {p_.shear_modulus, p_.L0, p_.k_area_global, p_.k_area_local, p_.k_volume, p_.vol_desired, p_.k_bend,
                p_.theta0};
    }
    RBCParameters<T>& params() { return p_; }
    const RBCParameters<T>& params() const { return p_; }
    RBCParameters<T>& params_ref() { return p_; }
    const RBCParameters<T>& params_ref() const { return p_; }
    // membrane forces of the current vertices (src/fsi/RBCParticle.hh:140-247)
    void compute_forces()
    {
        ParticleDevice& d = this->device();
        d.upload(this->vertices_, device_params());
        blCheck(bl_particle_membrane_forces(d.handle()), "RBCParticle::compute_forces");
        d.download(this->vertices_);
    }
    // the reference prints the membrane energy terms; the device computes
    // forces, not energies, so this prints the constrained quantities
    template <class Stream>
    void print_energies(Stream& out) const
    {
        out << "area " << this->area() << " (rest " << this->shape()->get_area() << "), volume " << this->volume()
            << " (desired " << p_.vol_desired << ")" << std::endl;
    }
    // one damped relaxation step (src/fsi/RBCParticle.hh:318-342)
    void relax_nodes(T damping)
    {
        ParticleDevice& d = this->device();
        d.upload(this->vertices_, device_params());
        blCheck(bl_rbc_relax_nodes(d.handle(), 0, damping), "RBCParticle::relax_nodes");
        d.download(this->vertices_);
        this->update();
    }

private:
    RBCParameters<T> p_;
};

// Particles without a device model here: declared so that progs using them
// compile; adding one to an immersed boundary is refused.
template <class T>
class SemiRigidParticle3D : public DeformableParticle3D<T> {
public:
    explicit SemiRigidParticle3D(const ParticleShape<T>* s) : DeformableParticle3D<T>(s) {}
    SemiRigidParticle3D(const ParticleShape<T>* s, const SemiRigidParticleParams<T>& p) : DeformableParticle3D<T>(s), p_(p) {}
    SemiRigidParticle3D* clone() const override { return new SemiRigidParticle3D(*this); }
    plint get_type_id() const override { return 3; }
    int model() const override { return -1; }
    std::vector<double> device_params() const override { return {}; }
    SemiRigidParticleParams<T>& params() { return p_; }
    const SemiRigidParticleParams<T>& params() const { return p_; }

private:
    SemiRigidParticleParams<T> p_;
};

// RigidParticle3D (src/fsi/RigidParticle.h:20-91): a scaled, rotated copy
// of its shape about the centre of mass.
template <class T>
class RigidParticle3D : public ParticleBase3D<T> {
public:
    explicit RigidParticle3D(const ParticleShape<T>* s) : ParticleBase3D<T>(s)
    {
        this->com_.resetToZero();
        update();
    }
    RigidParticle3D* clone() const override { return new RigidParticle3D(*this); }
    plint get_type_id() const override { return 2; }
    int model() const override { return -2; }
    std::vector<double> device_params() const override { return {}; }
    T& scale() { return scale_; }
    T& density() { return density_; }
    T& damping() { return damping_; }
    Quaternion<T>& orientation() { return orientation_; }
    const Quaternion<T>& orientation() const { return orientation_; }
    Array<T, 3>& velocity() { return velocity_; }
    Array<T, 3>& angular_velocity() { return ang_velocity_; }
    void set_center_of_mass(const Array<T, 3>& c) override
    {
        this->com_ = c;
        update();
    }
    // vertices = com + R (scale rest vertex): the shape's own origin is the
    // body's reference point (src/fsi/RigidParticle.hh:32-40)
    void update() override
    {
        if (!this->shape()) return;
        const Array<T, 3> com = this->com_;
        for (size_t k = 0; k < this->vertices_.size(); ++k)
            this->vertices_[k].pos = com + orientation_.apply_rotation(this->shape()->vertices[k] * scale_);
        ParticleBase3D<T>::update();
        this->com_ = com;
    }

private:
    Array<T, 3> velocity_, ang_velocity_;
    Quaternion<T> orientation_;
    T scale_ = 1, density_ = 1, damping_ = 0;
};

template <typename T, template <typename U> class Descriptor, class Periodicity = void>
class ImmersedBoundaryDynamics3D {
public:
    typedef std::map<plint, ParticleBase3D<T>*> ObjMapType;
    typedef typename ObjMapType::iterator ObjMapIterator;
    typedef typename ObjMapType::const_iterator ObjMapConstIterator;

    ImmersedBoundaryDynamics3D(MultiBlockLattice3D<T, Descriptor>& lattice, const ParticleShapeLibrary<T>& lib,
                               int kernel = BL_KERNEL_ROMA)
        : lattice_(lattice), lib_(lib), kernel_(kernel)
    {
    }
    ~ImmersedBoundaryDynamics3D()
    {
        if (g_) {
            lattice_.setBoundaryHandle(nullptr);
            bl_ibm_destroy(g_);
        }
    }
    ImmersedBoundaryDynamics3D(const ImmersedBoundaryDynamics3D&) = delete;
    ImmersedBoundaryDynamics3D& operator=(const ImmersedBoundaryDynamics3D&) = delete;

    // add_particle (src/fsi/ImmersedBoundaryDynamics.hh:508-525): a copy of
    // the particle joins with the next sequential id
    plint add_particle(const ParticleBase3D<T>* p)
    {
        if (g_) throw std::runtime_error("add_particle: particles must be added before init()");
        if (p->model() < 0)
            throw std::runtime_error("add_particle: this particle type has no device model (capsules and red cells do)");
        owned_.emplace_back(p->clone());
        ParticleBase3D<T>* c = owned_.back().get();
        c->set_id(next_id_);
        objs_[next_id_] = c;
        return next_id_++;
    }
    ObjMapIterator particles_begin() { return objs_.begin(); }
    ObjMapIterator particles_end() { return objs_.end(); }
    plint count_particles() const { return (plint)nv_.size(); }

    // load_checkpoint (src/fsi/ImmersedBoundaryDynamics_io.hh:75-158): the
    // particles of <name>.plb (and its rank file) replace the added ones at
    // init(); shape ids refer to the library's storage order
    void load_checkpoint(const std::string& name)
    {
        if (g_) throw std::runtime_error("load_checkpoint: call before init()");
        checkpoint_ = name;
    }
    // save_checkpoint (src/fsi/ImmersedBoundaryDynamics_io.hh:17-73):
    // <output dir><name>.plb and <output dir><name>/0.plb
    void save_checkpoint(const std::string& name) const
    {
        if (!g_) throw std::runtime_error("save_checkpoint: call init() first");
        lattice_.flushPendingBoundaryStep();
        const std::string dir = out_dir();
        makeDir(dir + name);
        blCheck(bl_ibm_save_checkpoint(g_, dir.empty() ? "." : dir.c_str(), name.c_str()), "save_checkpoint");
    }
    void init()
    {
        blCheck(bl_ibm_create(lattice_.handle(), kernel_, &g_), "ImmersedBoundaryDynamics3D");
        lattice_.setBoundaryHandle(g_);
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
            shape_of_.push_back(s);
        };
        for (plint k = 0; k < lib_.count(); ++k) add_shape(lib_.get_by_id(k));
        if (!checkpoint_.empty()) {
            const std::string in = global::directories().getInputDir();
            blCheck(bl_ibm_load_checkpoint(g_, in.empty() ? "." : in.c_str(), checkpoint_.c_str()), "load_checkpoint");
            int n = 0;
            blCheck(bl_ibm_count(g_, &n, nullptr), "load_checkpoint");
            objs_.clear();
            owned_.clear();
            for (int p = 0; p < n; ++p) {
                int s = -1, model = -1;
                long long id = -1;
                blCheck(bl_particle_describe(g_, p, &s, &model, &id), "load_checkpoint");
                nv_.push_back(shape_nv_.at((size_t)s));
                // a host copy to read the particle through (shape and model;
                // the parameters stay on the device)
                const ParticleShape<T>* sh = shape_of_.at((size_t)s);
                if (model == BL_MODEL_RBC) owned_.emplace_back(new RBCParticle<T>(sh));
                else owned_.emplace_back(new DeformableCapsuleParticle<T>(sh));
                ParticleBase3D<T>* c = owned_.back().get();
                c->set_id((plint)id);
                c->attach(g_, p);
                objs_[(plint)id] = c;
                index_[(plint)id] = p;
            }
        } else {
            int index = 0;
            for (auto& kv : objs_) {
                ParticleBase3D<T>* p = kv.second;
                const ParticleShape<T>* s = p->shape();
                add_shape(s);
                nv_.push_back(p->count_nodes());
                std::vector<double> pos;
                for (auto it = p->begin(); it != p->end(); ++it) pos.insert(pos.end(), {it->pos[0], it->pos[1], it->pos[2]});
                const std::vector<double> prm = p->device_params();
                int pid = -1;
                blCheck(bl_particle_add(g_, sid[s], p->model(), prm.data(), (int)prm.size(), pos.data(),
                                        p->bug_compatible(), &pid),
                        "add_particle");
                p->attach(g_, index);
                index_[kv.first] = index++;
            }
        }
        blCheck(bl_ibm_init(g_), "init");
        if (boundary_) set_boundary(boundary_);
        for (ForceDecorator<T>* d : decorators_) push_decorator(d);
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
    // synchronize_particle_states (src/fsi/ImmersedBoundaryDynamics.hh:
    // 416-450): the host copies' vertices (changed through the particles
    // after init) go to the device
    void synchronize_particle_states()
    {
        if (!g_) return;
        lattice_.flushPendingBoundaryStep();
        std::vector<double> pos;
        for (auto& kv : objs_)
            for (auto it = kv.second->begin(); it != kv.second->end(); ++it)
                pos.insert(pos.end(), {it->pos[0], it->pos[1], it->pos[2]});
        blCheck(bl_particle_set_positions(g_, -1, pos.data()), "synchronize_particle_states");
    }
    // move_vertices (src/fsi/ImmersedBoundaryDynamics.hh:461-493) outside
    // the coupled step: only the coupled step moves deformable particles on
    // the device
    void move_vertices()
    {
        throw std::runtime_error("move_vertices: particles move within the coupled step on the device "
                                 "(applyProcessingFunctional + collideAndStream)");
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

    // get_particle (src/fsi/ImmersedBoundaryDynamics.h:95): the particle of
    // that id, its host copy brought up to date with the device
    bool get_particle(plint id, ParticleBase3D<T>*& p)
    {
        auto it = objs_.find(id);
        if (it == objs_.end() || !g_) return false;
        lattice_.flushPendingBoundaryStep();
        const int index = index_.at(id);
        const plint nv = it->second->count_nodes();
        std::vector<double> buf(3 * nv);
        for (int what = 0; what < 3; ++what) {
            blCheck(bl_particle_get(g_, index, what, buf.data()), "get_particle");
            for (plint k = 0; k < nv; ++k) {
                Vertex<T>& v = it->second->get_node(k);
                Array<T, 3>& a = what == 0 ? v.pos : (what == 1 ? v.vel : v.force);
                a = Array<T, 3>(buf[3 * k], buf[3 * k + 1], buf[3 * k + 2]);
            }
        }
        it->second->update();
        p = it->second;
        return true;
    }
    bool get_particle(plint id, const ParticleBase3D<T>*& p)
    {
        ParticleBase3D<T>* q = nullptr;
        if (!get_particle(id, q)) return false;
        p = q;
        return true;
    }
    // current vertices of particle index i into a host copy
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
    // write_particles_as_vtk (src/fsi/ImmersedBoundaryDynamics_io.hh:160-176):
    // obj<id>-<it>.vtu per particle in the output directory (or `dir`)
    void write_particles_as_vtk(plint it, const std::string& dir)
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
    void write_particles_as_vtk(plint it) { write_particles_as_vtk(it, out_dir()); }
    // write_lightweight_particle_data (src/fsi/ImmersedBoundaryDynamics_io.hh:
    // 178-200): objsP0-<it>.txt in the output directory (or `dir`)
    void write_lightweight_particle_data(plint it, const std::string& dir)
    {
        lattice_.flushPendingBoundaryStep();
        std::ostringstream name;
        name << dir << "objsP0-" << it << ".txt";
        blCheck(bl_ibm_write_lightweight(g_, name.str().c_str(), (long long)it), "write_lightweight_particle_data");
    }
    void write_lightweight_particle_data(plint it) { write_lightweight_particle_data(it, out_dir()); }
    void schedule_step() { lattice_.setPendingBoundaryStep(g_); }

private:
    static std::string out_dir()
    {
        std::string d = global::directories().getOutputDir();
        if (d == "./") return "";
        if (!d.empty() && d.back() != '/') d.push_back('/');
        return d;
    }
    MultiBlockLattice3D<T, Descriptor>& lattice_;
    const ParticleShapeLibrary<T>& lib_;
    int kernel_;
    std::string checkpoint_;
    std::vector<plint> shape_nv_, nv_;  // vertices per shape id / per particle
    std::vector<const ParticleShape<T>*> shape_of_;
    void push_decorator(ForceDecorator<T>* d)
    {
        bl_boundary b;
        bl_potential u;
        if (!d->device_form(b, u)) {
            host_decorators_.push_back(d);
            lattice_.setPreStepHook([this]() { host_decorator_forces(); });
            return;
        }
        lattice_.flushPendingBoundaryStep();
        blCheck(bl_ibm_add_wall_interaction(g_, &b, &u), "add_force_decorator");
    }
    // the host decorators' vertex forces of the current state, in particle
    // order, for the next step (each decorator adds to forces that start at
    // zero, in the order the decorators were added)
    void host_decorator_forces()
    {
        int nv = 0;
        blCheck(bl_ibm_count(g_, nullptr, &nv), "force decorator");
        std::vector<double> pos(3 * (size_t)nv), out;
        blCheck(bl_particle_get(g_, -1, 0, pos.data()), "force decorator");
        out.reserve(pos.size());
        size_t k = 0;
        for (auto& kv : objs_) {
            ParticleBase3D<T>* p = kv.second;
            for (auto it = p->begin(); it != p->end(); ++it, ++k) {
                it->pos = Array<T, 3>(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]);
                it->force.resetToZero();
            }
            p->update();
            for (ForceDecorator<T>* d : host_decorators_) d->apply_force(p);
            for (auto it = p->begin(); it != p->end(); ++it) out.insert(out.end(), {it->force[0], it->force[1], it->force[2]});
        }
        blCheck(bl_ibm_set_vertex_forces(g_, out.data()), "force decorator");
    }
    void push_pp_force(const PPInteractionForce<T>* f)
    {
        const bl_potential u = f->desc();
        lattice_.flushPendingBoundaryStep();
        blCheck(bl_ibm_add_pp_force(g_, &u), "add_pp_force");
    }

    std::vector<std::unique_ptr<ParticleBase3D<T>>> owned_;
    ObjMapType objs_;
    std::map<plint, int> index_;  // particle id -> device index
    plint next_id_ = 0;
    std::vector<ForceDecorator<T>*> decorators_, host_decorators_;
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
inline std::ostream& operator<<(std::ostream& o, const Time& t) { return o << t.to_plint(); }

// Profile (src/fsi/Profile.h:20-90): named wall-clock timers; write_report
// puts "<name> <seconds>" lines in <output dir><prefix><rank>.txt.
class Profile {
public:
    static void start_timer(const std::string& name) { table()[name].start = std::chrono::steady_clock::now(); }
    static void stop_timer(const std::string& name)
    {
        Entry& e = table()[name];
        e.total += std::chrono::duration<double>(std::chrono::steady_clock::now() - e.start).count();
    }
    static double get_time(const std::string& name) { return table()[name].total; }
    static void write_report(const std::string& prefix)
    {
        std::ofstream out((global::directories().getOutputDir() + prefix + "0.txt").c_str());
        for (const auto& kv : table()) out << kv.first << " " << kv.second.total << "\n";
    }

private:
    struct Entry {
        std::chrono::steady_clock::time_point start;
        double total = 0;
    };
    static std::map<std::string, Entry>& table()
    {
        static std::map<std::string, Entry> t;
        return t;
    }
};

// Serialization helpers (src/fsi/utils.h) and the communication buffer of
// the reference's particle exchange (src/fsi/CommunicationBuffer.h): with
// one process nothing is sent and nothing is received.
namespace utils {
template <class V>
void pack(std::vector<char>& buf, const V& v)
{
    const char* p = reinterpret_cast<const char*>(&v);
    buf.insert(buf.end(), p, p + sizeof(V));
}
template <class V>
void unpack(char*& it, V& v)
{
    std::memcpy(reinterpret_cast<char*>(&v), it, sizeof(V));
    it += sizeof(V);
}
}  // namespace utils

class CommunicationBuffer {
public:
    void set_proc_list(const std::vector<plint>& procs)
    {
        if (!procs.empty()) throw std::runtime_error("CommunicationBuffer: one process, no peers");
    }
    template <class V>
    void pack(plint, const V&)
    {
        throw std::runtime_error("CommunicationBuffer: one process, no peers");
    }
    void send_and_receive_no_wait(bool = false) {}
    void finalize_send_and_receive() {}
    char* recv_buffer_begin() { return nullptr; }
    char* recv_buffer_end() { return nullptr; }
};

}  // namespace fsi

// VtkImageOutput3D (src/io/vtkDataOutput.h): the fields written with
// writeData go, on destruction, to <output dir><name>.vti, an XML ImageData
// file with raw appended Float32 arrays (point data; spacing dx).
template <typename T>
class VtkImageOutput3D {
public:
    explicit VtkImageOutput3D(const std::string& name, T dx = 1) : name_(name), dx_(dx) {}
    ~VtkImageOutput3D()
    {
        try {
            flush();
        } catch (...) {
        }
    }
    template <int nDim, typename F>
    void writeData(MultiTensorField3D<T, nDim>& f, const std::string& name, T scale = 1)
    {
        add(f.getBoundingBox(), name, nDim, f.data(), scale);
    }
    template <typename F>
    void writeData(MultiScalarField3D<T>& f, const std::string& name, T scale = 1)
    {
        add(f.getBoundingBox(), name, 1, f.data(), scale);
    }

private:
    struct Arr {
        std::string name;
        int ncomp;
        std::vector<float> v;
    };
    void add(const Box3D& b, const std::string& name, int ncomp, const std::vector<T>& v, T scale)
    {
        if (arrs_.empty()) box_ = b;
        else if (b.getNx() != box_.getNx() || b.getNy() != box_.getNy() || b.getNz() != box_.getNz())
            throw std::runtime_error("VtkImageOutput3D: fields of one file must have one size");
        // VTK image data runs x fastest; the fields here run z fastest
        const plint nx = b.getNx(), ny = b.getNy(), nz = b.getNz();
        Arr a{name, ncomp, std::vector<float>(v.size())};
        for (plint z = 0; z < nz; ++z)
            for (plint y = 0; y < ny; ++y)
                for (plint x = 0; x < nx; ++x)
                    for (int c = 0; c < ncomp; ++c)
                        a.v[((z * ny + y) * nx + x) * ncomp + c] = (float)(v[((x * ny + y) * nz + z) * ncomp + c] * scale);
        arrs_.push_back(std::move(a));
    }
    void flush()
    {
        if (arrs_.empty()) return;
        std::ofstream out((global::directories().getOutputDir() + name_ + ".vti").c_str(), std::ios::binary);
        if (!out) throw std::runtime_error("VtkImageOutput3D: cannot write " + name_);
        out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\" "
               "header_type=\"UInt64\">\n";
        out << "<ImageData WholeExtent=\"" << box_.x0 << " " << box_.x1 << " " << box_.y0 << " " << box_.y1 << " "
            << box_.z0 << " " << box_.z1 << "\" Origin=\"0 0 0\" Spacing=\"" << dx_ << " " << dx_ << " " << dx_ << "\">\n";
        out << "<Piece Extent=\"" << box_.x0 << " " << box_.x1 << " " << box_.y0 << " " << box_.y1 << " " << box_.z0
            << " " << box_.z1 << "\">\n<PointData>\n";
        unsigned long long off = 0;
        for (const Arr& a : arrs_) {
            out << "<DataArray type=\"Float32\" Name=\"" << a.name << "\" NumberOfComponents=\"" << a.ncomp
                << "\" format=\"appended\" offset=\"" << off << "\"/>\n";
            off += 8 + a.v.size() * sizeof(float);
        }
        out << "</PointData>\n</Piece>\n</ImageData>\n<AppendedData encoding=\"raw\">\n_";
        for (const Arr& a : arrs_) {
            const unsigned long long n = a.v.size() * sizeof(float);
            out.write(reinterpret_cast<const char*>(&n), sizeof(n));
            out.write(reinterpret_cast<const char*>(a.v.data()), (std::streamsize)n);
        }
        out << "\n</AppendedData>\n</VTKFile>\n";
        arrs_.clear();
    }
    std::string name_;
    T dx_;
    Box3D box_;
    std::vector<Arr> arrs_;
};

// XMLreader (src/io/xmlIO.h): the elements of an XML file as a tree;
// reader["a"]["b"].read(value) parses the text of <a><b>...</b></a> and
// throws PlbIOException when the element is missing or does not parse.
class XMLreaderProxy;
class XMLreader {
public:
    struct Node {
        std::string name, text;
        std::vector<Node> children;
    };
    explicit XMLreader(const std::string& file)
    {
        std::ifstream in(file.c_str());
        if (!in) throw PlbIOException("XMLreader: cannot open " + file);
        std::stringstream ss;
        ss << in.rdbuf();
        const std::string s = ss.str();
        size_t pos = 0;
        root_.name = "";
        parse_children(s, pos, root_, "");
    }
    XMLreaderProxy operator[](const std::string& name) const;
    const Node& root() const { return root_; }

private:
    static void skip_misc(const std::string& s, size_t& pos)
    {
        for (;;) {
            while (pos < s.size() && std::isspace((unsigned char)s[pos])) ++pos;
            if (s.compare(pos, 4, "<!--") == 0) {
                const size_t e = s.find("-->", pos);
                pos = e == std::string::npos ? s.size() : e + 3;
            } else if (s.compare(pos, 2, "<?") == 0) {
                const size_t e = s.find("?>", pos);
                pos = e == std::string::npos ? s.size() : e + 2;
            } else {
                return;
            }
        }
    }
    static void parse_children(const std::string& s, size_t& pos, Node& parent, const std::string& closing)
    {
        std::string text;
        while (pos < s.size()) {
            if (s[pos] != '<') {
                text.push_back(s[pos++]);
                continue;
            }
            if (s.compare(pos, 4, "<!--") == 0 || s.compare(pos, 2, "<?") == 0) {
                skip_misc(s, pos);
                continue;
            }
            if (s.compare(pos, 2, "</") == 0) {
                const size_t e = s.find('>', pos);
                if (e == std::string::npos) throw PlbIOException("XMLreader: unterminated closing tag");
                std::string name = s.substr(pos + 2, e - pos - 2);
                name.erase(name.find_last_not_of(" \t\r\n") + 1);
                if (name != closing) throw PlbIOException("XMLreader: </" + name + "> does not close <" + closing + ">");
                pos = e + 1;
                parent.text = trim(text);
                return;
            }
            const size_t e = s.find('>', pos);
            if (e == std::string::npos) throw PlbIOException("XMLreader: unterminated tag");
            std::string tag = s.substr(pos + 1, e - pos - 1);
            const bool empty = !tag.empty() && tag.back() == '/';
            if (empty) tag.pop_back();
            Node child;
            child.name = tag.substr(0, tag.find_first_of(" \t\r\n"));
            pos = e + 1;
            if (!empty) parse_children(s, pos, child, child.name);
            parent.children.push_back(std::move(child));
        }
        if (!closing.empty()) throw PlbIOException("XMLreader: <" + closing + "> is not closed");
        parent.text = trim(text);
    }
    static std::string trim(const std::string& t)
    {
        const size_t a = t.find_first_not_of(" \t\r\n");
        if (a == std::string::npos) return "";
        return t.substr(a, t.find_last_not_of(" \t\r\n") - a + 1);
    }
    Node root_;
};

class XMLreaderProxy {
public:
    XMLreaderProxy(const XMLreader::Node* n, const std::string& path) : n_(n), path_(path) {}
    XMLreaderProxy operator[](const std::string& name) const
    {
        if (n_)
            for (const auto& c : n_->children)
                if (c.name == name) return XMLreaderProxy(&c, path_ + "/" + name);
        return XMLreaderProxy(nullptr, path_ + "/" + name);
    }
    template <typename V>
    void read(V& v) const
    {
        if (!n_) throw PlbIOException("XMLreader: no element " + path_);
        std::istringstream is(n_->text);
        is >> v;
        if (is.fail()) throw PlbIOException("XMLreader: cannot read " + path_ + " from \"" + n_->text + "\"");
    }
    void read(std::string& v) const
    {
        if (!n_) throw PlbIOException("XMLreader: no element " + path_);
        v = n_->text;
    }
    void read(bool& v) const
    {
        if (!n_) throw PlbIOException("XMLreader: no element " + path_);
        const std::string& t = n_->text;
        if (t == "true" || t == "1") v = true;
        else if (t == "false" || t == "0") v = false;
        else throw PlbIOException("XMLreader: cannot read a boolean from " + path_);
    }

private:
    const XMLreader::Node* n_;
    std::string path_;
};

inline XMLreaderProxy XMLreader::operator[](const std::string& name) const
{
    return XMLreaderProxy(&root_, "")[name];
}

// The progs' field functionals over the immersed boundary's state
// (src/fsi/MacroProcessors.h): the phase (concentration of cell interior)
// and the external force of every cell, as host fields.
template <typename T, template <typename U> class Descriptor>
struct ComputeConcentrationFunctional {
};
template <typename T, template <typename U> class Descriptor>
struct ForceFieldFunctional {
};

template <typename T, template <typename U> class Descriptor>
void applyProcessingFunctional(ComputeConcentrationFunctional<T, Descriptor>* f, const Box3D&,
                               MultiBlockLattice3D<T, Descriptor>& lattice, MultiScalarField3D<T>& H)
{
    delete f;
    lattice.flushPendingBoundaryStep();
    std::vector<T>& d = H.data();
    if (!lattice.boundaryHandle()) {
        std::fill(d.begin(), d.end(), T(0));
        return;
    }
    std::vector<unsigned char> ph(d.size());
    blCheck(bl_ibm_get_phase(lattice.boundaryHandle(), ph.data()), "ComputeConcentrationFunctional");
    for (size_t k = 0; k < d.size(); ++k) d[k] = (T)ph[k];
}

template <typename T, template <typename U> class Descriptor>
void applyProcessingFunctional(ForceFieldFunctional<T, Descriptor>* f, const Box3D&,
                               MultiBlockLattice3D<T, Descriptor>& lattice, MultiTensorField3D<T, 3>& F)
{
    delete f;
    lattice.flushPendingBoundaryStep();
    F.fill_force(lattice.handle());
}

namespace fsi {
template <class T>
class PipeBoundary;
// PipeBoundary::writeVTK (src/fsi/Boundary.h): the pipe wall as an
// unstructured grid of quads (64 segments around, ends at x0 and x0 + length).
template <class T>
void write_pipe_vtk(const PipeBoundary<T>& b, const std::string& file)
{
    const int n = 64;
    std::ofstream out(file.c_str());
    if (!out) throw std::runtime_error("writeVTK: cannot write " + file);
    out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\">\n<UnstructuredGrid>\n"
        << "<Piece NumberOfPoints=\"" << 2 * n << "\" NumberOfCells=\"" << n << "\">\n<Points>\n"
        << "<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (int e = 0; e < 2; ++e)
        for (int k = 0; k < n; ++k) {
            const T a = 2 * M_PI * k / n;
            out << b.x0 + e * b.length << " " << b.y0 + b.radius * std::cos(a) << " " << b.z0 + b.radius * std::sin(a)
                << "\n";
        }
    out << "</DataArray>\n</Points>\n<Cells>\n<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    for (int k = 0; k < n; ++k) out << k << " " << (k + 1) % n << " " << n + (k + 1) % n << " " << n + k << "\n";
    out << "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    for (int k = 0; k < n; ++k) out << 4 * (k + 1) << "\n";
    out << "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (int k = 0; k < n; ++k) out << 9 << "\n";
    out << "</DataArray>\n</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
}
template <class T>
void PipeBoundary<T>::writeVTK(const std::string& file) const
{
    write_pipe_vtk(*this, file);
}
}  // namespace fsi

}  // namespace plb

#endif
