// On-disk formats of the immersed boundary and the lattice (host side).
//
//   particle pack      ParticleBase3D::pack (src/fsi/ParticleBase.hh:218-224):
//                      type id, shape id, particle id (plint each), should-
//                      voxelize (bool); DeformableParticle3D::pack
//                      (src/fsi/DeformableParticle3D.hh:277-287): pos, vel,
//                      force per vertex (3 doubles each); then
//                      DeformableCapsuleParticle (src/fsi/DeformableCapsule.hh:
//                      217-223): shear model (int, Skalak 0 / NeoHookean 1)
//                      and C, kb, nu_s, G; or RBCParticle
//                      (src/fsi/RBCParticle.hh:95-105,346-350): L0,
//                      shear_modulus, k_area_global, k_area_local, k_bend,
//                      k_volume, theta0, vol_desired.
//   fsi checkpoint     header <folder>/<name>.plb: particle count, file count
//                      (1) and the rank file's name <name>/0.plb; the rank
//                      file: particle count (plint) and the packs
//                      (src/fsi/ImmersedBoundaryDynamics_io.hh:17-73); loading
//                      re-creates the particles from the packs
//                      (src/fsi/ImmersedBoundaryDynamics_io.hh:75-158).
//   VTU mesh           ParticleBase3D::write_vtk (src/fsi/ParticleBase.hh:
//                      270-398): appended raw data, Float32 points and
//                      force_tot, Int32 connectivity / offsets / types, UInt32
//                      block sizes.
//   gmsh mesh          ParticleBase3D::write_gmsh (src/fsi/ParticleBase.hh:
//                      400-437), version 2.1 ASCII.
//   lightweight text   write_lightweight_particle_data / write_lightweight
//                      (src/fsi/ImmersedBoundaryDynamics_io.hh:178-200,
//                      src/fsi/DeformableParticle3D.hh:318-333): per particle
//                      tag:id:com:gyration (upper triangle):volume.
//   lattice checkpoint this library's own layout (the reference's Palabos
//                      parallelIO format is not restated): a text line, the
//                      sizes, storage, step count, the populations in their
//                      stored form and the external force of every cell.
//
// Type ids are the reference build's registration order
// (DeformableCapsuleParticle 0, RBCParticle 1).  A particle that neither
// type can hold (the Skalak + bending + area + volume membrane, or a capsule
// without the reference's c term) is written with type id kTypeOwn and its
// model, flag and eight parameters after the vertex block.
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ibm.hpp"

using namespace bl;

namespace {

constexpr long long kTypeCapsule = 0;
constexpr long long kTypeRbc = 1;
constexpr long long kTypeOwn = 100;

template <class T>
void put(std::vector<char>& b, const T& v)
{
    const char* p = reinterpret_cast<const char*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
}

template <class T>
bool take(const char*& it, const char* end, T& v)
{
    if (end - it < (long)sizeof(T)) return false;
    std::memcpy(&v, it, sizeof(T));
    it += sizeof(T);
    return true;
}

// pos / vel / force of particle pid from the device
int vertex_state(bl_ibm g, int pid, std::vector<double>& pos, std::vector<double>& vel, std::vector<double>& frc)
{
    const size_t n = 3 * (size_t)g->shapes[g->parts[pid].shape].nv;
    pos.resize(n);
    vel.resize(n);
    frc.resize(n);
    int rc;
    if ((rc = bl_particle_get(g, pid, 0, pos.data())) || (rc = bl_particle_get(g, pid, 1, vel.data())) ||
        (rc = bl_particle_get(g, pid, 2, frc.data())))
        return rc;
    return BL_OK;
}

int pack_particle(bl_ibm g, int pid, std::vector<char>& b)
{
    const auto& p = g->parts[pid];
    if (p.model == BL_MODEL_RIGID)  // RigidParticle3D::pack (src/fsi/RigidParticle.hh:114-130) is not restated
        return fail(BL_ERR_STATE, "checkpoint: rigid particles have no packed form here");
    std::vector<double> pos, vel, frc;
    int rc = vertex_state(g, pid, pos, vel, frc);
    if (rc) return rc;
    const bool capsule = (p.model == BL_MODEL_CAPSULE_SKALAK || p.model == BL_MODEL_CAPSULE_NEOHOOKEAN) && p.bug;
    const bool rbc = p.model == BL_MODEL_RBC;
    const long long type = capsule ? kTypeCapsule : (rbc ? kTypeRbc : kTypeOwn);
    put<long long>(b, type);
    put<long long>(b, (long long)p.shape);
    put<long long>(b, p.id);
    put<bool>(b, true);
    const int nv = g->shapes[p.shape].nv;
    for (int v = 0; v < nv; ++v) {
        for (int d = 0; d < 3; ++d) put<double>(b, pos[3 * v + d]);
        for (int d = 0; d < 3; ++d) put<double>(b, vel[3 * v + d]);
        for (int d = 0; d < 3; ++d) put<double>(b, frc[3 * v + d]);
    }
    const double* P = p.params;
    if (capsule) {  // params: G, C, nu_s, kb
        put<int>(b, p.model == BL_MODEL_CAPSULE_SKALAK ? 0 : 1);
        put<double>(b, P[1]);
        put<double>(b, P[3]);
        put<double>(b, P[2]);
        put<double>(b, P[0]);
    } else if (rbc) {  // params: G, L0, k_ag, k_al, k_v, V0, kb, theta0
        put<double>(b, P[1]);
        put<double>(b, P[0]);
        put<double>(b, P[2]);
        put<double>(b, P[3]);
        put<double>(b, P[6]);
        put<double>(b, P[4]);
        put<double>(b, P[7]);
        put<double>(b, P[5]);
    } else {
        put<int>(b, p.model);
        put<int>(b, p.bug);
        for (int k = 0; k < 8; ++k) put<double>(b, P[k]);
    }
    return BL_OK;
}

// One pack into a new particle (not yet initialised boundary); its id is
// returned in *id, its vertex velocities and forces kept for the upload at
// init.
int unpack_particle(bl_ibm g, const char*& it, const char* end, long long* id)
{
    long long type, shape;
    bool vox;
    if (!take(it, end, type) || !take(it, end, shape) || !take(it, end, *id) || !take(it, end, vox))
        return fail(BL_ERR_ARG, "checkpoint: truncated particle header");
    if (type != kTypeCapsule && type != kTypeRbc && type != kTypeOwn)
        return fail(BL_ERR_ARG, "checkpoint: unknown particle type " + std::to_string(type));
    if (shape < 0 || shape >= (long long)g->shapes.size())
        return fail(BL_ERR_ARG, "checkpoint: unknown shape id " + std::to_string(shape));
    const int nv = g->shapes[shape].nv;
    bl_ibm_s::PartHost p;
    p.shape = (int)shape;
    p.pos.resize(3 * (size_t)nv);
    p.vel.resize(3 * (size_t)nv);
    p.force.resize(3 * (size_t)nv);
    for (int v = 0; v < nv; ++v) {
        bool ok = true;
        for (int d = 0; d < 3; ++d) ok = ok && take(it, end, p.pos[3 * v + d]);
        for (int d = 0; d < 3; ++d) ok = ok && take(it, end, p.vel[3 * v + d]);
        for (int d = 0; d < 3; ++d) ok = ok && take(it, end, p.force[3 * v + d]);
        if (!ok) return fail(BL_ERR_ARG, "checkpoint: truncated vertex block");
    }
    double* P = p.params;
    bool ok = true;
    if (type == kTypeCapsule) {
        int sm;
        double C, kb, nu, G;
        ok = take(it, end, sm) && take(it, end, C) && take(it, end, kb) && take(it, end, nu) && take(it, end, G);
        p.model = sm == 0 ? BL_MODEL_CAPSULE_SKALAK : BL_MODEL_CAPSULE_NEOHOOKEAN;
        p.bug = 1;
        P[0] = G;
        P[1] = C;
        P[2] = nu;
        P[3] = kb;
        if (ok && sm != 0 && sm != 1) return fail(BL_ERR_ARG, "checkpoint: unknown shear model");
    } else if (type == kTypeRbc) {
        double L0, G, kag, kal, kb, kv, th, V0;
        ok = take(it, end, L0) && take(it, end, G) && take(it, end, kag) && take(it, end, kal) &&
             take(it, end, kb) && take(it, end, kv) && take(it, end, th) && take(it, end, V0);
        p.model = BL_MODEL_RBC;
        p.bug = 1;
        const double q[8] = {G, L0, kag, kal, kv, V0, kb, th};
        for (int k = 0; k < 8; ++k) P[k] = q[k];
    } else {
        ok = take(it, end, p.model) && take(it, end, p.bug);
        for (int k = 0; k < 8 && ok; ++k) ok = take(it, end, P[k]);
        if (ok && (p.model < 0 || p.model > 3)) return fail(BL_ERR_ARG, "checkpoint: unknown membrane model");
    }
    if (!ok) return fail(BL_ERR_ARG, "checkpoint: truncated particle parameters");
    p.id = *id;
    g->parts.push_back(std::move(p));
    return BL_OK;
}

// A regular file (not a directory, device or missing path): reading a
// directory through a filebuf throws, so every path a checkpoint names is
// checked before it is opened.
bool regular_file(const std::string& path)
{
    struct stat st;
    return ::stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string join(const std::string& folder, const std::string& file)
{
    if (folder.empty()) return file;
    return folder.back() == '/' ? folder + file : folder + "/" + file;
}

// upper triangle of the tensor of gyration about the centre of mass
// (src/fsi/DeformableParticle3D.hh:111-125)
void gyration(const std::vector<double>& pos, const double com[3], double A[6])
{
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const size_t nv = pos.size() / 3;
    for (size_t v = 0; v < nv; ++v)
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j) M[i][j] += (pos[3 * v + i] - com[i]) * (pos[3 * v + j] - com[j]);
    int k = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) A[k++] = M[i][j] / (double)nv;
}

}  // namespace

extern "C" {

int bl_shape_set_tag(bl_ibm g, int shape_id, const char* tag)
{
    if (!g || !tag) return fail(BL_ERR_ARG, "bl_shape_set_tag: null argument");
    if (shape_id < 0 || shape_id >= (int)g->shapes.size()) return fail(BL_ERR_ARG, "bl_shape_set_tag: unknown shape");
    g->shapes[shape_id].tag = tag;
    return BL_OK;
}

int bl_particle_describe(bl_ibm g, int pid, int* shape_id, int* model, long long* id)
{
    if (!g) return fail(BL_ERR_ARG, "null ibm handle");
    if (pid < 0 || pid >= (int)g->parts.size()) return fail(BL_ERR_ARG, "bl_particle_describe: bad particle index");
    const auto& p = g->parts[pid];
    if (shape_id) *shape_id = p.shape;
    if (model) *model = p.model;
    if (id) *id = p.id;
    return BL_OK;
}

int bl_particle_pack(bl_ibm g, int pid, char* buf, long long cap, long long* len)
{
    if (!g || !len) return fail(BL_ERR_ARG, "bl_particle_pack: null argument");
    if (!g->inited) return fail(BL_ERR_STATE, "bl_particle_pack: immersed boundary not initialised");
    if (pid < 0 || pid >= g->np) return fail(BL_ERR_ARG, "bl_particle_pack: bad particle id");
    std::vector<char> b;
    int rc = pack_particle(g, pid, b);
    if (rc) return rc;
    *len = (long long)b.size();
    if (buf) {
        if (cap < (long long)b.size()) return fail(BL_ERR_ARG, "bl_particle_pack: buffer too small");
        std::memcpy(buf, b.data(), b.size());
    }
    return BL_OK;
}

int bl_ibm_save_checkpoint(bl_ibm g, const char* folder, const char* name)
{
    if (!g || !folder || !name) return fail(BL_ERR_ARG, "bl_ibm_save_checkpoint: null argument");
    if (!g->inited) return fail(BL_ERR_STATE, "bl_ibm_save_checkpoint: immersed boundary not initialised");
    const std::string base(folder), nm(name);
    const std::string rank_file = nm + "/0.plb";
    {
        std::ofstream h(join(base, nm + ".plb"));
        if (!h) return fail(BL_ERR_ARG, "bl_ibm_save_checkpoint: cannot write the header in " + base);
        h << g->np << std::endl << 1 << std::endl << rank_file << std::endl;
    }
    std::vector<char> b;
    put<long long>(b, (long long)g->np);
    for (int p = 0; p < g->np; ++p) {
        int rc = pack_particle(g, p, b);
        if (rc) return rc;
    }
    std::ofstream out(join(base, rank_file), std::ios::binary);
    if (!out)
        return fail(BL_ERR_ARG, "bl_ibm_save_checkpoint: cannot write " + join(base, rank_file) +
                                    " (the folder " + join(base, nm) + " must exist)");
    out.write(b.data(), (std::streamsize)b.size());
    return out ? BL_OK : fail(BL_ERR_ARG, "bl_ibm_save_checkpoint: write failed");
}

namespace {
int load_checkpoint(bl_ibm g, const char* folder, const char* name);
}

int bl_ibm_load_checkpoint(bl_ibm g, const char* folder, const char* name)
{
    if (!g || !folder || !name) return fail(BL_ERR_ARG, "bl_ibm_load_checkpoint: null argument");
    if (g->inited) return fail(BL_ERR_STATE, "bl_ibm_load_checkpoint: particles must be loaded before bl_ibm_init");
    std::vector<bl_ibm_s::PartHost> keep = g->parts;
    try {
        return load_checkpoint(g, folder, name);
    } catch (const std::exception& e) {  // no exception leaves the C ABI
        g->parts.swap(keep);
        return fail(BL_ERR_ARG, std::string("bl_ibm_load_checkpoint: ") + e.what());
    }
}

}  // extern "C"

namespace {
int load_checkpoint(bl_ibm g, const char* folder, const char* name)
{
    const std::string base(folder), nm(name);
    if (!regular_file(join(base, nm + ".plb")))
        return fail(BL_ERR_ARG, "bl_ibm_load_checkpoint: cannot open " + join(base, nm + ".plb"));
    std::ifstream h(join(base, nm + ".plb"));
    if (!h) return fail(BL_ERR_ARG, "bl_ibm_load_checkpoint: cannot open " + join(base, nm + ".plb"));
    long long total = -1, nfiles = -1;
    h >> total >> nfiles;
    std::string line;
    std::getline(h, line);
    if (!h || total < 0 || nfiles < 1) return fail(BL_ERR_ARG, "bl_ibm_load_checkpoint: bad header");
    std::vector<bl_ibm_s::PartHost> keep;
    keep.swap(g->parts);  // the reference clears the particles first
    std::vector<long long> ids;
    for (long long f = 0; f < nfiles; ++f) {
        std::getline(h, line);
        std::ifstream in;
        if (regular_file(join(base, line))) in.open(join(base, line), std::ios::binary);
        if (!in.is_open() || !in) {
            g->parts.swap(keep);
            return fail(BL_ERR_ARG, "bl_ibm_load_checkpoint: cannot open " + join(base, line));
        }
        std::vector<char> data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        const char* it = data.data();
        const char* end = it + data.size();
        long long n = 0;
        if (!take(it, end, n) || n < 0) {
            g->parts.swap(keep);
            return fail(BL_ERR_ARG, "bl_ibm_load_checkpoint: bad rank file " + line);
        }
        for (long long k = 0; k < n; ++k) {
            long long id;
            int rc = unpack_particle(g, it, end, &id);
            if (rc) {
                g->parts.swap(keep);
                return rc;
            }
            ids.push_back(id);
        }
    }
    // particles are kept in increasing id order (the reference's particle
    // map); ids must be distinct
    const size_t n = g->parts.size();
    std::vector<size_t> order(n);
    for (size_t k = 0; k < n; ++k) order[k] = k;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ids[a] < ids[b]; });
    for (size_t k = 1; k < n; ++k)
        if (ids[order[k]] == ids[order[k - 1]]) {
            g->parts.swap(keep);
            return fail(BL_ERR_ARG, "bl_ibm_load_checkpoint: repeated particle id " + std::to_string(ids[order[k]]));
        }
    if ((long long)n != total) {
        g->parts.swap(keep);
        return fail(BL_ERR_ARG, "bl_ibm_load_checkpoint: the header's particle count does not match the files");
    }
    std::vector<bl_ibm_s::PartHost> sorted(n);
    for (size_t k = 0; k < n; ++k) sorted[k] = std::move(g->parts[order[k]]);
    g->parts.swap(sorted);
    return BL_OK;
}
}  // namespace

extern "C" {

int bl_particle_write_vtk(bl_ibm g, int pid, const char* path)
{
    if (!g || !path) return fail(BL_ERR_ARG, "bl_particle_write_vtk: null argument");
    if (!g->inited) return fail(BL_ERR_STATE, "bl_particle_write_vtk: immersed boundary not initialised");
    if (pid < 0 || pid >= g->np) return fail(BL_ERR_ARG, "bl_particle_write_vtk: bad particle id");
    std::vector<double> pos, vel, frc;
    int rc = vertex_state(g, pid, pos, vel, frc);
    if (rc) return rc;
    const ShapeHost& s = g->shapes[g->parts[pid].shape];
    const unsigned nv = (unsigned)s.nv, nt = (unsigned)s.nt;
    const unsigned sizes[5] = {3 * nv * 4u, 3 * nv * 4u, 3 * nt * 4u, nt * 4u, nt * 4u};
    unsigned long long off[5];
    off[0] = 0;
    for (int i = 1; i < 5; ++i) off[i] = off[i - 1] + 4 + sizes[i - 1];
    std::ofstream out(path, std::ios::out | std::ios::trunc | std::ios::binary);
    if (!out) return fail(BL_ERR_ARG, std::string("bl_particle_write_vtk: cannot write ") + path);
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\" >\n";
    out << "  <UnstructuredGrid>\n";
    out << "    <Piece NumberOfPoints=\"" << nv << "\" NumberOfCells=\"" << nt << "\" >" << std::endl;
    out << "      <Points>" << std::endl;
    out << "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" << off[0]
        << "\" />" << std::endl;
    out << "      </Points>" << std::endl;
    out << "      <PointData>" << std::endl;
    out << "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Name=\"force_tot\" format=\"appended\" offset=\""
        << off[1] << "\" />" << std::endl;
    out << "      </PointData>" << std::endl;
    out << "      <Cells>" << std::endl;
    out << "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"appended\" offset=\"" << off[2] << "\" />"
        << std::endl;
    out << "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"appended\" offset=\"" << off[3] << "\" />"
        << std::endl;
    out << "        <DataArray type=\"Int32\" Name=\"types\" format=\"appended\" offset=\"" << off[4] << "\" />"
        << std::endl;
    out << "      </Cells>" << std::endl;
    out << "    </Piece>" << std::endl;
    out << "  </UnstructuredGrid>" << std::endl;
    out << "  <AppendedData encoding=\"raw\">" << std::endl;
    out << "_";
    std::vector<char> b;
    put<unsigned>(b, sizes[0]);
    for (double x : pos) put<float>(b, (float)x);
    put<unsigned>(b, sizes[1]);
    for (double x : frc) put<float>(b, (float)x);
    put<unsigned>(b, sizes[2]);
    for (int k : s.tri) put<int>(b, k);
    put<unsigned>(b, sizes[3]);
    for (unsigned t = 0; t < nt; ++t) put<int>(b, (int)(3 * (t + 1)));
    put<unsigned>(b, sizes[4]);
    for (unsigned t = 0; t < nt; ++t) put<int>(b, 5);
    out.write(b.data(), (std::streamsize)b.size());
    out << std::endl << "  </AppendedData>" << std::endl;
    out << "</VTKFile>" << std::endl;
    return out ? BL_OK : fail(BL_ERR_ARG, "bl_particle_write_vtk: write failed");
}

int bl_particle_write_gmsh(bl_ibm g, int pid, const char* path)
{
    if (!g || !path) return fail(BL_ERR_ARG, "bl_particle_write_gmsh: null argument");
    if (!g->inited) return fail(BL_ERR_STATE, "bl_particle_write_gmsh: immersed boundary not initialised");
    if (pid < 0 || pid >= g->np) return fail(BL_ERR_ARG, "bl_particle_write_gmsh: bad particle id");
    const ShapeHost& s = g->shapes[g->parts[pid].shape];
    std::vector<double> pos(3 * (size_t)s.nv);
    int rc = bl_particle_get(g, pid, 0, pos.data());
    if (rc) return rc;
    std::ofstream out(path, std::ios::out | std::ios::trunc);
    if (!out) return fail(BL_ERR_ARG, std::string("bl_particle_write_gmsh: cannot write ") + path);
    out << "$MeshFormat" << std::endl << "2.1 0 " << sizeof(double) << std::endl << "$EndMeshFormat" << std::endl;
    out << "$Nodes" << std::endl << s.nv << std::endl;
    for (int v = 0; v < s.nv; ++v)
        out << v << " " << pos[3 * v] << " " << pos[3 * v + 1] << " " << pos[3 * v + 2] << std::endl;
    out << "$EndNodes" << std::endl;
    out << "$Elements" << std::endl << s.nt << std::endl;
    for (int t = 0; t < s.nt; ++t)
        out << t << " 2 3 0 1 0 " << s.tri[3 * t] << " " << s.tri[3 * t + 1] << " " << s.tri[3 * t + 2] << std::endl;
    out << "$EndElements";
    return out ? BL_OK : fail(BL_ERR_ARG, "bl_particle_write_gmsh: write failed");
}

int bl_ibm_write_lightweight(bl_ibm g, const char* path, long long iteration)
{
    if (!g || !path) return fail(BL_ERR_ARG, "bl_ibm_write_lightweight: null argument");
    if (!g->inited) return fail(BL_ERR_STATE, "bl_ibm_write_lightweight: immersed boundary not initialised");
    std::ofstream out(path, std::ios::out);
    if (!out) return fail(BL_ERR_ARG, std::string("bl_ibm_write_lightweight: cannot write ") + path);
    out << "iteration:" << iteration << std::endl;
    for (int p = 0; p < g->np; ++p) {
        double obs[14];
        int rc = bl_particle_observables(g, p, obs);
        if (rc) return rc;
        std::vector<double> pos(3 * (size_t)g->shapes[g->parts[p].shape].nv);
        if ((rc = bl_particle_get(g, p, 0, pos.data()))) return rc;
        double A[6];
        gyration(pos, obs + 2, A);
        out << g->shapes[g->parts[p].shape].tag << ":" << g->parts[p].id;
        for (int d = 0; d < 3; ++d) out << ":" << obs[2 + d];
        for (int k = 0; k < 6; ++k) out << ":" << A[k];
        out << ":" << obs[1];
        out << std::endl;
    }
    return out ? BL_OK : fail(BL_ERR_ARG, "bl_ibm_write_lightweight: write failed");
}

}  // extern "C"

// ------------------------------------------------------ lattice checkpoint
namespace {
const char kLatticeMagic[] = "biblbm_amd lattice checkpoint 1\n";
}

extern "C" {

int bl_lattice_save_checkpoint(bl_lattice L, const char* path)
{
    if (!L || !path) return fail(BL_ERR_ARG, "bl_lattice_save_checkpoint: null argument");
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) return fail(BL_ERR_ARG, std::string("bl_lattice_save_checkpoint: cannot write ") + path);
    out.write(kLatticeMagic, sizeof(kLatticeMagic) - 1);
    const int dims[4] = {L->nx, L->ny, L->nz, L->storage};
    const long long meta[3] = {L->steps, L->ncell, L->force_field ? 1 : 0};
    out.write((const char*)dims, sizeof(dims));
    out.write((const char*)meta, sizeof(meta));
    BL_HIP(hipStreamSynchronize(L->stream));
    std::vector<char> plane((size_t)L->ncell * L->pop_bytes);
    for (int i = 0; i < Q; ++i) {
        BL_HIP(hipMemcpy(plane.data(), (const char*)L->pops + (size_t)i * L->stride * L->pop_bytes, plane.size(),
                         hipMemcpyDeviceToHost));
        out.write(plane.data(), (std::streamsize)plane.size());
    }
    if (L->force_field) {
        std::vector<double> F(3 * (size_t)L->ncell);
        int rc = bl_lattice_get_force(L, F.data());
        if (rc) return rc;
        out.write((const char*)F.data(), (std::streamsize)(F.size() * sizeof(double)));
    }
    return out ? BL_OK : fail(BL_ERR_ARG, "bl_lattice_save_checkpoint: write failed");
}

int bl_lattice_load_checkpoint(bl_lattice L, const char* path)
{
    if (!L || !path) return fail(BL_ERR_ARG, "bl_lattice_load_checkpoint: null argument");
    if (!regular_file(path)) return fail(BL_ERR_ARG, std::string("bl_lattice_load_checkpoint: cannot open ") + path);
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(BL_ERR_ARG, std::string("bl_lattice_load_checkpoint: cannot open ") + path);
    char magic[sizeof(kLatticeMagic) - 1];
    int dims[4];
    long long meta[3];
    in.read(magic, sizeof(magic));
    in.read((char*)dims, sizeof(dims));
    in.read((char*)meta, sizeof(meta));
    if (!in || std::memcmp(magic, kLatticeMagic, sizeof(magic)) != 0)
        return fail(BL_ERR_ARG, "bl_lattice_load_checkpoint: not a lattice checkpoint");
    if (dims[0] != L->nx || dims[1] != L->ny || dims[2] != L->nz || dims[3] != L->storage || meta[1] != L->ncell)
        return fail(BL_ERR_ARG, "bl_lattice_load_checkpoint: the lattice's size or storage differs from the file's");
    std::vector<char> pops((size_t)Q * L->ncell * L->pop_bytes);
    in.read(pops.data(), (std::streamsize)pops.size());
    std::vector<double> F;
    if (meta[2]) {
        F.resize(3 * (size_t)L->ncell);
        in.read((char*)F.data(), (std::streamsize)(F.size() * sizeof(double)));
    }
    if (!in) return fail(BL_ERR_ARG, "bl_lattice_load_checkpoint: truncated file");
    BL_HIP(hipStreamSynchronize(L->stream));
    const size_t plane = (size_t)L->ncell * L->pop_bytes;
    for (int i = 0; i < Q; ++i)
        BL_HIP(hipMemcpy((char*)L->pops + (size_t)i * L->stride * L->pop_bytes, pops.data() + i * plane, plane,
                         hipMemcpyHostToDevice));
    L->steps = meta[0];
    ++L->gen;
    if (L->ibm) ibm_invalidate_graphs(L->ibm);
    if (meta[2]) return bl_lattice_set_force_field(L, F.data());
    return BL_OK;
}

}  // extern "C"
