// Host-side robustness of the library's parsing and set-up code, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (no GPU calls):
//   io.cpp    bl_ibm_load_checkpoint on a valid checkpoint made from the
//             reference's packed particles, on every truncation of its rank
//             file and header, and on randomly garbled copies;
//   mesh.cpp  build_shape on meshes with out-of-range, negative, repeated
//             and degenerate triangle indices and non-manifold edges;
//   init.cpp  bl_positions_* on empty, huge, NaN and overfull requests.
// Every call must return a status (BL_OK or an error) without a sanitizer
// report.  Usage: host_fuzz <golden dir> <scratch dir>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "ibm.hpp"
#include "mesh.hpp"

extern "C" {
int bl_ibm_load_checkpoint(bl_ibm g, const char* folder, const char* name);
}

namespace {

std::vector<char> read_file(const std::string& p)
{
    std::ifstream in(p, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

void write_file(const std::string& p, const std::vector<char>& d)
{
    std::ofstream out(p, std::ios::binary);
    out.write(d.data(), (std::streamsize)d.size());
}

// gmsh 2.2 ASCII: nodes, then triangles (element type 2)
bool read_msh(const std::string& p, std::vector<double>& v, std::vector<int>& t)
{
    std::ifstream in(p);
    std::string line;
    while (std::getline(in, line) && line != "$Nodes") {
    }
    int nv = 0, first = 0;
    in >> nv;
    v.resize(3 * (size_t)nv);
    for (int i = 0; i < nv; ++i) {
        int id;
        in >> id >> v[3 * i] >> v[3 * i + 1] >> v[3 * i + 2];
        if (i == 0) first = id;  // node numbering starts at 0 or 1
    }
    while (std::getline(in, line) && line != "$Elements") {
    }
    int ne = 0;
    in >> ne;
    for (int e = 0; e < ne; ++e) {
        int id, type, ntag;
        in >> id >> type >> ntag;
        for (int k = 0; k < ntag; ++k) {
            int tag;
            in >> tag;
        }
        const int nn = type == 2 ? 3 : (type == 15 ? 1 : (type == 1 ? 2 : 0));
        int q[3] = {0, 0, 0};
        for (int k = 0; k < nn; ++k) in >> q[k];
        if (type == 2)
            for (int k = 0; k < 3; ++k) t.push_back(q[k] - first);
    }
    return in.good() && nv > 0 && !t.empty();
}

struct Lcg {
    unsigned long long s;
    unsigned next() { return (unsigned)((s = s * 6364136223846793005ULL + 1442695040888963407ULL) >> 33); }
};

int checks = 0, errors = 0;
const char* stage = "";

void count(int rc)
{
    ++checks;
    if (rc != 0) ++errors;
}

// An exception that leaves a C entry point is a failure: name the stage.
[[noreturn]] void escaped(const char* what)
{
    std::fprintf(stderr, "host_fuzz: an exception left the C ABI in %s (call %d): %s\n", stage, checks, what);
    std::exit(1);
}

int load(bl_ibm_s& g, const std::string& dir)
{
    g.parts.clear();
    return bl_ibm_load_checkpoint(&g, dir.c_str(), "checkpoint_fsi");
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: host_fuzz <golden dir> <scratch dir>\n");
        return 2;
    }
    const std::string golden = argv[1], scratch = argv[2];

  try {
    // ---- mesh.cpp ----
    stage = "build_shape";
    std::vector<double> rest;
    std::vector<int> tri;
    if (!read_msh(golden + "/io/sphere626/rbc.msh", rest, tri)) {
        std::fprintf(stderr, "cannot read the golden mesh\n");
        return 2;
    }
    const int nv = (int)rest.size() / 3, nt = (int)tri.size() / 3;
    bl_ibm_s g;
    {
        bl::ShapeHost s;
        std::string err;
        if (!bl::build_shape(nv, nt, rest.data(), tri.data(), s, &err)) {
            std::fprintf(stderr, "golden mesh refused: %s\n", err.c_str());
            return 1;
        }
        g.shapes.push_back(s);
    }
    Lcg r{12345};
    for (int k = 0; k < 400; ++k) {
        std::vector<int> bad = tri;
        const int m = 1 + (int)(r.next() % 4);
        for (int j = 0; j < m; ++j) {
            const size_t at = r.next() % bad.size();
            switch (r.next() % 5) {
            case 0: bad[at] = -1 - (int)(r.next() % 3); break;          // negative
            case 1: bad[at] = nv + (int)(r.next() % 5); break;           // past the end
            case 2: bad[at] = bad[(at / 3) * 3 + (at + 1) % 3]; break;   // repeated in its triangle
            case 3: bad[at] = (int)(r.next() % nv); break;               // another vertex (non-manifold)
            default: bad[at] = std::numeric_limits<int>::max(); break;
            }
        }
        std::vector<double> pos = rest;
        if (k % 7 == 0) pos[3 * (r.next() % nv)] = std::nan("");
        bl::ShapeHost s;
        std::string err;
        const bool ok = bl::build_shape(nv, nt - (int)(k % 3), pos.data(), bad.data(), s, &err);
        count(ok ? 0 : 1);
    }

    // ---- io.cpp ----
    stage = "bl_ibm_load_checkpoint";
    const std::string names[3] = {"pack_skalak.bin", "pack_neohookean.bin", "pack_rbc.bin"};
    std::vector<char> rank;
    const long long n = 3;
    rank.insert(rank.end(), (const char*)&n, (const char*)&n + sizeof(n));
    for (int k = 0; k < 3; ++k) {
        std::vector<char> p = read_file(golden + "/io/sphere626/" + names[k]);
        if (p.empty()) {
            std::fprintf(stderr, "missing %s\n", names[k].c_str());
            return 2;
        }
        // each packed particle carries its id after type and shape: 0, 1, 2
        long long id = k;
        std::memcpy(p.data() + 16, &id, sizeof(id));
        rank.insert(rank.end(), p.begin(), p.end());
    }
    const std::string dir = scratch + "/ck";
    if (std::system(("mkdir -p " + dir + "/checkpoint_fsi").c_str()) != 0) return 2;
    const std::vector<char> header = {'3', ' ', '1', '\n', 'c', 'h', 'e', 'c', 'k', 'p', 'o', 'i', 'n', 't', '_', 'f',
                                      's', 'i', '/', '0', '.', 'p', 'l', 'b', '\n'};
    write_file(dir + "/checkpoint_fsi.plb", header);
    write_file(dir + "/checkpoint_fsi/0.plb", rank);
    if (load(g, dir) != 0 || g.parts.size() != 3) {
        std::fprintf(stderr, "the valid checkpoint did not load: %s\n", bl_last_error());
        return 1;
    }
    for (size_t cut = 0; cut < rank.size(); cut += (cut < 64 ? 1 : 97)) {  // truncations of the rank file
        write_file(dir + "/checkpoint_fsi/0.plb", std::vector<char>(rank.begin(), rank.begin() + cut));
        const int rc = load(g, dir);
        if (rc == 0) {
            std::fprintf(stderr, "a truncated rank file (%zu of %zu bytes) loaded\n", cut, rank.size());
            return 1;
        }
        count(rc);
    }
    for (int k = 0; k < 300; ++k) {  // garbled bytes
        std::vector<char> bad = rank;
        const int m = 1 + (int)(r.next() % 8);
        for (int j = 0; j < m; ++j) bad[r.next() % bad.size()] = (char)(r.next() & 0xff);
        write_file(dir + "/checkpoint_fsi/0.plb", bad);
        count(load(g, dir));
    }
    write_file(dir + "/checkpoint_fsi/0.plb", rank);
    for (size_t cut = 0; cut < header.size(); ++cut) {  // truncated and garbled headers
        write_file(dir + "/checkpoint_fsi.plb", std::vector<char>(header.begin(), header.begin() + cut));
        count(load(g, dir));
        std::vector<char> bad = header;
        bad[cut] = (char)(r.next() & 0xff);
        write_file(dir + "/checkpoint_fsi.plb", bad);
        count(load(g, dir));
    }

    // ---- init.cpp ----
    stage = "bl_positions";
    const int boxes[][6] = {{0, 10, 0, 10, 0, 10}, {0, 0, 0, 10, 0, 10}, {5, 1, 0, 10, 0, 10},
                            {0, 1000000000, 0, 1000000000, 0, 1000000000}, {-5, 5, -5, 5, -5, 5}};
    const double radii[] = {1.0, 0.0, -1.0, std::nan(""), 1e-9, 1e300, std::numeric_limits<double>::infinity()};
    for (const auto& b : boxes)
        for (double mr : radii) {
            bl_positions h = nullptr;
            const int rc = bl_positions_create(b, nullptr, mr, &h);
            count(rc);
            if (rc) continue;
            const double pr[] = {mr, mr / 2, 2 * mr, 0.0, std::nan("")};
            for (double rr : pr) {
                count(bl_positions_generate(h, 5, rr, 'R'));
                count(bl_positions_generate(h, -1, rr, 'R'));
            }
            long long cnt = 0;
            double buf[3 * 16];
            count(bl_positions_get(h, 'R', buf, 16, &cnt));
            count(bl_positions_get(h, 'R', buf, 0, &cnt));
            bl_positions_destroy(h);
        }
    std::printf("host_fuzz: %d calls, %d refused with an error, no sanitizer report\n", checks, errors);
    return 0;
  } catch (const std::exception& e) {
    escaped(e.what());
  }
}
