// Helpers shared by the golden-vector drivers that are compiled against the
// reference sources (Palabos + fsi in serial mode).  They only write .npy
// files (numpy format 1.0, little endian) and provide a small deterministic
// random generator so every fixture is reproducible from its seed.
#ifndef BIBLBM_REF_COMMON_H
#define BIBLBM_REF_COMMON_H

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include <sstream>
#include <sys/stat.h>

namespace refio {

inline void ensure_dir(const std::string& d) { mkdir(d.c_str(), 0755); }

template <class T> inline const char* npy_descr();
template <> inline const char* npy_descr<double>() { return "<f8"; }
template <> inline const char* npy_descr<long long>() { return "<i8"; }
template <> inline const char* npy_descr<int>() { return "<i4"; }
template <> inline const char* npy_descr<unsigned char>() { return "|u1"; }

template <class T>
void write_npy(const std::string& path, const T* data, const std::vector<long>& shape)
{
    std::ostringstream h;
    h << "{'descr': '" << npy_descr<T>() << "', 'fortran_order': False, 'shape': (";
    long n = 1;
    for (size_t i = 0; i < shape.size(); ++i) {
        h << shape[i] << (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
        n *= shape[i];
    }
    h << "), }";
    std::string hs = h.str();
    size_t total = 10 + hs.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hs.append(pad, ' ');
    hs.push_back('\n');
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(1); }
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    std::fwrite(magic, 1, 8, f);
    unsigned short hl = (unsigned short)hs.size();
    std::fwrite(&hl, 2, 1, f);
    std::fwrite(hs.data(), 1, hs.size(), f);
    std::fwrite(data, sizeof(T), (size_t)n, f);
    std::fclose(f);
}

template <class T>
void write_npy(const std::string& path, const std::vector<T>& v, const std::vector<long>& shape)
{
    write_npy(path, v.empty() ? (const T*)0 : &v[0], shape);
}

// 64-bit LCG (Knuth MMIX constants) -> doubles in [0, 1).
struct Rng {
    unsigned long long s;
    explicit Rng(unsigned long long seed) : s(seed * 2862933555777941757ULL + 3037000493ULL) {}
    double uniform()
    {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

}  // namespace refio

#endif
