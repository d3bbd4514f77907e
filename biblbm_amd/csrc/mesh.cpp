// Rest geometry and connectivity of a membrane shape (host, FP64, setup time).
//
// Reproduces the reference's ParticleShape construction
// (src/fsi/ParticleShape.hh:21-58): per-triangle area from the normal,
// signed volume, area-weighted centroid; links are the unique edges sorted by
// (i0, i1) with their first and second incident triangle
// (src/fsi/ParticleShape.hh:90-145); an adjacent pair is emitted per link
// with two triangles, in link order, rotated so the shared edge is (i0, i2)
// (src/fsi/ParticleShape.hh:147-180); the rest dihedral angle uses the
// normals of (v0,v1,v2) and (v0,v2,v3) (src/fsi/TriangleUtils.h:148-162);
// radius is the largest vertex distance from the origin
// (src/fsi/ParticleShape.hh:361-392).
#include <cmath>
#include <map>
#include <string>
#include <utility>

#include "mesh.hpp"

namespace bl {

namespace {

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

inline void unit_normal(V3 a, V3 b, V3 c, V3& n)
{
    n = cross(sub(b, a), sub(c, a));
    const double l = norm(n);
    n = {n.x / l, n.y / l, n.z / l};
}

}  // namespace

bool build_shape(int nv, int nt, const double* rest_pos, const int* tri, ShapeHost& s, std::string* err)
{
    auto V = [&](int i) { return V3{rest_pos[3 * i], rest_pos[3 * i + 1], rest_pos[3 * i + 2]}; };
    s = ShapeHost();
    s.nv = nv;
    s.nt = nt;
    s.rest.assign(rest_pos, rest_pos + 3 * (size_t)nv);
    s.tri.assign(tri, tri + 3 * (size_t)nt);
    for (int k = 0; k < 3 * nt; ++k)
        if (tri[k] < 0 || tri[k] >= nv) {
            if (err) *err = "triangle index out of range";
            return false;
        }

    // per-triangle rest quantities, totals
    s.tri_area0.resize(nt);
    s.tri_geo0.resize(4 * (size_t)nt);
    double cx = 0, cy = 0, cz = 0;
    for (int t = 0; t < nt; ++t) {
        const V3 a = V(tri[3 * t]), b = V(tri[3 * t + 1]), c = V(tri[3 * t + 2]);
        const V3 e0 = sub(b, a), e1 = sub(c, a);
        const V3 n = cross(e0, e1);
        const double nn = norm(n);
        if (!(nn > 0)) {
            if (err) *err = "degenerate triangle";
            return false;
        }
        const double area = 0.5 * nn;
        s.tri_area0[t] = area;
        const double l0 = norm(e0), l0p = norm(e1);
        s.tri_geo0[4 * t] = l0;
        s.tri_geo0[4 * t + 1] = l0p;
        s.tri_geo0[4 * t + 2] = dot(e0, e1) / (l0 * l0p);
        s.tri_geo0[4 * t + 3] = norm(cross(e0, e1)) / (l0 * l0p);
        s.area += area;
        s.volume += dot(a, cross(b, c)) / 6.;
        cx += ((a.x + b.x + c.x) / 3.0) * area;
        cy += ((a.y + b.y + c.y) / 3.0) * area;
        cz += ((a.z + b.z + c.z) / 3.0) * area;
    }
    s.center[0] = cx / s.area;
    s.center[1] = cy / s.area;
    s.center[2] = cz / s.area;
    double r2 = 0;
    for (int i = 0; i < nv; ++i) {
        const V3 p = V(i);
        r2 = std::max(r2, dot(p, p));
    }
    s.radius = std::sqrt(r2);

    // unique links with their incident triangles
    std::map<std::pair<int, int>, std::pair<int, int>> edges;  // (i0,i1) -> (t1, t2)
    for (int t = 0; t < nt; ++t) {
        const int ids[4] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2], tri[3 * t]};
        for (int e = 0; e < 3; ++e) {
            const int a = std::min(ids[e], ids[e + 1]), b = std::max(ids[e], ids[e + 1]);
            auto it = edges.find({a, b});
            if (it == edges.end()) {
                edges[{a, b}] = {t, -1};
            } else if (it->second.first != t) {
                if (it->second.second != -1 && it->second.second != t) {
                    if (err) *err = "non-manifold edge (more than two triangles)";
                    return false;
                }
                it->second.second = t;
            }
        }
    }
    s.nl = (int)edges.size();
    for (const auto& e : edges) {
        s.link.push_back(e.first.first);
        s.link.push_back(e.first.second);
        s.link_tri.push_back(e.second.first);
        s.link_tri.push_back(e.second.second);
        s.link_len.push_back(norm(sub(V(e.first.first), V(e.first.second))));
    }

    // adjacent triangle pairs, one per interior link, in link order
    for (int l = 0; l < s.nl; ++l) {
        const int ta = s.link_tri[2 * l], tb = s.link_tri[2 * l + 1];
        if (ta < 0 || tb < 0) continue;
        const int i0 = s.link[2 * l], i1 = s.link[2 * l + 1];
        int t1[3] = {tri[3 * ta], tri[3 * ta + 1], tri[3 * ta + 2]};
        int t2[3] = {tri[3 * tb], tri[3 * tb + 1], tri[3 * tb + 2]};
        auto rot = [](int* t) {  // (a, b, c) -> (c, a, b)
            const int c = t[2];
            t[2] = t[1];
            t[1] = t[0];
            t[0] = c;
        };
        int guard = 0;
        while (!((t1[0] == i0 && t1[2] == i1) || (t1[0] == i1 && t1[2] == i0)) && guard++ < 3) rot(t1);
        guard = 0;
        while (!((t2[0] == i0 && t2[1] == i1) || (t2[0] == i1 && t2[1] == i0)) && guard++ < 3) rot(t2);
        const int q[4] = {t1[0], t1[1], t1[2], t2[2]};
        for (int k = 0; k < 4; ++k) s.pair.push_back(q[k]);
        V3 n1, n2;
        unit_normal(V(q[0]), V(q[1]), V(q[2]), n1);
        unit_normal(V(q[0]), V(q[2]), V(q[3]), n2);
        const double cs = dot(n1, n2);
        const double sg = dot(sub(n1, n2), sub(V(q[1]), V(q[3]))) > 0 ? 1.0 : -1.0;
        s.pair_cs0.push_back(cs);
        s.pair_cs0.push_back(sg * norm(cross(n1, n2)));
    }
    s.np = (int)s.pair_cs0.size() / 2;

    // per-vertex incidence (CSR), element order preserved
    auto csr = [&](int nelem, int per, const std::vector<int>& idx, std::vector<int>& off, std::vector<int>& ent) {
        std::vector<int> cnt(nv + 1, 0);
        for (int e = 0; e < nelem; ++e)
            for (int k = 0; k < per; ++k) cnt[idx[per * e + k] + 1]++;
        off.assign(nv + 1, 0);
        for (int v = 0; v < nv; ++v) off[v + 1] = off[v] + cnt[v + 1];
        ent.assign(off[nv], 0);
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int e = 0; e < nelem; ++e)
            for (int k = 0; k < per; ++k) ent[fill[idx[per * e + k]]++] = e * 4 + k;
    };
    csr(s.nl, 2, s.link, s.v_link_off, s.v_link);
    csr(s.nt, 3, s.tri, s.v_tri_off, s.v_tri);
    csr(s.np, 4, s.pair, s.v_pair_off, s.v_pair);
    return true;
}

}  // namespace bl
