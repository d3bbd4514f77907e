// Host-side rest geometry and topology of a triangulated membrane shape.
#pragma once

#include <string>
#include <vector>

namespace bl {

struct ShapeHost {
    int nv = 0, nt = 0, nl = 0, np = 0;
    std::vector<double> rest;       // 3*nv
    std::vector<int> tri;           // 3*nt
    std::vector<double> tri_area0;  // nt
    std::vector<double> tri_geo0;   // 4*nt: l0, l0', cos(phi0), sin(phi0) of edges (v1-v0, v2-v0)
    std::vector<int> link;          // 2*nl, i0 < i1, sorted
    std::vector<int> link_tri;      // 2*nl, first / second incident triangle (-1 if none)
    std::vector<double> link_len;   // nl
    std::vector<int> pair;          // 4*np: (i0,i1,i2) and (i0,i2,i3) share edge (i0,i2)
    std::vector<double> pair_cs0;   // 2*np: cos, sin of the rest dihedral angle
    double area = 0, volume = 0, center[3] = {0, 0, 0}, radius = 0;
    std::string tag;  // shape descriptor (the reference's ParticleShape tag)

    // Per-vertex incidence lists (CSR), in the order the reference
    // accumulates: links, then triangles, then pairs (element index order).
    // Entries encode element * 4 + slot.
    std::vector<int> v_link_off, v_link;  // nv+1, entries
    std::vector<int> v_tri_off, v_tri;
    std::vector<int> v_pair_off, v_pair;
};

// Builds links, adjacent pairs and rest quantities as the reference
// ParticleShape does (src/fsi/ParticleShape.hh:21-212,361-392).
// Returns false (with a message) on a non-manifold or degenerate mesh.
bool build_shape(int nv, int nt, const double* rest_pos, const int* tri, ShapeHost& out,
                 std::string* err);

}  // namespace bl
