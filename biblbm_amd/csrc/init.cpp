// Initial particle positions (host side): ParticlePositionInitializer
// (src/fsi/ParticlePositionInitializer.hh:18-86).
//
// Points are drawn with the C library's rand(), three per attempt (x, y, z
// in that order), as box corner + radius + rand()/RAND_MAX * (extent -
// 2 radius); an attempt is kept when the sphere lies inside the boundary
// (margin = radius) and overlaps no sphere placed before (distance^2 <
// (r1 + r2)^2); at most 5e6 attempts per point.  A caller that seeds
// rand() as the reference's prog does gets the same points.  The overlap
// search goes through a uniform grid of cells no smaller than twice the
// largest radius (so only the 27 cells around a point can hold an overlap),
// the reference's Grid (src/fsi/Grid.hh:30-73) in spirit; the accepted
// points do not depend on it.
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "walls.hpp"

struct bl_positions_s {
    double lo[3], ext[3];  // box corner and extent (x1 - x0 per axis)
    bool has_boundary = false;
    bl_boundary boundary;
    double max_radius = 0;
    int nc[3] = {1, 1, 1};
    double cell[3] = {1, 1, 1};
    std::vector<std::vector<int>> grid;  // point indices per cell
    std::vector<double> pos, radius;
    std::vector<char> tag;

    long cell_of(const double* p, int c[3]) const
    {
        for (int a = 0; a < 3; ++a) {
            int k = (int)std::floor((p[a] - lo[a]) / cell[a]);
            c[a] = k < 0 ? 0 : (k >= nc[a] ? nc[a] - 1 : k);
        }
        return ((long)c[0] * nc[1] + c[1]) * nc[2] + c[2];
    }

    bool overlaps(const double* p, double r) const
    {
        int c[3];
        cell_of(p, c);
        for (int i = std::max(c[0] - 1, 0); i <= std::min(c[0] + 1, nc[0] - 1); ++i)
            for (int j = std::max(c[1] - 1, 0); j <= std::min(c[1] + 1, nc[1] - 1); ++j)
                for (int k = std::max(c[2] - 1, 0); k <= std::min(c[2] + 1, nc[2] - 1); ++k)
                    for (int q : grid[((long)i * nc[1] + j) * nc[2] + k]) {
                        const double dx = pos[3 * q] - p[0], dy = pos[3 * q + 1] - p[1], dz = pos[3 * q + 2] - p[2];
                        const double s = radius[q] + r;
                        if (dx * dx + dy * dy + dz * dz < s * s) return true;
                    }
        return false;
    }
};

using namespace bl;

extern "C" {

int bl_positions_create(const int box[6], const bl_boundary* boundary, double max_radius, bl_positions* out)
{
    if (!box || !out) return fail(BL_ERR_ARG, "bl_positions_create: null argument");
    if (!(max_radius > 0)) return fail(BL_ERR_ARG, "bl_positions_create: max_radius must be positive");
    // the search grid has cells of at least 2 max_radius: at most 2^24 of them
    double cells = 1.0;
    for (int a = 0; a < 3; ++a) {
        const double ext = (double)box[2 * a + 1] - (double)box[2 * a];
        if (!(ext > 0)) return fail(BL_ERR_ARG, "bl_positions_create: empty box");
        cells *= std::max(1.0, std::floor(ext / (2 * max_radius)));
    }
    if (!(cells <= (double)(1 << 24)))
        return fail(BL_ERR_CAPACITY, "bl_positions_create: the box holds more than 2^24 cells of 2 max_radius");
    bl_positions h = new bl_positions_s();
    for (int a = 0; a < 3; ++a) {
        h->lo[a] = box[2 * a];
        h->ext[a] = (double)box[2 * a + 1] - (double)box[2 * a];
        h->nc[a] = (int)std::max(1.0, std::floor(h->ext[a] / (2 * max_radius)));
        h->cell[a] = h->ext[a] / h->nc[a];
    }
    h->grid.resize((size_t)h->nc[0] * h->nc[1] * h->nc[2]);
    if (boundary) {
        h->has_boundary = true;
        h->boundary = *boundary;
    }
    h->max_radius = max_radius;
    *out = h;
    return BL_OK;
}

int bl_positions_generate(bl_positions h, long long n, double radius, char tag)
{
    if (!h) return fail(BL_ERR_ARG, "null positions handle");
    if (n < 0 || !(radius > 0)) return fail(BL_ERR_ARG, "bl_positions_generate: bad count or radius");
    if (2 * radius > 2 * h->max_radius)
        return fail(BL_ERR_ARG, "bl_positions_generate: particle radius is larger than the grid spacing");
    for (long long k = 0; k < n; ++k) {
        double p[3];
        bool ok = false;
        for (long i = 0; i < 5000000L && !ok; ++i) {
            for (int a = 0; a < 3; ++a)
                p[a] = h->lo[a] + radius + ((double)std::rand() / (double)RAND_MAX) * (h->ext[a] - 2 * radius);
            if (h->has_boundary && !bnd_contains(h->boundary, p, radius)) continue;
            ok = !h->overlaps(p, radius);
        }
        if (!ok)
            return fail(BL_ERR_CAPACITY, "bl_positions_generate: no room for point " + std::to_string(k) + " of " +
                                             std::to_string(n));
        const int idx = (int)h->radius.size();
        h->pos.insert(h->pos.end(), p, p + 3);
        h->radius.push_back(radius);
        h->tag.push_back(tag);
        int c[3];
        h->grid[h->cell_of(p, c)].push_back(idx);
    }
    return BL_OK;
}

int bl_positions_get(bl_positions h, char tag, double* out, long long cap, long long* n)
{
    if (!h || !n) return fail(BL_ERR_ARG, "bl_positions_get: null argument");
    long long m = 0;
    for (size_t q = 0; q < h->tag.size(); ++q) {
        if (h->tag[q] != tag) continue;
        if (out) {
            if (m >= cap) return fail(BL_ERR_ARG, "bl_positions_get: buffer too small");
            for (int a = 0; a < 3; ++a) out[3 * m + a] = h->pos[3 * q + a];
        }
        ++m;
    }
    *n = m;
    return BL_OK;
}

int bl_positions_destroy(bl_positions h)
{
    delete h;
    return BL_OK;
}

}  // extern "C"
