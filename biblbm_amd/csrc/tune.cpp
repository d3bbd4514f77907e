// The project's autotuner: launch shapes of the collide-and-stream kernel
// per AA step parity, the launch structure of coupled steps and the
// immersed-boundary kernels' launch parameters; results
// kept in a per-process table that can be written to and read from a file
// (one JSON object per line) and that a new lattice of a tuned size picks up
// at creation.  The reference has no tuner: its block decomposition is fixed
// (src/multiBlock/defaultMultiBlockPolicy3D.h:119-129).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <string>
#include <vector>

#include "ibm.hpp"
#include "lattice.hpp"

namespace bl {
namespace {

struct Launch {
    int block = 256, xcd = 1, nt = 0;
    double ms = 0.0;
    int trials = 0;  // tuning trials that timed it
};

struct CollideEntry {
    std::string device;
    int nx = 0, ny = 0, nz = 0, storage = 0;
    Launch par[2];
};

struct CoupledEntry {
    std::string device;
    int nx = 0, ny = 0, nz = 0, storage = 0;
    long vertices = 0;
    int fused = -1, layered = -1;
    double ms = 0.0;
    int lp[BL_IBM_NLAUNCH] = {256, 256, 4, 256, 4, 32, 0};  // bl_ibm_set_launch
    bool has_lp = false;
};

std::vector<CollideEntry> g_collide;
std::vector<CoupledEntry> g_coupled;

std::string device_arch(int device)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
    return std::string(p.gcnArchName);
}

template <class E>
E* find_entry(std::vector<E>& v, const std::string& dev, const bl_lattice_s* L)
{
    for (E& e : v)
        if (e.device == dev && e.nx == L->nx && e.ny == L->ny && e.nz == L->nz && e.storage == L->storage) return &e;
    return nullptr;
}

void put_report(const std::string& rep, char* report, int cap)
{
    if (report && cap > 0) {
        const size_t n = std::min<size_t>(rep.size(), (size_t)cap - 1);
        rep.copy(report, n);
        report[n] = '\0';
    }
}

// The table entry of a boundary (device, lattice size, storage, vertex
// count), created with the boundary's current settings if absent.
CoupledEntry& coupled_entry(bl_ibm g)
{
    const bl_lattice L = g->lat;
    int nv = 0;
    bl_ibm_count(g, nullptr, &nv);
    const std::string dev = device_arch(L->device);
    for (CoupledEntry& c : g_coupled)
        if (c.device == dev && c.nx == L->nx && c.ny == L->ny && c.nz == L->nz && c.storage == L->storage &&
            c.vertices == nv)
            return c;
    g_coupled.emplace_back();
    CoupledEntry& e = g_coupled.back();
    e.device = dev;
    e.nx = L->nx; e.ny = L->ny; e.nz = L->nz; e.storage = L->storage;
    e.vertices = nv;
    e.fused = g->fused_staging;
    e.layered = g->layered;
    return e;
}

}  // namespace

// A lattice of a tuned size on a tuned device starts with the tuned shapes.
void tune_cache_apply(bl_lattice_s* L)
{
    const std::string dev = device_arch(L->device);
    const CollideEntry* e = find_entry(g_collide, dev, L);
    if (!e) return;
    for (int p = 0; p < 2; ++p) {
        L->block[p] = e->par[p].block;
        L->xcd_remap[p] = e->par[p].xcd;
        L->nontemporal[p] = e->par[p].nt;
    }
    ++L->gen;
}

}  // namespace bl

using namespace bl;

extern "C" {

int bl_tune_collide(bl_lattice L, char* report, int cap)
{
    if (!L) return fail(BL_ERR_ARG, "bl_tune_collide: null lattice");
    // candidates, per parity: threads per block x block order x nontemporal
    // hints
    static const int blocks[] = {64, 128, 192, 256};
    static const int orders[] = {0, 1, BL_ORDER_ROWS, 4, 16, 64, 256, 1024};  // see bl_lattice_set_launch_parity
    std::vector<Launch> cand[2];
    for (int xcd : orders)
        for (int nt = 0; nt < 3; ++nt)  // 2: the odd step's z-aligned planes only
            for (int b : blocks) {
                Launch l;
                l.block = b;
                l.xcd = xcd;
                l.nt = nt;
                if (nt < 2) cand[0].push_back(l);
                cand[1].push_back(l);
            }
    // the odd step's z-segment kernel, for the segment lengths the lattice
    // takes
    static const int seg_orders[] = {0, 4, 16, 64};
    for (int b : blocks)
        if (zseg_ok(L, b))
            for (int xcd : seg_orders)
                for (int nt = 0; nt < 2; ++nt) {
                    Launch l;
                    l.block = -b;
                    l.xcd = xcd;
                    l.nt = nt;
                    cand[1].push_back(l);
                }
    const size_t sz = L->pop_bytes;
    const size_t bytes = (size_t)Q * L->stride * sz;
    void* saved = nullptr;
    BL_HIP(hipMalloc(&saved, bytes));
    BL_HIP(hipMemcpyAsync(saved, L->pops, bytes, hipMemcpyDeviceToDevice, L->stream));
    const long long steps0 = L->steps;
    int keep_block[2] = {L->block[0], L->block[1]}, keep_xcd[2] = {L->xcd_remap[0], L->xcd_remap[1]},
        keep_nt[2] = {L->nontemporal[0], L->nontemporal[1]};
    hipEvent_t ev[3];
    for (auto& e : ev) BL_HIP(hipEventCreate(&e));
    int rc = BL_OK;
    if (L->steps & 1) rc = collide_stream_launch(L);  // trials start on an even step
    // trial k: even candidate k mod (even candidates), odd candidate k mod
    // (odd candidates); each candidate's time is the mean over its trials
    const size_t ntrial = std::max(cand[0].size(), cand[1].size());
    const int reps = 4;
    for (size_t k = 0; k < ntrial && rc == BL_OK; ++k) {
        const Launch& e = cand[0][k % cand[0].size()];
        const Launch& o = cand[1][k % cand[1].size()];
        L->block[0] = e.block; L->xcd_remap[0] = e.xcd; L->nontemporal[0] = e.nt;
        L->block[1] = o.block; L->xcd_remap[1] = o.xcd; L->nontemporal[1] = o.nt;
        for (int w = 0; w < 2 && rc == BL_OK; ++w) rc = collide_stream_launch(L);
        double te = 0, to = 0;
        for (int r = 0; r < reps && rc == BL_OK; ++r) {
            hipEventRecord(ev[0], L->stream);
            rc = collide_stream_launch(L);
            hipEventRecord(ev[1], L->stream);
            if (rc == BL_OK) rc = collide_stream_launch(L);
            hipEventRecord(ev[2], L->stream);
            if (rc) break;
            hipEventSynchronize(ev[2]);
            float a = 0.f, b = 0.f;
            hipEventElapsedTime(&a, ev[0], ev[1]);
            hipEventElapsedTime(&b, ev[1], ev[2]);
            te += a;
            to += b;
        }
        Launch& ce = cand[0][k % cand[0].size()];
        Launch& co = cand[1][k % cand[1].size()];
        ce.ms += te / reps;
        co.ms += to / reps;
        ++ce.trials;
        ++co.trials;
    }
    // Refinement: the three fastest candidates of each parity by their sweep
    // means, timed again (2 warm-up launches, `reps2` pairs) with the other
    // parity at its fastest; the pick is the fastest of these.
    const int reps2 = 8;
    std::vector<Launch> top[2];
    for (int p = 0; p < 2; ++p) {
        std::vector<Launch> c = cand[p];
        for (Launch& l : c) l.ms /= l.trials > 0 ? l.trials : 1;
        std::stable_sort(c.begin(), c.end(), [](const Launch& a, const Launch& b) { return a.ms < b.ms; });
        c.resize(std::min<size_t>(3, c.size()));
        top[p] = c;
    }
    for (int p = 0; p < 2 && rc == BL_OK; ++p) {
        const Launch other = top[1 - p][0];
        for (Launch& t : top[p]) {
            const Launch& e = p == 0 ? t : other;
            const Launch& o = p == 0 ? other : t;
            L->block[0] = e.block; L->xcd_remap[0] = e.xcd; L->nontemporal[0] = e.nt;
            L->block[1] = o.block; L->xcd_remap[1] = o.xcd; L->nontemporal[1] = o.nt;
            for (int w = 0; w < 2 && rc == BL_OK; ++w) rc = collide_stream_launch(L);
            double tp = 0;
            for (int r = 0; r < reps2 && rc == BL_OK; ++r) {
                hipEventRecord(ev[0], L->stream);
                rc = collide_stream_launch(L);
                hipEventRecord(ev[1], L->stream);
                if (rc == BL_OK) rc = collide_stream_launch(L);
                hipEventRecord(ev[2], L->stream);
                if (rc) break;
                hipEventSynchronize(ev[2]);
                float a = 0.f, b = 0.f;
                hipEventElapsedTime(&a, ev[0], ev[1]);
                hipEventElapsedTime(&b, ev[1], ev[2]);
                tp += p == 0 ? a : b;
            }
            t.ms = tp / reps2;
        }
    }
    for (auto& e : ev) hipEventDestroy(e);
    // the trials advanced the state: put the populations and the step count back
    hipMemcpyAsync(L->pops, saved, bytes, hipMemcpyDeviceToDevice, L->stream);
    hipStreamSynchronize(L->stream);
    hipFree(saved);
    L->steps = steps0;
    if (rc) {
        for (int p = 0; p < 2; ++p) {
            L->block[p] = keep_block[p];
            L->xcd_remap[p] = keep_xcd[p];
            L->nontemporal[p] = keep_nt[p];
        }
        return rc;
    }
    Launch best[2];
    std::string rep = "{";
    for (int p = 0; p < 2; ++p) {
        rep += p == 0 ? "\"even\": [" : "], \"odd\": [";
        for (size_t k = 0; k < cand[p].size(); ++k) {
            Launch c = cand[p][k];
            c.ms /= c.trials > 0 ? c.trials : 1;
            char buf[128];
            snprintf(buf, sizeof(buf), "%s{\"block\": %d, \"xcd_remap\": %d, \"nontemporal\": %d, \"ms\": %.6f}",
                     k ? ", " : "", c.block, c.xcd, c.nt, c.ms);
            rep += buf;
        }
    }
    rep += "], \"refined\": {";
    for (int p = 0; p < 2; ++p) {
        best[p].ms = 1e30;
        rep += p == 0 ? "\"even\": [" : "], \"odd\": [";
        for (size_t k = 0; k < top[p].size(); ++k) {
            const Launch& c = top[p][k];
            char buf[128];
            snprintf(buf, sizeof(buf), "%s{\"block\": %d, \"xcd_remap\": %d, \"nontemporal\": %d, \"ms\": %.6f}",
                     k ? ", " : "", c.block, c.xcd, c.nt, c.ms);
            rep += buf;
            if (c.ms < best[p].ms) best[p] = c;
        }
        L->block[p] = best[p].block;
        L->xcd_remap[p] = best[p].xcd;
        L->nontemporal[p] = best[p].nt;
    }
    rep += "]}";
    ++L->gen;  // captured steps hold the old launch shapes
    char buf[256];
    snprintf(buf, sizeof(buf),
             ", \"best\": {\"even\": {\"block\": %d, \"xcd_remap\": %d, \"nontemporal\": %d, \"ms\": %.6f}, "
             "\"odd\": {\"block\": %d, \"xcd_remap\": %d, \"nontemporal\": %d, \"ms\": %.6f}}}",
             best[0].block, best[0].xcd, best[0].nt, best[0].ms, best[1].block, best[1].xcd, best[1].nt, best[1].ms);
    rep += buf;
    // remember the result for lattices of this size
    const std::string dev = device_arch(L->device);
    CollideEntry* e = find_entry(g_collide, dev, L);
    if (!e) {
        g_collide.emplace_back();
        e = &g_collide.back();
        e->device = dev;
        e->nx = L->nx; e->ny = L->ny; e->nz = L->nz; e->storage = L->storage;
    }
    e->par[0] = best[0];
    e->par[1] = best[1];
    put_report(rep, report, cap);
    return BL_OK;
}

int bl_tune_coupled(bl_ibm g, int n_steps, char* report, int cap)
{
    if (!g || n_steps < 2 || n_steps % 2) return fail(BL_ERR_ARG, "bl_tune_coupled: n_steps must be even and >= 2");
    int rc = need_stencil(g);
    if (rc) return rc;
    bl_lattice L = g->lat;
    const int keep_fused = g->fused_staging, keep_layered = g->layered;
    const bool phase = L->phase != nullptr && L->omega_c0 != L->omega_c1;
    struct Trial {
        int fused, layered;
        double ms;
    };
    std::vector<Trial> trials;
    for (int f = 0; f < (phase ? 1 : 2); ++f)
        for (int l = 0; l < 2; ++l) trials.push_back(Trial{f, l, 0.0});
    hipEvent_t a, b;
    BL_HIP(hipEventCreate(&a));
    BL_HIP(hipEventCreate(&b));
    std::string rep = "{\"trials\": [";
    Trial best{keep_fused, keep_layered, 1e30};
    for (size_t k = 0; k < trials.size() && rc == BL_OK; ++k) {
        Trial& t = trials[k];
        if ((rc = bl_ibm_set_fused_staging(g, t.fused)) || (rc = bl_ibm_set_layered(g, t.layered))) break;
        // a first run captures this layout's graphs (and settles pending state)
        if ((rc = bl_coupled_steps(g, n_steps, 1))) break;
        hipEventRecord(a, L->stream);
        if ((rc = bl_coupled_steps(g, n_steps, 1))) break;
        hipEventRecord(b, L->stream);
        hipEventSynchronize(b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, a, b);
        t.ms = (double)ms / n_steps;
        char buf[128];
        snprintf(buf, sizeof(buf), "%s{\"fused_staging\": %d, \"layered\": %d, \"ms_per_step\": %.6f}", k ? ", " : "",
                 t.fused, t.layered, t.ms);
        rep += buf;
        if (t.ms < best.ms) best = t;
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc) {
        bl_ibm_set_fused_staging(g, keep_fused);
        bl_ibm_set_layered(g, keep_layered);
        return rc;
    }
    if ((rc = bl_ibm_set_fused_staging(g, best.fused)) || (rc = bl_ibm_set_layered(g, best.layered))) return rc;
    char buf[160];
    snprintf(buf, sizeof(buf), "], \"steps_run\": %d, \"best\": {\"fused_staging\": %d, \"layered\": %d, \"ms_per_step\": %.6f}}",
             (int)(2 * trials.size()) * n_steps, best.fused, best.layered, best.ms);
    rep += buf;
    CoupledEntry& e = coupled_entry(g);
    e.fused = best.fused;
    e.layered = best.layered;
    e.ms = best.ms;
    put_report(rep, report, cap);
    return BL_OK;
}

int bl_tune_ibm(bl_ibm g, int n_steps, char* report, int cap)
{
    if (!g || n_steps < 2 || n_steps % 2) return fail(BL_ERR_ARG, "bl_tune_ibm: n_steps must be even and >= 2");
    int rc = need_stencil(g);
    if (rc) return rc;
    bl_lattice L = g->lat;
    static const std::vector<int> values[BL_IBM_NLAUNCH] = {
        {64, 128, 256}, {64, 128, 256}, {2, 4, 8}, {64, 128, 256}, {1, 2, 4, 8}, {16, 32, 64, 128}, {0, 2, 4, 6}};
    static const char* names[BL_IBM_NLAUNCH] = {"bin_block", "elem_block", "gather_lanes", "interp_block",
                                                "spread_small", "spread_big", "xcd"};
    int keep[BL_IBM_NLAUNCH], best[BL_IBM_NLAUNCH];
    if ((rc = bl_ibm_get_launch(g, keep))) return rc;
    std::copy(keep, keep + BL_IBM_NLAUNCH, best);
    hipEvent_t a, b;
    BL_HIP(hipEventCreate(&a));
    BL_HIP(hipEventCreate(&b));
    // time of n_steps with parameters p (a first run captures the graphs)
    auto time_of = [&](const int* p, double& ms) -> int {
        int r;
        if ((r = bl_ibm_set_launch(g, p)) || (r = bl_coupled_steps(g, n_steps, 1))) return r;
        hipEventRecord(a, L->stream);
        if ((r = bl_coupled_steps(g, n_steps, 1))) return r;
        hipEventRecord(b, L->stream);
        hipEventSynchronize(b);
        float t = 0.f;
        hipEventElapsedTime(&t, a, b);
        ms = (double)t / n_steps;
        return BL_OK;
    };
    double best_ms = 0.0;
    int runs = 1;
    rc = time_of(best, best_ms);
    std::string rep = "{\"trials\": [";
    bool first = true;
    for (int pass = 0; pass < 2 && rc == BL_OK; ++pass)
        for (int k = 0; k < BL_IBM_NLAUNCH && rc == BL_OK; ++k) {
            // every other value of the parameter once per pass, whichever
            // wins on the way (the number of trials is always the same)
            const int cur = best[k];
            for (int v : values[k]) {
                if (v == cur) continue;
                int p[BL_IBM_NLAUNCH];
                std::copy(best, best + BL_IBM_NLAUNCH, p);
                p[k] = v;
                double ms = 0.0;
                if ((rc = time_of(p, ms))) break;
                ++runs;
                char buf[128];
                snprintf(buf, sizeof(buf), "%s{\"%s\": %d, \"ms_per_step\": %.6f}", first ? "" : ", ", names[k], v, ms);
                rep += buf;
                first = false;
                if (ms < best_ms) {
                    best_ms = ms;
                    best[k] = v;
                }
            }
        }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc) {
        bl_ibm_set_launch(g, keep);
        return rc;
    }
    if ((rc = bl_ibm_set_launch(g, best))) return rc;
    rep += "], \"steps_run\": " + std::to_string(2 * runs * n_steps) + ", \"best\": {";
    for (int k = 0; k < BL_IBM_NLAUNCH; ++k) {
        char buf[64];
        snprintf(buf, sizeof(buf), "\"%s\": %d, ", names[k], best[k]);
        rep += buf;
    }
    char buf[64];
    snprintf(buf, sizeof(buf), "\"ms_per_step\": %.6f}}", best_ms);
    rep += buf;
    CoupledEntry& e = coupled_entry(g);
    std::copy(best, best + BL_IBM_NLAUNCH, e.lp);
    e.has_lp = true;
    put_report(rep, report, cap);
    return BL_OK;
}

int bl_tune_apply_coupled(bl_ibm g)
{
    if (!g || !g->lat) return fail(BL_ERR_ARG, "bl_tune_apply_coupled: no lattice");
    const bl_lattice L = g->lat;
    int nv = 0;
    bl_ibm_count(g, nullptr, &nv);
    const std::string dev = device_arch(L->device);
    for (const CoupledEntry& c : g_coupled)
        if (c.device == dev && c.nx == L->nx && c.ny == L->ny && c.nz == L->nz && c.storage == L->storage &&
            c.vertices == nv) {
            int rc = bl_ibm_set_fused_staging(g, c.fused);
            if (!rc) rc = bl_ibm_set_layered(g, c.layered);
            if (!rc && c.has_lp) rc = bl_ibm_set_launch(g, c.lp);
            return rc;
        }
    return BL_OK;
}

int bl_tune_cache_save(const char* path)
{
    if (!path) return fail(BL_ERR_ARG, "bl_tune_cache_save: null path");
    std::ofstream out(path);
    if (!out) return fail(BL_ERR_ARG, std::string("bl_tune_cache_save: cannot write ") + path);
    for (const CollideEntry& e : g_collide) {
        char buf[512];
        snprintf(buf, sizeof(buf),
                 "{\"kind\": \"collide\", \"device\": \"%s\", \"nx\": %d, \"ny\": %d, \"nz\": %d, \"storage\": %d, "
                 "\"even\": [%d, %d, %d, %.6f], \"odd\": [%d, %d, %d, %.6f]}",
                 e.device.c_str(), e.nx, e.ny, e.nz, e.storage, e.par[0].block, e.par[0].xcd, e.par[0].nt, e.par[0].ms,
                 e.par[1].block, e.par[1].xcd, e.par[1].nt, e.par[1].ms);
        out << buf << "\n";
    }
    for (const CoupledEntry& e : g_coupled) {
        char buf[512];
        snprintf(buf, sizeof(buf),
                 "{\"kind\": \"coupled\", \"device\": \"%s\", \"nx\": %d, \"ny\": %d, \"nz\": %d, \"storage\": %d, "
                 "\"vertices\": %ld, \"fused_staging\": %d, \"layered\": %d, \"ms_per_step\": %.6f, "
                 "\"ibm_launch\": [%d, %d, %d, %d, %d, %d, %d]}",
                 e.device.c_str(), e.nx, e.ny, e.nz, e.storage, e.vertices, e.fused, e.layered, e.ms,
                 e.has_lp ? e.lp[0] : 0, e.lp[1], e.lp[2], e.lp[3], e.lp[4], e.lp[5], e.lp[6]);
        out << buf << "\n";
    }
    return out ? BL_OK : fail(BL_ERR_ARG, std::string("bl_tune_cache_save: write failed: ") + path);
}

int bl_tune_cache_load(const char* path)
{
    if (!path) return fail(BL_ERR_ARG, "bl_tune_cache_load: null path");
    std::ifstream in(path);
    if (!in) return fail(BL_ERR_ARG, std::string("bl_tune_cache_load: cannot read ") + path);
    std::string line;
    int bad = 0;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        char dev[128] = {0};
        CollideEntry c;
        CoupledEntry d;
        if (sscanf(line.c_str(),
                   "{\"kind\": \"collide\", \"device\": \"%127[^\"]\", \"nx\": %d, \"ny\": %d, \"nz\": %d, "
                   "\"storage\": %d, \"even\": [%d, %d, %d, %lf], \"odd\": [%d, %d, %d, %lf]}",
                   dev, &c.nx, &c.ny, &c.nz, &c.storage, &c.par[0].block, &c.par[0].xcd, &c.par[0].nt, &c.par[0].ms,
                   &c.par[1].block, &c.par[1].xcd, &c.par[1].nt, &c.par[1].ms) == 13) {
            bool ok = true;
            for (int p = 0; p < 2; ++p) {
                const int lanes = c.par[p].block < 0 ? -c.par[p].block : c.par[p].block;
                ok = ok && lanes >= 64 && lanes <= 256 && lanes % 64 == 0 &&
                     (c.par[p].block > 0 || c.par[p].xcd != BL_ORDER_ROWS) &&
                     c.par[p].xcd >= 0 && c.par[p].xcd <= (1 << 20) &&
                     (c.par[p].xcd <= 1 || c.par[p].xcd == BL_ORDER_ROWS || (c.par[p].xcd & (c.par[p].xcd - 1)) == 0) &&
                     c.par[p].nt >= 0 && c.par[p].nt <= 2;
            }
            if (!ok) {
                ++bad;
                continue;
            }
            c.device = dev;
            bool found = false;
            for (CollideEntry& e : g_collide)
                if (e.device == c.device && e.nx == c.nx && e.ny == c.ny && e.nz == c.nz && e.storage == c.storage) {
                    e = c;
                    found = true;
                }
            if (!found) g_collide.push_back(c);
        } else if (sscanf(line.c_str(),
                          "{\"kind\": \"coupled\", \"device\": \"%127[^\"]\", \"nx\": %d, \"ny\": %d, \"nz\": %d, "
                          "\"storage\": %d, \"vertices\": %ld, \"fused_staging\": %d, \"layered\": %d, "
                          "\"ms_per_step\": %lf, \"ibm_launch\": [%d, %d, %d, %d, %d, %d, %d]}",
                          dev, &d.nx, &d.ny, &d.nz, &d.storage, &d.vertices, &d.fused, &d.layered, &d.ms, &d.lp[0],
                          &d.lp[1], &d.lp[2], &d.lp[3], &d.lp[4], &d.lp[5], &d.lp[6]) == 16) {
            if (d.fused < -1 || d.fused > 1 || d.layered < -1 || d.layered > 1) {
                ++bad;
                continue;
            }
            // ibm_launch [0, ...]: no launch parameters tuned for this entry
            d.has_lp = d.lp[0] != 0;
            auto in = [](int v, std::initializer_list<int> ok) {
                return std::find(ok.begin(), ok.end(), v) != ok.end();
            };
            if (d.has_lp && !(in(d.lp[0], {64, 128, 256}) && in(d.lp[1], {64, 128, 256}) && in(d.lp[2], {2, 4, 8}) &&
                              in(d.lp[3], {64, 128, 256}) && in(d.lp[4], {1, 2, 4, 8}) &&
                              in(d.lp[5], {16, 32, 64, 128}) && in(d.lp[6], {0, 2, 4, 6}))) {
                ++bad;
                continue;
            }
            if (!d.has_lp) {
                const int dflt[BL_IBM_NLAUNCH] = {256, 256, 4, 256, 4, 32, 0};
                std::copy(dflt, dflt + BL_IBM_NLAUNCH, d.lp);
            }
            d.device = dev;
            bool found = false;
            for (CoupledEntry& e : g_coupled)
                if (e.device == d.device && e.nx == d.nx && e.ny == d.ny && e.nz == d.nz && e.storage == d.storage &&
                    e.vertices == d.vertices) {
                    e = d;
                    found = true;
                }
            if (!found) g_coupled.push_back(d);
        } else {
            ++bad;
        }
    }
    if (bad) return fail(BL_ERR_ARG, std::string("bl_tune_cache_load: ") + std::to_string(bad) + " unreadable line(s) in " + path);
    return BL_OK;
}

}  // extern "C"
