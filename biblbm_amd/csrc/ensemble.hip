// Ensembles of independent boxes advanced together.
//
// An ensemble groups lattices (each optionally with its immersed boundary)
// created through the ordinary entry points.  One hipGraph holds two
// coupled steps (both AA parities) of every box.  Batched (the default
// whenever the boxes allow it: same storage and parity, paired-stage IBM
// steps, no Guo walls or phase fields): every stage of a step is one launch
// over all boxes, each box's workgroups running with its own arguments
// (ensemble_batch_build, ibm.hip).  Otherwise each box is a parallel branch
// forked from and joined into the ensemble's stream.  The reference runs
// such boxes as separate programs.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "ibm.hpp"
#include "lattice.hpp"

struct bl_ensemble_s {
    std::vector<bl_lattice> lats;
    std::vector<bl_ibm> ibms;  // nullptr entries: lattice-only boxes
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t graph = nullptr;
    // what the graph was captured for, per box: AA parity, tile-list buffer,
    // pending list reset, lattice and boundary generations
    struct Key {
        long long state;
        unsigned long long lgen, igen;
        bool operator==(const Key& o) const { return state == o.state && lgen == o.lgen && igen == o.igen; }
        bool operator!=(const Key& o) const { return !(*this == o); }
    };
    std::vector<Key> parity;
    bl::EnsembleBatch* batch = nullptr;  // the tables of the captured batched steps
    int mode = -1;                       // -1 batched when possible, 0 branches, 1 batched only
    int order[2] = {0, 0};               // batched collide workgroup order per AA parity
    int nt[2] = {0, 0};                  // and nontemporal level
};

namespace bl {
namespace {

int box_step(bl_ensemble e, size_t i)
{
    if (e->ibms[i]) return ibm_step_launch(e->ibms[i], false);  // the boxes are the parallel branches
    int rc = collide_stream_launch(e->lats[i]);
    return rc ? rc : velocity_walls_launch(e->lats[i]);
}

// the ensemble stream waits for each box's pending work, and the box
// streams wait for the ensemble stream afterwards
int order_before(bl_ensemble e)
{
    for (size_t i = 0; i < e->lats.size(); ++i) {
        BL_HIP(hipEventRecord(e->join[i], e->lats[i]->stream));
        BL_HIP(hipStreamWaitEvent(e->stream, e->join[i], 0));
    }
    return BL_OK;
}

int order_after(bl_ensemble e)
{
    BL_HIP(hipEventRecord(e->done, e->stream));
    for (bl_lattice L : e->lats) BL_HIP(hipStreamWaitEvent(L->stream, e->done, 0));
    return BL_OK;
}

// what a captured step depends on: the AA parity of the lattice, the
// tile-list buffer and pending list reset of its immersed boundary, and the
// generations that every setter of either bumps
bl_ensemble_s::Key capture_key(bl_ensemble e, size_t i)
{
    const bl_ibm g = e->ibms[i];
    const long long st = (e->lats[i]->steps & 1) | (g ? (long long)g->alt << 1 : 0) |
                         (g ? (long long)g->prepped << 2 : 0);
    return bl_ensemble_s::Key{st, e->lats[i]->gen, g ? g->gen : 0ull};
}

bool parity_matches(bl_ensemble e)
{
    if (!e->graph || e->parity.size() != e->lats.size()) return false;
    for (size_t i = 0; i < e->lats.size(); ++i)
        if (capture_key(e, i) != e->parity[i]) return false;
    return true;
}

int capture_batched(bl_ensemble e)
{
    BL_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = ensemble_batch_launch(e->batch, e->stream);
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(e->stream, &graph);
    if (rc != BL_OK || ee != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return rc ? rc : fail(BL_ERR_HIP, std::string("ensemble capture: ") + hipGetErrorString(ee));
    }
    const hipError_t ie = hipGraphInstantiate(&e->graph, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(BL_ERR_HIP, std::string("ensemble graph: ") + hipGetErrorString(ie));
    e->parity.resize(e->lats.size());
    for (size_t i = 0; i < e->lats.size(); ++i) e->parity[i] = capture_key(e, i);
    return BL_OK;
}

int capture(bl_ensemble e)
{
    // a replay still running may read the old tables
    if (e->graph || e->batch) BL_HIP(hipStreamSynchronize(e->stream));
    if (e->graph) {
        hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    ensemble_batch_free(e->batch);
    e->batch = nullptr;
    if (e->mode != 0) {
        int rc = ensemble_batch_build(e->lats, e->ibms, e->order, e->nt, &e->batch);
        if (rc) return rc;
        if (e->batch) return capture_batched(e);
        if (e->mode == 1)
            return fail(BL_ERR_STATE, "ensemble: the boxes cannot run batched (storage, parity, step layout, "
                                      "walls or pending state differ)");
    }
    std::vector<long long> s0(e->lats.size());
    for (size_t i = 0; i < e->lats.size(); ++i) s0[i] = e->lats[i]->steps;
    BL_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = BL_OK;
    hipError_t he = hipEventRecord(e->fork, e->stream);
    for (size_t i = 0; i < e->lats.size() && rc == BL_OK && he == hipSuccess; ++i) {
        hipStream_t s = e->lats[i]->stream;
        he = hipStreamWaitEvent(s, e->fork, 0);
        for (int k = 0; k < 2 && rc == BL_OK && he == hipSuccess; ++k) rc = box_step(e, i);
        if (rc == BL_OK && he == hipSuccess) he = hipEventRecord(e->join[i], s);
        if (rc == BL_OK && he == hipSuccess) he = hipStreamWaitEvent(e->stream, e->join[i], 0);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(e->stream, &graph);
    for (size_t i = 0; i < e->lats.size(); ++i) e->lats[i]->steps = s0[i];  // capture ran nothing
    if (rc != BL_OK) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (he != hipSuccess || ee != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return fail(BL_ERR_HIP, std::string("ensemble capture: ") + hipGetErrorString(he != hipSuccess ? he : ee));
    }
    const hipError_t ie = hipGraphInstantiate(&e->graph, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(BL_ERR_HIP, std::string("ensemble graph: ") + hipGetErrorString(ie));
    e->parity.resize(e->lats.size());
    for (size_t i = 0; i < e->lats.size(); ++i) e->parity[i] = capture_key(e, i);
    return BL_OK;
}

bool needs_settle(bl_ensemble e)
{
    for (bl_ibm g : e->ibms)
        if (g && ibm_needs_settle(g)) return true;
    return false;
}

int run_steps(bl_ensemble e, int n, int use_graph)
{
    int rc;
    int done = 0;
    // force fields loaded into a lattice are consumed over two eager steps
    for (; done < n && needs_settle(e); ++done)
        for (size_t i = 0; i < e->lats.size(); ++i)
            if ((rc = box_step(e, i))) return rc;
    if (use_graph && n - done >= 2) {
        if (!parity_matches(e) && (rc = capture(e))) return rc;
        if ((rc = order_before(e))) return rc;
        for (; done + 2 <= n; done += 2) {
            BL_HIP(hipGraphLaunch(e->graph, e->stream));
            for (bl_lattice L : e->lats) L->steps += 2;
        }
        if ((rc = order_after(e))) return rc;
    }
    for (; done < n; ++done)
        for (size_t i = 0; i < e->lats.size(); ++i)
            if ((rc = box_step(e, i))) return rc;
    return BL_OK;
}

}  // namespace
}  // namespace bl

using namespace bl;

extern "C" {

int bl_ensemble_create(int n_boxes, const bl_lattice* lats, const bl_ibm* ibms, int flags, bl_ensemble* out)
{
    if (n_boxes <= 0 || !lats || !out) return fail(BL_ERR_ARG, "bl_ensemble_create: bad argument");
    if (flags != 0) return fail(BL_ERR_ARG, "bl_ensemble_create: flags must be 0");
    for (int i = 0; i < n_boxes; ++i) {
        if (!lats[i]) return fail(BL_ERR_ARG, "bl_ensemble_create: null lattice");
        for (int j = 0; j < i; ++j)
            if (lats[j] == lats[i]) return fail(BL_ERR_ARG, "bl_ensemble_create: a lattice appears twice");
        if (ibms && ibms[i]) {
            if (ibms[i]->lat != lats[i])
                return fail(BL_ERR_ARG, "bl_ensemble_create: immersed boundary belongs to another lattice");
            if (!ibms[i]->inited) return fail(BL_ERR_STATE, "bl_ensemble_create: immersed boundary not initialised");
            int rc = need_stencil(ibms[i]);
            if (rc) return rc;
        }
    }
    bl_ensemble e = new bl_ensemble_s();
    e->lats.assign(lats, lats + n_boxes);
    e->ibms.assign((size_t)n_boxes, nullptr);
    if (ibms)
        for (int i = 0; i < n_boxes; ++i) e->ibms[i] = ibms[i];
    e->join.assign((size_t)n_boxes, nullptr);
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e->done, hipEventDisableTiming);
    for (int i = 0; i < n_boxes && he == hipSuccess; ++i) he = hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming);
    if (he != hipSuccess) {
        bl_ensemble_destroy(e);
        return fail(BL_ERR_HIP, std::string("bl_ensemble_create: ") + hipGetErrorString(he));
    }
    *out = e;
    return BL_OK;
}

int bl_ensemble_destroy(bl_ensemble e)
{
    if (!e) return BL_OK;
    if (e->stream) hipStreamSynchronize(e->stream);
    if (e->graph) hipGraphExecDestroy(e->graph);
    ensemble_batch_free(e->batch);
    for (hipEvent_t ev : e->join)
        if (ev) hipEventDestroy(ev);
    if (e->fork) hipEventDestroy(e->fork);
    if (e->done) hipEventDestroy(e->done);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return BL_OK;
}

int bl_ensemble_set_batched(bl_ensemble e, int mode)
{
    if (!e || mode < -1 || mode > 1) return fail(BL_ERR_ARG, "bl_ensemble_set_batched: mode must be -1, 0 or 1");
    if (mode != e->mode && e->graph) {
        if (e->stream) hipStreamSynchronize(e->stream);
        hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
        ensemble_batch_free(e->batch);
        e->batch = nullptr;
    }
    e->mode = mode;
    return BL_OK;
}

int bl_ensemble_tune(bl_ensemble e, int n_steps, char* report, int cap)
{
    if (!e || n_steps < 2 || n_steps % 2) return fail(BL_ERR_ARG, "bl_ensemble_tune: n_steps must be even and >= 2");
    // the batched collide's workgroup order, then its nontemporal level, per
    // AA parity in turn (the other held), over real steps
    static const int orders[] = {0, 1, 4, 16, 64, 256};
    const int keep[4] = {e->order[0], e->order[1], e->nt[0], e->nt[1]};
    int best[4] = {keep[0], keep[1], keep[2], keep[3]};  // order even, odd; nt even, odd
    double best_ms = 1e30;
    int steps_run = 0, rc = BL_OK;
    std::string rep = "{\"trials\": [";
    bool first = true;
    auto set = [&](const int t[4]) -> int {
        const bool same = e->order[0] == t[0] && e->order[1] == t[1] && e->nt[0] == t[2] && e->nt[1] == t[3];
        e->order[0] = t[0];
        e->order[1] = t[1];
        e->nt[0] = t[2];
        e->nt[1] = t[3];
        if (e->graph && !same) {  // recapture with the new launch shape
            BL_HIP(hipStreamSynchronize(e->stream));
            hipGraphExecDestroy(e->graph);
            e->graph = nullptr;
        }
        return BL_OK;
    };
    auto trial = [&](const int t[4], double& ms) -> int {
        int r;
        if ((r = set(t))) return r;
        if ((r = bl_ensemble_steps(e, n_steps, 1))) return r;
        if ((r = bl_ensemble_time_steps(e, n_steps, 1, &ms))) return r;
        steps_run += 2 * n_steps;
        ms /= n_steps;
        return BL_OK;
    };
    auto run = [&](const int t[4]) -> int {
        double ms = 0;
        const int r = trial(t, ms);
        if (r) return r;
        char buf[160];
        snprintf(buf, sizeof(buf), "%s{\"even\": %d, \"odd\": %d, \"nt_even\": %d, \"nt_odd\": %d, "
                 "\"ms_per_step\": %.6f}", first ? "" : ", ", t[0], t[1], t[2], t[3], ms);
        rep += buf;
        first = false;
        if (ms < best_ms) {
            best_ms = ms;
            for (int k = 0; k < 4; ++k) best[k] = t[k];
        }
        return BL_OK;
    };
    for (int p = 0; p < 2 && rc == BL_OK; ++p)
        for (int o : orders) {
            int t[4] = {best[0], best[1], best[2], best[3]};
            t[p] = o;
            if ((rc = run(t))) break;
        }
    for (int p = 0; p < 2 && rc == BL_OK; ++p)
        for (int nt = 0; nt <= 2; ++nt) {
            int t[4] = {best[0], best[1], best[2], best[3]};
            if (t[2 + p] == nt) continue;  // timed already
            t[2 + p] = nt;
            if ((rc = run(t))) break;
        }
    if (rc) {
        set(keep);
        return rc;
    }
    if ((rc = set(best))) return rc;
    char buf[256];
    snprintf(buf, sizeof(buf), "], \"steps_run\": %d, \"batched\": %d, \"best\": {\"even\": %d, \"odd\": %d, "
             "\"nt_even\": %d, \"nt_odd\": %d, \"ms_per_step\": %.6f}}", steps_run, e->batch ? 1 : 0, best[0],
             best[1], best[2], best[3], best_ms);
    rep += buf;
    if (report && cap > 0) {
        const size_t n = std::min<size_t>(rep.size(), (size_t)cap - 1);
        rep.copy(report, n);
        report[n] = '\0';
    }
    return BL_OK;
}

int bl_ensemble_set_launch(bl_ensemble e, int parity, int order, int nontemporal)
{
    if (!e) return fail(BL_ERR_ARG, "null ensemble");
    if (parity != 0 && parity != 1) return fail(BL_ERR_ARG, "bl_ensemble_set_launch: parity must be 0 or 1");
    if (order < 0 || order > (1 << 20) || order == BL_ORDER_ROWS || (order > 1 && (order & (order - 1))))
        return fail(BL_ERR_ARG, "bl_ensemble_set_launch: order must be 0, 1 or a power of two up to 2^20");
    if (nontemporal < 0 || nontemporal > 2)
        return fail(BL_ERR_ARG, "bl_ensemble_set_launch: nontemporal must be 0, 1 or 2");
    if (e->order[parity] == order && e->nt[parity] == nontemporal) return BL_OK;
    if (e->graph) {  // the next steps capture with the new launch
        BL_HIP(hipStreamSynchronize(e->stream));
        hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    e->order[parity] = order;
    e->nt[parity] = nontemporal;
    return BL_OK;
}

int bl_ensemble_is_batched(bl_ensemble e, int* batched)
{
    if (!e || !batched) return fail(BL_ERR_ARG, "bl_ensemble_is_batched: null argument");
    *batched = e->graph && e->batch ? 1 : 0;
    return BL_OK;
}

int bl_ensemble_steps(bl_ensemble e, int n_steps, int use_graph)
{
    if (!e || n_steps < 0) return fail(BL_ERR_ARG, "bl_ensemble_steps: bad argument");
    int rc = run_steps(e, n_steps, use_graph);
    if (rc) return rc;
    for (bl_ibm g : e->ibms)
        if (g && (rc = ibm_check_overflow(g))) return rc;
    return BL_OK;
}

int bl_ensemble_time_steps(bl_ensemble e, int n_steps, int use_graph, double* ms)
{
    if (!e || n_steps <= 0 || !ms) return fail(BL_ERR_ARG, "bl_ensemble_time_steps: bad argument");
    int rc;
    hipEvent_t a, b;
    BL_HIP(hipEventCreate(&a));
    BL_HIP(hipEventCreate(&b));
    if ((rc = order_before(e))) return rc;
    BL_HIP(hipEventRecord(a, e->stream));
    if ((rc = order_after(e))) return rc;
    if ((rc = run_steps(e, n_steps, use_graph))) return rc;
    if ((rc = order_before(e))) return rc;
    BL_HIP(hipEventRecord(b, e->stream));
    BL_HIP(hipEventSynchronize(b));
    float t = 0.f;
    BL_HIP(hipEventElapsedTime(&t, a, b));
    *ms = (double)t;
    hipEventDestroy(a);
    hipEventDestroy(b);
    for (bl_ibm g : e->ibms)
        if (g && (rc = ibm_check_overflow(g))) return rc;
    return BL_OK;
}

}  // extern "C"
