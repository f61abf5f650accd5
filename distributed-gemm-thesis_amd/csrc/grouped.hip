// Grouped batches of FP64 GEMMs: one launch for many uneven problems, plans
// that keep their descriptor tables resident, HIP-graph replay and an
// autotuner that picks the tile scheduling per shape class.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>

#include "dg_internal.hpp"

struct dg_grouped_plan {
    struct Group {
        char opA, opB;
        bool vec;
        int tiles[2], np;
        dg::Prob* dprobs = nullptr;
        dg::Seg* dsegs = nullptr;
    };
    std::vector<Group> groups;
    int mode = DG_PLAN_AUTOTUNE;   // requested
    int variant = -1;              // DG_GEMM_MODE_* in use (-1: not tuned yet)
    int cfg = 0;                   // DG_GEMM_TILE_* in use
    std::string shape_key;
    int* counter = nullptr;        // device work counter (persistent mode)
    hipGraphExec_t graph = nullptr;
    int graph_variant = -1;
    hipStream_t capture_stream = nullptr;
    int device = 0;
    std::vector<std::pair<double*, size_t>> c_spans;  // bytes of every output, for the tuning snapshot
};

namespace dg {

namespace {

// Shape class of a batch: the op pairs and, per problem, log2 buckets of
// m, n, k, sorted so that permutations of one batch share a class.
std::string shape_class(const dg_gemm_desc* d, int count) {
    std::vector<std::string> keys;
    keys.reserve(count);
    auto bucket = [](int64_t x) { return x <= 0 ? 0 : (int)std::ceil(std::log2((double)x) * 2.0); };
    for (int i = 0; i < count; ++i) {
        char buf[64];
        snprintf(buf, sizeof(buf), "%c%c%d.%d.%d", norm_op(d[i].opA), norm_op(d[i].opB), bucket(d[i].m),
                 bucket(d[i].n), bucket(d[i].k));
        keys.push_back(buf);
    }
    std::sort(keys.begin(), keys.end());
    std::string k;
    for (auto& s : keys) { k += s; k += ','; }
    return k;
}

struct TuneCache {
    std::mutex mu;
    std::map<std::string, int> best;
    bool loaded = false;
};

TuneCache& tune_cache() {
    static TuneCache* c = new TuneCache();
    return *c;
}

// Optional persistence: DG_AUTOTUNE_CACHE names a text file of "key variant" lines.
void tune_load_locked(TuneCache& c) {
    if (c.loaded) return;
    c.loaded = true;
    const char* path = getenv("DG_AUTOTUNE_CACHE");
    if (!path) return;
    FILE* f = fopen(path, "r");
    if (!f) return;
    char key[8192];
    int v;
    while (fscanf(f, "%8191s %d", key, &v) == 2) c.best[key] = v;
    fclose(f);
}

void tune_store_locked(TuneCache& c, const std::string& key, int v) {
    c.best[key] = v;
    const char* path = getenv("DG_AUTOTUNE_CACHE");
    if (!path) return;
    FILE* f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s %d\n", key.c_str(), v);
    fclose(f);
}

int validate_descs(const dg_gemm_desc* d, int count) {
    if (count < 0 || (count > 0 && !d)) return fail(DG_ERR_INVALID, "grouped: bad descriptor array");
    for (int i = 0; i < count; ++i) {
        const char oa = norm_op(d[i].opA), ob = norm_op(d[i].opB);
        if (!oa || !ob) return fail(DG_ERR_INVALID, "grouped: problem %d: invalid op", i);
        if (d[i].m < 0 || d[i].n < 0 || d[i].k < 0) return fail(DG_ERR_INVALID, "grouped: problem %d: negative size", i);
        const int64_t ra = oa == 'N' ? d[i].m : d[i].k, rb = ob == 'N' ? d[i].k : d[i].n;
        if (d[i].lda < (ra > 1 ? ra : 1) || d[i].ldb < (rb > 1 ? rb : 1) || d[i].ldc < (d[i].m > 1 ? d[i].m : 1))
            return fail(DG_ERR_INVALID, "grouped: problem %d: leading dimension too small", i);
        if (d[i].m > 0 && d[i].n > 0 && (!d[i].C || (d[i].k > 0 && (!d[i].A || !d[i].B))))
            return fail(DG_ERR_INVALID, "grouped: problem %d: null pointer", i);
    }
    return DG_OK;
}

// Problems of one op pair, largest k first (their tiles are dealt first).
void build_group(const dg_gemm_desc* d, const std::vector<int>& idx, std::vector<Prob>& probs, std::vector<Seg>& segs) {
    std::vector<int> order = idx;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a].k > d[b].k; });
    for (int i : order) {
        const dg_gemm_desc& g = d[i];
        if (g.m == 0 || g.n == 0 || g.k == 0) continue;
        Prob p{};
        p.m = g.m; p.n = g.n; p.c = g.C; p.ldc = g.ldc; p.alpha = g.alpha; p.beta = g.beta;
        p.seg0 = (int32_t)segs.size(); p.nseg = 1;
        segs.push_back(Seg{g.A, g.B, g.lda, g.ldb, g.k});
        probs.push_back(p);
    }
}

// Problems with k == 0 or alpha == 0 only scale C.
int scale_only(const dg_gemm_desc* d, int count, hipStream_t st) {
    for (int i = 0; i < count; ++i)
        if (d[i].m > 0 && d[i].n > 0 && (d[i].k == 0 || d[i].alpha == 0.0)) {
            int rc = launch_scale(d[i].m, d[i].n, d[i].beta, d[i].C, d[i].ldc, st);
            if (rc) return rc;
        }
    return DG_OK;
}

std::map<std::pair<char, char>, std::vector<int>> by_ops(const dg_gemm_desc* d, int count) {
    std::map<std::pair<char, char>, std::vector<int>> m;
    for (int i = 0; i < count; ++i)
        if (d[i].alpha != 0.0) m[{norm_op(d[i].opA), norm_op(d[i].opB)}].push_back(i);
    return m;
}

int plan_launch(dg_grouped_plan* p, int variant, int cfg, hipStream_t st) {
    for (auto& g : p->groups) {
        int rc = launch_gemm_f64_dev(g.opA, g.opB, g.vec, variant, cfg, g.dprobs, g.np, g.dsegs, g.tiles, p->counter, st);
        if (rc) return rc;
    }
    return DG_OK;
}

// Times each scheduling variant (best of three) on the plan's own operands
// and keeps the fastest.  The outputs are saved before the trials and
// restored after them, so tuning leaves C as it found it.
int autotune(dg_grouped_plan* p, hipStream_t st) {
    TuneCache& c = tune_cache();
    {
        std::lock_guard<std::mutex> lock(c.mu);
        tune_load_locked(c);
        auto it = c.best.find(p->shape_key);
        if (it != c.best.end()) { p->variant = it->second % 4; p->cfg = it->second / 4; return DG_OK; }
    }
    const int variants[3] = {DG_GEMM_MODE_XCD, DG_GEMM_MODE_ROUND_ROBIN, DG_GEMM_MODE_PERSISTENT};
    const int cfgs[2] = {DG_GEMM_TILE_128x128, DG_GEMM_TILE_256x128};
    size_t snap_bytes = 0;
    for (auto& sp : p->c_spans) snap_bytes += sp.second;
    char* snap = nullptr;
    if (snap_bytes) {
        if (hipMalloc(&snap, snap_bytes) != hipSuccess) return fail(DG_ERR_ALLOC, "autotune: snapshot of %zu bytes", snap_bytes);
        size_t off = 0;
        for (auto& sp : p->c_spans) {
            DG_HIP(hipMemcpyAsync(snap + off, sp.first, sp.second, hipMemcpyDeviceToDevice, st));
            off += sp.second;
        }
    }
    hipEvent_t e0, e1;
    DG_HIP(hipEventCreate(&e0));
    DG_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    int best = DG_GEMM_MODE_XCD;  // encoded cfg * 4 + scheduling mode
    for (int cfg : cfgs)
        for (int v : variants) {
            float vbest = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {
                DG_HIP(hipEventRecord(e0, st));
                int rc = plan_launch(p, v, cfg, st);
                if (rc) return rc;
                DG_HIP(hipEventRecord(e1, st));
                DG_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                DG_HIP(hipEventElapsedTime(&ms, e0, e1));
                vbest = std::min(vbest, ms);
            }
            if (vbest < best_ms) { best_ms = vbest; best = cfg * 4 + v; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (snap) {
        size_t off = 0;
        for (auto& sp : p->c_spans) {
            DG_HIP(hipMemcpyAsync(sp.first, snap + off, sp.second, hipMemcpyDeviceToDevice, st));
            off += sp.second;
        }
        DG_HIP(hipStreamSynchronize(st));
        (void)hipFree(snap);
    }
    p->variant = best % 4;
    p->cfg = best / 4;
    std::lock_guard<std::mutex> lock(c.mu);
    tune_store_locked(c, p->shape_key, best);
    return DG_OK;
}

}  // namespace

}  // namespace dg

extern "C" {

int dg_grouped_dgemm(const dg_gemm_desc* descs, int count, void* stream) {
    using namespace dg;
    int rc = validate_descs(descs, count);
    if (rc) return rc;
    if (count == 0) return DG_OK;
    rc = require_device();
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    rc = scale_only(descs, count, st);
    if (rc) return rc;
    for (auto& kv : by_ops(descs, count)) {
        std::vector<Prob> probs;
        std::vector<Seg> segs;
        build_group(descs, kv.second, probs, segs);
        if (probs.empty()) continue;
        rc = launch_gemm_f64(kv.first.first, kv.first.second, probs, segs, st, DG_GEMM_MODE_ROUND_ROBIN);
        if (rc) return rc;
    }
    return DG_OK;
}

int dg_grouped_plan_create(const dg_gemm_desc* descs, int count, int mode, dg_grouped_plan** out) {
    using namespace dg;
    if (!out) return fail(DG_ERR_INVALID, "dg_grouped_plan_create: null output");
    *out = nullptr;
    if (mode != DG_PLAN_AUTOTUNE && mode != DG_PLAN_NONPERSISTENT && mode != DG_PLAN_PERSISTENT)
        return fail(DG_ERR_INVALID, "dg_grouped_plan_create: mode %d", mode);
    int rc = validate_descs(descs, count);
    if (rc) return rc;
    for (int i = 0; i < count; ++i)
        if (descs[i].m > 0 && descs[i].n > 0 && (descs[i].k == 0 || descs[i].alpha == 0.0))
            return fail(DG_ERR_UNSUPPORTED, "dg_grouped_plan_create: problem %d has no product (k == 0 or alpha == 0)", i);
    rc = require_device();
    if (rc) return rc;
    dg_grouped_plan* p = new dg_grouped_plan();
    p->mode = mode;
    (void)hipGetDevice(&p->device);
    p->shape_key = shape_class(descs, count);
    if (mode == DG_PLAN_NONPERSISTENT) p->variant = DG_GEMM_MODE_ROUND_ROBIN;
    if (mode == DG_PLAN_PERSISTENT) p->variant = DG_GEMM_MODE_PERSISTENT;
    p->cfg = default_tile_cfg();
    auto cleanup = [&](int code) { dg_grouped_plan_destroy(p); return code; };
    if (hipMalloc(&p->counter, sizeof(int)) != hipSuccess) return cleanup(fail(DG_ERR_ALLOC, "plan: counter alloc"));
    for (int i = 0; i < count; ++i)
        if (descs[i].m > 0 && descs[i].n > 0)
            p->c_spans.push_back({descs[i].C, sizeof(double) * (size_t)(descs[i].ldc * (descs[i].n - 1) + descs[i].m)});
    for (auto& kv : by_ops(descs, count)) {
        std::vector<Prob> probs0, probs;
        std::vector<Seg> segs;
        build_group(descs, kv.second, probs0, segs);
        bool vec = true;
        int tiles[2] = {0, 0};
        rc = prepare_gemm(probs0, segs, probs, &vec, tiles);
        if (rc) return cleanup(rc);
        if (probs.empty()) continue;
        dg_grouped_plan::Group g;
        g.opA = kv.first.first;
        g.opB = kv.first.second;
        g.vec = vec;
        g.tiles[0] = tiles[0];
        g.tiles[1] = tiles[1];
        g.np = (int)probs.size();
        if (hipMalloc(&g.dprobs, probs.size() * sizeof(Prob)) != hipSuccess ||
            hipMalloc(&g.dsegs, segs.size() * sizeof(Seg)) != hipSuccess) {
            p->groups.push_back(g);
            return cleanup(fail(DG_ERR_ALLOC, "plan: table alloc"));
        }
        p->groups.push_back(g);
        if (hipMemcpy(g.dprobs, probs.data(), probs.size() * sizeof(Prob), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(g.dsegs, segs.data(), segs.size() * sizeof(Seg), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(DG_ERR_HIP, "plan: table upload"));
    }
    *out = p;
    return DG_OK;
}

int dg_grouped_plan_execute(dg_grouped_plan* p, void* stream) {
    using namespace dg;
    if (!p) return fail(DG_ERR_INVALID, "dg_grouped_plan_execute: null plan");
    hipStream_t st = as_stream(stream);
    if (p->variant < 0) {
        int rc = autotune(p, st);
        if (rc) return rc;
    }
    return plan_launch(p, p->variant, p->cfg, st);
}

int dg_grouped_plan_graph_launch(dg_grouped_plan* p, void* stream) {
    using namespace dg;
    if (!p) return fail(DG_ERR_INVALID, "dg_grouped_plan_graph_launch: null plan");
    hipStream_t st = as_stream(stream);
    if (p->variant < 0) {
        int rc = autotune(p, st);
        if (rc) return rc;
    }
    if (!p->graph || p->graph_variant != p->variant * 4 + p->cfg) {
        if (p->graph) { (void)hipGraphExecDestroy(p->graph); p->graph = nullptr; }
        // Capture on a private stream (the legacy default stream cannot be
        // captured); the graph is then launched on the caller's stream.
        if (!p->capture_stream) DG_HIP(hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking));
        DG_HIP(hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal));
        int rc = plan_launch(p, p->variant, p->cfg, p->capture_stream);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(p->capture_stream, &graph);
        if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        if (e != hipSuccess) return fail(DG_ERR_HIP, "plan capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&p->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) { p->graph = nullptr; return fail(DG_ERR_HIP, "plan instantiate: %s", hipGetErrorString(e)); }
        p->graph_variant = p->variant * 4 + p->cfg;
    }
    DG_HIP(hipGraphLaunch(p->graph, st));
    return DG_OK;
}

int dg_grouped_plan_variant(const dg_grouped_plan* p) {
    if (!p) return -1;
    if (p->variant == dg::DG_GEMM_MODE_PERSISTENT) return DG_PLAN_PERSISTENT;
    if (p->variant < 0) return DG_PLAN_AUTOTUNE;
    return DG_PLAN_NONPERSISTENT;
}

int dg_grouped_plan_destroy(dg_grouped_plan* p) {
    if (!p) return DG_OK;
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    if (p->capture_stream) (void)hipStreamDestroy(p->capture_stream);
    for (auto& g : p->groups) {
        if (g.dprobs) (void)hipFree(g.dprobs);
        if (g.dsegs) (void)hipFree(g.dsegs);
    }
    if (p->counter) (void)hipFree(p->counter);
    delete p;
    return DG_OK;
}

}  // extern "C"
