// Grouped batches of FP64 GEMMs: one launch for many uneven problems, plans
// that keep their descriptor tables resident, HIP-graph replay and an
// autotuner that picks the tile scheduling per shape class and, in the Ozaki
// form, the arithmetic (FP64 MFMA or INT8 CRT emulation) per problem.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "dg_internal.hpp"

struct dg_grouped_plan {
    struct Group {
        char opA, opB;
        dg::GemmBatchDev b;         // device tables owned by the plan
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
    // Ozaki form (DG_FORM_OZAKI_INT8): per-problem choice made on first use.
    int form = DG_FORM_FP64;
    int moduli_req = 0;                 // 0: per problem, FP64-level count
    bool forms_done = true;
    std::vector<dg_gemm_desc> descs;    // the batch, kept for the deferred build
    std::vector<int> prob_form;         // DG_FORM_* per problem (-1 before the choice)
    std::vector<int> prob_moduli;       // moduli per Ozaki problem (0 for FP64)
    std::vector<int> oz_idx;            // problems run in the Ozaki form, in order
    void* oz_ws = nullptr;
    int64_t oz_ws_bytes = 0;
    // Operands shared by Ozaki problems (a tile row of A read by every tile of
    // that row, ...): each distinct line set is prepared once per launch into
    // oz_shared at `off`, then every product reads it.  Empty when no
    // operand repeats or the buffers do not fit; each problem then runs
    // dg_ozaki_dgemm in oz_ws.
    struct OzOperand {
        const double* x;
        int64_t rows, k, rs, ks;
        int nmod;
        int64_t off;
    };
    std::vector<OzOperand> oz_ops;
    std::vector<int> oz_a, oz_b;   // per entry of oz_idx: its op(A) rows and op(B) columns in oz_ops
    void* oz_shared = nullptr;
    int64_t oz_scan_off = 0, oz_digit_off = 0;
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

int launch_ozaki_problem(const dg_grouped_plan* p, int i, hipStream_t st) {
    const dg_gemm_desc& d = p->descs[i];
    return dg_ozaki_dgemm(d.opA, d.opB, d.m, d.n, d.k, d.alpha, d.A, d.lda, d.B, d.ldb, d.beta, d.C, d.ldc,
                          p->prob_moduli[i], DG_OZAKI_INT8_CRT, p->oz_ws, p->oz_ws_bytes, nullptr, st);
}

int plan_launch(dg_grouped_plan* p, int variant, int cfg, hipStream_t st, bool with_ozaki = true) {
    for (auto& g : p->groups) {
        int rc = launch_gemm_f64_dev(g.opA, g.opB, variant, cfg, g.b, p->counter, st);
        if (rc) return rc;
    }
    if (!with_ozaki) return DG_OK;
    if (p->oz_shared) {
        char* base = static_cast<char*>(p->oz_shared);
        for (const auto& o : p->oz_ops) {
            int rc = crt_prepare(o.x, o.rows, o.k, o.rs, o.ks, o.nmod, base + o.off, base + p->oz_scan_off, st);
            if (rc) return rc;
        }
        for (size_t t = 0; t < p->oz_idx.size(); ++t) {
            const int i = p->oz_idx[t];
            const dg_gemm_desc& d = p->descs[i];
            int rc = crt_multiply(d.m, d.n, d.k, p->prob_moduli[i], d.alpha, base + p->oz_ops[p->oz_a[t]].off,
                                  base + p->oz_ops[p->oz_b[t]].off, d.beta, d.C, d.ldc, base + p->oz_digit_off, st);
            if (rc) return rc;
        }
        return DG_OK;
    }
    for (int i : p->oz_idx) {
        int rc = launch_ozaki_problem(p, i, st);
        if (rc) return rc;
    }
    return DG_OK;
}

// Finds the operands that several Ozaki problems read (same lines: pointer,
// line count, k, strides, moduli) and, when any repeats, no emulated
// problem's C overlaps an operand and the prepared copies fit in half of the
// free device memory, allocates one buffer for all of them plus one scan and
// one digit buffer; otherwise leaves the per-problem path in place.  The products are bit for bit those of
// dg_ozaki_dgemm: a prepared operand depends only on its own lines.
int share_operands(dg_grouped_plan* p) {
    p->oz_ops.clear();
    p->oz_a.clear();
    p->oz_b.clear();
    std::map<std::tuple<uintptr_t, int64_t, int64_t, int64_t, int64_t, int>, int> index;
    int64_t off = 0, scan = 0, digits = 0;
    auto add = [&](const double* x, int64_t rows, int64_t k, int64_t rs, int64_t ks, int nmod) {
        const auto key = std::make_tuple(reinterpret_cast<uintptr_t>(x), rows, k, rs, ks, nmod);
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        const int id = (int)p->oz_ops.size();
        p->oz_ops.push_back({x, rows, k, rs, ks, nmod, off});
        off += (crt_operand_bytes(rows, k, nmod) + 255) / 256 * 256;
        scan = std::max(scan, crt_scan_bytes(rows, k));
        index[key] = id;
        return id;
    };
    for (int i : p->oz_idx) {
        const dg_gemm_desc& d = p->descs[i];
        const char oa = norm_op(d.opA), ob = norm_op(d.opB);
        const int nm = p->prob_moduli[i];
        // op(A)(i, k) at A[i*rs + k*ks]; op(B) as the rows of its transpose
        p->oz_a.push_back(add(d.A, d.m, d.k, oa == 'N' ? 1 : d.lda, oa == 'N' ? d.lda : 1, nm));
        p->oz_b.push_back(add(d.B, d.n, d.k, ob == 'N' ? d.ldb : 1, ob == 'N' ? 1 : d.ldb, nm));
        digits = std::max(digits, crt_digit_bytes(d.m, d.n, nm));
    }
    // Preparing every operand before the first product is only the order of
    // the per-problem path when no emulated problem writes memory that an
    // operand reads (a batch whose outputs feed later problems keeps that path).
    bool disjoint = true;
    for (int i : p->oz_idx) {
        const dg_gemm_desc& d = p->descs[i];
        const uintptr_t c0 = reinterpret_cast<uintptr_t>(d.C);
        const uintptr_t c1 = c0 + 8 * (uintptr_t)((d.n - 1) * d.ldc + d.m);
        for (const auto& o : p->oz_ops) {
            const uintptr_t x0 = reinterpret_cast<uintptr_t>(o.x);
            const uintptr_t x1 = x0 + 8 * (uintptr_t)((o.rows - 1) * o.rs + (o.k - 1) * o.ks + 1);
            if (c0 < x1 && x0 < c1) disjoint = false;
        }
    }
    size_t free_b = 0, total_b = 0;
    const int64_t bytes = off + (scan + 255) / 256 * 256 + digits;
    const bool repeats = p->oz_ops.size() < 2 * p->oz_idx.size();
    if (!repeats || !disjoint || hipMemGetInfo(&free_b, &total_b) != hipSuccess || bytes > (int64_t)(free_b / 2) ||
        hipMalloc(&p->oz_shared, (size_t)bytes) != hipSuccess) {
        (void)hipGetLastError();  // a refused allocation is not an error here
        p->oz_shared = nullptr;
        p->oz_ops.clear();
        p->oz_a.clear();
        p->oz_b.clear();
        return DG_OK;
    }
    p->oz_scan_off = off;
    p->oz_digit_off = off + (scan + 255) / 256 * 256;
    if (p->oz_ws) {  // the per-problem workspace served the form trials only
        (void)hipFree(p->oz_ws);
        p->oz_ws = nullptr;
        p->oz_ws_bytes = 0;
    }
    return DG_OK;
}

// Device tables of the FP64 problems listed in `use` (all when empty).
int build_groups(dg_grouped_plan* p, const dg_gemm_desc* descs, int count, const std::vector<char>* use) {
    std::vector<dg_gemm_desc> sel;
    sel.reserve(count);
    for (int i = 0; i < count; ++i)
        if (!use || (*use)[i]) sel.push_back(descs[i]);
    for (auto& kv : by_ops(sel.data(), (int)sel.size())) {
        std::vector<Prob> probs0;
        std::vector<Seg> segs;
        build_group(sel.data(), kv.second, probs0, segs);
        GemmBatch hb;
        int rc = prepare_gemm(kv.first.first, kv.first.second, probs0, segs, hb);
        if (rc) return rc;
        if (hb.dma.empty() && hb.dmae.empty() && hb.gen.empty() && hb.nar[0].empty() && hb.nar[1].empty()) continue;
        dg_grouped_plan::Group g;
        g.opA = kv.first.first;
        g.opB = kv.first.second;
        p->groups.push_back(g);
        GemmBatchDev& b = p->groups.back().b;  // owned by the plan from here on (freed on failure too)
        auto upload = [&](const auto& v, auto** dst) -> bool {
            using T = typename std::decay_t<decltype(v)>::value_type;
            *dst = nullptr;
            if (v.empty()) return true;
            void* d = nullptr;
            if (hipMalloc(&d, v.size() * sizeof(T)) != hipSuccess) return false;
            *dst = static_cast<const T*>(d);
            return hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
        };
        if (!upload(hb.dma, &b.dma) || !upload(hb.dmae, &b.dmae) || !upload(hb.gen, &b.gen) ||
            !upload(hb.nar[0], &b.nar[0]) || !upload(hb.nar[1], &b.nar[1]) || !upload(hb.segs, &b.segs))
            return fail(DG_ERR_ALLOC, "plan: table upload");
        for (int w = 0; w < 2; ++w) {
            b.nnar[w] = (int)hb.nar[w].size();
            b.tiles_nar[w] = hb.tiles_nar[w];
            b.vec_nar[w] = hb.vec_nar[w];
        }
        b.ndma = (int)hb.dma.size();
        b.tiles_dma = hb.tiles_dma;
        b.ndmae = (int)hb.dmae.size();
        b.tiles_dmae = hb.tiles_dmae;
        b.ngen = (int)hb.gen.size();
        b.tiles_gen[0] = hb.tiles_gen[0];
        b.tiles_gen[1] = hb.tiles_gen[1];
        b.vec = hb.vec;
    }
    return DG_OK;
}

// Output bytes of problem i (its C footprint).
size_t c_bytes(const dg_gemm_desc& d) { return sizeof(double) * (size_t)(d.ldc * (d.n - 1) + d.m); }

// Timing trials write the outputs; this scope saves them first and, on every
// exit path (a failed launch included), writes them back, waits, frees the
// snapshot and destroys the two timing events.  `saved` is set only once the
// whole snapshot is enqueued, so a failure while saving restores nothing.
struct TrialScope {
    hipStream_t st = nullptr;
    std::vector<std::pair<double*, size_t>> spans;
    char* snap = nullptr;
    bool saved = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    int save(hipStream_t stream, std::vector<std::pair<double*, size_t>> outs) {
        st = stream;
        spans = std::move(outs);
        size_t bytes = 0;
        for (auto& sp : spans) bytes += sp.second;
        if (bytes) {
            if (hipMalloc(&snap, bytes) != hipSuccess) {
                snap = nullptr;
                return fail(DG_ERR_ALLOC, "autotune: snapshot of %zu bytes", bytes);
            }
            size_t off = 0;
            for (auto& sp : spans) {
                DG_HIP(hipMemcpyAsync(snap + off, sp.first, sp.second, hipMemcpyDeviceToDevice, st));
                off += sp.second;
            }
            saved = true;
        }
        DG_HIP(hipEventCreate(&e0));
        DG_HIP(hipEventCreate(&e1));
        return DG_OK;
    }
    ~TrialScope() {
        if (saved) {
            size_t off = 0;
            for (auto& sp : spans) {
                (void)hipMemcpyAsync(sp.first, snap + off, sp.second, hipMemcpyDeviceToDevice, st);
                off += sp.second;
            }
        }
        if (snap) {
            (void)hipStreamSynchronize(st);
            (void)hipFree(snap);
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// Times one problem alone (best of three): FP64 MFMA, or Ozaki with the
// plan's workspace.
int time_problem(const dg_grouped_plan* p, int i, bool ozaki, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                 float* out) {
    const dg_gemm_desc& d = p->descs[i];
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        DG_HIP(hipEventRecord(e0, st));
        int rc = ozaki ? launch_ozaki_problem(p, i, st)
                       : dg_dgemm(d.opA, d.opB, d.m, d.n, d.k, d.alpha, d.A, d.lda, d.B, d.ldb, d.beta, d.C, d.ldc, st);
        if (rc) return rc;
        DG_HIP(hipEventRecord(e1, st));
        DG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        DG_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    *out = best;
    return DG_OK;
}

// Ozaki form: picks each problem's form, allocates the shared workspace and
// uploads the tables of the FP64 problems.  Outputs are restored afterwards.
int choose_forms(dg_grouped_plan* p, hipStream_t st) {
    const int count = (int)p->descs.size();
    p->prob_form.assign(count, DG_FORM_FP64);
    p->prob_moduli.assign(count, 0);
    std::vector<int> cand;
    int64_t ws = 0;
    for (int i = 0; i < count; ++i) {
        const dg_gemm_desc& d = p->descs[i];
        if (d.m <= 0 || d.n <= 0 || d.k <= 0 || d.alpha == 0.0 || d.k > 131071) continue;
        const double flops = 2.0 * (double)d.m * (double)d.n * (double)d.k;
        if (flops < 4294967296.0) continue;
        const int nm = p->moduli_req > 0 ? p->moduli_req : crt_moduli_for_fp64(d.k);
        p->prob_moduli[i] = nm;
        cand.push_back(i);
        ws = std::max<int64_t>(ws, dg_ozaki_workspace(d.m, d.n, d.k, nm, DG_OZAKI_INT8_CRT));
    }
    if (!cand.empty()) {
        if (hipMalloc(&p->oz_ws, (size_t)ws) != hipSuccess) return fail(DG_ERR_ALLOC, "plan: Ozaki workspace of %lld bytes", (long long)ws);
        p->oz_ws_bytes = ws;
        TuneCache& c = tune_cache();
        for (int i : cand) {
            const dg_gemm_desc& d = p->descs[i];
            char key[160];
            snprintf(key, sizeof(key), "form:%c%c.%lld.%lld.%lld.%d", norm_op(d.opA), norm_op(d.opB), (long long)d.m,
                     (long long)d.n, (long long)d.k, p->prob_moduli[i]);
            int choice = -1;
            {
                std::lock_guard<std::mutex> lock(c.mu);
                tune_load_locked(c);
                auto it = c.best.find(key);
                if (it != c.best.end()) choice = it->second;
            }
            if (choice < 0) {
                // time both forms on this problem's own operands; C is restored
                // when `trial` ends, whether or not the trials succeed
                float t64 = 0.f, toz = 0.f;
                {
                    TrialScope trial;
                    int rc = trial.save(st, {{d.C, c_bytes(d)}});
                    if (!rc) rc = time_problem(p, i, false, st, trial.e0, trial.e1, &t64);
                    if (!rc) rc = time_problem(p, i, true, st, trial.e0, trial.e1, &toz);
                    if (rc) return rc;
                }
                choice = toz < t64 ? DG_FORM_OZAKI_INT8 : DG_FORM_FP64;
                std::lock_guard<std::mutex> lock(c.mu);
                tune_store_locked(c, key, choice);
            }
            p->prob_form[i] = choice;
        }
    }
    std::vector<char> use(count, 1);
    p->oz_idx.clear();
    for (int i = 0; i < count; ++i)
        if (p->prob_form[i] == DG_FORM_OZAKI_INT8) {
            use[i] = 0;
            p->oz_idx.push_back(i);
        } else {
            p->prob_moduli[i] = 0;
        }
    if (p->oz_idx.empty() && p->oz_ws) {
        (void)hipFree(p->oz_ws);
        p->oz_ws = nullptr;
        p->oz_ws_bytes = 0;
    }
    if (!p->oz_idx.empty()) {
        int rc = share_operands(p);
        if (rc) return rc;
    }
    int rc = build_groups(p, p->descs.data(), count, &use);
    if (rc) return rc;
    p->forms_done = true;
    return DG_OK;
}

// Times each scheduling variant (best of three) on the plan's own operands
// and keeps the fastest.  The outputs are saved before the trials and
// restored after them, also when a trial fails, so tuning leaves C as it
// found it.
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
    float best_ms = 1e30f;
    int best = DG_GEMM_MODE_XCD;  // encoded cfg * 4 + scheduling mode
    {
        // the outputs are restored when `trial` ends, whether or not the trials succeed
        TrialScope trial;
        int rc = trial.save(st, p->c_spans);
        if (rc) return rc;
        for (int cfg : cfgs)
            for (int v : variants) {
                float vbest = 1e30f;
                for (int rep = 0; rep < 3; ++rep) {
                    DG_HIP(hipEventRecord(trial.e0, st));
                    rc = plan_launch(p, v, cfg, st, false);
                    if (rc) return rc;
                    DG_HIP(hipEventRecord(trial.e1, st));
                    DG_HIP(hipEventSynchronize(trial.e1));
                    float ms = 0.f;
                    DG_HIP(hipEventElapsedTime(&ms, trial.e0, trial.e1));
                    vbest = std::min(vbest, ms);
                }
                if (vbest < best_ms) { best_ms = vbest; best = cfg * 4 + v; }
            }
    }
    p->variant = best % 4;
    p->cfg = best / 4;
    std::lock_guard<std::mutex> lock(c.mu);
    tune_store_locked(c, p->shape_key, best);
    return DG_OK;
}

// ------------------------------------------------------------ no allocation
// The allocation-free form of the one-shot grouped GEMM keeps its descriptor
// tables in this file-scope device array instead of the table cache.  Every
// GPU holds its own copy until the process ends; its address on a device
// comes from hipGetSymbolAddress, taken once per device with that device
// current, and is the only handle host code uses for it.  Calls on one GPU
// share the array: calls that overlap on one GPU race.
constexpr size_t kScratchBytes = 256 * 1024;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_device(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(DG_ERR_NO_DEVICE, "scratch: device %d", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(DG_ERR_HIP, "scratch: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return DG_OK;
}

// Tables reach the scratch array through kernel arguments (copied at launch,
// and kept by value in a captured graph), in chunks of kStoreWords 16-byte
// words: no host buffer has to outlive the call.
constexpr int kStoreWords = 224;
struct StoreChunk {
    uint4 w[kStoreWords];
    int32_t n;
};

__global__ void store_chunk_kernel(uint4* __restrict__ dst, StoreChunk c) {
    const int t = threadIdx.x;
    if (t < c.n) dst[t] = c.w[t];
}

int store_bytes(void* dst, const std::vector<unsigned char>& bytes, hipStream_t st) {
    const size_t words = (bytes.size() + 15) / 16;
    for (size_t w0 = 0; w0 < words; w0 += kStoreWords) {
        StoreChunk c{};
        c.n = (int32_t)std::min<size_t>(kStoreWords, words - w0);
        const size_t b0 = w0 * 16, nb = std::min<size_t>((size_t)c.n * 16, bytes.size() - b0);
        memcpy(c.w, bytes.data() + b0, nb);
        hipLaunchKernelGGL(store_chunk_kernel, dim3(1), dim3(256), 0, st, static_cast<uint4*>(dst) + w0, c);
        DG_LAUNCH_CHECK();
    }
    return DG_OK;
}

}  // namespace

}  // namespace dg

extern "C" {

int dg_grouped_dgemm_noalloc(const dg_gemm_desc* descs, int count, void* stream) {
    using namespace dg;
    int rc = validate_descs(descs, count);
    if (rc) return rc;
    if (count == 0) return DG_OK;
    rc = require_device();
    if (rc) return rc;
    int dev = 0;
    DG_HIP(hipGetDevice(&dev));
    void* base = nullptr;
    rc = scratch_for_device(dev, &base);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    // Lay out every op group's tables (problems routed to each kernel, then
    // segments) in one host image, 256-byte aligned pieces, and check the size
    // before anything is launched.
    struct Piece { char opA, opB; GemmBatch hb; std::vector<Seg> segs; size_t off_dma, off_dmae, off_gen, off_seg; };
    std::vector<Piece> pieces;
    size_t total = 0;
    auto take = [&](size_t bytes) { const size_t o = total; total = (total + bytes + 255) / 256 * 256; return o; };
    for (auto& kv : by_ops(descs, count)) {
        std::vector<Prob> probs;
        std::vector<Seg> segs;
        build_group(descs, kv.second, probs, segs);
        if (probs.empty()) continue;
        Piece pc{kv.first.first, kv.first.second, GemmBatch{}, std::move(segs), 0, 0, 0, 0};
        // no sub-problems here (so no narrow strips either): the tables stay
        // within the documented capacity
        rc = prepare_gemm(pc.opA, pc.opB, probs, pc.segs, pc.hb, false);
        if (rc) return rc;
        pc.off_dma = take(pc.hb.dma.size() * sizeof(Prob));
        pc.off_dmae = take(pc.hb.dmae.size() * sizeof(Prob));
        pc.off_gen = take(pc.hb.gen.size() * sizeof(Prob));
        pc.off_seg = take(pc.hb.segs.size() * sizeof(Seg));
        pieces.push_back(std::move(pc));
    }
    if (total > kScratchBytes)
        return fail(DG_ERR_WORKSPACE,
                    "dg_grouped_dgemm_noalloc: the batch's descriptor tables need %zu bytes, more than the %zu-byte "
                    "scratch array (about %zu problems of one segment each)",
                    total, kScratchBytes, kScratchBytes / (2 * sizeof(Prob) + sizeof(Seg)));
    rc = scale_only(descs, count, st);
    if (rc) return rc;
    if (pieces.empty()) return DG_OK;
    std::vector<unsigned char> image(total, 0);
    for (const Piece& pc : pieces) {
        if (!pc.hb.dma.empty()) memcpy(image.data() + pc.off_dma, pc.hb.dma.data(), pc.hb.dma.size() * sizeof(Prob));
        if (!pc.hb.dmae.empty())
            memcpy(image.data() + pc.off_dmae, pc.hb.dmae.data(), pc.hb.dmae.size() * sizeof(Prob));
        if (!pc.hb.gen.empty()) memcpy(image.data() + pc.off_gen, pc.hb.gen.data(), pc.hb.gen.size() * sizeof(Prob));
        memcpy(image.data() + pc.off_seg, pc.hb.segs.data(), pc.hb.segs.size() * sizeof(Seg));
    }
    rc = store_bytes(base, image, st);
    if (rc) return rc;
    unsigned char* b = static_cast<unsigned char*>(base);
    for (const Piece& pc : pieces) {
        GemmBatchDev d;
        d.dma = pc.hb.dma.empty() ? nullptr : reinterpret_cast<const Prob*>(b + pc.off_dma);
        d.ndma = (int)pc.hb.dma.size();
        d.tiles_dma = pc.hb.tiles_dma;
        d.dmae = pc.hb.dmae.empty() ? nullptr : reinterpret_cast<const Prob*>(b + pc.off_dmae);
        d.ndmae = (int)pc.hb.dmae.size();
        d.tiles_dmae = pc.hb.tiles_dmae;
        d.gen = pc.hb.gen.empty() ? nullptr : reinterpret_cast<const Prob*>(b + pc.off_gen);
        d.ngen = (int)pc.hb.gen.size();
        d.tiles_gen[0] = pc.hb.tiles_gen[0];
        d.tiles_gen[1] = pc.hb.tiles_gen[1];
        d.vec = pc.hb.vec;
        d.segs = reinterpret_cast<const Seg*>(b + pc.off_seg);
        rc = launch_gemm_f64_dev(pc.opA, pc.opB, DG_GEMM_MODE_ROUND_ROBIN, default_tile_cfg(), d, nullptr, st);
        if (rc) return rc;
    }
    return DG_OK;
}

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
    return dg_grouped_plan_create_form(descs, count, mode, DG_FORM_FP64, 0, out);
}

int dg_grouped_plan_create_form(const dg_gemm_desc* descs, int count, int mode, int form, int moduli,
                                dg_grouped_plan** out) {
    using namespace dg;
    if (!out) return fail(DG_ERR_INVALID, "dg_grouped_plan_create: null output");
    *out = nullptr;
    if (mode != DG_PLAN_AUTOTUNE && mode != DG_PLAN_NONPERSISTENT && mode != DG_PLAN_PERSISTENT)
        return fail(DG_ERR_INVALID, "dg_grouped_plan_create: mode %d", mode);
    if (form != DG_FORM_FP64 && form != DG_FORM_OZAKI_INT8)
        return fail(DG_ERR_INVALID, "dg_grouped_plan_create: form %d", form);
    if (form == DG_FORM_OZAKI_INT8 && moduli != 0 && (moduli < 2 || moduli > 16))
        return fail(DG_ERR_INVALID, "dg_grouped_plan_create: moduli %d (0 or 2..16)", moduli);
    int rc = validate_descs(descs, count);
    if (rc) return rc;
    for (int i = 0; i < count; ++i)
        if (descs[i].m > 0 && descs[i].n > 0 && (descs[i].k == 0 || descs[i].alpha == 0.0))
            return fail(DG_ERR_UNSUPPORTED, "dg_grouped_plan_create: problem %d has no product (k == 0 or alpha == 0)", i);
    rc = require_device();
    if (rc) return rc;
    dg_grouped_plan* p = new dg_grouped_plan();
    p->mode = mode;
    p->form = form;
    p->moduli_req = moduli;
    (void)hipGetDevice(&p->device);
    p->shape_key = shape_class(descs, count);
    if (form == DG_FORM_OZAKI_INT8) p->shape_key = "oz" + std::to_string(moduli) + ":" + p->shape_key;
    if (mode == DG_PLAN_NONPERSISTENT) p->variant = DG_GEMM_MODE_ROUND_ROBIN;
    if (mode == DG_PLAN_PERSISTENT) p->variant = DG_GEMM_MODE_PERSISTENT;
    p->cfg = default_tile_cfg();
    auto cleanup = [&](int code) { dg_grouped_plan_destroy(p); return code; };
    if (hipMalloc(&p->counter, sizeof(int)) != hipSuccess) return cleanup(fail(DG_ERR_ALLOC, "plan: counter alloc"));
    for (int i = 0; i < count; ++i)
        if (descs[i].m > 0 && descs[i].n > 0) p->c_spans.push_back({descs[i].C, c_bytes(descs[i])});
    if (form == DG_FORM_OZAKI_INT8) {
        // the per-problem forms (and with them the FP64 tables) are chosen on first use
        p->descs.assign(descs, descs + count);
        p->forms_done = false;
    } else {
        rc = build_groups(p, descs, count, nullptr);
        if (rc) return cleanup(rc);
        p->prob_form.assign(count, DG_FORM_FP64);
        p->prob_moduli.assign(count, 0);
    }
    *out = p;
    return DG_OK;
}

int dg_grouped_plan_execute(dg_grouped_plan* p, void* stream) {
    using namespace dg;
    if (!p) return fail(DG_ERR_INVALID, "dg_grouped_plan_execute: null plan");
    hipStream_t st = as_stream(stream);
    if (!p->forms_done) {
        int rc = choose_forms(p, st);
        if (rc) return rc;
    }
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
    if (!p->forms_done) {
        int rc = choose_forms(p, st);
        if (rc) return rc;
    }
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

int dg_grouped_plan_problem_form(const dg_grouped_plan* p, int index, int* moduli) {
    if (moduli) *moduli = 0;
    if (!p || !p->forms_done || index < 0 || index >= (int)p->prob_form.size()) return -1;
    if (moduli) *moduli = p->prob_moduli[index];
    return p->prob_form[index];
}

int dg_grouped_plan_destroy(dg_grouped_plan* p) {
    if (!p) return DG_OK;
    if (p->oz_ws) (void)hipFree(p->oz_ws);
    if (p->oz_shared) (void)hipFree(p->oz_shared);
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    if (p->capture_stream) (void)hipStreamDestroy(p->capture_stream);
    for (auto& g : p->groups) {
        if (g.b.dma) (void)hipFree(const_cast<dg::Prob*>(g.b.dma));
        if (g.b.dmae) (void)hipFree(const_cast<dg::Prob*>(g.b.dmae));
        if (g.b.gen) (void)hipFree(const_cast<dg::Prob*>(g.b.gen));
        for (const dg::Prob* t : g.b.nar)
            if (t) (void)hipFree(const_cast<dg::Prob*>(t));
        if (g.b.segs) (void)hipFree(const_cast<dg::Seg*>(g.b.segs));
    }
    if (p->counter) (void)hipFree(p->counter);
    delete p;
    return DG_OK;
}

}  // extern "C"
