// Virtual-grid executors: 2D/2.5D Cannon and SUMMA on one GPU.
//
// Every virtual rank becomes one problem of the grouped FP64 GEMM.  What the
// reference does with messages becomes segment pointers:
//   * Cannon skew + rotations (src/25DCannon.cpp:303-366): rank (i,j,k) at
//     step r multiplies op(A)_{i,m} op(B)_{m,j}, m = (i + j + off[k][r]) mod p;
//     the segment for step r points at those tiles where they already sit.
//   * Stack broadcast (src/25DCannon.cpp:174-194): all layers read the same
//     A/B tiles, so nothing is copied.
//   * Layer reduction (src/25DCannon.cpp:203-228): layers k > 0 write
//     partial C tiles into the workspace; one HBM pass adds them into layer 0.
//   * SUMMA panel broadcasts: panel t of A is read from process column t mod Q
//     and panel t of B from process row t mod P.
// All problems of one call run in a single launch.
#include <algorithm>

#include "dg_internal.hpp"

namespace dg {

namespace {

// Reference rotation offsets.  After the skew (target column
// (j - i + k*p/c) mod p, src/25DCannon.cpp:303-304) rank (i,j,k) holds
// A_{i,m} with m = i + j - k*p/c; rotation r sends to column j + r
// (src/25DCannon.cpp:333, 364), so after r rotations the cumulative shift is
// 1 + 2 + ... + r = r(r+1)/2.
int64_t ref_offset(int p, int c, int k, int r) { return -(int64_t)k * (p / c) - (int64_t)r * (r + 1) / 2; }
// Textbook Cannon: shift by one per step.
int64_t std_offset(int p, int c, int k, int r) { return -(int64_t)k * (p / c) - r; }

bool covers(int p, int c, int sched) {
    std::vector<int> seen(p, 0);
    for (int k = 0; k < c; ++k)
        for (int r = 0; r < p / c; ++r) {
            int64_t off = sched == DG_SCHED_REFERENCE ? ref_offset(p, c, k, r) : std_offset(p, c, k, r);
            int m = (int)(((off % p) + p) % p);
            if (seen[m]++) return false;
        }
    return true;
}

int resolve_schedule(int p, int c, int schedule, int* out) {
    if (schedule == DG_SCHED_AUTO) {
        *out = covers(p, c, DG_SCHED_REFERENCE) ? DG_SCHED_REFERENCE : DG_SCHED_STANDARD;
        return DG_OK;
    }
    if (schedule == DG_SCHED_REFERENCE) {
        if (!covers(p, c, DG_SCHED_REFERENCE))
            return fail(DG_ERR_SCHEDULE,
                        "reference rotation schedule does not visit every k tile for p=%d, c=%d "
                        "(it does for power-of-two p or p/c <= 2); use DG_SCHED_STANDARD",
                        p, c);
        *out = schedule;
        return DG_OK;
    }
    if (schedule == DG_SCHED_STANDARD) { *out = schedule; return DG_OK; }
    return fail(DG_ERR_INVALID, "unknown schedule %d", schedule);
}

int m_index(int p, int c, int sched, int i, int j, int k, int r) {
    int64_t off = sched == DG_SCHED_REFERENCE ? ref_offset(p, c, k, r) : std_offset(p, c, k, r);
    return (int)((((int64_t)i + j + off) % p + p) % p);
}

}  // namespace

}  // namespace dg

extern "C" {

int64_t dg_cannon25d_workspace(int64_t M, int64_t N, int p, int c) {
    if (M < 0 || N < 0 || p < 1 || c < 1) return 0;
    return (int64_t)(c - 1) * M * N;
}

int dg_cannon_m_index(int p, int c, int schedule, int i, int j, int k, int r) {
    using namespace dg;
    if (p < 1 || c < 1 || p % c != 0 || i < 0 || i >= p || j < 0 || j >= p || k < 0 || k >= c || r < 0 || r >= p / c)
        return -1;
    int s = 0;
    if (resolve_schedule(p, c, schedule, &s) != DG_OK) return -1;
    return m_index(p, c, s, i, j, k, r);
}

int dg_cannon25d(char opA, char opB, int64_t M, int64_t N, int64_t K, double alpha, const double* A_packed,
                 const double* B_packed, double beta, double* C_packed, int p, int c, int schedule, double* workspace,
                 int64_t workspace_elems, void* stream) {
    using namespace dg;
    const char oa = norm_op(opA), ob = norm_op(opB);
    if (!oa || !ob) return fail(DG_ERR_INVALID, "dg_cannon25d: invalid op '%c%c'", opA, opB);
    if (p < 1 || c < 1) return fail(DG_ERR_INVALID, "dg_cannon25d: grid %dx%dx%d", p, p, c);
    // The reference needs p*p*c ranks with pc3 = sqrt(size/c^3) integral (src/25DCannon.cpp:41).
    if (p % c != 0) return fail(DG_ERR_INVALID, "dg_cannon25d: c=%d must divide p=%d", c, p);
    if (M < 0 || N < 0 || K < 0) return fail(DG_ERR_INVALID, "dg_cannon25d: negative dimension");
    if (M % p || N % p || K % p)
        return fail(DG_ERR_INVALID, "dg_cannon25d: p=%d must divide M=%lld, N=%lld and K=%lld", p, (long long)M,
                    (long long)N, (long long)K);
    int sched = 0;
    int rc = resolve_schedule(p, c, schedule, &sched);
    if (rc) return rc;
    if (M == 0 || N == 0) return DG_OK;
    const int64_t need = dg_cannon25d_workspace(M, N, p, c);
    if (workspace_elems < need)
        return fail(DG_ERR_WORKSPACE, "dg_cannon25d: workspace %lld < %lld doubles", (long long)workspace_elems,
                    (long long)need);
    if (!A_packed || !B_packed || !C_packed || (need > 0 && !workspace))
        return fail(DG_ERR_INVALID, "dg_cannon25d: null pointer");
    rc = require_device();
    if (rc) return rc;
    hipStream_t st = as_stream(stream);

    const int64_t mt = M / p, nt = N / p, kt = K / p;
    const int steps = p / c;
    const int64_t tileA = mt * kt, tileB = kt * nt, tileC = mt * nt;
    if (K == 0 || alpha == 0.0) {
        // No products: layer 0 is beta*C, the other layers add nothing.
        return launch_scale(M * N, 1, beta, C_packed, M * N, st);
    }
    std::vector<Prob> probs;
    std::vector<Seg> segs;
    probs.reserve((size_t)p * p * c);
    segs.reserve((size_t)p * p * c * steps);
    for (int k = 0; k < c; ++k)
        for (int i = 0; i < p; ++i)
            for (int j = 0; j < p; ++j) {
                Prob pr{};
                pr.m = mt;
                pr.n = nt;
                const int64_t toff = ((int64_t)i * p + j) * tileC;
                pr.c = k == 0 ? C_packed + toff : workspace + (int64_t)(k - 1) * M * N + toff;
                pr.ldc = mt;
                pr.alpha = alpha;
                pr.beta = k == 0 ? beta : 0.0;  // src/25DCannon.cpp:322: only layer 0 carries beta*C
                pr.seg0 = (int32_t)segs.size();
                pr.nseg = steps;
                for (int r = 0; r < steps; ++r) {
                    const int m = m_index(p, c, sched, i, j, k, r);
                    Seg s{};
                    if (oa == 'N') { s.a = A_packed + ((int64_t)i * p + m) * tileA; s.lda = mt; }  // stored tile (i,m)
                    else           { s.a = A_packed + ((int64_t)m * p + i) * tileA; s.lda = kt; }  // stored tile (m,i)^T
                    if (ob == 'N') { s.b = B_packed + ((int64_t)m * p + j) * tileB; s.ldb = kt; }  // stored tile (m,j)
                    else           { s.b = B_packed + ((int64_t)j * p + m) * tileB; s.ldb = nt; }  // stored tile (j,m)^T
                    s.klen = kt;
                    segs.push_back(s);
                }
                probs.push_back(pr);
            }
    rc = launch_gemm_f64(oa, ob, probs, segs, st, DG_GEMM_MODE_XCD);
    if (rc) return rc;
    if (c > 1) return dg_layer_reduce(C_packed, workspace, M * N, c - 1, stream);
    return DG_OK;
}

int dg_summa(char opA, char opB, int64_t M, int64_t N, int64_t K, double alpha, const double* A_packed,
             const double* B_packed, double beta, double* C_packed, int64_t mb, int64_t nb, int64_t kb, int P, int Q,
             void* stream) {
    using namespace dg;
    const char oa = norm_op(opA), ob = norm_op(opB);
    if (!oa || !ob) return fail(DG_ERR_INVALID, "dg_summa: invalid op '%c%c'", opA, opB);
    if (M < 0 || N < 0 || K < 0) return fail(DG_ERR_INVALID, "dg_summa: negative dimension");
    if (mb <= 0 || nb <= 0 || kb <= 0) return fail(DG_ERR_INVALID, "dg_summa: block sizes must be positive");
    if (P < 1 || Q < 1 || P > 256 || Q > 256) return fail(DG_ERR_INVALID, "dg_summa: grid %dx%d", P, Q);
    if (M == 0 || N == 0) return DG_OK;
    if (!A_packed || !B_packed || !C_packed) return fail(DG_ERR_INVALID, "dg_summa: null pointer");
    int rc = require_device();
    if (rc) return rc;
    hipStream_t st = as_stream(stream);

    // Local extents and rank-order base offsets of the packed buffers: the
    // local matrix of rank pr*Q + pc starts at rowpre[pr]*cols + lrows[pr]*colpre[pc].
    auto extents = [](int64_t n, int64_t b, int np, std::vector<int64_t>& loc, std::vector<int64_t>& pre) {
        loc.resize(np);
        pre.resize(np + 1);
        int64_t acc = 0;
        for (int x = 0; x < np; ++x) { loc[x] = dg_numroc(n, b, x, 0, np); pre[x] = acc; acc += loc[x]; }
        pre[np] = acc;
    };
    // Stored operands: A is M x K (opA = N) or K x M (opA = T) with blocks
    // mb x kb or kb x mb; B is K x N or N x K with blocks kb x nb or nb x kb;
    // all distributed block-cyclically over the same P x Q grid.
    std::vector<int64_t> rA, rApre, cA, cApre, rB, rBpre, cB, cBpre, rC, rCpre, cC, cCpre;
    if (oa == 'N') { extents(M, mb, P, rA, rApre); extents(K, kb, Q, cA, cApre); }
    else           { extents(K, kb, P, rA, rApre); extents(M, mb, Q, cA, cApre); }
    if (ob == 'N') { extents(K, kb, P, rB, rBpre); extents(N, nb, Q, cB, cBpre); }
    else           { extents(N, nb, P, rB, rBpre); extents(K, kb, Q, cB, cBpre); }
    extents(M, mb, P, rC, rCpre);
    extents(N, nb, Q, cC, cCpre);
    const int64_t colsA = oa == 'N' ? K : M, colsB = ob == 'N' ? N : K;
    auto baseA = [&](int pr, int pc) { return rApre[pr] * colsA + rA[pr] * cApre[pc]; };
    auto baseB = [&](int pr, int pc) { return rBpre[pr] * colsB + rB[pr] * cBpre[pc]; };
    auto baseC = [&](int pr, int pc) { return rCpre[pr] * N + rC[pr] * cCpre[pc]; };

    if (K == 0 || alpha == 0.0) {
        for (int pr = 0; pr < P; ++pr)
            for (int pc = 0; pc < Q; ++pc) {
                rc = launch_scale(rC[pr], cC[pc], beta, C_packed + baseC(pr, pc), rC[pr] > 0 ? rC[pr] : 1, st);
                if (rc) return rc;
            }
        return DG_OK;
    }
    // One problem per rank's local C.  A transposed operand's panel is not one
    // block of one owner (the global row block ib of op(A) is the stored
    // column block ib, owned by process column ib % Q), so with opA = T the
    // local C splits into its row blocks, and with opB = T into its column
    // blocks; each piece then reads one block of one owner per panel.
    // Every element still sums the panels in order 0, 1, ...
    const int64_t npanels = (K + kb - 1) / kb;
    std::vector<Prob> probs;
    std::vector<Seg> segs;
    for (int pr = 0; pr < P; ++pr)
        for (int pc = 0; pc < Q; ++pc) {
            if (rC[pr] == 0 || cC[pc] == 0) continue;
            const int64_t nmb = oa == 'N' ? 1 : (rC[pr] + mb - 1) / mb;  // row pieces
            const int64_t nnb = ob == 'N' ? 1 : (cC[pc] + nb - 1) / nb;  // column pieces
            for (int64_t jm = 0; jm < nmb; ++jm)
                for (int64_t jn = 0; jn < nnb; ++jn) {
                    const int64_t r0 = oa == 'N' ? 0 : jm * mb, c0 = ob == 'N' ? 0 : jn * nb;
                    const int64_t rows = oa == 'N' ? rC[pr] : std::min<int64_t>(mb, rC[pr] - r0);
                    const int64_t cols = ob == 'N' ? cC[pc] : std::min<int64_t>(nb, cC[pc] - c0);
                    const int64_t ib = (int64_t)pr + jm * P;   // global row block of op(A) (opA = T)
                    const int64_t jb = (int64_t)pc + jn * Q;   // global column block of op(B) (opB = T)
                    Prob p{};
                    p.m = rows;
                    p.n = cols;
                    p.c = C_packed + baseC(pr, pc) + r0 + c0 * rC[pr];
                    p.ldc = rC[pr];
                    p.alpha = alpha;
                    p.beta = beta;
                    p.seg0 = (int32_t)segs.size();
                    p.nseg = (int32_t)npanels;
                    for (int64_t t = 0; t < npanels; ++t) {
                        Seg s{};
                        if (oa == 'N') {
                            // local rows of rank (pr, t % Q), its local column block t / Q
                            const int qa = (int)(t % Q);
                            s.a = A_packed + baseA(pr, qa) + (t / Q) * kb * rA[pr];
                            s.lda = rA[pr];
                        } else {
                            // stored row block t (owner row t % P), column block ib (owner column ib % Q)
                            const int pa = (int)(t % P), qa = (int)(ib % Q);
                            s.a = A_packed + baseA(pa, qa) + (t / P) * kb + (ib / Q) * mb * rA[pa];
                            s.lda = rA[pa];
                        }
                        if (ob == 'N') {
                            const int pb = (int)(t % P);
                            s.b = B_packed + baseB(pb, pc) + (t / P) * kb;
                            s.ldb = rB[pb];
                        } else {
                            // stored row block jb (owner row jb % P), column block t (owner column t % Q)
                            const int pb = (int)(jb % P), qb = (int)(t % Q);
                            s.b = B_packed + baseB(pb, qb) + (jb / P) * nb + (t / Q) * kb * rB[pb];
                            s.ldb = rB[pb];
                        }
                        s.klen = (K - t * kb) < kb ? (K - t * kb) : kb;
                        segs.push_back(s);
                    }
                    probs.push_back(p);
                }
        }
    if (probs.size() > (size_t)INT32_MAX || segs.size() > (size_t)INT32_MAX)
        return fail(DG_ERR_UNSUPPORTED, "dg_summa: too many blocks");
    return launch_gemm_f64(oa, ob, probs, segs, st, DG_GEMM_MODE_XCD);
}

}  // extern "C"
