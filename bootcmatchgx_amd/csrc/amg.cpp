// amg.cpp -- BCMG hierarchy build, preconditioner application and the FCG
// driver, all on one HIP stream of the context.
//
// Hierarchy: adaptiveCoarsening / matchingAggregation
// (src/preconditioner/bcmg/matchingAggregation.cu:106-466): per level,
// aggrsweeps+1 pairwise steps, each suitor matching -> P_i -> A_{i+1} =
// R_i (A_i P_i) and w_{i+1} = R_i w_i, composing P <- P P_i.
// Cycle: GAMG_cycle (GAMG_cycle.cu:58-236) with l1-Jacobi smoothing; the
// first sweep from a zero guess is the diagonal scaling f/M (bit-identical to
// the reference's full sweep on zero), and the ping-pong buffer holding a
// level's result is tracked instead of copied back.
// FCG: flexibileConjugateGradients_v3 (src/solver/fcg/FCG.cu:27-175); one
// iteration pair is captured in a hipGraph and replayed, the host polls the
// device stop flag every `check_every` iterations.
#include "amg.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <type_traits>

namespace bcmgx {

// ------------------------------------------------------------ autotune ----
int tune_lanes(Context* c, const DCsr& A, bool fp32) {
    if (A.n < 65536) return default_lanes(A.nnz, A.n);
    if (fp32 && !A.valf) throw Error(BCMGX_ERR_STATE, "tune_lanes: FP32 values not built");
    std::string sig = std::to_string(A.n) + ":" + std::to_string(A.nnz) + (fp32 ? "f" : "d") + (A.tiled() ? "t" : "");
    auto it = c->tuned_lanes.find(sig);
    if (it != c->tuned_lanes.end()) return it->second;
    int base = default_lanes(A.nnz, A.n);
    // vector kernels (L lanes per row), CSR-stream kernels (-TPR lanes per row
    // reduction), their paired-load form (-200 - TPR), and both with rows dealt
    // to the XCDs in super-blocks of 32 blocks (-6000 - ...); with a row-tile
    // index the tiled kernel (-404), also in super-blocks of 16 and 64 (-5404, -7404)
    std::vector<int> cands = {std::max(1, base / 2), base, std::min(64, base * 2), -2, -4, -8, -204, -208, -6004, -6204};
    if (A.tiled()) cands.insert(cands.end(), {-404, -5404, -7404});
    hipEvent_t e0, e1;
    BCMGX_HIP(hipEventCreate(&e0));
    BCMGX_HIP(hipEventCreate(&e1));
    int best = base;
    float bestms = 1e30f;
    if (fp32) {
        DBuf<float> x(c, A.n), y(c, A.n);
        fill<float>(c, x.get(), A.n, 1.0f);
        for (int L : cands) {
            SpmvArgs a;
            a.n = A.n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.valf.get(); a.nnz = A.nnz;
            set_tiles(a, A);
            a.x = x.get(); a.y = y.get(); a.lanes = L;
            spmv<float, float>(c, SpOp::AX, a);
            BCMGX_HIP(hipEventRecord(e0, c->stream));
            for (int r = 0; r < 5; r++) spmv<float, float>(c, SpOp::AX, a);
            BCMGX_HIP(hipEventRecord(e1, c->stream));
            BCMGX_HIP(hipEventSynchronize(e1));
            float ms;
            BCMGX_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < bestms) { bestms = ms; best = L; }
        }
    } else {
        DBuf<double> x(c, A.n), y(c, A.n);
        fill<double>(c, x.get(), A.n, 1.0);
        for (int L : cands) {
            SpmvArgs a;
            a.n = A.n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
            set_tiles(a, A);
            a.x = x.get(); a.y = y.get(); a.lanes = L;
            spmv<double, double>(c, SpOp::AX, a);
            BCMGX_HIP(hipEventRecord(e0, c->stream));
            for (int r = 0; r < 5; r++) spmv<double, double>(c, SpOp::AX, a);
            BCMGX_HIP(hipEventRecord(e1, c->stream));
            BCMGX_HIP(hipEventSynchronize(e1));
            float ms;
            BCMGX_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < bestms) { bestms = ms; best = L; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->tuned_lanes[sig] = best;
    return best;
}

int cached_lanes(Context* c, const DCsr& A, bool fp32) {
    if (A.n >= 65536) {
        auto it = c->tuned_lanes.find(std::to_string(A.n) + ":" + std::to_string(A.nnz) + (fp32 ? "f" : "d") +
                                      (A.tiled() ? "t" : ""));
        if (it != c->tuned_lanes.end()) return it->second;
    }
    return default_lanes(A.nnz, A.n);
}

template <typename V>
static float time_spmm(Context* c, const DCsr& A, const void* val, int k, int chunk, hipEvent_t e0, hipEvent_t e1) {
    DBuf<V> x(c, (size_t)A.m * k), y(c, (size_t)A.n * k);
    fill<V>(c, x.get(), (long long)A.m * k, (V)1);
    SpmvArgs a;
    a.n = A.n; a.row = A.row.get(); a.col = A.col.get(); a.val = val; a.nnz = A.nnz;
    a.x = x.get(); a.y = y.get(); a.k = k; a.kchunk = chunk;
    spmv<V, V>(c, SpOp::AX, a);
    BCMGX_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < 3; r++) spmv<V, V>(c, SpOp::AX, a);
    BCMGX_HIP(hipEventRecord(e1, c->stream));
    BCMGX_HIP(hipEventSynchronize(e1));
    float ms;
    BCMGX_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms;
}

int tune_chunk(Context* c, const DCsr& A, bool fp32, int k) {
    if (k <= 4 || A.n < 65536) return k;
    if (fp32 && !A.valf) throw Error(BCMGX_ERR_STATE, "tune_chunk: FP32 values not built");
    std::string sig = std::to_string(A.n) + ":" + std::to_string(A.nnz) + (fp32 ? "f" : "d") + ":k" + std::to_string(k);
    auto it = c->tuned_lanes.find(sig);
    if (it != c->tuned_lanes.end()) return it->second;
    hipEvent_t e0, e1;
    BCMGX_HIP(hipEventCreate(&e0));
    BCMGX_HIP(hipEventCreate(&e1));
    int best = k;
    float bestms = 1e30f;
    // full width through the pipelined (k), one-pass (-1) or pipelined kernel with
    // streamed matrix loads (-2); the pipelined kernel with rows dealt to the
    // XCDs in super-groups of 512, 2048 or 8192 rows (-16 - log2 of the
    // super-group in row groups); then narrower column chunks
    std::vector<int> cands = {k, -1, -2};
    int cpl = fp32 ? 4 : 2;
    if (k % cpl == 0 && k >= 2 * cpl && k <= 64) {
        int KP = 2;
        while (KP < k / cpl) KP <<= 1;
        int rpb = 256 / KP;  // rows per row group of the pipelined kernel
        for (int rows : {512, 2048, 8192}) {
            int lg = 0;
            while ((rpb << (lg + 1)) <= rows) lg++;
            cands.push_back(-16 - lg);
        }
    }
    for (int ch = (k + 1) / 2; ch >= 4; ch = (ch + 1) / 2) {
        cands.push_back(ch);
        if (ch == 4) break;
    }
    for (int ch : cands) {
        float ms = fp32 ? time_spmm<float>(c, A, A.valf.get(), k, ch, e0, e1)
                        : time_spmm<double>(c, A, A.val.get(), k, ch, e0, e1);
        if (ms < bestms) { bestms = ms; best = ch; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->tuned_lanes[sig] = best;
    return best;
}

// ----------------------------------------------------------- hierarchy ----
// setMaxCoarseSize (AMG.cu:238-241), truncation included
static int reference_max_coarse_size(long long n) { return (int)(40 * std::pow((double)n, (double)1. / 3.)); }

// Setup work queued on the context's second stream, behind everything queued
// on the first so far, so that it overlaps the latency-bound matching of the
// next levels.  Blocks the queued work frees are held back by the pool until
// join() has made the first stream wait for it; the destructor waits for the
// second stream when a setup ends early (an error between run() and join()).
struct SideQueue {
    Context* c;
    bool pending = false;
    explicit SideQueue(Context* c_) : c(c_) {}
    SideQueue(const SideQueue&) = delete;
    SideQueue& operator=(const SideQueue&) = delete;
    ~SideQueue() {
        if (pending) {
            (void)hipStreamSynchronize(c->side);
            c->pool->release_deferred();
        }
    }
    static void order(hipStream_t before, hipStream_t after) {
        hipEvent_t e;
        BCMGX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        BCMGX_HIP(hipEventRecord(e, before));
        hipError_t r = hipStreamWaitEvent(after, e, 0);
        (void)hipEventDestroy(e);
        BCMGX_HIP(r);
    }
    template <typename F>
    void run(F&& f) {
        hipStream_t side = c->side_stream(), main = c->stream;
        order(main, side);
        pending = true;
        c->stream = side;
        c->pool->set_defer(true);
        try {
            f();
        } catch (...) {
            c->stream = main;
            c->pool->set_defer(false);
            throw;
        }
        c->stream = main;
        c->pool->set_defer(false);
    }
    void join() {
        if (!pending) return;
        order(c->side, c->stream);
        c->pool->release_deferred();
        pending = false;
    }
};

// A's diagonal positions, every row checked to have one
static DBuf<int> check_diagonals(Context* c, const DCsr& A) {
    DBuf<int> dind(c, std::max(A.n, 1));
    diag_index(c, A.n, A.row.get(), A.col.get(), dind.get());
    if (count_missing_diagonals(c, A.n, dind.get()) == 0) return dind;
    std::vector<int> h(A.n);  // name the first offending row
    c->d2h(h.data(), dind.get(), A.n);
    for (int i = 0; i < A.n; i++)
        BCMGX_CHECK(h[i] >= 0, BCMGX_ERR_ARG, "row " + std::to_string(i) + " has no diagonal entry (required by the matching)");
    return dind;
}

// w0: the smooth vector the matching weights start from (nullptr: w = 1, as
// BootBuildData::initByParams fills it, AMG.cu:383-384).  Once level l is final
// (its matrix and aggregates) its l1 row sums M and restriction R are computed and then
// level_done(l) is called, both on the side queue when one is given (the
// levels below are still being built meanwhile) or else in order.
static void build_hierarchy(Context* c, const DCsr& A0, const bcmgx_params& p, Hierarchy& H, const double* w0 = nullptr,
                            SideQueue* side = nullptr, const std::function<void(int)>& level_done = {}) {
    BCMGX_CHECK(A0.n == A0.m, BCMGX_ERR_ARG, "the matrix must be square");
    DBuf<int> dind0 = check_diagonals(c, A0);
    int maxcs = p.max_coarse_size > 0 ? p.max_coarse_size : reference_max_coarse_size(A0.n);
    H.maxcs = maxcs;
    int sweeps = p.aggrsweeps + 1;
    H.L.clear();
    H.L.emplace_back(new Level());
    H.L[0]->Aref = &A0;
    auto final_level = [&](int l) {
        auto work = [&] {
            Level& lv = *H.L[l];
            const DCsr& M = lv.mat();
            lv.M.reset(c, std::max(M.n, 1));
            abs_row_sum(c, M.n, M.row.get(), M.val.get(), lv.M.get());
            if (l + 1 < H.levels())  // the restriction of the composed prolongator
                transpose_prolongator(c, M.n, H.L[l + 1]->mat().n, lv.agg.get(), lv.pval.get(), lv.R);
            if (level_done) level_done(l);
        };
        if (side) side->run(work);
        else work();
    };

    DBuf<double> w(c, A0.n);
    if (w0) copy_d2d(c, w.get(), w0, sizeof(double) * (size_t)A0.n);
    else fill<double>(c, w.get(), A0.n, 1.0);
    int ftcoarse = 1;
    double avc = 0.0;
    // the diagonal positions of the matrix the next sweep matches (written by the
    // Galerkin product that made it, when its one-launch path ran)
    DBuf<int> dind_next;
    for (int level = 1; level < p.max_levels;) {
        Level& prev = *H.L[level - 1];
        const DCsr* Ai = &prev.mat();
        DCsr Aown;  // intermediate sweep operator
        DBuf<double> wi = std::move(w);
        DBuf<int> agg;
        DBuf<double> pv;
        DCsr Afinal;
        for (int s = 0; s < sweeps; s++) {
            int n = Ai->n;
            DBuf<int> M(c, n), agg_i(c, n);
            DBuf<double> pval_i(c, n);
            // A Galerkin product keeps structural entries, so every row of R (A P) holds a
            // diagonal when every row of A does (aggregate c's row gets a_ii of its members):
            // only the fine matrix needs the check
            DBuf<int> dind_i = std::move(dind_next);
            if (Ai != &A0 && !dind_i) {
                dind_i.reset(c, std::max(n, 1));
                diag_index(c, n, Ai->row.get(), Ai->col.get(), dind_i.get());
            }
            suitor_match(c, *Ai, wi.get(), M.get(), Ai == &A0 ? dind0.get() : dind_i.get());
            int npairs = 0;
            int nc = make_prolongator(c, n, M.get(), wi.get(), agg_i.get(), pval_i.get(), &npairs);
            BCMGX_CHECK(nc > 0, BCMGX_ERR_NUMERIC, "coarsening produced an empty level");
            // every vertex in a good pair or a singleton of its own: aggregates of at most 2
            // (a zero-weight singleton joins the aggregate before it, makeP's :41-44)
            bool pairs_only = (long long)n == (long long)npairs + nc;
            DCsr Ri, Anew;
            if (pairs_only) pair_restriction(c, n, nc, npairs, agg_i.get(), M.get(), pval_i.get(), Ri);
            else transpose_prolongator(c, n, nc, agg_i.get(), pval_i.get(), Ri);
            DBuf<int> dind_c(c, std::max(nc, 1));
            if (galerkin(c, *Ai, nc, agg_i.get(), pval_i.get(), Ri, Anew, pairs_only ? 2 : -1, dind_c.get()))
                dind_next = std::move(dind_c);
            DBuf<double> wn(c, nc);
            {
                SpmvArgs a;  // w_{i+1} = R_i w_i, one lane per row (sequential sums)
                a.n = nc; a.row = Ri.row.get(); a.col = Ri.col.get(); a.val = Ri.val.get(); a.nnz = Ri.nnz;
                a.x = wi.get(); a.y = wn.get(); a.lanes = 1;
                spmv<double, double>(c, SpOp::AX, a);
            }
            double size_pre = n, size_c = nc;
            if (size_pre / size_c <= 1.2) ftcoarse = 100;  // COARSERATIO_THRSLD, FTCOARSE_INC
            if (s == 0) {
                agg = std::move(agg_i);
                pv = std::move(pval_i);
            } else {
                compose_prolongators(c, (int)agg.size(), agg.get(), pv.get(), agg_i.get(), pval_i.get());
            }
            Aown = std::move(Anew);
            Ai = &Aown;
            wi = std::move(wn);
            if (size_c <= (double)ftcoarse * maxcs) break;
        }
        w = std::move(wi);
        std::unique_ptr<Level> nl(new Level());
        nl->A = std::move(Aown);
        int nc = nl->A.n;
        prev.agg = std::move(agg);
        prev.pval = std::move(pv);
        avc += (double)prev.mat().n / (double)nc;
        H.L.push_back(std::move(nl));
        level++;
        final_level(level - 2);
        if ((double)nc <= (double)ftcoarse * maxcs) break;
    }
    int nlv = H.levels();
    final_level(nlv - 1);
    double nnz0 = (double)H.L[0]->mat().nnz, cx = 0, cw = 0;
    for (int i = 0; i < nlv; i++) {
        cx += H.L[i]->mat().nnz;
        cw += std::pow(2, i) * H.L[i]->mat().nnz;
    }
    H.op_cmplx = cx / nnz0;
    H.op_wcmplx = cw / nnz0;
    H.avg_cratio = nlv > 1 ? avc / (nlv - 1) : 1;
}

// A_l P_l of every level above the coarsest, values in FP32 only: with them the
// first post-smoothing sweep forms f - A (x + P e) as r_pre - (A P) e
static void make_level_product(Context* c, Hierarchy& H, int l) {
    Level& lv = *H.L[l];
    const DCsr& A = lv.mat();
    DCsr P;
    prolongator_csr(c, A.n, H.L[l + 1]->mat().n, lv.agg.get(), lv.pval.get(), P);
    spgemm(c, A, P, lv.AP);
    lv.AP.valf.reset(c, std::max(lv.AP.nnz, 1));
    convert<double, float>(c, lv.AP.val.get(), lv.AP.valf.get(), lv.AP.nnz);
    lv.AP.val.release();
    lv.lanesAP = default_lanes(lv.AP.nnz, lv.AP.n);
}
// The same product queued with no host synchronisation (false when A_l's rows
// are too long for it); level_product_lengths reads the lengths afterwards
static bool level_product_async(Context* c, Hierarchy& H, int l) {
    Level& lv = *H.L[l];
    return unit_product_f32_async(c, lv.mat(), H.L[l + 1]->mat().n, lv.agg.get(), lv.pval.get(), lv.AP);
}
static void level_product_lengths(Context* c, const std::vector<Level*>& lvs) {
    if (lvs.empty()) return;
    int* h = (int*)c->host_scratch(lvs.size() * sizeof(int));
    for (size_t t = 0; t < lvs.size(); t++)
        BCMGX_HIP(hipMemcpyAsync(h + t, lvs[t]->AP.row.get() + lvs[t]->AP.n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    c->sync();  // one round trip for every level
    for (size_t t = 0; t < lvs.size(); t++) {
        Level& lv = *lvs[t];
        lv.AP.nnz = h[t];
        lv.lanesAP = default_lanes(lv.AP.nnz, lv.AP.n);
    }
}

static void level_fp32_copy(Context* c, Level& l) {
    const DCsr& A = l.mat();
    // the FP32 values live beside the FP64 ones; level 0's (the caller's
    // matrix) are converted again by every setup, as the coarse levels' are
    DCsr& Am = const_cast<DCsr&>(A);
    if (!Am.valf) Am.valf.reset(c, std::max(A.nnz, 1));
    convert<double, float>(c, A.val.get(), Am.valf.get(), A.nnz);
    l.Mf.reset(c, std::max(A.n, 1));
    convert<double, float>(c, l.M.get(), l.Mf.get(), A.n);
    if (l.pval) {
        l.pvalf.reset(c, l.pval.size());
        convert<double, float>(c, l.pval.get(), l.pvalf.get(), (long long)l.pval.size());
        l.R.valf.reset(c, std::max(l.R.nnz, 1));
        convert<double, float>(c, l.R.val.get(), l.R.valf.get(), l.R.nnz);
    }
}


// ------------------------------------------------------------ prec ----
struct SolveWork {
    int k = 0, n = 0;
    DBuf<double> B, X, r, v, w, d, q;
    DBuf<double> partials;
    DBuf<FcgState> st;
    DBuf<int> flags;  // [0] all_done, [1] columns done
    FcgBufs fb;
    FcgEntry entry;     // level-0 entry of the FP32 cycle filled by the updates
    bool fused = false;
    // FP32 z: the cycle's result kept in FP32 and widened inside the fused product and the update
    bool z32 = false;
    DBuf<float> zf;
    int z32_blocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t cg_graph = nullptr;  // one CG iteration
    hipGraphExec_t cg_exec = nullptr;
    int lanesA = 0;
    int chunkA = 0;
    int mv_blocks = 0;  // grid of the fused SpMM + dots kernel (0: not used)
    // the stop rule baked into the captured graphs (the finishing kernels take
    // rtol and itnlim by value): a solve with another rule captures again
    double cap_rtol = -1;
    int cap_itnlim = -1;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cg_exec) (void)hipGraphExecDestroy(cg_exec);
        if (cg_graph) (void)hipGraphDestroy(cg_graph);
        exec = nullptr;
        graph = nullptr;
        cg_exec = nullptr;
        cg_graph = nullptr;
    }
    ~SolveWork() { drop_graph(); }
};

Prec::~Prec() {}

void Prec::drop_graphs() {
    for (auto& s : solves) s->drop_graph();
}

void Prec::ensure_work(int k) {
    if (ptype == BCMGX_PREC_BCMG && k > 1 && k != tuned_k) {
        // SpMM column chunk for this batch width (tuned outside any graph capture)
        bool f32 = p.precision == BCMGX_PRECISION_FP32;
        for (int h = 0; h < n_hrc(); h++)
            for (auto& lp : hrc(h).L) lp->chunkA = p.autotune ? tune_chunk(ctx, lp->mat(), f32, k) : k;
        tuned_k = k;
    }
    if (k <= kcap) return;
    kcap = k;
    // the cycle buffers are about to be reallocated: graphs captured by earlier
    // solves hold their addresses, so they are dropped and captured again
    drop_graphs();
    if (ptype == BCMGX_PREC_BCMG) {
        // the hierarchies of a bootstrap run one after another and share the
        // cycle buffers: level l is sized for the largest level l among them
        int nl = 0, ncoarse = 1;
        for (int h = 0; h < n_hrc(); h++) {
            nl = std::max(nl, hrc(h).levels());
            ncoarse = std::max(ncoarse, hrc(h).L.back()->mat().n);
        }
        std::vector<size_t> lsz(nl, 1);
        for (int h = 0; h < n_hrc(); h++)
            for (int l = 0; l < hrc(h).levels(); l++) lsz[l] = std::max(lsz[l], (size_t)hrc(h).L[l]->mat().n);
        if (p.precision == BCMGX_PRECISION_FP32) {
            wf.clear();
            wf.resize(nl);
            for (int l = 0; l < nl; l++) {
                size_t sz = lsz[l] * k;
                wf[l].F.reset(ctx, sz); wf[l].X.reset(ctx, sz); wf[l].X2.reset(ctx, sz); wf[l].Res.reset(ctx, sz);
            }
            size_t cs = (size_t)ncoarse * k;
            cd_F.reset(ctx, cs); cd_X.reset(ctx, cs); cd_X2.reset(ctx, cs);
        } else {
            wd.clear();
            wd.resize(nl);
            for (int l = 0; l < nl; l++) {
                size_t sz = lsz[l] * k;
                if (l > 0) wd[l].F.reset(ctx, sz);
                wd[l].X.reset(ctx, sz); wd[l].X2.reset(ctx, sz); wd[l].Res.reset(ctx, sz);
            }
        }
    }
    tmp.reset(ctx, (size_t)std::max(A->n, 1) * k);
    if (n_hrc() > 1 && p.bootstrap_type == BCMGX_BOOT_ADDITIVE) {
        tmp2.reset(ctx, (size_t)std::max(A->n, 1) * k);
        tmp3.reset(ctx, (size_t)std::max(A->n, 1) * k);
    }
}

template <typename V>
static const V* mat_val(const DCsr& A) {
    if constexpr (std::is_same<V, float>::value) return A.valf.get();
    else return A.val.get();
}

// k sweeps of l1-Jacobi; (cur, oth) is the ping-pong pair, cur holds the
// guess (ignored when xzero).  Returns the buffer holding the result.  With
// `out64` (FP32 sweeps only) the last sweep (one from a nonzero guess) writes
// its result widened to FP64 there instead, and nullptr is returned; with
// `outv` it writes it there (same type) and returns outv.
template <typename V>
static V* jacobi_sweeps(Context* c, const DCsr& A, int lanes, const V* M, const V* F, V* cur, V* oth, int s, bool xzero,
                        int k, const int* stop, int chunk = 0, double* out64 = nullptr, V* outv = nullptr) {
    int done = 0;
    if (s <= 0) {
        if (xzero) fill<V>(c, cur, (long long)A.n * k, (V)0);
        return cur;
    }
    if (xzero) {
        jacobi_from_zero<V>(c, A.n, k, F, M, cur, stop);
        done = 1;
    }
    for (; done < s; done++) {
        SpmvArgs a;
        a.n = A.n; a.row = A.row.get(); a.col = A.col.get(); a.val = mat_val<V>(A); a.nnz = A.nnz;
        set_tiles(a, A);
        a.x = cur; a.y = oth; a.b = F; a.D = M; a.k = k; a.lanes = lanes; a.stop = stop; a.kchunk = chunk;
        bool last = done + 1 == s;
        if (last && out64) {
            a.y64 = out64;
            spmv<V, V>(c, SpOp::JACOBI, a);
            return nullptr;
        }
        if (last && outv) {
            a.y = outv;
            spmv<V, V>(c, SpOp::JACOBI, a);
            return outv;
        }
        spmv<V, V>(c, SpOp::JACOBI, a);
        std::swap(cur, oth);
    }
    return cur;
}

template <typename V>
struct CycleRun {
    Context* c;
    Prec* pr;
    Hierarchy* H;  // the hierarchy cycled (one of a bootstrap's)
    std::vector<LevelWork<V>>* W;
    int k;
    const int* stop;
    const double* f0 = nullptr;  // level 0's right-hand side in FP64 when its F is not written
};

template <typename V>
static const V* lvl_M(const Level& l) {
    if constexpr (std::is_same<V, float>::value) return l.Mf.get();
    else return l.M.get();
}
template <typename V>
static const V* lvl_pval(const Level& l) {
    if constexpr (std::is_same<V, float>::value) return l.pvalf.get();
    else return l.pval.get();
}

// GAMG_cycle (GAMG_cycle.cu:58-236), 0-based level l.  Xin: initial guess
// buffer (one of W[l].X / W[l].X2) or nullptr for zero.  Returns the buffer
// holding the level's result.
template <typename V>
static V* gamg_cycle(CycleRun<V>& R, int l, const V* F, V* Xin, bool first_sweep_done = false, double* out64 = nullptr,
                     V* outv = nullptr) {
    Context* c = R.c;
    Prec* pr = R.pr;
    const bcmgx_params& p = pr->p;
    Hierarchy& H = *R.H;
    int nl = H.levels();
    Level& lv = *H.L[l];
    const DCsr& A = lv.mat();
    LevelWork<V>& w = (*R.W)[l];
    V* cur = Xin ? Xin : w.X.get();
    V* oth = (cur == w.X.get()) ? w.X2.get() : w.X.get();
    bool xzero = Xin == nullptr;
    int k = R.k;
    if (l == nl - 1) {
        if constexpr (std::is_same<V, float>::value) {
            // the coarsest-level sweeps run in FP64 on the FP64 operator
            long long nk = (long long)A.n * k;
            double* res;
            if (xzero && p.relaxnumber_coarse >= 1) {  // widening and the first sweep from zero in one pass
                jacobi_from_zero_f32in(c, A.n, k, F, lv.M.get(), pr->cd_F.get(), pr->cd_X.get(), R.stop);
                res = jacobi_sweeps<double>(c, A, lv.lanesA, lv.M.get(), pr->cd_F.get(), pr->cd_X.get(), pr->cd_X2.get(),
                                            p.relaxnumber_coarse - 1, false, k, R.stop);
            } else {
                convert_stop<float, double>(c, F, pr->cd_F.get(), nk, R.stop);
                if (!xzero) convert_stop<float, double>(c, cur, pr->cd_X.get(), nk, R.stop);
                res = jacobi_sweeps<double>(c, A, lv.lanesA, lv.M.get(), pr->cd_F.get(), pr->cd_X.get(), pr->cd_X2.get(),
                                            p.relaxnumber_coarse, xzero, k, R.stop);
            }
            convert_stop<double, float>(c, res, cur, nk, R.stop);
            return cur;
        } else {
            return jacobi_sweeps<V>(c, A, lv.lanesA, lvl_M<V>(lv), F, cur, oth, p.relaxnumber_coarse, xzero, k, R.stop,
                                    lv.chunkA);
        }
    }
    int coeff = p.cycle_type == BCMGX_CYCLE_VARIABLE_V ? (1 << l) : 1;
    if (first_sweep_done) {
        // w.X already holds the first sweep from zero, F / M
        cur = w.X.get();
        oth = w.X2.get();
        cur = jacobi_sweeps<V>(c, A, lv.lanesA, lvl_M<V>(lv), F, cur, oth, p.prerelax_sweeps * coeff - 1, false, k, R.stop,
                               lv.chunkA);
    } else {
        cur = jacobi_sweeps<V>(c, A, lv.lanesA, lvl_M<V>(lv), F, cur, oth, p.prerelax_sweeps * coeff, xzero, k, R.stop,
                               lv.chunkA);
    }
    oth = (cur == w.X.get()) ? w.X2.get() : w.X.get();
    // Res = F - A x ; F_{l+1} = R Res
    {
        SpmvArgs a;
        a.n = A.n; a.row = A.row.get(); a.col = A.col.get(); a.val = mat_val<V>(A); a.nnz = A.nnz;
        set_tiles(a, A);
        a.x = cur; a.y = w.Res.get(); a.b = F; a.k = k; a.lanes = lv.lanesA; a.stop = R.stop; a.kchunk = lv.chunkA;
        if (l == 0 && R.f0) {  // F = (float) r rounded as it is read (the same values)
            a.b = nullptr;
            a.b64 = R.f0;
        }
        spmv<V, V>(c, SpOp::RES, a);
    }
    LevelWork<V>& wc = (*R.W)[l + 1];
    // a coarse level below the coarsest starts with a sweep from zero, F / M:
    // the restriction writes it along with F
    bool sweep0 = l + 1 < nl - 1 && p.prerelax_sweeps >= 1;
    {
        SpmvArgs a;
        a.n = lv.R.n; a.row = lv.R.row.get(); a.col = lv.R.col.get(); a.val = mat_val<V>(lv.R); a.nnz = lv.R.nnz;
        a.x = w.Res.get(); a.y = wc.F.get(); a.k = k; a.lanes = lv.lanesR; a.stop = R.stop;
        if (sweep0) {
            a.x0 = wc.X.get();
            a.D = lvl_M<V>(*H.L[l + 1]);
        }
        spmv<V, V>(c, SpOp::AX, a);
    }
    V* xc = nullptr;
    if (nl > 2 || p.relaxnumber_coarse > 0) {
        int gs = 1;  // setGridSweeps (AMG.cu:306-331)
        if (p.cycle_type == BCMGX_CYCLE_W && l < nl - 2) gs = 2;
        if (p.cycle_type == BCMGX_CYCLE_H && (l % 2) == 0) gs = 2;
        for (int i = 1; i <= gs; i++) {
            xc = gamg_cycle<V>(R, l + 1, wc.F.get(), xc, i == 1 && sweep0);
            if (l + 1 == nl - 1) break;
        }
    }
    oth = (cur == w.X.get()) ? w.X2.get() : w.X.get();
    int post = p.postrelax_sweeps * coeff;
    if constexpr (std::is_same<V, float>::value) {
        // the prolongation and the first post-smoothing sweep from A P in one pass:
        // f - A (x + P e) = (f - A x) - (A P) e, the first term still in w.Res
        if (xc && post >= 1 && lv.AP.valf) {
            V* y = post == 1 && outv ? outv : oth;
            SpmvArgs a;
            a.n = lv.AP.n; a.row = lv.AP.row.get(); a.col = lv.AP.col.get(); a.val = lv.AP.valf.get();
            a.nnz = lv.AP.nnz; a.x = xc; a.y = y; a.b = w.Res.get(); a.D = lvl_M<V>(lv); a.k = k;
            a.lanes = lv.lanesAP; a.stop = R.stop;
            galerkin_post_sweep<V>(c, a, lv.agg.get(), lvl_pval<V>(lv), cur);
            if (post == 1) {
                if (!out64) return y;
                convert_stop<V, double>(c, y, out64, (long long)A.n * k, R.stop);  // the result widened
                return nullptr;
            }
            cur = y;
            oth = (cur == w.X.get()) ? w.X2.get() : w.X.get();
            return jacobi_sweeps<V>(c, A, lv.lanesA, lvl_M<V>(lv), F, cur, oth, post - 1, false, k, R.stop, lv.chunkA,
                                    out64, outv);
        }
    }
    if (xc) prolong_add<V>(c, A.n, k, lv.agg.get(), lvl_pval<V>(lv), xc, cur, R.stop);
    cur = jacobi_sweeps<V>(c, A, lv.lanesA, lvl_M<V>(lv), F, cur, oth, post, false, k, R.stop, lv.chunkA, out64, outv);
    return cur;
}

void Prec::refresh_level0() {
    if (ptype == BCMGX_PREC_L1_JACOBI) {
        l1_diag(ctx, A->n, A->row.get(), A->col.get(), A->val.get(), pl1j.get());
    } else if (ptype == BCMGX_PREC_BCMG && !H.L.empty()) {
        for (int h = 0; h < n_hrc(); h++) {  // every hierarchy's level 0 is A
            Level& l0 = *hrc(h).L[0];
            abs_row_sum(ctx, A->n, A->row.get(), A->val.get(), l0.M.get());
            if (l0.Mf) convert<double, float>(ctx, l0.M.get(), l0.Mf.get(), A->n);
            if (l0.AP.valf) make_level_product(ctx, hrc(h), 0);  // A P with the new values of A
        }
    }
}

bool Prec::fused_entry(int k, FcgEntry* e) {
    f0_from_r = false;
    if (ptype != BCMGX_PREC_BCMG || p.precision != BCMGX_PRECISION_FP32 || H.levels() < 2 || p.prerelax_sweeps < 1 ||
        n_hrc() > 1)
        return false;
    ensure_work(k);
    // level 0's F is read only by the residual after the first sweep when that sweep
    // is the only pre-smoothing, the first post-smoothing sweep comes from A P and is
    // the only one, and a coarse correction exists; the residual then takes r (FP64)
    // when it runs on the pipelined k-column SpMM
    const Level& l0 = *H.L[0];
    bool pipelined = k % 4 == 0 && k >= 8 && k <= 64 && l0.chunkA != -1 && (l0.chunkA <= 0 || l0.chunkA >= k);
    f0_from_r = p.prerelax_sweeps == 1 && p.postrelax_sweeps == 1 && p.post_galerkin && l0.AP.valf &&
                (H.levels() > 2 || p.relaxnumber_coarse > 0) && pipelined;
    e->F = f0_from_r ? nullptr : wf[0].F.get();
    e->X = wf[0].X.get();
    e->M = H.L[0]->Mf.get();
    return true;
}

void Prec::apply_entry_f32(const double* r, float* z, int k, const int* stop) {
    BCMGX_CHECK(ptype == BCMGX_PREC_BCMG && p.precision == BCMGX_PRECISION_FP32 && H.levels() >= 2 && p.prerelax_sweeps >= 1 &&
                    n_hrc() == 1,
                BCMGX_ERR_STATE, "level-0 entry not applicable");
    ensure_work(k);
    CycleRun<float> R{ctx, this, &H, &wf, k, stop};
    if (f0_from_r) R.f0 = r;
    float* res = gamg_cycle<float>(R, 0, wf[0].F.get(), nullptr, true, nullptr, z);
    if (res != z) copy_d2d(ctx, z, res, sizeof(float) * (size_t)A->n * k);
}

void Prec::apply_hrc(int h, const double* r, double* u, int k, bool x_zero, const int* stop, bool entry_ready) {
    ensure_work(k);
    Context* c = ctx;
    long long nk = (long long)A->n * k;
    Hierarchy& Hh = hrc(h);
    if (p.precision == BCMGX_PRECISION_FP32) {
        CycleRun<float> R{c, this, &Hh, &wf, k, stop};
        float* res;
        // the last level-0 sweep writes u in FP64 (res == nullptr) when the cycle ends in one
        if (entry_ready) {
            BCMGX_CHECK(x_zero && h == 0 && Hh.levels() >= 2 && p.prerelax_sweeps >= 1, BCMGX_ERR_STATE,
                        "level-0 entry not applicable");
            if (f0_from_r) R.f0 = r;
            res = gamg_cycle<float>(R, 0, wf[0].F.get(), nullptr, true, u);
        } else {
            convert_stop<double, float>(c, r, wf[0].F.get(), nk, stop);
            float* xin = nullptr;
            if (!x_zero) {
                convert_stop<double, float>(c, u, wf[0].X.get(), nk, stop);
                xin = wf[0].X.get();
            }
            res = gamg_cycle<float>(R, 0, wf[0].F.get(), xin, false, u);
        }
        if (res) convert_stop<float, double>(c, res, u, nk, stop);
    } else {
        BCMGX_CHECK(!entry_ready, BCMGX_ERR_STATE, "level-0 entry not applicable");
        CycleRun<double> R{c, this, &Hh, &wd, k, stop};
        double* xin = nullptr;
        if (!x_zero) {
            copy_d2d(c, wd[0].X.get(), u, sizeof(double) * nk);
            xin = wd[0].X.get();
        }
        // the last level-0 sweep writes u directly when the cycle ends in one
        double* res = gamg_cycle<double>(R, 0, r, xin, false, nullptr, u);
        if (res != u) convert_stop<double, double>(c, res, u, nk, stop);
    }
}

void Prec::apply(const double* r, double* u, int k, bool x_zero, const int* stop, bool entry_ready) {
    ensure_work(k);
    Context* c = ctx;
    long long nk = (long long)A->n * k;
    switch (ptype) {
    case BCMGX_PREC_NONE:
        copy_d2d(c, u, r, sizeof(double) * nk);
        break;
    case BCMGX_PREC_L1_JACOBI: {  // l1jacobi_apply (l1jacobi.cu:513-523)
        double* res = jacobi_sweeps<double>(c, *A, H.L.empty() ? 0 : 0, pl1j.get(), r, u, tmp.get(), p.l1jacsweeps, x_zero, k, stop);
        if (res != u) copy_d2d(c, u, res, sizeof(double) * nk);
        break;
    }
    case BCMGX_PREC_AFSAI: {  // afsai_apply (afsai.cu:1092-1103): u = G^T (G r)
        SpmvArgs a;
        a.n = G.n; a.row = G.row.get(); a.col = G.col.get(); a.val = G.val.get(); a.nnz = G.nnz;
        a.x = r; a.y = tmp.get(); a.k = k; a.lanes = lanesG; a.stop = stop;
        spmv<double, double>(c, SpOp::AX, a);
        SpmvArgs b;
        b.n = GT.n; b.row = GT.row.get(); b.col = GT.col.get(); b.val = GT.val.get(); b.nnz = GT.nnz;
        b.x = tmp.get(); b.y = u; b.k = k; b.lanes = lanesGT; b.stop = stop;
        spmv<double, double>(c, SpOp::AX, b);
        break;
    }
    case BCMGX_PREC_BCMG: {  // bcmg_apply (bcmg.cu:51-90)
        int nh = n_hrc();
        if (nh == 1) {
            apply_hrc(0, r, u, k, x_zero, stop, entry_ready);
            break;
        }
        BCMGX_CHECK(!entry_ready, BCMGX_ERR_STATE, "level-0 entry not applicable");
        if (p.bootstrap_type == BCMGX_BOOT_ADDITIVE) {
            // u <- u + sum_h B_h (r - A u), the corrections added in hierarchy order
            const double* res = r;
            if (!x_zero) {
                SpmvArgs a;
                a.n = A->n; a.row = A->row.get(); a.col = A->col.get(); a.val = A->val.get(); a.nnz = A->nnz;
                set_tiles(a, *A);
                a.x = u; a.y = tmp.get(); a.b = r; a.k = k; a.lanes = cached_lanes(c, *A, false); a.stop = stop;
                spmv<double, double>(c, SpOp::RES, a);
                res = tmp.get();
            }
            // from a zero guess the sum is accumulated in u itself (0 + sum = sum)
            double* acc = x_zero ? u : tmp2.get();
            double* t = x_zero ? tmp2.get() : tmp3.get();
            apply_hrc(0, res, acc, k, true, stop);
            for (int h = 1; h < nh; h++) {
                apply_hrc(h, res, t, k, true, stop);
                add_to(c, acc, t, nk, stop);
            }
            if (!x_zero) add_to(c, u, acc, nk, stop);
        } else {
            // MULTIPLICATIVE (bcmg.cu:65-85): each hierarchy's cycle starts from the
            // previous result; SYMMETRIZED_MULTIPLICATIVE adds the way back, h = n-2 .. 0
            for (int h = 0; h < nh; h++) apply_hrc(h, r, u, k, x_zero && h == 0, stop);
            if (p.bootstrap_type == BCMGX_BOOT_SYMMETRIZED_MULTIPLICATIVE)
                for (int h = nh - 2; h >= 0; h--) apply_hrc(h, r, u, k, false, stop);
        }
        break;
    }
    default:
        throw Error(BCMGX_ERR_PARAM, "unhandled preconditioner type");
    }
}

// bootstrap (bootstrap.cu:11-46): after the first hierarchy, while fewer than
// max_hrc exist and the composition of those built reduces the slowest error
// by no better than conv_ratio per application, one more hierarchy from that
// error.  The reference leaves the estimate at Boot::init's 1.0 and asserts a
// single hierarchy (bootstrap.cu:37-38); the estimate here is BootCMatch's:
// solver_it = 15 (AMG.cu:342) applications of the composition to A x = 0 from a
// random x, ratio = sqrt(x_15' A x_15 / x_14' A x_14), and the next smooth
// vector x_15 / max_i |x_15,i|.  The applications are the preconditioner's own
// (right-hand side 0, guess x); the dot products are fixed-order block sums.
template <typename Build>
static void bootstrap(Context* c, const DCsr& A, Prec* pr, Build build_next) {
    const bcmgx_params& p = pr->p;
    const int solver_it = 15;
    int n = A.n;
    while (pr->n_hrc() < p.max_hrc) {
        pr->ensure_work(1);
        DBuf<double> x(c, std::max(n, 1)), ax(c, std::max(n, 1)), zero(c, std::max(n, 1)), dots(c, solver_it + 1);
        int nblk = fcg_partial_blocks(n);
        DBuf<double> part(c, std::max(nblk, 1));
        fill<double>(c, zero.get(), n, 0.0);
        boot_random(c, x.get(), n, (unsigned long long)pr->n_hrc());
        auto anorm2 = [&](int slot) {
            SpmvArgs a;
            a.n = n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
            set_tiles(a, A);
            a.x = x.get(); a.y = ax.get(); a.lanes = cached_lanes(c, A, false);
            spmv<double, double>(c, SpOp::AX, a);
            block_dots(c, n, 1, x.get(), ax.get(), 0, part.get(), nblk, dots.get() + slot);
        };
        anorm2(0);
        for (int it = 1; it <= solver_it; it++) {
            pr->apply(zero.get(), x.get(), 1, false, nullptr);  // x <- (I - B A) x
            anorm2(it);
        }
        double a[2];
        c->d2h(a, dots.get() + solver_it - 1, 2);
        pr->boot_ratio = a[0] > 0.0 ? std::sqrt(a[1] / a[0]) : 0.0;
        if (!(pr->boot_ratio > p.conv_ratio)) break;
        scale_by_absmax(c, x.get(), n);
        build_next(x.get());
        pr->kcap = 0;  // the cycle buffers are sized again for the new hierarchy
        pr->tuned_k = 0;
    }
}

std::unique_ptr<Prec> prec_setup(Context* c, const DCsr& A, const bcmgx_params& p) {
    double t0 = now_s();
    std::unique_ptr<Prec> pr(new Prec());
    pr->ctx = c;
    pr->A = &A;
    pr->p = p;
    pr->ptype = p.preconditioner;
    // the row-tile index of the caller's matrix (its pattern only: built once, kept
    // with the matrix) serves the one-column products with it
    if (A.n >= 65536 && p.autotune) build_tile_index(c, const_cast<DCsr&>(A));
    switch (p.preconditioner) {
    case BCMGX_PREC_NONE:
        break;
    case BCMGX_PREC_L1_JACOBI:
        pr->pl1j.reset(c, std::max(A.n, 1));
        l1_diag(c, A.n, A.row.get(), A.col.get(), A.val.get(), pr->pl1j.get());
        break;
    case BCMGX_PREC_AFSAI:
        afsai_build(c, A, pr->G, pr->GT);
        pr->lanesG = default_lanes(pr->G.nnz, pr->G.n);
        pr->lanesGT = default_lanes(pr->GT.nnz, pr->GT.n);
        break;
    case BCMGX_PREC_BCMG: {
        BCMGX_CHECK(p.aggr_type == 0, BCMGX_ERR_UNSUPPORTED, "aggr_type != 0 is not supported");
        BCMGX_CHECK(p.bootstrap_type >= 0 && p.bootstrap_type < BCMGX_BOOT_INVALID, BCMGX_ERR_PARAM,
                    "invalid bootstrap_type");
        BCMGX_CHECK(p.max_hrc >= 1, BCMGX_ERR_PARAM, "max_hrc must be at least 1");
        BCMGX_CHECK(p.cycle_type >= 0 && p.cycle_type < BCMGX_CYCLE_INVALID, BCMGX_ERR_PARAM, "invalid cycle_type");
        bool f32 = p.precision == BCMGX_PRECISION_FP32;
        // a level's FP32 copies and A_l P_l go to the second stream as soon as
        // the level is final; finish() joins them
        SideQueue side(c);
        std::vector<char> async_product;
        auto level_done = [&](Hierarchy& H) {
            async_product.assign(64, 0);
            Hierarchy* hp = &H;
            return [&, hp](int l) {  // runs on the side queue
                if (!f32) return;
                level_fp32_copy(c, *hp->L[l]);
                if (p.post_galerkin && l + 1 < hp->levels() && l < (int)async_product.size())
                    async_product[l] = level_product_async(c, *hp, l) ? 1 : 0;
            };
        };
        auto finish = [&](Hierarchy& H) {
            side.join();
            if (f32 && p.post_galerkin) {
                std::vector<Level*> queued;
                for (int l = 0; l + 1 < H.levels(); l++) {
                    if (l < (int)async_product.size() && async_product[l]) queued.push_back(H.L[l].get());
                    else make_level_product(c, H, l);
                }
                level_product_lengths(c, queued);
            }
            for (auto& lp : H.L) {
                Level& l = *lp;
                const DCsr& M = l.mat();
                l.lanesA = p.autotune ? tune_lanes(c, M, f32) : default_lanes(M.nnz, M.n);
                l.lanesR = l.R.n > 0 ? default_lanes(l.R.nnz, l.R.n) : 1;
            }
        };
        build_hierarchy(c, A, p, pr->H, nullptr, &side, level_done(pr->H));
        finish(pr->H);
        bootstrap(c, A, pr.get(), [&](const double* w) {
            pr->HB.emplace_back(new Hierarchy());
            Hierarchy& H = *pr->HB.back();
            build_hierarchy(c, A, p, H, w, &side, level_done(H));
            finish(H);
        });
        break;
    }
    default:
        throw Error(BCMGX_ERR_PARAM, "invalid preconditioner type");
    }
    pr->ensure_work(1);
    c->sync();
    pr->setup_seconds = now_s() - t0;
    return pr;
}

// ---------------------------------------------------------------- FCG ----

// The same iteration with z in FP32: the direction d and its image q are updated
// in place (no ping-pong), w holds A z.  The same operations in the same order as
// the widened-FP64 iteration; only the dot products group their sums differently.
static void fcg_iteration_z32(Context* c, const DCsr& A, Prec* prec, SolveWork& s, int idx) {
    const int* stop = s.flags.get();
    prec->apply_entry_f32(s.r.get(), s.zf.get(), s.k, stop);
    SpmvArgs a;
    a.n = A.n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
    set_tiles(a, A);
    a.x = s.zf.get(); a.y = s.w.get(); a.k = s.k; a.lanes = s.lanesA; a.stop = stop; a.kchunk = s.chunkA;
    BCMGX_CHECK(spmm_dots_z32(c, a, s.r.get(), s.q.get(), s.fb.partials, s.z32_blocks, s.fb.all_done), BCMGX_ERR_STATE,
                "fused FP32-z product not applicable");
    fcg_scalars3(c, s.fb, s.z32_blocks, idx);
    fcg_update_z32(c, s.fb, s.zf.get(), s.d.get(), s.w.get(), s.q.get(), s.X.get(), s.r.get(), idx, s.entry);
    fcg_finish_iteration(c, s.fb, idx);
}

static void fcg_iteration(Context* c, const DCsr& A, Prec* prec, SolveWork& s, int idx) {
    if (s.z32) {
        fcg_iteration_z32(c, A, prec, s, idx);
        return;
    }
    double* P0 = idx == 0 ? s.v.get() : s.d.get();
    double* W0 = idx == 0 ? s.w.get() : s.q.get();
    double* P1 = idx == 0 ? s.d.get() : s.v.get();
    double* W1 = idx == 0 ? s.q.get() : s.w.get();
    const int* stop = s.flags.get();
    prec->apply(s.r.get(), P0, s.k, true, stop, s.fused);  // P0 zero-filled in the reference (FCG.cu:106,119)
    SpmvArgs a;
    a.n = A.n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
    set_tiles(a, A);
    a.x = P0; a.y = W0; a.k = s.k; a.lanes = s.lanesA; a.stop = stop; a.kchunk = s.chunkA;
    if (s.mv_blocks > 0 && spmm_dots(c, a, s.r.get(), W1, s.fb.partials, s.mv_blocks, s.fb.all_done)) {
        fcg_scalars3(c, s.fb, s.mv_blocks, idx);
    } else {
        spmv<double, double>(c, SpOp::AX, a);
        fcg_dots3(c, s.fb, s.r.get(), P0, W0, W1, idx);
    }
    fcg_update(c, s.fb, P0, P1, W0, W1, s.X.get(), s.r.get(), idx, s.entry);
    fcg_finish_iteration(c, s.fb, idx);
}

// the Krylov workspace for k columns of A (kept per k: graphs are bound to it)
static SolveWork& solve_work(Context* c, const DCsr& A, Prec* prec, const bcmgx_params& p, int k) {
    int n = A.n;
    SolveWork* sp = nullptr;
    for (auto& w : prec->solves)
        if (w->k == k && w->n == n) sp = w.get();
    if (!sp) {
        prec->solves.emplace_back(new SolveWork());
        sp = prec->solves.back().get();
        SolveWork& s = *sp;
        s.k = k;
        s.n = n;
        size_t nk = (size_t)std::max(n, 1) * k;
        s.B.reset(c, nk); s.X.reset(c, nk); s.r.reset(c, nk); s.v.reset(c, nk);
        s.w.reset(c, nk); s.d.reset(c, nk); s.q.reset(c, nk);
        s.fb.n = n;
        s.fb.k = k;
        s.fb.nblk = fcg_partial_blocks(n);
        s.mv_blocks = spmm_dots_blocks(c, n, k);
        s.z32_blocks = spmm_dots_z32_blocks(c, n, k);
        s.partials.reset(c, (size_t)std::max({s.fb.nblk, s.mv_blocks, s.z32_blocks}) * k * 3);
        s.st.reset(c, k);
        s.flags.reset(c, 3);
        s.fb.partials = s.partials.get();
        s.fb.st = s.st.get();
        s.fb.all_done = s.flags.get();
        s.lanesA = p.autotune ? tune_lanes(c, A, false) : default_lanes(A.nnz, A.n);
        s.chunkA = p.autotune ? tune_chunk(c, A, false, k) : k;
    }
    return *sp;
}

SolveResult fcg_solve(Context* c, const DCsr& A, Prec* prec, const bcmgx_params& p, const double* b, double* x, int k,
                      const SolveIo* io) {
    BCMGX_CHECK(k >= 1 && k <= 64, BCMGX_ERR_ARG, "nrhs must be in [1, 64]");
    BCMGX_CHECK(p.solver == BCMGX_SOLVER_FCG, BCMGX_ERR_STATE, "fcg_solve called for another solver");
    double t0 = now_s();
    int n = A.n;
    SolveWork& s = solve_work(c, A, prec, p, k);
    if (s.cap_rtol != p.rtol || s.cap_itnlim != p.itnlim) s.drop_graph();
    s.fb.rtol = s.cap_rtol = p.rtol;
    s.fb.itnlim = s.cap_itnlim = p.itnlim;
    prec->ensure_work(k);
    s.entry = FcgEntry();
    s.fused = prec->fused_entry(k, &s.entry);
    // z in FP32 when the FP32 cycle is entered from the update and the fused FP32-z product applies
    bool z32 = s.fused && p.fcg_z32 && s.z32_blocks > 0 && s.z32_blocks % 8 == 0 && (s.chunkA <= 0 || s.chunkA >= k);
    if (z32 != s.z32) s.drop_graph();  // the captured iterations differ
    s.z32 = z32;
    if (s.z32 && !s.zf) s.zf.reset(c, (size_t)std::max(n, 1) * k);
    long long nk = (long long)n * k;
    int* xzero = s.flags.get() + 2;  // the initial guess holds only zeros (set while staging it)
    if (io) {
        fill<int>(c, xzero, 1, 1);
        colmajor_to_interleaved(c, io->x, io->ld, s.X.get(), n, k, xzero);
        // a zero guess: every row sum of A x is +0, so r = b - A x is b itself
        // and b goes straight to r
        colmajor_to_interleaved(c, io->b, io->ld, s.B.get(), n, k, nullptr, xzero, s.r.get());
    } else {
        copy_d2d(c, s.B.get(), b, sizeof(double) * nk);
        copy_d2d(c, s.X.get(), x, sizeof(double) * nk);
    }
    fill<int>(c, s.flags.get(), 2, 0);
    // ---- initial step (FCG.cu:44-99)
    {
        SpmvArgs a;
        a.n = n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
        set_tiles(a, A);
        a.x = s.X.get(); a.y = s.r.get(); a.b = s.B.get(); a.k = k; a.lanes = s.lanesA; a.kchunk = s.chunkA;
        a.stop = io ? xzero : nullptr;
        spmv<double, double>(c, SpOp::RES, a);  // r = b - A x
    }
    if (s.fused) fcg_norm_entry(c, s.fb, s.r.get(), 0, s.entry);  // delta0 = ||r||, and the FP32 cycle's entry
    else fcg_norm(c, s.fb, s.r.get(), 0);                         // delta0 = ||r||
    if (s.z32) {
        // v = B r kept in FP32; w = A v with (r.v, w.v) in the fused product (its third
        // dot, against r here, is not used)
        prec->apply_entry_f32(s.r.get(), s.zf.get(), k, nullptr);
        SpmvArgs a;
        a.n = n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
        set_tiles(a, A);
        a.x = s.zf.get(); a.y = s.w.get(); a.k = k; a.lanes = s.lanesA; a.kchunk = s.chunkA;
        BCMGX_CHECK(spmm_dots_z32(c, a, s.r.get(), s.r.get(), s.fb.partials, s.z32_blocks, nullptr), BCMGX_ERR_STATE,
                    "fused FP32-z product not applicable");
        fcg_init_scalars(c, s.fb, s.z32_blocks);
        fcg_init_update_z32(c, s.fb, s.X.get(), s.r.get(), s.zf.get(), s.w.get(), s.d.get(), s.q.get(), s.entry);
    } else {
        prec->apply(s.r.get(), s.v.get(), k, true, nullptr, s.fused);
        SpmvArgs a;
        a.n = n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
        set_tiles(a, A);
        a.x = s.v.get(); a.y = s.w.get(); a.k = k; a.lanes = s.lanesA; a.kchunk = s.chunkA;
        spmv<double, double>(c, SpOp::AX, a);  // w = A v
        fcg_init_dots(c, s.fb, s.r.get(), s.v.get(), s.w.get());
        fcg_init_update(c, s.fb, s.X.get(), s.r.get(), s.v.get(), s.w.get(), s.d.get(), s.q.get(), s.entry);
    }
    // ---- iterations, two at a time (the buffer roles alternate)
    int check = std::max(2, p.check_every + (p.check_every & 1));
    bool use_graph = p.use_graph != 0;
    if (use_graph && !s.exec) {
        BCMGX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            fcg_iteration(c, A, prec, s, 0);
            fcg_iteration(c, A, prec, s, 1);
        } catch (...) {
            hipGraph_t g;
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        BCMGX_HIP(hipStreamEndCapture(c->stream, &s.graph));
        BCMGX_HIP(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
    }
    int launched = 0;
    int* hflag = (int*)c->host_scratch(sizeof(int) * 2);
    while (true) {
        for (int j = 0; j < check; j += 2) {
            if (use_graph) BCMGX_HIP(hipGraphLaunch(s.exec, c->stream));
            else {
                fcg_iteration(c, A, prec, s, 0);
                fcg_iteration(c, A, prec, s, 1);
            }
            launched += 2;
        }
        BCMGX_HIP(hipMemcpyAsync(hflag, s.flags.get(), sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
        c->sync();
        if (hflag[0]) break;
        if (launched > p.itnlim + 2) break;  // the device stops every column at itnlim
    }
    // every column stopped after the first iteration of the last pair: its x step waits in pend
    // and its direction is still in v (the first iteration's p0), or in d with z in FP32
    const double* pw = s.z32 ? s.d.get() : s.v.get();
    if (io) {  // the flush rides on the way out
        interleaved_to_colmajor(c, s.X.get(), io->x, io->ld, n, k, s.fb.st, pw);
        fcg_clear_pending(c, s.fb);
    } else {
        fcg_flush_pending(c, s.fb, pw, s.X.get());
        copy_d2d(c, x, s.X.get(), sizeof(double) * nk);
    }
    std::vector<FcgState> st(k);
    c->d2h(st.data(), s.st.get(), k);
    SolveResult res;
    prec->last_niter.assign(k, 0);
    prec->last_exit.assign(k, 0);
    prec->last_rel.assign(k, 0);
    for (int j = 0; j < k; j++) {
        int niter = st[j].iter + 1;  // FCG.cu:164
        double rel = st[j].delta0 > 0 ? st[j].l2 / st[j].delta0 : 0.0;
        prec->last_niter[j] = niter;
        prec->last_exit[j] = st[j].l2;
        prec->last_rel[j] = rel;
        res.niter = std::max(res.niter, niter);
        res.exit_res = std::max(res.exit_res, st[j].l2);
        res.rel_res = std::max(res.rel_res, rel);
        if (st[j].iter >= p.itnlim) res.retv = -1;  // FCG.cu:160-162
        if (!std::isfinite(st[j].l2)) res.retv = BCMGX_ERR_NUMERIC;
    }
    res.del0 = st[0].delta0;
    res.seconds = now_s() - t0;
    return res;
}

// cg_hs (src/solver/cghs/CG_HS.cu:24-172) on k interleaved columns: the FCG
// workspace, preconditioners and vector kernels, every scalar on the device
// (per-column masks), the general iteration replayed from a hipGraph.
SolveResult cghs_solve(Context* c, const DCsr& A, Prec* prec, const bcmgx_params& p, const double* b, double* x, int k) {
    BCMGX_CHECK(k >= 1 && k <= 64, BCMGX_ERR_ARG, "nrhs must be in [1, 64]");
    BCMGX_CHECK(p.solver == BCMGX_SOLVER_CGHS, BCMGX_ERR_STATE, "cghs_solve called for another solver");
    double t0 = now_s();
    int n = A.n;
    SolveWork& s = solve_work(c, A, prec, p, k);
    if (s.cap_rtol != p.rtol || s.cap_itnlim != p.itnlim) s.drop_graph();
    s.fb.rtol = s.cap_rtol = p.rtol;
    s.fb.itnlim = s.cap_itnlim = p.itnlim;
    prec->ensure_work(k);
    long long nk = (long long)n * k;
    copy_d2d(c, s.B.get(), b, sizeof(double) * nk);
    copy_d2d(c, s.X.get(), x, sizeof(double) * nk);
    fill<int>(c, s.flags.get(), 2, 0);
    double* U = s.v.get();  // preconditioned residual
    double* S = s.w.get();  // A p
    double* P = s.d.get();  // direction
    double* R = s.r.get();
    const int* stop = s.flags.get();
    auto spmm = [&](const double* xin, double* yout, SpOp op, const int* st) {
        SpmvArgs a;
        a.n = n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
        set_tiles(a, A);
        a.x = xin; a.y = yout; a.b = s.B.get(); a.k = k; a.lanes = s.lanesA; a.kchunk = s.chunkA; a.stop = st;
        spmv<double, double>(c, op, a);
    };
    spmm(s.X.get(), R, SpOp::RES, nullptr);  // r = b - A x  (CG_HS.cu:53-57)
    prec->apply(R, U, k, true, nullptr);    // u = B r, p = u  (CG_HS.cu:59-68)
    copy_d2d(c, P, U, sizeof(double) * nk);
    cg_init(c, s.fb, R, p.stop_criterion);
    if (p.itnlim >= 1) {
        // iteration 0 (CG_HS.cu:110-135): no preconditioner application, rho = u.r
        spmm(P, S, SpOp::AX, stop);
        cg_alpha(c, s.fb, P, S, U, R, true);
        cg_update_finish(c, s.fb, P, S, s.X.get(), R);
        auto iteration = [&]() {  // CG_HS.cu:87-135
            prec->apply(R, U, k, true, stop);
            cg_beta_direction(c, s.fb, R, U, P);
            spmm(P, S, SpOp::AX, stop);
            cg_alpha(c, s.fb, P, S, U, R, false);
            cg_update_finish(c, s.fb, P, S, s.X.get(), R);
        };
        bool use_graph = p.use_graph != 0;
        if (use_graph && !s.cg_exec) {
            BCMGX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            try {
                iteration();
            } catch (...) {
                hipGraph_t g;
                (void)hipStreamEndCapture(c->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            BCMGX_HIP(hipStreamEndCapture(c->stream, &s.cg_graph));
            BCMGX_HIP(hipGraphInstantiate(&s.cg_exec, s.cg_graph, nullptr, nullptr, 0));
        }
        int check = std::max(1, p.check_every);
        int launched = 1;
        int* hflag = (int*)c->host_scratch(sizeof(int) * 2);
        while (launched < p.itnlim) {
            for (int j = 0; j < check && launched < p.itnlim; j++, launched++) {
                if (use_graph) BCMGX_HIP(hipGraphLaunch(s.cg_exec, c->stream));
                else iteration();
            }
            BCMGX_HIP(hipMemcpyAsync(hflag, s.flags.get(), sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
            c->sync();
            if (hflag[0]) break;
        }
    }
    copy_d2d(c, x, s.X.get(), sizeof(double) * nk);
    std::vector<FcgState> st(k);
    c->d2h(st.data(), s.st.get(), k);
    SolveResult res;
    prec->last_niter.assign(k, 0);
    prec->last_exit.assign(k, 0);
    prec->last_rel.assign(k, 0);
    for (int j = 0; j < k; j++) {
        bool conv = st[j].pad0 != 0;
        // out->niter = iter + 1 + excnlim (CG_HS.cu:165-171): iter + 1 when it converged, else itnlim
        int niter = p.itnlim < 1 ? 0 : (conv ? st[j].iter + 1 : st[j].iter);
        double rel = st[j].delta0 > 0 ? st[j].l2 / st[j].delta0 : 0.0;
        prec->last_niter[j] = niter;
        prec->last_exit[j] = st[j].l2;
        prec->last_rel[j] = rel;
        res.niter = std::max(res.niter, niter);
        res.exit_res = std::max(res.exit_res, st[j].l2);
        res.rel_res = std::max(res.rel_res, rel);
        if (!conv) res.retv = -1;
        if (!std::isfinite(st[j].l2)) res.retv = BCMGX_ERR_NUMERIC;
    }
    res.del0 = st[0].delta0;
    res.seconds = now_s() - t0;
    return res;
}

}  // namespace bcmgx
