// amg.h -- hierarchy, preconditioners and the FCG driver (host side).
#pragma once

#include <chrono>
#include <memory>
#include <vector>

#include "kernels.h"

namespace bcmgx {

struct Level {
    DCsr A;                 // owned (levels >= 1); level 0 borrows the user's matrix
    const DCsr* Aref = nullptr;
    DBuf<double> M;         // l1 row sums sum_j |a_ij| (relaxPrepare)
    DBuf<float> Mf;
    DBuf<int> agg;          // prolongator to the next level: one entry per row
    DBuf<double> pval;
    DBuf<float> pvalf;
    DCsr R;                 // restrictor P^T (rows sorted by fine index)
    DCsr AP;                // A P, FP32 values only (valf): the first post-smoothing sweep of an FP32 cycle
    int lanesA = 0, lanesR = 0, lanesAP = 0;
    int chunkA = 0;         // columns per SpMM launch for the current batch width
    const DCsr& mat() const { return Aref ? *Aref : A; }
};

struct Hierarchy {
    std::vector<std::unique_ptr<Level>> L;
    double op_cmplx = 1, op_wcmplx = 1, avg_cratio = 1;
    int maxcs = 0;
    int levels() const { return (int)L.size(); }
};

// cycle work vectors of one level, k columns interleaved
template <typename V>
struct LevelWork {
    DBuf<V> F, X, X2, Res;
};

struct SolveWork;

struct Prec {
    Context* ctx = nullptr;
    const DCsr* A = nullptr;
    bcmgx_params p{};
    int ptype = BCMGX_PREC_NONE;
    Hierarchy H;                                   // the first hierarchy (boot->H_array[0])
    std::vector<std::unique_ptr<Hierarchy>> HB;    // bootstrap hierarchies 1 .. n_hrc-1
    double boot_ratio = 1.0;                       // boot->estimated_ratio (1 until estimated)
    int n_hrc() const { return 1 + (int)HB.size(); }
    Hierarchy& hrc(int h) { return h == 0 ? H : *HB[h - 1]; }
    // one-level l1-Jacobi
    DBuf<double> pl1j;
    // AFSAI
    DCsr G, GT;
    int lanesG = 0, lanesGT = 0;
    // work for the preconditioner application, sized for `kcap` columns
    int kcap = 0;
    int tuned_k = 0;
    std::vector<LevelWork<double>> wd;
    std::vector<LevelWork<float>> wf;
    DBuf<double> cd_F, cd_X, cd_X2;  // FP64 coarsest-level buffers of the FP32 cycle
    DBuf<double> tmp;                // AFSAI intermediate / prec_apply staging
    DBuf<double> tmp2, tmp3;         // additive bootstrap composition: the sum and one hierarchy's correction
    double setup_seconds = 0;
    // solve workspaces per k (graphs are bound to their buffers)
    std::vector<std::unique_ptr<SolveWork>> solves;
    // per-column results of the last solve
    std::vector<int> last_niter;
    std::vector<double> last_exit, last_rel;

    ~Prec();
    void ensure_work(int k);
    void drop_graphs();
    // u <- B r for k interleaved columns; `x_zero`: u holds zeros on entry (the FCG case).
    // `entry_ready`: the level-0 entry of the FP32 cycle (see fused_entry) already
    // holds F = (float) r (unless the cycle reads r itself) and the first sweep X = F / M.
    void apply(const double* r, double* u, int k, bool x_zero, const int* stop, bool entry_ready = false);
    // one cycle of bootstrap hierarchy h alone: u <- cycle_h(r) from the guess u
    void apply_hrc(int h, const double* r, double* u, int k, bool x_zero, const int* stop, bool entry_ready = false);
    // the level-0 buffers an FCG update can fill for the next apply (FP32 BCMG
    // with at least two levels and a pre-smoothing sweep); false otherwise.  When
    // the level's right-hand side is read only by the residual after the first
    // sweep and that runs on the pipelined SpMM, F is left out (e->F = null) and
    // the cycle rounds r there itself (f0_from_r)
    bool fused_entry(int k, FcgEntry* e);
    bool f0_from_r = false;
    // the FP32 cycle from that entry for the residual r, its result z left in FP32
    // (k interleaved columns): the values the FP64 output would hold, before the exact widening
    void apply_entry_f32(const double* r, float* z, int k, const int* stop);
    // recompute what depends on level 0's values after they changed in place:
    // the l1 row sums M (and Mf) of BCMG, the l1 diagonal of one-level l1-Jacobi
    void refresh_level0();
};

std::unique_ptr<Prec> prec_setup(Context* c, const DCsr& A, const bcmgx_params& p);

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct SolveResult {
    int retv = 0, niter = 0;
    double exit_res = 0, rel_res = 0, del0 = 0, seconds = 0;
};
// FCG on k interleaved columns (device pointers b, x; x in/out)
// the caller's column-major device blocks (leading dimension ld), staged
// straight into and out of the solver's interleaved buffers
struct SolveIo {
    const double* b = nullptr;
    double* x = nullptr;
    int ld = 0;
};
// with io, b and x are ignored and the blocks come from io
SolveResult fcg_solve(Context* c, const DCsr& A, Prec* prec, const bcmgx_params& p, const double* b, double* x, int k,
                      const SolveIo* io = nullptr);
// CG (Hestenes-Stiefel) on k interleaved columns, same conventions
SolveResult cghs_solve(Context* c, const DCsr& A, Prec* prec, const bcmgx_params& p, const double* b, double* x, int k);
// s-step CG (Chronopoulos-Gear; p.solver CGS or CGS_CUBLAS) or its pipelined
// form (PIPELINED_CGS), s = p.sstep, the columns one after another
SolveResult cgs_solve(Context* c, const DCsr& A, Prec* prec, const bcmgx_params& p, const double* b, double* x, int k);

// autotune: fastest lanes per row for SpMV with A (k = 1)
int tune_lanes(Context* c, const DCsr& A, bool fp32);
// the lanes tune_lanes chose for A if it ran, else the nnz/row heuristic (no timing, no allocation)
int cached_lanes(Context* c, const DCsr& A, bool fp32);
// autotune: fastest column chunk for SpMM with A on k interleaved columns
int tune_chunk(Context* c, const DCsr& A, bool fp32, int k);

}  // namespace bcmgx
