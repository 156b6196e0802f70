// kernels.h -- host-side launchers of the HIP kernels.  All launches go to
// the context stream; none synchronises unless it returns a host value.
#pragma once

#include "common.h"

namespace bcmgx {

// ---- allocation-free SpMM (noalloc.hip) ------------------------------------------
// device address of `device`'s copy of the fixed scratch array (taken once per
// device; `device` must be current), and the array's size in bytes
void* scratch_address(int device);
size_t scratch_bytes();
// bcmgx_spmm on device blocks with both operands staged in c->scratch
// (no allocation); lanes: lanes per row for k == 1
void spmm_noalloc(Context* c, const DCsr& A, const double* x, double* y, int k, int ld, double alpha, double beta, int op,
                  int kchunk, int lanes);

// ---- primitives (prims.hip) -------------------------------------------------
// out[0..n) = exclusive prefix sums of in, out[n] = total; returns the total.
int scan_exclusive(Context* c, const int* in, int* out, long long n);
long long scan_exclusive(Context* c, const long long* in, long long* out, long long n);
long long scan_exclusive_i32_to_i64(Context* c, const int* in, long long* out, long long n);
// the same, the total left in out[n] on the device (no host synchronisation)
void scan_exclusive_async(Context* c, const int* in, int* out, long long n);
void scan_exclusive_async(Context* c, const long long* in, long long* out, long long n);
template <typename T>
void fill(Context* c, T* p, long long n, T v);
template <typename A, typename B>
void convert(Context* c, const A* s, B* d, long long n);
void copy_d2d(Context* c, void* dst, const void* src, size_t bytes);
// bootstrap test vectors: x_i = 2 u_i - 1, u_i the top 53 bits of
// splitmix64(seed + (i + 1) * 0x9E3779B97F4A7C15) scaled to [0, 1)
void boot_random(Context* c, double* x, long long n, unsigned long long seed);
// x_i <- x_i / max_j |x_j| (the maximum is order-independent; x unchanged when it is 0)
void scale_by_absmax(Context* c, double* x, long long n);
// y_i <- y_i + x_i
void add_to(Context* c, double* y, const double* x, long long n, const int* stop);
double stream_copy_gbs(Context* c, size_t bytes, int reps);

// ---- sparse matrix-vector family (spmv.hip) ---------------------------------
// Vectors with k columns are interleaved row-major: x[i*k + c].
enum class SpOp {
    AX,      // y = A x                                   (CSR.cu:452-492, OP 1)
    RES,     // y = b - A x                               (OP 2 applied to a copy of b)
    AXPBY,   // y = alpha A x + beta y                    (OP 0)
    JACOBI   // y = ((-(A x) + b) / D) + x                (_jacobi_it_full<0>, l1jacobi.cu:344-379)
};

struct SpmvArgs {
    int n = 0;
    const int* row = nullptr;
    const int* col = nullptr;
    const void* val = nullptr;  // T*
    long long nnz = 0;
    const void* x = nullptr;    // V*
    void* y = nullptr;          // V*
    const void* b = nullptr;    // V*  (RES, JACOBI)
    const double* b64 = nullptr;  // RES with V = float on the pipelined k-column SpMM: b held in FP64,
                                  // rounded to V as it is read (in place of b)
    const void* D = nullptr;    // V*  (JACOBI)
    double alpha = 1.0, beta = 0.0;
    int k = 1;                  // number of vector columns
    int lanes = 0;              // lanes per row for k == 1 (0 = heuristic, < 0 = CSR-stream kernel)
    int kchunk = 0;             // k > 1: columns per launch (0 = all k, pipelined kernel; -1 = all k, one-pass
                                // kernel; -2 = all k, pipelined kernel with streamed matrix loads;
                                // -16-s = all k, pipelined kernel with rows dealt to the XCDs in
                                // super-groups of 2^s row groups)
    const int* stop = nullptr;  // device flag: skip the launch's work when *stop != 0
    double* y64 = nullptr;      // JACOBI with V = float: write the result widened to FP64 here (y is scratch)
    void* x0 = nullptr;         // AX: also write y / D here (V*; a first l1-Jacobi sweep from zero on y)
    const int* tu = nullptr;    // the matrix's row-tile index (DCsr::tu, tucol, tlidx), or null
    const int* tucol = nullptr;
    const unsigned short* tlidx = nullptr;
};
// the row-tile index of A into a (k = 1 launches whose lanes code asks for the tiled kernel use it)
inline void set_tiles(SpmvArgs& a, const DCsr& A) {
    a.tu = A.tu.get();
    a.tucol = A.tucol.get();
    a.tlidx = A.tlidx.get();
}
// build A's row-tile index (once: a built index is kept)
void build_tile_index(Context* c, DCsr& A);
// T = matrix value type, V = vector type: (double,double), (float,float), (float,double)
template <typename T, typename V>
void spmv(Context* c, SpOp op, const SpmvArgs& a);

// FCG: w = A p (FP64, k interleaved columns) fused with the per-column dot
// products (r.p, w.p, wold.p) written as partials[(blk*k + c)*3 + m] of
// `nblk` = spmm_dots_blocks(n, k) blocks.  Returns false (nothing launched)
// when the shape does not fit the fused kernel (k odd or < 4, column chunks,
// unaligned vectors); the caller then runs the SpMM and the dots apart.
int spmm_dots_blocks(Context* c, int n, int k);
bool spmm_dots(Context* c, const SpmvArgs& a, const double* r, const double* wold, double* partials, int nblk,
               const int* all_done);

// lanes-per-row heuristic from nnz/row (wave64 adaptation of CSR.cu:614-652)
int default_lanes(long long nnz, int n);

// x_i = f_i / D_i  (first l1-Jacobi sweep from a zero guess)
template <typename V>
void jacobi_from_zero(Context* c, int n, int k, const V* f, const V* D, V* x, const int* stop);
// mixed precision: F = (float) r; X = F / D   (level-0 entry of the FP32 cycle)
void jacobi_from_zero_f64in(Context* c, int n, int k, const double* r, const float* D, float* F, float* X, const int* stop);
// F = (double) f and the first sweep from zero X = F / D, both in FP64
void jacobi_from_zero_f32in(Context* c, int n, int k, const float* f, const double* D, double* F, double* X, const int* stop);
// x_i = p_i * xc[agg_i] + x_i   (prolongation, OP 0 with alpha = beta = 1)
template <typename V>
void prolong_add(Context* c, int n, int k, const int* agg, const V* pval, const V* xc, V* x, const int* stop);
template <typename A, typename B>
void convert_stop(Context* c, const A* s, B* d, long long n, const int* stop);
// the first post-smoothing sweep of a level from its product A P: a describes
// A P (n x nc) with a.x = the coarse correction e (nc x k), a.b = the
// pre-smoothing residual r_pre = f - A xpre, a.D = the l1 row sums; writes
// y = ((-(A P) e + r_pre) / D) + (p_i e[agg_i] + xpre), i.e. one l1-Jacobi sweep
// on f from xpre + P e.  y must differ from xpre and b.
template <typename V>
void galerkin_post_sweep(Context* c, const SpmvArgs& a, const int* agg, const V* pval, const V* xpre);

// the same with z (the SpMM's x) in FP32, widened exactly on load (four columns
// per lane); 0 / false unless k is a multiple of 4 in [8, 64]
int spmm_dots_z32_blocks(Context* c, int n, int k);
bool spmm_dots_z32(Context* c, const SpmvArgs& a, const double* r, const double* wold, double* partials, int nblk,
                   const int* all_done);

// ---- FCG kernels (fcg.hip) ----------------------------------------------------
struct FcgState {
    double delta0, delta, l2, alpha, beta, gamma, theta, theta2;
    double pend;  // FCG: theta2 of the first iteration of a pair, applied to x by the second
    int iter, done, pad0, pad1;  // FCG: pad1 = 1 while pend waits; CG: pad0 = converged
};
// column-major blocks (leading dimension ld) <-> k interleaved columns.
// zero_flag (may be null) is cleared when the block holds a nonzero; the block
// goes to alt instead of d when alt_if (may be null) is set and nonzero; st/p
// (may be null) add the waiting FCG x step (pad1, pend) of each column on the way out.
void colmajor_to_interleaved(Context* c, const double* s, int ld, double* d, int n, int k, int* zero_flag = nullptr,
                             const int* alt_if = nullptr, double* alt = nullptr);
void interleaved_to_colmajor(Context* c, const double* s, double* d, int ld, int n, int k, const FcgState* st = nullptr,
                             const double* p = nullptr);
struct FcgBufs {
    int n = 0, k = 1;
    int nblk = 0;              // partial-sum blocks (fixed for a given n)
    double* partials = nullptr;  // nblk * k * 3
    FcgState* st = nullptr;      // k states
    int* all_done = nullptr;     // 1 when every column stopped
    double rtol = 1e-6;
    int itnlim = 2000;
};
int fcg_partial_blocks(int n);
// ||v||^2 partials of each column -> st[c].l2 = sqrt(sum); mode 0: delta0 = l2 (initial)
void fcg_norm(Context* c, const FcgBufs& f, const double* v, int mode);
// Optional level-0 entry of an FP32 V-cycle written by the FCG updates: with
// the new residual r, F = (float) r and X = F / M (the first l1-Jacobi sweep
// from a zero guess), so the next preconditioner application starts there.
// F may be null: then only X is written and the cycle rounds r itself where it
// needs the level's right-hand side.
struct FcgEntry {
    float* F = nullptr;
    float* X = nullptr;
    const float* M = nullptr;
};
// fcg_norm of r that also writes the FP32 cycle's level-0 entry from r
void fcg_norm_entry(Context* c, const FcgBufs& f, const double* r, int mode, const FcgEntry& e);
// alpha = r.v, beta = w.v -> delta = beta, theta = alpha / delta  (FCG.cu:56-86)
void fcg_init_dots(Context* c, const FcgBufs& f, const double* r, const double* v, const double* w);
// the initial scalars from the partials of nblk blocks (alpha, beta in slots 0, 1)
void fcg_init_scalars(Context* c, const FcgBufs& f, int nblk);
// fcg_init_update with v the FP32 cycle result z, widened exactly
void fcg_init_update_z32(Context* c, const FcgBufs& f, double* x, double* r, const float* z, const double* w, double* d,
                         double* q, const FcgEntry& e = FcgEntry());
// x += theta v; r -= theta w  (FCG.cu:88-89); d = v, q = w
void fcg_init_update(Context* c, const FcgBufs& f, double* x, double* r, const double* v, const double* w, double* d, double* q,
                     const FcgEntry& e = FcgEntry());
// (alpha, beta, gamma) = (r.p0, w0.p0, w1.p0), then theta, delta, theta2 (FCG.cu:113-136)
void fcg_dots3(Context* c, const FcgBufs& f, const double* r, const double* p0, const double* w0, const double* w1, int phase);
// the scalar step of fcg_dots3 alone, over `nblk` partials already written.
// phase: 0 for the first iteration of a pair (its x step waits), 1 for the second
void fcg_scalars3(Context* c, const FcgBufs& f, int nblk, int phase);
// double_merged_axpy x2 (FCG.cu:138-144) fused with the ||r||^2 partials.  Phase 0
// leaves x alone (its step waits in pend); phase 1 adds both steps to x in order.
void fcg_update(Context* c, const FcgBufs& f, double* p0, const double* p1, double* w0, const double* w1, double* x, double* r,
                int phase, const FcgEntry& e = FcgEntry());
// ||r|| -> convergence test and iteration count (FCG.cu:146-156)
void fcg_finish_iteration(Context* c, const FcgBufs& f, int phase);
// the update with the preconditioned residual z in FP32 and the direction p and
// its image w updated in place: t = z - theta p, u = w0 - theta w; p = t, w = u
// (the same operations, in the same order, as fcg_update)
void fcg_update_z32(Context* c, const FcgBufs& f, const float* z, double* p, const double* w0, double* w, double* x, double* r,
                    int phase, const FcgEntry& e);
// a step still waiting after the last pair (every column stopped after its
// first iteration): x += pend t, t the first iteration's direction
// clears every column's waiting x step (after it was applied elsewhere)
void fcg_clear_pending(Context* c, const FcgBufs& f);
void fcg_flush_pending(Context* c, const FcgBufs& f, const double* t, double* x);

// ---- CG (Hestenes-Stiefel) kernels (fcg.hip), CG_HS.cu:24-172 ----------------
// ||r|| -> delta0 (||r|| for stop_criterion 1, else 1), iteration count 0
void cg_init(Context* c, const FcgBufs& f, const double* r, int stopcrit);
// rho = r.u, beta = rho / rho_prev, p = u + beta p
void cg_beta_direction(Context* c, const FcgBufs& f, const double* r, const double* u, double* p);
// gamma = p.s (first iteration also rho = u.r), alpha = rho / gamma
void cg_alpha(Context* c, const FcgBufs& f, const double* p, const double* s, const double* u, const double* r, bool first);
// x += alpha p, r -= alpha s, ||r|| -> convergence test, iteration count, itnlim
void cg_update_finish(Context* c, const FcgBufs& f, const double* p, const double* s, double* x, double* r);

// ---- block-vector kernels of the s-step solvers (krylov.hip) -------------------
// out[j] = A[:, j] . B[:, j] (B column j at B + j*ldb; ldb = 0: one vector), fixed-order sums;
// partials: nblk * m doubles
void block_dots(Context* c, int n, int m, const double* A, const double* B, long long ldb, double* partials, int nblk,
                double* out);
// C = C + A B (A n x s, B s x s device), E = E + f (C al)   (cdmmv)
void block_dmmv(Context* c, int n, int s, const double* A, const double* B, double* C, const double* al, double* E, double f);
// C = C + (A B)   (cdmm)
void block_dmm(Context* c, int n, int s, const double* A, const double* B, double* C);
// D = C + f (A b)   (cdmv2v)
void block_dmv2v(Context* c, int n, int s, const double* A, const double* bv, const double* C, double f, double* D);
// dst[i * ds] = src[i * ss] for i < n
void strided_copy(Context* c, double* dst, int ds, const double* src, int ss, int n);

// ---- setup kernels (setup.hip) -------------------------------------------------
// position of the first diagonal entry of each row (-1 if none)
void diag_index(Context* c, int n, const int* row, const int* col, int* dind);
// rows with no diagonal entry (dind < 0), counted on the device
int count_missing_diagonals(Context* c, int n, const int* dind);
// M_i = sum_j |a_ij| in column order (CSRm::absoluteRowSum, CSR.cu:1362-1397)
void abs_row_sum(Context* c, int n, const int* row, const double* val, double* M);
// l1 diagonal of the one-level preconditioner: a_ii + sum_{j!=i} |a_ij| (l1jacobi.cu:484-493)
void l1_diag(Context* c, int n, const int* row, const int* col, const double* val, double* d);

struct MatchWork;
// suitor(): makeAH + toMaximumProductMatrix + _write_T_warp + Suitor (matching.cu:253-288)
// dind_checked: A's diagonal positions, already checked to be present in every
// row (nullptr: found and checked here)
void suitor_match(Context* c, const DCsr& A, const double* w, int* M, const int* dind_checked = nullptr);
// makeP_GPU with canonical singleton order; returns the coarse size (npairs_out: the good pairs,
// aggregates 0 .. npairs-1; the coarse size minus it counts the singletons of nonzero weight)
int make_prolongator(Context* c, int n, const int* M, const double* w, int* agg, double* pval, int* npairs_out = nullptr);
// P^T of a pairwise aggregation whose every vertex is in a good pair or a singleton of nonzero
// weight (n = npairs + coarse size): at most 2 entries a row
void pair_restriction(Context* c, int n, int nc, int npairs, const int* agg, const int* M, const double* pval, DCsr& R);
// R = P^T for a one-entry-per-row P given as (agg, pval): rows sorted by fine index
void transpose_prolongator(Context* c, int n, int nc, const int* agg, const double* pval, DCsr& R);
// general transpose (CSRm::transpose / Transpose_local)
void transpose(Context* c, const DCsr& A, DCsr& T);
// C = A B, deterministic ascending-k accumulation, sorted columns
void spgemm(Context* c, const DCsr& A, const DCsr& B, DCsr& C);
// A.maxrow and A.unit_rows from A's row pointers on the device (one read-back)
void csr_shape(Context* c, DCsr& A);
// C = R (A P) for the prolongator P = (agg, pval) onto nc, the same bits as spgemm(R, spgemm(A, P));
// A P is not compacted when A's rows fit the one-launch product.  rmax: a bound on the entries of
// R's rows (-1: none known).  Returns true when it also wrote each row's diagonal position into
// dind (nc ints, when given)
bool galerkin(Context* c, const DCsr& A, int nc, const int* agg, const double* pval, const DCsr& R, DCsr& C, int rmax = -1,
              int* dind = nullptr);
// A P with FP32 values for the prolongator (agg, pval) onto nc columns, queued with no host
// synchronisation (C.nnz = -1: the length is at C.row[n]); false when A's rows are too long
bool unit_product_f32_async(Context* c, const DCsr& A, int nc, const int* agg, const double* pval, DCsr& C);
// P <- P P_i for one-entry-per-row operators: agg[k] = agg_i[agg[k]], pval[k] *= pval_i[agg[k]]
void compose_prolongators(Context* c, int n, int* agg, double* pval, const int* agg_i, const double* pval_i);
// csr of a one-entry-per-row prolongator (row = identity)
void prolongator_csr(Context* c, int n, int nc, const int* agg, const double* pval, DCsr& P);
// g[k] = -sum_c lambda[row(k), c] x[col(k), c]
void sparse_outer_grad(Context* c, const DCsr& A, const double* lambda, const double* x, int k, double* g);

// ---- AFSAI (afsai.hip) ------------------------------------------------------------
void afsai_build(Context* c, const DCsr& A, DCsr& G, DCsr& GT);

}  // namespace bcmgx
