// capi.cpp -- the extern "C" boundary (include/bcmgx.h).  Every entry point
// catches bcmgx::Error / std::exception and returns a status code; the
// message is kept on the context for bcmgx_last_error().
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "amg.h"

using namespace bcmgx;

struct bcmgx_matrix_s;
struct bcmgx_prec_s;
// A context owns the device memory of every handle created on it.  Destroying
// the context releases that memory and orphans the handles (ctx = nullptr);
// their own destroy calls then only free the host structures.  Destroying a
// matrix orphans the preconditioners set up on it the same way (their level 0
// is that matrix).  Any destruction order is therefore safe, and an orphaned
// handle answers BCMGX_ERR_STATE.
struct bcmgx_context_s {
    Context* c = nullptr;
    std::set<bcmgx_matrix_s*> mats;
    std::set<bcmgx_prec_s*> precs;
};
struct bcmgx_matrix_s {
    bcmgx_context_s* ctx = nullptr;
    DCsr A;
};
struct bcmgx_prec_s {
    bcmgx_context_s* ctx = nullptr;
    bcmgx_matrix_s* A = nullptr;
    std::unique_ptr<Prec> P;
};

static std::string g_last_error;
static std::mutex g_err_mu;

static int record(bcmgx_context_s* ctx, int code, const char* msg) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_error = msg;
    if (ctx && ctx->c) ctx->c->last_error = msg;
    return code;
}

#define API_BEGIN try {
#define API_END(ctxp)                                                  \
    }                                                                  \
    catch (const bcmgx::Error& e) { return record(ctxp, e.code, e.what()); } \
    catch (const std::bad_alloc&) { return record(ctxp, BCMGX_ERR_ALLOC, "host allocation failed"); } \
    catch (const std::exception& e) { return record(ctxp, BCMGX_ERR_ARG, e.what()); }

// copy an n x k block (column-major, leading dimension ld; host or device)
// into a device buffer with k interleaved columns
static void stage_in(Context* c, const double* src, int ld, int mem, int n, int k, double* dst_interleaved, DBuf<double>& tmp) {
    size_t nk = (size_t)n * k;
    if (k == 1) {
        if (mem == BCMGX_MEM_HOST) BCMGX_HIP(hipMemcpyAsync(dst_interleaved, src, nk * 8, hipMemcpyHostToDevice, c->stream));
        else copy_d2d(c, dst_interleaved, src, nk * 8);
        return;
    }
    const double* dsrc = src;
    if (mem == BCMGX_MEM_HOST) {
        tmp.reset(c, (size_t)ld * k);
        BCMGX_HIP(hipMemcpyAsync(tmp.get(), src, (size_t)ld * k * 8, hipMemcpyHostToDevice, c->stream));
        dsrc = tmp.get();
    }
    colmajor_to_interleaved(c, dsrc, ld, dst_interleaved, n, k);
}
static void stage_out(Context* c, const double* src_interleaved, double* dst, int ld, int mem, int n, int k, DBuf<double>& tmp) {
    size_t nk = (size_t)n * k;
    if (k == 1) {
        if (mem == BCMGX_MEM_HOST) {
            BCMGX_HIP(hipMemcpyAsync(dst, src_interleaved, nk * 8, hipMemcpyDeviceToHost, c->stream));
            c->sync();
        } else copy_d2d(c, dst, src_interleaved, nk * 8);
        return;
    }
    if (mem == BCMGX_MEM_HOST) {
        tmp.reset(c, (size_t)ld * k);
        if (ld > n) {  // keep the caller's padding rows: fetch them first
            BCMGX_HIP(hipMemcpyAsync(tmp.get(), dst, (size_t)ld * k * 8, hipMemcpyHostToDevice, c->stream));
        }
        interleaved_to_colmajor(c, src_interleaved, tmp.get(), ld, n, k);
        BCMGX_HIP(hipMemcpyAsync(dst, tmp.get(), (size_t)ld * k * 8, hipMemcpyDeviceToHost, c->stream));
        c->sync();
    } else {
        interleaved_to_colmajor(c, src_interleaved, dst, ld, n, k);
    }
}

extern "C" {

const char* bcmgx_error_string(int s) {
    switch (s) {
    case BCMGX_OK: return "ok";
    case BCMGX_ERR_ARG: return "invalid argument";
    case BCMGX_ERR_HIP: return "HIP runtime error";
    case BCMGX_ERR_ALLOC: return "device allocation failed";
    case BCMGX_ERR_IO: return "file could not be opened or parsed";
    case BCMGX_ERR_PARAM: return "invalid settings value";
    case BCMGX_ERR_UNSUPPORTED: return "unsupported option";
    case BCMGX_ERR_NUMERIC: return "non-finite value or breakdown";
    case BCMGX_ERR_STATE: return "handle in the wrong state";
    default: return "unknown status";
    }
}

const char* bcmgx_last_error(bcmgx_context ctx) {
    if (ctx && ctx->c) return ctx->c->last_error.c_str();
    return g_last_error.c_str();
}

int bcmgx_context_create(int device, bcmgx_context* out) {
    if (!out) return BCMGX_ERR_ARG;
    *out = nullptr;
    API_BEGIN
    int cnt = 0;
    BCMGX_HIP(hipGetDeviceCount(&cnt));
    BCMGX_CHECK(device >= 0 && device < cnt, BCMGX_ERR_ARG, "no such HIP device");
    auto* h = new bcmgx_context_s();
    try {
        h->c = new Context(device);
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    return BCMGX_OK;
    API_END(nullptr)
}

int bcmgx_context_destroy(bcmgx_context ctx) {
    if (!ctx) return BCMGX_OK;
    API_BEGIN
    (void)hipSetDevice(ctx->c->device);
    ctx->c->sync();
    for (auto* p : ctx->precs) {
        p->P.reset();
        p->ctx = nullptr;
    }
    for (auto* m : ctx->mats) {
        m->A = DCsr();
        m->ctx = nullptr;
    }
    delete ctx->c;
    delete ctx;
    return BCMGX_OK;
    API_END(nullptr)
}

int bcmgx_context_synchronize(bcmgx_context ctx) {
    if (!ctx) return BCMGX_ERR_ARG;
    API_BEGIN
    ctx->c->sync();
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_device_alloc(bcmgx_context ctx, size_t bytes, void** ptr) {
    if (!ctx || !ptr) return BCMGX_ERR_ARG;
    *ptr = nullptr;
    API_BEGIN
    BCMGX_HIP(hipSetDevice(ctx->c->device));
    *ptr = ctx->c->pool->alloc(bytes ? bytes : 1);
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_device_free(bcmgx_context ctx, void* ptr) {
    if (!ctx) return BCMGX_ERR_ARG;
    API_BEGIN
    ctx->c->sync();
    ctx->c->pool->free(ptr);
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_memcpy(bcmgx_context ctx, void* dst, const void* src, size_t bytes, int kind) {
    if (!ctx || (bytes && (!dst || !src)) || kind < 0 || kind > 2) return BCMGX_ERR_ARG;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    if (bytes) BCMGX_HIP(hipMemcpyAsync(dst, src, bytes, k, c->stream));
    c->sync();
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_csr_create(bcmgx_context ctx, int n, int m, int nnz, const int* row, const int* col, const double* val, int mem,
                     bcmgx_matrix* out) {
    if (!ctx || !out || n < 0 || m < 0 || nnz < 0 || !row || (nnz > 0 && (!col || !val))) return BCMGX_ERR_ARG;
    *out = nullptr;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    auto* h = new bcmgx_matrix_s();
    h->ctx = ctx;
    try {
        h->A.alloc(c, n, m, nnz);
        hipMemcpyKind k = mem == BCMGX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
        BCMGX_HIP(hipMemcpyAsync(h->A.row.get(), row, sizeof(int) * ((size_t)n + 1), k, c->stream));
        if (nnz > 0) {
            BCMGX_HIP(hipMemcpyAsync(h->A.col.get(), col, sizeof(int) * (size_t)nnz, k, c->stream));
            BCMGX_HIP(hipMemcpyAsync(h->A.val.get(), val, sizeof(double) * (size_t)nnz, k, c->stream));
        }
        // shape facts the kernels use: the longest row, one entry per row
        if (mem == BCMGX_MEM_HOST) {
            int mx = 0;
            for (int i = 0; i < n; i++) mx = std::max(mx, row[i + 1] - row[i]);
            h->A.maxrow = mx;
            if (nnz == n && n > 0) {
                bool unit = true;
                for (int i = 0; i <= n && unit; i++) unit = row[i] == i;
                h->A.unit_rows = unit;
            }
            c->sync();
        } else {
            csr_shape(c, h->A);  // reads back after the copies (same stream)
        }
    } catch (...) {
        delete h;
        throw;
    }
    ctx->mats.insert(h);
    *out = h;
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_csr_destroy(bcmgx_matrix A) {
    if (!A) return BCMGX_OK;
    API_BEGIN
    if (A->ctx) {
        (void)hipSetDevice(A->ctx->c->device);
        A->ctx->c->sync();
        auto& precs = A->ctx->precs;
        for (auto it = precs.begin(); it != precs.end();) {
            bcmgx_prec_s* p = *it;
            if (p->A == A) {
                p->P.reset();
                p->A = nullptr;
                p->ctx = nullptr;
                it = precs.erase(it);
            } else {
                ++it;
            }
        }
        A->ctx->mats.erase(A);
    }
    delete A;
    return BCMGX_OK;
    API_END(nullptr)
}

int bcmgx_csr_info(bcmgx_matrix A, int* n, int* m, int* nnz) {
    if (!A) return BCMGX_ERR_ARG;
    if (n) *n = A->A.n;
    if (m) *m = A->A.m;
    if (nnz) *nnz = A->A.nnz;
    return BCMGX_OK;
}

int bcmgx_csr_analyse(bcmgx_matrix A, double* seconds) {
    if (!A) return BCMGX_ERR_ARG;
    if (!A->ctx) return BCMGX_ERR_STATE;
    API_BEGIN
    Context* c = A->ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    c->sync();
    double t0 = now_s();
    build_tile_index(c, A->A);
    c->sync();
    if (seconds) *seconds = now_s() - t0;
    return BCMGX_OK;
    API_END(A->ctx)
}

int bcmgx_csr_update_values(bcmgx_matrix A, const double* val, int mem) {
    if (!A || (A->A.nnz > 0 && !val)) return BCMGX_ERR_ARG;
    if (!A->ctx) return BCMGX_ERR_STATE;
    API_BEGIN
    Context* c = A->ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    DCsr& M = A->A;
    if (M.nnz > 0) {
        hipMemcpyKind k = mem == BCMGX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
        BCMGX_HIP(hipMemcpyAsync(M.val.get(), val, sizeof(double) * (size_t)M.nnz, k, c->stream));
        // an FP32 copy (built by an FP32 hierarchy on this matrix) follows the new values
        if (M.valf) convert<double, float>(c, M.val.get(), M.valf.get(), M.nnz);
        // so do the level-0 divisors of the preconditioners set up on it
        for (auto* p : A->ctx->precs)
            if (p->A == A && p->P) p->P->refresh_level0();
    }
    c->sync();
    return BCMGX_OK;
    API_END(A->ctx)
}

int bcmgx_csr_copy_to_host(bcmgx_matrix A, int* row, int* col, double* val) {
    if (!A) return BCMGX_ERR_ARG;
    if (!A->ctx) return BCMGX_ERR_STATE;
    API_BEGIN
    Context* c = A->ctx->c;
    if (row) c->d2h(row, A->A.row.get(), (size_t)A->A.n + 1);
    if (col && A->A.nnz) c->d2h(col, A->A.col.get(), A->A.nnz);
    if (val && A->A.nnz) c->d2h(val, A->A.val.get(), A->A.nnz);
    return BCMGX_OK;
    API_END(A->ctx)
}

int bcmgx_prec_setup(bcmgx_context ctx, bcmgx_matrix A, const bcmgx_params* p, bcmgx_prec* out) {
    if (!ctx || !A || !p || !out) return BCMGX_ERR_ARG;
    *out = nullptr;
    API_BEGIN
    BCMGX_HIP(hipSetDevice(ctx->c->device));
    BCMGX_CHECK(p->preconditioner >= 0 && p->preconditioner < BCMGX_PREC_INVALID, BCMGX_ERR_PARAM,
                "Unhandled value for enum class PreconditionerType in prec_setup()");
    BCMGX_CHECK(A->A.n > 0 && A->A.n == A->A.m, BCMGX_ERR_ARG, "the preconditioner needs a nonempty square matrix");
    auto* h = new bcmgx_prec_s();
    h->ctx = ctx;
    h->A = A;
    try {
        h->P = prec_setup(ctx->c, A->A, *p);
    } catch (...) {
        delete h;
        throw;
    }
    ctx->precs.insert(h);
    *out = h;
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_prec_apply(bcmgx_prec prec, const double* r, double* u, int nrhs, int ld, int mem) {
    if (!prec || !r || !u || nrhs < 1 || nrhs > 64) return BCMGX_ERR_ARG;
    if (!prec->ctx) return BCMGX_ERR_STATE;
    API_BEGIN
    Context* c = prec->ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    int n = prec->A->A.n;
    if (ld < n) ld = n;
    size_t nk = (size_t)n * nrhs;
    DBuf<double> R(c, std::max<size_t>(nk, 1)), U(c, std::max<size_t>(nk, 1)), tmp;
    stage_in(c, r, ld, mem, n, nrhs, R.get(), tmp);
    stage_in(c, u, ld, mem, n, nrhs, U.get(), tmp);
    prec->P->apply(R.get(), U.get(), nrhs, false, nullptr);
    stage_out(c, U.get(), u, ld, mem, n, nrhs, tmp);
    c->sync();
    return BCMGX_OK;
    API_END(prec->ctx)
}

int bcmgx_prec_finalize(bcmgx_prec prec) {
    if (!prec) return BCMGX_OK;
    API_BEGIN
    if (prec->ctx) {
        (void)hipSetDevice(prec->ctx->c->device);
        prec->ctx->c->sync();
        prec->ctx->precs.erase(prec);
    }
    delete prec;
    return BCMGX_OK;
    API_END(nullptr)
}

int bcmgx_prec_get_info(bcmgx_prec prec, bcmgx_prec_info* info) {
    if (!prec || !info) return BCMGX_ERR_ARG;
    if (!prec->ctx) return BCMGX_ERR_STATE;
    std::memset(info, 0, sizeof(*info));
    Prec* P = prec->P.get();
    info->ptype = P->ptype;
    info->num_levels = P->ptype == BCMGX_PREC_BCMG ? P->H.levels() : 1;
    info->op_cmplx = P->H.op_cmplx;
    info->op_wcmplx = P->H.op_wcmplx;
    info->avg_cratio = P->H.avg_cratio;
    info->max_coarse_size = P->H.maxcs;
    info->setup_seconds = P->setup_seconds;
    info->g_nnz = P->G.nnz;
    info->num_hrc = P->ptype == BCMGX_PREC_BCMG ? P->n_hrc() : 0;
    info->estimated_ratio = P->boot_ratio;
    return BCMGX_OK;
}

int bcmgx_prec_hrc_levels(bcmgx_prec prec, int h, int* num_levels) {
    if (!prec || !num_levels) return BCMGX_ERR_ARG;
    if (!prec->ctx) return BCMGX_ERR_STATE;
    Prec* P = prec->P.get();
    if (P->ptype != BCMGX_PREC_BCMG || h < 0 || h >= P->n_hrc()) return BCMGX_ERR_ARG;
    *num_levels = P->hrc(h).levels();
    return BCMGX_OK;
}

int bcmgx_prec_hrc_level_size(bcmgx_prec prec, int h, int level, int* n, int* nnz) {
    if (!prec || level < 0) return BCMGX_ERR_ARG;
    if (!prec->ctx) return BCMGX_ERR_STATE;
    Prec* P = prec->P.get();
    if (P->ptype != BCMGX_PREC_BCMG || h < 0 || h >= P->n_hrc() || level >= P->hrc(h).levels()) return BCMGX_ERR_ARG;
    const DCsr& A = P->hrc(h).L[level]->mat();
    if (n) *n = A.n;
    if (nnz) *nnz = A.nnz;
    return BCMGX_OK;
}

int bcmgx_prec_level_size(bcmgx_prec prec, int level, int* n, int* nnz) {
    return bcmgx_prec_hrc_level_size(prec, 0, level, n, nnz);
}

int bcmgx_prec_level_matrix(bcmgx_prec prec, int level, int* row, int* col, double* val) {
    return bcmgx_prec_hrc_level_matrix(prec, 0, level, row, col, val);
}

int bcmgx_prec_level_prolongator(bcmgx_prec prec, int level, int* agg, double* pval) {
    return bcmgx_prec_hrc_level_prolongator(prec, 0, level, agg, pval);
}

int bcmgx_prec_hrc_level_matrix(bcmgx_prec prec, int h, int level, int* row, int* col, double* val) {
    if (!prec || level < 0) return BCMGX_ERR_ARG;
    if (!prec->ctx) return BCMGX_ERR_STATE;
    API_BEGIN
    Prec* P = prec->P.get();
    BCMGX_CHECK(P->ptype == BCMGX_PREC_BCMG && h >= 0 && h < P->n_hrc() && level < P->hrc(h).levels(), BCMGX_ERR_ARG,
                "no such level");
    const DCsr& A = P->hrc(h).L[level]->mat();
    Context* c = prec->ctx->c;
    if (row) c->d2h(row, A.row.get(), (size_t)A.n + 1);
    if (col && A.nnz) c->d2h(col, A.col.get(), A.nnz);
    if (val && A.nnz) c->d2h(val, A.val.get(), A.nnz);
    return BCMGX_OK;
    API_END(prec->ctx)
}

int bcmgx_prec_hrc_level_prolongator(bcmgx_prec prec, int h, int level, int* agg, double* pval) {
    if (!prec || level < 0) return BCMGX_ERR_ARG;
    if (!prec->ctx) return BCMGX_ERR_STATE;
    API_BEGIN
    Prec* P = prec->P.get();
    BCMGX_CHECK(P->ptype == BCMGX_PREC_BCMG && h >= 0 && h < P->n_hrc() && level < P->hrc(h).levels() - 1, BCMGX_ERR_ARG,
                "no prolongator at this level");
    Level& l = *P->hrc(h).L[level];
    Context* c = prec->ctx->c;
    int n = l.mat().n;
    if (agg) c->d2h(agg, l.agg.get(), n);
    if (pval) c->d2h(pval, l.pval.get(), n);
    return BCMGX_OK;
    API_END(prec->ctx)
}

int bcmgx_prec_afsai_factor(bcmgx_prec prec, int* n, int* nnz, int* row, int* col, double* val) {
    if (!prec) return BCMGX_ERR_ARG;
    if (!prec->ctx) return BCMGX_ERR_STATE;
    API_BEGIN
    Prec* P = prec->P.get();
    BCMGX_CHECK(P->ptype == BCMGX_PREC_AFSAI, BCMGX_ERR_STATE, "not an AFSAI preconditioner");
    if (n) *n = P->G.n;
    if (nnz) *nnz = P->G.nnz;
    Context* c = prec->ctx->c;
    if (row) c->d2h(row, P->G.row.get(), (size_t)P->G.n + 1);
    if (col && P->G.nnz) c->d2h(col, P->G.col.get(), P->G.nnz);
    if (val && P->G.nnz) c->d2h(val, P->G.val.get(), P->G.nnz);
    return BCMGX_OK;
    API_END(prec->ctx)
}

int bcmgx_solve(bcmgx_context ctx, bcmgx_matrix A, bcmgx_prec prec, const bcmgx_params* p, const double* b, double* x, int nrhs,
                int ld, int mem, bcmgx_solver_out* out) {
    if (!ctx || !A || !p || !b || !x || nrhs < 1 || nrhs > 64) return BCMGX_ERR_ARG;
    if (A->ctx != ctx || (prec && prec->ctx != ctx)) return BCMGX_ERR_STATE;
    API_BEGIN
    // solve() dispatches on the solver type (src/solver/solve.cu:19-40).  CGS_CUBLAS
    // (cgs_cublas/CGscublas.cu) is the s-step method with cuBLAS doing the block
    // products: the same s-step kernels serve it.
    BCMGX_CHECK(p->solver >= BCMGX_SOLVER_CGHS && p->solver <= BCMGX_SOLVER_CGS_CUBLAS, BCMGX_ERR_UNSUPPORTED,
                "Unhandled value for enum class SolverType");
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    int n = A->A.n;
    BCMGX_CHECK(n > 0 && A->A.m == n, BCMGX_ERR_ARG, "the solver needs a nonempty square matrix");
    if (ld < n) ld = n;
    std::unique_ptr<Prec> own;
    Prec* P = prec ? prec->P.get() : nullptr;
    if (!P || P->A != &A->A) {
        // solve with no preconditioner handle: PreconditionerType::NONE
        BCMGX_CHECK(!prec, BCMGX_ERR_ARG, "preconditioner was set up for another matrix");
        bcmgx_params q = *p;
        q.preconditioner = BCMGX_PREC_NONE;
        own = prec_setup(c, A->A, q);
        P = own.get();
    }
    SolveResult r;
    if (p->solver == BCMGX_SOLVER_FCG && mem != BCMGX_MEM_HOST && nrhs > 1 && n > 0) {
        // device blocks: FCG stages them into its own buffers
        SolveIo io;
        io.b = b;
        io.x = x;
        io.ld = ld;
        r = fcg_solve(c, A->A, P, *p, nullptr, nullptr, nrhs, &io);
    } else {
        size_t nk = (size_t)std::max(n, 1) * nrhs;
        DBuf<double> B(c, nk), X(c, nk), tmp;
        stage_in(c, b, ld, mem, n, nrhs, B.get(), tmp);
        stage_in(c, x, ld, mem, n, nrhs, X.get(), tmp);
        r = p->solver == BCMGX_SOLVER_CGHS  ? cghs_solve(c, A->A, P, *p, B.get(), X.get(), nrhs)
            : p->solver != BCMGX_SOLVER_FCG ? cgs_solve(c, A->A, P, *p, B.get(), X.get(), nrhs)
                                            : fcg_solve(c, A->A, P, *p, B.get(), X.get(), nrhs);
        stage_out(c, X.get(), x, ld, mem, n, nrhs, tmp);
    }
    c->sync();
    if (out) {
        out->retv = r.retv;
        out->niter = r.niter;
        out->exit_res = r.exit_res;
        out->rel_res = r.rel_res;
        out->del0 = r.del0;
        out->solve_seconds = r.seconds;
        out->full_n = n;
        out->local_n = n;
    }
    if (prec && own == nullptr) {
        // results kept on the handle for bcmgx_solve_columns
    }
    return r.retv == BCMGX_ERR_NUMERIC ? record(ctx, BCMGX_ERR_NUMERIC, "non-finite residual norm") : BCMGX_OK;
    API_END(ctx)
}

int bcmgx_solve_columns(bcmgx_prec prec, int nrhs, int* niter, double* exit_res, double* rel_res) {
    if (!prec) return BCMGX_ERR_ARG;
    if (!prec->ctx) return BCMGX_ERR_STATE;
    Prec* P = prec->P.get();
    if ((int)P->last_niter.size() < nrhs) return BCMGX_ERR_STATE;
    for (int j = 0; j < nrhs; j++) {
        if (niter) niter[j] = P->last_niter[j];
        if (exit_res) exit_res[j] = P->last_exit[j];
        if (rel_res) rel_res[j] = P->last_rel[j];
    }
    return BCMGX_OK;
}

// SolverOut::dump (src/solver/SolverOut.cu:9-97) + Hierarchy::printInfo (AMG.cu:167-180)
int bcmgx_dump_info(const char* path, const bcmgx_params* p, bcmgx_prec prec, const bcmgx_solver_out* out) {
    if (!path || !p || !out) return BCMGX_ERR_ARG;
    FILE* fp = std::fopen(path, "w");
    if (!fp) return BCMGX_ERR_IO;
    const char* solver = bcmgx_enum_to_string("solver", p->solver);
    std::fprintf(fp, "nprocs                                       : %d\n", 1);
    std::fprintf(fp, "Solver                                       : %s\n", solver);
    std::fprintf(fp, "Tolerance                                    : %g\n", p->rtol);
    std::fprintf(fp, "Preconditioner                               : %s\n", bcmgx_enum_to_string("preconditioner", p->preconditioner));
    if (p->preconditioner == BCMGX_PREC_L1_JACOBI)
        std::fprintf(fp, "l1-jacobi sweeps                             : %d\n", p->l1jacsweeps);
    if (out->retv < 0) std::fprintf(fp, "*** Exceeded max iter ***\n");
    if (out->retv > 0) std::fprintf(fp, "*** Non positive definite matrix found ***\n");
    std::fprintf(fp, "Total iterations                             : %d\n", out->niter);
    std::fprintf(fp, "Global dimension (number of rows)            : %d\n", out->full_n);
    std::fprintf(fp, "Local dimension (number of rows)             : %d\n", out->local_n);
    std::fprintf(fp, "Initial residual                             : %.10lf\n", out->del0);
    std::fprintf(fp, "Final residual                               : %.10lf\n", out->exit_res);
    if (p->preconditioner == BCMGX_PREC_BCMG && prec && prec->P) {
        Prec* P = prec->P.get();
        std::fprintf(fp, "Number of hierarchies                        : %d\n", P->n_hrc());
        std::fprintf(fp, "Estimated ratio                              : %f\n", P->boot_ratio);
        std::fprintf(fp, "max_hrc                                      : %d\n", p->max_hrc);
        std::fprintf(fp, "conv_ratio                                   : %f\n", p->conv_ratio);
        std::fprintf(fp, "bootstrap_composition_type                   : %s\n", bcmgx_enum_to_string("bootstrap_type", p->bootstrap_type));
        std::fprintf(fp, "solver_it                                    : %d\n", 15);
        std::fprintf(fp, "ncycle_type                                  : %s\n", bcmgx_enum_to_string("cycle_type", p->cycle_type));
        std::fprintf(fp, "relax_type                                   : %s\n", bcmgx_enum_to_string("relax_type", p->relax_type));
        std::fprintf(fp, "relaxnumber_coarse                           : %d\n", p->relaxnumber_coarse);
        std::fprintf(fp, "prerelax_number                              : %d\n", p->prerelax_sweeps);
        std::fprintf(fp, "postrelax_number                             : %d\n", p->postrelax_sweeps);
        std::fprintf(fp, "relax_weight                                 : %f\n", 1.0);
        std::fprintf(fp, "Number of levels                             : %d\n", P->H.levels());
        for (int i = 0; i < P->H.levels(); i++) {
            const DCsr& A = P->H.L[i]->mat();
            std::fprintf(fp, "A%-44d: n: %d nnz: %d avg_nnz: %f\n", i, A.n, A.nnz, (double)((float)A.nnz / (float)A.n));
        }
        std::fprintf(fp, "Current cmplx for V-cycle                    : %lf\n", P->H.op_cmplx);
        std::fprintf(fp, "Current cmplx for W-cycle                    : %lf\n", P->H.op_wcmplx);
        std::fprintf(fp, "Average Coarsening Ratio                     : %lf\n", P->H.avg_cratio);
    }
    std::fclose(fp);
    return BCMGX_OK;
}

// ------------------------------------------------------ kernel-level API ---
int bcmgx_spmv(bcmgx_context ctx, bcmgx_matrix A, const double* x, double* y, double alpha, double beta, int op, int lanes, int mem) {
    if (!ctx || !A || !x || !y || op < 0 || op > 2) return BCMGX_ERR_ARG;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    DCsr& M = A->A;
    if (lanes < 0 && (-lanes) % 1000 >= 400) build_tile_index(c, M);  // the row-tile kernel asked for
    DBuf<double> X(c, std::max(M.m, 1)), Y(c, std::max(M.n, 1));
    hipMemcpyKind kin = mem == BCMGX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    BCMGX_HIP(hipMemcpyAsync(X.get(), x, sizeof(double) * M.m, kin, c->stream));
    if (op != 1) BCMGX_HIP(hipMemcpyAsync(Y.get(), y, sizeof(double) * M.n, kin, c->stream));
    SpmvArgs a;
    a.n = M.n; a.row = M.row.get(); a.col = M.col.get(); a.val = M.val.get(); a.nnz = M.nnz;
    set_tiles(a, M);
    a.x = X.get(); a.y = Y.get(); a.b = Y.get(); a.alpha = alpha; a.beta = beta; a.lanes = lanes;
    SpOp o = op == 1 ? SpOp::AX : (op == 2 ? SpOp::RES : SpOp::AXPBY);
    if (lanes == 0) a.lanes = tune_lanes(c, M, false);
    spmv<double, double>(c, o, a);
    hipMemcpyKind kout = mem == BCMGX_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    BCMGX_HIP(hipMemcpyAsync(y, Y.get(), sizeof(double) * M.n, kout, c->stream));
    c->sync();
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_spgemm(bcmgx_context ctx, bcmgx_matrix A, bcmgx_matrix B, bcmgx_matrix* C) {
    if (!ctx || !A || !B || !C) return BCMGX_ERR_ARG;
    *C = nullptr;
    API_BEGIN
    BCMGX_HIP(hipSetDevice(ctx->c->device));
    auto* h = new bcmgx_matrix_s();
    h->ctx = ctx;
    try {
        spgemm(ctx->c, A->A, B->A, h->A);
    } catch (...) {
        delete h;
        throw;
    }
    ctx->mats.insert(h);
    *C = h;
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_transpose(bcmgx_context ctx, bcmgx_matrix A, bcmgx_matrix* T) {
    if (!ctx || !A || !T) return BCMGX_ERR_ARG;
    *T = nullptr;
    API_BEGIN
    BCMGX_HIP(hipSetDevice(ctx->c->device));
    auto* h = new bcmgx_matrix_s();
    h->ctx = ctx;
    try {
        transpose(ctx->c, A->A, h->A);
        ctx->c->sync();
    } catch (...) {
        delete h;
        throw;
    }
    ctx->mats.insert(h);
    *T = h;
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_matching(bcmgx_context ctx, bcmgx_matrix A, const double* w, int* M) {
    if (!ctx || !A || !w || !M) return BCMGX_ERR_ARG;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    int n = A->A.n;
    DBuf<double> W(c, std::max(n, 1));
    DBuf<int> DM(c, std::max(n, 1));
    c->h2d(W.get(), w, n);
    suitor_match(c, A->A, W.get(), DM.get());
    c->d2h(M, DM.get(), n);
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_make_prolongator(bcmgx_context ctx, int n, const int* M, const double* w, int* agg, double* pval, int* nc) {
    if (!ctx || n < 0 || !M || !w || !agg || !pval || !nc) return BCMGX_ERR_ARG;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    DBuf<int> DM(c, std::max(n, 1)), DA(c, std::max(n, 1));
    DBuf<double> DW(c, std::max(n, 1)), DP(c, std::max(n, 1));
    c->h2d(DM.get(), M, n);
    c->h2d(DW.get(), w, n);
    *nc = make_prolongator(c, n, DM.get(), DW.get(), DA.get(), DP.get());
    c->d2h(agg, DA.get(), n);
    c->d2h(pval, DP.get(), n);
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_spmm(bcmgx_context ctx, bcmgx_matrix A, const double* x, double* y, int k, int ld, double alpha,
               double beta, int op, int kchunk, int mem) {
    if (!ctx || !A || !x || !y || op < 0 || op > 2 || k < 1 || k > 64 || (kchunk < -2 && (kchunk > -16 || kchunk < -28))) return BCMGX_ERR_ARG;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    const DCsr& M = A->A;
    if (ld < std::max(M.n, M.m)) throw Error(BCMGX_ERR_ARG, "leading dimension smaller than the matrix");
    DBuf<double> X(c, (size_t)std::max(M.m, 1) * k), Y(c, (size_t)std::max(M.n, 1) * k), tmp;
    stage_in(c, x, ld, mem, M.m, k, X.get(), tmp);
    if (op != 1) stage_in(c, y, ld, mem, M.n, k, Y.get(), tmp);
    SpmvArgs a;
    a.n = M.n; a.row = M.row.get(); a.col = M.col.get(); a.val = M.val.get(); a.nnz = M.nnz;
    set_tiles(a, M);
    a.x = X.get(); a.y = Y.get(); a.b = Y.get(); a.alpha = alpha; a.beta = beta; a.k = k; a.kchunk = kchunk;
    a.lanes = k == 1 ? tune_lanes(c, M, false) : 0;
    SpOp o = op == 1 ? SpOp::AX : (op == 2 ? SpOp::RES : SpOp::AXPBY);
    spmv<double, double>(c, o, a);
    stage_out(c, Y.get(), y, ld, mem, M.n, k, tmp);
    c->sync();
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_spmm_noalloc(bcmgx_context ctx, bcmgx_matrix A, const double* x, double* y, int k, int ld, double alpha,
                       double beta, int op, int kchunk) {
    if (!ctx || !A || !x || !y || op < 0 || op > 2 || k < 1 || k > 64 || (kchunk < -2 && (kchunk > -16 || kchunk < -28)))
        return BCMGX_ERR_ARG;
    if (A->ctx != ctx) return BCMGX_ERR_STATE;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    spmm_noalloc(c, A->A, x, y, k, ld, alpha, beta, op, kchunk, k == 1 ? cached_lanes(c, A->A, false) : 0);
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_time_spmv(bcmgx_context ctx, bcmgx_matrix A, int precision, int k, int lanes, int reps, double* ms) {
    if (!ctx || !A || !ms || k < 1 || k > 64 || reps < 1 || (precision != 0 && precision != 1)) return BCMGX_ERR_ARG;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    DCsr& M = A->A;
    size_t nx = (size_t)std::max(M.m, 1) * k, ny = (size_t)std::max(M.n, 1) * k;
    SpmvArgs a;
    a.n = M.n; a.row = M.row.get(); a.col = M.col.get(); a.nnz = M.nnz; a.k = k;
    set_tiles(a, M);
    if (precision == 1 && !M.valf) {  // the FP32 copy must exist before any FP32 launch (tuning included)
        M.valf.reset(c, std::max(M.nnz, 1));
        convert<double, float>(c, M.val.get(), M.valf.get(), M.nnz);
    }
    if (k == 1 && M.n >= 65536) {
        build_tile_index(c, M);
        set_tiles(a, M);
    }
    if (k == 1) a.lanes = lanes != 0 ? lanes : tune_lanes(c, M, precision == 1);
    else a.kchunk = (lanes >= -2 || (lanes <= -16 && lanes >= -28)) ? lanes : 0;
    DBuf<double> xd, yd;
    DBuf<float> xf, yf;
    if (precision == 1) {
        xf.reset(c, nx); yf.reset(c, ny);
        fill<float>(c, xf.get(), (long long)nx, 1.0f);
        a.val = M.valf.get(); a.x = xf.get(); a.y = yf.get();
    } else {
        xd.reset(c, nx); yd.reset(c, ny);
        fill<double>(c, xd.get(), (long long)nx, 1.0);
        a.val = M.val.get(); a.x = xd.get(); a.y = yd.get();
    }
    auto launch = [&]() {
        if (precision == 1) spmv<float, float>(c, SpOp::AX, a);
        else spmv<double, double>(c, SpOp::AX, a);
    };
    launch();
    hipEvent_t e0, e1;
    BCMGX_HIP(hipEventCreate(&e0));
    BCMGX_HIP(hipEventCreate(&e1));
    BCMGX_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) launch();
    BCMGX_HIP(hipEventRecord(e1, c->stream));
    BCMGX_HIP(hipEventSynchronize(e1));
    float t = 0;
    BCMGX_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / reps;
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_time_smoother(bcmgx_context ctx, bcmgx_matrix A, int precision, int k, int reps, double* ms) {
    if (!ctx || !A || !ms || k < 1 || k > 64 || reps < 1 || (precision != 0 && precision != 1)) return BCMGX_ERR_ARG;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    DCsr& M = A->A;
    BCMGX_CHECK(M.n == M.m && M.n > 0, BCMGX_ERR_ARG, "the smoother needs a nonempty square matrix");
    bool f32 = precision == 1;
    if (f32 && !M.valf) {
        M.valf.reset(c, std::max(M.nnz, 1));
        convert<double, float>(c, M.val.get(), M.valf.get(), M.nnz);
    }
    int n = M.n;
    size_t nk = (size_t)n * k;
    SpmvArgs a;
    a.n = n; a.row = M.row.get(); a.col = M.col.get(); a.nnz = M.nnz; a.k = k;
    set_tiles(a, M);
    // the launch shape the V-cycle picks for this operator and batch width
    if (k == 1) a.lanes = tune_lanes(c, M, f32);
    else a.kchunk = tune_chunk(c, M, f32, k);
    DBuf<double> d64(c, n), xd, yd, bd;
    DBuf<float> df, xf, yf, bf;
    l1_diag(c, n, M.row.get(), M.col.get(), M.val.get(), d64.get());
    if (f32) {
        df.reset(c, n); xf.reset(c, nk); yf.reset(c, nk); bf.reset(c, nk);
        convert<double, float>(c, d64.get(), df.get(), n);
        fill<float>(c, xf.get(), (long long)nk, 1.0f);
        fill<float>(c, bf.get(), (long long)nk, 1.0f);
        a.val = M.valf.get(); a.x = xf.get(); a.y = yf.get(); a.b = bf.get(); a.D = df.get();
    } else {
        xd.reset(c, nk); yd.reset(c, nk); bd.reset(c, nk);
        fill<double>(c, xd.get(), (long long)nk, 1.0);
        fill<double>(c, bd.get(), (long long)nk, 1.0);
        a.val = M.val.get(); a.x = xd.get(); a.y = yd.get(); a.b = bd.get(); a.D = d64.get();
    }
    auto launch = [&]() {
        if (f32) spmv<float, float>(c, SpOp::JACOBI, a);
        else spmv<double, double>(c, SpOp::JACOBI, a);
    };
    launch();
    hipEvent_t e0, e1;
    BCMGX_HIP(hipEventCreate(&e0));
    BCMGX_HIP(hipEventCreate(&e1));
    BCMGX_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) launch();
    BCMGX_HIP(hipEventRecord(e1, c->stream));
    BCMGX_HIP(hipEventSynchronize(e1));
    float t = 0;
    BCMGX_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / reps;
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_time_spgemm(bcmgx_context ctx, bcmgx_matrix A, bcmgx_matrix B, int reps, double* ms, int* nnz_c) {
    if (!ctx || !A || !B || !ms || reps < 1) return BCMGX_ERR_ARG;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    BCMGX_CHECK(A->A.m == B->A.n, BCMGX_ERR_ARG, "spgemm: inner dimensions differ");
    {
        DCsr C;
        spgemm(c, A->A, B->A, C);  // warm-up (and the size of the result)
        if (nnz_c) *nnz_c = C.nnz;
    }
    c->sync();
    hipEvent_t e0, e1;
    BCMGX_HIP(hipEventCreate(&e0));
    BCMGX_HIP(hipEventCreate(&e1));
    BCMGX_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) {
        DCsr C;
        spgemm(c, A->A, B->A, C);
    }
    BCMGX_HIP(hipEventRecord(e1, c->stream));
    BCMGX_HIP(hipEventSynchronize(e1));
    float t = 0;
    BCMGX_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / reps;
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_stream_copy_gbs(bcmgx_context ctx, size_t bytes, int reps, double* gbs) {
    if (!ctx || !gbs || bytes < 1024 || reps < 1) return BCMGX_ERR_ARG;
    API_BEGIN
    BCMGX_HIP(hipSetDevice(ctx->c->device));
    *gbs = stream_copy_gbs(ctx->c, bytes, reps);
    return BCMGX_OK;
    API_END(ctx)
}

int bcmgx_sparse_outer_grad(bcmgx_context ctx, bcmgx_matrix A, const double* lambda, const double* x, int nrhs, double* g, int mem) {
    if (!ctx || !A || !lambda || !x || !g || nrhs < 1 || nrhs > 64) return BCMGX_ERR_ARG;
    API_BEGIN
    Context* c = ctx->c;
    BCMGX_HIP(hipSetDevice(c->device));
    const DCsr& M = A->A;
    int n = M.n;
    size_t nk = (size_t)std::max(n, 1) * nrhs;
    DBuf<double> L(c, nk), X(c, nk), G(c, std::max(M.nnz, 1)), tmp;
    stage_in(c, lambda, n, mem, n, nrhs, L.get(), tmp);
    stage_in(c, x, n, mem, n, nrhs, X.get(), tmp);
    sparse_outer_grad(c, M, L.get(), X.get(), nrhs, G.get());
    hipMemcpyKind kout = mem == BCMGX_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    BCMGX_HIP(hipMemcpyAsync(g, G.get(), sizeof(double) * M.nnz, kout, c->stream));
    c->sync();
    return BCMGX_OK;
    API_END(ctx)
}

}  // extern "C"
