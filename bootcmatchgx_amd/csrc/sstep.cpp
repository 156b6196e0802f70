// sstep.cpp -- s-step CG (Chronopoulos-Gear, solver/cgs/CGs.cu:43-236) and its
// pipelined form (solver/pipelined_cgs/pipelinedCGs.cu:56-300) on the device
// kernels.  The Krylov basis (mpk, solver/krylov_base/mpk.cu:10-78) is
// built with the library's SpMV and preconditioner applications into
// column-major n x 2s blocks, the 2s moments (computevm,
// krylov_base/computevm.cu) are reduced in a fixed order and read back, the
// s x s recurrences (scalarWorkMO, op/scalarWorkMO.cu) run on the host as in
// the reference, and the block updates (mydmmv, op/basic.cu) run on the
// device.  Right-hand sides are solved one after another.
#include <algorithm>
#include <cmath>
#include <vector>

#include "amg.h"
#include "kernels.h"

namespace bcmgx {

namespace {

// Cholesky L L^T of the s x s column-major matrix M (lower part), then nrhs
// solves in place on B (column-major s x nrhs).  Returns j+1 when the j-th
// pivot is not positive (dpotrf's info), 0 otherwise.
int cholesky_solve(std::vector<double>& M, double* B, int s, int nrhs) {
    for (int j = 0; j < s; j++) {
        double d = M[j + j * s];
        for (int k = 0; k < j; k++) d -= M[j + k * s] * M[j + k * s];
        if (!(d > 0.0)) return j + 1;
        d = std::sqrt(d);
        M[j + j * s] = d;
        for (int i = j + 1; i < s; i++) {
            double t = M[i + j * s];
            for (int k = 0; k < j; k++) t -= M[i + k * s] * M[j + k * s];
            M[i + j * s] = t / d;
        }
    }
    for (int c = 0; c < nrhs; c++) {
        double* b = B + (size_t)c * s;
        for (int i = 0; i < s; i++) {
            double t = b[i];
            for (int k = 0; k < i; k++) t -= M[i + k * s] * b[k];
            b[i] = t / M[i + i * s];
        }
        for (int i = s - 1; i >= 0; i--) {
            double t = b[i];
            for (int k = i + 1; k < s; k++) t -= M[k + i * s] * b[k];
            b[i] = t / M[i + i * s];
        }
    }
    return 0;
}

// scalarWorkMO (op/scalarWorkMO.cu:8-85): from the moments vm (2s), the step
// coefficients alpha (s) and the basis recombination beta (s x s); W (s x s)
// carries over between calls.  Nonzero: the s x s system is not positive definite.
struct SstepScalars {
    int s = 1;
    std::vector<double> W, alpha, beta;
    explicit SstepScalars(int s_) : s(s_), W(s_ * s_, 0.0), alpha(s_, 0.0), beta(s_ * s_, 0.0) {}
    int step(const std::vector<double>& vm, int iter) {
        if (iter == 0) {
            for (int i = 0; i < s; i++)
                for (int j = 0; j < s; j++) W[i * s + j] = vm[i + j + 1];  // Hankel matrix of the moments
            std::vector<double> Wc(W);
            std::copy(vm.begin(), vm.begin() + s, alpha.begin());
            return cholesky_solve(Wc, alpha.data(), s, 1);
        }
        std::vector<double> h(2 * s - 1, 0.0), b1(s * s);
        for (int i = 0; i < s; i++) {
            double t = vm[i];
            for (int j = 0; j < i; j++) t += h[s - 1 + j] * alpha[s - i + j - 1];
            h[s - 1 + i] = -t / alpha[s - 1];
        }
        for (int i = 0; i < s; i++)
            for (int j = 0; j < s; j++) b1[i + j * s] = beta[i + j * s] = -h[i + j];
        if (int info = cholesky_solve(W, beta.data(), s, s)) return info;  // the previous W
        for (int i = 0; i < s; i++)
            for (int j = 0; j < s; j++) W[i * s + j] = vm[i + j + 1];
        for (int i = 0; i < s; i++)  // W -= b1^T beta
            for (int j = 0; j < s; j++) {
                double t = 0.0;
                for (int k = 0; k < s; k++) t += b1[k + i * s] * beta[k + j * s];
                W[i + j * s] = -1.0 * t + 1.0 * W[i + j * s];
            }
        std::vector<double> Wc(W);
        std::copy(vm.begin(), vm.begin() + s, alpha.begin());
        return cholesky_solve(Wc, alpha.data(), s, 1);
    }
};

struct SstepRun {
    Context* c;
    const DCsr& A;
    Prec* prec;
    bool precon;
    int n, s, nblk, lanes;
    DBuf<double> partials, dvm, dab;  // moments; alpha (s) then beta (s x s)
    SstepRun(Context* c_, const DCsr& A_, Prec* p_, int s_, int lanes_)
        : c(c_), A(A_), prec(p_), precon(p_->ptype != BCMGX_PREC_NONE), n(A_.n), s(s_), lanes(lanes_) {
        nblk = fcg_partial_blocks(n);
        partials.reset(c, (size_t)nblk * 2 * s + 1);
        dvm.reset(c, 2 * s + 1);
        dab.reset(c, s + s * s);
    }
    void ax(const double* x, double* y, SpOp op = SpOp::AX, const double* b = nullptr) {
        SpmvArgs a;
        a.n = n; a.row = A.row.get(); a.col = A.col.get(); a.val = A.val.get(); a.nnz = A.nnz;
        set_tiles(a, A);
        a.x = x; a.y = y; a.b = b; a.k = 1; a.lanes = lanes;
        spmv<double, double>(c, op, a);
    }
    void precond(const double* r, double* u) { prec->apply(r, u, 1, true, nullptr); }
    double* colp(double* base, int j) { return base + (size_t)j * n; }
    // mpk (mpk.cu:10-78)
    void mpk(const double* x, double* sP) {
        if (!precon) {
            copy_d2d(c, sP, x, sizeof(double) * n);
            for (int i = 0; i < s; i++) ax(colp(sP, i), colp(sP, i + 1));
            return;
        }
        precond(x, sP);
        int pin = 0, pout = s;
        for (int i = 0; i < 2 * s - 1; i++) {
            if (i % 2 == 0) ax(colp(sP, pin), colp(sP, pout));
            else precond(colp(sP, pin), colp(sP, pout));
            int t = pin;
            pin = pout;
            pout = t + 1;
        }
    }
    // computevm (computevm.cu:6-31), read back
    void moments(double* sP, const double* r, std::vector<double>& vm) {
        double* out = dvm.get();
        if (!precon) {
            block_dots(c, n, s + 1, sP, r, 0, partials.get(), nblk, out);
            block_dots(c, n, s - 1, colp(sP, 1), colp(sP, s), 0, partials.get(), nblk, out + s + 1);
        } else {
            block_dots(c, n, s, sP, r, 0, partials.get(), nblk, out);
            block_dots(c, n, s, sP, colp(sP, 2 * s - 1), 0, partials.get(), nblk, out + s);
        }
        vm.assign(2 * s, 0.0);
        c->d2h(vm.data(), out, 2 * s);
    }
    double norm(const double* r) {
        block_dots(c, n, 1, r, r, 0, partials.get(), nblk, dvm.get());
        return std::sqrt(c->read1(dvm.get()));
    }
    void upload(const SstepScalars& sc) {
        std::vector<double> h(s + s * s);
        std::copy(sc.alpha.begin(), sc.alpha.end(), h.begin());
        std::copy(sc.beta.begin(), sc.beta.end(), h.begin() + s);
        c->h2d(dab.get(), h.data(), h.size());
    }
};

struct ColumnResult {
    int retv = -1, niter = 0;
    double l2 = 0, delta0 = 1;
};

// cgsstep (CGs.cu:43-236) on one right-hand side, x in/out (device)
ColumnResult cgs_column(SstepRun& R, const bcmgx_params& p, const double* b, double* x) {
    Context* c = R.c;
    int n = R.n, s = R.s;
    bool cp = !R.precon && p.ru_res;
    DBuf<double> r(c, n), sP1(c, (size_t)n * 2 * s), sP2(c, (size_t)n * 2 * s), sCP1, sCP2;
    fill<double>(c, sP1.get(), (long long)n * 2 * s, 0.0);
    if (cp) {
        sCP1.reset(c, (size_t)n * s);
        sCP2.reset(c, (size_t)n * s);
        fill<double>(c, sCP1.get(), (long long)n * s, 0.0);
    }
    R.ax(x, r.get(), SpOp::RES, b);  // r = b - A x  (CGs.cu:71-74)
    SstepScalars sc(s);
    ColumnResult res;
    if (p.stop_criterion == 1) res.delta0 = R.norm(r.get());
    std::vector<double> vm;
    int iter, info = 0, excnlim = -1;
    const double* al = R.dab.get();
    const double* be = R.dab.get() + s;
    for (iter = 0; iter < p.itnlim; iter++) {
        if (iter > 0 && info == 0 && p.ru_res && p.rec_res_int > 0 && iter % p.rec_res_int == 0)
            R.ax(x, r.get(), SpOp::RES, b);  // CGs.cu:121-131
        double* cur = iter % 2 == 0 ? sP2.get() : sP1.get();
        double* old = iter % 2 == 0 ? sP1.get() : sP2.get();
        double* ccur = iter % 2 == 0 ? sCP2.get() : sCP1.get();
        double* cold = iter % 2 == 0 ? sCP1.get() : sCP2.get();
        R.mpk(r.get(), cur);
        R.moments(cur, r.get(), vm);
        info = sc.step(vm, iter);
        if (info != 0) {
            res.retv = info;
            break;
        }
        R.upload(sc);
        if (cp) copy_d2d(c, ccur, cur + n, sizeof(double) * (size_t)n * s);
        block_dmmv(c, n, s, old, be, cur, al, x, 1.0);  // basis += old basis beta; x += basis alpha
        if (p.ru_res) {
            if (R.precon) block_dmmv(c, n, s, old + (size_t)s * n, be, cur + (size_t)s * n, al, r.get(), -1.0);
            else block_dmmv(c, n, s, cold, be, ccur, al, r.get(), -1.0);
        } else {
            R.ax(x, r.get(), SpOp::RES, b);
        }
        res.l2 = R.norm(r.get());
        if (res.l2 < p.rtol * res.delta0) {
            excnlim = 0;
            res.retv = 0;
            break;
        }
    }
    res.niter = iter + 1 + excnlim;  // CGs.cu:216-222
    return res;
}

// pipecomputevm (pipelinedCGs.cu:7-27), read back
void pipe_moments(SstepRun& R, const double* UR, const double* UQ, const double* r, std::vector<double>& vm) {
    Context* c = R.c;
    int n = R.n, s = R.s;
    double* out = R.dvm.get();
    const double* q = UQ + (size_t)(s - 1) * n;
    block_dots(c, n, s, UR, r, 0, R.partials.get(), R.nblk, out);
    if (R.precon) {
        block_dots(c, n, s, UR, q, 0, R.partials.get(), R.nblk, out + s);
    } else {
        block_dots(c, n, 1, q, r, 0, R.partials.get(), R.nblk, out + s);
        block_dots(c, n, s - 1, UQ, q, 0, R.partials.get(), R.nblk, out + s + 1);
    }
    vm.assign(2 * s, 0.0);
    c->d2h(vm.data(), out, 2 * s);
}

// pipelinedcgsstep (pipelinedCGs.cu:56-300) on one right-hand side.  URm/UQm
// hold s+1 blocks of n x s (block j at j*s*n); sets 1 and 2 swap roles
// between even and odd iterations.
ColumnResult pipecgs_column(SstepRun& R, const bcmgx_params& p, const double* b, double* x) {
    Context* c = R.c;
    int n = R.n, s = R.s;
    size_t bs = (size_t)n * s;
    DBuf<double> r(c, n), sP2(c, (size_t)n * 2 * s), sT(c, (size_t)n * 2 * s), ams(c, n);
    DBuf<double> UR1m(c, bs * (s + 1)), UR2m(c, bs * (s + 1)), UQ1m(c, bs * (s + 1)), UQ2m(c, bs * (s + 1));
    DBuf<double> UR1(c, bs), UR2(c, bs), UQ1(c, bs), UQ2(c, bs);
    for (DBuf<double>* z : {&UR1m, &UR2m, &UQ1m, &UQ2m}) fill<double>(c, z->get(), (long long)bs * (s + 1), 0.0);
    for (DBuf<double>* z : {&UR1, &UR2, &UQ1, &UQ2}) fill<double>(c, z->get(), (long long)bs, 0.0);
    R.ax(x, r.get(), SpOp::RES, b);  // r = b - A x  (pipelinedCGs.cu:101-103)
    R.mpk(r.get(), sP2.get());
    copy_d2d(c, UR2m.get(), sP2.get(), sizeof(double) * bs);
    copy_d2d(c, UQ2m.get(), sP2.get() + (R.precon ? bs : (size_t)n), sizeof(double) * bs);
    SstepScalars sc(s);
    ColumnResult res;
    if (p.stop_criterion == 1) res.delta0 = R.norm(r.get());
    std::vector<double> vm;
    const double* al = R.dab.get();
    const double* be = R.dab.get() + s;
    int iter, excnlim = -1;
    for (iter = 0; iter < p.itnlim; iter++) {
        bool even = iter % 2 == 0;
        double* URn = even ? UR2m.get() : UR1m.get();  // the set this iteration builds
        double* UQn = even ? UQ2m.get() : UQ1m.get();
        double* URo = even ? UR1m.get() : UR2m.get();  // the previous iteration's
        double* UQo = even ? UQ1m.get() : UQ2m.get();
        double* URs = even ? UR1.get() : UR2.get();    // its block 0, saved
        double* UQs = even ? UQ1.get() : UQ2.get();
        double* URsn = even ? UR2.get() : UR1.get();
        double* UQsn = even ? UQ2.get() : UQ1.get();
        if (iter > 0) {  // pipelinedCGs.cu:150-160 / 218-228
            for (int j = 1; j < s + 1; j++) {
                block_dmm(c, n, s, URn + j * bs, be, URo + j * bs);
                block_dmm(c, n, s, UQn + j * bs, be, UQo + j * bs);
            }
            for (int j = 0; j < s; j++) {
                block_dmv2v(c, n, s, URo + (j + 1) * bs, al, URs + (size_t)j * n, -1.0, URn + (size_t)j * n);
                block_dmv2v(c, n, s, UQo + (j + 1) * bs, al, UQs + (size_t)j * n, -1.0, UQn + (size_t)j * n);
            }
        }
        pipe_moments(R, URn, UQn, r.get(), vm);
        copy_d2d(c, ams.get(), UQn + (size_t)(s - 1) * n, sizeof(double) * n);
        R.mpk(ams.get(), sT.get());
        int info = sc.step(vm, iter);
        if (info != 0) {
            res.retv = info;
            break;
        }
        R.upload(sc);
        size_t qo = R.precon ? bs : (size_t)n;
        for (int j = 1; j < s + 1; j++) {  // block j: columns j..s-1 of block 0, then the first j of the new basis
            copy_d2d(c, URn + j * bs, URn + (size_t)j * n, sizeof(double) * (size_t)(s - j) * n);
            copy_d2d(c, URn + j * bs + (size_t)(s - j) * n, sT.get(), sizeof(double) * (size_t)j * n);
            copy_d2d(c, UQn + j * bs, UQn + (size_t)j * n, sizeof(double) * (size_t)(s - j) * n);
            copy_d2d(c, UQn + j * bs + (size_t)(s - j) * n, sT.get() + qo, sizeof(double) * (size_t)j * n);
        }
        copy_d2d(c, URsn, URn, sizeof(double) * bs);
        copy_d2d(c, UQsn, UQn, sizeof(double) * bs);
        block_dmmv(c, n, s, URo, be, URn, al, x, 1.0);
        block_dmmv(c, n, s, UQo, be, UQn, al, r.get(), -1.0);
        res.l2 = R.norm(r.get());
        if (res.l2 < p.rtol * res.delta0) {
            excnlim = 0;
            res.retv = 0;
            break;
        }
    }
    res.niter = iter + 1 + excnlim;
    return res;
}

}  // namespace

SolveResult cgs_solve(Context* c, const DCsr& A, Prec* prec, const bcmgx_params& p, const double* b, double* x, int k) {
    BCMGX_CHECK(k >= 1 && k <= 64, BCMGX_ERR_ARG, "nrhs must be in [1, 64]");
    BCMGX_CHECK(p.solver == BCMGX_SOLVER_CGS || p.solver == BCMGX_SOLVER_CGS_CUBLAS || p.solver == BCMGX_SOLVER_PIPELINED_CGS,
                BCMGX_ERR_STATE, "cgs_solve called for another solver");
    BCMGX_CHECK(p.sstep >= 1 && p.sstep <= 16, BCMGX_ERR_ARG, "sstep must be in [1, 16]");
    double t0 = now_s();
    int n = A.n;
    prec->ensure_work(1);
    int lanes = p.autotune ? tune_lanes(c, A, false) : default_lanes(A.nnz, A.n);
    SstepRun R(c, A, prec, p.sstep, lanes);
    // the columns arrive interleaved (x[i*k + j]); each is solved on its own
    DBuf<double> bc(c, std::max(n, 1)), xc(c, std::max(n, 1));
    SolveResult res;
    prec->last_niter.assign(k, 0);
    prec->last_exit.assign(k, 0);
    prec->last_rel.assign(k, 0);
    for (int j = 0; j < k; j++) {
        if (k == 1) {
            copy_d2d(c, bc.get(), b, sizeof(double) * n);
            copy_d2d(c, xc.get(), x, sizeof(double) * n);
        } else {
            strided_copy(c, bc.get(), 1, b + j, k, n);
            strided_copy(c, xc.get(), 1, x + j, k, n);
        }
        ColumnResult cr = p.solver == BCMGX_SOLVER_PIPELINED_CGS ? pipecgs_column(R, p, bc.get(), xc.get())
                                                                 : cgs_column(R, p, bc.get(), xc.get());
        if (k == 1) copy_d2d(c, x, xc.get(), sizeof(double) * n);
        else strided_copy(c, x + j, k, xc.get(), 1, n);
        double rel = cr.delta0 > 0 ? cr.l2 / cr.delta0 : 0.0;
        prec->last_niter[j] = cr.niter;
        prec->last_exit[j] = cr.l2;
        prec->last_rel[j] = rel;
        res.niter = std::max(res.niter, cr.niter);
        res.exit_res = std::max(res.exit_res, cr.l2);
        res.rel_res = std::max(res.rel_res, rel);
        if (cr.retv != 0 && res.retv <= 0) res.retv = cr.retv;  // a breakdown (> 0) outranks itnlim (-1)
        if (!std::isfinite(cr.l2)) res.retv = BCMGX_ERR_NUMERIC;
        if (j == 0) res.del0 = cr.delta0;
    }
    c->sync();
    res.seconds = now_s() - t0;
    return res;
}

}  // namespace bcmgx
