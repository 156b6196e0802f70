// fcg.hip -- flexible CG vector kernels (flexibileConjugateGradients_v3,
// src/solver/fcg/FCG.cu:27-175) with the scalar recurrences kept on the
// device.  The reference reads every dot product back to the host
// (triple_innerproduct, src/op/triple_inner_product.cu:64-101; norm_MPI,
// src/datastruct/vector.cu:492-516); here the scalars, the convergence test
// and the iteration count live in FcgState so a whole iteration can be
// replayed from a hipGraph and the host only polls a flag.
//
// Reductions are deterministic: a fixed number of partial blocks (a function
// of n only), a fixed in-block tree, then one block per column summing the
// partials in a fixed tree.  k vectors are interleaved (x[i*k + c]); every
// column has its own scalars and stop flag.
#include "kernels.h"

#include <algorithm>

namespace bcmgx {

// partial-sum blocks of the streaming FCG kernels: a function of n only, so
// the reduction order never changes
int fcg_partial_blocks(int n) {
    long long b = ((long long)n + 2047) / 2048;
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (int)b;
}

static inline int kpad(int k) {
    int p = 1;
    while (p < k) p <<= 1;
    return p;
}

// NV dot products per column over a row range; partials[(blk*k + c)*3 + m]
template <int NV>
__global__ __launch_bounds__(WG) void k_dots(int n, int k, int KP, const double* __restrict__ a0, const double* __restrict__ b0,
                                             const double* __restrict__ a1, const double* __restrict__ b1,
                                             const double* __restrict__ a2, const double* __restrict__ b2,
                                             double* __restrict__ partials, const int* __restrict__ all_done) {
    if (all_done && *all_done) return;
    __shared__ double s[NV][WG];
    int tid = threadIdx.x;
    int R = WG / KP;
    int rl = tid / KP, c = tid % KP;
    double acc[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) acc[m] = 0.0;
    if (c < k) {
        for (long long i = (long long)blockIdx.x * R + rl; i < n; i += (long long)gridDim.x * R) {
            long long o = i * k + c;
            acc[0] += a0[o] * b0[o];
            if (NV > 1) acc[1] += a1[o] * b1[o];
            if (NV > 2) acc[2] += a2[o] * b2[o];
        }
    }
#pragma unroll
    for (int m = 0; m < NV; m++) s[m][tid] = acc[m];
    __syncthreads();
    for (int st = WG / 2; st >= KP; st >>= 1) {
        if (tid < st) {
#pragma unroll
            for (int m = 0; m < NV; m++) s[m][tid] += s[m][tid + st];
        }
        __syncthreads();
    }
    if (tid < k) {
#pragma unroll
        for (int m = 0; m < NV; m++) partials[((long long)blockIdx.x * k + tid) * 3 + m] = s[m][tid];
    }
}

// sum partials of column blockIdx.x in a fixed tree; thread 0 gets the totals
template <int NV>
__device__ void sum_partials(const double* __restrict__ partials, int nblk, int k, int c, double* out, double (*s)[WG]) {
    int tid = threadIdx.x;
    double acc[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) acc[m] = 0.0;
    for (int b = tid; b < nblk; b += WG)
#pragma unroll
        for (int m = 0; m < NV; m++) acc[m] += partials[((long long)b * k + c) * 3 + m];
#pragma unroll
    for (int m = 0; m < NV; m++) s[m][tid] = acc[m];
    __syncthreads();
    for (int st = WG / 2; st > 0; st >>= 1) {
        if (tid < st)
#pragma unroll
            for (int m = 0; m < NV; m++) s[m][tid] += s[m][tid + st];
        __syncthreads();
    }
    if (tid == 0)
#pragma unroll
        for (int m = 0; m < NV; m++) out[m] = s[m][0];
}

// mode 0: initial residual norm (delta0); mode 1: post-initial-step norm
__global__ __launch_bounds__(WG) void k_norm_final(const double* __restrict__ partials, int nblk, int k, FcgState* st,
                                                   int mode, int* all_done, int* ndone) {
    __shared__ double s[1][WG];
    double tot[1];
    int c = blockIdx.x;
    sum_partials<1>(partials, nblk, k, c, tot, s);
    if (threadIdx.x == 0) {
        double l2 = sqrt(tot[0]);
        st[c].l2 = l2;
        if (mode == 0) {
            st[c].delta0 = l2;
            st[c].iter = 0;
            st[c].done = 0;
            st[c].pend = 0.0;
            st[c].pad1 = 0;
            if (l2 == 0.0) {  // b == A x already: nothing to iterate on
                st[c].done = 1;
                if (atomicAdd(ndone, 1) + 1 == k) *all_done = 1;
            }
        }
    }
}

// ||r||^2 partials in k_dots<1>'s blocks and order, with the FP32 cycle's
// level-0 entry (F = (float) r, X = F/M) written on the way, as k_update does
__global__ __launch_bounds__(WG) void k_norm_entry(int n, int k, int KP, const double* __restrict__ r,
                                                   double* __restrict__ partials, FcgEntry e) {
    __shared__ double s[WG];
    int tid = threadIdx.x;
    int R = WG / KP;
    int rl = tid / KP, c = tid % KP;
    double acc = 0.0;
    if (c < k) {
        for (long long i = (long long)blockIdx.x * R + rl; i < n; i += (long long)gridDim.x * R) {
            long long o = i * k + c;
            double rr = r[o];
            acc += rr * rr;
            float fv = (float)rr;
            if (e.F) e.F[o] = fv;
            e.X[o] = fv / e.M[i];
        }
    }
    s[tid] = acc;
    __syncthreads();
    for (int st = WG / 2; st >= KP; st >>= 1) {
        if (tid < st) s[tid] += s[tid + st];
        __syncthreads();
    }
    if (tid < k) partials[((long long)blockIdx.x * k + tid) * 3] = s[tid];
}

void fcg_norm_entry(Context* c, const FcgBufs& f, const double* r, int mode, const FcgEntry& e) {
    int KP = kpad(f.k);
    k_norm_entry<<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, r, f.partials, e);
    BCMGX_LAUNCHED();
    k_norm_final<<<f.k, WG, 0, c->stream>>>(f.partials, f.nblk, f.k, f.st, mode, f.all_done, f.all_done + 1);
    BCMGX_LAUNCHED();
}

void fcg_norm(Context* c, const FcgBufs& f, const double* v, int mode) {
    int KP = kpad(f.k);
    k_dots<1><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, v, v, nullptr, nullptr, nullptr, nullptr, f.partials, nullptr);
    BCMGX_LAUNCHED();
    k_norm_final<<<f.k, WG, 0, c->stream>>>(f.partials, f.nblk, f.k, f.st, mode, f.all_done, f.all_done + 1);
    BCMGX_LAUNCHED();
}

__global__ __launch_bounds__(WG) void k_init_scalars(const double* __restrict__ partials, int nblk, int k, FcgState* st) {
    __shared__ double s[2][WG];
    double tot[2];
    int c = blockIdx.x;
    sum_partials<2>(partials, nblk, k, c, tot, s);
    if (threadIdx.x == 0) {
        FcgState& S = st[c];
        S.alpha = tot[0];
        S.beta = tot[1];
        S.delta = S.beta;                // FCG.cu:84
        S.theta = S.done ? 0.0 : S.alpha / S.delta;  // FCG.cu:85
        S.gamma = 0.0;
    }
}

void fcg_init_dots(Context* c, const FcgBufs& f, const double* r, const double* v, const double* w) {
    int KP = kpad(f.k);
    k_dots<2><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, r, v, w, v, nullptr, nullptr, f.partials, nullptr);
    BCMGX_LAUNCHED();
    fcg_init_scalars(c, f, f.nblk);
}

void fcg_init_scalars(Context* c, const FcgBufs& f, int nblk) {
    k_init_scalars<<<f.k, WG, 0, c->stream>>>(f.partials, nblk, f.k, f.st);
    BCMGX_LAUNCHED();
}

// v in FP64, or the FP32 cycle result widened exactly
template <typename VT>
__global__ __launch_bounds__(WG) void k_init_update(int n, int k, const FcgState* __restrict__ st, double* __restrict__ x,
                                                    double* __restrict__ r, const VT* __restrict__ v,
                                                    const double* __restrict__ w, double* __restrict__ d,
                                                    double* __restrict__ q, FcgEntry e) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * k) return;
    int i = (int)(t / k);
    int c = (int)(t - (long long)i * k);
    double th = st[c].theta;
    double vv = (double)v[t], ww = w[t];
    double rr = r[t];
    if (!st[c].done) {
        x[t] += th * vv;      // my_axpby(v, x, theta, 1): y += a x  (basic.cu:114-130)
        rr += -th * ww;       // my_axpby(w, r, -theta, 1)
        r[t] = rr;
    }
    d[t] = vv;                // d = clone(v), q = clone(w)  (FCG.cu:98-99)
    q[t] = ww;
    if (e.X) {
        float f = (float)rr;
        if (e.F) e.F[t] = f;
        e.X[t] = f / e.M[i];
    }
}

void fcg_init_update(Context* c, const FcgBufs& f, double* x, double* r, const double* v, const double* w, double* d, double* q,
                     const FcgEntry& e) {
    long long nk = (long long)f.n * f.k;
    if (!nk) return;
    k_init_update<double><<<grid_for(nk, WG), WG, 0, c->stream>>>(f.n, f.k, f.st, x, r, v, w, d, q, e);
    BCMGX_LAUNCHED();
}

void fcg_init_update_z32(Context* c, const FcgBufs& f, double* x, double* r, const float* z, const double* w, double* d,
                         double* q, const FcgEntry& e) {
    long long nk = (long long)f.n * f.k;
    if (!nk) return;
    k_init_update<float><<<grid_for(nk, WG), WG, 0, c->stream>>>(f.n, f.k, f.st, x, r, z, w, d, q, e);
    BCMGX_LAUNCHED();
}

__global__ __launch_bounds__(WG) void k_scalars3(const double* __restrict__ partials, int nblk, int k, FcgState* st,
                                                 const int* __restrict__ all_done, int phase) {
    if (*all_done) return;
    __shared__ double s[3][WG];
    double tot[3];
    int c = blockIdx.x;
    sum_partials<3>(partials, nblk, k, c, tot, s);
    if (threadIdx.x == 0) {
        FcgState& S = st[c];
        if (S.done) return;
        S.alpha = tot[0];
        S.beta = tot[1];
        S.gamma = tot[2];
        S.theta = S.gamma / S.delta;                        // FCG.cu:132
        S.delta = S.beta - ((S.gamma * S.gamma) / S.delta);  // FCG.cu:135
        S.theta2 = S.alpha / S.delta;                       // FCG.cu:136
        if (phase == 0) {  // the x step of this iteration waits for the pair's second update
            S.pend = S.theta2;
            S.pad1 = 1;
        }
    }
}

void fcg_scalars3(Context* c, const FcgBufs& f, int nblk, int phase) {
    k_scalars3<<<f.k, WG, 0, c->stream>>>(f.partials, nblk, f.k, f.st, f.all_done, phase);
    BCMGX_LAUNCHED();
}

void fcg_dots3(Context* c, const FcgBufs& f, const double* r, const double* p0, const double* w0, const double* w1, int phase) {
    int KP = kpad(f.k);
    k_dots<3><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, r, p0, w0, p0, w1, p0, f.partials, f.all_done);
    BCMGX_LAUNCHED();
    k_scalars3<<<f.k, WG, 0, c->stream>>>(f.partials, f.nblk, f.k, f.st, f.all_done, phase);
    BCMGX_LAUNCHED();
}

// double_merged_axpy(P1, P0, x, -theta, theta2); double_merged_axpy(W1, W0, r, -theta, -theta2)
// (double_merged_axpy.cu:28-39), fused, with the partial sums of r.r.  The x
// step is taken every second iteration: phase 0 leaves x alone (its theta2
// waits in pend, its direction stays in the buffer that is p1 next time),
// phase 1 forms x = theta2' t' + (pend p1 + x) -- the two steps in the
// reference's order, so x has the same bits -- and saves one x read and write.
template <int PHASE>
__global__ __launch_bounds__(WG) void k_update(int n, int k, int KP, const FcgState* __restrict__ st,
                                               double* __restrict__ p0, const double* __restrict__ p1,
                                               double* __restrict__ w0, const double* __restrict__ w1,
                                               double* __restrict__ x, double* __restrict__ r,
                                               double* __restrict__ partials, const int* __restrict__ all_done,
                                               FcgEntry e) {
    if (*all_done) return;
    __shared__ double s[WG];
    int tid = threadIdx.x;
    int R = WG / KP;
    int rl = tid / KP, c = tid % KP;
    double acc = 0.0;
    if (c < k) {
        const FcgState S = st[c];
        double mth = -S.theta, th2 = S.theta2, mth2 = -S.theta2;
        bool waiting = PHASE == 1 && S.pad1;  // the pair's first step, still to be added
        double pend = S.pend;
        for (long long i = (long long)blockIdx.x * R + rl; i < n; i += (long long)gridDim.x * R) {
            long long o = i * k + c;
            double rr;
            if (!S.done) {
                double p1o = p1[o];
                double t = mth * p1o + p0[o];
                if constexpr (PHASE == 1) {
                    double xo = x[o];
                    double xw = pend * p1o + xo;
                    x[o] = th2 * t + (waiting ? xw : xo);
                }
                p0[o] = t;
                double u = mth * w1[o] + w0[o];
                rr = mth2 * u + r[o];
                r[o] = rr;
                w0[o] = u;
            } else {
                rr = r[o];
                if (waiting) x[o] = pend * p1[o] + x[o];
            }
            acc += rr * rr;
            if (e.X) {
                float f = (float)rr;
                if (e.F) e.F[o] = f;
                e.X[o] = f / e.M[i];
            }
        }
    }
    s[tid] = acc;
    __syncthreads();
    for (int stp = WG / 2; stp >= KP; stp >>= 1) {
        if (tid < stp) s[tid] += s[tid + stp];
        __syncthreads();
    }
    if (tid < k) partials[((long long)blockIdx.x * k + tid) * 3] = s[tid];
}

void fcg_update(Context* c, const FcgBufs& f, double* p0, const double* p1, double* w0, const double* w1, double* x, double* r,
                int phase, const FcgEntry& e) {
    int KP = kpad(f.k);
    if (phase == 0) k_update<0><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, f.st, p0, p1, w0, w1, x, r, f.partials, f.all_done, e);
    else k_update<1><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, f.st, p0, p1, w0, w1, x, r, f.partials, f.all_done, e);
    BCMGX_LAUNCHED();
}

// k_update with z in FP32 (widened exactly) and p, w updated in place
template <int PHASE>
__global__ __launch_bounds__(WG) void k_update_z32(int n, int k, int KP, const FcgState* __restrict__ st,
                                                   const float* __restrict__ z, double* __restrict__ p,
                                                   const double* __restrict__ w0, double* __restrict__ w,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ partials, const int* __restrict__ all_done,
                                                   FcgEntry e) {
    if (*all_done) return;
    __shared__ double s[WG];
    int tid = threadIdx.x;
    int R = WG / KP;
    int rl = tid / KP, c = tid % KP;
    double acc = 0.0;
    if (c < k) {
        const FcgState S = st[c];
        double mth = -S.theta, th2 = S.theta2, mth2 = -S.theta2;
        bool waiting = PHASE == 1 && S.pad1;
        double pend = S.pend;
        for (long long i = (long long)blockIdx.x * R + rl; i < n; i += (long long)gridDim.x * R) {
            long long o = i * k + c;
            double rr;
            if (!S.done) {
                double po = p[o];
                double t = mth * po + (double)z[o];
                if constexpr (PHASE == 1) {
                    double xo = x[o];
                    double xw = pend * po + xo;
                    x[o] = th2 * t + (waiting ? xw : xo);
                }
                p[o] = t;
                double u = mth * w[o] + w0[o];
                rr = mth2 * u + r[o];
                r[o] = rr;
                w[o] = u;
            } else {
                rr = r[o];
                if (waiting) x[o] = pend * p[o] + x[o];
            }
            acc += rr * rr;
            if (e.X) {
                float f = (float)rr;
                if (e.F) e.F[o] = f;
                e.X[o] = f / e.M[i];
            }
        }
    }
    s[tid] = acc;
    __syncthreads();
    for (int stp = WG / 2; stp >= KP; stp >>= 1) {
        if (tid < stp) s[tid] += s[tid + stp];
        __syncthreads();
    }
    if (tid < k) partials[((long long)blockIdx.x * k + tid) * 3] = s[tid];
}

void fcg_update_z32(Context* c, const FcgBufs& f, const float* z, double* p, const double* w0, double* w, double* x, double* r,
                    int phase, const FcgEntry& e) {
    int KP = kpad(f.k);
    if (phase == 0) k_update_z32<0><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, f.st, z, p, w0, w, x, r, f.partials, f.all_done, e);
    else k_update_z32<1><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, f.st, z, p, w0, w, x, r, f.partials, f.all_done, e);
    BCMGX_LAUNCHED();
}

// l2 = ||r||; iter++; stop when !(l2 > rtol*delta0 && iter < itnlim)  (FCG.cu:146-156)
__global__ __launch_bounds__(WG) void k_finish(const double* __restrict__ partials, int nblk, int k, FcgState* st,
                                               double rtol, int itnlim, int* all_done, int* ndone, int phase) {
    if (*all_done) return;
    __shared__ double s[1][WG];
    double tot[1];
    int c = blockIdx.x;
    sum_partials<1>(partials, nblk, k, c, tot, s);
    if (threadIdx.x == 0) {
        FcgState& S = st[c];
        if (phase == 1) S.pad1 = 0;  // this pair's update added the waiting step
        if (S.done) return;
        S.l2 = sqrt(tot[0]);
        S.iter += 1;
        if (!(S.l2 > rtol * S.delta0 && S.iter < itnlim)) {
            S.done = 1;
            if (atomicAdd(ndone, 1) + 1 == k) *all_done = 1;
        }
    }
}

void fcg_finish_iteration(Context* c, const FcgBufs& f, int phase) {
    k_finish<<<f.k, WG, 0, c->stream>>>(f.partials, f.nblk, f.k, f.st, f.rtol, f.itnlim, f.all_done, f.all_done + 1, phase);
    BCMGX_LAUNCHED();
}

__global__ __launch_bounds__(WG) void k_flush_pending(int n, int k, int KP, FcgState* st, const double* __restrict__ t,
                                                      double* __restrict__ x) {
    int tid = threadIdx.x;
    int R = WG / KP;
    int rl = tid / KP, c = tid % KP;
    if (c >= k || !st[c].pad1) return;
    double pend = st[c].pend;
    for (long long i = (long long)blockIdx.x * R + rl; i < n; i += (long long)gridDim.x * R) {
        long long o = i * k + c;
        x[o] = pend * t[o] + x[o];
    }
}

__global__ void k_clear_pending(int k, FcgState* st) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < k) st[c].pad1 = 0;
}

void fcg_clear_pending(Context* c, const FcgBufs& f) {
    k_clear_pending<<<1, 64, 0, c->stream>>>(f.k, f.st);
    BCMGX_LAUNCHED();
}

void fcg_flush_pending(Context* c, const FcgBufs& f, const double* t, double* x) {
    int KP = kpad(f.k);
    k_flush_pending<<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, f.st, t, x);
    BCMGX_LAUNCHED();
    k_clear_pending<<<1, 64, 0, c->stream>>>(f.k, f.st);
    BCMGX_LAUNCHED();
}

// ---- CG (Hestenes-Stiefel), cg_hs (src/solver/cghs/CG_HS.cu:24-172) ---------
// FcgState fields as used by CG: alpha = rho = r.u of the current iteration,
// delta = rho of the previous one, theta = beta, gamma = p.s, theta2 = alpha
// (the step), pad0 = 1 once the column converged.

// delta0 = ||r|| (stop_criterion 1) or 1 (CG_HS.cu:74-79); a zero initial
// residual stops the column at once (the reference would divide 0 by 0)
__global__ __launch_bounds__(WG) void k_cg_init(const double* __restrict__ partials, int nblk, int k, FcgState* st,
                                                int stopcrit, int* all_done, int* ndone) {
    __shared__ double s[1][WG];
    double tot[1];
    int c = blockIdx.x;
    sum_partials<1>(partials, nblk, k, c, tot, s);
    if (threadIdx.x == 0) {
        FcgState& S = st[c];
        double l2 = sqrt(tot[0]);
        S.l2 = l2;
        S.delta0 = stopcrit == 1 ? l2 : 1.0;
        S.iter = 0;
        S.done = 0;
        S.pad0 = 0;
        S.delta = 0.0;
        if (l2 == 0.0) {
            S.done = 1;
            S.pad0 = 1;
            if (atomicAdd(ndone, 1) + 1 == k) *all_done = 1;
        }
    }
}

void cg_init(Context* c, const FcgBufs& f, const double* r, int stopcrit) {
    int KP = kpad(f.k);
    k_dots<1><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, r, r, nullptr, nullptr, nullptr, nullptr, f.partials, nullptr);
    BCMGX_LAUNCHED();
    k_cg_init<<<f.k, WG, 0, c->stream>>>(f.partials, f.nblk, f.k, f.st, stopcrit, f.all_done, f.all_done + 1);
    BCMGX_LAUNCHED();
}

// rho = r.u, beta = rho / rho_prev  (CG_HS.cu:89-101)
__global__ __launch_bounds__(WG) void k_cg_beta(const double* __restrict__ partials, int nblk, int k, FcgState* st,
                                                const int* __restrict__ all_done) {
    if (*all_done) return;
    __shared__ double s[1][WG];
    double tot[1];
    int c = blockIdx.x;
    sum_partials<1>(partials, nblk, k, c, tot, s);
    if (threadIdx.x == 0) {
        FcgState& S = st[c];
        if (S.done) return;
        S.alpha = tot[0];
        S.theta = S.alpha / S.delta;
    }
}

// p = 1 u + beta p  (my_axpby, basic.cu:114-130; CG_HS.cu:103-107)
__global__ __launch_bounds__(WG) void k_cg_direction(int n, int k, int KP, const FcgState* __restrict__ st,
                                                     const double* __restrict__ u, double* __restrict__ p,
                                                     const int* __restrict__ all_done) {
    if (*all_done) return;
    int tid = threadIdx.x;
    int R = WG / KP;
    int rl = tid / KP, c = tid % KP;
    if (c >= k || st[c].done) return;
    double be = st[c].theta;
    for (long long i = (long long)blockIdx.x * R + rl; i < n; i += (long long)gridDim.x * R) {
        long long o = i * k + c;
        p[o] = 1.0 * u[o] + be * p[o];
    }
}

void cg_beta_direction(Context* c, const FcgBufs& f, const double* r, const double* u, double* p) {
    int KP = kpad(f.k);
    k_dots<1><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, r, u, nullptr, nullptr, nullptr, nullptr, f.partials, f.all_done);
    BCMGX_LAUNCHED();
    k_cg_beta<<<f.k, WG, 0, c->stream>>>(f.partials, f.nblk, f.k, f.st, f.all_done);
    BCMGX_LAUNCHED();
    k_cg_direction<<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, f.st, u, p, f.all_done);
    BCMGX_LAUNCHED();
}

// (first iteration: rho = u.r) gamma = p.s, alpha = rho / gamma; rho_prev = rho  (CG_HS.cu:112-129)
template <int NV>
__global__ __launch_bounds__(WG) void k_cg_alpha(const double* __restrict__ partials, int nblk, int k, FcgState* st,
                                                 const int* __restrict__ all_done) {
    if (*all_done) return;
    __shared__ double s[NV][WG];
    double tot[NV];
    int c = blockIdx.x;
    sum_partials<NV>(partials, nblk, k, c, tot, s);
    if (threadIdx.x == 0) {
        FcgState& S = st[c];
        if (S.done) return;
        if (NV == 2) S.alpha = tot[1];
        S.gamma = tot[0];
        S.theta2 = S.alpha / S.gamma;
        S.delta = S.alpha;
    }
}

void cg_alpha(Context* c, const FcgBufs& f, const double* p, const double* s, const double* u, const double* r, bool first) {
    int KP = kpad(f.k);
    if (first) {
        k_dots<2><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, p, s, u, r, nullptr, nullptr, f.partials, f.all_done);
        BCMGX_LAUNCHED();
        k_cg_alpha<2><<<f.k, WG, 0, c->stream>>>(f.partials, f.nblk, f.k, f.st, f.all_done);
    } else {
        k_dots<1><<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, p, s, nullptr, nullptr, nullptr, nullptr, f.partials, f.all_done);
        BCMGX_LAUNCHED();
        k_cg_alpha<1><<<f.k, WG, 0, c->stream>>>(f.partials, f.nblk, f.k, f.st, f.all_done);
    }
    BCMGX_LAUNCHED();
}

// x = alpha p + 1 x; r = -alpha s + 1 r (CG_HS.cu:131-132), with the ||r||^2 partials
__global__ __launch_bounds__(WG) void k_cg_update(int n, int k, int KP, const FcgState* __restrict__ st,
                                                  const double* __restrict__ p, const double* __restrict__ sv,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ partials, const int* __restrict__ all_done) {
    if (*all_done) return;
    __shared__ double s[WG];
    int tid = threadIdx.x;
    int R = WG / KP;
    int rl = tid / KP, c = tid % KP;
    double acc = 0.0;
    if (c < k) {
        const FcgState S = st[c];
        double al = S.theta2;
        for (long long i = (long long)blockIdx.x * R + rl; i < n; i += (long long)gridDim.x * R) {
            long long o = i * k + c;
            double rr;
            if (!S.done) {
                x[o] = al * p[o] + 1.0 * x[o];
                rr = -al * sv[o] + 1.0 * r[o];
                r[o] = rr;
            } else {
                rr = r[o];
            }
            acc += rr * rr;
        }
    }
    s[tid] = acc;
    __syncthreads();
    for (int stp = WG / 2; stp >= KP; stp >>= 1) {
        if (tid < stp) s[tid] += s[tid + stp];
        __syncthreads();
    }
    if (tid < k) partials[((long long)blockIdx.x * k + tid) * 3] = s[tid];
}

// l2 = ||r||; converged when l2 < rtol delta0, else the iteration count
// advances and itnlim stops the column (CG_HS.cu:134-163)
__global__ __launch_bounds__(WG) void k_cg_finish(const double* __restrict__ partials, int nblk, int k, FcgState* st,
                                                  double rtol, int itnlim, int* all_done, int* ndone) {
    if (*all_done) return;
    __shared__ double s[1][WG];
    double tot[1];
    int c = blockIdx.x;
    sum_partials<1>(partials, nblk, k, c, tot, s);
    if (threadIdx.x == 0) {
        FcgState& S = st[c];
        if (S.done) return;
        S.l2 = sqrt(tot[0]);
        bool stop = false;
        if (S.l2 < rtol * S.delta0) {
            S.pad0 = 1;
            stop = true;
        } else if (++S.iter >= itnlim) {
            stop = true;
        }
        if (stop) {
            S.done = 1;
            if (atomicAdd(ndone, 1) + 1 == k) *all_done = 1;
        }
    }
}

void cg_update_finish(Context* c, const FcgBufs& f, const double* p, const double* s, double* x, double* r) {
    int KP = kpad(f.k);
    k_cg_update<<<f.nblk, WG, 0, c->stream>>>(f.n, f.k, KP, f.st, p, s, x, r, f.partials, f.all_done);
    BCMGX_LAUNCHED();
    k_cg_finish<<<f.k, WG, 0, c->stream>>>(f.partials, f.nblk, f.k, f.st, f.rtol, f.itnlim, f.all_done, f.all_done + 1);
    BCMGX_LAUNCHED();
}

}  // namespace bcmgx
