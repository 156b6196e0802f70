// krylov.hip -- block-vector kernels of the s-step CG solvers (CGs.cu,
// pipelinedCGs.cu): dot products of several basis vectors with one vector
// (mynddot, op/basic.cu:316-367) reduced in a fixed order, and the tall-skinny
// block updates of op/basic.cu (cdmmv, cdmm, cdmv2v), one thread per row,
// with the column sums taken in column order (the reference adds them with
// atomics).  Blocks are column-major n x m arrays.
#include "kernels.h"

namespace bcmgx {

// partials[(blk * m + j)] = sum over the block's rows of A[:, j] . B[:, j*ldb/n]
__global__ __launch_bounds__(WG) void k_ndot(int n, int m, const double* __restrict__ A, const double* __restrict__ B,
                                             long long ldb, double* __restrict__ partials) {
    __shared__ double s[WG];
    for (int j = 0; j < m; j++) {
        const double* a = A + (long long)j * n;
        const double* b = B + (long long)j * ldb;
        double acc = 0.0;
        for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < n; i += (long long)gridDim.x * WG) acc += a[i] * b[i];
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int st = WG / 2; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[(long long)blockIdx.x * m + j] = s[0];
        __syncthreads();
    }
}

// out[j] = sum of the nblk partials of column j, fixed tree
__global__ __launch_bounds__(WG) void k_ndot_final(int nblk, int m, const double* __restrict__ partials,
                                                   double* __restrict__ out) {
    __shared__ double s[WG];
    int j = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += WG) acc += partials[(long long)b * m + j];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int st = WG / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = s[0];
}

void block_dots(Context* c, int n, int m, const double* A, const double* B, long long ldb, double* partials, int nblk,
                double* out) {
    if (m <= 0) return;
    k_ndot<<<nblk, WG, 0, c->stream>>>(n, m, A, B, ldb, partials);
    BCMGX_LAUNCHED();
    k_ndot_final<<<m, WG, 0, c->stream>>>(nblk, m, partials, out);
    BCMGX_LAUNCHED();
}

// cdmmv (op/basic.cu:387-447): C[:, j] = C[:, j] + A B[:, j]; E += f (C alpha)
__global__ __launch_bounds__(WG) void k_dmmv(int n, int s, const double* __restrict__ A, const double* __restrict__ B,
                                             double* __restrict__ C, const double* __restrict__ al, double* __restrict__ E,
                                             double f) {
    long long i = (long long)blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int j = 0; j < s; j++) {
        double ro = C[(long long)j * n + i];
        for (int k = 0; k < s; k++) ro += A[(long long)k * n + i] * B[k + s * j];
        C[(long long)j * n + i] = ro;
        double t = ro * al[j];
        acc = j == 0 ? t : acc + t;
    }
    E[i] += f * acc;
}

void block_dmmv(Context* c, int n, int s, const double* A, const double* B, double* C, const double* al, double* E, double f) {
    if (n <= 0) return;
    k_dmmv<<<grid_for(n, WG), WG, 0, c->stream>>>(n, s, A, B, C, al, E, f);
    BCMGX_LAUNCHED();
}

// cdmm (op/basic.cu:163-183): C = C + (A B)
__global__ __launch_bounds__(WG) void k_dmm(int n, int s, const double* __restrict__ A, const double* __restrict__ B,
                                            double* __restrict__ C) {
    long long i = (long long)blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    for (int j = 0; j < s; j++) {
        double ro = 0.0;
        for (int k = 0; k < s; k++) ro += A[(long long)k * n + i] * B[k + s * j];
        C[(long long)j * n + i] += ro;
    }
}

void block_dmm(Context* c, int n, int s, const double* A, const double* B, double* C) {
    if (n <= 0) return;
    k_dmm<<<grid_for(n, WG), WG, 0, c->stream>>>(n, s, A, B, C);
    BCMGX_LAUNCHED();
}

// cdmv2v (op/basic.cu:23-34): D = C + f (A b)
__global__ __launch_bounds__(WG) void k_dmv2v(int n, int s, const double* __restrict__ A, const double* __restrict__ bv,
                                              const double* __restrict__ C, double f, double* __restrict__ D) {
    long long i = (long long)blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    double ro = 0.0;
    for (int k = 0; k < s; k++) ro += A[(long long)k * n + i] * bv[k];
    D[i] = C[i] + f * ro;
}

void block_dmv2v(Context* c, int n, int s, const double* A, const double* bv, const double* C, double f, double* D) {
    if (n <= 0) return;
    k_dmv2v<<<grid_for(n, WG), WG, 0, c->stream>>>(n, s, A, bv, C, f, D);
    BCMGX_LAUNCHED();
}

// dst[i * ds] = src[i * ss], i < n (one column of an interleaved block in or out)
__global__ __launch_bounds__(WG) void k_strided_copy(double* __restrict__ dst, int ds, const double* __restrict__ src, int ss,
                                                     int n) {
    long long i = (long long)blockIdx.x * WG + threadIdx.x;
    if (i < n) dst[i * ds] = src[i * ss];
}

void strided_copy(Context* c, double* dst, int ds, const double* src, int ss, int n) {
    if (n <= 0) return;
    k_strided_copy<<<grid_for(n, WG), WG, 0, c->stream>>>(dst, ds, src, ss, n);
    BCMGX_LAUNCHED();
}

}  // namespace bcmgx
