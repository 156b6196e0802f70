// afsai.hip -- adaptive factorized sparse approximate inverse, M^-1 = G^T G.
//
// Restates afsai_setup (src/preconditioner/afsai/afsai.cu:287-1090) for one
// GPU.  The reference runs its loop step by step over all rows with separate
// kernels (d_kapGrad_merge, kapGradCUDA_WiP.cu:189-368; d_gatherFullSysAdapt,
// d_gFSA.cu:8-79; cudacholdcMix, choldc.cu:24-77; multiscalar_product;
// check_done, afsai.cu:261-280).  Rows never interact, so here each half of a
// wave64 owns one row for all steps and keeps its pattern, coefficients and
// dense system in LDS:
//   * Kaporin gradient: lane r walks row r of {i} u P (factor 1, then the
//     current coefficients); columns < i are merged in ascending order with a
//     wave min, a column's value is summed in lane order; the largest
//     |gradient| outside P wins, ties to the smallest column; a zero maximum
//     adds nothing and the row is done.
//   * dense system A_s[P,P] (packed upper, column-major) and rhs -A_s[i,P];
//     Cholesky with the reference's operation order (diagonal stored as its
//     reciprocal), lane j computing element (i,j) of step i; triangular
//     solves and DKap = b.x on one lane.
//   * stop when |DKap - DKap_old| < 1e-3 DKap_old or DKap == 0; at most 30
//     entries (AFSAISTEP = 30, AFSAILFIL = 1, AFSAIEPSILON, afsai.cu:41-45).
// A_s = D^-1/2 A D^-1/2 (scalediag + CSRscaleA_0).  createprc (:165-202):
// row i of G = (a_s,ii - DKap_i)^-1/2 [x, 1].
#include "kernels.h"

#include <climits>

namespace bcmgx {

constexpr int AF_MMAX = 30;
constexpr int AF_PACK = AF_MMAX * (AF_MMAX + 1) / 2;
constexpr double AF_EPS = 0.001;
constexpr int AF_WAVES = 2;  // waves per workgroup

__global__ void k_scale_diag(int n, const double* __restrict__ val, const int* __restrict__ dind, double* __restrict__ d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 1.0 / sqrt(val[dind[i]]);
}
__global__ void k_scale_A(int n, const int* __restrict__ row, const int* __restrict__ col, const double* __restrict__ val,
                          const double* __restrict__ d, double* __restrict__ cs) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double di = d[i];
    for (int j = row[i]; j < row[i + 1]; j++) cs[j] = (val[j] * d[col[j]]) * di;
}

// order LDS accesses of the lanes of one wave (no hardware barrier needed)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// min over the 32 lanes of a half wave
__device__ __forceinline__ int half_min_int(int v) {
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
        int u = __shfl_xor(v, o, 32);
        v = u < v ? u : v;
    }
    return v;
}

// Two rows per wave, one per half (32 lanes: the row and its at most 30 pattern
// rows in the gradient): the serial parts (Cholesky diagonals, triangular solves,
// DKap) run for both rows in one instruction stream.  Cross-lane operations are
// always executed by the whole wave; a half whose row is done (or past n) only
// idles through them.  Every floating-point operation of a row is the one of the
// one-row formulation, in the same order.
__global__ __launch_bounds__(64 * AF_WAVES) void k_afsai_rows(int n, const int* __restrict__ row, const int* __restrict__ col,
                                                            const double* __restrict__ cs, int* __restrict__ pat_out,
                                                            double* __restrict__ x_out, int* __restrict__ mrow_out,
                                                            double* __restrict__ dk_out) {
    __shared__ int pat_s[2 * AF_WAVES][AF_MMAX];
    __shared__ double x_s[2 * AF_WAVES][AF_MMAX];
    __shared__ double b_s[2 * AF_WAVES][AF_MMAX];
    __shared__ double L_s[2 * AF_WAVES][AF_PACK];
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int half = lane >> 5, hl = lane & 31;
    int slot = 2 * wid + half;
    long long ri = 2 * ((long long)blockIdx.x * AF_WAVES + wid) + half;
    bool live = ri < n;
    if (__ballot(live) == 0) return;  // whole wave (LDS is per half; only wave-level sync below)
    int i = live ? (int)ri : 0;
    int* pat = pat_s[slot];
    double* xs = x_s[slot];
    double* bb = b_s[slot];
    double* Lp = L_s[slot];
    int mrow = 0;
    double dk = 0.0, dko = 0.0;
    bool active = live;  // the half's row still grows (uniform within the half)
    for (int step = 0; step < AF_MMAX; step++) {
        if (__ballot(active) == 0) break;
        // ---------------- Kaporin gradient over columns < i
        int rr = -1;
        double fac = 0.0;
        if (active) {
            if (hl == 0) { rr = i; fac = 1.0; }
            else if (hl <= mrow) { rr = pat[hl - 1]; fac = xs[hl - 1]; }
        }
        int ptr = 0, end = 0;
        if (rr >= 0) { ptr = row[rr]; end = row[rr + 1]; }
        int pp = 0;  // pattern cursor (uniform within the half)
        int best = -1;
        double bestabs = -1.0, bestval = 0.0;
        while (true) {
            int cc = (rr >= 0 && ptr < end) ? col[ptr] : INT_MAX;
            if (cc >= i) cc = INT_MAX;
            int jmin = half_min_int(cc);
            bool more = jmin != INT_MAX;
            if (__ballot(more) == 0) break;
            bool hit = more && cc == jmin;
            double contrib = hit ? cs[ptr] * fac : 0.0;
            if (hit) ptr++;
            unsigned mh = (unsigned)(__ballot(hit) >> (32 * half));
            // the column's value: the contributions summed in lane order (both halves
            // take as many shuffles as the larger count)
            int cnt = __popc(mh);
            int c0 = __shfl(cnt, 0, 64), c1 = __shfl(cnt, 32, 64);
            int cmax = c0 > c1 ? c0 : c1;
            double acc = 0.0;
            for (int q = 0; q < cmax; q++) {
                int b = mh ? __ffs(mh) - 1 : 0;
                double v = __shfl(contrib, b, 32);
                if (mh) {
                    acc += v;
                    mh &= mh - 1;
                }
            }
            if (more) {
                while (pp < mrow && pat[pp] < jmin) pp++;
                bool inpat = pp < mrow && pat[pp] == jmin;
                double a = inpat ? 0.0 : fabs(acc);
                if (a > bestabs) {
                    bestabs = a;
                    best = jmin;
                    bestval = inpat ? 0.0 : acc;
                }
            }
        }
        if (active && (best < 0 || bestval == 0.0)) active = false;  // no new entry: the row is done
        if (__ballot(active) == 0) break;
        // insert `best` into the sorted pattern (the half's lane 0)
        if (active && hl == 0) {
            int q = mrow;
            while (q > 0 && pat[q - 1] > best) {
                pat[q] = pat[q - 1];
                q--;
            }
            pat[q] = best;
        }
        wave_sync();
        if (active) mrow++;
        // ---------------- dense system A_s[P,P] and rhs
        if (active) {
            for (int t = hl; t < mrow * (mrow + 1) / 2; t += 32) Lp[t] = 0.0;
            if (hl < mrow) bb[hl] = 0.0;
        }
        wave_sync();
        if (active && hl < mrow) {
            int a = hl;
            int r2 = pat[a];
            int jj = row[r2], je = row[r2 + 1];
            for (int b = a; b < mrow; b++) {
                while (jj < je && col[jj] < pat[b]) jj++;
                if (jj == je) break;
                if (col[jj] == pat[b]) Lp[b * (b + 1) / 2 + a] = cs[jj];
            }
        }
        if (active && hl == 0) {
            int jj = row[i], je = row[i + 1], ii = 0;
            while (jj < je && col[jj] < i && ii < mrow) {
                if (pat[ii] > col[jj]) jj++;
                else if (pat[ii] == col[jj]) { bb[ii] = -cs[jj]; jj++; ii++; }
                else ii++;
            }
        }
        wave_sync();
        // ---------------- Cholesky (cudacholdcMix order)
        bool fail = false;
        int mh0 = __shfl(active ? mrow : 0, 0, 64), mh1 = __shfl(active ? mrow : 0, 32, 64);
        int mmax = mh0 > mh1 ? mh0 : mh1;
        for (int ci = 0; ci < mmax; ci++) {
            bool on = active && !fail && ci < mrow;  // uniform within the half
            // diagonal
            double sum = 0.0;
            if (on && hl == 0) {
                sum = Lp[ci * (ci + 1) / 2 + ci];
                for (int k = 0; k < ci; k++) sum -= Lp[ci * (ci + 1) / 2 + k] * Lp[ci * (ci + 1) / 2 + k];
                if (!(sum <= 0.0)) Lp[ci * (ci + 1) / 2 + ci] = 1. / sqrt(sum);
            }
            sum = __shfl(sum, 0, 32);
            if (on && sum <= 0.0) {
                fail = true;
                on = false;
            }
            wave_sync();
            int j = ci + 1 + hl;
            if (on && j < mrow) {
                double s2 = Lp[j * (j + 1) / 2 + ci];
                for (int k = 0; k < ci; k++) s2 -= Lp[ci * (ci + 1) / 2 + k] * Lp[j * (j + 1) / 2 + k];
                Lp[j * (j + 1) / 2 + ci] = s2 * Lp[ci * (ci + 1) / 2 + ci];
            }
            wave_sync();
        }
        // b kept for DKap; x solved in x_s (on breakdown x = b, as the reference leaves it)
        if (active && hl == 0) {
            for (int t = 0; t < mrow; t++) xs[t] = bb[t];
            if (!fail) {
                for (int a = 0; a < mrow; a++) {
                    double s = xs[a];
                    for (int b = a - 1; b >= 0; b--) s -= Lp[a * (a + 1) / 2 + b] * xs[b];
                    xs[a] = s * Lp[a * (a + 1) / 2 + a];
                }
                for (int a = mrow - 1; a >= 0; a--) {
                    double s = xs[a];
                    for (int b = a + 1; b < mrow; b++) s -= Lp[b * (b + 1) / 2 + a] * xs[b];
                    xs[a] = s * Lp[a * (a + 1) / 2 + a];
                }
            }
            double s = 0.0;
            for (int e = 0; e < mrow; e++) s += bb[e] * xs[e];
            dk = s;
        }
        wave_sync();
        dk = __shfl(dk, 0, 32);
        if (active) {
            bool stop = (fabs(dk - dko) < AF_EPS * dko) || dk == 0.0;
            dko = dk;
            if (stop) active = false;
        }
    }
    if (live && hl < mrow) {
        pat_out[(long long)i * AF_MMAX + hl] = pat[hl];
        x_out[(long long)i * AF_MMAX + hl] = xs[hl];
    }
    if (live && hl == 0) {
        mrow_out[i] = mrow;
        dk_out[i] = dk;
    }
}

__global__ void k_g_rowlen(int n, const int* __restrict__ mrow, int* __restrict__ len) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) len[i] = mrow[i] + 1;
}

// createprc (afsai.cu:165-202)
__global__ void k_createprc(int n, const double* __restrict__ cs, const int* __restrict__ dind, const int* __restrict__ pat,
                            const double* __restrict__ xs, const int* __restrict__ mrow, const double* __restrict__ dk,
                            const int* __restrict__ grow, int* __restrict__ gcol, double* __restrict__ gval) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = grow[i], m = mrow[i];
    double sc = cs[dind[i]] - dk[i];
    if (sc > 0.0) {
        sc = 1.0 / sqrt(sc);
        for (int k = 0; k < m; k++) {
            gval[o + k] = sc * xs[(long long)i * AF_MMAX + k];
            gcol[o + k] = pat[(long long)i * AF_MMAX + k];
        }
    } else {
        // negative scaling factor (afsai.cu:187): off-diagonals left at their zero initialisation
        sc = 1.0;
        for (int k = 0; k < m; k++) { gval[o + k] = 0.0; gcol[o + k] = 0; }
    }
    gval[o + m] = sc;
    gcol[o + m] = i;
}

void afsai_build(Context* c, const DCsr& A, DCsr& G, DCsr& GT) {
    int n = A.n;
    BCMGX_CHECK(A.n == A.m, BCMGX_ERR_ARG, "AFSAI needs a square matrix");
    DBuf<int> dind(c, std::max(n, 1));
    diag_index(c, n, A.row.get(), A.col.get(), dind.get());
    {
        std::vector<int> h(n);
        c->d2h(h.data(), dind.get(), n);
        for (int i = 0; i < n; i++) BCMGX_CHECK(h[i] >= 0, BCMGX_ERR_ARG, "AFSAI: row without diagonal entry");
    }
    DBuf<double> d(c, std::max(n, 1)), cs(c, std::max(A.nnz, 1));
    k_scale_diag<<<grid_for(n, WG), WG, 0, c->stream>>>(n, A.val.get(), dind.get(), d.get());
    BCMGX_LAUNCHED();
    k_scale_A<<<grid_for(n, WG), WG, 0, c->stream>>>(n, A.row.get(), A.col.get(), A.val.get(), d.get(), cs.get());
    BCMGX_LAUNCHED();
    DBuf<int> pat(c, (size_t)std::max(n, 1) * AF_MMAX), mrow(c, std::max(n, 1));
    DBuf<double> xs(c, (size_t)std::max(n, 1) * AF_MMAX), dk(c, std::max(n, 1));
    k_afsai_rows<<<grid_for(n, 2 * AF_WAVES), 64 * AF_WAVES, 0, c->stream>>>(n, A.row.get(), A.col.get(), cs.get(), pat.get(), xs.get(),
                                                                      mrow.get(), dk.get());
    BCMGX_LAUNCHED();
    DBuf<int> len(c, std::max(n, 1));
    k_g_rowlen<<<grid_for(n, WG), WG, 0, c->stream>>>(n, mrow.get(), len.get());
    BCMGX_LAUNCHED();
    G.n = G.m = n;
    G.row.reset(c, (size_t)n + 1);
    long long nnz = scan_exclusive(c, len.get(), G.row.get(), n);
    G.nnz = (int)nnz;
    G.col.reset(c, std::max<long long>(nnz, 1));
    G.val.reset(c, std::max<long long>(nnz, 1));
    k_createprc<<<grid_for(n, WG), WG, 0, c->stream>>>(n, cs.get(), dind.get(), pat.get(), xs.get(), mrow.get(), dk.get(),
                                                      G.row.get(), G.col.get(), G.val.get());
    BCMGX_LAUNCHED();
    transpose(c, G, GT);
    c->sync();
}

}  // namespace bcmgx
