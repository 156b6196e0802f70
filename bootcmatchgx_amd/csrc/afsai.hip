// afsai.hip -- adaptive factorized sparse approximate inverse, M^-1 = G^T G.
//
// Restates afsai_setup (src/preconditioner/afsai/afsai.cu:287-1090) for one
// GPU.  The reference runs its loop step by step over all rows with separate
// kernels (d_kapGrad_merge, kapGradCUDA_WiP.cu:189-368; d_gatherFullSysAdapt,
// d_gFSA.cu:8-79; cudacholdcMix, choldc.cu:24-77; multiscalar_product;
// check_done, afsai.cu:261-280).  Rows never interact, so here each half of a
// wave64 owns one row for all steps and keeps its pattern, coefficients and
// dense system in LDS:
//   * Kaporin gradient over the columns < i of the rows {i} u P (factor 1, then
//     the current coefficients), each column's value summed in lane order (row
//     i, then the pattern rows in column order); the largest |gradient| outside
//     P wins, ties to the smallest column; a zero maximum adds nothing and the
//     row is done.  Two ways to the same sums:
//       - column table (first launch): a row's columns get slots of a 256-slot
//         hash table in LDS once, when the row joins; each step then adds each
//         row's products into the slots, one row after the other in lane order,
//         and reduces the table for the largest entry.  Rows whose entries left
//         of the diagonal are more than 32, not strictly increasing, or would
//         take the table past 192 columns are handed to
//       - the merge (second launch, those rows only): lane r walks row r, the
//         columns merged in ascending order with a wave min, a column's
//         contributions summed over the lanes in order.
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
constexpr int AF_TS = 256;   // column-table slots per row (slot indices are bytes)
constexpr int AF_TLIM = 192; // columns a table takes before its row goes to the merge
constexpr int AF_EMAX = 32;  // entries left of the diagonal a row may bring into the table
static_assert(AF_TS == 256 && AF_TLIM < AF_TS && AF_TS <= AF_PACK, "table slots: bytes, never full, sums in L_s");

// 8-bit multiplicative hash of a column
__device__ __forceinline__ int af_hash(int c) { return (int)(((unsigned)c * 2654435761u) >> 24); }

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

// s - P(0) Q(0) - P(1) Q(1) - ... (s + ... when ADD) over cnt terms, each product rounded and
// then subtracted (added) in term order; the operands of eight terms are loaded before their
// eight dependent updates, so the chain waits for the LDS once per eight terms, not per term
template <bool ADD, typename FP, typename FQ>
__device__ __forceinline__ double af_chain(double s, int cnt, FP P, FQ Q) {
    int t = 0;
    for (; t + 8 <= cnt; t += 8) {
        double p[8], q[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            p[u] = P(t + u);
            q[u] = Q(t + u);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) s = ADD ? s + p[u] * q[u] : s - p[u] * q[u];
    }
    for (; t < cnt; t++) s = ADD ? s + P(t) * Q(t) : s - P(t) * Q(t);
    return s;
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
// one-row formulation, in the same order.  TABLE: the gradient from the column
// table, rows it cannot take appended to redo; otherwise the merge, over the rows
// of rows_in (all rows when null).
template <bool TABLE>
__global__ __launch_bounds__(64 * AF_WAVES) void k_afsai_rows(int n, const int* __restrict__ rows_in, int nrows,
                                                            const int* __restrict__ row, const int* __restrict__ col,
                                                            const double* __restrict__ cs, int* __restrict__ pat_out,
                                                            double* __restrict__ x_out, int* __restrict__ mrow_out,
                                                            double* __restrict__ dk_out, int* __restrict__ redo,
                                                            int* __restrict__ nredo) {
    constexpr int NH = 2 * AF_WAVES, TH = TABLE ? NH : 1;
    __shared__ int pat_s[NH][AF_MMAX];
    __shared__ double x_s[NH][AF_MMAX];
    __shared__ double b_s[NH][AF_MMAX];
    __shared__ double L_s[NH][AF_PACK];  // TABLE: also the slots' sums during the gradient
    // TABLE: slot columns (-1 empty, sign bit set once in the pattern), each row's slots and
    // (first entry, entries left of the diagonal) by the order rows joined (0: row i), the
    // joining index of each pattern position, the table's column count
    __shared__ int hcol_s[TH][TABLE ? AF_TS : 1];
    __shared__ unsigned char eslot_s[TH][TABLE ? AF_MMAX + 1 : 1][TABLE ? AF_EMAX : 1];
    __shared__ int2 rbc_s[TH][TABLE ? AF_MMAX + 1 : 1];
    __shared__ unsigned char ord_s[TH][TABLE ? AF_MMAX : 1];
    __shared__ int hcnt_s[TH];
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int half = lane >> 5, hl = lane & 31;
    int slot = 2 * wid + half;
    long long ri = 2 * ((long long)blockIdx.x * AF_WAVES + wid) + half;
    bool live = ri < (rows_in ? (long long)nrows : (long long)n);
    if (__ballot(live) == 0) return;  // whole wave (LDS is per half; only wave-level sync below)
    int i = live ? (rows_in ? rows_in[ri] : (int)ri) : 0;
    int* pat = pat_s[slot];
    double* xs = x_s[slot];
    double* bb = b_s[slot];
    double* Lp = L_s[slot];
    int ts = TABLE ? slot : 0;
    int* hcol = hcol_s[ts];
    int2* rbc = rbc_s[ts];
    unsigned char* ord = ord_s[ts];
    int mrow = 0;
    double dk = 0.0, dko = 0.0;
    bool active = live;  // the half's row still grows (uniform within the half)
    bool handed = false;  // TABLE: the row goes to the merge
    // TABLE: row r's entries left of the diagonal into the table as joining index q (whole
    // wave; `on` uniform per half); false when the table cannot take them
    auto table_row = [&](bool on, int q, int r) -> bool {
        int beg = 0, end = 0;
        if (on) {
            beg = row[r];
            end = row[r + 1];
        }
        int cc = (on && beg + hl < end) ? col[beg + hl] : INT_MAX;
        unsigned ge = (unsigned)(__ballot(cc >= i) >> (32 * half));
        int cnt = ge ? __ffs(ge) - 1 : 32;
        int nx = __shfl_down(cc, 1, 32);
        unsigned bad = (unsigned)(__ballot(hl + 1 < cnt && nx <= cc) >> (32 * half));
        bool ok = on && bad == 0 && hcnt_s[ts] + cnt <= AF_TLIM;
        if (ok && cnt == 32 && beg + 32 < end && col[beg + 32] < i) ok = false;  // more than 32
        if (ok && hl < cnt) {
            // an empty slot always remains (at most AF_TLIM < AF_TS columns): the probe ends
            int s = af_hash(cc);
            while (true) {
                int v = hcol[s];
                if ((v & INT_MAX) == cc) break;
                if (v == -1) {
                    int old = atomicCAS(&hcol[s], -1, cc);
                    if (old == -1) {
                        atomicAdd(&hcnt_s[ts], 1);
                        break;
                    }
                    if ((old & INT_MAX) == cc) break;
                }
                s = (s + 1) & (AF_TS - 1);
            }
            eslot_s[ts][q][hl] = (unsigned char)s;
        }
        if (ok && hl == 0) rbc[q] = make_int2(beg, cnt);
        wave_sync();
        return ok;
    };
    if constexpr (TABLE) {
        for (int u = hl; u < AF_TS; u += 32) hcol[u] = -1;
        if (hl == 0) hcnt_s[ts] = 0;
        wave_sync();
        if (!table_row(live, 0, i) && live) {
            handed = true;
            active = false;
        }
    }
    for (int step = 0; step < AF_MMAX; step++) {
        if (__ballot(active) == 0) break;
        int best = -1, bslot = -1;
        if constexpr (TABLE) {
            // ---------------- Kaporin gradient from the column table
            double* hv = Lp;
            if (active)
                for (int u = hl; u < AF_TS; u += 32) hv[u] = 0.0;
            wave_sync();
            int m0 = __shfl(active ? mrow : -1, 0, 64), m1 = __shfl(active ? mrow : -1, 32, 64);
            int lmax = m0 > m1 ? m0 : m1;
            // each row's products into its columns' slots, the rows in lane order; the adds
            // to one slot keep that order (LDS operations of a wave complete in order, and a
            // lane's own adds may alias, so the compiler keeps them in program order)
            for (int l0 = 0; l0 <= lmax; l0 += 8) {
                double v[8];
                int sl[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    int L = l0 + u;
                    sl[u] = -1;
                    v[u] = 0.0;
                    if (active && L <= mrow) {
                        int q = L == 0 ? 0 : ord[L - 1];
                        int2 bc = rbc[q];
                        if (hl < bc.y) {
                            double fac = L == 0 ? 1.0 : xs[L - 1];
                            v[u] = cs[bc.x + hl] * fac;
                            sl[u] = eslot_s[ts][q][hl];
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (sl[u] >= 0) hv[sl[u]] += v[u];
            }
            wave_sync();
            // the largest |sum| outside the pattern, ties to the smallest column (a NaN never wins)
            double ba = -1.0;
            int bc = INT_MAX, bs = -1;
            if (active) {
                for (int u = hl; u < AF_TS; u += 32) {
                    int cv = hcol[u];
                    if (cv >= 0) {
                        double a = fabs(hv[u]);
                        if (a > ba || (a == ba && cv < bc)) {
                            ba = a;
                            bc = cv;
                            bs = u;
                        }
                    }
                }
            }
#pragma unroll
            for (int o = 16; o > 0; o >>= 1) {
                double oa = __shfl_xor(ba, o, 32);
                int oc = __shfl_xor(bc, o, 32), os = __shfl_xor(bs, o, 32);
                if (oa > ba || (oa == ba && oc < bc)) {
                    ba = oa;
                    bc = oc;
                    bs = os;
                }
            }
            if (active && ba > 0.0) {
                best = bc;
                bslot = bs;
            }
            if (active && best < 0) active = false;  // no new entry: the row is done
        } else {
            // ---------------- Kaporin gradient over columns < i, by merging the rows
            int rr = -1;
            double fac = 0.0;
            if (active) {
                if (hl == 0) { rr = i; fac = 1.0; }
                else if (hl <= mrow) { rr = pat[hl - 1]; fac = xs[hl - 1]; }
            }
            int ptr = 0, end = 0;
            if (rr >= 0) { ptr = row[rr]; end = row[rr + 1]; }
            int pp = 0;  // pattern cursor (uniform within the half)
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
        }
        if (__ballot(active) == 0) break;
        // insert `best` into the sorted pattern (the half's lane 0)
        if (active && hl == 0) {
            int q = mrow;
            while (q > 0 && pat[q - 1] > best) {
                pat[q] = pat[q - 1];
                if constexpr (TABLE) ord[q] = ord[q - 1];
                q--;
            }
            pat[q] = best;
            if constexpr (TABLE) {
                ord[q] = (unsigned char)(mrow + 1);
                hcol[bslot] |= INT_MIN;
            }
        }
        wave_sync();
        if (active) mrow++;
        if constexpr (TABLE) {
            if (!table_row(active, mrow, best) && active) {
                handed = true;
                active = false;
            }
        }
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
            const double* lc = Lp + ci * (ci + 1) / 2;
            double sum = 0.0;
            if (on && hl == 0) {
                auto Lc = [&](int k) { return lc[k]; };
                sum = af_chain<false>(lc[ci], ci, Lc, Lc);
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
                const double* lj = Lp + j * (j + 1) / 2;
                double s2 = af_chain<false>(lj[ci], ci, [&](int k) { return lc[k]; }, [&](int k) { return lj[k]; });
                Lp[j * (j + 1) / 2 + ci] = s2 * lc[ci];
            }
            wave_sync();
        }
        // b kept for DKap; x solved in x_s (on breakdown x = b, as the reference leaves it)
        if (active && hl == 0) {
            for (int t = 0; t < mrow; t++) xs[t] = bb[t];
            if (!fail) {
                // L y = b, row a's terms b = a-1 down to 0; then L^T x = y, terms b = a+1 up
                for (int a = 0; a < mrow; a++) {
                    const double* la = Lp + a * (a + 1) / 2;
                    double s = af_chain<false>(xs[a], a, [&](int t) { return la[a - 1 - t]; },
                                               [&](int t) { return xs[a - 1 - t]; });
                    xs[a] = s * la[a];
                }
                for (int a = mrow - 1; a >= 0; a--) {
                    double s = af_chain<false>(xs[a], mrow - 1 - a,
                                               [&](int t) { return Lp[(a + 1 + t) * (a + 2 + t) / 2 + a]; },
                                               [&](int t) { return xs[a + 1 + t]; });
                    xs[a] = s * Lp[a * (a + 1) / 2 + a];
                }
            }
            dk = af_chain<true>(0.0, mrow, [&](int e) { return bb[e]; }, [&](int e) { return xs[e]; });
        }
        wave_sync();
        dk = __shfl(dk, 0, 32);
        if (active) {
            bool stop = (fabs(dk - dko) < AF_EPS * dko) || dk == 0.0;
            dko = dk;
            if (stop) active = false;
        }
    }
    if (handed) {
        if (hl == 0) redo[atomicAdd(nredo, 1)] = i;
        return;
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
    DBuf<int> redo(c, std::max(n, 1)), nredo(c, 1);
    fill<int>(c, nredo.get(), 1, 0);
    k_afsai_rows<true><<<grid_for(n, 2 * AF_WAVES), 64 * AF_WAVES, 0, c->stream>>>(
        n, nullptr, 0, A.row.get(), A.col.get(), cs.get(), pat.get(), xs.get(), mrow.get(), dk.get(), redo.get(), nredo.get());
    BCMGX_LAUNCHED();
    int nr = c->read1(nredo.get());  // rows the column table could not take
    if (nr > 0) {
        k_afsai_rows<false><<<grid_for(nr, 2 * AF_WAVES), 64 * AF_WAVES, 0, c->stream>>>(
            n, redo.get(), nr, A.row.get(), A.col.get(), cs.get(), pat.get(), xs.get(), mrow.get(), dk.get(), nullptr, nullptr);
        BCMGX_LAUNCHED();
    }
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
