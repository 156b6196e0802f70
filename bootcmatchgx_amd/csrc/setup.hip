// setup.hip -- AMG setup kernels: matching weights, Suitor matching,
// prolongator, transposes and the Galerkin SpGEMM.
//
// Everything here is bit-reproducible and agrees bit-for-bit with the CPU
// oracle: floating-point expressions are written in the reference's operation
// order and compiled without FMA contraction, and no floating-point value is
// ever accumulated through atomics.
#include "kernels.h"

#include <algorithm>
#include <cfloat>
#include <climits>

namespace bcmgx {

// ------------------------------------------------------------ diagonals ---
// Row-local passes over a CSR matrix with coalesced loads: a block stages the
// entries of its RS_ROWS consecutive rows in LDS (when they fit), then thread r
// walks row r in column order.  Rows of a block holding more than RS_CAP
// entries are walked straight from global memory.
constexpr int RS_ROWS = 128;
constexpr int RS_CAP = 4096;

// dst[0, cnt) = src[0, cnt) by the block's RS_ROWS threads, each with
// RS_UNROLL loads in flight (loads of clamped, always valid indices, stored
// only when in range); a strided copy loop waits on every load before its
// LDS store
constexpr int RS_UNROLL = 8;
template <typename T>
__device__ __forceinline__ void stage_rows(const T* __restrict__ src, T* dst, int cnt) {
    for (int t0 = 0; t0 < cnt; t0 += RS_ROWS * RS_UNROLL) {
        T v[RS_UNROLL];
#pragma unroll
        for (int u = 0; u < RS_UNROLL; u++) {
            int t = t0 + u * RS_ROWS + (int)threadIdx.x;
            v[u] = src[t < cnt ? t : cnt - 1];
        }
#pragma unroll
        for (int u = 0; u < RS_UNROLL; u++) {
            int t = t0 + u * RS_ROWS + (int)threadIdx.x;
            if (t < cnt) dst[t] = v[u];
        }
    }
}

__global__ __launch_bounds__(RS_ROWS) void k_diag_index_s(int n, const int* __restrict__ row, const int* __restrict__ col,
                                                     int* __restrict__ dind) {
    __shared__ int sc[RS_CAP];
    int r0 = blockIdx.x * RS_ROWS;
    int rend = r0 + RS_ROWS < n ? r0 + RS_ROWS : n;
    int base = row[r0], cnt = row[rend] - base;
    bool staged = cnt <= RS_CAP;
    if (staged) stage_rows(col + base, sc, cnt);
    __syncthreads();
    int i = r0 + threadIdx.x;
    if (i >= n) return;
    int d = -1, s = row[i], e = row[i + 1];
    if (staged) {
        for (int j = s; j < e; j++)
            if (sc[j - base] == i) { d = j; break; }
    } else {
        for (int j = s; j < e; j++)
            if (col[j] == i) { d = j; break; }
    }
    dind[i] = d;
}

__global__ __launch_bounds__(RS_ROWS) void k_abs_row_sum_s(int n, const int* __restrict__ row, const double* __restrict__ val,
                                                      double* __restrict__ M) {
    __shared__ double sv[RS_CAP];
    int r0 = blockIdx.x * RS_ROWS;
    int rend = r0 + RS_ROWS < n ? r0 + RS_ROWS : n;
    int base = row[r0], cnt = row[rend] - base;
    bool staged = cnt <= RS_CAP;
    if (staged) stage_rows(val + base, sv, cnt);
    __syncthreads();
    int i = r0 + threadIdx.x;
    if (i >= n) return;
    double t = 0.0;
    int s = row[i], e = row[i + 1];
    if (staged) {
        for (int j = s; j < e; j++) t += fabs(sv[j - base]);
    } else {
        for (int j = s; j < e; j++) t += fabs(val[j]);
    }
    M[i] = t;
}

void diag_index(Context* c, int n, const int* row, const int* col, int* dind) {
    if (n <= 0) return;
    k_diag_index_s<<<grid_for(n, RS_ROWS), RS_ROWS, 0, c->stream>>>(n, row, col, dind);
    BCMGX_LAUNCHED();
}

// CSRm::absoluteRowSum (CSR.cu:1362-1397), column order (k_abs_row_sum_s above)
void abs_row_sum(Context* c, int n, const int* row, const double* val, double* M) {
    if (n <= 0) return;
    k_abs_row_sum_s<<<grid_for(n, RS_ROWS), RS_ROWS, 0, c->stream>>>(n, row, val, M);
    BCMGX_LAUNCHED();
}

// set_l1j (l1jacobi.cu:484-493): mydiag then add_absrowsum_nodiag
__global__ __launch_bounds__(RS_ROWS) void k_l1_diag(int n, const int* __restrict__ row, const int* __restrict__ col,
                                                     const double* __restrict__ val, double* __restrict__ d) {
    __shared__ int sc[RS_CAP];
    __shared__ double sv[RS_CAP];
    int r0 = blockIdx.x * RS_ROWS;
    int rend = r0 + RS_ROWS < n ? r0 + RS_ROWS : n;
    int base = row[r0], cnt = row[rend] - base;
    bool staged = cnt <= RS_CAP;
    if (staged) {
        stage_rows(col + base, sc, cnt);
        stage_rows(val + base, sv, cnt);
    }
    __syncthreads();
    int i = r0 + threadIdx.x;
    if (i >= n) return;
    double dg = 0.0, s = 0.0;
    for (int j = row[i]; j < row[i + 1]; j++)
        if ((staged ? sc[j - base] : col[j]) == i) dg = staged ? sv[j - base] : val[j];
    for (int j = row[i]; j < row[i + 1]; j++)
        if ((staged ? sc[j - base] : col[j]) != i) s += fabs(staged ? sv[j - base] : val[j]);
    d[i] = dg + s;
}
void l1_diag(Context* c, int n, const int* row, const int* col, const double* val, double* d) {
    if (n <= 0) return;
    k_l1_diag<<<grid_for(n, RS_ROWS), RS_ROWS, 0, c->stream>>>(n, row, col, val, d);
    BCMGX_LAUNCHED();
}

// ------------------------------------------------------ matching weights ---
// _makeC (matching.cu:113-135): C_i = a_ii w_i w_i
__global__ void k_make_C(int n, const double* __restrict__ val, const int* __restrict__ dind, const double* __restrict__ w,
                         double* __restrict__ C) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    C[i] = val[dind[i]] * w[i] * w[i];
}

// The matching graph W in one pass over A: for every off-diagonal entry, AH
// (_makeAH_warp, matching.cu:66-111), its weight (_make_w, :219-229) and the
// symmetrisation W(i, c) <- W(c, i) for c < i (_write_T_warp, :29-64); W lies
// on A's pattern without the diagonal (row i of W starts at row[i] - i).
//
//   AH_ij = 1 - 2 a_ij w_i w_j / (C_i + C_j)   (DBL_EPSILON when C_i + C_j <= DBL_EPSILON)
//   W_ij  = log |AH_ij|, or log(SUITOR_EPS / (0.999 m)) when AH_ij = 0,
//
// where m is the least |AH| over all entries, clamped up to DBL_EPSILON
// (toMaximumProductMatrix, :231-250).  m only enters the weight of an entry
// whose AH is 0, and then m = 0 itself, so that weight is always
// log(SUITOR_EPS / (0.999 DBL_EPSILON)): no pass over AH for its minimum.
//
// A lower entry (c < i) takes the weight of the mirrored entry (c, i), whose AH
// is computed here from a_ci: C_i + C_c and w_i w_c are the same bits in either
// operand order, so it equals what row c computes for that entry.  A mirrored
// pair of positive weight -- only such a pair can carry a proposal
// (find_Candidate, suitor.cu:91-94) -- is recorded both ways in mir (filled
// with -1 beforehand: other entries keep it).
//
// Preference ranks.  Vertex i orders its neighbours c by (W(i,c), c) -- the
// Suitor tie rule "heavier, then higher id" (suitor.cu:108,115,136-140).
// rank[e] = position of entry e of row i in that order (0: the least liked).
// A row of at most L off-diagonal entries keeps one entry per lane, staged in
// LDS, where the group reads every entry of the row by broadcast; a longer row
// reads its weights back from W after the block's barrier (the group wrote them).
constexpr double SUITOR_EPS = 3.885780586188048e-10;  // suitor.h:8

__device__ __forceinline__ double ah_of(double a, double wt, double norm) {
    return norm > DBL_EPSILON ? 1. - ((2. * a * wt) / norm) : DBL_EPSILON;
}

__device__ __forceinline__ double weight_of(double v) {
    double s = fabs(v);
    return log(s != 0.0 ? s : SUITOR_EPS / (0.999 * DBL_EPSILON));
}

__device__ __forceinline__ int lower_bound_dev(const int* a, int len, int key) {
    int lo = 0, hi = len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <int L>
__global__ __launch_bounds__(WG) void k_weights(int n, const int* __restrict__ row, const int* __restrict__ col,
                                                const double* __restrict__ val, const int* __restrict__ dind,
                                                const double* __restrict__ w, const double* __restrict__ C,
                                                int* __restrict__ Wcol, double* W, int* __restrict__ mir,
                                                int* __restrict__ rank) {
    __shared__ double sw[WG];
    __shared__ int sc[WG];
    int tid = (int)threadIdx.x;
    int i = (blockIdx.x * WG + tid) / L;
    int lane = tid % L;
    int gb = tid - lane;
    int s = 0, e = 1, d = 0;
    double Ci = 0.0, wi = 0.0;
    if (i < n) {
        s = row[i];
        e = row[i + 1];
        d = dind[i];
        Ci = C[i];
        wi = w[i];
    }
    int si = s - i;     // W's row start
    int len = e - s - 1;  // off-diagonal entries
    double wa = 0.0;    // this lane's first entry (the only one when len <= L)
    int ca = 0;
    for (int t = lane; t < len; t += L) {
        int j = s + t + (s + t >= d ? 1 : 0);  // A's position of the t-th off-diagonal entry
        int o = si + t;
        int cc = col[j];
        double a = val[j];
        int src = o;
        if (cc < i) {
            int b = row[cc], lc = row[cc + 1] - b;
            // a symmetric stencil puts the mirror at the reflected position of a
            // row of the same length: one probe, whose value is read with it
            // (b + p < row[cc + 1] <= j: in range), else a binary search
            int p = lc - 1 - (j - s);
            bool inr = lc == e - s && p >= 0 && p < lc;
            int q = inr ? b + p : j;
            int cq = col[q];
            double aq = val[q];
            if (!(inr && cq == i)) {
                p = lower_bound_dev(col + b, lc, i);
                q = p < lc && col[b + p] == i ? b + p : -1;
                aq = q >= 0 ? val[q] : a;
            }
            if (q >= 0) {
                a = aq;
                src = q - cc - (q > dind[cc] ? 1 : 0);  // the mirror's place in W
            }
        }
        double wv = weight_of(ah_of(a, w[cc] * wi, C[cc] + Ci));
        Wcol[o] = cc;
        W[o] = wv;
        if (src != o && wv > 0) {
            mir[o] = src;
            mir[src] = o;
        }
        if (t == lane) {
            wa = wv;
            ca = cc;
        }
    }
    sw[tid] = wa;
    sc[tid] = ca;
    __syncthreads();
    if (i >= n) return;
    if (len <= L) {  // uniform over the group
        if (lane < len) {
            int r = 0;
            for (int b = 0; b < len; b++) {
                double wb = sw[gb + b];
                int cb = sc[gb + b];
                r += (wb < wa || (wb == wa && cb < ca)) ? 1 : 0;
            }
            rank[si + lane] = r;
        }
        return;
    }
    for (int a = si + lane; a < si + len; a += L) {
        double wv = W[a];
        int cw = Wcol[a];
        int r = 0;
        for (int b = si; b < si + len; b++) {
            double wb = W[b];
            r += (wb < wv || (wb == wv && Wcol[b] < cw)) ? 1 : 0;
        }
        rank[a] = r;
    }
}

// Preference lists.  pref[Wrow[u] + j] = {y, key} for the j-th best neighbour
// y of u in u's own order (the reverse of rank: heaviest first, ties to the
// higher id), with key = 1 + the rank of u among y's neighbours.  key = 0 marks
// a neighbour u never proposes to: non-positive weight (find_Candidate,
// suitor.cu:91-94) or no mirrored entry (u is unknown to y); mir holds only
// the pairs of positive weight.
template <int G>
__global__ __launch_bounds__(WG) void k_pref(int n, const int* __restrict__ Wrow, const int* __restrict__ Wcol,
                                             const int* __restrict__ rank, const int* __restrict__ mir,
                                             int2* __restrict__ pref) {
    int u = (blockIdx.x * WG + (int)threadIdx.x) / G;
    int lane = (int)threadIdx.x % G;
    if (u >= n) return;
    int s = Wrow[u], e = Wrow[u + 1];
    for (int a = s + lane; a < e; a += G) {
        int y = Wcol[a];
        int q = mir[a];  // u's entry in y's row, -1 when absent or not positive
        int key = (y >= 0 && y < n && q >= 0) ? rank[q] + 1 : 0;
        pref[e - 1 - rank[a]] = make_int2(y, key);
    }
}

// lanes per proposing vertex (the candidates one ballot looks at)
constexpr int SUITOR_LANES = 4;

// Lock-free Suitor.  S[y] = (1 + rank of y's current suitor) << 32 | suitor
// (0: none).  A vertex proposes to its best neighbour -- positive weight,
// heaviest, ties to the higher id (find_Candidate, suitor.cu:80-147) -- that
// would prefer it to its current suitor; atomicMax installs the offer and
// returns the displaced suitor, which this thread then re-proposes
// (set_suitor, suitor.cu:51-78, without the spin locks).  Every failed
// atomicMax means another offer raised S[y], and offers only rise, so the grid
// drains.  The fixed point is the unique stable (greedy) matching of the total
// order (W, max id, min id), whatever the interleaving.
//
// One launch is one round.  Each G-lane group takes a vertex from the round's
// work list (all vertices in round 0) and follows its proposal chain for at
// most `maxsteps` proposals; a vertex still unsettled then goes to the next
// round's list.  Bounding the steps keeps one long displacement chain from
// serialising a whole launch: every round works on all pending vertices in
// parallel with fresh offers.
//
// The rounds are bound by the throughput of the device-scope suitor-word
// reads, so each vertex keeps a cursor into its preference list: pos[v] is the
// list position of v's last offer.  A neighbour that would not accept v (its
// suitor word beats v's key there) never will, since suitor words only rise;
// so v resumes after its cursor and reads only the suitor words of neighbours
// it has not yet been refused by.  The group reads G list entries and G suitor
// words at a time and takes the first that would accept (one ballot).  A stale
// read of a cursor (another XCD's write not yet visible) only lowers it, which
// costs reads, never a candidate.
template <int G>
__global__ __launch_bounds__(WG) void k_suitor(int n, const int* __restrict__ Wrow, const int2* __restrict__ pref,
                                               int* pos, unsigned long long* S, const int* __restrict__ in_list,
                                               const int* __restrict__ in_count, int* __restrict__ out_list,
                                               int* __restrict__ out_count, int maxsteps) {
    static_assert(G <= 32, "group mask");
    int lane = (int)threadIdx.x % G;
    int gbase = ((int)threadIdx.x & 63) - lane;  // the group's first lane in its wave
    int ngroups = gridDim.x * (WG / G);
    int total = in_list ? *in_count : n;
    for (int it = (blockIdx.x * WG + (int)threadIdx.x) / G; it < total; it += ngroups) {
        // round 0 visits the vertices from the highest id down: a suitee prefers
        // the higher suitor, so offers made in this order are rarely displaced
        int cur = in_list ? in_list[it] : n - 1 - it;
        int p = in_list ? pos[cur] + 1 : Wrow[cur];  // next list position to try
        int e = Wrow[cur + 1];
        int steps = 0;
        while (cur >= 0) {  // uniform inside the group
            if (steps == maxsteps) {
                if (lane == 0) {
                    pos[cur] = p - 1;
                    out_list[atomicAdd(out_count, 1)] = cur;
                }
                break;
            }
            steps++;
            int best = -1, bkey = 0, bidx = -1;
            for (; p < e; p += G) {  // G candidates at a time, best first
                int2 pk = make_int2(-1, 0);
                if (p + lane < e) pk = pref[p + lane];
                bool ok = false;
                // (the range test only matters for rows with repeated columns,
                // which leave slots of the list unwritten)
                if (pk.y > 0 && (unsigned)pk.x < (unsigned)n) {
                    unsigned long long sy = __hip_atomic_load(&S[pk.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = (int)(sy >> 32) < pk.y;
                }
                unsigned long long m = (__ballot(ok) >> gbase) & ((1ull << G) - 1);
                if (m) {
                    int f = __ffsll(m) - 1;
                    best = __shfl(pk.x, f, G);
                    bkey = __shfl(pk.y, f, G);
                    bidx = p + f;
                    break;
                }
            }
            if (best < 0) break;  // no neighbour left that would accept cur
            unsigned olo = 0, ohi = 0;
            if (lane == 0) {
                pos[cur] = bidx;
                unsigned long long offer = ((unsigned long long)(unsigned)bkey << 32) | (unsigned)cur;
                unsigned long long old = atomicMax(&S[best], offer);
                olo = (unsigned)old;
                ohi = (unsigned)(old >> 32);
            }
            olo = __shfl(olo, 0, G);
            ohi = __shfl(ohi, 0, G);
            if ((int)ohi < bkey) {  // installed; the displaced suitor (if any) resumes after its cursor
                cur = ohi > 0 ? (int)olo : -1;
                if (cur >= 0) {
                    p = pos[cur] + 1;
                    e = Wrow[cur + 1];
                }
            } else {
                p = bidx + 1;  // refused: a better offer holds that neighbour
            }
        }
    }
}

// cursors before round 0: every list is untried
__global__ void k_pos_init(int n, const int* __restrict__ Wrow, int* __restrict__ pos) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) pos[v] = Wrow[v] - 1;
}

__global__ void k_suitor_decode(int n, const unsigned long long* __restrict__ S, int* __restrict__ M) {
    int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= n) return;
    unsigned long long s = S[y];
    M[y] = (s >> 32) ? (int)(unsigned)s : -1;
}


__global__ void k_count_missing_diag(int n, const int* __restrict__ dind, int* __restrict__ cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && dind[i] < 0) atomicAdd(cnt, 1);
}

__global__ void k_make_wrow(int n, const int* __restrict__ row, int* __restrict__ Wrow) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) Wrow[i] = row[i] - i;
}

// Rounds of k_suitor until no vertex is pending.  The list length is read
// back every few rounds (an empty round costs one tiny launch).
template <int G>
static void suitor_rounds(Context* c, int n, const int* Wrow, const int2* pref, int* pos, unsigned long long* S) {
    // proposals per vertex per launch: short while most vertices are pending
    // (the launch is throughput-bound), longer once only chains remain (each
    // launch is then latency-bound and fewer launches finish them sooner)
    auto maxsteps_of = [](int round) { return round < 8 ? 8 : round < 16 ? 16 : round < 24 ? 32 : 64; };
    // round r counts its pending vertices in cnt[r % RING]: a ring of counters
    // cleared together every RING rounds, not one clear per round
    constexpr int RING = 64;
    DBuf<int> la(c, n), lb(c, n), cnt(c, RING);
    fill<int>(c, cnt.get(), RING, 0);
    unsigned grid = (unsigned)std::min<long long>(grid_for(n, WG / G), (long long)c->num_cus * 16);
    k_suitor<G><<<grid, WG, 0, c->stream>>>(n, Wrow, pref, pos, S, nullptr, nullptr, la.get(), cnt.get(),
                                             maxsteps_of(0));
    BCMGX_LAUNCHED();
    int* lists[2] = {la.get(), lb.get()};
    int cur = 0;
    for (int round = 1;; round++) {
        // lists[cur] holds the pending vertices, counted in cnt[(round - 1) % RING]
        if (round % RING == 0) fill<int>(c, cnt.get(), RING - 1, 0);  // every counter but this round's input
        const int* in_cnt = cnt.get() + (round - 1) % RING;
        int* out_cnt = cnt.get() + round % RING;
        k_suitor<G><<<grid, WG, 0, c->stream>>>(n, Wrow, pref, pos, S, lists[cur], in_cnt, lists[cur ^ 1],
                                                 out_cnt, maxsteps_of(round));
        BCMGX_LAUNCHED();
        cur ^= 1;
        if (round % 8 == 0 && c->read1(out_cnt) == 0) break;
        // every round settles at least one proposal; offers are bounded by the edge count
        BCMGX_CHECK(round < 4 * n + 1024, BCMGX_ERR_NUMERIC, "matching did not settle");
    }
}

static int lanes_for(long long nnz, int n) {
    double d = n > 0 ? (double)nnz / n : 1;
    if (d <= 4) return 4;
    if (d <= 8) return 8;
    if (d <= 16) return 16;
    return 32;
}

int count_missing_diagonals(Context* c, int n, const int* dind) {
    if (n <= 0) return 0;
    DBuf<int> miss(c, 1);
    fill<int>(c, miss.get(), 1, 0);
    k_count_missing_diag<<<grid_for(n, WG), WG, 0, c->stream>>>(n, dind, miss.get());
    BCMGX_LAUNCHED();
    return c->read1(miss.get());
}

void suitor_match(Context* c, const DCsr& A, const double* w, int* M, const int* dind_checked) {
    int n = A.n;
    if (n <= 0) return;
    long long m = (long long)A.nnz - n;  // off-diagonal entries (one diagonal per row, checked by the caller)
    size_t mm = (size_t)std::max<long long>(m, 1);
    DBuf<int> dind_own;
    const int* dind = dind_checked;
    if (!dind) {
        dind_own.reset(c, n);
        diag_index(c, n, A.row.get(), A.col.get(), dind_own.get());
        // makeAH requires a diagonal entry in every row (matching.cu:155)
        BCMGX_CHECK(count_missing_diagonals(c, n, dind_own.get()) == 0, BCMGX_ERR_ARG,
                    "matching: every row needs a diagonal entry");
        dind = dind_own.get();
    }
    DBuf<double> C(c, n);
    k_make_C<<<grid_for(n, WG), WG, 0, c->stream>>>(n, A.val.get(), dind, w, C.get());
    BCMGX_LAUNCHED();
    DBuf<int> Wrow(c, (size_t)n + 1), Wcol(c, mm);
    k_make_wrow<<<grid_for(n + 1, WG), WG, 0, c->stream>>>(n, A.row.get(), Wrow.get());
    BCMGX_LAUNCHED();
    int L = lanes_for(A.nnz, n);
    DBuf<int> mir(c, mm);
    fill<int>(c, mir.get(), (long long)mm, -1);
    DBuf<double> Wv(c, mm);  // the weights
    DBuf<int> rank(c, mm);
#define BCMGX_W(LL)                                                                                                   \
    case LL:                                                                                                          \
        k_weights<LL><<<grid_for(n, WG / LL), WG, 0, c->stream>>>(n, A.row.get(), A.col.get(), A.val.get(), dind, w,   \
                                                                  C.get(), Wcol.get(), Wv.get(), mir.get(), rank.get()); \
        break;
    switch (L) { BCMGX_W(4) BCMGX_W(8) BCMGX_W(16) BCMGX_W(32) }
#undef BCMGX_W
    BCMGX_LAUNCHED();
    DBuf<int2> pref(c, mm);
    DBuf<int> pos(c, n);
    DBuf<long long> S(c, n);
    fill<long long>(c, S.get(), n, 0);
    // the cursors start below every list: a read that misses a later write
    // (not yet visible across XCDs) still sees a valid lower bound
    k_pos_init<<<grid_for(n, WG), WG, 0, c->stream>>>(n, Wrow.get(), pos.get());
    BCMGX_LAUNCHED();
#define BCMGX_MT(LL)                                                                                                   \
    case LL:                                                                                                           \
        k_pref<LL><<<grid_for(n, WG / LL), WG, 0, c->stream>>>(n, Wrow.get(), Wcol.get(), rank.get(), mir.get(),       \
                                                               pref.get());                                            \
        BCMGX_LAUNCHED();                                                                                              \
        break;
    switch (L) { BCMGX_MT(4) BCMGX_MT(8) BCMGX_MT(16) BCMGX_MT(32) }
#undef BCMGX_MT
    auto* Su = reinterpret_cast<unsigned long long*>(S.get());
    // proposals are latency-bound: the first SUITOR_LANES candidates of the
    // preference list usually hold the taken one, and fewer lanes per vertex
    // keep more proposal chains in flight per wave
    int LS = std::min(L, SUITOR_LANES);
    switch (LS) {
    case 2: suitor_rounds<2>(c, n, Wrow.get(), pref.get(), pos.get(), Su); break;
    case 4: suitor_rounds<4>(c, n, Wrow.get(), pref.get(), pos.get(), Su); break;
    case 8: suitor_rounds<8>(c, n, Wrow.get(), pref.get(), pos.get(), Su); break;
    default: suitor_rounds<16>(c, n, Wrow.get(), pref.get(), pos.get(), Su); break;
    }
    k_suitor_decode<<<grid_for(n, WG), WG, 0, c->stream>>>(n, Su, M);
    BCMGX_LAUNCHED();
}

// ------------------------------------------------------------ prolongator ---
// makeP_GPU (matchingPairAggregation.cu:109-170).  Pair (v < u = M[v]) is good
// when ||(w_v, w_u)|| > eps (_aggregate_symmetric_unsorted, :47-79); good
// pairs take ids in master order (DeviceSelect::If, :127-148 +
// _aggregate_symmetric_sort, :81-93).  Singletons are numbered in ascending
// index (the reference uses atomicAdd order, :38): good -> next id with
// P = w/|w|, bad -> (running count) - 1 with P = 0 (:41-44).
__global__ void k_pair_flags(int n, const int* __restrict__ M, const double* __restrict__ w, int* __restrict__ good,
                             double* __restrict__ pval) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int u = M[v];
    int g = 0;
    if (u >= 0 && v < u) {
        double wv = w[v], wu = w[u];
        double nrm = sqrt(wv * wv + wu * wu);
        if (nrm > DBL_EPSILON) {
            g = 1;
            pval[v] = wv / nrm;
            pval[u] = wu / nrm;
        }
    }
    good[v] = g;
}
__global__ void k_assign_pairs(int n, const int* __restrict__ M, const int* __restrict__ good, const int* __restrict__ pid,
                               int* __restrict__ agg) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (good[v]) {
        agg[v] = pid[v];
        agg[M[v]] = pid[v];
    }
}
__global__ void k_single_flags(int n, const int* __restrict__ M, const int* __restrict__ good, const double* __restrict__ w,
                               int* __restrict__ single, int* __restrict__ inpair) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int u = M[v];
    int ip = good[v] || (u >= 0 && u < v && good[u]);
    inpair[v] = ip;
    single[v] = (!ip && fabs(w[v]) > DBL_EPSILON) ? 1 : 0;
}
__global__ void k_assign_singles(int n, const int* __restrict__ npairs_d, const int* __restrict__ inpair, const int* __restrict__ sid,
                                 const double* __restrict__ w, int* __restrict__ agg, double* __restrict__ pval) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (inpair[v]) return;
    int npairs = *npairs_d;
    double wv = w[v];
    if (fabs(wv) > DBL_EPSILON) {
        agg[v] = npairs + sid[v];
        pval[v] = wv / fabs(wv);
    } else {
        int a = npairs + sid[v] - 1;
        agg[v] = a > 0 ? a : 0;
        pval[v] = 0.0;
    }
}

int make_prolongator(Context* c, int n, const int* M, const double* w, int* agg, double* pval, int* npairs_out) {
    if (npairs_out) *npairs_out = 0;
    if (n <= 0) return 0;
    DBuf<int> good(c, n), pid(c, (size_t)n + 1), single(c, n), inpair(c, n), sid(c, (size_t)n + 1);
    k_pair_flags<<<grid_for(n, WG), WG, 0, c->stream>>>(n, M, w, good.get(), pval);
    BCMGX_LAUNCHED();
    scan_exclusive_async(c, good.get(), pid.get(), n);  // pid[n] = number of pairs, read by the kernels
    k_assign_pairs<<<grid_for(n, WG), WG, 0, c->stream>>>(n, M, good.get(), pid.get(), agg);
    BCMGX_LAUNCHED();
    k_single_flags<<<grid_for(n, WG), WG, 0, c->stream>>>(n, M, good.get(), w, single.get(), inpair.get());
    BCMGX_LAUNCHED();
    scan_exclusive_async(c, single.get(), sid.get(), n);
    k_assign_singles<<<grid_for(n, WG), WG, 0, c->stream>>>(n, pid.get() + n, inpair.get(), sid.get(), w, agg, pval);
    BCMGX_LAUNCHED();
    int npairs, nsingles;
    c->d2h2(&npairs, pid.get() + n, 1, &nsingles, sid.get() + n, 1);
    if (npairs_out) *npairs_out = npairs;
    return npairs + nsingles;
}

// The restriction P^T of a pairwise aggregation with no zero-weight singleton
// (n = 2 npairs + singletons): aggregate c < npairs holds its pair (v < u = M[v]),
// the others one vertex each, so row c of P^T is [v, u] or [i] -- the rows the
// general transpose produces (source rows ascending), written in one pass.
__global__ void k_pair_restriction(int n, int nc, int npairs, const int* __restrict__ agg, const int* __restrict__ M,
                                   const double* __restrict__ pval, int* __restrict__ rrow, int* __restrict__ rcol,
                                   double* __restrict__ rval) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= nc) rrow[i] = i < npairs ? 2 * i : npairs + i;
    if (i >= n) return;
    int a = agg[i];
    int pos = a < npairs ? 2 * a + (M[i] < i ? 1 : 0) : npairs + a;
    rcol[pos] = i;
    rval[pos] = pval[i];
}
void pair_restriction(Context* c, int n, int nc, int npairs, const int* agg, const int* M, const double* pval, DCsr& R) {
    R.alloc(c, nc, n, n);
    R.maxrow = npairs > 0 ? 2 : 1;
    k_pair_restriction<<<grid_for(std::max(n, nc + 1), WG), WG, 0, c->stream>>>(n, nc, npairs, agg, M, pval, R.row.get(),
                                                                             R.col.get(), R.val.get());
    BCMGX_LAUNCHED();
}

// P <- P P_i for one-entry-per-row operators (nsparseMGPU(P, P_i), matchingAggregation.cu:175-191)
__global__ void k_compose(int n, int* __restrict__ agg, double* __restrict__ pval, const int* __restrict__ agg_i,
                          const double* __restrict__ pval_i) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int a = agg[k];
    pval[k] = pval[k] * pval_i[a];
    agg[k] = agg_i[a];
}
void compose_prolongators(Context* c, int n, int* agg, double* pval, const int* agg_i, const double* pval_i) {
    if (n <= 0) return;
    k_compose<<<grid_for(n, WG), WG, 0, c->stream>>>(n, agg, pval, agg_i, pval_i);
    BCMGX_LAUNCHED();
}

__global__ void k_iota(int n, int* __restrict__ r) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) r[i] = i;
}
void prolongator_csr(Context* c, int n, int nc, const int* agg, const double* pval, DCsr& P) {
    P.alloc(c, n, nc, n);
    P.unit_rows = true;
    P.maxrow = 1;
    k_iota<<<grid_for(n + 1, WG), WG, 0, c->stream>>>(n, P.row.get());
    BCMGX_LAUNCHED();
    copy_d2d(c, P.col.get(), agg, sizeof(int) * n);
    copy_d2d(c, P.val.get(), pval, sizeof(double) * n);
}

// ------------------------------------------------------------- transpose ---
// CSRm::Transpose_local (CSR.cu:1752-1822): entries ordered by (col, row).
// Count per column, scan, scatter with atomic cursors, then sort each output
// row by source row so the result is deterministic.
__global__ void k_count_cols(long long nnz, const int* __restrict__ col, int* __restrict__ cnt) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nnz) atomicAdd(&cnt[col[t]], 1);
}
__global__ void k_scatter_T(int n, const int* __restrict__ row, const int* __restrict__ col, const double* __restrict__ val,
                            const int* __restrict__ trow, int* __restrict__ cursor, int* __restrict__ tcol,
                            double* __restrict__ tval) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = row[i]; j < row[i + 1]; j++) {
        int cc = col[j];
        int p = trow[cc] + atomicAdd(&cursor[cc], 1);
        tcol[p] = i;
        tval[p] = val[j];
    }
}
__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m, int w) {
    return (unsigned long long)__shfl_xor((long long)v, m, w);
}

// Each row sorted by column in place, equal columns keeping their order; one wave a row.
// Rows of at most 64 entries: a bitonic sort of the (column, position) keys across the
// lanes, each value then fetched from the lane it came from.  Longer rows: one lane's
// insertion sort.
__global__ __launch_bounds__(WG) void k_sort_rows(int n, const int* __restrict__ row, int* __restrict__ col,
                                                  double* __restrict__ val) {
    int lane = threadIdx.x & 63;
    long long i = ((long long)blockIdx.x * WG + threadIdx.x) >> 6;
    if (i >= n) return;  // whole wave
    int s = row[i], e = row[i + 1], len = e - s;
    if (len <= 1) return;
    if (len <= 64) {
        unsigned long long key = lane < len ? ((unsigned long long)(unsigned)col[s + lane] << 32) | (unsigned)lane : ~0ull;
        double v = lane < len ? val[s + lane] : 0.0;
#pragma unroll
        for (int kk = 2; kk <= 64; kk <<= 1) {
#pragma unroll
            for (int j = kk >> 1; j > 0; j >>= 1) {
                unsigned long long pk = shfl_xor_u64(key, j, 64);
                bool up = (lane & kk) == 0;
                bool lower = (lane & j) == 0;
                if ((lower == up) ? (pk < key) : (pk > key)) key = pk;
            }
        }
        double sv = __shfl(v, (int)(key & 63u), 64);
        if (lane < len) {
            col[s + lane] = (int)(key >> 32);
            val[s + lane] = sv;
        }
        return;
    }
    if (lane != 0) return;
    for (int a = s + 1; a < e; a++) {
        int kc = col[a];
        double kv = val[a];
        int b = a - 1;
        while (b >= s && col[b] > kc) {
            col[b + 1] = col[b];
            val[b + 1] = val[b];
            b--;
        }
        col[b + 1] = kc;
        val[b + 1] = kv;
    }
}
static void transpose_raw(Context* c, int n, int m, long long nnz, const int* row, const int* col, const double* val, DCsr& T) {
    T.alloc(c, m, n, (int)nnz);
    DBuf<int> cnt(c, (size_t)m + 1), cursor(c, (size_t)m + 1);
    fill<int>(c, cnt.get(), m + 1, 0);
    fill<int>(c, cursor.get(), m + 1, 0);
    if (nnz > 0) {
        k_count_cols<<<grid_for(nnz, WG), WG, 0, c->stream>>>(nnz, col, cnt.get());
        BCMGX_LAUNCHED();
    }
    scan_exclusive_async(c, cnt.get(), T.row.get(), m);
    if (n > 0) {
        k_scatter_T<<<grid_for(n, WG), WG, 0, c->stream>>>(n, row, col, val, T.row.get(), cursor.get(), T.col.get(), T.val.get());
        BCMGX_LAUNCHED();
    }
    if (m > 0) {
        k_sort_rows<<<grid_for(m, WG / 64), WG, 0, c->stream>>>(m, T.row.get(), T.col.get(), T.val.get());
        BCMGX_LAUNCHED();
    }
}
void transpose(Context* c, const DCsr& A, DCsr& T) { transpose_raw(c, A.n, A.m, A.nnz, A.row.get(), A.col.get(), A.val.get(), T); }
void transpose_prolongator(Context* c, int n, int nc, const int* agg, const double* pval, DCsr& R) {
    DBuf<int> row(c, (size_t)n + 1);
    k_iota<<<grid_for(n + 1, WG), WG, 0, c->stream>>>(n, row.get());
    BCMGX_LAUNCHED();
    transpose_raw(c, n, nc, n, row.get(), agg, pval, R);
}

// ---------------------------------------------------------------- SpGEMM ---
// C = A B row by row (replaces nsparse's two-phase hash SpGEMM,
// kernel_spgemm_hash_d.cu:1196-1464).  Each row's products are generated in
// order (A entry, then B entry), tagged key = (column << 32) | sequence, sorted
// (bitonic) in LDS, and every run of equal columns is summed left to right:
// C_ij = sum over k in ascending position, the order of the oracle's
// Gustavson loop.  Rows are binned by product count:
//   <= 64 products     : one wave64 per row, sort in registers via shuffles;
//   <= 2048 products   : one workgroup per row, sort in LDS;
//   larger             : one workgroup per row, sort in global scratch.
constexpr int SMALL_P = 64;
constexpr int MED_P = 2048;  // 2048 x (8 + 8 + 4) B + scan = 41 KiB of LDS

// the largest of n ints into *mx (which holds 0 or a smaller value): a grid-stride
// pass with one atomic per block, since device-scope atomics on one address serialise
constexpr int MAX_INT_BLOCKS = 512;
__global__ __launch_bounds__(WG) void k_max_int(int n, const int* __restrict__ v, int* __restrict__ mx) {
    __shared__ int part[WG / 64];
    int x = 0;
    for (int i = blockIdx.x * WG + (int)threadIdx.x; i < n; i += gridDim.x * WG) x = v[i] > x ? v[i] : x;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int y = __shfl_xor(x, o, 64);
        x = y > x ? y : x;
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < WG / 64; w++) x = part[w] > x ? part[w] : x;
        if (x > 0) atomicMax(mx, x);
    }
}

// the longest row into out[0] and the number of rows i whose pointer is not i into
// out[1] (both hold 0 on entry): a grid-stride pass, one atomic of each per block
__global__ __launch_bounds__(WG) void k_row_shape(int n, const int* __restrict__ row, int* __restrict__ out) {
    __shared__ int part[2][WG / 64];
    int mx = 0, bad = 0;
    for (int i = blockIdx.x * WG + (int)threadIdx.x; i < n; i += gridDim.x * WG) {
        int a = row[i], b = row[i + 1];
        mx = b - a > mx ? b - a : mx;
        bad += a != i ? 1 : 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int y = __shfl_xor(mx, o, 64);
        mx = y > mx ? y : mx;
        bad += __shfl_xor(bad, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = mx;
        part[1][threadIdx.x >> 6] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < WG / 64; w++) {
            mx = part[0][w] > mx ? part[0][w] : mx;
            bad += part[1][w];
        }
        if (mx > 0) atomicMax(out, mx);
        if (bad > 0) atomicAdd(out + 1, bad);
    }
}

void csr_shape(Context* c, DCsr& A) {
    A.maxrow = 0;
    A.unit_rows = false;
    if (A.n <= 0) return;
    DBuf<int> out(c, 2);
    fill<int>(c, out.get(), 2, 0);
    k_row_shape<<<std::min<unsigned>(grid_for(A.n, WG), MAX_INT_BLOCKS), WG, 0, c->stream>>>(A.n, A.row.get(), out.get());
    BCMGX_LAUNCHED();
    int h[2];
    c->d2h(h, out.get(), 2);
    A.maxrow = h[0];
    A.unit_rows = A.nnz == A.n && h[1] == 0;
}

__global__ void k_row_lengths(int n, const int* __restrict__ arow, long long* __restrict__ cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = arow[i + 1] - arow[i];
}

// B's row k is [brow[k], bend[k]) in every product kernel: bend = brow + 1 for a
// CSR matrix, or the ends of rows kept in place with gaps after them
__global__ __launch_bounds__(RS_ROWS) void k_count_products(int n, const int* __restrict__ arow, const int* __restrict__ acol,
                                                            const int* __restrict__ brow, const int* __restrict__ bend,
                                                            long long* __restrict__ cnt) {
    __shared__ int sc[RS_CAP];
    int r0 = blockIdx.x * RS_ROWS;
    int rend = r0 + RS_ROWS < n ? r0 + RS_ROWS : n;
    int base = arow[r0], ne = arow[rend] - base;
    bool staged = ne <= RS_CAP;
    if (staged) stage_rows(acol + base, sc, ne);
    __syncthreads();
    int i = r0 + threadIdx.x;
    if (i >= n) return;
    long long s = 0;
    int j = arow[i], je = arow[i + 1];
    if (staged) {
        // eight row-length gathers in flight (integer sums: any order is exact)
        for (; j + 8 <= je; j += 8) {
            int d[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                int k = sc[j + u - base];
                d[u] = bend[k] - brow[k];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) s += d[u];
        }
    }
    for (; j < je; j++) {
        int k = staged ? sc[j - base] : acol[j];
        s += bend[k] - brow[k];
    }
    cnt[i] = s;
}

// bins: 0..2 rows of <= 16 / 32 / 64 products (W-lane sorts), 3 medium, 4 large
constexpr int NBINS = 5;
// bin rows by product count (order inside a bin is irrelevant): the waves of a
// block take slots from LDS counters over BIN_CHUNKS consecutive chunks of WG
// rows, then one global atomic per non-empty bin per block reserves the
// block's range (all blocks add to the same few counters, so fewer, larger
// blocks contend less)
constexpr int BIN_CHUNKS = 8;
__global__ __launch_bounds__(WG) void k_bin_rows(int n, const int* __restrict__ arow, const long long* __restrict__ cnt,
                                                 int* __restrict__ bins, int* __restrict__ bincount, int cap) {
    __shared__ int lcount[NBINS];
    __shared__ int lbase[NBINS];
    int lane = threadIdx.x & 63;
    if (threadIdx.x < NBINS) lcount[threadIdx.x] = 0;
    long long i0 = (long long)blockIdx.x * BIN_CHUNKS * WG + threadIdx.x;
    int bq[BIN_CHUNKS];
#pragma unroll
    for (int q = 0; q < BIN_CHUNKS; q++) {  // every chunk's loads first
        long long i = i0 + (long long)q * WG;
        int b = -1;
        if (i < n) {
            long long p = cnt[i];
            int alen = arow[i + 1] - arow[i];
            long long m = p > alen ? p : alen;
            b = m <= 16 ? 0 : m <= 32 ? 1 : m <= SMALL_P ? 2 : (m <= MED_P ? 3 : 4);
        }
        bq[q] = b;
    }
    __syncthreads();
    unsigned long long below = (1ull << lane) - 1ull;
    int slot[BIN_CHUNKS];
#pragma unroll
    for (int q = 0; q < BIN_CHUNKS; q++) {
        int myslot = 0;
        for (int bb = 0; bb < NBINS; bb++) {
            unsigned long long msk = __ballot(bq[q] == bb);
            if (!msk) continue;
            int leader = __ffsll((long long)msk) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(&lcount[bb], __popcll(msk));
            base = __shfl(base, leader, 64);
            if (bq[q] == bb) myslot = base + __popcll(msk & below);
        }
        slot[q] = myslot;
    }
    __syncthreads();
    if (threadIdx.x < NBINS) lbase[threadIdx.x] = lcount[threadIdx.x] ? atomicAdd(&bincount[threadIdx.x], lcount[threadIdx.x]) : 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BIN_CHUNKS; q++)
        if (bq[q] >= 0) bins[(long long)bq[q] * cap + lbase[bq[q]] + slot[q]] = (int)(i0 + (long long)q * WG);
}

// W lanes per row (W = 16, 32 or 64), products <= W: 64 / W rows per wave.
// K32: B has fewer than 2^26 - 1 columns, so (column, position) packs into one
// 32-bit key; the sort then moves one 32-bit word per lane and stage, and each
// lane fetches its product's value by position afterwards (same order, same
// sums as the 64-bit key with the value carried along).
// UNITB: B has exactly one entry per row (a prolongator: row k of B is entry k),
// so each entry of A makes one product, found without B's row pointers.
template <int W, bool K32, bool UNITB = false>
__global__ __launch_bounds__(WG) void k_spgemm_small(int nrows, const int* __restrict__ rows, const int* __restrict__ arow,
                                                     const int* __restrict__ acol, const double* __restrict__ aval,
                                                     const int* __restrict__ brow, const int* __restrict__ bend,
                                                     const int* __restrict__ bcol, const double* __restrict__ bval,
                                                     const long long* __restrict__ poff, int* __restrict__ tcol,
                                                     double* __restrict__ tval, int* __restrict__ uniq,
                                                     const int* __restrict__ poff32, int* __restrict__ rend) {
    // rows == nullptr: the w-th group takes row w; poff32 (if given) holds the product offsets
    // instead of poff; uniq (if given) receives each row's result length, rend (if given) its end
    // in tcol / tval (with poff32)
    constexpr int G = WG / W;  // rows per block
    __shared__ int pre_s[G][W + 1];
    __shared__ double v_s[G][W];
    __shared__ double p_s[K32 ? G : 1][K32 ? W : 1];  // K32: products by position
    int gid = threadIdx.x / W, lane = threadIdx.x % W;
    int wl = threadIdx.x & 63;             // lane in the wave
    int gshift = wl - lane;                // first wave lane of this row's group
    int w = blockIdx.x * G + gid;
    bool active = w < nrows;  // inactive groups still take part in the barriers
    int i = active ? (rows ? rows[w] : w) : 0;
    // one row per wave (W = 64): the row, and so its bounds, are the wave's own (scalar loads)
    if constexpr (W == 64) i = __builtin_amdgcn_readfirstlane(i);
    int s = active ? arow[i] : 0, alen = active ? arow[i + 1] - s : 0;
    // lane a holds A's entry a (at most W of them: rows are binned by max(products, entries)):
    // its column, and for a general B its value and B row start, which the product lanes then
    // take by shuffle instead of loading them again (with unit rows each lane's product is its
    // own entry, read directly)
    int k = -1, cntl = 0, kb = 0;
    double ka_v = 0.0;
    if (lane < alen) {
        k = acol[s + lane];
        if constexpr (!UNITB) {
            ka_v = aval[s + lane];
            kb = brow[k];
        }
        cntl = UNITB ? 1 : bend[k] - kb;
    }
    // inclusive scan of B-row lengths across the group (with unit rows: lane + 1, up to alen)
    int x = cntl;
    if constexpr (UNITB) {
        x = lane < alen ? lane + 1 : alen;
    } else {
#pragma unroll
        for (int d = 1; d < W; d <<= 1) {
            int y = __shfl_up(x, d, W);
            if (lane >= d) x += y;
        }
    }
    int total = UNITB ? alen : __shfl(x, W - 1, W);
    pre_s[gid][lane + 1] = x;
    if (lane == 0) pre_s[gid][0] = 0;
    __syncthreads();
    // A row of A with at most two entries (the rows of R in P^T (A P)) has at
    // most two runs of products, each in column order when B's rows are.  Run 2
    // then goes reversed into the top lanes (run 1 ascending, the padding,
    // run 2 descending: a bitonic sequence), which the sort's last merge alone
    // orders.  Keys carry their product position, so any layout sorts the same.
    bool two = K32 && alen <= 2;  // uniform over the group
    int n1 = two && alen == 2 ? pre_s[gid][1] : total, n2 = total - n1;
    int pos = lane;  // the product this lane holds (W: padding)
    if (two && lane >= n1) pos = lane >= W - n2 ? n1 + (W - 1 - lane) : W;
    unsigned long long key = ~0ull;
    unsigned key32 = ~0u;
    double v = 0.0;
    // owner: largest a with pre[a] <= pos (with unit rows: pos itself); the shuffles run on
    // every lane, outside the branch, so each source lane is active
    int a = 0;
    if (pos < total) {
        int lo = 0, hi = alen;  // pre[0] = 0 <= pos
        if constexpr (UNITB) {
            lo = pos;
        } else {
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (pre_s[gid][mid] <= pos) lo = mid;
                else hi = mid;
            }
        }
        a = lo;
    }
    int ka = 0, kba = 0;
    double ava = 0.0;
    if constexpr (!UNITB) {
        ka = __shfl(k, a, W);
        kba = __shfl(kb, a, W);
        ava = __shfl(ka_v, a, W);
    }
    if (pos < total) {
        if constexpr (UNITB) {
            ka = acol[s + a];
            ava = aval[s + a];
        }
        int jb = UNITB ? ka : kba + (pos - pre_s[gid][a]);
        int cc = bcol[jb];
        v = ava * bval[jb];
        key = ((unsigned long long)(unsigned)cc << 32) | (unsigned)pos;
        key32 = ((unsigned)cc << 6) | (unsigned)pos;
    }
    if constexpr (K32) {
        if (pos < total) p_s[gid][pos] = v;
        bool merge_only = false, sorted = false;
        unsigned prv = (unsigned)__shfl_up((int)key32, 1, W);
        if (two) {  // both runs in order (run 2 descending over the lanes)?
            bool bad = (lane >= 1 && lane < n1 && !(prv < key32)) || (lane > W - n2 && !(prv > key32));
            unsigned long long bm = (__ballot(bad) >> gshift) & (W == 64 ? ~0ull : ((1ull << W) - 1ull));
            merge_only = bm == 0;
        } else {
            // the products already in key order over the lanes (the padding's all-ones keys last):
            // the case of a prolongator that numbers its aggregates in the order of their rows
            bool bad = lane >= 1 && lane < total && !(prv < key32);
            unsigned long long bm = (__ballot(bad) >> gshift) & (W == 64 ? ~0ull : ((1ull << W) - 1ull));
            sorted = bm == 0;
        }
        if (sorted) {
        } else if (merge_only) {
            // the last merge of the bitonic sort (all lanes ascending)
#pragma unroll
            for (int j = W >> 1; j > 0; j >>= 1) {
                unsigned pk = (unsigned)__shfl_xor((int)key32, j, W);
                bool lower = (lane & j) == 0;
                bool take = lower ? (pk < key32) : (pk > key32);
                if (take) key32 = pk;
            }
        } else {
            // bitonic sort of W 32-bit keys across the group
#pragma unroll
            for (int kk = 2; kk <= W; kk <<= 1) {
#pragma unroll
                for (int j = kk >> 1; j > 0; j >>= 1) {
                    unsigned pk = (unsigned)__shfl_xor((int)key32, j, W);
                    bool up = (lane & kk) == 0;
                    bool lower = (lane & j) == 0;
                    bool take = (lower == up) ? (pk < key32) : (pk > key32);
                    if (take) key32 = pk;
                }
            }
        }
        __syncthreads();  // p_s complete
        if (lane < total) v = p_s[gid][key32 & 63u];
    } else {
        // bitonic sort of W (key, v) pairs across the group
#pragma unroll
        for (int kk = 2; kk <= W; kk <<= 1) {
#pragma unroll
            for (int j = kk >> 1; j > 0; j >>= 1) {
                unsigned long long pk = shfl_xor_u64(key, j, W);
                double pv = __shfl_xor(v, j, W);
                bool up = (lane & kk) == 0;
                bool lower = (lane & j) == 0;
                bool take = (lower == up) ? (pk < key) : (pk > key);
                if (take) {
                    key = pk;
                    v = pv;
                }
            }
        }
    }
    v_s[gid][lane] = v;
    int colv = K32 ? (int)(key32 >> 6) : (int)(key >> 32);
    int prevc = __shfl_up(colv, 1, W);
    bool head = lane < total && (lane == 0 || prevc != colv);
    unsigned long long hm = __ballot(head);
    hm = (hm >> gshift) & (W == 64 ? ~0ull : ((1ull << W) - 1ull));  // this group's lanes
    __syncthreads();
    if (head && active) {
        double acc = v;
        // sum the run left to right
        for (int t = lane + 1; t < total; t++) {
            if (!((hm >> t) & 1ull)) acc += v_s[gid][t];
            else break;
        }
        int u = __popcll(hm & ((1ull << lane) - 1ull));
        // an in-place product writes at A's own row start, which this group holds
        long long o = (poff32 ? (poff32 == arow ? (long long)s : (long long)poff32[i]) : poff[i]) + u;
        tcol[o] = colv;
        tval[o] = acc;
    }
    if (lane == 0 && active) {
        int u = __popcll(hm);
        if (uniq) uniq[i] = u;
        if (rend) rend[i] = (poff32 == arow ? s : poff32[i]) + u;
    }
}

// one workgroup per row; keys/vals/pre point at LDS (medium) or global scratch (large)
__device__ void spgemm_row_wg(int i, const int* __restrict__ arow, const int* __restrict__ acol, const double* __restrict__ aval,
                              const int* __restrict__ brow, const int* __restrict__ bend, const int* __restrict__ bcol,
                              const double* __restrict__ bval,
                              long long o0, int* __restrict__ tcol, double* __restrict__ tval, int* __restrict__ uniq,
                              unsigned long long* keys, double* vals, int* pre, int* lscan) {
    int tid = threadIdx.x;
    int s = arow[i], alen = arow[i + 1] - s;
    // prefix of B-row lengths over the A row (chunks of WG with carry)
    int carry = 0;
    for (int base = 0; base < alen; base += WG) {
        int a = base + tid;
        int cl = 0;
        if (a < alen) {
            int k = acol[s + a];
            cl = bend[k] - brow[k];
        }
        lscan[tid] = cl;
        __syncthreads();
        for (int d = 1; d < WG; d <<= 1) {
            int y = tid >= d ? lscan[tid - d] : 0;
            __syncthreads();
            lscan[tid] += y;
            __syncthreads();
        }
        if (a < alen) pre[a + 1] = carry + lscan[tid];
        int tot = lscan[WG - 1];
        __syncthreads();
        carry += tot;
    }
    if (tid == 0) pre[0] = 0;
    __syncthreads();
    int total = carry;
    int P = 1;
    while (P < total) P <<= 1;
    for (int p = tid; p < P; p += WG) {
        unsigned long long key = ~0ull;
        double v = 0.0;
        if (p < total) {
            int lo = 0, hi = alen;
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (pre[mid] <= p) lo = mid;
                else hi = mid;
            }
            int ka = acol[s + lo];
            int jb = brow[ka] + (p - pre[lo]);
            v = aval[s + lo] * bval[jb];
            key = ((unsigned long long)(unsigned)bcol[jb] << 32) | (unsigned)p;
        }
        keys[p] = key;
        vals[p] = v;
    }
    __syncthreads();
    for (int kk = 2; kk <= P; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int p = tid; p < P; p += WG) {
                int q = p ^ j;
                if (q > p) {
                    bool up = (p & kk) == 0;
                    unsigned long long a = keys[p], b = keys[q];
                    if ((a > b) == up) {
                        keys[p] = b;
                        keys[q] = a;
                        double t = vals[p];
                        vals[p] = vals[q];
                        vals[q] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
    // heads, their order, sums
    int ucarry = 0;
    for (int base = 0; base < total; base += WG) {
        int p = base + tid;
        int hd = 0;
        if (p < total) hd = (p == 0 || (keys[p] >> 32) != (keys[p - 1] >> 32)) ? 1 : 0;
        lscan[tid] = hd;
        __syncthreads();
        for (int d = 1; d < WG; d <<= 1) {
            int y = tid >= d ? lscan[tid - d] : 0;
            __syncthreads();
            lscan[tid] += y;
            __syncthreads();
        }
        if (hd) {
            unsigned cc = (unsigned)(keys[p] >> 32);
            double acc = vals[p];
            for (int t = p + 1; t < total && (unsigned)(keys[t] >> 32) == cc; t++) acc += vals[t];
            long long o = o0 + ucarry + lscan[tid] - 1;
            tcol[o] = (int)cc;
            tval[o] = acc;
        }
        int tot = lscan[WG - 1];
        __syncthreads();
        ucarry += tot;
    }
    if (tid == 0) uniq[i] = ucarry;
}

__global__ __launch_bounds__(WG) void k_spgemm_med(int nrows, const int* __restrict__ rows, const int* __restrict__ arow,
                                                   const int* __restrict__ acol, const double* __restrict__ aval,
                                                   const int* __restrict__ brow, const int* __restrict__ bend,
                                                   const int* __restrict__ bcol, const double* __restrict__ bval,
                                                   const long long* __restrict__ poff, int* __restrict__ tcol,
                                                   double* __restrict__ tval, int* __restrict__ uniq) {
    __shared__ unsigned long long keys[MED_P];
    __shared__ double vals[MED_P];
    __shared__ int pre[MED_P + 1];
    __shared__ int lscan[WG];
    int i = rows[blockIdx.x];
    spgemm_row_wg(i, arow, acol, aval, brow, bend, bcol, bval, poff[i], tcol, tval, uniq, keys, vals, pre, lscan);
}

__global__ __launch_bounds__(WG) void k_spgemm_large(int nrows, const int* __restrict__ rows, const int* __restrict__ arow,
                                                     const int* __restrict__ acol, const double* __restrict__ aval,
                                                     const int* __restrict__ brow, const int* __restrict__ bend,
                                                     const int* __restrict__ bcol, const double* __restrict__ bval,
                                                     const long long* __restrict__ poff, int* __restrict__ tcol,
                                                     double* __restrict__ tval, int* __restrict__ uniq, unsigned long long* gkeys, double* gvals, int* gpre,
                                                     const long long* __restrict__ soff, const long long* __restrict__ aoff) {
    __shared__ int lscan[WG];
    int r = blockIdx.x;
    int i = rows[r];
    spgemm_row_wg(i, arow, acol, aval, brow, bend, bcol, bval, poff[i], tcol, tval, uniq, gkeys + soff[r], gvals + soff[r],
                  gpre + aoff[r], lscan);
}

// dind (if given, -1 on entry): the position of each row's diagonal entry
template <typename O, typename V>
__global__ void k_spgemm_compact(int n, const O* __restrict__ poff, const int* __restrict__ crow,
                                 const int* __restrict__ tcol, const double* __restrict__ tval, int* __restrict__ ccol,
                                 V* __restrict__ cval, int* __restrict__ dind) {
    int i = (blockIdx.x * WG + (int)threadIdx.x) / 16;
    int lane = (int)threadIdx.x % 16;
    if (i >= n) return;
    long long src = poff[i];
    int dst = crow[i], len = crow[i + 1] - crow[i];
    for (int t = lane; t < len; t += 16) {
        int cc = tcol[src + t];
        ccol[dst + t] = cc;
        cval[dst + t] = (V)tval[src + t];
        if (dind && cc == i) dind[i] = dst + t;
    }
}

// sizes of the large-bin scratch for each large row
__global__ void k_large_sizes(int nl, const int* __restrict__ rows, const long long* __restrict__ cnt,
                              const int* __restrict__ arow, long long* __restrict__ psz, long long* __restrict__ asz) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nl) return;
    int i = rows[r];
    long long p = cnt[i], P = 1;
    while (P < p) P <<= 1;
    psz[r] = P;
    asz[r] = (long long)(arow[i + 1] - arow[i]) + 1;
}

// C's row pointers from the per-row result sizes, its nnz and largest row read back in
// one round trip, then the rows moved from their product positions (poff) into place
template <typename O>
static void spgemm_finish(Context* c, int n, const int* uniq, const O* poff, const int* tcol, const double* tval, DCsr& C,
                          int* dind = nullptr) {
    C.row.reset(c, (size_t)n + 1);
    DBuf<int> mx(c, 1);
    fill<int>(c, mx.get(), 1, 0);
    k_max_int<<<std::min<unsigned>(grid_for(n, WG), MAX_INT_BLOCKS), WG, 0, c->stream>>>(n, uniq, mx.get());
    BCMGX_LAUNCHED();
    scan_exclusive_async(c, uniq, C.row.get(), n);
    int nnz = 0, maxrow = 0;
    c->d2h2(&nnz, C.row.get() + n, 1, &maxrow, mx.get(), 1);
    BCMGX_CHECK(nnz >= 0, BCMGX_ERR_UNSUPPORTED, "spgemm: result has more than 2^31 entries");
    C.nnz = nnz;
    C.maxrow = maxrow;
    C.col.reset(c, (size_t)std::max(nnz, 1));
    C.val.reset(c, (size_t)std::max(nnz, 1));
    if (dind) fill<int>(c, dind, n, -1);
    k_spgemm_compact<O, double><<<grid_for(n, WG / 16), WG, 0, c->stream>>>(n, poff, C.row.get(), tcol, tval, C.col.get(),
                                                                           C.val.get(), dind);
    BCMGX_LAUNCHED();
}

// The right operand of a product: row k is [row[k], end[k]) of col / val
struct BRows {
    int n, m;
    const int* row;
    const int* end;
    const int* col;
    const double* val;
    bool unit;  // row k holds exactly entry k
};

// A B by the sort-based kernels (count, bin by product count, numeric, compact)
static void spgemm_rows(Context* c, const DCsr& A, const BRows& B, DCsr& C) {
    int n = A.n;
    bool unitb = B.unit;
    bool k32 = B.m < (1 << 26);
    DBuf<long long> cnt(c, n), poff(c, (size_t)n + 1);
    if (unitb) k_row_lengths<<<grid_for(n, WG), WG, 0, c->stream>>>(n, A.row.get(), cnt.get());
    else k_count_products<<<grid_for(n, RS_ROWS), RS_ROWS, 0, c->stream>>>(n, A.row.get(), A.col.get(), B.row, B.end, cnt.get());
    BCMGX_LAUNCHED();
    scan_exclusive_async(c, cnt.get(), poff.get(), n);
    DBuf<int> bins(c, (size_t)NBINS * n), bincount(c, NBINS);
    fill<int>(c, bincount.get(), NBINS, 0);
    k_bin_rows<<<grid_for(n, WG * BIN_CHUNKS), WG, 0, c->stream>>>(n, A.row.get(), cnt.get(), bins.get(), bincount.get(), n);
    BCMGX_LAUNCHED();
    long long total;
    int bc[NBINS];
    c->d2h2(&total, poff.get() + n, 1, bc, bincount.get(), NBINS);  // product count and bin sizes, one round trip
    size_t tsz = (size_t)std::max<long long>(total, 1);
    DBuf<int> tcol(c, tsz);
    DBuf<double> tval(c, tsz);
    DBuf<int> uniq(c, (size_t)n + 1);
#define BCMGX_SMALL(BIN, W)                                                                                                 \
    if (bc[BIN] > 0) {                                                                                                      \
        auto kern = k32 ? (unitb ? k_spgemm_small<W, true, true> : k_spgemm_small<W, true>) : k_spgemm_small<W, false>;   \
        kern<<<grid_for(bc[BIN], WG / W), WG, 0, c->stream>>>(bc[BIN], bins.get() + (size_t)(BIN) * n, A.row.get(),       \
                                                              A.col.get(), A.val.get(), B.row, B.end, B.col, B.val,        \
                                                              poff.get(), tcol.get(), tval.get(), uniq.get(), nullptr,     \
                                                              nullptr);                                                    \
        BCMGX_LAUNCHED();                                                                                                   \
    }
    BCMGX_SMALL(0, 16)
    BCMGX_SMALL(1, 32)
    BCMGX_SMALL(2, 64)
#undef BCMGX_SMALL
    if (bc[3] > 0) {
        k_spgemm_med<<<bc[3], WG, 0, c->stream>>>(bc[3], bins.get() + 3 * (size_t)n, A.row.get(), A.col.get(), A.val.get(),
                                                  B.row, B.end, B.col, B.val, poff.get(), tcol.get(), tval.get(), uniq.get());
        BCMGX_LAUNCHED();
    }
    if (bc[4] > 0) {
        int nl = bc[4];
        DBuf<long long> psz(c, nl), asz(c, nl), soff(c, (size_t)nl + 1), aoff(c, (size_t)nl + 1);
        k_large_sizes<<<grid_for(nl, WG), WG, 0, c->stream>>>(nl, bins.get() + 4 * (size_t)n, cnt.get(), A.row.get(), psz.get(), asz.get());
        BCMGX_LAUNCHED();
        long long ptot = scan_exclusive(c, psz.get(), soff.get(), nl);
        long long atot = scan_exclusive(c, asz.get(), aoff.get(), nl);
        DBuf<unsigned long long> gk(c, (size_t)ptot);
        DBuf<double> gv(c, (size_t)ptot);
        DBuf<int> gp(c, (size_t)atot);
        k_spgemm_large<<<nl, WG, 0, c->stream>>>(nl, bins.get() + 4 * (size_t)n, A.row.get(), A.col.get(), A.val.get(),
                                                 B.row, B.end, B.col, B.val, poff.get(), tcol.get(), tval.get(), uniq.get(),
                                                 gk.get(), gv.get(), gp.get(), soff.get(), aoff.get());
        BCMGX_LAUNCHED();
        // the scratch goes back to the pool here; the pool hands blocks out again only to
        // work queued behind this kernel on the same stream
    }
    spgemm_finish<long long>(c, n, uniq.get(), poff.get(), tcol.get(), tval.get(), C);
}

// A P with P a prolongator (one entry per row) and every row of A within one
// 32-lane group: the products of row i sit at A's own entry positions, one
// launch over all rows with no counting, binning or read-back.  Row i of the
// result is left at [A.row[i], rend[i]) of tcol / tval.
// P given by its aggregates and values (row k of P is the entry (k, agg[k]) = pval[k])
static bool unit_product_fits(const DCsr& A, int nc) { return nc < (1 << 26) && A.maxrow >= 0 && A.maxrow <= 32; }
static void unit_product_in_place(Context* c, const DCsr& A, const int* agg, const double* pval, int* tcol, double* tval,
                                  int* uniq, int* rend) {
    int n = A.n;  // (one product per entry: B's row pointers are not read)
    if (A.maxrow <= 16)
        k_spgemm_small<16, true, true><<<grid_for(n, WG / 16), WG, 0, c->stream>>>(
            n, nullptr, A.row.get(), A.col.get(), A.val.get(), nullptr, nullptr, agg, pval, nullptr, tcol, tval, uniq,
            A.row.get(), rend);
    else
        k_spgemm_small<32, true, true><<<grid_for(n, WG / 32), WG, 0, c->stream>>>(
            n, nullptr, A.row.get(), A.col.get(), A.val.get(), nullptr, nullptr, agg, pval, nullptr, tcol, tval, uniq,
            A.row.get(), rend);
    BCMGX_LAUNCHED();
}

void spgemm(Context* c, const DCsr& A, const DCsr& B, DCsr& C) {
    BCMGX_CHECK(A.m == B.n, BCMGX_ERR_ARG, "spgemm: inner dimensions differ");
    int n = A.n;
    C.n = n;
    C.m = B.m;
    C.unit_rows = false;
    if (n == 0) {
        C.alloc(c, 0, B.m, 0);
        fill<int>(c, C.row.get(), 1, 0);
        return;
    }
    if (B.unit_rows && B.nnz == B.n && unit_product_fits(A, B.m)) {
        size_t tsz = (size_t)std::max(A.nnz, 1);
        DBuf<int> tcol(c, tsz), uniq(c, (size_t)n + 1);
        DBuf<double> tval(c, tsz);
        unit_product_in_place(c, A, B.col.get(), B.val.get(), tcol.get(), tval.get(), uniq.get(), nullptr);
        spgemm_finish<int>(c, n, uniq.get(), A.row.get(), tcol.get(), tval.get(), C);
        return;
    }
    bool unitb = B.unit_rows && B.nnz == B.n;  // a prolongator: one product per entry of A
    spgemm_rows(c, A, BRows{B.n, B.m, B.row.get(), B.row.get() + 1, B.col.get(), B.val.get(), unitb}, C);
}

// A P for the prolongator (agg, pval) of A's n rows onto nc, when every row of A
// fits the one-launch product: row pointers, columns and values rounded to FP32
// (C.valf; no FP64 values) queued on c's stream with no host synchronisation.
// The result's length stays on the device at C.row[n] (C.nnz = -1 until read);
// the columns and values are sized for A's entries.  False (nothing queued)
// when A does not fit.
bool unit_product_f32_async(Context* c, const DCsr& A, int nc, const int* agg, const double* pval, DCsr& C) {
    int n = A.n;
    if (n <= 0 || !unit_product_fits(A, nc)) return false;
    size_t tsz = (size_t)std::max(A.nnz, 1);
    DBuf<int> tcol(c, tsz), uniq(c, (size_t)n + 1);
    DBuf<double> tval(c, tsz);
    unit_product_in_place(c, A, agg, pval, tcol.get(), tval.get(), uniq.get(), nullptr);
    C.n = n;
    C.m = nc;
    C.nnz = -1;
    C.unit_rows = false;
    C.maxrow = -1;
    C.row.reset(c, (size_t)n + 1);
    scan_exclusive_async(c, uniq.get(), C.row.get(), n);
    C.col.reset(c, tsz);
    C.val.release();
    C.valf.reset(c, tsz);
    k_spgemm_compact<int, float><<<grid_for(n, WG / 16), WG, 0, c->stream>>>(n, A.row.get(), C.row.get(), tcol.get(),
                                                                           tval.get(), C.col.get(), C.valf.get(), nullptr);
    BCMGX_LAUNCHED();
    return true;
}

__global__ void k_stride_offsets(int n, int stride, int* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i * stride;
}

// The Galerkin product R (A P).  When A P fits the one-launch product its rows
// stay where that launch leaves them (in A's entry positions, gaps after each)
// and R reads them there: A P is never compacted or measured on the host.  Its
// rows then hold at most A.maxrow entries, so when R's rows hold at most rmax
// (the caller's bound: 2 for a pairwise aggregation's restriction) and
// rmax A.maxrow <= 64, every row of the product fits one 64-lane group: one
// launch over all rows, 64 product slots per row, no counting or binning.
bool galerkin(Context* c, const DCsr& A, int nc, const int* agg, const double* pval, const DCsr& R, DCsr& C, int rmax,
              int* dind) {
    BCMGX_CHECK(R.m == A.n && R.n == nc, BCMGX_ERR_ARG, "galerkin: dimensions differ");
    if (A.n > 0 && R.n > 0 && unit_product_fits(A, nc)) {
        size_t tsz = (size_t)std::max(A.nnz, 1);
        DBuf<int> tcol(c, tsz), rend(c, (size_t)A.n);
        DBuf<double> tval(c, tsz);
        unit_product_in_place(c, A, agg, pval, tcol.get(), tval.get(), nullptr, rend.get());
        C.n = R.n;
        C.m = nc;
        C.unit_rows = false;
        if (rmax >= 0 && rmax * A.maxrow <= SMALL_P && (long long)R.n * SMALL_P < (1ll << 31)) {
            size_t slots = (size_t)R.n * SMALL_P;
            DBuf<int> off(c, (size_t)R.n), ccol(c, slots), uniq(c, (size_t)R.n + 1);
            DBuf<double> cval(c, slots);
            k_stride_offsets<<<grid_for(R.n, WG), WG, 0, c->stream>>>(R.n, SMALL_P, off.get());
            BCMGX_LAUNCHED();
            k_spgemm_small<SMALL_P, true, false><<<grid_for(R.n, WG / SMALL_P), WG, 0, c->stream>>>(
                R.n, nullptr, R.row.get(), R.col.get(), R.val.get(), A.row.get(), rend.get(), tcol.get(), tval.get(), nullptr,
                ccol.get(), cval.get(), uniq.get(), off.get(), nullptr);
            BCMGX_LAUNCHED();
            spgemm_finish<int>(c, R.n, uniq.get(), off.get(), ccol.get(), cval.get(), C, dind);
            return dind != nullptr;
        }
        spgemm_rows(c, R, BRows{A.n, nc, A.row.get(), rend.get(), tcol.get(), tval.get(), false}, C);
        return false;
    }
    DCsr P, AP;
    prolongator_csr(c, A.n, nc, agg, pval, P);
    spgemm(c, A, P, AP);
    spgemm(c, R, AP, C);
    return false;
}

// ------------------------------------------------------- row-tile index ---
// One workgroup per tile of TILE_ROWS rows: its columns sorted in LDS (bitonic,
// padded to a power of two), the distinct ones counted (phase 0) or listed and
// every entry given its position in the list (phase 1).  A tile is indexed when
// it has entries, its paired loads from an even start fit one pass
// (cnt + (base & 1) <= TILE_MAX_ENTRIES) and it has at most TILE_MAX_COLS
// distinct columns; others count 0.
__global__ __launch_bounds__(WG) void k_tile_cols(int n, const int* __restrict__ row, const int* __restrict__ col,
                                                  int* __restrict__ ucnt, const int* __restrict__ tu,
                                                  int* __restrict__ tucol, unsigned short* __restrict__ tlidx, int phase) {
    __shared__ int keys[TILE_MAX_ENTRIES];
    __shared__ int uniq[TILE_MAX_COLS];
    __shared__ int wsum[WG / 64];
    int t = blockIdx.x, tid = threadIdx.x;
    int r0 = t * TILE_ROWS;
    int rend = r0 + TILE_ROWS < n ? r0 + TILE_ROWS : n;
    int base = row[r0], cnt = row[rend] - base;
    bool fits = cnt > 0 && cnt + (base & 1) <= TILE_MAX_ENTRIES;
    if (phase == 1 && tu[t + 1] == tu[t]) return;  // not indexed
    if (!fits) {
        if (phase == 0 && tid == 0) ucnt[t] = 0;
        return;
    }
    int P = 1;
    while (P < cnt) P <<= 1;
    for (int p = tid; p < P; p += WG) keys[p] = p < cnt ? col[base + p] : INT_MAX;
    __syncthreads();
    for (int kk = 2; kk <= P; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int p = tid; p < P; p += WG) {
                int q = p ^ j;
                if (q > p) {
                    int a = keys[p], b = keys[q];
                    bool up = (p & kk) == 0;
                    if ((a > b) == up) {
                        keys[p] = b;
                        keys[q] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    // distinct columns in order: heads of runs, ranked by a block scan per chunk of WG
    int lane = tid & 63, w = tid >> 6;
    int carry = 0;
    for (int c0 = 0; c0 < cnt; c0 += WG) {
        int p = c0 + tid;
        bool head = p < cnt && (p == 0 || keys[p] != keys[p - 1]);
        unsigned long long m = __ballot(head);
        int below = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int off = carry;
        for (int q = 0; q < w; q++) off += wsum[q];
        if (phase == 1 && head && off + below < TILE_MAX_COLS) uniq[off + below] = keys[p];
        int tot = 0;
        for (int q = 0; q < WG / 64; q++) tot += wsum[q];
        carry += tot;
        __syncthreads();
    }
    int U = carry <= TILE_MAX_COLS ? carry : 0;
    if (phase == 0) {
        if (tid == 0) ucnt[t] = U;
        return;
    }
    int u0 = tu[t];
    for (int u = tid; u < U; u += WG) tucol[u0 + u] = uniq[u];
    for (int j = tid; j < cnt; j += WG) {
        int c = col[base + j];
        int lo = 0, hi = U - 1;  // uniq[lo] <= c <= uniq[hi], c is in the list
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (uniq[mid] < c) lo = mid + 1;
            else hi = mid;
        }
        tlidx[base + j] = (unsigned short)lo;
    }
}

void build_tile_index(Context* c, DCsr& A) {
    if (A.tiled() || A.n <= 0) return;
    int nt = (A.n + TILE_ROWS - 1) / TILE_ROWS;
    DBuf<int> cnt(c, nt), tu(c, (size_t)nt + 1);
    k_tile_cols<<<nt, WG, 0, c->stream>>>(A.n, A.row.get(), A.col.get(), cnt.get(), nullptr, nullptr, nullptr, 0);
    BCMGX_LAUNCHED();
    long long tot = scan_exclusive(c, cnt.get(), tu.get(), nt);
    DBuf<int> tucol(c, (size_t)std::max<long long>(tot, 1));
    DBuf<unsigned short> tlidx(c, (size_t)std::max(A.nnz, 2));
    k_tile_cols<<<nt, WG, 0, c->stream>>>(A.n, A.row.get(), A.col.get(), cnt.get(), tu.get(), tucol.get(), tlidx.get(), 1);
    BCMGX_LAUNCHED();
    A.tu = std::move(tu);
    A.tucol = std::move(tucol);
    A.tlidx = std::move(tlidx);
}

// ------------------------------------------------------ adjoint gradient ---
// g[e] = -sum_c lambda[row(e), c] x[col(e), c] on the stored pattern of A
template <int L>
__global__ __launch_bounds__(WG) void k_outer_grad(int n, int k, const int* __restrict__ row, const int* __restrict__ col,
                                                   const double* __restrict__ lam, const double* __restrict__ x,
                                                   double* __restrict__ g) {
    int i = (blockIdx.x * WG + (int)threadIdx.x) / L;
    int lane = (int)threadIdx.x % L;
    if (i >= n) return;
    for (int j = row[i] + lane; j < row[i + 1]; j += L) {
        long long cc = col[j];
        double s = 0.0;
        for (int q = 0; q < k; q++) s += lam[(long long)i * k + q] * x[cc * k + q];
        g[j] = -s;
    }
}
void sparse_outer_grad(Context* c, const DCsr& A, const double* lambda, const double* x, int k, double* g) {
    if (A.n <= 0) return;
    k_outer_grad<8><<<grid_for(A.n, WG / 8), WG, 0, c->stream>>>(A.n, k, A.row.get(), A.col.get(), lambda, x, g);
    BCMGX_LAUNCHED();
}

}  // namespace bcmgx
