// spmv.hip -- CSR SpMV / SpMM, l1-Jacobi sweeps and grid transfers.
//
// One vector (k == 1): a group of L lanes (L | 64) owns one row and strides
// over its nonzeros, then reduces with xor-shuffles inside the group -- the
// wave64 counterpart of the reference's mini-warp kernels
// (_CSR_vector_mul_mini_warp<OP>, src/datastruct/CSR.cu:452-492;
// _jacobi_it_full, src/preconditioner/l1jacobi/l1jacobi.cu:344-379).
// k vectors (k > 1): a group of KP >= k lanes owns one row, lane c handles
// column c, so every (val, col) pair is loaded once for all k right-hand sides
// and the x gather of a nonzero is one contiguous k-wide segment.
// Matrix values and columns are streamed with nontemporal loads so the L2
// keeps the gathered x; blocks are remapped so each XCD walks one contiguous
// range of rows.
#include "kernels.h"

namespace bcmgx {

template <typename T, typename V, int L, SpOp OP>
__global__ __launch_bounds__(WG) void k_spmv1(int n, const int* __restrict__ row, const int* __restrict__ col,
                                              const T* __restrict__ val, const V* __restrict__ x, V* __restrict__ y,
                                              const V* __restrict__ b, const V* __restrict__ D, V alpha, V beta,
                                              const int* __restrict__ stop) {
    if (stop && *stop) return;
    constexpr int RPB = WG / L;
    int blk = xcd_remap(blockIdx.x, gridDim.x);
    int r = blk * RPB + (int)threadIdx.x / L;
    int lane = (int)threadIdx.x % L;
    if (r >= n) return;  // a whole L-lane group leaves together
    int s = row[r], e = row[r + 1];
    V t = 0;
    for (int j = s + lane; j < e; j += L) {
        V a = (V)ldg_nt(val + j);
        int cc = ldg_nt(col + j);
        if (OP == SpOp::AXPBY) t += (alpha * a) * x[cc];
        else t += a * x[cc];
    }
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, L);
    if (lane == 0) {
        if (OP == SpOp::AX) y[r] = t;
        else if (OP == SpOp::RES) y[r] = b[r] - t;
        else if (OP == SpOp::AXPBY) y[r] = t + beta * y[r];
        else y[r] = ((-t + b[r]) / D[r]) + x[r];
    }
}

// order LDS accesses of the lanes of one wave (compiler fence; no hardware barrier)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, typename V, int KP, SpOp OP>
__global__ __launch_bounds__(WG) void k_spmmk(int n, int k, int c0, int kc, const int* __restrict__ row,
                                              const int* __restrict__ col, const T* __restrict__ val,
                                              const V* __restrict__ x, V* __restrict__ y, const V* __restrict__ b,
                                              const V* __restrict__ D, V alpha, V beta, const int* __restrict__ stop) {
    if (stop && *stop) return;
    constexpr int RPB = WG / KP;
    constexpr int U = KP < 16 ? KP : 16;  // independent x gathers in flight per lane
    // one (col, val) slot per thread: the KP lanes of a row load KP entries of
    // it with one coalesced access, then read them back as LDS broadcasts, so
    // no x gather waits on a matrix load
    __shared__ int scol[WG];
    __shared__ T sval[WG];
    int blk = xcd_remap(blockIdx.x, gridDim.x);
    int tid = threadIdx.x;
    int r = blk * RPB + tid / KP;
    int cl = tid % KP;
    int g0 = tid - cl;  // first slot of this row's group
    if (r >= n) return;  // whole group (groups never straddle a wave)
    int c = c0 + cl;  // this lane's column among the k interleaved ones
    bool active = cl < kc;
    int s = row[r], e = row[r + 1];
    const V* xc = x + c;
    V t = 0;
    for (int base = s; base < e; base += KP) {
        int j = base + cl;
        scol[tid] = j < e ? col[j] : 0;
        sval[tid] = j < e ? val[j] : (T)0;
        wave_lds_sync();
        int cnt = e - base < KP ? e - base : KP;
        if (active) {
            // U divides KP, so q + u < KP stays inside this group's slots
            for (int q = 0; q < cnt; q += U) {
                V xv[U];
#pragma unroll
                for (int u = 0; u < U; u++) xv[u] = q + u < cnt ? xc[(long long)scol[g0 + q + u] * k] : (V)0;
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (q + u < cnt) {
                        V a = (V)sval[g0 + q + u];
                        if (OP == SpOp::AXPBY) t += (alpha * a) * xv[u];
                        else t += a * xv[u];
                    }
                }
            }
        }
        wave_lds_sync();
    }
    if (!active) return;
    long long o = (long long)r * k + c;
    if (OP == SpOp::AX) y[o] = t;
    else if (OP == SpOp::RES) y[o] = b[o] - t;
    else if (OP == SpOp::AXPBY) y[o] = t + beta * y[o];
    else y[o] = ((-t + b[o]) / D[r]) + x[o];
}

// k-vector SpMM with CPL consecutive columns per lane (one 8-16 byte vector
// gather per nonzero): a group of KP lanes owns a row, stages M*KP (col, val)
// entries of it in LDS with coalesced loads, then every lane gathers its CPL
// columns of the x rows, U gathers in flight.  Each column's sum runs over the
// row in order, as in k_spmmk, so the results are the same bits.  Needs
// k % CPL == 0 and KP * CPL >= k; lanes past k / CPL idle.
template <typename T, typename V, int KP, int CPL, int M, SpOp OP, bool OUT64 = false>
__global__ __launch_bounds__(WG) void k_spmmv(int n, int k, const int* __restrict__ row, const int* __restrict__ col,
                                              const T* __restrict__ val, const V* __restrict__ x, V* __restrict__ y,
                                              const V* __restrict__ b, const V* __restrict__ D, V alpha, V beta,
                                              const int* __restrict__ stop, double* __restrict__ y64 = nullptr) {
    typedef V VV __attribute__((ext_vector_type(CPL)));
    typedef double VD __attribute__((ext_vector_type(CPL)));
    if (stop && *stop) return;
    constexpr int RPB = WG / KP;
    constexpr int E = KP * M;
    constexpr int U = 16;
    __shared__ int scol[WG * M];
    __shared__ T sval[WG * M];
    int blk = xcd_remap(blockIdx.x, gridDim.x);
    int tid = threadIdx.x;
    int cl = tid % KP;
    int g = tid / KP;
    int r = blk * RPB + g;
    if (r >= n) return;  // whole group (groups never straddle a wave)
    bool active = cl * CPL < k;
    int s = row[r], e = row[r + 1];
    int* sc = scol + g * E;
    T* sv = sval + g * E;
    const V* xc = x + cl * CPL;
    VV t = (VV)(V)0;
    for (int base = s; base < e; base += E) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            int j = base + m * KP + cl;
            sc[m * KP + cl] = j < e ? col[j] : 0;
            sv[m * KP + cl] = j < e ? val[j] : (T)0;
        }
        wave_lds_sync();
        int cnt = e - base < E ? e - base : E;
        if (active) {
            for (int q = 0; q < cnt; q += U) {
                VV xv[U];
#pragma unroll
                for (int u = 0; u < U; u++)
                    xv[u] = q + u < cnt ? *(const VV*)(xc + (long long)sc[q + u] * k) : (VV)(V)0;
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (q + u < cnt) {
                        V a = (V)sv[q + u];
                        if (OP == SpOp::AXPBY) t += (alpha * a) * xv[u];
                        else t += a * xv[u];
                    }
                }
            }
        }
        wave_lds_sync();
    }
    if (!active) return;
    long long o = (long long)r * k + cl * CPL;
    if constexpr (OUT64) {  // JACOBI result widened to FP64 (the last sweep of an FP32 cycle)
        VV res = ((-t + *(const VV*)(b + o)) / D[r]) + *(const VV*)(x + o);
        *(VD*)(y64 + o) = __builtin_convertvector(res, VD);
        return;
    }
    VV* yo = (VV*)(y + o);
    if (OP == SpOp::AX) *yo = t;
    else if (OP == SpOp::RES) *yo = *(const VV*)(b + o) - t;
    else if (OP == SpOp::AXPBY) *yo = t + beta * *yo;
    else *yo = ((-t + *(const VV*)(b + o)) / D[r]) + *(const VV*)(x + o);
}

// Operands of the pipelined SpMM's optional epilogues.
struct SpmmExtra {
    double* y64 = nullptr;         // MODE 1: JACOBI result widened to FP64
    void* x0 = nullptr;            // MODE 3 (AX): y / D, a first l1-Jacobi sweep from zero on y
    const double* r = nullptr;     // MODE 2: FCG residual
    const double* wold = nullptr;  // MODE 2: previous A-direction
    double* partials = nullptr;    // MODE 2: per-block dot partials
    const int* agg = nullptr;      // MODE 4: each row's coarse index (x is the coarse correction e)
    const void* pv = nullptr;      // MODE 4: the prolongator's value of each row (V*)
    const void* xpre = nullptr;    // MODE 4: the pre-smoothed iterate (V*)
    const double* b64 = nullptr;   // RES (MODE 0): b held in FP64, rounded to V as it is read
};

// Software-pipelined persistent k-vector SpMM, CPL consecutive columns per
// lane (16-byte gathers), KP lanes per row.  Rows go to XCDs (hardware
// dispatch puts block b on XCD b % 8) in one of two ways: sg == 0 cuts them
// into 8 contiguous segments, one per XCD; sg > 0 deals super-groups of sg
// row groups round-robin to the XCDs, so that all XCDs sweep the same band of
// rows at once and a row's far neighbours (a stencil's next plane) fall on
// the XCD, and into the L2, that loaded them as its own rows' x a short while
// before.  The blocks of an XCD stride through its rows in order.  Each KP-lane
// group walks its own sequence of rows with a three-stage pipeline: the row
// bounds of row i+2 and the first M*KP (col, val) entries of row i+1 are in
// flight while row i's x gathers (16 per lane) are.  Every column's sum runs
// over the row in order, so results are bit-identical to the one-vector
// kernels' sequential order.
//   MODE 0: y = op(A x);  MODE 1: JACOBI written widened to FP64 (ex.y64);
//   MODE 2 (FP64, AX): w = A p and FCG's triple dot product (r.p, w.p, wold.p)
//          per column (triple_innerproduct, src/op/triple_inner_product.cu:7-101),
//          reduced per block in a fixed order into ex.partials[(blk*k + c)*3 + m].
//   MODE 4: the first post-smoothing sweep from the product A P (the matrix
//          here, n x nc) and the coarse correction e (x here): with
//          z1 = p_i e[agg_i] + xpre_i (the prolongation), y = ((-t + b) / D) + z1,
//          t = (A P) e, b the pre-smoothing residual.
template <typename T, typename V, int KP, int CPL, int M, SpOp OP, int MODE, bool NT = false, typename XT = V>
__global__ __launch_bounds__(WG) void k_spmmp(int n, int k, const int* __restrict__ row, const int* __restrict__ col,
                                              const T* __restrict__ val, const XT* __restrict__ x, V* __restrict__ y,
                                              const V* __restrict__ b, const V* __restrict__ D, V alpha, V beta,
                                              const int* __restrict__ stop, SpmmExtra ex, int sg) {
    typedef V VV __attribute__((ext_vector_type(CPL)));
    typedef XT XV __attribute__((ext_vector_type(CPL)));  // x may be narrower than V: widened exactly on load
    typedef double VD __attribute__((ext_vector_type(CPL)));
    if (stop && *stop) return;
    constexpr int RPB = WG / KP;
    constexpr int E = KP * M;
    // gathers in flight per lane: 16, or 8 for widened 4-column FP64 accumulators (registers)
    constexpr int U = (sizeof(XT) < sizeof(V) && CPL >= 4) ? 8 : 16;
    __shared__ int scol[WG * M];
    __shared__ T sval[WG * M];
    int tid = threadIdx.x;
    int cl = tid % KP;
    int g = tid / KP;
    bool active = cl * CPL < k;
    const int X = 8;
    int xcd = blockIdx.x % X, local = blockIdx.x / X, nloc = gridDim.x / X;
    int ngroups = (n + RPB - 1) / RPB;
    int g0 = (int)((long long)ngroups * xcd / X), g1 = (int)((long long)ngroups * (xcd + 1) / X);
    int* sc = scol + g * E;
    T* sv = sval + g * E;
    const XT* xc = x + cl * CPL;
    VD da = (VD)0.0, db = (VD)0.0, dc = (VD)0.0;
    // the t-th row group of this XCD, -1 past its last (increasing in t)
    auto group_of = [&](int q) -> int {
        if (sg <= 0) return g0 + q < g1 ? g0 + q : -1;
        long long gq = ((long long)(q / sg) * X + xcd) * sg + q % sg;
        return gq < ngroups ? (int)gq : -1;
    };
    // row grp*RPB + g (-1 past the end) and its bounds
    auto bounds = [&](int grp, int& ri, int& rs, int& re) {
        int pos = grp * RPB + g;
        if (grp >= 0 && pos < n) {
            ri = pos;
            rs = row[ri];
            re = row[ri + 1];
        } else {
            ri = -1;
            rs = 0;
            re = 0;
        }
    };
    auto chunk = [&](int base, int re, int* pc, T* pv) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            int j = base + m * KP + cl;
            if constexpr (NT) {  // matrix entries streamed past the caches, which keep x
                pc[m] = j < re ? __builtin_nontemporal_load(col + j) : 0;
                pv[m] = j < re ? __builtin_nontemporal_load(val + j) : (T)0;
            } else {
                pc[m] = j < re ? col[j] : 0;
                pv[m] = j < re ? val[j] : (T)0;
            }
        }
    };
    auto gather = [&](int cnt, VV& t) {
        if (!active) return;
        for (int q = 0; q < cnt; q += U) {
            XV xv[U];
            if constexpr (std::is_same<XT, V>::value) {
#pragma unroll
                for (int u = 0; u < U; u++) xv[u] = q + u < cnt ? *(const XV*)(xc + (long long)sc[q + u] * k) : (XV)(XT)0;
            } else {
                // narrow x: unconditional loads (past the row, row 0, weighted by nothing),
                // so that the widening stays out of the load branches and all U loads are
                // in flight together
#pragma unroll
                for (int u = 0; u < U; u++) {
                    int cc = q + u < cnt ? sc[q + u] : 0;
                    xv[u] = *(const XV*)(xc + (long long)cc * k);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (q + u < cnt) {
                    V a = (V)sv[q + u];
                    VV xw = __builtin_convertvector(xv[u], VV);
                    if (OP == SpOp::AXPBY) t += (alpha * a) * xw;
                    else t += a * xw;
                }
            }
        }
    };
    int it = local;  // uniform over the block
    int grp = group_of(it);
    int r0, s0, e0, r1, s1, e1;
    bounds(grp, r0, s0, e0);
    bounds(group_of(it + nloc), r1, s1, e1);
    int pc[M];
    T pv[M];
    chunk(s0, e0, pc, pv);
    while (grp >= 0) {
        int r2, s2, e2;
        bounds(group_of(it + 2 * nloc), r2, s2, e2);
#pragma unroll
        for (int m = 0; m < M; m++) {
            sc[m * KP + cl] = pc[m];
            sv[m * KP + cl] = pv[m];
        }
        wave_lds_sync();
        int pc1[M];
        T pv1[M];
        chunk(s1, e1, pc1, pv1);  // next row's first entries, in flight during the gathers
        VV t = (VV)(V)0;
        gather(e0 - s0 < E ? e0 - s0 : E, t);
        wave_lds_sync();
        for (int base = s0 + E; base < e0; base += E) {  // rows longer than E: the rest unpipelined
            int qc[M];
            T qv[M];
            chunk(base, e0, qc, qv);
#pragma unroll
            for (int m = 0; m < M; m++) {
                sc[m * KP + cl] = qc[m];
                sv[m * KP + cl] = qv[m];
            }
            wave_lds_sync();
            gather(e0 - base < E ? e0 - base : E, t);
            wave_lds_sync();
        }
        int ri = r0;
        if (ri >= 0 && active) {
            long long o = (long long)ri * k + cl * CPL;
            if constexpr (MODE == 1) {
                VV res = ((-t + *(const VV*)(b + o)) / D[ri]) + __builtin_convertvector(*(const XV*)(x + o), VV);
                *(VD*)(ex.y64 + o) = __builtin_convertvector(res, VD);
            } else if constexpr (MODE == 3) {
                *(VV*)(y + o) = t;
                *(VV*)((V*)ex.x0 + o) = t / D[ri];
            } else if constexpr (MODE == 4) {
                VV e = __builtin_convertvector(*(const XV*)(x + (long long)ex.agg[ri] * k + cl * CPL), VV);
                VV z1 = ((const V*)ex.pv)[ri] * e + *(const VV*)((const V*)ex.xpre + o);
                *(VV*)(y + o) = ((-t + *(const VV*)(b + o)) / D[ri]) + z1;
            } else if constexpr (MODE == 2) {
                *(VV*)(y + o) = t;
                VD pw = __builtin_convertvector(*(const XV*)(x + o), VD);
                da += *(const VD*)(ex.r + o) * pw;
                db += t * pw;
                dc += *(const VD*)(ex.wold + o) * pw;
            } else {
                VV* yo = (VV*)(y + o);
                if (OP == SpOp::AX) *yo = t;
                else if (OP == SpOp::RES)
                    *yo = (ex.b64 ? __builtin_convertvector(*(const VD*)(ex.b64 + o), VV) : *(const VV*)(b + o)) - t;
                else if (OP == SpOp::AXPBY) *yo = t + beta * *yo;
                else *yo = ((-t + *(const VV*)(b + o)) / D[ri]) + __builtin_convertvector(*(const XV*)(x + o), VV);
            }
        }
        it += nloc;
        grp = group_of(it);
        r0 = r1; s0 = s1; e0 = e1;
        r1 = r2; s1 = s2; e1 = e2;
#pragma unroll
        for (int m = 0; m < M; m++) {
            pc[m] = pc1[m];
            pv[m] = pv1[m];
        }
    }
    if constexpr (MODE == 2) {
        // fixed-order block reduction over the row groups, per column
        __shared__ double red[3 * CPL][WG];
#pragma unroll
        for (int q = 0; q < CPL; q++) {
            red[3 * q + 0][tid] = da[q];
            red[3 * q + 1][tid] = db[q];
            red[3 * q + 2][tid] = dc[q];
        }
        __syncthreads();
        for (int st = WG / 2; st >= KP; st >>= 1) {
            if (tid < st) {
#pragma unroll
                for (int m = 0; m < 3 * CPL; m++) red[m][tid] += red[m][tid + st];
            }
            __syncthreads();
        }
        if (tid < KP && active) {
#pragma unroll
            for (int q = 0; q < CPL; q++) {
                double* pp = ex.partials + ((long long)blockIdx.x * k + tid * CPL + q) * 3;
                pp[0] = red[3 * q + 0][tid];
                pp[1] = red[3 * q + 1][tid];
                pp[2] = red[3 * q + 2][tid];
            }
        }
    }
}

// resident blocks of one pipelined-SpMM instantiation on the whole device, a multiple of 8
template <typename K>
static int persistent_grid(Context* c, K kernel) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, WG, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    int g = c->num_cus * per_cu;
    return std::max(8, g / 8 * 8);
}

// super-group size (row groups) of the XCD-interleaved row order that a column
// choice asks for: kchunk <= -16 is the pipelined kernel with super-groups of
// 2^(-16 - kchunk) groups; other choices keep the per-XCD segments (0)
static int kchunk_sg(int kchunk) { return kchunk <= -16 && kchunk >= -28 ? 1 << (-16 - kchunk) : 0; }

template <typename T, typename V, int KP, int CPL, int M, SpOp OP, int MODE, bool NT, typename XT = V>
static void launch_spmmp(Context* c, const SpmvArgs& a, const SpmmExtra& ex, int grid) {
    auto kern = k_spmmp<T, V, KP, CPL, M, OP, MODE, NT, XT>;
    if (grid <= 0) {
        static int cached = 0;  // per instantiation; identical devices assumed within a process
        if (!cached) cached = persistent_grid(c, kern);
        grid = cached;
    }
    kern<<<grid, WG, 0, c->stream>>>(a.n, a.k, a.row, a.col, (const T*)a.val, (const XT*)a.x, (V*)a.y, (const V*)a.b,
                                     (const V*)a.D, (V)a.alpha, (V)a.beta, a.stop, ex, kchunk_sg(a.kchunk));
    BCMGX_LAUNCHED();
}

// lanes per row of the pipelined SpMM for k columns of CPL per lane, or 0 when it does not apply
static int spmmp_lanes(int k, int cpl) {
    if (k % cpl || k < 2 * cpl || k > 64) return 0;
    int lanes = k / cpl, KP = 2;
    while (KP < lanes) KP <<= 1;
    return KP;
}

template <typename T, typename V, SpOp OP, int MODE, bool NT = false>
static bool dispatch_spmmp(Context* c, const SpmvArgs& a, const SpmmExtra& ex, int grid) {
    constexpr int CPL = 16 / (int)sizeof(V);
    switch (spmmp_lanes(a.k, CPL)) {
    case 2: launch_spmmp<T, V, 2, CPL, 8, OP, MODE, NT>(c, a, ex, grid); return true;
    case 4: launch_spmmp<T, V, 4, CPL, 8, OP, MODE, NT>(c, a, ex, grid); return true;
    case 8: launch_spmmp<T, V, 8, CPL, 4, OP, MODE, NT>(c, a, ex, grid); return true;
    case 16: launch_spmmp<T, V, 16, CPL, 2, OP, MODE, NT>(c, a, ex, grid); return true;
    case 32: launch_spmmp<T, V, 32, CPL, 1, OP, MODE, NT>(c, a, ex, grid); return true;
    default: return false;
    }
}

template <typename T, typename V, SpOp OP, int MODE>
static int spmmp_grid(Context* c, int k) {
    constexpr int CPL = 16 / (int)sizeof(V);
    switch (spmmp_lanes(k, CPL)) {
    case 2: return persistent_grid(c, k_spmmp<T, V, 2, CPL, 8, OP, MODE>);
    case 4: return persistent_grid(c, k_spmmp<T, V, 4, CPL, 8, OP, MODE>);
    case 8: return persistent_grid(c, k_spmmp<T, V, 8, CPL, 4, OP, MODE>);
    case 16: return persistent_grid(c, k_spmmp<T, V, 16, CPL, 2, OP, MODE>);
    case 32: return persistent_grid(c, k_spmmp<T, V, 32, CPL, 1, OP, MODE>);
    default: return 0;
    }
}

// grid of the fused FCG kernel for k columns: the device's resident block
// count, fixed for a device, so the partial count never changes between calls
int spmm_dots_blocks(Context* c, int n, int k) {
    (void)n;
    return spmmp_grid<double, double, SpOp::AX, 2>(c, k);
}

bool spmm_dots(Context* c, const SpmvArgs& a, const double* r, const double* wold, double* partials, int nblk,
               const int* all_done) {
    bool aligned = (((uintptr_t)a.x | (uintptr_t)a.y | (uintptr_t)r | (uintptr_t)wold) & 15) == 0;
    if (nblk <= 0 || nblk % 8 || !aligned || !spmmp_lanes(a.k, 2) || (a.kchunk > 0 && a.kchunk < a.k)) return false;
    SpmmExtra ex;
    ex.r = r;
    ex.wold = wold;
    ex.partials = partials;
    SpmvArgs b = a;
    b.stop = all_done;
    if (a.kchunk == -2) return dispatch_spmmp<double, double, SpOp::AX, 2, true>(c, b, ex, nblk);
    return dispatch_spmmp<double, double, SpOp::AX, 2>(c, b, ex, nblk);
}

// The fused FCG product with the preconditioned residual z held in FP32 -- the
// values an FP32 V-cycle computes, before any widening.  Each lane gathers four
// columns as one 16-byte float4 and widens them exactly into FP64 accumulators:
// W0 has the bits of the widened-FP64 path (each column's row sum in order), from
// half the gathered bytes and half the load instructions per row; the three dot
// products are summed over this kernel's own (fixed) row groups.
template <int KP, int M>
static int z32_grid(Context* c) {
    return persistent_grid(c, k_spmmp<double, double, KP, 4, M, SpOp::AX, 2, false, float>);
}

int spmm_dots_z32_blocks(Context* c, int n, int k) {
    (void)n;
    switch (spmmp_lanes(k, 4)) {
    case 2: return z32_grid<2, 8>(c);
    case 4: return z32_grid<4, 8>(c);
    case 8: return z32_grid<8, 4>(c);
    case 16: return z32_grid<16, 2>(c);
    default: return 0;
    }
}

bool spmm_dots_z32(Context* c, const SpmvArgs& a, const double* r, const double* wold, double* partials, int nblk,
                   const int* all_done) {
    bool aligned = (((uintptr_t)a.x | (uintptr_t)a.y | (uintptr_t)r | (uintptr_t)wold) & 15) == 0;
    if (nblk <= 0 || nblk % 8 || !aligned || (a.kchunk > 0 && a.kchunk < a.k)) return false;
    SpmmExtra ex;
    ex.r = r;
    ex.wold = wold;
    ex.partials = partials;
    SpmvArgs b = a;
    b.stop = all_done;
    switch (spmmp_lanes(a.k, 4)) {
    case 2: launch_spmmp<double, double, 2, 4, 8, SpOp::AX, 2, false, float>(c, b, ex, nblk); return true;
    case 4: launch_spmmp<double, double, 4, 4, 8, SpOp::AX, 2, false, float>(c, b, ex, nblk); return true;
    case 8: launch_spmmp<double, double, 8, 4, 4, SpOp::AX, 2, false, float>(c, b, ex, nblk); return true;
    case 16: launch_spmmp<double, double, 16, 4, 2, SpOp::AX, 2, false, float>(c, b, ex, nblk); return true;
    default: return false;
    }
}

// CSR-stream SpMV for one vector: a workgroup owns RB = 256/TPR consecutive
// rows; the products of their (contiguous) nonzeros are formed with fully
// coalesced streaming loads into LDS (eight independent val/col/x loads per
// thread in flight), then TPR lanes reduce each row from LDS.  The LDS tile
// holds 32 nonzeros per row on average (so a block stays small enough for
// full occupancy); blocks with more reduce straight from global memory.
template <typename T, typename V, int TPR, SpOp OP>
__global__ __launch_bounds__(WG) void k_spmv_stream(int n, const int* __restrict__ row, const int* __restrict__ col,
                                                    const T* __restrict__ val, const V* __restrict__ x, V* __restrict__ y,
                                                    const V* __restrict__ b, const V* __restrict__ D, V alpha, V beta,
                                                    const int* __restrict__ stop, int sb) {
    if (stop && *stop) return;
    constexpr int RB = WG / TPR;
    constexpr int CAP = RB * 32 < 4096 ? RB * 32 : 4096;
    constexpr int U = 8;
    __shared__ V prod[CAP];
    __shared__ int rp[RB + 1];
    int blk = xcd_order(blockIdx.x, gridDim.x, sb);
    int r0 = blk * RB;
    int tid = threadIdx.x;
    // the block's nonzero range from two uniform (scalar) loads; the per-row
    // pointers of the reduction load alongside the val/col stream below
    int rend = r0 + RB < n ? r0 + RB : n;
    int base = row[r0], cnt = row[rend] - base;
    int myrp = 0;
    if (tid < RB) myrp = row[r0 + tid < n ? r0 + tid : n];
    bool staged = cnt <= CAP;
    if (staged) {
        for (int t0 = tid; t0 < cnt; t0 += U * WG) {
            V a[U], xv[U];
            int cc[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                int t = t0 + q * WG;
                a[q] = 0;
                cc[q] = 0;
                if (t < cnt) {
                    a[q] = (V)ldg_nt(val + base + t);
                    cc[q] = ldg_nt(col + base + t);
                }
            }
#pragma unroll
            for (int q = 0; q < U; q++) xv[q] = (t0 + q * WG < cnt) ? x[cc[q]] : (V)0;
#pragma unroll
            for (int q = 0; q < U; q++) {
                int t = t0 + q * WG;
                if (t < cnt) prod[t] = (OP == SpOp::AXPBY ? alpha * a[q] : a[q]) * xv[q];
            }
        }
    }
    if (tid < RB) rp[tid] = myrp;
    if (tid == 0) rp[RB] = base + cnt;  // = row[rend]
    __syncthreads();
    int lr = tid / TPR, sub = tid % TPR;
    int r = r0 + lr;
    if (r >= n) return;  // whole TPR-lane group
    int s = rp[lr], e = rp[lr + 1];
    V acc = 0;
    if (staged) {
        for (int q = s + sub; q < e; q += TPR) acc += prod[q - base];
    } else {
        for (int q = s + sub; q < e; q += TPR) {
            V a = (V)val[q];
            if (OP == SpOp::AXPBY) a = alpha * a;
            acc += a * x[col[q]];
        }
    }
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, TPR);
    if (sub == 0) {
        if (OP == SpOp::AX) y[r] = acc;
        else if (OP == SpOp::RES) y[r] = b[r] - acc;
        else if (OP == SpOp::AXPBY) y[r] = acc + beta * y[r];
        else y[r] = ((-acc + b[r]) / D[r]) + x[r];
    }
}

// The CSR-stream SpMV with the matrix streamed in pairs: every thread loads
// two consecutive (val, col) entries with one 2-wide load each (16 + 8 bytes
// for FP64), from an even-aligned start, so a block issues half as many
// memory instructions.  Same chunks, staging and per-row reduction as
// k_spmv_stream, hence the same bits.  The final pair of the matrix is read
// element by element when nnz is odd (no read past the arrays).
template <typename T, typename V, int TPR, SpOp OP>
__global__ __launch_bounds__(WG) void k_spmv_stream2(int n, long long nnz, const int* __restrict__ row,
                                                     const int* __restrict__ col, const T* __restrict__ val,
                                                     const V* __restrict__ x, V* __restrict__ y, const V* __restrict__ b,
                                                     const V* __restrict__ D, V alpha, V beta,
                                                     const int* __restrict__ stop, int sb) {
    typedef T T2 __attribute__((ext_vector_type(2)));
    typedef int I2 __attribute__((ext_vector_type(2)));
    if (stop && *stop) return;
    constexpr int RB = WG / TPR;
    constexpr int CAP = RB * 32 < 4096 ? RB * 32 : 4096;
    constexpr int UP = 4;  // pairs per thread per pass
    __shared__ V prod[CAP];
    __shared__ int rp[RB + 1];
    int blk = xcd_order(blockIdx.x, gridDim.x, sb);
    int r0 = blk * RB;
    int tid = threadIdx.x;
    int rend = r0 + RB < n ? r0 + RB : n;
    int base = row[r0], cnt = row[rend] - base;
    int myrp = 0;
    if (tid < RB) myrp = row[r0 + tid < n ? r0 + tid : n];
    bool staged = cnt <= CAP;
    if (staged) {
        int abase = base & ~1;             // even global start: 2-wide loads stay aligned
        int acnt = base + cnt - abase;     // entries from abase to the chunk end
        const T2* v2 = (const T2*)(val + abase);
        const I2* c2 = (const I2*)(col + abase);
        for (int p0 = tid; 2 * p0 < acnt; p0 += UP * WG) {
            T2 a[UP];
            I2 cc[UP];
#pragma unroll
            for (int q = 0; q < UP; q++) {
                int pp = p0 + q * WG;
                a[q] = (T2)(T)0;
                cc[q] = (I2)0;
                if (2 * pp < acnt) {
                    long long g = (long long)abase + 2 * pp;
                    if (g + 1 < nnz) {
                        a[q] = __builtin_nontemporal_load(v2 + pp);
                        cc[q] = __builtin_nontemporal_load(c2 + pp);
                    } else {  // the matrix's last entry, unpaired
                        a[q].x = val[g];
                        cc[q].x = col[g];
                    }
                }
            }
            V x0[UP], x1[UP];
#pragma unroll
            for (int q = 0; q < UP; q++) {
                int t = abase + 2 * (p0 + q * WG) - base;  // chunk position of the pair's first entry (may be -1)
                x0[q] = (t >= 0 && t < cnt) ? x[cc[q].x] : (V)0;
                x1[q] = (t + 1 >= 0 && t + 1 < cnt) ? x[cc[q].y] : (V)0;
            }
#pragma unroll
            for (int q = 0; q < UP; q++) {
                int t = abase + 2 * (p0 + q * WG) - base;
                V a0 = (V)a[q].x, a1 = (V)a[q].y;
                if (OP == SpOp::AXPBY) {
                    a0 = alpha * a0;
                    a1 = alpha * a1;
                }
                if (t >= 0 && t < cnt) prod[t] = a0 * x0[q];
                if (t + 1 >= 0 && t + 1 < cnt) prod[t + 1] = a1 * x1[q];
            }
        }
    }
    if (tid < RB) rp[tid] = myrp;
    if (tid == 0) rp[RB] = base + cnt;  // = row[rend]
    __syncthreads();
    int lr = tid / TPR, sub = tid % TPR;
    int r = r0 + lr;
    if (r >= n) return;  // whole TPR-lane group
    int s = rp[lr], e = rp[lr + 1];
    V acc = 0;
    if (staged) {
        for (int q = s + sub; q < e; q += TPR) acc += prod[q - base];
    } else {
        for (int q = s + sub; q < e; q += TPR) {
            V a = (V)val[q];
            if (OP == SpOp::AXPBY) a = alpha * a;
            acc += a * x[col[q]];
        }
    }
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, TPR);
    if (sub == 0) {
        if (OP == SpOp::AX) y[r] = acc;
        else if (OP == SpOp::RES) y[r] = b[r] - acc;
        else if (OP == SpOp::AXPBY) y[r] = acc + beta * y[r];
        else y[r] = ((-acc + b[r]) / D[r]) + x[r];
    }
}

// CSR-stream SpMV over the matrix's row-tile index: a workgroup owns one tile of
// TILE_ROWS = 64 rows (4 lanes per row in the reduction).  It loads the x
// entries of the tile's distinct columns once into LDS (the columns sorted, so
// neighbouring lanes read neighbouring x), streams the (value, 16-bit local
// column) pairs, forms each product from LDS and reduces the rows as
// k_spmv_stream2 does (same products, same order: the same bits).  Fewer bytes
// than CSR (2-byte column positions, plus the distinct-column list: about 9
// per row on the 27-point grid) and a handful of coalesced x loads per tile
// instead of one gather per entry.  A tile the index leaves out is read by
// column index from global memory.
template <typename T, typename V, SpOp OP>
__global__ __launch_bounds__(WG) void k_spmv_tiled(int n, long long nnz, const int* __restrict__ row,
                                                   const int* __restrict__ col, const T* __restrict__ val,
                                                   const int* __restrict__ tu, const int* __restrict__ tucol,
                                                   const unsigned short* __restrict__ tlidx, const V* __restrict__ x,
                                                   V* __restrict__ y, const V* __restrict__ b, const V* __restrict__ D,
                                                   V alpha, V beta, const int* __restrict__ stop, int sb) {
    constexpr int TPR = WG / TILE_ROWS, RB = TILE_ROWS, UP = 4, XP = TILE_MAX_COLS / WG;
    static_assert(2 * UP * WG == TILE_MAX_ENTRIES, "one pass of paired loads covers an indexed tile");
    typedef T T2 __attribute__((ext_vector_type(2)));
    typedef unsigned short U2 __attribute__((ext_vector_type(2)));
    if (stop && *stop) return;
    // the products overwrite the x entries once every product is formed: one LDS region
    static_assert(TILE_MAX_COLS <= TILE_MAX_ENTRIES && (TILE_MAX_COLS & (TILE_MAX_COLS - 1)) == 0,
                  "x fits the products' space; positions mask to it");
    __shared__ V prod[TILE_MAX_ENTRIES];
    V* xs = prod;
    __shared__ int rp[RB + 1];
    int blk = xcd_order(blockIdx.x, gridDim.x, sb);
    int r0 = blk * RB;
    int tid = threadIdx.x;
    int rend = r0 + RB < n ? r0 + RB : n;
    int base = row[r0], cnt = row[rend] - base;
    int u0 = tu[blk], U = tu[blk + 1] - u0;
    int myrp = 0;
    if (tid < RB) myrp = row[r0 + tid < n ? r0 + tid : n];
    bool indexed = U > 0;  // an indexed tile has entries and at most TILE_MAX_* of both
    if (indexed) {
        // x of the distinct columns (index clamped into the list: every load unconditional and in flight)
        int uc[XP];
#pragma unroll
        for (int p = 0; p < XP; p++) {
            int u = tid + p * WG;
            uc[p] = tucol[u0 + (u < U ? u : U - 1)];
        }
        V xv[XP];
#pragma unroll
        for (int p = 0; p < XP; p++) xv[p] = x[uc[p]];
        // the tile's (value, local column) pairs from an even-aligned start
        int abase = base & ~1, acnt = base + cnt - abase;
        const T2* v2 = (const T2*)(val + abase);
        const U2* l2 = (const U2*)(tlidx + abase);
        T2 a[UP];
        U2 ll[UP];
#pragma unroll
        for (int q = 0; q < UP; q++) {
            int pp = tid + q * WG;
            a[q] = (T2)(T)0;
            ll[q] = (U2)0;
            if (2 * pp < acnt) {
                long long g = (long long)abase + 2 * pp;
                if (g + 1 < nnz) {
                    a[q] = __builtin_nontemporal_load(v2 + pp);
                    ll[q] = l2[pp];  // a plain load: streamed (nontemporal) 4-byte pairs measured 8% slower
                } else {  // the matrix's last entry, unpaired
                    a[q].x = val[g];
                    ll[q].x = tlidx[g];
                }
            }
        }
#pragma unroll
        for (int p = 0; p < XP; p++) xs[tid + p * WG] = xv[p];
        __syncthreads();
        V p0[UP], p1[UP];
#pragma unroll
        for (int q = 0; q < UP; q++) {
            V a0 = (V)a[q].x, a1 = (V)a[q].y;
            if (OP == SpOp::AXPBY) {
                a0 = alpha * a0;
                a1 = alpha * a1;
            }
            // a pair half outside the tile holds a neighbour tile's position, or none (an
            // unindexed tile's): masked into the x region, its product is never stored
            p0[q] = a0 * xs[ll[q].x & (TILE_MAX_COLS - 1)];
            p1[q] = a1 * xs[ll[q].y & (TILE_MAX_COLS - 1)];
        }
        __syncthreads();  // every x entry read: the products take the space
#pragma unroll
        for (int q = 0; q < UP; q++) {
            int t = abase + 2 * (tid + q * WG) - base;  // chunk position of the pair's first entry (may be -1)
            if (t >= 0 && t < cnt) prod[t] = p0[q];
            if (t + 1 >= 0 && t + 1 < cnt) prod[t + 1] = p1[q];
        }
    }
    if (tid < RB) rp[tid] = myrp;
    if (tid == 0) rp[RB] = base + cnt;
    __syncthreads();
    int lr = tid / TPR, sub = tid % TPR;
    int r = r0 + lr;
    if (r >= n) return;  // whole TPR-lane group
    int s = rp[lr], e = rp[lr + 1];
    V acc = 0;
    if (indexed) {
        for (int q = s + sub; q < e; q += TPR) acc += prod[q - base];
    } else {
        for (int q = s + sub; q < e; q += TPR) {
            V av = (V)val[q];
            if (OP == SpOp::AXPBY) av = alpha * av;
            acc += av * x[col[q]];
        }
    }
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, TPR);
    if (sub == 0) {
        if (OP == SpOp::AX) y[r] = acc;
        else if (OP == SpOp::RES) y[r] = b[r] - acc;
        else if (OP == SpOp::AXPBY) y[r] = acc + beta * y[r];
        else y[r] = ((-acc + b[r]) / D[r]) + x[r];
    }
}

int default_lanes(long long nnz, int n) {
    if (n <= 0) return 1;
    double d = (double)nnz / n;
    if (d <= 1.5) return 1;
    if (d <= 3) return 2;
    if (d <= 6) return 4;
    if (d <= 12) return 8;
    if (d <= 24) return 16;
    if (d <= 48) return 32;
    return 64;
}

template <typename T, typename V, SpOp OP>
static void launch1(Context* c, const SpmvArgs& a, int L) {
    const T* val = (const T*)a.val;
    const V* x = (const V*)a.x;
    V* y = (V*)a.y;
    const V* b = (const V*)a.b;
    const V* D = (const V*)a.D;
    V al = (V)a.alpha, be = (V)a.beta;
    // L < 0: CSR-stream kernels, L = -(1000 s + P + T): T lanes per row
    // reduction, P = 200 for the (val, col) pair stream, P = 400 for the
    // row-tile kernel (T = 4, A's tile index in a), s = 0 for per-XCD
    // row segments or s >= 1 for XCD-interleaved super-blocks of 2^(s-1) blocks
    int sb = 0;
    if (L < 0) {
        int s = -L / 1000;
        if (s > 13) throw Error(BCMGX_ERR_ARG, "stream super-block code out of range");
        sb = s > 0 ? 1 << (s - 1) : 0;
        L = -(-L % 1000);
    }
    if (L <= -400) {  // the tiled CSR-stream kernel (TILE_ROWS rows per workgroup), when A's row-tile index is there
        bool ok = a.tu && a.tucol && a.tlidx && (((uintptr_t)a.val) & (2 * sizeof(T) - 1)) == 0 &&
                  (((uintptr_t)a.tlidx) & 3) == 0;
        if (ok) {
            k_spmv_tiled<T, V, OP><<<grid_for(a.n, TILE_ROWS), WG, 0, c->stream>>>(
                a.n, a.nnz, a.row, a.col, val, a.tu, a.tucol, a.tlidx, x, y, b, D, al, be, a.stop, sb);
            BCMGX_LAUNCHED();
            return;
        }
        L = -200 - WG / TILE_ROWS;  // else the paired-load stream kernel with the same lanes per row
    }
    if (L <= -200) {  // CSR-stream variant streaming (val, col) pairs, TPR = -L - 200 lanes per row
        bool aligned = (((uintptr_t)a.val) & (2 * sizeof(T) - 1)) == 0 && (((uintptr_t)a.col) & 7) == 0;
        if (!aligned) L += 200;  // the single-entry stream kernel serves unaligned arrays
    }
    if (L <= -200) {
#define BCMGX_S2(TT)                                                                                           \
    case TT:                                                                                                   \
        k_spmv_stream2<T, V, TT, OP><<<grid_for(a.n, WG / TT), WG, 0, c->stream>>>(a.n, a.nnz, a.row, a.col, val, x, \
                                                                                  y, b, D, al, be, a.stop, sb); \
        break;
        switch (-L - 200) {
            BCMGX_S2(1) BCMGX_S2(2) BCMGX_S2(4) BCMGX_S2(8) BCMGX_S2(16)
        default: throw Error(BCMGX_ERR_ARG, "stream lanes per row must be a power of two <= 16");
        }
#undef BCMGX_S2
        BCMGX_LAUNCHED();
        return;
    }
    if (L < 0) {  // CSR-stream variant, TPR = -L lanes per row
#define BCMGX_S(TT)                                                                                                 \
    case TT:                                                                                                        \
        k_spmv_stream<T, V, TT, OP><<<grid_for(a.n, WG / TT), WG, 0, c->stream>>>(a.n, a.row, a.col, val, x, y, b, D, \
                                                                                 al, be, a.stop, sb);               \
        break;
        switch (-L) {
            BCMGX_S(1) BCMGX_S(2) BCMGX_S(4) BCMGX_S(8) BCMGX_S(16)
        default: throw Error(BCMGX_ERR_ARG, "stream lanes per row must be a power of two <= 16");
        }
#undef BCMGX_S
        BCMGX_LAUNCHED();
        return;
    }
#define BCMGX_L(LL)                                                                                            \
    case LL:                                                                                                   \
        k_spmv1<T, V, LL, OP><<<grid_for(a.n, WG / LL), WG, 0, c->stream>>>(a.n, a.row, a.col, val, x, y, b, D, \
                                                                          al, be, a.stop);                   \
        break;
    switch (L) {
        BCMGX_L(1) BCMGX_L(2) BCMGX_L(4) BCMGX_L(8) BCMGX_L(16) BCMGX_L(32) BCMGX_L(64)
    default: throw Error(BCMGX_ERR_ARG, "lanes per row must be a power of two <= 64");
    }
#undef BCMGX_L
    BCMGX_LAUNCHED();
}

template <typename T, typename V, SpOp OP>
static void launchk(Context* c, const SpmvArgs& a) {
    const T* val = (const T*)a.val;
    const V* x = (const V*)a.x;
    V* y = (V*)a.y;
    const V* b = (const V*)a.b;
    const V* D = (const V*)a.D;
    V al = (V)a.alpha, be = (V)a.beta;
    // columns are processed in chunks of kc: a narrower chunk keeps the
    // gathered x rows of a stencil neighbourhood inside the L2
    int kc = a.kchunk > 0 && a.kchunk < a.k ? a.kchunk : a.k;
    // whole rows of x as 16-byte vectors per lane: CPL = 16 / sizeof(V) columns
    constexpr int CPL = 16 / (int)sizeof(V);
    bool aligned = (((uintptr_t)x | (uintptr_t)y | (uintptr_t)b | (uintptr_t)a.y64 | (uintptr_t)a.b64) & 15) == 0;
    if (a.b64 && !(kc == a.k && a.kchunk != -1 && aligned && spmmp_lanes(a.k, CPL) && OP == SpOp::RES))
        throw Error(BCMGX_ERR_STATE, "spmv: an FP64 right-hand side only on the pipelined k-column residual");
    if (a.kchunk == -1 && aligned && spmmp_lanes(a.k, CPL)) {
        // one launch over all rows (no persistent pipeline): the autotuner's other full-width choice
        int KP = spmmp_lanes(a.k, CPL);
        bool out64 = OP == SpOp::JACOBI && a.y64 != nullptr;
#define BCMGX_V(KK, MM)                                                                                               \
    case KK:                                                                                                          \
        if (out64)                                                                                                    \
            k_spmmv<T, V, KK, CPL, MM, OP, true><<<grid_for(a.n, WG / KK), WG, 0, c->stream>>>(                       \
                a.n, a.k, a.row, a.col, val, x, y, b, D, al, be, a.stop, a.y64);                                      \
        else                                                                                                          \
            k_spmmv<T, V, KK, CPL, MM, OP><<<grid_for(a.n, WG / KK), WG, 0, c->stream>>>(a.n, a.k, a.row, a.col, val, \
                                                                                        x, y, b, D, al, be, a.stop);  \
        break;
        switch (KP) {
            BCMGX_V(2, 8) BCMGX_V(4, 8) BCMGX_V(8, 4) BCMGX_V(16, 2) BCMGX_V(32, 1)
        default: throw Error(BCMGX_ERR_ARG, "at most 64 right-hand sides");
        }
#undef BCMGX_V
        BCMGX_LAUNCHED();
        return;
    }
    if (kc == a.k && aligned && spmmp_lanes(a.k, CPL)) {
        SpmmExtra ex;
        ex.b64 = a.b64;
        bool nt = a.kchunk == -2;
        if (OP == SpOp::JACOBI && a.y64) {
            ex.y64 = a.y64;
            if (nt) dispatch_spmmp<T, V, OP, 1, true>(c, a, ex, 0);
            else dispatch_spmmp<T, V, OP, 1>(c, a, ex, 0);
        } else {
            if (nt) dispatch_spmmp<T, V, OP, 0, true>(c, a, ex, 0);
            else dispatch_spmmp<T, V, OP, 0>(c, a, ex, 0);
        }
        return;
    }
    for (int c0 = 0; c0 < a.k; c0 += kc) {
        int w = a.k - c0 < kc ? a.k - c0 : kc;
        int KP = 2;
        while (KP < w) KP <<= 1;
#define BCMGX_K(KK)                                                                                              \
    case KK:                                                                                                     \
        k_spmmk<T, V, KK, OP><<<grid_for(a.n, WG / KK), WG, 0, c->stream>>>(a.n, a.k, c0, w, a.row, a.col, val, x, y, \
                                                                            b, D, al, be, a.stop);               \
        break;
        switch (KP) {
            BCMGX_K(2) BCMGX_K(4) BCMGX_K(8) BCMGX_K(16) BCMGX_K(32) BCMGX_K(64)
        default: throw Error(BCMGX_ERR_ARG, "at most 64 right-hand sides");
        }
#undef BCMGX_K
        BCMGX_LAUNCHED();
    }
}

template <typename T, typename V>
void spmv(Context* c, SpOp op, const SpmvArgs& a) {
    if (a.n <= 0) return;
    if (a.y64) {
        // JACOBI into y (V) and widened into y64: the vector SpMM writes y64
        // directly; every other shape sweeps into y and converts
        if (op != SpOp::JACOBI || !std::is_same<V, float>::value || !a.y)
            throw Error(BCMGX_ERR_STATE, "spmv: FP64 output only for an FP32 Jacobi sweep");
        constexpr int CPL = 16 / (int)sizeof(V);
        int kc = a.kchunk > 0 && a.kchunk < a.k ? a.kchunk : a.k;
        bool aligned = (((uintptr_t)a.x | (uintptr_t)a.y | (uintptr_t)a.b | (uintptr_t)a.y64) & 15) == 0;
        if (!(a.k > 1 && kc == a.k && a.k % CPL == 0 && a.k >= 2 * CPL && aligned)) {
            SpmvArgs b = a;
            b.y64 = nullptr;
            spmv<T, V>(c, op, b);
            convert_stop<V, double>(c, (const V*)a.y, a.y64, (long long)a.n * a.k, a.stop);
            return;
        }
    }
    if (a.x0) {
        // y = A x and x0 = y / D (a first sweep from zero on y): in one pass through
        // the pipelined SpMM where it applies, else the product and then the sweep
        if (op != SpOp::AX || !a.D || !a.y) throw Error(BCMGX_ERR_STATE, "spmv: x0 output only for y = A x with D");
        constexpr int CPL = 16 / (int)sizeof(V);
        bool aligned = (((uintptr_t)a.x | (uintptr_t)a.y | (uintptr_t)a.x0) & 15) == 0;
        bool full = a.kchunk == 0 || a.kchunk >= a.k;
        if (a.k > 1 && full && aligned && spmmp_lanes(a.k, CPL) && a.row && (a.nnz == 0 || (a.col && a.val && a.x))) {
            SpmmExtra ex;
            ex.x0 = a.x0;
            SpmvArgs b = a;
            b.kchunk = 0;
            dispatch_spmmp<T, V, SpOp::AX, 3>(c, b, ex, 0);
        } else {
            SpmvArgs b = a;
            b.x0 = nullptr;
            spmv<T, V>(c, op, b);
            jacobi_from_zero<V>(c, a.n, a.k, (const V*)a.y, (const V*)a.D, (V*)a.x0, a.stop);
        }
        return;
    }
    // operands the kernels dereference: a null one is a library bug, reported instead of launched
    if (!a.row || (a.nnz > 0 && (!a.col || !a.val || !a.x)) || !a.y || a.k < 1 ||
        (op == SpOp::RES && !a.b && !a.b64) || (op == SpOp::JACOBI && (!a.b || !a.D || !a.x)))
        throw Error(BCMGX_ERR_STATE, "spmv: null operand");
    if (a.b64 && (a.k == 1 || op != SpOp::RES || !std::is_same<V, float>::value))
        throw Error(BCMGX_ERR_STATE, "spmv: an FP64 right-hand side only on the pipelined k-column residual");
    if (a.k == 1) {
        int L = a.lanes != 0 ? a.lanes : default_lanes(a.nnz, a.n);  // < 0: CSR-stream kernel
        switch (op) {
        case SpOp::AX: launch1<T, V, SpOp::AX>(c, a, L); break;
        case SpOp::RES: launch1<T, V, SpOp::RES>(c, a, L); break;
        case SpOp::AXPBY: launch1<T, V, SpOp::AXPBY>(c, a, L); break;
        case SpOp::JACOBI: launch1<T, V, SpOp::JACOBI>(c, a, L); break;
        }
    } else {
        switch (op) {
        case SpOp::AX: launchk<T, V, SpOp::AX>(c, a); break;
        case SpOp::RES: launchk<T, V, SpOp::RES>(c, a); break;
        case SpOp::AXPBY: launchk<T, V, SpOp::AXPBY>(c, a); break;
        case SpOp::JACOBI: launchk<T, V, SpOp::JACOBI>(c, a); break;
        }
    }
}
template void spmv<double, double>(Context*, SpOp, const SpmvArgs&);
template void spmv<float, float>(Context*, SpOp, const SpmvArgs&);

// ---------------------------------------------------------------- vectors ---
template <typename V>
__global__ void k_jacobi0(long long nk, int k, const V* __restrict__ f, const V* __restrict__ D, V* __restrict__ x,
                          const int* __restrict__ stop) {
    if (stop && *stop) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk) return;
    // ((-(A 0) + f) / D) + 0 == f / D
    x[t] = f[t] / D[t / k];
}
// 32-bit indexing (n k < 2^31); k == 1 needs no division
template <typename V, bool K1>
__global__ void k_jacobi0_32(int nk, int k, const V* __restrict__ f, const V* __restrict__ D, V* __restrict__ x,
                             const int* __restrict__ stop) {
    if (stop && *stop) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk) return;
    x[t] = f[t] / D[K1 ? t : t / k];
}
template <typename V>
void jacobi_from_zero(Context* c, int n, int k, const V* f, const V* D, V* x, const int* stop) {
    long long nk = (long long)n * k;
    if (!nk) return;
    if (nk < (1ll << 31)) {
        if (k == 1) k_jacobi0_32<V, true><<<grid_for(nk, WG), WG, 0, c->stream>>>((int)nk, k, f, D, x, stop);
        else k_jacobi0_32<V, false><<<grid_for(nk, WG), WG, 0, c->stream>>>((int)nk, k, f, D, x, stop);
    } else {
        k_jacobi0<V><<<grid_for(nk, WG), WG, 0, c->stream>>>(nk, k, f, D, x, stop);
    }
    BCMGX_LAUNCHED();
}
template void jacobi_from_zero<double>(Context*, int, int, const double*, const double*, double*, const int*);
template void jacobi_from_zero<float>(Context*, int, int, const float*, const float*, float*, const int*);

__global__ void k_jacobi0_f64in(long long nk, int k, const double* __restrict__ r, const float* __restrict__ D,
                                float* __restrict__ F, float* __restrict__ X, const int* __restrict__ stop) {
    if (stop && *stop) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk) return;
    float f = (float)r[t];
    F[t] = f;
    X[t] = f / D[t / k];
}
void jacobi_from_zero_f64in(Context* c, int n, int k, const double* r, const float* D, float* F, float* X, const int* stop) {
    long long nk = (long long)n * k;
    if (!nk) return;
    k_jacobi0_f64in<<<grid_for(nk, WG), WG, 0, c->stream>>>(nk, k, r, D, F, X, stop);
    BCMGX_LAUNCHED();
}

// FP32 f widened to FP64 F and the first sweep from zero X = F / D in FP64
// (convert_stop<float, double> then jacobi_from_zero<double>, in one pass)
__global__ void k_jacobi0_f32in(long long nk, int k, const float* __restrict__ f, const double* __restrict__ D,
                                double* __restrict__ F, double* __restrict__ X, const int* __restrict__ stop) {
    if (stop && *stop) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk) return;
    double v = (double)f[t];
    F[t] = v;
    X[t] = v / D[t / k];
}
void jacobi_from_zero_f32in(Context* c, int n, int k, const float* f, const double* D, double* F, double* X, const int* stop) {
    long long nk = (long long)n * k;
    if (!nk) return;
    k_jacobi0_f32in<<<grid_for(nk, WG), WG, 0, c->stream>>>(nk, k, f, D, F, X, stop);
    BCMGX_LAUNCHED();
}

template <typename V>
__global__ void k_prolong(long long nk, int k, const int* __restrict__ agg, const V* __restrict__ pval,
                          const V* __restrict__ xc, V* __restrict__ x, const int* __restrict__ stop) {
    if (stop && *stop) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk) return;
    long long i = t / k;
    int cc = (int)(t - i * k);
    V tt = pval[i] * xc[(long long)agg[i] * k + cc];
    x[t] = tt + x[t];
}
// the same with CPL consecutive columns per thread (16-byte accesses), 32-bit indexing
template <typename V, int CPL>
__global__ void k_prolong_v(int nkv, int kv, const int* __restrict__ agg, const V* __restrict__ pval,
                            const V* __restrict__ xc, V* __restrict__ x, const int* __restrict__ stop) {
    typedef V VV __attribute__((ext_vector_type(CPL)));
    if (stop && *stop) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nkv) return;
    int i = t / kv;
    int cv = t - i * kv;
    VV xv = *((const VV*)xc + ((long long)agg[i] * kv + cv));
    VV* xo = (VV*)x + t;
    *xo = pval[i] * xv + *xo;
}
template <typename V>
void prolong_add(Context* c, int n, int k, const int* agg, const V* pval, const V* xc, V* x, const int* stop) {
    long long nk = (long long)n * k;
    if (!nk) return;
    constexpr int CPL = 16 / (int)sizeof(V);
    bool aligned = (((uintptr_t)xc | (uintptr_t)x) & 15) == 0;
    if (k % CPL == 0 && aligned && nk / CPL < (1ll << 31)) {
        int nkv = (int)(nk / CPL);
        k_prolong_v<V, CPL><<<grid_for(nkv, WG), WG, 0, c->stream>>>(nkv, k / CPL, agg, pval, xc, x, stop);
    } else {
        k_prolong<V><<<grid_for(nk, WG), WG, 0, c->stream>>>(nk, k, agg, pval, xc, x, stop);
    }
    BCMGX_LAUNCHED();
}
template void prolong_add<double>(Context*, int, int, const int*, const double*, const double*, double*, const int*);
template void prolong_add<float>(Context*, int, int, const int*, const float*, const float*, float*, const int*);

// the epilogue of the A P post-smoothing sweep on t = (A P) e already in y:
// y = ((-t + b) / D) + (p_i e[agg_i] + xpre)
template <typename V>
__global__ void k_post_galerkin(long long nk, int k, const int* __restrict__ agg, const V* __restrict__ pval,
                                const V* __restrict__ e, const V* __restrict__ xpre, const V* __restrict__ b,
                                const V* __restrict__ D, V* __restrict__ y, const int* __restrict__ stop) {
    if (stop && *stop) return;
    long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nk) return;
    long long i = o / k;
    int c = (int)(o - i * k);
    V z1 = pval[i] * e[(long long)agg[i] * k + c] + xpre[o];
    y[o] = ((-y[o] + b[o]) / D[i]) + z1;
}

template <typename V>
void galerkin_post_sweep(Context* c, const SpmvArgs& a, const int* agg, const V* pval, const V* xpre) {
    if (a.n <= 0) return;
    if (!a.row || (a.nnz > 0 && (!a.col || !a.val)) || !a.x || !a.y || !a.b || !a.D || !agg || !pval || !xpre ||
        a.y == xpre || a.y == a.b)
        throw Error(BCMGX_ERR_STATE, "galerkin_post_sweep: bad operand");
    constexpr int CPL = 16 / (int)sizeof(V);
    bool aligned = (((uintptr_t)a.x | (uintptr_t)a.y | (uintptr_t)a.b | (uintptr_t)xpre) & 15) == 0;
    bool full = a.kchunk <= 0 || a.kchunk >= a.k;
    if (a.k > 1 && full && aligned && spmmp_lanes(a.k, CPL)) {
        SpmmExtra ex;
        ex.agg = agg;
        ex.pv = pval;
        ex.xpre = xpre;
        SpmvArgs b = a;
        b.kchunk = a.kchunk == -2 || a.kchunk <= -16 ? a.kchunk : 0;
        if (b.kchunk == -2) dispatch_spmmp<V, V, SpOp::AX, 4, true>(c, b, ex, 0);
        else dispatch_spmmp<V, V, SpOp::AX, 4>(c, b, ex, 0);
        return;
    }
    SpmvArgs b = a;
    b.b = nullptr;
    b.D = nullptr;
    spmv<V, V>(c, SpOp::AX, b);  // t = (A P) e into y
    long long nk = (long long)a.n * a.k;
    k_post_galerkin<V><<<grid_for(nk, WG), WG, 0, c->stream>>>(nk, a.k, agg, pval, (const V*)a.x, xpre, (const V*)a.b,
                                                              (const V*)a.D, (V*)a.y, a.stop);
    BCMGX_LAUNCHED();
}
template void galerkin_post_sweep<double>(Context*, const SpmvArgs&, const int*, const double*, const double*);
template void galerkin_post_sweep<float>(Context*, const SpmvArgs&, const int*, const float*, const float*);

template <typename A, typename B>
__global__ void k_convert_stop(const A* __restrict__ s, B* __restrict__ d, long long n, const int* __restrict__ stop) {
    if (stop && *stop) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = (B)s[t];
}
template <typename A, typename B>
void convert_stop(Context* c, const A* s, B* d, long long n, const int* stop) {
    if (!n) return;
    k_convert_stop<A, B><<<grid_for(n, WG), WG, 0, c->stream>>>(s, d, n, stop);
    BCMGX_LAUNCHED();
}
template void convert_stop<float, double>(Context*, const float*, double*, long long, const int*);
template void convert_stop<double, float>(Context*, const double*, float*, long long, const int*);
template void convert_stop<double, double>(Context*, const double*, double*, long long, const int*);

}  // namespace bcmgx
