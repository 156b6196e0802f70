/*
 * bcmgx.h -- C ABI of the MI355X-native BootCMatchGX single-GPU solve path.
 *
 * Every entry point returns an int status (BCMGX_OK == 0, negative on error,
 * bcmgx_error_string() names it); nothing calls exit().  Handles are opaque.
 * All state lives in a context (one HIP device, one HIP stream), so distinct
 * contexts are independent.  Matrices are CSR with int32 row pointers and
 * column indices and FP64 values; columns must be sorted inside each row and
 * every row must hold its diagonal entry, as the reference requires
 * (BCMGX/src/preconditioner/bcmg/matching.cu:155, makeAH).
 *
 * Citations are relative to the reference tree (BootCMatchGX v1.1).
 */
#ifndef BCMGX_H
#define BCMGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    BCMGX_OK = 0,
    BCMGX_ERR_ARG = -1,        /* invalid argument / shape */
    BCMGX_ERR_HIP = -2,        /* HIP runtime error */
    BCMGX_ERR_ALLOC = -3,      /* device allocation failed */
    BCMGX_ERR_IO = -4,         /* file could not be opened or parsed */
    BCMGX_ERR_PARAM = -5,      /* invalid settings value (INVALID enum) */
    BCMGX_ERR_UNSUPPORTED = -6,/* option outside the single-GPU path */
    BCMGX_ERR_NUMERIC = -7,    /* non-finite value / breakdown */
    BCMGX_ERR_STATE = -8       /* handle used in the wrong state */
};

/* memory-space flag for pointers crossing the boundary */
enum { BCMGX_MEM_HOST = 0, BCMGX_MEM_DEVICE = 1 };

/* ---- settings ---------------------------------------------------------
 * Mirrors `struct params` (src/config/Params.h:238-291) field by field, with
 * the same defaults, plus extension keys after `rec_res_int`.
 * Enumerations use the reference's enumerator order.
 */
enum { BCMGX_SOLVER_CGHS = 0, BCMGX_SOLVER_FCG = 1, BCMGX_SOLVER_CGS = 2,
       BCMGX_SOLVER_PIPELINED_CGS = 3, BCMGX_SOLVER_CGS_CUBLAS = 4, BCMGX_SOLVER_INVALID = 5 };
enum { BCMGX_BOOT_MULTIPLICATIVE = 0, BCMGX_BOOT_SYMMETRIZED_MULTIPLICATIVE = 1,
       BCMGX_BOOT_ADDITIVE = 2, BCMGX_BOOT_INVALID = 3 };
enum { BCMGX_MATCH_SUITOR = 0, BCMGX_MATCH_INVALID = 1 };
enum { BCMGX_CYCLE_V = 0, BCMGX_CYCLE_H = 1, BCMGX_CYCLE_W = 2, BCMGX_CYCLE_VARIABLE_V = 3,
       BCMGX_CYCLE_INVALID = 4 };
enum { BCMGX_COARSE_L1_JACOBI = 0, BCMGX_COARSE_INVALID = 1 };
enum { BCMGX_RELAX_L1_JACOBI = 0, BCMGX_RELAX_INVALID = 1 };
enum { BCMGX_PREC_NONE = 0, BCMGX_PREC_L1_JACOBI = 1, BCMGX_PREC_BCMG = 2, BCMGX_PREC_AFSAI = 3,
       BCMGX_PREC_INVALID = 4 };
/* extension: precision of the AMG hierarchy values and of the V-cycle */
enum { BCMGX_PRECISION_FP64 = 0, BCMGX_PRECISION_FP32 = 1 };

typedef struct bcmgx_params {
    int mem_alloc_size;     /* GLOB_MEM_ALLOC_SIZE (unused by this build) */
    int error;
    int solver;             /* key "solver", default CGHS */
    int itnlim;             /* default 2000 */
    double rtol;            /* default 1e-6 */
    int dispnorm;           /* default 1 */
    int preconditioner;     /* key "preconditioner", default NONE */
    int bootstrap_type;     /* default MULTIPLICATIVE */
    int max_hrc;            /* default 1 */
    double conv_ratio;      /* default 0.8 */
    int matchtype;          /* default SUITOR */
    int aggrsweeps;         /* default 2 */
    int aggr_type;          /* default 0 */
    int max_levels;         /* default 39 */
    int cycle_type;         /* default V_CYCLE */
    int coarse_solver;      /* default L1_JACOBI */
    int relax_type;         /* default L1_JACOBI */
    int relaxnumber_coarse; /* default 20 */
    int prerelax_sweeps;    /* default 4 */
    int postrelax_sweeps;   /* default 4 */
    int l1jacsweeps;        /* default 4 */
    int sstep;              /* default 1 */
    int stop_criterion;     /* default 1 */
    int ru_res;             /* default 1 */
    int rec_res_int;        /* default 0 */
    /* ---- extension keys (no counterpart in the reference) ---- */
    int max_coarse_size;    /* "max_coarse_size": 0 = reference formula (int)(40*cbrt(N)) */
    int precision;          /* "precision": FP64 (default) or FP32 hierarchy + V-cycle */
    int use_graph;          /* "use_graph": capture the FCG iteration in a hipGraph (default 1) */
    int autotune;           /* "autotune": pick SpMV/SpMM lanes-per-row by timing (default 1) */
    int check_every;        /* "check_every": FCG iterations between host convergence polls (default 8) */
    int fcg_z32;            /* "fcg_z32": with an FP32 hierarchy, FCG keeps the cycle's result z in FP32 and
                               widens it exactly inside its kernels (fewer bytes; the dot products are blocked
                               differently, so iterates may differ in the last bits; default 1) */
    int post_galerkin;      /* "post_galerkin": with an FP32 hierarchy, the first post-smoothing sweep of each
                               level takes its residual as r_pre - (A P) e, from the pre-smoothing residual and
                               the product A P kept from setup, in place of f - A (x + P e) (the same value in
                               exact arithmetic, from the sparser A P; default 1) */
} bcmgx_params;

/* Defaults of src/config/Params.h:238-291. */
int bcmgx_params_default(bcmgx_params* p);
/* Params::initFromPropertiesFile (src/config/Params.cu:326-404): `key = value`
 * lines, '#' comments, unknown keys ignored, enum strings as in the reference.
 * A line without '=' is an error (BCMGX_ERR_IO), as Params.cu:306-309. */
int bcmgx_params_from_properties(const char* path, bcmgx_params* p);
/* Set one key (same parsing as a properties line). Unknown key: ignored, returns BCMGX_OK. */
int bcmgx_params_set(bcmgx_params* p, const char* key, const char* value);
/* Enum <-> string helpers (Params.cu:16-245). `kind` is one of "solver",
 * "preconditioner", "cycle_type", "bootstrap_type", "matchtype", "coarse_solver", "relax_type". */
int bcmgx_enum_from_string(const char* kind, const char* str);
const char* bcmgx_enum_to_string(const char* kind, int value);

/* ---- context ------------------------------------------------------------ */
typedef struct bcmgx_context_s* bcmgx_context;
typedef struct bcmgx_matrix_s* bcmgx_matrix;
typedef struct bcmgx_prec_s* bcmgx_prec;

/* Replaces Handles::init + cudaSetDevice (src/utility/handles.cu, driverSolve.cu:200-205,259). */
int bcmgx_context_create(int device, bcmgx_context* ctx);
int bcmgx_context_destroy(bcmgx_context ctx);
int bcmgx_context_synchronize(bcmgx_context ctx);
const char* bcmgx_error_string(int status);
/* Last error message recorded on this context (or globally if ctx == NULL). */
const char* bcmgx_last_error(bcmgx_context ctx);

/* ---- matrices -----------------------------------------------------------
 * Replaces CSRm::init + CSRm::copyToDevice (src/datastruct/CSR.h:297,347).
 * The library copies the arrays into its own device memory.
 */
int bcmgx_csr_create(bcmgx_context ctx, int n, int m, int nnz, const int* row, const int* col,
                     const double* val, int mem, bcmgx_matrix* A);
/* Destroying a matrix also releases the preconditioners set up on it (their
 * level 0 is this matrix): their later calls return BCMGX_ERR_STATE, and
 * bcmgx_prec_finalize still frees their handles.  Destroying the context
 * releases every matrix and preconditioner on it the same way. */
int bcmgx_csr_destroy(bcmgx_matrix A);
int bcmgx_csr_info(bcmgx_matrix A, int* n, int* m, int* nnz);
/* copy CSR arrays back to host buffers (row: n+1, col/val: nnz) */
int bcmgx_csr_copy_to_host(bcmgx_matrix A, int* row, int* col, double* val);
/* Overwrite the nnz values of A (same pattern).  No counterpart in the
 * reference, whose matrices are immutable after copyToDevice (CSR.h:347);
 * used by the autograd shim when the values it differentiates change.  A
 * preconditioner set up on A keeps its hierarchy: its level 0 is A itself and
 * sees the new values, and the level-0 divisors are recomputed from them (the
 * l1 row sums of BCMG, the l1 diagonal of one-level l1-Jacobi), as is the
 * level-0 product A P of an FP32 hierarchy with post_galerkin; the coarse
 * levels, the prolongators and an AFSAI factor are not rebuilt. */
int bcmgx_csr_update_values(bcmgx_matrix A, const double* val, int mem);
/* Build A's row-tile index now (it is built on first need otherwise) and
 * report the seconds it took.  No counterpart in the reference (its SpMV reads
 * CSR only).  For every tile of 64 rows the index lists the distinct columns
 * its entries touch and gives each entry the 16-bit position of its column in
 * that list; one-column products with A then stage those x entries in LDS once
 * per tile.  It depends on the pattern only, so it is kept across value
 * updates and preconditioner setups on A; it costs 2 bytes per entry plus 4
 * per listed column of device memory. */
int bcmgx_csr_analyse(bcmgx_matrix A, double* seconds);

/* Host-side problem generators: src/generator/laplacian.cu:119-334 (7-point),
 * :336-777 (27-point), block ordering internal_index (:14-25).  Output arrays
 * are malloc'ed by the library and released with bcmgx_free. */
int bcmgx_gen_laplacian3d(int points, int nx, int ny, int nz, int P, int Q, int R,
                          int* n, int* nnz, int** row, int** col, double** val);
/* Extension (no reference counterpart): variable-coefficient 7-point diffusion.
 * kind 0: face conductivity exp(sigma*g), g ~ N(0,1); kind 1: uniform[lo, hi].
 * RNG std::mt19937_64(seed). */
int bcmgx_gen_diffusion7p(int nx, int ny, int nz, int kind, double p1, double p2, uint64_t seed,
                          int* n, int* nnz, int** row, int** col, double** val);
/* readMTXDouble + check_and_fix_order (src/datastruct/CSR.cu:1973-2152, :1895-1918). */
int bcmgx_read_mtx(const char* path, int* n, int* m, int* nnz, int** row, int** col, double** val);
void bcmgx_free(void* p);

/* ---- device vectors ------------------------------------------------------
 * Device buffers on the context's device (Vector::init<T>(n, true, true) and
 * Vector::copyTo / copyToDevice, src/datastruct/vector.h) so callers without
 * their own allocator can keep inputs resident in HBM.  kind of bcmgx_memcpy:
 * 0 host->device, 1 device->host, 2 device->device; it completes before returning. */
int bcmgx_device_alloc(bcmgx_context ctx, size_t bytes, void** ptr);
int bcmgx_device_free(bcmgx_context ctx, void* ptr);
int bcmgx_memcpy(bcmgx_context ctx, void* dst, const void* src, size_t bytes, int kind);

/* ---- preconditioner -----------------------------------------------------
 * prec_setup (src/preconditioner/prec_setup.h:55): the preconditioner type
 * comes from params->preconditioner (the reference's cgsprec::ptype).
 * BCMG with max_hrc > 1 is the bootstrap (bootstrap.cu:11-46): after the
 * first hierarchy, while fewer than max_hrc exist and the composition of those
 * built so far has an estimated convergence ratio above conv_ratio, one more
 * hierarchy is built from the slowest error of that composition (15
 * applications to A x = 0 from a random x, whose last A-norm ratio is the
 * estimate).  bootstrap_type composes them in prec_apply: MULTIPLICATIVE
 * (bcmg.cu:65-85), SYMMETRIZED_MULTIPLICATIVE (forward then back, a symmetric
 * preconditioner) or ADDITIVE (corrections summed from one residual).  The
 * reference stops at one hierarchy (bootstrap.cu:37-38) and applies
 * MULTIPLICATIVE only (bcmg.cu:86-89).  A matrix with no rows, or not square,
 * is refused with BCMGX_ERR_ARG before any device work.
 */
int bcmgx_prec_setup(bcmgx_context ctx, bcmgx_matrix A, const bcmgx_params* p, bcmgx_prec* prec);
/* prec_apply (src/preconditioner/prec_apply.h:32): u <- B r.  As in the
 * reference, u is also the initial guess of the V-cycle; FCG zero-fills it.
 * nrhs columns, column-major with leading dimension ld. */
int bcmgx_prec_apply(bcmgx_prec prec, const double* r, double* u, int nrhs, int ld, int mem);
/* prec_finalize (src/preconditioner/prec_finalize.h:29). */
int bcmgx_prec_finalize(bcmgx_prec prec);

typedef struct bcmgx_prec_info {
    int ptype;
    int num_levels;
    double op_cmplx;        /* AMG::Hierarchy::finalize_cmplx (AMG.cu:143-152) */
    double op_wcmplx;       /* finalize_wcmplx (AMG.cu:154-163) */
    double avg_cratio;      /* adaptiveCoarsening (matchingAggregation.cu:437-441) */
    int max_coarse_size;    /* limit used (reference formula or extension key) */
    double setup_seconds;   /* device time of the setup (hipEvents) */
    long long g_nnz;        /* AFSAI: nnz of G */
    int num_hrc;            /* BCMG: hierarchies built by the bootstrap (boot::n_hrc, AMG.h:91) */
    double estimated_ratio; /* BCMG: boot::estimated_ratio (AMG.h:92), 1 until estimated */
} bcmgx_prec_info;
int bcmgx_prec_get_info(bcmgx_prec prec, bcmgx_prec_info* info);
/* level sizes: n and nnz of A_level */
int bcmgx_prec_level_size(bcmgx_prec prec, int level, int* n, int* nnz);
/* copy A_level to host (row: n+1, col/val: nnz) */
int bcmgx_prec_level_matrix(bcmgx_prec prec, int level, int* row, int* col, double* val);
/* prolongator of level -> level+1: one entry per fine row: coarse index and value */
int bcmgx_prec_level_prolongator(bcmgx_prec prec, int level, int* agg, double* pval);
/* The same for hierarchy h of a bootstrap (boot::H_array[h], AMG.h:88-93);
 * h = 0 is the hierarchy the three calls above read. */
int bcmgx_prec_hrc_levels(bcmgx_prec prec, int h, int* num_levels);
int bcmgx_prec_hrc_level_size(bcmgx_prec prec, int h, int level, int* n, int* nnz);
int bcmgx_prec_hrc_level_matrix(bcmgx_prec prec, int h, int level, int* row, int* col, double* val);
int bcmgx_prec_hrc_level_prolongator(bcmgx_prec prec, int h, int level, int* agg, double* pval);
/* AFSAI factor G (lower triangular, row-sorted) */
int bcmgx_prec_afsai_factor(bcmgx_prec prec, int* n, int* nnz, int* row, int* col, double* val);

/* ---- solver ---------------------------------------------------------------
 * solve (src/solver/solve.h:33, dispatch solve.cu:19-40) on p->solver:
 *   FCG           solve_fcg (src/solver/fcg/FCG.cu:27-215)
 *   CGHS          solve_cg_hs (src/solver/cghs/CG_HS.cu:24-192)
 *   CGS           solve_cgs, s-step CG with s = p->sstep (src/solver/cgs/CGs.cu:43-236)
 *   CGS_CUBLAS    solve_cgs_cublas, the same s-step method (src/solver/cgs_cublas/CGscublas.cu)
 *   PIPELINED_CGS solve_pipelined_cgs (src/solver/pipelined_cgs/pipelinedCGs.cu:56-329)
 * x is in/out (initial guess, then solution), as the reference's x0.
 * nrhs >= 1 columns (column-major, leading dimension ld) share one
 * preconditioner.  FCG and CGHS solve them together, each column with its own
 * scalars and stopping test; the s-step forms solve them one after another.
 * Stopping: FCG ||r_k|| <= rtol ||r_0|| (FCG.cu:93,156); the others
 * ||r_k|| < rtol delta0 with delta0 = ||r_0|| for stop_criterion 1, else 1
 * (CG_HS.cu:74-79,141).  A matrix with no rows, or not square, is refused
 * with BCMGX_ERR_ARG.
 */
typedef struct bcmgx_solver_out {
    int retv;               /* 0 converged, -1 itnlim reached (FCG.cu:160-162), > 0 the s-step
                               s x s system was not positive definite (scalarWorkMO); worst column */
    int niter;              /* the reference's counting (FCG.cu:164, CG_HS.cu:168); max over columns */
    double exit_res;        /* ||r|| at exit (FCG.cu:165); max over columns */
    double rel_res;         /* exit_res / delta0; max over columns */
    double del0;            /* delta0 of column 0 (||r_0||, or 1 for an absolute test) */
    double solve_seconds;   /* device time of the solve */
    int full_n;
    int local_n;
} bcmgx_solver_out;
int bcmgx_solve(bcmgx_context ctx, bcmgx_matrix A, bcmgx_prec prec, const bcmgx_params* p,
                const double* b, double* x, int nrhs, int ld, int mem, bcmgx_solver_out* out);
/* per-column results of the last bcmgx_solve on this preconditioner */
int bcmgx_solve_columns(bcmgx_prec prec, int nrhs, int* niter, double* exit_res, double* rel_res);

/* info file of src/solver/SolverOut.cu:9-97 with AMG::Hierarchy::printInfo (AMG.cu:167-180) */
int bcmgx_dump_info(const char* path, const bcmgx_params* p, bcmgx_prec prec, const bcmgx_solver_out* out);

/* ---- kernel-level entry points (parity tests and micro-benchmarks) --------
 * y = alpha*A*x + beta*y, op as _CSR_vector_mul_mini_warp<OP> (CSR.cu:452-492):
 * op 1: y = A x; op 2: y = y - A x; op 0: y = alpha A x + beta y.
 * lanes: lanes per row (0 = autotuned choice; -T: CSR-stream kernel with T lanes
 * per row reduction; -200-T: the same streaming (val, col) pairs; -404: the
 * row-tile kernel, 64 rows per workgroup with the x entries of each tile's
 * distinct columns staged once (builds the matrix's row-tile index on first
 * use; it is kept with the matrix); -1000 s - code: rows dealt to the XCDs in
 * super-blocks of 2^(s-1) workgroups).  Every lanes choice of the stream
 * family (-T, -200-T, -404) reduces a row in the same order for a given T. */
int bcmgx_spmv(bcmgx_context ctx, bcmgx_matrix A, const double* x, double* y, double alpha,
               double beta, int op, int lanes, int mem);
/* The same on k right-hand sides at once (x, y column-major with leading
 * dimension ld >= n; the device works on k interleaved columns).  No
 * counterpart in the reference, which multiplies one vector at a time: it is
 * the multi-vector kernel of the batched solve.  kchunk: columns per launch
 * (0 = all k through the pipelined persistent kernel, -1 = all k through the
 * one-pass kernel, -2 = pipelined with streamed matrix loads, -16-s (s in
 * 0..12) = pipelined with rows dealt to the XCDs in super-groups of 2^s row
 * groups instead of one contiguous segment per XCD).  Each column is the sequential row sum of that column
 * (the oracle's order) bit for bit. */
int bcmgx_spmm(bcmgx_context ctx, bcmgx_matrix A, const double* x, double* y, int k, int ld, double alpha,
               double beta, int op, int kchunk, int mem);
/* The same product without allocating: x and y are device blocks, and the two
 * operands are staged in the library's fixed per-device scratch array of
 * 64 MiB ((rows + columns) x nrhs x 8 bytes, each part rounded up to 256
 * bytes, must fit; a larger call returns BCMGX_ERR_ARG with the sizes in the
 * message).  No counterpart in the reference: it serves callers that cannot
 * allocate device memory, such as code under HIP graph capture and tight
 * serving loops.  It calls no hipMalloc/hipFree and does not autotune (k = 1
 * uses the lanes per row bcmgx_spmv/bcmgx_spmm tuned for A, if they ran, else
 * the nnz-per-row heuristic; kchunk as in bcmgx_spmm).  It returns once the
 * work is queued on the context's stream; bcmgx_context_synchronize waits for
 * it.  Every call on one device shares that device's array: calls that
 * overlap on one device (the streams of two contexts, two host threads) race. */
int bcmgx_spmm_noalloc(bcmgx_context ctx, bcmgx_matrix A, const double* x, double* y, int k, int ld, double alpha,
                       double beta, int op, int kchunk);
/* C = A*B (replaces nsparse spgemm_csrseg_kernel_hash, kernel_spgemm_hash_d.cu:1196-1247),
 * sorted columns; each C_ij summed over k in ascending order. */
int bcmgx_spgemm(bcmgx_context ctx, bcmgx_matrix A, bcmgx_matrix B, bcmgx_matrix* C);
/* C = A^T (CSRm::Transpose_local, CSR.cu:1752-1822) */
int bcmgx_transpose(bcmgx_context ctx, bcmgx_matrix A, bcmgx_matrix* T);
/* Suitor matching of suitor() (matching.cu:253-288): M[i] = partner or -1. */
int bcmgx_matching(bcmgx_context ctx, bcmgx_matrix A, const double* w, int* M);
/* makeP_GPU (matchingPairAggregation.cu:109-170), canonical singleton order:
 * agg[i] coarse index, pval[i] value; returns the coarse size in *nc. */
int bcmgx_make_prolongator(bcmgx_context ctx, int n, const int* M, const double* w, int* agg,
                           double* pval, int* nc);
/* Micro-benchmark of one SpMV/SpMM launch (driverSpMV, src/example/driverSpMV.cu):
 * y = A x on k interleaved columns, FP64 (precision 0) or FP32 matrix and
 * vectors (precision 1); `reps` timed launches on the context stream after one
 * warm-up, bracketed by HIP events.  *ms = average milliseconds per launch.
 * `lanes`: k == 1: lanes per row (0 = autotuned, -T = CSR-stream kernel, -200-T = paired-load CSR-stream);
 * k > 1: columns per launch (0 = all k, pipelined kernel; -1 = all k, one-pass kernel;
 * -2 = all k, pipelined kernel with streamed matrix loads; -16-s = all k, pipelined
 * kernel with XCD-interleaved super-groups of 2^s row groups). */
int bcmgx_time_spmv(bcmgx_context ctx, bcmgx_matrix A, int precision, int k, int lanes, int reps, double* ms);
/* Micro-benchmark of the smoother: one l1-Jacobi sweep y = x + (b - A x) / D
 * (_jacobi_it_full<0>, l1jacobi.cu:344-379) on k interleaved columns, D = the
 * l1 diagonal of A, FP64 (precision 0) or FP32 matrix and vectors (precision 1),
 * with the launch shape the V-cycle picks for this operator and width; `reps`
 * timed launches after one warm-up, bracketed by HIP events. */
int bcmgx_time_smoother(bcmgx_context ctx, bcmgx_matrix A, int precision, int k, int reps, double* ms);
/* Micro-benchmark of the SpGEMM C = A*B (the Galerkin products of the setup,
 * in place of nsparse's spgemm_csrseg_kernel_hash, kernel_spgemm_hash_d.cu:1196-1247):
 * `reps` whole products (row counting, binning, numeric kernels, compaction)
 * after one warm-up, bracketed by HIP events; *nnz_c = nonzeros of C. */
int bcmgx_time_spgemm(bcmgx_context ctx, bcmgx_matrix A, bcmgx_matrix B, int reps, double* ms, int* nnz_c);
/* Sustained device-to-device copy bandwidth (GB/s) over `bytes` bytes. */
int bcmgx_stream_copy_gbs(bcmgx_context ctx, size_t bytes, int reps, double* gbs);

/* ---- adjoint support for the differentiable solve -------------------------
 * Extension (no reference counterpart): g[k] = -sum_c lambda[row(k), c] x[col(k), c]
 * for every stored entry k of A (the sparsity-restricted outer product that
 * is dL/dA_ij for x = A^{-1} b). lambda, x: n x nrhs column-major. */
int bcmgx_sparse_outer_grad(bcmgx_context ctx, bcmgx_matrix A, const double* lambda,
                            const double* x, int nrhs, double* g, int mem);

#ifdef __cplusplus
}
#endif
#endif
