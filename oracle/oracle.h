/*
 * oracle.h -- C ABI of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++17 (+ optional OpenMP)
 * restatement of BootCMatchGX's single-GPU solve path, used as the checker
 * for the HIP product and as the timed CPU baseline.  Nothing in the product
 * (bootcmatchgx_amd/) may link or call it.
 *
 * Pinned by the reference's fixtures and by mathematics (tests/test_oracle.py):
 * generators (MATLAB .mtx fixtures, closed forms), MTX reader (bcsstk09),
 * spmv/spgemm/transpose (dense products), matching weights (2x2 quadratic-form
 * identity), suitor (independent greedy matching + invariants), make_p (known
 * answer), hierarchy (dense P^T A P, coarse-size truncation), whole solves of
 * every preconditioner (the reference's solution fixtures at 1e-5), AFSAI
 * (triangularity, unit diagonal of G A G^T).
 *
 * Parity unpinned: the reference binary's own FCG iteration counts and level
 * shapes (none published; counts are checked against this restatement only);
 * its atomic-order singleton numbering (canonicalised here to ascending
 * index); its lock-based Suitor reaching the greedy matching; nsparse/cuBLAS
 * round-off; the last ulp of log() between glibc and the GPU; AFSAI argmax
 * ties (resolved here to the smallest column).  The diffusion generator is
 * build-defined and has no reference counterpart.
 */
#ifndef BCMGX_ORACLE_H
#define BCMGX_ORACLE_H

#include "../include/bcmgx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ocsr ocsr;
typedef struct oprec oprec;

ocsr* or_csr_new(int n, int m, int nnz, const int* row, const int* col, const double* val);
void or_csr_free(ocsr* A);
void or_csr_dims(const ocsr* A, int* n, int* m, int* nnz);
void or_csr_copy(const ocsr* A, int* row, int* col, double* val);

ocsr* or_gen_laplacian3d(int points, int nx, int ny, int nz, int P, int Q, int R);
ocsr* or_gen_diffusion7p(int nx, int ny, int nz, int kind, double p1, double p2, unsigned long long seed);
ocsr* or_read_mtx(const char* path);

void or_spmv(const ocsr* A, const double* x, double* y, double alpha, double beta, int op);
ocsr* or_spgemm(const ocsr* A, const ocsr* B);
ocsr* or_transpose(const ocsr* A);

/* matching weights W (log |AH|, symmetrised) laid out on A's off-diagonal pattern */
int or_matching_weights(const ocsr* A, const double* w, int* Wrow, int* Wcol, double* Wval);
void or_suitor(const ocsr* A, const double* w, int* M);
void or_greedy_matching(const ocsr* A, const double* w, int* M);
int or_make_p(int n, const int* M, const double* w, int* agg, double* pval);

oprec* or_prec_setup(const ocsr* A, const bcmgx_params* p);
void or_prec_free(oprec* pr);
int or_prec_levels(const oprec* pr);
const ocsr* or_prec_level_A(const oprec* pr, int level);
int or_prec_level_P(const oprec* pr, int level, int* agg, double* pval);
void or_prec_hierarchy_stats(const oprec* pr, double* op_cmplx, double* op_wcmplx, double* avg_cratio, int* maxcoarse);
const ocsr* or_prec_afsai_G(const oprec* pr);
void or_prec_apply(const oprec* pr, const double* r, double* u);
/* bootstrap: number of hierarchies and the last estimated convergence ratio
 * (1 when none was estimated); hierarchy h's levels, matrices, prolongators */
int or_prec_num_hrc(const oprec* pr, double* estimated_ratio);
int or_prec_hrc_levels(const oprec* pr, int h);
const ocsr* or_prec_hrc_level_A(const oprec* pr, int h, int level);
int or_prec_hrc_level_P(const oprec* pr, int h, int level, int* agg, double* pval);
/* one cycle of hierarchy h alone (u in/out, the initial guess) */
void or_prec_apply_hrc(const oprec* pr, int h, const double* r, double* u);
/* a one-hierarchy BCMG preconditioner whose matching starts from the smooth vector w */
oprec* or_prec_setup_w(const ocsr* A, const bcmgx_params* p, const double* w);
/* the bootstrap's random start vectors: x_i = 2 u_i - 1, u_i from splitmix64 */
void or_boot_random(int n, unsigned long long seed, double* x);

/* FCG (FCG.cu:27-175).  x in/out.  Returns retv. */
int or_fcg(const ocsr* A, const oprec* pr, const bcmgx_params* p, const double* b, double* x,
           int* niter, double* exit_res, double* del0);
// cg_hs (CG_HS.cu:24-172); same outputs as or_fcg (niter counted as the reference does)
int or_cghs(const ocsr* A, const oprec* pr, const bcmgx_params* p, const double* b, double* x,
            int* niter, double* exit_res, double* del0);
// cgsstep (CGs.cu:43-236), s = the sstep setting; same outputs
int or_cgs(const ocsr* A, const oprec* pr, const bcmgx_params* p, const double* b, double* x,
           int* niter, double* exit_res, double* del0);
// pipelinedcgsstep (pipelinedCGs.cu:56-300); same outputs
int or_pipecgs(const ocsr* A, const oprec* pr, const bcmgx_params* p, const double* b, double* x,
               int* niter, double* exit_res, double* del0);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
