/*
 * CPU oracle for the Schwarz-(F)GMRES solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain restatement of the
 * reference algorithms (izbassarn7/schwarz-gmres-dashboard, C++17) used as
 * the checker for the HIP product path.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  The product library
 * (libsgamd.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  All arithmetic
 * is compiled with -ffp-contract=off so that sums and products round exactly
 * as the reference's g++ build does on x86-64 (no FMA contraction).
 *
 * Pinning status (what checks each function besides retyping it):
 *   or_spmv, or_dot            closed forms (Poisson A*1 row sums, sum 1..n),
 *                              brute-force dense products, golden vectors from
 *                              the reference built under oracle/_ref.
 *   or_ilu0_*                  golden factors from the reference ILU0Precond
 *                              (bit-exact), (L*U)_ij == A_ij on the pattern,
 *                              tridiagonal exact-LU, Poisson d-recurrence.
 *   or_overlap                 brute-force BFS over dense adjacency in tests.
 *   or_precond_* (RAS/ASM)     nparts=1 == ILU0, delta=0 == block-ILU0, golden
 *                              apply outputs from reference ILU0 on reference
 *                              extract_submatrix blocks (oracle/_ref).
 *   coarse space (sparse R0)   dense R0*A*R0^T brute force in numpy.
 *   or_lu_*                    golden LU/pivots from the reference CoarseSolver
 *                              (oracle/_ref), scipy.linalg.lu_factor.
 *   or_gmres / or_fgmres       golden iteration counts and residual histories
 *                              from the reference gmres.cpp / fgmres.cpp
 *                              (oracle/_ref), known answers 201 (Poisson 100^2)
 *                              and 301 (150^2) of the compiled reference
 *                              (tol 1e-10, b = A*1), identity and SPD exactness.
 *   or_metrics_json            golden strings from reference Metrics::to_json.
 *   or_ilut_new                golden apply outputs, l_nnz / u_nnz and GMRES
 *                              iteration counts from the reference ILUTPrecond
 *                              (oracle/_ref), bit for bit; drop_tol = 0 with
 *                              unbounded fill on a tridiagonal == exact LU.
 *   or_bicgstab                golden iteration counts, residual histories and
 *                              solutions from the reference bicgstab.h
 *                              (oracle/_ref), bit for bit; identity -> 1 step.
 *   or_lanczos / or_tridiag_eigs / or_condition / or_spectral
 *                              golden Lanczos coefficients, Ritz extremes and
 *                              power-iteration radii from the reference
 *                              analysis headers (oracle/_ref); tridiagonal
 *                              eigenvalues against numpy.linalg.eigvalsh;
 *                              Poisson 1D/2D closed-form spectra.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generators (src/mesh/poisson.h:10-74) ------------------------------- */
int or_poisson2d(int nx, int ny, int* row_ptr, int* col_idx, double* vals);
int or_poisson3d(int nx, int ny, int nz, int* row_ptr, int* col_idx, double* vals);

/* ---- CSR / BLAS-1 (src/core/sparse_matrix.h:30-43, src/core/vector.h) ---- */
void or_spmv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y);
double or_dot(int n, const double* a, const double* b);
int or_extract(int n, const int* rp, const int* ci, const double* v, int m, const int* idx,
               int* sub_rp, int* sub_ci, double* sub_v);

/* ---- ILU(0) (src/precond/ilu0.h:13-79) ----------------------------------- */
void or_ilu0_factor(int n, const int* rp, const int* ci, const double* a, double* lu, int* diag);
void or_ilu0_solve(int n, const int* rp, const int* ci, const double* lu, const int* diag,
                   const double* x, double* y);

/* ---- decomposition given a partition (src/decomp/domain_decomp.h:61-93) --- */
int or_overlap(int n, const int* rp, const int* ci, const int* part, int nparts, int overlap,
               int* sub_ptr, int* sub_idx);

/* ---- preconditioners (src/precond/{jacobi,ilu0,asm,ras,two_level_schwarz}.h) ----------------------------------- */
enum { OR_PC_NONE = 0, OR_PC_JACOBI = 1, OR_PC_ILU0 = 2, OR_PC_ASM = 3, OR_PC_RAS = 4,
       OR_PC_TWOLEVEL_ASM = 5, OR_PC_TWOLEVEL_RAS = 6 };
/* precision: 0 = FP64 (reference), 1 = FP32 factors and FP32 subdomain/coarse
 * apply (the mixed mode of the product, restated on the CPU). */
void* or_precond_new(int kind, int n, const int* rp, const int* ci, const double* v,
                     int nparts, int overlap, const int* part, int precision);
void or_precond_apply(void* h, const double* x, double* y);
/* ILUTPrecond(drop_tol, fill_factor) (src/precond/ilut.h:21-193); l_nnz / u_nnz
 * as its l_nnz() / u_nnz(); freed with or_precond_free. */
void* or_ilut_new(int n, const int* rp, const int* ci, const double* v, double drop_tol,
                  int fill_factor, int precision, long long* l_nnz, long long* u_nnz);
/* Longest per-row index list: the reference's elimination loop is defined only
 * while it stays within the 2 * fill_factor * (nnz / n) entries it reserves. */
long long or_ilut_max_list(void* h);
int or_precond_coarse(void* h, double* A0, double* LU, int* piv);
void or_precond_free(void* h);

/* ---- dense coarse LU (src/precond/coarse_solve.h:15-67) ------------------ */
int or_lu_factor(int n, double* A, int* piv);
void or_lu_solve(int n, const double* LU, const int* piv, const double* b, double* x);

/* ---- Krylov drivers (src/core/gmres.cpp:34-160, src/core/fgmres.cpp:28-172) */
/* method 0 = left-preconditioned GMRES, 1 = FGMRES.  Returns iterations. */
int or_solve(int method, int n, const int* rp, const int* ci, const double* v,
             const double* b, double* x, void* pc, int restart, int max_iter, double tol,
             double* residual_norm, int* converged, double* hist, int hist_cap, int* hist_len);

/* ---- BiCGSTAB (src/core/bicgstab.h:57-204): tol relative to ||r0||.  Returns iterations. */
int or_bicgstab(int n, const int* rp, const int* ci, const double* v, const double* b, double* x,
                void* pc, int max_iter, double tol, double* residual_norm, int* converged,
                double* hist, int hist_cap, int* hist_len);

/* ---- analysis (src/analysis/condition_estimator.h, spectral_analysis.h) --- */
int or_lanczos(int n, const int* rp, const int* ci, const double* v, void* pc, int k,
               unsigned seed, double* alpha, double* beta);
void or_tridiag_eigs(int n, const double* d, const double* e, double* out);
int or_condition(int n, const int* rp, const int* ci, const double* v, void* pc, int k,
                 unsigned seed, double* out, int* valid);
int or_spectral(int n, const int* rp, const int* ci, const double* v, void* pc, int max_iter,
                double tol, unsigned seed, double* rho);

/* ---- JSONL record (src/io/metrics_logger.h:28-51) ------------------------ */
int or_metrics_json(const char* matrix, const char* method, int n, int nnz, int ranks,
                    int threads, int gpus, int overlap, int restart, int iterations,
                    double tol, double residual_norm, double time_setup_s,
                    double time_solve_s, double speedup, double efficiency,
                    char* out, int cap);

#ifdef __cplusplus
}
#endif
