/*
 * sg_capi.h -- C ABI of the MI355X Schwarz-(F)GMRES solver (libsgamd.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an sg_status; the
 * message of the last failure on a context is available from sg_last_error().
 * Device pointers ("d_" prefix) are HIP device allocations (from sg_malloc or
 * any other HIP allocation on the same device); host pointers are ordinary
 * memory.  A context owns one HIP stream; every call made through a context
 * and the objects created from it is ordered on that stream and is
 * single-threaded.  Separate contexts may be used from separate threads.
 *
 * Each declaration cites the reference interface it replaces
 * (izbassarn7/schwarz-gmres-dashboard, paths relative to its root).
 */
#ifndef SG_CAPI_H
#define SG_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SG_OK = 0,
    SG_E_ARG = 1,        /* invalid argument / shape                           */
    SG_E_HIP = 2,        /* HIP runtime failure                                */
    SG_E_SINGULAR = 3,   /* "Coarse matrix is singular" (coarse_solve.h:31-32) */
    SG_E_PARTITION = 4,  /* partitioning failed (domain_decomp.h:54-55)        */
    SG_E_IO = 5,
    SG_E_NOCONV = 6,     /* reserved: non-convergence is reported, not raised  */
    SG_E_STATE = 7       /* object used before setup / wrong order             */
} sg_status;

typedef struct sg_context sg_context;
typedef struct sg_csr sg_csr;
typedef struct sg_precond sg_precond;

/* ---- context, memory, errors ----------------------------------------------
 * Replaces gpu::CudaHandles (src/gpu/cuda_common.h:42-66), DeviceBuffer
 * (src/gpu/cuda_common.h:69-107) and the *_CHECK error macros (:13-40). */
sg_status sg_context_create(int device, sg_context** out);
sg_status sg_context_destroy(sg_context* ctx);
sg_status sg_context_synchronize(sg_context* ctx);
const char* sg_last_error(const sg_context* ctx); /* ctx may be NULL: last global error */
sg_status sg_malloc(sg_context* ctx, size_t bytes, void** d_ptr);
sg_status sg_free(sg_context* ctx, void* d_ptr);
sg_status sg_memcpy_h2d(sg_context* ctx, void* d_dst, const void* h_src, size_t bytes);
sg_status sg_memcpy_d2h(sg_context* ctx, void* h_dst, const void* d_src, size_t bytes);
sg_status sg_memcpy_d2d(sg_context* ctx, void* d_dst, const void* d_src, size_t bytes);
sg_status sg_memset_zero(sg_context* ctx, void* d_ptr, size_t bytes);

/* ---- CSR matrix -----------------------------------------------------------
 * Replaces CSRMatrix + DeviceCSRMatrix::upload (src/core/sparse_matrix.h:13-28,
 * src/gpu/cuda_sparse.h:15-51).  Copies the host arrays (0-based, int32
 * indices) to the device and keeps a host copy for the preconditioner setup.
 * Columns must be sorted within each row for ILU(0)-based preconditioners
 * (src/precond/ilu0.h:31-51 relies on it). */
sg_status sg_csr_create(sg_context* ctx, int nrows, int ncols, int nnz, const int* row_ptr,
                        const int* col_idx, const double* vals, sg_csr** out);
sg_status sg_csr_destroy(sg_csr* A);
sg_status sg_csr_info(const sg_csr* A, int* nrows, int* ncols, int* nnz);

/* y = alpha*A*x + beta*y (FP64).  Replaces CSRMatrix::spmv
 * (src/core/sparse_matrix.h:30-43) and DeviceCSRMatrix::spmv via cusparseSpMV
 * (src/gpu/cuda_sparse.h:53-72).  Each row is summed sequentially in CSR order,
 * so with alpha=1, beta=0 the result equals the reference's serial SpMV bit for
 * bit. */
sg_status sg_spmv(sg_csr* A, double alpha, const double* d_x, double beta, double* d_y);

/* ---- Matrix Market files (host only; no GPU needed) -------------------------
 * Replaces read_matrix_market / write_matrix_market (src/io/matrix_market.h:14-103).
 * flags = 0 reads as the reference does: '%' lines whose text contains
 * "symmetric" / "pattern" set those flags, symmetric files are mirrored,
 * entries are sorted by (row, col) and duplicates stay separate entries
 * (in file order).  Inputs on which the reference is undefined (unreadable
 * sizes or indices, indices out of range) fail with SG_E_IO.  The arrays are
 * allocated by the library (row_ptr[nrows+1], col_idx[nnz], vals[nnz]) and
 * released with sg_mm_free. */
typedef enum {
    SG_MM_SUM_DUPLICATES = 1, /* merge entries with the same (row, col) by summing  */
    SG_MM_STRICT = 2          /* header flags from the %%MatrixMarket banner only;
                                 skew-symmetric mirrored with negation; blank and
                                 comment lines skipped anywhere; malformed values,
                                 array / complex files and extra entries rejected   */
} sg_mm_flags;
sg_status sg_mm_read(const char* path, int flags, int* nrows, int* ncols, int* nnz,
                     int** row_ptr, int** col_idx, double** vals);
void sg_mm_free(void* p);
/* "general" banner, sizes, one "row col value" line per stored entry (1-based),
 * values in %g style with `digits` significant digits (6 = the reference's
 * stream default; 17 round-trips). */
sg_status sg_mm_write(const char* path, int nrows, int ncols, const int* row_ptr,
                      const int* col_idx, const double* vals, int digits);

/* ---- BLAS-1 (FP64, device vectors) ----------------------------------------
 * Replace schwarz::dot/norm2/axpy/scale (src/core/vector.h:11-48) and
 * gpu::cublas_dot/nrm2/axpy/scal (src/gpu/cuda_blas.h:10-39).  Reductions use a
 * fixed tree, so results are deterministic run to run. */
sg_status sg_dot(sg_context* ctx, int n, const double* d_a, const double* d_b, double* result);
sg_status sg_nrm2(sg_context* ctx, int n, const double* d_a, double* result);
sg_status sg_axpy(sg_context* ctx, int n, double alpha, const double* d_x, double* d_y);
sg_status sg_scal(sg_context* ctx, int n, double alpha, double* d_x);

/* ---- preconditioners --------------------------------------------------------
 * Replace the Preconditioner hierarchy (src/precond/preconditioner.h:7-20):
 * JacobiPrecond (jacobi.h:10-27), ILU0Precond (ilu0.h:11-93), ASMPrecond
 * (asm.h:15-71), RASPrecond (ras.h:16-82), TwoLevelSchwarzPrecond
 * (two_level_schwarz.h:21-75) and the device ASMGPUPrecond (src/gpu/asm_gpu.h:30-186). */
typedef enum {
    SG_PC_NONE = 0,
    SG_PC_JACOBI = 1,
    SG_PC_ILU0 = 2,
    SG_PC_ASM = 3,
    SG_PC_RAS = 4,
    SG_PC_TWOLEVEL_ASM = 5,
    SG_PC_TWOLEVEL_RAS = 6,
    SG_PC_LEARNED = 7, /* made by sg_precond_create_learned only */
    SG_PC_ILUT = 8     /* ILUTPrecond(drop_tol, fill_factor) (src/precond/ilut.h:21-193):
                          dual-threshold factorisation on the host, as the reference's;
                          forward / backward sweeps on the device (FP64 or FP32)      */
} sg_precond_kind;

typedef enum { SG_FP64 = 0, SG_FP32 = 1 } sg_precision;

typedef struct {
    int nparts;            /* subdomains (RASPrecond(nparts, overlap), ras.h:18-19)      */
    int overlap;           /* BFS overlap layers (domain_decomp.h:72-86)                  */
    int precision;         /* sg_precision of ILU factors, subdomain and coarse apply    */
    const int* partition;  /* host int[nrows] subdomain of each row, or NULL: built-in
                              partitioner (stands in for partition_metis,
                              domain_decomp.h:23-58)                                      */
    int ufactor;           /* allowed imbalance in 1/1000 (METIS_OPTION_UFACTOR); 0 = 30 */
    int coarse_solve;      /* two-level only: 0 = LU forward/back solve (coarse_solve.h
                              order), 1 = precomputed inverse (one dense mat-vec)         */
    int weighted_partition; /* built-in partitioner only: 1 weighs edge (i,j) by |a_ij| so
                              cuts follow weak couplings; 0 = unit weights as the
                              reference's METIS call (domain_decomp.h:48-52)              */
    double drop_tol;       /* ILUT: relative drop tolerance (entries below drop_tol times
                              the row's max |a_ij| are dropped)                           */
    int fill_factor;       /* ILUT: keep at most fill_factor * (nnz / n) entries in each
                              row of L and of U; drop_tol = fill_factor = 0 selects the
                              reference's defaults 1e-4 and 10                            */
} sg_precond_opts;

sg_status sg_precond_create(sg_context* ctx, sg_precond_kind kind, const sg_precond_opts* opts,
                            sg_precond** out);
/* Host decomposition (partition, overlap, extract), device ILU(0) factorisation
 * and coarse-matrix LU.  Replaces Preconditioner::setup. */
sg_status sg_precond_setup(sg_precond* M, const sg_csr* A);
/* d_y = M^{-1} d_x.  Replaces Preconditioner::apply (host pointers) with device
 * pointers on the context stream. */
sg_status sg_precond_apply(sg_precond* M, const double* d_x, double* d_y);
sg_status sg_precond_destroy(sg_precond* M);

/* Host callback preconditioner: fn(user, x, y, n) computes y = M^{-1} x.  With
 * device_pointers = 0 the library stages x and y through host memory, so any
 * host-side Preconditioner::apply can be plugged in (FlexApply,
 * src/core/fgmres.h:37). */
typedef void (*sg_apply_fn)(void* user, const double* x, double* y, int n);
sg_status sg_precond_create_callback(sg_context* ctx, int n, sg_apply_fn fn, void* user,
                                     int device_pointers, sg_precond** out);

/* Host-only graph partition of an n x n CSR matrix into nparts parts
 * (part[n] out).  Multilevel recursive bisection; deterministic.  Stands in
 * for partition_metis (src/decomp/domain_decomp.h:23-58): nparts <= 1 gives
 * all zeros, self loops are ignored.  ufactor: allowed imbalance in 1/1000
 * (0 = 30, METIS's default).  weighted = 1 weighs edges by |a_ij| (vals
 * required); 0 ignores vals (may be NULL).  Needs no GPU. */
sg_status sg_partition(int n, const int* row_ptr, const int* col_idx, const double* vals,
                       int nparts, int ufactor, int weighted, int* part);

/* Sizes of a set-up preconditioner, for traffic models: number of subdomains,
 * sum of overlapped subdomain sizes (N_ext), stored local nonzeros (nnz_ext),
 * forward and backward sweep levels summed over subdomains, and whether the
 * subdomain vector lives in LDS. */
typedef struct {
    int nsub;
    long long n_ext;
    long long nnz_ext;
    long long levels_fwd;
    long long levels_bwd;
    int max_levels;
    int z_in_lds;
    int nc;
    int max_ns;       /* rows of the largest (overlapped) subdomain            */
    int sweep_shape;  /* sweep kernel shape in use (see sg_context_set_autotune:
                         0..2 barrier per level with 256/512/1024 threads,
                         3..4 sync-free with 512/1024, 5..6 barrier with deeper
                         prefetch, 7..9 staged through LDS with 1024/512/256),
                         -1 before the first apply */
} sg_precond_info;
sg_status sg_precond_stats(const sg_precond* M, sg_precond_info* info);

/* Diagnostic with no counterpart in the reference: one apply d_y = M^{-1} d_x
 * with the clock of the staged subdomain sweep armed.  Every workgroup of that
 * sweep writes 12 values to rec (cap >= 12 * nsub entries), at its launch slot:
 * ticks of the device wall clock at its start, when the right-hand side is
 * gathered, when both sweeps are done and at its end; subdomain, hardware id
 * (HIP's __smid: XCC, shader engine and CU ids packed), overlapped rows,
 * forward levels, backward levels, staged chunks, owned rows, 0.  *nrec
 * counts the workgroups that recorded (0: the sweep in use is not the staged
 * kernel; empty or guarded-off subdomains leave zeros).  *rate_khz is the
 * wall clock's rate. */
sg_status sg_precond_sweep_clock(sg_precond* M, const double* d_x, double* d_y,
                                 long long* rec, int cap, int* nrec, double* rate_khz);

/* Introspection for parity tests and tooling (all host outputs). */
sg_status sg_precond_partition(const sg_precond* M, int* part /* int[nrows] */);
/* Total overlapped size sum_s |subdomain_dofs[s]|, then the sets (sorted global
 * ids, concatenated; ptr has nparts+1 entries).  Pass NULL arrays to query. */
sg_status sg_precond_subdomains(const sg_precond* M, int* total, int* ptr, int* idx);
/* ILU(0) factor values of subdomain s in that subdomain's local CSR order
 * (ILU0Precond::lu_vals, ilu0.h:83); s = 0 for the global ILU(0). */
sg_status sg_precond_ilu_values(const sg_precond* M, int s, int* nnz, double* vals);
/* ILUTPrecond::l_nnz() / u_nnz() (src/precond/ilut.h:175-176). */
sg_status sg_precond_ilut_nnz(const sg_precond* M, long long* l_nnz, long long* u_nnz);
/* Coarse matrix A0 (row-major nc x nc) and its LU factors / pivots. */
sg_status sg_precond_coarse(const sg_precond* M, int* nc, double* A0, double* LU, int* piv);

/* ---- Krylov solve ------------------------------------------------------------
 * Replaces gmres (src/core/gmres.h:28-33), fgmres / fgmres_flex
 * (src/core/fgmres.h:42-55) and gpu::gmres_gpu / solve_gpu
 * (src/gpu/gmres_gpu.h:25-72). */
typedef enum {
    SG_GMRES_LEFT = 0,
    SG_FGMRES = 1,
    /* Right-preconditioned BiCGSTAB, replaces bicgstab (src/core/bicgstab.h:57-204).
     * tol_abs is then the reference's RELATIVE tolerance (stop when ||s|| or
     * ||r|| < tol * ||r0||, bicgstab.h:140,180); max_iter bounds iterations;
     * restart and ortho are ignored; rtol must be 0.  hist receives ||r|| after
     * every full iteration (and ||s|| at a half-step exit).  Batches, graphs
     * (SG_SOLVE_GRAPH / SG_SOLVE_NO_GRAPH) and lookahead work as for GMRES. */
    SG_BICGSTAB = 2
} sg_method;
typedef enum { SG_ORTHO_MGS = 0, SG_ORTHO_CGS2 = 1 } sg_ortho;

typedef struct {
    int method;      /* sg_method                                                     */
    int restart;     /* m (GMRESParams::restart, default 30)                          */
    int max_iter;    /* GMRESParams::max_iter (cycle bound; total steps checked per
                        cycle, so iterations may exceed it by up to m-1)              */
    double tol_abs;  /* GMRESParams::tol: absolute test on the residual estimate
                        (||M^-1 r|| for left GMRES, ||r|| for FGMRES)                  */
    double rtol;     /* > 0: stop only when the FP64 true residual ||b-Ax||/||b|| <= rtol,
                        verified with an FP64 SpMV (tol_abs is then ignored)           */
    int ortho;       /* sg_ortho: MGS (reference order) or CGS2                       */
    int lookahead;   /* Arnoldi steps the host may run ahead of the device (0 = 2)     */
    int flags;       /* sg_solve_flags, or 0                                          */
} sg_solve_params;

typedef enum {
    /* sg_solve: capture the restart cycle (restart residual, m Arnoldi steps,
     * back substitution, x update) once as a hipGraph and replay it every
     * cycle; steps after convergence are launched and return at once. */
    SG_SOLVE_GRAPH = 1,
    /* sg_solve_batch: launch every kernel directly instead of the graph. */
    SG_SOLVE_NO_GRAPH = 2
} sg_solve_flags;

typedef struct {
    int iterations;           /* GMRESResult::iterations                              */
    int converged;            /* GMRESResult::converged                               */
    double residual_norm;     /* GMRESResult::residual_norm (last estimate)           */
    double rel_true_residual; /* ||b - A x|| / ||b|| in FP64 at exit                  */
    double time_solve_s;      /* GMRESResult::time_solve_s (wall, includes the sync)  */
    int cycles;               /* restart cycles run (0 for BiCGSTAB)                  */
    int precond_applies;
    int hist_len;             /* residual-history entries produced (GMRES: iterations;
                                 BiCGSTAB: full iterations, plus one at a half-step
                                 exit); hist holds the first min(hist_len, hist_cap)  */
} sg_solve_result;

/* x (device, in/out) is the initial guess and the solution.  hist (host, may be
 * NULL) receives the residual estimate after every Arnoldi step
 * (GMRESResult::residual_history), up to hist_cap entries. */
sg_status sg_solve(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                   const sg_solve_params* params, sg_solve_result* result, double* hist,
                   int hist_cap);

/* No-allocation variant of sg_solve.  No counterpart in the reference: it is
 * for callers that cannot allocate device or pinned memory per call (tight
 * serving loops, harnesses that forbid it).  Same arguments, arithmetic and
 * results as sg_solve on a single system (iterations, x and history equal bit
 * for bit), with these limits: GMRES or FGMRES on a single-system matrix; a
 * device preconditioner (not a host callback or the learned stage); no graph
 * (SG_SOLVE_GRAPH is rejected) and no lookahead (every step of a cycle is
 * launched, finished systems' kernels return at once); kernel shapes are not
 * timed (the shapes in place are used; every shape computes identical values).
 * Its whole workspace -- state, reduction tables, Hessenberg and Gram storage,
 * history, partial sums and the Krylov basis, about 8 n (2m + 6) bytes for
 * FGMRES with a preconditioner, 8 n (m + 6) otherwise, plus 8 (max_iter + m)
 * -- lives in one static device array of 64 MiB on each GPU the library runs
 * on, held until the process ends; a larger system fails with SG_E_ARG before
 * anything is launched.  Calls that overlap on one GPU (from any contexts or
 * threads) race on that array; each call starts from a state it initialises
 * itself, so the contents left by an earlier call do not matter. */
sg_status sg_solve_noalloc(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                           const sg_solve_params* params, sg_solve_result* result, double* hist,
                           int hist_cap);

/* Verification hook, no counterpart in the reference: the Krylov basis of the
 * last restart cycle that ran Arnoldi steps in the last single-system
 * GMRES/FGMRES sg_solve on this context (the workspace is kept by the context
 * between solves; the record is replaced by the next such solve).  n, m
 * (restart), k (steps of that cycle) and kv (basis vectors formed: k + 1, or k
 * when the cycle ended at its m-th step or broke down) are always written; V
 * (host, n x kv, column i at V + i*n), Z (host, n x k, FGMRES with a
 * preconditioner only) and Hbar (host, (m+1) x m column-major, the Hessenberg
 * columns as Gram-Schmidt produced them, before the Givens rotations) may be
 * NULL.  With them, FGMRES satisfies A Z_j = V_{j+1} Hbar_j and left GMRES
 * M^-1 A V_j = V_{j+1} Hbar_j for j < kv (the Arnoldi loops of
 * src/core/fgmres.cpp:77-107 and src/core/gmres.cpp:84-128). */
sg_status sg_solve_last_basis(sg_context* ctx, int* n, int* m, int* k, int* kv, double* V,
                              double* Z, double* Hbar);

/* ---- batched solve (many independent systems) -----------------------------
 * No counterpart in the reference (it solves one system per run,
 * src/main.cpp:238-243).  The systems are stored as one block-diagonal CSR:
 * rows [sys_rows[i], sys_rows[i+1]) of system i (sys_rows[0] = 0,
 * sys_rows[nsys] = n) may only reference columns in that range, with global
 * column indices.  Every SpMV, preconditioner sweep and reduction covers all
 * systems in one launch; a preconditioner built on this matrix must not let a
 * subdomain span two systems (give each system its own parts in
 * sg_precond_opts.partition) for per-system early exit. */
sg_status sg_csr_create_batch(sg_context* ctx, int nsys, const int* sys_rows, int nnz,
                              const int* row_ptr, const int* col_idx, const double* vals,
                              sg_csr** out);
sg_status sg_csr_systems(const sg_csr* A, int* nsys);
/* Solves all systems of a batch matrix together: each system runs its own
 * GMRES/FGMRES state machine (own Hessenberg, Givens, tolerance, restart and
 * convergence decisions) with the same arithmetic as sg_solve on that system
 * alone, so its x and iteration count equal a separate sg_solve bit for bit.
 * result[nsys]; hist (may be NULL) is nsys rows of hist_cap entries.  The
 * restart cycle runs as a captured hipGraph unless SG_SOLVE_NO_GRAPH. */
sg_status sg_solve_batch(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                         const sg_solve_params* params, sg_solve_result* result, double* hist,
                         int hist_cap);

/* ---- learned stage: residual-to-error network -------------------------------
 * No counterpart in the reference.  A fixed 4-level U-Net (channels
 * 16-32-64-64, 3x3 convolutions, ReLU, 2x2 average pooling, nearest
 * upsampling with skip concatenation, 1x1 head; 302705 parameters packed as
 * documented in oracle/unet_oracle.cpp) with bfloat16 weights and activations
 * and fp32 accumulation on the matrix cores.  Inputs [B][2][H][W] fp32
 * (residual / rms(residual), log kappa), output [B][H][W] fp32; H and W must
 * be multiples of 8.  Loss kind 0 (default): mean over the batch of
 * ||out - t||^2 / ||t||^2.  Loss kind 1 (no target): with s = input channel 0
 * and A the 2D five-point finite-volume diffusion operator of
 * kappa = exp(input channel 1) (harmonic faces, Dirichlet boundary), the mean
 * of 1 - (s . A out)^2 / (|s|^2 |A out|^2): the share of |s|^2 that the best
 * multiple of out fails to remove. */
typedef struct sg_unet sg_unet;
int sg_unet_nparams(void);
sg_status sg_unet_create(sg_context* ctx, sg_unet** out);
sg_status sg_unet_destroy(sg_unet* net);
/* host fp32 parameter vectors (the fp32 master copy; Adam state is reset) */
sg_status sg_unet_set_params(sg_unet* net, const float* params);
sg_status sg_unet_get_params(const sg_unet* net, float* params);
/* device buffers */
sg_status sg_unet_forward(sg_unet* net, int B, int H, int W, const float* d_in, float* d_out);
/* loss kind 0: relative squared error against a target; 1: residual loss
 * 1 - (s.Au)^2 / (|s|^2 |Au|^2) with u the output, A rebuilt from the log kappa
 * channel (no target); 2: kind 1 plus 0.01 ln(|Au|^2 / |s|^2)^2, which holds
 * the output's scale */
sg_status sg_unet_set_loss(sg_unet* net, int kind);
/* forward + backward; loss and grad (host, nparams floats) may be NULL;
 * d_target may be NULL with loss kind 1 */
sg_status sg_unet_loss_grad(sg_unet* net, int B, int H, int W, const float* d_in,
                            const float* d_target, double* loss, float* grad);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) with the last gradient; with a
 * weight decay wd > 0 the step also shrinks every parameter by lr * wd * p
 * (decoupled, AdamW) */
sg_status sg_unet_adam_step(sg_unet* net, float lr);
/* decoupled weight decay of the Adam step (default 0) */
sg_status sg_unet_set_weight_decay(sg_unet* net, float wd);

/* ---- learned FGMRES stage ------------------------------------------------------
 * No counterpart in the reference: a flexible preconditioner for FGMRES
 * (fgmres_flex's FlexApply role, src/core/fgmres.h:37-55) on a 2D grid.  Each
 * application, on the device:  z_b = base(x);  s = x - A z_b;
 * e = rms(s) * net(s / rms(s), log kappa);  alpha = (s . A e) / (A e . A e);
 * z = z_b + alpha e when that removes at least min_gain of ||s||^2, else
 * z = z_b (the base two-level Schwarz step is the fallback).  Because z
 * changes from step to step it is valid with SG_FGMRES only.  base and net
 * are not owned and must outlive the stage; sg_precond_setup(stage, A) needs
 * base already set up on A. */
typedef struct {
    int nx, ny;      /* grid: row y*nx + x; nx and ny multiples of 8            */
    double min_gain; /* 0 -> 0.01                                               */
    int every;       /* run the network on every every-th application (0 -> 1) */
    int coarsen;     /* 0 or 1: the network sees the nx x ny grid; 2 or 4: it sees
                        the c x c block averages of s and log kappa (nx/c and ny/c
                        multiples of 8) and its output is interpolated bilinearly
                        back to the grid before the A-scaling above              */
} sg_learned_opts;
sg_status sg_precond_create_learned(sg_context* ctx, sg_precond* base, sg_unet* net,
                                    const double* log_kappa /* host, nx*ny */,
                                    const sg_learned_opts* opts, sg_precond** out);
/* applications so far, and how many of them kept the network's correction */
sg_status sg_precond_learned_stats(const sg_precond* M, long long* calls, long long* accepted);

/* ---- convergence analytics ----------------------------------------------------
 * Replace src/analysis/condition_estimator.h, spectral_analysis.h and
 * convergence_analysis.h (called by src/main.cpp:211-220,255-261 under
 * --analyze). */
typedef struct {
    double lambda_min, lambda_max, kappa; /* extreme Ritz values, lambda_max / lambda_min
                                             (1e30 when |lambda_min| <= 1e-14)        */
    int lanczos_steps;
    int valid;                            /* lambda_min > 0 (meaningful for SPD)      */
} sg_condition_estimate;
/* ConditionEstimate estimate_condition(A, M, k, seed) (condition_estimator.h:94-191):
 * k Lanczos steps on M^{-1}A (A when M is NULL) from the reference's LCG start
 * vector, stopping early on an invariant subspace; the Ritz extremes of the
 * tridiagonal.  alpha[k] / beta[k-1] (host, may be NULL) receive its
 * coefficients.  All zero when min(k, n) < 2 or a coefficient is not finite. */
sg_status sg_estimate_condition(sg_csr* A, sg_precond* M, int k, unsigned seed,
                                sg_condition_estimate* out, double* alpha, double* beta);
typedef struct {
    double spectral_radius; /* |mu| of the step before the stopping one, as the reference */
    int convergence_ok;     /* spectral_radius < 1                                        */
    int power_iters;
} sg_spectral_result;
/* estimate_spectral_radius(A, M, max_iter, tol, seed) (spectral_analysis.h:33-95):
 * power iteration on I - M^{-1}A with Rayleigh quotients, stopping when two
 * successive quotients agree to tol * (1 + |mu|) or ||E v|| < 1e-14.
 * d_eigvec (device, n, may be NULL) receives the last iterate. */
sg_status sg_estimate_spectral_radius(sg_csr* A, sg_precond* M, int max_iter, double tol,
                                      unsigned seed, sg_spectral_result* out, double* d_eigvec);
typedef struct {
    double initial_residual, final_residual, asymptotic_rate;
    int stagnation_start; /* -1 = none */
    int converged;
    int iterations;       /* history length */
} sg_convergence;
/* ConvergenceMetrics::compute (convergence_analysis.h:29-66); rates (may be
 * NULL) receives the n per-step rates.  Host only. */
sg_status sg_convergence_analyze(const double* hist, int n, int converged, sg_convergence* out,
                                 double* rates);
/* ConvergenceMetrics::write_jsonl (convergence_analysis.h:69-93): appends one
 * "type":"convergence" line to path.  Host only. */
sg_status sg_convergence_write_jsonl(const char* path, const char* matrix, const char* precond,
                                     int nparts, int overlap, const double* hist, int n,
                                     int converged);

/* ---- kernel timing (HIP events on the context stream) ------------------------
 * With profiling on, every launch of the listed kernel families is bracketed
 * by HIP events; totals are collected at each synchronisation. */
typedef enum {
    SG_K_SPMV = 0,     /* CSR SpMV (Arnoldi and residual)          */
    SG_K_PRECOND = 1,  /* whole preconditioner apply                */
    SG_K_TRSV = 2,     /* fused gather + ILU(0) sweeps + scatter    */
    SG_K_ORTHO = 3,    /* Gram-Schmidt + norm + Givens              */
    SG_K_UPDATE = 4,   /* restart residual, basis scaling, x update */
    SG_K_GRAM_DOT = 5,    /* CGS2 pass 1 alone (V^T w, V^T v_j); also in ORTHO */
    SG_K_GRAM_UPDATE = 6, /* CGS2 pass 2 alone (w -= V h, ||w||^2); also in ORTHO */
    SG_K_COUNT = 7
} sg_kernel_family;
sg_status sg_context_set_profiling(sg_context* ctx, int enable);
/* Kernel-shape selection.  mode 1 (default): the first SpMV on a matrix and
 * the first apply of a preconditioner time each shape and keep the fastest
 * (SpMV: CSR with 256, 512 or 1024 rows per workgroup, or the SELL-64 copy
 * (slices of 64 rows stored column by column, built when its padding is below
 * half the nonzeros) with 32-bit columns or, when every column is within 32767
 * of its row, 16-bit offsets; subdomain sweeps: 256, 512 or 1024 threads with a
 * barrier per level and row data two levels ahead, 512 / 1024 threads
 * sync-free with per-row flags in LDS, 256 / 512 threads with a barrier per
 * level and row data five levels ahead, or 1024 / 512 / 256 threads with the
 * factor data staged through LDS a chunk of levels ahead).  mode 0: fixed
 * defaults (SpMV CSR 256, sweep 512).  mode 2..11: force sweep shape mode-2 in
 * that order and SpMV shape (mode-2) mod 5 (3 = SELL-64, 4 = SELL-64 with 16-bit
 * offsets; CSR 256 when that copy was not built).
 * All shapes compute identical values, so the mode never changes results.  No counterpart in the reference (its kernels have fixed
 * launch shapes, src/gpu/cuda_sparse.h, src/gpu/local_solve_kernel.cu). */
sg_status sg_context_set_autotune(sg_context* ctx, int mode);
/* Tuning cache: the autotuner (mode 1) first looks a matrix / preconditioner
 * up by a key (sizes plus a hash of its sparsity structure) and uses the
 * stored shape without timing; new choices are appended to the file at path
 * (one "key shape" per line), which this call loads.  Makes the selection
 * deterministic across runs.  NULL or "" keeps the cache in memory only. */
sg_status sg_context_set_tune_cache(sg_context* ctx, const char* path);
sg_status sg_context_kernel_stats(sg_context* ctx, int family, long long* launches,
                                  double* total_ms);
sg_status sg_context_reset_stats(sg_context* ctx);
/* Work units of a family's timed launches since the last reset: for
 * SG_K_GRAM_DOT / SG_K_GRAM_UPDATE the basis columns j + 1 of each launch
 * (so algorithmic bytes are 8 n (units + launches) and 8 n (units + 2 launches)),
 * one per launch for the other families.  Host only. */
sg_status sg_context_kernel_units(sg_context* ctx, int family, double* units);

#ifdef __cplusplus
}
#endif
#endif /* SG_CAPI_H */
