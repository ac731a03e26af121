/*
 * libsds_hip -- the MI355X (gfx950) numeric phase of SparseDirectSolver.jl's
 * multifrontal solver behind a plain C ABI.
 *
 * SparseDirectSolver.jl has no C ABI of its own: the numeric phase is reached
 * from `execute!` through `_numeric_phase!` (src/solver.jl:717-741) and the
 * solve phase through `_solve_phase!` (src/solver.jl:810-871), both on Julia
 * objects (`Symbolic`, `Numeric`, `SolveWorkspace`). Each entry point below
 * replaces one of those Julia calls; a Julia package extension binds them with
 * `ccall` (INTEGRATION.md shows the binding).
 *
 * Conventions
 *  - One handle per solver, bound to one HIP stream; every call is
 *    asynchronous on that stream except where a host result is returned.
 *    One host thread per handle (a `DirectSolver` is an unlocked mutable
 *    handle, src/solver.jl:187).
 *  - Integer arrays of the analysis are passed exactly as the Julia
 *    `Symbolic` holds them: 1-based, element offsets, Int32 or Int64
 *    (`index_bits`).
 *  - Device pointers (nzval, B, X) are borrowed for the duration of the call;
 *    the library owns the factor, the contribution-block pool and the solve
 *    and refinement workspaces (`allocate_numeric`, src/numeric/storage.jl:205-256;
 *    `allocate_solve`, src/solve/sweeps.jl:164-172).
 *  - Return value: SDS_SUCCESS or an error code mapping to the reference's
 *    exception types (src/errors.jl:10-57); sds_last_error() has the message.
 *    A numerical failure is not an error: it is reported through `info`
 *    (src/numeric/factorize.jl:300-302).
 */
#ifndef SDS_HIP_H
#define SDS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sds_handle* sds_handle_t;

enum {
    SDS_SUCCESS = 0,
    SDS_INVALID_VALUE = 1,       /* InvalidValueError  (src/errors.jl:18) */
    SDS_NOT_SUPPORTED = 2,       /* NotSupportedError  (src/errors.jl:30) */
    SDS_FACTORIZATION_ERROR = 3, /* FactorizationError (src/errors.jl:42), e.g. phase order */
    SDS_DEVICE_ERROR = 4,        /* a HIP runtime error */
    SDS_INTERRUPTED = 5          /* InterruptedError   (src/logging.jl:59-63) */
};

/* element types of the factor (the input values are always Float64) */
enum { SDS_R64 = 0, SDS_R32 = 1 };

/* structures: cudssMatrixType_t values (src/types.jl:17-23) */
enum { SDS_STRUCTURE_G = 0, SDS_STRUCTURE_S = 1, SDS_STRUCTURE_H = 2, SDS_STRUCTURE_SPD = 3, SDS_STRUCTURE_HPD = 4 };

/* solve phases: cudssPhase_t bits (src/types.jl:56-71) */
enum {
    SDS_PHASE_SOLVE_FWD_PERM = 16,
    SDS_PHASE_SOLVE_FWD = 32,
    SDS_PHASE_SOLVE_DIAG = 64,
    SDS_PHASE_SOLVE_BWD = 128,
    SDS_PHASE_SOLVE_BWD_PERM = 256,
    SDS_PHASE_SOLVE_REFINEMENT = 512,
    SDS_PHASE_SOLVE = 1008
};

/*
 * Create a handle on `stream` (a hipStream_t; NULL = the default stream) and
 * the current HIP device. Replaces the backend binding that `DirectSolver`
 * makes at construction (src/solver.jl:226-244).
 */
int sds_create(sds_handle_t* handle, void* stream);
int sds_destroy(sds_handle_t handle);
const char* sds_last_error(sds_handle_t handle);
int sds_set_stream(sds_handle_t handle, void* stream);

/*
 * The device maps of an analysis: `DEVICE_MAPS` of src/symbolic/maps.jl:16-18
 * (the entries the numeric and solve phases read), 1-based as in `Symbolic`,
 * each array of `index_bits` (32 or 64) integers; `amap` entries are signed
 * (negative = mirrored entry: conjugate for "H"/"HPD", the U^T structure for
 * "G"). The refinement map (`refinement_map`, src/solve/refinement.jl:67-101:
 * CSR of the full matrix over the user's nzval, src < 0 = conjugated) may be
 * NULL when no refinement is wanted; for "G" it holds 2n rows (M, then M^T).
 */
typedef struct {
    int64_t n, nnz, ns;
    int32_t structure; /* SDS_STRUCTURE_* */
    int32_t index_bits;
    const void *perm, *super_ptr, *snparent, *rowptr, *rowval;
    const void *front_ptr, *front_nrows, *front_ncols;
    const void *child_ptr, *child_list, *relind_ptr, *relind;
    const void *amap, *amap_ptr, *amap_src;
    int64_t full_nrows, full_nnz;
    const void *full_rowptr, *full_colval, *full_src;
} sds_symbolic_desc;

/*
 * Upload an analysis and build the device schedule (tree levels, size
 * classes, contribution-block placement). Replaces `adapt(backend, S, INT)`
 * (src/symbolic/maps.jl:288-294) plus the launch plan of `NumericPlan`
 * (src/numeric/storage.jl:36-104).
 */
int sds_upload_symbolic(sds_handle_t handle, const sds_symbolic_desc* desc);

/*
 * Allocate the numeric storage for factors of element type `dtype`
 * (SDS_R64, or SDS_R32 = FP32 factor with FP64 refinement) and `nbatch`
 * members (must be 1: a uniform batch is uploaded as one block-diagonal forest of
 * copies of the member's analysis, whose members are the blocks). Replaces
 * `allocate_numeric` (src/numeric/storage.jl:205-256).
 */
int sds_alloc_numeric(sds_handle_t handle, int32_t dtype, int32_t nbatch);

/* pivoting parameters of one factorization (src/options.jl:155-222) */
typedef struct {
    int32_t pivot_type;        /* cudssPivotType_t value (src/types.jl:84-92) */
    int32_t pivot_epsilon_alg; /* 0 default, 1 "algo1" (scaled by max |a_ij|), 2 "algo2" */
    double pivot_threshold;
    double pivot_epsilon;      /* < 0: the default for the factor precision (src/options.jl:57-61) */
    const int8_t* pivot_sign;  /* host, n entries (original rows), or NULL */
    /* host flag polled between launch groups, or NULL: one byte, nonzero = interrupt (the layout of
     * the reference's `user_host_interrupt`, a Threads.Atomic{Bool}: src/options.jl:201) */
    const volatile uint8_t* interrupt;
} sds_factor_params;

/*
 * Numeric factorization of the values `nzval` (device, Float64, nnz entries in
 * the order of the user's CSR). Replaces `factorize!` / `factorize_cholesky!`
 * / `factorize_ldlt!` / `factorize_lu!` (src/numeric/factorize.jl:315-360,
 * src/numeric/ldlt.jl:1447-1475, src/numeric/lu.jl:543-555), including the
 * one device-to-host read of the reduced status: `*info` is 0 or the original
 * column (1-based) of the first failed pivot (src/numeric/factorize.jl:383-417).
 */
int sds_factorize(sds_handle_t handle, const double* nzval, const sds_factor_params* params, int64_t* info);

/*
 * Solve phases with the current factor: `phase` is SDS_PHASE_SOLVE or an OR
 * of its sub-phases, run in order fwd_perm, fwd, diag, bwd, bwd_perm,
 * refinement (src/solver.jl:846-869). B and X are device, column-major,
 * Float64, n x nrhs with leading dimensions ldb/ldx (X == B allowed except
 * with refinement). `solve_mode` 0 solves A x = b, 1 A^T x = b
 * (src/solve/refinement.jl:10-15). Refinement (`refine!`,
 * src/solve/refinement.jl:359-404) runs `ir_n_steps` steps of
 * X += A^-1 (B - A X) with the residual in Float64 on the values `nzval`
 * (device; may be NULL without refinement), stopping early when
 * ir_tol > 0 and max_k ||R_k||_2 / ||B_k||_2 <= ir_tol (||R_k||_2 for a zero B_k,
 * src/solve/refinement.jl:284-291); the steps done go to *ir_steps (may be NULL). Replaces permute_rhs!, forward_sweep!,
 * diagonal_sweep!, backward_sweep!, unpermute_solution! and refine!
 * (src/solve/permute.jl:117-150, src/solve/sweeps.jl:908-1039).
 */
int sds_solve(sds_handle_t handle, int32_t phase, int32_t solve_mode, const double* B, int64_t ldb, double* X,
              int64_t ldx, int32_t nrhs, const double* nzval, int32_t ir_n_steps, double ir_tol, int32_t* ir_steps);

/*
 * Schur complement mode (`schur_mode = 1` with `user_schur_indices`, src/numeric/schur.jl:1-10,
 * src/solver.jl:400-416): after sds_upload_symbolic of an analysis that ordered the Schur
 * indices last as one dense root supernode of nschur columns, sds_set_schur(h, nschur) makes that
 * root the Schur block (0 turns the mode off); call sds_alloc_numeric afterwards. sds_factorize
 * then factors every other front and only assembles the root (`assemble_schur!`,
 * src/numeric/schur.jl:52-78): its panel holds S = A22 - A21 A11^-1 A12, its columns get the
 * identity pivot order, unit 1x1 pivots and zero statistics. The solve phases read B and write X
 * in factor order unless they permute (the last nschur entries are the Schur block, in increasing
 * original order): SDS_PHASE_SOLVE_FWD_SCHUR leaves the condensed right-hand side
 * b2 - A21 A11^-1 b1 there, SDS_PHASE_SOLVE_DIAG applies D^-1 to the factored part, and
 * SDS_PHASE_SOLVE_BWD_SCHUR, with the solution x2 of S x2 = b2' in those entries of B, gives the
 * solution of A x = b. A solve phase with both sweeps, or with refinement, is SDS_NOT_SUPPORTED.
 */
#define SDS_PHASE_SOLVE_FWD_SCHUR 48   /* fwd_perm | fwd */
#define SDS_PHASE_SOLVE_BWD_SCHUR 384  /* bwd | bwd_perm */
int sds_set_schur(sds_handle_t handle, int64_t nschur);

/*
 * Matching and scaling (`matching_alg` = "algo1".."algo6", src/solver.jl:543-604,
 * src/matching/mc64.jl:463-470): after sds_upload_symbolic of the analysis of the matched matrix
 * M = (Dr A Dc)[:, q] ("G": the stored pattern with column c relabeled q^-1(c); the symmetric
 * structures: D A D, the pattern unchanged), weights (host, nnz: r_i c_j of each stored entry
 * (i, j) in the original numbering), rscale and cscale (host, n: Dr and Dc in the original
 * numbering; equal for the symmetric structures) and colperm (host, n, 1-based: q, NULL for the
 * symmetric structures). sds_factorize then factors the values w_p a_p (its statistics, inertia
 * and diag are those of M); the solve phases solve A x = b as M y = Dr b, x = Dc q y (A^T x = b
 * as M^T y = q^T Dc b, x = Dr y), and the refinement residual is that of A with the caller's
 * values (the full-pattern map must then be that of A, in the original numbering). weights =
 * NULL clears the matching. Not with the Schur complement mode.
 */
int sds_set_matching(sds_handle_t handle, const double* weights, const double* rscale, const double* cscale,
                     const int64_t* colperm);

/*
 * The Schur complement S of the last factorization (Schur complement mode) into the device
 * buffer S: nschur x nschur, column-major, leading dimension lds, Float64 (the full matrix: the
 * symmetric structures mirror the lower triangle, "G" assembles the whole square).
 * Replaces `schur_matrix!` with a dense destination (src/numeric/schur.jl:112-150).
 * Asynchronous on the handle's stream.
 */
int sds_schur_export(sds_handle_t handle, double* S, int64_t lds);

/*
 * sds_solve without device allocation. No counterpart in the reference: it is for callers that
 * cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). Same arguments, kernels and results as sds_solve (bitwise), but the solve
 * workspace (permuted right-hand sides, update vectors, partial sums, residual, a copy of B,
 * column norms) lives in a fixed array of SDS_NOALLOC_SCRATCH_MIB MiB per GPU inside the library,
 * reserved on a device at the first sds_solve_noalloc call there and held until the process
 * ends; the call itself allocates nothing. Needs about (2 s + 16) n nrhs bytes plus the update
 * vectors (s = 8 for an SDS_R64 factor, 4 for SDS_R32); a call that needs more returns
 * SDS_INVALID_VALUE, and FGMRES refinement (whose Krylov basis is allocated) returns
 * SDS_NOT_SUPPORTED. Every call on one GPU shares that GPU's array and leaves its contents
 * behind: calls that overlap on one GPU (several handles, several streams or threads) race and
 * must be serialized by the caller. Not profiled by sds_profile.
 */
#define SDS_NOALLOC_SCRATCH_MIB 128
int sds_solve_noalloc(sds_handle_t handle, int32_t phase, int32_t solve_mode, const double* B, int64_t ldb, double* X,
                      int64_t ldx, int32_t nrhs, const double* nzval, int32_t ir_n_steps, double ir_tol,
                      int32_t* ir_steps);

/*
 * Refinement mode of sds_solve: SDS_IR_PLAIN (0, the steps above) or SDS_IR_FGMRES (1):
 * `ir_mode = "ir" | "fgmres"` (src/types.jl:185, :306; src/solver.jl:385-392, :907).
 * With FGMRES the refinement phase replaces `fgmres_refine!` (src/solve/refinement.jl:554-620):
 * R0 = B - op(A) X; when ir_tol > 0 and max_k ||R0_k||_2 / ||B_k||_2 <= ir_tol nothing more is
 * done; else FGMRES on op(A) D = R0 over the n nrhs stacked columns, right-preconditioned by the
 * factorization, for at most ir_n_steps iterations (one SpMV and one solve each), stopping once
 * ||R0 - op(A) D||_2 <= ir_tol min_k ||B_k||_2 (1 for a zero B_k; ir_tol = 0: all iterations),
 * then X += D;
 * *ir_steps gets the iterations. It synchronizes with the host once per iteration, so
 * sds_graph_capture refuses it (SDS_NOT_SUPPORTED) when refinement is captured.
 */
#define SDS_IR_PLAIN 0
#define SDS_IR_FGMRES 1
int sds_set_ir_mode(sds_handle_t handle, int32_t mode);

/*
 * Tile-shape autotuner (no counterpart in the reference, whose kernels have fixed shapes;
 * the chosen form of level and batch launches whose tile shapes come from an autotuner):
 * times `reps` factorizations of the values nzval for each candidate shape of the front
 * kernels (child rows per extend-add task: 512, 1024, 2048, 4096; for Cholesky also the K
 * chunk of the contribution-block update: 16, 32 or 64, the number of trailing-update tiles of a
 * step from which they get a launch of their own: never, 256, 512, 1024, the number of
 * workgroups from which a step launch uses the build for two workgroups per CU: never, 1, 64,
 * 128, 256, 512, the small-front kernels: 0 = fronts of at most 32 rows one wave each and
 * the others in blocked panels, 1 = one wave each / column by column, 2 = all column by column,
 * and the LDS budget of the leaf-subtree kernel (regime A: one workgroup walks a leaf subtree with
 * its fronts and contribution blocks in LDS, src/numeric/subtree.jl:274-327): 0 = none, 16, 32 or
 * 64 KiB, and the panel block of the large fronts: 0 = the 64-column step chain runs over the
 * whole panel, else 128, 256 or 512 columns: the chain runs inside each block and the columns
 * right of a block get its update in one pass with K = the block), keeps the fastest, and leaves
 * the numeric storage reallocated for it (factor contents are not kept: factorize again). tiles
 * (may be NULL, else 8 entries) receives {extend-add rows, K chunk, trailing split, two-per-CU
 * steps, small-front kernels, subtree KiB, panel block, pivot panels (not tuned: the setting in
 * place)}. Except for the panel block, which
 * changes the order in which a panel entry receives its updates, every candidate gives the same
 * factor bit for bit. Synchronous; needs sds_alloc_numeric first.
 */
int sds_autotune(sds_handle_t handle, const double* nzval, const sds_factor_params* params, int32_t reps,
                 int32_t* tiles);

/*
 * Set the tile shapes of the front kernels directly, as sds_autotune reports them: tiles =
 * {extend-add rows (512, 1024, 2048 or 4096), K chunk (16, 32 or 64), trailing split (0 = never,
 * else >= 1), two-per-CU steps (0 = never, else >= 1), small-front kernels (0, 1 or 2), subtree
 * KiB (0, 16, 32 or 64), panel block (0 or a multiple of 64 up to 4096), pivot panels (LDL^T large
 * fronts: 0 = one pivot per pair of launches, the default; 1 = panels of 32 columns with deferred
 * updates, one workgroup per front; both give the same factor bit for bit)}. No
 * counterpart in the reference (whose kernels have fixed shapes); it restores a saved tuning
 * without timing again. Takes effect at the next sds_alloc_numeric (the launch plan depends on it).
 */
int sds_set_tiles(sds_handle_t handle, const int32_t* tiles);

/*
 * HIP graph of one refactorization followed by one solve (no counterpart in the
 * reference, which replays nothing; it is this library's form of the
 * "refactorize + solve captured in a graph" of uneven batches and serving loops).
 * sds_graph_capture records, on a private stream, exactly what sds_factorize and
 * sds_solve would launch for these arguments (values nzval, params, phase,
 * solve_mode, B, X, nrhs, ir_n_steps with ir_tol = 0: no host synchronization),
 * without running it; the pointers are baked into the graph, so each replay reads
 * the current contents of nzval and B and writes X. pivot_sign (if any) is copied
 * once at capture. sds_graph_launch replays it on the handle's stream
 * (asynchronous); sds_graph_info waits and returns the replay's info, as
 * sds_factorize does. Capturing again replaces the graph; sds_alloc_numeric and
 * sds_destroy free it.
 */
int sds_graph_capture(sds_handle_t handle, const double* nzval, const sds_factor_params* params, int32_t phase,
                      int32_t solve_mode, const double* B, int64_t ldb, double* X, int64_t ldx, int32_t nrhs,
                      int32_t ir_n_steps);
int sds_graph_launch(sds_handle_t handle);
int sds_graph_info(sds_handle_t handle, int64_t* info);
int sds_graph_destroy(sds_handle_t handle);

/*
 * Adjoint of X = A^-1 B with respect to the stored values of A (no counterpart in
 * the reference, whose Project.toml has no ChainRules; the differentiation rule is
 * this library's addition): grad[p] = -sum_r lambda[i, r] X[j, r] for the stored
 * entry p at (i, j) of the user's matrix, where lambda = A^-T Xbar comes from
 * sds_solve with solve_mode = 1 on the same factor. A mirrored entry of a symmetric
 * structure ("S", "H", "SPD", "HPD") stands for (i, j) and (j, i) and receives both
 * terms; an entry the solver ignores (the other triangle of a full view) receives 0.
 * lambda, X (n x nrhs, column-major) and grad (nnz) are device Float64 buffers;
 * asynchronous on the handle's stream. Needs the full-pattern map at upload.
 */
int sds_adjoint_scatter(sds_handle_t handle, const double* lambda, int64_t ldl, const double* X, int64_t ldx,
                        int32_t nrhs, double* grad);

/*
 * Data parameters after a factorization (src/solver.jl getparam, the
 * `Numeric` fields of src/numeric/storage.jl:161-179):
 *  stats: 6 * ns Int64 per-front (npos, nneg, nzero, nperturbed, n2x2, info);
 *  totals: 6 Int64 (sums; totals[5] = fronts with a failure);
 *  inertia = (totals[0], totals[1]); npivots = totals[3]
 *  (src/reference/ldlt.jl:570-616).
 */
int sds_get_stats(sds_handle_t handle, int64_t* stats, int64_t* totals);
/* local pivot order (n Int32, 1-based factor columns, `numeric.piv`) and pivot kinds (n Int8) */
int sds_get_pivots(sds_handle_t handle, int32_t* piv, int8_t* kind);
/* D (2n entries of the factor element type: diagonal, then 2x2 subdiagonals); "diag" */
int sds_get_diag(sds_handle_t handle, void* d);
/* host copy of the factor panels (factor_len entries of the factor element type);
 * which = 0: `factor` (L), 1: `ufactor` (U^T of "G") */
int sds_get_factor(sds_handle_t handle, int32_t which, void* out);
/* "memory_estimates" (src/symbolic/maps.jl:338-356): 16 Int64, bytes */
int sds_memory_estimates(sds_handle_t handle, int64_t* est);

/*
 * Per-kernel-class timing. No counterpart in the reference, which profiles from
 * outside (CUPTI traces of bench/profile_phases.jl). With profiling on, every
 * launch of the phases is bracketed by HIP events on the handle's stream and
 * its algorithmic flops and bytes are recorded; sds_profile_read() waits for
 * the recorded launches and returns the totals of one class since
 * sds_profile(h, 1) (which clears them). sds_profile(h, 0) stops recording.
 */
enum {
    SDS_K_FRONT_SMALL = 0, /* fused small-front assembly + factorization */
    SDS_K_ZERO_SCATTER,    /* large fronts: zero + assembly of A */
    SDS_K_EXTEND_ADD,      /* large fronts: extend-add of one child rank */
    SDS_K_POTRF,           /* large fronts: diagonal-block Cholesky */
    SDS_K_TRSM,            /* large fronts: panel triangular solve */
    SDS_K_UPDATE_PANEL,    /* large fronts: Schur update of the panel columns */
    SDS_K_UPDATE_CB,       /* large fronts: contribution-block update (K = w) */
    SDS_K_STATS,           /* statistics / info reduction */
    SDS_K_PERMUTE,         /* right-hand-side permutation in / out */
    SDS_K_FWD,             /* forward sweep (one launch per tree level) */
    SDS_K_BWD,             /* backward sweep (one launch per tree level) */
    SDS_K_DIAG,            /* D^-1 */
    SDS_K_RESIDUAL,        /* refinement residual (FP64 SpMV) */
    SDS_K_PIVOT_PANEL,     /* large fronts: pivoted panel factorization (LDL^T, LU) */
    SDS_K_INVERSE,         /* Cholesky: inverses of the diagonal blocks (for the sweeps) */
    SDS_K_SUBTREE,         /* Cholesky regime A: leaf subtrees, one workgroup each */
    SDS_K_COUNT
};
int sds_profile(sds_handle_t handle, int32_t enable);
int sds_profile_read(sds_handle_t handle, int32_t kind, double* ms, int64_t* launches, double* flops, double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* SDS_HIP_H */
