// Launchers of the device kernels (defined in k_factor.hip, k_ldlt.hip, k_lu.hip, k_solve.hip).
#pragma once
#include "sds_common.h"

namespace sds {

template <typename T>
void launch_subtree_chol(hipStream_t st, const DevSym& S, const int32_t* sub_ptr, const int32_t* sub_nodes, int first,
                         int count, int lds, const int32_t* lfront, const int32_t* lcb, const double* nz, T* factor,
                         T* cbpool, T* DB, int32_t* status, double eps);
// Schur complement mode: the root's pivot order, unit pivots and zero statistics; the export of S
template <typename T>
void launch_schur_meta(hipStream_t st, const DevSym& S, int s, int w, int32_t* piv, int8_t* kind, T* dvec,
                       int64_t* stats);
template <typename T>
void launch_schur_export(hipStream_t st, const DevSym& S, int s, int ns, const T* factor, int full, double* out,
                         int64_t lds);
// its rows of the forward sweep: the children's update vectors only
template <typename T>
void launch_schur_pull(hipStream_t st, const DevSym& S, int s, T* Y, const T* U, int nrhs);
template <typename T>
void launch_front_col_chol(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, int fmax,
                           const double* nz, T* factor, T* cbpool, T* DB, int32_t* status, double eps);
template <typename T>
void launch_front_tiny_chol(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, const double* nz,
                            T* factor, T* cbpool, T* DB, int32_t* status, double eps);
template <typename T>
void launch_front_small_chol(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, int fmax,
                             const double* nz, T* factor, T* cbpool, T* DB, int32_t* status, double eps);
template <typename T>
void launch_big_zero_scatter(hipStream_t st, const DevSym& S, const Task* tasks, int n, const double* nz, T* factor,
                             T* cbpool);
template <typename T>
void launch_big_extend_add(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* cbpool);
template <typename T>
void launch_big_diag(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* DB, int32_t* status,
                     double eps);
// pb: panel block width (0: the step chain runs over the whole panel)
template <typename T>
void launch_big_step(hipStream_t st, const DevSym& S, const Task* tasks, int n, int nstep, int ncopy, T* factor,
                     T* cbpool, T* DB, T* LN, int32_t* status, double eps, int occ, int pb);
// the next strip of a panel block (first ndiag tasks: the next diagonal blocks, factored too)
template <typename T>
void launch_big_block(hipStream_t st, const DevSym& S, const Task* tasks, int n, int ndiag, T* factor, T* cbpool,
                      T* DB, int32_t* status, double eps, int pb);
template <typename T>
void launch_big_update(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* cbpool,
                       const int32_t* status, int mode, int kc, int pb);
void launch_chol_stats(hipStream_t st, const DevSym& S, const int32_t* status, int64_t* stats, int64_t* info);
// info[0] = the smallest of the `blocks` block minima at info + 8 (0 for none)
void launch_info_final(hipStream_t st, int blocks, int64_t* info);
void launch_fill_i32(hipStream_t st, int32_t* p, int64_t n, int32_t v);

// LDL^T (k_ldlt.hip)
template <typename T>
void launch_front_small_ldlt(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, int fmax,
                             const double* nz, T* factor, T* cbpool, T* dvec, int32_t* piv, int8_t* kind,
                             int64_t* stats, int pivot, double u, double eps, int scaled, const double* aux,
                             const int8_t* psign);
template <typename T>
void launch_ldlt_pivot(hipStream_t st, const DevSym& S, const Task* tasks, int n, int t, T* factor, T* dvec,
                       int32_t* piv, int8_t* kind, int64_t* stats, int32_t* lstep, int pivot, double u, double eps,
                       int scaled, const double* aux, const int8_t* psign);
template <typename T>
void launch_ldlt_update(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, const int32_t* lstep);
// the upper triangle of each large front's fully summed square from the lower (task (s, I0, J0),
// a 64 x 64 tile with I0 >= J0), before its pivot launches
template <typename T>
void launch_ldlt_mirror(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor);
// deferred-update panels of the large fronts: panel e of every front (task (s, scratch offset)),
// then its update of the columns after it (task (s, I0, J0, scratch offset))
template <typename T>
void launch_ldlt_panel(hipStream_t st, const DevSym& S, const Task* tasks, int n, int e, T* factor, T* dvec,
                       int32_t* piv, int8_t* kind, int64_t* stats, int32_t* lstep, T* scratch, int pivot, double u,
                       double eps, int scaled, const double* aux, const int8_t* psign);
template <typename T>
void launch_ldlt_trail(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, const int8_t* kind,
                       const int32_t* lstep, const T* scratch);
// scratch elements of one front's panels, and the columns of a panel
int64_t ldlt_panel_scratch(int f, int w);
int ldlt_panel_cols();
template <typename T>
void launch_ldlt_finish(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, int32_t* piv);
template <typename T>
void launch_ldlt_cb(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, T* cbpool,
                    const T* dvec, const int8_t* kind);
void launch_absmax(hipStream_t st, int64_t nnz, const double* nz, double* aux);

// LU (k_lu.hip)
template <typename T>
void launch_front_small_lu(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, int fmax,
                           const double* nz, T* factor, T* ufactor, T* cbpool, T* dvec, int32_t* piv, int8_t* kind,
                           int64_t* stats, int pivoting, double u, double eps, int scaled, const double* aux);
template <typename T>
void launch_big_lu_zero(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* ufactor, T* cbpool);
template <typename T>
void launch_big_lu_scatter(hipStream_t st, const DevSym& S, const Task* tasks, int n, const double* nz, T* factor,
                           T* ufactor);
template <typename T>
void launch_big_lu_extend_add(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* ufactor,
                              T* cbpool);
template <typename T>
void launch_lu_cb(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* ufactor,
                  T* cbpool, const T* dvec);
template <typename T>
void launch_lu_pivot(hipStream_t st, const DevSym& S, const Task* tasks, int n, int k, T* factor, T* ufactor,
                     T* dvec, int32_t* piv, int8_t* kind, int64_t* stats, int pivoting, double u, double eps,
                     int scaled, const double* aux);
template <typename T>
void launch_lu_update(hipStream_t st, const DevSym& S, const Task* tasks, int n, int k, T* factor, T* ufactor,
                      const T* dvec);
template <typename T>
void launch_lu_finish(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* ufactor,
                      const T* dvec, int32_t* piv);
void launch_info_from_stats(hipStream_t st, const DevSym& S, const int64_t* stats, int64_t* info);

template <typename T>
void launch_permute_in(hipStream_t st, int n, int nrhs, const int32_t* perm, const double* B, int64_t ldb, T* Y,
                       const int32_t* cq = nullptr, const double* sc = nullptr);
template <typename T>
void launch_permute_out(hipStream_t st, int n, int nrhs, const int32_t* perm, const T* Y, double* X, int64_t ldx,
                        int add, const int32_t* cq = nullptr, const double* sc = nullptr);
// out = w .* nz (the matching's scaled values)
void launch_scale_values(hipStream_t st, int64_t nnz, const double* w, const double* nz, double* out);
template <typename T>
void launch_fwd(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, const T* factor, T* Y, T* U, T* Z,
                int nrhs, int unit, const int32_t* piv);
template <typename T>
void launch_bwd(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, const T* factor, T* Y, T* Z,
                int nrhs, int unit, const int32_t* piv);
template <typename T>
void launch_diag(hipStream_t st, int n, int nrhs, const T* d, const int8_t* kind, T* Y);
template <typename T>
void launch_fwd_pull(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* Y, T* U, T* Z, int nrhs,
                     const int32_t* piv);
template <typename T>
void launch_fwd_step(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, T* Y, T* U, T* Z,
                     int nrhs, int unit);
template <typename T>
void launch_finish(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* Y, const T* Z, int nrhs,
                   const int32_t* piv);
template <typename T>
void launch_bwd_l21(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* Y, T* part,
                    int nrhs);
template <typename T>
void launch_bwd_l21_sum(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* Y, const T* part, int nrhs);
int bwd_l21_chunk();  // rows per partial sum of launch_bwd_l21
template <typename T>
void launch_bwd_step(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, T* Y, T* Z, int nrhs,
                     int unit);
// Cholesky sweeps with inverted diagonal blocks (k_solve_inv.hip)
template <typename T>
void launch_block_inverse(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* DB, int bs);
template <typename T>
void launch_fwd_diag(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, int threads,
                     const T* factor, const T* DB, T* Y, T* U, int nrhs);
template <typename T>
void launch_fwd_rows(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n,
                     const T* factor, T* Y, const T* Z, T* U, int nrhs);
template <typename T>
void launch_bwd_l21t(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n,
                     const T* factor, const T* Y, T* part, int nrhs);
template <typename T>
void launch_bwd_diag(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, int threads,
                     const T* factor, const T* DB, T* Y, const T* part, int nrhs);
template <typename T>
void launch_step_inverse(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n,
                         const T* factor, const T* DB, T* XB);
template <typename T>
void launch_fwd_inv(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, const T* XB,
                    const T* Y, T* Z, const T* U, int nrhs);
template <typename T>
void launch_bwd_inv(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, const T* XB,
                    T* Y, const T* Z, int nrhs);
// one launch for a step group's narrow steps and wide steps (forward: k_fwd_diag + k_fwd_inv tasks;
// backward: k_bwd_diag + k_sub_partials tasks), 256 threads
template <typename T>
void launch_fwd_group(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tdiag, int ndiag,
                      const Task* tinv, int ninv, const T* factor, const T* DB, const T* XB, T* Y, T* Z, T* U,
                      int nrhs);
template <typename T>
void launch_bwd_group(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tdiag, int ndiag,
                      const Task* tsub, int nsub, const T* factor, const T* DB, T* Y, T* Z, const T* part,
                      int nrhs);
template <typename T>
void launch_sub_partials(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n,
                         const T* Y, T* Z, const T* part, int nrhs);
template <typename T>
void launch_copy_rows(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, T* Y,
                      const T* Z, int nrhs);
void launch_adjoint_scatter(hipStream_t st, int64_t nnz, int nrhs, const int32_t* ai, const int32_t* aj,
                            const int8_t* aflag, const double* lam, int64_t ldl, const double* X, int64_t ldx,
                            double* grad);
// FGMRES refinement vectors (k_solve.hip): out[0] = x . y (part: DOT_BLOCKS partial sums), y -= a[0] x,
// y *= s, X(:, r) += sum_j c_j Z_j(:, r)
constexpr int DOT_BLOCKS = 256;
void launch_dot(hipStream_t st, int64_t len, const double* x, const double* y, double* part, double* out);
void launch_axpy_neg(hipStream_t st, int64_t len, double* y, const double* x, const double* a);
void launch_scale(hipStream_t st, int64_t len, double* y, double s);
void launch_add_combination(hipStream_t st, int n, int nrhs, double* X, int64_t ldx, const double* Z, int64_t zs, int k,
                            const double* c);
void launch_residual(hipStream_t st, int n, int nrhs, const int64_t* rowptr, const int32_t* colval, const int64_t* src,
                     const double* nz, const double* X, int64_t ldx, const double* B, int64_t ldb, double* R);

}  // namespace sds
