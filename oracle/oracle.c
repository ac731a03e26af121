/*
 * CPU oracle of the numeric phase -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of SparseDirectSolver.jl's CPU reference factorizations
 * (src/reference/cholesky.jl, src/reference/ldlt.jl, src/reference/lu.jl): the
 * same postorder traversal over the supernodes, the same assembly of A through
 * amap_ptr/amap_src/amap, the same extend-add of the children's contribution
 * blocks through relind in child_list order, and the same dense in-front
 * decision rules (potrf acceptance, Bunch-Kaufman 1x1/2x2 with the threshold
 * test, threshold partial pivoting for LU, static perturbation). Only the
 * tests, the package's smoke check and the CPU-baseline leg of the benchmark
 * may load it; the product path never does.
 *
 * All index arrays are the 1-based arrays of the analysis (the Julia
 * `Symbolic`, src/symbolic/maps.jl:62-95). Dense fronts are column-major f x f.
 * Built with -ffp-contract=off so that no multiply-add is fused (the Julia
 * reference does not fuse them either).
 *
 * Real element types only (Float64); the complex variants of the reference are
 * not restated.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int64_t n, nnz, ns, factor_len;
    const int64_t *perm, *super_ptr, *snparent, *snpost, *rowptr, *rowval;
    const int64_t *front_ptr, *front_nrows, *front_ncols;
    const int64_t *child_ptr, *child_list, *relind_ptr, *relind;
    const int64_t *amap, *amap_ptr, *amap_src;
} oracle_sym;

enum { NF = 6, ST_NPOS = 0, ST_NNEG = 1, ST_NZERO = 2, ST_NPERT = 3, ST_N2X2 = 4, ST_INFO = 5 };
enum { K_1X1 = 1, K_2X2_FIRST = 2, K_2X2_SECOND = 3, K_PERTURBED = 4 };
enum { PV_NONE = 0, PV_DIAG = 1, PV_BK = 2 };

#define F_(i, j) F[(int64_t)(j) * f + (i)]

/* contribution blocks of the traversal (full m x m, column-major), freed once consumed */
typedef struct {
    double **cb;
    int64_t *m;
} cbstore;

static void cb_init(cbstore *c, int64_t ns) {
    c->cb = (double **)calloc((size_t)ns, sizeof(double *));
    c->m = (int64_t *)calloc((size_t)ns, sizeof(int64_t));
}
static void cb_free(cbstore *c, int64_t ns) {
    for (int64_t s = 0; s < ns; ++s) free(c->cb[s]);
    free(c->cb);
    free(c->m);
}

/* zero the front, add A through the owner-grouped amap, extend-add the children
 * (src/reference/cholesky.jl:73-92; ldlt.jl:176-190; lu.jl:114-131). `full`:
 * the children's whole m x m blocks are added (LU), else their lower triangles;
 * `lu`: a negative amap offset is an entry of the U^T structure (placed at (j,i)). */
static void assemble(const oracle_sym *S, int64_t s, const double *nz, double *F, cbstore *cbs, int full, int lu) {
    int64_t f = S->front_nrows[s], base = S->front_ptr[s];
    memset(F, 0, sizeof(double) * (size_t)(f * f));
    for (int64_t k = S->amap_ptr[s]; k < S->amap_ptr[s + 1]; ++k) {
        int64_t p = S->amap_src[k - 1] - 1;
        int64_t off = S->amap[p];
        int64_t loc = (off < 0 ? -off : off) - base;
        int64_t i = loc % f, j = loc / f;
        if (lu && off < 0)
            F_(j, i) += nz[p];
        else
            F_(i, j) += nz[p]; /* real: the conjugate of a mirrored entry is itself */
    }
    for (int64_t k = S->child_ptr[s]; k < S->child_ptr[s + 1]; ++k) {
        int64_t c = S->child_list[k - 1] - 1;
        double *C = cbs->cb[c];
        if (!C) continue;
        int64_t m = cbs->m[c];
        const int64_t *ri = S->relind + (S->relind_ptr[c] - 1);
        for (int64_t jj = 0; jj < m; ++jj)
            for (int64_t ii = full ? 0 : jj; ii < m; ++ii) F_(ri[ii] - 1, ri[jj] - 1) += C[jj * m + ii];
        free(C);
        cbs->cb[c] = NULL;
    }
}

static void keep_cb(const oracle_sym *S, int64_t s, const double *F, cbstore *cbs) {
    int64_t f = S->front_nrows[s], w = S->front_ncols[s], m = f - w;
    if (m <= 0 || S->snparent[s] == 0) return;
    double *C = (double *)malloc(sizeof(double) * (size_t)(m * m));
    for (int64_t jj = 0; jj < m; ++jj)
        for (int64_t ii = 0; ii < m; ++ii) C[jj * m + ii] = F_(w + ii, w + jj);
    cbs->cb[s] = C;
    cbs->m[s] = m;
}

/* ------------------------------------------------------------------------- */
/* Cholesky: ref_factorize! (src/reference/cholesky.jl:44-111)                */

/* dense Cholesky of F11 (left-looking, as LAPACK potf2 'L'); 0 or the failed column (1-based) */
static int64_t potrf_lower(double *F, int64_t f, int64_t w) {
    for (int64_t j = 0; j < w; ++j) {
        for (int64_t k = 0; k < j; ++k) {
            double ljk = F_(j, k);
            for (int64_t i = j; i < w; ++i) F_(i, j) -= F_(i, k) * ljk;
        }
        double ajj = F_(j, j);
        if (!(ajj > 0.0)) return j + 1;
        ajj = sqrt(ajj);
        F_(j, j) = ajj;
        double r = 1.0 / ajj;
        for (int64_t i = j + 1; i < w; ++i) F_(i, j) *= r;
    }
    return 0;
}

static int64_t cholesky_range(const oracle_sym *S, const double *nz, double *factor, int64_t *stats, int64_t q1);

int64_t oracle_cholesky(const oracle_sym *S, const double *nz, double *factor, int64_t *stats) {
    return cholesky_range(S, nz, factor, stats, S->ns);
}

/* the first q1 supernodes of the postorder only (a union of whole subtrees): the bounded
 * sample the CPU baseline times */
int64_t oracle_cholesky_prefix(const oracle_sym *S, const double *nz, double *factor, int64_t *stats, int64_t q1) {
    return cholesky_range(S, nz, factor, stats, q1 < S->ns ? q1 : S->ns);
}

static int64_t cholesky_range(const oracle_sym *S, const double *nz, double *factor, int64_t *stats, int64_t q1) {
    int64_t ns = S->ns;
    memset(factor, 0, sizeof(double) * (size_t)S->factor_len);
    memset(stats, 0, sizeof(int64_t) * (size_t)(NF * ns));
    cbstore cbs;
    cb_init(&cbs, ns);
    int64_t fmax = 0;
    for (int64_t s = 0; s < ns; ++s) fmax = S->front_nrows[s] > fmax ? S->front_nrows[s] : fmax;
    double *F = (double *)malloc(sizeof(double) * (size_t)(fmax * fmax + 1));
    int64_t ret = 0;
    for (int64_t q = 0; q < q1; ++q) {
        int64_t s = S->snpost[q] - 1;
        int64_t f = S->front_nrows[s], w = S->front_ncols[s], m = f - w, base = S->front_ptr[s] - 1;
        assemble(S, s, nz, F, &cbs, 0, 0);
        int64_t linfo = potrf_lower(F, f, w);
        if (linfo > 0) {
            stats[s * NF + ST_NPOS] = linfo - 1;
            stats[s * NF + ST_INFO] = linfo;
            ret = S->perm[S->super_ptr[s] + linfo - 2];
            break;
        }
        if (m > 0) {
            /* trsm('R','L','C','N'): F21 <- F21 L11^-T, column by column */
            for (int64_t k = 0; k < w; ++k) {
                for (int64_t j = 0; j < k; ++j) {
                    double lkj = F_(k, j);
                    for (int64_t i = w; i < f; ++i) F_(i, k) -= F_(i, j) * lkj;
                }
                double dk = F_(k, k);
                for (int64_t i = w; i < f; ++i) F_(i, k) /= dk;
            }
            /* syrk('L','N',-1,F21,1,F22) */
            for (int64_t jj = w; jj < f; ++jj)
                for (int64_t k = 0; k < w; ++k) {
                    double ljk = F_(jj, k);
                    for (int64_t ii = jj; ii < f; ++ii) F_(ii, jj) -= F_(ii, k) * ljk;
                }
            keep_cb(S, s, F, &cbs);
        }
        for (int64_t j = 0; j < w; ++j)
            for (int64_t i = j; i < f; ++i) factor[base + j * f + i] = F_(i, j);
        stats[s * NF + ST_NPOS] = w;
    }
    free(F);
    cb_free(&cbs, ns);
    return ret;
}

/* ------------------------------------------------------------------------- */
/* solves: ref_solve! (cholesky.jl:127-170), ref_solve_ldlt! (ldlt.jl:467-539)
 * and the LU solves (lu.jl:283-294, restated supernode by supernode).
 * Y is n x nrhs column-major in supernodal numbering. */

/* forward sweep with the lower panels `P` (unit diagonal when `unit`);
 * `piv` (nullable): reorder each supernode's slice by its local pivot order first */
static void fwd_sweep(const oracle_sym *S, const double *P, int unit, const int32_t *piv, double *Y,
                      int64_t nrhs, double *blk) {
    int64_t n = S->n;
    for (int64_t q = 0; q < S->ns; ++q) {
        int64_t s = S->snpost[q] - 1;
        int64_t f = S->front_nrows[s], w = S->front_ncols[s], c0 = S->super_ptr[s] - 1;
        const double *L = P + (S->front_ptr[s] - 1);
        const int64_t *rows = S->rowval + (S->rowptr[s] - 1);
        for (int64_t r = 0; r < nrhs; ++r) {
            double *y = Y + r * n;
            if (piv) {
                for (int64_t k = 0; k < w; ++k) blk[k] = y[piv[c0 + k] - 1];
                for (int64_t k = 0; k < w; ++k) y[c0 + k] = blk[k];
            }
            for (int64_t j = 0; j < w; ++j) {
                double x = y[c0 + j];
                if (!unit) x /= L[j * f + j];
                y[c0 + j] = x;
                for (int64_t i = j + 1; i < w; ++i) y[c0 + i] -= L[j * f + i] * x;
            }
            for (int64_t i = w; i < f; ++i) {
                double t = 0.0;
                for (int64_t j = 0; j < w; ++j) t += L[j * f + i] * y[c0 + j];
                y[rows[i] - 1] -= t;
            }
        }
    }
}

/* backward sweep with the transposes of the lower panels `P`; `piv` (nullable):
 * undo each supernode's local pivot order afterwards */
static void bwd_sweep(const oracle_sym *S, const double *P, int unit, const int32_t *piv, double *Y,
                      int64_t nrhs, double *blk) {
    int64_t n = S->n;
    for (int64_t q = S->ns - 1; q >= 0; --q) {
        int64_t s = S->snpost[q] - 1;
        int64_t f = S->front_nrows[s], w = S->front_ncols[s], c0 = S->super_ptr[s] - 1;
        const double *L = P + (S->front_ptr[s] - 1);
        const int64_t *rows = S->rowval + (S->rowptr[s] - 1);
        for (int64_t r = 0; r < nrhs; ++r) {
            double *y = Y + r * n;
            for (int64_t j = 0; j < w; ++j) {
                double t = 0.0;
                for (int64_t i = w; i < f; ++i) t += L[j * f + i] * y[rows[i] - 1];
                y[c0 + j] -= t;
            }
            for (int64_t j = w - 1; j >= 0; --j) {
                double x = y[c0 + j];
                for (int64_t i = j + 1; i < w; ++i) x -= L[j * f + i] * y[c0 + i];
                if (!unit) x /= L[j * f + j];
                y[c0 + j] = x;
            }
            if (piv) {
                for (int64_t k = 0; k < w; ++k) blk[k] = y[c0 + k];
                for (int64_t k = 0; k < w; ++k) y[piv[c0 + k] - 1] = blk[k];
            }
        }
    }
}

static double *permute_in(const oracle_sym *S, const double *B, int64_t nrhs) {
    int64_t n = S->n;
    double *Y = (double *)malloc(sizeof(double) * (size_t)(n * nrhs + 1));
    for (int64_t r = 0; r < nrhs; ++r)
        for (int64_t k = 0; k < n; ++k) Y[r * n + k] = B[r * n + S->perm[k] - 1];
    return Y;
}

static void permute_out(const oracle_sym *S, const double *Y, double *X, int64_t nrhs) {
    int64_t n = S->n;
    for (int64_t r = 0; r < nrhs; ++r)
        for (int64_t k = 0; k < n; ++k) X[r * n + S->perm[k] - 1] = Y[r * n + k];
}

static double *blk_alloc(const oracle_sym *S) {
    int64_t wmax = 1;
    for (int64_t s = 0; s < S->ns; ++s) wmax = S->front_ncols[s] > wmax ? S->front_ncols[s] : wmax;
    return (double *)malloc(sizeof(double) * (size_t)wmax);
}

void oracle_solve_cholesky(const oracle_sym *S, const double *factor, const double *B, double *X, int64_t nrhs) {
    double *Y = permute_in(S, B, nrhs), *blk = blk_alloc(S);
    fwd_sweep(S, factor, 0, NULL, Y, nrhs, blk);
    bwd_sweep(S, factor, 0, NULL, Y, nrhs, blk);
    permute_out(S, Y, X, nrhs);
    free(Y);
    free(blk);
}

/* D^-1 with 1x1 and 2x2 blocks (ldlt.jl:505-524) */
static void diag_sweep(const oracle_sym *S, const double *d, const int8_t *kind, double *Y, int64_t nrhs, int two) {
    int64_t n = S->n;
    for (int64_t k = 0; k < n;) {
        if (two && kind[k] == K_2X2_FIRST) {
            double a = d[k], b = d[n + k], c = d[k + 1];
            double det = a * c - b * b;
            for (int64_t r = 0; r < nrhs; ++r) {
                double y1 = Y[r * n + k], y2 = Y[r * n + k + 1];
                Y[r * n + k] = (c * y1 - b * y2) / det;
                Y[r * n + k + 1] = (a * y2 - b * y1) / det;
            }
            k += 2;
        } else {
            for (int64_t r = 0; r < nrhs; ++r) Y[r * n + k] /= d[k];
            k += 1;
        }
    }
}

void oracle_solve_ldlt(const oracle_sym *S, const double *factor, const double *d, const int32_t *piv,
                       const int8_t *kind, const double *B, double *X, int64_t nrhs) {
    double *Y = permute_in(S, B, nrhs), *blk = blk_alloc(S);
    fwd_sweep(S, factor, 1, piv, Y, nrhs, blk);
    diag_sweep(S, d, kind, Y, nrhs, 1);
    bwd_sweep(S, factor, 1, piv, Y, nrhs, blk);
    permute_out(S, Y, X, nrhs);
    free(Y);
    free(blk);
}

/* LU solve of A X = B (transpose = 0) or A^T X = B (transpose = 1): P A' = L D U with the
 * local row orders `piv`; L = factor, U^T = ufactor, both unit lower panels */
void oracle_solve_lu(const oracle_sym *S, const double *factor, const double *ufactor, const double *d,
                     const int32_t *piv, const double *B, double *X, int64_t nrhs, int transpose) {
    double *Y = permute_in(S, B, nrhs), *blk = blk_alloc(S);
    if (!transpose) {
        fwd_sweep(S, factor, 1, piv, Y, nrhs, blk);
        diag_sweep(S, d, NULL, Y, nrhs, 0);
        bwd_sweep(S, ufactor, 1, NULL, Y, nrhs, blk);
    } else {
        fwd_sweep(S, ufactor, 1, NULL, Y, nrhs, blk);
        diag_sweep(S, d, NULL, Y, nrhs, 0);
        bwd_sweep(S, factor, 1, piv, Y, nrhs, blk);
    }
    permute_out(S, Y, X, nrhs);
    free(Y);
    free(blk);
}

/* ------------------------------------------------------------------------- */
/* LDL^T with Bunch-Kaufman: ref_ldlt! (src/reference/ldlt.jl:121-202) and the
 * dense in-front factorization _ldlt_front! (ldlt.jl:384-454). Real symmetric. */

typedef struct {
    int pivot;
    double u, eps;
} ldlt_params;

/* Julia's max(a, b) propagates NaN; keep that */
static inline double jmax(double a, double b) {
    if (a != a || b != b) return a + b;
    return a > b ? a : b;
}

/* d >= u max_i |F[i, c]| with the early exit (ldlt.jl:252-260) */
static int threshold_ok(const double *F, int64_t f, int64_t c, int64_t k, double u, double d) {
    double m = 0.0;
    for (int64_t i = k; i < f; ++i) {
        if (i == c) continue;
        m = jmax(m, fabs(F_(i, c)));
        if (!(d >= u * m)) return 0;
    }
    return d >= u * m;
}

static int accept_1x1(const double *F, int64_t f, int64_t c, int64_t k, const ldlt_params *p) {
    double d = fabs(F_(c, c));
    return d >= p->eps && threshold_ok(F, f, c, k, p->u, d);
}

static int nonsingular_2x2(const double *F, int64_t f, int64_t k, int64_t r, const ldlt_params *p) {
    double det = F_(k, k) * F_(r, r) - F_(k, r) * F_(r, k);
    double mx = jmax(jmax(fabs(F_(k, k)), fabs(F_(r, k))), fabs(F_(r, r)));
    return det != 0.0 && isfinite(det) && mx >= p->eps;
}

/* largest |F[i, c]| over the rows k..f-1 but c and skip (ldlt.jl:240-248) */
static double colmax_j(const double *F, int64_t f, int64_t c, int64_t k, int64_t skip) {
    double m = 0.0;
    for (int64_t i = k; i < f; ++i) {
        if (i == c || i == skip) continue;
        m = jmax(m, fabs(F_(i, c)));
    }
    return m;
}

static int accept_2x2(const double *F, int64_t f, int64_t k, int64_t r, const ldlt_params *p) {
    if (!nonsingular_2x2(F, f, k, r, p)) return 0;
    double det = F_(k, k) * F_(r, r) - F_(k, r) * F_(r, k);
    double e11 = fabs(F_(r, r) / det), e12 = fabs(F_(k, r) / det), e21 = fabs(F_(r, k) / det),
           e22 = fabs(F_(k, k) / det);
    double m1 = colmax_j(F, f, k, k, r), m2 = colmax_j(F, f, r, k, k);
    return p->u * (e11 * m1 + e21 * m2) <= 1.0 && p->u * (e12 * m1 + e22 * m2) <= 1.0;
}

/* acceptable 1x1 pivot of the block k..w-1 with the largest |a_jj| (-1 if none) (ldlt.jl:264-276) */
static int64_t best_1x1(const double *F, int64_t f, int64_t k, int64_t w, const ldlt_params *p) {
    int64_t best = -1;
    double bv = 0.0;
    for (int64_t j = k; j < w; ++j) {
        double a = fabs(F_(j, j));
        if (!(a >= p->eps)) continue;
        if (best >= 0 && !(a > bv)) continue;
        if (!threshold_ok(F, f, j, k, p->u, a)) continue;
        best = j;
        bv = a;
    }
    return best;
}

/* pivot choice at column k (ldlt.jl:280-325): *c = pivot column, *r = 2x2 partner or -1 */
static void choose_pivot(const double *F, int64_t f, int64_t k, int64_t w, const ldlt_params *p, int64_t *c,
                         int64_t *r) {
    *r = -1;
    if (p->pivot == PV_NONE) {
        *c = k;
        return;
    }
    if (p->pivot == PV_DIAG) {
        if (accept_1x1(F, f, k, k, p)) {
            *c = k;
            return;
        }
        int64_t best = best_1x1(F, f, k, w, p);
        if (best >= 0) {
            *c = best;
            return;
        }
        int64_t big = k;
        for (int64_t j = k + 1; j < w; ++j)
            if (fabs(F_(j, j)) > fabs(F_(big, big))) big = j;
        *c = fabs(F_(big, big)) >= p->eps ? big : k;
        return;
    }
    const double alpha = (1.0 + sqrt(17.0)) / 8.0; /* Bunch-Kaufman constant, ldlt.jl:22 */
    double lambda = 0.0;
    int64_t rr = -1;
    for (int64_t i = k + 1; i < w; ++i) {
        double a = fabs(F_(i, k));
        if (a > lambda) {
            lambda = a;
            rr = i;
        }
    }
    double akk = fabs(F_(k, k));
    int64_t bc = k, br = -1;
    if (rr < 0 || akk >= alpha * lambda) {
        bc = k;
    } else {
        double sigma = 0.0;
        for (int64_t i = k; i < w; ++i)
            if (i != rr) sigma = jmax(sigma, fabs(F_(i, rr)));
        if (akk * sigma >= alpha * lambda * lambda)
            bc = k;
        else if (fabs(F_(rr, rr)) >= alpha * sigma)
            bc = rr;
        else {
            bc = k;
            br = rr;
        }
    }
    int ok = br < 0 ? accept_1x1(F, f, bc, k, p) : accept_2x2(F, f, k, br, p);
    if (ok) {
        *c = bc;
        *r = br;
        return;
    }
    int64_t best = best_1x1(F, f, k, w, p);
    if (best >= 0) {
        *c = best;
        return;
    }
    if (br >= 0 && nonsingular_2x2(F, f, k, br, p)) {
        *c = bc;
        *r = br;
        return;
    }
    *c = bc;
}

/* symmetric interchange of columns/rows p, q of the first w columns (ldlt.jl:328-340) */
static void swap_front(double *F, int64_t f, int64_t p, int64_t q, int64_t w, int64_t *lp) {
    if (p == q) return;
    for (int64_t i = 0; i < f; ++i) {
        double t = F_(i, p);
        F_(i, p) = F_(i, q);
        F_(i, q) = t;
    }
    for (int64_t j = 0; j < w; ++j) {
        double t = F_(p, j);
        F_(p, j) = F_(q, j);
        F_(q, j) = t;
    }
    int64_t t = lp[p];
    lp[p] = lp[q];
    lp[q] = t;
}

static void ldlt_front(double *F, int64_t f, int64_t w, const ldlt_params *p, const int8_t *sign_req,
                       int64_t *lp, int8_t *kind, double *dd, double *ee, int64_t *st) {
    for (int64_t k = 0; k < w; ++k) {
        lp[k] = k;
        kind[k] = 0;
        dd[k] = ee[k] = 0.0;
    }
    for (int q = 0; q < NF; ++q) st[q] = 0;
    int64_t k = 0;
    while (k < w) {
        int64_t c, r;
        choose_pivot(F, f, k, w, p, &c, &r);
        if (r < 0) {
            swap_front(F, f, k, c, w, lp);
            double d = F_(k, k);
            kind[k] = K_1X1;
            if (!(fabs(d) >= p->eps) || d == 0.0) {
                if (d == 0.0) st[ST_NZERO]++;
                int req = sign_req ? sign_req[lp[k]] : 0;
                double sg = req != 0 ? (double)req : (d < 0.0 ? -1.0 : 1.0);
                d = sg * p->eps;
                kind[k] = K_PERTURBED;
                st[ST_NPERT]++;
                if (d == 0.0 && st[ST_INFO] == 0) st[ST_INFO] = lp[k] + 1;
            }
            F_(k, k) = d;
            dd[k] = d;
            if (d > 0.0)
                st[ST_NPOS]++;
            else if (d < 0.0)
                st[ST_NNEG]++;
            for (int64_t i = k + 1; i < f; ++i) {
                double l = F_(i, k) / d;
                for (int64_t j = k + 1; j < w; ++j) F_(i, j) -= l * F_(k, j);
                F_(i, k) = l;
            }
            k += 1;
        } else {
            swap_front(F, f, k + 1, r, w, lp);
            double a = F_(k, k), b = F_(k + 1, k), c2 = F_(k + 1, k + 1);
            double up = b;
            F_(k, k + 1) = up;
            double det = a * c2 - up * b;
            double e11 = c2 / det, e12 = -up / det, e21 = -b / det, e22 = a / det;
            kind[k] = K_2X2_FIRST;
            kind[k + 1] = K_2X2_SECOND;
            dd[k] = a;
            dd[k + 1] = c2;
            ee[k] = b;
            st[ST_N2X2]++;
            if (det < 0.0) {
                st[ST_NPOS]++;
                st[ST_NNEG]++;
            } else if (a > 0.0) {
                st[ST_NPOS] += 2;
            } else {
                st[ST_NNEG] += 2;
            }
            for (int64_t i = k + 2; i < f; ++i) {
                double x1 = F_(i, k), x2 = F_(i, k + 1);
                double l1 = x1 * e11 + x2 * e21;
                double l2 = x1 * e12 + x2 * e22;
                for (int64_t j = k + 2; j < w; ++j) F_(i, j) -= l1 * F_(k, j) + l2 * F_(k + 1, j);
                F_(i, k) = l1;
                F_(i, k + 1) = l2;
            }
            F_(k + 1, k) = 0.0;
            k += 2;
        }
    }
}

/* pivot: 0 'N', 1 'D', 2 Bunch-Kaufman; eps: the effective perturbation (scaled by the
 * caller for pivot_epsilon_alg = "algo1"); psign (nullable): requested signs by original row */
int64_t oracle_ldlt(const oracle_sym *S, const double *nz, int pivot, double u, double eps, const int8_t *psign,
                    double *factor, double *d, int32_t *piv, int8_t *pkind, int64_t *stats) {
    int64_t ns = S->ns, n = S->n;
    ldlt_params prm = {pivot, u, eps};
    memset(factor, 0, sizeof(double) * (size_t)S->factor_len);
    memset(d, 0, sizeof(double) * (size_t)(2 * n));
    memset(stats, 0, sizeof(int64_t) * (size_t)(NF * ns));
    memset(piv, 0, sizeof(int32_t) * (size_t)n);
    memset(pkind, 0, (size_t)n);
    cbstore cbs;
    cb_init(&cbs, ns);
    int64_t fmax = 1, wmax = 1;
    for (int64_t s = 0; s < ns; ++s) {
        fmax = S->front_nrows[s] > fmax ? S->front_nrows[s] : fmax;
        wmax = S->front_ncols[s] > wmax ? S->front_ncols[s] : wmax;
    }
    double *F = (double *)malloc(sizeof(double) * (size_t)(fmax * fmax));
    double *W = (double *)malloc(sizeof(double) * (size_t)(fmax * wmax));
    int64_t *lp = (int64_t *)malloc(sizeof(int64_t) * (size_t)wmax);
    int8_t *kind = (int8_t *)malloc((size_t)wmax);
    int8_t *sreq = (int8_t *)malloc((size_t)wmax);
    double *dd = (double *)malloc(sizeof(double) * (size_t)wmax);
    double *ee = (double *)malloc(sizeof(double) * (size_t)wmax);
    for (int64_t q = 0; q < ns; ++q) {
        int64_t s = S->snpost[q] - 1;
        int64_t f = S->front_nrows[s], w = S->front_ncols[s], m = f - w;
        int64_t c0 = S->super_ptr[s] - 1, base = S->front_ptr[s] - 1;
        assemble(S, s, nz, F, &cbs, 0, 0);
        for (int64_t j = 0; j < w; ++j)
            for (int64_t i = j + 1; i < w; ++i) F_(j, i) = F_(i, j);
        for (int64_t k = 0; k < w; ++k) sreq[k] = psign ? psign[S->perm[c0 + k] - 1] : 0;
        int64_t *st = stats + s * NF;
        ldlt_front(F, f, w, &prm, psign ? sreq : NULL, lp, kind, dd, ee, st);
        for (int64_t k = 0; k < w; ++k) {
            int64_t g = c0 + k;
            piv[g] = (int32_t)(c0 + lp[k] + 1);
            pkind[g] = kind[k];
            d[g] = dd[k];
            d[n + g] = ee[k];
        }
        if (m > 0) {
            /* W = L21 D (ldlt.jl:206-231), then F22 <- F22 - W L21^T */
            for (int64_t k = 0; k < w;) {
                if (kind[k] == K_2X2_FIRST) {
                    double a = dd[k], b = ee[k], c = dd[k + 1];
                    for (int64_t i = 0; i < m; ++i) {
                        double x1 = F_(w + i, k), x2 = F_(w + i, k + 1);
                        W[k * m + i] = x1 * a + x2 * b;
                        W[(k + 1) * m + i] = x1 * b + x2 * c;
                    }
                    k += 2;
                } else {
                    for (int64_t i = 0; i < m; ++i) W[k * m + i] = F_(w + i, k) * dd[k];
                    k += 1;
                }
            }
            for (int64_t jj = 0; jj < m; ++jj)
                for (int64_t k = 0; k < w; ++k) {
                    double l = F_(w + jj, k);
                    for (int64_t ii = jj; ii < m; ++ii) F_(w + ii, w + jj) -= W[k * m + ii] * l;
                }
            keep_cb(S, s, F, &cbs);
        }
        for (int64_t j = 0; j < w; ++j) {
            factor[base + j * f + j] = 1.0;
            for (int64_t i = j + 1; i < f; ++i) factor[base + j * f + i] = F_(i, j);
        }
    }
    free(F);
    free(W);
    free(lp);
    free(kind);
    free(sreq);
    free(dd);
    free(ee);
    cb_free(&cbs, ns);
    /* the original column of the failed pivot with the smallest factor column (ldlt.jl:208-216) */
    int64_t best = 0;
    for (int64_t s = 0; s < ns; ++s) {
        int64_t fi = stats[s * NF + ST_INFO];
        if (fi > 0) {
            int64_t col = S->super_ptr[s] + fi - 1;
            best = best == 0 ? col : (col < best ? col : best);
        }
    }
    return best == 0 ? 0 : S->perm[best - 1];
}

/* ------------------------------------------------------------------------- */
/* LU with in-block threshold partial pivoting: ref_lu! (src/reference/lu.jl:84-155)
 * and _lu_front! (lu.jl:163-214) */

static int64_t lu_choose_row(const double *F, int64_t f, int64_t k, int64_t w, double u) {
    double g = 0.0;
    for (int64_t i = k; i < f; ++i) g = jmax(g, fabs(F_(i, k)));
    double akk = fabs(F_(k, k));
    if (akk >= u * g) return k;
    int64_t r = k;
    double best = akk;
    for (int64_t i = k + 1; i < w; ++i) {
        double a = fabs(F_(i, k));
        if (a > best) {
            r = i;
            best = a;
        }
    }
    return r;
}

/* dense L D U of the fully summed columns 0..w-1 of the full f x f front F with row
 * interchanges among the rows 0..w-1: _lu_front! (lu.jl:163-214). On return the strict lower
 * part of F[:, 0:w] is L and the strict upper part of F[0:w, :] is U (both unit), F[k,k] = D[k];
 * lp[k] = the assembled row eliminated at step k (0-based). The contribution block F22 is left
 * to the caller. */
static void lu_front(double *F, int64_t f, int64_t w, int pivoting, double u, double eps, int64_t *lp, int8_t *kind,
                     double *dd, int64_t *st) {
    for (int64_t k = 0; k < w; ++k) {
        lp[k] = k;
        kind[k] = K_1X1;
    }
    for (int q = 0; q < NF; ++q) st[q] = 0;
    for (int64_t k = 0; k < w; ++k) {
        int64_t r = pivoting ? lu_choose_row(F, f, k, w, u) : k;
        if (r != k) {
            for (int64_t j = 0; j < f; ++j) {
                double t = F_(k, j);
                F_(k, j) = F_(r, j);
                F_(r, j) = t;
            }
            int64_t t = lp[k];
            lp[k] = lp[r];
            lp[r] = t;
        }
        double dk = F_(k, k);
        if (!(fabs(dk) >= eps) || dk == 0.0) {
            if (dk == 0.0) st[ST_NZERO]++;
            dk = (dk < 0.0 ? -1.0 : 1.0) * eps;
            kind[k] = K_PERTURBED;
            st[ST_NPERT]++;
            if (dk == 0.0 && st[ST_INFO] == 0) st[ST_INFO] = k + 1;
        }
        F_(k, k) = dk;
        dd[k] = dk;
        for (int64_t j = k + 1; j < w; ++j)
            for (int64_t i = k + 1; i < f; ++i) F_(i, j) -= F_(i, k) / dk * F_(k, j);
        for (int64_t j = w; j < f; ++j)
            for (int64_t i = k + 1; i < w; ++i) F_(i, j) -= F_(i, k) / dk * F_(k, j);
        for (int64_t i = k + 1; i < f; ++i) F_(i, k) /= dk;
        for (int64_t j = k + 1; j < f; ++j) F_(k, j) /= dk;
    }
}

int64_t oracle_lu(const oracle_sym *S, const double *nz, int pivoting, double u, double eps, double *factor,
                  double *ufactor, double *d, int32_t *piv, int8_t *pkind, int64_t *stats) {
    int64_t ns = S->ns, n = S->n;
    memset(factor, 0, sizeof(double) * (size_t)S->factor_len);
    memset(ufactor, 0, sizeof(double) * (size_t)S->factor_len);
    memset(d, 0, sizeof(double) * (size_t)(2 * n));
    memset(stats, 0, sizeof(int64_t) * (size_t)(NF * ns));
    memset(piv, 0, sizeof(int32_t) * (size_t)n);
    memset(pkind, 0, (size_t)n);
    cbstore cbs;
    cb_init(&cbs, ns);
    int64_t fmax = 1, wmax = 1;
    for (int64_t s = 0; s < ns; ++s) {
        fmax = S->front_nrows[s] > fmax ? S->front_nrows[s] : fmax;
        wmax = S->front_ncols[s] > wmax ? S->front_ncols[s] : wmax;
    }
    double *F = (double *)malloc(sizeof(double) * (size_t)(fmax * fmax));
    int64_t *lp = (int64_t *)malloc(sizeof(int64_t) * (size_t)wmax);
    double *dd = (double *)malloc(sizeof(double) * (size_t)wmax);
    for (int64_t q = 0; q < ns; ++q) {
        int64_t s = S->snpost[q] - 1;
        int64_t f = S->front_nrows[s], w = S->front_ncols[s], m = f - w;
        int64_t c0 = S->super_ptr[s] - 1, base = S->front_ptr[s] - 1;
        assemble(S, s, nz, F, &cbs, 1, 1);
        int64_t *st = stats + s * NF;
        lu_front(F, f, w, pivoting, u, eps, lp, pkind + c0, dd, st);
        for (int64_t k = 0; k < w; ++k) {
            piv[c0 + k] = (int32_t)(c0 + lp[k] + 1);
            d[c0 + k] = dd[k];
        }
        if (m > 0) {
            /* F22 <- F22 - (L21 D) U12 (lu.jl:137-145) */
            for (int64_t jj = 0; jj < m; ++jj)
                for (int64_t ii = 0; ii < m; ++ii) {
                    double acc = 0.0;
                    for (int64_t k = 0; k < w; ++k) acc += F_(w + ii, k) * dd[k] * F_(k, w + jj);
                    F_(w + ii, w + jj) -= acc;
                }
            keep_cb(S, s, F, &cbs);
        }
        for (int64_t j = 0; j < w; ++j) {
            factor[base + j * f + j] = 1.0;
            ufactor[base + j * f + j] = 1.0;
            for (int64_t i = j + 1; i < f; ++i) {
                factor[base + j * f + i] = F_(i, j);
                ufactor[base + j * f + i] = F_(j, i);
            }
        }
    }
    free(F);
    free(lp);
    free(dd);
    cb_free(&cbs, ns);
    int64_t best = 0;
    for (int64_t s = 0; s < ns; ++s) {
        int64_t fi = stats[s * NF + ST_INFO];
        if (fi > 0) {
            int64_t col = S->super_ptr[s] + fi - 1;
            best = best == 0 ? col : (col < best ? col : best);
        }
    }
    return best == 0 ? 0 : S->perm[best - 1];
}

/* ------------------------------------------------------------------------- */
/* The dense in-front kernels on their own, for pinning the pivot rules on fronts
 * with contribution-block rows (f > w). F is f x f column-major; for LDL^T both
 * triangles of the leading w x w block and rows w..f-1 of columns 0..w-1 are
 * read. The per-column outputs have w entries; st has 6. */

void oracle_ldlt_front(double *F, int64_t f, int64_t w, int pivot, double u, double eps, const int8_t *sign_req,
                       int64_t *lp, int8_t *kind, double *dd, double *ee, int64_t *st) {
    ldlt_params prm = {pivot, u, eps};
    ldlt_front(F, f, w, &prm, sign_req, lp, kind, dd, ee, st);
}

void oracle_lu_front(double *F, int64_t f, int64_t w, int pivoting, double u, double eps, int64_t *lp, int8_t *kind,
                     double *dd, int64_t *st) {
    lu_front(F, f, w, pivoting, u, eps, lp, kind, dd, st);
}

/* ------------------------------------------------------------------------- */
/* Front-by-front pieces for a driver that does the dense work of each front with
 * LAPACK/BLAS, as ref_factorize! does (src/reference/cholesky.jl:44-111: potrf, trsm,
 * syrk per front): the context keeps the children's contribution blocks between the
 * calls. assemble: zero the f x f front F (column-major) and add A and the children's
 * blocks (cholesky.jl:73-92); keep: store the contribution block F22 + lower(C) of
 * front s (C = -F21 F21^T, m x m column-major) for its parent. */

typedef struct {
    oracle_sym S;
    cbstore cbs;
} oracle_front_ctx;

void *oracle_front_begin(const oracle_sym *S) {
    oracle_front_ctx *c = (oracle_front_ctx *)malloc(sizeof(oracle_front_ctx));
    c->S = *S;
    cb_init(&c->cbs, S->ns);
    return c;
}

void oracle_front_assemble(void *ctx, int64_t s, const double *nz, double *F) {
    oracle_front_ctx *c = (oracle_front_ctx *)ctx;
    assemble(&c->S, s, nz, F, &c->cbs, 0, 0);
}

void oracle_front_keep(void *ctx, int64_t s, const double *F, const double *C) {
    oracle_front_ctx *c = (oracle_front_ctx *)ctx;
    const oracle_sym *S = &c->S;
    int64_t f = S->front_nrows[s], w = S->front_ncols[s], m = f - w;
    if (m <= 0 || S->snparent[s] == 0) return;
    double *B = (double *)malloc(sizeof(double) * (size_t)(m * m));
    for (int64_t jj = 0; jj < m; ++jj)
        for (int64_t ii = 0; ii < m; ++ii)
            B[jj * m + ii] = ii >= jj ? F_(w + ii, w + jj) + C[jj * m + ii] : 0.0;
    c->cbs.cb[s] = B;
    c->cbs.m[s] = m;
}

void oracle_front_end(void *ctx) {
    oracle_front_ctx *c = (oracle_front_ctx *)ctx;
    cb_free(&c->cbs, c->S.ns);
    free(c);
}
