// LDL^T factorization kernels (gfx950) with diagonal and Bunch-Kaufman pivoting inside
// each front's fully summed block, threshold test against the whole remaining column,
// best-1x1 fallback and static perturbation (+-eps, optionally with requested signs).
//
// Every floating-point operation of the factorization is the one the CPU reference
// performs, in the same order, with no fused multiply-add (the file is compiled with
// contraction off): a front's pivot decisions, factor, D and contribution block are
// bitwise those of src/reference/ldlt.jl in FP64. The pivot search reductions
// (column maxima, argmax with the first index on ties) are order-independent, so they
// run as parallel reductions.
//
// Small fronts (f x f fits the LDS) are assembled, factored and their contribution
// block formed in one workgroup; large fronts are assembled in global memory by the
// shared column-chunk and extend-add kernels and factored on the panel in place (the full
// w x w fully summed block lives in the panel's top rows, the upper triangle being the
// mirror the reference keeps), one pivot per pair of launches: the pivot choice by one
// workgroup per front, its update by 64 x 64 tiles over the GPU; then their contribution
// block is updated by 64 x 64 tiles, each entry subtracting the terms W[i, k] L[j, k] in
// the order k = 0, 1, ... of the reference.
//
// Reference counterparts: ref_ldlt! and _ldlt_front! (src/reference/ldlt.jl:121-202,
// :384-454), _choose_pivot (:280-325), _best_1x1 (:264-276), the threshold test
// (:252-260), the 2x2 test (:240-248), W = L21 D and the Schur update (:206-231);
// device statistics src/numeric/ldlt.jl:1276-1356.
#pragma clang fp contract(off)
#include "k_front.h"
#include "sds_common.h"
#include "sds_kernels.h"

namespace sds {

enum { PV_NONE = 0, PV_DIAG = 1, PV_BK = 2 };
enum { ST_NPOS = 0, ST_NNEG = 1, ST_NZERO = 2, ST_NPERT = 3, ST_N2X2 = 4, ST_INFO = 5 };

// max that propagates NaN, as Julia's max
__device__ inline double jmax(double a, double b) {
    if (a != a || b != b) return a + b;
    return a > b ? a : b;
}

// (value, index) pairs of the argmax reductions: index < 0 = no candidate; a larger value
// wins, equal values keep the smaller index (the first one a sequential scan meets)
__device__ inline bool vi_better(double ov, int oi, double v, int i) {
    if (oi < 0) return false;
    if (i < 0) return true;
    return ov > v || (ov == v && oi < i);
}

template <int NT>
struct RedBuf {
    double v[NT / 64];
    int i[NT / 64];
};

__device__ inline double wave_jmax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = jmax(v, __shfl_xor(v, o));
    return v;
}

__device__ inline void wave_argmax(double& v, int& i) {
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o);
        const int oi = __shfl_xor(i, o);
        if (vi_better(ov, oi, v, i)) {
            v = ov;
            i = oi;
        }
    }
}

// block-wide reductions: every thread gets the same result
template <int NT>
__device__ double block_jmax(double v, RedBuf<NT>& rb) {
    v = wave_jmax(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) rb.v[wv] = v;
    __syncthreads();
    double r = rb.v[0];
    for (int q = 1; q < NT / 64; ++q) r = jmax(r, rb.v[q]);
    __syncthreads();
    return r;
}

template <int NT>
__device__ void block_argmax(double& v, int& i, RedBuf<NT>& rb) {
    wave_argmax(v, i);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        rb.v[wv] = v;
        rb.i[wv] = i;
    }
    __syncthreads();
    v = rb.v[0];
    i = rb.i[0];
    for (int q = 1; q < NT / 64; ++q)
        if (vi_better(rb.v[q], rb.i[q], v, i)) {
            v = rb.v[q];
            i = rb.i[q];
        }
    __syncthreads();
}

struct LdltPrm {
    int pivot;
    double u, eps;
};

// the front view of one factorization: F is f x f column-major (LDS or the panel), only
// columns < w are touched; lp is the local pivot order (original positions)
template <typename T, int NT>
struct Front {
    T* F;
    int f, w;
    int* lp;
    RedBuf<NT>* rb;
    double* skey = nullptr;  // LDS scratch of best_1x1_sorted (scap entries), or none
    int* sidx = nullptr;
    int scap = 0;
    __device__ double at(int i, int j) const { return (double)F[(int64_t)j * f + i]; }
    __device__ T& ref(int i, int j) const { return F[(int64_t)j * f + i]; }

    // largest |F[i, c]| over the rows k..f-1 but s1 and s2, NaN-propagating (ldlt.jl:240-260)
    __device__ double colmax(int c, int k, int s1, int s2) const {
        double m = 0.0;
        for (int i = k + threadIdx.x; i < f; i += NT)
            if (i != s1 && i != s2) m = jmax(m, fabs(at(i, c)));
        return block_jmax<NT>(m, *rb);
    }
    // d >= u max_{i in k..f-1, i != c} |F[i, c]| (the early exit of the reference is the same
    // predicate: the running maximum only grows)
    __device__ bool accept_1x1(int c, int k, const LdltPrm& p) const {
        const double d = fabs(at(c, c));
        if (!(d >= p.eps)) return false;
        return d >= p.u * colmax(c, k, c, -1);
    }
    __device__ bool nonsingular_2x2(int k, int r, const LdltPrm& p) const {
        const double det = at(k, k) * at(r, r) - at(k, r) * at(r, k);
        const double mx = jmax(jmax(fabs(at(k, k)), fabs(at(r, k))), fabs(at(r, r)));
        return det != 0.0 && isfinite(det) && mx >= p.eps;
    }
    __device__ bool accept_2x2(int k, int r, const LdltPrm& p) const {
        if (!nonsingular_2x2(k, r, p)) return false;
        const double det = at(k, k) * at(r, r) - at(k, r) * at(r, k);
        const double e11 = fabs(at(r, r) / det), e12 = fabs(at(k, r) / det), e21 = fabs(at(r, k) / det),
                     e22 = fabs(at(k, k) / det);
        const double m1 = colmax(k, k, k, r), m2 = colmax(r, k, r, k);
        return p.u * (e11 * m1 + e21 * m2) <= 1.0 && p.u * (e12 * m1 + e22 * m2) <= 1.0;
    }
    // acceptable 1x1 pivot of the block k..w-1 with the largest |a_jj|, first on ties; -1 if
    // none (ldlt.jl:264-276). One wave per candidate column; a column that cannot beat its
    // wave's current best is not tested (it cannot be the sequential scan's choice either).
    __device__ int best_1x1(int k, const LdltPrm& p) const {
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        double bv = 0.0;
        int bi = -1;
        for (int j = k + wv; j < w; j += NT / 64) {
            const double a = fabs(at(j, j));
            if (!(a >= p.eps)) continue;
            if (bi >= 0 && !(a > bv)) continue;
            double m = 0.0;
            for (int i = k + lane; i < f; i += 64)
                if (i != j) m = jmax(m, fabs(at(i, j)));
            m = wave_jmax(m);
            if (!(a >= p.u * m)) continue;
            bv = a;
            bi = j;
        }
        // each wave's candidate is wave-uniform; combine across waves
        if (lane == 0) {
            rb->v[wv] = bv;
            rb->i[wv] = bi;
        }
        __syncthreads();
        double v = rb->v[0];
        int i = rb->i[0];
        for (int q = 1; q < NT / 64; ++q)
            if (vi_better(rb->v[q], rb->i[q], v, i)) {
                v = rb->v[q];
                i = rb->i[q];
            }
        __syncthreads();
        return i;
    }
    // best_1x1 for a large front: the candidates sorted by |a_jj| (largest first, the smaller
    // index first on ties; those below eps last), then tested sixteen at a time in that order
    // (one wave per column, the column maximum over all rows below): the first acceptable one
    // in the order is the one the sequential scan of the reference picks. Usually the first
    // round decides, where the scan of every column reads (w - k) f entries.
    __device__ int best_1x1_sorted(int k, const LdltPrm& p) const {
        const int n = w - k, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
        int np = 1;
        while (np < n) np <<= 1;
        for (int q = tid; q < np; q += NT) {
            double key = -1.0;
            if (q < n) {
                const double a = fabs(at(k + q, k + q));
                key = a >= p.eps ? a : -1.0;
            }
            skey[q] = key;
            sidx[q] = k + q;
        }
        __syncthreads();
        // bitonic sort into the order "before(x, y)": key larger, or equal key and smaller index
        for (int size = 2; size <= np; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = tid; i < np; i += NT) {
                    const int l = i ^ stride;
                    if (l <= i) continue;
                    const bool lbefore = skey[l] > skey[i] || (skey[l] == skey[i] && sidx[l] < sidx[i]);
                    const bool ibefore = skey[i] > skey[l] || (skey[i] == skey[l] && sidx[i] < sidx[l]);
                    if ((i & size) == 0 ? lbefore : ibefore) {
                        const double tk = skey[i];
                        skey[i] = skey[l];
                        skey[l] = tk;
                        const int ti = sidx[i];
                        sidx[i] = sidx[l];
                        sidx[l] = ti;
                    }
                }
                __syncthreads();
            }
        for (int base = 0; base < n && skey[base] >= 0.0; base += NT / 64) {
            const int q = base + wv;
            int found = 0x7fffffff;
            if (q < n && skey[q] >= 0.0) {
                const int j = sidx[q];
                const double a = skey[q];
                double m = 0.0;
                for (int i = k + lane; i < f; i += 64)
                    if (i != j) m = jmax(m, fabs(at(i, j)));
                m = wave_jmax(m);
                if (a >= p.u * m) found = q;
            }
            if (lane == 0) rb->i[wv] = found;
            __syncthreads();
            int best = rb->i[0];
            for (int w2 = 1; w2 < NT / 64; ++w2) best = min(best, rb->i[w2]);
            __syncthreads();
            if (best != 0x7fffffff) return sidx[best];
        }
        return -1;
    }
    __device__ int best_1x1_any(int k, const LdltPrm& p) const {
        if (!skey || w - k > scap) return best_1x1(k, p);
        // the first candidate in the sorted order (largest |a_jj| >= eps, the smaller index on
        // ties) tested by the whole block first: it usually is the answer
        double v = -1.0;
        int bi = -1;
        for (int j = k + threadIdx.x; j < w; j += NT) {
            const double a = fabs(at(j, j));
            if (a >= p.eps && vi_better(a, j, v, bi)) {
                v = a;
                bi = j;
            }
        }
        block_argmax<NT>(v, bi, *rb);
        if (bi < 0) return -1;
        if (v >= p.u * colmax(bi, k, bi, -1)) return bi;
        return best_1x1_sorted(k, p);
    }
    // pivot choice at column k (ldlt.jl:280-325): c = pivot column, r = 2x2 partner or -1
    __device__ void choose(int k, const LdltPrm& p, int& c, int& r) const {
        r = -1;
        if (p.pivot == PV_NONE) {
            c = k;
            return;
        }
        if (p.pivot == PV_DIAG) {
            if (accept_1x1(k, k, p)) {
                c = k;
                return;
            }
            const int best = best_1x1_any(k, p);
            if (best >= 0) {
                c = best;
                return;
            }
            // the largest diagonal entry, first on ties; a NaN |a_kk| keeps k
            double v = fabs(at(k, k));
            int bi = k;
            for (int j = k + 1 + threadIdx.x; j < w; j += NT) {
                const double a = fabs(at(j, j));
                if (a > v) {
                    v = a;
                    bi = j;
                }
            }
            block_argmax<NT>(v, bi, *rb);
            c = fabs(at(bi, bi)) >= p.eps ? bi : k;
            return;
        }
        const double alpha = (1.0 + sqrt(17.0)) / 8.0;  // Bunch-Kaufman constant (ldlt.jl:22)
        double lambda = 0.0;
        int rr = -1;
        for (int i = k + 1 + threadIdx.x; i < w; i += NT) {
            const double a = fabs(at(i, k));
            if (a > lambda) {
                lambda = a;
                rr = i;
            }
        }
        block_argmax<NT>(lambda, rr, *rb);
        if (rr < 0) lambda = 0.0;
        const double akk = fabs(at(k, k));
        int bc = k, br = -1;
        if (rr < 0 || akk >= alpha * lambda) {
            bc = k;
        } else {
            double sigma = 0.0;
            for (int i = k + threadIdx.x; i < w; i += NT)
                if (i != rr) sigma = jmax(sigma, fabs(at(i, rr)));
            sigma = block_jmax<NT>(sigma, *rb);
            if (akk * sigma >= alpha * lambda * lambda)
                bc = k;
            else if (fabs(at(rr, rr)) >= alpha * sigma)
                bc = rr;
            else {
                bc = k;
                br = rr;
            }
        }
        const bool ok = br < 0 ? accept_1x1(bc, k, p) : accept_2x2(k, br, p);
        if (ok) {
            c = bc;
            r = br;
            return;
        }
        const int best = best_1x1_any(k, p);
        if (best >= 0) {
            c = best;
            return;
        }
        if (br >= 0 && nonsingular_2x2(k, br, p)) {
            c = bc;
            r = br;
            return;
        }
        c = bc;
    }
    // symmetric interchange of positions a and b: columns over all f rows, then rows over the
    // w fully summed columns (ldlt.jl:328-340)
    __device__ void swap(int a, int b) const {
        __syncthreads();  // the pivot choice's last reads of F are done
        if (a == b) return;
        for (int i = threadIdx.x; i < f; i += NT) {
            const T t = ref(i, a);
            ref(i, a) = ref(i, b);
            ref(i, b) = t;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < w; j += NT) {
            const T t = ref(a, j);
            ref(a, j) = ref(b, j);
            ref(b, j) = t;
        }
        if (threadIdx.x == 0) {
            const int t = lp[a];
            lp[a] = lp[b];
            lp[b] = t;
        }
        __syncthreads();
    }
};

// Per-front outputs of the pivot loop.
template <typename T>
struct PivOut {
    int8_t* kind;  // w
    T* dd;         // w: D (1x1 value, or the 2x2 diagonal entries)
    T* ee;         // w: 2x2 subdiagonal at the first position
    int* st;       // NF statistics (shared)
};

// The pivot loop of one front (_ldlt_front!, ldlt.jl:384-454) run by all NT threads.
// sreq (nullable): requested sign of each original local position (psign of its row).
template <typename T, int NT>
__device__ void ldlt_front(const Front<T, NT>& Fr, const LdltPrm& p, const int8_t* sreq_row,
                           const int32_t* perm_c0, PivOut<T> out, double* sh) {
    const int f = Fr.f, w = Fr.w, tid = threadIdx.x;
    int k = 0;
    while (k < w) {
        int c, r;
        Fr.choose(k, p, c, r);
        if (r < 0) {
            Fr.swap(k, c);
            if (tid == 0) {
                double d = Fr.at(k, k);
                int8_t kd = KIND_1X1;
                if (!(fabs(d) >= p.eps) || d == 0.0) {
                    if (d == 0.0) out.st[ST_NZERO]++;
                    const int req = sreq_row ? (int)sreq_row[perm_c0[Fr.lp[k]]] : 0;
                    const double sg = req != 0 ? (double)req : (d < 0.0 ? -1.0 : 1.0);
                    d = sg * p.eps;
                    kd = KIND_PERTURBED;
                    out.st[ST_NPERT]++;
                    if (d == 0.0 && out.st[ST_INFO] == 0) out.st[ST_INFO] = Fr.lp[k] + 1;
                }
                Fr.ref(k, k) = T(d);
                out.kind[k] = kd;
                out.dd[k] = T(d);
                out.ee[k] = T(0);
                if (d > 0.0)
                    out.st[ST_NPOS]++;
                else if (d < 0.0)
                    out.st[ST_NNEG]++;
                sh[0] = d;
            }
            __syncthreads();
            const T d = T(sh[0]);
            for (int i = k + 1 + tid; i < f; i += NT) Fr.ref(i, k) = Fr.ref(i, k) / d;
            __syncthreads();
            const int R = f - k - 1, C = w - k - 1;
            for (int q = tid; q < R * C; q += NT) {
                const int j = k + 1 + q / R, i = k + 1 + q % R;
                Fr.ref(i, j) = Fr.ref(i, j) - Fr.ref(i, k) * Fr.ref(k, j);
            }
            __syncthreads();
            k += 1;
        } else {
            Fr.swap(k + 1, r);
            if (tid == 0) {
                const T a = Fr.ref(k, k), b = Fr.ref(k + 1, k), c2 = Fr.ref(k + 1, k + 1);
                const T up = b;
                Fr.ref(k, k + 1) = up;
                const T det = a * c2 - up * b;
                sh[0] = (double)(c2 / det);
                sh[1] = (double)(-up / det);
                sh[2] = (double)(-b / det);
                sh[3] = (double)(a / det);
                out.kind[k] = KIND_2X2_FIRST;
                out.kind[k + 1] = KIND_2X2_SECOND;
                out.dd[k] = a;
                out.dd[k + 1] = c2;
                out.ee[k] = b;
                out.ee[k + 1] = T(0);
                out.st[ST_N2X2]++;
                if (det < T(0)) {
                    out.st[ST_NPOS]++;
                    out.st[ST_NNEG]++;
                } else if (a > T(0)) {
                    out.st[ST_NPOS] += 2;
                } else {
                    out.st[ST_NNEG] += 2;
                }
                Fr.ref(k + 1, k) = T(0);  // the reference clears it; nobody reads it below
            }
            __syncthreads();
            const T e11 = T(sh[0]), e12 = T(sh[1]), e21 = T(sh[2]), e22 = T(sh[3]);
            for (int i = k + 2 + tid; i < f; i += NT) {
                const T x1 = Fr.ref(i, k), x2 = Fr.ref(i, k + 1);
                Fr.ref(i, k) = x1 * e11 + x2 * e21;
                Fr.ref(i, k + 1) = x1 * e12 + x2 * e22;
            }
            __syncthreads();
            const int R = f - k - 2, C = w - k - 2;
            for (int q = tid; q < R * C; q += NT) {
                const int j = k + 2 + q / R, i = k + 2 + q % R;
                Fr.ref(i, j) = Fr.ref(i, j) - (Fr.ref(i, k) * Fr.ref(k, j) + Fr.ref(i, k + 1) * Fr.ref(k + 1, j));
            }
            __syncthreads();
            k += 2;
        }
    }
}

// W[i, k] = (L21 D)[i, k] for row i of L (column-major, ld f) given the pivot kinds and D
// (ldlt.jl:206-231); k is a factor column of the front
template <typename T>
__device__ inline T w_entry(const T* L, int64_t f, int i, int k, const int8_t* kind, const T* dd, const T* ee) {
    const int8_t kd = kind[k];
    if (kd == KIND_2X2_FIRST) {
        const T x1 = L[(int64_t)k * f + i], x2 = L[(int64_t)(k + 1) * f + i];
        return x1 * dd[k] + x2 * ee[k];
    }
    if (kd == KIND_2X2_SECOND) {
        const T x1 = L[(int64_t)(k - 1) * f + i], x2 = L[(int64_t)k * f + i];
        return x1 * ee[k - 1] + x2 * dd[k];
    }
    return L[(int64_t)k * f + i] * dd[k];
}

template <typename T>
__device__ inline double eps_eff(double eps, int scaled, const double* aux) {
    if (!scaled) return eps;
    const double a = aux[0];
    return a > 0 ? eps * a : eps;
}

// ---------------------------------------------------------------------------
// small fronts: one workgroup, the dense front in LDS

constexpr int LDLT_SMALL_W = 128;  // largest front of the small class (f <= this)

template <typename T>
__global__ void __launch_bounds__(256) k_front_small_ldlt(DevSym S, const int32_t* __restrict__ fronts,
                                                         const double* __restrict__ nz, T* __restrict__ factor,
                                                         T* __restrict__ cbpool, T* __restrict__ dvec,
                                                         int32_t* __restrict__ piv, int8_t* __restrict__ pkind,
                                                         int64_t* __restrict__ stats, LdltPrm prm, int scaled,
                                                         const double* __restrict__ aux,
                                                         const int8_t* __restrict__ psign) {
    constexpr int NT = 256;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    __shared__ RedBuf<NT> rb;
    __shared__ int lp[LDLT_SMALL_W];
    __shared__ int8_t kind[LDLT_SMALL_W];
    __shared__ T dd[LDLT_SMALL_W], ee[LDLT_SMALL_W];
    __shared__ int st[NF];
    __shared__ double sh[4];
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x;
    prm.eps = eps_eff<T>(prm.eps, scaled, aux);
    for (int k = tid; k < w; k += NT) {
        lp[k] = k;
        kind[k] = 0;
        dd[k] = ee[k] = T(0);
    }
    if (tid < NF) st[tid] = 0;
    assemble_dense<T>(S, s, nz, cbpool, F, f, 0);
    // the upper triangle of the fully summed block mirrors the lower one
    for (int q = tid; q < w * w; q += NT) {
        const int j = q / w, i = q - j * w;
        if (i > j) F[i * f + j] = F[j * f + i];
    }
    __syncthreads();
    Front<T, NT> Fr{F, f, w, lp, &rb};
    ldlt_front<T, NT>(Fr, prm, psign, S.perm + c0, PivOut<T>{kind, dd, ee, st}, sh);
    // W = L21 D, then the lower triangle of F22 <- F22 - W L21^T, k in order
    const int64_t co = S.cb_off[s];
    if (co >= 0) {
        T* C = cbpool + co;
        for (int q = tid; q < m * m; q += NT) {
            const int jj = q / m, ii = q - jj * m;
            if (ii < jj) continue;
            T acc = F[(w + jj) * f + w + ii];
            for (int k = 0; k < w; ++k) acc = acc - w_entry<T>(F, f, w + ii, k, kind, dd, ee) * F[k * f + w + jj];
            C[q] = acc;
        }
    }
    T* P = factor + S.panel_off[s];
    for (int q = tid; q < f * w; q += NT) {
        const int j = q / f, i = q - j * f;
        P[q] = i > j ? F[q] : (i == j ? T(1) : T(0));
    }
    for (int k = tid; k < w; k += NT) {
        const int g = c0 + k;
        piv[g] = c0 + lp[k];
        pkind[g] = kind[k];
        dvec[g] = dd[k];
        dvec[S.n + g] = ee[k];
    }
    if (tid < NF) stats[(int64_t)s * NF + tid] = st[tid];
}

// ---------------------------------------------------------------------------
// large fronts, one pivot per pair of launches over the whole GPU: k_ldlt_pivot (one
// workgroup per front) chooses the pivot of the front's next column, swaps, sets D and
// scales the pivot column(s); k_ldlt_update (64 x 64 tiles over every front of the level)
// applies that pivot's rank-1 or rank-2 update to the remaining fully summed columns (all
// rows). Each entry goes through the operations of ldlt_front above, in the same order. A
// 2x2 pivot takes two columns in one pair, so the number of pairs a front needs is only
// known on the device: the launch pairs run to the level's widest front and a front whose
// columns are done skips them. lstep[3 s ..]: the step's first column, its size (1, 2, or 0
// when the front is done) and the next column.

// launch pair t: the next pivot of every large front. task: (s)
template <typename T>
__global__ void __launch_bounds__(1024) k_ldlt_pivot(DevSym S, const Task* __restrict__ tasks, int t,
                                                    T* __restrict__ factor, T* __restrict__ dvec,
                                                    int32_t* __restrict__ piv, int8_t* __restrict__ pkind,
                                                    int64_t* __restrict__ stats, int32_t* __restrict__ lstep,
                                                    LdltPrm prm, int scaled, const double* __restrict__ aux,
                                                    const int8_t* __restrict__ psign) {
    constexpr int NT = 1024;
    constexpr int SCAP = 2048;  // candidates of the sorted best-1x1 scan
    __shared__ RedBuf<NT> rb;
    __shared__ double sh[4];
    __shared__ double skey[SCAP];
    __shared__ int sidx[SCAP];
    const int s = tasks[blockIdx.x].s;
    const int f = S.fnrows[s], w = S.fncols[s];
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x;
    prm.eps = eps_eff<T>(prm.eps, scaled, aux);
    T* P = factor + S.panel_off[s];
    int* lp = piv + c0;
    int8_t* kind = pkind + c0;
    T* dd = dvec + c0;
    T* ee = dvec + S.n + c0;
    int64_t* st = stats + (int64_t)s * NF;
    int32_t* ls = lstep + 3 * s;
    int k;
    if (t == 0) {
        for (int q = tid; q < w; q += NT) {
            lp[q] = q;
            kind[q] = 0;
            dd[q] = ee[q] = T(0);
        }
        if (tid < NF) st[tid] = 0;
        // (the upper triangle of the fully summed square, the reference's full square, was
        // mirrored from the lower one by k_ldlt_mirror)
        k = 0;
    } else {
        k = ls[2];
    }
    __syncthreads();
    if (k >= w) {
        if (tid == 0) ls[1] = 0;
        return;
    }
    Front<T, NT> Fr{P, f, w, lp, &rb, skey, sidx, SCAP};
    // The usual step in one pass over column k: its entries k..f-1 in registers (the loads
    // issued together), one combined reduction giving the Bunch-Kaufman partner search (the
    // largest |a_ik| over rows k+1..w-1, first index on ties) and the threshold test's column
    // maximum (rows k+1..f-1, NaN-propagating). When column k is its own acceptable 1x1 pivot
    // (the first outcome of choose: no partner needed, the threshold test passes) the step
    // takes it and scales the column from the registers; otherwise choose decides from the
    // panel with the same predicates.
    constexpr int CR = 8;  // column entries per thread held in registers
    __shared__ double smx[NT / 64];
    T x[CR];
    bool fast = false;
    if (f - k <= CR * NT) {
        const T* col = P + (int64_t)k * f;
#pragma unroll
        for (int u = 0; u < CR; ++u) x[u] = col[min(k + tid + NT * u, f - 1)];
        const double akk = fabs((double)col[k]);
        __builtin_amdgcn_sched_barrier(0);
        double lam = 0.0, m = 0.0;
        int rr = -1;
#pragma unroll
        for (int u = 0; u < CR; ++u) {
            const int i = k + tid + NT * u;
            if (i > k && i < f) {
                const double a = fabs((double)x[u]);
                m = jmax(m, a);
                if (i < w && a > lam) {
                    lam = a;
                    rr = i;
                }
            }
        }
        wave_argmax(lam, rr);
        m = wave_jmax(m);
        const int lane = tid & 63, wv = tid >> 6;
        if (lane == 0) {
            rb.v[wv] = lam;
            rb.i[wv] = rr;
            smx[wv] = m;
        }
        __syncthreads();
        lam = rb.v[0];
        rr = rb.i[0];
        m = smx[0];
        for (int q = 1; q < NT / 64; ++q) {
            if (vi_better(rb.v[q], rb.i[q], lam, rr)) {
                lam = rb.v[q];
                rr = rb.i[q];
            }
            m = jmax(m, smx[q]);
        }
        __syncthreads();
        if (rr < 0) lam = 0.0;
        if (prm.pivot == PV_NONE) {
            fast = true;
        } else {
            const double alpha = (1.0 + sqrt(17.0)) / 8.0;  // Bunch-Kaufman constant (ldlt.jl:22)
            const bool self = prm.pivot == PV_DIAG || rr < 0 || akk >= alpha * lam;
            fast = self && akk >= prm.eps && akk >= prm.u * m;
        }
    }
    int c = k, r = -1;
    if (!fast) Fr.choose(k, prm, c, r);
    const int8_t* sreq_row = psign;
    const int32_t* perm_c0 = S.perm + c0;
    if (r < 0) {
        Fr.swap(k, c);
        if (tid == 0) {
            // (thread 0 holds a_kk in x[0] on the one-pass step; the sign counts are added by
            // atomics whose result is not waited for)
            double d = fast ? (double)x[0] : Fr.at(k, k);
            int8_t kd = KIND_1X1;
            if (!(fabs(d) >= prm.eps) || d == 0.0) {
                if (d == 0.0) st[ST_NZERO]++;
                const int req = sreq_row ? (int)sreq_row[perm_c0[lp[k]]] : 0;
                const double sg = req != 0 ? (double)req : (d < 0.0 ? -1.0 : 1.0);
                d = sg * prm.eps;
                kd = KIND_PERTURBED;
                st[ST_NPERT]++;
                if (d == 0.0 && st[ST_INFO] == 0) st[ST_INFO] = lp[k] + 1;
            }
            Fr.ref(k, k) = T(d);
            kind[k] = kd;
            dd[k] = T(d);
            ee[k] = T(0);
            if (d > 0.0)
                atomicAdd(reinterpret_cast<unsigned long long*>(st + ST_NPOS), 1ull);
            else if (d < 0.0)
                atomicAdd(reinterpret_cast<unsigned long long*>(st + ST_NNEG), 1ull);
            sh[0] = d;
            ls[0] = k;
            ls[1] = 1;
            ls[2] = k + 1;
        }
        __syncthreads();
        const T d = T(sh[0]);
        if (fast) {
            T* col = P + (int64_t)k * f;
#pragma unroll
            for (int u = 0; u < CR; ++u) {
                const int i = k + tid + NT * u;
                if (i > k && i < f) col[i] = x[u] / d;
            }
        } else {
            for (int i = k + 1 + tid; i < f; i += NT) Fr.ref(i, k) = Fr.ref(i, k) / d;
        }
    } else {
        Fr.swap(k + 1, r);
        if (tid == 0) {
            const T a = Fr.ref(k, k), b = Fr.ref(k + 1, k), c2 = Fr.ref(k + 1, k + 1);
            const T up = b;
            Fr.ref(k, k + 1) = up;
            const T det = a * c2 - up * b;
            sh[0] = (double)(c2 / det);
            sh[1] = (double)(-up / det);
            sh[2] = (double)(-b / det);
            sh[3] = (double)(a / det);
            kind[k] = KIND_2X2_FIRST;
            kind[k + 1] = KIND_2X2_SECOND;
            dd[k] = a;
            dd[k + 1] = c2;
            ee[k] = b;
            ee[k + 1] = T(0);
            st[ST_N2X2]++;
            if (det < T(0)) {
                st[ST_NPOS]++;
                st[ST_NNEG]++;
            } else if (a > T(0)) {
                st[ST_NPOS] += 2;
            } else {
                st[ST_NNEG] += 2;
            }
            Fr.ref(k + 1, k) = T(0);
            ls[0] = k;
            ls[1] = 2;
            ls[2] = k + 2;
        }
        __syncthreads();
        const T e11 = T(sh[0]), e12 = T(sh[1]), e21 = T(sh[2]), e22 = T(sh[3]);
        for (int i = k + 2 + tid; i < f; i += NT) {
            const T x1 = Fr.ref(i, k), x2 = Fr.ref(i, k + 1);
            Fr.ref(i, k) = x1 * e11 + x2 * e21;
            Fr.ref(i, k + 1) = x1 * e12 + x2 * e22;
        }
    }
}

// before the pivot launches: the upper triangle of each large front's fully summed square
// mirrors the lower one (the reference's full square). task: (s, I0, J0), the 64 x 64 tile of
// rows I0.. and columns J0.. with I0 >= J0, written transposed through LDS (both sides
// coalesced)
template <typename T>
__global__ void __launch_bounds__(256) k_ldlt_mirror(DevSym S, const Task* __restrict__ tasks, T* __restrict__ factor) {
    __shared__ T tl[64][65];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    const int I0 = t.a, J0 = t.b, ni = min(64, w - I0), nj = min(64, w - J0);
    T* P = factor + S.panel_off[s];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    {
        T v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)
            v[q] = P[(int64_t)(J0 + min(wv + 4 * q, nj - 1)) * f + I0 + min(lane, ni - 1)];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 16; ++q) tl[wv + 4 * q][lane] = v[q];  // tl[j][i] = A(I0 + i, J0 + j)
    }
    __syncthreads();
    // A(J0 + j, I0 + i) = A(I0 + i, J0 + j) for I0 + i > J0 + j: lanes along j
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = wv + 4 * q, j = lane;
        if (i < ni && j < nj && I0 + i > J0 + j) P[(int64_t)(I0 + i) * f + J0 + j] = tl[j][i];
    }
}

// launch pair t: the update of each front's pivot of that pair on the fully summed columns
// after it (every row below). task: (s, I0, J0), a 64 x 64 tile, lanes along the rows
template <typename T>
__global__ void __launch_bounds__(256) k_ldlt_update(DevSym S, const Task* __restrict__ tasks, T* __restrict__ factor,
                                                     const int32_t* __restrict__ lstep) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    const int32_t* ls = lstep + 3 * s;
    const int sz = ls[1];
    if (sz == 0) return;
    const int k = ls[0], k2 = k + sz;
    const int i = t.a + (threadIdx.x & 63);
    if (i < k2 || i >= f) return;
    T* P = factor + S.panel_off[s];
    const T* ck = P + (int64_t)k * f;
    const int j1 = min(t.b + 64, w);
    if (sz == 1) {
        // the thread's sixteen columns: every load in flight together (clamped to the tile's
        // last column), then the updates and the stores; row k is not written here
        const int jb = max(t.b, k2) + (threadIdx.x >> 6);
        if (jb >= j1) return;
        const T lk = ck[i];
        T a[16], u[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t o = (int64_t)min(jb + 4 * q, j1 - 1) * f;
            a[q] = P[o + i];
            u[q] = P[o + k];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (jb + 4 * q < j1) P[(int64_t)(jb + 4 * q) * f + i] = a[q] - lk * u[q];
    } else {
        const T l1 = ck[i], l2 = ck[f + i];
        for (int j = max(t.b, k2) + (threadIdx.x >> 6); j < j1; j += 4) {
            T& a = P[(int64_t)j * f + i];
            a = a - (l1 * P[(int64_t)j * f + k] + l2 * P[(int64_t)j * f + k + 1]);
        }
    }
}

// ---------------------------------------------------------------------------
// large fronts with deferred updates: k_ldlt_panel (one workgroup per front) takes the pivots
// of a panel of at least LP_NB columns one after the other. The front's columns from the current
// pivot on keep, in the panel memory, their values of the panel's start; the columns the pivot
// choice reads are formed when it reads them, by applying the panel's pivots to them in the
// reference's order: x - L(i, p) U(p, j) for a 1x1 pivot p, x - (L(i, p) U(p, j) + L(i, p + 1)
// U(p + 1, j)) for a 2x2 one, pivot after pivot. U(p, j), the row of pivot p in column j at
// the time of pivot p (the upper mirror entry the reference's update reads), is formed once per
// pivot for every later column. k_ldlt_trail then applies the panel's pivots to the columns
// after it in the same way over the GPU. Every entry thus goes through the operations of
// ldlt_front in the same order, later: the factor is bitwise that of one pivot at a time. The
// symmetric interchanges act on the stored values, as in the reference; a row interchange of
// two positions at or after the current pivot commutes with the pending updates, since it
// exchanges the L rows of the panel's pivots too.

constexpr int LP_NB = 32;           // columns per panel (a 2x2 pivot at its end takes one more)
constexpr int LP_MAXP = LP_NB + 1;  // pivot rows of one panel

// scratch of one front during a panel (element offsets from its base): the current columns X0,
// X1, X2 (f entries each), the pivot rows U (LP_MAXP x w, U[q w + j] = U(k0 + q, j)) and the
// current diagonal DG (w)
__host__ __device__ inline int64_t lp_scratch(int f, int w) { return 3 * (int64_t)f + (int64_t)(LP_MAXP + 1) * w; }

template <typename T, int NT>
struct PanelFront {
    T* P;
    int f, w, k0;
    int* lp;
    RedBuf<NT>* rb;
    T *X0, *X1, *X2, *U, *DG;
    // LDS: per column q of the panel (k0 + q) the kind of its pivot step: 0 none yet, 1 a 1x1,
    // 2 the first and 3 the second column of a 2x2
    int* ctype;
    T* Uc;  // U(., j) of the column being formed
    T* Lr;  // L(r, .) of the pivot row whose U row is being formed
    double* skey;
    int* sidx;
    int scap;

    // x minus the panel's pivots in order, lv[q] = L(row, k0 + q) and u(q) = U(k0 + q, column)
    // (the operations and their order of ldlt_front's updates)
    template <class UF>
    __device__ __attribute__((always_inline)) T apply(T x, const T (&lv)[LP_MAXP], UF u) const {
#pragma unroll
        for (int q = 0; q < LP_MAXP; ++q) {
            const int ct = ctype[q];
            if (ct == 1) {
                x = x - lv[q] * u(q);
            } else if (ct == 2 && q + 1 < LP_MAXP) {
                x = x - (lv[q] * u(q) + lv[q + 1] * u(q + 1));
            }
        }
        return x;
    }
    // L(i, k0 .. k0 + np - 1) into registers, all loads issued before any is used
    __device__ __attribute__((always_inline)) void load_l(int i, int np, T (&lv)[LP_MAXP]) const {
#pragma unroll
        for (int q = 0; q < LP_MAXP; ++q) lv[q] = q < np ? P[(int64_t)(k0 + q) * f + i] : T(0);
    }
    // current values of column j at rows k..f-1 into X (all threads; ends with a barrier)
    __device__ __attribute__((always_inline)) void form_col(int j, int k, T* X) const {
        const int np = k - k0;
        for (int q = threadIdx.x; q < np; q += NT) Uc[q] = U[(int64_t)q * w + j];
        __syncthreads();
        const T* cj = P + (int64_t)j * f;
        for (int i = k + threadIdx.x; i < f; i += NT) {
            T lv[LP_MAXP];
            load_l(i, np, lv);
            X[i] = apply(cj[i], lv, [&](int q) { return Uc[q]; });
        }
        __syncthreads();
    }
    // current diagonal entries of the columns k..w-1 into DG (all threads; ends with a barrier)
    __device__ __attribute__((always_inline)) void form_diag(int k) const {
        const int np = k - k0;
        for (int j = k + threadIdx.x; j < w; j += NT) {
            T lv[LP_MAXP], uv[LP_MAXP];
            load_l(j, np, lv);
#pragma unroll
            for (int q = 0; q < LP_MAXP; ++q) uv[q] = q < np ? U[(int64_t)q * w + j] : T(0);
            DG[j] = apply(P[(int64_t)j * f + j], lv, [&](int q) { return uv[q]; });
        }
        __syncthreads();
    }
    // U row of the pivot row r (a position at or after the current pivot, before the current
    // pivot's own step is recorded) for the columns j0..w-1 (all threads; ends with a barrier)
    __device__ __attribute__((always_inline)) void form_urow(int r, int j0) const {
        const int np = r - k0;
        for (int q = threadIdx.x; q < LP_MAXP; q += NT) Lr[q] = q < np ? P[(int64_t)(k0 + q) * f + r] : T(0);
        __syncthreads();
        T lv[LP_MAXP];
#pragma unroll
        for (int q = 0; q < LP_MAXP; ++q) lv[q] = Lr[q];
        T* ur = U + (int64_t)(r - k0) * w;
        for (int j = j0 + threadIdx.x; j < w; j += NT) {
            T uv[LP_MAXP];
#pragma unroll
            for (int q = 0; q < LP_MAXP; ++q) uv[q] = q < np ? U[(int64_t)q * w + j] : T(0);
            ur[j] = apply(P[(int64_t)j * f + r], lv, [&](int q) { return uv[q]; });
        }
        __syncthreads();
    }
    // largest |X[i]| over the rows k..f-1 but s1 and s2, NaN-propagating (ldlt.jl:240-260)
    __device__ __attribute__((always_inline)) double colmax(const T* X, int k, int s1, int s2) const {
        double m = 0.0;
        for (int i = k + threadIdx.x; i < f; i += NT)
            if (i != s1 && i != s2) m = jmax(m, fabs((double)X[i]));
        return block_jmax<NT>(m, *rb);
    }
    // the same for the current column j, formed by one wave (lanes along the rows)
    __device__ __attribute__((always_inline)) double wave_colmax(int j, int k, int s1) const {
        const int lane = threadIdx.x & 63, np = k - k0;
        T uv[LP_MAXP];
#pragma unroll
        for (int q = 0; q < LP_MAXP; ++q) uv[q] = q < np ? U[(int64_t)q * w + j] : T(0);
        double m = 0.0;
        const T* cj = P + (int64_t)j * f;
        for (int i = k + lane; i < f; i += 64) {
            T lv[LP_MAXP];
            load_l(i, np, lv);
            const T x = apply(cj[i], lv, [&](int q) { return uv[q]; });
            if (i != s1) m = jmax(m, fabs((double)x));
        }
        return wave_jmax(m);
    }
    __device__ __attribute__((always_inline)) bool accept_1x1(const T* X, int c, int k, const LdltPrm& p) const {
        const double d = fabs((double)X[c]);
        if (!(d >= p.eps)) return false;
        return d >= p.u * colmax(X, k, c, -1);
    }
    // the 2x2 block (k, r) from the current columns k (Xk) and r (Xr)
    __device__ __attribute__((always_inline)) bool nonsingular_2x2(const T* Xk, const T* Xr, int k, int r, const LdltPrm& p) const {
        const double akk = Xk[k], arr = Xr[r], akr = Xr[k], ark = Xk[r];
        const double det = akk * arr - akr * ark;
        const double mx = jmax(jmax(fabs(akk), fabs(ark)), fabs(arr));
        return det != 0.0 && isfinite(det) && mx >= p.eps;
    }
    __device__ __attribute__((always_inline)) bool accept_2x2(const T* Xk, const T* Xr, int k, int r, const LdltPrm& p) const {
        if (!nonsingular_2x2(Xk, Xr, k, r, p)) return false;
        const double akk = Xk[k], arr = Xr[r], akr = Xr[k], ark = Xk[r];
        const double det = akk * arr - akr * ark;
        const double e11 = fabs(arr / det), e12 = fabs(akr / det), e21 = fabs(ark / det), e22 = fabs(akk / det);
        const double m1 = colmax(Xk, k, k, r), m2 = colmax(Xr, k, r, k);
        return p.u * (e11 * m1 + e21 * m2) <= 1.0 && p.u * (e12 * m1 + e22 * m2) <= 1.0;
    }
    // best_1x1 (ldlt.jl:264-276) on the current diagonal DG (formed here): the acceptable 1x1
    // pivot of k..w-1 with the largest |a_jj|, first on ties, -1 if none. The largest
    // candidate is tested first by the whole block (its column formed in X2, which holds it on
    // return); else the candidates in sorted order, sixteen at a time, one wave per column.
    __device__ __attribute__((always_inline)) int best_1x1(int k, const LdltPrm& p) const {
        form_diag(k);
        const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
        double v = -1.0;
        int bi = -1;
        for (int j = k + tid; j < w; j += NT) {
            const double a = fabs((double)DG[j]);
            if (a >= p.eps && vi_better(a, j, v, bi)) {
                v = a;
                bi = j;
            }
        }
        block_argmax<NT>(v, bi, *rb);
        if (bi < 0) return -1;
        form_col(bi, k, X2);
        if (v >= p.u * colmax(X2, k, bi, -1)) return bi;
        const int n = w - k;
        if (n <= scap) {
            int np = 1;
            while (np < n) np <<= 1;
            for (int q = tid; q < np; q += NT) {
                double key = -1.0;
                if (q < n) {
                    const double a = fabs((double)DG[k + q]);
                    key = a >= p.eps ? a : -1.0;
                }
                skey[q] = key;
                sidx[q] = k + q;
            }
            __syncthreads();
            for (int size = 2; size <= np; size <<= 1)
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int i = tid; i < np; i += NT) {
                        const int l = i ^ stride;
                        if (l <= i) continue;
                        const bool lbefore = skey[l] > skey[i] || (skey[l] == skey[i] && sidx[l] < sidx[i]);
                        const bool ibefore = skey[i] > skey[l] || (skey[i] == skey[l] && sidx[i] < sidx[l]);
                        if ((i & size) == 0 ? lbefore : ibefore) {
                            const double tk = skey[i];
                            skey[i] = skey[l];
                            skey[l] = tk;
                            const int ti = sidx[i];
                            sidx[i] = sidx[l];
                            sidx[l] = ti;
                        }
                    }
                    __syncthreads();
                }
            // the first entry is bi, already rejected
            for (int base = 1; base < n && skey[base] >= 0.0; base += NT / 64) {
                const int q = base + wv;
                int found = 0x7fffffff;
                if (q < n && skey[q] >= 0.0) {
                    const int j = sidx[q];
                    const double m = wave_colmax(j, k, j);
                    if (skey[q] >= p.u * m) found = q;
                }
                if (lane == 0) rb->i[wv] = found;
                __syncthreads();
                int best = rb->i[0];
                for (int w2 = 1; w2 < NT / 64; ++w2) best = min(best, rb->i[w2]);
                __syncthreads();
                if (best != 0x7fffffff) {
                    const int j = sidx[best];
                    form_col(j, k, X2);
                    return j;
                }
            }
            return -1;
        }
        // too many candidates to sort: every column, one wave each; a column that cannot beat
        // its wave's current best is not tested
        double bv = 0.0;
        int bj = -1;
        for (int j = k + wv; j < w; j += NT / 64) {
            const double a = fabs((double)DG[j]);
            if (!(a >= p.eps)) continue;
            if (bj >= 0 && !(a > bv)) continue;
            if (!(a >= p.u * wave_colmax(j, k, j))) continue;
            bv = a;
            bj = j;
        }
        if (lane == 0) {
            rb->v[wv] = bv;
            rb->i[wv] = bj;
        }
        __syncthreads();
        double rv = rb->v[0];
        int ri = rb->i[0];
        for (int q = 1; q < NT / 64; ++q)
            if (vi_better(rb->v[q], rb->i[q], rv, ri)) {
                rv = rb->v[q];
                ri = rb->i[q];
            }
        __syncthreads();
        if (ri >= 0) form_col(ri, k, X2);
        return ri;
    }
    // pivot choice at column k (ldlt.jl:280-325), as Front::choose on the current columns:
    // c = pivot column, r = 2x2 partner or -1; on return the current column c is in *Xc (1x1),
    // or the columns k and r are in X0 and X1 (2x2)
    __device__ __attribute__((always_inline)) void choose(int k, const LdltPrm& p, int& c, int& r, T*& Xc) const {
        r = -1;
        form_col(k, k, X0);
        Xc = X0;
        if (p.pivot == PV_NONE) {
            c = k;
            return;
        }
        if (p.pivot == PV_DIAG) {
            if (accept_1x1(X0, k, k, p)) {
                c = k;
                return;
            }
            const int best = best_1x1(k, p);
            if (best >= 0) {
                c = best;
                Xc = X2;
                return;
            }
            // the largest diagonal entry, first on ties; a NaN |a_kk| keeps k (DG was formed by
            // best_1x1)
            double v = fabs((double)X0[k]);
            int bi = k;
            for (int j = k + 1 + threadIdx.x; j < w; j += NT) {
                const double a = fabs((double)DG[j]);
                if (a > v) {
                    v = a;
                    bi = j;
                }
            }
            block_argmax<NT>(v, bi, *rb);
            c = fabs((double)DG[bi]) >= p.eps ? bi : k;
            if (c != k) {
                form_col(c, k, X2);
                Xc = X2;
            }
            return;
        }
        const double alpha = (1.0 + sqrt(17.0)) / 8.0;  // Bunch-Kaufman constant (ldlt.jl:22)
        double lambda = 0.0;
        int rr = -1;
        for (int i = k + 1 + threadIdx.x; i < w; i += NT) {
            const double a = fabs((double)X0[i]);
            if (a > lambda) {
                lambda = a;
                rr = i;
            }
        }
        block_argmax<NT>(lambda, rr, *rb);
        if (rr < 0) lambda = 0.0;
        const double akk = fabs((double)X0[k]);
        int bc = k, br = -1;
        T* Xb = X0;
        if (rr < 0 || akk >= alpha * lambda) {
            bc = k;
        } else {
            form_col(rr, k, X1);
            double sigma = 0.0;
            for (int i = k + threadIdx.x; i < w; i += NT)
                if (i != rr) sigma = jmax(sigma, fabs((double)X1[i]));
            sigma = block_jmax<NT>(sigma, *rb);
            if (akk * sigma >= alpha * lambda * lambda) {
                bc = k;
            } else if (fabs((double)X1[rr]) >= alpha * sigma) {
                bc = rr;
                Xb = X1;
            } else {
                bc = k;
                br = rr;
            }
        }
        const bool ok = br < 0 ? accept_1x1(Xb, bc, k, p) : accept_2x2(X0, X1, k, br, p);
        if (ok) {
            c = bc;
            r = br;
            Xc = Xb;
            return;
        }
        const int best = best_1x1(k, p);
        if (best >= 0) {
            c = best;
            Xc = X2;
            return;
        }
        if (br >= 0 && nonsingular_2x2(X0, X1, k, br, p)) {
            c = bc;
            r = br;
            return;
        }
        c = bc;
        Xc = Xb;
    }
    // symmetric interchange of positions a and b (ldlt.jl:328-340) on the stored values, the
    // pivot rows' entries of those columns, and the rows of the current columns held in Xa, Xb
    __device__ __attribute__((always_inline)) void swap(int a, int b, int k, T* Xa, T* Xb) const {
        __syncthreads();  // the pivot choice's last reads are done
        if (a == b) return;
        for (int i = threadIdx.x; i < f; i += NT) {
            const T t = P[(int64_t)a * f + i];
            P[(int64_t)a * f + i] = P[(int64_t)b * f + i];
            P[(int64_t)b * f + i] = t;
        }
        for (int q = threadIdx.x; q < k - k0; q += NT) {
            const T t = U[(int64_t)q * w + a];
            U[(int64_t)q * w + a] = U[(int64_t)q * w + b];
            U[(int64_t)q * w + b] = t;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < w; j += NT) {
            const T t = P[(int64_t)j * f + a];
            P[(int64_t)j * f + a] = P[(int64_t)j * f + b];
            P[(int64_t)j * f + b] = t;
        }
        if (threadIdx.x == 0) {
            const int t = lp[a];
            lp[a] = lp[b];
            lp[b] = t;
            if (Xa) {
                const T x = Xa[a];
                Xa[a] = Xa[b];
                Xa[b] = x;
            }
            if (Xb) {
                const T x = Xb[a];
                Xb[a] = Xb[b];
                Xb[b] = x;
            }
        }
        __syncthreads();
    }
};

// panel e of every large front of a level: its pivots until LP_NB columns are taken or the
// front is done (the first panel also sets up the front). task: (s, scratch offset)
template <typename T>
__global__ void __launch_bounds__(512) k_ldlt_panel(DevSym S, const Task* __restrict__ tasks, int e,
                                                    T* __restrict__ factor, T* __restrict__ dvec,
                                                    int32_t* __restrict__ piv, int8_t* __restrict__ pkind,
                                                    int64_t* __restrict__ stats, int32_t* __restrict__ lstep,
                                                    T* __restrict__ scratch, LdltPrm prm, int scaled,
                                                    const double* __restrict__ aux, const int8_t* __restrict__ psign) {
    constexpr int NT = 512;
    constexpr int SCAP = 2048;
    __shared__ RedBuf<NT> rb;
    __shared__ double sh[4];
    __shared__ double skey[SCAP];
    __shared__ int sidx[SCAP];
    __shared__ int ctype[LP_MAXP];
    __shared__ T Uc[LP_MAXP], Lr[LP_MAXP];
    const Task tk = tasks[blockIdx.x];
    const int s = tk.s;
    const int f = S.fnrows[s], w = S.fncols[s];
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x;
    prm.eps = eps_eff<T>(prm.eps, scaled, aux);
    T* P = factor + S.panel_off[s];
    int* lp = piv + c0;
    int8_t* kind = pkind + c0;
    T* dd = dvec + c0;
    T* ee = dvec + S.n + c0;
    int64_t* st = stats + (int64_t)s * NF;
    int32_t* ls = lstep + 3 * s;
    int k;
    if (e == 0) {
        for (int q = tid; q < w; q += NT) {
            lp[q] = q;
            kind[q] = 0;
            dd[q] = ee[q] = T(0);
        }
        if (tid < NF) st[tid] = 0;
        // the upper triangle of the fully summed square mirrors the lower (the reference's
        // full square)
        for (int q = tid; q < w * w; q += NT) {
            const int j = q / w, i = q - j * w;
            if (i > j) P[(int64_t)i * f + j] = P[(int64_t)j * f + i];
        }
        k = 0;
    } else {
        k = ls[2];
    }
    __syncthreads();
    if (k >= w) {
        if (tid == 0) ls[1] = 0;
        return;
    }
    const int k0 = k;
    for (int q = tid; q < LP_MAXP; q += NT) ctype[q] = 0;
    T* base = scratch + tk.a;
    PanelFront<T, NT> Z{P, f, w, k0, lp, &rb, base, base + f, base + 2 * (int64_t)f, base + 3 * (int64_t)f,
                        base + 3 * (int64_t)f + (int64_t)LP_MAXP * w, ctype, Uc, Lr, skey, sidx, SCAP};
    const int8_t* sreq_row = psign;
    const int32_t* perm_c0 = S.perm + c0;
    __syncthreads();
    while (k < w && k - k0 < LP_NB) {
        int c, r;
        T* Xc;
        Z.choose(k, prm, c, r, Xc);
        if (r < 0) {
            Z.swap(k, c, k, Xc, nullptr);
            if (tid == 0) {
                double d = (double)Xc[k];
                int8_t kd = KIND_1X1;
                if (!(fabs(d) >= prm.eps) || d == 0.0) {
                    if (d == 0.0) st[ST_NZERO]++;
                    const int req = sreq_row ? (int)sreq_row[perm_c0[lp[k]]] : 0;
                    const double sg = req != 0 ? (double)req : (d < 0.0 ? -1.0 : 1.0);
                    d = sg * prm.eps;
                    kd = KIND_PERTURBED;
                    st[ST_NPERT]++;
                    if (d == 0.0 && st[ST_INFO] == 0) st[ST_INFO] = lp[k] + 1;
                }
                P[(int64_t)k * f + k] = T(d);
                kind[k] = kd;
                dd[k] = T(d);
                ee[k] = T(0);
                if (d > 0.0)
                    st[ST_NPOS]++;
                else if (d < 0.0)
                    st[ST_NNEG]++;
                sh[0] = d;
            }
            __syncthreads();
            const T d = T(sh[0]);
            for (int i = k + 1 + tid; i < f; i += NT) P[(int64_t)k * f + i] = Xc[i] / d;
            __syncthreads();
            Z.form_urow(k, k + 1);
            if (tid == 0) ctype[k - k0] = 1;
            __syncthreads();
            k += 1;
        } else {
            Z.swap(k + 1, r, k, Z.X0, Z.X1);
            if (tid == 0) {
                const T a = Z.X0[k], b = Z.X0[k + 1], c2 = Z.X1[k + 1];
                const T up = b;
                const T det = a * c2 - up * b;
                sh[0] = (double)(c2 / det);
                sh[1] = (double)(-up / det);
                sh[2] = (double)(-b / det);
                sh[3] = (double)(a / det);
                kind[k] = KIND_2X2_FIRST;
                kind[k + 1] = KIND_2X2_SECOND;
                dd[k] = a;
                dd[k + 1] = c2;
                ee[k] = b;
                ee[k + 1] = T(0);
                st[ST_N2X2]++;
                if (det < T(0)) {
                    st[ST_NPOS]++;
                    st[ST_NNEG]++;
                } else if (a > T(0)) {
                    st[ST_NPOS] += 2;
                } else {
                    st[ST_NNEG] += 2;
                }
                P[(int64_t)k * f + k] = a;
                P[(int64_t)(k + 1) * f + k + 1] = c2;
                P[(int64_t)k * f + k + 1] = T(0);  // the reference clears it
            }
            __syncthreads();
            const T e11 = T(sh[0]), e12 = T(sh[1]), e21 = T(sh[2]), e22 = T(sh[3]);
            for (int i = k + 2 + tid; i < f; i += NT) {
                const T x1 = Z.X0[i], x2 = Z.X1[i];
                P[(int64_t)k * f + i] = x1 * e11 + x2 * e21;
                P[(int64_t)(k + 1) * f + i] = x1 * e12 + x2 * e22;
            }
            __syncthreads();
            Z.form_urow(k, k + 2);
            Z.form_urow(k + 1, k + 2);
            if (tid == 0) {
                ctype[k - k0] = 2;
                ctype[k + 1 - k0] = 3;
            }
            __syncthreads();
            k += 2;
        }
    }
    if (tid == 0) {
        ls[0] = k0;
        ls[1] = k - k0;
        ls[2] = k;
    }
}

// the update of panel e's pivots on the columns after it: the tile rows [I0, I0 + 64) x columns
// [J0, J0 + 64) of the entries at or after the panel's end (every row below, the fully summed
// square's upper mirror included), each entry applying the panel's pivots in order as ldlt_front
// does. task: (s, I0, J0, scratch offset)
template <typename T>
__global__ void __launch_bounds__(256) k_ldlt_trail(DevSym S, const Task* __restrict__ tasks, T* __restrict__ factor,
                                                   const int8_t* __restrict__ pkind, const int32_t* __restrict__ lstep,
                                                   const T* __restrict__ scratch) {
    __shared__ T Ls[LP_MAXP][TILE];
    __shared__ T Us[LP_MAXP][TILE];
    __shared__ int stp[LP_MAXP], ssz[LP_MAXP], nsh;
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    const int32_t* ls = lstep + 3 * s;
    const int np = ls[1];
    if (np == 0) return;
    const int k0 = ls[0], kend = ls[2];
    const int I0 = t.a, J0 = t.b;
    if (I0 + TILE <= kend || J0 + TILE <= kend || I0 >= f || J0 >= w) return;
    const T* P = factor + S.panel_off[s];
    T* Pw = factor + S.panel_off[s];
    const int8_t* kind = pkind + S.super_ptr[s];
    const T* U = scratch + t.c + 3 * (int64_t)f;
    const int tid = threadIdx.x;
    for (int q = tid; q < np * TILE; q += 256) {
        const int pp = q / TILE, c = q - pp * TILE;
        const int i = I0 + c, j = J0 + c;
        Ls[pp][c] = i < f ? P[(int64_t)(k0 + pp) * f + i] : T(0);
        Us[pp][c] = j < w ? U[(int64_t)pp * w + j] : T(0);
    }
    if (tid == 0) {
        int n = 0;
        for (int p = k0; p < kend; ++p) {
            if (kind[p] == KIND_2X2_SECOND) continue;
            stp[n] = p - k0;
            ssz[n] = kind[p] == KIND_2X2_FIRST ? 2 : 1;
            ++n;
        }
        nsh = n;
    }
    __syncthreads();
    const int ns = nsh;
    const int tx = tid & 15, ty = tid >> 4;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int r = 4 * ty + a, i = I0 + r;
        if (i < kend || i >= f) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int c = 4 * tx + b, j = J0 + c;
            if (j < kend || j >= w) continue;
            T x = P[(int64_t)j * f + i];
            for (int q = 0; q < ns; ++q) {
                const int pp = stp[q];
                if (ssz[q] == 1)
                    x = x - Ls[pp][r] * Us[pp][c];
                else
                    x = x - (Ls[pp][r] * Us[pp][c] + Ls[pp + 1][r] * Us[pp + 1][c]);
            }
            Pw[(int64_t)j * f + i] = x;
        }
    }
}

// after the last pair: unit diagonal, empty upper triangle and the global pivot order.
// task: (s)
template <typename T>
__global__ void __launch_bounds__(1024) k_ldlt_finish(DevSym S, const Task* __restrict__ tasks,
                                                     T* __restrict__ factor, int32_t* __restrict__ piv) {
    constexpr int NT = 1024;
    const int s = tasks[blockIdx.x].s;
    const int f = S.fnrows[s], w = S.fncols[s], c0 = S.super_ptr[s], tid = threadIdx.x;
    T* P = factor + S.panel_off[s];
    for (int q = tid; q < w * w; q += NT) {
        const int j = q / w, i = q - j * w;
        if (i <= j) P[(int64_t)j * f + i] = i == j ? T(1) : T(0);
    }
    for (int q = tid; q < w; q += NT) piv[c0 + q] = c0 + piv[c0 + q];
}

// Contribution block of a large front: the 64 x 64 tile (I0, J0) (front rows, I0 >= J0 >= w)
// of the lower triangle of F22 <- F22 - W L21^T, each entry subtracting W[i, k] L[j, k] for
// k = 0, 1, ..., w - 1 in order. Chunks of 16 columns of W and L21 are staged in LDS.
// Thread (ty, tx) owns rows I0 + 4 ty + a and columns J0 + 4 tx + b (a, b < 4).
template <typename T>
__global__ void __launch_bounds__(256) k_ldlt_cb(DevSym S, const Task* __restrict__ tasks,
                                                const T* __restrict__ factor, T* __restrict__ cbpool,
                                                const T* __restrict__ dvec, const int8_t* __restrict__ pkind) {
    constexpr int KC = 16;
    __shared__ T Ws[KC][TILE];
    __shared__ T Ls[KC][TILE];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int I0 = t.a, J0 = t.b;
    const int c0 = S.super_ptr[s];
    const T* L = factor + S.panel_off[s];
    T* C = cbpool + S.cb_off[s];
    const int8_t* kind = pkind + c0;
    const T* dd = dvec + c0;
    const T* ee = dvec + S.n + c0;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = min(I0 + 4 * ty + a, f - 1), j = min(J0 + 4 * tx + b, f - 1);
            acc[a][b] = C[(int64_t)(j - w) * m + (i - w)];  // (clamped; entries j > i are not stored)
        }
    // the thread's four (kk, c) staging slots, kk = (tid >> 6) + 4 u, c = tid & 63: the raw
    // values of chunk k0 + KC (the L entries W(i, k) may combine, D, the pivot kinds, L(j, k))
    // are loaded, clamped, while chunk k0 is multiplied; W(i, k) is formed as w_entry forms it
    const int c = tid & 63, ic = min(I0 + c, f - 1), jc = min(J0 + c, f - 1);
    T l0[4], lm[4], lq[4], d0[4], e0[4], em[4], lj[4];
    int kd[4];
    auto load_raw = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kc = min(k0 + (tid >> 6) + 4 * u, w - 1), km = max(kc - 1, 0), kq = min(kc + 1, w - 1);
            kd[u] = kind[kc];
            l0[u] = L[(int64_t)kc * f + ic];
            lm[u] = L[(int64_t)km * f + ic];
            lq[u] = L[(int64_t)kq * f + ic];
            d0[u] = dd[kc];
            e0[u] = ee[kc];
            em[u] = ee[km];
            lj[u] = L[(int64_t)kc * f + jc];
        }
    };
    load_raw(0);
    for (int k0 = 0; k0 < w; k0 += KC) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kk = (tid >> 6) + 4 * u, k = k0 + kk;
            const T wv = kd[u] == KIND_2X2_FIRST    ? l0[u] * d0[u] + lq[u] * e0[u]
                         : kd[u] == KIND_2X2_SECOND ? lm[u] * em[u] + l0[u] * d0[u]
                                                    : l0[u] * d0[u];
            Ws[kk][c] = (k < w && I0 + c < f) ? wv : T(0);
            Ls[kk][c] = (k < w && J0 + c < f) ? lj[u] : T(0);
        }
        __syncthreads();
        if (k0 + KC < w) load_raw(k0 + KC);
        const int kn = min(KC, w - k0);
        for (int kk = 0; kk < kn; ++kk) {
            T wv[4], lv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) wv[a] = Ws[kk][4 * ty + a];
#pragma unroll
            for (int b = 0; b < 4; ++b) lv[b] = Ls[kk][4 * tx + b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = acc[a][b] - wv[a] * lv[b];
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = I0 + 4 * ty + a, j = J0 + 4 * tx + b;
            if (i < f && j <= i) C[(int64_t)(j - w) * m + (i - w)] = acc[a][b];
        }
}

// aux[0] = max |nzval| (the scale of pivot_epsilon_alg = "algo1"): non-negative doubles
// order like their bit patterns, so an integer atomic max is exact; a NaN wins, as the
// reference's NaN-propagating max. aux[0] must be zero on entry.
__global__ void k_absmax(int64_t nnz, const double* __restrict__ nz, unsigned long long* __restrict__ aux) {
    unsigned long long m = 0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long b = __double_as_longlong(fabs(nz[q]));
        m = b > m ? b : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(m, o);
        m = x > m ? x : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(aux, m);
}

// info[0] = the smallest failed factor column (1-based; 0 = none) over the fronts'
// statistics (ldlt.jl:208-216)
__global__ void __launch_bounds__(256) k_info_from_stats(DevSym S, const int64_t* __restrict__ stats,
                                                        int64_t* __restrict__ part) {
    __shared__ int64_t best[256];
    int64_t mn = INT64_MAX;
    for (int64_t s = blockIdx.x * 256 + threadIdx.x; s < S.ns; s += (int64_t)gridDim.x * 256) {
        const int64_t fi = stats[s * NF + ST_INFO];
        if (fi > 0) mn = min(mn, (int64_t)S.super_ptr[s] + fi);
    }
    best[threadIdx.x] = mn;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) best[threadIdx.x] = min(best[threadIdx.x], best[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = best[0];
}

// ---------------------------------------------------------------------------

template <typename T>
void launch_front_small_ldlt(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, int fmax,
                             const double* nz, T* factor, T* cbpool, T* dvec, int32_t* piv, int8_t* kind,
                             int64_t* stats, int pivot, double u, double eps, int scaled, const double* aux,
                             const int8_t* psign) {
    size_t lds = (size_t)fmax * fmax * sizeof(T);
    LdltPrm p{pivot, u, eps};
    hipLaunchKernelGGL(k_front_small_ldlt<T>, dim3(nfronts), dim3(256), lds, st, S, fronts, nz, factor, cbpool, dvec,
                       piv, kind, stats, p, scaled, aux, psign);
}

template <typename T>
void launch_ldlt_pivot(hipStream_t st, const DevSym& S, const Task* tasks, int n, int t, T* factor, T* dvec,
                       int32_t* piv, int8_t* kind, int64_t* stats, int32_t* lstep, int pivot, double u, double eps,
                       int scaled, const double* aux, const int8_t* psign) {
    LdltPrm p{pivot, u, eps};
    hipLaunchKernelGGL(k_ldlt_pivot<T>, dim3(n), dim3(1024), 0, st, S, tasks, t, factor, dvec, piv, kind, stats,
                       lstep, p, scaled, aux, psign);
}

template <typename T>
void launch_ldlt_panel(hipStream_t st, const DevSym& S, const Task* tasks, int n, int e, T* factor, T* dvec,
                       int32_t* piv, int8_t* kind, int64_t* stats, int32_t* lstep, T* scratch, int pivot, double u,
                       double eps, int scaled, const double* aux, const int8_t* psign) {
    LdltPrm p{pivot, u, eps};
    hipLaunchKernelGGL(k_ldlt_panel<T>, dim3(n), dim3(512), 0, st, S, tasks, e, factor, dvec, piv, kind, stats,
                       lstep, scratch, p, scaled, aux, psign);
}

template <typename T>
void launch_ldlt_trail(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, const int8_t* kind,
                       const int32_t* lstep, const T* scratch) {
    hipLaunchKernelGGL(k_ldlt_trail<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, kind, lstep, scratch);
}

int64_t ldlt_panel_scratch(int f, int w) { return lp_scratch(f, w); }
int ldlt_panel_cols() { return LP_NB; }

template <typename T>
void launch_ldlt_mirror(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor) {
    hipLaunchKernelGGL(k_ldlt_mirror<T>, dim3(n), dim3(256), 0, st, S, tasks, factor);
}

template <typename T>
void launch_ldlt_update(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, const int32_t* lstep) {
    hipLaunchKernelGGL(k_ldlt_update<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, lstep);
}

template <typename T>
void launch_ldlt_finish(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, int32_t* piv) {
    hipLaunchKernelGGL(k_ldlt_finish<T>, dim3(n), dim3(1024), 0, st, S, tasks, factor, piv);
}

template <typename T>
void launch_ldlt_cb(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, T* cbpool,
                    const T* dvec, const int8_t* kind) {
    hipLaunchKernelGGL(k_ldlt_cb<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, dvec, kind);
}

void launch_absmax(hipStream_t st, int64_t nnz, const double* nz, double* aux) {
    hipLaunchKernelGGL(k_absmax, dim3((unsigned)std::min<int64_t>(std::max<int64_t>((nnz + 255) / 256, 1), 1024)),
                       dim3(256), 0, st, nnz, nz, reinterpret_cast<unsigned long long*>(aux));
}

void launch_info_from_stats(hipStream_t st, const DevSym& S, const int64_t* stats, int64_t* info) {
    // block minima at info + 8, then the smallest into info[0]
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(INFO_BLOCKS, (S.ns + 255) / 256));
    hipLaunchKernelGGL(k_info_from_stats, dim3(blocks), dim3(256), 0, st, S, stats, info + 8);
    launch_info_final(st, blocks, info);
}

#define INST(T)                                                                                                    \
    template void launch_front_small_ldlt<T>(hipStream_t, const DevSym&, const int32_t*, int, int, const double*, \
                                             T*, T*, T*, int32_t*, int8_t*, int64_t*, int, double, double, int,   \
                                             const double*, const int8_t*);                                        \
    template void launch_ldlt_cb<T>(hipStream_t, const DevSym&, const Task*, int, const T*, T*, const T*,         \
                                    const int8_t*);                                                                \
    template void launch_ldlt_pivot<T>(hipStream_t, const DevSym&, const Task*, int, int, T*, T*, int32_t*,        \
                                       int8_t*, int64_t*, int32_t*, int, double, double, int, const double*,       \
                                       const int8_t*);                                                             \
    template void launch_ldlt_update<T>(hipStream_t, const DevSym&, const Task*, int, T*, const int32_t*);        \
    template void launch_ldlt_mirror<T>(hipStream_t, const DevSym&, const Task*, int, T*);                        \
    template void launch_ldlt_panel<T>(hipStream_t, const DevSym&, const Task*, int, int, T*, T*, int32_t*,        \
                                       int8_t*, int64_t*, int32_t*, T*, int, double, double, int, const double*,   \
                                       const int8_t*);                                                             \
    template void launch_ldlt_trail<T>(hipStream_t, const DevSym&, const Task*, int, T*, const int8_t*,           \
                                       const int32_t*, const T*);                                                  \
    template void launch_ldlt_finish<T>(hipStream_t, const DevSym&, const Task*, int, T*, int32_t*);
INST(double)
INST(float)
#undef INST

}  // namespace sds
