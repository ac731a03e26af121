// Shared device-side definitions of libsds_hip: the symbolic maps as the
// kernels see them (0-based, converted once at upload from the 1-based arrays
// of the analysis), task records of the level-batched launches, and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sds {

// Device view of the analysis (src/symbolic/maps.jl:62-95), 0-based.
struct DevSym {
    int32_t n, ns;
    const int32_t* perm;        // n: original row of factor column k
    const int32_t* super_ptr;   // ns + 1
    const int32_t* fnrows;      // ns: f
    const int32_t* fncols;      // ns: w
    const int64_t* panel_off;   // ns: element offset of the f x w panel in the factor buffer
    const int64_t* rowptr;      // ns + 1: rows of supernode s are rowval[rowptr[s] .. rowptr[s+1])
    const int32_t* rowval;      // supernodal row numbers
    const int32_t* child_ptr;   // ns + 1
    const int32_t* child_list;  // children in increasing id (owner-pull order)
    const int64_t* relind_ptr;  // ns + 1
    const int32_t* relind;      // positions of a child's CB rows among its parent's rows
    const int64_t* asm_ptr;     // ns + 1: entries of A assembled into front s
    const int64_t* asm_src;     // position of the entry in nzval
    const int64_t* asm_dst;     // local panel offset j f + i; -(offset + 1) for a mirrored (U^T) entry
    const int64_t* acol_ptr;    // n + 1: the entries of global factor column j start at acol_ptr[j]
    const int64_t* cb_off;      // ns: contribution block (m x m, ld m) in the CB pool, -1 = none
    const int64_t* uv_off;      // ns: solve update vector (m entries per right-hand side), -1 = none
    // Cholesky sweeps, the forward pull map: row i of front s has the update-vector rows
    // pm_src[pm_ptr[pm_row[s] + i] .. pm_ptr[pm_row[s] + i + 1]) of its children (child_list order)
    const int64_t* pm_row;      // ns: first map row of the front
    const int32_t* pm_ptr;      // sum of f + 1
    const int32_t* pm_src;      // uv_off[c] + ii for the child rows ii that land there
    int32_t unsym;              // structure "G": full fronts, L and U^T panels
};

// One task of a level-batched launch.
struct Task {
    int32_t s;  // front
    int32_t a, b, c;
};

// pivot kinds (src/numeric/storage.jl:182-185)
enum { KIND_1X1 = 1, KIND_2X2_FIRST = 2, KIND_2X2_SECOND = 3, KIND_PERTURBED = 4 };

constexpr int NB = 64;     // pivot block of the blocked front factorization
// widest small front whose diagonal-block inverse its factorization kernel forms (wider ones: the
// per-block inverse kernel, whose register footprint would cost the front kernels their occupancy)
constexpr int FUSED_INV_MAX = 16;
constexpr int INFO_BLOCKS = 256;
constexpr int TINY_F = 32;  // fronts of at most this many rows: one wave each (Cholesky)  // block minima of the info reduction (after the 8 info words)
constexpr int TILE = 64;   // Schur-update tile
// flag in field c of a panel update tile (c = the first column of K): K spans a whole panel
// block (the block width is an argument of the launch) instead of one NB-column step
constexpr int32_t BIGK = 1 << 30;
constexpr int PANEL_ROWS = 256;  // panel rows per workgroup of the fused block step
constexpr int EA_COLS = 8;      // child columns per extend-add task
constexpr int EA_ROWS = 1024;   // child rows per extend-add task (default; the autotuner may pick another)
constexpr int NF = 6;      // per-front statistics fields
constexpr int MAX_NRHS = 64;  // right-hand sides per solve call
// Cholesky sweeps with inverted diagonal blocks: a front's columns are solved in steps of at
// most SOLVE_STEP columns by one workgroup each; narrow fronts (at most NB columns) with at most
// FUSE_ROWS rows below their columns form their update vector in the same workgroup, the others
// in separate row-tile launches; L21^T partial sums are formed per L21T_CHUNK rows
constexpr int SOLVE_STEP = 512;
constexpr int FUSE_ROWS = 256;
constexpr int L21T_CHUNK = 512;

// one solve step: columns [c0, c1) of front s; L(rows below c1, step)^T x partial sums at poff
// (units of nrhs), nch chunks of L21T_CHUNK rows; a wide step (more than NB columns) keeps the
// inverse of its diagonal block at xoff (column-major, ld c1 - c0; -1: none)
struct SolveStep {
    int32_t s, c0, c1, nch;
    int64_t poff, xoff;
};

// value of `v` in lane `lane` (wave-uniform) of the calling wave, as a register broadcast
__device__ inline double lane_bcast(double v, int lane) {
    int2 x = __builtin_bit_cast(int2, v);
    x.x = __builtin_amdgcn_readlane(x.x, lane);
    x.y = __builtin_amdgcn_readlane(x.y, lane);
    return __builtin_bit_cast(double, x);
}
__device__ inline float lane_bcast(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// row[r] = row[r] + a[r] (or - a[r] when sub) for r < nr (at most 8): the loads issued together
// (clamped into the row), the results held in registers by empty asm statements, then the
// stores. Written as one read-modify-write per r, the compiler waited for every memory
// operation before each store.
template <typename T>
__device__ __attribute__((always_inline)) inline void rmw8(T* row, const T (&a)[8], int nr, bool sub) {
    T o[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) o[r] = row[min(r, nr - 1)];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        o[r] = sub ? o[r] - a[r] : o[r] + a[r];
        asm volatile("" : "+v"(o[r]));
    }
#pragma unroll
    for (int r = 0; r < 8; ++r)
        if (r < nr) row[r] = o[r];
}

#define SDS_HIP_CHECK(x)                                      \
    do {                                                      \
        hipError_t e_ = (x);                                  \
        if (e_ != hipSuccess) return ::sds::hip_fail(e_, #x); \
    } while (0)

int hip_fail(hipError_t e, const char* what);

// the 16 x 16 x 4 matrix-core products (v_mfma_f64_16x16x4f64, v_mfma_f32_16x16x4f32): lane l
// gives A(l % 16, l / 16) and B(l / 16, l % 16); element q of its accumulator is
// C(row(l / 16, q), l % 16)
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ inline acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int l4, int q) { return 4 * q + l4; }
};
template <>
struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ inline acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int l4, int q) { return 4 * l4 + q; }
};

}  // namespace sds
