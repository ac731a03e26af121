// Host side of libsds_hip: the C ABI of include/sds_hip.h, the device schedule
// built from an uploaded analysis, and the drivers of the numeric and solve
// phases. All device memory is allocated at upload/allocation time; the phases
// themselves allocate nothing (except growing the solve workspace when a call
// brings more right-hand sides than any before it).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <chrono>
#include <map>
#include <mutex>
#include <set>
#include <vector>

#include "../../include/sds_hip.h"
#include "sds_common.h"
#include "sds_kernels.h"

// The solve workspace of sds_solve_noalloc: one array per GPU (each device holds its own
// copy, at its own address, until the process ends), shared by every call on that GPU.
// Host code never uses the name `scratch` as an address: it is a host-side stand-in; the
// device address comes from hipGetSymbolAddress, once per device (scratch_base).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(SDS_NOALLOC_SCRATCH_MIB) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace sds {

int hip_fail(hipError_t e, const char* what) {
    (void)e;
    (void)what;
    return SDS_DEVICE_ERROR;
}

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// launch kinds of the large-front plans
enum LaunchKind {
    L_ZERO_SCATTER,
    L_EXTEND_ADD,
    L_POTRF,
    L_TRSM,
    L_UPDATE_PANEL,
    L_STEP,          // k_big_step: a step, with the previous step's copies and trailing update
    L_BLOCK,         // k_big_block: a panel block's update of the next strip, the next diagonal blocks
    L_UPDATE_CB,
    L_LDLT_CB,
    L_LU_ZERO,
    L_LU_SCATTER,
    L_LU_EXTEND,
    L_LU_CB,
    L_LDLT_MIRROR,   // k_ldlt_mirror: the upper triangle of the fully summed squares
    L_LDLT_PIVOT,    // k_ldlt_pivot: launch pair k's pivot of every large front of the level
    L_LDLT_UPDATE,   // k_ldlt_update: that pivot's update tiles
    L_LDLT_FINISH,   // k_ldlt_finish: unit diagonals and the pivot order
    L_LDLT_PANEL,    // k_ldlt_panel: a panel of pivots of every large front of the level
    L_LDLT_TRAIL,    // k_ldlt_trail: that panel's update of the columns after it
    L_LU_PIVOT,      // k_lu_pivot: step k's pivot of every large front of the level
    L_LU_UPDATE,     // k_lu_update: step k's rank-1 update tiles
    L_LU_FINISH,     // k_lu_finish: the last deferred scaling, U^T and unit diagonals
    S_FWD_PULL,
    S_FWD_STEP,
    S_FWD_FINISH,
    S_BWD_L21,
    S_BWD_L21_SUM,
    S_BWD_STEP,
    S_BWD_FINISH
};

// a front the solve gives one workgroup (no more than one diagonal block, moderate height)
inline bool solve_small(int f, int w) { return w <= NB && f <= 2048; }

struct Launch {
    int kind;
    int64_t off;
    int n;
    double flops, bytes;  // algorithmic work of the launch
    int nstep = 0, ncopy = 0;  // k_big_step launches: step tasks, then copies, then update tiles
    int k = 0;                 // pivot step of the k_lu_pivot / k_lu_update launches
};

struct Level {
    std::vector<int32_t> small, big, all;
    std::vector<int32_t> ssmall;        // fronts the solve handles one workgroup each
    int fmax_small = 0;
    int64_t off_small = 0, off_all = 0, off_ssmall = 0;  // offsets into the device front lists
    int nsub = 0;  // fronts factored by the subtree kernel (Cholesky regime A)
    // Cholesky: the small fronts of at most TINY_F rows (one wave each) and the others
    std::vector<int32_t> tiny, mid;
    int64_t off_tiny = 0, off_mid = 0;
    int fmax_mid = 0;
    double tiny_flops = 0, tiny_bytes = 0;
    std::vector<Launch> plan;           // large-front launches of this level
    std::vector<Launch> fwd_plan, bwd_plan;  // large-front solve launches (per right-hand side work)
    double small_flops = 0, small_bytes = 0;
    double fwd_bytes = 0, fwd_flops = 0;  // factor bytes read by a sweep, flops per right-hand side
    double rhs_rows = 0;                  // rows of the level's fronts (right-hand-side entries touched)
    // Cholesky sweeps with inverted diagonal blocks, per solve step g of the level's fronts:
    // forward diagonal (one workgroup per front step, `threads` each) and row-tile launches,
    // backward partial-sum and diagonal launches
    struct StepGroup {
        Launch fdiag, frows, bl21t, bdiag;  // narrow steps (one block), rows below, partial sums
        Launch fwide, bwide, copy;           // wide steps: the inverse products, the copies of their rows
        Launch fcopy;                        // forward copies of the wide steps with no rows below
        int threads;
    };
    std::vector<StepGroup> groups;
};

// plan launch kind -> profiling kind of include/sds_hip.h
int prof_kind(int kind) {
    switch (kind) {
        case L_ZERO_SCATTER: return SDS_K_ZERO_SCATTER;
        case L_EXTEND_ADD: return SDS_K_EXTEND_ADD;
        case L_POTRF: return SDS_K_POTRF;
        case L_TRSM: return SDS_K_TRSM;
        case L_UPDATE_PANEL: return SDS_K_UPDATE_PANEL;
        case L_STEP: return SDS_K_TRSM;
        case L_BLOCK: return SDS_K_UPDATE_PANEL;
        case L_LU_PIVOT: return SDS_K_PIVOT_PANEL;
        case L_LDLT_MIRROR: return SDS_K_PIVOT_PANEL;
        case L_LDLT_PIVOT: return SDS_K_PIVOT_PANEL;
        case L_LDLT_UPDATE: return SDS_K_PIVOT_PANEL;
        case L_LDLT_FINISH: return SDS_K_PIVOT_PANEL;
        case L_LDLT_PANEL: return SDS_K_PIVOT_PANEL;
        case L_LDLT_TRAIL: return SDS_K_UPDATE_PANEL;
        case L_LU_UPDATE: return SDS_K_PIVOT_PANEL;
        case L_LU_FINISH: return SDS_K_PIVOT_PANEL;
        case L_LU_ZERO: return SDS_K_ZERO_SCATTER;
        case L_LU_SCATTER: return SDS_K_ZERO_SCATTER;
        case L_LU_EXTEND: return SDS_K_EXTEND_ADD;
        default: return SDS_K_UPDATE_CB;
    }
}

// factorization method of a structure
enum Method { M_CHOLESKY = 0, M_LDLT = 1, M_LU = 2 };

struct ProfRec {
    int kind, e0, e1;
    double flops, bytes;
};

// The solve workspace of one call: Y and Z (n x nrhs of the factor type, row-interleaved), the
// update vectors U, the backward sweep's partial sums, the residual R and the copy Bc of B
// (Float64), 2 MAX_NRHS slots for the column norms followed by DOT_BLOCKS partial sums of the
// dot products, and the host copy of the norms: the handle's buffers (sds_solve).
struct SolveWork {
    void *Y = nullptr, *Z = nullptr, *U = nullptr, *part = nullptr;
    double *R = nullptr, *Bc = nullptr, *norms = nullptr;
    double* hnorms = nullptr;
};

}  // namespace
}  // namespace sds

using namespace sds;

struct sds_handle {
    hipStream_t stream = nullptr;
    int device = 0;
    std::string err;
    // analysis (host copies, 0-based)
    bool have_sym = false;
    int64_t n = 0, nnz = 0, ns = 0, factor_len = 0;
    int structure = 3;
    int method = M_CHOLESKY;
    std::vector<int32_t> perm, super_ptr, fnrows, fncols, snparent, child_ptr, child_list;
    std::vector<int64_t> panel_off, asm_count;
    std::vector<int32_t> cb_split;  // per child: its contribution-block columns that land in the parent's panel
    std::vector<int64_t> relind_ptr;  // per child: positions of its contribution-block rows among its parent's
    std::vector<int32_t> relind;      // rows (increasing), as on the device
    // inverses of the diagonal blocks (Cholesky): blocks of at most 16, 32 and NB columns
    Launch inv_launch[3]{{L_POTRF, 0, 0, 0, 0}, {L_POTRF, 0, 0, 0, 0}, {L_POTRF, 0, 0, 0, 0}};
    SolveStep* d_steps = nullptr;            // solve steps of the Cholesky sweeps
    Launch sinv_launch{L_POTRF, 0, 0, 0, 0};  // inverses of the wide steps' diagonal blocks
    void* xbuf = nullptr;                    // those inverses
    int64_t part_units = 0;                  // partial-sum buffer, units of nrhs
    DevBuf part;
    std::vector<DevBuf> symbufs;  // device maps
    // matching and scaling (matching_alg, sds_set_matching): the entry weights r_i c_j (nnz), the
    // scaled values the factorization reads, the row and column scalings (n) and the column
    // permutation q of "G" (0-based, none for the symmetric structures)
    std::vector<DevBuf> mtbufs;
    double *mt_w = nullptr, *mt_vals = nullptr, *mt_rs = nullptr, *mt_cs = nullptr;
    int32_t* mt_q = nullptr;
    DevSym dsym{};
    int64_t full_nrows = 0;
    int64_t* full_rowptr = nullptr;
    int32_t* full_colval = nullptr;
    int64_t* full_src = nullptr;
    // adjoint map: stored entry p at (row, col) of the permuted-back matrix, 0 unused /
    // 1 one position / 2 a mirrored symmetric entry (both positions)
    int32_t* adj_i = nullptr;
    int32_t* adj_j = nullptr;
    int8_t* adj_flag = nullptr;
    // schedule + numeric
    int dtype = -1;
    std::vector<Level> levels;
    std::vector<DevBuf> numbufs;
    int32_t* d_fronts = nullptr;  // per-level front lists (small lists, then all lists)
    Task* d_tasks = nullptr;
    int64_t cb_len = 0, uv_len = 0;
    int64_t* d_cb_off = nullptr;
    int64_t* d_uv_off = nullptr;
    int64_t* d_pm_row = nullptr;  // the forward pull map of the Cholesky sweeps (DevSym)
    int32_t* d_pm_ptr = nullptr;
    int32_t* d_pm_src = nullptr;
    bool placed = false;                      // cb_place / uv_place hold this analysis' placements
    int place_key = 0;                        // ... for this element type and subtree budget
    // regime A (Cholesky): subtree budget (KiB of LDS, 0 = off), subtrees (node ranges of
    // sub_nodes, in postorder), per supernode the stack offsets of its front and of its
    // contribution block, and one launch per LDS class {first subtree, count, LDS bytes}
    int sub_kib = 0;
    // Schur complement mode: the root supernode holding the Schur block (assembled, not factored),
    // -1 = off, and its size
    int32_t schur_s = -1;
    int64_t nschur = 0;
    std::vector<int32_t> sub_ptr, sub_nodes, sub_lfront, sub_lcb;
    struct SubLaunch {
        int first, count, lds;
        double flops, bytes;  // algorithmic work: the fronts' flops; panels written, A read, roots' blocks
    };
    std::vector<SubLaunch> sub_launch;
    int32_t *d_sub_ptr = nullptr, *d_sub_nodes = nullptr, *d_sub_lfront = nullptr, *d_sub_lcb = nullptr;
    std::vector<int64_t> cb_place, uv_place;  // element offsets of the contribution blocks / update vectors
    void* factor = nullptr;
    void* ufactor = nullptr;  // U^T panels of "G"
    void* cbpool = nullptr;
    void* dbuf = nullptr;   // inverse diagonal blocks (NB entries per column)
    void* lnbuf = nullptr;  // a large-front step's solved rows of the next block (NB entries per column)
    int32_t* lstep = nullptr;  // LDL^T large fronts: per front the current step (column, size, next column)
    // tile shapes chosen by sds_autotune: child rows per extend-add task, K chunk of the
    // contribution-block update (32: four workgroups per CU, 64: two)
    int ea_rows = EA_ROWS, upd_kc = 32;
    int rest_split = 0;  // trailing tiles of a step from which they get a launch of their own (0: never)
    int max_rest = 0;    // most trailing tiles of one step in the current plan
    int step_occ2 = 0;   // step launches of at least this many workgroups use the two-per-CU build (0: none)
    int small_kind = 0;  // small-front kernels (Cholesky): 0 one wave per tiny front + blocked,
                         // 1 one wave per tiny front + column by column, 2 column by column
    int panel_block = 0;  // Cholesky large fronts: panel block width (a multiple of NB; 0 = one block)
    int pivot_panels = 0;  // LDL^T large fronts: 1 deferred-update panels, 0 one pivot per launch pair
    int64_t ldlt_scr = 0;  // ... the panels' scratch (elements)
    void* lscratch = nullptr;
    int32_t* status = nullptr;
    void* dvec = nullptr;        // D: 2n entries of the factor type (diagonal, 2x2 subdiagonals)
    int32_t* piv = nullptr;      // n: local pivot order as 0-based factor columns
    int8_t* pkind = nullptr;     // n: pivot kinds
    int8_t* psign = nullptr;     // n: requested pivot signs by original row (pivot_sign)
    double* aux = nullptr;       // max |a_ij| of the values (pivot_epsilon_alg = "algo1")
    int64_t* stats = nullptr;
    int64_t* d_info = nullptr;
    int64_t* h_info = nullptr;  // pinned
    bool factorized = false;
    // solve workspace
    int cap_nrhs = 0;
    DevBuf Y, Z, U, R, Bc, norms;
    double* h_norms = nullptr;  // pinned
    // refinement: 0 plain steps, 1 FGMRES (sds_set_ir_mode); the FGMRES Krylov vectors
    int ir_mode = 0;
    DevBuf kry;
    double* h_kry = nullptr;  // pinned, h_kry_cap entries
    size_t h_kry_cap = 0;
    int fsmall = 0;
    // captured refactorization + solve (sds_graph_capture), on a private capture stream
    hipStream_t cstream = nullptr;

    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // profiling (HIP events around every launch on the handle's stream)
    bool prof = false;
    std::vector<hipEvent_t> evpool;
    size_t evnext = 0;
    std::vector<ProfRec> recs;
    double acc_ms[SDS_K_COUNT] = {}, acc_flops[SDS_K_COUNT] = {}, acc_bytes[SDS_K_COUNT] = {};
    int64_t acc_n[SDS_K_COUNT] = {};
    int64_t nnz_full = 0;
};

namespace {

int fail(sds_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define HIPCHK(h, x)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess)                                                                 \
            return fail(h, SDS_DEVICE_ERROR, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

int dev_alloc(sds_handle* h, std::vector<DevBuf>& owner, size_t bytes, void** out) {
    void* p = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
        return fail(h, SDS_DEVICE_ERROR,
                    "hipMalloc of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    owner.push_back({p, bytes});
    *out = p;
    return SDS_SUCCESS;
}

void free_all(std::vector<DevBuf>& bufs) {
    for (auto& b : bufs) (void)hipFree(b.p);
    bufs.clear();
}

int next_event(sds_handle* h) {
    if (h->evnext == h->evpool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        h->evpool.push_back(e);
    }
    return (int)h->evnext++;
}

// run `fn` (one kernel launch) bracketed by events when profiling is on
template <class F>
void timed(sds_handle* h, int kind, double flops, double bytes, F&& fn) {
    if (!h->prof) {
        fn();
        return;
    }
    int e0 = next_event(h), e1 = next_event(h);
    if (e0 < 0 || e1 < 0) {
        fn();
        return;
    }
    (void)hipEventRecord(h->evpool[e0], h->stream);
    fn();
    (void)hipEventRecord(h->evpool[e1], h->stream);
    h->recs.push_back({kind, e0, e1, flops, bytes});
}

// wait for the recorded launches and accumulate their durations
void harvest(sds_handle* h) {
    if (h->recs.empty()) return;
    (void)hipStreamSynchronize(h->stream);
    for (const ProfRec& r : h->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, h->evpool[r.e0], h->evpool[r.e1]) != hipSuccess) continue;
        h->acc_ms[r.kind] += ms;
        h->acc_flops[r.kind] += r.flops;
        h->acc_bytes[r.kind] += r.bytes;
        h->acc_n[r.kind] += 1;
    }
    h->recs.clear();
    h->evnext = 0;
}

template <typename V>
int upload(sds_handle* h, std::vector<DevBuf>& owner, const std::vector<V>& v, V** out) {
    void* p;
    int rc = dev_alloc(h, owner, v.size() * sizeof(V), &p);
    if (rc) return rc;
    if (!v.empty()) HIPCHK(h, hipMemcpy(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
    *out = static_cast<V*>(p);
    return SDS_SUCCESS;
}

struct IdxReader {
    const void* p;
    int bits;
    int64_t operator[](int64_t i) const {
        return bits == 32 ? (int64_t) static_cast<const int32_t*>(p)[i] : static_cast<const int64_t*>(p)[i];
    }
};

// best-fit placement of blocks with lifetimes [first, last] (levels) on an unbounded buffer:
// level by level, release the blocks whose last level has passed (merging adjacent holes), then
// place the new ones, largest first, each in the smallest hole that holds it (or on top).
// Holes are kept by offset (for merging) and by (length, offset) (for the fit): O(b log b) for
// b blocks. Returns the high-water mark.
int64_t place_blocks(const std::vector<int64_t>& len, const std::vector<int>& first, const std::vector<int>& last,
                     int nlev, std::vector<int64_t>& off) {
    int64_t nb = (int64_t)len.size();
    off.assign(nb, -1);
    std::vector<std::vector<int64_t>> born(nlev + 1), dies(nlev + 2);
    for (int64_t b = 0; b < nb; ++b)
        if (len[b] > 0) {
            born[first[b]].push_back(b);
            dies[last[b] + 1].push_back(b);
        }
    std::map<int64_t, int64_t> holes;              // offset -> length
    std::set<std::pair<int64_t, int64_t>> bylen;   // (length, offset)
    int64_t top = 0;
    auto drop = [&](std::map<int64_t, int64_t>::iterator it) {
        bylen.erase({it->second, it->first});
        holes.erase(it);
    };
    auto release = [&](int64_t o, int64_t l) {
        auto nx = holes.find(o + l);
        if (nx != holes.end()) {
            l += nx->second;
            drop(nx);
        }
        auto it = holes.lower_bound(o);
        if (it != holes.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == o) {
                o = pv->first;
                l += pv->second;
                drop(pv);
            }
        }
        if (o + l == top) {
            top = o;
            return;
        }
        holes[o] = l;
        bylen.insert({l, o});
    };
    int64_t hw = 0;
    for (int L = 0; L <= nlev; ++L) {
        for (int64_t b : dies[L]) release(off[b], len[b]);
        std::vector<int64_t>& nbk = born[L];
        std::sort(nbk.begin(), nbk.end(), [&](int64_t a, int64_t b) { return len[a] > len[b]; });
        for (int64_t b : nbk) {
            const int64_t l = len[b];
            auto it = bylen.lower_bound({l, INT64_MIN});
            if (it != bylen.end()) {
                const int64_t hl = it->first, ho = it->second;
                bylen.erase(it);
                holes.erase(ho);
                off[b] = ho;
                if (hl > l) {
                    holes[ho + l] = hl - l;
                    bylen.insert({hl - l, ho + l});
                }
            } else {
                off[b] = top;
                top += l;
            }
            hw = std::max(hw, top);
        }
    }
    return hw;
}

void graph_free(sds_handle* h) {
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->gexec = nullptr;
    h->graph = nullptr;
}

void free_numeric(sds_handle* h) {
    graph_free(h);  // a captured graph refers to the buffers
    free_all(h->numbufs);
    h->ufactor = nullptr;
    if (h->h_info) (void)hipHostFree(h->h_info);
    if (h->h_norms) (void)hipHostFree(h->h_norms);
    h->h_info = nullptr;
    h->h_norms = nullptr;
    if (h->h_kry) (void)hipHostFree(h->h_kry);
    h->h_kry = nullptr;
    h->h_kry_cap = 0;
    for (DevBuf* b : {&h->Y, &h->Z, &h->U, &h->R, &h->Bc, &h->norms, &h->part, &h->kry}) {
        if (b->p) (void)hipFree(b->p);
        *b = DevBuf{};
    }
    h->cap_nrhs = 0;
    h->levels.clear();
    h->dtype = -1;
    h->factorized = false;
}

int build_schedule(sds_handle* h, int dtype) {
    const int64_t ns = h->ns;
    const size_t es = dtype == SDS_R64 ? 8 : 4;
    // largest front whose dense f x f square fits 64 KiB of LDS next to the kernel's static words
    // (the pivoting kernels keep per-column pivot data in LDS too)
    int fsmall = h->method == M_CHOLESKY ? (int)std::floor(std::sqrt((65536.0 - 256.0) / (double)es))
                                         : std::min(128, (int)std::floor(std::sqrt((65536.0 - 4096.0) / (double)es)));
    h->fsmall = fsmall;
    // tree levels by height: leaves 0, a parent one above its highest child
    std::vector<int> lev(ns, 0);
    for (int64_t s = 0; s < ns; ++s)
        if (h->snparent[s] >= 0) lev[h->snparent[s]] = std::max(lev[h->snparent[s]], lev[s] + 1);
    int nlev = 0;
    for (int64_t s = 0; s < ns; ++s) nlev = std::max(nlev, lev[s] + 1);
    h->levels.assign(nlev, Level{});
    h->max_rest = 0;
    h->ldlt_scr = 0;
    std::vector<char> big(ns, 0);
    // regime A (Cholesky, sub_kib > 0): leaf subtrees of small fronts whose LDS stack fits the
    // budget, one workgroup each, walked in postorder before the levels. A front sits above its
    // children's contribution blocks (each an m x m square), which it consumes; its own block then
    // moves down to where its first child's block began. peak(v) = max(max_i(base_i + peak(c_i)),
    // sum_c m_c^2 + f^2) with base_i the blocks of the earlier children.
    std::vector<int32_t> in_sub(ns, -1);
    h->sub_ptr.assign(1, 0);
    h->sub_nodes.clear();
    h->sub_lfront.assign(ns, 0);
    h->sub_lcb.assign(ns, 0);
    h->sub_launch.clear();
    if (h->method == M_CHOLESKY && h->sub_kib > 0) {
        // elements of the stack: the budget less the kernel's static words (at most 64 KiB in all)
        const int64_t cap = ((int64_t)h->sub_kib * 1024 - 256) / (int64_t)es;
        std::vector<int64_t> peak(ns, 0);
        std::vector<char> ok(ns, 0);
        for (int64_t s = 0; s < ns; ++s) {
            const int64_t f = h->fnrows[s];
            bool good = f <= fsmall && s != h->schur_s;
            int64_t base = 0, pk = 0;
            for (int k = h->child_ptr[s]; k < h->child_ptr[s + 1]; ++k) {
                const int32_t c = h->child_list[k];
                const int64_t mc = h->fnrows[c] - h->fncols[c];
                good = good && ok[c];
                pk = std::max(pk, base + peak[c]);
                base += mc * mc;
            }
            pk = std::max(pk, base + f * f);
            peak[s] = pk;
            ok[s] = good && pk <= cap;
        }
        // subtree roots (at least two fronts: a lone front gains nothing), their postorder and the
        // stack offsets, grouped by the LDS their peak needs (16, 32 or 64 KiB classes)
        std::vector<std::vector<int32_t>> byclass(3);
        for (int64_t s = 0; s < ns; ++s) {
            const int32_t p = h->snparent[s];
            if (!ok[s] || (p >= 0 && ok[p]) || h->child_ptr[s + 1] == h->child_ptr[s]) continue;
            const int64_t bytes = peak[s] * (int64_t)es + 256;
            byclass[bytes <= 16384 ? 0 : bytes <= 32768 ? 1 : 2].push_back((int32_t)s);
        }
        for (int cls = 0; cls < 3; ++cls) {
            int64_t pmax = 0;
            const int first_sub = (int)h->sub_ptr.size() - 1;
            for (int32_t root : byclass[cls]) {
                pmax = std::max(pmax, peak[root]);
                // postorder with children in child_list order (iterative), offsets on the way down
                std::vector<std::pair<int32_t, int>> stk{{root, 0}};
                h->sub_lcb[root] = 0;
                while (!stk.empty()) {
                    auto& [v, ci] = stk.back();
                    const int nc = h->child_ptr[v + 1] - h->child_ptr[v];
                    if (ci == 0) {
                        int64_t b = h->sub_lcb[v];
                        for (int k = h->child_ptr[v]; k < h->child_ptr[v + 1]; ++k) {
                            const int32_t c = h->child_list[k];
                            const int64_t mc = h->fnrows[c] - h->fncols[c];
                            h->sub_lcb[c] = (int32_t)b;
                            b += mc * mc;
                        }
                        h->sub_lfront[v] = (int32_t)b;
                    }
                    if (ci < nc) {
                        const int32_t c = h->child_list[h->child_ptr[v] + ci];
                        ++ci;
                        stk.push_back({c, 0});
                    } else {
                        h->sub_nodes.push_back(v);
                        in_sub[v] = (int32_t)h->sub_ptr.size() - 1;
                        stk.pop_back();
                    }
                }
                h->sub_ptr.push_back((int32_t)h->sub_nodes.size());
            }
            const int cnt = (int)h->sub_ptr.size() - 1 - first_sub;
            if (cnt) h->sub_launch.push_back({first_sub, cnt, (int)(pmax * (int64_t)es), 0.0, 0.0});
        }
    }
    for (int64_t s = 0; s < ns; ++s) {
        if (s == h->schur_s) {
            big[s] = 1;  // assembled by the last pseudo-level below; in no sweep plan
            continue;
        }
        Level& L = h->levels[lev[s]];
        L.all.push_back((int32_t)s);
        if (in_sub[s] >= 0) {
            // factored by its subtree's workgroup; still a small front for the sweeps
            L.nsub += 1;
        } else if (h->fnrows[s] <= fsmall) {
            L.small.push_back((int32_t)s);
            L.fmax_small = std::max(L.fmax_small, h->fnrows[s]);
            if (h->fnrows[s] <= TINY_F) {
                L.tiny.push_back((int32_t)s);
            } else {
                L.mid.push_back((int32_t)s);
                L.fmax_mid = std::max(L.fmax_mid, h->fnrows[s]);
            }
        } else {
            L.big.push_back((int32_t)s);
            big[s] = 1;
        }
    }
    // contribution blocks (m x m) and solve update vectors (m per rhs): live from
    // their front's level to their parent's level
    std::vector<int64_t> cblen(ns, 0), uvlen(ns, 0);
    std::vector<int> first(ns), last(ns);
    for (int64_t s = 0; s < ns; ++s) {
        int64_t m = h->fnrows[s] - h->fncols[s];
        first[s] = lev[s];
        last[s] = h->snparent[s] >= 0 ? lev[h->snparent[s]] : lev[s];
        if (m > 0 && (h->snparent[s] >= 0 || big[s])) cblen[s] = m * m;
        if (m > 0 && h->snparent[s] >= 0) uvlen[s] = m;
        if (in_sub[s] >= 0) {
            // inside a subtree the blocks stay in LDS; a subtree root's block is written before the
            // levels begin
            const int32_t p = h->snparent[s];
            if (p >= 0 && in_sub[p] == in_sub[s]) cblen[s] = 0;
            else first[s] = 0;
        }
    }
    // the placements depend on the analysis, the big/small split (element type) and the subtree
    // budget: computed again only when one of them changes
    const int place_key = (int)dtype * 1000 + h->sub_kib;
    if (!h->placed || h->place_key != place_key) {
        h->cb_len = place_blocks(cblen, first, last, nlev, h->cb_place);
        h->uv_len = place_blocks(uvlen, first, last, nlev, h->uv_place);
        h->placed = true;
        h->place_key = place_key;
    }
    const std::vector<int64_t>& cb_off = h->cb_place;
    const std::vector<int64_t>& uv_off = h->uv_place;
    // front lists
    std::vector<int32_t> flist;
    for (auto& L : h->levels) {
        L.off_small = (int64_t)flist.size();
        flist.insert(flist.end(), L.small.begin(), L.small.end());
        L.off_tiny = (int64_t)flist.size();
        flist.insert(flist.end(), L.tiny.begin(), L.tiny.end());
        L.off_mid = (int64_t)flist.size();
        flist.insert(flist.end(), L.mid.begin(), L.mid.end());
    }
    for (auto& L : h->levels) {
        L.off_all = (int64_t)flist.size();
        flist.insert(flist.end(), L.all.begin(), L.all.end());
    }
    for (auto& L : h->levels) {
        for (int32_t s : L.all)
            if (solve_small(h->fnrows[s], h->fncols[s])) L.ssmall.push_back(s);
        L.off_ssmall = (int64_t)flist.size();
        flist.insert(flist.end(), L.ssmall.begin(), L.ssmall.end());
    }
    // algorithmic work: Cholesky flops of a front sum_{k<w} (f-k)^2 (the CHOLMOD
    // convention of the reference, src/symbolic/etree.jl:153-160), bytes = what the
    // front must move at least: its panel and contribution block written once, A read
    // (value + map), its children's contribution blocks read once
    const double dbl_es = (double)es;
    auto front_flops = [&](int32_t s) {
        double f = h->fnrows[s], w = h->fncols[s], fl = 0;
        for (int k = 0; k < (int)w; ++k) fl += (f - k) * (f - k);
        return fl;
    };
    auto front_bytes = [&](int32_t s) {
        double f = h->fnrows[s], w = h->fncols[s], m = f - w;
        double b = dbl_es * (f * w - w * (w - 1) / 2 + m * (m + 1) / 2);
        b += 16.0 * (double)(h->asm_count[s]);
        for (int k = h->child_ptr[s]; k < h->child_ptr[s + 1]; ++k) {
            double mc = h->fnrows[h->child_list[k]] - h->fncols[h->child_list[k]];
            b += dbl_es * mc * (mc + 1) / 2;
        }
        return b;
    };
    for (auto& sl : h->sub_launch)
        for (int t = sl.first; t < sl.first + sl.count; ++t)
            for (int k = h->sub_ptr[t]; k < h->sub_ptr[t + 1]; ++k) {
                const int32_t v = h->sub_nodes[k];
                const double f = h->fnrows[v], w = h->fncols[v], m = f - w;
                sl.flops += front_flops(v);
                sl.bytes += dbl_es * (f * w - w * (w - 1) / 2) + 16.0 * (double)h->asm_count[v];
                if (k == h->sub_ptr[t + 1] - 1) sl.bytes += dbl_es * m * (m + 1) / 2;
            }
    for (auto& L : h->levels) {
        for (int32_t s : L.small) {
            L.small_flops += front_flops(s);
            L.small_bytes += front_bytes(s);
            if (h->fnrows[s] <= TINY_F) {
                L.tiny_flops += front_flops(s);
                L.tiny_bytes += front_bytes(s);
            }
        }
        for (int32_t s : L.ssmall) {
            double f = h->fnrows[s], w = h->fncols[s];
            L.fwd_bytes += dbl_es * (f * w - w * (w - 1) / 2);
            L.fwd_flops += 2.0 * (f * w - w * (w - 1) / 2);
            L.rhs_rows += f;
        }
    }
    // entries (i, j) of a tile with j <= i, rows < f, columns in [jmin, jlim)
    auto tile_entries = [](int I0, int J0, int f, int jmin, int jlim) {
        double c = 0;
        for (int j = std::max(J0, jmin); j < std::min(J0 + TILE, jlim); ++j) {
            int i0 = std::max(I0, j), i1 = std::min(I0 + TILE, f);
            if (i1 > i0) c += i1 - i0;
        }
        return c;
    };
    // task lists of the large fronts
    std::vector<Task> tasks;
    auto push_launch = [&](Level& L, int kind, const std::vector<Task>& t, double flops, double bytes) {
        if (t.empty()) return;
        L.plan.push_back({kind, (int64_t)tasks.size(), (int)t.size(), flops, bytes});
        tasks.insert(tasks.end(), t.begin(), t.end());
    };
    // Cholesky big fronts: the contribution block is written whole by its update (no zeroing),
    // so the children's entries are added in two parts: the columns that land in the panel
    // before the factorization, the contribution-block columns after the update
    const bool chol = h->method == M_CHOLESKY;
    auto push_extend = [&](Level& L, bool cb_part, int kind) {
        if (kind == L_EXTEND_ADD) {
            // owner-pull, one launch: a task owns the parent entries of EA_COLS parent columns
            // (never across w) and ea_rows parent rows from the block's first column, and adds
            // the children's entries that land there in child_list order. Its entry list (after
            // the launch's task records, in the same task array): (child, first child column of
            // the block, child rows [ii0, ii1)). Cholesky: the panel columns [0, w) before the
            // factorization, the contribution-block columns [w, f) after the update; otherwise
            // all of them before the factorization.
            std::vector<Task> rec, ent;
            double by = 0;
            auto lb = [&](int32_t c, int v) {
                const int32_t* r0 = h->relind.data() + h->relind_ptr[c];
                const int32_t* r1 = h->relind.data() + h->relind_ptr[c + 1];
                return (int)(std::lower_bound(r0, r1, v) - r0);
            };
            for (int32_t s : L.big) {
                const int f = h->fnrows[s], w = h->fncols[s];
                const int p0 = chol && cb_part ? w : 0, p1 = chol && !cb_part ? w : f;
                for (int kc = h->child_ptr[s]; kc < h->child_ptr[s + 1]; ++kc) {
                    const int32_t c = h->child_list[kc];
                    if (cb_off[c] < 0) continue;
                    const double mc = h->fnrows[c] - h->fncols[c];
                    const double j0 = lb(c, p0), j1 = lb(c, p1);
                    // read the child's block, read + write the parent
                    by += 3.0 * dbl_es * ((mc - j0) * (mc - j0 + 1) - (mc - j1) * (mc - j1 + 1)) / 2;
                }
                // column blocks never straddle w (the kernel ends a block at w)
                for (int pc0 = p0; pc0 < p1; pc0 = pc0 < w && pc0 + EA_COLS > w ? w : pc0 + EA_COLS) {
                    const int pc1 = std::min(pc0 + EA_COLS, pc0 < w ? w : p1);
                    for (int pr0 = pc0; pr0 < f; pr0 += h->ea_rows) {
                        const int pr1 = std::min(pr0 + h->ea_rows, f);
                        const int32_t e0 = (int32_t)ent.size();
                        for (int kc = h->child_ptr[s]; kc < h->child_ptr[s + 1]; ++kc) {
                            const int32_t c = h->child_list[kc];
                            if (cb_off[c] < 0) continue;
                            const int jlo = lb(c, pc0), jhi = lb(c, pc1);
                            if (jlo == jhi) continue;
                            const int ilo = std::max(jlo, lb(c, pr0)), ihi = lb(c, pr1);
                            if (ilo < ihi) ent.push_back({c, jlo, ilo, ihi});
                        }
                        if ((int32_t)ent.size() > e0) rec.push_back({s, pc0, e0, (int32_t)ent.size() - e0});
                    }
                }
            }
            if (rec.empty()) return;
            for (Task& r : rec) r.b += (int32_t)rec.size();  // entry offsets from the launch's first record
            L.plan.push_back({kind, (int64_t)tasks.size(), (int)rec.size(), 0, by});
            tasks.insert(tasks.end(), rec.begin(), rec.end());
            tasks.insert(tasks.end(), ent.begin(), ent.end());
            return;
        }
        int maxch = 0;
        for (int32_t s : L.big) maxch = std::max(maxch, h->child_ptr[s + 1] - h->child_ptr[s]);
        for (int r = 0; r < maxch; ++r) {  // one launch per child rank: each parent entry has one writer
            std::vector<Task> t;
            double by = 0;
            for (int32_t s : L.big) {
                if (h->child_ptr[s] + r >= h->child_ptr[s + 1]) continue;
                int32_t c = h->child_list[h->child_ptr[s] + r];
                if (cb_off[c] < 0) continue;
                const int mc = h->fnrows[c] - h->fncols[c];
                for (int jj = 0; jj < mc; jj += 4) t.push_back({s, c, jj, std::min(4, mc - jj)});
                by += 3.0 * dbl_es * (double)mc * mc;  // read the child's block, read + write the parent
            }
            push_launch(L, kind, t, 0, by);
        }
    };
    // Cholesky: zeroing a large front's panel and adding A into it reads nothing the
    // factorization writes, and nothing else writes that panel before the front's own level: the
    // panels of every level are prepared by one launch at the head of the first level with
    // large fronts (one launch instead of one per level on the level chain)
    bool zs_done = false;
    for (auto& L : h->levels) {
        if (L.big.empty()) continue;
        std::vector<Task> t;
        double fl = 0, by = 0;
        for (auto& L2 : h->levels) {
            if (!(chol && !zs_done) && &L2 != &L) continue;
            for (int32_t s : L2.big) {
                double f = h->fnrows[s], w = h->fncols[s], m = f - w;
                // Cholesky: the panel columns only (a task past w would zero the head of the
                // front's contribution-block slot, which another front's live block may hold
                // before this front's level)
                const int cols = chol ? h->fncols[s] : h->fnrows[s];
                for (int c = 0; c < cols; c += 8) t.push_back({s, c, std::min(8, cols - c), 0});
                by += dbl_es * (f * w + (chol ? 0.0 : m * (m + 1) / 2)) + 16.0 * (double)h->asm_count[s];
            }
        }
        const bool lu = h->method == M_LU;
        if (lu) {
            push_launch(L, L_LU_ZERO, t, 0, by);
            t.clear();
            for (int32_t s : L.big)
                for (int c = 0; c < h->fncols[s]; c += 8) t.push_back({s, c, 8, 0});
            push_launch(L, L_LU_SCATTER, t, 0, 0);
        } else if (!chol || !zs_done) {
            push_launch(L, L_ZERO_SCATTER, t, 0, by);
            zs_done = chol;
        }
        push_extend(L, false, lu ? L_LU_EXTEND : L_EXTEND_ADD);
        int wmax = 0;
        for (int32_t s : L.big) wmax = std::max(wmax, h->fncols[s]);
        if (lu) {
            // pivoted LU of the large fronts, one step at a time over all of them: the pivot
            // (one workgroup per front) and the rank-1 update tiles (every front's remaining
            // fully summed columns and U12, 64 x 64 tiles; one tile list per 64 steps)
            std::vector<Task> tp;
            for (int32_t s : L.big) tp.push_back({s, 0, 0, 0});
            const int64_t off_p = (int64_t)tasks.size();
            tasks.insert(tasks.end(), tp.begin(), tp.end());
            for (int e = 0; e * 64 < wmax; ++e) {
                std::vector<Task> tu;
                for (int32_t s : L.big) {
                    const int f = h->fnrows[s], w = h->fncols[s];
                    if (64 * e >= w) continue;
                    for (int J0 = 64 * e; J0 < w; J0 += 64)
                        for (int I0 = 64 * e; I0 < f; I0 += 64) tu.push_back({s, I0, J0, 0});
                    for (int J0 = w; J0 < f; J0 += 64)
                        for (int I0 = 64 * e; I0 < w; I0 += 64) tu.push_back({s, I0, J0, 0});
                }
                const int64_t off_u = (int64_t)tasks.size();
                tasks.insert(tasks.end(), tu.begin(), tu.end());
                for (int k = 64 * e; k < std::min(wmax, 64 * e + 64); ++k) {
                    double fl1 = 0, by1 = 0;
                    for (int32_t s : L.big) {
                        const double f = h->fnrows[s], w = h->fncols[s];
                        if (k >= w) continue;
                        const double ent = (f - k - 1) * (w - k - 1) + (w - k - 1) * (f - w);
                        fl1 += 3.0 * ent;
                        by1 += 2.0 * dbl_es * ent;
                    }
                    L.plan.push_back({L_LU_PIVOT, off_p, (int)tp.size(), 0.0, 0.0});
                    L.plan.back().k = k;
                    L.plan.push_back({L_LU_UPDATE, off_u, (int)tu.size(), fl1, by1});
                    L.plan.back().k = k;
                }
            }
            L.plan.push_back({L_LU_FINISH, off_p, (int)tp.size(), 0.0, 0.0});
            wmax = 0;
        } else if (h->method != M_CHOLESKY && h->pivot_panels) {
            // LDL^T of the large fronts in panels of deferred updates: per panel e one launch takes
            // each front's next pivots until ldlt_panel_cols() columns are done (one workgroup per
            // front, its scratch at the offset in the task), then one launch applies them to the
            // entries after the panel (64 x 64 tiles). A panel takes at least that many columns,
            // so ceil(wmax / cols) panels finish every front of the level.
            const int nbp = ldlt_panel_cols();
            std::vector<Task> tp;
            int64_t off = 0;
            for (int32_t s : L.big) {
                tp.push_back({s, (int32_t)off, 0, 0});
                off += ldlt_panel_scratch(h->fnrows[s], h->fncols[s]);
            }
            h->ldlt_scr = std::max(h->ldlt_scr, off);
            const int64_t off_p = (int64_t)tasks.size();
            tasks.insert(tasks.end(), tp.begin(), tp.end());
            const int npan = (wmax + nbp - 1) / nbp;
            for (int e = 0; e < npan; ++e) {
                std::vector<Task> tu;
                double fl1 = 0, by1 = 0;
                const int kmin = (e + 1) * nbp, lo = kmin / TILE * TILE;
                for (size_t q = 0; q < L.big.size(); ++q) {
                    const int32_t s = L.big[q];
                    const int f = h->fnrows[s], w = h->fncols[s];
                    if (kmin >= w) continue;
                    for (int J0 = lo; J0 < w; J0 += TILE)
                        for (int I0 = lo; I0 < f; I0 += TILE) tu.push_back({s, I0, J0, tp[q].a});
                    // a panel of nbp 1x1 pivots on the (f - kmin) x (w - kmin) entries after it
                    const double ent = (double)(f - kmin) * (w - kmin);
                    fl1 += 2.0 * nbp * ent;
                    by1 += 2.0 * dbl_es * ent;
                }
                L.plan.push_back({L_LDLT_PANEL, off_p, (int)tp.size(), 0.0, 0.0});
                L.plan.back().k = e;
                push_launch(L, L_LDLT_TRAIL, tu, fl1, by1);
            }
            L.plan.push_back({L_LDLT_FINISH, off_p, (int)tp.size(), 0.0, 0.0});
            wmax = 0;
        } else if (h->method != M_CHOLESKY) {
            // LDL^T of the large fronts, one pivot (1x1 or 2x2) per launch pair over all of them:
            // the pivot (one workgroup per front) and its update tiles (the remaining fully summed
            // columns, all rows, 64 x 64 tiles; one tile list per 64 pairs; a pair takes at least
            // one column, so pair t works on columns >= t), after the mirror tiles of the fully
            // summed squares
            std::vector<Task> tp, tm;
            for (int32_t s : L.big) {
                tp.push_back({s, 0, 0, 0});
                const int w = h->fncols[s];
                for (int J0 = 0; J0 < w; J0 += 64)
                    for (int I0 = J0; I0 < w; I0 += 64) tm.push_back({s, I0, J0, 0});
            }
            push_launch(L, L_LDLT_MIRROR, tm, 0.0, 0.0);
            const int64_t off_p = (int64_t)tasks.size();
            tasks.insert(tasks.end(), tp.begin(), tp.end());
            for (int e = 0; e * 64 < wmax; ++e) {
                std::vector<Task> tu;
                for (int32_t s : L.big) {
                    const int f = h->fnrows[s], w = h->fncols[s];
                    if (64 * e >= w) continue;
                    for (int J0 = 64 * e; J0 < w; J0 += 64)
                        for (int I0 = 64 * e; I0 < f; I0 += 64) tu.push_back({s, I0, J0, 0});
                }
                const int64_t off_u = (int64_t)tasks.size();
                tasks.insert(tasks.end(), tu.begin(), tu.end());
                for (int k = 64 * e; k < std::min(wmax, 64 * e + 64); ++k) {
                    double fl1 = 0, by1 = 0;
                    for (int32_t s : L.big) {
                        const double f = h->fnrows[s], w = h->fncols[s];
                        if (k >= w) continue;
                        const double ent = (f - k - 1) * (w - k - 1);
                        fl1 += 2.0 * ent;
                        by1 += 2.0 * dbl_es * ent;
                    }
                    L.plan.push_back({L_LDLT_PIVOT, off_p, (int)tp.size(), 0.0, 0.0});
                    L.plan.back().k = k;
                    L.plan.push_back({L_LDLT_UPDATE, off_u, (int)tu.size(), fl1, by1});
                }
            }
            L.plan.push_back({L_LDLT_FINISH, off_p, (int)tp.size(), 0.0, 0.0});
            wmax = 0;
        }
        // per 64-column step b (k = 64 b): one k_big_step launch with the step's row tiles
        // (row solve, look-ahead update of block b + 1, the next diagonal block), the previous
        // step's copies and its trailing update of blocks b + 2 on (the first step's diagonal
        // block by k_big_diag before; the last step's copies and update in a launch of their own).
        // With panel blocks of pb columns the chain and the trailing tiles stay inside a block
        // [K0, K0 + pb); at each block start K1 = K0 + pb a k_big_block launch updates the next
        // strip [K1, K1 + pb) with K = pb (and factors its diagonal block), and the tiles of the
        // block's update right of that strip ride in the next block's step launches, a share in
        // each (they touch no column those launches read or write).
        const int pb = h->panel_block;
        std::vector<Task> pc, pr;  // the previous step's copies and update tiles
        std::vector<Task> bigq;    // the previous block's update tiles right of the current block
        double pfr = 0, pbr = 0, bigf = 0, bigb = 0;
        int big_left = 0;          // step launches of the current block still to take a share
        // rest_split: a step whose trailing tiles number at least rest_split sends them to a
        // launch of their own right after it (k_big_update_mfma, several workgroups per CU)
        // instead of riding in the step launch (one workgroup per CU beside the chain)
        auto push_step = [&](std::vector<Task>& st, double fs, double bs) {
            std::vector<Task> rest(pr);
            double rf = pfr, rb = pbr;
            if (big_left > 0) {  // this launch's share of the previous block's update
                const size_t take = (bigq.size() + big_left - 1) / big_left;
                const double frac = bigq.empty() ? 0.0 : (double)take / (double)bigq.size();
                rest.insert(rest.end(), bigq.begin(), bigq.begin() + take);
                rf += bigf * frac;
                rb += bigb * frac;
                bigf -= bigf * frac;
                bigb -= bigb * frac;
                bigq.erase(bigq.begin(), bigq.begin() + take);
                --big_left;
            }
            if (st.empty() && pc.empty() && rest.empty()) return;
            h->max_rest = std::max(h->max_rest, (int)rest.size());
            const bool split = h->rest_split > 0 && (int)rest.size() >= h->rest_split;
            std::vector<Task> all(st);
            all.insert(all.end(), pc.begin(), pc.end());
            if (!split) all.insert(all.end(), rest.begin(), rest.end());
            push_launch(L, L_STEP, all, fs + (split ? 0.0 : rf), bs + (split ? 0.0 : rb) + dbl_es * 2.0 * NB * NB * (double)pc.size());
            L.plan.back().nstep = (int)st.size();
            L.plan.back().ncopy = (int)pc.size();
            if (split) push_launch(L, L_UPDATE_PANEL, rest, rf, rb);
        };
        for (int k = 0; k < wmax; k += NB) {
            if (pb > 0 && k > 0 && k % pb == 0) {
                // block [k - pb, k) done: the next strip and diagonal blocks, and the tiles right of
                // the strip for the next block's step launches
                std::vector<Task> none;
                push_step(none, 0, 0);  // (nothing is left over at a block's end)
                std::vector<Task> tdg, tst;
                double fb = 0, bb = 0;
                bigq.clear();
                bigf = bigb = 0;
                const int32_t kc = (int32_t)(k - pb) | BIGK;
                for (int32_t s : L.big) {
                    const int f = h->fnrows[s], w = h->fncols[s];
                    if (k >= w) continue;
                    const int nb = std::min(NB, w - k);
                    tdg.push_back({s, k, k, kc});
                    fb += (double)nb * nb * nb / 3.0 * 2.0;
                    bb += dbl_es * ((double)nb * (nb + 1) + (double)NB * nb);
                    for (int J0 = k; J0 < w; J0 += TILE)
                        for (int I0 = J0; I0 < f; I0 += TILE) {
                            const double e = tile_entries(I0, J0, f, k, w);
                            const double fl2 = 2.0 * e * pb, by2 = 2.0 * dbl_es * e;
                            if (J0 < k + pb) {
                                fb += fl2;
                                bb += by2;
                                if (I0 != k || J0 != k) tst.push_back({s, I0, J0, kc});
                            } else {
                                bigq.push_back({s, I0, J0, kc});
                                bigf += fl2;
                                bigb += by2;
                            }
                        }
                    bb += dbl_es * (double)(f - k + pb) * pb;  // the block's columns read once
                }
                const int ndg = (int)tdg.size();
                tdg.insert(tdg.end(), tst.begin(), tst.end());
                push_launch(L, L_BLOCK, tdg, fb, bb);
                if (!tdg.empty()) L.plan.back().nstep = ndg;
                big_left = std::max(1, (std::min(wmax, k + pb) - k + NB - 1) / NB);
            }
            std::vector<Task> td, tt, tr, tc;
            double fp = 0, ft = 0, fr = 0, bp = 0, bt = 0, br = 0;
            for (int32_t s : L.big) {
                int f = h->fnrows[s], w = h->fncols[s];
                if (k >= w) continue;
                const int K0 = pb > 0 ? k - k % pb : 0, K1 = pb > 0 ? std::min(w, K0 + pb) : w;
                int nb = std::min(NB, w - k);
                if (k == 0) td.push_back({s, 0, 0, 0});
                // the diagonal block's factorization and inverse
                fp += (double)nb * nb * nb / 3.0 * 2.0;
                bp += dbl_es * ((double)nb * (nb + 1) + (double)NB * nb);
                for (int r0 = k + nb; r0 < f; r0 += 64) tt.push_back({s, k, r0, 0});
                ft += 2.0 * (double)(f - k - nb) * nb * nb;
                bt += dbl_es * 2.0 * (double)(f - k - nb) * nb;
                const bool prev = k >= K0 + NB;
                if (k + NB < K1) {  // the look-ahead update of block b + 1 (and b's predecessor's part)
                    const int J = k + NB;
                    tc.push_back({s, k, 0, 0});
                    for (int I0 = J; I0 < f; I0 += TILE) {
                        double e = tile_entries(I0, J, f, J, K1);
                        ft += 2.0 * e * (prev ? 2 * NB : NB);
                        bt += 2.0 * dbl_es * e;
                    }
                    if (J + NB < K1) {  // the diagonal tile after next, by the row tile of its rows
                        double e = tile_entries(J + NB, J + NB, f, J + NB, K1);
                        ft += 2.0 * e * (prev ? 2 * NB : NB);
                        bt += 2.0 * dbl_es * e;
                    }
                }
                const int base = k + 3 * NB;
                for (int I0 = base; I0 < f; I0 += TILE)
                    for (int J0 = base; J0 <= I0 && J0 < K1; J0 += TILE) {
                        // the first diagonal tile gets this step's update from the next step's
                        // row tile of the same rows (with that step's own update)
                        if (I0 == base && J0 == base) continue;
                        tr.push_back({s, I0, J0, k});
                        double e = tile_entries(I0, J0, f, base, K1);
                        fr += 2.0 * e * nb;
                        br += 2.0 * dbl_es * e;
                    }
            }
            push_launch(L, L_POTRF, td, fp, bp);
            push_step(tt, ft, bt);
            pc = tc;
            pr = tr;
            pfr = fr;
            pbr = br;
        }
        {
            std::vector<Task> none;
            push_step(none, 0, 0);
            pc.clear();
            pr.clear();
            pfr = pbr = 0;
            while (big_left > 0 && !bigq.empty()) push_step(none, 0, 0);
        }
        t.clear();
        fl = by = 0;
        for (int32_t s : L.big) {
            int f = h->fnrows[s], w = h->fncols[s];
            if (f == w || cb_off[s] < 0 || h->snparent[s] < 0) continue;
            for (int I0 = w; I0 < f; I0 += TILE)
                for (int J0 = w; J0 < (lu ? f : I0 + 1); J0 += TILE) {
                    t.push_back({s, I0, J0, 0});
                    double e = lu ? (double)std::min(TILE, f - I0) * std::min(TILE, f - J0)
                                  : tile_entries(I0, J0, f, w, f);
                    fl += 2.0 * e * w;
                    by += 2.0 * dbl_es * e;
                }
            by += dbl_es * (double)(f - w) * w;  // L21 read once
        }
        push_launch(L, h->method == M_CHOLESKY ? L_UPDATE_CB : lu ? L_LU_CB : L_LDLT_CB, t, fl, by);
        if (chol) push_extend(L, true, L_EXTEND_ADD);
    }
    if (h->schur_s >= 0) {
        // Schur complement mode: one more level with the root's assembly only (zero, A, the
        // children's blocks; for LU the L and U^T structures), after every factored front
        // (src/numeric/schur.jl:52-78). Its front lists stay empty: the sweeps skip it.
        Level LS;
        const int32_t s = h->schur_s;
        const double f = h->fnrows[s], w = h->fncols[s];
        LS.big.push_back(s);
        std::vector<Task> t;
        const bool lu = h->method == M_LU;
        for (int c = 0; c < h->fnrows[s]; c += 8) t.push_back({s, c, 8, 0});
        const double by = dbl_es * f * w * (lu ? 2.0 : 1.0) + 16.0 * (double)h->asm_count[s];
        if (lu) {
            push_launch(LS, L_LU_ZERO, t, 0, by);
            t.clear();
            for (int c = 0; c < h->fncols[s]; c += 8) t.push_back({s, c, 8, 0});
            push_launch(LS, L_LU_SCATTER, t, 0, 0);
        } else {
            push_launch(LS, L_ZERO_SCATTER, t, 0, by);
        }
        push_extend(LS, false, lu ? L_LU_EXTEND : L_EXTEND_ADD);
        LS.big.clear();
        h->levels.push_back(LS);
    }
    void* p;
    int rc;
    // solve plans of the fronts that are not solve_small: per level, the forward sweep is
    // pull, NB-row block steps (row tiles of 256 below the block), finish; the backward sweep
    // is the L21 gather, NB-row block steps from the last (column tiles of 64), finish.
    // Flops and bytes are per right-hand side (bytes: the factor entries read).
    auto push_solve = [&](std::vector<Launch>& plan, int kind, const std::vector<Task>& t, double flops,
                          double bytes) {
        if (t.empty()) return;
        plan.push_back({kind, (int64_t)tasks.size(), (int)t.size(), flops, bytes});
        tasks.insert(tasks.end(), t.begin(), t.end());
    };
    int64_t l21_units = 0;  // the largest level's backward partial sums, units of nrhs
    for (auto& L : h->levels) {
        std::vector<int32_t> bs;
        for (int32_t s : L.all)
            if (!solve_small(h->fnrows[s], h->fncols[s])) bs.push_back(s);
        if (bs.empty()) continue;
        std::vector<Task> t;
        int wmax = 0;
        for (int32_t s : bs) {
            t.push_back({s, 0, 0, 0});
            wmax = std::max(wmax, h->fncols[s]);
        }
        push_solve(L.fwd_plan, S_FWD_PULL, t, 0, 0);
        for (int k = 0; k < wmax; k += NB) {
            t.clear();
            double fl = 0, by = 0;
            for (int32_t s : bs) {
                int f = h->fnrows[s], w = h->fncols[s];
                if (k >= w) continue;
                int nb = std::min(NB, w - k);
                int first = 1;
                for (int r0 = k + nb; r0 < f; r0 += 256) {
                    t.push_back({s, k, r0, first});
                    first = 0;
                }
                if (first) t.push_back({s, k, f, 1});
                fl += 2.0 * ((double)(f - k - nb) * nb + (double)nb * nb / 2);
                by += dbl_es * ((double)(f - k - nb) * nb + (double)nb * (nb + 1) / 2);
            }
            push_solve(L.fwd_plan, S_FWD_STEP, t, fl, by);
        }
        t.clear();
        for (int32_t s : bs)
            for (int c = 0; c < h->fncols[s]; c += 256) t.push_back({s, c, 0, 0});
        push_solve(L.fwd_plan, S_FWD_FINISH, t, 0, 0);
        // backward: the partial sums of L21^T x per chunk of rows (into the partial-sum buffer,
        // a front's chunks from poff on), then their sum per column in chunk order
        t.clear();
        std::vector<Task> tsum;
        double fl = 0, by = 0;
        int64_t poff = 0;
        const int rc = bwd_l21_chunk();
        for (int32_t s : bs) {
            int f = h->fnrows[s], w = h->fncols[s];
            if (f == w) continue;
            const int nch = (f - w + rc - 1) / rc;
            for (int c = 0; c < w; c += 64) {
                for (int q = 0; q < nch; ++q) t.push_back({s, c, q, (int32_t)poff});
                tsum.push_back({s, c, nch, (int32_t)poff});
            }
            poff += (int64_t)nch * w;
            fl += 2.0 * (double)(f - w) * w;
            by += dbl_es * (double)(f - w) * w;
        }
        l21_units = std::max(l21_units, poff);
        push_solve(L.bwd_plan, S_BWD_L21, t, fl, by);
        push_solve(L.bwd_plan, S_BWD_L21_SUM, tsum, 0, 0);
        for (int k = ((wmax - 1) / NB) * NB; k >= 0; k -= NB) {
            t.clear();
            fl = by = 0;
            for (int32_t s : bs) {
                int w = h->fncols[s];
                if (k >= w) continue;
                int nb = std::min(NB, w - k);
                int first = 1;
                for (int c = 0; c < k; c += 64) {
                    t.push_back({s, k, c, first});
                    first = 0;
                }
                if (first) t.push_back({s, k, 0, 1});
                fl += 2.0 * ((double)k * nb + (double)nb * nb / 2);
                by += dbl_es * ((double)k * nb + (double)nb * (nb + 1) / 2);
            }
            push_solve(L.bwd_plan, S_BWD_STEP, t, fl, by);
        }
        t.clear();
        for (int32_t s : bs)
            for (int c = 0; c < h->fncols[s]; c += 256) t.push_back({s, c, 0, 0});
        push_solve(L.bwd_plan, S_BWD_FINISH, t, 0, 0);
    }
    // Cholesky: the diagonal-block inverses and the inverse-block sweep plans
    h->part_units = h->method == M_CHOLESKY ? 0 : l21_units;
    if (h->method == M_CHOLESKY) {
        std::vector<Task> t[3];
        double fl[3] = {0, 0, 0}, by[3] = {0, 0, 0};
        // big fronts: done by k_big_diag; small fronts of at most FUSED_INV_MAX columns: by their
        // factorization
        for (int64_t s = 0; s < ns; ++s)
            for (int k = 0; k < h->fncols[s] && !big[s] && h->fncols[s] > FUSED_INV_MAX; k += NB) {
                const int nbi = std::min(NB, h->fncols[s] - k), cls = nbi <= 16 ? 0 : nbi <= 32 ? 1 : 2;
                t[cls].push_back({(int32_t)s, k, 0, 0});
                const double nb = nbi;
                fl[cls] += nb * nb * nb / 3.0;
                by[cls] += dbl_es * (nb * (nb + 1) / 2 + (double)NB * nb);
            }
        for (int cls = 0; cls < 3; ++cls) {
            h->inv_launch[cls] = {L_POTRF, (int64_t)tasks.size(), (int)t[cls].size(), fl[cls], by[cls]};
            tasks.insert(tasks.end(), t[cls].begin(), t[cls].end());
        }
        std::vector<SolveStep> steps;
        std::vector<Task> ts;  // (step, column block) of the wide steps' inverses
        int64_t xlen = 0;
        double sfl = 0, sby = 0;
        auto put = [&](std::vector<Task>& v, int kind, double fl2, double by2) {
            Launch ln{kind, (int64_t)tasks.size(), (int)v.size(), fl2, by2};
            tasks.insert(tasks.end(), v.begin(), v.end());
            return ln;
        };
        for (auto& L : h->levels) {
            int wmax = 0;
            for (int32_t s : L.all) wmax = std::max(wmax, h->fncols[s]);
            for (int g = 0; g * SOLVE_STEP < wmax; ++g) {
                std::vector<Task> td, tr, tt, tfw, tbw, tcp, tcpf;
                double e11s = 0, efused = 0, erows = 0;  // factor entries read (see below)
                int threads = 64;  // narrow steps: one diagonal block (more threads for fused rows)
                int64_t units = 0;
                for (int32_t s : L.all) {
                    const int f = h->fnrows[s], w = h->fncols[s], m = f - w;
                    const int c0 = g * SOLVE_STEP, c1 = std::min(w, c0 + SOLVE_STEP);
                    if (c0 >= w) continue;
                    const int nw = c1 - c0;
                    const bool fused = w <= NB && m <= FUSE_ROWS;
                    const int mb = f - c1;  // rows below the step
                    const int nch = mb > 0 ? (mb + L21T_CHUNK - 1) / L21T_CHUNK : 0;
                    const int32_t id = (int32_t)steps.size();
                    const bool wide = nw > NB;
                    steps.push_back({s, c0, c1, nch, units, wide ? xlen : -1});
                    units += (int64_t)nch * nw;
                    // a wide step's rows below read y_step from Z and the first of their tasks copies it
                    // to Y (no copy launch in the forward sweep); without rows below, a copy task
                    const bool has_rows = !fused && (c1 < w || (m > 0 && uv_off[s] >= 0));
                    if (wide) {
                        for (int b = 0; b * NB < nw; ++b) {
                            ts.push_back({id, b, 0, 0});
                            tcp.push_back({id, b * NB, std::min(NB, nw - b * NB), 0});
                            if (!has_rows) tcpf.push_back(tcp.back());
                        }
                        for (int r = 0; r < nw; r += 16) tfw.push_back({id, r, 0, 0});
                        for (int c = 0; c < nw; c += 8) tbw.push_back({id, c, 0, 0});
                        xlen += (int64_t)nw * nw;
                        const double w3 = (double)nw * nw * nw;
                        sfl += w3 / 3.0;
                        sby += dbl_es * ((double)nw * nw / 2 + (double)nw * nw);
                    } else {
                        td.push_back({id, fused ? 1 : 0, 0, 0});
                        if (f > 64) threads = 256;
                    }
                    // factor entries read: the step's triangle and inverse blocks, the rows below
                    e11s += (double)nw * (nw + 1) / 2 + (double)NB * nw;
                    (fused ? efused : erows) += (double)mb * nw;
                    if (!fused) {
                        const size_t t0 = tr.size();
                        for (int r0 = c1; r0 < w; r0 += 64) tr.push_back({id, r0, 0, wide ? 1 : 0});
                        if (m > 0 && uv_off[s] >= 0)
                            for (int r0 = w; r0 < f; r0 += 64) tr.push_back({id, r0, 0, wide ? 1 : 0});
                        if (wide && tr.size() > t0) tr[t0].b = 1;
                    }
                    for (int c = c0; c < c1 && nch; c += 8)
                        for (int ch = 0; ch < nch; ++ch) tt.push_back({id, c, ch, 0});
                }
                h->part_units = std::max(h->part_units, units);
                Level::StepGroup G{};
                G.threads = threads;
                // flops per right-hand side and factor bytes of each launch
                G.fdiag = put(td, S_FWD_STEP, 2.0 * (e11s + efused), dbl_es * (e11s + efused));
                G.frows = put(tr, S_FWD_STEP, 2.0 * erows, dbl_es * erows);
                G.bl21t = put(tt, S_BWD_L21, 2.0 * (efused + erows), dbl_es * (efused + erows));
                G.bdiag = put(td, S_BWD_STEP, 2.0 * e11s, dbl_es * e11s);
                G.fwide = put(tfw, S_FWD_STEP, 0, 0);
                G.bwide = put(tbw, S_BWD_STEP, 0, 0);
                G.copy = put(tcp, S_FWD_FINISH, 0, 0);
                G.fcopy = put(tcpf, S_FWD_FINISH, 0, 0);
                L.groups.push_back(G);
            }
        }
        if ((rc = upload(h, h->numbufs, steps, &h->d_steps))) return rc;
        h->sinv_launch = put(ts, L_POTRF, sfl, sby);
        if ((rc = dev_alloc(h, h->numbufs, es * std::max<int64_t>(xlen, 1), &h->xbuf))) return rc;
    }
    if ((rc = upload(h, h->numbufs, flist, &h->d_fronts))) return rc;
    if ((rc = upload(h, h->numbufs, tasks, &h->d_tasks))) return rc;
    if (!h->sub_launch.empty()) {
        if ((rc = upload(h, h->numbufs, h->sub_ptr, &h->d_sub_ptr))) return rc;
        if ((rc = upload(h, h->numbufs, h->sub_nodes, &h->d_sub_nodes))) return rc;
        if ((rc = upload(h, h->numbufs, h->sub_lfront, &h->d_sub_lfront))) return rc;
        if ((rc = upload(h, h->numbufs, h->sub_lcb, &h->d_sub_lcb))) return rc;
    }
    if ((rc = upload(h, h->numbufs, cb_off, &h->d_cb_off))) return rc;
    if ((rc = upload(h, h->numbufs, uv_off, &h->d_uv_off))) return rc;
    h->dsym.cb_off = h->d_cb_off;
    h->dsym.uv_off = h->d_uv_off;
    if (h->method == M_CHOLESKY) {
        // the forward pull map: per front row, its children's update-vector rows in child_list
        // order (one gather per row in the forward sweep, no search and no barrier per child)
        std::vector<int64_t> pm_row(ns + 1, 0);
        for (int64_t s = 0; s < ns; ++s) pm_row[s + 1] = pm_row[s] + h->fnrows[s];
        std::vector<int64_t> cnt(pm_row[ns] + 1, 0);
        for (int64_t s = 0; s < ns; ++s)
            for (int kc = h->child_ptr[s]; kc < h->child_ptr[s + 1]; ++kc) {
                const int32_t c = h->child_list[kc];
                if (uv_off[c] < 0) continue;
                for (int64_t q = h->relind_ptr[c]; q < h->relind_ptr[c + 1]; ++q) cnt[pm_row[s] + h->relind[q] + 1] += 1;
            }
        for (int64_t r = 0; r < pm_row[ns]; ++r) cnt[r + 1] += cnt[r];
        if (cnt[pm_row[ns]] >= ((int64_t)1 << 31) || h->uv_len >= ((int64_t)1 << 31))
            return fail(h, SDS_NOT_SUPPORTED, "solve pull map beyond 2^31 entries");
        std::vector<int32_t> pm_ptr(cnt.begin(), cnt.end()), pm_src(std::max<int64_t>(cnt[pm_row[ns]], 1));
        for (int64_t s = 0; s < ns; ++s)
            for (int kc = h->child_ptr[s]; kc < h->child_ptr[s + 1]; ++kc) {
                const int32_t c = h->child_list[kc];
                if (uv_off[c] < 0) continue;
                for (int64_t q = h->relind_ptr[c]; q < h->relind_ptr[c + 1]; ++q)
                    pm_src[cnt[pm_row[s] + h->relind[q]]++] = (int32_t)(uv_off[c] + (q - h->relind_ptr[c]));
            }
        pm_row.resize(ns);
        if ((rc = upload(h, h->numbufs, pm_row, &h->d_pm_row))) return rc;
        if ((rc = upload(h, h->numbufs, pm_ptr, &h->d_pm_ptr))) return rc;
        if ((rc = upload(h, h->numbufs, pm_src, &h->d_pm_src))) return rc;
        h->dsym.pm_row = h->d_pm_row;
        h->dsym.pm_ptr = h->d_pm_ptr;
        h->dsym.pm_src = h->d_pm_src;
    }
    (void)p;
    return SDS_SUCCESS;
}

template <typename T>
int factorize_cholesky(sds_handle* h, const double* nz, double eps, const volatile uint8_t* interrupt) {
    hipStream_t st = h->stream;
    T* factor = static_cast<T*>(h->factor);
    T* cb = static_cast<T*>(h->cbpool);
    launch_fill_i32(st, h->status, h->ns, 0);
    // regime A: the leaf subtrees, one workgroup each, one launch per LDS class
    for (const auto& sl : h->sub_launch)
        timed(h, SDS_K_SUBTREE, sl.flops, sl.bytes, [&] {
            launch_subtree_chol<T>(st, h->dsym, h->d_sub_ptr, h->d_sub_nodes, sl.first, sl.count, sl.lds,
                                   h->d_sub_lfront, h->d_sub_lcb, nz, factor, cb, static_cast<T*>(h->dbuf), h->status,
                                   eps);
        });
    for (auto& L : h->levels) {
        if (interrupt && *interrupt) return fail(h, SDS_INTERRUPTED, "factorization interrupted");
        if (!L.small.empty()) {
            // small_kind 0: fronts of at most TINY_F rows one wave each, the others blocked;
            // 1: one wave each / column by column; 2: all column by column
            T* DB = static_cast<T*>(h->dbuf);
            const bool tiny = h->small_kind != 2;
            if (tiny && !L.tiny.empty())
                timed(h, SDS_K_FRONT_SMALL, L.tiny_flops, L.tiny_bytes, [&] {
                    launch_front_tiny_chol<T>(st, h->dsym, h->d_fronts + L.off_tiny, (int)L.tiny.size(), nz, factor, cb,
                                              DB, h->status, eps);
                });
            const int32_t* fl = h->d_fronts + (tiny ? L.off_mid : L.off_small);
            const int nf = (int)(tiny ? L.mid.size() : L.small.size());
            const int fm = tiny ? L.fmax_mid : L.fmax_small;
            if (nf)
                timed(h, SDS_K_FRONT_SMALL, tiny ? L.small_flops - L.tiny_flops : L.small_flops,
                      tiny ? L.small_bytes - L.tiny_bytes : L.small_bytes, [&] {
                          if (h->small_kind == 0)
                              launch_front_small_chol<T>(st, h->dsym, fl, nf, fm, nz, factor, cb, DB, h->status, eps);
                          else
                              launch_front_col_chol<T>(st, h->dsym, fl, nf, fm, nz, factor, cb, DB, h->status, eps);
                      });
        }
        for (const Launch& ln : L.plan) {
            const Task* t = h->d_tasks + ln.off;
            timed(h, prof_kind(ln.kind), ln.flops, ln.bytes, [&] {
                switch (ln.kind) {
                    case L_ZERO_SCATTER: launch_big_zero_scatter<T>(st, h->dsym, t, ln.n, nz, factor, cb); break;
                    case L_EXTEND_ADD: launch_big_extend_add<T>(st, h->dsym, t, ln.n, factor, cb); break;
                    case L_POTRF:
                        launch_big_diag<T>(st, h->dsym, t, ln.n, factor, static_cast<T*>(h->dbuf), h->status, eps);
                        break;
                    case L_STEP:
                        launch_big_step<T>(st, h->dsym, t, ln.n, ln.nstep, ln.ncopy, factor, cb,
                                           static_cast<T*>(h->dbuf), static_cast<T*>(h->lnbuf), h->status, eps,
                                           h->step_occ2 > 0 && ln.n >= h->step_occ2 ? 2 : 1, h->panel_block);
                        break;
                    case L_BLOCK:
                        launch_big_block<T>(st, h->dsym, t, ln.n, ln.nstep, factor, cb, static_cast<T*>(h->dbuf),
                                            h->status, eps, h->panel_block);
                        break;
                    case L_UPDATE_PANEL:
                        launch_big_update<T>(st, h->dsym, t, ln.n, factor, cb, h->status, 0, h->upd_kc,
                                             h->panel_block);
                        break;
                    case L_UPDATE_CB:
                        launch_big_update<T>(st, h->dsym, t, ln.n, factor, cb, h->status, 1, h->upd_kc,
                                             h->panel_block);
                        break;
                }
            });
        }
    }
    timed(h, SDS_K_STATS, 0, 0, [&] { launch_chol_stats(st, h->dsym, h->status, h->stats, h->d_info); });
    if (h->schur_s >= 0)
        launch_schur_meta<T>(st, h->dsym, h->schur_s, (int)h->nschur, nullptr, nullptr, static_cast<T*>(h->dvec),
                             h->stats);
    // the inverses of the diagonal blocks for the sweeps (a failed factorization gives
    // inverses of whatever the panels hold; its solves are not meaningful anyway)
    for (int cls = 0; cls < 3; ++cls) {
        const Launch& il = h->inv_launch[cls];
        if (il.n)
            timed(h, SDS_K_INVERSE, il.flops, il.bytes, [&] {
                launch_block_inverse<T>(st, h->dsym, h->d_tasks + il.off, il.n, factor, static_cast<T*>(h->dbuf),
                                        cls == 0 ? 16 : cls == 1 ? 32 : NB);
            });
    }
    if (h->sinv_launch.n)
        timed(h, SDS_K_INVERSE, h->sinv_launch.flops, h->sinv_launch.bytes, [&] {
            launch_step_inverse<T>(st, h->dsym, h->d_steps, h->d_tasks + h->sinv_launch.off, h->sinv_launch.n, factor,
                                   static_cast<const T*>(h->dbuf), static_cast<T*>(h->xbuf));
        });
    HIPCHK(h, hipGetLastError());
    return SDS_SUCCESS;
}

// LDL^T (src/numeric/ldlt.jl:1447-1475 drives the same steps): pivot 0 none, 1 diagonal,
// 2 Bunch-Kaufman; `scaled`: eps times max |a_ij| (pivot_epsilon_alg = "algo1")
template <typename T>
int factorize_ldlt(sds_handle* h, const double* nz, int pivot, double u, double eps, int scaled,
                   const int8_t* psign, const volatile uint8_t* interrupt) {
    hipStream_t st = h->stream;
    T* factor = static_cast<T*>(h->factor);
    T* cb = static_cast<T*>(h->cbpool);
    T* dv = static_cast<T*>(h->dvec);
    if (scaled) {
        launch_fill_i32(st, reinterpret_cast<int32_t*>(h->aux), 2, 0);
        launch_absmax(st, h->nnz, nz, h->aux);
    }
    for (auto& L : h->levels) {
        if (interrupt && *interrupt) return fail(h, SDS_INTERRUPTED, "factorization interrupted");
        if (!L.small.empty())
            timed(h, SDS_K_FRONT_SMALL, L.small_flops, L.small_bytes, [&] {
                launch_front_small_ldlt<T>(st, h->dsym, h->d_fronts + L.off_small, (int)L.small.size(), L.fmax_small,
                                           nz, factor, cb, dv, h->piv, h->pkind, h->stats, pivot, u, eps, scaled,
                                           h->aux, psign);
            });
        for (const Launch& ln : L.plan) {
            const Task* t = h->d_tasks + ln.off;
            timed(h, prof_kind(ln.kind), ln.flops, ln.bytes, [&] {
                switch (ln.kind) {
                    case L_ZERO_SCATTER: launch_big_zero_scatter<T>(st, h->dsym, t, ln.n, nz, factor, cb); break;
                    case L_EXTEND_ADD: launch_big_extend_add<T>(st, h->dsym, t, ln.n, factor, cb); break;
                    case L_LDLT_PIVOT:
                        launch_ldlt_pivot<T>(st, h->dsym, t, ln.n, ln.k, factor, dv, h->piv, h->pkind, h->stats,
                                             h->lstep, pivot, u, eps, scaled, h->aux, psign);
                        break;
                    case L_LDLT_UPDATE: launch_ldlt_update<T>(st, h->dsym, t, ln.n, factor, h->lstep); break;
                    case L_LDLT_MIRROR: launch_ldlt_mirror<T>(st, h->dsym, t, ln.n, factor); break;
                    case L_LDLT_PANEL:
                        launch_ldlt_panel<T>(st, h->dsym, t, ln.n, ln.k, factor, dv, h->piv, h->pkind, h->stats,
                                             h->lstep, static_cast<T*>(h->lscratch), pivot, u, eps, scaled, h->aux,
                                             psign);
                        break;
                    case L_LDLT_TRAIL:
                        launch_ldlt_trail<T>(st, h->dsym, t, ln.n, factor, h->pkind, h->lstep,
                                             static_cast<const T*>(h->lscratch));
                        break;
                    case L_LDLT_FINISH: launch_ldlt_finish<T>(st, h->dsym, t, ln.n, factor, h->piv); break;
                    case L_LDLT_CB: launch_ldlt_cb<T>(st, h->dsym, t, ln.n, factor, cb, dv, h->pkind); break;
                }
            });
        }
    }
    if (h->schur_s >= 0)
        launch_schur_meta<T>(st, h->dsym, h->schur_s, (int)h->nschur, h->piv, h->pkind, static_cast<T*>(h->dvec),
                             h->stats);
    timed(h, SDS_K_STATS, 0, 0, [&] { launch_info_from_stats(st, h->dsym, h->stats, h->d_info); });
    HIPCHK(h, hipGetLastError());
    return SDS_SUCCESS;
}

// LU of "G" (src/numeric/lu.jl:543-555 drives the same steps)
template <typename T>
int factorize_lu(sds_handle* h, const double* nz, int pivoting, double u, double eps, int scaled,
                 const volatile uint8_t* interrupt) {
    hipStream_t st = h->stream;
    T* factor = static_cast<T*>(h->factor);
    T* uf = static_cast<T*>(h->ufactor);
    T* cb = static_cast<T*>(h->cbpool);
    T* dv = static_cast<T*>(h->dvec);
    if (scaled) {
        launch_fill_i32(st, reinterpret_cast<int32_t*>(h->aux), 2, 0);
        launch_absmax(st, h->nnz, nz, h->aux);
    }
    for (auto& L : h->levels) {
        if (interrupt && *interrupt) return fail(h, SDS_INTERRUPTED, "factorization interrupted");
        if (!L.small.empty())
            timed(h, SDS_K_FRONT_SMALL, L.small_flops, L.small_bytes, [&] {
                launch_front_small_lu<T>(st, h->dsym, h->d_fronts + L.off_small, (int)L.small.size(), L.fmax_small,
                                         nz, factor, uf, cb, dv, h->piv, h->pkind, h->stats, pivoting, u, eps, scaled,
                                         h->aux);
            });
        for (const Launch& ln : L.plan) {
            const Task* t = h->d_tasks + ln.off;
            timed(h, prof_kind(ln.kind), ln.flops, ln.bytes, [&] {
                switch (ln.kind) {
                    case L_LU_ZERO: launch_big_lu_zero<T>(st, h->dsym, t, ln.n, factor, uf, cb); break;
                    case L_LU_SCATTER: launch_big_lu_scatter<T>(st, h->dsym, t, ln.n, nz, factor, uf); break;
                    case L_LU_EXTEND: launch_big_lu_extend_add<T>(st, h->dsym, t, ln.n, factor, uf, cb); break;
                    case L_LU_PIVOT:
                        launch_lu_pivot<T>(st, h->dsym, t, ln.n, ln.k, factor, uf, dv, h->piv, h->pkind, h->stats,
                                           pivoting, u, eps, scaled, h->aux);
                        break;
                    case L_LU_UPDATE: launch_lu_update<T>(st, h->dsym, t, ln.n, ln.k, factor, uf, dv); break;
                    case L_LU_FINISH: launch_lu_finish<T>(st, h->dsym, t, ln.n, factor, uf, dv, h->piv); break;
                    case L_LU_CB: launch_lu_cb<T>(st, h->dsym, t, ln.n, factor, uf, cb, dv); break;
                }
            });
        }
    }
    if (h->schur_s >= 0)
        launch_schur_meta<T>(st, h->dsym, h->schur_s, (int)h->nschur, h->piv, h->pkind, static_cast<T*>(h->dvec),
                             h->stats);
    timed(h, SDS_K_STATS, 0, 0, [&] { launch_info_from_stats(st, h->dsym, h->stats, h->d_info); });
    HIPCHK(h, hipGetLastError());
    return SDS_SUCCESS;
}

int ensure_solve_workspace(sds_handle* h, int nrhs) {
    if (nrhs <= h->cap_nrhs) return SDS_SUCCESS;
    size_t es = h->dtype == SDS_R64 ? 8 : 4;
    for (DevBuf* b : {&h->Y, &h->Z, &h->U, &h->R, &h->Bc, &h->norms, &h->part}) {
        if (b->p) (void)hipFree(b->p);
        *b = DevBuf{};
    }
    auto get = [&](DevBuf& b, size_t bytes) -> int {
        bytes = std::max<size_t>(bytes, 8);
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) return fail(h, SDS_DEVICE_ERROR, std::string("hipMalloc: ") + hipGetErrorString(e));
        b.bytes = bytes;
        return SDS_SUCCESS;
    };
    int rc;
    if ((rc = get(h->Y, es * h->n * nrhs))) return rc;
    if ((rc = get(h->Z, es * h->n * nrhs))) return rc;
    if ((rc = get(h->U, es * std::max<int64_t>(h->uv_len, 1) * nrhs))) return rc;
    if ((rc = get(h->R, 8 * h->n * nrhs))) return rc;
    if ((rc = get(h->Bc, 8 * h->n * nrhs))) return rc;
    if ((rc = get(h->norms, 8 * (2 * MAX_NRHS + DOT_BLOCKS)))) return rc;
    if ((rc = get(h->part, es * std::max<int64_t>(h->part_units, 1) * nrhs))) return rc;
    h->cap_nrhs = nrhs;
    return SDS_SUCCESS;
}

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};

// Device address of the handle's GPU's scratch array, looked up with that GPU current on the
// first call for the device.
int scratch_base(sds_handle* h, void** out) {
    if (h->device < 0 || h->device >= MAX_DEVICES) return fail(h, SDS_NOT_SUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[h->device]) {
        int cur = -1;
        HIPCHK(h, hipGetDevice(&cur));
        if (cur != h->device) HIPCHK(h, hipSetDevice(h->device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != h->device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(h, SDS_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        scratch_ptr[h->device] = p;
    }
    *out = scratch_ptr[h->device];
    return SDS_SUCCESS;
}

// The solve workspace for nrhs right-hand sides laid out from `base` (256-byte aligned
// pieces, the sizes of ensure_solve_workspace); returns the bytes it spans.
size_t noalloc_layout(const sds_handle* h, int nrhs, char* base, SolveWork* W) {
    const size_t es = h->dtype == SDS_R64 ? 8 : 4;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base + off;
        off += (std::max<size_t>(bytes, 8) + 255) & ~size_t(255);
        return static_cast<void*>(p);
    };
    SolveWork w;
    w.Y = take(es * h->n * nrhs);
    w.Z = take(es * h->n * nrhs);
    w.U = take(es * std::max<int64_t>(h->uv_len, 1) * nrhs);
    w.R = static_cast<double*>(take(8 * h->n * nrhs));
    w.Bc = static_cast<double*>(take(8 * h->n * nrhs));
    w.norms = static_cast<double*>(take(8 * (2 * MAX_NRHS + DOT_BLOCKS)));
    w.part = take(es * std::max<int64_t>(h->part_units, 1) * nrhs);
    w.hnorms = h->h_norms;
    if (W) *W = w;
    return off;
}

SolveWork handle_work(sds_handle* h) {
    SolveWork W;
    W.Y = h->Y.p;
    W.Z = h->Z.p;
    W.U = h->U.p;
    W.part = h->part.p;
    W.R = static_cast<double*>(h->R.p);
    W.Bc = static_cast<double*>(h->Bc.p);
    W.norms = static_cast<double*>(h->norms.p);
    W.hnorms = h->h_norms;
    return W;
}

// The forward sweep runs with the lower panels `ffac` (the local row order `fpiv` applied
// first), the backward sweep with the transposes of `bfac` (`bpiv` undone last):
// Cholesky L / L; LDL^T L + piv / L + piv; LU L + piv / U^T; LU transposed U^T / L + piv.
template <typename T>
void sweeps(sds_handle* h, const SolveWork& W, int nrhs, int fwd, int diag, int bwd, int transposed) {
    hipStream_t st = h->stream;
    const T* L = static_cast<const T*>(h->factor);
    const T* Ut = h->ufactor ? static_cast<const T*>(h->ufactor) : L;
    T* Y = static_cast<T*>(W.Y);
    T* Z = static_cast<T*>(W.Z);
    T* U = static_cast<T*>(W.U);
    const int unit = h->method == M_CHOLESKY ? 0 : 1;
    const int32_t* pv = h->method == M_CHOLESKY ? nullptr : h->piv;
    const bool lu = h->method == M_LU;
    const T* ffac = lu && transposed ? Ut : L;
    const T* bfac = lu && !transposed ? Ut : L;
    const int32_t* fpiv = lu && transposed ? nullptr : pv;
    const int32_t* bpiv = lu && !transposed ? nullptr : pv;
    const double es = sizeof(T);
    if (h->method == M_CHOLESKY) {
        // inverted diagonal blocks (DB) and level launches over all fronts
        const T* Li = static_cast<const T*>(h->dbuf);
        T* part = static_cast<T*>(W.part);
        const SolveStep* stp = h->d_steps;
        const T* Xb = static_cast<const T*>(h->xbuf);
        const Task* tk = h->d_tasks;
        if (fwd)
            for (auto& Lv : h->levels)
                for (auto& G : Lv.groups) {
                    if (G.fdiag.n && G.fwide.n) {
                        // narrow and wide steps of the group (different fronts) in one launch
                        timed(h, SDS_K_FWD, G.fdiag.flops * nrhs, G.fdiag.bytes, [&] {
                            launch_fwd_group<T>(st, h->dsym, stp, tk + G.fdiag.off, G.fdiag.n, tk + G.fwide.off,
                                                G.fwide.n, L, Li, Xb, Y, Z, U, nrhs);
                        });
                    } else if (G.fdiag.n) {
                        timed(h, SDS_K_FWD, G.fdiag.flops * nrhs, G.fdiag.bytes, [&] {
                            launch_fwd_diag<T>(st, h->dsym, stp, tk + G.fdiag.off, G.fdiag.n, G.threads, L, Li, Y, U,
                                               nrhs);
                        });
                    } else if (G.fwide.n) {
                        timed(h, SDS_K_FWD, 0, 0, [&] {
                            launch_fwd_inv<T>(st, h->dsym, stp, tk + G.fwide.off, G.fwide.n, Xb, Y, Z, U, nrhs);
                        });
                    }
                    if (G.fcopy.n)
                        timed(h, SDS_K_FWD, 0, 0, [&] {
                            launch_copy_rows<T>(st, h->dsym, stp, tk + G.fcopy.off, G.fcopy.n, Y, Z, nrhs);
                        });
                    if (G.frows.n)
                        timed(h, SDS_K_FWD, G.frows.flops * nrhs, G.frows.bytes, [&] {
                            launch_fwd_rows<T>(st, h->dsym, stp, tk + G.frows.off, G.frows.n, L, Y, Z, U, nrhs);
                        });
                }
        if (fwd && h->schur_s >= 0)
            timed(h, SDS_K_FWD, 0, 0, [&] { launch_schur_pull<T>(st, h->dsym, h->schur_s, Y, U, nrhs); });
        if (bwd)
            for (auto it = h->levels.rbegin(); it != h->levels.rend(); ++it)
                for (auto G = it->groups.rbegin(); G != it->groups.rend(); ++G) {
                    if (G->bl21t.n)
                        timed(h, SDS_K_BWD, G->bl21t.flops * nrhs, G->bl21t.bytes, [&] {
                            launch_bwd_l21t<T>(st, h->dsym, stp, tk + G->bl21t.off, G->bl21t.n, L, Y, part, nrhs);
                        });
                    // narrow steps, and the wide steps' t = y_step less the partial sums into Z (one
                    // launch when the group has both), then the wide steps' x = inverse^T t into y
                    if (G->bdiag.n && G->bwide.n) {
                        timed(h, SDS_K_BWD, G->bdiag.flops * nrhs, G->bdiag.bytes, [&] {
                            launch_bwd_group<T>(st, h->dsym, stp, tk + G->bdiag.off, G->bdiag.n, tk + G->copy.off,
                                                G->copy.n, L, Li, Y, Z, part, nrhs);
                        });
                    } else if (G->bdiag.n) {
                        timed(h, SDS_K_BWD, G->bdiag.flops * nrhs, G->bdiag.bytes, [&] {
                            launch_bwd_diag<T>(st, h->dsym, stp, tk + G->bdiag.off, G->bdiag.n, 64, L, Li, Y, part,
                                               nrhs);
                        });
                    } else if (G->bwide.n) {
                        timed(h, SDS_K_BWD, 0, 0, [&] {
                            launch_sub_partials<T>(st, h->dsym, stp, tk + G->copy.off, G->copy.n, Y, Z, part,
                                                   nrhs);
                        });
                    }
                    if (G->bwide.n) {
                        timed(h, SDS_K_BWD, 0, 0, [&] {
                            launch_bwd_inv<T>(st, h->dsym, stp, tk + G->bwide.off, G->bwide.n, Xb, Y, Z, nrhs);
                        });
                    }
                }
        return;
    }
    auto run = [&](const Launch& ln, int kind) {
        const Task* t = h->d_tasks + ln.off;
        timed(h, kind, ln.flops * nrhs, ln.bytes + (ln.flops > 0 ? es * nrhs * ln.flops / 16 : 0), [&] {
            switch (ln.kind) {
                case S_FWD_PULL: launch_fwd_pull<T>(st, h->dsym, t, ln.n, Y, U, Z, nrhs, fpiv); break;
                case S_FWD_STEP: launch_fwd_step<T>(st, h->dsym, t, ln.n, ffac, Y, U, Z, nrhs, unit); break;
                case S_FWD_FINISH: launch_finish<T>(st, h->dsym, t, ln.n, Y, Z, nrhs, nullptr); break;
                case S_BWD_L21:
                    launch_bwd_l21<T>(st, h->dsym, t, ln.n, bfac, Y, static_cast<T*>(W.part), nrhs);
                    break;
                case S_BWD_L21_SUM:
                    launch_bwd_l21_sum<T>(st, h->dsym, t, ln.n, Y, static_cast<const T*>(W.part), nrhs);
                    break;
                case S_BWD_STEP: launch_bwd_step<T>(st, h->dsym, t, ln.n, bfac, Y, Z, nrhs, unit); break;
                case S_BWD_FINISH: launch_finish<T>(st, h->dsym, t, ln.n, Y, Z, nrhs, bpiv); break;
            }
        });
    };
    if (fwd)
        for (auto& L : h->levels) {
            if (!L.ssmall.empty())
                timed(h, SDS_K_FWD, L.fwd_flops * nrhs, L.fwd_bytes + es * nrhs * L.rhs_rows, [&] {
                    launch_fwd<T>(st, h->dsym, h->d_fronts + L.off_ssmall, (int)L.ssmall.size(), ffac, Y, U, Z,
                                  nrhs, unit, fpiv);
                });
            for (const Launch& ln : L.fwd_plan) run(ln, SDS_K_FWD);
        }
    if (fwd && h->schur_s >= 0)
        timed(h, SDS_K_FWD, 0, 0, [&] { launch_schur_pull<T>(st, h->dsym, h->schur_s, Y, U, nrhs); });
    if (diag && h->method != M_CHOLESKY)
        timed(h, SDS_K_DIAG, 2.0 * h->n * nrhs, es * (2.0 * h->n + 2.0 * h->n * nrhs) + h->n, [&] {
            launch_diag<T>(st, (int)h->n, nrhs, static_cast<const T*>(h->dvec), lu ? nullptr : h->pkind, Y);
        });
    if (bwd)
        for (auto it = h->levels.rbegin(); it != h->levels.rend(); ++it) {
            if (!it->ssmall.empty())
                timed(h, SDS_K_BWD, it->fwd_flops * nrhs, it->fwd_bytes + es * nrhs * it->rhs_rows, [&] {
                    launch_bwd<T>(st, h->dsym, h->d_fronts + it->off_ssmall, (int)it->ssmall.size(), bfac, Y, Z,
                                  nrhs, unit, bpiv);
                });
            for (const Launch& ln : it->bwd_plan) run(ln, SDS_K_BWD);
        }
}

// FGMRES refinement (ir_mode "fgmres", src/solve/refinement.jl:554-620 with the Krylov.jl FGMRES
// it drives): R0 = B - op(A) X; with tol > 0, nothing to do when every column has
// ||R0_k|| <= tol ||B_k||; else FGMRES on op(A) D = R0 over the n nrhs stacked columns, right
// preconditioned by the factorization (one full solve per iteration), at most nsteps iterations,
// stopping once ||R0 - op(A) D|| <= tol min_k ||B_k|| (tol = 0: all nsteps, or an exact
// residual); then X += D. Arnoldi by modified Gram-Schmidt with the coefficients on the device;
// the Hessenberg least-squares problem by Givens rotations on the host, one synchronization per
// iteration (the reference's dot products synchronize too). Returns the iterations performed.
template <typename T>
int fgmres_t(sds_handle* h, const SolveWork& W, int transposed, const double* Bsrc, int64_t ldbs, double* X,
             int64_t ldx, int nrhs, const double* nz, int m, double tol, int* iters) {
    // the matching's sides (src/solver.jl:583-604): A x = b solves M y = Dr b with M = (Dr A Dc)[:, q]
    // and x = Dc q y; A^T x = b the transposed way round (no matching: none of them)
    const int32_t* IN_CQ = transposed ? h->mt_q : nullptr;
    const double* IN_SC = transposed ? h->mt_cs : h->mt_rs;
    const int32_t* OUT_CQ = transposed ? nullptr : h->mt_q;
    const double* OUT_SC = transposed ? h->mt_rs : h->mt_cs;
    hipStream_t st = h->stream;
    const int n = (int)h->n;
    const int64_t N = (int64_t)n * nrhs;
    *iters = 0;
    // V_0..V_m, Z_0..Z_{m-1}, the Hessenberg columns (m + 2 per column), DOT_BLOCKS partials, then
    // ns scalars: the 2 nrhs column norms, later the m solution coefficients
    const int hc = m + 2, ns = std::max(2 * nrhs, m) + 8;
    const size_t need = sizeof(double) * ((size_t)(2 * m + 1) * N + (size_t)hc * (m + 1) + DOT_BLOCKS + ns);
    if (h->kry.bytes < need) {
        if (h->kry.p) (void)hipFree(h->kry.p);
        h->kry = DevBuf{};
        hipError_t e = hipMalloc(&h->kry.p, need);
        if (e != hipSuccess) return fail(h, SDS_DEVICE_ERROR, std::string("hipMalloc: ") + hipGetErrorString(e));
        h->kry.bytes = need;
    }
    const size_t hcap = (size_t)std::max(ns, hc) + 8;
    if (h->h_kry_cap < hcap) {
        if (h->h_kry) (void)hipHostFree(h->h_kry);
        h->h_kry = nullptr;
        h->h_kry_cap = 0;
        HIPCHK(h, hipHostMalloc((void**)&h->h_kry, sizeof(double) * hcap, hipHostMallocDefault));
        h->h_kry_cap = hcap;
    }
    double* V = static_cast<double*>(h->kry.p);
    double* Zk = V + (int64_t)(m + 1) * N;
    double* H = Zk + (int64_t)m * N;  // column j at H + j hc
    double* part = H + (int64_t)hc * (m + 1);
    double* cn = part + DOT_BLOCKS;   // ns scalars
    const int64_t* rp = h->full_rowptr + (transposed && h->method == M_LU ? h->n : 0);
    T* Y = static_cast<T*>(W.Y);
    const double es = sizeof(T), pbytes = (double)n * nrhs * (8.0 + es + 4.0);
    auto spmv = [&](const double* x, int64_t ldxx, const double* b, int64_t ldb, double* out) {
        timed(h, SDS_K_RESIDUAL, 2.0 * h->nnz_full * nrhs, (double)h->nnz_full * (8 + 4 + 8) + 3.0 * 8 * n * nrhs,
              [&] { launch_residual(st, n, nrhs, rp, h->full_colval, h->full_src, nz, x, ldxx, b, ldb, out); });
    };
    auto read = [&](const double* d, int cnt, double* hst) -> int {
        HIPCHK(h, hipMemcpyAsync(h->h_kry, d, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        std::copy(h->h_kry, h->h_kry + cnt, hst);
        return SDS_SUCCESS;
    };
    int rc;
    spmv(X, ldx, Bsrc, ldbs, V);  // R0
    double atol = 0.0;
    if (tol > 0) {
        for (int r = 0; r < nrhs; ++r) {
            launch_dot(st, n, V + (int64_t)r * n, V + (int64_t)r * n, part, cn + r);
            launch_dot(st, n, Bsrc + r * ldbs, Bsrc + r * ldbs, part, cn + nrhs + r);
        }
        std::vector<double> nr(2 * nrhs);
        if ((rc = read(cn, 2 * nrhs, nr.data()))) return rc;
        // the largest relative residual (||R_k|| for B_k = 0) and the smallest ||B_k|| (1 for
        // B_k = 0): _max_relative_residual and _min_rhs_norm, src/solve/refinement.jl:284, :544
        double rel = 0.0, bmin = INFINITY;
        for (int r = 0; r < nrhs; ++r) {
            const double rk = std::sqrt(nr[r]), bk = std::sqrt(nr[nrhs + r]);
            rel = std::max(rel, bk > 0 ? rk / bk : rk);
            bmin = std::min(bmin, bk > 0 ? bk : 1.0);
        }
        if (rel <= tol) return SDS_SUCCESS;
        atol = tol * bmin;
    }
    launch_dot(st, N, V, V, part, cn);
    double beta;
    if ((rc = read(cn, 1, &beta))) return rc;
    beta = std::sqrt(beta);
    if (!(beta > atol) || beta == 0.0) return SDS_SUCCESS;
    launch_scale(st, N, V, 1.0 / beta);
    // the rotated Hessenberg matrix (column-major, m + 1 rows), the rotations and g = Q^T beta e1
    std::vector<double> Rm((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1, 0.0), col(hc);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < m; ++j) {
        double* Vj = V + (int64_t)j * N;
        double* Vn = V + (int64_t)(j + 1) * N;
        double* Zj = Zk + (int64_t)j * N;
        // Z_j = op(A)^-1 V_j by the factorization, then V_{j+1} = op(A) Z_j, orthogonalized
        timed(h, SDS_K_PERMUTE, 0, pbytes, [&] { launch_permute_in<T>(st, n, nrhs, h->dsym.perm, Vj, n, Y, IN_CQ, IN_SC); });
        sweeps<T>(h, W, nrhs, 1, 1, 1, transposed);
        timed(h, SDS_K_PERMUTE, 0, pbytes, [&] { launch_permute_out<T>(st, n, nrhs, h->dsym.perm, Y, Zj, n, 0, OUT_CQ, OUT_SC); });
        spmv(Zj, n, nullptr, 0, Vn);
        for (int i = 0; i <= j; ++i) {
            launch_dot(st, N, Vn, V + (int64_t)i * N, part, H + (int64_t)j * hc + i);
            launch_axpy_neg(st, N, Vn, V + (int64_t)i * N, H + (int64_t)j * hc + i);
        }
        launch_dot(st, N, Vn, Vn, part, H + (int64_t)j * hc + j + 1);
        if ((rc = read(H + (int64_t)j * hc, j + 2, col.data()))) return rc;
        const double hn = std::sqrt(col[j + 1]);
        if (hn > 0) launch_scale(st, N, Vn, 1.0 / hn);
        // apply the earlier rotations to the new column, then the one that zeroes hn
        double* rj = Rm.data() + (size_t)j * (m + 1);
        for (int i = 0; i <= j; ++i) rj[i] = col[i];
        for (int i = 0; i < j; ++i) {
            const double a = cs[i] * rj[i] + sn[i] * rj[i + 1];
            rj[i + 1] = -sn[i] * rj[i] + cs[i] * rj[i + 1];
            rj[i] = a;
        }
        const double rr = std::hypot(rj[j], hn);
        cs[j] = rr > 0 ? rj[j] / rr : 1.0;
        sn[j] = rr > 0 ? hn / rr : 0.0;
        rj[j] = rr;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        k = j + 1;
        if (std::fabs(g[j + 1]) <= atol || hn == 0.0) break;
    }
    // y = R^-1 g, then X += sum_j y_j Z_j
    std::vector<double> y(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
        double v = g[i];
        for (int l = i + 1; l < k; ++l) v -= Rm[(size_t)l * (m + 1) + i] * y[l];
        y[i] = Rm[(size_t)i * (m + 1) + i] != 0 ? v / Rm[(size_t)i * (m + 1) + i] : 0.0;
    }
    std::copy(y.begin(), y.end(), h->h_kry);
    HIPCHK(h, hipMemcpyAsync(cn, h->h_kry, sizeof(double) * k, hipMemcpyHostToDevice, st));
    launch_add_combination(st, n, nrhs, X, ldx, Zk, N, k, cn);
    HIPCHK(h, hipStreamSynchronize(st));  // h_kry is read by the copy above
    *iters = k;
    return SDS_SUCCESS;
}

// Largest relative residual max_k ||R_k|| / ||B_k|| (||R_k|| for B_k = 0) from the 2 nrhs squared
// column norms (R's, then B's): _max_relative_residual, src/solve/refinement.jl:284-291.
double max_relative_residual(const double* nr, int nrhs) {
    double rel = 0.0;
    for (int r = 0; r < nrhs; ++r) {
        const double rk = std::sqrt(nr[r]), bk = std::sqrt(nr[nrhs + r]);
        rel = std::max(rel, bk > 0 ? rk / bk : rk);
    }
    return rel;
}

template <typename T>
int solve_t(sds_handle* h, const SolveWork& W, int32_t phase, int transposed, const double* B, int64_t ldb, double* X,
            int64_t ldx, int nrhs, const double* nz, int32_t ir_n_steps, double ir_tol, int32_t* ir_steps) {
    // the matching's sides (src/solver.jl:583-604): A x = b solves M y = Dr b with M = (Dr A Dc)[:, q]
    // and x = Dc q y; A^T x = b the transposed way round (no matching: none of them)
    const int32_t* IN_CQ = transposed ? h->mt_q : nullptr;
    const double* IN_SC = transposed ? h->mt_cs : h->mt_rs;
    const int32_t* OUT_CQ = transposed ? nullptr : h->mt_q;
    const double* OUT_SC = transposed ? h->mt_rs : h->mt_cs;
    hipStream_t st = h->stream;
    T* Y = static_cast<T*>(W.Y);
    const int n = (int)h->n;
    const bool refine = (phase & SDS_PHASE_SOLVE_REFINEMENT) && ir_n_steps > 0;
    const double* Bsrc = B;
    if (refine && X == B) {
        HIPCHK(h, hipMemcpy2DAsync(W.Bc, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                                   hipMemcpyDeviceToDevice, st));
        Bsrc = W.Bc;
    }
    const int64_t ldbs = Bsrc == B ? ldb : n;
    const double es = sizeof(T), pbytes = (double)n * nrhs * (8.0 + es + 4.0);
    // Schur complement mode (src/solver.jl:400-416): every phase reads B and writes X, in factor
    // order unless it permutes (the last nschur entries are the Schur block)
    const bool schur = h->schur_s >= 0;
    if (phase & SDS_PHASE_SOLVE_FWD_PERM)
        timed(h, SDS_K_PERMUTE, 0, pbytes, [&] { launch_permute_in<T>(st, n, nrhs, h->dsym.perm, B, ldb, Y, IN_CQ, IN_SC); });
    else if (schur)
        timed(h, SDS_K_PERMUTE, 0, pbytes, [&] { launch_permute_in<T>(st, n, nrhs, nullptr, B, ldb, Y); });
    // Schur mode, LU: the diagonal goes with the backward phase and the diagonal phase is the
    // identity, as for Cholesky (src/solver.jl:409-411)
    const int diag = schur && h->method == M_LU ? (phase & SDS_PHASE_SOLVE_BWD) : (phase & SDS_PHASE_SOLVE_DIAG);
    sweeps<T>(h, W, nrhs, phase & SDS_PHASE_SOLVE_FWD, diag, phase & SDS_PHASE_SOLVE_BWD, transposed);
    if (phase & SDS_PHASE_SOLVE_BWD_PERM)
        timed(h, SDS_K_PERMUTE, 0, pbytes, [&] { launch_permute_out<T>(st, n, nrhs, h->dsym.perm, Y, X, ldx, 0, OUT_CQ, OUT_SC); });
    else if (schur)
        timed(h, SDS_K_PERMUTE, 0, pbytes, [&] { launch_permute_out<T>(st, n, nrhs, nullptr, Y, X, ldx, 0); });
    int done = 0;
    if (refine) {
        if (!h->full_rowptr) return fail(h, SDS_INVALID_VALUE, "refinement needs the full-pattern map at upload");
        if (!nz) return fail(h, SDS_INVALID_VALUE, "refinement needs the matrix values");
        if (h->ir_mode == SDS_IR_FGMRES) {
            int rc;
            if ((rc = fgmres_t<T>(h, W, transposed, Bsrc, ldbs, X, ldx, nrhs, nz, ir_n_steps, ir_tol, &done)))
                return rc;
            if (ir_steps) *ir_steps = done;
            HIPCHK(h, hipGetLastError());
            return SDS_SUCCESS;
        }
        double* R = W.R;
        double* dotp = W.norms + 2 * MAX_NRHS;  // DOT_BLOCKS partial sums
        // with ir_tol > 0: B's squared column norms once, on the device (they do not change)
        if (ir_tol > 0)
            for (int r = 0; r < nrhs; ++r)
                launch_dot(st, n, Bsrc + r * ldbs, Bsrc + r * ldbs, dotp, W.norms + nrhs + r);
        for (int it = 0; it < ir_n_steps; ++it) {
            // one SpMV: values + column + source index per entry, X gathered, B read, R written
            timed(h, SDS_K_RESIDUAL, 2.0 * h->nnz_full * nrhs,
                  (double)h->nnz_full * (8 + 4 + 8) + 3.0 * 8 * n * nrhs, [&] {
                      // the rows n..2n-1 of a "G" map are those of M^T
                      launch_residual(st, n, nrhs, h->full_rowptr + (transposed && h->method == M_LU ? n : 0),
                                      h->full_colval, h->full_src, nz, X, ldx, Bsrc, ldbs, R);
                  });
            if (ir_tol > 0) {
                // the residual's squared column norms on the device; 2 nrhs scalars come back
                // (one synchronization per step, as residual_norms!, src/solve/refinement.jl:378-383)
                for (int r = 0; r < nrhs; ++r) launch_dot(st, n, R + (int64_t)r * n, R + (int64_t)r * n, dotp, W.norms + r);
                HIPCHK(h, hipMemcpyAsync(W.hnorms, W.norms, sizeof(double) * 2 * nrhs, hipMemcpyDeviceToHost, st));
                HIPCHK(h, hipStreamSynchronize(st));
                if (max_relative_residual(W.hnorms, nrhs) <= ir_tol) break;
            }
            timed(h, SDS_K_PERMUTE, 0, pbytes, [&] { launch_permute_in<T>(st, n, nrhs, h->dsym.perm, R, n, Y, IN_CQ, IN_SC); });
            sweeps<T>(h, W, nrhs, 1, 1, 1, transposed);
            timed(h, SDS_K_PERMUTE, 0, pbytes + 8.0 * n * nrhs,
                  [&] { launch_permute_out<T>(st, n, nrhs, h->dsym.perm, Y, X, ldx, 1, OUT_CQ, OUT_SC); });
            ++done;
        }
    }
    if (ir_steps) *ir_steps = done;
    HIPCHK(h, hipGetLastError());
    return SDS_SUCCESS;
}

}  // namespace

extern "C" {

int sds_create(sds_handle_t* out, void* stream) {
    if (!out) return SDS_INVALID_VALUE;
    auto h = new sds_handle();
    h->stream = static_cast<hipStream_t>(stream);
    (void)hipGetDevice(&h->device);
    *out = h;
    return SDS_SUCCESS;
}

int sds_destroy(sds_handle_t h) {
    if (!h) return SDS_SUCCESS;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    else (void)hipDeviceSynchronize();
    graph_free(h);
    if (h->cstream) (void)hipStreamDestroy(h->cstream);

    free_numeric(h);
    free_all(h->symbufs);
    free_all(h->mtbufs);
    for (hipEvent_t e : h->evpool) (void)hipEventDestroy(e);
    delete h;
    return SDS_SUCCESS;
}

const char* sds_last_error(sds_handle_t h) { return h ? h->err.c_str() : "null handle"; }

int sds_set_stream(sds_handle_t h, void* stream) {
    if (!h) return SDS_INVALID_VALUE;
    h->stream = static_cast<hipStream_t>(stream);
    return SDS_SUCCESS;
}

int sds_set_ir_mode(sds_handle_t h, int32_t mode) {
    if (!h) return SDS_INVALID_VALUE;
    if (mode != SDS_IR_PLAIN && mode != SDS_IR_FGMRES) return fail(h, SDS_INVALID_VALUE, "ir_mode must be 0 or 1");
    h->ir_mode = mode;
    return SDS_SUCCESS;
}

int sds_upload_symbolic(sds_handle_t h, const sds_symbolic_desc* d) {
    if (!h || !d) return SDS_INVALID_VALUE;
    if (d->index_bits != 32 && d->index_bits != 64) return fail(h, SDS_INVALID_VALUE, "index_bits must be 32 or 64");
    if (d->structure < SDS_STRUCTURE_G || d->structure > SDS_STRUCTURE_HPD)
        return fail(h, SDS_INVALID_VALUE, "unknown structure");
    free_numeric(h);
    free_all(h->symbufs);
    free_all(h->mtbufs);
    h->mt_w = h->mt_vals = h->mt_rs = h->mt_cs = nullptr;
    h->mt_q = nullptr;
    h->have_sym = false;
    h->placed = false;
    h->schur_s = -1;
    h->nschur = 0;
    const int b = d->index_bits;
    const int64_t n = d->n, ns = d->ns, nnz = d->nnz;
    h->n = n;
    h->ns = ns;
    h->nnz = nnz;
    h->structure = d->structure;
    h->method = (d->structure == SDS_STRUCTURE_SPD || d->structure == SDS_STRUCTURE_HPD) ? M_CHOLESKY
                : d->structure == SDS_STRUCTURE_G                                      ? M_LU
                                                                                       : M_LDLT;
    IdxReader perm{d->perm, b}, sptr{d->super_ptr, b}, spar{d->snparent, b}, rptr{d->rowptr, b}, rval{d->rowval, b},
        fptr{d->front_ptr, b}, fnr{d->front_nrows, b}, fnc{d->front_ncols, b}, cptr{d->child_ptr, b},
        clist{d->child_list, b}, riptr{d->relind_ptr, b}, rind{d->relind, b}, amap{d->amap, b},
        aptr{d->amap_ptr, b}, asrc{d->amap_src, b};
    h->perm.resize(n);
    for (int64_t k = 0; k < n; ++k) h->perm[k] = (int32_t)(perm[k] - 1);
    h->super_ptr.resize(ns + 1);
    for (int64_t s = 0; s <= ns; ++s) h->super_ptr[s] = (int32_t)(sptr[s] - 1);
    h->fnrows.resize(ns);
    h->fncols.resize(ns);
    h->snparent.resize(ns);
    h->panel_off.resize(ns);
    for (int64_t s = 0; s < ns; ++s) {
        h->fnrows[s] = (int32_t)fnr[s];
        h->fncols[s] = (int32_t)fnc[s];
        h->snparent[s] = (int32_t)(spar[s] - 1);
        h->panel_off[s] = fptr[s] - 1;
        if (h->snparent[s] >= 0 && h->snparent[s] <= s)
            return fail(h, SDS_INVALID_VALUE, "snparent must point to a later supernode");
    }
    h->factor_len = fptr[ns] - 1;
    std::vector<int64_t> rowptr(ns + 1);
    for (int64_t s = 0; s <= ns; ++s) rowptr[s] = rptr[s] - 1;
    std::vector<int32_t> rowval(rowptr[ns]);
    for (int64_t q = 0; q < rowptr[ns]; ++q) rowval[q] = (int32_t)(rval[q] - 1);
    h->child_ptr.resize(ns + 1);
    for (int64_t s = 0; s <= ns; ++s) h->child_ptr[s] = (int32_t)(cptr[s] - 1);
    h->child_list.resize(h->child_ptr[ns]);
    for (int64_t q = 0; q < h->child_ptr[ns]; ++q) h->child_list[q] = (int32_t)(clist[q] - 1);
    std::vector<int64_t>& relind_ptr = h->relind_ptr;
    relind_ptr.assign(ns + 1, 0);
    for (int64_t s = 0; s <= ns; ++s) relind_ptr[s] = riptr[s] - 1;
    std::vector<int32_t>& relind = h->relind;
    relind.assign(relind_ptr[ns], 0);
    for (int64_t q = 0; q < relind_ptr[ns]; ++q) relind[q] = (int32_t)(rind[q] - 1);
    // relind is increasing: a child's first cb_split columns land in the parent's panel
    // columns, the rest in its contribution block
    h->cb_split.assign(ns, 0);
    for (int64_t c = 0; c < ns; ++c) {
        const int32_t p = h->snparent[c];
        if (p < 0) continue;
        int64_t q = relind_ptr[c];
        while (q < relind_ptr[c + 1] && relind[q] < h->fncols[p]) ++q;
        h->cb_split[c] = (int32_t)(q - relind_ptr[c]);
    }
    // assembly entries grouped by front (owner-pull), with local panel offsets
    std::vector<int64_t> asm_ptr(ns + 1), asm_src, asm_dst, acol_ptr(n + 1, 0);
    for (int64_t s = 0; s <= ns; ++s) asm_ptr[s] = aptr[s] - 1;
    h->asm_count.resize(ns);
    for (int64_t s = 0; s < ns; ++s) h->asm_count[s] = asm_ptr[s + 1] - asm_ptr[s];
    asm_src.resize(asm_ptr[ns]);
    asm_dst.resize(asm_ptr[ns]);
    for (int64_t s = 0; s < ns; ++s) {
        const int64_t f = h->fnrows[s];
        const int32_t c0 = h->super_ptr[s], w = h->fncols[s];
        int32_t col = 0;
        for (int64_t k = asm_ptr[s]; k < asm_ptr[s + 1]; ++k) {
            int64_t p = asrc[k] - 1;
            if (p < 0 || p >= nnz) return fail(h, SDS_INVALID_VALUE, "amap_src out of range");
            int64_t off = amap[p];
            int64_t loc = (off < 0 ? -off : off) - 1 - h->panel_off[s];
            if (loc < 0 || loc >= f * w) return fail(h, SDS_INVALID_VALUE, "amap offset outside its front");
            asm_src[k] = p;
            asm_dst[k] = off > 0 ? loc : -(loc + 1);
            int32_t j = (int32_t)(loc / f);
            while (col <= j) acol_ptr[c0 + col++] = k;
        }
        while (col <= w) {
            acol_ptr[c0 + col] = asm_ptr[s + 1];
            ++col;
        }
    }
    acol_ptr[n] = asm_ptr[ns];
    int rc;
    DevSym& D = h->dsym;
    D = DevSym{};
    D.n = (int32_t)n;
    D.ns = (int32_t)ns;
    D.unsym = d->structure == SDS_STRUCTURE_G;
    int32_t *p_perm, *p_sptr, *p_fnr, *p_fnc, *p_rval, *p_cptr, *p_clist, *p_rind;
    int64_t *p_poff, *p_rptr, *p_riptr, *p_aptr, *p_asrc, *p_adst, *p_acol;
    if ((rc = upload(h, h->symbufs, h->perm, &p_perm))) return rc;
    if ((rc = upload(h, h->symbufs, h->super_ptr, &p_sptr))) return rc;
    if ((rc = upload(h, h->symbufs, h->fnrows, &p_fnr))) return rc;
    if ((rc = upload(h, h->symbufs, h->fncols, &p_fnc))) return rc;
    if ((rc = upload(h, h->symbufs, h->panel_off, &p_poff))) return rc;
    if ((rc = upload(h, h->symbufs, rowptr, &p_rptr))) return rc;
    if ((rc = upload(h, h->symbufs, rowval, &p_rval))) return rc;
    if ((rc = upload(h, h->symbufs, h->child_ptr, &p_cptr))) return rc;
    if ((rc = upload(h, h->symbufs, h->child_list, &p_clist))) return rc;
    if ((rc = upload(h, h->symbufs, relind_ptr, &p_riptr))) return rc;
    if ((rc = upload(h, h->symbufs, relind, &p_rind))) return rc;
    if ((rc = upload(h, h->symbufs, asm_ptr, &p_aptr))) return rc;
    if ((rc = upload(h, h->symbufs, asm_src, &p_asrc))) return rc;
    if ((rc = upload(h, h->symbufs, asm_dst, &p_adst))) return rc;
    if ((rc = upload(h, h->symbufs, acol_ptr, &p_acol))) return rc;
    D.perm = p_perm;
    D.super_ptr = p_sptr;
    D.fnrows = p_fnr;
    D.fncols = p_fnc;
    D.panel_off = p_poff;
    D.rowptr = p_rptr;
    D.rowval = p_rval;
    D.child_ptr = p_cptr;
    D.child_list = p_clist;
    D.relind_ptr = p_riptr;
    D.relind = p_rind;
    D.asm_ptr = p_aptr;
    D.asm_src = p_asrc;
    D.asm_dst = p_adst;
    D.acol_ptr = p_acol;
    h->full_rowptr = nullptr;
    h->full_colval = nullptr;
    h->full_src = nullptr;
    h->full_nrows = 0;
    h->adj_i = h->adj_j = nullptr;
    h->adj_flag = nullptr;
    if (d->full_rowptr && d->full_colval && d->full_src) {
        IdxReader frp{d->full_rowptr, b}, fcv{d->full_colval, b}, fsrc{d->full_src, b};
        int64_t nr = d->full_nrows, fnnz = d->full_nnz;
        std::vector<int64_t> rp(nr + 1), src(fnnz);
        std::vector<int32_t> cv(fnnz);
        for (int64_t i = 0; i <= nr; ++i) rp[i] = frp[i] - 1;
        for (int64_t e = 0; e < fnnz; ++e) {
            cv[e] = (int32_t)(fcv[e] - 1);
            int64_t sv = fsrc[e];
            src[e] = sv > 0 ? sv - 1 : sv;
        }
        if ((rc = upload(h, h->symbufs, rp, &h->full_rowptr))) return rc;
        if ((rc = upload(h, h->symbufs, cv, &h->full_colval))) return rc;
        if ((rc = upload(h, h->symbufs, src, &h->full_src))) return rc;
        h->full_nrows = nr;
        h->nnz_full = fnnz;
        // the adjoint map from the rows of M (the first n rows of the map)
        std::vector<int32_t> ai(nnz, 0), aj(nnz, 0);
        std::vector<int8_t> af(nnz, 0);
        for (int64_t r = 0; r < std::min<int64_t>(nr, n); ++r)
            for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
                const int64_t sv = src[e], p = sv >= 0 ? sv : -(sv + 1);
                if (p < 0 || p >= nnz) return fail(h, SDS_INVALID_VALUE, "full_src out of range");
                if (af[p] == 0) {
                    ai[p] = (int32_t)r;
                    aj[p] = cv[e];
                    af[p] = 1;
                } else if (ai[p] == cv[e] && aj[p] == (int32_t)r && ai[p] != aj[p]) {
                    af[p] = 2;
                }
            }
        if ((rc = upload(h, h->symbufs, ai, &h->adj_i))) return rc;
        if ((rc = upload(h, h->symbufs, aj, &h->adj_j))) return rc;
        if ((rc = upload(h, h->symbufs, af, &h->adj_flag))) return rc;
    }
    h->have_sym = true;
    return SDS_SUCCESS;
}

int sds_alloc_numeric(sds_handle_t h, int32_t dtype, int32_t nbatch) {
    if (!h) return SDS_INVALID_VALUE;
    if (!h->have_sym) return fail(h, SDS_FACTORIZATION_ERROR, "no analysis uploaded");
    if (dtype != SDS_R64 && dtype != SDS_R32) return fail(h, SDS_INVALID_VALUE, "dtype must be SDS_R64 or SDS_R32");
    if (nbatch != 1)
        return fail(h, SDS_NOT_SUPPORTED,
                    "nbatch must be 1: upload a uniform batch as the block-diagonal forest of its members");
    free_numeric(h);
    int rc = build_schedule(h, dtype);
    if (rc) return rc;

    size_t es = dtype == SDS_R64 ? 8 : 4;
    void* p;
    if ((rc = dev_alloc(h, h->numbufs, es * std::max<int64_t>(h->factor_len, 1), &h->factor))) return rc;
    if ((rc = dev_alloc(h, h->numbufs, es * std::max<int64_t>(h->cb_len, 1), &h->cbpool))) return rc;
    h->ufactor = nullptr;
    if (h->method == M_LU &&
        (rc = dev_alloc(h, h->numbufs, es * std::max<int64_t>(h->factor_len, 1), &h->ufactor)))
        return rc;
    if ((rc = dev_alloc(h, h->numbufs, es * NB * std::max<int64_t>(h->n, 1), &h->dbuf))) return rc;
    if ((rc = dev_alloc(h, h->numbufs, es * NB * std::max<int64_t>(h->n, 1), &h->lnbuf))) return rc;
    if ((rc = dev_alloc(h, h->numbufs, sizeof(int32_t) * 3 * (h->ns + 1), &p))) return rc;
    h->lstep = static_cast<int32_t*>(p);
    h->lscratch = nullptr;
    if (h->ldlt_scr > 0 && (rc = dev_alloc(h, h->numbufs, es * h->ldlt_scr, &h->lscratch))) return rc;
    if ((rc = dev_alloc(h, h->numbufs, sizeof(int32_t) * (h->ns + 1), &p))) return rc;
    h->status = static_cast<int32_t*>(p);
    if ((rc = dev_alloc(h, h->numbufs, es * 2 * std::max<int64_t>(h->n, 1), &h->dvec))) return rc;
    if ((rc = dev_alloc(h, h->numbufs, sizeof(int32_t) * std::max<int64_t>(h->n, 1), &p))) return rc;
    h->piv = static_cast<int32_t*>(p);
    if ((rc = dev_alloc(h, h->numbufs, std::max<int64_t>(h->n, 1), &p))) return rc;
    h->pkind = static_cast<int8_t*>(p);
    if ((rc = dev_alloc(h, h->numbufs, std::max<int64_t>(h->n, 1), &p))) return rc;
    h->psign = static_cast<int8_t*>(p);
    if ((rc = dev_alloc(h, h->numbufs, sizeof(double) * 2, &p))) return rc;
    h->aux = static_cast<double*>(p);
    if ((rc = dev_alloc(h, h->numbufs, sizeof(int64_t) * NF * (h->ns + 1), &p))) return rc;
    h->stats = static_cast<int64_t*>(p);
    if ((rc = dev_alloc(h, h->numbufs, sizeof(int64_t) * (8 + INFO_BLOCKS), &p))) return rc;
    h->d_info = static_cast<int64_t*>(p);
    HIPCHK(h, hipHostMalloc((void**)&h->h_info, sizeof(int64_t) * 8, hipHostMallocDefault));
    HIPCHK(h, hipHostMalloc((void**)&h->h_norms, sizeof(double) * 2 * MAX_NRHS, hipHostMallocDefault));
    h->dtype = dtype;
    return SDS_SUCCESS;
}

}  // extern "C"

namespace {

// the launches of one factorization on the handle's stream, without the host read of info.
// sign_ready: pivot_sign is already in h->psign (sds_graph_capture copies it before the capture
// begins: a blocking copy cannot be recorded); else it is copied asynchronously on the stream.
int factorize_enqueue(sds_handle* h, const double* nzval, const sds_factor_params* prm, bool sign_ready) {
    if (h->dtype < 0) return fail(h, SDS_FACTORIZATION_ERROR, "numeric storage not allocated");
    if (!nzval && h->nnz > 0) return fail(h, SDS_INVALID_VALUE, "nzval is NULL");
    double eps = h->dtype == SDS_R64 ? 1e-13 : 1e-5;  // default_pivot_epsilon (src/options.jl:57-61)
    const volatile uint8_t* intr = nullptr;
    int ptype = 0, scaled = 0;
    double u = 0.01;
    const int8_t* psign = nullptr;
    if (prm) {
        if (prm->pivot_epsilon >= 0) eps = prm->pivot_epsilon;
        intr = prm->interrupt;
        ptype = prm->pivot_type;
        u = prm->pivot_threshold;
        scaled = prm->pivot_epsilon_alg == 1;
        if (prm->pivot_epsilon_alg < 0 || prm->pivot_epsilon_alg > 2)
            return fail(h, SDS_INVALID_VALUE, "pivot_epsilon_alg must be 0, 1 or 2");
        if (!(u >= 0.0 && u <= 1.0)) return fail(h, SDS_INVALID_VALUE, "pivot_threshold must be in [0, 1]");
        if (prm->pivot_sign) {
            if (!sign_ready)
                HIPCHK(h, hipMemcpyAsync(h->psign, prm->pivot_sign, h->n, hipMemcpyHostToDevice, h->stream));
            psign = h->psign;
        }
    }
    h->factorized = false;
    int rc;
    if (h->mt_w) {
        // the matching's scaled values r_i a_ij c_j are what gets factored (src/solver.jl:595-601)
        launch_scale_values(h->stream, h->nnz, h->mt_w, nzval, h->mt_vals);
        nzval = h->mt_vals;
    }
    if (h->method == M_LDLT) {
        // pivot_type (src/types.jl:84-92): 'N' none, 'D' diagonal, 'A'/'L'/'B' Bunch-Kaufman
        // inside the fully summed block (src/reference/ldlt.jl:44-51)
        int pivot;
        switch (ptype) {
            case 1: pivot = 0; break;
            case 4: pivot = 1; break;
            case 0: case 5: case 6: pivot = 2; break;
            default: return fail(h, SDS_NOT_SUPPORTED, "pivot_type is not supported by LDL^T");
        }
        rc = h->dtype == SDS_R64 ? factorize_ldlt<double>(h, nzval, pivot, u, eps, scaled, psign, intr)
                                 : factorize_ldlt<float>(h, nzval, pivot, u, eps, scaled, psign, intr);
    } else if (h->method == M_LU) {
        // 'N' and 'D' keep the diagonal pivots, 'A'/'L'/'B' pivot by rows (src/reference/lu.jl:26-38)
        int pivoting;
        switch (ptype) {
            case 1: case 4: pivoting = 0; break;
            case 0: case 5: case 6: pivoting = 1; break;
            default: return fail(h, SDS_NOT_SUPPORTED, "pivot_type is not supported by LU");
        }
        rc = h->dtype == SDS_R64 ? factorize_lu<double>(h, nzval, pivoting, u, eps, scaled, intr)
                                 : factorize_lu<float>(h, nzval, pivoting, u, eps, scaled, intr);
    } else {
        rc = h->dtype == SDS_R64 ? factorize_cholesky<double>(h, nzval, eps, intr)
                                 : factorize_cholesky<float>(h, nzval, eps, intr);
    }
    return rc;
}

// info of the last factorization: the one device-to-host read (src/numeric/factorize.jl:383-417)
int read_info(sds_handle* h, int64_t* info) {
    HIPCHK(h, hipMemcpyAsync(h->h_info, h->d_info, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int64_t col = h->h_info[0];
    if (info) *info = col > 0 ? (int64_t)h->perm[col - 1] + 1 : 0;
    return SDS_SUCCESS;
}

int check_solve_args(sds_handle* h, int32_t phase, int32_t solve_mode, int64_t ldb, int64_t ldx, int32_t nrhs) {
    if (nrhs < 1 || nrhs > MAX_NRHS) return fail(h, SDS_NOT_SUPPORTED, "nrhs must be in 1..64");
    if (ldb < h->n || ldx < h->n) return fail(h, SDS_INVALID_VALUE, "leading dimension smaller than n");
    if (solve_mode != 0 && solve_mode != 1 && solve_mode != 2)
        return fail(h, SDS_INVALID_VALUE, "solve_mode must be 0, 1 or 2");
    if ((phase & ~SDS_PHASE_SOLVE) != 0 || phase == 0) return fail(h, SDS_INVALID_VALUE, "not a solve phase");
    if (h->schur_s >= 0 && ((phase & SDS_PHASE_SOLVE_REFINEMENT) ||
                            ((phase & SDS_PHASE_SOLVE_FWD) && (phase & SDS_PHASE_SOLVE_BWD))))
        return fail(h, SDS_NOT_SUPPORTED,
                    "Schur complement mode: the factorization is partial; solve in the phases fwd_perm | fwd, diag, "
                    "bwd | bwd_perm");
    return SDS_SUCCESS;
}

}  // namespace

extern "C" {

int sds_factorize(sds_handle_t h, const double* nzval, const sds_factor_params* prm, int64_t* info) {
    if (!h) return SDS_INVALID_VALUE;
    h->factorized = false;
    int rc = factorize_enqueue(h, nzval, prm, false);
    if (rc) return rc;
    if ((rc = read_info(h, info))) return rc;
    h->factorized = true;
    return SDS_SUCCESS;
}

int sds_autotune(sds_handle_t h, const double* nzval, const sds_factor_params* prm, int32_t reps,
                 int32_t* tiles) {
    if (!h) return SDS_INVALID_VALUE;
    if (h->dtype < 0) return fail(h, SDS_FACTORIZATION_ERROR, "numeric storage not allocated");
    if (reps < 1) return fail(h, SDS_INVALID_VALUE, "reps must be positive");
    const int dtype = h->dtype;
    // candidate tile shapes; the K chunk only matters to the Cholesky contribution-block update
    struct Cand {
        int ea, kc, rs, oc, sk, sb, pb;
    };
    std::vector<Cand> cand;
    for (int ea : {512, 1024, 2048, 4096})
        for (int kc : {16, 32, 64})
            if (kc == 32 || h->method == M_CHOLESKY) cand.push_back({ea, kc, 0, 0, 0, 0, h->panel_block});
    if (h->method == M_CHOLESKY) {  // with the best of the above
        for (int pb : {0, 128, 256, 512})
            if (pb != h->panel_block) cand.push_back({-1, -1, 0, 0, 0, 0, pb});
        for (int rs : {256, 512, 1024})
            if (rs <= h->max_rest) cand.push_back({-1, -1, rs, 0, 0, 0, -1});
        for (int oc : {1, 64, 128, 256, 512}) cand.push_back({-1, -1, -1, oc, 0, 0, -1});
        for (int sk : {1, 2}) cand.push_back({-1, -1, -1, -1, sk, 0, -1});
        for (int sb : {16, 32, 64}) cand.push_back({-1, -1, -1, -1, -1, sb, -1});
    }
    double best = 1e300;
    Cand pick{h->ea_rows, h->upd_kc, h->rest_split, h->step_occ2, h->small_kind, h->sub_kib, h->panel_block};
    int64_t info = 0;
    int rc;
    for (Cand c : cand) {
        if (c.ea < 0) {
            c.ea = pick.ea;
            c.kc = pick.kc;
        }
        if (c.rs < 0) c.rs = pick.rs;
        if (c.oc < 0) c.oc = pick.oc;
        if (c.sk < 0) c.sk = pick.sk;
        if (c.pb < 0) c.pb = pick.pb;
        h->ea_rows = c.ea;
        h->upd_kc = c.kc;
        h->rest_split = c.rs;
        h->step_occ2 = c.oc;
        h->small_kind = c.sk;
        h->sub_kib = c.sb;
        h->panel_block = c.pb;
        if ((rc = sds_alloc_numeric(h, dtype, 1))) return rc;
        if ((rc = sds_factorize(h, nzval, prm, &info))) return rc;  // warm-up
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; ++r)
            if ((rc = sds_factorize(h, nzval, prm, &info))) return rc;
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (t < best) {
            best = t;
            pick = c;
        }
    }
    h->ea_rows = pick.ea;
    h->upd_kc = pick.kc;
    h->rest_split = pick.rs;
    h->step_occ2 = pick.oc;
    h->small_kind = pick.sk;
    h->sub_kib = pick.sb;
    h->panel_block = pick.pb;
    if ((rc = sds_alloc_numeric(h, dtype, 1))) return rc;
    if (tiles) {
        tiles[0] = pick.ea;
        tiles[1] = pick.kc;
        tiles[2] = pick.rs;
        tiles[3] = pick.oc;
        tiles[4] = pick.sk;
        tiles[5] = pick.sb;
        tiles[6] = pick.pb;
        tiles[7] = h->pivot_panels;
    }
    return SDS_SUCCESS;
}

int sds_set_tiles(sds_handle_t h, const int32_t* tiles) {
    if (!h) return SDS_INVALID_VALUE;
    if (!tiles) return fail(h, SDS_INVALID_VALUE, "tiles is NULL");
    const int ea = tiles[0], kc = tiles[1], rs = tiles[2], oc = tiles[3], sk = tiles[4], sb = tiles[5],
              pb = tiles[6], pp = tiles[7];
    if (ea != 512 && ea != 1024 && ea != 2048 && ea != 4096)
        return fail(h, SDS_INVALID_VALUE, "extend-add rows must be 512, 1024, 2048 or 4096");
    if (kc != 16 && kc != 32 && kc != 64) return fail(h, SDS_INVALID_VALUE, "the K chunk must be 16, 32 or 64");
    if (rs < 0 || oc < 0) return fail(h, SDS_INVALID_VALUE, "the trailing split and two-per-CU steps must be >= 0");
    if (sk < 0 || sk > 2) return fail(h, SDS_INVALID_VALUE, "the small-front kernel must be 0, 1 or 2");
    if (sb != 0 && sb != 16 && sb != 32 && sb != 64)
        return fail(h, SDS_INVALID_VALUE, "the subtree budget must be 0, 16, 32 or 64 (KiB)");
    if (pb < 0 || pb % NB != 0 || pb > 4096)
        return fail(h, SDS_INVALID_VALUE, "the panel block must be 0 or a multiple of 64 up to 4096");
    if (pp != 0 && pp != 1) return fail(h, SDS_INVALID_VALUE, "pivot panels must be 0 or 1");
    h->ea_rows = ea;
    h->upd_kc = kc;
    h->rest_split = rs;
    h->step_occ2 = oc;
    h->small_kind = sk;
    h->sub_kib = sb;
    h->panel_block = pb;
    h->pivot_panels = pp;
    return SDS_SUCCESS;
}

int sds_graph_capture(sds_handle_t h, const double* nzval, const sds_factor_params* prm, int32_t phase,
                      int32_t solve_mode, const double* B, int64_t ldb, double* X, int64_t ldx, int32_t nrhs,
                      int32_t ir_n_steps) {
    if (!h) return SDS_INVALID_VALUE;
    if (h->dtype < 0) return fail(h, SDS_FACTORIZATION_ERROR, "numeric storage not allocated");
    int rc = check_solve_args(h, phase, solve_mode, ldb, ldx, nrhs);
    if (rc) return rc;
    if (!B || !X) return fail(h, SDS_INVALID_VALUE, "B and X are needed");
    if (h->ir_mode == SDS_IR_FGMRES && (phase & SDS_PHASE_SOLVE_REFINEMENT) && ir_n_steps > 0)
        return fail(h, SDS_NOT_SUPPORTED, "FGMRES refinement synchronizes with the host: it cannot be captured");
    // every allocation happens before the capture
    if ((rc = ensure_solve_workspace(h, nrhs))) return rc;
    graph_free(h);
    if (!h->cstream) HIPCHK(h, hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // pivot_sign goes to the device now: the graph reads it from there at every replay
    if (prm && prm->pivot_sign) HIPCHK(h, hipMemcpy(h->psign, prm->pivot_sign, h->n, hipMemcpyHostToDevice));
    hipStream_t user = h->stream;
    const bool prof = h->prof;
    h->prof = false;
    h->stream = h->cstream;
    const SolveWork W = handle_work(h);
    hipError_t e = hipStreamBeginCapture(h->cstream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = factorize_enqueue(h, nzval, prm, true);
        const int tr = h->method == M_LU && solve_mode != 0;
        if (!rc)
            rc = h->dtype == SDS_R64 ? solve_t<double>(h, W, phase, tr, B, ldb, X, ldx, nrhs, nzval, ir_n_steps, 0.0,
                                                       nullptr)
                                     : solve_t<float>(h, W, phase, tr, B, ldb, X, ldx, nrhs, nzval, ir_n_steps, 0.0,
                                                      nullptr);
        hipError_t e2 = hipStreamEndCapture(h->cstream, &h->graph);
        if (!rc && e2 != hipSuccess) rc = fail(h, SDS_DEVICE_ERROR, std::string("hipStreamEndCapture: ") +
                                                                        hipGetErrorString(e2));
    } else {
        rc = fail(h, SDS_DEVICE_ERROR, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    h->stream = user;
    h->prof = prof;
    if (rc) {
        graph_free(h);
        return rc;
    }
    HIPCHK(h, hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
    return SDS_SUCCESS;
}

int sds_graph_launch(sds_handle_t h) {
    if (!h) return SDS_INVALID_VALUE;
    if (!h->gexec) return fail(h, SDS_INVALID_VALUE, "no captured graph");
    HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
    h->factorized = true;
    return SDS_SUCCESS;
}

int sds_graph_info(sds_handle_t h, int64_t* info) {
    if (!h) return SDS_INVALID_VALUE;
    return read_info(h, info);
}

int sds_graph_destroy(sds_handle_t h) {
    if (!h) return SDS_INVALID_VALUE;
    graph_free(h);
    return SDS_SUCCESS;
}

int sds_solve(sds_handle_t h, int32_t phase, int32_t solve_mode, const double* B, int64_t ldb, double* X, int64_t ldx,
              int32_t nrhs, const double* nzval, int32_t ir_n_steps, double ir_tol, int32_t* ir_steps) {
    if (!h) return SDS_INVALID_VALUE;
    if (!h->factorized) return fail(h, SDS_FACTORIZATION_ERROR, "solve before a factorization");
    int rc = check_solve_args(h, phase, solve_mode, ldb, ldx, nrhs);
    if (rc) return rc;
    if ((rc = ensure_solve_workspace(h, nrhs))) return rc;
    const SolveWork W = handle_work(h);
    // symmetric structures: A^T = A; real values: A^H = A^T (src/solve/refinement.jl:10-15)
    const int tr = h->method == M_LU && solve_mode != 0;
    return h->dtype == SDS_R64
               ? solve_t<double>(h, W, phase, tr, B, ldb, X, ldx, nrhs, nzval, ir_n_steps, ir_tol, ir_steps)
               : solve_t<float>(h, W, phase, tr, B, ldb, X, ldx, nrhs, nzval, ir_n_steps, ir_tol, ir_steps);
}

int sds_solve_noalloc(sds_handle_t h, int32_t phase, int32_t solve_mode, const double* B, int64_t ldb, double* X,
                      int64_t ldx, int32_t nrhs, const double* nzval, int32_t ir_n_steps, double ir_tol,
                      int32_t* ir_steps) {
    if (!h) return SDS_INVALID_VALUE;
    if (!h->factorized) return fail(h, SDS_FACTORIZATION_ERROR, "solve before a factorization");
    int rc = check_solve_args(h, phase, solve_mode, ldb, ldx, nrhs);
    if (rc) return rc;
    if (h->ir_mode == SDS_IR_FGMRES && (phase & SDS_PHASE_SOLVE_REFINEMENT) && ir_n_steps > 0)
        return fail(h, SDS_NOT_SUPPORTED, "sds_solve_noalloc: FGMRES refinement needs a Krylov basis; use sds_solve");
    const size_t need = noalloc_layout(h, nrhs, nullptr, nullptr);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(h, SDS_INVALID_VALUE,
                    "sds_solve_noalloc: n = " + std::to_string(h->n) + " with nrhs = " + std::to_string(nrhs) +
                        " needs " + std::to_string(need) + " bytes of solve workspace, more than its " +
                        std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte scratch; use sds_solve or fewer right-hand sides");
    void* base = nullptr;
    if ((rc = scratch_base(h, &base))) return rc;
    SolveWork W;
    noalloc_layout(h, nrhs, static_cast<char*>(base), &W);
    // profiling records HIP events, which it may create: off for this call
    const bool prof = h->prof;
    h->prof = false;
    const int tr = h->method == M_LU && solve_mode != 0;
    rc = h->dtype == SDS_R64
             ? solve_t<double>(h, W, phase, tr, B, ldb, X, ldx, nrhs, nzval, ir_n_steps, ir_tol, ir_steps)
             : solve_t<float>(h, W, phase, tr, B, ldb, X, ldx, nrhs, nzval, ir_n_steps, ir_tol, ir_steps);
    h->prof = prof;
    return rc;
}

int sds_set_schur(sds_handle_t h, int64_t nschur) {
    if (!h) return SDS_INVALID_VALUE;
    if (!h->have_sym) return fail(h, SDS_FACTORIZATION_ERROR, "no analysis uploaded");
    if (nschur < 0) return fail(h, SDS_INVALID_VALUE, "nschur must be >= 0");
    free_numeric(h);  // the launch plan changes: sds_alloc_numeric again
    h->placed = false;
    if (nschur == 0) {
        h->schur_s = -1;
        h->nschur = 0;
        return SDS_SUCCESS;
    }
    const int64_t s = h->ns - 1;
    if (s < 0 || h->fncols[s] != nschur || h->fnrows[s] != nschur || h->snparent[s] >= 0)
        return fail(h, SDS_INVALID_VALUE,
                    "the last supernode is not a dense root of nschur columns: analyse with the Schur indices");
    h->schur_s = (int32_t)s;
    h->nschur = nschur;
    return SDS_SUCCESS;
}

int sds_set_matching(sds_handle_t h, const double* weights, const double* rscale, const double* cscale,
                     const int64_t* colperm) {
    if (!h) return SDS_INVALID_VALUE;
    if (!h->have_sym) return fail(h, SDS_FACTORIZATION_ERROR, "no analysis uploaded");
    free_all(h->mtbufs);
    h->mt_w = h->mt_vals = h->mt_rs = h->mt_cs = nullptr;
    h->mt_q = nullptr;
    h->factorized = false;
    if (!weights) return SDS_SUCCESS;
    if (!rscale || !cscale) return fail(h, SDS_INVALID_VALUE, "a matching needs its row and column scalings");
    if (h->schur_s >= 0) return fail(h, SDS_NOT_SUPPORTED, "matching is not supported with schur_mode");
    const int64_t n = h->n, nnz = h->nnz;
    std::vector<double> w(weights, weights + nnz), rs(rscale, rscale + n), cs(cscale, cscale + n);
    int rc;
    void* p;
    if ((rc = upload(h, h->mtbufs, w, &h->mt_w))) return rc;
    if ((rc = dev_alloc(h, h->mtbufs, sizeof(double) * std::max<int64_t>(nnz, 1), &p))) return rc;
    h->mt_vals = static_cast<double*>(p);
    if ((rc = upload(h, h->mtbufs, rs, &h->mt_rs))) return rc;
    if ((rc = upload(h, h->mtbufs, cs, &h->mt_cs))) return rc;
    if (colperm) {
        std::vector<int32_t> q(n);
        std::vector<char> seen(n, 0);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t c = colperm[i] - 1;
            if (c < 0 || c >= n || seen[c]) return fail(h, SDS_INVALID_VALUE, "colperm is not a permutation");
            seen[c] = 1;
            q[i] = (int32_t)c;
        }
        if ((rc = upload(h, h->mtbufs, q, &h->mt_q))) return rc;
    }
    return SDS_SUCCESS;
}

int sds_schur_export(sds_handle_t h, double* S, int64_t lds) {
    if (!h) return SDS_INVALID_VALUE;
    if (h->schur_s < 0) return fail(h, SDS_INVALID_VALUE, "not in Schur complement mode");
    if (!h->factorized) return fail(h, SDS_FACTORIZATION_ERROR, "schur_matrix before a factorization");
    if (!S || lds < h->nschur) return fail(h, SDS_INVALID_VALUE, "null buffer or leading dimension below nschur");
    const int full = h->method == M_LU;
    if (h->dtype == SDS_R64)
        launch_schur_export<double>(h->stream, h->dsym, h->schur_s, (int)h->nschur, static_cast<double*>(h->factor),
                                    full, S, lds);
    else
        launch_schur_export<float>(h->stream, h->dsym, h->schur_s, (int)h->nschur, static_cast<float*>(h->factor),
                                   full, S, lds);
    HIPCHK(h, hipGetLastError());
    return SDS_SUCCESS;
}

int sds_adjoint_scatter(sds_handle_t h, const double* lambda, int64_t ldl, const double* X, int64_t ldx, int32_t nrhs,
                        double* grad) {
    if (!h) return SDS_INVALID_VALUE;
    if (!h->adj_flag) return fail(h, SDS_INVALID_VALUE, "the adjoint needs the full-pattern map at upload");
    if (!lambda || !X || !grad || nrhs < 1) return fail(h, SDS_INVALID_VALUE, "null buffer or nrhs < 1");
    if (ldl < h->n || ldx < h->n) return fail(h, SDS_INVALID_VALUE, "leading dimension smaller than n");
    launch_adjoint_scatter(h->stream, h->nnz, nrhs, h->adj_i, h->adj_j, h->adj_flag, lambda, ldl, X, ldx, grad);
    HIPCHK(h, hipGetLastError());
    return SDS_SUCCESS;
}

int sds_get_stats(sds_handle_t h, int64_t* stats, int64_t* totals) {
    if (!h || h->dtype < 0) return SDS_INVALID_VALUE;
    std::vector<int64_t> st((size_t)NF * h->ns);
    HIPCHK(h, hipMemcpyAsync(st.data(), h->stats, sizeof(int64_t) * st.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (stats) std::memcpy(stats, st.data(), sizeof(int64_t) * st.size());
    if (totals) {
        for (int q = 0; q < NF; ++q) totals[q] = 0;
        for (int64_t s = 0; s < h->ns; ++s) {
            for (int q = 0; q < NF - 1; ++q) totals[q] += st[s * NF + q];
            totals[NF - 1] += st[s * NF + NF - 1] != 0;
        }
    }
    return SDS_SUCCESS;
}

int sds_get_pivots(sds_handle_t h, int32_t* piv, int8_t* kind) {
    if (!h || h->dtype < 0) return SDS_INVALID_VALUE;
    if (h->method != M_CHOLESKY) {
        std::vector<int32_t> p(h->n);
        HIPCHK(h, hipStreamSynchronize(h->stream));
        HIPCHK(h, hipMemcpy(p.data(), h->piv, sizeof(int32_t) * h->n, hipMemcpyDeviceToHost));
        if (piv)
            for (int64_t k = 0; k < h->n; ++k) piv[k] = p[k] + 1;
        if (kind) HIPCHK(h, hipMemcpy(kind, h->pkind, h->n, hipMemcpyDeviceToHost));
        return SDS_SUCCESS;
    }
    // Cholesky: no pivoting, every pivot a 1x1 in place
    for (int64_t k = 0; k < h->n; ++k) {
        if (piv) piv[k] = (int32_t)(k + 1);
        if (kind) kind[k] = KIND_1X1;
    }
    return SDS_SUCCESS;
}

int sds_get_diag(sds_handle_t h, void* d) {
    if (!h || h->dtype < 0 || !d) return SDS_INVALID_VALUE;
    size_t es = h->dtype == SDS_R64 ? 8 : 4;
    if (h->method != M_CHOLESKY) {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        HIPCHK(h, hipMemcpy(d, h->dvec, es * 2 * h->n, hipMemcpyDeviceToHost));
        return SDS_SUCCESS;
    }
    // Cholesky: the diagonal of L, read from the panels
    std::memset(d, 0, es * 2 * h->n);
    for (int64_t s = 0; s < h->ns; ++s) {
        int64_t f = h->fnrows[s];
        for (int32_t j = 0; j < h->fncols[s]; ++j)
            HIPCHK(h, hipMemcpy(static_cast<char*>(d) + es * (h->super_ptr[s] + j),
                                static_cast<char*>(h->factor) + es * (h->panel_off[s] + (int64_t)j * f + j), es,
                                hipMemcpyDeviceToHost));
    }
    return SDS_SUCCESS;
}

int sds_get_factor(sds_handle_t h, int32_t which, void* out) {
    if (!h || h->dtype < 0 || !out) return SDS_INVALID_VALUE;
    if (which != 0 && (which != 1 || !h->ufactor)) return fail(h, SDS_NOT_SUPPORTED, "no U^T factor for this structure");
    size_t es = h->dtype == SDS_R64 ? 8 : 4;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(out, which ? h->ufactor : h->factor, es * h->factor_len, hipMemcpyDeviceToHost));
    return SDS_SUCCESS;
}

int sds_profile(sds_handle_t h, int32_t enable) {
    if (!h) return SDS_INVALID_VALUE;
    harvest(h);
    if (enable) {
        for (int k = 0; k < SDS_K_COUNT; ++k) {
            h->acc_ms[k] = h->acc_flops[k] = h->acc_bytes[k] = 0;
            h->acc_n[k] = 0;
        }
    }
    h->prof = enable != 0;
    return SDS_SUCCESS;
}

int sds_profile_read(sds_handle_t h, int32_t kind, double* ms, int64_t* launches, double* flops, double* bytes) {
    if (!h || kind < 0 || kind >= SDS_K_COUNT) return SDS_INVALID_VALUE;
    harvest(h);
    if (ms) *ms = h->acc_ms[kind];
    if (launches) *launches = h->acc_n[kind];
    if (flops) *flops = h->acc_flops[kind];
    if (bytes) *bytes = h->acc_bytes[kind];
    return SDS_SUCCESS;
}

int sds_memory_estimates(sds_handle_t h, int64_t* est) {
    if (!h || !est) return SDS_INVALID_VALUE;
    for (int q = 0; q < 16; ++q) est[q] = 0;
    size_t es = h->dtype == SDS_R32 ? 4 : 8;
    int64_t maps = 0;
    for (auto& b : h->symbufs) maps += (int64_t)b.bytes;
    est[6] = (int64_t)es * h->factor_len;
    est[7] = (int64_t)es * 2 * h->n;
    est[8] = (int64_t)es * h->cb_len;
    est[9] = maps;
    est[10] = (int64_t)(h->ns + 1) * NF * 8 + (h->ns + 1) * 4 + h->n * 6;
    est[0] = est[6] + est[7] + est[9] + est[10];
    est[1] = est[0] + est[8];
    est[4] = est[1];
    return SDS_SUCCESS;
}

}  // extern "C"
