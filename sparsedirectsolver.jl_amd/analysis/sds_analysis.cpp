// Host symbolic analysis of SparseDirectSolver.jl, restated in C++ so that the
// numeric phase can be driven without a Julia runtime. The analysis is the
// input producer of the hot path: it turns a CSR pattern and an ordering into
// the supernodal partition, the panel layout and the assembly/extend-add maps
// that the device numeric phase and the CPU oracle both consume.
//
// Every output array is 1-based, exactly as the Julia `Symbolic` holds it
// (src/symbolic/maps.jl:62-95), so the arrays can be handed to the C ABI of the
// numeric library unchanged.
//
// Steps and the reference functions they follow:
//   symmetric pattern           src/symbolic/pattern.jl:138-164 (view filter :113-120)
//   elimination tree            src/symbolic/etree.jl:14-36
//   postorder                   src/symbolic/etree.jl:45-78
//   column counts (GNP)         src/symbolic/etree.jl:81-147
//   fundamental supernodes      src/symbolic/supernodes.jl:90-112
//   relaxed amalgamation        src/symbolic/supernodes.jl:142-228
//   supernodal rows / flops     src/symbolic/supernodes.jl:285-339
//   panel layout (panel_ptr)    src/symbolic/layout.jl:271-276
//   relind                      src/symbolic/maps.jl:110-131
//   amap + owner grouping       src/symbolic/maps.jl:145-187
//   child lists                 src/symbolic/maps.jl:204-217
//   full-pattern residual map   src/symbolic/pattern.jl:243-290, src/solve/refinement.jl:67-101
// Orderings: natural, user permutation, geometric nested dissection of grids
// (METIS/AMD are not available in this environment) and an approximate minimum
// degree of our own (src/symbolic/ordering.jl picks AMD.jl/Metis.jl there).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using i64 = int64_t;
using V = std::vector<i64>;

enum Structure { ST_G = 0, ST_S = 1, ST_H = 2, ST_SPD = 3, ST_HPD = 4 };
enum View { VW_F = 0, VW_L = 1, VW_U = 2 };

struct Analysis {
    i64 n = 0, nnz = 0, ns = 0;
    int structure = 0, view = 0;
    i64 nnz_stored = 0, nnz_L = 0;
    i64 nschur = 0;  // Schur complement mode: the last supernode's columns (0 = off)
    i64 pair_rounds = 0;  // rounds of the "default" pair search that added pairs
    double flops = 0.0, flops_true = 0.0;
    std::map<std::string, V> f;  // named 1-based integer arrays
    std::string err;
};

thread_local std::string g_err;

// ----------------------------------------------------------------------------
// symmetric pattern (0-based adjacency lists, sorted, unique, no diagonal)

bool keep_entry(int structure, int view, i64 r, i64 c) {
    if (structure == ST_G) return true;
    if (view == VW_U) return c >= r;
    return c <= r;  // 'L' and 'F' read the lower triangle
}

struct Pattern {
    i64 n;
    V colptr, rowval;
    const i64* nb(i64 j) const { return rowval.data() + colptr[j]; }
    i64 deg(i64 j) const { return colptr[j + 1] - colptr[j]; }
};

Pattern symmetric_pattern(i64 n, const V& rp, const V& cv, int structure, int view) {
    V cnt(n + 1, 0);
    for (i64 r = 0; r < n; ++r)
        for (i64 p = rp[r]; p < rp[r + 1]; ++p) {
            i64 c = cv[p];
            if (c == r || !keep_entry(structure, view, r, c)) continue;
            cnt[r + 1]++;
            cnt[c + 1]++;
        }
    for (i64 j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
    V rv(cnt[n]);
    V next(cnt.begin(), cnt.end() - 1);
    for (i64 r = 0; r < n; ++r)
        for (i64 p = rp[r]; p < rp[r + 1]; ++p) {
            i64 c = cv[p];
            if (c == r || !keep_entry(structure, view, r, c)) continue;
            rv[next[c]++] = r;
            rv[next[r]++] = c;
        }
    Pattern P;
    P.n = n;
    P.colptr.assign(n + 1, 0);
    i64 out = 0;
    for (i64 j = 0; j < n; ++j) {
        auto b = rv.begin() + cnt[j], e = rv.begin() + cnt[j + 1];
        std::sort(b, e);
        i64 prev = -1;
        for (auto it = b; it != e; ++it) {
            if (*it == prev) continue;
            rv[out++] = *it;
            prev = *it;
        }
        P.colptr[j + 1] = out;
    }
    rv.resize(out);
    P.rowval = std::move(rv);
    return P;
}

// ----------------------------------------------------------------------------
// elimination tree, postorder, column counts (0-based, -1 = root)

V etree(const Pattern& P, const V& perm, const V& iperm) {
    i64 n = P.n;
    V parent(n, -1), anc(n, -1);
    for (i64 k = 0; k < n; ++k) {
        const i64* nb = P.nb(perm[k]);
        for (i64 t = 0; t < P.deg(perm[k]); ++t) {
            i64 i = iperm[nb[t]];
            if (i >= k) continue;
            while (i != -1 && i < k) {  // climb with path compression towards k
                i64 inext = anc[i];
                anc[i] = k;
                if (inext == -1) parent[i] = k;
                i = inext;
            }
        }
    }
    return parent;
}

V postorder(const V& parent) {
    i64 n = (i64)parent.size();
    V head(n, -1), next(n, -1);
    for (i64 j = n - 1; j >= 0; --j) {  // reverse insertion: increasing child lists
        i64 p = parent[j];
        if (p < 0) continue;
        next[j] = head[p];
        head[p] = j;
    }
    V post(n);
    V stack;
    i64 k = 0;
    for (i64 root = 0; root < n; ++root) {
        if (parent[root] >= 0) continue;
        stack.push_back(root);
        while (!stack.empty()) {
            i64 p = stack.back();
            i64 c = head[p];
            if (c < 0) {
                stack.pop_back();
                post[k++] = p;
            } else {
                head[p] = next[c];
                stack.push_back(c);
            }
        }
    }
    if (k != n) throw std::runtime_error("parent does not describe a forest");
    return post;
}

V colcounts(const Pattern& P, const V& perm, const V& iperm, const V& parent, const V& post) {
    i64 n = P.n;
    V delta(n, 0), first(n, -1), maxfirst(n, -1), prevleaf(n, -1), anc(n);
    std::iota(anc.begin(), anc.end(), 0);
    for (i64 k = 0; k < n; ++k) {
        i64 j = post[k];
        delta[j] = first[j] == -1 ? 1 : 0;
        while (j != -1 && first[j] == -1) {
            first[j] = k;
            j = parent[j];
        }
    }
    for (i64 k = 0; k < n; ++k) {
        i64 j = post[k];
        if (parent[j] != -1) delta[parent[j]]--;
        const i64* nb = P.nb(perm[j]);
        for (i64 t = 0; t < P.deg(perm[j]); ++t) {
            i64 i = iperm[nb[t]];
            // Gilbert-Ng-Peyton leaf test of row subtree i
            if (i <= j || first[j] <= maxfirst[i]) continue;
            maxfirst[i] = first[j];
            i64 jprev = prevleaf[i];
            prevleaf[i] = j;
            delta[j]++;
            if (jprev == -1) continue;  // first leaf of the row subtree
            i64 q = jprev;
            while (q != anc[q]) q = anc[q];
            for (i64 s = jprev; s != q;) {
                i64 sp = anc[s];
                anc[s] = q;
                s = sp;
            }
            delta[q]--;  // subsequent leaf: q is the least common ancestor
        }
        if (parent[j] != -1) anc[j] = parent[j];
    }
    for (i64 k = 0; k < n; ++k) {
        i64 j = post[k];
        if (parent[j] != -1) delta[parent[j]] += delta[j];
    }
    return delta;
}

// ----------------------------------------------------------------------------
// supernodes

struct ColumnPartition {
    V order;      // supernodal column k = etree column order[k]
    V super_ptr;  // 0-based, ns + 1
    V snparent;   // -1 = root
};

V col2sn_of(const V& sptr) {
    i64 ns = (i64)sptr.size() - 1;
    V c2s(sptr.back());
    for (i64 s = 0; s < ns; ++s)
        for (i64 j = sptr[s]; j < sptr[s + 1]; ++j) c2s[j] = s;
    return c2s;
}

V relabel_parent(const V& parent, const V& order) {
    i64 n = (i64)order.size();
    V inv(n);
    for (i64 k = 0; k < n; ++k) inv[order[k]] = k;
    V par(n);
    for (i64 k = 0; k < n; ++k) par[k] = parent[order[k]] < 0 ? -1 : inv[parent[order[k]]];
    return par;
}

i64 trapezoid(i64 w, i64 f) { return w * f - (w * (w - 1)) / 2; }

// nschur > 0: the last nschur columns of the order (the Schur complement's) form one supernode of
// their own, and no other supernode reaches into them
ColumnPartition fundamental_supernodes(const V& parent, const V& post, const V& counts, i64 nschur = 0) {
    i64 n = (i64)parent.size();
    V par = relabel_parent(parent, post);
    V cnt(n);
    for (i64 k = 0; k < n; ++k) cnt[k] = counts[post[k]];
    V nchild(n, 0);
    for (i64 k = 0; k < n; ++k)
        if (par[k] >= 0) nchild[par[k]]++;
    ColumnPartition cp;
    cp.order = post;
    cp.super_ptr.push_back(0);
    for (i64 k = 1; k < n; ++k) {
        bool joins = par[k - 1] == k && nchild[k] == 1 && cnt[k - 1] == cnt[k] + 1;
        if (nschur > 0 && k >= n - nschur) joins = k > n - nschur;
        if (!joins) cp.super_ptr.push_back(k);
    }
    if (n > 0) cp.super_ptr.push_back(n);
    V c2s = col2sn_of(cp.super_ptr);
    i64 ns = (i64)cp.super_ptr.size() - 1;
    cp.snparent.assign(ns, -1);
    for (i64 s = 0; s < ns; ++s) {
        i64 pl = par[cp.super_ptr[s + 1] - 1];
        cp.snparent[s] = pl < 0 ? -1 : c2s[pl];
    }
    return cp;
}

// schur_root: the last supernode is the Schur complement's, which takes no children's columns
ColumnPartition amalgamate(const ColumnPartition& sn, const V& parent, const V& counts, i64 max_width,
                           double zero_fraction, i64 min_width, bool schur_root = false) {
    if (max_width < 1) throw std::runtime_error("amalgamate: max_width must be >= 1");
    i64 n = (i64)sn.order.size();
    V cnt(n);
    for (i64 k = 0; k < n; ++k) cnt[k] = counts[sn.order[k]];
    const V& uptr = sn.super_ptr;
    i64 nu = (i64)uptr.size() - 1;
    for (i64 u = 0; u < nu; ++u)
        for (i64 j = uptr[u]; j < uptr[u + 1]; ++j)
            if (cnt[j] != cnt[uptr[u]] - (j - uptr[u]))
                throw std::runtime_error("amalgamate needs exact supernodes");
    V par = relabel_parent(parent, sn.order);
    V c2u = col2sn_of(uptr);
    V uparent(nu);
    for (i64 u = 0; u < nu; ++u) {
        i64 pl = par[uptr[u + 1] - 1];
        uparent[u] = pl < 0 ? -1 : c2u[pl];
        if (uparent[u] >= 0 && uparent[u] <= u) throw std::runtime_error("order is not topological");
    }
    V width(nu), rows(nu), truennz(nu), zeros(nu, 0), into(nu, -1);
    i64 total = 0;
    for (i64 u = 0; u < nu; ++u) {
        width[u] = uptr[u + 1] - uptr[u];
        rows[u] = cnt[uptr[u]];
        truennz[u] = trapezoid(width[u], rows[u]);
        total += truennz[u];
    }
    i64 budget = (i64)std::floor(zero_fraction * (double)total);
    i64 used = 0;
    std::vector<V> children(nu);
    for (i64 u = 0; u < nu; ++u)
        if (uparent[u] >= 0) children[uparent[u]].push_back(u);
    for (i64 p = 0; p < nu; ++p) {
        if (schur_root && p == nu - 1) continue;
        V& kids = children[p];
        // most shared rows first, then narrow first, then by index
        std::sort(kids.begin(), kids.end(), [&](i64 a, i64 b) {
            i64 ka = -(rows[a] - width[a]), kb = -(rows[b] - width[b]);
            if (ka != kb) return ka < kb;
            if (width[a] != width[b]) return width[a] < width[b];
            return a < b;
        });
        for (i64 c : kids) {
            i64 w = width[c] + width[p];
            if (w > max_width) continue;
            i64 added = width[c] * (width[c] + rows[p] - rows[c]);
            if (used + added > budget) continue;
            i64 z = zeros[c] + zeros[p] + added;
            i64 t = truennz[c] + truennz[p];
            if (!((double)z <= zero_fraction * (double)t || w <= min_width)) continue;
            into[c] = p;
            used += added;
            width[p] = w;
            rows[p] += width[c];
            zeros[p] = z;
            truennz[p] = t;
        }
    }
    V rep(nu);
    std::iota(rep.begin(), rep.end(), 0);
    for (i64 u = nu - 1; u >= 0; --u)
        if (into[u] >= 0) rep[u] = rep[into[u]];
    V ids(nu, -1);
    i64 ns = 0;
    for (i64 u = 0; u < nu; ++u)
        if (into[u] < 0) ids[u] = ns++;
    V gparent(ns, -1);
    std::vector<V> members(ns);
    for (i64 u = 0; u < nu; ++u) {
        members[ids[rep[u]]].push_back(u);
        if (into[u] < 0 && uparent[u] >= 0) gparent[ids[u]] = ids[rep[uparent[u]]];
    }
    V gpost = postorder(gparent);
    ColumnPartition cp;
    V neworder;
    neworder.reserve(n);
    cp.super_ptr.push_back(0);
    for (i64 g : gpost) {
        for (i64 u : members[g])
            for (i64 j = uptr[u]; j < uptr[u + 1]; ++j) neworder.push_back(j);
        cp.super_ptr.push_back((i64)neworder.size());
    }
    V gpos(ns);
    for (i64 k = 0; k < ns; ++k) gpos[gpost[k]] = k;
    cp.snparent.resize(ns);
    for (i64 k = 0; k < ns; ++k) {
        i64 g = gpost[k];
        cp.snparent[k] = gparent[g] < 0 ? -1 : gpos[gparent[g]];
    }
    cp.order.resize(n);
    for (i64 k = 0; k < n; ++k) cp.order[k] = sn.order[neworder[k]];
    return cp;
}

// ----------------------------------------------------------------------------
// geometric nested dissection of an nx x ny x nz grid (natural row-major
// numbering x + nx (y + ny z)): the middle plane of the longest side is the
// separator, ordered after both halves; boxes of at most `leaf` points are
// numbered naturally

void nd_box(i64 nx, i64 ny, i64 x0, i64 x1, i64 y0, i64 y1, i64 z0, i64 z1, i64 leaf, V& out) {
    i64 lx = x1 - x0, ly = y1 - y0, lz = z1 - z0;
    if (lx <= 0 || ly <= 0 || lz <= 0) return;
    if (lx * ly * lz <= leaf || (lx <= 2 && ly <= 2 && lz <= 2)) {
        for (i64 z = z0; z < z1; ++z)
            for (i64 y = y0; y < y1; ++y)
                for (i64 x = x0; x < x1; ++x) out.push_back(x + nx * (y + ny * z));
        return;
    }
    if (lx >= ly && lx >= lz) {
        i64 m = x0 + lx / 2;
        nd_box(nx, ny, x0, m, y0, y1, z0, z1, leaf, out);
        nd_box(nx, ny, m + 1, x1, y0, y1, z0, z1, leaf, out);
        nd_box(nx, ny, m, m + 1, y0, y1, z0, z1, 1 << 30, out);
    } else if (ly >= lz) {
        i64 m = y0 + ly / 2;
        nd_box(nx, ny, x0, x1, y0, m, z0, z1, leaf, out);
        nd_box(nx, ny, x0, x1, m + 1, y1, z0, z1, leaf, out);
        nd_box(nx, ny, x0, x1, m, m + 1, z0, z1, 1 << 30, out);
    } else {
        i64 m = z0 + lz / 2;
        nd_box(nx, ny, x0, x1, y0, y1, z0, m, leaf, out);
        nd_box(nx, ny, x0, x1, y0, y1, m + 1, z1, leaf, out);
        nd_box(nx, ny, x0, x1, y0, y1, m, m + 1, 1 << 30, out);
    }
}

// ----------------------------------------------------------------------------
// approximate minimum degree on the quotient graph (element absorption,
// approximate external degrees, mass elimination of indistinguishable nodes
// by hashing). A compact implementation of the method of Amestoy, Davis and
// Duff (SIAM J. Matrix Anal. Appl. 17, 1996); no dense-row detection and no
// aggressive absorption.

V amd_order(const Pattern& P) {
    const i64 n = P.n;
    if (n == 0) return V();
    // element/variable adjacency: for variable i, the list of elements (first
    // elen[i] entries) followed by the variables it is adjacent to
    std::vector<V> adj(n);
    for (i64 j = 0; j < n; ++j) adj[j].assign(P.nb(j), P.nb(j) + P.deg(j));
    V elen(n, 0);           // number of elements in adj[i]
    V nv(n, 1);             // supervariable size (0 = absorbed into another)
    V degree(n);            // approximate external degree
    std::vector<char> status(n, 0);  // 0 variable, 1 element, 2 absorbed/dead
    std::vector<V> lelem(n);          // variables of element e (live members)
    V w(n, 0);
    i64 wflg = 1;
    for (i64 i = 0; i < n; ++i) degree[i] = P.deg(i);
    // degree buckets
    V head(n + 1, -1), next(n, -1), prev(n, -1);
    auto bucket_insert = [&](i64 i) {
        i64 d = std::min(degree[i], n);
        next[i] = head[d];
        prev[i] = -1;
        if (head[d] >= 0) prev[head[d]] = i;
        head[d] = i;
    };
    auto bucket_remove = [&](i64 i) {
        i64 d = std::min(degree[i], n);
        if (prev[i] >= 0) next[prev[i]] = next[i];
        else head[d] = next[i];
        if (next[i] >= 0) prev[next[i]] = prev[i];
        next[i] = prev[i] = -1;
    };
    for (i64 i = 0; i < n; ++i) bucket_insert(i);
    V order;
    order.reserve(n);
    std::vector<V> merged(n);  // variables merged into supervariable i
    i64 mindeg = 0, nel = 0;
    V Lme;
    auto bump = [&]() {
        if (++wflg > (i64)1 << 60) {
            std::fill(w.begin(), w.end(), 0);
            wflg = 1;
        }
    };
    while (nel < n) {
        while (mindeg <= n && head[mindeg] < 0) ++mindeg;
        i64 me = head[mindeg];
        bucket_remove(me);
        // the new element me: union of its variables and its elements' variables
        bump();
        w[me] = wflg;
        Lme.clear();
        for (i64 t = 0; t < (i64)adj[me].size(); ++t) {
            i64 e = adj[me][t];
            if (t < elen[me]) {
                if (status[e] != 1) continue;
                for (i64 v : lelem[e])
                    if (status[v] == 0 && nv[v] > 0 && w[v] != wflg) {
                        w[v] = wflg;
                        Lme.push_back(v);
                    }
                status[e] = 2;  // absorbed into me
                lelem[e].clear();
            } else if (status[e] == 0 && nv[e] > 0 && w[e] != wflg) {
                w[e] = wflg;
                Lme.push_back(e);
            }
        }
        status[me] = 1;
        order.push_back(me);
        nel += nv[me];
        for (i64 v : merged[me]) order.push_back(v);
        merged[me].clear();
        adj[me].clear();
        elen[me] = 0;
        lelem[me] = Lme;
        // update the variables of me: prune absorbed elements and me's members
        for (i64 v : Lme) bucket_remove(v);
        // |Le \ Lme| for every element e adjacent to a variable of Lme
        bump();
        i64 tag = wflg;
        for (i64 v : Lme) w[v] = tag;  // mark Lme
        std::vector<std::pair<i64, i64>> ext;  // not needed beyond this scope
        std::map<i64, i64> we;                  // element -> |Le \ Lme|
        for (i64 v : Lme)
            for (i64 t = 0; t < elen[v]; ++t) {
                i64 e = adj[v][t];
                if (status[e] != 1 || e == me) continue;
                auto it = we.find(e);
                if (it == we.end()) {
                    i64 c = 0;
                    for (i64 u : lelem[e])
                        if (status[u] == 0 && nv[u] > 0) c += nv[u];
                    it = we.emplace(e, c).first;
                }
                it->second -= nv[v];
            }
        i64 lme_size = 0;
        for (i64 v : Lme) lme_size += nv[v];
        for (i64 v : Lme) {
            // rebuild adjacency: elements (me first), then variables not in Lme
            V nadj;
            nadj.push_back(me);
            i64 deg_e = 0;
            for (i64 t = 0; t < elen[v]; ++t) {
                i64 e = adj[v][t];
                if (status[e] != 1 || e == me) continue;
                i64 ext_e = we.count(e) ? we[e] : 0;
                if (ext_e <= 0) {  // element e is covered by me: absorb
                    status[e] = 2;
                    lelem[e].clear();
                    continue;
                }
                nadj.push_back(e);
                deg_e += ext_e;
            }
            i64 ne = (i64)nadj.size();
            i64 deg_v = 0;
            for (i64 t = elen[v]; t < (i64)adj[v].size(); ++t) {
                i64 u = adj[v][t];
                if (status[u] != 0 || nv[u] == 0 || w[u] == tag || u == me) continue;
                nadj.push_back(u);
                deg_v += nv[u];
            }
            adj[v] = std::move(nadj);
            elen[v] = ne;
            // approximate external degree (Amestoy-Davis-Duff): the smallest of the old degree
            // grown by the new element, the element-sum bound, and the variables left
            i64 d = std::min(degree[v] + (lme_size - nv[v]), (lme_size - nv[v]) + deg_e + deg_v);
            d = std::min(d, (n - nel) - nv[v]);
            degree[v] = std::max<i64>(d, 0);
        }
        // supervariable detection among Lme by hashing the adjacency lists
        std::map<std::pair<i64, i64>, V> hbuckets;
        for (i64 v : Lme) {
            if (nv[v] == 0) continue;
            V a = adj[v];
            std::sort(a.begin() + 1, a.begin() + elen[v]);
            std::sort(a.begin() + elen[v], a.end());
            adj[v] = a;
            i64 h = 0;
            for (i64 x : a) h = h * 1000003 + x;
            hbuckets[{(i64)a.size(), h}].push_back(v);
        }
        for (auto& kv : hbuckets) {
            V& vs = kv.second;
            for (size_t a = 0; a < vs.size(); ++a) {
                i64 i = vs[a];
                if (nv[i] == 0) continue;
                for (size_t b = a + 1; b < vs.size(); ++b) {
                    i64 j = vs[b];
                    if (nv[j] == 0 || elen[i] != elen[j] || adj[i] != adj[j]) continue;
                    nv[i] += nv[j];
                    nv[j] = 0;
                    status[j] = 2;
                    merged[i].push_back(j);
                    for (i64 x : merged[j]) merged[i].push_back(x);
                    merged[j].clear();
                    adj[j].clear();
                    elen[j] = 0;
                }
            }
        }
        V live;
        for (i64 v : Lme)
            if (nv[v] > 0) {
                degree[v] = std::max<i64>(0, std::min(degree[v], (n - nel) - nv[v]));
                live.push_back(v);
                bucket_insert(v);
                mindeg = std::min(mindeg, std::min(degree[v], n));
            }
        lelem[me] = live;
    }
    return order;
}

// ----------------------------------------------------------------------------
// nested dissection of a general graph (reordering_alg "algo4"; the reference takes a library
// nested dissection there, src/symbolic/ordering.jl:198-204): recursive bisection by level-
// structure separators (George & Liu, "An automatic nested dissection algorithm for irregular
// finite element problems", SIAM J. Numer. Anal. 15, 1978). From a pseudo-peripheral vertex, the
// breadth-first levels; the level where half the vertices are reached, less its vertices with no
// neighbour in the next level, separates the levels before it from those after; both parts are
// dissected in turn and the separator is numbered after them. Parts of at most ND_LEAF vertices,
// and parts whose level structure is too shallow to split, are ordered by minimum degree.

constexpr i64 ND_LEAF = 128;

struct NdWork {
    const Pattern& P;
    V loc;  // global vertex -> index in the current part, -1 outside
    V lev;  // breadth-first level within the current part, -1 unreached
    V out;  // the ordering, first eliminated first
};

// minimum-degree order of the part verts (loc must mark it), appended to out
void nd_leaf(NdWork& W, const V& verts) {
    const i64 m = (i64)verts.size();
    Pattern Q;
    Q.n = m;
    Q.colptr.assign(m + 1, 0);
    for (i64 a = 0; a < m; ++a) {
        const i64 v = verts[a];
        for (i64 t = 0; t < W.P.deg(v); ++t) {
            const i64 u = W.P.nb(v)[t];
            if (W.loc[u] >= 0) Q.rowval.push_back(W.loc[u]);
        }
        Q.colptr[a + 1] = (i64)Q.rowval.size();
        std::sort(Q.rowval.begin() + Q.colptr[a], Q.rowval.end());
    }
    for (i64 p : amd_order(Q)) W.out.push_back(verts[p]);
}

// breadth-first levels of the part from root (lev must be -1 on the part); returns the vertices
// in level order and the level starts
void nd_bfs(NdWork& W, i64 root, V& order, V& lstart) {
    order.assign(1, root);
    lstart.assign(1, 0);
    W.lev[root] = 0;
    for (size_t h = 0; h < order.size(); ++h) {
        const i64 v = order[h];
        if (W.lev[v] >= (i64)lstart.size()) lstart.push_back((i64)h);
        for (i64 t = 0; t < W.P.deg(v); ++t) {
            const i64 u = W.P.nb(v)[t];
            if (W.loc[u] >= 0 && W.lev[u] < 0) {
                W.lev[u] = W.lev[v] + 1;
                order.push_back(u);
            }
        }
    }
    lstart.push_back((i64)order.size());
}

i64 nd_part_degree(const NdWork& W, i64 v) {
    i64 d = 0;
    for (i64 t = 0; t < W.P.deg(v); ++t) d += W.loc[W.P.nb(v)[t]] >= 0;
    return d;
}

void nd_rec(NdWork& W, V verts, int depth) {
    const i64 m = (i64)verts.size();
    if (m == 0) return;
    for (i64 a = 0; a < m; ++a) W.loc[verts[a]] = a;
    auto release = [&](const V& vs) {
        for (i64 v : vs) {
            W.loc[v] = -1;
            W.lev[v] = -1;
        }
    };
    if (m <= ND_LEAF || depth > 200) {
        nd_leaf(W, verts);
        release(verts);
        return;
    }
    // connected components: dissected one after the other
    V order, lstart;
    nd_bfs(W, verts[0], order, lstart);
    if ((i64)order.size() < m) {
        std::vector<V> comps;
        comps.push_back(order);
        for (i64 v : verts)
            if (W.lev[v] < 0) {
                nd_bfs(W, v, order, lstart);
                comps.push_back(order);
            }
        release(verts);
        for (auto& c : comps) nd_rec(W, c, depth + 1);
        return;
    }
    // pseudo-peripheral root: from a vertex of smallest degree, the last level's vertex of smallest
    // degree, while the eccentricity grows
    for (i64 v : verts) W.lev[v] = -1;
    i64 root = verts[0], best = nd_part_degree(W, root);
    for (i64 v : verts) {
        const i64 d = nd_part_degree(W, v);
        if (d < best) {
            best = d;
            root = v;
        }
    }
    nd_bfs(W, root, order, lstart);
    for (int it = 0; it < 8; ++it) {
        const i64 ecc = (i64)lstart.size() - 2;
        i64 cand = order[lstart[ecc]], cd = nd_part_degree(W, cand);
        for (i64 h = lstart[ecc]; h < lstart[ecc + 1]; ++h)
            if (nd_part_degree(W, order[h]) < cd) {
                cand = order[h];
                cd = nd_part_degree(W, cand);
            }
        for (i64 v : verts) W.lev[v] = -1;
        V o2, l2;
        nd_bfs(W, cand, o2, l2);
        if ((i64)l2.size() - 2 <= ecc) {
            for (i64 v : verts) W.lev[v] = -1;
            nd_bfs(W, root, order, lstart);
            break;
        }
        root = cand;
        order.swap(o2);
        lstart.swap(l2);
    }
    const i64 nlev = (i64)lstart.size() - 1;
    if (nlev < 3) {
        nd_leaf(W, verts);
        release(verts);
        return;
    }
    // the level reaching half the vertices, kept off the first and the last level
    i64 mid = 1;
    while (mid < nlev - 2 && lstart[mid + 1] < (m + 1) / 2) ++mid;
    V A, B, sep;
    for (i64 h = 0; h < m; ++h) {
        const i64 v = order[h], l = W.lev[v];
        if (l < mid) {
            A.push_back(v);
        } else if (l > mid) {
            B.push_back(v);
        } else {
            bool up = false;
            for (i64 t = 0; t < W.P.deg(v) && !up; ++t) {
                const i64 u = W.P.nb(v)[t];
                up = W.loc[u] >= 0 && W.lev[u] == mid + 1;
            }
            (up ? sep : A).push_back(v);
        }
    }
    release(verts);
    nd_rec(W, A, depth + 1);
    nd_rec(W, B, depth + 1);
    for (i64 v : sep) W.out.push_back(v);
}

V nd_order(const Pattern& P) {
    NdWork W{P, V(P.n, -1), V(P.n, -1), {}};
    W.out.reserve(P.n);
    V all(P.n);
    std::iota(all.begin(), all.end(), 0);
    nd_rec(W, all, 0);
    return W.out;
}

// ordering_cost (src/symbolic/ordering.jl:86-129): flops (sum of the squared column counts) plus
// 1e8 flops per level of the amalgamated supernodal schedule, the automatic choice's model
double ordering_cost(const Pattern& Q, const V& perm, i64 max_width, double zero_fraction, i64 min_width,
                     int use_superpanels) {
    const i64 n = Q.n;
    V iperm(n);
    for (i64 k = 0; k < n; ++k) iperm[perm[k]] = k;
    const V parent = etree(Q, perm, iperm);
    const V post = postorder(parent);
    const V counts = colcounts(Q, perm, iperm, parent, post);
    double fl = 0.0;
    for (i64 c : counts) fl += (double)c * (double)c;
    ColumnPartition cp = fundamental_supernodes(parent, post, counts);
    if (use_superpanels) cp = amalgamate(cp, parent, counts, max_width, zero_fraction, min_width);
    const i64 ns = (i64)cp.snparent.size();
    V h(ns, 1);
    i64 depth = ns > 0 ? 1 : 0;
    for (i64 s = 0; s < ns; ++s) {  // a parent comes after its children
        depth = std::max(depth, h[s]);
        if (cp.snparent[s] >= 0) h[cp.snparent[s]] = std::max(h[cp.snparent[s]], h[s] + 1);
    }
    return fl + 1e8 * (double)depth;
}

// ----------------------------------------------------------------------------
// 2x2 pivot pairs of "S"/"H" (src/symbolic/pairs.jl): a row whose diagonal is absent or
// negligible is matched with a neighbour; the ordering runs on the graph compressed by the pairs
// (the partner eliminated right before its candidate) and the symbolic factorization on the pair
// pattern, where both columns of a pair carry the union structure, so that the pair lies in one
// supernode and in-front Bunch-Kaufman can take it as a 2x2 pivot.

// pivot_pairs settings (src/types.jl:196): none, "default" (the structurally zero pivots),
// "all" (every candidate)
enum PairMode { PAIRS_NONE = 0, PAIRS_DEFAULT = 1, PAIRS_ALL = 2 };
constexpr int PAIR_MAX_ROUNDS = 8;  // src/symbolic/pairs.jl:164

using Pairs = std::vector<std::pair<i64, i64>>;  // (partner, candidate), 0-based

// position of row r in column c's sorted adjacency (r must be a neighbour)
i64 adj_pos(const Pattern& P, i64 c, i64 r) {
    const i64* b = P.nb(c);
    return P.colptr[c] + (std::lower_bound(b, b + P.deg(c), r) - b);
}

struct PairCands {
    std::vector<char> cand;    // absent diagonal or |a_ii| <= tol max_{j != i} |a_ij|
    std::vector<double> w;     // |a_ij| aligned with P.rowval, duplicates summed first
    std::vector<double> amax;  // max_{j != i} |a_ij| (0 without off-diagonal entries)
};

// pivot_candidates (pairs.jl:60-99): the values read with the view rules of the pattern
PairCands pivot_candidates(const Pattern& P, const V& rp, const V& cv, const double* nz, int structure, int view,
                           double tol) {
    const i64 n = P.n;
    std::vector<double> dval(n, 0.0), wval(P.rowval.size(), 0.0);
    std::vector<char> dpres(n, 0);
    for (i64 r = 0; r < n; ++r)
        for (i64 p = rp[r]; p < rp[r + 1]; ++p) {
            const i64 c = cv[p];
            if (!keep_entry(structure, view, r, c)) continue;
            if (r == c) {
                dval[r] += nz[p];
                dpres[r] = 1;
            } else {
                wval[adj_pos(P, c, r)] += nz[p];
                wval[adj_pos(P, r, c)] += nz[p];
            }
        }
    PairCands pc;
    pc.w.resize(wval.size());
    for (size_t p = 0; p < wval.size(); ++p) pc.w[p] = std::fabs(wval[p]);
    pc.amax.assign(n, 0.0);
    pc.cand.assign(n, 0);
    for (i64 i = 0; i < n; ++i) {
        for (i64 p = P.colptr[i]; p < P.colptr[i + 1]; ++p) pc.amax[i] = std::max(pc.amax[i], pc.w[p]);
        pc.cand[i] = !dpres[i] || std::fabs(dval[i]) <= tol * pc.amax[i];
    }
    return pc;
}

// pivot_pairs "all" (pairs.jl:126-151): a greedy matching of the candidates to non-candidate
// neighbours with a_ij != 0, candidates by increasing count of such neighbours (then index),
// each taking the free neighbour of largest |a_ij| (ties: smaller degree, then smaller index)
Pairs pivot_pairs_all(const Pattern& P, const PairCands& pc) {
    const i64 n = P.n;
    auto eligible = [&](i64 p) { return !pc.cand[P.rowval[p]] && pc.w[p] > 0; };
    std::vector<std::pair<i64, i64>> order;  // (eligible count, candidate)
    for (i64 i = 0; i < n; ++i) {
        if (!pc.cand[i]) continue;
        i64 c = 0;
        for (i64 p = P.colptr[i]; p < P.colptr[i + 1]; ++p) c += eligible(p);
        order.push_back({c, i});
    }
    std::sort(order.begin(), order.end());
    std::vector<char> matched(n, 0);
    Pairs pairs;
    for (auto [cnt, i] : order) {
        (void)cnt;
        i64 best = -1;
        for (i64 p = P.colptr[i]; p < P.colptr[i + 1]; ++p) {
            const i64 j = P.rowval[p];
            if (!eligible(p) || matched[j]) continue;
            if (best < 0 || pc.w[p] > pc.w[best] || (pc.w[p] == pc.w[best] && P.deg(j) < P.deg(P.rowval[best])))
                best = p;
        }
        if (best < 0) continue;
        matched[P.rowval[best]] = 1;
        pairs.push_back({P.rowval[best], i});
    }
    std::sort(pairs.begin(), pairs.end(), [](auto& a, auto& b) { return a.second < b.second; });
    return pairs;
}

// structural_zero_pivots (pairs.jl:184-249): the candidates, not in a pair, whose pivot is
// structurally zero when A[perm, perm] is factored column by column, in elimination order. A
// perfect matching of the leading block is kept: a non-candidate takes its own diagonal, a pair
// (a, c) with c right after a is matched crosswise, a candidate searches an augmenting path from
// its row to its column through earlier columns (|a_ij| > 0 edges, and the diagonal of a
// non-candidate row that an earlier path moved off it); one without is reported (and matched to
// itself, as if a partner fixed it).
V structural_zero_pivots(const PairCands& pc, const Pattern& P, const V& perm, const Pairs& pairs) {
    const i64 n = P.n;
    V pos(n);
    for (i64 k = 0; k < n; ++k) pos[perm[k]] = k;
    V mate(n, -1);
    for (auto& [a, b] : pairs) {
        mate[a] = b;
        mate[b] = a;
    }
    V colm(n, -1), seen(n, -1), out;
    V srow, sptr, scol;
    for (i64 k = 0; k < n; ++k) {
        const i64 v = perm[k];
        const i64 m = mate[v];
        if (m >= 0 && pos[m] == k - 1) {  // the second column of a pair: the 2x2 block, crosswise
            colm[v] = m;
            colm[m] = v;
            continue;
        }
        if (!pc.cand[v] || m >= 0) {  // a nonzero diagonal, or the first column of a pair
            colm[v] = v;
            continue;
        }
        srow.assign(1, v);
        sptr.assign(1, P.colptr[v]);
        scol.assign(1, -1);
        bool found = false;
        while (!srow.empty()) {
            const i64 r = srow.back(), p = sptr.back();
            if (p > P.colptr[r + 1]) {  // row r exhausted: backtrack
                srow.pop_back();
                sptr.pop_back();
                scol.pop_back();
                continue;
            }
            sptr.back() = p + 1;
            i64 j;
            if (p == P.colptr[r + 1]) {  // the last slot: the diagonal of a non-candidate row
                if (pc.cand[r]) continue;
                j = r;
            } else {
                j = P.rowval[p];
                if (!(pc.w[p] > 0)) continue;
            }
            if (!(pos[j] <= k && seen[j] != k)) continue;
            seen[j] = k;
            scol.back() = j;
            if (j == v) {  // reached the free column: flip the path
                for (size_t t = 0; t < srow.size(); ++t) colm[scol[t]] = srow[t];
                found = true;
                break;
            }
            srow.push_back(colm[j]);
            sptr.push_back(P.colptr[colm[j]]);
            scol.push_back(-1);
        }
        if (!found) {
            out.push_back(v);
            colm[v] = v;
        }
    }
    return out;
}

// one round of pivot_pairs "default" (pairs.jl:283-318): pair the structurally zero pivots under
// perm with a free neighbour j, |a_ij| >= u max_k |a_ik| (another candidate allowed), taken by
// increasing count of such partners, then by position in perm, each the one of largest |a_ij|
// (ties: earliest in perm). Returns the pairs added.
i64 zero_pivot_pairs(Pairs& pairs, const PairCands& pc, const Pattern& P, const V& perm, double u) {
    const V zs = structural_zero_pivots(pc, P, perm, pairs);
    const i64 n = P.n;
    std::vector<char> inpair(n, 0);
    for (auto& [a, b] : pairs) inpair[a] = inpair[b] = 1;
    V pos(n);
    for (i64 k = 0; k < n; ++k) pos[perm[k]] = k;
    auto eligible = [&](i64 i, i64 p) { return !inpair[P.rowval[p]] && pc.w[p] > 0 && pc.w[p] >= u * pc.amax[i]; };
    std::vector<std::pair<std::pair<i64, i64>, i64>> order;  // ((free count, position), candidate)
    for (i64 i : zs) {
        i64 c = 0;
        for (i64 p = P.colptr[i]; p < P.colptr[i + 1]; ++p) c += eligible(i, p);
        order.push_back({{c, pos[i]}, i});
    }
    std::sort(order.begin(), order.end());
    i64 added = 0;
    for (auto& o : order) {
        const i64 i = o.second;
        if (inpair[i]) continue;  // taken as the partner of an earlier candidate
        i64 best = -1;
        for (i64 p = P.colptr[i]; p < P.colptr[i + 1]; ++p) {
            if (!eligible(i, p)) continue;
            if (best < 0 || pc.w[p] > pc.w[best] || (pc.w[p] == pc.w[best] && pos[P.rowval[p]] < pos[P.rowval[best]]))
                best = p;
        }
        if (best < 0) continue;
        const i64 j = P.rowval[best];
        inpair[j] = inpair[i] = 1;
        pairs.push_back({j, i});
        ++added;
    }
    std::sort(pairs.begin(), pairs.end(), [](auto& a, auto& b) { return a.second < b.second; });
    return added;
}

// symmetric pattern from (row, column) pairs (both triangles given), sorted and unique
Pattern pattern_from_edges(i64 n, std::vector<std::pair<i64, i64>>& e) {
    std::sort(e.begin(), e.end(), [](auto& a, auto& b) { return a.second != b.second ? a.second < b.second : a.first < b.first; });
    e.erase(std::unique(e.begin(), e.end()), e.end());
    Pattern Q;
    Q.n = n;
    Q.colptr.assign(n + 1, 0);
    Q.rowval.reserve(e.size());
    for (auto& [r, c] : e) {
        Q.colptr[c + 1]++;
        Q.rowval.push_back(r);
    }
    for (i64 j = 0; j < n; ++j) Q.colptr[j + 1] += Q.colptr[j];
    return Q;
}

// the groups of the compressed graph (pairs.jl:388-411): a pair is one vertex (partner first),
// every other column one vertex, numbered by increasing first column
void pair_groups(i64 n, const Pairs& pairs, V& gptr, V& members) {
    V mate(n, -1);
    std::vector<char> paired(n, 0);
    for (auto& [a, b] : pairs) {
        if (a < 0 || a >= n || b < 0 || b >= n || a == b || paired[a] || paired[b])
            throw std::runtime_error("pivot pairs must be disjoint pairs of distinct indices");
        mate[a] = b;
        paired[a] = paired[b] = 1;
    }
    gptr.assign(1, 0);
    members.clear();
    for (i64 v = 0; v < n; ++v) {
        if (mate[v] >= 0) {
            members.push_back(v);
            members.push_back(mate[v]);
        } else if (!paired[v]) {
            members.push_back(v);
        } else {
            continue;  // a candidate: emitted with its partner
        }
        gptr.push_back((i64)members.size());
    }
}

// compressed_pattern (pairs.jl:421-438): a pair's adjacency is the union of its rows
Pattern compressed_pattern(const Pattern& P, const Pairs& pairs, V& gptr, V& members) {
    pair_groups(P.n, pairs, gptr, members);
    const i64 ng = (i64)gptr.size() - 1;
    V group(P.n);
    for (i64 g = 0; g < ng; ++g)
        for (i64 k = gptr[g]; k < gptr[g + 1]; ++k) group[members[k]] = g;
    std::vector<std::pair<i64, i64>> e;
    for (i64 j = 0; j < P.n; ++j)
        for (i64 t = 0; t < P.deg(j); ++t) {
            const i64 gi = group[P.nb(j)[t]], gj = group[j];
            if (gi != gj) e.push_back({gi, gj});
        }
    return pattern_from_edges(ng, e);
}

// expand_permutation (pairs.jl:447-456): the columns of each vertex in turn, partner first
V expand_permutation(const V& cperm, const V& gptr, const V& members) {
    V perm;
    perm.reserve(members.size());
    for (i64 g : cperm)
        for (i64 k = gptr[g]; k < gptr[g + 1]; ++k) perm.push_back(members[k]);
    return perm;
}

// pair_pattern (pairs.jl:469-485): the expansion of the compressed pattern, both columns of a pair
// with the union structure and adjacent to each other
Pattern pair_pattern(const Pattern& P, const Pairs& pairs) {
    if (pairs.empty()) return P;
    V gptr, members;
    const Pattern Pc = compressed_pattern(P, pairs, gptr, members);
    std::vector<std::pair<i64, i64>> e;
    for (i64 g = 0; g < Pc.n; ++g) {
        for (i64 a = gptr[g]; a < gptr[g + 1]; ++a)
            for (i64 b = gptr[g]; b < gptr[g + 1]; ++b)
                if (a != b) e.push_back({members[a], members[b]});
        for (i64 t = 0; t < Pc.deg(g); ++t) {
            const i64 h = Pc.nb(g)[t];
            for (i64 a = gptr[g]; a < gptr[g + 1]; ++a)
                for (i64 b = gptr[h]; b < gptr[h + 1]; ++b) e.push_back({members[b], members[a]});
        }
    }
    return pattern_from_edges(P.n, e);
}

// ----------------------------------------------------------------------------
// MC64-style matching and scaling (src/matching/mc64.jl, Duff & Koster 2001): the rows of the
// stored matrix matched to its columns, on the host, from the values

// the binary heap of (key, vertex) with lazy deletion of the successive-shortest-path search
struct MinHeap {
    std::vector<double> k;
    V v;
    void push(double key, i64 x) {
        k.push_back(key);
        v.push_back(x);
        size_t i = k.size() - 1;
        while (i > 0) {
            const size_t p = (i - 1) / 2;
            if (k[p] <= k[i]) break;
            std::swap(k[p], k[i]);
            std::swap(v[p], v[i]);
            i = p;
        }
    }
    void pop(double& key, i64& x) {
        key = k[0];
        x = v[0];
        const double lk = k.back();
        const i64 lv = v.back();
        k.pop_back();
        v.pop_back();
        const size_t m = k.size();
        if (m == 0) return;
        k[0] = lk;
        v[0] = lv;
        size_t i = 0;
        for (;;) {
            const size_t l = 2 * i + 1;
            if (l >= m) break;
            const size_t c = (l + 1 < m && k[l + 1] < k[l]) ? l + 1 : l;
            if (k[i] <= k[c]) break;
            std::swap(k[i], k[c]);
            std::swap(v[i], v[c]);
            i = c;
        }
    }
};

// min_cost_matching (mc64.jl:101-181): minimum-cost matching of maximum cardinality of the left
// vertices 0..n-1 (adjacency adj[ptr[i] .. ptr[i+1]), costs >= 0) to the right ones, by successive
// shortest augmenting paths (Dijkstra on the reduced costs cost - u_i - v_j >= 0); lmate[i] the
// right vertex of i (-1 none), u and v the duals (u_i + v_j <= cost, equality on matched edges)
void min_cost_matching(const V& ptr, const V& adj, const std::vector<double>& cost, i64 n, V& lmate,
                       std::vector<double>& u, std::vector<double>& v) {
    lmate.assign(n, -1);
    V rmate(n, -1);
    u.assign(n, 0.0);
    v.assign(n, 0.0);
    for (i64 i = 0; i < n; ++i) {  // feasible start: u = the row minimum; the tight edges greedily
        double mn = 0.0;
        bool first = true;
        for (i64 p = ptr[i]; p < ptr[i + 1]; ++p) {
            mn = first ? cost[p] : std::min(mn, cost[p]);
            first = false;
        }
        u[i] = first ? 0.0 : std::min(0.0, mn);  // the minimum with the initial value 0
        for (i64 p = ptr[i]; p < ptr[i + 1]; ++p) {
            const i64 j = adj[p];
            if (rmate[j] < 0 && cost[p] - u[i] <= 0) {
                lmate[i] = j;
                rmate[j] = i;
                break;
            }
        }
    }
    std::vector<double> dist(n, INFINITY);
    V pred(n, -1), touched, finals;
    std::vector<char> done(n, 0);
    MinHeap hp;
    for (i64 i0 = 0; i0 < n; ++i0) {
        if (lmate[i0] >= 0) continue;
        touched.clear();
        finals.clear();
        hp.k.clear();
        hp.v.clear();
        i64 cur = i0, jfree = -1;
        double dcur = 0.0, L = INFINITY;
        for (;;) {
            for (i64 p = ptr[cur]; p < ptr[cur + 1]; ++p) {
                const i64 j = adj[p];
                if (done[j]) continue;
                const double nd = dcur + std::max(cost[p] - u[cur] - v[j], 0.0);
                if (nd < dist[j]) {
                    if (std::isinf(dist[j])) touched.push_back(j);
                    dist[j] = nd;
                    pred[j] = cur;
                    hp.push(nd, j);
                }
            }
            i64 j = -1;
            while (!hp.k.empty()) {
                double d;
                i64 jj;
                hp.pop(d, jj);
                if (done[jj] || d > dist[jj]) continue;
                j = jj;
                break;
            }
            if (j < 0) break;  // no augmenting path: i0 stays unmatched
            done[j] = 1;
            finals.push_back(j);
            if (rmate[j] < 0) {
                jfree = j;
                L = dist[j];
                break;
            }
            cur = rmate[j];
            dcur = dist[j];  // the matched edge has reduced cost 0
        }
        if (jfree >= 0) {
            u[i0] += L;
            for (i64 j : finals) {
                if (j == jfree) continue;
                u[rmate[j]] += L - dist[j];
                v[j] -= L - dist[j];
            }
            i64 j = jfree;
            for (;;) {  // flip the path
                const i64 i = pred[j];
                const i64 nj = lmate[i];
                lmate[i] = j;
                rmate[j] = i;
                if (i == i0) break;
                j = nj;
            }
        }
        for (i64 j : touched) {
            dist[j] = INFINITY;
            done[j] = 0;
        }
    }
}

// the full matrix (both triangles of a symmetric structure) as rows -> columns with |a_ij|,
// duplicates summed first, explicit zeros kept (mc64.jl:184-207)
void abs_csr(i64 n, const V& rp, const V& cv, const double* nz, int structure, int view, V& ptr, V& adj,
             std::vector<double>& w) {
    std::vector<std::map<i64, double>> rows(n);
    for (i64 r = 0; r < n; ++r)
        for (i64 p = rp[r]; p < rp[r + 1]; ++p) {
            const i64 c = cv[p];
            if (structure == ST_G) {
                rows[r][c] += nz[p];
                continue;
            }
            if (!keep_entry(structure, view, r, c)) continue;
            rows[r][c] += nz[p];
            if (r != c) rows[c][r] += nz[p];
        }
    ptr.assign(n + 1, 0);
    adj.clear();
    w.clear();
    for (i64 r = 0; r < n; ++r) {
        for (auto& [c, x] : rows[r]) {
            adj.push_back(c);
            w.push_back(std::fabs(x));
        }
        ptr[r + 1] = (i64)adj.size();
    }
}

// maximum cardinality matching on the kept edges (zero costs)
V cardinality_matching(const V& ptr, const V& adj, const std::vector<char>& keep, i64 n) {
    V sp(n + 1, 0), sa;
    for (i64 i = 0; i < n; ++i) {
        for (i64 p = ptr[i]; p < ptr[i + 1]; ++p)
            if (keep[p]) sa.push_back(adj[p]);
        sp[i + 1] = (i64)sa.size();
    }
    V lm;
    std::vector<double> u, v;
    min_cost_matching(sp, sa, std::vector<double>(sa.size(), 0.0), n, lm, u, v);
    return lm;
}

i64 matched_count(const V& lm) {
    i64 c = 0;
    for (i64 x : lm) c += x >= 0;
    return c;
}

// mc64 (mc64.jl:265-312): job 1 maximum cardinality; jobs 2 and 3 the largest smallest matched
// |a_ij| (bisection over the distinct values); job 4 the largest sum; job 5 the largest product
// with the scalings r_i = exp(u_i) / max_j |a_ij|, c_j = exp(v_j) (|r_i a_ij c_j| <= 1, equality
// on the matched entries). Jobs 2-5 use the nonzero entries only.
void mc64(const V& ptr, const V& adj, const std::vector<double>& w, i64 n, int job, V& lmate,
          std::vector<double>& rs, std::vector<double>& cs) {
    rs.assign(n, 1.0);
    cs.assign(n, 1.0);
    if (job == 1) {
        lmate = cardinality_matching(ptr, adj, std::vector<char>(w.size(), 1), n);
        return;
    }
    if (job == 2 || job == 3) {
        const V full = cardinality_matching(ptr, adj, std::vector<char>(w.size(), 1), n);
        const i64 target = matched_count(full);
        std::vector<double> vals;
        for (double x : w)
            if (x > 0) vals.push_back(x);
        std::sort(vals.begin(), vals.end());
        vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
        if (vals.empty()) {
            lmate = full;
            return;
        }
        auto with = [&](double t) {
            std::vector<char> keep(w.size());
            for (size_t p = 0; p < w.size(); ++p) keep[p] = w[p] >= t;
            return cardinality_matching(ptr, adj, keep, n);
        };
        V best = with(vals[0]);
        if (matched_count(best) != target) {
            lmate = full;
            return;
        }
        i64 lo = 0, hi = (i64)vals.size() - 1;
        while (lo < hi) {
            const i64 mid = (lo + hi + 1) / 2;
            V m = with(vals[mid]);
            if (matched_count(m) == target) {
                lo = mid;
                best = m;
            } else {
                hi = mid - 1;
            }
        }
        lmate = best;
        return;
    }
    if (job != 4 && job != 5) throw std::runtime_error("unknown matching job; expected 1 to 5");
    std::vector<double> rowmax(n, 0.0);
    for (i64 i = 0; i < n; ++i)
        for (i64 p = ptr[i]; p < ptr[i + 1]; ++p) rowmax[i] = std::max(rowmax[i], w[p]);
    V sp(n + 1, 0), sa;
    std::vector<double> cost;
    for (i64 i = 0; i < n; ++i) {
        for (i64 p = ptr[i]; p < ptr[i + 1]; ++p) {
            if (!(w[p] > 0)) continue;
            sa.push_back(adj[p]);
            cost.push_back(job == 5 ? std::log(rowmax[i]) - std::log(w[p]) : rowmax[i] - w[p]);
        }
        sp[i + 1] = (i64)sa.size();
    }
    std::vector<double> u, v;
    min_cost_matching(sp, sa, cost, n, lmate, u, v);
    if (job == 5) {
        for (i64 i = 0; i < n; ++i) {
            rs[i] = rowmax[i] > 0 ? std::exp(u[i]) / rowmax[i] : 1.0;
            cs[i] = std::exp(v[i]);
        }
        for (auto* s : {&rs, &cs})
            for (double& x : *s)
                if (!(std::isfinite(x) && x > 0)) x = 1.0;
    }
}

// compute_matching (mc64.jl:370-392): "G" the row/column scalings of the job; the symmetric
// structures d_i = sqrt(r_i c_i) (Duff & Pralet 2005), an unmatched index (structurally
// singular) getting 1 / max_j |a_ij| d_j over its matched neighbours; the matching completed to a
// permutation (unmatched rows take the unmatched columns in increasing order)
void compute_matching(i64 n, const V& rp, const V& cv, const double* nz, int structure, int view, int job, V& perm,
                      std::vector<char>& matched, std::vector<double>& rs, std::vector<double>& cs) {
    V ptr, adj, lmate;
    std::vector<double> w;
    abs_csr(n, rp, cv, nz, structure, view, ptr, adj, w);
    mc64(ptr, adj, w, n, job, lmate, rs, cs);
    matched.assign(n, 0);
    std::vector<char> used(n, 0);
    for (i64 i = 0; i < n; ++i)
        if (lmate[i] >= 0) {
            matched[i] = 1;
            used[lmate[i]] = 1;
        }
    perm = lmate;
    i64 nxt = 0;
    for (i64 i = 0; i < n; ++i) {
        if (perm[i] >= 0) continue;
        while (used[nxt]) ++nxt;
        perm[i] = nxt;
        used[nxt] = 1;
    }
    if (structure == ST_G) return;
    std::vector<double> d(n);
    for (i64 i = 0; i < n; ++i) d[i] = std::sqrt(rs[i] * cs[i]);
    for (i64 i = 0; i < n; ++i) {
        if (matched[i]) continue;
        double m = 0.0;
        for (i64 p = ptr[i]; p < ptr[i + 1]; ++p)
            if (matched[adj[p]]) m = std::max(m, w[p] * d[adj[p]]);
        d[i] = m > 0 ? 1.0 / m : 1.0;
    }
    for (double& x : d)
        if (!(std::isfinite(x) && x > 0)) x = 1.0;
    rs = d;
    cs = d;
}

// matching_pairs (mc64.jl:413-461): the 2x2 pivot pairs of "S"/"H" from the cycles of the
// symmetric matching (matched indices only): a 2-cycle is a pair, a longer cycle is split into
// pairs of consecutive indices (even length: the alternative with the larger product of scaled
// |d_i a_ij d_j|; odd: leaving out the index with the largest scaled diagonal); "default" keeps a
// pair only when one of its indices is a 2x2 candidate of the scaled matrix; the candidate is the
// index with the smaller relative scaled diagonal
Pairs matching_pairs(i64 n, const V& rp, const V& cv, const double* nz, int structure, int view, const V& perm,
                     const std::vector<char>& matched, const std::vector<double>& d, int mode, double tol) {
    Pairs pairs;
    if (mode == PAIRS_NONE) return pairs;
    V ptr, adj;
    std::vector<double> w;
    abs_csr(n, rp, cv, nz, structure, view, ptr, adj, w);
    std::vector<double> dg(n, 0.0), offmax(n, 0.0);
    for (i64 i = 0; i < n; ++i)
        for (i64 p = ptr[i]; p < ptr[i + 1]; ++p) {
            const double x = d[i] * w[p] * d[adj[p]];
            if (adj[p] == i) dg[i] = x;
            else offmax[i] = std::max(offmax[i], x);
        }
    auto entry = [&](i64 i, i64 j) {
        for (i64 p = ptr[i]; p < ptr[i + 1]; ++p)
            if (adj[p] == j) return d[i] * w[p] * d[j];
        return 0.0;
    };
    auto ratio = [&](i64 i) { return offmax[i] > 0 ? dg[i] / offmax[i] : INFINITY; };
    auto candidate = [&](i64 i) { return dg[i] <= tol * offmax[i]; };
    std::vector<char> seen(n, 0);
    V cyc;
    for (i64 i0 = 0; i0 < n; ++i0) {
        if (seen[i0] || !matched[i0]) continue;
        cyc.clear();
        i64 i = i0;
        bool ok = true;
        while (!seen[i]) {
            seen[i] = 1;
            cyc.push_back(i);
            if (!matched[i]) ok = false;
            i = perm[i];
        }
        const i64 L = (i64)cyc.size();
        if (!(ok && i == i0 && L >= 2)) continue;
        i64 start;  // 0-based position of the first pair's first index
        if (L % 2 == 0) {
            double p0 = 1.0, p1 = 1.0;
            for (i64 k = 0; k < L; k += 2) {
                p0 *= entry(cyc[k], cyc[k + 1]);
                p1 *= entry(cyc[k + 1], cyc[(k + 2) % L]);
            }
            start = p1 > p0 ? 1 : 0;
        } else {
            i64 am = 0;  // the largest scaled diagonal, first on ties
            for (i64 k = 1; k < L; ++k)
                if (dg[cyc[k]] > dg[cyc[am]]) am = k;
            start = (am + 1) % L;  // the index after the singleton
        }
        for (i64 t = 0; t + 1 < L; t += 2) {
            const i64 a = cyc[(start + t) % L], b = cyc[(start + t + 1) % L];
            if (mode != PAIRS_ALL && !candidate(a) && !candidate(b)) continue;
            pairs.push_back(ratio(a) < ratio(b) ? std::make_pair(b, a) : std::make_pair(a, b));
        }
    }
    std::sort(pairs.begin(), pairs.end(), [](auto& a, auto& b) { return a.second < b.second; });
    return pairs;
}

// ----------------------------------------------------------------------------

Analysis* analyze(i64 n, const i64* rowptr_in, const i64* colval_in, int structure, int view, int index_base,
                  const i64* perm_in, int ordering, i64 max_width, double zero_fraction, i64 min_width,
                  int use_superpanels, const int8_t* schur, const double* values, int pair_mode, double pair_tol,
                  double pivot_u, const i64* fixed_pairs, i64 nfixed) {
    auto A = new Analysis();
    A->n = n;
    A->structure = structure;
    A->view = view;
    i64 base = index_base;  // 0 or 1
    V rp(n + 1), cv;
    for (i64 i = 0; i <= n; ++i) rp[i] = rowptr_in[i] - base;
    if (rp[0] != 0) throw std::runtime_error("rowptr must start at the index base");
    for (i64 r = 0; r < n; ++r)
        if (rp[r] > rp[r + 1]) throw std::runtime_error("rowptr is not nondecreasing");
    i64 nnz = rp[n];
    A->nnz = nnz;
    cv.resize(nnz);
    for (i64 p = 0; p < nnz; ++p) {
        cv[p] = colval_in[p] - base;
        if (cv[p] < 0 || cv[p] >= n) throw std::runtime_error("column index out of range");
    }
    Pattern P = symmetric_pattern(n, rp, cv, structure, view);

    // ordering: etree numbering column k = original perm0[k]
    V perm0(n);
    Pairs pairs;
    int pair_rounds = 0;
    if (perm_in) {
        std::vector<char> seen(n, 0);
        for (i64 k = 0; k < n; ++k) {
            perm0[k] = perm_in[k] - 1;
            if (perm0[k] < 0 || perm0[k] >= n || seen[perm0[k]]) throw std::runtime_error("user_perm is not a permutation");
            seen[perm0[k]] = 1;
        }
    } else if (ordering == 3 || ordering == 4 || ordering == 6) {
        // 3 approximate minimum degree, 4 nested dissection, 6 the automatic choice between the two
        // (src/symbolic/ordering.jl:279-326): per algorithm its ordering, with the 2x2 pivot pairs
        // of "S"/"H" (not with the Schur complement mode; pairs.jl:27-35): fixed pairs (those of a
        // matching) order the compressed graph; "all" pairs every candidate first; "default" pairs
        // the structurally zero pivots of the ordering and re-orders the compressed graph until a
        // round adds no pair. The choice takes the fewest structurally zero pivots left, then the
        // lower cost (ties keep the minimum degree).
        const bool sym = structure == ST_S || structure == ST_H;
        const bool fixedp = sym && !schur && nfixed > 0;
        const bool search = sym && !schur && !fixedp && pair_mode != PAIRS_NONE && values;
        PairCands pc;
        if (search) pc = pivot_candidates(P, rp, cv, values, structure, view, pair_tol);
        if (fixedp)
            for (i64 t = 0; t < nfixed; ++t) pairs.push_back({fixed_pairs[2 * t] - 1, fixed_pairs[2 * t + 1] - 1});
        auto base = [&](const Pattern& G, int alg) { return alg == 4 ? nd_order(G) : amd_order(G); };
        auto compressed_order = [&](const Pairs& prs, int alg) {
            V gptr, members;
            const Pattern Pc = compressed_pattern(P, prs, gptr, members);
            return expand_permutation(base(Pc, alg), gptr, members);
        };
        struct Cand {
            V perm;
            Pairs prs;
            int rounds = 0;
        };
        auto order_on = [&](int alg) {
            Cand c;
            if (fixedp) {
                c.prs = pairs;
                c.perm = compressed_order(c.prs, alg);
            } else if (search && pair_mode == PAIRS_ALL) {
                c.prs = pivot_pairs_all(P, pc);
                c.perm = compressed_order(c.prs, alg);
            } else {
                c.perm = base(P, alg);
                while (search && c.rounds < PAIR_MAX_ROUNDS) {
                    if (zero_pivot_pairs(c.prs, pc, P, c.perm, pivot_u) == 0) break;
                    ++c.rounds;
                    c.perm = compressed_order(c.prs, alg);
                }
            }
            return c;
        };
        const std::vector<int> algs = ordering == 6 ? std::vector<int>{3, 4} : std::vector<int>{ordering};
        Cand best;
        size_t best_zeros = 0;
        double best_cost = 0.0;
        for (size_t q = 0; q < algs.size(); ++q) {
            Cand c = order_on(algs[q]);
            if (algs.size() > 1) {
                const size_t zeros = search ? structural_zero_pivots(pc, P, c.perm, c.prs).size() : 0;
                const double cost = ordering_cost(pair_pattern(P, c.prs), c.perm, max_width, zero_fraction, min_width,
                                                  use_superpanels);
                if (q > 0 && !(zeros < best_zeros || (zeros == best_zeros && cost < best_cost))) continue;
                best_zeros = zeros;
                best_cost = cost;
            }
            best = std::move(c);
        }
        perm0 = std::move(best.perm);
        pairs = std::move(best.prs);
        pair_rounds = best.rounds;
    } else {
        std::iota(perm0.begin(), perm0.end(), 0);
    }
    // Schur complement mode (src/numeric/schur.jl:1-10): the flagged rows and columns go last, in
    // increasing original order, after the others in the ordering's order
    i64 nschur = 0;
    if (schur) {
        V rest, last;
        for (i64 k = 0; k < n; ++k)
            if (!schur[perm0[k]]) rest.push_back(perm0[k]);
        for (i64 i = 0; i < n; ++i)
            if (schur[i]) last.push_back(i);
        nschur = (i64)last.size();
        if (nschur == 0) throw std::runtime_error("schur_mode = 1 needs user_schur_indices with at least one flag");
        rest.insert(rest.end(), last.begin(), last.end());
        perm0 = rest;
    }
    A->nschur = nschur;
    V iperm0(n);
    for (i64 k = 0; k < n; ++k) iperm0[perm0[k]] = k;
    A->pair_rounds = pair_rounds;
    {
        V pv;
        for (auto& [a, b] : pairs) {
            pv.push_back(a + 1);
            pv.push_back(b + 1);
        }
        A->f["pairs"] = pv;
    }
    // the symbolic factorization runs on the pair pattern (the pattern itself without pairs)
    if (!pairs.empty()) P = pair_pattern(P, pairs);

    V parent = etree(P, perm0, iperm0);
    V post = postorder(parent);
    V counts = colcounts(P, perm0, iperm0, parent, post);
    if (nschur > 0) {
        // the Schur columns (etree numbering n - nschur .. n - 1) after every other column, in that
        // numbering, with the counts of a dense lower triangle: children before parents still holds
        // (a parent's etree number is larger than its child's)
        V p2;
        p2.reserve(n);
        for (i64 k = 0; k < n; ++k)
            if (post[k] < n - nschur) p2.push_back(post[k]);
        for (i64 j = n - nschur; j < n; ++j) {
            p2.push_back(j);
            counts[j] = n - j;
        }
        post = p2;
    }
    ColumnPartition cp = fundamental_supernodes(parent, post, counts, nschur);
    if (use_superpanels) cp = amalgamate(cp, parent, counts, max_width, zero_fraction, min_width, nschur > 0);

    // supernodal symbolic factorization
    V fperm(n), fiperm(n);
    for (i64 k = 0; k < n; ++k) fperm[k] = perm0[cp.order[k]];
    for (i64 k = 0; k < n; ++k) fiperm[fperm[k]] = k;
    V par = relabel_parent(parent, cp.order);
    V cnt(n);
    for (i64 k = 0; k < n; ++k) cnt[k] = counts[cp.order[k]];
    i64 ns = (i64)cp.super_ptr.size() - 1;
    A->ns = ns;
    const V& sptr = cp.super_ptr;
    V col2sn = col2sn_of(sptr);
    const V& snparent = cp.snparent;
    std::vector<V> children(ns);
    for (i64 s = 0; s < ns; ++s)
        if (snparent[s] >= 0) children[snparent[s]].push_back(s);
    V snrowptr(ns + 1, 0), snrowval;
    V mark(n, -1), below;
    i64 nnz_stored = 0;
    double flops = 0;
    for (i64 s = 0; s < ns; ++s) {
        i64 first = sptr[s], last = sptr[s + 1] - 1;
        below.clear();
        for (i64 j = first; j <= last; ++j) {
            const i64* nb = P.nb(fperm[j]);
            for (i64 t = 0; t < P.deg(fperm[j]); ++t) {
                i64 i = fiperm[nb[t]];
                if (i > last && mark[i] != s) {
                    mark[i] = s;
                    below.push_back(i);
                }
            }
        }
        for (i64 c : children[s])
            for (i64 q = snrowptr[c]; q < snrowptr[c + 1]; ++q) {
                i64 i = snrowval[q];
                if (i > last && mark[i] != s) {
                    mark[i] = s;
                    below.push_back(i);
                }
            }
        std::sort(below.begin(), below.end());
        for (i64 j = first; j <= last; ++j) snrowval.push_back(j);
        for (i64 i : below) snrowval.push_back(i);
        snrowptr[s + 1] = (i64)snrowval.size();
        i64 w = last - first + 1, f = w + (i64)below.size();
        nnz_stored += trapezoid(w, f);
        for (i64 k = 0; k < w; ++k) flops += (double)(f - k) * (double)(f - k);
    }
    V snpost = postorder(snparent);
    A->nnz_stored = nnz_stored;
    A->flops = flops;
    i64 nnzl = 0;
    double ft = 0;
    for (i64 k = 0; k < n; ++k) {
        nnzl += cnt[k];
        ft += (double)cnt[k] * (double)cnt[k];
    }
    A->nnz_L = nnzl;
    A->flops_true = ft;

    // layout: panel_ptr (1-based element offsets)
    V frows(ns), fcols(ns), panel_ptr(ns + 1);
    panel_ptr[0] = 1;
    for (i64 s = 0; s < ns; ++s) {
        fcols[s] = sptr[s + 1] - sptr[s];
        frows[s] = snrowptr[s + 1] - snrowptr[s];
        panel_ptr[s + 1] = panel_ptr[s] + frows[s] * fcols[s];
    }

    // relind: positions (1-based) of a child's rows below its width among its parent's rows
    V relind_ptr(ns + 1), relind;
    relind_ptr[0] = 1;
    for (i64 c = 0; c < ns; ++c) {
        i64 p = snparent[c];
        if (p >= 0) {
            i64 k = snrowptr[p], kend = snrowptr[p + 1];
            for (i64 q = snrowptr[c] + fcols[c]; q < snrowptr[c + 1]; ++q) {
                i64 i = snrowval[q];
                while (k < kend && snrowval[k] < i) ++k;
                if (k >= kend || snrowval[k] != i) throw std::runtime_error("child row missing from its parent");
                relind.push_back(k - snrowptr[p] + 1);
            }
        }
        relind_ptr[c + 1] = (i64)relind.size() + 1;
    }

    // amap: factor offset of every stored entry; negative = conjugate / U^T structure
    bool neg = structure == ST_H || structure == ST_HPD || structure == ST_G;
    V amap(nnz, 0);
    for (i64 r = 0; r < n; ++r)
        for (i64 p = rp[r]; p < rp[r + 1]; ++p) {
            i64 c = cv[p];
            if (!keep_entry(structure, view, r, c)) continue;
            i64 i = fiperm[r], j = fiperm[c];
            bool flip = i < j;
            if (flip) std::swap(i, j);
            i64 s = col2sn[j];
            auto b = snrowval.begin() + snrowptr[s], e = snrowval.begin() + snrowptr[s + 1];
            auto it = std::lower_bound(b, e, i);
            if (it == e || *it != i) throw std::runtime_error("entry outside the symbolic factor");
            i64 pos = it - b;
            i64 off = panel_ptr[s] + (j - sptr[s]) * frows[s] + pos;
            amap[p] = (flip && neg) ? -off : off;
        }
    // owner-pull grouping: sorted by destination (L before U^T for "G"), then position
    bool signed_ = structure == ST_G;
    V order;
    order.reserve(nnz);
    for (i64 p = 0; p < nnz; ++p)
        if (amap[p] != 0) order.push_back(p);
    std::stable_sort(order.begin(), order.end(), [&](i64 a, i64 b) {
        i64 oa = std::llabs(amap[a]), ob = std::llabs(amap[b]);
        if (oa != ob) return oa < ob;
        bool na = signed_ && amap[a] < 0, nb_ = signed_ && amap[b] < 0;
        if (na != nb_) return !na;
        return a < b;
    });
    V amap_ptr(ns + 1, 0);
    for (i64 p : order) {
        i64 off = std::llabs(amap[p]);
        i64 owner = std::upper_bound(panel_ptr.begin(), panel_ptr.end(), off) - panel_ptr.begin() - 1;
        amap_ptr[owner + 1]++;
    }
    amap_ptr[0] = 1;
    for (i64 s = 0; s < ns; ++s) amap_ptr[s + 1] += amap_ptr[s];

    // children in increasing id
    V child_ptr(ns + 1, 0), child_list;
    for (i64 s = 0; s < ns; ++s)
        if (snparent[s] >= 0) child_ptr[snparent[s] + 1]++;
    child_ptr[0] = 1;
    for (i64 s = 0; s < ns; ++s) child_ptr[s + 1] += child_ptr[s];
    child_list.resize(child_ptr[ns] - 1);
    {
        V nxt(child_ptr.begin(), child_ptr.end() - 1);
        for (i64 s = 0; s < ns; ++s)
            if (snparent[s] >= 0) child_list[(nxt[snparent[s]]++) - 1] = s + 1;
    }

    // full-pattern residual map of the stored matrix: row-major, (row, col, src) order,
    // src negative for a conjugated (mirrored Hermitian) contribution; duplicates kept.
    // For "G" the rows of the transpose follow as rows n+1..2n.
    struct Contrib { i64 r, c, p; bool cj; };
    std::vector<Contrib> contrib;
    bool herm = structure == ST_H || structure == ST_HPD;
    for (i64 r = 0; r < n; ++r)
        for (i64 p = rp[r]; p < rp[r + 1]; ++p) {
            i64 c = cv[p];
            if (!keep_entry(structure, view, r, c)) continue;
            contrib.push_back({r, c, p, false});
            if (structure != ST_G && r != c) contrib.push_back({c, r, p, herm});
        }
    std::sort(contrib.begin(), contrib.end(), [](const Contrib& a, const Contrib& b) {
        if (a.r != b.r) return a.r < b.r;
        if (a.c != b.c) return a.c < b.c;
        return a.p < b.p;
    });
    i64 nrows_map = structure == ST_G ? 2 * n : n;
    V full_rowptr(nrows_map + 1, 0), full_colval, full_src;
    for (auto& t : contrib) {
        full_rowptr[t.r + 1]++;
        full_colval.push_back(t.c + 1);
        full_src.push_back(t.cj ? -(t.p + 1) : (t.p + 1));
    }
    if (structure == ST_G) {
        // rows of M^T: the entries (i, j) of M by column j, rows i in order
        std::vector<Contrib> tr(contrib);
        std::stable_sort(tr.begin(), tr.end(), [](const Contrib& a, const Contrib& b) { return a.c < b.c; });
        for (auto& t : tr) {
            full_rowptr[n + t.c + 1]++;
            full_colval.push_back(t.r + 1);
            full_src.push_back(t.cj ? -(t.p + 1) : (t.p + 1));
        }
    }
    full_rowptr[0] = 1;
    for (i64 i = 0; i < nrows_map; ++i) full_rowptr[i + 1] += full_rowptr[i];

    // 1-based outputs
    auto one = [](const V& v) {
        V o(v.size());
        for (size_t k = 0; k < v.size(); ++k) o[k] = v[k] + 1;
        return o;
    };
    auto& F = A->f;
    F["perm"] = one(fperm);
    F["iperm"] = one(fiperm);
    F["parent"] = V(par.size());
    for (size_t k = 0; k < par.size(); ++k) F["parent"][k] = par[k] + 1;  // 0 = root
    F["counts"] = cnt;
    F["super_ptr"] = one(sptr);
    F["col2sn"] = one(col2sn);
    F["snparent"] = V(ns);
    for (i64 s = 0; s < ns; ++s) F["snparent"][s] = snparent[s] + 1;
    F["snpost"] = one(snpost);
    F["rowptr"] = one(snrowptr);
    F["rowval"] = one(snrowval);
    F["front_ptr"] = panel_ptr;
    F["front_nrows"] = frows;
    F["front_ncols"] = fcols;
    F["child_ptr"] = child_ptr;
    F["child_list"] = child_list;
    F["relind_ptr"] = relind_ptr;
    F["relind"] = relind;
    F["amap"] = amap;
    F["amap_ptr"] = amap_ptr;
    F["amap_src"] = one(order);
    F["full_rowptr"] = full_rowptr;
    F["full_colval"] = full_colval;
    F["full_src"] = full_src;
    F["etree_perm"] = one(perm0);
    return A;
}

}  // namespace

extern "C" {

// Run the analysis. ordering: 0 natural, 1 user permutation (perm must be
// given), 3 approximate minimum degree, 4 nested dissection, 6 the automatic choice between 3 and
// 4 by the reference's cost model. Returns NULL on error
// (sdsa_last_error() holds the message).
// schur: NULL, or n flags (original rows) of the Schur complement mode, whose rows and columns go
// last and form the last supernode.
// values (nzval of the first member, may be NULL), pair_mode (0 none, 1 "default", 2 "all"),
// pair_tol (pivot_pair_tolerance) and pivot_u (pivot_threshold, the partners' acceptance):
// 2x2 pivot pairs of "S"/"H" with the AMD ordering; fixed_pairs (nfixed (partner, candidate)
// pairs, 1-based, e.g. those of a matching) replace the search. The pairs used are the field
// "pairs" (partner, candidate, ... 1-based).
void* sdsa_analyze(int64_t n, const int64_t* rowptr, const int64_t* colval, int structure, int view,
                   int index_base, const int64_t* perm, int ordering, int64_t max_width, double zero_fraction,
                   int64_t min_width, int use_superpanels, const int8_t* schur, const double* values, int pair_mode,
                   double pair_tol, double pivot_u, const int64_t* fixed_pairs, int64_t nfixed) {
    try {
        return analyze(n, rowptr, colval, structure, view, index_base, perm, ordering, max_width, zero_fraction,
                       min_width, use_superpanels, schur, values, pair_mode, pair_tol, pivot_u, fixed_pairs, nfixed);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

const char* sdsa_last_error(void) { return g_err.c_str(); }

void sdsa_free(void* h) { delete static_cast<Analysis*>(h); }

// length of a named array, -1 if unknown
int64_t sdsa_field_len(void* h, const char* name) {
    auto A = static_cast<Analysis*>(h);
    auto it = A->f.find(name);
    return it == A->f.end() ? -1 : (int64_t)it->second.size();
}

int sdsa_field_copy(void* h, const char* name, int64_t* dst) {
    auto A = static_cast<Analysis*>(h);
    auto it = A->f.find(name);
    if (it == A->f.end()) return -1;
    std::memcpy(dst, it->second.data(), it->second.size() * sizeof(int64_t));
    return 0;
}

// scalars: 0 n, 1 nnz, 2 ns, 3 nnz_stored, 4 nnz_L, 5 stored-panel flops, 6 true flops (sum c_j^2),
// 7 Schur complement size (0 = no Schur mode)
double sdsa_scalar(void* h, int which) {
    auto A = static_cast<Analysis*>(h);
    switch (which) {
        case 0: return (double)A->n;
        case 1: return (double)A->nnz;
        case 2: return (double)A->ns;
        case 3: return (double)A->nnz_stored;
        case 4: return (double)A->nnz_L;
        case 5: return A->flops;
        case 6: return A->flops_true;
        case 7: return (double)A->nschur;
        case 8: return (double)A->pair_rounds;
    }
    return -1;
}

// MC64-style matching (src/matching/mc64.jl:370-392) of the CSR matrix (index_base 0 or 1) with
// job 1..5: perm (1-based, row i matched to column perm[i], completed to a permutation), matched
// (n flags), rscale / cscale (n; equal for the symmetric structures). 0 on success, -1 on error.
int sdsa_matching(int64_t n, const int64_t* rowptr, const int64_t* colval, const double* nzval, int structure,
                  int view, int index_base, int job, int64_t* perm, int8_t* matched, double* rscale, double* cscale) {
    try {
        V rp(n + 1), cv;
        for (i64 i = 0; i <= n; ++i) rp[i] = rowptr[i] - index_base;
        cv.resize(rp[n]);
        for (i64 p = 0; p < rp[n]; ++p) {
            cv[p] = colval[p] - index_base;
            if (cv[p] < 0 || cv[p] >= n) throw std::runtime_error("column index out of range");
        }
        V pm;
        std::vector<char> mt;
        std::vector<double> rs, cs;
        compute_matching(n, rp, cv, nzval, structure, view, job, pm, mt, rs, cs);
        for (i64 i = 0; i < n; ++i) {
            perm[i] = pm[i] + 1;
            matched[i] = mt[i];
            rscale[i] = rs[i];
            cscale[i] = cs[i];
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// matching_pairs (src/matching/mc64.jl:413-461): the 2x2 pivot pairs (partner, candidate; 1-based,
// at most n entries into pairs) from the cycles of a symmetric matching (perm 1-based, matched,
// d the symmetric scaling); mode 0 none, 1 "default", 2 "all". Returns the number of pairs, -1 on error.
int64_t sdsa_matching_pairs(int64_t n, const int64_t* rowptr, const int64_t* colval, const double* nzval,
                            int structure, int view, int index_base, const int64_t* perm, const int8_t* matched,
                            const double* d, int mode, double tol, int64_t* pairs) {
    try {
        V rp(n + 1), cv, pm(n);
        for (i64 i = 0; i <= n; ++i) rp[i] = rowptr[i] - index_base;
        cv.resize(rp[n]);
        for (i64 p = 0; p < rp[n]; ++p) cv[p] = colval[p] - index_base;
        for (i64 i = 0; i < n; ++i) pm[i] = perm[i] - 1;
        std::vector<char> mt(matched, matched + n);
        std::vector<double> dv(d, d + n);
        const Pairs pr = matching_pairs(n, rp, cv, nzval, structure, view, pm, mt, dv, mode, tol);
        for (size_t t = 0; t < pr.size(); ++t) {
            pairs[2 * t] = pr[t].first + 1;
            pairs[2 * t + 1] = pr[t].second + 1;
        }
        return (int64_t)pr.size();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// geometric nested dissection of an nx x ny x nz grid; perm (1-based) of length nx ny nz
int sdsa_grid_nd(int64_t nx, int64_t ny, int64_t nz, int64_t leaf, int64_t* perm) {
    V out;
    out.reserve(nx * ny * nz);
    nd_box(nx, ny, 0, nx, 0, ny, 0, nz, leaf < 1 ? 1 : leaf, out);
    if ((int64_t)out.size() != nx * ny * nz) return -1;
    for (size_t k = 0; k < out.size(); ++k) perm[k] = out[k] + 1;
    return 0;
}

}  // extern "C"
