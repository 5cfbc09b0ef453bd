// symbolic.cpp -- host symbolic analysis (see symbolic.h).
#include "symbolic.h"

#include <algorithm>
#include <chrono>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <utility>

namespace sle {
namespace {

// Angle references: in every island of closed branches the SLACK-typed bus of
// highest index (the reference keeps the last SLACK bus added,
// network_model.cu:85-87).  States: the angles of all other buses in bus order,
// then all magnitudes (sparse_matrix.cu:1677-1685 for a single island).
std::vector<int32_t> island_reference(const ModelView& mv, std::vector<int32_t>* roots) {
    std::vector<int32_t> up(mv.n);
    std::iota(up.begin(), up.end(), 0);
    auto find = [&](int a) {
        while (up[a] != a) a = up[a] = up[up[a]];
        return a;
    };
    for (int k = 0; k < mv.nb; ++k)
        if (mv.status[k]) {
            int a = find(mv.from[k]), b = find(mv.to[k]);
            if (a != b) up[std::max(a, b)] = std::min(a, b);
        }
    for (int b = 0; b < mv.n; ++b) up[b] = find(b);  // every bus points at its island's smallest bus
    std::vector<int32_t> ref(mv.n, -1);
    for (int b = 0; b < mv.n; ++b)
        if (mv.bus_type[b] == 2) ref[up[b]] = b;
    std::vector<int32_t> out(mv.n);
    for (int b = 0; b < mv.n; ++b) out[b] = ref[up[b]];
    if (roots) *roots = std::move(up);
    return out;  // reference bus of each bus's island (-1: none)
}

void angle_references(const ModelView& mv, Symbolic& S, std::vector<int32_t>* roots = nullptr) {
    const std::vector<int32_t> ref = island_reference(mv, roots);
    S.th_state.assign(mv.n, -1);
    S.nth = 0;
    S.n_refs = 0;
    for (int b = 0; b < mv.n; ++b) {
        if (ref[b] != b)
            S.th_state[b] = S.nth++;
        else
            S.n_refs++;
    }
    S.ns = S.nth + mv.n;
}

// Ybus pattern and per-entry contribution lists.  Entry (i,i) collects the
// from-end (ff) or to-end (tt) self admittance of each closed incident branch,
// off-diagonal (i,j) the ft / tf transfer admittances; contributions keep
// branch order.  Self-loops only touch the diagonal.
void build_ybus(const ModelView& mv, Symbolic& S) {
    const int n = mv.n;
    std::vector<int32_t> cnt(n + 1, 0);
    for (int i = 0; i < n; ++i) cnt[i + 1]++;  // diagonal marker
    for (int k = 0; k < mv.nb; ++k) {
        if (mv.present && !mv.present[k]) continue;
        int i = mv.from[k], j = mv.to[k];
        if (i == j) {
            cnt[i + 1] += 2;
        } else {
            cnt[i + 1] += 2;
            cnt[j + 1] += 2;
        }
    }
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    // (col, contribution) per row; contribution -1 marks the bare diagonal
    std::vector<std::pair<int32_t, int32_t>> trip(cnt[n]);
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; ++i) trip[fill[i]++] = {i, -1};
    for (int k = 0; k < mv.nb; ++k) {
        if (mv.present && !mv.present[k]) continue;
        int i = mv.from[k], j = mv.to[k];
        if (i == j) {
            trip[fill[i]++] = {i, k * 4 + 0};
            trip[fill[i]++] = {i, k * 4 + 1};
        } else {
            trip[fill[i]++] = {i, k * 4 + 0};
            trip[fill[i]++] = {j, k * 4 + 2};
            trip[fill[j]++] = {j, k * 4 + 1};
            trip[fill[j]++] = {i, k * 4 + 3};
        }
    }
    S.yrow.assign(n + 1, 0);
    S.ycol.clear();
    S.ycptr.assign(1, 0);
    S.ycon.clear();
    S.ydiag.assign(n, -1);
    for (int i = 0; i < n; ++i) {
        auto b = trip.begin() + cnt[i], e = trip.begin() + cnt[i + 1];
        std::stable_sort(b, e, [](const auto& a, const auto& c) { return a.first < c.first; });
        for (auto it = b; it != e; ++it) {
            if (S.ycol.size() == (size_t)S.yrow[i] || S.ycol.back() != it->first) {
                S.ycol.push_back(it->first);
                S.ycptr.push_back(S.ycptr.back());
                if (it->first == i) S.ydiag[i] = (int32_t)S.ycol.size() - 1;
            }
            if (it->second >= 0) {
                S.ycon.push_back(it->second);
                S.ycptr.back()++;
            }
        }
        S.yrow[i + 1] = (int32_t)S.ycol.size();
    }
}

// Exact H row structure in natural state order, sorted.
void build_h(const ModelView& mv, Symbolic& S) {
    const std::vector<int32_t>& ts = S.th_state;
    auto theta_state = [&](int b) { return ts[b]; };
    auto v_state = [&](int b) { return S.nth + b; };
    auto has_th = [&](int b) { return ts[b] >= 0; };
    S.hrow.assign(mv.m + 1, 0);
    auto row_len = [&](int i) -> int {
        if (mv.m_present && !mv.m_present[i]) return 0;
        int t = mv.mtype[i], l = mv.mloc[i];
        if (t == MT_V_MAG) return 1;
        if (t == MT_V_ANGLE) return has_th(l) ? 1 : 0;
        if (is_bus_measurement(t)) {
            int len = 0;
            for (int e = S.yrow[l]; e < S.yrow[l + 1]; ++e) len += has_th(S.ycol[e]) ? 2 : 1;
            return len;
        }
        int a = mv.from[l], b = mv.to[l];
        if (a == b) return has_th(a) ? 2 : 1;
        return 2 + (has_th(a) ? 1 : 0) + (has_th(b) ? 1 : 0);
    };
    for (int i = 0; i < mv.m; ++i) S.hrow[i + 1] = S.hrow[i] + row_len(i);
    S.hcol.resize(S.hrow[mv.m]);
    for (int i = 0; i < mv.m; ++i) {
        int q = S.hrow[i];
        if (q == S.hrow[i + 1]) continue;
        int t = mv.mtype[i], l = mv.mloc[i];
        if (t == MT_V_MAG) {
            S.hcol[q] = v_state(l);
        } else if (t == MT_V_ANGLE) {
            S.hcol[q] = theta_state(l);
        } else if (is_bus_measurement(t)) {
            for (int e = S.yrow[l]; e < S.yrow[l + 1]; ++e)
                if (has_th(S.ycol[e])) S.hcol[q++] = theta_state(S.ycol[e]);
            for (int e = S.yrow[l]; e < S.yrow[l + 1]; ++e) S.hcol[q++] = v_state(S.ycol[e]);
        } else {
            int a = std::min(mv.from[l], mv.to[l]), b = std::max(mv.from[l], mv.to[l]);
            if (has_th(a)) S.hcol[q++] = theta_state(a);
            if (b != a && has_th(b)) S.hcol[q++] = theta_state(b);
            S.hcol[q++] = v_state(a);
            if (b != a) S.hcol[q++] = v_state(b);
        }
    }
}

// Bus coupling graph of G: two buses are adjacent when some active measurement
// row holds a state of each.
void bus_graph(const ModelView& mv, const Symbolic& S, std::vector<int32_t>& aptr, std::vector<int32_t>& adj) {
    const int n = mv.n;
    // bus sets of rows
    auto bus_set = [&](int i, std::vector<int>& out) {
        out.clear();
        if (S.hrow[i] == S.hrow[i + 1]) return;
        int t = mv.mtype[i], l = mv.mloc[i];
        if (t == MT_V_MAG || t == MT_V_ANGLE) {
            out.push_back(l);
        } else if (is_bus_measurement(t)) {
            for (int e = S.yrow[l]; e < S.yrow[l + 1]; ++e) out.push_back(S.ycol[e]);
        } else {
            out.push_back(mv.from[l]);
            if (mv.to[l] != mv.from[l]) out.push_back(mv.to[l]);
        }
    };
    // rows incident to each bus
    std::vector<int32_t> rptr(n + 1, 0);
    std::vector<int> bs;
    for (int i = 0; i < mv.m; ++i) {
        bus_set(i, bs);
        if (bs.size() > 1)
            for (int b : bs) rptr[b + 1]++;
    }
    for (int b = 0; b < n; ++b) rptr[b + 1] += rptr[b];
    std::vector<int32_t> rows(rptr[n]);
    std::vector<int32_t> f(rptr.begin(), rptr.end() - 1);
    for (int i = 0; i < mv.m; ++i) {
        bus_set(i, bs);
        if (bs.size() > 1)
            for (int b : bs) rows[f[b]++] = i;
    }
    std::vector<int32_t> mark(n, -1);
    aptr.assign(n + 1, 0);
    adj.clear();
    std::vector<int32_t> tmp;
    for (int b = 0; b < n; ++b) {
        tmp.clear();
        mark[b] = b;
        for (int q = rptr[b]; q < rptr[b + 1]; ++q) {
            bus_set(rows[q], bs);
            for (int u : bs)
                if (mark[u] != b) {
                    mark[u] = b;
                    tmp.push_back(u);
                }
        }
        std::sort(tmp.begin(), tmp.end());
        adj.insert(adj.end(), tmp.begin(), tmp.end());
        aptr[b + 1] = (int32_t)adj.size();
    }
}

// Exact minimum degree on the explicit elimination graph.  The bus of
// smallest current degree is eliminated next, ties broken by the lowest bus
// index; its neighbourhood becomes a clique.  Lazy-deletion binary heap.
void minimum_degree(int n, const std::vector<int32_t>& aptr, const std::vector<int32_t>& adj0,
                    std::vector<int32_t>& order) {
    std::vector<std::vector<int32_t>> adj(n);
    for (int b = 0; b < n; ++b) adj[b].assign(adj0.begin() + aptr[b], adj0.begin() + aptr[b + 1]);
    using Item = std::pair<int32_t, int32_t>;  // (degree, bus)
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    for (int b = 0; b < n; ++b) pq.push({(int32_t)adj[b].size(), b});
    std::vector<uint8_t> done(n, 0);
    order.clear();
    order.reserve(n);
    std::vector<int32_t> merged;
    while (!pq.empty()) {
        Item it = pq.top();
        pq.pop();
        int v = it.second;
        if (done[v] || it.first != (int32_t)adj[v].size()) continue;
        done[v] = 1;
        order.push_back(v);
        const std::vector<int32_t> nb = std::move(adj[v]);
        adj[v].clear();
        for (int u : nb) {
            // adj[u] := (adj[u] U nb) \ {u, v}
            const std::vector<int32_t>& au = adj[u];
            merged.clear();
            merged.reserve(au.size() + nb.size());
            size_t p = 0, q = 0;
            while (p < au.size() || q < nb.size()) {
                int x;
                if (q == nb.size() || (p < au.size() && au[p] < nb[q])) {
                    x = au[p++];
                } else if (p == au.size() || nb[q] < au[p]) {
                    x = nb[q++];
                } else {
                    x = au[p++];
                    q++;
                }
                if (x != u && x != v) merged.push_back(x);
            }
            adj[u].swap(merged);
            pq.push({(int32_t)adj[u].size(), u});
        }
    }
}

}  // namespace

std::vector<int32_t> angle_states(const ModelView& mv, std::vector<int32_t>* islands) {
    Symbolic S;
    angle_references(mv, S, islands);
    return S.th_state;
}

void analyze(const ModelView& mv, Symbolic& S) {
    auto t0 = std::chrono::steady_clock::now();
    if (mv.n < 1 || mv.slack < 0 || mv.slack >= mv.n) throw std::runtime_error("model has no valid slack bus");
    S = Symbolic();
    S.n = mv.n;
    S.nb = mv.nb;
    S.m = mv.m;
    S.slack = mv.slack;
    angle_references(mv, S);
    const int n = mv.n, ns = S.ns;

    build_ybus(mv, S);
    build_h(mv, S);

    // ---- ordering -------------------------------------------------------
    std::vector<int32_t> aptr, adj;
    bus_graph(mv, S, aptr, adj);
    minimum_degree(n, aptr, adj, S.bus_perm);
    S.perm.clear();
    S.perm.reserve(ns);
    for (int b : S.bus_perm) {
        if (S.th_state[b] >= 0) S.perm.push_back(S.th_state[b]);
        S.perm.push_back(S.nth + b);
    }
    S.iperm.assign(ns, 0);
    for (int k = 0; k < ns; ++k) S.iperm[S.perm[k]] = k;
    S.bus_pth.assign(n, -1);
    S.bus_pv.assign(n, -1);
    for (int b = 0; b < n; ++b) {
        if (S.th_state[b] >= 0) S.bus_pth[b] = S.iperm[S.th_state[b]];
        S.bus_pv[b] = S.iperm[S.nth + b];
    }

    // ---- H^T in permuted state order -------------------------------------
    const int64_t nnzh = (int64_t)S.hcol.size();
    S.tptr.assign(ns + 1, 0);
    for (int64_t q = 0; q < nnzh; ++q) S.tptr[S.iperm[S.hcol[q]] + 1]++;
    for (int k = 0; k < ns; ++k) S.tptr[k + 1] += S.tptr[k];
    S.tpos.resize(nnzh);
    S.trow.resize(nnzh);
    {
        std::vector<int32_t> f(S.tptr.begin(), S.tptr.end() - 1);
        for (int i = 0; i < mv.m; ++i)
            for (int q = S.hrow[i]; q < S.hrow[i + 1]; ++q) {
                int k = S.iperm[S.hcol[q]];
                S.tpos[f[k]] = q;
                S.trow[f[k]] = i;
                f[k]++;
            }
    }

    // ---- symbolic SpGEMM: lower triangle of P G P^T with product lists --------
    {
        std::vector<int32_t> mark(ns, -1), local(ns, 0), rows, cnt, fillp;
        S.g_row.clear();
        S.g_col.clear();
        S.gpair_ptr.assign(1, 0);
        S.gpair.clear();
        for (int j = 0; j < ns; ++j) {
            rows.clear();
            for (int p = S.tptr[j]; p < S.tptr[j + 1]; ++p) {
                int i = S.trow[p];
                for (int q = S.hrow[i]; q < S.hrow[i + 1]; ++q) {
                    int r = S.iperm[S.hcol[q]];
                    if (r < j) continue;
                    if (mark[r] != j) {
                        mark[r] = j;
                        rows.push_back(r);
                    }
                }
            }
            std::sort(rows.begin(), rows.end());
            cnt.assign(rows.size(), 0);
            for (size_t e = 0; e < rows.size(); ++e) local[rows[e]] = (int32_t)e;
            for (int p = S.tptr[j]; p < S.tptr[j + 1]; ++p) {
                int i = S.trow[p];
                for (int q = S.hrow[i]; q < S.hrow[i + 1]; ++q) {
                    int r = S.iperm[S.hcol[q]];
                    if (r >= j) cnt[local[r]]++;
                }
            }
            int64_t base = S.gpair.size() / 2;
            fillp.assign(rows.size(), 0);
            int64_t acc = base;
            for (size_t e = 0; e < rows.size(); ++e) {
                fillp[e] = (int32_t)acc;
                acc += cnt[e];
                S.g_row.push_back(rows[e]);
                S.g_col.push_back(j);
                if (acc > INT32_MAX) throw std::runtime_error("gain product list exceeds 2^31 terms");
                S.gpair_ptr.push_back((int32_t)acc);
            }
            S.gpair.resize(2 * acc);
            for (int p = S.tptr[j]; p < S.tptr[j + 1]; ++p) {
                int i = S.trow[p];
                int qj = S.tpos[p];
                for (int q = S.hrow[i]; q < S.hrow[i + 1]; ++q) {
                    int r = S.iperm[S.hcol[q]];
                    if (r < j) continue;
                    int32_t at = fillp[local[r]]++;
                    S.gpair[2 * (int64_t)at] = qj;
                    S.gpair[2 * (int64_t)at + 1] = q;
                }
            }
        }
        S.nnz_g = (int64_t)S.g_row.size();
    }

    // ---- elimination tree (Liu) ----------------------------------------------
    std::vector<int32_t> rp(ns + 1, 0), rc;  // row-wise strictly lower pattern
    for (int64_t e = 0; e < S.nnz_g; ++e)
        if (S.g_row[e] > S.g_col[e]) rp[S.g_row[e] + 1]++;
    for (int k = 0; k < ns; ++k) rp[k + 1] += rp[k];
    rc.resize(rp[ns]);
    {
        std::vector<int32_t> f(rp.begin(), rp.end() - 1);
        for (int64_t e = 0; e < S.nnz_g; ++e)
            if (S.g_row[e] > S.g_col[e]) rc[f[S.g_row[e]]++] = S.g_col[e];
    }
    S.parent.assign(ns, -1);
    {
        std::vector<int32_t> anc(ns, -1);
        for (int i = 0; i < ns; ++i)
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                int r = rc[p];
                while (anc[r] != -1 && anc[r] != i) {
                    int t = anc[r];
                    anc[r] = i;
                    r = t;
                }
                if (anc[r] == -1) {
                    anc[r] = i;
                    S.parent[r] = i;
                }
            }
    }

    // ---- column structures of L ----------------------------------------------
    std::vector<int32_t> cptr(ns + 1, 0), cch;  // children lists of the column etree
    for (int j = 0; j < ns; ++j)
        if (S.parent[j] >= 0) cptr[S.parent[j] + 1]++;
    for (int j = 0; j < ns; ++j) cptr[j + 1] += cptr[j];
    cch.resize(cptr[ns]);
    {
        std::vector<int32_t> f(cptr.begin(), cptr.end() - 1);
        for (int j = 0; j < ns; ++j)
            if (S.parent[j] >= 0) cch[f[S.parent[j]]++] = j;
    }
    std::vector<int64_t> lptr(ns + 1, 0);
    std::vector<int32_t> lrow;  // strictly-below-diagonal rows per column
    {
        std::vector<int32_t> mark(ns, -1), tmp;
        // G column j: entries with g_col == j are contiguous (column order)
        std::vector<int64_t> gcp(ns + 1, 0);
        for (int64_t e = 0; e < S.nnz_g; ++e) gcp[S.g_col[e] + 1]++;
        for (int k = 0; k < ns; ++k) gcp[k + 1] += gcp[k];
        for (int j = 0; j < ns; ++j) {
            tmp.clear();
            mark[j] = j;
            for (int64_t e = gcp[j]; e < gcp[j + 1]; ++e) {
                int r = S.g_row[e];
                if (mark[r] != j) {
                    mark[r] = j;
                    tmp.push_back(r);
                }
            }
            for (int p = cptr[j]; p < cptr[j + 1]; ++p) {
                int ch = cch[p];
                for (int64_t q = lptr[ch]; q < lptr[ch + 1]; ++q) {
                    int r = lrow[q];
                    if (r > j && mark[r] != j) {
                        mark[r] = j;
                        tmp.push_back(r);
                    }
                }
            }
            std::sort(tmp.begin(), tmp.end());
            lrow.insert(lrow.end(), tmp.begin(), tmp.end());
            lptr[j + 1] = (int64_t)lrow.size();
        }
    }
    S.colcount.resize(ns);
    S.nnz_l = 0;
    S.flops = 0;
    for (int j = 0; j < ns; ++j) {
        S.colcount[j] = 1 + (int32_t)(lptr[j + 1] - lptr[j]);
        S.nnz_l += S.colcount[j];
        S.flops += (int64_t)S.colcount[j] * S.colcount[j];
    }

    // ---- fundamental supernodes ------------------------------------------------
    std::vector<int32_t> snode_of(ns, 0);
    S.sn_first.clear();
    for (int j = 0; j < ns; ++j) {
        bool merge = j > 0 && S.parent[j - 1] == j && (cptr[j + 1] - cptr[j]) == 1 &&
                     S.colcount[j - 1] == S.colcount[j] + 1;
        if (!merge) S.sn_first.push_back(j);
        snode_of[j] = (int32_t)S.sn_first.size() - 1;
    }
    S.nsn = (int)S.sn_first.size();
    const int nsn = S.nsn;
    S.sn_k.resize(nsn);
    S.sn_c.resize(nsn);
    S.sn_rows_off.assign(nsn + 1, 0);
    S.sn_rows.clear();
    S.max_c = 0;
    for (int s = 0; s < nsn; ++s) {
        int f = S.sn_first[s];
        int last = (s + 1 < nsn) ? S.sn_first[s + 1] : ns;
        S.sn_k[s] = last - f;
        S.sn_c[s] = S.colcount[f];
        S.sn_rows.push_back(f);
        S.sn_rows.insert(S.sn_rows.end(), lrow.begin() + lptr[f], lrow.begin() + lptr[f + 1]);
        S.sn_rows_off[s + 1] = (int64_t)S.sn_rows.size();
        S.max_c = std::max(S.max_c, S.sn_c[s]);
        for (int q = 0; q < S.sn_k[s]; ++q)
            if (S.sn_rows[S.sn_rows_off[s] + q] != f + q) throw std::runtime_error("supernode rows inconsistent");
    }
    S.sn_parent.assign(nsn, -1);
    for (int s = 0; s < nsn; ++s) {
        int last = S.sn_first[s] + S.sn_k[s] - 1;
        if (S.parent[last] >= 0) S.sn_parent[s] = snode_of[S.parent[last]];
    }
    S.sn_child_ptr.assign(nsn + 1, 0);
    for (int s = 0; s < nsn; ++s)
        if (S.sn_parent[s] >= 0) S.sn_child_ptr[S.sn_parent[s] + 1]++;
    for (int s = 0; s < nsn; ++s) S.sn_child_ptr[s + 1] += S.sn_child_ptr[s];
    S.sn_child.resize(S.sn_child_ptr[nsn]);
    {
        std::vector<int32_t> f(S.sn_child_ptr.begin(), S.sn_child_ptr.end() - 1);
        for (int s = 0; s < nsn; ++s)
            if (S.sn_parent[s] >= 0) S.sn_child[f[S.sn_parent[s]]++] = s;
    }
    // levels: height above the leaves (children precede parents in column order)
    S.sn_level.assign(nsn, 0);
    for (int s = 0; s < nsn; ++s)
        if (S.sn_parent[s] >= 0)
            S.sn_level[S.sn_parent[s]] = std::max(S.sn_level[S.sn_parent[s]], S.sn_level[s] + 1);
    S.n_levels = 0;
    for (int s = 0; s < nsn; ++s) S.n_levels = std::max(S.n_levels, S.sn_level[s] + 1);
    S.level_ptr.assign(S.n_levels + 1, 0);
    for (int s = 0; s < nsn; ++s) S.level_ptr[S.sn_level[s] + 1]++;
    for (int l = 0; l < S.n_levels; ++l) S.level_ptr[l + 1] += S.level_ptr[l];
    S.level_sn.resize(nsn);
    {
        std::vector<int32_t> f(S.level_ptr.begin(), S.level_ptr.end() - 1);
        for (int s = 0; s < nsn; ++s) S.level_sn[f[S.sn_level[s]]++] = s;
    }
    // front / update-vector / relative-map offsets
    S.sn_foff.assign(nsn + 1, 0);
    S.sn_moff.assign(nsn + 1, 0);
    S.sn_uoff.assign(nsn + 1, 0);
    S.sn_rmoff.assign(nsn + 1, 0);
    for (int s = 0; s < nsn; ++s) {
        int64_t c = S.sn_c[s], u = S.sn_c[s] - S.sn_k[s];
        S.sn_foff[s + 1] = S.sn_foff[s] + c * c;
        S.sn_moff[s + 1] = S.sn_moff[s] + c * S.sn_k[s];
        S.sn_uoff[s + 1] = S.sn_uoff[s] + u;
        S.sn_rmoff[s + 1] = S.sn_rmoff[s] + u;
    }
    S.ftotal = S.sn_foff[nsn];
    S.mtotal = S.sn_moff[nsn];
    S.utotal = S.sn_uoff[nsn];
    S.rmap.assign(S.sn_rmoff[nsn], 0);
    for (int s = 0; s < nsn; ++s) {
        int p = S.sn_parent[s];
        int u = S.sn_c[s] - S.sn_k[s];
        if (u == 0) continue;
        if (p < 0) throw std::runtime_error("root supernode with off-diagonal rows");
        const int32_t* pr = S.sn_rows.data() + S.sn_rows_off[p];
        const int32_t* pe = pr + S.sn_c[p];
        for (int a = 0; a < u; ++a) {
            int r = S.sn_rows[S.sn_rows_off[s] + S.sn_k[s] + a];
            const int32_t* it = std::lower_bound(pr, pe, r);
            if (it == pe || *it != r) throw std::runtime_error("child front row missing in parent");
            S.rmap[S.sn_rmoff[s] + a] = (int32_t)(it - pr);
        }
    }
    // G entries -> front positions
    S.g_sn_ptr.assign(nsn + 1, 0);
    S.gloc.resize(S.nnz_g);
    S.gdiag.assign(ns, -1);
    for (int64_t e = 0; e < S.nnz_g; ++e) {
        int j = S.g_col[e], i = S.g_row[e];
        int s = snode_of[j];
        S.g_sn_ptr[s + 1]++;
        const int32_t* pr = S.sn_rows.data() + S.sn_rows_off[s];
        const int32_t* pe = pr + S.sn_c[s];
        const int32_t* it = std::lower_bound(pr, pe, i);
        if (it == pe || *it != i) throw std::runtime_error("gain entry outside the factor pattern");
        S.gloc[e] = (j - S.sn_first[s]) * S.sn_c[s] + (int32_t)(it - pr);
        if (i == j) S.gdiag[j] = (int32_t)e;
    }
    for (int s = 0; s < nsn; ++s) S.g_sn_ptr[s + 1] += S.g_sn_ptr[s];
    S.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void natural_gain_pattern(const Symbolic& S, std::vector<int32_t>& colptr, std::vector<int32_t>& row,
                          std::vector<int64_t>* entry_of) {
    std::vector<std::pair<int64_t, int64_t>> key(S.nnz_g);
    for (int64_t e = 0; e < S.nnz_g; ++e) {
        int a = S.perm[S.g_row[e]], b = S.perm[S.g_col[e]];
        int r = std::max(a, b), c = std::min(a, b);
        key[e] = {(int64_t)c * S.ns + r, e};
    }
    std::sort(key.begin(), key.end());
    colptr.assign(S.ns + 1, 0);
    row.resize(S.nnz_g);
    if (entry_of) entry_of->resize(S.nnz_g);
    for (int64_t q = 0; q < S.nnz_g; ++q) {
        int c = (int)(key[q].first / S.ns);
        row[q] = (int32_t)(key[q].first % S.ns);
        colptr[c + 1]++;
        if (entry_of) (*entry_of)[q] = key[q].second;
    }
    for (int k = 0; k < S.ns; ++k) colptr[k + 1] += colptr[k];
}

}  // namespace sle
