// wls_oracle.cpp -- FP64 host restatement of the reference's WLS Gauss-Newton
// state estimator.  TEST INFRASTRUCTURE ONLY: it is the checker the HIP engine
// is compared against (tests/, __graft_entry__.smoke() and the cpu_baseline
// leg of bench.py).  Nothing in the shipped library links, loads or calls it.
//
// What it restates (reference file:line; the reference's numerical defects are
// corrected to the textbook pi-model, Abur & Exposito 2004 / MATPOWER):
//   admittances        network_model.cu:595-612 (FP64, eps clamp 1e-12 as in
//                      sle_types.cuh:62)
//   weights            network_model.cu:142-143  w = 1/max(sigma,eps)^2
//   flat start         network_model.cu:614-623, kernels.cu:799-817
//   Ybus (pi-model)    sparse_matrix.cu:410-492 (sorted columns, parallel
//                      branches merged, closed branches only)
//   injections         kernels.cu:129-230
//   branch flows       kernels.cu:241-333 (Q_from shunt /a^2, phase shift with
//                      the same sign in Ybus and flows)
//   h(x) by type       kernels.cu:381-489 (pseudo = injection; I = |S|/V)
//   residual, J        kernels.cu:495-549 (J at the final state)
//   Huber weights      kernels.cu:634-671
//   state layout       sparse_matrix.cu:1677-1685, kernels.cu:721-750
//   convergence        kernels.cu:752-782, wls_solver.cu:493-517
//   GN loop            wls_solver.cu:221-431 (modes, gain caching, retries)
//   Jacobian           sparse_matrix.cu:1627-2147, exact derivatives
//   ordering/etree/
//   supernodes         cholesky_optimized.cu:281-387, 978-1024 (design
//                      target; here: exact minimum degree on the bus graph of
//                      G, ties to the lowest bus index, Liu etree, fundamental
//                      supernodes)
//   factor + solve     sparse_matrix.cu:793-895 (cusolverSp csrlsvchol is
//                      replaced by a left-looking sparse Cholesky)
//
//   bad data (LNR)     sle_engine.cu:959-977 thresholds |r|/sigma; here the
//                      normalized residual |r|/sqrt(Omega_ii) of Abur &
//                      Exposito ch. 5, Omega from the sparse inverse subset
//
// Pins (tests/test_oracle_*.py): closed-form 2-bus cases, injection = sum of
// branch flows, central finite differences of h(x) for H, dense H^T W H for G,
// brute-force symbolic elimination for etree/column counts/supernodes, dense
// Cholesky, noise-free recovery, first-order optimality and the
// scipy.optimize.least_squares golden vectors under tests/golden/.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <set>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kEps = 1e-12;
constexpr double kNewtonSwitch = 1e-2;  // Huber: step size below which the outliers' gain weights drop
enum { V_MAG = 0, V_ANGLE = 1, P_INJ = 2, Q_INJ = 3, P_FLOW = 4, Q_FLOW = 5, I_MAG = 6, P_PSEUDO = 7, Q_PSEUDO = 8 };

struct Case {
    int n = 0, nb = 0, m = 0, slack = 0, ns = 0;
    int nth = 0;           // angle states: every bus except the angle reference of its island
    std::vector<int> thi;  // angle state of each bus (-1 for a reference)
    std::vector<int> bus_type;
    std::vector<double> vset;
    std::vector<int> from, to;
    std::vector<double> r, x, bc, tap, shift;
    std::vector<uint8_t> status;
    std::vector<int> mtype, mloc, mend;
    std::vector<double> sigma, z;
    std::vector<uint8_t> active;

    // admittances
    std::vector<double> gs, bs;
    // Ybus
    std::vector<int> yrow, ycol;
    std::vector<double> yG, yB;
    // H pattern (natural state order)
    std::vector<int> hrow, hcol;
    // G lower pattern, CSC by column (natural order): rows >= col
    std::vector<int> gcolptr, grow;
    // ordering
    std::vector<int> bus_perm;   // elimination position -> bus
    std::vector<int> perm;       // new state index -> natural state index
    std::vector<int> iperm;      // natural -> new
    std::vector<int> parent;     // etree on permuted states
    std::vector<int> colcount;   // nnz of L column incl. diagonal
    std::vector<int> snode_ptr;  // fundamental supernodes
    // L structure (permuted), CSC incl. diagonal first
    std::vector<int64_t> lptr;
    std::vector<int> lrow;
    std::vector<double> lval;
    bool factor_valid = false;

    // state
    std::vector<double> V, th;
    // last results at the final state
    std::vector<double> h, res, P, Q, Pf, Qf, Pt, Qt;
    double J = 0;
};

inline bool is_bus_meas(int t) { return t == V_MAG || t == V_ANGLE || t == P_INJ || t == Q_INJ || t == P_PSEUDO || t == Q_PSEUDO; }
inline int th_idx(const Case& c, int bus) { return c.thi[bus]; }
inline int v_idx(const Case& c, int bus) { return c.nth + bus; }
inline bool has_th(const Case& c, int bus) { return c.thi[bus] >= 0; }

// Angle references: in every island of closed branches the SLACK-typed bus of
// highest index (the reference keeps the last SLACK bus added,
// network_model.cu:85-87; with one island this is that bus).  An island without
// one keeps all its angles as states (its gain matrix is then singular, as in
// the reference).  State layout: angles of the other buses in bus order, then
// all magnitudes (sparse_matrix.cu:1677-1685 for one island).
void compute_references(Case& c) {
    std::vector<int> up(c.n);
    std::iota(up.begin(), up.end(), 0);
    auto find = [&](int a) {
        while (up[a] != a) a = up[a] = up[up[a]];
        return a;
    };
    for (int k = 0; k < c.nb; ++k)
        if (c.status[k]) {
            int a = find(c.from[k]), b = find(c.to[k]);
            if (a != b) up[std::max(a, b)] = std::min(a, b);
        }
    std::vector<int> ref(c.n, -1);
    for (int b = 0; b < c.n; ++b)
        if (c.bus_type[b] == 2) ref[find(b)] = b;  // ascending: the last one wins
    c.thi.assign(c.n, -1);
    c.nth = 0;
    for (int b = 0; b < c.n; ++b)
        if (ref[find(b)] != b) c.thi[b] = c.nth++;
    c.ns = c.nth + c.n;
}

void compute_admittances(Case& c) {
    c.gs.resize(c.nb);
    c.bs.resize(c.nb);
    for (int k = 0; k < c.nb; ++k) {
        double d = c.r[k] * c.r[k] + c.x[k] * c.x[k];
        if (d < kEps) d = kEps;
        c.gs[k] = c.r[k] / d;
        c.bs[k] = -c.x[k] / d;
    }
}

void recompute_active(Case& c) {
    for (int i = 0; i < c.m; ++i) {
        bool a = true;
        if (!is_bus_meas(c.mtype[i])) a = c.status[c.mloc[i]] != 0;
        c.active[i] = a;
    }
}

// Ybus: rows sorted by column, parallel branches merged.  The pattern holds
// every branch, open or closed; an open branch contributes zero admittance, so
// a breaker operation changes values only.
void build_ybus(Case& c) {
    std::vector<std::vector<std::pair<int, std::pair<double, double>>>> rows(c.n);
    for (int i = 0; i < c.n; ++i) rows[i].push_back({i, {0.0, 0.0}});
    for (int k = 0; k < c.nb; ++k) {
        const double on = c.status[k] ? 1.0 : 0.0;
        int i = c.from[k], j = c.to[k];
        double g = on * c.gs[k], b = on * c.bs[k], a = c.tap[k], ph = c.shift[k];
        double hb = on * 0.5 * c.bc[k];
        double a2 = a * a;
        double cp = std::cos(ph), sp = std::sin(ph);
        // Yff = (g + j(b+hb))/a^2 ; Ytt = g + j(b+hb)
        // Yft = -(g+jb) e^{+j ph}/a ; Ytf = -(g+jb) e^{-j ph}/a
        double yff_r = g / a2, yff_i = (b + hb) / a2;
        double ytt_r = g, ytt_i = b + hb;
        double yft_r = -(g * cp - b * sp) / a, yft_i = -(g * sp + b * cp) / a;
        double ytf_r = -(g * cp + b * sp) / a, ytf_i = -(b * cp - g * sp) / a;
        rows[i][0].second.first += yff_r;
        rows[i][0].second.second += yff_i;
        rows[j][0].second.first += ytt_r;
        rows[j][0].second.second += ytt_i;
        if (i == j) continue;
        rows[i].push_back({j, {yft_r, yft_i}});
        rows[j].push_back({i, {ytf_r, ytf_i}});
    }
    c.yrow.assign(c.n + 1, 0);
    c.ycol.clear();
    c.yG.clear();
    c.yB.clear();
    for (int i = 0; i < c.n; ++i) {
        auto& R = rows[i];
        // stable sort keeps branch order for merging parallel branches
        std::stable_sort(R.begin(), R.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t q = 0; q < R.size(); ++q) {
            if (!c.ycol.empty() && (int)c.ycol.size() > c.yrow[i] && c.ycol.back() == R[q].first) {
                c.yG.back() += R[q].second.first;
                c.yB.back() += R[q].second.second;
            } else {
                c.ycol.push_back(R[q].first);
                c.yG.push_back(R[q].second.first);
                c.yB.push_back(R[q].second.second);
            }
        }
        c.yrow[i + 1] = (int)c.ycol.size();
    }
}

// Injections and flows at state (V, th).
void eval_power(const Case& c, const double* V, const double* th, double* P, double* Q, double* Pf, double* Qf,
                double* Pt, double* Qt) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c.n; ++i) {
        double p = 0, q = 0;
        for (int e = c.yrow[i]; e < c.yrow[i + 1]; ++e) {
            int j = c.ycol[e];
            double t = th[i] - th[j];
            double ct = std::cos(t), st = std::sin(t);
            p += V[j] * (c.yG[e] * ct + c.yB[e] * st);
            q += V[j] * (c.yG[e] * st - c.yB[e] * ct);
        }
        P[i] = V[i] * p;
        Q[i] = V[i] * q;
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < c.nb; ++k) {
        if (!c.status[k]) {
            Pf[k] = Qf[k] = Pt[k] = Qt[k] = 0;
            continue;
        }
        int i = c.from[k], j = c.to[k];
        double g = c.gs[k], b = c.bs[k], a = c.tap[k], bsh = b + 0.5 * c.bc[k];
        double t = th[i] - th[j] - c.shift[k];
        double ct = std::cos(t), st = std::sin(t);
        double A = V[i] * V[j] / a;
        Pf[k] = V[i] * V[i] * g / (a * a) - A * (g * ct + b * st);
        Qf[k] = -V[i] * V[i] * bsh / (a * a) - A * (g * st - b * ct);
        Pt[k] = V[j] * V[j] * g - A * (g * ct - b * st);
        Qt[k] = -V[j] * V[j] * bsh + A * (g * st + b * ct);
    }
}

void eval_h(const Case& c, const double* V, const double* th, const double* P, const double* Q, const double* Pf,
            const double* Qf, const double* Pt, const double* Qt, double* h) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c.m; ++i) {
        int l = c.mloc[i];
        double v = 0;
        switch (c.mtype[i]) {
            case V_MAG: v = V[l]; break;
            case V_ANGLE: v = th[l]; break;
            case P_INJ:
            case P_PSEUDO: v = P[l]; break;
            case Q_INJ:
            case Q_PSEUDO: v = Q[l]; break;
            case P_FLOW: v = c.mend[i] ? Pt[l] : Pf[l]; break;
            case Q_FLOW: v = c.mend[i] ? Qt[l] : Qf[l]; break;
            case I_MAG: {
                double p = c.mend[i] ? Pt[l] : Pf[l], q = c.mend[i] ? Qt[l] : Qf[l];
                double ve = c.mend[i] ? V[c.to[l]] : V[c.from[l]];
                v = std::sqrt(p * p + q * q) / ve;
                break;
            }
        }
        h[i] = c.active[i] ? v : 0.0;
    }
}

// Structural H row (natural state order, sorted); a measurement on an open
// branch keeps its row (its values and weight are zero).
void h_row_pattern(const Case& c, int i, std::vector<int>& cols) {
    cols.clear();
    int t = c.mtype[i], l = c.mloc[i];
    if (t == V_MAG) {
        cols.push_back(v_idx(c, l));
    } else if (t == V_ANGLE) {
        if (has_th(c, l)) cols.push_back(th_idx(c, l));
    } else if (is_bus_meas(t)) {
        for (int e = c.yrow[l]; e < c.yrow[l + 1]; ++e) {
            int j = c.ycol[e];
            if (has_th(c, j)) cols.push_back(th_idx(c, j));
        }
        for (int e = c.yrow[l]; e < c.yrow[l + 1]; ++e) cols.push_back(v_idx(c, c.ycol[e]));
    } else {
        int a = std::min(c.from[l], c.to[l]), b = std::max(c.from[l], c.to[l]);
        if (has_th(c, a)) cols.push_back(th_idx(c, a));
        if (has_th(c, b) && b != a) cols.push_back(th_idx(c, b));
        cols.push_back(v_idx(c, a));
        if (b != a) cols.push_back(v_idx(c, b));
    }
}

void build_h_pattern(Case& c) {
    c.hrow.assign(c.m + 1, 0);
    c.hcol.clear();
    std::vector<int> cols;
    for (int i = 0; i < c.m; ++i) {
        h_row_pattern(c, i, cols);
        c.hcol.insert(c.hcol.end(), cols.begin(), cols.end());
        c.hrow[i + 1] = (int)c.hcol.size();
    }
}

// Jacobian values into the H pattern (exact derivatives, App. B formulas).
void eval_jacobian(const Case& c, const double* V, const double* th, const double* P, const double* Q,
                   const double* Pf, const double* Qf, const double* Pt, const double* Qt, double* hv) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c.m; ++i) {
        int r0 = c.hrow[i], r1 = c.hrow[i + 1];
        if (r0 == r1) continue;
        if (!c.active[i]) {  // inactive rows are zero
            for (int q = r0; q < r1; ++q) hv[q] = 0;
            continue;
        }
        int t = c.mtype[i], l = c.mloc[i];
        auto put = [&](int col, double v) {
            for (int q = r0; q < r1; ++q)
                if (c.hcol[q] == col) {
                    hv[q] += v;
                    return;
                }
        };
        for (int q = r0; q < r1; ++q) hv[q] = 0;
        if (t == V_MAG) {
            put(v_idx(c, l), 1.0);
        } else if (t == V_ANGLE) {
            if (has_th(c, l)) put(th_idx(c, l), 1.0);
        } else if (is_bus_meas(t)) {
            bool isP = (t == P_INJ || t == P_PSEUDO);
            int ii = l;
            double Gii = 0, Bii = 0;
            for (int e = c.yrow[ii]; e < c.yrow[ii + 1]; ++e) {
                int j = c.ycol[e];
                if (j == ii) {
                    Gii = c.yG[e];
                    Bii = c.yB[e];
                    continue;
                }
                double tt = th[ii] - th[j], ct = std::cos(tt), st = std::sin(tt);
                double G = c.yG[e], B = c.yB[e];
                double dth, dv;
                if (isP) {
                    dth = V[ii] * V[j] * (G * st - B * ct);
                    dv = V[ii] * (G * ct + B * st);
                } else {
                    dth = -V[ii] * V[j] * (G * ct + B * st);
                    dv = V[ii] * (G * st - B * ct);
                }
                if (has_th(c, j)) put(th_idx(c, j), dth);
                put(v_idx(c, j), dv);
            }
            double Vi = V[ii];
            double dthi, dvi;
            if (isP) {
                dthi = -Q[ii] - Vi * Vi * Bii;
                dvi = P[ii] / Vi + Vi * Gii;
            } else {
                dthi = P[ii] - Vi * Vi * Gii;
                dvi = Q[ii] / Vi - Vi * Bii;
            }
            if (has_th(c, ii)) put(th_idx(c, ii), dthi);
            put(v_idx(c, ii), dvi);
        } else {
            int fi = c.from[l], tj = c.to[l];
            double g = c.gs[l], b = c.bs[l], a = c.tap[l], bsh = b + 0.5 * c.bc[l];
            double tt = th[fi] - th[tj] - c.shift[l], ct = std::cos(tt), st = std::sin(tt);
            double Vi = V[fi], Vj = V[tj], A = Vi * Vj / a;
            // partials wrt (th_i, th_j, V_i, V_j) for P and Q at the measured end
            double dP[4], dQ[4];
            if (c.mend[i] == 0) {
                dP[0] = A * (g * st - b * ct);
                dP[1] = -dP[0];
                dP[2] = 2 * Vi * g / (a * a) - (Vj / a) * (g * ct + b * st);
                dP[3] = -(Vi / a) * (g * ct + b * st);
                dQ[0] = -A * (g * ct + b * st);
                dQ[1] = -dQ[0];
                dQ[2] = -2 * Vi * bsh / (a * a) - (Vj / a) * (g * st - b * ct);
                dQ[3] = -(Vi / a) * (g * st - b * ct);
            } else {
                dP[0] = A * (g * st + b * ct);
                dP[1] = -dP[0];
                dP[2] = -(Vj / a) * (g * ct - b * st);
                dP[3] = 2 * Vj * g - (Vi / a) * (g * ct - b * st);
                dQ[0] = A * (g * ct - b * st);
                dQ[1] = -dQ[0];
                dQ[2] = (Vj / a) * (g * st + b * ct);
                dQ[3] = -2 * Vj * bsh + (Vi / a) * (g * st + b * ct);
            }
            double d[4];
            if (t == P_FLOW) {
                for (int q = 0; q < 4; ++q) d[q] = dP[q];
            } else if (t == Q_FLOW) {
                for (int q = 0; q < 4; ++q) d[q] = dQ[q];
            } else {  // I_MAG = S / V_end
                double p = c.mend[i] ? Pt[l] : Pf[l], q = c.mend[i] ? Qt[l] : Qf[l];
                double ve = c.mend[i] ? Vj : Vi;
                double S = std::sqrt(p * p + q * q);
                double inv = S > kEps ? 1.0 / (S * ve) : 0.0;
                for (int k = 0; k < 4; ++k) d[k] = (p * dP[k] + q * dQ[k]) * inv;
                d[c.mend[i] ? 3 : 2] -= S / (ve * ve);
            }
            if (has_th(c, fi)) put(th_idx(c, fi), d[0]);
            if (has_th(c, tj)) put(th_idx(c, tj), d[1]);
            put(v_idx(c, fi), d[2]);
            put(v_idx(c, tj), d[3]);
        }
    }
}

// Transpose of the H pattern: for each state column, the (row, position) list.
void h_transpose(const Case& c, std::vector<int>& tptr, std::vector<int>& trow, std::vector<int>& tpos) {
    tptr.assign(c.ns + 1, 0);
    for (int q : c.hcol) tptr[q + 1]++;
    for (int j = 0; j < c.ns; ++j) tptr[j + 1] += tptr[j];
    trow.resize(c.hcol.size());
    tpos.resize(c.hcol.size());
    std::vector<int> fill(tptr.begin(), tptr.end() - 1);
    for (int i = 0; i < c.m; ++i)
        for (int q = c.hrow[i]; q < c.hrow[i + 1]; ++q) {
            int j = c.hcol[q];
            trow[fill[j]] = i;
            tpos[fill[j]] = q;
            fill[j]++;
        }
}

// G = H^T W H, lower triangle (row >= col), CSC in natural order.  With
// hv == nullptr only the pattern is built.
void gain(Case& c, const double* hv, const double* w, std::vector<double>* gval) {
    std::vector<int> tptr, trow, tpos;
    h_transpose(c, tptr, trow, tpos);
    std::vector<std::vector<int>> rows(c.ns);
    std::vector<std::vector<double>> vals(c.ns);
#pragma omp parallel
    {
        std::vector<double> acc(c.ns, 0.0);
        std::vector<int> mark(c.ns, -1), touched;
#pragma omp for schedule(dynamic, 512)
        for (int j = 0; j < c.ns; ++j) {
            touched.clear();
            for (int k = tptr[j]; k < tptr[j + 1]; ++k) {
                int i = trow[k];
                double hij = hv ? hv[tpos[k]] * w[i] : 0.0;
                for (int q = c.hrow[i]; q < c.hrow[i + 1]; ++q) {
                    int rr = c.hcol[q];
                    if (rr < j) continue;
                    if (mark[rr] != j) {
                        mark[rr] = j;
                        acc[rr] = 0;
                        touched.push_back(rr);
                    }
                    if (hv) acc[rr] += hij * hv[q];
                }
            }
            std::sort(touched.begin(), touched.end());
            rows[j] = touched;
            if (hv) {
                vals[j].resize(touched.size());
                for (size_t q = 0; q < touched.size(); ++q) vals[j][q] = acc[touched[q]];
            }
        }
    }
    c.gcolptr.assign(c.ns + 1, 0);
    for (int j = 0; j < c.ns; ++j) c.gcolptr[j + 1] = c.gcolptr[j] + (int)rows[j].size();
    c.grow.resize(c.gcolptr[c.ns]);
    if (gval) gval->resize(c.gcolptr[c.ns]);
    for (int j = 0; j < c.ns; ++j) {
        std::copy(rows[j].begin(), rows[j].end(), c.grow.begin() + c.gcolptr[j]);
        if (gval) std::copy(vals[j].begin(), vals[j].end(), gval->begin() + c.gcolptr[j]);
    }
}

// Exact minimum degree on the bus graph of G: repeatedly eliminate the bus of
// smallest current degree (ties: lowest bus index), turning its neighbourhood
// into a clique.  Explicit elimination graph with std::set adjacency.
void order_buses(Case& c) {
    std::vector<std::set<int>> adj(c.n);
    std::vector<int> sbus(c.ns);
    for (int b = 0; b < c.n; ++b) {
        if (has_th(c, b)) sbus[th_idx(c, b)] = b;
        sbus[v_idx(c, b)] = b;
    }
    for (int j = 0; j < c.ns; ++j)
        for (int q = c.gcolptr[j]; q < c.gcolptr[j + 1]; ++q) {
            int a = sbus[c.grow[q]], b = sbus[j];
            if (a != b) {
                adj[a].insert(b);
                adj[b].insert(a);
            }
        }
    std::set<std::pair<int, int>> pq;
    for (int b = 0; b < c.n; ++b) pq.insert({(int)adj[b].size(), b});
    c.bus_perm.clear();
    while (!pq.empty()) {
        int v = pq.begin()->second;
        pq.erase(pq.begin());
        c.bus_perm.push_back(v);
        std::vector<int> nb(adj[v].begin(), adj[v].end());
        for (int u : nb) {
            pq.erase({(int)adj[u].size(), u});
            adj[u].erase(v);
            for (int w : nb)
                if (w != u) adj[u].insert(w);
            pq.insert({(int)adj[u].size(), u});
        }
        adj[v].clear();
    }
}

// State permutation, etree, L structure, fundamental supernodes.
void symbolic(Case& c) {
    c.perm.clear();
    for (int b : c.bus_perm) {
        if (has_th(c, b)) c.perm.push_back(th_idx(c, b));
        c.perm.push_back(v_idx(c, b));
    }
    c.iperm.assign(c.ns, 0);
    for (int k = 0; k < c.ns; ++k) c.iperm[c.perm[k]] = k;
    // permuted lower pattern: column j' holds rows i' > j'
    std::vector<std::vector<int>> acol(c.ns);
    for (int j = 0; j < c.ns; ++j)
        for (int q = c.gcolptr[j]; q < c.gcolptr[j + 1]; ++q) {
            int a = c.iperm[c.grow[q]], b = c.iperm[j];
            if (a == b) continue;
            acol[std::min(a, b)].push_back(std::max(a, b));
        }
    // symbolic elimination: struct(L_j) = A_j U children's structs, minus j
    std::vector<std::vector<int>> lst(c.ns);
    std::vector<std::vector<int>> children(c.ns);
    c.parent.assign(c.ns, -1);
    std::vector<int> mark(c.ns, -1);
    for (int j = 0; j < c.ns; ++j) {
        std::vector<int>& s = lst[j];
        mark[j] = j;
        for (int i : acol[j])
            if (mark[i] != j) {
                mark[i] = j;
                s.push_back(i);
            }
        for (int ch : children[j])
            for (int i : lst[ch])
                if (i > j && mark[i] != j) {
                    mark[i] = j;
                    s.push_back(i);
                }
        std::sort(s.begin(), s.end());
        if (!s.empty()) {
            c.parent[j] = s[0];
            children[s[0]].push_back(j);
        }
    }
    c.colcount.resize(c.ns);
    c.lptr.assign(c.ns + 1, 0);
    for (int j = 0; j < c.ns; ++j) {
        c.colcount[j] = 1 + (int)lst[j].size();
        c.lptr[j + 1] = c.lptr[j] + c.colcount[j];
    }
    c.lrow.resize(c.lptr[c.ns]);
    for (int j = 0; j < c.ns; ++j) {
        c.lrow[c.lptr[j]] = j;
        std::copy(lst[j].begin(), lst[j].end(), c.lrow.begin() + c.lptr[j] + 1);
    }
    // fundamental supernodes
    std::vector<int> nchild(c.ns, 0);
    for (int j = 0; j < c.ns; ++j)
        if (c.parent[j] >= 0) nchild[c.parent[j]]++;
    c.snode_ptr.assign(1, 0);
    for (int j = 1; j < c.ns; ++j) {
        bool merge = c.parent[j - 1] == j && nchild[j] == 1 && c.colcount[j - 1] == c.colcount[j] + 1;
        if (!merge) c.snode_ptr.push_back(j);
    }
    c.snode_ptr.push_back(c.ns);
    c.factor_valid = false;
}

// Left-looking numeric Cholesky of P G P^T on the symbolic structure.
// Returns false on a non-positive pivot.
bool factor(Case& c, const std::vector<double>& gval) {
    int ns = c.ns;
    c.lval.assign(c.lrow.size(), 0.0);
    // scatter G into L positions
    std::vector<std::vector<std::pair<int, double>>> acol(ns);
    for (int j = 0; j < ns; ++j)
        for (int q = c.gcolptr[j]; q < c.gcolptr[j + 1]; ++q) {
            int a = c.iperm[c.grow[q]], b = c.iperm[j];
            acol[std::min(a, b)].push_back({std::max(a, b), gval[q]});
        }
    std::vector<double> xw(ns, 0.0);
    std::vector<int64_t> next(ns);  // next unused row position per column
    std::vector<int> link(ns, -1), head(ns, -1);
    // head[j]: list of columns k whose next row is j
    for (int j = 0; j < ns; ++j) {
        for (int64_t p = c.lptr[j]; p < c.lptr[j + 1]; ++p) xw[c.lrow[p]] = 0.0;
        for (auto& e : acol[j]) xw[e.first] += e.second;
        int k = head[j];
        while (k >= 0) {
            int nk = link[k];
            double ljk = c.lval[next[k]];
            for (int64_t p = next[k]; p < c.lptr[k + 1]; ++p) xw[c.lrow[p]] -= c.lval[p] * ljk;
            next[k]++;
            if (next[k] < c.lptr[k + 1]) {
                int r = c.lrow[next[k]];
                link[k] = head[r];
                head[r] = k;
            }
            k = nk;
        }
        double d = xw[j];
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        d = std::sqrt(d);
        c.lval[c.lptr[j]] = d;
        for (int64_t p = c.lptr[j] + 1; p < c.lptr[j + 1]; ++p) c.lval[p] = xw[c.lrow[p]] / d;
        next[j] = c.lptr[j] + 1;
        if (next[j] < c.lptr[j + 1]) {
            int r = c.lrow[next[j]];
            link[j] = head[r];
            head[r] = j;
        }
    }
    return true;
}

// Solve (P G P^T) y = P b, return dx = P^T y in natural order.
void solve_factor(const Case& c, const std::vector<double>& b, std::vector<double>& dx) {
    int ns = c.ns;
    std::vector<double> y(ns);
    for (int k = 0; k < ns; ++k) y[k] = b[c.perm[k]];
    for (int j = 0; j < ns; ++j) {
        y[j] /= c.lval[c.lptr[j]];
        double yj = y[j];
        for (int64_t p = c.lptr[j] + 1; p < c.lptr[j + 1]; ++p) y[c.lrow[p]] -= c.lval[p] * yj;
    }
    for (int j = ns - 1; j >= 0; --j) {
        double s = y[j];
        for (int64_t p = c.lptr[j] + 1; p < c.lptr[j + 1]; ++p) s -= c.lval[p] * y[c.lrow[p]];
        y[j] = s / c.lval[c.lptr[j]];
    }
    dx.resize(ns);
    for (int k = 0; k < ns; ++k) dx[c.perm[k]] = y[k];
}

// Sparse inverse subset Z = (P G P^T)^-1 on the structure of L (entries of Z
// at every position of L's pattern), by the column recurrences of Takahashi,
// Fagan and Chin (1973) in the form of Erisman and Tinney (1975), from the last
// column to the first:
//   Z_ij = -(1/L_jj) sum_{k in struct(j)} L_kj Z_ik   (i in struct(j))
//   Z_jj = (1/L_jj) (1/L_jj - sum_{k in struct(j)} L_kj Z_kj)
// struct(j) is a clique of the filled graph, so every Z_ik needed lies in the
// pattern of an already finished column.
double z_at(const Case& c, const std::vector<double>& zval, int i, int j) {
    if (i < j) std::swap(i, j);
    const int* b = c.lrow.data() + c.lptr[j];
    const int* e = c.lrow.data() + c.lptr[j + 1];
    const int* p = std::lower_bound(b, e, i);
    if (p == e || *p != i) return std::nan("");  // outside the filled pattern: not in the subset
    return zval[p - c.lrow.data()];
}

void selected_inverse(const Case& c, std::vector<double>& zval) {
    zval.assign(c.lrow.size(), 0.0);
    for (int j = c.ns - 1; j >= 0; --j) {
        const int64_t d = c.lptr[j], e = c.lptr[j + 1];
        const double ljj = c.lval[d];
        for (int64_t q = d + 1; q < e; ++q) {
            const int i = c.lrow[q];
            double s = 0;
            for (int64_t t = d + 1; t < e; ++t) s += c.lval[t] * z_at(c, zval, i, c.lrow[t]);
            zval[q] = -s / ljj;
        }
        double s = 0;
        for (int64_t t = d + 1; t < e; ++t) s += c.lval[t] * zval[t];
        zval[d] = (1.0 / ljj - s) / ljj;
    }
}

void prepare(Case& c) {
    compute_references(c);
    compute_admittances(c);
    recompute_active(c);
    build_ybus(c);
    build_h_pattern(c);
    gain(c, nullptr, nullptr, nullptr);
    order_buses(c);
    symbolic(c);
}

}  // namespace

// ---------------------------------------------------------------------------
// C interface (ctypes from the tests and the bench cpu_baseline leg)
// ---------------------------------------------------------------------------
extern "C" {

struct OracleModel {
    int32_t n_bus, n_branch, n_meas, slack;
    const int32_t* bus_type;
    const double* vset;
    const int32_t* br_from;
    const int32_t* br_to;
    const double* br_r;
    const double* br_x;
    const double* br_b;
    const double* br_tap;
    const double* br_shift;
    const uint8_t* br_status;
    const int32_t* m_type;
    const int32_t* m_loc;
    const int32_t* m_end;
    const double* m_sigma;
    const double* m_z;
};

struct OracleSolveParams {
    int32_t mode;  // 0 REALTIME (reuse a valid factor), 1 PRECISION (refactor every iteration)
    int32_t max_iter;
    double tol;
    int32_t robust;
    double gamma;
    int32_t flat_start;
    double vmin, vmax;
};

struct OracleSolveResult {
    int32_t status;  // SLE_ConvergenceStatus values
    int32_t iterations;
    int32_t refactors;
    double max_mismatch;
    double objective;
    double seconds;
};

void* or_create(const OracleModel* md) {
    Case* c = new Case();
    c->n = md->n_bus;
    c->nb = md->n_branch;
    c->m = md->n_meas;
    c->slack = md->slack;
    c->bus_type.assign(md->bus_type, md->bus_type + c->n);
    c->vset.assign(md->vset, md->vset + c->n);
    c->from.assign(md->br_from, md->br_from + c->nb);
    c->to.assign(md->br_to, md->br_to + c->nb);
    c->r.assign(md->br_r, md->br_r + c->nb);
    c->x.assign(md->br_x, md->br_x + c->nb);
    c->bc.assign(md->br_b, md->br_b + c->nb);
    c->tap.assign(md->br_tap, md->br_tap + c->nb);
    c->shift.assign(md->br_shift, md->br_shift + c->nb);
    c->status.assign(md->br_status, md->br_status + c->nb);
    c->mtype.assign(md->m_type, md->m_type + c->m);
    c->mloc.assign(md->m_loc, md->m_loc + c->m);
    c->mend.assign(md->m_end, md->m_end + c->m);
    c->sigma.resize(c->m);
    for (int i = 0; i < c->m; ++i) c->sigma[i] = std::max(md->m_sigma[i], kEps);
    c->z.assign(md->m_z, md->m_z + c->m);
    c->active.assign(c->m, 1);
    prepare(*c);
    c->V.assign(c->n, 1.0);
    c->th.assign(c->n, 0.0);
    for (int i = 0; i < c->n; ++i)
        if (c->bus_type[i] != 0) c->V[i] = c->vset[i];
    return c;
}

void or_destroy(void* h) { delete (Case*)h; }

void or_set_z(void* h, const double* z) {
    Case* c = (Case*)h;
    std::copy(z, z + c->m, c->z.begin());
}

// Re-run the symbolic phase after a change of branch status.
void or_set_branch_status(void* h, int32_t k, int32_t st) {
    Case* c = (Case*)h;
    c->status[k] = (uint8_t)st;
    prepare(*c);
}

void or_set_state(void* h, const double* V, const double* th) {
    Case* c = (Case*)h;
    std::copy(V, V + c->n, c->V.begin());
    std::copy(th, th + c->n, c->th.begin());
}

void or_get_state(void* h, double* V, double* th) {
    Case* c = (Case*)h;
    std::copy(c->V.begin(), c->V.end(), V);
    std::copy(c->th.begin(), c->th.end(), th);
}

// sizes: [ns, nnzY, nnzH, nnzG_lower, nnzL, n_supernodes]
void or_sizes(void* h, int64_t* out) {
    Case* c = (Case*)h;
    out[0] = c->ns;
    out[1] = (int64_t)c->ycol.size();
    out[2] = (int64_t)c->hcol.size();
    out[3] = (int64_t)c->grow.size();
    out[4] = (int64_t)c->lrow.size();
    out[5] = (int64_t)c->snode_ptr.size() - 1;
}

void or_get_ybus(void* h, int32_t* rowptr, int32_t* col, double* G, double* B) {
    Case* c = (Case*)h;
    std::copy(c->yrow.begin(), c->yrow.end(), rowptr);
    std::copy(c->ycol.begin(), c->ycol.end(), col);
    std::copy(c->yG.begin(), c->yG.end(), G);
    std::copy(c->yB.begin(), c->yB.end(), B);
}

void or_get_h_pattern(void* h, int32_t* rowptr, int32_t* col) {
    Case* c = (Case*)h;
    std::copy(c->hrow.begin(), c->hrow.end(), rowptr);
    std::copy(c->hcol.begin(), c->hcol.end(), col);
}

void or_get_g_pattern(void* h, int32_t* colptr, int32_t* row) {
    Case* c = (Case*)h;
    std::copy(c->gcolptr.begin(), c->gcolptr.end(), colptr);
    std::copy(c->grow.begin(), c->grow.end(), row);
}

// bus_perm[n], perm[ns], parent[ns], colcount[ns], snode_ptr[nsn+1]
void or_get_symbolic(void* h, int32_t* bus_perm, int32_t* perm, int32_t* parent, int32_t* colcount,
                     int32_t* snode_ptr) {
    Case* c = (Case*)h;
    std::copy(c->bus_perm.begin(), c->bus_perm.end(), bus_perm);
    std::copy(c->perm.begin(), c->perm.end(), perm);
    std::copy(c->parent.begin(), c->parent.end(), parent);
    std::copy(c->colcount.begin(), c->colcount.end(), colcount);
    std::copy(c->snode_ptr.begin(), c->snode_ptr.end(), snode_ptr);
}

// Evaluate h(x), P/Q injections and branch flows at the given state.
void or_eval(void* h, const double* V, const double* th, double* hx, double* P, double* Q, double* Pf, double* Qf,
             double* Pt, double* Qt) {
    Case* c = (Case*)h;
    eval_power(*c, V, th, P, Q, Pf, Qf, Pt, Qt);
    eval_h(*c, V, th, P, Q, Pf, Qf, Pt, Qt, hx);
}

// Jacobian values in the H pattern at the given state.
void or_jacobian(void* h, const double* V, const double* th, double* hv) {
    Case* c = (Case*)h;
    std::vector<double> P(c->n), Q(c->n), Pf(c->nb), Qf(c->nb), Pt(c->nb), Qt(c->nb);
    eval_power(*c, V, th, P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data());
    eval_jacobian(*c, V, th, P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data(), hv);
}

// Gauss-Newton right-hand side b = H^T W r at the given state (WLS weights,
// natural state order), as in the GN loop (kernels.cu:823-862 semantics, sums
// per state over the transposed H pattern).
void or_rhs(void* h, const double* V, const double* th, double* b) {
    Case* c = (Case*)h;
    std::vector<double> P(c->n), Q(c->n), Pf(c->nb), Qf(c->nb), Pt(c->nb), Qt(c->nb), hx(c->m);
    std::vector<double> hv(c->hcol.size());
    eval_power(*c, V, th, P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data());
    eval_h(*c, V, th, P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data(), hx.data());
    eval_jacobian(*c, V, th, P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data(), hv.data());
    std::vector<int> tptr, trow, tpos;
    h_transpose(*c, tptr, trow, tpos);
    for (int j = 0; j < c->ns; ++j) {
        double s = 0;
        for (int k = tptr[j]; k < tptr[j + 1]; ++k) {
            const int i = trow[k];
            if (!c->active[i]) continue;
            s += hv[tpos[k]] * (c->z[i] - hx[i]) / (c->sigma[i] * c->sigma[i]);
        }
        b[j] = s;
    }
}

// G lower values (pattern order of or_get_g_pattern) for state and weights w.
void or_gain(void* h, const double* V, const double* th, const double* w, double* gv) {
    Case* c = (Case*)h;
    std::vector<double> hv(c->hcol.size());
    or_jacobian(h, V, th, hv.data());
    std::vector<double> g;
    gain(*c, hv.data(), w, &g);
    std::copy(g.begin(), g.end(), gv);
}

// Numeric factor of G (pattern order) -> L values in the symbolic structure.
// lptr[ns+1] (int64), lrow[nnzL], lval[nnzL].  Returns 0 on success.
int32_t or_factor(void* h, const double* gv, int64_t* lptr, int32_t* lrow, double* lval) {
    Case* c = (Case*)h;
    std::vector<double> g(gv, gv + c->grow.size());
    if (!factor(*c, g)) return -1;
    std::copy(c->lptr.begin(), c->lptr.end(), lptr);
    std::copy(c->lrow.begin(), c->lrow.end(), lrow);
    std::copy(c->lval.begin(), c->lval.end(), lval);
    c->factor_valid = true;
    return 0;
}

// The Gauss-Newton / Huber loop (wls_solver.cu:221-431 semantics).
void or_solve(void* hh, const OracleSolveParams* prm, OracleSolveResult* out) {
    Case& c = *(Case*)hh;
    auto t0 = std::chrono::steady_clock::now();
    OracleSolveResult R{};
    R.status = 5;
    std::vector<double> hx(c.m), P(c.n), Q(c.n), Pf(c.nb), Qf(c.nb), Pt(c.nb), Qt(c.nb), r(c.m), w(c.m), wg(c.m);
    std::vector<double> hv(c.hcol.size()), b(c.ns), dx, g;
    std::vector<int> tptr, trow, tpos;
    h_transpose(c, tptr, trow, tpos);
    auto flat = [&]() {
        for (int i = 0; i < c.n; ++i) {
            c.V[i] = c.bus_type[i] != 0 ? c.vset[i] : 1.0;
            c.th[i] = 0.0;
        }
    };
    if (prm->flat_start) flat();
    int retries = 0;
    bool retry = false;
    do {
        if (retry) {
            flat();
            c.factor_valid = false;
            retry = false;
        }
        bool converged = false, diverged = false, singular = false;
        R.iterations = 0;
        double last_step = 1e300;
        for (int it = 0; it < prm->max_iter; ++it) {
            // robust == 2: IRLS gain weights until the steps are small, then reduced ones
            const bool newton = prm->robust == 2 && last_step < kNewtonSwitch;
            eval_power(c, c.V.data(), c.th.data(), P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data());
            eval_h(c, c.V.data(), c.th.data(), P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data(),
                   hx.data());
            for (int i = 0; i < c.m; ++i) {
                r[i] = c.active[i] ? c.z[i] - hx[i] : 0.0;
                double s = c.sigma[i];
                double wi = 1.0 / (s * s);
                if (prm->robust) {
                    double u = std::fabs(r[i]) / s;
                    if (u > prm->gamma) wi = prm->gamma / (s * std::fabs(r[i]));
                }
                w[i] = c.active[i] ? wi : 0.0;
                // robust == 2: near the solution the gain weights of the outliers
                // fall from the IRLS weight gamma/(sigma^2 |u|) to
                // (gamma/|u|)^6 / sigma^2: nearer the Huber loss's zero curvature
                // beyond gamma, yet continuous at gamma (a weight that jumps to
                // ~0 there lets rows near the threshold flip in and out and the
                // iteration cycle); the right-hand side stays H^T psi(u)/sigma,
                // so the minimiser is the IRLS one
                wg[i] = w[i];
                if (newton) {
                    const double u = std::fabs(r[i]) / s;
                    if (u > prm->gamma && c.active[i]) {
                        const double x = prm->gamma / u;
                        wg[i] = x * x * x * x * x * x / (s * s);
                    }
                }
            }
            eval_jacobian(c, c.V.data(), c.th.data(), P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(),
                          Qt.data(), hv.data());
            bool refactor = prm->robust || !c.factor_valid || prm->mode == 1;
            if (refactor) {
                gain(c, hv.data(), wg.data(), &g);
                if (!factor(c, g)) {
                    singular = true;
                    c.factor_valid = false;
                    break;
                }
                c.factor_valid = true;
                R.refactors++;
            }
            for (int j = 0; j < c.ns; ++j) {
                double s = 0;
                for (int k = tptr[j]; k < tptr[j + 1]; ++k) s += hv[tpos[k]] * w[trow[k]] * r[trow[k]];
                b[j] = s;
            }
            solve_factor(c, b, dx);
            double mx = 0;
            bool finite = true;
            for (int j = 0; j < c.ns; ++j) {
                if (!std::isfinite(dx[j])) finite = false;
                mx = std::max(mx, std::fabs(dx[j]));
            }
            R.iterations = it + 1;
            R.max_mismatch = mx;
            last_step = mx;
            if (!finite) {
                diverged = true;
                break;
            }
            for (int bI = 0; bI < c.n; ++bI) {
                if (has_th(c, bI)) c.th[bI] += dx[th_idx(c, bI)];
                double v = c.V[bI] + dx[v_idx(c, bI)];
                c.V[bI] = std::min(std::max(v, prm->vmin), prm->vmax);
            }
            if (mx < prm->tol) {
                converged = true;
                break;
            }
        }
        if (singular)
            R.status = 3;
        else if (diverged)
            R.status = 2;
        else if (converged)
            R.status = 0;
        else
            R.status = 1;
        if ((diverged) && retries < 2 && prm->mode != 0) {
            retry = true;
            retries++;
        }
    } while (retry);
    // results at the final state
    eval_power(c, c.V.data(), c.th.data(), P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data());
    eval_h(c, c.V.data(), c.th.data(), P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data(), hx.data());
    double J = 0;
    c.res.assign(c.m, 0.0);
    for (int i = 0; i < c.m; ++i) {
        if (!c.active[i]) continue;
        c.res[i] = c.z[i] - hx[i];
        J += c.res[i] * c.res[i] / (c.sigma[i] * c.sigma[i]);
    }
    c.J = J;
    c.h = hx;
    R.objective = J;
    R.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out = R;
}

// Largest-normalized-residual statistics at the current state (Abur &
// Exposito 2004, ch. 5): with R = diag(sigma^2), G = H^T R^-1 H and the
// residual covariance Omega = R - H G^-1 H^T, omega[i] = Omega_ii and
// rn[i] = |r_i| / sqrt(Omega_ii).  Inactive rows and critical measurements
// (Omega_ii <= 1e-10 sigma^2: the residual is structurally zero) get rn = 0.
// The reference thresholds |r|/sigma instead (sle_engine.cu:959-977).
// Returns 0, or 3 when the gain is singular.
int32_t or_normalized_residuals(void* hh, double* rn, double* omega) {
    Case& c = *(Case*)hh;
    std::vector<double> hx(c.m), P(c.n), Q(c.n), Pf(c.nb), Qf(c.nb), Pt(c.nb), Qt(c.nb), w(c.m);
    std::vector<double> hv(c.hcol.size()), g, zval;
    eval_power(c, c.V.data(), c.th.data(), P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data());
    eval_h(c, c.V.data(), c.th.data(), P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data(), hx.data());
    eval_jacobian(c, c.V.data(), c.th.data(), P.data(), Q.data(), Pf.data(), Qf.data(), Pt.data(), Qt.data(),
                  hv.data());
    for (int i = 0; i < c.m; ++i) w[i] = c.active[i] ? 1.0 / (c.sigma[i] * c.sigma[i]) : 0.0;
    gain(c, hv.data(), w.data(), &g);
    if (!factor(c, g)) {
        c.factor_valid = false;
        return 3;
    }
    c.factor_valid = true;
    selected_inverse(c, zval);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c.m; ++i) {
        rn[i] = 0.0;
        omega[i] = 0.0;
        if (!c.active[i]) continue;
        const double s2 = c.sigma[i] * c.sigma[i];
        double lev = 0;  // h_i G^-1 h_i^T / sigma_i^2
        for (int a = c.hrow[i]; a < c.hrow[i + 1]; ++a)
            for (int b = c.hrow[i]; b < c.hrow[i + 1]; ++b)
                lev += hv[a] * hv[b] * z_at(c, zval, c.iperm[c.hcol[a]], c.iperm[c.hcol[b]]);
        lev /= s2;
        const double om = s2 * (1.0 - lev);
        omega[i] = om;
        const double r = c.z[i] - hx[i];
        rn[i] = om > 1e-10 * s2 ? std::fabs(r) / std::sqrt(om) : 0.0;
    }
    return 0;
}

void or_get_residuals(void* h, double* res) {
    Case* c = (Case*)h;
    std::copy(c->res.begin(), c->res.end(), res);
}

int32_t or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Threads of the following oracle calls (the CPU baseline is timed on all
// host cores and on one).
void or_set_num_threads(int32_t n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

}  // extern "C"
