"""Pins of the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped
shunt term, a wrong sign, a transposed operand, a wrong slack index) fails:
closed-form 2-bus admittances, the complex-phasor measurement function of the
grid generator, injection = sum of branch flows, central finite differences of
h(x), dense H^T W H, brute-force symbolic elimination, dense Cholesky,
noise-free recovery, first-order optimality, a chi-square check, and golden
vectors from scipy.optimize.least_squares (tests/golden/make_golden.py).
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle.oracle import Oracle
from state_estimation_with_cuda_amd import gridgen as gg

HERE = os.path.dirname(os.path.abspath(__file__))


def two_bus(r=0.02, x=0.1, B=0.04, tap=1.0, shift=0.0):
    g = gg.grid14(0)
    g.n = 2
    g.bus_type = np.array([2, 0], np.int32)
    g.vset = np.ones(2)
    g.br_from = np.array([0], np.int32)
    g.br_to = np.array([1], np.int32)
    g.r, g.x, g.b = np.array([r]), np.array([x]), np.array([B])
    g.tap, g.shift = np.array([tap]), np.array([shift])
    g.is_xfmr = np.array([int(tap != 1.0)], np.int32)
    g.status = np.ones(1, np.uint8)
    g.m_type = np.array([0, 0, 2, 3, 4, 5, 4, 5, 6, 6, 1], np.int32)
    g.m_loc = np.array([0, 1, 1, 1, 0, 0, 0, 0, 0, 0, 1], np.int32)
    g.m_end = np.array([0, 0, 0, 0, 0, 0, 1, 1, 0, 1, 0], np.int32)
    g.sigma = np.full(11, 0.01)
    g.V_true = np.array([1.02, 0.97])
    g.th_true = np.array([0.0, -0.08])
    g.z_true = gg.measurement_function(g, g.V_true, g.th_true)
    g.z = g.z_true.copy()
    return g


def variant14():
    """14-bus grid with a phase shifter, a tap transformer with charging, and
    I_MAG / V_ANGLE / pseudo measurements."""
    g = gg.grid14(3)
    k = int(np.nonzero(g.is_xfmr)[0][0])
    g.shift = g.shift.copy()
    g.shift[k] = 0.05
    g.b = g.b.copy()
    g.b[k] = 0.02
    extra_t = np.array([6, 6, 1, 1, 7, 8], np.int32)
    extra_l = np.array([2, 5, 3, 9, 4, 4], np.int32)
    extra_e = np.array([0, 1, 0, 0, 0, 0], np.int32)
    g.m_type = np.concatenate([g.m_type, extra_t])
    g.m_loc = np.concatenate([g.m_loc, extra_l])
    g.m_end = np.concatenate([g.m_end, extra_e])
    g.sigma = np.concatenate([g.sigma, np.full(6, 0.01)])
    g.z_true = gg.measurement_function(g, g.V_true, g.th_true)
    g.z = g.z_true + np.random.default_rng(1).normal(0, 1, g.m) * g.sigma
    return g


def dense_from_csr(rp, col, val, shape):
    return sp.csr_matrix((val, col, rp), shape=shape).toarray()


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("tap,shift", [(1.0, 0.0), (0.97, 0.0), (1.03, 0.1)])
def test_two_bus_admittance_closed_form(tap, shift):
    r, x, B = 0.02, 0.1, 0.04
    g = two_bus(r, x, B, tap, shift)
    o = Oracle(g)
    rp, col, G, Bm = o.ybus()
    Y = dense_from_csr(rp, col, G + 1j * Bm, (2, 2))
    y = 1.0 / complex(r, x)
    t = tap * np.exp(1j * shift)
    expect = np.array([[(y + 0.5j * B) / tap**2, -y / np.conj(t)], [-y / t, y + 0.5j * B]])
    assert np.allclose(Y, expect, rtol=1e-14, atol=1e-14)
    if shift == 0.0:
        assert np.allclose(Y, Y.T)
    if tap == 1.0 and shift == 0.0:
        assert np.allclose(Y.sum(axis=1), 0.5j * B)


@pytest.mark.parametrize("tap,shift", [(1.0, 0.0), (0.95, 0.0), (1.04, -0.07)])
def test_two_bus_flows_closed_form(tap, shift):
    g = two_bus(0.03, 0.12, 0.05, tap, shift)
    o = Oracle(g)
    V, th = g.V_true, g.th_true
    e = o.eval(V, th)
    y = 1.0 / complex(0.03, 0.12)
    t = tap * np.exp(1j * shift)
    E = V * np.exp(1j * th)
    If = (y + 0.025j) / tap**2 * E[0] - y / np.conj(t) * E[1]
    It = -y / t * E[0] + (y + 0.025j) * E[1]
    Sf, St = E[0] * np.conj(If), E[1] * np.conj(It)
    assert abs(e["Pf"][0] - Sf.real) < 1e-14 and abs(e["Qf"][0] - Sf.imag) < 1e-14
    assert abs(e["Pt"][0] - St.real) < 1e-14 and abs(e["Qt"][0] - St.imag) < 1e-14


@pytest.mark.parametrize("maker", [lambda: gg.grid14(0), variant14, lambda: gg.geometric(1500, seed=3, flow_meas=3000)])
def test_h_matches_phasor_form_and_conservation(maker):
    g = maker()
    o = Oracle(g)
    rng = np.random.default_rng(2)
    V = g.V_true + rng.normal(0, 0.03, g.n)
    th = g.th_true + rng.normal(0, 0.03, g.n)
    th[g.slack] = 0
    e = o.eval(V, th)
    hp = gg.measurement_function(g, V, th)
    assert np.max(np.abs(e["h"] - hp)) < 1e-12 * max(1.0, np.abs(hp).max())
    # injection = sum of branch end flows (no bus shunts)
    P = np.zeros(g.n)
    Q = np.zeros(g.n)
    np.add.at(P, g.br_from, e["Pf"])
    np.add.at(P, g.br_to, e["Pt"])
    np.add.at(Q, g.br_from, e["Qf"])
    np.add.at(Q, g.br_to, e["Qt"])
    assert np.max(np.abs(P - e["P"])) < 1e-12 and np.max(np.abs(Q - e["Q"])) < 1e-12


@pytest.mark.parametrize("maker", [lambda: gg.grid14(0), variant14, lambda: two_bus(0.02, 0.1, 0.04, 0.97, 0.05)])
def test_jacobian_central_differences(maker):
    g = maker()
    o = Oracle(g)
    rng = np.random.default_rng(4)
    V = g.V_true + rng.normal(0, 0.02, g.n)
    th = g.th_true + rng.normal(0, 0.02, g.n)
    th[g.slack] = 0
    rp, col = o.h_pattern()
    Hs = dense_from_csr(rp, col, o.jacobian(V, th), (g.m, o.ns))
    n, s = g.n, g.slack
    Hfd = np.zeros((g.m, o.ns))
    step = 1e-6
    for j in range(o.ns):
        Vp, Vm, tp, tm = V.copy(), V.copy(), th.copy(), th.copy()
        if j < n - 1:
            bus = j if j < s else j + 1
            tp[bus] += step
            tm[bus] -= step
        else:
            bus = j - (n - 1)
            Vp[bus] += step
            Vm[bus] -= step
        Hfd[:, j] = (o.eval(Vp, tp)["h"] - o.eval(Vm, tm)["h"]) / (2 * step)
    err = np.abs(Hs - Hfd).max() / np.abs(Hfd).max()
    assert err < 1e-7, err
    # structural pattern covers every nonzero derivative
    assert np.all((np.abs(Hfd) > 1e-9) <= (dense_from_csr(rp, col, np.ones(col.shape[0]), (g.m, o.ns)) > 0))


def test_gain_is_dense_htwh():
    g = variant14()
    o = Oracle(g)
    V, th = g.V_true, g.th_true
    rp, col = o.h_pattern()
    H = dense_from_csr(rp, col, o.jacobian(V, th), (g.m, o.ns))
    w = np.random.default_rng(0).uniform(0.5, 2.0, g.m)
    Gd = H.T @ (w[:, None] * H)
    cp, rw = o.g_pattern()
    Gl = sp.csc_matrix((o.gain(V, th, w), rw, cp), shape=(o.ns, o.ns)).toarray()
    assert np.allclose(np.tril(Gd), Gl, rtol=1e-13, atol=1e-12 * np.abs(Gd).max())
    # the lower pattern is exactly the structural product
    S = (np.abs(H) > 0).astype(float)
    struct = np.tril(S.T @ S) > 0
    assert np.array_equal(struct, sp.csc_matrix((np.ones(rw.size), rw, cp), shape=(o.ns, o.ns)).toarray() > 0)


def brute_symbolic(A):
    """Boolean elimination of a symmetric pattern in the given order."""
    n = A.shape[0]
    L = np.tril(A.copy().astype(bool))
    for j in range(n):
        rows = np.nonzero(L[j + 1:, j])[0] + j + 1
        for a in rows:
            L[a, rows[rows <= a]] = True
    parent = np.full(n, -1)
    for j in range(n):
        r = np.nonzero(L[j + 1:, j])[0]
        if r.size:
            parent[j] = j + 1 + r[0]
    colcount = L.sum(axis=0) + (~np.diag(L)).astype(int) * 0
    return L, parent, colcount


def test_symbolic_matches_brute_force():
    for g in (gg.grid14(0), variant14(), gg.geometric(60, seed=1, flow_meas="all")):
        o = Oracle(g)
        cp, rw = o.g_pattern()
        Gp = sp.csc_matrix((np.ones(rw.size), rw, cp), shape=(o.ns, o.ns)).toarray() > 0
        Gp = Gp | Gp.T
        s = o.symbolic()
        perm = s["perm"]
        A = Gp[np.ix_(perm, perm)]
        np.fill_diagonal(A, True)
        L, parent, cc = brute_symbolic(A)
        assert np.array_equal(parent, s["parent"])
        assert np.array_equal(cc, s["colcount"])
        assert L.sum() == o.nnz_l
        # fundamental supernodes: j joins j-1 iff parent(j-1)=j, j has one child, cc(j-1)=cc(j)+1
        nchild = np.bincount(parent[parent >= 0], minlength=o.ns)
        starts = [0] + [j for j in range(1, o.ns)
                        if not (parent[j - 1] == j and nchild[j] == 1 and cc[j - 1] == cc[j] + 1)]
        assert np.array_equal(np.array(starts + [o.ns]), s["snode_ptr"])


def exact_min_degree(adj_sets):
    """Exact minimum degree by definition (ties: lowest index)."""
    adj = [set(a) for a in adj_sets]
    alive = set(range(len(adj)))
    order = []
    while alive:
        v = min(alive, key=lambda u: (len(adj[u]), u))
        order.append(v)
        nb = adj[v]
        for u in nb:
            adj[u] |= nb
            adj[u].discard(u)
            adj[u].discard(v)
        alive.discard(v)
        adj[v] = set()
    return order


def test_ordering_is_exact_minimum_degree_and_reduces_fill():
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    for g in (gg.grid14(0), gg.geometric(300, seed=6, flow_meas="all")):
        o = Oracle(g)
        cp, rw = o.g_pattern()
        n = g.n
        sbus = np.empty(o.ns, int)
        s = g.slack
        for b in range(n):
            if b != s:
                sbus[b if b < s else b - 1] = b
            sbus[n - 1 + b] = b
        adj = [set() for _ in range(n)]
        for j in range(o.ns):
            for q in range(cp[j], cp[j + 1]):
                a, b = sbus[rw[q]], sbus[j]
                if a != b:
                    adj[a].add(b)
                    adj[b].add(a)
        assert exact_min_degree(adj) == list(o.symbolic()["bus_perm"])
        # fill: minimum degree beats natural order and RCM on the bus graph
        A = np.zeros((n, n), bool)
        for a in range(n):
            A[a, list(adj[a])] = True
        np.fill_diagonal(A, True)

        def fill(order):
            return brute_symbolic(A[np.ix_(order, order)])[0].sum()

        md = fill(np.array(o.symbolic()["bus_perm"]))
        rcm = fill(reverse_cuthill_mckee(sp.csr_matrix(A), symmetric_mode=True))
        # minimum degree is a heuristic: it must clearly beat the natural order
        # and stay within a small factor of RCM (which can win on tree-like grids)
        assert md <= fill(np.arange(n)) and md <= 1.25 * rcm


def test_factor_matches_dense_cholesky():
    g = gg.geometric(200, seed=2, flow_meas="all")
    o = Oracle(g)
    w = 1 / g.sigma**2
    gv = o.gain(g.V_true, g.th_true, w)
    cp, rw = o.g_pattern()
    Gl = sp.csc_matrix((gv, rw, cp), shape=(o.ns, o.ns)).toarray()
    G = Gl + Gl.T - np.diag(np.diag(Gl))
    perm = o.symbolic()["perm"]
    Ld = np.linalg.cholesky(G[np.ix_(perm, perm)])
    lp, lr, lv = o.factor(gv)
    L = sp.csc_matrix((lv, lr, lp), shape=(o.ns, o.ns)).toarray()
    assert np.abs(L - Ld).max() < 1e-10 * np.abs(Ld).max()


@pytest.mark.parametrize("maker", [lambda: gg.grid14(0), variant14, lambda: gg.geometric(2000, seed=5, flow_meas=5000)])
def test_noise_free_recovery(maker):
    g = maker()
    g.z = g.z_true.copy()
    o = Oracle(g)
    r = o.solve(mode=1, max_iter=50, tol=1e-12, flat_start=True)
    V, th = o.state()
    assert r["status"] == 0
    assert np.abs(V - g.V_true).max() < 1e-10 and np.abs(th - g.th_true).max() < 1e-10
    assert r["objective"] < 1e-16 * g.m


def test_first_order_optimality_and_objective():
    g = gg.geometric(2000, seed=7, flow_meas=4000)
    o = Oracle(g)
    r = o.solve(mode=1, max_iter=50, tol=1e-12, flat_start=True)
    V, th = o.state()
    rp, col = o.h_pattern()
    H = sp.csr_matrix((o.jacobian(V, th), col, rp), shape=(g.m, o.ns))
    res = o.residuals()
    w = 1 / g.sigma**2
    grad = H.T @ (w * res)
    scale = np.abs(H.T @ (w * g.z)).max()
    assert np.abs(grad).max() < 1e-9 * scale
    assert abs(r["objective"] - np.sum(w * res**2)) < 1e-9 * r["objective"]


def test_chi_square_sanity():
    Js = []
    for seed in range(30):
        g = gg.grid14(seed + 100)
        o = Oracle(g)
        r = o.solve(mode=1, max_iter=50, tol=1e-10, flat_start=True)
        assert r["status"] == 0
        Js.append(r["objective"])
    dof = 122 - 27
    assert abs(np.mean(Js) - dof) < 4 * np.sqrt(2 * dof / 30)


def test_golden_vectors_scipy_least_squares():
    d = json.load(open(os.path.join(HERE, "golden", "grid14_golden.json")))
    g = gg.grid14(0)
    o = Oracle(g)
    r = o.solve(mode=1, tol=1e-12, flat_start=True)
    V, th = o.state()
    assert np.abs(V - d["grid14_wls"]["V"]).max() < 1e-8
    assert np.abs(th - d["grid14_wls"]["theta"]).max() < 1e-8
    assert abs(r["objective"] - d["grid14_wls"]["J"]) < 1e-9 * d["grid14_wls"]["J"]
    gh = gg.grid14(0)
    gh.z = gh.z.copy()
    rows = d["grid14_huber"]["gross_rows"]
    gh.z[rows] += 20 * gh.sigma[rows]
    o = Oracle(gh)
    r = o.solve(mode=1, tol=1e-12, flat_start=True, robust=True, gamma=d["grid14_huber"]["gamma"], max_iter=500)
    assert r["status"] == 0
    V, th = o.state()
    assert np.abs(V - d["grid14_huber"]["V"]).max() < 1e-8
    assert np.abs(th - d["grid14_huber"]["theta"]).max() < 1e-8


def test_huber_reduces_to_wls_for_large_gamma():
    g = gg.grid14(2)
    a, b = Oracle(g), Oracle(g)
    a.solve(mode=1, tol=1e-12, flat_start=True)
    b.solve(mode=1, tol=1e-12, flat_start=True, robust=True, gamma=1e6)
    Va, ta = a.state()
    Vb, tb = b.state()
    assert np.abs(Va - Vb).max() < 1e-12 and np.abs(ta - tb).max() < 1e-12


def test_realtime_mode_reuses_factor_and_reaches_same_optimum():
    g = gg.geometric(1000, seed=12, flow_meas=2500)
    o = Oracle(g)
    r0 = o.solve(mode=1, tol=1e-11, flat_start=True)
    V0, t0 = o.state()
    z = g.z.copy()
    rng = np.random.default_rng(3)
    idx = rng.choice(g.m, g.m // 10, replace=False)
    z[idx] += rng.normal(0, 1, idx.size) * g.sigma[idx]
    o.set_z(z)
    r1 = o.solve(mode=0, max_iter=50, tol=1e-11)
    assert r1["status"] == 0 and r1["refactors"] == 0
    g2 = gg.Grid(**{**g.__dict__})
    g2.z = z
    o2 = Oracle(g2)
    o2.solve(mode=1, tol=1e-11, flat_start=True)
    V1, t1 = o.state()
    V2, t2 = o2.state()
    assert np.abs(V1 - V2).max() < 1e-9 and np.abs(t1 - t2).max() < 1e-9
    del r0, V0, t0


def test_independent_areas_solve_like_separate_models():
    """Areas side by side (one SLACK bus each, no branch between them) must give
    each area's own WLS optimum: a wrong angle-reference rule (a single global
    reference, a shifted state index) makes the gain singular or moves the
    angles of the later areas."""
    areas = [gg.geometric(300, seed=31, flow_meas=700), gg.grid14(3), gg.geometric(200, seed=32, flow_meas="all")]
    g = gg.concat(areas)
    o = Oracle(g)
    r = o.solve(mode=1, tol=1e-11, flat_start=True)
    assert r["status"] == 0
    V, th = o.state()
    J = 0.0
    b0 = 0
    for a in areas:
        oa = Oracle(a)
        ra = oa.solve(mode=1, tol=1e-11, flat_start=True)
        Va, tha = oa.state()
        assert np.abs(V[b0:b0 + a.n] - Va).max() < 1e-9
        assert np.abs(th[b0:b0 + a.n] - tha).max() < 1e-9
        assert th[b0 + a.slack] == 0.0
        J += ra["objective"]
        b0 += a.n
    assert abs(r["objective"] - J) < 1e-10 * J


def test_huber_newton_reaches_the_irls_minimiser():
    """Huber with reduced gain weights near the solution (outliers at
    (gamma/|u|)^6 of their weight in the gain once the steps are small)
    converges to the same M-estimate as the reference's IRLS weights, in at most
    two thirds of the iterations."""
    g = gg.geometric(3000, seed=23, flow_meas="all", dup_meas=600, gross_frac=0.01)
    o = Oracle(g)
    ri = o.solve(mode=1, max_iter=400, tol=1e-11, robust=1, gamma=1.5, flat_start=True)
    Vi, ti = o.state()
    rn = o.solve(mode=1, max_iter=400, tol=1e-11, robust=2, gamma=1.5, flat_start=True)
    Vn, tn = o.state()
    assert ri["status"] == 0 and rn["status"] == 0
    assert np.abs(Vi - Vn).max() < 1e-9 and np.abs(ti - tn).max() < 1e-9
    print(f"IRLS {ri['iterations']} iterations, reduced gain weights {rn['iterations']}")
    assert 3 * rn["iterations"] <= 2 * ri["iterations"]


# ---------------------------------------------------------------------------
# Largest normalized residual (bad-data identification)
@pytest.mark.parametrize("maker", [lambda: gg.grid14(0), variant14, lambda: gg.geometric(60, seed=9, flow_meas=120)])
def test_normalized_residuals_match_dense_covariance(maker):
    """Omega = R - H G^-1 H^T computed densely at the oracle's optimum equals the
    oracle's sparse-inverse-subset diagonal; rn = |r| / sqrt(Omega_ii); the
    leverages 1 - Omega_ii / sigma_i^2 sum to the number of states."""
    g = maker()
    o = Oracle(g)
    assert o.solve(mode=1, max_iter=50, tol=1e-12, flat_start=True)["status"] == 0
    V, th = o.state()
    rn, om = o.normalized_residuals()
    rp, col = o.h_pattern()
    H = dense_from_csr(rp, col, o.jacobian(V, th), (g.m, o.ns))
    R = g.sigma ** 2
    G = H.T @ (H / R[:, None])
    Om = R - np.einsum("ij,ji->i", H, np.linalg.solve(G, H.T))
    assert np.allclose(om, Om, rtol=1e-8, atol=1e-12 * R.max())
    r = g.z - o.eval(V, th)["h"]
    test = Om > 1e-10 * R
    assert np.allclose(rn[test], np.abs(r[test]) / np.sqrt(Om[test]), rtol=1e-7)
    assert np.all(rn[~test] == 0)
    assert abs((1.0 - om / R).sum() - o.ns) < 1e-8 * o.ns


def test_largest_normalized_residual_finds_a_gross_error():
    """One reading corrupted by +20 sigma in a redundant grid: it has the
    largest normalized residual, far above 3; without it none exceeds 4."""
    g = gg.geometric(200, seed=12, flow_meas="all")
    bad = 137
    z = g.z.copy()
    z[bad] += 20 * g.sigma[bad]
    g2 = gg.Grid(**{**g.__dict__})
    g2.z = z
    o = Oracle(g2)
    assert o.solve(mode=1, max_iter=50, tol=1e-12, flat_start=True)["status"] == 0
    rn, _ = o.normalized_residuals()
    assert int(np.argmax(rn)) == bad and rn[bad] > 10
    o = Oracle(g)
    assert o.solve(mode=1, max_iter=50, tol=1e-12, flat_start=True)["status"] == 0
    rn, _ = o.normalized_residuals()
    assert rn.max() < 4.5


@pytest.mark.parametrize("maker", [lambda: gg.grid14(0), variant14])
def test_rhs_is_half_the_negative_objective_gradient(maker):
    """b = H^T W r equals -1/2 dJ/dx for J = sum w (z - h(x))^2, by central
    differences of the oracle's h(x) (state layout: angles, then magnitudes)."""
    g = maker()
    o = Oracle(g)
    rng = np.random.default_rng(6)
    V = g.V_true + rng.normal(0, 0.02, g.n)
    th = g.th_true + rng.normal(0, 0.02, g.n)
    th[g.slack] = 0
    b = o.rhs(V, th)
    w = 1.0 / g.sigma ** 2

    def J(Vx, tx):
        r = g.z - o.eval(Vx, tx)["h"]
        return float(np.sum(w * r * r))

    n, s, step = g.n, g.slack, 1e-6
    fd = np.zeros(o.ns)
    for j in range(o.ns):
        Vp, Vm, tp, tm = V.copy(), V.copy(), th.copy(), th.copy()
        if j < n - 1:
            bus = j if j < s else j + 1
            tp[bus] += step
            tm[bus] -= step
        else:
            bus = j - (n - 1)
            Vp[bus] += step
            Vm[bus] -= step
        fd[j] = -0.5 * (J(Vp, tp) - J(Vm, tm)) / (2 * step)
    assert np.max(np.abs(b - fd)) < 1e-5 * max(1.0, np.abs(b).max())
