"""Seeded synthetic power grids and measurement sets for the WLS workloads.

Workloads (sizes as named by the estimator's benchmark plan):
  * ``grid14(seed)``        14-bus meshed grid, 20 branches, 3 tap transformers,
                            V + P/Q injections at all buses + P/Q flows at both
                            ends of every branch (m = 122).
  * ``geometric(n, ...)``   random geometric mesh: n buses in the unit square,
                            Euclidean MST over the Delaunay edges plus the
                            shortest remaining Delaunay edges up to ~1.35 n
                            branches (mean degree ~2.7), 10 % tap transformers.
                            Used for the 10k-bus, 200k-bus, batched-area and
                            serving workloads.

Branch parameters: r ~ U(0.01, 0.1), x ~ U(0.05, 0.3) pu, line charging
B ~ U(0, 0.05) on lines, transformers tap ~ U(0.95, 1.05) with B = 0 and no
phase shift.  True state V ~ U(0.95, 1.05), theta ~ U(-0.3, 0.3) rad, bus 0 is
the slack (theta = 0, V setpoint 1.0).  Measurements are h(x_true) plus
Gaussian noise, sigma = 0.004 pu for |V| and 0.01 pu for P/Q.

h(x) here is written in complex phasor form (branch end currents I = Y V, end
flows S = V conj(I), injections = sum of the end flows of the closed branches
at a bus) so that it is an independent restatement of the trigonometric Ybus
forms used by the engine and the oracle; noise-free recovery tests rely on
that independence.

Measurement type codes and the pi-model follow the reference ABI
(sle_api.h:183-193) and MATPOWER's branch model.
"""
from __future__ import annotations

import dataclasses

import numpy as np

V_MAG, V_ANGLE, P_INJ, Q_INJ, P_FLOW, Q_FLOW, I_MAG, P_PSEUDO, Q_PSEUDO = range(9)
BUS_PQ, BUS_PV, BUS_SLACK = 0, 1, 2
SIGMA_V, SIGMA_PQ = 0.004, 0.01


@dataclasses.dataclass
class Grid:
    n: int
    slack: int
    bus_type: np.ndarray  # int32 [n]
    vset: np.ndarray  # f64 [n]
    br_from: np.ndarray  # int32 [b]
    br_to: np.ndarray  # int32 [b]
    r: np.ndarray
    x: np.ndarray
    b: np.ndarray  # total line charging
    tap: np.ndarray
    shift: np.ndarray
    is_xfmr: np.ndarray  # int32 [b]
    status: np.ndarray  # uint8 [b]
    m_type: np.ndarray  # int32 [m]
    m_loc: np.ndarray  # int32 [m]
    m_end: np.ndarray  # int32 [m]
    sigma: np.ndarray  # f64 [m]
    z: np.ndarray  # f64 [m]
    V_true: np.ndarray
    th_true: np.ndarray
    z_true: np.ndarray  # h(x_true) without noise
    name: str = ""

    @property
    def nb(self) -> int:
        return int(self.br_from.shape[0])

    @property
    def m(self) -> int:
        return int(self.m_type.shape[0])

    @property
    def ns(self) -> int:
        # one angle reference per island; generated grids have one SLACK bus per island
        return 2 * self.n - int(np.count_nonzero(self.bus_type == BUS_SLACK))


# ---------------------------------------------------------------------------
# h(x) in phasor form
# ---------------------------------------------------------------------------
def branch_admittances(g: Grid):
    d = g.r * g.r + g.x * g.x
    d = np.maximum(d, 1e-12)
    ys = (g.r - 1j * g.x) / d
    t = g.tap * np.exp(1j * g.shift)
    on = g.status.astype(bool)
    yff = np.where(on, (ys + 0.5j * g.b) / (g.tap * g.tap), 0)
    ytt = np.where(on, ys + 0.5j * g.b, 0)
    yft = np.where(on, -ys / np.conj(t), 0)
    ytf = np.where(on, -ys / t, 0)
    return yff, ytt, yft, ytf


def power_flow_quantities(g: Grid, V: np.ndarray, th: np.ndarray):
    """End flows S = V conj(I) of every branch and bus injections as their sums."""
    yff, ytt, yft, ytf = branch_admittances(g)
    f, t = g.br_from, g.br_to
    E = V * np.exp(1j * th)
    If = yff * E[f] + yft * E[t]
    It = ytf * E[f] + ytt * E[t]
    Sf = E[f] * np.conj(If)
    St = E[t] * np.conj(It)
    # injection = sum of flows leaving the bus through each closed branch
    n = g.n
    Sinj = np.zeros(n, dtype=complex)
    np.add.at(Sinj, f, Sf)
    np.add.at(Sinj, t, St)
    return Sinj, Sf, St


def measurement_function(g: Grid, V: np.ndarray, th: np.ndarray, m_type=None, m_loc=None, m_end=None):
    m_type = g.m_type if m_type is None else m_type
    m_loc = g.m_loc if m_loc is None else m_loc
    m_end = g.m_end if m_end is None else m_end
    Sinj, Sf, St = power_flow_quantities(g, V, th)
    h = np.zeros(m_type.shape[0])
    mt, ml, me = m_type, m_loc, m_end
    sel = mt == V_MAG
    h[sel] = V[ml[sel]]
    sel = mt == V_ANGLE
    h[sel] = th[ml[sel]]
    sel = (mt == P_INJ) | (mt == P_PSEUDO)
    h[sel] = Sinj[ml[sel]].real
    sel = (mt == Q_INJ) | (mt == Q_PSEUDO)
    h[sel] = Sinj[ml[sel]].imag
    Send = np.where(me == 1, St[np.minimum(ml, g.nb - 1)], Sf[np.minimum(ml, g.nb - 1)]) if g.nb else np.zeros(0)
    sel = mt == P_FLOW
    h[sel] = Send[sel].real
    sel = mt == Q_FLOW
    h[sel] = Send[sel].imag
    sel = mt == I_MAG
    if sel.any():
        vend = np.where(me == 1, V[g.br_to[np.minimum(ml, g.nb - 1)]], V[g.br_from[np.minimum(ml, g.nb - 1)]])
        h[sel] = np.abs(Send[sel]) / vend[sel]
    # flows on open branches are inactive (h = 0 by convention)
    if g.nb:
        flow = (mt == P_FLOW) | (mt == Q_FLOW) | (mt == I_MAG)
        open_ = np.zeros(mt.shape[0], dtype=bool)
        open_[flow] = g.status[ml[flow]] == 0
        h[open_] = 0.0
    return h


# ---------------------------------------------------------------------------
# Builders
# ---------------------------------------------------------------------------
def _branch_params(rng, nb, xfmr_idx):
    r = rng.uniform(0.01, 0.1, nb)
    x = rng.uniform(0.05, 0.3, nb)
    b = rng.uniform(0.0, 0.05, nb)
    tap = np.ones(nb)
    is_x = np.zeros(nb, dtype=np.int32)
    is_x[xfmr_idx] = 1
    tap[xfmr_idx] = rng.uniform(0.95, 1.05, len(xfmr_idx))
    b[xfmr_idx] = 0.0
    return r, x, b, tap, is_x


def _finish(rng, n, f, t, r, x, b, tap, is_x, meas, name, gross_frac=0.0, gross_sigmas=20.0):
    m_type, m_loc, m_end = (np.asarray(a, dtype=np.int32) for a in meas)
    sigma = np.where(m_type == V_MAG, SIGMA_V, SIGMA_PQ).astype(np.float64)
    V = rng.uniform(0.95, 1.05, n)
    th = rng.uniform(-0.3, 0.3, n)
    th[0] = 0.0
    bus_type = np.zeros(n, dtype=np.int32)
    bus_type[0] = BUS_SLACK
    vset = np.ones(n)
    g = Grid(
        n=n,
        slack=0,
        bus_type=bus_type,
        vset=vset,
        br_from=np.asarray(f, dtype=np.int32),
        br_to=np.asarray(t, dtype=np.int32),
        r=r,
        x=x,
        b=b,
        tap=tap,
        shift=np.zeros(len(f)),
        is_xfmr=is_x,
        status=np.ones(len(f), dtype=np.uint8),
        m_type=m_type,
        m_loc=m_loc,
        m_end=m_end,
        sigma=sigma,
        z=np.zeros(len(m_type)),
        V_true=V,
        th_true=th,
        z_true=np.zeros(len(m_type)),
        name=name,
    )
    g.z_true = measurement_function(g, V, th)
    g.z = g.z_true + rng.normal(0.0, 1.0, g.m) * sigma
    if gross_frac > 0:
        k = int(round(gross_frac * g.m))
        idx = rng.choice(g.m, size=k, replace=False)
        g.z[idx] += gross_sigmas * sigma[idx]
    return g


def grid14(seed: int = 0) -> Grid:
    """14-bus meshed grid: random spanning tree + 6 extra distinct edges."""
    rng = np.random.default_rng(seed)
    n = 14
    order = rng.permutation(n)
    edges = set()
    f, t = [], []
    for k in range(1, n):
        a = int(order[k])
        bb = int(order[rng.integers(0, k)])
        edges.add((min(a, bb), max(a, bb)))
        f.append(a)
        t.append(bb)
    while len(f) < 20:
        a, bb = (int(v) for v in rng.integers(0, n, 2))
        if a == bb or (min(a, bb), max(a, bb)) in edges:
            continue
        edges.add((min(a, bb), max(a, bb)))
        f.append(a)
        t.append(bb)
    nb = len(f)
    xf = rng.choice(nb, size=3, replace=False)
    r, x, b, tap, is_x = _branch_params(rng, nb, xf)
    mt, ml, me = [], [], []
    for tp in (V_MAG, P_INJ, Q_INJ):
        mt += [tp] * n
        ml += list(range(n))
        me += [0] * n
    for k in range(nb):
        for end in (0, 1):
            for tp in (P_FLOW, Q_FLOW):
                mt.append(tp)
                ml.append(k)
                me.append(end)
    return _finish(rng, n, f, t, r, x, b, tap, is_x, (mt, ml, me), "grid14")


def _geometric_topology(rng, n, target_b):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree
    from scipy.spatial import Delaunay

    pts = rng.random((n, 2))
    tri = Delaunay(pts)
    s = tri.simplices
    e = np.concatenate([s[:, [0, 1]], s[:, [1, 2]], s[:, [0, 2]]])
    e.sort(axis=1)
    e = np.unique(e, axis=0)
    ln = np.linalg.norm(pts[e[:, 0]] - pts[e[:, 1]], axis=1)
    A = coo_matrix((ln, (e[:, 0], e[:, 1])), shape=(n, n)).tocsr()
    T = minimum_spanning_tree(A).tocoo()
    tree = np.stack([np.minimum(T.row, T.col), np.maximum(T.row, T.col)], axis=1)
    key = e[:, 0].astype(np.int64) * n + e[:, 1]
    tkey = tree[:, 0].astype(np.int64) * n + tree[:, 1]
    extra_mask = ~np.isin(key, tkey)
    extra = e[extra_mask]
    extra = extra[np.argsort(ln[extra_mask], kind="stable")]
    need = max(0, target_b - tree.shape[0])
    edges = np.concatenate([tree, extra[:need]])
    # randomise orientation and order
    flip = rng.random(edges.shape[0]) < 0.5
    edges[flip] = edges[flip][:, ::-1]
    edges = edges[rng.permutation(edges.shape[0])]
    return edges[:, 0].astype(np.int32), edges[:, 1].astype(np.int32)


def geometric(
    n: int,
    seed: int,
    flow_meas: str | int = "all",
    dup_meas: int = 0,
    gross_frac: float = 0.0,
    branches_per_bus: float = 1.35,
    name: str = "",
) -> Grid:
    """Random geometric mesh grid.

    flow_meas: "all" (P/Q at both ends of every branch) or an integer count of
    flow measurements drawn without replacement from the 4*b candidates.
    dup_meas: extra duplicate meters drawn uniformly (with replacement) from the
    injection and flow candidates, each with independent noise.
    """
    rng = np.random.default_rng(seed)
    f, t = _geometric_topology(rng, n, int(round(branches_per_bus * n)))
    nb = f.shape[0]
    xf = rng.choice(nb, size=int(round(0.1 * nb)), replace=False)
    r, x, b, tap, is_x = _branch_params(rng, nb, xf)
    buses = np.arange(n, dtype=np.int32)
    mt = [np.full(n, V_MAG), np.full(n, P_INJ), np.full(n, Q_INJ)]
    ml = [buses, buses, buses]
    me = [np.zeros(n, np.int32)] * 3
    cand_t = np.tile(np.array([P_FLOW, Q_FLOW, P_FLOW, Q_FLOW], np.int32), nb)
    cand_l = np.repeat(np.arange(nb, dtype=np.int32), 4)
    cand_e = np.tile(np.array([0, 0, 1, 1], np.int32), nb)
    if flow_meas == "all":
        sel = np.arange(4 * nb)
    else:
        sel = np.sort(rng.choice(4 * nb, size=int(flow_meas), replace=False))
    mt.append(cand_t[sel])
    ml.append(cand_l[sel])
    me.append(cand_e[sel])
    if dup_meas:
        # candidates: P/Q injections (2n) then flows (4 nb)
        pick = rng.integers(0, 2 * n + 4 * nb, dup_meas)
        inj = pick < 2 * n
        dt = np.where(inj, np.where(pick % 2 == 0, P_INJ, Q_INJ), 0).astype(np.int32)
        dl = np.where(inj, pick // 2, 0).astype(np.int32)
        de = np.zeros(dup_meas, np.int32)
        fl = ~inj
        q = pick[fl] - 2 * n
        dt[fl] = cand_t[q]
        dl[fl] = cand_l[q]
        de[fl] = cand_e[q]
        mt.append(dt)
        ml.append(dl)
        me.append(de)
    meas = (np.concatenate(mt), np.concatenate(ml), np.concatenate(me))
    return _finish(rng, n, f, t, r, x, b, tap, is_x, meas, name or f"geometric{n}", gross_frac=gross_frac)


# Named workloads ------------------------------------------------------------
def workload_grid14() -> Grid:
    return grid14(seed=0)


def workload_10k() -> Grid:
    """10,000 buses, ~13,500 branches, ~60,000 measurements (seed 1)."""
    return geometric(10_000, seed=1, flow_meas=30_000, name="geometric10k")


def workload_200k() -> Grid:
    """200,000 buses, ~270,000 branches, 2,000,000 measurements, 1 % gross errors (seed 2)."""
    n = 200_000
    return geometric(n, seed=2, flow_meas="all", dup_meas=320_000, gross_frac=0.01, name="geometric200k")


def area_sizes(count: int = 256, lo: int = 500, hi: int = 20_000, seed: int = 3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)


def workload_area(n: int, seed: int) -> Grid:
    """One area estimator: V + P/Q injections at all buses + ~3n flows (redundancy ~3)."""
    nb_est = int(round(1.35 * n))
    return geometric(int(n), seed=seed, flow_meas=min(3 * int(n), 4 * nb_est), name=f"area{n}")


def hot_start(g: Grid, V: np.ndarray, th: np.ndarray, sigma: float = 0.01, seed: int = 0):
    """Previous solution perturbed by N(0, sigma^2) per state element (reference angles fixed)."""
    rng = np.random.default_rng(seed)
    V2 = V + rng.normal(0, sigma, V.shape)
    th2 = th + rng.normal(0, sigma, th.shape)
    ref = g.bus_type == BUS_SLACK
    th2[ref] = th[ref]
    return V2, th2


def concat(grids, name: str = "") -> Grid:
    """Independent areas side by side as one model (block-diagonal gain matrix):
    buses, branches and measurements of area k follow those of areas < k; each
    area keeps its own SLACK bus, which is the angle reference of its island."""
    bo = np.cumsum([0] + [g.n for g in grids])
    ko = np.cumsum([0] + [g.nb for g in grids])
    cat = lambda f: np.concatenate([getattr(g, f) for g in grids])
    m_loc = []
    for g, b0, k0 in zip(grids, bo, ko):
        bus_meas = np.isin(g.m_type, [V_MAG, V_ANGLE, P_INJ, Q_INJ, P_PSEUDO, Q_PSEUDO])
        m_loc.append(np.where(bus_meas, g.m_loc + b0, g.m_loc + k0).astype(np.int32))
    return Grid(
        n=int(bo[-1]), slack=int(grids[0].slack), bus_type=cat("bus_type"), vset=cat("vset"),
        br_from=np.concatenate([g.br_from + b0 for g, b0 in zip(grids, bo)]).astype(np.int32),
        br_to=np.concatenate([g.br_to + b0 for g, b0 in zip(grids, bo)]).astype(np.int32),
        r=cat("r"), x=cat("x"), b=cat("b"), tap=cat("tap"), shift=cat("shift"), is_xfmr=cat("is_xfmr"),
        status=cat("status"), m_type=cat("m_type"), m_loc=np.concatenate(m_loc), m_end=cat("m_end"),
        sigma=cat("sigma"), z=cat("z"), V_true=cat("V_true"), th_true=cat("th_true"), z_true=cat("z_true"),
        name=name or f"concat{len(grids)}")
