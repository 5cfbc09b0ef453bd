"""Host symbolic analysis of the library (no GPU needed) against the oracle.

The library's symbolic phase (Ybus / H / G patterns, minimum-degree ordering,
elimination tree, column counts, fundamental supernodes) is an independent
implementation from the oracle's; both must agree bit-exactly.  The oracle's
own symbolic results are pinned by brute force in test_oracle_pins.py.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from state_estimation_with_cuda_amd import gridgen as gg
from state_estimation_with_cuda_amd import sle


def host_engine(g):
    e = sle.Engine(double=True, initialize=False)
    e.load_grid(g)
    return e


CASES = {
    "grid14": lambda: gg.grid14(0),
    "grid14_seed5": lambda: gg.grid14(5),
    "geo500_partial_flows": lambda: gg.geometric(500, seed=4, flow_meas=700),
    "geo3000": lambda: gg.geometric(3000, seed=9, flow_meas="all", dup_meas=500),
    # independent areas in one model: one angle reference per island
    "areas3": lambda: gg.concat([gg.geometric(300, seed=21, flow_meas=600), gg.grid14(2),
                                 gg.geometric(500, seed=22, flow_meas="all")]),
}


@pytest.mark.parametrize("name", list(CASES))
def test_patterns_and_ordering_match_oracle(name):
    g = CASES[name]()
    e = host_engine(g)
    o = Oracle(g)
    rp, col = e.h_pattern()
    orp, ocol = o.h_pattern()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    gc, gr = e.g_pattern()
    ogc, ogr = o.g_pattern()
    assert np.array_equal(gc, ogc) and np.array_equal(gr, ogr)
    es, os_ = e.ordering(), o.symbolic()
    for k in ("bus_perm", "perm", "parent", "colcount", "snode_ptr"):
        assert np.array_equal(es[k], os_[k]), k
    sz = e.sizes()
    assert sz["ns"] == g.ns == 2 * g.n - int(np.count_nonzero(g.bus_type == gg.BUS_SLACK))
    assert sz["nnz_l"] == o.nnz_l
    assert sz["n_snodes"] == o.n_snodes
    # Ybus pattern (values need the device)
    yrp, ycol = e.ybus(values=False)
    orp_y, ocol_y, _, _ = o.ybus()
    assert np.array_equal(yrp, orp_y) and np.array_equal(ycol, ocol_y)
    e.close()


def test_open_branch_keeps_structure_like_oracle():
    """Opening a breaker inside an island changes values only: the patterns,
    ordering and supernodes stay those of the closed network (the open branch
    contributes zero admittance, its flow rows zero weight), in the library and
    in the oracle alike."""
    g = gg.geometric(400, seed=8, flow_meas="all")
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    k = None
    for cand in range(g.nb):
        keep = np.ones(g.nb, bool)
        keep[cand] = False
        A = sp.coo_matrix((np.ones(keep.sum()), (g.br_from[keep], g.br_to[keep])), shape=(g.n, g.n))
        if cg.connected_components(A, directed=False)[0] == 1:
            k = cand
            break
    assert k is not None
    e = host_engine(g)
    before = (e.h_pattern(), e.g_pattern(), e.ordering())
    e.set_branch_status(k, 0)
    o = Oracle(g)
    o.set_branch_status(k, 0)
    rp, col = e.h_pattern()
    orp, ocol = o.h_pattern()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(rp, before[0][0]) and np.array_equal(col, before[0][1])
    gc, gr = e.g_pattern()
    assert np.array_equal(gc, before[1][0]) and np.array_equal(gr, before[1][1])
    es, os_ = e.ordering(), o.symbolic()
    for key in ("bus_perm", "perm", "parent", "colcount", "snode_ptr"):
        assert np.array_equal(es[key], os_[key]), key
        assert np.array_equal(es[key], before[2][key]), key
    e.close()


def test_splitting_an_island_moves_the_angle_reference_like_oracle():
    """Two areas joined by one tie branch form one island (one angle reference,
    the SLACK bus of higher index); opening the tie splits them and each keeps
    its own reference: one state less, and a new symbolic structure."""
    a, b = gg.geometric(200, seed=51, flow_meas="all"), gg.geometric(150, seed=52, flow_meas="all")
    g = gg.concat([a, b])
    # tie branch between bus 0 of area b and bus 1 of area a
    g.br_from = np.append(g.br_from, a.n).astype(np.int32)
    g.br_to = np.append(g.br_to, 1).astype(np.int32)
    for f, v in (("r", 0.02), ("x", 0.1), ("b", 0.0), ("tap", 1.0), ("shift", 0.0)):
        setattr(g, f, np.append(getattr(g, f), v))
    g.is_xfmr = np.append(g.is_xfmr, 0).astype(np.int32)
    g.status = np.append(g.status, 1).astype(np.uint8)
    e = host_engine(g)
    o = Oracle(g)
    assert e.sizes()["ns"] == 2 * g.n - 1
    tie = g.nb - 1
    e.set_branch_status(tie, 0)
    o.set_branch_status(tie, 0)
    assert e.sizes()["ns"] == 2 * g.n - 2 == o.ns
    es, os_ = e.ordering(), o.symbolic()
    for key in ("bus_perm", "perm", "parent", "colcount", "snode_ptr"):
        assert np.array_equal(es[key], os_[key]), key
    rp, col = e.h_pattern()
    orp, ocol = o.h_pattern()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    e.close()


def test_symbolic_sizes_200k_scale_structure():
    """Structure of a 20k-bus mesh: fronts stay small under minimum degree."""
    g = gg.geometric(20_000, seed=2, flow_meas="all")
    e = host_engine(g)
    sz = e.sizes()
    assert sz["ns"] == 2 * g.n - 1
    assert sz["max_front"] <= 64
    assert sz["nnz_l"] < 3 * sz["nnz_g"]
    e.close()


def test_breaker_sequences_track_islands_like_oracle():
    """A run of breaker operations, each checked against the oracle's full
    island analysis: splitting and re-joining two areas (one angle reference
    each, then one for both), opening and closing a branch inside an island
    (same structure), and cutting off a bus without a SLACK bus (new islands,
    same references)."""
    a, b = gg.geometric(200, seed=51, flow_meas="all"), gg.geometric(150, seed=52, flow_meas="all")
    g = gg.concat([a, b])
    g.br_from = np.append(g.br_from, a.n).astype(np.int32)
    g.br_to = np.append(g.br_to, 1).astype(np.int32)
    for f, v in (("r", 0.02), ("x", 0.1), ("b", 0.0), ("tap", 1.0), ("shift", 0.0)):
        setattr(g, f, np.append(getattr(g, f), v))
    g.is_xfmr = np.append(g.is_xfmr, 0).astype(np.int32)
    g.status = np.append(g.status, 1).astype(np.uint8)
    tie = g.nb - 1
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    def parts(skip):
        keep = np.ones(g.nb, bool)
        keep[list(skip)] = False
        A = sp.coo_matrix((np.ones(keep.sum()), (g.br_from[keep], g.br_to[keep])), shape=(g.n, g.n))
        return cg.connected_components(A, directed=False)[0]

    inner = next(k for k in range(a.nb) if parts([k]) == 1)  # a branch of area a on a cycle
    deg = np.bincount(np.concatenate([g.br_from, g.br_to]), minlength=g.n)
    leaf = next((k for k in range(a.nb) if (deg[g.br_from[k]] == 1 or deg[g.br_to[k]] == 1)
                 and g.bus_type[g.br_from[k] if deg[g.br_from[k]] == 1 else g.br_to[k]] != 2), None)
    ops = [(tie, 0), (tie, 1), (inner, 0), (inner, 1), (tie, 0), (inner, 0), (tie, 1), (inner, 1)]
    if leaf is not None:
        ops += [(leaf, 0), (inner, 0), (leaf, 1), (inner, 1)]
    e = host_engine(g)
    o = Oracle(g)
    for k, st in ops:
        e.set_branch_status(k, st)
        o.set_branch_status(k, st)
        assert e.sizes()["ns"] == o.ns, (k, st)
        es, os_ = e.ordering(), o.symbolic()
        for key in ("perm", "parent", "snode_ptr"):
            assert np.array_equal(es[key], os_[key]), (k, st, key)
    e.close()
