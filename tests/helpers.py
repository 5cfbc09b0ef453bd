"""Shared helpers for the parity tests (engine + oracle on the same grid)."""
from __future__ import annotations

import numpy as np

from oracle.oracle import Oracle
from state_estimation_with_cuda_amd import gridgen as gg
from state_estimation_with_cuda_amd import sle


def engine_for(g, double_precision: bool, abi_double: bool = True, **cfg):
    e = sle.Engine(double=abi_double)
    e.load_grid(g)
    e.set_config(use_double_precision=int(double_precision), **cfg)
    e.upload()
    e.set_telemetry_f64(g.z)
    return e


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def random_state(g, seed=7, spread=0.05):
    rng = np.random.default_rng(seed)
    V = g.V_true + rng.normal(0, spread, g.n)
    th = g.th_true + rng.normal(0, spread, g.n)
    th[g.slack] = 0.0
    return V, th


GRIDS = {
    "grid14": lambda: gg.grid14(0),
    "geo2k": lambda: gg.geometric(2000, seed=11, flow_meas=6000),
    "geo10k": gg.workload_10k,
}
