"""Writes the WLS / Huber golden vectors for the 14-bus workload.

Independent of the engine and of the oracle: the measurement function is the
phasor form of state_estimation_with_cuda_amd.gridgen, and the estimate is
computed by scipy.optimize.least_squares (trust-region reflective / lm on the
weighted residuals (z - h(x)) / sigma; loss='huber' with f_scale = gamma for
the Huber M-estimator, whose stationarity condition is the IRLS fixed point).

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
from scipy.optimize import least_squares

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from state_estimation_with_cuda_amd import gridgen as gg  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def solve(g, loss="linear", gamma=1.5):
    n, s = g.n, g.slack
    keep = np.array([b for b in range(n) if b != s])

    def unpack(x):
        th = np.zeros(n)
        th[keep] = x[: n - 1]
        V = x[n - 1:]
        return V, th

    def fun(x):
        V, th = unpack(x)
        return (g.z - gg.measurement_function(g, V, th)) / g.sigma

    x0 = np.concatenate([np.zeros(n - 1), np.ones(n)])
    kw = dict(xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=20000)
    if loss == "linear":
        r = least_squares(fun, x0, method="lm", **kw)
    else:
        r = least_squares(fun, x0, method="trf", loss="huber", f_scale=gamma, **kw)
    V, th = unpack(r.x)
    res = g.z - gg.measurement_function(g, V, th)
    J = float(np.sum((res / g.sigma) ** 2))
    return V, th, J, r


def grid14_with_gross_errors():
    g = gg.grid14(0)
    g.z = g.z.copy()
    g.z[[20, 57]] += 20 * g.sigma[[20, 57]]
    return g


def main():
    out = {}
    g = gg.grid14(0)
    V, th, J, r = solve(g)
    out["grid14_wls"] = dict(V=V.tolist(), theta=th.tolist(), J=J, nfev=int(r.nfev), status=int(r.status))
    gh = grid14_with_gross_errors()
    V, th, J, r = solve(gh, loss="huber", gamma=1.5)
    out["grid14_huber"] = dict(V=V.tolist(), theta=th.tolist(), J=J, gross_rows=[20, 57], gamma=1.5,
                               nfev=int(r.nfev), status=int(r.status))
    with open(os.path.join(HERE, "grid14_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "grid14_golden.json"))


if __name__ == "__main__":
    main()
