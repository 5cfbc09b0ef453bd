#!/usr/bin/env python3
"""WLS state estimation on one MI355X: the 200k-bus incremental-update metric,
with the Huber, batched-area and serving workloads beside it.

Headline (`value`): the 200,000-bus / ~270,000-branch / 2,000,000-measurement
seeded geometric grid (1 % gross errors), resident on the GPU.  One step is
one incremental update: 10 % of the measurement values are replaced (a
telemetry batch already staged in HBM is scattered into z on the device) and a
REALTIME Gauss-Newton solve runs to ||dx||_inf < 1e-6 on the cached gain
factor.  Two precision forms run the same sequence of updates: plain FP64, and
the mixed form (FP64 h(x), residual, Jacobian, right-hand side, state and
factorization; the cached factor's solve data rounded to FP32 once per
factorization, so the iterations on it solve in FP32).  `value` is the mixed
form's rate when it is the faster (config.form names the form), with p50/p99
step latency; both forms are reported in full, with their errors against the
FP64 CPU oracle, a roofline entry for the dominant kernel (HIP-event timed) and
the oracle's CPU baseline.  Each form also reports `host_path`: the same
updates pushed from host memory (sle_ext_UpdateTelemetryIndexed through pinned
staging) with the estimate read back to the host inside the timed step.

Beside the headline, unless --no-extras:
  huber200k   flat-start Huber solve (gamma 1.5 sigma, to 1e-8) of the same
              grid, both forms: time, iterations, supernode-size histogram,
              error vs the oracle (warm-started at the GPU estimate and
              iterated to 1e-10: the Huber objective is convex, with one
              minimiser);
  lnr200k     the largest-normalized-residual test at the estimate: time of
              one pass (Omega_ii of all 2M rows), largest rn, rows above 3;
  areas256    256 independent areas (500-20,000 buses) batched in one model,
              each re-solved from its previous solution + N(0, 0.01), against
              one engine per area;
  serve200k   the serving loop on the resident 200k model, both forms: Poisson
              updates at 200/s, a breaker toggle per 50 updates, new 5k-50k-bus
              areas at 1/s within a 64 GiB budget; p50/p99/max latency from
              arrival to the estimate on the host, and the error of a sampled
              update against the oracle.

--workload areas / serve run one of those alone with its own options.
--copy-calibration runs a plain device copy of known size first; under
`rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` it calibrates the counters that
give the HBM traffic per launch recorded in profiles/hbm_traffic_200k.md.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "p99 update latency + solves/s, 200k-bus WLS, mixed-prec vs FP64 at 1e-6"
HBM_PEAK_GBS = 8000.0
TOL = 1e-6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_200k.md")
AREA_MAX_ITER = 25  # a batch iterates until its slowest area converges
HUBER_GAMMA = 1.5
HUBER_TOL = 1e-8  # the Huber iteration contracts linearly: a 1e-6 step still leaves ~1e-6 to the minimiser
AREA_MODE = 1  # PRECISION: a perturbed start refactors (the cached factor of the old point may not contract)
COPY_BYTES = 512 << 20


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_updates(g, count, frac, seed):
    """Requests: each replaces a random 10 % of the base telemetry with fresh noisy
    readings.  Returned as staged batches that also restore the previous
    request's entries, so that applying them in order leaves z = base + request."""
    rng = np.random.default_rng(seed)
    k = int(round(frac * g.m))
    reqs = []
    for _ in range(count):
        idx = np.sort(rng.choice(g.m, size=k, replace=False)).astype(np.int32)
        val = g.z[idx] + rng.normal(0.0, 1.0, k) * g.sigma[idx]
        reqs.append((idx, val))
    batches = []
    prev = np.zeros(0, np.int32)
    for idx, val in reqs:
        restore = np.setdiff1d(prev, idx, assume_unique=True)
        bi = np.concatenate([restore, idx]).astype(np.int32)
        bv = np.concatenate([g.z[restore], val])
        batches.append((bi, bv))
        prev = idx
    return reqs, batches


def request_z(g, req):
    z = g.z.copy()
    z[req[0]] = req[1]
    return z


def algorithmic_bytes(sz, g, prec_bytes, direct=True, plan=None):
    """Compulsory HBM bytes per launch of each kernel class (each element once),
    for an iteration that reuses the factor.  prec_bytes is the factor/solve
    vector width (4 mixed, 8 plain FP64).  direct: no CG refinement.  plan: the
    engine's solve-plan sizes (sle_ext_GetSolvePlanSizes): the products of the
    partitioned-inverse solve and, when the right-hand side is formed by
    location, the bytes of those two kernels."""
    n, nb, m, ns = g.n, g.nb, g.m, sz["ns"]
    nnz_y, nnz_h = sz["nnz_y"], sz["nnz_h"]
    s = prec_bytes
    plan = plan or {}
    ybus = (n + 1) * 4 + nnz_y * (4 + 8 * 4)  # yrow, ycol, yG, yB, Pij, Qij
    jac = nnz_h * (4 + 8) if direct else nnz_h * 8
    out = {
        # V, th read; P, Q write
        "power": n * 8 * 4 + ybus,
        # per row: code (1 B), loc, z, sigma read, scaled residual write, hrow;
        # Jacobian products / values; bus V/th/P/Q, Ybus rows and branch data once
        "meas": m * (1 + 4 + 8 + 8 + 8) + (m + 1) * 4 + jac + n * 8 * 4 + ybus + nb * (4 + 4 + 8 * 5),
        # tptr, tout, the products, b write
        "rhs": (ns + 1) * 4 + ns * 4 + nnz_h * 8 + ns * s,
        # update: dx read, V/theta read+write, pth/pv
        "update": ns * s + n * (8 * 4 + 8),
    }
    if plan.get("rhs_by_location") and direct:
        # k_meas_loc (with the injections): Ybus rows; bus V/th read, P/Q
        # written; per row code, z, sigma; row ranges; branch data; per-bus and
        # per-branch sums written
        out["meas"] = (n + 1) * 4 + nnz_y * (4 + 8 + 8) + n * 8 * 4 + m * (1 + 8 + 8) \
            + 2 * (n + 1) * 4 + (2 * nb + 1) * 4 + nb * (4 + 4 + 8 * 5) + 32 * n + 32 * nb
        out["power"] = 0  # fused into the measurement pass on this path
        # k_rhs_loc: Ybus rows with the transposed admittances, bus V/th/P/Q
        # and sums, incident branches and their sums, state maps, b
        out["rhs"] = (n + 1) * 4 + nnz_y * (4 + 8 + 8) + n * 8 * 4 + 32 * n + (n + 1) * 4 \
            + 2 * nb * 4 + 32 * nb + 8 * n + ns * s
    if plan.get("subtrees"):
        def sub(ent, us, nsub, nc):
            # M_t once; column map; right-hand side in, solution / update rows out; headers
            fwd = s * ent + 4 * nc + s * nc + s * nc + s * us + 28 * nsub
            bwd = s * ent + 4 * nc + s * (nc + us) + 4 * us + s * nc + 28 * nsub
            return fwd, bwd
        nc_all = ns - plan["top_columns"]  # columns of all dense subtrees
        out["sub_fwd"], out["sub_bwd"] = sub(plan["bottom_entries"], plan["bottom_update_rows"],
                                             plan["bottom_subtrees"], plan["bottom_columns"])
        nsub1 = plan["subtrees"] - plan["bottom_subtrees"]
        if nsub1:
            nc1 = nc_all - plan["bottom_columns"]
            f1, b1 = sub(plan["subtree_entries"] - plan["bottom_entries"],
                         plan["subtree_update_rows"] - plan["bottom_update_rows"], nsub1, nc1)
            out["mid_fwd"] = f1 + nc1 * (2 * s + 8)  # with the in-place assembly of the middle columns
            out["mid_bwd"] = b1
    if plan.get("top_columns"):
        ntc, nq, nseg = plan["top_columns"], plan["top_row_entries"], plan["segments"]
        # row-grouped L_TT^-1 entries (each row one contiguous range of c_T), c_T
        # once, y_T out, segment headers
        out["pi_fwd"] = s * nq + 2 * s * ntc + 36 * nseg
        # block entries, path rows once, y_T once, per-column metadata, x_T out
        out["pi_bwd"] = s * plan["top_block_entries"] + 4 * plan["top_path_rows"] + 2 * s * ntc + 24 * ntc
    return out


def measured_traffic(form, kernel):
    """HBM bytes per launch of a kernel class of one form from the committed
    counter summary (lines `traffic <form>.<class> <bytes>`), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for ln in f:
                parts = ln.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == f"{form}.{kernel}":
                    return int(parts[2])
    except OSError:
        pass
    return None


def copy_calibration():
    """A plain device-to-device copy of COPY_BYTES (read COPY_BYTES, write COPY_BYTES)."""
    import torch

    a = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    del a, b
    log(f"copy calibration: 3 x {COPY_BYTES} B read + write")


def roofline_for(form, kt, ab):
    cand = {k: v for k, v in (kt or {}).items() if k in ab and v["launches"]}
    if not cand:
        return None
    dom = max(cand, key=lambda k: cand[k]["ms"])
    # the byte models are per run of a class (one launch, but for the middle
    # forward class two: assembly and product)
    avg_ms = cand[dom]["ms"] / max(cand[dom]["calls"], 1)
    traffic = measured_traffic(form, dom)
    if traffic is None and dom in ("meas", "power"):  # FP64 kernels in every form
        traffic = measured_traffic("fp64", dom)
    achieved = ab[dom] / (avg_ms * 1e-3) / 1e9
    return dict(kernel=dom, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, bytes_per_launch=int(ab[dom]),
                avg_launch_ms=round(avg_ms, 5))


def engine_options(args) -> dict:
    """--set KEY=VALUE pairs (sle_ext_SetOption keys, or use_cuda_graphs) applied to every engine."""
    out = {}
    for kv in args.set or []:
        k, _, v = kv.partition("=")
        out[k.strip()] = float(v)
    return out


CONFIG_KEYS = ("use_cuda_graphs",)  # --set keys that are SLE_Config fields rather than engine options


def configure(e, fp64, pcg_steps, options):
    e.set_config(use_double_precision=int(fp64), enable_profiling=0,
                 **{k: int(v) for k, v in (options or {}).items() if k in CONFIG_KEYS})
    e.set_option("pcg_steps", 0 if fp64 else pcg_steps)
    for k, v in (options or {}).items():
        if k not in CONFIG_KEYS:
            e.set_option(k, v)


def supernode_histogram(e):
    """Supernode sizes (columns per supernode) of the engine's symbolic factor."""
    sizes = np.diff(e.ordering()["snode_ptr"])
    edges = [1, 2, 3, 5, 9, 17, 33, 65]
    top = max(int(sizes.max()) + 1, edges[-1] + 1)
    counts = np.histogram(sizes, bins=edges + [top])[0]
    return {f"{lo}-{hi - 1}": int(c) for lo, hi, c in zip(edges, edges[1:] + [top], counts) if c}


def run_form(g, fp64, reqs, batches, warmup, steps, profile_steps, pcg_steps=0, options=None, autotune=False,
             huber=False):
    from state_estimation_with_cuda_amd import sle

    name = "fp64" if fp64 else "mixed"
    t0 = time.perf_counter()
    e = sle.Engine(double=True)
    e.load_grid(g)
    configure(e, fp64, pcg_steps, options)
    e.upload()
    e.set_telemetry_f64(g.z)
    setup_s = time.perf_counter() - t0
    sz = e.sizes()
    log(f"{name}: model loaded + symbolic in {setup_s:.1f}s; sizes {sz}")
    log(f"{name}: plan: {e.describe()}")
    t0 = time.perf_counter()
    r0 = e.solve(mode=1, max_iterations=30, tolerance=TOL, flat_start=True)
    init_ms = (time.perf_counter() - t0) * 1e3
    log(f"{name}: initial flat-start PRECISION solve {init_ms:.1f} ms, iters={r0['iterations']} status={r0['status']}")
    tune = ""
    if autotune:
        t0 = time.perf_counter()
        tune = e.autotune(3)
        log(f"{name}: autotune {time.perf_counter() - t0:.1f}s: {tune}")
    for k in range(warmup + steps):
        e.stage_telemetry(k, *batches[k])
    lat = []
    iters = []
    flags = set()
    ok = 0

    def step(k):
        e.apply_staged(k)
        return e.solve(mode=0, max_iterations=10, tolerance=TOL)

    for k in range(warmup):
        step(k)
    t_start = time.perf_counter()
    for k in range(warmup, warmup + steps):
        ts = time.perf_counter()
        r = step(k)
        lat.append((time.perf_counter() - ts) * 1e3)
        iters.append(r["iterations"])
        flags.add(r["flags"])
        ok += r["status"] == 0
    total = time.perf_counter() - t_start
    # the same kind of update from host memory, the estimate read back inside the step
    hlat, h_ok = [], 0
    V, th = np.empty(g.n), np.empty(g.n)  # the client's result arrays, reused across requests
    # one untimed pass first (the pinned staging and read-back buffers are made on first use);
    # it re-applies the last staged request, so z is unchanged
    e.set_telemetry_indexed(*batches[warmup + steps - 1])
    e.solve(mode=0, max_iterations=10, tolerance=TOL)
    e.state(out=(V, th))
    t_start = time.perf_counter()
    for k in range(warmup + steps, warmup + 2 * steps):
        ts = time.perf_counter()
        e.set_telemetry_indexed(*batches[k])
        r = e.solve(mode=0, max_iterations=10, tolerance=TOL)
        e.state(out=(V, th))
        hlat.append((time.perf_counter() - ts) * 1e3)
        h_ok += r["status"] == 0
    h_total = time.perf_counter() - t_start
    objective = r["objective"]
    # largest-normalized-residual test at this estimate: the first call builds the
    # row-to-front plan on the host, the second is the steady-state cost
    lnr = None
    if huber:
        e.normalized_residuals()
        t0 = time.perf_counter()
        rn, _ = e.normalized_residuals()
        lnr = dict(ms=round((time.perf_counter() - t0) * 1e3, 2), max_rn=round(float(rn.max()), 3),
                   above_3=int(np.sum(rn > 3.0)), above_6=int(np.sum(rn > 6.0)),
                   gross_errors_planted=int(round(0.01 * g.m)))
        log(f"{name}: normalized residuals {lnr}")
    e_plan = e.describe()
    plan_sizes = e.solve_plan_sizes()
    last_z = request_z(g, reqs[warmup + 2 * steps - 1])
    # profiled pass (separate, so the events do not perturb the timed loop)
    kt = None
    if profile_steps:
        e.set_config(enable_profiling=1)
        e.reset_kernel_times()
        for k in range(warmup, warmup + profile_steps):
            step(k)
        kt = e.kernel_times()
        e.set_config(enable_profiling=0)
    hub = None
    if huber:
        e.set_telemetry_f64(g.z)
        t0 = time.perf_counter()
        rh = e.solve(mode=1, max_iterations=150, tolerance=HUBER_TOL, robust=True, gamma=HUBER_GAMMA,
                     flat_start=True)
        wall = (time.perf_counter() - t0) * 1e3
        hV, hth = e.state()
        hub = dict(tolerance=HUBER_TOL, status=rh["status"], iterations=rh["iterations"], refactors=rh["refactors"],
                   gpu_ms=round(rh["gpu_ms"], 2), wall_ms=round(wall, 2), objective=rh["objective"],
                   supernodes=supernode_histogram(e), V=hV, th=hth)
        log(f"{name}: Huber flat start {wall:.1f} ms, iters={rh['iterations']} status={rh['status']}")
    e.close()
    lat = np.array(lat)
    hlat = np.array(hlat)
    return dict(
        name=name,
        solves_per_s=steps / total,
        ms_per_step=total / steps * 1e3,
        p50_ms=float(np.percentile(lat, 50)),
        p99_ms=float(np.percentile(lat, 99)),
        max_ms=float(lat.max()),
        host_path=dict(solves_per_s=round(steps / h_total, 2), p50_ms=round(float(np.percentile(hlat, 50)), 4),
                       p99_ms=round(float(np.percentile(hlat, 99)), 4), max_ms=round(float(hlat.max()), 4),
                       converged_steps=h_ok),
        mean_iterations=float(np.mean(iters)),
        converged=ok,
        initial_solve_ms=init_ms,
        initial_iterations=r0["iterations"],
        setup_s=setup_s,
        V=V,
        th=th,
        last_z=last_z,
        objective=objective,
        kernel_times=kt,
        sizes=sz,
        plan_sizes=plan_sizes,
        flags=sorted(flags),
        plan=e_plan,
        autotune=tune,
        huber=hub,
        lnr=lnr,
    )


def areas_result(args, form, steps, warmup, per_instance_steps):
    """Batched uneven areas: `--areas` independent area estimators (bus counts
    log-uniform in [500, 20000], redundancy ~3) solved as one model -- every
    kernel launch covers all areas -- against one engine per area solved one
    after the other.  A step re-solves every area from its previous solution
    perturbed by N(0, 0.01) per state element (PRECISION: refactor every iteration, tol 1e-6).
    Every iteration of the batch -- gain product, factorization, solve data,
    solves and update for all areas -- is one replay of a captured HIP graph,
    and the batch's solve and factorization plans are the autotuner's."""
    from state_estimation_with_cuda_amd import gridgen, sle

    t0 = time.perf_counter()
    sizes = gridgen.area_sizes(args.areas, seed=3)
    areas = [gridgen.workload_area(int(n), seed=1000 + k) for k, n in enumerate(sizes)]
    g = gridgen.concat(areas, name=f"areas{args.areas}")
    log(f"areas: {len(areas)} areas, n={g.n} nb={g.nb} m={g.m}, generated in {time.perf_counter() - t0:.1f}s")
    fp64 = form == "fp64"
    offs = np.cumsum([0] + [a.n for a in areas])

    def engine(grid):
        e = sle.Engine(double=True)
        e.load_grid(grid)
        configure(e, fp64, args.pcg_steps, engine_options(args))
        e.upload()
        e.set_telemetry_f64(grid.z)
        return e

    # batched: one model
    t0 = time.perf_counter()
    eb = engine(g)
    r = eb.solve(mode=1, max_iterations=30, tolerance=TOL, flat_start=True)
    V0, th0 = eb.state()
    log(f"batched: setup + first solve {time.perf_counter() - t0:.1f}s, status {r['status']} iters {r['iterations']}")
    # tile and supernode-block sizes of the batch chosen by the autotuner (timed on the device)
    t0 = time.perf_counter()
    tune = eb.autotune(3) if not args.no_autotune_areas else ""
    if tune:
        log(f"batched: autotune {time.perf_counter() - t0:.1f}s: {tune}")
    starts = [gridgen.hot_start(g, V0, th0, 0.01, seed=50 + k) for k in range(warmup + steps)]
    for k in range(warmup):
        eb.set_state(*starts[k])
        eb.solve(mode=AREA_MODE, max_iterations=AREA_MAX_ITER, tolerance=TOL)
    tb, iters, ok = 0.0, [], 0
    for k in range(warmup, warmup + steps):
        eb.set_state(*starts[k])  # host upload of the perturbed start, outside the timed solve
        ts = time.perf_counter()
        r = eb.solve(mode=AREA_MODE, max_iterations=AREA_MAX_ITER, tolerance=TOL)
        tb += time.perf_counter() - ts
        iters.append(r["iterations"])
        ok += r["status"] == 0
    Vb, thb = eb.state()
    plan = eb.describe()
    eb.close()
    batched_rate = len(areas) * steps / tb
    # per-instance: one engine per area, solved one after the other
    t0 = time.perf_counter()
    engines = [engine(a) for a in areas]
    for e in engines:
        e.solve(mode=1, max_iterations=30, tolerance=TOL, flat_start=True)
    log(f"per-area engines: setup + first solves {time.perf_counter() - t0:.1f}s")
    tp, p_iters, p_ok = 0.0, [], 0
    worst = {"V": 0.0, "theta": 0.0}
    ksteps = max(1, min(steps, per_instance_steps))
    for k in range(warmup + steps - ksteps, warmup + steps):  # ends on the batch's last start
        V, th = starts[k]
        for a, e, b0 in zip(areas, engines, offs):
            e.set_state(V[b0:b0 + a.n], th[b0:b0 + a.n])
            ts = time.perf_counter()
            ra = e.solve(mode=AREA_MODE, max_iterations=AREA_MAX_ITER, tolerance=TOL)
            tp += time.perf_counter() - ts
            p_iters.append(ra["iterations"])
            p_ok += ra["status"] == 0
    # the same last start solved per area vs inside the batch (tolerance 1e-6)
    for a, e, b0 in zip(areas, engines, offs):
        Va, tha = e.state()
        worst["V"] = max(worst["V"], float(np.abs(Va - Vb[b0:b0 + a.n]).max()))
        worst["theta"] = max(worst["theta"], float(np.abs(tha - thb[b0:b0 + a.n]).max()))
    for e in engines:
        e.close()
    per_rate = len(areas) * ksteps / tp
    out = {
        "metric": "area solves/s, batched uneven areas (one launch sequence for all areas)",
        "value": round(batched_rate, 1),
        "unit": "area-solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(tb / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if fp64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{len(areas)} independent areas, n log-uniform [500, 20000] (sum {g.n}), m {g.m}; "
                               "hot start = previous solution + N(0, 0.01); PRECISION to 1e-6",
                   "form": form},
        "per_instance_solves_per_s": round(per_rate, 1),
        "per_instance_mean_iterations": float(np.mean(p_iters)),
        "per_instance_converged": f"{p_ok}/{len(p_iters)}",
        "batched_vs_per_instance": round(batched_rate / per_rate, 2),
        "mean_iterations": float(np.mean(iters)),
        "converged_steps": ok,
        "plan": plan,
        "autotune": tune,
        "batched_vs_alone_state_diff": worst,
    }
    return out


def run_areas(args):
    print(json.dumps(areas_result(args, args.form, args.steps, args.warmup, 3)), flush=True)


def serve_result(args, form, duration, g=None, oracle_cache=None):
    """Serving loop on the resident 200k-bus model: Poisson arrivals of
    incremental updates (10 % of z re-measured), a breaker operation after every
    50th update (alternately opening and closing a non-bridge branch), and full
    solves of new 5k-50k-bus areas admitted by an HBM budget; latency is
    arrival to the estimate on the host."""
    from state_estimation_with_cuda_amd import gridgen, serve, sle

    sys.setswitchinterval(1e-4)  # the request threads hand the interpreter over quickly
    rng = np.random.default_rng(4)
    g = g if g is not None else gridgen.workload_200k()
    fp64 = form == "fp64"

    def engine(grid):
        e = sle.Engine(double=True)
        e.load_grid(grid)
        configure(e, fp64, args.pcg_steps, engine_options(args))
        e.upload()
        e.set_telemetry_f64(grid.z)
        return e

    main_e = engine(g)
    r = main_e.solve(mode=1, max_iterations=30, tolerance=TOL, flat_start=True)
    log(f"serve: resident model ready, first solve status {r['status']}")
    # areas: a pool of new area models, log-uniform 5k-50k buses
    pool_n = np.exp(rng.uniform(np.log(5_000), np.log(50_000), args.area_pool)).astype(int)
    pool = [gridgen.workload_area(int(n), seed=7000 + k) for k, n in enumerate(pool_n)]
    cal = engine(pool[0])
    bytes_per_bus = 1.25 * cal.gpu_bytes() / pool[0].n
    cal.close()
    log(f"serve: {len(pool)} area models, {bytes_per_bus:.0f} B/bus estimate")
    # request stream
    wl = serve.make_workload(g, rng, args.rate, duration, toggle_every=args.toggle_every,
                             area_rate=args.area_rate, area_pool=pool, record_every=args.record_every,
                             replays=args.replays)
    reqs, replay, status = wl.requests, wl.replay, wl.breakers
    srv = serve.Server(main_e, engine, hbm_budget=int(args.hbm_budget_gib * (1 << 30)), area_slots=args.area_slots,
                       bytes_per_bus=bytes_per_bus, tol=TOL, max_iterations=10,
                       toggle_mode=args.toggle_mode)
    t0 = time.perf_counter()
    done = serve.run(srv, reqs)
    wall = time.perf_counter() - t0
    summ = serve.summarize(done, warmup=args.serve_warmup)
    bad = sorted(r.arrival for r in done if r.kind == serve.UPDATE and r.status != 0)
    summ["first_unconverged_update_s"] = round(bad[0], 3) if bad else None
    slow = sorted((r for r in done if r.kind != serve.AREA and r.done is not None), key=lambda r: -r.latency)[:8]
    summ["long_service"] = [{"kind": r.kind, "rid": r.rid, "arrival_s": round(r.arrival, 4),
                             "service_ms": round((r.done - r.start) * 1e3, 2), "iterations": r.iterations,
                             "status": r.status} for r in done
                            if r.kind != serve.AREA and r.done is not None and r.done - r.start > 0.05][:10]
    worst = max((r for r in done if r.kind != serve.AREA and r.done is not None), key=lambda r: r.latency, default=None)
    if worst is not None:  # the neighbourhood of the slowest request, for diagnosis
        near = sorted((r for r in done if r.kind != serve.AREA and abs(r.arrival - worst.arrival) < 0.5),
                      key=lambda r: r.arrival)
        summ["around_slowest"] = [[r.kind[0], r.rid, round(r.arrival, 4), round(r.submitted, 4), round(r.start, 4),
                                   round(r.done, 4)] for r in near]
    summ["slowest"] = [{"kind": r.kind, "rid": r.rid, "arrival_s": round(r.arrival, 4),
                        "wait_ms": round((r.start - r.arrival) * 1e3, 2), "service_ms": round((r.done - r.start) * 1e3, 2),
                        "iterations": r.iterations} for r in slow]
    ok = [r for r in done if r.error is None and r.status == 0 and r.arrival >= args.serve_warmup]
    rate = len(ok) / max(wall - args.serve_warmup, 1e-9)
    # isolation: the same request alone gives the same result
    iso = {"update": 0.0, "area": 0.0}
    for r in done:
        if r.kind == serve.UPDATE and r.rid in replay and r.result is not None and r.error is None:
            zr, st = replay[r.rid]
            for kb, sv in status.items():  # breakers back to their state at that request
                main_e.set_branch_status(kb, st.get(kb, 1))
            main_e.set_telemetry_f64(zr)
            main_e.set_state(*r.before[:2])
            main_e.solve(mode=0, max_iterations=10, tolerance=TOL)
            V, th = main_e.state()
            iso["update"] = max(iso["update"], float(np.abs(V - r.result[0]).max()),
                                float(np.abs(th - r.result[1]).max()))
    for r in done:
        if r.kind == serve.AREA and r.record and r.result is not None:
            e = engine(r.payload)
            e.solve(mode=1, max_iterations=50, tolerance=TOL, flat_start=True)
            V, th = e.state()
            e.close()
            iso["area"] = max(iso["area"], float(np.abs(V - r.result[0]).max()), float(np.abs(th - r.result[1]).max()))
    err = None
    if not args.no_cpu_baseline and replay:
        from oracle.oracle import Oracle

        q0 = min(replay)
        r0 = [r for r in done if r.kind == serve.UPDATE and r.rid == q0][0]
        zr, st = replay[q0]
        key = (q0, tuple(sorted(st.items())), float(zr.sum()))
        if oracle_cache is None or key not in oracle_cache:
            g2 = gridgen.Grid(**{**g.__dict__})
            g2.z = zr
            g2.status = g.status.copy()
            for kb, sv in st.items():
                g2.status[kb] = sv
            o = Oracle(g2)
            o.solve(mode=1, max_iter=30, tol=1e-10, flat_start=True)
            ref = o.state()
            if oracle_cache is not None:
                oracle_cache[key] = ref
        else:
            ref = oracle_cache[key]
        oV, oth = ref
        if r0.result is not None:
            err = {"request": q0, "breakers": {str(k): v for k, v in st.items()},
                   "V": float(np.abs(r0.result[0] - oV).max()), "theta": float(np.abs(r0.result[1] - oth).max())}
    main_e.close()
    u = summ["update"]
    out = {
        "metric": "serving: p99 incremental-update latency (arrival to result) on the resident 200k-bus model",
        "value": round(u["p99_ms"], 3) if u["p99_ms"] is not None else None,
        "unit": "ms",
        "n_gpus": 1,
        "steps": len(done),
        "warmup": args.serve_warmup,
        "ms_per_step": round(u["p50_ms"], 3) if u["p50_ms"] is not None else None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if fp64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"Poisson {args.rate}/s updates (10% of z) for {duration}s, breaker toggle per "
                               f"{args.toggle_every} updates, new 5k-50k-bus areas at {args.area_rate}/s within "
                               f"{args.hbm_budget_gib} GiB",
                   "form": form, "toggle_mode": "REALTIME" if args.toggle_mode == 0 else "PRECISION"},
        "latency": summ,
        "solves_per_s": round(rate, 1),
        "hbm_peak_gib": round(srv.peak_used / (1 << 30), 2),
        "estimate_misses": srv.estimate_misses,
        "isolation_max_diff": iso,
        "error_vs_oracle": err,
    }
    return out


def run_serve(args):
    print(json.dumps(serve_result(args, args.form, args.duration)), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--buses", type=int, default=200_000, help="grid size (200000 is the metric's workload)")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=20, help="incremental updates timed on the CPU oracle")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true", help="time the mixed form only")
    ap.add_argument("--skip-mixed", action="store_true", help="time the plain FP64 form only")
    ap.add_argument("--pcg-steps", type=int, default=0, help="mixed form: CG steps per Newton step (engine default 0)")
    ap.add_argument("--copy-calibration", action="store_true", help="run a plain device copy first")
    ap.add_argument("--workload", choices=["update200k", "areas", "serve"], default="update200k",
                    help="update200k: the headline metric; areas: batched uneven areas; serve: serving loop")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: update arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--serve-warmup", type=float, default=5.0, help="serve: seconds excluded from the statistics")
    ap.add_argument("--area-rate", type=float, default=1.0, help="serve: new areas per second")
    ap.add_argument("--area-pool", type=int, default=6)
    ap.add_argument("--area-slots", type=int, default=3,
                    help="serve: concurrent area engines (with the resident model at most 4 streams, one hardware queue each)")
    ap.add_argument("--hbm-budget-gib", type=float, default=64.0)
    ap.add_argument("--record-every", type=int, default=500)
    ap.add_argument("--toggle-every", type=int, default=50, help="serve: updates between breaker operations (0: none)")
    ap.add_argument("--toggle-mode", type=int, default=0, choices=[0, 1],
                    help="serve: solve mode after a breaker operation (1 PRECISION, 0 REALTIME)")
    ap.add_argument("--replays", type=int, default=4)
    ap.add_argument("--areas", type=int, default=256)
    ap.add_argument("--form", choices=["fp64", "mixed"], default="fp64", help="areas workload: precision form")
    ap.add_argument("--autotune", action="store_true", help="autotune the solve and factorization plans first")
    ap.add_argument("--no-autotune-areas", action="store_true",
                    help="batched areas: keep the default plans instead of the autotuner's")
    ap.add_argument("--set", action="append", metavar="KEY=VALUE", help="engine option (sle_ext_SetOption), repeatable")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only: no Huber, batched-area or serving legs in the default line")
    ap.add_argument("--extra-serve-duration", type=float, default=20.0,
                    help="default line: seconds of arrivals of each form's serving leg")
    ap.add_argument("--extra-serve-warmup", type=float, default=2.0,
                    help="default line: seconds of each serving leg excluded from its statistics")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs (--gpus > 1) are not supported", file=sys.stderr)
        sys.exit(2)

    from state_estimation_with_cuda_amd import gridgen

    if args.copy_calibration:
        copy_calibration()
    if args.workload == "areas":
        run_areas(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return
    t0 = time.perf_counter()
    if args.buses == 200_000:
        g = gridgen.workload_200k()
    else:
        g = gridgen.geometric(args.buses, seed=2, flow_meas="all", dup_meas=int(1.6 * args.buses), gross_frac=0.01)
    log(f"grid: n={g.n} nb={g.nb} m={g.m} generated in {time.perf_counter() - t0:.1f}s")
    # staged-HBM steps, then as many host-path steps on the following requests
    reqs, batches = make_updates(g, args.warmup + max(2 * args.steps, args.profile_steps), 0.10, seed=4)

    opts = engine_options(args)
    extras = not args.no_extras
    mixed = None if args.skip_mixed else run_form(g, False, reqs, batches, args.warmup, args.steps,
                                                  args.profile_steps, args.pcg_steps, opts, args.autotune, extras)
    fp64 = None if args.skip_fp64 else run_form(g, True, reqs, batches, args.warmup, args.steps, args.profile_steps,
                                                0, opts, args.autotune, extras)
    forms = [f for f in (mixed, fp64) if f]
    best = max(forms, key=lambda f: f["solves_per_s"])
    sz = best["sizes"]

    def summary(f):
        kt = f["kernel_times"] or {}
        # the mixed form solves with an FP64 factor once it switched to one (flag 2)
        wide = f["name"] == "fp64" or any(fl & 2 for fl in f["flags"])
        ab = algorithmic_bytes(f["sizes"], g, 8 if wide else 4, direct=f["name"] == "fp64" or args.pcg_steps == 0,
                               plan=f["plan_sizes"])
        return {
            "solves_per_s": round(f["solves_per_s"], 2),
            "ms_per_step": round(f["ms_per_step"], 4),
            "p50_ms": round(f["p50_ms"], 4),
            "p99_ms": round(f["p99_ms"], 4),
            "max_ms": round(f["max_ms"], 4),
            "mean_iterations": f["mean_iterations"],
            "converged_steps": f["converged"],
            "initial_solve_ms": round(f["initial_solve_ms"], 2),
            "initial_iterations": f["initial_iterations"],
            "solve_flags": f["flags"],
            "host_path": f["host_path"],
            "plan": f["plan"],
            "plan_sizes": f["plan_sizes"],
            "autotune": f["autotune"],
            "kernel_ms_per_step": {k: round(v["ms"] / max(args.profile_steps, 1), 4) for k, v in kt.items()},
            "kernel_launches_per_step": {k: round(v["launches"] / max(args.profile_steps, 1), 2)
                                         for k, v in kt.items()},
            # the mixed form on the FP64 factor runs the FP64 kernels: their traffic
            "roofline": roofline_for("fp64" if wide else f["name"], kt, ab),
        }

    out = {
        "metric": METRIC,
        "value": round(best["solves_per_s"], 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best["name"] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "incremental WLS update of a 200k-bus grid: 10% of z replaced (staged in HBM), REALTIME "
                        "Gauss-Newton to ||dx||inf<1e-6 with the cached gain factor (host_path: telemetry from "
                        "host memory and the estimate read back, inside the step)",
            "form": best["name"],
            "buses": g.n,
            "branches": g.nb,
            "measurements": g.m,
            "tolerance": TOL,
            "options": opts,
        },
        "p50_ms": round(best["p50_ms"], 4),
        "p99_ms": round(best["p99_ms"], 4),
        "symbolic": {k: sz[k] for k in ("ns", "nnz_h", "nnz_g", "nnz_l", "n_snodes", "n_levels", "max_front")},
    }
    for f in forms:
        out[f["name"]] = summary(f)
    if fp64 and mixed:
        out["mixed_vs_fp64"] = round(mixed["solves_per_s"] / fp64["solves_per_s"], 3)
        out["mixed_vs_fp64_state_diff"] = {
            "V": float(np.abs(mixed["V"] - fp64["V"]).max()),
            "theta": float(np.abs(mixed["th"] - fp64["th"]).max()),
        }
    out["roofline"] = out[best["name"]]["roofline"]

    if not args.no_cpu_baseline:
        from oracle.oracle import Oracle, cpu_model, num_threads, set_num_threads

        t0 = time.perf_counter()
        o = Oracle(g)
        o.solve(mode=1, max_iter=30, tol=TOL, flat_start=True)
        log(f"oracle: setup + initial solve {time.perf_counter() - t0:.1f}s")
        cores = num_threads()
        times = []
        for k in range(args.cpu_sample):
            o.set_z(request_z(g, reqs[args.warmup + args.steps - args.cpu_sample + k]))
            r = o.solve(mode=0, max_iter=10, tol=TOL)
            times.append(r["seconds"])
        cpu_rate = len(times) / sum(times)
        # the same updates on one core (a smaller sample)
        set_num_threads(1)
        times1 = []
        for k in range(min(3, args.cpu_sample)):
            o.set_z(request_z(g, reqs[args.warmup + args.steps - args.cpu_sample + k]))
            r = o.solve(mode=0, max_iter=10, tol=TOL)
            times1.append(r["seconds"])
        set_num_threads(cores)
        # reference optimum for the last request, then the GPU errors against it
        o.set_z(best["last_z"])
        o.solve(mode=1, max_iter=30, tol=1e-10)
        oV, oth = o.state()
        oJ = o.solve(mode=1, max_iter=1, tol=1.0)["objective"]
        err = {}
        for f in forms:
            err[f["name"]] = {"V": float(np.abs(f["V"] - oV).max()), "theta": float(np.abs(f["th"] - oth).max()),
                              "J_rel": abs(f["objective"] - oJ) / abs(oJ)}
        out["error_vs_oracle"] = err
        out["oracle_objective"] = oJ
        out["cpu_baseline"] = {
            "value": round(cpu_rate, 4),
            "unit": "solves/s",
            "cores": cores,
            "kind": "port",
            "sample": f"{len(times)} incremental updates (10% of z replaced, REALTIME WLS with cached factor, "
                      f"tol {TOL}) of the same 200k-bus model on the FP64 oracle: {sum(times):.1f}s",
            "cpu_model": cpu_model(),
            "one_core_solves_per_s": round(len(times1) / sum(times1), 4),
            "one_core_sample": f"{len(times1)} of those updates on 1 thread: {sum(times1):.1f}s",
        }
        # Huber: the oracle iterated from each form's estimate to 1e-10 (one minimiser)
        for f in forms:
            h = f["huber"]
            if not h:
                continue
            o.set_z(g.z)
            o.set_state(h["V"], h["th"])
            t0 = time.perf_counter()
            ro = o.solve(mode=1, max_iter=60, tol=1e-10, robust=2, gamma=HUBER_GAMMA)
            hV, hth = o.state()
            h["error_vs_oracle"] = {"V": float(np.abs(h["V"] - hV).max()), "theta": float(np.abs(h["th"] - hth).max()),
                                    "J_rel": abs(h["objective"] - ro["objective"]) / abs(ro["objective"]),
                                    "oracle_iterations": ro["iterations"], "oracle_status": ro["status"]}
            log(f"oracle: Huber check of {f['name']} in {time.perf_counter() - t0:.1f}s, {ro['iterations']} iterations")
        del o
    for f in forms:
        h = f["huber"]
        if h:
            out[f["name"]]["huber200k"] = {k: v for k, v in h.items() if k not in ("V", "th")}
        if f["lnr"]:
            out[f["name"]]["lnr200k"] = f["lnr"]
    if extras:
        # the serving loop on the resident 200k model, both forms (the same request stream)
        cache = {}
        sargs = argparse.Namespace(**{**vars(args), "serve_warmup": args.extra_serve_warmup, "replays": 1})
        serving = {}
        for f in forms:
            t0 = time.perf_counter()
            res = serve_result(sargs, f["name"], args.extra_serve_duration, g=g, oracle_cache=cache)
            log(f"serve {f['name']}: {time.perf_counter() - t0:.1f}s")
            serving[f["name"]] = {"update": res["latency"]["update"], "toggle": res["latency"]["toggle"],
                                  "area": res["latency"]["area"], "solves_per_s": res["solves_per_s"],
                                  "hbm_peak_gib": res["hbm_peak_gib"], "isolation_max_diff": res["isolation_max_diff"],
                                  "error_vs_oracle": res["error_vs_oracle"]}
        out["serve200k"] = {"workload": res["config"]["workload"], "latency": "arrival to the estimate on the host",
                            **serving}
        # batched uneven areas (one model) against one engine per area
        t0 = time.perf_counter()
        ares = areas_result(args, "fp64", 5, 2, 1)
        log(f"areas: {time.perf_counter() - t0:.1f}s")
        out["areas256"] = {k: ares[k] for k in ("value", "unit", "ms_per_step", "config", "per_instance_solves_per_s",
                                                "batched_vs_per_instance", "mean_iterations", "converged_steps",
                                                "batched_vs_alone_state_diff")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
