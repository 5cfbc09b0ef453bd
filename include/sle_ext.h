/*
 * sle_ext.h -- extension entry points of libsle beyond the reference ABI.
 *
 * None of these has a counterpart in the reference's sle_api.h.  They exist
 * for callers that need FP64 values regardless of the SLE_Real width, for
 * index-based (string-free) telemetry and topology updates on large models,
 * and for parity checks of the symbolic and numeric phases against a host
 * oracle.  All values are FP64 and all indices are the insertion-order
 * indices of the elements (the order of the sle_Add* calls).
 */
#ifndef SLE_EXT_H
#define SLE_EXT_H

#include "sle_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FP64 solve result (SLE_Result carries SLE_Real). */
typedef struct SLE_ExtResult {
    int32_t status;      /* SLE_ConvergenceStatus */
    int32_t iterations;
    int32_t refactors;   /* numeric factorizations performed */
    int32_t fp64_refactors; /* mixed form: pivot-triggered FP64 refactors */
    double max_mismatch;
    double objective;
    double largest_residual; /* max |z-h|/sigma over active rows */
    int32_t largest_residual_idx;
    int32_t flags;       /* bit 0: dependency-driven factor/solve sweeps; bit 1: FP64 factor in use;
                            bit 2: mixed form on the FP64 factor (FP32 solve data) */
    double gpu_ms;       /* time of the solve: first launch to the final status on the host */
    double host_ms;      /* wall time of the call */
    int32_t pivots_substituted; /* mixed form: iterations whose FP32 factor had pivots substituted */
    int32_t pcg_steps;   /* conjugate-gradient steps per Newton step (0: direct solve) */
    int32_t graph_captures; /* iteration graphs captured during this solve (0 when replays reused them) */
    int32_t seq_fallbacks;  /* replayed iterations whose status the host waited for by a stream
                               synchronization after its spin budget (expected 0) */
    int32_t retries;        /* flat-start restarts after divergence (PRECISION; wls_solver.cu:260-334) */
    int32_t ahead;          /* iterations launched before the previous iteration's status was seen */
    int32_t ahead_idle;     /* of those, iterations that found the solve converged and changed nothing */
} SLE_ExtResult;

/* Engine options outside SLE_Config:
 *   "pcg_steps"       mixed form: conjugate-gradient steps refining each Newton
 *                     step with the FP32 factor as preconditioner (0..16, default 0:
 *                     the Gauss-Newton iteration itself refines the FP32 solve)
 *   "task_sweeps"     1 (default): walk whole bottom subtrees and the narrow
 *                     top of the supernodal tree inside single workgroups
 *                     (fewer launches); 0: one launch per level
 *   "lds_solve"       1 (default): triangular solves walk subtrees / top
 *                     chunks with the L panels resident in LDS
 *   "lds_waves_bottom", "lds_waves_tail"  waves per workgroup of the LDS
 *                     subtree tasks and top chunks (default 16 each)
 *   "pinv"            1 (default): the top of the supernodal tree is solved with
 *                     explicit blocks of L^-1 (partitioned inverse, recomputed
 *                     with every factorization) instead of level sweeps
 *   "pinv_cut"        fronts per bottom subtree under it (0: default, 32)
 *   "pinv_waves"      waves per bottom-subtree workgroup under it (default 4)
 *   "pinv_dense"      1 (default): each bottom subtree (at most 64 columns, 128
 *                     rows with its root's update rows) is solved with one
 *                     dense solve matrix [L_BB^-1; -L_UB L_BB^-1] made after
 *                     every factorization, one product per subtree and
 *                     direction; 0: LDS tasks walk the subtree's fronts
 *   "pinv_mid"        1 (default, with "pinv_dense"): the forest left above the
 *                     bottom subtrees is cut the same way once more (middle
 *                     subtrees, same limits), so that only the rest is solved
 *                     through L_TT^-1 blocks; 0: two levels
 *   "pinv_wg_fwd"     waves per middle subtree in the forward product: 1 (one
 *                     wave per subtree), 4, 8 or 16 (one workgroup per subtree,
 *                     columns split over its waves); 0 (default): 4
 *   "pinv_wg_bwd"     the same for the backward product (rows split over the
 *                     waves); 0 (default): 8 for FP32 solve data, 1 for FP64
 *   "pinv_pack"       1: the bottom subtree products run bottom subtrees of at
 *                     most 32 rows two to a wave and of at most 16 rows four
 *                     to a wave; 0: one subtree per wave; -1 (default): packed
 *                     for FP32 solve data, one per wave for FP64
 *   "queue_plain"     REALTIME iterations on the cached factor queued one ahead
 *                     of the running one: 2 (default) as plain kernel launches,
 *                     the solve's first iteration replaying its captured graph;
 *                     1: every such iteration as plain launches; 0: graph
 *                     replays only (with use_cuda_graphs)
 *   "fork_sub_m"      1 (default): a factorization forms the solve matrices of
 *                     the subtrees inside its bottom tasks on a second stream
 *                     while the top of the tree factors in one workgroup; 2:
 *                     from the end of the bottom tasks, beside the middle
 *                     levels too; 0: in order on the instance's stream
 *   "xcd_remap"       1 (default): the measurement, power, right-hand-side and
 *                     top-product kernels run contiguous block ranges per XCD
 *                     (process-wide on the device)
 *   "meas_by_type"    1 (default): the device keeps the measurement rows grouped
 *                     by evaluation class (bus values, injections, flows), each
 *                     class by anchor bus; 0: by anchor bus only
 *   "mixed_fp64_factor" 1 (default): the mixed form (use_double_precision 0)
 *                     factors G in FP64 and rounds the factor's solve data
 *                     (panels, subtree solve matrices, top inverse blocks) to
 *                     FP32 once per factorization; the iterations on the cached
 *                     factor then solve in FP32 with FP64 residuals, RHS and
 *                     state; 0: FP32 factorization (switching to an FP64 factor
 *                     when it contracts poorly)
 *   "rhs_by_location" 1 (default): iterations on a cached factor form H^T W r
 *                     from the summed weighted residuals of each bus and
 *                     branch end, gathered per state (no per-entry products);
 *                     needs "meas_by_type"; 0: per Jacobian entry
 *   "lds_flags"       0 (default): the fronts of a subtree task are solved
 *                     level by level with a workgroup barrier per level;
 *                     1: each wave waits on LDS flags of just the fronts it
 *                     depends on
 *   "huber_newton"    1 (default): Huber estimation takes IRLS gain weights
 *                     (the reference's, kernels.cu:634-671) until the steps fall
 *                     below 1e-2, then gain weights (gamma/|u|)^6 / sigma^2 for the
 *                     outliers (|u| = |r|/sigma > gamma), nearer the Huber loss's
 *                     zero curvature there yet continuous at gamma; the right-hand
 *                     side keeps the IRLS weights, so the minimiser is the same,
 *                     in fewer iterations; 0: IRLS throughout
 *   "level_schedule"  1 (default): factor and solve level by level (kernel
 *                     boundaries order the fronts); 0: one dependency-driven
 *                     launch per sweep with release/acquire flags per front */
SLE_API SLE_StatusCode SLE_CALL sle_ext_SetOption(SLE_Handle handle, const char* key, double value);

/* Autotuner of the triangular-solve plan: on the current factor (solve once
 * first), time `repetitions` forward + backward solves for each candidate
 * bottom-subtree size, workgroup width and in-task synchronisation of the
 * partitioned-inverse solve and keep the fastest, then time the factorization
 * plans (subtree and top-chunk sizes) the same way (the measurements appear in
 * sle_ext_Describe); the choice becomes the "pinv_cut" / "pinv_waves" /
 * "lds_flags" options. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_AutotuneSolve(SLE_Handle handle, int32_t repetitions);

/* Largest-normalized-residual statistics at the current estimate (after a
 * solve): rn[i] = |r_i| / sqrt(Omega_ii) and omega[i] = Omega_ii, where
 * Omega = R - H G^-1 H^T is the covariance of the WLS residuals (R =
 * diag(sigma^2), G = H^T R^-1 H at the estimate), insertion order, count
 * entries at most; 0 for inactive measurements and for critical ones (Omega_ii
 * ~ 0: their residual is structurally zero).  Omega_ii comes from the sparse
 * inverse subset of G on the supernodal fronts, computed on the GPU from a fresh
 * FP64 factorization at the estimate (which becomes the cached factor).  The
 * reference has no counterpart: it tests |r|/sigma (sle_engine.cu:959-977,
 * sle_api.cu:1295-1316).  Returns SLE_ERROR_SINGULAR_MATRIX for a singular gain.
 * sle_IdentifyBadData uses rn (option "bad_data_lnr", default 1); with
 * SLE_Config.auto_bad_data_rejection every converged solve repeats the
 * largest-normalized-residual test, setting the measurement aside and
 * re-estimating while rn exceeds bad_data_threshold (at most option
 * "max_eliminations", default 10, per solve), and SLE_Result.largest_residual /
 * largest_residual_idx / bad_data_count report the last test and the number set
 * aside (without it -- unless option "lnr_in_result" is 1 -- largest_residual is
 * max |r|/sigma and bad_data_count 0).  Each solve with auto rejection starts
 * with every measurement active again. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_NormalizedResiduals(SLE_Handle handle, double* rn, double* omega,
                                                           int32_t count);

/* Set measurement `index` (insertion order) aside (active = 0) or back in use
 * (active = 1): a value-only change (the structure is kept); the next solve
 * refactors.  No counterpart in the reference. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_SetMeasurementActive(SLE_Handle handle, int32_t index, int32_t active);

/* Solve with an explicit solver configuration; FP64 outputs. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_Solve(SLE_Handle handle, const SLE_SolverConfig* config,
                                              SLE_ExtResult* result);

/* Allocation-free variant of sle_ext_Solve.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It runs the REALTIME Gauss-Newton iterations on the handle's cached gain
 * factor exactly as sle_ext_Solve does, but makes no device allocation or
 * free: the iteration status record and the per-block partials of the
 * reductions live in a static device array of 1 MiB per GPU (enough for about
 * 12 million measurement rows), and no HIP graph is instantiated.  Every such
 * call on one GPU shares that array, so calls that overlap on one GPU (two
 * handles on two threads) race; the array keeps the last call's contents and
 * each call initializes what it reads.  Returns SLE_ERROR_INVALID_ARGUMENT,
 * before launching anything, when the request would need an allocation: a
 * mode other than SLE_MODE_REALTIME, use_robust, no cached factor (solve once
 * with sle_ext_Solve first), pending model / parameter / breaker changes, or a
 * model whose partials exceed the array. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_SolveNoAlloc(SLE_Handle handle, const SLE_SolverConfig* config,
                                                     SLE_ExtResult* result);

/* Telemetry: values[q] for measurement idx[q] (count may be any size). */
SLE_API SLE_StatusCode SLE_CALL sle_ext_UpdateTelemetryIndexed(SLE_Handle handle, const int32_t* idx,
                                                               const double* values, int32_t count);
/* Full telemetry vector in FP64, insertion order, count == measurement count. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_UpdateTelemetryF64(SLE_Handle handle, const double* values,
                                                           int32_t count);

/* State in FP64 (hot start in / estimate out). count == bus count. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_SetState(SLE_Handle handle, const double* v, const double* theta,
                                                 int32_t count);
SLE_API SLE_StatusCode SLE_CALL sle_ext_GetState(SLE_Handle handle, double* v, double* theta, int32_t count);

/* Results at the final state in FP64; any pointer may be NULL. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_GetResults(SLE_Handle handle, double* p_inj, double* q_inj,
                                                   double* p_from, double* q_from, double* p_to, double* q_to,
                                                   double* h, double* residual);

/* Device-resident telemetry batches.  Stage copies (idx, values) into HBM
 * slot `slot` ahead of time; Apply scatters the slot into the measurement
 * vector on the device (asynchronously, ordered before the next solve).  A
 * slot can be applied any number of times. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_StageTelemetry(SLE_Handle handle, int32_t slot, const int32_t* idx,
                                                       const double* values, int32_t count);
SLE_API SLE_StatusCode SLE_CALL sle_ext_ApplyStagedTelemetry(SLE_Handle handle, int32_t slot);

/* Per-kernel-class device times (HIP events on the engine's stream), recorded
 * while SLE_Config.enable_profiling is set.  Classes, in order: power
 * injections, measurement pass, gain SpGEMM, supernodal factor, right-hand
 * side, forward solve, backward solve, state update, CG refinement, and inside
 * the two solves: bottom-subtree forward product, top forward product, top
 * backward product, bottom-subtree backward product, middle-subtree forward
 * (assembly and product) and backward products; then the measurement pass
 * that opens a solve's first iteration on the cached factor, and the status
 * finish of iterations that a measurement pass closes (on the cached factor
 * with the right-hand side by location, the measurement pass class times the
 * closing pass).  ms[c] is the accumulated
 * time, launches[c] the kernel launches, calls[c] the phase runs. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_GetKernelTimes(SLE_Handle handle, double* ms, int64_t* launches,
                                                       int64_t* calls, int32_t count);
SLE_API SLE_StatusCode SLE_CALL sle_ext_ResetKernelTimes(SLE_Handle handle);

/* Sizes of the current solve plan (count entries at most, after a solve):
 * [0] bottom subtrees, [1] entries of their solve matrices (one copy),
 * [2] sum of their root update rows, [3] top supernodes, [4] top columns,
 * [5] top block entries, [6] top row entries (forward product),
 * [7] forward segments, [8] path rows of the top, [9] 1 when the right-hand
 * side is formed by location; [0]-[2] cover bottom and middle subtrees, and
 * [10]-[13] the bottom ones alone: count, entries, update rows, columns.  No
 * counterpart in the reference. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_GetSolvePlanSizes(SLE_Handle handle, int64_t* sizes, int32_t count);

/* One-line description of the engine's current execution plan (precision
 * form, factor/solve sweep schedule); valid until the next call on the thread. */
SLE_API const char* SLE_CALL sle_ext_Describe(SLE_Handle handle);

/* Breaker status of branch k (insertion index); applied before the next solve. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_SetBranchStatus(SLE_Handle handle, int32_t branch_index,
                                                        SLE_SwitchStatus status);

/* --- symbolic introspection (after sle_UploadModel) --------------------- */
/* sizes[0..11]: n_states, nnz(Ybus), nnz(H), nnz(G lower), nnz(L),
 * n_supernodes, n_levels, max front size, product terms of G, front storage
 * (elements), factor flops (sum of column counts squared), symbolic seconds*1e6 */
SLE_API SLE_StatusCode SLE_CALL sle_ext_GetSizes(SLE_Handle handle, int64_t* sizes);
SLE_API SLE_StatusCode SLE_CALL sle_ext_GetYbus(SLE_Handle handle, int32_t* row_ptr, int32_t* col, double* g,
                                                double* b);
SLE_API SLE_StatusCode SLE_CALL sle_ext_GetHPattern(SLE_Handle handle, int32_t* row_ptr, int32_t* col);
/* natural state order, lower triangle, CSC */
SLE_API SLE_StatusCode SLE_CALL sle_ext_GetGPattern(SLE_Handle handle, int32_t* col_ptr, int32_t* row);
/* bus_perm[n], perm[n_states], parent[n_states], colcount[n_states], snode_first[n_supernodes+1] */
SLE_API SLE_StatusCode SLE_CALL sle_ext_GetOrdering(SLE_Handle handle, int32_t* bus_perm, int32_t* perm,
                                                    int32_t* parent, int32_t* colcount, int32_t* snode_first);

/* --- numeric introspection at the current state ------------------------- */
/* The Gauss-Newton right-hand side b = H^T W r (WLS weights 1/sigma^2, r = z -
 * h(x)) at the current state, natural state order (n_states entries), from the
 * device kernels of the iterations on a cached factor (by_location = 1: summed
 * weighted residuals per bus and branch end, gathered per state) or of the
 * refactoring iterations (0: per Jacobian entry, summed per state).  Returns
 * SLE_ERROR_INVALID_ARGUMENT for by_location = 1 when that form is off. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_GainRhs(SLE_Handle handle, int32_t by_location, double* b);
SLE_API SLE_StatusCode SLE_CALL sle_ext_Evaluate(SLE_Handle handle, double* h, double* p_inj, double* q_inj,
                                                 double* p_from, double* q_from, double* p_to, double* q_to);
/* H values in the H pattern order and G = H^T H (unit weights) in the natural G pattern order */
SLE_API SLE_StatusCode SLE_CALL sle_ext_JacobianGain(SLE_Handle handle, double* h_values, double* g_values);
/* Factor P G P^T (current state, unit Huber weights) and return L as CSC in
 * permuted order: col_ptr[n_states+1] (int64), row[nnz(L)], val[nnz(L)].
 * Returns SLE_ERROR_SINGULAR_MATRIX when a pivot falls below threshold. */
SLE_API SLE_StatusCode SLE_CALL sle_ext_Factor(SLE_Handle handle, int64_t* col_ptr, int32_t* row, double* val);
/* Solve G dx = b with the factor of the last sle_ext_Factor (natural order). */
SLE_API SLE_StatusCode SLE_CALL sle_ext_SolveFactor(SLE_Handle handle, const double* b, double* dx);

#ifdef __cplusplus
}
#endif

#endif /* SLE_EXT_H */
