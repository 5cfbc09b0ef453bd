// instance.h -- one estimator instance resident on the GPU (no HIP types here).
//
// An Instance owns every device buffer of one network: branch parameters,
// Ybus values, state, measurements, the Jacobian values, the gain entries, the
// dense supernodal fronts of the factor and the triangular-solve vectors.  It
// runs the Gauss-Newton / Huber loop (wls_solver.cu:221-431 semantics) with all
// arithmetic on the device; the host only reads one small status record per
// iteration to test convergence.
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "symbolic.h"

namespace sle {

// Host-side branch / bus / measurement arrays in internal indices (FP64).
struct HostArrays {
    int n = 0, nb = 0, m = 0, slack = 0;
    std::vector<int32_t> bus_type;
    std::vector<double> vset;
    std::vector<int32_t> from, to;
    std::vector<double> gs, bs, bc, tap, shift;  // series g, b; total charging; tap; phase shift
    std::vector<uint8_t> status;   // 1 = closed
    std::vector<uint8_t> present;  // 1 = branch in the structure (not removed)
    std::vector<int32_t> mtype, mloc, mend;
    std::vector<uint8_t> active;     // 1 = measurement used now (its branch closed)
    std::vector<uint8_t> m_present;  // 1 = measurement in the structure (not removed)
    std::vector<uint8_t> rejected;   // 1 = set aside as bad data (inactive; empty = none)
    std::vector<double> sigma, z;
};

struct SolveParams {
    int mode = 1;          // 0 REALTIME (reuse a valid factor), 1 PRECISION (refactor every iteration)
    int max_iter = 100;
    double tol = 1e-4;
    int robust = 0;        // 0 WLS, 1 Huber IRLS (the reference's weights), 2 Huber with Newton gain weights
    double gamma = 1.5;
    int flat_start = 0;
    double vmin = 0.5, vmax = 1.5;
    double time_limit_ms = 0;  // REALTIME only, 0 = none
    int retries = 2;           // flat-start retries after divergence (not in REALTIME)
    int pcg_steps = 0;         // >0: conjugate-gradient refinement of each Newton step (factor as preconditioner)
};

struct SolveOutcome {
    int status = 5;  // SLE_ConvergenceStatus
    int iterations = 0;
    int refactors = 0;
    int fp64_refactors = 0;  // mixed form without CG: pivots that fell below threshold
    int pivots_substituted = 0;  // mixed form with CG: iterations whose FP32 factor had substituted pivots
    double max_mismatch = 0;
    double objective = 0;
    double largest_residual = 0;  // max |r|/sigma over active rows
    int largest_residual_idx = -1;
    double ms = 0;  // wall time from the solve's start to its final status
    int flags = 0;  // bit 0: dependency-driven sweeps; bit 1: FP64 factor active (mixed form fell back)
    int newton_fallbacks = 0;  // Huber-Newton gains that lost rank (IRLS used instead)
    int graph_captures = 0;    // iteration graphs captured and instantiated during the solve
    int seq_fallbacks = 0;     // replayed iterations whose status the host did not see within the spin budget
    int retries = 0;           // flat-start restarts after divergence (PRECISION only)
    int ahead = 0;             // iterations launched before the previous iteration's status was seen
    int ahead_idle = 0;        // of those, iterations queued behind a converged one (they change nothing)
};

// Device bytes held by an instance, and a count of its allocations and
// releases: a captured launch graph holds buffer addresses, so one captured
// before the instance's latest (re)allocation is retired rather than trusted.
// Other instances' allocations do not touch it.
struct Ledger {
    size_t bytes = 0;
    uint64_t epoch = 0;
    bool frozen = false;  // set during an allocation-free solve: any (re)allocation throws NoAllocError
};

// Kernel classes timed with HIP events on the instance stream (profiling mode).
// PH_FWD / PH_BWD span a whole solve direction; PH_SUB_* and PH_PI_* time the
// bottom-subtree and top products inside them (partitioned-inverse solve)
enum Phase : int {
    PH_POWER = 0,
    PH_MEAS,
    PH_SPGEMM,
    PH_FACTOR,
    PH_RHS,
    PH_FWD,
    PH_BWD,
    PH_UPDATE,
    PH_PCG,
    PH_SUB_FWD,
    PH_PI_FWD,
    PH_PI_BWD,
    PH_SUB_BWD,
    PH_MID_FWD,
    PH_MID_BWD,
    PH_MEAS_LEAD,  // the measurement pass that opens the first iteration of a solve on the cached factor
    PH_FINISH,     // the iteration's status (max |dx|, J) when a measurement pass closes the iteration
    PH_COUNT
};

struct KernelTimes {
    double ms[PH_COUNT] = {0};       // accumulated device time between the phase's first launch and last completion
    int64_t launches[PH_COUNT] = {0};  // kernel launches issued in the phase
    int64_t calls[PH_COUNT] = {0};     // times the phase ran
};

// Raised for any HIP runtime failure (mapped to SLE_ERROR_CUDA_FAILURE).
struct GpuError : std::runtime_error {
    explicit GpuError(const std::string& s) : std::runtime_error(s) {}
};

// Raised when the allocation-free solve cannot run without allocating (no
// cached factor, a mode that refactors, a model whose reduction partials
// exceed the scratch array); mapped to SLE_ERROR_INVALID_ARGUMENT.  Thrown
// before any device allocation is made.
struct NoAllocError : std::runtime_error {
    explicit NoAllocError(const std::string& s) : std::runtime_error(s) {}
};

class Instance {
   public:
    struct Impl;
    // precision: 1 = plain FP64 form, 0 = mixed (FP32 H/G/factor, FP64 residual/RHS/state)
    Instance(int device, int precision);
    ~Instance();
    Instance(const Instance&) = delete;
    Instance& operator=(const Instance&) = delete;

    // Allocate and upload everything for this symbolic analysis.
    void setup(const HostArrays& h, std::shared_ptr<const Symbolic> S);
    // Re-upload branch parameters + status and rebuild Ybus values (pattern unchanged).
    // Branch parameters and breaker status (values only: the structure holds open
    // branches too); refreshes Ybus and the measurement activity.
    void update_branches(const HostArrays& h);
    // Breaker operations without a full upload: stage any number of them, then
    // refresh_topology() uploads status and activity and recomputes Ybus.
    void set_branch_closed(int k, bool closed);
    void refresh_topology();
    // Bad-data rejection of measurement i (insertion order): a value-only
    // change of its activity, like a breaker operation; refresh_topology()
    // applies it (and invalidates the factor).
    void set_measurement_rejected(int i, bool rejected);
    // Leverages s_i = h_i G^-1 h_i^T / sigma_i^2 and residuals r_i = z_i - h_i(x)
    // of every measurement (insertion order) at the current state, with the
    // WLS gain at that state: the gain is refactored in FP64, the sparse inverse
    // subset Z = G^-1 is formed front by front from the root down, and each
    // row's leverage comes from the one front holding all of its columns.  The
    // factor left behind is the WLS gain's at this state (a cached factor for
    // REALTIME iterations).  Returns false when the gain is singular.
    bool leverages(double* lev, double* r);
    void set_z(const double* z);
    void set_z_indexed(const int32_t* idx, const double* val, int count);
    void set_state(const double* V, const double* th);
    void get_state(double* V, double* th);
    void flat_start();
    void invalidate_factor() { factor_valid_ = false; }
    bool factor_valid() const { return factor_valid_; }

    SolveOutcome solve(const SolveParams& p);
    // The same REALTIME iterations on the cached factor without any device
    // allocation: the status record and the reductions' per-block partials
    // live in the device's static scratch array (kScratchBytes), shared by all
    // instances on the device, so two such calls that overlap on one device
    // race.  No captured graph is used (instantiating one allocates).  Throws
    // NoAllocError, before launching anything, when the request needs an
    // allocation: a mode other than REALTIME WLS, no valid cached factor, or
    // partials larger than the scratch.
    SolveOutcome solve_no_alloc(const SolveParams& p);

    // Results of the last solve at the final state (downloaded on demand).
    void get_results(double* P, double* Q, double* Pf, double* Qf, double* Pt, double* Qt, double* hx, double* r);

    // Introspection for parity tests: evaluate at the current state.
    void debug_eval(double* hx, double* P, double* Q, double* Pf, double* Qf, double* Pt, double* Qt);
    void debug_ybus(double* G, double* B);
    // H values (unscaled, natural pattern order) and G entries (symbolic order),
    // computed with unit Huber weights, in the instance precision.
    void debug_jacobian_gain(double* hv, double* gv);
    // Factor the current G and return the dense fronts (symbolic order).
    int debug_factor(double* fronts);
    // Solve G dx = b with the current factor (b, dx in natural state order).
    void debug_solve(const double* b, double* dx);
    // The Gauss-Newton right-hand side H^T W r at the current state (WLS
    // weights, natural state order) from the device kernels of either form:
    // by location (k_meas_loc + k_rhs_loc) or per Jacobian entry (k_meas +
    // k_rhs_sum).  Returns false when the by-location form is not set up.
    bool debug_rhs(bool by_location, double* b);

    size_t device_bytes() const { return bytes_.bytes; }
    // while on, any device (re)allocation of this instance throws NoAllocError
    void freeze_allocations(bool on) { bytes_.frozen = on; }
    int precision() const { return precision_; }
    const Symbolic& symbolic() const { return *S_; }
    KernelTimes& kernel_times() { return ktimes_; }
    void reset_kernel_times() { ktimes_ = KernelTimes(); }
    void set_profiling(bool on) { profiling_ = on; }
    bool profiling() const { return profiling_; }
    void set_use_graphs(bool on) { use_graphs_ = on; }
    // iterations on the cached factor that close with the measurement pass:
    // launched as plain kernels (queued one ahead like graph replays) instead
    // of replaying a captured graph; 0: none, 1: all, 2: those queued ahead
    void set_queue_plain(int mode) { queue_plain_ = mode; }
    // dependency-driven sweeps (default when every front fits in LDS) or level-by-level launches
    void set_level_schedule(bool on);
    bool dataflow_sweeps() const;
    // task sweeps (bottom subtrees per workgroup, narrow top in one workgroup)
    void set_task_sweeps(bool on);
    // factorization plan: bottom subtrees of at most `cut` fronts per workgroup,
    // top levels of at most `wcut` fronts in the single-workgroup tail (0: the
    // cost model's choice; takes effect at setup)
    void set_factor_plan(int cut, int wcut);
    // LDS-resident task triangular solves
    void set_lds_solve(bool on);
    // partitioned-inverse solve of the top of the tree; cut = fronts per bottom
    // subtree (0: default), waves = waves per bottom-subtree workgroup (cut and
    // waves take effect at setup; -1 keeps the current value)
    void set_partitioned_inverse(bool on, int cut, int waves);
    // waves per workgroup of the bottom-subtree tasks and of the top chunks (takes effect at setup)
    void set_lds_waves(int bottom, int tail);
    // LDS-resident task solves: each wave waits on per-front LDS flags of the
    // fronts it depends on (true) or the workgroup meets at a barrier per level
    void set_lds_flags(bool on);
    // partitioned-inverse solve: bottom subtrees as one dense solve matrix each
    // (true) or as LDS tasks walking their fronts (takes effect at setup)
    void set_dense_subtrees(bool on);
    // the bottom subtrees' solve matrices formed on a second stream while the
    // rest of the tree factors: 0 in order, 1 (default) beside the top of the
    // tree, 2 beside the middle levels and the top
    void set_fork_subtree_matrices(int mode);
    // with dense subtrees: a second level of them inside the forest above the
    // bottom ones, leaving a smaller top to the partitioned inverse (setup)
    void set_middle_subtrees(bool on);
    // middle subtree products: waves per subtree, forward and backward (1: one
    // wave each; 4, 8 or 16: one workgroup each; 0: by precision)
    void set_subtree_workgroups(int fwd, int bwd);
    // bottom subtree products: subtrees of at most 32 (16) rows two (four) to a
    // wave (1), one per wave (0), or by precision (-1: packed for FP32 solve data)
    void set_subtree_packing(int mode);
    // XCD-aware block order of the gather kernels (process-wide on this device)
    void set_xcd_remap(bool on);
    // device order of the measurement rows: grouped by evaluation class, each
    // class by anchor bus (true), or by anchor bus only (takes effect at setup)
    void set_meas_by_type(bool on);
    // iterations on a cached factor form the right-hand side by location
    // (summed weighted residuals per bus and branch end, gathered per state)
    // instead of per Jacobian entry (takes effect at setup; needs the class order)
    void set_rhs_by_location(bool on);
    // Mixed form (precision 0): refactor in FP64 and round the solve data to
    // FP32 for the iterations on the cached factor (on), or factor in FP32 (off).
    void set_mixed_fp64_factor(bool on);
    // sizes of the current solve plan: [0] bottom subtrees, [1] entries of their
    // solve matrices (one copy), [2] sum of their root update rows, [3] top
    // supernodes, [4] top columns, [5] top block entries, [6] top row entries
    // (forward product), [7] forward segments, [8] path rows of the top
    // (sum of P_s), [9] 1 if the right-hand side is formed by location, and of
    // the bottom subtrees alone: [10] count, [11] entries, [12] update rows,
    // [13] columns
    void solve_plan_sizes(int64_t* out, int count) const;
    bool lds_flags() const;
    std::string sweep_plan() const;
    // Autotuner of the solve plan: times the forward + backward solves of the
    // current factor for each candidate (bottom-subtree size x waves per
    // workgroup) and keeps the fastest; returns what it measured.
    std::string autotune_solve(int reps);
    int partitioned_inverse_cut() const;
    int partitioned_inverse_waves() const;

    // Device-resident telemetry batches: stage (host -> HBM) ahead of time,
    // apply (scatter into z on the device) right before a solve.
    void stage_telemetry(int slot, const int32_t* idx, const double* val, int count);
    void apply_staged(int slot);
    int staged_count(int slot) const;

   private:
    Impl* d_ = nullptr;
    std::shared_ptr<const Symbolic> S_;
    int device_ = 0;
    int precision_ = 0;
    bool factor_valid_ = false;
    bool fp64_factor_active_ = false;  // mixed form fell back to FP64 after a pivot trigger
    bool profiling_ = false;
    bool use_graphs_ = false;
    int queue_plain_ = 0;
    Ledger bytes_;
    KernelTimes ktimes_;
};

// Looks up the address of the allocation-free solve's static scratch array on
// `device` (once per device; called at engine initialization).
void init_device_scratch(int device);
// Size of that array in bytes (per device).
size_t device_scratch_bytes();

// Device helpers exposed to the engine.
int device_count();
int runtime_version();
const char* device_name(int dev);

}  // namespace sle
