// engine.h -- host engine behind the sle_* C-ABI (FP64 internally).
//
// Replaces sle::SLEEngine (sle_engine.h:288-1081, sle_engine.cu:55-1773) and
// sle::NetworkModel (network_model.h:79-668, network_model.cu:47-734).  The
// engine keeps the host network model and id maps, runs the symbolic analysis
// when the topology changes, and drives one GPU Instance.  Status codes and
// enum values are those of sle_api.h (as plain ints here, so this file does not
// depend on the width of SLE_Real).
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "instance.h"
#include "symbolic.h"

namespace sle {

enum Status : int {
    ST_OK = 0,
    ST_INVALID_HANDLE = -1,
    ST_NOT_INITIALIZED = -2,
    ST_GPU_FAILURE = -3,
    ST_INVALID_ARGUMENT = -4,
    ST_MODEL_INCOMPLETE = -5,
    ST_NOT_OBSERVABLE = -6,
    ST_SINGULAR = -7,
    ST_DIVERGED = -8,
    ST_MAX_ITERATIONS = -9,
    ST_OUT_OF_MEMORY = -10,
    ST_NOT_FOUND = -11,
    ST_DUPLICATE = -12,
    ST_CONFIG_FILE = -13,
    ST_UNKNOWN = -99
};

// sle_engine.config keys (sle_config.h:81-113) with the reference defaults
// (sle_engine.config:33-73).
struct EngineConfig {
    int device_id = 0;
    size_t max_buses = 200000, max_branches = 400000, max_measurements = 2000000, max_switching_devices = 100000;
    bool enable_robust = false;
    double huber_gamma = 1.5;
    double convergence_tolerance = 1e-4;
    int max_realtime_iterations = 5;
    int max_precision_iterations = 100;
    double realtime_time_limit_ms = 20.0;
    bool use_flat_start_default = false;
    bool enable_profiling = false;
    bool use_cuda_graphs = true;
    int block_size_standard = 256;
    int block_size_reduction = 512;
    bool use_double_precision = false;
    double voltage_min_pu = 0.5, voltage_max_pu = 1.5;
    double bad_data_threshold = 3.0;
    bool auto_bad_data_rejection = false;
};

bool load_config_file(const std::string& path, EngineConfig& cfg, std::string& err);
bool save_config_file(const std::string& path, const EngineConfig& cfg, std::string& err);

struct SolverConfigIn {
    int mode = 0;
    int max_iterations = 5;
    double tolerance = 1e-4;
    double huber_gamma = 1.5;
    bool use_robust = false;
    bool use_flat_start = false;
    double time_limit_ms = 0;
};

struct ResultOut {
    int status = 5;
    int iterations = 0;
    double max_mismatch = 0, objective = 0, ms = 0, largest_residual = 0;
    int largest_residual_idx = -1, bad_data_count = 0;
    int refactors = 0, fp64_refactors = 0, flags = 0, pivots_substituted = 0, pcg_steps = 0;
    int graph_captures = 0, seq_fallbacks = 0, retries = 0;
    int ahead = 0, ahead_idle = 0;
};

// Engine options beyond SLE_Config (sle_ext_SetOption).
struct EngineOptions {
    int pcg_steps = 0;         // mixed form: CG steps refining each Newton step (0 = FP32 factor solve, GN refines)
    bool level_schedule = true;   // level-by-level sweeps (0: single dependency-driven launch per sweep)
    bool task_sweeps = true;      // group bottom subtrees / the narrow top into single-workgroup tasks
    bool lds_solve = true;        // LDS-resident task triangular solves
    int lds_waves_bottom = 16, lds_waves_tail = 16;
    bool pinv = true;             // partitioned-inverse solve of the top of the supernodal tree
    bool huber_newton = true;     // Huber: reduced outlier gain weights near the solution (IRLS before)
    int pinv_cut = 0;             // fronts per bottom subtree under it (0: default)
    int pinv_waves = 4;           // waves per bottom-subtree workgroup under it
    bool lds_flags = false;       // task solves: per-front LDS flags instead of per-level barriers
    bool pinv_dense = true;       // partitioned inverse: bottom subtrees as dense solve matrices
    bool pinv_mid = true;         // and a middle level of dense subtrees under a smaller top
    int pinv_wg_fwd = 0;          // middle subtree products: waves per subtree, forward (1: one wave
    int pinv_wg_bwd = 0;          // each; 4, 8, 16: a workgroup each; 0: by precision) and backward
    int pinv_pack = -1;           // bottom subtree products: small subtrees two or four to a wave
                                  // (1), one a wave (0), by precision (-1: packed for FP32 solve data)
    int fork_sub_m = 1;           // bottom subtree matrices overlap the top (1) or the middle and top (2)
    int queue_plain = 2;          // iterations on the cached factor as plain launches, not graphs
                                  // (0: none, 1: all, 2: those queued ahead)
    bool xcd_remap = true;        // XCD-aware block order of the gather kernels
    bool meas_by_type = true;     // measurement rows grouped by evaluation class on the device
    bool rhs_by_location = true;  // cached-factor iterations: RHS from per-location sums
    bool mixed_fp64_factor = true;  // mixed form: FP64 factorization, FP32 solve data
    bool bad_data_lnr = true;     // bad data by normalized residuals |r|/sqrt(Omega_ii) (0: |r|/sigma, the reference's)
    bool lnr_in_result = false;   // every solve reports the largest normalized residual (an extra refactor)
    int max_eliminations = 10;    // auto_bad_data_rejection: measurements set aside per solve at most
    int factor_cut = 0, factor_wcut = 0;  // factorization plan (0: the cost model's)
};

class Engine {
   public:
    Engine();
    ~Engine();

    // lifecycle / config
    int initialize();
    bool initialized() const { return initialized_; }
    EngineConfig& config() { return cfg_; }
    int set_option(const char* key, double value);
    const EngineOptions& options() const { return opt_; }
    void set_config(const EngineConfig& c);
    std::string config_path;

    // model building (returns Status)
    int add_bus(const char* id, double base_kv, int type, double vset, double p, double q);
    int add_branch(const char* id, const char* from, const char* to, double r, double x, double b, double g,
                   double tap, double shift, int is_xfmr, const char* sd_id);
    int add_measurement(const char* id, int type, const char* loc, int end, double sigma, double pt, double ct,
                        int pseudo);
    int add_switch(const char* id, const char* branch_id, int status);
    int remove_bus(const char* id);
    int remove_branch(const char* id);
    int update_branch(const char* id, double r, double x, double b, double g, double tap, double shift, int is_xfmr);
    int set_tap(const char* id, double tap);
    int set_phase_shift(const char* id, double shift);
    int set_impedance(const char* id, double r, double x, double b, double g);
    bool get_tap(const char* id, double& tap) const;
    bool get_phase_shift(const char* id, double& s) const;

    // meters (network_model.cu:312-455 semantics: channels map to measurements)
    int add_meter(const char* id, int type, const char* bus_id, const char* branch_id, int end, double pt, double ct,
                  double sv, double sp, double si);
    int meter_reading(const char* meter, const char* channel, double* value, int which) const;  // 0 z, 1 est, 2 res
    int update_meter_reading(const char* meter, const char* channel, double value);

    // data
    int upload_model();
    int update_telemetry(const double* values, int count);
    int update_telemetry_indexed(const int32_t* idx, const double* values, int count);
    int update_measurement(const char* id, double value);
    int sync_measurements();
    // device-resident telemetry batches (see Instance::stage_telemetry)
    int stage_telemetry(int slot, const int32_t* idx, const double* values, int count);
    int apply_staged(int slot);
    void flush_staged_host();

    // switches
    int set_switch_status(const char* sd_id, int status);
    int get_switch_status(const char* sd_id) const;
    int apply_switch_changes();
    int apply_topology_values();
    int set_branch_status_index(int k, int status);

    // solve
    int solve(const SolverConfigIn& sc, ResultOut& out);
    // The same REALTIME solve without any device allocation (Instance::solve_no_alloc);
    // throws NoAllocError when the request or the engine's state needs one.
    int solve_no_alloc(const SolverConfigIn& sc, ResultOut& out);
    SolverConfigIn mode_config(int mode) const;
    int check_observability(int& observable, int& islands);
    int apply_flat_start();

    // results
    const std::vector<double>& V() {
        ensure_state();
        return V_;
    }
    const std::vector<double>& th() {
        ensure_state();
        return th_;
    }
    void ensure_state();  // download the estimate after a solve (on first use)
    // the estimate into caller arrays of count entries (fetched straight from
    // the device after a solve when the caller takes all of it)
    void state_out(double* v, double* th, int count);
    int set_state(const double* V, const double* th, int count);
    void ensure_results();
    std::vector<double> P_, Q_, Pf_, Qf_, Pt_, Qt_, h_, r_;
    double objective() const { return last_.objective; }
    double last_solve_ms() const { return last_.ms; }
    const ResultOut& last_result() const { return last_; }
    std::vector<int32_t> identify_bad_data(double threshold);
    // Normalized residuals rn_i = |r_i| / sqrt(Omega_ii) and Omega_ii at the
    // current state (insertion order; 0 for inactive and critical rows).
    int normalized_residuals(std::vector<double>& rn, std::vector<double>& omega);
    // Set a measurement aside as bad data (inactive, structure unchanged) or back.
    // apply = false: staged, applied by the next refresh of the device topology
    int set_measurement_rejected(int i, bool rejected, bool apply = true);

    // lookups
    int bus_index(const char* id) const;
    int branch_index(const char* id) const;
    int meas_index(const char* id) const;
    int switch_index(const char* id) const;
    int n_bus() const { return (int)bus_id_.size(); }
    int n_branch() const { return (int)br_id_.size(); }
    int n_meas() const { return (int)m_id_.size(); }
    int n_meter() const { return (int)meter_id_.size(); }
    size_t gpu_bytes() const;

    // measurement metadata (for sle_GetMeasurementDetails)
    const std::string& meas_id(int i) const { return m_id_[i]; }
    int meas_type(int i) const { return m_type_[i]; }
    int meas_loc(int i) const { return m_loc_[i]; }
    int meas_end(int i) const { return m_end_[i]; }
    double meas_sigma(int i) const { return m_sigma_[i]; }
    double meas_pt(int i) const { return m_pt_[i]; }
    double meas_ct(int i) const { return m_ct_[i]; }
    int meas_pseudo(int i) const { return m_pseudo_[i]; }
    double meas_z(int i) {
        flush_staged_host();
        return m_z_[i];
    }
    bool meas_active(int i) const;
    const std::string& bus_id(int i) const { return bus_id_[i]; }
    const std::string& branch_id(int i) const { return br_id_[i]; }

    // introspection for parity tests
    Instance* instance() { return inst_.get(); }
    // Autotune the solve plan on the current factor (solves once first if needed).
    int autotune(int reps, std::string& report);
    const std::string& autotune_report() const { return tune_report_; }
    std::shared_ptr<const Symbolic> symbolic() const { return sym_; }
    int ensure_ready();     // upload/rebuild as needed without solving
    int ensure_symbolic();  // host symbolic analysis only (no device needed)
    void load_ieee14();

   private:
    SolveParams solve_params(const SolverConfigIn& sc) const;
    void take_outcome(const SolveOutcome& o, const SolveParams& p, ResultOut& out);
    void lnr_result(const SolverConfigIn& sc, ResultOut& out);
    void build_host_arrays(HostArrays& h) const;
    void admittance(int k, double& g, double& b) const;
    int rebuild();
    // Breaker operations since the last island analysis that keep every island
    // (a closed branch inside one island, an opened branch whose ends stay
    // connected): then the angle references, and with them the state layout,
    // are unchanged without a pass over the whole network.
    bool islands_kept(const uint8_t* status);
    bool ends_connected(int a, int b, const uint8_t* status);
    void set_islands(const ModelView& mv);

    EngineConfig cfg_;
    bool topo_dirty_ = false;       // breaker status changed since the last upload to the device
    std::vector<int> topo_changed_;  // branches operated since then
    // islands of the last full analysis: per bus its island's smallest bus; the
    // present branches incident to each bus; search marks
    std::vector<int32_t> isl_, adj_ptr_, adj_br_, isl_q0_, isl_q1_;
    std::vector<uint32_t> isl_mark_;
    uint32_t isl_epoch_ = 0;
    std::string tune_report_;
    EngineOptions opt_;
    bool initialized_ = false;
    // buses
    std::vector<std::string> bus_id_;
    std::vector<double> bus_kv_, bus_vset_, bus_p_, bus_q_;
    std::vector<int32_t> bus_type_;
    std::vector<uint8_t> bus_removed_;
    int slack_ = -1;
    // branches
    std::vector<std::string> br_id_;
    std::vector<int32_t> br_from_, br_to_, br_xfmr_, br_sd_;
    std::vector<double> br_r_, br_x_, br_b_, br_g_, br_tap_, br_shift_;
    std::vector<uint8_t> br_status_, br_removed_;
    // measurements
    std::vector<std::string> m_id_;
    std::vector<int32_t> m_type_, m_loc_, m_end_, m_pseudo_;
    std::vector<double> m_sigma_, m_pt_, m_ct_, m_z_;
    std::vector<uint8_t> m_removed_;
    std::vector<uint8_t> m_rejected_;  // set aside as bad data (sle_ext_SetMeasurementActive, auto rejection)
    // switches
    std::vector<std::string> sd_id_;
    std::vector<int32_t> sd_branch_, sd_status_, sd_pending_;
    // meters
    struct Meter {
        int type, bus, branch, end;
        int mv = -1, mp = -1, mq = -1, mi = -1;
    };
    std::vector<std::string> meter_id_;
    std::vector<Meter> meters_;
    std::unordered_map<std::string, int> bus_map_, br_map_, m_map_, sd_map_, meter_map_;

    // state
    bool model_dirty_ = true;     // structure changed: symbolic + full upload
    bool sym_dirty_ = true;       // host symbolic analysis stale
    bool params_dirty_ = false;   // branch parameters changed: Ybus numeric
    bool z_dirty_ = true;         // telemetry changed on host
    bool results_dirty_ = true;   // host copies of P/Q/flows/h/r stale
    bool state_dirty_ = false;    // host copy of V/theta older than the device state
    std::vector<std::pair<std::vector<int32_t>, std::vector<double>>> staged_host_;
    std::vector<int> staged_pending_;  // slots applied on the device, not yet mirrored into m_z_
    std::vector<double> V_, th_;
    std::shared_ptr<Symbolic> sym_;
    std::unique_ptr<Instance> inst_;
    int inst_precision_ = -1;
    ResultOut last_;
};

}  // namespace sle
