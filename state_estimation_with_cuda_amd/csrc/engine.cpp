// engine.cpp -- host engine (see engine.h).
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace sle {

namespace {
constexpr double kEps = 1e-12;  // sle_types.cuh:62 (SLE_REAL_EPSILON, double build)

bool is_bus_meas(int t) { return is_bus_measurement(t); }

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

bool parse_bool(const std::string& v) {
    std::string l = v;
    for (auto& c : l) c = (char)tolower(c);
    return l == "1" || l == "true" || l == "yes" || l == "on";
}
}  // namespace

// ---------------------------------------------------------------------------
// Config file: INI sections [device] [capacity] [solver] [performance]
// [numerical] [bad_data] with the keys of sle_config.h:81-113.
// ---------------------------------------------------------------------------
bool load_config_file(const std::string& path, EngineConfig& c, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open config file " + path;
        return false;
    }
    std::string line;
    int ln = 0;
    while (std::getline(f, line)) {
        ++ln;
        std::string s = trim(line);
        if (s.empty() || s[0] == '#' || s[0] == ';' || s[0] == '[') continue;
        size_t eq = s.find('=');
        if (eq == std::string::npos) {
            err = "config line " + std::to_string(ln) + ": expected key = value";
            return false;
        }
        std::string k = trim(s.substr(0, eq)), v = trim(s.substr(eq + 1));
        size_t hash = v.find('#');
        if (hash != std::string::npos) v = trim(v.substr(0, hash));
        try {
            if (k == "device_id") c.device_id = std::stoi(v);
            else if (k == "max_buses") c.max_buses = std::stoull(v);
            else if (k == "max_branches") c.max_branches = std::stoull(v);
            else if (k == "max_measurements") c.max_measurements = std::stoull(v);
            else if (k == "max_switching_devices") c.max_switching_devices = std::stoull(v);
            else if (k == "enable_robust_estimation") c.enable_robust = parse_bool(v);
            else if (k == "huber_gamma") c.huber_gamma = std::stod(v);
            else if (k == "convergence_tolerance") c.convergence_tolerance = std::stod(v);
            else if (k == "max_realtime_iterations") c.max_realtime_iterations = std::stoi(v);
            else if (k == "max_precision_iterations") c.max_precision_iterations = std::stoi(v);
            else if (k == "realtime_time_limit_ms") c.realtime_time_limit_ms = std::stod(v);
            else if (k == "use_flat_start_default") c.use_flat_start_default = parse_bool(v);
            else if (k == "enable_profiling") c.enable_profiling = parse_bool(v);
            else if (k == "use_cuda_graphs") c.use_cuda_graphs = parse_bool(v);
            else if (k == "block_size_standard") c.block_size_standard = std::stoi(v);
            else if (k == "block_size_reduction") c.block_size_reduction = std::stoi(v);
            else if (k == "use_double_precision") c.use_double_precision = parse_bool(v);
            else if (k == "voltage_min_pu") c.voltage_min_pu = std::stod(v);
            else if (k == "voltage_max_pu") c.voltage_max_pu = std::stod(v);
            else if (k == "bad_data_threshold") c.bad_data_threshold = std::stod(v);
            else if (k == "auto_bad_data_rejection") c.auto_bad_data_rejection = parse_bool(v);
            // unknown keys are ignored, as the reference loader does
        } catch (...) {
            err = "config line " + std::to_string(ln) + ": bad value for " + k;
            return false;
        }
    }
    return true;
}

bool save_config_file(const std::string& path, const EngineConfig& c, std::string& err) {
    std::ofstream f(path);
    if (!f) {
        err = "cannot write config file " + path;
        return false;
    }
    auto b = [](bool v) { return v ? "true" : "false"; };
    f << "# WLS state estimation engine configuration\n";
    f << "[device]\ndevice_id = " << c.device_id << "\n\n";
    f << "[capacity]\nmax_buses = " << c.max_buses << "\nmax_branches = " << c.max_branches
      << "\nmax_measurements = " << c.max_measurements << "\nmax_switching_devices = " << c.max_switching_devices
      << "\n\n";
    f << "[solver]\nenable_robust_estimation = " << b(c.enable_robust) << "\nhuber_gamma = " << c.huber_gamma
      << "\nconvergence_tolerance = " << c.convergence_tolerance
      << "\nmax_realtime_iterations = " << c.max_realtime_iterations
      << "\nrealtime_time_limit_ms = " << c.realtime_time_limit_ms
      << "\nmax_precision_iterations = " << c.max_precision_iterations
      << "\nuse_flat_start_default = " << b(c.use_flat_start_default) << "\n\n";
    f << "[performance]\nenable_profiling = " << b(c.enable_profiling) << "\nuse_cuda_graphs = " << b(c.use_cuda_graphs)
      << "\nblock_size_standard = " << c.block_size_standard << "\nblock_size_reduction = " << c.block_size_reduction
      << "\n\n";
    f << "[numerical]\nuse_double_precision = " << b(c.use_double_precision) << "\nvoltage_min_pu = " << c.voltage_min_pu
      << "\nvoltage_max_pu = " << c.voltage_max_pu << "\n\n";
    f << "[bad_data]\nbad_data_threshold = " << c.bad_data_threshold
      << "\nauto_bad_data_rejection = " << b(c.auto_bad_data_rejection) << "\n";
    return (bool)f;
}

// ---------------------------------------------------------------------------
Engine::Engine() {
    bus_map_.reserve(1024);
    br_map_.reserve(2048);
    m_map_.reserve(8192);
}

Engine::~Engine() = default;

int Engine::initialize() {
    if (initialized_) return ST_OK;
    int nd = device_count();
    if (nd <= 0) return ST_GPU_FAILURE;
    if (cfg_.device_id < 0 || cfg_.device_id >= nd) return ST_INVALID_ARGUMENT;
    init_device_scratch(cfg_.device_id);  // the allocation-free solve's scratch address on this device
    initialized_ = true;
    return ST_OK;
}

void Engine::set_config(const EngineConfig& c) {
    bool prec_change = c.use_double_precision != cfg_.use_double_precision;
    cfg_ = c;
    if (inst_) {
        inst_->set_profiling(cfg_.enable_profiling);
        inst_->set_use_graphs(cfg_.use_cuda_graphs);
    }
    if (prec_change) model_dirty_ = true;
}

int Engine::set_option(const char* key, double value) {
    if (!key) return ST_INVALID_ARGUMENT;
    std::string k = key;
    if (k == "pcg_steps") {
        if (value < 0 || value > 16) return ST_INVALID_ARGUMENT;
        opt_.pcg_steps = (int)value;
    } else if (k == "lds_solve") {
        opt_.lds_solve = value != 0;
        if (inst_) inst_->set_lds_solve(opt_.lds_solve);
    } else if (k == "lds_waves_bottom" || k == "lds_waves_tail") {
        if (value < 1 || value > 16) return ST_INVALID_ARGUMENT;
        (k == "lds_waves_bottom" ? opt_.lds_waves_bottom : opt_.lds_waves_tail) = (int)value;
        model_dirty_ = true;  // the plan is rebuilt at the next upload
    } else if (k == "huber_newton") {
        opt_.huber_newton = value != 0;
    } else if (k == "pinv") {
        opt_.pinv = value != 0;
        if (inst_) inst_->set_partitioned_inverse(opt_.pinv, -1, -1);
    } else if (k == "pinv_cut") {
        if (value < 0 || value > 1e6) return ST_INVALID_ARGUMENT;
        opt_.pinv_cut = (int)value;
        model_dirty_ = true;  // the plan is rebuilt at the next upload
    } else if (k == "pinv_waves") {
        if (value < 1 || value > 16) return ST_INVALID_ARGUMENT;
        opt_.pinv_waves = (int)value;
        model_dirty_ = true;
    } else if (k == "mixed_fp64_factor") {
        opt_.mixed_fp64_factor = value != 0;
        if (inst_) inst_->set_mixed_fp64_factor(opt_.mixed_fp64_factor);
    } else if (k == "rhs_by_location") {
        opt_.rhs_by_location = value != 0;
        model_dirty_ = true;  // the device layout is rebuilt at the next upload
    } else if (k == "meas_by_type") {
        opt_.meas_by_type = value != 0;
        model_dirty_ = true;  // the device layout is rebuilt at the next upload
    } else if (k == "xcd_remap") {
        opt_.xcd_remap = value != 0;
        if (inst_) inst_->set_xcd_remap(opt_.xcd_remap);
    } else if (k == "pinv_mid") {
        opt_.pinv_mid = value != 0;
        model_dirty_ = true;  // the plan is rebuilt at the next upload
    } else if (k == "pinv_wg_fwd" || k == "pinv_wg_bwd") {
        if (value != 0 && value != 1 && value != 4 && value != 8 && value != 16) return ST_INVALID_ARGUMENT;
        (k == "pinv_wg_fwd" ? opt_.pinv_wg_fwd : opt_.pinv_wg_bwd) = (int)value;
        if (inst_) inst_->set_subtree_workgroups(opt_.pinv_wg_fwd, opt_.pinv_wg_bwd);
    } else if (k == "pinv_pack") {
        if (value != -1 && value != 0 && value != 1) return ST_INVALID_ARGUMENT;
        opt_.pinv_pack = (int)value;
        if (inst_) inst_->set_subtree_packing(opt_.pinv_pack);
    } else if (k == "fork_sub_m") {
        if (value != 0 && value != 1 && value != 2) return ST_INVALID_ARGUMENT;
        opt_.fork_sub_m = (int)value;
        if (inst_) inst_->set_fork_subtree_matrices(opt_.fork_sub_m);
    } else if (k == "queue_plain") {
        if (value != 0 && value != 1 && value != 2) return ST_INVALID_ARGUMENT;
        opt_.queue_plain = (int)value;
        if (inst_) inst_->set_queue_plain(opt_.queue_plain);
    } else if (k == "pinv_dense") {
        opt_.pinv_dense = value != 0;
        model_dirty_ = true;  // the plan is rebuilt at the next upload
    } else if (k == "lds_flags") {
        opt_.lds_flags = value != 0;
        if (inst_) inst_->set_lds_flags(opt_.lds_flags);
    } else if (k == "task_sweeps") {
        opt_.task_sweeps = value != 0;
        if (inst_) inst_->set_task_sweeps(opt_.task_sweeps);
    } else if (k == "level_schedule") {
        opt_.level_schedule = value != 0;
        if (inst_) inst_->set_level_schedule(opt_.level_schedule);
    } else if (k == "factor_cut" || k == "factor_wcut") {
        if (value < 0 || value > 1e6) return ST_INVALID_ARGUMENT;
        (k == "factor_cut" ? opt_.factor_cut : opt_.factor_wcut) = (int)value;
        model_dirty_ = true;  // the plan is rebuilt at the next upload
    } else if (k == "bad_data_lnr") {
        opt_.bad_data_lnr = value != 0;
    } else if (k == "lnr_in_result") {
        opt_.lnr_in_result = value != 0;
    } else if (k == "max_eliminations") {
        if (value < 0 || value > 1e6) return ST_INVALID_ARGUMENT;
        opt_.max_eliminations = (int)value;
    } else {
        return ST_INVALID_ARGUMENT;
    }
    return ST_OK;
}

int Engine::bus_index(const char* id) const {
    if (!id) return -1;
    auto it = bus_map_.find(id);
    return it == bus_map_.end() ? -1 : it->second;
}
int Engine::branch_index(const char* id) const {
    if (!id) return -1;
    auto it = br_map_.find(id);
    return it == br_map_.end() ? -1 : it->second;
}
int Engine::meas_index(const char* id) const {
    if (!id) return -1;
    auto it = m_map_.find(id);
    return it == m_map_.end() ? -1 : it->second;
}
int Engine::switch_index(const char* id) const {
    if (!id) return -1;
    auto it = sd_map_.find(id);
    return it == sd_map_.end() ? -1 : it->second;
}

// network_model.cu:71-92
int Engine::add_bus(const char* id, double kv, int type, double vset, double p, double q) {
    if (!id) return ST_INVALID_ARGUMENT;
    if (type < 0 || type > 2) return ST_INVALID_ARGUMENT;
    if (bus_map_.count(id)) return ST_DUPLICATE;
    if (bus_id_.size() >= cfg_.max_buses) return ST_OUT_OF_MEMORY;
    int idx = (int)bus_id_.size();
    bus_map_.emplace(id, idx);
    bus_id_.emplace_back(id);
    bus_kv_.push_back(kv);
    bus_type_.push_back(type);
    bus_vset_.push_back(vset);
    bus_p_.push_back(p);
    bus_q_.push_back(q);
    bus_removed_.push_back(0);
    if (type == 2) slack_ = idx;
    V_.push_back(type != 0 ? vset : 1.0);
    th_.push_back(0.0);
    model_dirty_ = true;
    sym_dirty_ = true;
    return ST_OK;
}

// network_model.cu:94-118
int Engine::add_branch(const char* id, const char* from, const char* to, double r, double x, double b, double g,
                       double tap, double shift, int is_xfmr, const char* sd_id) {
    if (!id || !from || !to) return ST_INVALID_ARGUMENT;
    if (br_map_.count(id)) return ST_DUPLICATE;
    int f = bus_index(from), t = bus_index(to);
    if (f < 0 || t < 0) return ST_NOT_FOUND;
    if (!(tap > 0)) tap = 1.0;
    if (br_id_.size() >= cfg_.max_branches) return ST_OUT_OF_MEMORY;
    int idx = (int)br_id_.size();
    br_map_.emplace(id, idx);
    br_id_.emplace_back(id);
    br_from_.push_back(f);
    br_to_.push_back(t);
    br_r_.push_back(r);
    br_x_.push_back(x);
    br_b_.push_back(b);
    br_g_.push_back(g);
    br_tap_.push_back(tap);
    br_shift_.push_back(shift);
    br_xfmr_.push_back(is_xfmr);
    br_sd_.push_back(-1);
    br_status_.push_back(1);
    br_removed_.push_back(0);
    (void)sd_id;  // a switching device is attached with sle_AddSwitch
    model_dirty_ = true;
    sym_dirty_ = true;
    return ST_OK;
}

// network_model.cu:120-149
int Engine::add_measurement(const char* id, int type, const char* loc, int end, double sigma, double pt, double ct,
                            int pseudo) {
    if (!id || !loc) return ST_INVALID_ARGUMENT;
    if (type < 0 || type > 8 || end < 0 || end > 1) return ST_INVALID_ARGUMENT;
    if (m_map_.count(id)) return ST_DUPLICATE;
    int l = is_bus_meas(type) ? bus_index(loc) : branch_index(loc);
    if (l < 0) return ST_NOT_FOUND;
    if (m_id_.size() >= cfg_.max_measurements) return ST_OUT_OF_MEMORY;
    int idx = (int)m_id_.size();
    m_map_.emplace(id, idx);
    m_id_.emplace_back(id);
    m_type_.push_back(type);
    m_loc_.push_back(l);
    m_end_.push_back(end);
    m_sigma_.push_back(std::max(sigma, kEps));
    m_pt_.push_back(pt);
    m_ct_.push_back(ct);
    m_pseudo_.push_back(pseudo);
    m_z_.push_back(0.0);
    m_removed_.push_back(0);
    m_rejected_.push_back(0);
    model_dirty_ = true;
    sym_dirty_ = true;
    return ST_OK;
}

// network_model.cu:151-180
int Engine::add_switch(const char* id, const char* branch_id, int status) {
    if (!id || !branch_id) return ST_INVALID_ARGUMENT;
    if (sd_map_.count(id)) return ST_DUPLICATE;
    int k = branch_index(branch_id);
    if (k < 0) return ST_NOT_FOUND;
    if (sd_id_.size() >= cfg_.max_switching_devices) return ST_OUT_OF_MEMORY;
    int idx = (int)sd_id_.size();
    sd_map_.emplace(id, idx);
    sd_id_.emplace_back(id);
    sd_branch_.push_back(k);
    sd_status_.push_back(status ? 1 : 0);
    sd_pending_.push_back(-1);
    br_sd_[k] = idx;
    br_status_[k] = status ? 1 : 0;
    model_dirty_ = true;
    sym_dirty_ = true;
    return ST_OK;
}

// Removal takes the element out of service (the reference only erased the id,
// network_model.cu:186-214): a removed branch is permanently open, a removed
// bus loses its branches and measurements.
int Engine::remove_branch(const char* id) {
    int k = branch_index(id);
    if (k < 0) return ST_NOT_FOUND;
    br_removed_[k] = 1;
    br_status_[k] = 0;
    br_map_.erase(id);
    model_dirty_ = true;
    sym_dirty_ = true;
    return ST_OK;
}

int Engine::remove_bus(const char* id) {
    int b = bus_index(id);
    if (b < 0) return ST_NOT_FOUND;
    if (b == slack_) return ST_INVALID_ARGUMENT;
    bus_removed_[b] = 1;
    for (size_t k = 0; k < br_id_.size(); ++k)
        if (br_from_[k] == b || br_to_[k] == b) {
            br_removed_[k] = 1;
            br_status_[k] = 0;
            br_map_.erase(br_id_[k]);
        }
    for (size_t i = 0; i < m_id_.size(); ++i)
        if (is_bus_meas(m_type_[i]) && m_loc_[i] == b) m_removed_[i] = 1;
    bus_map_.erase(id);
    model_dirty_ = true;
    sym_dirty_ = true;
    return ST_OK;
}

int Engine::update_branch(const char* id, double r, double x, double b, double g, double tap, double shift,
                          int is_xfmr) {
    int k = branch_index(id);
    if (k < 0) return ST_NOT_FOUND;
    br_r_[k] = r;
    br_x_[k] = x;
    br_b_[k] = b;
    br_g_[k] = g;
    br_tap_[k] = tap > 0 ? tap : 1.0;
    br_shift_[k] = shift;
    br_xfmr_[k] = is_xfmr;
    params_dirty_ = true;
    return ST_OK;
}

int Engine::set_tap(const char* id, double tap) {
    int k = branch_index(id);
    if (k < 0) return ST_NOT_FOUND;
    if (!(tap > 0)) return ST_INVALID_ARGUMENT;
    br_tap_[k] = tap;
    params_dirty_ = true;
    return ST_OK;
}

int Engine::set_phase_shift(const char* id, double s) {
    int k = branch_index(id);
    if (k < 0) return ST_NOT_FOUND;
    br_shift_[k] = s;
    params_dirty_ = true;
    return ST_OK;
}

int Engine::set_impedance(const char* id, double r, double x, double b, double g) {
    int k = branch_index(id);
    if (k < 0) return ST_NOT_FOUND;
    br_r_[k] = r;
    br_x_[k] = x;
    br_b_[k] = b;
    br_g_[k] = g;
    params_dirty_ = true;
    return ST_OK;
}

bool Engine::get_tap(const char* id, double& tap) const {
    int k = branch_index(id);
    if (k < 0) return false;
    tap = br_tap_[k];
    return true;
}

bool Engine::get_phase_shift(const char* id, double& s) const {
    int k = branch_index(id);
    if (k < 0) return false;
    s = br_shift_[k];
    return true;
}

// network_model.cu:312-383: a meter expands into measurements "<id>.V",
// "<id>.kW", "<id>.kVAR", "<id>.A" depending on its type.
int Engine::add_meter(const char* id, int type, const char* bus_id, const char* branch_id, int end, double pt,
                      double ct, double sv, double sp, double si) {
    if (!id || !bus_id) return ST_INVALID_ARGUMENT;
    if (meter_map_.count(id)) return ST_DUPLICATE;
    int b = bus_index(bus_id);
    if (b < 0) return ST_NOT_FOUND;
    int k = -1;
    if (type != 0) {
        k = branch_index(branch_id);
        if (k < 0) return ST_NOT_FOUND;
    }
    if (type < 0 || type > 3) return ST_INVALID_ARGUMENT;
    Meter mt{type, b, k, end};
    std::string base = id;
    auto add = [&](const char* suffix, int mtype, const char* loc, int e, double sigma, double ptr, double ctr) {
        std::string mid = base + suffix;
        int rc = add_measurement(mid.c_str(), mtype, loc, e, sigma, ptr, ctr, 0);
        return rc == ST_OK ? (int)m_id_.size() - 1 : -1;
    };
    if (type == 0) {
        mt.mv = add(".V", 0, bus_id, 0, sv, pt, 1.0);
    } else if (type == 1) {
        mt.mv = add(".V", 0, bus_id, 0, sv, pt, 1.0);
        mt.mp = add(".kW", 4, branch_id, end, sp, pt, ct);
        mt.mq = add(".kVAR", 5, branch_id, end, sp, pt, ct);
        mt.mi = add(".A", 6, branch_id, end, si, pt, ct);
    } else if (type == 2) {
        mt.mi = add(".A", 6, branch_id, end, si, pt, ct);
    } else {
        mt.mp = add(".kW", 4, branch_id, end, sp, pt, ct);
        mt.mq = add(".kVAR", 5, branch_id, end, sp, pt, ct);
    }
    meter_map_.emplace(id, (int)meter_id_.size());
    meter_id_.emplace_back(id);
    meters_.push_back(mt);
    return ST_OK;
}

namespace {
int channel_meas(int mv, int mp, int mq, int mi, const std::string& ch) {
    if (ch == "V" || ch == "kV" || ch == "voltage") return mv;
    if (ch == "kW" || ch == "MW" || ch == "P" || ch == "power") return mp;
    if (ch == "kVAR" || ch == "MVAR" || ch == "Q" || ch == "reactive") return mq;
    if (ch == "A" || ch == "kA" || ch == "I" || ch == "current") return mi;
    return -1;
}
}  // namespace

int Engine::meter_reading(const char* meter, const char* channel, double* value, int which) const {
    if (!meter || !channel || !value) return ST_INVALID_ARGUMENT;
    auto it = meter_map_.find(meter);
    if (it == meter_map_.end()) return ST_NOT_FOUND;
    const Meter& mt = meters_[it->second];
    int i = channel_meas(mt.mv, mt.mp, mt.mq, mt.mi, channel);
    if (i < 0) return ST_NOT_FOUND;
    if (which == 0) {
        const_cast<Engine*>(this)->flush_staged_host();
        *value = m_z_[i];
    } else {
        const_cast<Engine*>(this)->ensure_results();
        if (which == 1)
            *value = i < (int)h_.size() ? h_[i] : 0.0;
        else
            *value = i < (int)r_.size() ? r_[i] : 0.0;
    }
    return ST_OK;
}

int Engine::update_meter_reading(const char* meter, const char* channel, double value) {
    if (!meter || !channel) return ST_INVALID_ARGUMENT;
    auto it = meter_map_.find(meter);
    if (it == meter_map_.end()) return ST_NOT_FOUND;
    const Meter& mt = meters_[it->second];
    int i = channel_meas(mt.mv, mt.mp, mt.mq, mt.mi, channel);
    if (i < 0) return ST_NOT_FOUND;
    flush_staged_host();
    m_z_[i] = value;
    z_dirty_ = true;
    return ST_OK;
}

// ---------------------------------------------------------------------------
void Engine::admittance(int k, double& g, double& b) const {
    // network_model.cu:595-612 in FP64
    double r = br_r_[k], x = br_x_[k];
    double d = r * r + x * x;
    if (d < kEps) d = kEps;
    g = r / d;
    b = -x / d;
}

bool Engine::meas_active(int i) const {
    if (m_removed_[i] || m_rejected_[i]) return false;
    int t = m_type_[i];
    if (is_bus_meas(t)) return !bus_removed_[m_loc_[i]];
    return br_status_[m_loc_[i]] != 0 && !br_removed_[m_loc_[i]];
}

void Engine::build_host_arrays(HostArrays& h) const {
    h.n = n_bus();
    h.nb = n_branch();
    h.m = n_meas();
    h.slack = slack_;
    h.bus_type = bus_type_;
    h.vset = bus_vset_;
    h.from = br_from_;
    h.to = br_to_;
    h.gs.resize(h.nb);
    h.bs.resize(h.nb);
    for (int k = 0; k < h.nb; ++k) admittance(k, h.gs[k], h.bs[k]);
    h.bc = br_b_;
    h.tap = br_tap_;
    h.shift = br_shift_;
    h.status.resize(h.nb);
    h.present.resize(h.nb);
    for (int k = 0; k < h.nb; ++k) {
        h.status[k] = (br_status_[k] && !br_removed_[k]) ? 1 : 0;
        h.present[k] = br_removed_[k] ? 0 : 1;
    }
    h.mtype = m_type_;
    h.mloc = m_loc_;
    h.mend = m_end_;
    h.active.resize(h.m);
    h.m_present.resize(h.m);
    for (int i = 0; i < h.m; ++i) {
        h.active[i] = meas_active(i) ? 1 : 0;
        const int l = m_loc_[i];
        h.m_present[i] = !m_removed_[i] && (is_bus_meas(m_type_[i]) ? !bus_removed_[l] : !br_removed_[l]);
    }
    h.rejected = m_rejected_;
    h.sigma = m_sigma_;
    h.z = m_z_;
}

int Engine::ensure_symbolic() {
    if (bus_id_.empty() || slack_ < 0) return ST_MODEL_INCOMPLETE;
    if (sym_ && topo_dirty_ && !sym_dirty_) {
        // breaker operations keep the structure unless an island's angle reference changed
        HostArrays h;
        build_host_arrays(h);
        ModelView mv;
        mv.n = h.n;
        mv.nb = h.nb;
        mv.slack = h.slack;
        mv.bus_type = h.bus_type.data();
        mv.from = h.from.data();
        mv.to = h.to.data();
        mv.status = h.status.data();
        if (!islands_kept(mv.status)) {
            if (angle_states(mv) != sym_->th_state) sym_dirty_ = true;
            set_islands(mv);
        }
    }
    if (sym_ && !sym_dirty_) return ST_OK;
    HostArrays h;
    build_host_arrays(h);
    ModelView mv;
    mv.n = h.n;
    mv.nb = h.nb;
    mv.m = h.m;
    mv.slack = h.slack;
    mv.bus_type = h.bus_type.data();
    mv.from = h.from.data();
    mv.to = h.to.data();
    mv.status = h.status.data();
    mv.mtype = h.mtype.data();
    mv.mloc = h.mloc.data();
    mv.mend = h.mend.data();
    mv.active = h.active.data();
    mv.present = h.present.data();
    mv.m_present = h.m_present.data();
    auto S = std::make_shared<Symbolic>();
    analyze(mv, *S);
    sym_ = S;
    sym_dirty_ = false;
    set_islands(mv);
    return ST_OK;
}

void Engine::set_islands(const ModelView& mv) {
    (void)angle_states(mv, &isl_);
    // incident present branches per bus (open ones included: status decides at search time)
    const int n = mv.n;
    adj_ptr_.assign(n + 1, 0);
    for (int k = 0; k < mv.nb; ++k) {
        if (mv.present && !mv.present[k]) continue;
        adj_ptr_[mv.from[k] + 1]++;
        adj_ptr_[mv.to[k] + 1]++;
    }
    for (int b = 0; b < n; ++b) adj_ptr_[b + 1] += adj_ptr_[b];
    adj_br_.resize(adj_ptr_[n]);
    std::vector<int32_t> f(adj_ptr_.begin(), adj_ptr_.end() - 1);
    for (int k = 0; k < mv.nb; ++k) {
        if (mv.present && !mv.present[k]) continue;
        adj_br_[f[mv.from[k]]++] = k;
        adj_br_[f[mv.to[k]]++] = k;
    }
    isl_mark_.assign(n, 0);
    isl_epoch_ = 0;
}

bool Engine::islands_kept(const uint8_t* status) {
    if (isl_.size() != (size_t)n_bus() || adj_ptr_.size() != (size_t)n_bus() + 1) return false;
    for (int k : topo_changed_) {
        const int a = br_from_[k], b = br_to_[k];
        if (status[k] ? isl_[a] != isl_[b] : !ends_connected(a, b, status)) return false;
    }
    return true;
}

// Whether buses a and b are joined by closed branches: two breadth-first
// searches, one from each end, the smaller frontier advancing first, until
// they meet or one side runs out.  In a meshed network an opened branch's ends
// meet again within a few hops.
bool Engine::ends_connected(int a, int b, const uint8_t* status) {
    if (a == b) return true;
    if (isl_epoch_ > 0xfffffff0u) {
        std::fill(isl_mark_.begin(), isl_mark_.end(), 0u);
        isl_epoch_ = 0;
    }
    const uint32_t ma = ++isl_epoch_, mb = ++isl_epoch_;
    isl_q0_.assign(1, a);
    isl_q1_.assign(1, b);
    isl_mark_[a] = ma;
    isl_mark_[b] = mb;
    size_t h0 = 0, h1 = 0;
    while (h0 < isl_q0_.size() && h1 < isl_q1_.size()) {
        const bool first = isl_q0_.size() - h0 <= isl_q1_.size() - h1;
        std::vector<int32_t>& q = first ? isl_q0_ : isl_q1_;
        size_t& h = first ? h0 : h1;
        const uint32_t me = first ? ma : mb, other = first ? mb : ma;
        const int u = q[h++];
        for (int p = adj_ptr_[u]; p < adj_ptr_[u + 1]; ++p) {
            const int k = adj_br_[p];
            if (!status[k]) continue;
            const int v = br_from_[k] == u ? br_to_[k] : br_from_[k];
            if (isl_mark_[v] == other) return true;
            if (isl_mark_[v] != me) {
                isl_mark_[v] = me;
                q.push_back(v);
            }
        }
    }
    return false;
}

int Engine::rebuild() {
    if (!initialized_) return ST_NOT_INITIALIZED;
    ensure_state();
    flush_staged_host();
    int rc = ensure_symbolic();
    if (rc != ST_OK) return rc;
    HostArrays h;
    build_host_arrays(h);
    int prec = cfg_.use_double_precision ? 1 : 0;
    if (!inst_ || inst_precision_ != prec) {
        inst_.reset();
        inst_ = std::make_unique<Instance>(cfg_.device_id, prec);
        inst_precision_ = prec;
    }
    inst_->set_profiling(cfg_.enable_profiling);
    inst_->set_use_graphs(cfg_.use_cuda_graphs);
    inst_->set_level_schedule(opt_.level_schedule);
    inst_->set_task_sweeps(opt_.task_sweeps);
    inst_->set_factor_plan(opt_.factor_cut, opt_.factor_wcut);
    inst_->set_lds_solve(opt_.lds_solve);
    inst_->set_lds_waves(opt_.lds_waves_bottom, opt_.lds_waves_tail);
    inst_->set_partitioned_inverse(opt_.pinv, opt_.pinv_cut, opt_.pinv_waves);
    inst_->set_lds_flags(opt_.lds_flags);
    inst_->set_dense_subtrees(opt_.pinv_dense);
    inst_->set_middle_subtrees(opt_.pinv_mid);
    inst_->set_subtree_workgroups(opt_.pinv_wg_fwd, opt_.pinv_wg_bwd);
    inst_->set_subtree_packing(opt_.pinv_pack);
    inst_->set_queue_plain(opt_.queue_plain);
    inst_->set_fork_subtree_matrices(opt_.fork_sub_m);
    inst_->set_meas_by_type(opt_.meas_by_type);
    inst_->set_rhs_by_location(opt_.rhs_by_location);
    inst_->set_mixed_fp64_factor(opt_.mixed_fp64_factor);
    inst_->setup(h, sym_);
    inst_->set_xcd_remap(opt_.xcd_remap);
    inst_->set_state(V_.data(), th_.data());
    model_dirty_ = false;
    params_dirty_ = false;
    z_dirty_ = false;
    results_dirty_ = true;
    return ST_OK;
}

int Engine::ensure_ready() {
    if (!initialized_) return ST_NOT_INITIALIZED;
    apply_topology_values();
    if (model_dirty_ || !inst_) {
        int rc = rebuild();
        if (rc != ST_OK) return rc;
    }
    if (params_dirty_) {
        HostArrays h;
        build_host_arrays(h);
        inst_->update_branches(h);
        params_dirty_ = false;
    }
    if (z_dirty_) {
        inst_->set_z(m_z_.data());
        z_dirty_ = false;
    }
    return ST_OK;
}

// sle_engine.cu:450-497
int Engine::upload_model() {
    if (!initialized_) return ST_NOT_INITIALIZED;
    model_dirty_ = true;
    return ensure_ready();
}

// sle_engine.cu:508-524: values in insertion order, count == measurement count
int Engine::update_telemetry(const double* values, int count) {
    if (!values || count != n_meas()) return ST_INVALID_ARGUMENT;
    staged_pending_.clear();
    std::copy(values, values + count, m_z_.begin());
    if (inst_ && !model_dirty_) {
        inst_->set_z(m_z_.data());
        z_dirty_ = false;
    } else {
        z_dirty_ = true;
    }
    return ST_OK;
}

int Engine::update_telemetry_indexed(const int32_t* idx, const double* values, int count) {
    if (!idx || !values || count < 0) return ST_INVALID_ARGUMENT;
    for (int q = 0; q < count; ++q)
        if (idx[q] < 0 || idx[q] >= n_meas()) return ST_INVALID_ARGUMENT;
    flush_staged_host();
    for (int q = 0; q < count; ++q) m_z_[idx[q]] = values[q];
    if (inst_ && !model_dirty_ && !z_dirty_) {
        inst_->set_z_indexed(idx, values, count);
    } else {
        z_dirty_ = true;
    }
    return ST_OK;
}

int Engine::update_measurement(const char* id, double value) {
    int i = meas_index(id);
    if (i < 0) return ST_NOT_FOUND;
    flush_staged_host();
    m_z_[i] = value;
    z_dirty_ = true;
    return ST_OK;
}

int Engine::stage_telemetry(int slot, const int32_t* idx, const double* values, int count) {
    if (slot < 0 || slot > 1 << 20 || count < 0 || (count && (!idx || !values))) return ST_INVALID_ARGUMENT;
    for (int q = 0; q < count; ++q)
        if (idx[q] < 0 || idx[q] >= n_meas()) return ST_INVALID_ARGUMENT;
    int rc = ensure_ready();
    if (rc != ST_OK) return rc;
    if ((int)staged_host_.size() <= slot) staged_host_.resize(slot + 1);
    staged_host_[slot].first.assign(idx, idx + count);
    staged_host_[slot].second.assign(values, values + count);
    inst_->stage_telemetry(slot, idx, values, count);
    return ST_OK;
}

int Engine::apply_staged(int slot) {
    if (slot < 0 || slot >= (int)staged_host_.size()) return ST_INVALID_ARGUMENT;
    if (!inst_ || model_dirty_ || z_dirty_) {
        // no current device copy of z: mirror on the host, upload at the next solve
        flush_staged_host();
        for (size_t q = 0; q < staged_host_[slot].first.size(); ++q)
            m_z_[staged_host_[slot].first[q]] = staged_host_[slot].second[q];
        z_dirty_ = true;
        return ST_OK;
    }
    inst_->apply_staged(slot);
    staged_pending_.push_back(slot);
    return ST_OK;
}

void Engine::flush_staged_host() {
    for (int slot : staged_pending_)
        for (size_t q = 0; q < staged_host_[slot].first.size(); ++q)
            m_z_[staged_host_[slot].first[q]] = staged_host_[slot].second[q];
    staged_pending_.clear();
}

void Engine::ensure_state() {
    if (!state_dirty_) return;
    if (inst_) inst_->get_state(V_.data(), th_.data());
    state_dirty_ = false;
}

void Engine::state_out(double* v, double* th, int count) {
    const int n = std::min(count, n_bus());
    if (state_dirty_ && inst_ && n == n_bus() && v && th) {
        inst_->get_state(v, th);  // the host copy stays stale until a getter needs it
        return;
    }
    ensure_state();
    if (v) std::copy(V_.begin(), V_.begin() + n, v);
    if (th) std::copy(th_.begin(), th_.begin() + n, th);
}

int Engine::sync_measurements() {
    if (!inst_ || model_dirty_) return ensure_ready();
    if (z_dirty_) {
        inst_->set_z(m_z_.data());
        z_dirty_ = false;
    }
    return ST_OK;
}

// sle_engine.cu:568-580, 1745-1773: status changes are queued and applied
// before the next solve; applying one changes the topology (symbolic redo).
int Engine::set_switch_status(const char* sd_id, int status) {
    int s = switch_index(sd_id);
    if (s < 0) return ST_NOT_FOUND;
    sd_pending_[s] = status ? 1 : 0;
    return ST_OK;
}

int Engine::get_switch_status(const char* sd_id) const {
    int s = switch_index(sd_id);
    if (s < 0) return 0;
    return sd_pending_[s] >= 0 ? sd_pending_[s] : sd_status_[s];
}

int Engine::apply_switch_changes() {
    for (size_t s = 0; s < sd_id_.size(); ++s) {
        if (sd_pending_[s] < 0) continue;
        int st = sd_pending_[s];
        sd_pending_[s] = -1;
        if (st == sd_status_[s]) continue;
        sd_status_[s] = st;
        int k = sd_branch_[s];
        if (!br_removed_[k]) {
            br_status_[k] = (uint8_t)st;
            topo_dirty_ = true;
            topo_changed_.push_back(k);
        }
    }
    return ST_OK;
}

// A breaker operation changes values only (the structure holds open branches)
// unless it changes which bus is the angle reference of some island: then the
// state layout, and with it the symbolic phase, is redone.
int Engine::apply_topology_values() {
    if (!topo_dirty_) return ST_OK;
    topo_dirty_ = false;
    std::vector<uint8_t> status(n_branch());
    for (int k = 0; k < n_branch(); ++k) status[k] = (br_status_[k] && !br_removed_[k]) ? 1 : 0;
    ModelView mv;
    mv.n = n_bus();
    mv.nb = n_branch();
    mv.m = n_meas();
    mv.slack = slack_;
    mv.bus_type = bus_type_.data();
    mv.from = br_from_.data();
    mv.to = br_to_.data();
    mv.status = status.data();
    if (!sym_) {
        model_dirty_ = true;
        sym_dirty_ = true;
        topo_changed_.clear();
        return ST_OK;
    }
    const std::vector<int> changed = topo_changed_;
    if (!islands_kept(mv.status)) {
        if (angle_states(mv) != sym_->th_state) {  // new state layout: symbolic redo
            model_dirty_ = true;
            sym_dirty_ = true;
            topo_changed_.clear();
            return ST_OK;
        }
        set_islands(mv);  // new islands, same references
    }
    topo_changed_ = changed;
    if (model_dirty_ || !inst_) {  // the next upload carries the status anyway
        topo_changed_.clear();
        return ST_OK;
    }
    for (int k : topo_changed_) inst_->set_branch_closed(k, status[k] != 0);
    topo_changed_.clear();
    inst_->refresh_topology();
    results_dirty_ = true;
    return ST_OK;
}

int Engine::set_branch_status_index(int k, int status) {
    if (k < 0 || k >= n_branch()) return ST_INVALID_ARGUMENT;
    if (br_removed_[k]) return ST_INVALID_ARGUMENT;
    uint8_t st = status ? 1 : 0;
    if (br_status_[k] != st) {
        br_status_[k] = st;
        if (br_sd_[k] >= 0) sd_status_[br_sd_[k]] = st;
        topo_dirty_ = true;
        topo_changed_.push_back(k);
    }
    return ST_OK;
}

SolverConfigIn Engine::mode_config(int mode) const {
    // sle_engine.cu:609-628
    SolverConfigIn sc;
    sc.mode = mode;
    if (mode == 0) {
        sc.max_iterations = cfg_.max_realtime_iterations;
        sc.time_limit_ms = cfg_.realtime_time_limit_ms;
    } else {
        sc.max_iterations = cfg_.max_precision_iterations;
        sc.time_limit_ms = 0;
    }
    sc.tolerance = cfg_.convergence_tolerance;
    sc.use_robust = cfg_.enable_robust;
    sc.huber_gamma = cfg_.huber_gamma;
    sc.use_flat_start = cfg_.use_flat_start_default;
    return sc;
}

// sle_engine.cu:630-682
int Engine::solve(const SolverConfigIn& sc, ResultOut& out) {
    out = ResultOut();
    if (!initialized_) return ST_NOT_INITIALIZED;
    apply_switch_changes();
    int rc = ensure_ready();
    if (rc != ST_OK) {
        out.status = 4;  // NOT_OBSERVABLE
        return rc;
    }
    SolveParams p = solve_params(sc);
    if (cfg_.auto_bad_data_rejection) {  // every solve starts from the full telemetry
        bool any = false;
        for (int i = 0; i < n_meas(); ++i)
            if (m_rejected_[i]) {
                set_measurement_rejected(i, false, false);
                any = true;
            }
        if (any && inst_ && !model_dirty_) inst_->refresh_topology();
    }
    SolveOutcome o = inst_->solve(p);
    take_outcome(o, p, out);
    if ((cfg_.auto_bad_data_rejection || opt_.lnr_in_result) && o.status == 0) lnr_result(sc, out);
    return ST_OK;
}

// The largest normalized residual of a converged estimate into the result;
// with auto_bad_data_rejection, the largest-normalized-residual test repeated:
// while it exceeds the threshold, that measurement is set aside and the state
// re-estimated (at most max_eliminations per solve).
void Engine::lnr_result(const SolverConfigIn& sc, ResultOut& out) {
    const double th = cfg_.bad_data_threshold;
    int eliminated = 0;
    std::vector<double> rn, om;
    for (;;) {
        if (normalized_residuals(rn, om) != ST_OK) return;
        int k = -1;
        for (int i = 0; i < n_meas(); ++i)
            if (rn[i] > 0 && (k < 0 || rn[i] > rn[k])) k = i;
        out.largest_residual = k >= 0 ? rn[k] : 0.0;
        out.largest_residual_idx = k;
        out.bad_data_count = eliminated;
        if (!cfg_.auto_bad_data_rejection || k < 0 || rn[k] <= th || eliminated >= opt_.max_eliminations) break;
        set_measurement_rejected(k, true);
        ++eliminated;
        SolveParams p = solve_params(sc);
        p.flat_start = 0;
        SolveOutcome o = inst_->solve(p);
        ResultOut again;
        take_outcome(o, p, again);
        again.iterations += out.iterations;
        again.refactors += out.refactors;
        out = again;
        out.bad_data_count = eliminated;
        if (o.status != 0) break;
    }
    last_ = out;
}

SolveParams Engine::solve_params(const SolverConfigIn& sc) const {
    SolveParams p;
    p.mode = sc.mode;
    p.max_iter = std::max(1, sc.max_iterations);
    p.tol = sc.tolerance;
    p.robust = sc.use_robust ? (opt_.huber_newton ? 2 : 1) : 0;
    p.gamma = sc.huber_gamma;
    p.flat_start = sc.use_flat_start ? 1 : 0;
    p.vmin = cfg_.voltage_min_pu;
    p.vmax = cfg_.voltage_max_pu;
    p.time_limit_ms = sc.mode == 0 ? sc.time_limit_ms : 0.0;
    p.pcg_steps = cfg_.use_double_precision ? 0 : opt_.pcg_steps;
    return p;
}

int Engine::solve_no_alloc(const SolverConfigIn& sc, ResultOut& out) {
    out = ResultOut();
    if (!initialized_) return ST_NOT_INITIALIZED;
    bool pending = topo_dirty_ || !topo_changed_.empty();
    for (int s : sd_pending_) pending = pending || s >= 0;
    if (!inst_ || model_dirty_ || params_dirty_ || pending)
        throw NoAllocError("the allocation-free solve needs the uploaded and solved model without pending model, "
                           "parameter or breaker changes (solve with sle_ext_Solve first)");
    if (cfg_.auto_bad_data_rejection || opt_.lnr_in_result)
        throw NoAllocError("the allocation-free solve does not run the normalized-residual test "
                           "(auto_bad_data_rejection / lnr_in_result)");
    struct Unfreeze {
        Instance* i;
        ~Unfreeze() { i->freeze_allocations(false); }
    } unfreeze{inst_.get()};
    inst_->freeze_allocations(true);
    if (z_dirty_) {
        inst_->set_z(m_z_.data());
        z_dirty_ = false;
    }
    SolveParams p = solve_params(sc);
    SolveOutcome o = inst_->solve_no_alloc(p);
    take_outcome(o, p, out);
    return ST_OK;
}

void Engine::take_outcome(const SolveOutcome& o, const SolveParams& p, ResultOut& out) {
    state_dirty_ = true;
    results_dirty_ = true;
    out.status = o.status;
    out.iterations = o.iterations;
    out.max_mismatch = o.max_mismatch;
    out.objective = o.objective;
    out.ms = o.ms;
    out.largest_residual = o.largest_residual;
    out.largest_residual_idx = o.largest_residual_idx;
    out.refactors = o.refactors;
    out.fp64_refactors = o.fp64_refactors;
    out.flags = o.flags;
    out.pivots_substituted = o.pivots_substituted;
    out.pcg_steps = p.pcg_steps;
    out.graph_captures = o.graph_captures;
    out.seq_fallbacks = o.seq_fallbacks;
    out.retries = o.retries;
    out.ahead = o.ahead;
    out.ahead_idle = o.ahead_idle;
    out.bad_data_count = 0;
    last_ = out;
}

int Engine::autotune(int reps, std::string& report) {
    if (!initialized_) return ST_NOT_INITIALIZED;
    if (!inst_ || model_dirty_ || !inst_->factor_valid()) return ST_INVALID_ARGUMENT;  // needs a solved model
    report = inst_->autotune_solve(std::max(1, reps));
    tune_report_ = report;
    opt_.pinv_cut = inst_->partitioned_inverse_cut();
    opt_.pinv_waves = inst_->partitioned_inverse_waves();
    opt_.lds_flags = inst_->lds_flags();
    return ST_OK;
}

int Engine::check_observability(int& observable, int& islands) {
    // Topological check (sle_engine.cu:684-753): union-find islands over closed
    // branches, then numerical observability through the gain factor pivots.
    int n = n_bus();
    std::vector<int> par(n);
    std::iota(par.begin(), par.end(), 0);
    std::function<int(int)> find = [&](int a) {
        while (par[a] != a) {
            par[a] = par[par[a]];
            a = par[a];
        }
        return a;
    };
    for (int k = 0; k < n_branch(); ++k)
        if (br_status_[k] && !br_removed_[k]) {
            int a = find(br_from_[k]), b = find(br_to_[k]);
            if (a != b) par[a] = b;
        }
    islands = 0;
    for (int b = 0; b < n; ++b)
        if (!bus_removed_[b] && find(b) == b) islands++;
    observable = 0;
    if (islands != 1) return ST_OK;
    int rc = ensure_ready();
    if (rc != ST_OK) return rc;
    int piv = inst_->debug_factor(nullptr);
    observable = piv == 0 ? 1 : 0;
    inst_->invalidate_factor();
    return ST_OK;
}

int Engine::apply_flat_start() {
    state_dirty_ = false;
    for (int b = 0; b < n_bus(); ++b) {
        V_[b] = bus_type_[b] != 0 ? bus_vset_[b] : 1.0;
        th_[b] = 0.0;
    }
    if (inst_ && !model_dirty_) inst_->set_state(V_.data(), th_.data());
    return ST_OK;
}

int Engine::set_state(const double* V, const double* th, int count) {
    if (!V || !th || count != n_bus()) return ST_INVALID_ARGUMENT;
    state_dirty_ = false;
    std::copy(V, V + count, V_.begin());
    std::copy(th, th + count, th_.begin());
    if (inst_ && !model_dirty_) inst_->set_state(V_.data(), th_.data());
    return ST_OK;
}

void Engine::ensure_results() {
    if (!results_dirty_) return;
    int n = n_bus(), nb = n_branch(), m = n_meas();
    P_.assign(n, 0);
    Q_.assign(n, 0);
    Pf_.assign(nb, 0);
    Qf_.assign(nb, 0);
    Pt_.assign(nb, 0);
    Qt_.assign(nb, 0);
    h_.assign(m, 0);
    r_.assign(m, 0);
    if (inst_ && !model_dirty_)
        inst_->get_results(P_.data(), Q_.data(), Pf_.data(), Qf_.data(), Pt_.data(), Qt_.data(), h_.data(),
                           r_.data());
    results_dirty_ = false;
}

// Active measurements whose normalized residual |r| / sqrt(Omega_ii) exceeds
// the threshold (Abur & Exposito 2004, ch. 5), in index order; with option
// bad_data_lnr = 0 the reference's test |r| / sigma > threshold
// (sle_engine.cu:959-977).
std::vector<int32_t> Engine::identify_bad_data(double threshold) {
    double th = threshold > 0 ? threshold : cfg_.bad_data_threshold;
    std::vector<int32_t> out;
    if (opt_.bad_data_lnr) {
        std::vector<double> rn, om;
        if (normalized_residuals(rn, om) != ST_OK) return out;
        for (int i = 0; i < n_meas(); ++i)
            if (meas_active(i) && rn[i] > th) out.push_back(i);
        return out;
    }
    ensure_results();
    flush_staged_host();
    for (int i = 0; i < n_meas(); ++i)
        if (meas_active(i) && std::fabs(r_[i]) / m_sigma_[i] > th) out.push_back(i);
    return out;
}

int Engine::normalized_residuals(std::vector<double>& rn, std::vector<double>& omega) {
    if (!initialized_) return ST_NOT_INITIALIZED;
    int rc = ensure_ready();
    if (rc != ST_OK) return rc;
    const int m = n_meas();
    std::vector<double> lev(m, 0.0), r(m, 0.0);
    if (!inst_->leverages(lev.data(), r.data())) return ST_SINGULAR;
    rn.assign(m, 0.0);
    omega.assign(m, 0.0);
    for (int i = 0; i < m; ++i) {
        if (!meas_active(i)) continue;
        const double s2 = m_sigma_[i] * m_sigma_[i];
        const double om = s2 * (1.0 - lev[i]);
        omega[i] = om;
        // a critical measurement (Omega_ii ~ 0) has a structurally zero residual: not testable
        rn[i] = om > 1e-10 * s2 ? std::fabs(r[i]) / std::sqrt(om) : 0.0;
    }
    results_dirty_ = true;
    return ST_OK;
}

int Engine::set_measurement_rejected(int i, bool rejected, bool apply) {
    if (i < 0 || i >= n_meas()) return ST_INVALID_ARGUMENT;
    if ((m_rejected_[i] != 0) == rejected) return ST_OK;
    m_rejected_[i] = rejected ? 1 : 0;
    results_dirty_ = true;
    if (inst_ && !model_dirty_) {
        inst_->set_measurement_rejected(i, rejected);
        if (apply) inst_->refresh_topology();
    }
    return ST_OK;
}

size_t Engine::gpu_bytes() const { return inst_ ? inst_->device_bytes() : 0; }

// ---------------------------------------------------------------------------
// IEEE 14-bus case (public test system; measurement set of data/ieee14bus.h:68-270)
// ---------------------------------------------------------------------------
void Engine::load_ieee14() {
    struct B {
        const char* id;
        double kv;
        int type;
        double v, p, q;
    };
    static const B buses[] = {
        {"Bus1", 69.0, 2, 1.060, 0.0, 0.0},      {"Bus2", 69.0, 1, 1.045, 0.40, 0.0},
        {"Bus3", 69.0, 1, 1.010, 0.0, 0.0},      {"Bus4", 69.0, 0, 1.0, -0.478, -0.039},
        {"Bus5", 69.0, 0, 1.0, -0.076, -0.016},  {"Bus6", 13.8, 1, 1.070, 0.0, 0.0},
        {"Bus7", 13.8, 0, 1.0, 0.0, 0.0},        {"Bus8", 18.0, 1, 1.090, 0.0, 0.0},
        {"Bus9", 13.8, 0, 1.0, -0.295, -0.166},  {"Bus10", 13.8, 0, 1.0, -0.090, -0.058},
        {"Bus11", 13.8, 0, 1.0, -0.035, -0.018}, {"Bus12", 13.8, 0, 1.0, -0.061, -0.016},
        {"Bus13", 13.8, 0, 1.0, -0.135, -0.058}, {"Bus14", 13.8, 0, 1.0, -0.149, -0.050},
    };
    struct L {
        const char *id, *f, *t;
        double r, x, b, tap;
        int xf;
    };
    static const L lines[] = {
        {"Line1-2", "Bus1", "Bus2", 0.01938, 0.05917, 0.0528, 1.0, 0},
        {"Line1-5", "Bus1", "Bus5", 0.05403, 0.22304, 0.0492, 1.0, 0},
        {"Line2-3", "Bus2", "Bus3", 0.04699, 0.19797, 0.0438, 1.0, 0},
        {"Line2-4", "Bus2", "Bus4", 0.05811, 0.17632, 0.0340, 1.0, 0},
        {"Line2-5", "Bus2", "Bus5", 0.05695, 0.17388, 0.0346, 1.0, 0},
        {"Line3-4", "Bus3", "Bus4", 0.06701, 0.17103, 0.0128, 1.0, 0},
        {"Line4-5", "Bus4", "Bus5", 0.01335, 0.04211, 0.0, 1.0, 0},
        {"Line6-11", "Bus6", "Bus11", 0.09498, 0.19890, 0.0, 1.0, 0},
        {"Line6-12", "Bus6", "Bus12", 0.12291, 0.25581, 0.0, 1.0, 0},
        {"Line6-13", "Bus6", "Bus13", 0.06615, 0.13027, 0.0, 1.0, 0},
        {"Line7-8", "Bus7", "Bus8", 0.0, 0.17615, 0.0, 1.0, 0},
        {"Line7-9", "Bus7", "Bus9", 0.11001, 0.20640, 0.0, 1.0, 0},
        {"Line9-10", "Bus9", "Bus10", 0.03181, 0.08450, 0.0, 1.0, 0},
        {"Line9-14", "Bus9", "Bus14", 0.12711, 0.27038, 0.0, 1.0, 0},
        {"Line10-11", "Bus10", "Bus11", 0.08205, 0.19207, 0.0, 1.0, 0},
        {"Line12-13", "Bus12", "Bus13", 0.22092, 0.19988, 0.0, 1.0, 0},
        {"Line13-14", "Bus13", "Bus14", 0.17093, 0.34802, 0.0, 1.0, 0},
        {"Xfmr4-7", "Bus4", "Bus7", 0.0, 0.20912, 0.0, 0.978, 1},
        {"Xfmr4-9", "Bus4", "Bus9", 0.0, 0.55618, 0.0, 0.969, 1},
        {"Xfmr5-6", "Bus5", "Bus6", 0.0, 0.25202, 0.0, 0.932, 1},
    };
    for (const B& b : buses) add_bus(b.id, b.kv, b.type, b.v, b.p, b.q);
    for (const L& l : lines) add_branch(l.id, l.f, l.t, l.r, l.x, l.b, 0.0, l.tap, 0.0, l.xf, nullptr);
    for (int i = 1; i <= 14; ++i) {
        std::string bid = "Bus" + std::to_string(i), mid = "V" + std::to_string(i);
        add_measurement(mid.c_str(), 0, bid.c_str(), 0, 0.004, 1.0, 1.0, 0);
    }
    const int pin[] = {1, 2, 3, 6, 8}, qin[] = {1, 2, 3};
    for (int b : pin) {
        std::string bid = "Bus" + std::to_string(b), mid = "P" + std::to_string(b);
        add_measurement(mid.c_str(), 2, bid.c_str(), 0, 0.01, 1.0, 1.0, 0);
    }
    for (int b : qin) {
        std::string bid = "Bus" + std::to_string(b), mid = "Q" + std::to_string(b);
        add_measurement(mid.c_str(), 3, bid.c_str(), 0, 0.01, 1.0, 1.0, 0);
    }
    struct F {
        const char *id, *br;
        int type;
    };
    static const F flows[] = {{"P1-2", "Line1-2", 4}, {"Q1-2", "Line1-2", 5}, {"P1-5", "Line1-5", 4},
                              {"Q1-5", "Line1-5", 5}, {"P2-3", "Line2-3", 4}, {"P4-7", "Xfmr4-7", 4},
                              {"Q4-7", "Xfmr4-7", 5}};
    for (const F& f : flows) add_measurement(f.id, f.type, f.br, 0, 0.008, 1.0, 1.0, 0);
    add_measurement("P7_pseudo", 7, "Bus7", 0, 0.001, 1.0, 1.0, 1);
    add_measurement("Q7_pseudo", 8, "Bus7", 0, 0.001, 1.0, 1.0, 1);
}

}  // namespace sle
