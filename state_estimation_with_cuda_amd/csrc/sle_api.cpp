// sle_api.cpp -- the sle_* C-ABI (include/sle_api.h) and its extensions
// (include/sle_ext.h) over sle::Engine.
//
// Replaces src/sle_api.cu of the reference: same exports, argument checks,
// status mapping (sle_api.cu:971-984), "copy min(count, n)" output convention
// (sle_api.cu:1055-1056), sentinel returns of scalar getters and a
// thread-local last-error string (sle_api.cu:124).  Compiled twice: with
// SLE_Real = float (libsle.so) and with -DSLE_USE_DOUBLE (libsle_f64.so).  No
// C++ exception crosses this boundary.
#define SLE_EXPORTS
#include "sle_api.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>

#include "engine.h"
#include "sle_ext.h"

struct SLE_EngineImpl {
    sle::Engine eng;
    std::string cfg_path_buf;
};

namespace {

thread_local std::string g_last_error;
thread_local std::string g_str_cache;
thread_local std::string g_id_cache;

const char* kDefaultConfig = "sle_engine.config";

inline sle::Engine& E(SLE_Handle h) { return h->eng; }

SLE_StatusCode fail(SLE_StatusCode c, const std::string& msg) {
    g_last_error = msg;
    return c;
}

SLE_StatusCode from_int(int rc, const char* what) {
    if (rc != sle::ST_OK) g_last_error = std::string(what) + " failed (status " + std::to_string(rc) + ")";
    return (SLE_StatusCode)rc;
}

template <typename F>
SLE_StatusCode guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(SLE_ERROR_OUT_OF_MEMORY, "host or device allocation failed");
    } catch (const sle::NoAllocError& e) {
        return fail(SLE_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const sle::GpuError& e) {
        return fail(SLE_ERROR_CUDA_FAILURE, e.what());
    } catch (const std::exception& e) {
        return fail(SLE_ERROR_UNKNOWN, e.what());
    } catch (...) {
        return fail(SLE_ERROR_UNKNOWN, "unknown error");
    }
}

void to_engine_config(const SLE_Config* c, sle::EngineConfig& e) {
    e.device_id = c->device_id;
    e.max_buses = c->max_buses;
    e.max_branches = c->max_branches;
    e.max_measurements = c->max_measurements;
    e.max_switching_devices = c->max_switching_devices;
    e.enable_robust = c->enable_robust != 0;
    e.huber_gamma = c->huber_gamma;
    e.convergence_tolerance = c->convergence_tolerance;
    e.max_realtime_iterations = c->max_realtime_iterations;
    e.max_precision_iterations = c->max_precision_iterations;
    e.realtime_time_limit_ms = c->realtime_time_limit_ms;
    e.use_flat_start_default = c->use_flat_start_default != 0;
    e.enable_profiling = c->enable_profiling != 0;
    e.use_cuda_graphs = c->use_cuda_graphs != 0;
    e.block_size_standard = c->block_size_standard;
    e.block_size_reduction = c->block_size_reduction;
    e.use_double_precision = c->use_double_precision != 0;
    e.voltage_min_pu = c->voltage_min_pu;
    e.voltage_max_pu = c->voltage_max_pu;
    e.bad_data_threshold = c->bad_data_threshold;
    e.auto_bad_data_rejection = c->auto_bad_data_rejection != 0;
}

void from_engine_config(const sle::EngineConfig& e, SLE_Config* c) {
    std::memset(c, 0, sizeof(*c));
    c->device_id = e.device_id;
    c->max_buses = e.max_buses;
    c->max_branches = e.max_branches;
    c->max_measurements = e.max_measurements;
    c->max_switching_devices = e.max_switching_devices;
    c->enable_robust = e.enable_robust;
    c->huber_gamma = (SLE_Real)e.huber_gamma;
    c->convergence_tolerance = (SLE_Real)e.convergence_tolerance;
    c->max_realtime_iterations = e.max_realtime_iterations;
    c->max_precision_iterations = e.max_precision_iterations;
    c->realtime_time_limit_ms = (SLE_Real)e.realtime_time_limit_ms;
    c->use_flat_start_default = e.use_flat_start_default;
    c->enable_profiling = e.enable_profiling;
    c->use_cuda_graphs = e.use_cuda_graphs;
    c->block_size_standard = e.block_size_standard;
    c->block_size_reduction = e.block_size_reduction;
    c->use_double_precision = e.use_double_precision;
    c->voltage_min_pu = (SLE_Real)e.voltage_min_pu;
    c->voltage_max_pu = (SLE_Real)e.voltage_max_pu;
    c->bad_data_threshold = (SLE_Real)e.bad_data_threshold;
    c->auto_bad_data_rejection = e.auto_bad_data_rejection;
}

void fill_result(const sle::ResultOut& r, SLE_Result* out) {
    if (!out) return;
    out->status = (SLE_ConvergenceStatus)r.status;
    out->iterations = r.iterations;
    out->max_mismatch = (SLE_Real)r.max_mismatch;
    out->objective = (SLE_Real)r.objective;
    out->computation_time_ms = (SLE_Real)r.ms;
    out->largest_residual = (SLE_Real)r.largest_residual;
    out->largest_residual_idx = r.largest_residual_idx;
    out->bad_data_count = r.bad_data_count;
}

sle::SolverConfigIn from_solver_config(const SLE_SolverConfig* c) {
    sle::SolverConfigIn s;
    s.mode = c->mode;
    s.max_iterations = c->max_iterations;
    s.tolerance = c->tolerance;
    s.huber_gamma = c->huber_gamma;
    s.use_robust = c->use_robust != 0;
    s.use_flat_start = c->use_flat_start != 0;
    s.time_limit_ms = c->time_limit_ms;
    return s;
}

template <typename Src>
void copy_out(SLE_Real* dst, const Src& src, int32_t count) {
    int32_t n = std::min<int32_t>(count, (int32_t)src.size());
    for (int32_t i = 0; i < n; ++i) dst[i] = (SLE_Real)src[i];
}

SLE_StatusCode create_common(SLE_Handle* handle, const SLE_Config* cfg, const char* path) {
    if (!handle) return fail(SLE_ERROR_INVALID_ARGUMENT, "handle pointer is NULL");
    *handle = nullptr;
    return guarded([&]() -> SLE_StatusCode {
        SLE_EngineImpl* impl = new SLE_EngineImpl();
        if (cfg) {
            to_engine_config(cfg, impl->eng.config());
        } else {
            const char* p = path ? path : kDefaultConfig;
            std::string err;
            FILE* f = std::fopen(p, "r");
            if (f) {
                std::fclose(f);
                if (!sle::load_config_file(p, impl->eng.config(), err)) {
                    delete impl;
                    return fail(SLE_ERROR_CONFIG_FILE, err);
                }
                impl->eng.config_path = p;
            } else if (path) {
                delete impl;
                return fail(SLE_ERROR_CONFIG_FILE, std::string("cannot open config file ") + path);
            }
        }
        *handle = impl;
        return SLE_OK;
    });
}

}  // namespace

extern "C" {

// ---- lifecycle -------------------------------------------------------------
SLE_API SLE_StatusCode SLE_CALL sle_Create(SLE_Handle* handle) { return create_common(handle, nullptr, nullptr); }

SLE_API SLE_StatusCode SLE_CALL sle_CreateWithConfig(SLE_Handle* handle, const SLE_Config* config) {
    if (!config) return fail(SLE_ERROR_INVALID_ARGUMENT, "config is NULL");
    return create_common(handle, config, nullptr);
}

SLE_API SLE_StatusCode SLE_CALL sle_CreateFromConfigFile(SLE_Handle* handle, const char* config_path) {
    if (!config_path) return fail(SLE_ERROR_INVALID_ARGUMENT, "config path is NULL");
    return create_common(handle, nullptr, config_path);
}

SLE_API SLE_StatusCode SLE_CALL sle_Initialize(SLE_Handle handle) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        int rc = E(handle).initialize();
        if (rc == sle::ST_GPU_FAILURE) return fail(SLE_ERROR_CUDA_FAILURE, "no HIP device available");
        return from_int(rc, "initialize");
    });
}

SLE_API void SLE_CALL sle_Destroy(SLE_Handle handle) {
    if (!handle) return;
    try {
        delete handle;
    } catch (...) {
    }
}

// ---- configuration ---------------------------------------------------------
SLE_API SLE_StatusCode SLE_CALL sle_LoadConfigFile(const char* config_path, SLE_Config* config) {
    if (!config_path || !config) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        sle::EngineConfig e;
        std::string err;
        if (!sle::load_config_file(config_path, e, err)) return fail(SLE_ERROR_CONFIG_FILE, err);
        from_engine_config(e, config);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_SaveConfigFile(const char* config_path, const SLE_Config* config) {
    if (!config_path || !config) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        sle::EngineConfig e;
        to_engine_config(config, e);
        std::string err;
        if (!sle::save_config_file(config_path, e, err)) return fail(SLE_ERROR_CONFIG_FILE, err);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ReloadConfig(SLE_Handle handle) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    std::string p = E(handle).config_path.empty() ? kDefaultConfig : E(handle).config_path;
    return sle_ReloadConfigFromFile(handle, p.c_str());
}

SLE_API SLE_StatusCode SLE_CALL sle_ReloadConfigFromFile(SLE_Handle handle, const char* config_path) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    if (!config_path) return fail(SLE_ERROR_INVALID_ARGUMENT, "config path is NULL");
    return guarded([&]() -> SLE_StatusCode {
        sle::EngineConfig e = E(handle).config();
        std::string err;
        if (!sle::load_config_file(config_path, e, err)) return fail(SLE_ERROR_CONFIG_FILE, err);
        E(handle).set_config(e);
        E(handle).config_path = config_path;
        return SLE_OK;
    });
}

SLE_API const char* SLE_CALL sle_GetDefaultConfigPath(void) { return kDefaultConfig; }

SLE_API const char* SLE_CALL sle_GetConfigPath(SLE_Handle handle) {
    if (!handle) return nullptr;
    handle->cfg_path_buf = E(handle).config_path.empty() ? kDefaultConfig : E(handle).config_path;
    return handle->cfg_path_buf.c_str();
}

SLE_API SLE_StatusCode SLE_CALL sle_EnsureConfigFile(const char* config_path) {
    const char* p = config_path ? config_path : kDefaultConfig;
    return guarded([&]() -> SLE_StatusCode {
        FILE* f = std::fopen(p, "r");
        if (f) {
            std::fclose(f);
            return SLE_OK;
        }
        std::string err;
        if (!sle::save_config_file(p, sle::EngineConfig(), err)) return fail(SLE_ERROR_CONFIG_FILE, err);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_GetConfig(SLE_Handle handle, SLE_Config* config) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    if (!config) return fail(SLE_ERROR_INVALID_ARGUMENT, "config is NULL");
    from_engine_config(E(handle).config(), config);
    return SLE_OK;
}

SLE_API SLE_StatusCode SLE_CALL sle_SetConfig(SLE_Handle handle, const SLE_Config* config) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    if (!config) return fail(SLE_ERROR_INVALID_ARGUMENT, "config is NULL");
    return guarded([&]() -> SLE_StatusCode {
        sle::EngineConfig e;
        to_engine_config(config, e);
        E(handle).set_config(e);
        return SLE_OK;
    });
}

// ---- errors / device -------------------------------------------------------
SLE_API const char* SLE_CALL sle_GetLastError(SLE_Handle handle) {
    (void)handle;
    return g_last_error.c_str();
}

SLE_API const char* SLE_CALL sle_GetDeviceInfo(SLE_Handle handle) {
    int dev = handle ? E(handle).config().device_id : 0;
    try {
        if (sle::device_count() <= 0) {
            g_str_cache = "no HIP device";
        } else {
            g_str_cache = sle::device_name(dev);
        }
    } catch (...) {
        g_str_cache = "unavailable";
    }
    return g_str_cache.c_str();
}

// ---- model -----------------------------------------------------------------
SLE_API SLE_StatusCode SLE_CALL sle_AddBus(SLE_Handle handle, const SLE_BusInfo* info) {
    if (!handle || !info || !info->id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).add_bus(info->id, info->base_kv, (int)info->type, info->v_setpoint,
                                          info->p_scheduled, info->q_scheduled),
                        "add bus");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_AddBuses(SLE_Handle handle, const SLE_BusInfo* buses, int32_t count) {
    if (!handle || !buses || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    for (int32_t i = 0; i < count; ++i) {
        SLE_StatusCode s = sle_AddBus(handle, &buses[i]);
        if (s != SLE_OK) return s;
    }
    return SLE_OK;
}

SLE_API SLE_StatusCode SLE_CALL sle_AddBranch(SLE_Handle handle, const SLE_BranchInfo* info) {
    if (!handle || !info || !info->id || !info->from_bus || !info->to_bus)
        return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).add_branch(info->id, info->from_bus, info->to_bus, info->resistance,
                                             info->reactance, info->susceptance, info->conductance,
                                             info->tap_ratio, info->phase_shift, info->is_transformer, info->sd_id),
                        "add branch");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_AddBranches(SLE_Handle handle, const SLE_BranchInfo* branches, int32_t count) {
    if (!handle || !branches || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    for (int32_t i = 0; i < count; ++i) {
        SLE_StatusCode s = sle_AddBranch(handle, &branches[i]);
        if (s != SLE_OK) return s;
    }
    return SLE_OK;
}

SLE_API SLE_StatusCode SLE_CALL sle_AddMeasurement(SLE_Handle handle, const SLE_MeasurementInfo* info) {
    if (!handle || !info || !info->id || !info->location) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).add_measurement(info->id, (int)info->type, info->location, (int)info->branch_end,
                                                  info->sigma, info->pt_ratio, info->ct_ratio, info->is_pseudo),
                        "add measurement");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_AddMeasurements(SLE_Handle handle, const SLE_MeasurementInfo* meas,
                                                    int32_t count) {
    if (!handle || !meas || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    for (int32_t i = 0; i < count; ++i) {
        SLE_StatusCode s = sle_AddMeasurement(handle, &meas[i]);
        if (s != SLE_OK) return s;
    }
    return SLE_OK;
}

SLE_API SLE_StatusCode SLE_CALL sle_AddSwitch(SLE_Handle handle, const SLE_SwitchInfo* info) {
    if (!handle || !info || !info->id || !info->branch_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).add_switch(info->id, info->branch_id, (int)info->status), "add switch");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_RemoveBus(SLE_Handle handle, const char* bus_id) {
    if (!handle || !bus_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).remove_bus(bus_id), "remove bus"); });
}

SLE_API SLE_StatusCode SLE_CALL sle_RemoveBranch(SLE_Handle handle, const char* branch_id) {
    if (!handle || !branch_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).remove_branch(branch_id), "remove branch"); });
}

SLE_API SLE_StatusCode SLE_CALL sle_UpdateBranch(SLE_Handle handle, const SLE_BranchInfo* info) {
    if (!handle || !info || !info->id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).update_branch(info->id, info->resistance, info->reactance, info->susceptance,
                                                info->conductance, info->tap_ratio, info->phase_shift,
                                                info->is_transformer),
                        "update branch");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_SetTransformerTap(SLE_Handle handle, const char* branch_id, SLE_Real tap_ratio) {
    if (!handle || !branch_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).set_tap(branch_id, tap_ratio), "set tap"); });
}

SLE_API SLE_StatusCode SLE_CALL sle_SetTransformerPhaseShift(SLE_Handle handle, const char* branch_id,
                                                             SLE_Real phase_shift) {
    if (!handle || !branch_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).set_phase_shift(branch_id, phase_shift), "set phase shift");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_SetBranchImpedance(SLE_Handle handle, const char* branch_id, SLE_Real r,
                                                       SLE_Real x, SLE_Real b, SLE_Real g) {
    if (!handle || !branch_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).set_impedance(branch_id, r, x, b, g), "set impedance");
    });
}

SLE_API SLE_Real SLE_CALL sle_GetTransformerTap(SLE_Handle handle, const char* branch_id) {
    if (!handle || !branch_id) return (SLE_Real)1.0;
    double t = 1.0;
    return E(handle).get_tap(branch_id, t) ? (SLE_Real)t : (SLE_Real)1.0;
}

SLE_API SLE_Real SLE_CALL sle_GetTransformerPhaseShift(SLE_Handle handle, const char* branch_id) {
    if (!handle || !branch_id) return (SLE_Real)0.0;
    double s = 0.0;
    return E(handle).get_phase_shift(branch_id, s) ? (SLE_Real)s : (SLE_Real)0.0;
}

// ---- meters ------------------------------------------------------------------
SLE_API SLE_StatusCode SLE_CALL sle_AddMeter(SLE_Handle handle, const SLE_MeterInfo* info) {
    if (!handle || !info || !info->id || !info->bus_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).add_meter(info->id, (int)info->type, info->bus_id, info->branch_id,
                                            (int)info->branch_end, info->pt_ratio, info->ct_ratio, info->sigma_v,
                                            info->sigma_p, info->sigma_i),
                        "add meter");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_UpdateMeterReading(SLE_Handle handle, const char* meter_id, const char* channel,
                                                       SLE_Real value) {
    if (!handle || !meter_id || !channel) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).update_meter_reading(meter_id, channel, value), "update meter reading");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_UpdateMeterReadings(SLE_Handle handle, const SLE_MeterReading* readings,
                                                        int32_t count, int32_t sync_to_device) {
    if (!handle || !readings || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded([&]() -> SLE_StatusCode {
        for (int32_t i = 0; i < count; ++i) {
            int rc = E(handle).update_meter_reading(readings[i].meter_id, readings[i].channel, readings[i].value);
            if (rc != sle::ST_OK) return from_int(rc, "update meter readings");
        }
        if (sync_to_device) return from_int(E(handle).sync_measurements(), "sync measurements");
        return SLE_OK;
    });
}

static SLE_StatusCode meter_get(SLE_Handle handle, const char* meter_id, const char* channel, SLE_Real* value,
                                int which) {
    if (!handle || !meter_id || !channel || !value) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        double v = 0;
        int rc = E(handle).meter_reading(meter_id, channel, &v, which);
        if (rc == sle::ST_OK) *value = (SLE_Real)v;
        return from_int(rc, "meter reading");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_GetMeterReading(SLE_Handle handle, const char* meter_id, const char* channel,
                                                    SLE_Real* value) {
    return meter_get(handle, meter_id, channel, value, 0);
}

SLE_API SLE_StatusCode SLE_CALL sle_GetMeterEstimate(SLE_Handle handle, const char* meter_id, const char* channel,
                                                     SLE_Real* value) {
    return meter_get(handle, meter_id, channel, value, 1);
}

SLE_API SLE_StatusCode SLE_CALL sle_GetMeterResidual(SLE_Handle handle, const char* meter_id, const char* channel,
                                                     SLE_Real* residual) {
    return meter_get(handle, meter_id, channel, residual, 2);
}

// ---- data --------------------------------------------------------------------
SLE_API SLE_StatusCode SLE_CALL sle_UploadModel(SLE_Handle handle) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).upload_model(), "upload model"); });
}

SLE_API SLE_StatusCode SLE_CALL sle_UpdateTelemetry(SLE_Handle handle, const SLE_Real* values, int32_t count) {
    if (!handle || !values || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded([&]() -> SLE_StatusCode {
        if (count != E(handle).n_meas())
            return fail(SLE_ERROR_INVALID_ARGUMENT, "telemetry count must equal the measurement count");
#ifdef SLE_USE_DOUBLE
        return from_int(E(handle).update_telemetry(values, count), "update telemetry");
#else
        std::vector<double> z(values, values + count);
        return from_int(E(handle).update_telemetry(z.data(), count), "update telemetry");
#endif
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_UpdateMeasurement(SLE_Handle handle, const char* meas_id, SLE_Real value) {
    if (!handle || !meas_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).update_measurement(meas_id, value), "update measurement");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_UpdateMeasurementBatch(SLE_Handle handle, const SLE_MeasurementUpdate* updates,
                                                           int32_t count, int32_t sync_to_device) {
    if (!handle || !updates || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded([&]() -> SLE_StatusCode {
        for (int32_t i = 0; i < count; ++i) {
            int rc = E(handle).update_measurement(updates[i].meas_id, updates[i].value);
            if (rc != sle::ST_OK) return from_int(rc, "update measurement batch");
        }
        if (sync_to_device) return from_int(E(handle).sync_measurements(), "sync measurements");
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_SyncMeasurementsToDevice(SLE_Handle handle) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).sync_measurements(), "sync measurements"); });
}

SLE_API SLE_StatusCode SLE_CALL sle_GetMeasurementDetails(SLE_Handle handle, const char* meas_id,
                                                          SLE_MeasurementDetails* d) {
    if (!handle || !meas_id || !d) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        sle::Engine& e = E(handle);
        int i = e.meas_index(meas_id);
        if (i < 0) return fail(SLE_ERROR_ELEMENT_NOT_FOUND, "measurement not found");
        e.ensure_results();
        d->id = e.meas_id(i).c_str();
        d->type = (SLE_MeasurementType)e.meas_type(i);
        d->location_id = sle::is_bus_measurement(e.meas_type(i)) ? e.bus_id(e.meas_loc(i)).c_str()
                                                                  : e.branch_id(e.meas_loc(i)).c_str();
        d->branch_end = (SLE_BranchEnd)e.meas_end(i);
        d->sigma = (SLE_Real)e.meas_sigma(i);
        d->pt_ratio = (SLE_Real)e.meas_pt(i);
        d->ct_ratio = (SLE_Real)e.meas_ct(i);
        d->is_pseudo = e.meas_pseudo(i);
        d->internal_index = i;
        d->current_value = (SLE_Real)e.meas_z(i);
        d->estimated_value = i < (int)e.h_.size() ? (SLE_Real)e.h_[i] : (SLE_Real)0;
        d->residual = i < (int)e.r_.size() ? (SLE_Real)e.r_[i] : (SLE_Real)0;
        d->is_active = e.meas_active(i) ? 1 : 0;
        return SLE_OK;
    });
}

SLE_API const char* SLE_CALL sle_GetMeasurementIdByIndex(SLE_Handle handle, int32_t index) {
    if (!handle || index < 0 || index >= E(handle).n_meas()) return nullptr;
    g_id_cache = E(handle).meas_id(index);
    return g_id_cache.c_str();
}

// ---- switches ------------------------------------------------------------------
SLE_API SLE_StatusCode SLE_CALL sle_SetSwitchStatus(SLE_Handle handle, const char* sd_id, SLE_SwitchStatus status) {
    if (!handle || !sd_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).set_switch_status(sd_id, (int)status), "set switch status");
    });
}

SLE_API SLE_SwitchStatus SLE_CALL sle_GetSwitchStatus(SLE_Handle handle, const char* sd_id) {
    if (!handle || !sd_id) return SLE_SWITCH_OPEN;
    return E(handle).get_switch_status(sd_id) ? SLE_SWITCH_CLOSED : SLE_SWITCH_OPEN;
}

SLE_API SLE_StatusCode SLE_CALL sle_ApplySwitchChanges(SLE_Handle handle) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).apply_switch_changes(), "apply switches"); });
}

// ---- solve -------------------------------------------------------------------------
static SLE_StatusCode map_status(int st) {
    switch (st) {
        case 0: return SLE_OK;
        case 1: return SLE_ERROR_MAX_ITERATIONS;
        case 2: return SLE_ERROR_DIVERGED;
        case 3: return SLE_ERROR_SINGULAR_MATRIX;
        case 4: return SLE_ERROR_NOT_OBSERVABLE;
        default: return SLE_ERROR_UNKNOWN;
    }
}

SLE_API SLE_StatusCode SLE_CALL sle_Solve(SLE_Handle handle, SLE_EstimationMode mode, SLE_Result* result) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        sle::ResultOut r;
        int rc = E(handle).solve(E(handle).mode_config((int)mode), r);
        fill_result(r, result);
        if (rc == sle::ST_NOT_INITIALIZED) return fail(SLE_ERROR_NOT_INITIALIZED, "engine not initialized");
        if (rc != sle::ST_OK) return fail(SLE_ERROR_NOT_OBSERVABLE, "model upload failed");
        SLE_StatusCode s = map_status(r.status);
        if (s != SLE_OK) g_last_error = "solve did not converge (status " + std::to_string(r.status) + ")";
        return s;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_SolveWithConfig(SLE_Handle handle, const SLE_SolverConfig* config,
                                                    SLE_Result* result) {
    if (!handle || !config) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        sle::ResultOut r;
        int rc = E(handle).solve(from_solver_config(config), r);
        fill_result(r, result);
        if (rc == sle::ST_NOT_INITIALIZED) return fail(SLE_ERROR_NOT_INITIALIZED, "engine not initialized");
        // sle_api.cu:1011: only OK or MAX_ITERATIONS
        return r.status == 0 ? SLE_OK : SLE_ERROR_MAX_ITERATIONS;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_CheckObservability(SLE_Handle handle, int32_t* is_observable,
                                                       int32_t* island_count) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        int obs = 0, isl = 0;
        int rc = E(handle).check_observability(obs, isl);
        if (is_observable) *is_observable = obs;
        if (island_count) *island_count = isl;
        return rc == sle::ST_OK ? SLE_OK : from_int(rc, "check observability");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ApplyFlatStart(SLE_Handle handle) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).apply_flat_start(), "flat start"); });
}

// ---- results -------------------------------------------------------------------------
SLE_API SLE_StatusCode SLE_CALL sle_GetVoltageMagnitudes(SLE_Handle handle, SLE_Real* values, int32_t count) {
    if (!handle || !values || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    copy_out(values, E(handle).V(), count);
    return SLE_OK;
}

SLE_API SLE_StatusCode SLE_CALL sle_GetVoltageAngles(SLE_Handle handle, SLE_Real* values, int32_t count) {
    if (!handle || !values || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    copy_out(values, E(handle).th(), count);
    return SLE_OK;
}

SLE_API SLE_Real SLE_CALL sle_GetBusVoltage(SLE_Handle handle, const char* bus_id) {
    if (!handle || !bus_id) return (SLE_Real)0;
    int b = E(handle).bus_index(bus_id);
    return b < 0 ? (SLE_Real)0 : (SLE_Real)E(handle).V()[b];
}

SLE_API SLE_Real SLE_CALL sle_GetBusAngle(SLE_Handle handle, const char* bus_id) {
    if (!handle || !bus_id) return (SLE_Real)0;
    int b = E(handle).bus_index(bus_id);
    return b < 0 ? (SLE_Real)0 : (SLE_Real)E(handle).th()[b];
}

SLE_API SLE_StatusCode SLE_CALL sle_GetPowerInjections(SLE_Handle handle, SLE_Real* p_values, SLE_Real* q_values,
                                                       int32_t count) {
    if (!handle || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded([&]() -> SLE_StatusCode {
        E(handle).ensure_results();
        if (p_values) copy_out(p_values, E(handle).P_, count);
        if (q_values) copy_out(q_values, E(handle).Q_, count);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_GetBranchFlows(SLE_Handle handle, SLE_Real* p_from, SLE_Real* q_from,
                                                   SLE_Real* p_to, SLE_Real* q_to, int32_t count) {
    if (!handle || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded([&]() -> SLE_StatusCode {
        E(handle).ensure_results();
        if (p_from) copy_out(p_from, E(handle).Pf_, count);
        if (q_from) copy_out(q_from, E(handle).Qf_, count);
        if (p_to) copy_out(p_to, E(handle).Pt_, count);
        if (q_to) copy_out(q_to, E(handle).Qt_, count);
        return SLE_OK;
    });
}

static SLE_Real branch_value(SLE_Handle handle, const char* branch_id, SLE_BranchEnd end, bool p) {
    if (!handle || !branch_id) return (SLE_Real)0;
    try {
        int k = E(handle).branch_index(branch_id);
        if (k < 0) return (SLE_Real)0;
        E(handle).ensure_results();
        const auto& v = p ? (end == SLE_BRANCH_TO ? E(handle).Pt_ : E(handle).Pf_)
                          : (end == SLE_BRANCH_TO ? E(handle).Qt_ : E(handle).Qf_);
        return k < (int)v.size() ? (SLE_Real)v[k] : (SLE_Real)0;
    } catch (...) {
        return (SLE_Real)0;
    }
}

SLE_API SLE_Real SLE_CALL sle_GetBranchP(SLE_Handle handle, const char* branch_id, SLE_BranchEnd end) {
    return branch_value(handle, branch_id, end, true);
}

SLE_API SLE_Real SLE_CALL sle_GetBranchQ(SLE_Handle handle, const char* branch_id, SLE_BranchEnd end) {
    return branch_value(handle, branch_id, end, false);
}

SLE_API SLE_StatusCode SLE_CALL sle_GetBusResiduals(SLE_Handle handle, const char* bus_id, SLE_Real* v_residual,
                                                    SLE_Real* p_residual, SLE_Real* q_residual, int32_t* count) {
    if (!handle || !bus_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        sle::Engine& e = E(handle);
        int b = e.bus_index(bus_id);
        if (b < 0) return fail(SLE_ERROR_ELEMENT_NOT_FOUND, "bus not found");
        const SLE_Real nan = std::numeric_limits<SLE_Real>::quiet_NaN();
        if (v_residual) *v_residual = nan;
        if (p_residual) *p_residual = nan;
        if (q_residual) *q_residual = nan;
        e.ensure_results();
        int found = 0;
        for (int i = 0; i < e.n_meas(); ++i) {
            int t = e.meas_type(i);
            if (!sle::is_bus_measurement(t) || e.meas_loc(i) != b) continue;
            SLE_Real r = (SLE_Real)e.r_[i];
            if (t == 0) {
                if (v_residual) *v_residual = r;
                found++;
            } else if (t == 2 || t == 7) {
                if (p_residual) *p_residual = r;
                found++;
            } else if (t == 3 || t == 8) {
                if (q_residual) *q_residual = r;
                found++;
            }
        }
        if (count) *count = found;
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_GetBranchResiduals(SLE_Handle handle, const char* branch_id, SLE_BranchEnd end,
                                                       SLE_Real* p_residual, SLE_Real* q_residual,
                                                       SLE_Real* i_residual, int32_t* count) {
    if (!handle || !branch_id) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        sle::Engine& e = E(handle);
        int k = e.branch_index(branch_id);
        if (k < 0) return fail(SLE_ERROR_ELEMENT_NOT_FOUND, "branch not found");
        const SLE_Real nan = std::numeric_limits<SLE_Real>::quiet_NaN();
        if (p_residual) *p_residual = nan;
        if (q_residual) *q_residual = nan;
        if (i_residual) *i_residual = nan;
        e.ensure_results();
        int found = 0;
        for (int i = 0; i < e.n_meas(); ++i) {
            int t = e.meas_type(i);
            if (sle::is_bus_measurement(t) || e.meas_loc(i) != k || e.meas_end(i) != (int)end) continue;
            SLE_Real r = (SLE_Real)e.r_[i];
            if (t == 4) {
                if (p_residual) *p_residual = r;
                found++;
            } else if (t == 5) {
                if (q_residual) *q_residual = r;
                found++;
            } else if (t == 6) {
                if (i_residual) *i_residual = r;
                found++;
            }
        }
        if (count) *count = found;
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_GetResiduals(SLE_Handle handle, SLE_Real* values, int32_t count) {
    if (!handle || !values || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded([&]() -> SLE_StatusCode {
        E(handle).ensure_results();
        copy_out(values, E(handle).r_, count);
        return SLE_OK;
    });
}

SLE_API SLE_Real SLE_CALL sle_GetObjective(SLE_Handle handle) {
    if (!handle) return (SLE_Real)0;
    return (SLE_Real)E(handle).objective();
}

SLE_API SLE_StatusCode SLE_CALL sle_IdentifyBadData(SLE_Handle handle, SLE_Real threshold, int32_t* indices,
                                                    int32_t max_count, int32_t* actual_count) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        std::vector<int32_t> bad = E(handle).identify_bad_data(threshold);
        int32_t n = std::min<int32_t>(max_count, (int32_t)bad.size());
        if (indices && n > 0) std::memcpy(indices, bad.data(), n * sizeof(int32_t));
        if (actual_count) *actual_count = (int32_t)bad.size();
        return SLE_OK;
    });
}

// ---- info ------------------------------------------------------------------------------
SLE_API int32_t SLE_CALL sle_GetBusCount(SLE_Handle handle) { return handle ? E(handle).n_bus() : 0; }
SLE_API int32_t SLE_CALL sle_GetBranchCount(SLE_Handle handle) { return handle ? E(handle).n_branch() : 0; }
SLE_API int32_t SLE_CALL sle_GetMeasurementCount(SLE_Handle handle) { return handle ? E(handle).n_meas() : 0; }
SLE_API int32_t SLE_CALL sle_GetMeterCount(SLE_Handle handle) { return handle ? E(handle).n_meter() : 0; }
SLE_API size_t SLE_CALL sle_GetGPUMemoryUsage(SLE_Handle handle) { return handle ? E(handle).gpu_bytes() : 0; }
SLE_API SLE_Real SLE_CALL sle_GetLastSolveTime(SLE_Handle handle) {
    return handle ? (SLE_Real)E(handle).last_solve_ms() : (SLE_Real)0;
}

SLE_API SLE_StatusCode SLE_CALL sle_LoadIEEE14Bus(SLE_Handle handle) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        E(handle).load_ieee14();
        return SLE_OK;
    });
}

SLE_API const char* SLE_CALL sle_GetVersion(void) { return SLE_VERSION_STRING; }

SLE_API int32_t SLE_CALL sle_GetCUDAVersion(void) {
    try {
        return sle::runtime_version();
    } catch (...) {
        return 0;
    }
}

SLE_API int32_t SLE_CALL sle_IsGPUAvailable(void) {
    try {
        return sle::device_count() > 0 ? 1 : 0;
    } catch (...) {
        return 0;
    }
}

// =========================================================================
// Extensions (include/sle_ext.h)
// =========================================================================
namespace {
SLE_StatusCode ext_solve(SLE_Handle handle, const SLE_SolverConfig* config, SLE_ExtResult* result, bool no_alloc) {
    if (!handle || !config) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        auto t0 = std::chrono::steady_clock::now();
        sle::ResultOut r;
        int rc = no_alloc ? E(handle).solve_no_alloc(from_solver_config(config), r)
                          : E(handle).solve(from_solver_config(config), r);
        if (result) {
            std::memset(result, 0, sizeof(*result));
            result->status = r.status;
            result->iterations = r.iterations;
            result->refactors = r.refactors;
            result->fp64_refactors = r.fp64_refactors;
            result->max_mismatch = r.max_mismatch;
            result->objective = r.objective;
            result->largest_residual = r.largest_residual;
            result->largest_residual_idx = r.largest_residual_idx;
            result->flags = r.flags;
            result->pivots_substituted = r.pivots_substituted;
            result->pcg_steps = r.pcg_steps;
            result->graph_captures = r.graph_captures;
            result->seq_fallbacks = r.seq_fallbacks;
            result->retries = r.retries;
            result->ahead = r.ahead;
            result->ahead_idle = r.ahead_idle;
            result->gpu_ms = r.ms;
            result->host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        if (rc != sle::ST_OK) return from_int(rc, "solve");
        return map_status(r.status);
    });
}
}  // namespace

SLE_API SLE_StatusCode SLE_CALL sle_ext_Solve(SLE_Handle handle, const SLE_SolverConfig* config,
                                              SLE_ExtResult* result) {
    return ext_solve(handle, config, result, false);
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_SolveNoAlloc(SLE_Handle handle, const SLE_SolverConfig* config,
                                                     SLE_ExtResult* result) {
    return ext_solve(handle, config, result, true);
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_NormalizedResiduals(SLE_Handle handle, double* rn, double* omega,
                                                           int32_t count) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    if (count < 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded([&]() -> SLE_StatusCode {
        std::vector<double> r, om;
        int rc = E(handle).normalized_residuals(r, om);
        if (rc != sle::ST_OK) return from_int(rc, "normalized residuals");
        const size_t n = std::min<size_t>((size_t)count, r.size());
        if (rn) std::copy(r.begin(), r.begin() + n, rn);
        if (omega) std::copy(om.begin(), om.begin() + n, omega);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_SetMeasurementActive(SLE_Handle handle, int32_t index, int32_t active) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).set_measurement_rejected(index, active == 0), "set measurement active");
    });
}

SLE_API const char* SLE_CALL sle_ext_Describe(SLE_Handle handle) {
    if (!handle) return nullptr;
    try {
        sle::Engine& e = E(handle);
        std::string d = e.config().use_double_precision ? "plain FP64" : "mixed: FP64 factor, FP32 solve data, FP64 residuals and state";
        d += "; pcg_steps=" + std::to_string(e.config().use_double_precision ? 0 : e.options().pcg_steps);
        if (e.instance()) d += "; sweeps=" + e.instance()->sweep_plan();
        if (!e.autotune_report().empty()) d += "; autotune: " + e.autotune_report();
        g_str_cache = d;
    } catch (...) {
        g_str_cache = "unavailable";
    }
    return g_str_cache.c_str();
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_AutotuneSolve(SLE_Handle handle, int32_t repetitions) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        std::string report;
        int rc = E(handle).autotune(repetitions, report);
        if (rc == sle::ST_INVALID_ARGUMENT) return fail(SLE_ERROR_INVALID_ARGUMENT, "autotune needs a solved model");
        return from_int(rc, "autotune");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_SetOption(SLE_Handle handle, const char* key, double value) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).set_option(key, value), "set option"); });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_UpdateTelemetryIndexed(SLE_Handle handle, const int32_t* idx,
                                                               const double* values, int32_t count) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).update_telemetry_indexed(idx, values, count), "indexed telemetry");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_UpdateTelemetryF64(SLE_Handle handle, const double* values, int32_t count) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).update_telemetry(values, count), "update telemetry");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_SetState(SLE_Handle handle, const double* v, const double* theta,
                                                 int32_t count) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).set_state(v, theta, count), "set state"); });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GetState(SLE_Handle handle, double* v, double* theta, int32_t count) {
    if (!handle || count <= 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded([&]() -> SLE_StatusCode {
        E(handle).state_out(v, theta, count);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GetResults(SLE_Handle handle, double* p_inj, double* q_inj, double* p_from,
                                                   double* q_from, double* p_to, double* q_to, double* h,
                                                   double* residual) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        sle::Engine& e = E(handle);
        e.ensure_results();
        auto cp = [](double* dst, const std::vector<double>& v) {
            if (dst) std::copy(v.begin(), v.end(), dst);
        };
        cp(p_inj, e.P_);
        cp(q_inj, e.Q_);
        cp(p_from, e.Pf_);
        cp(q_from, e.Qf_);
        cp(p_to, e.Pt_);
        cp(q_to, e.Qt_);
        cp(h, e.h_);
        cp(residual, e.r_);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_SetBranchStatus(SLE_Handle handle, int32_t branch_index,
                                                        SLE_SwitchStatus status) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).set_branch_status_index(branch_index, (int)status), "set branch status");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_StageTelemetry(SLE_Handle handle, int32_t slot, const int32_t* idx,
                                                       const double* values, int32_t count) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        return from_int(E(handle).stage_telemetry(slot, idx, values, count), "stage telemetry");
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_ApplyStagedTelemetry(SLE_Handle handle, int32_t slot) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode { return from_int(E(handle).apply_staged(slot), "apply staged"); });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GetKernelTimes(SLE_Handle handle, double* ms, int64_t* launches,
                                                       int64_t* calls, int32_t count) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    sle::Instance* inst = E(handle).instance();
    for (int32_t c = 0; c < count && c < sle::PH_COUNT; ++c) {
        if (ms) ms[c] = inst ? inst->kernel_times().ms[c] : 0.0;
        if (launches) launches[c] = inst ? inst->kernel_times().launches[c] : 0;
        if (calls) calls[c] = inst ? inst->kernel_times().calls[c] : 0;
    }
    return SLE_OK;
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GetSolvePlanSizes(SLE_Handle handle, int64_t* sizes, int32_t count) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    if (!sizes || count < 0) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL sizes");
    for (int32_t i = 0; i < count; ++i) sizes[i] = 0;
    sle::Instance* inst = E(handle).instance();
    if (inst) inst->solve_plan_sizes(sizes, count);
    return SLE_OK;
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_ResetKernelTimes(SLE_Handle handle) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    if (E(handle).instance()) E(handle).instance()->reset_kernel_times();
    return SLE_OK;
}

static SLE_StatusCode need_symbolic(SLE_Handle handle, std::shared_ptr<const sle::Symbolic>& S,
                                    bool device = true) {
    int rc = (device && E(handle).initialized()) ? E(handle).ensure_ready() : E(handle).ensure_symbolic();
    if (rc != sle::ST_OK) return from_int(rc, "prepare model");
    S = E(handle).symbolic();
    if (!S) return fail(SLE_ERROR_NOT_INITIALIZED, "model not uploaded");
    return SLE_OK;
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GetSizes(SLE_Handle handle, int64_t* sizes) {
    if (!handle || !sizes) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        std::shared_ptr<const sle::Symbolic> S;
        SLE_StatusCode rc = need_symbolic(handle, S, false);
        if (rc != SLE_OK) return rc;
        sizes[0] = S->ns;
        sizes[1] = (int64_t)S->ycol.size();
        sizes[2] = (int64_t)S->hcol.size();
        sizes[3] = S->nnz_g;
        sizes[4] = S->nnz_l;
        sizes[5] = S->nsn;
        sizes[6] = S->n_levels;
        sizes[7] = S->max_c;
        sizes[8] = (int64_t)S->gpair.size() / 2;
        sizes[9] = S->ftotal;
        sizes[10] = S->flops;
        sizes[11] = (int64_t)(S->seconds * 1e6);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GetYbus(SLE_Handle handle, int32_t* row_ptr, int32_t* col, double* g,
                                                double* b) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        std::shared_ptr<const sle::Symbolic> S;
        SLE_StatusCode rc = need_symbolic(handle, S, false);
        if (rc != SLE_OK) return rc;
        if (row_ptr) std::copy(S->yrow.begin(), S->yrow.end(), row_ptr);
        if (col) std::copy(S->ycol.begin(), S->ycol.end(), col);
        if (g && b) {
            if (!E(handle).instance()) return fail(SLE_ERROR_NOT_INITIALIZED, "Ybus values need an initialized engine");
            E(handle).instance()->debug_ybus(g, b);
        }
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GetHPattern(SLE_Handle handle, int32_t* row_ptr, int32_t* col) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        std::shared_ptr<const sle::Symbolic> S;
        SLE_StatusCode rc = need_symbolic(handle, S, false);
        if (rc != SLE_OK) return rc;
        if (row_ptr) std::copy(S->hrow.begin(), S->hrow.end(), row_ptr);
        if (col) std::copy(S->hcol.begin(), S->hcol.end(), col);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GetGPattern(SLE_Handle handle, int32_t* col_ptr, int32_t* row) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        std::shared_ptr<const sle::Symbolic> S;
        SLE_StatusCode rc = need_symbolic(handle, S, false);
        if (rc != SLE_OK) return rc;
        std::vector<int32_t> cp, rw;
        sle::natural_gain_pattern(*S, cp, rw);
        if (col_ptr) std::copy(cp.begin(), cp.end(), col_ptr);
        if (row) std::copy(rw.begin(), rw.end(), row);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GetOrdering(SLE_Handle handle, int32_t* bus_perm, int32_t* perm,
                                                    int32_t* parent, int32_t* colcount, int32_t* snode_first) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        std::shared_ptr<const sle::Symbolic> S;
        SLE_StatusCode rc = need_symbolic(handle, S, false);
        if (rc != SLE_OK) return rc;
        if (bus_perm) std::copy(S->bus_perm.begin(), S->bus_perm.end(), bus_perm);
        if (perm) std::copy(S->perm.begin(), S->perm.end(), perm);
        if (parent) std::copy(S->parent.begin(), S->parent.end(), parent);
        if (colcount) std::copy(S->colcount.begin(), S->colcount.end(), colcount);
        if (snode_first) {
            std::copy(S->sn_first.begin(), S->sn_first.end(), snode_first);
            snode_first[S->nsn] = S->ns;
        }
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_Evaluate(SLE_Handle handle, double* h, double* p_inj, double* q_inj,
                                                 double* p_from, double* q_from, double* p_to, double* q_to) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        std::shared_ptr<const sle::Symbolic> S;
        if (!E(handle).initialized()) return fail(SLE_ERROR_NOT_INITIALIZED, "engine not initialized");
        SLE_StatusCode rc = need_symbolic(handle, S);
        if (rc != SLE_OK) return rc;
        E(handle).instance()->debug_eval(h, p_inj, q_inj, p_from, q_from, p_to, q_to);
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_GainRhs(SLE_Handle handle, int32_t by_location, double* b) {
    if (!handle || !b) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        if (!E(handle).initialized()) return fail(SLE_ERROR_NOT_INITIALIZED, "engine not initialized");
        int rc = E(handle).ensure_ready();
        if (rc != sle::ST_OK) return from_int(rc, "gain rhs");
        if (!E(handle).instance()->debug_rhs(by_location != 0, b))
            return fail(SLE_ERROR_INVALID_ARGUMENT, "the right-hand side by location is off (option rhs_by_location)");
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_JacobianGain(SLE_Handle handle, double* h_values, double* g_values) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        std::shared_ptr<const sle::Symbolic> S;
        if (!E(handle).initialized()) return fail(SLE_ERROR_NOT_INITIALIZED, "engine not initialized");
        SLE_StatusCode rc = need_symbolic(handle, S);
        if (rc != SLE_OK) return rc;
        std::vector<double> gsym(std::max<int64_t>(1, S->nnz_g));
        E(handle).instance()->debug_jacobian_gain(h_values, gsym.data());
        if (g_values) {
            std::vector<int32_t> cp, rw;
            std::vector<int64_t> ent;
            sle::natural_gain_pattern(*S, cp, rw, &ent);
            for (int64_t q = 0; q < S->nnz_g; ++q) g_values[q] = gsym[ent[q]];
        }
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_Factor(SLE_Handle handle, int64_t* col_ptr, int32_t* row, double* val) {
    if (!handle) return fail(SLE_ERROR_INVALID_HANDLE, "NULL handle");
    return guarded([&]() -> SLE_StatusCode {
        std::shared_ptr<const sle::Symbolic> S;
        if (!E(handle).initialized()) return fail(SLE_ERROR_NOT_INITIALIZED, "engine not initialized");
        SLE_StatusCode rc = need_symbolic(handle, S);
        if (rc != SLE_OK) return rc;
        std::vector<double> fr(std::max<int64_t>(1, S->ftotal));
        int piv = E(handle).instance()->debug_factor(fr.data());
        int64_t q = 0;
        if (col_ptr) col_ptr[0] = 0;
        for (int s = 0; s < S->nsn; ++s) {
            int c = S->sn_c[s], k = S->sn_k[s];
            const int32_t* rows = S->sn_rows.data() + S->sn_rows_off[s];
            const double* F = fr.data() + S->sn_foff[s];
            for (int jl = 0; jl < k; ++jl) {
                for (int i = jl; i < c; ++i) {
                    if (row) row[q] = rows[i];
                    if (val) val[q] = F[(int64_t)jl * c + i];
                    ++q;
                }
                if (col_ptr) col_ptr[S->sn_first[s] + jl + 1] = q;
            }
        }
        if (piv) return fail(SLE_ERROR_SINGULAR_MATRIX, "pivot below threshold");
        return SLE_OK;
    });
}

SLE_API SLE_StatusCode SLE_CALL sle_ext_SolveFactor(SLE_Handle handle, const double* b, double* dx) {
    if (!handle || !b || !dx) return fail(SLE_ERROR_INVALID_ARGUMENT, "NULL argument");
    return guarded([&]() -> SLE_StatusCode {
        std::shared_ptr<const sle::Symbolic> S;
        if (!E(handle).initialized()) return fail(SLE_ERROR_NOT_INITIALIZED, "engine not initialized");
        SLE_StatusCode rc = need_symbolic(handle, S);
        if (rc != SLE_OK) return rc;
        E(handle).instance()->debug_solve(b, dx);
        return SLE_OK;
    });
}

}  // extern "C"
