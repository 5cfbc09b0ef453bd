"""ctypes binding of the sle_* C-ABI (include/sle_api.h, include/sle_ext.h).

This is the Python view of the same drop-in boundary a C/C++ host links
against: one ``Engine`` per ``SLE_Handle``.  Models are loaded with the bulk
``sle_AddBuses`` / ``sle_AddBranches`` / ``sle_AddMeasurements`` calls from
numpy structured arrays laid out exactly like the C structs, so 2M
measurements load without per-element Python work.

The shared library must have been built (``python -m
state_estimation_with_cuda_amd.build``); there is no fallback: a missing
library raises ``ImportError`` and a missing GPU makes ``sle_Initialize``
return SLE_ERROR_CUDA_FAILURE.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")

# SLE_StatusCode (sle_api.h:145-161)
OK = 0
ERR = {
    -1: "INVALID_HANDLE",
    -2: "NOT_INITIALIZED",
    -3: "CUDA_FAILURE",
    -4: "INVALID_ARGUMENT",
    -5: "MODEL_INCOMPLETE",
    -6: "NOT_OBSERVABLE",
    -7: "SINGULAR_MATRIX",
    -8: "DIVERGED",
    -9: "MAX_ITERATIONS",
    -10: "OUT_OF_MEMORY",
    -11: "ELEMENT_NOT_FOUND",
    -12: "DUPLICATE_ID",
    -13: "CONFIG_FILE",
    -99: "UNKNOWN",
}
MODE_REALTIME, MODE_PRECISION = 0, 1
CONVERGED, MAX_ITERATIONS, DIVERGED, SINGULAR, NOT_OBSERVABLE, IN_PROGRESS = range(6)


class SLEError(RuntimeError):
    def __init__(self, code, where, msg=""):
        super().__init__(f"{where}: {ERR.get(code, code)} ({code}) {msg}")
        self.code = code


def _structs(real):
    class BusInfo(C.Structure):
        _fields_ = [("id", C.c_char_p), ("base_kv", real), ("type", C.c_int), ("v_setpoint", real),
                    ("p_scheduled", real), ("q_scheduled", real)]

    class BranchInfo(C.Structure):
        _fields_ = [("id", C.c_char_p), ("from_bus", C.c_char_p), ("to_bus", C.c_char_p), ("resistance", real),
                    ("reactance", real), ("susceptance", real), ("conductance", real), ("tap_ratio", real),
                    ("phase_shift", real), ("is_transformer", C.c_int32), ("sd_id", C.c_char_p)]

    class MeasurementInfo(C.Structure):
        _fields_ = [("id", C.c_char_p), ("type", C.c_int), ("location", C.c_char_p), ("branch_end", C.c_int),
                    ("sigma", real), ("pt_ratio", real), ("ct_ratio", real), ("is_pseudo", C.c_int32)]

    class SwitchInfo(C.Structure):
        _fields_ = [("id", C.c_char_p), ("branch_id", C.c_char_p), ("status", C.c_int)]

    class MeterInfo(C.Structure):
        _fields_ = [("id", C.c_char_p), ("type", C.c_int), ("bus_id", C.c_char_p), ("branch_id", C.c_char_p),
                    ("branch_end", C.c_int), ("pt_ratio", real), ("ct_ratio", real), ("sigma_v", real),
                    ("sigma_p", real), ("sigma_i", real)]

    class MeterReading(C.Structure):
        _fields_ = [("meter_id", C.c_char_p), ("channel", C.c_char_p), ("value", real)]

    class Result(C.Structure):
        _fields_ = [("status", C.c_int), ("iterations", C.c_int32), ("max_mismatch", real), ("objective", real),
                    ("computation_time_ms", real), ("largest_residual", real), ("largest_residual_idx", C.c_int32),
                    ("bad_data_count", C.c_int32)]

    class Config(C.Structure):
        _fields_ = [("device_id", C.c_int32), ("max_buses", C.c_size_t), ("max_branches", C.c_size_t),
                    ("max_measurements", C.c_size_t), ("max_switching_devices", C.c_size_t),
                    ("enable_robust", C.c_int32), ("huber_gamma", real), ("convergence_tolerance", real),
                    ("max_realtime_iterations", C.c_int32), ("max_precision_iterations", C.c_int32),
                    ("realtime_time_limit_ms", real), ("use_flat_start_default", C.c_int32),
                    ("enable_profiling", C.c_int32), ("use_cuda_graphs", C.c_int32),
                    ("block_size_standard", C.c_int32), ("block_size_reduction", C.c_int32),
                    ("use_double_precision", C.c_int32), ("voltage_min_pu", real), ("voltage_max_pu", real),
                    ("bad_data_threshold", real), ("auto_bad_data_rejection", C.c_int32)]

    class SolverConfig(C.Structure):
        _fields_ = [("mode", C.c_int), ("max_iterations", C.c_int32), ("tolerance", real), ("huber_gamma", real),
                    ("use_robust", C.c_int32), ("use_flat_start", C.c_int32), ("time_limit_ms", real)]

    class MeasurementUpdate(C.Structure):
        _fields_ = [("meas_id", C.c_char_p), ("value", real)]

    class MeasurementDetails(C.Structure):
        _fields_ = [("id", C.c_char_p), ("type", C.c_int), ("location_id", C.c_char_p), ("branch_end", C.c_int),
                    ("sigma", real), ("pt_ratio", real), ("ct_ratio", real), ("is_pseudo", C.c_int32),
                    ("internal_index", C.c_int32), ("current_value", real), ("estimated_value", real),
                    ("residual", real), ("is_active", C.c_int32)]

    return dict(BusInfo=BusInfo, BranchInfo=BranchInfo, MeasurementInfo=MeasurementInfo, SwitchInfo=SwitchInfo,
                MeterInfo=MeterInfo, MeterReading=MeterReading, Result=Result, Config=Config,
                SolverConfig=SolverConfig, MeasurementUpdate=MeasurementUpdate,
                MeasurementDetails=MeasurementDetails)


class ExtResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("refactors", C.c_int32),
                ("fp64_refactors", C.c_int32), ("max_mismatch", C.c_double), ("objective", C.c_double),
                ("largest_residual", C.c_double), ("largest_residual_idx", C.c_int32), ("flags", C.c_int32),
                ("gpu_ms", C.c_double), ("host_ms", C.c_double), ("pivots_substituted", C.c_int32),
                ("pcg_steps", C.c_int32), ("graph_captures", C.c_int32), ("seq_fallbacks", C.c_int32),
                ("retries", C.c_int32), ("ahead", C.c_int32), ("ahead_idle", C.c_int32)]


# every function of the reference ABI (sle_api.h:447-1391) and the extensions
API_FUNCTIONS = [
    "sle_Create", "sle_CreateWithConfig", "sle_CreateFromConfigFile", "sle_Initialize", "sle_Destroy",
    "sle_LoadConfigFile", "sle_SaveConfigFile", "sle_ReloadConfig", "sle_ReloadConfigFromFile",
    "sle_GetDefaultConfigPath", "sle_GetConfigPath", "sle_EnsureConfigFile", "sle_GetConfig", "sle_SetConfig",
    "sle_GetLastError", "sle_GetDeviceInfo", "sle_AddBus", "sle_AddBuses", "sle_AddBranch", "sle_AddBranches",
    "sle_AddMeasurement", "sle_AddMeasurements", "sle_AddSwitch", "sle_RemoveBus", "sle_RemoveBranch",
    "sle_UpdateBranch", "sle_SetTransformerTap", "sle_SetTransformerPhaseShift", "sle_SetBranchImpedance",
    "sle_GetTransformerTap", "sle_GetTransformerPhaseShift", "sle_AddMeter", "sle_UpdateMeterReading",
    "sle_UpdateMeterReadings", "sle_GetMeterReading", "sle_GetMeterEstimate", "sle_GetMeterResidual",
    "sle_UploadModel", "sle_UpdateTelemetry", "sle_UpdateMeasurement", "sle_UpdateMeasurementBatch",
    "sle_SyncMeasurementsToDevice", "sle_GetMeasurementDetails", "sle_GetMeasurementIdByIndex",
    "sle_SetSwitchStatus", "sle_GetSwitchStatus", "sle_ApplySwitchChanges", "sle_Solve", "sle_SolveWithConfig",
    "sle_CheckObservability", "sle_ApplyFlatStart", "sle_GetVoltageMagnitudes", "sle_GetVoltageAngles",
    "sle_GetBusVoltage", "sle_GetBusAngle", "sle_GetPowerInjections", "sle_GetBranchFlows", "sle_GetBranchP",
    "sle_GetBranchQ", "sle_GetBusResiduals", "sle_GetBranchResiduals", "sle_GetResiduals", "sle_GetObjective",
    "sle_IdentifyBadData", "sle_GetBusCount", "sle_GetBranchCount", "sle_GetMeasurementCount", "sle_GetMeterCount",
    "sle_GetGPUMemoryUsage", "sle_GetLastSolveTime", "sle_LoadIEEE14Bus", "sle_GetVersion", "sle_GetCUDAVersion",
    "sle_IsGPUAvailable",
]
EXT_FUNCTIONS = [
    "sle_ext_Solve", "sle_ext_UpdateTelemetryIndexed", "sle_ext_UpdateTelemetryF64", "sle_ext_SetState",
    "sle_ext_GetState", "sle_ext_GetResults", "sle_ext_SetBranchStatus", "sle_ext_GetSizes", "sle_ext_GetYbus",
    "sle_ext_GetHPattern", "sle_ext_GetGPattern", "sle_ext_GetOrdering", "sle_ext_Evaluate",
    "sle_ext_JacobianGain", "sle_ext_Factor", "sle_ext_SolveFactor", "sle_ext_StageTelemetry",
    "sle_ext_ApplyStagedTelemetry", "sle_ext_GetKernelTimes", "sle_ext_ResetKernelTimes", "sle_ext_SetOption",
    "sle_ext_Describe", "sle_ext_AutotuneSolve", "sle_ext_GetSolvePlanSizes", "sle_ext_SolveNoAlloc",
    "sle_ext_NormalizedResiduals", "sle_ext_SetMeasurementActive", "sle_ext_GainRhs",
]
KERNEL_CLASSES = ["power", "meas", "spgemm", "factor", "rhs", "fwd", "bwd", "update", "pcg",
                  "sub_fwd", "pi_fwd", "pi_bwd", "sub_bwd", "mid_fwd", "mid_bwd", "meas_lead", "finish"]
PLAN_SIZES = ["subtrees", "subtree_entries", "subtree_update_rows", "top_supernodes", "top_columns",
              "top_block_entries", "top_row_entries", "segments", "top_path_rows", "rhs_by_location",
              "bottom_subtrees", "bottom_entries", "bottom_update_rows", "bottom_columns"]


class Library:
    """One loaded flavour of the library (float or double SLE_Real)."""

    def __init__(self, double: bool = True):
        name = "libsle_f64.so" if double else "libsle.so"
        path = os.path.join(LIB_DIR, name)
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built; run python -m state_estimation_with_cuda_amd.build")
        self.path = path
        self.double = double
        self.real = C.c_double if double else C.c_float
        self.np_real = np.float64 if double else np.float32
        self.L = C.CDLL(path)
        self.S = _structs(self.real)
        self._sign()

    def _sign(self):
        L, S, R = self.L, self.S, self.real
        H = C.c_void_p
        P = C.POINTER
        i32, i64 = C.c_int32, C.c_int64
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        sig = {
            "sle_Create": (C.c_int, [P(H)]),
            "sle_CreateWithConfig": (C.c_int, [P(H), P(S["Config"])]),
            "sle_CreateFromConfigFile": (C.c_int, [P(H), C.c_char_p]),
            "sle_Initialize": (C.c_int, [H]),
            "sle_Destroy": (None, [H]),
            "sle_LoadConfigFile": (C.c_int, [C.c_char_p, P(S["Config"])]),
            "sle_SaveConfigFile": (C.c_int, [C.c_char_p, P(S["Config"])]),
            "sle_ReloadConfig": (C.c_int, [H]),
            "sle_ReloadConfigFromFile": (C.c_int, [H, C.c_char_p]),
            "sle_GetDefaultConfigPath": (C.c_char_p, []),
            "sle_GetConfigPath": (C.c_char_p, [H]),
            "sle_EnsureConfigFile": (C.c_int, [C.c_char_p]),
            "sle_GetConfig": (C.c_int, [H, P(S["Config"])]),
            "sle_SetConfig": (C.c_int, [H, P(S["Config"])]),
            "sle_GetLastError": (C.c_char_p, [H]),
            "sle_GetDeviceInfo": (C.c_char_p, [H]),
            "sle_AddBus": (C.c_int, [H, P(S["BusInfo"])]),
            "sle_AddBuses": (C.c_int, [H, C.c_void_p, i32]),
            "sle_AddBranch": (C.c_int, [H, P(S["BranchInfo"])]),
            "sle_AddBranches": (C.c_int, [H, C.c_void_p, i32]),
            "sle_AddMeasurement": (C.c_int, [H, P(S["MeasurementInfo"])]),
            "sle_AddMeasurements": (C.c_int, [H, C.c_void_p, i32]),
            "sle_AddSwitch": (C.c_int, [H, P(S["SwitchInfo"])]),
            "sle_RemoveBus": (C.c_int, [H, C.c_char_p]),
            "sle_RemoveBranch": (C.c_int, [H, C.c_char_p]),
            "sle_UpdateBranch": (C.c_int, [H, P(S["BranchInfo"])]),
            "sle_SetTransformerTap": (C.c_int, [H, C.c_char_p, R]),
            "sle_SetTransformerPhaseShift": (C.c_int, [H, C.c_char_p, R]),
            "sle_SetBranchImpedance": (C.c_int, [H, C.c_char_p, R, R, R, R]),
            "sle_GetTransformerTap": (R, [H, C.c_char_p]),
            "sle_GetTransformerPhaseShift": (R, [H, C.c_char_p]),
            "sle_AddMeter": (C.c_int, [H, P(S["MeterInfo"])]),
            "sle_UpdateMeterReading": (C.c_int, [H, C.c_char_p, C.c_char_p, R]),
            "sle_UpdateMeterReadings": (C.c_int, [H, P(S["MeterReading"]), i32, i32]),
            "sle_GetMeterReading": (C.c_int, [H, C.c_char_p, C.c_char_p, P(R)]),
            "sle_GetMeterEstimate": (C.c_int, [H, C.c_char_p, C.c_char_p, P(R)]),
            "sle_GetMeterResidual": (C.c_int, [H, C.c_char_p, C.c_char_p, P(R)]),
            "sle_UploadModel": (C.c_int, [H]),
            "sle_UpdateTelemetry": (C.c_int, [H, C.c_void_p, i32]),
            "sle_UpdateMeasurement": (C.c_int, [H, C.c_char_p, R]),
            "sle_UpdateMeasurementBatch": (C.c_int, [H, P(S["MeasurementUpdate"]), i32, i32]),
            "sle_SyncMeasurementsToDevice": (C.c_int, [H]),
            "sle_GetMeasurementDetails": (C.c_int, [H, C.c_char_p, P(S["MeasurementDetails"])]),
            "sle_GetMeasurementIdByIndex": (C.c_char_p, [H, i32]),
            "sle_SetSwitchStatus": (C.c_int, [H, C.c_char_p, C.c_int]),
            "sle_GetSwitchStatus": (C.c_int, [H, C.c_char_p]),
            "sle_ApplySwitchChanges": (C.c_int, [H]),
            "sle_Solve": (C.c_int, [H, C.c_int, P(S["Result"])]),
            "sle_SolveWithConfig": (C.c_int, [H, P(S["SolverConfig"]), P(S["Result"])]),
            "sle_CheckObservability": (C.c_int, [H, P(i32), P(i32)]),
            "sle_ApplyFlatStart": (C.c_int, [H]),
            "sle_GetVoltageMagnitudes": (C.c_int, [H, C.c_void_p, i32]),
            "sle_GetVoltageAngles": (C.c_int, [H, C.c_void_p, i32]),
            "sle_GetBusVoltage": (R, [H, C.c_char_p]),
            "sle_GetBusAngle": (R, [H, C.c_char_p]),
            "sle_GetPowerInjections": (C.c_int, [H, C.c_void_p, C.c_void_p, i32]),
            "sle_GetBranchFlows": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, i32]),
            "sle_GetBranchP": (R, [H, C.c_char_p, C.c_int]),
            "sle_GetBranchQ": (R, [H, C.c_char_p, C.c_int]),
            "sle_GetBusResiduals": (C.c_int, [H, C.c_char_p, P(R), P(R), P(R), P(i32)]),
            "sle_GetBranchResiduals": (C.c_int, [H, C.c_char_p, C.c_int, P(R), P(R), P(R), P(i32)]),
            "sle_GetResiduals": (C.c_int, [H, C.c_void_p, i32]),
            "sle_GetObjective": (R, [H]),
            "sle_IdentifyBadData": (C.c_int, [H, R, C.c_void_p, i32, P(i32)]),
            "sle_GetBusCount": (i32, [H]),
            "sle_GetBranchCount": (i32, [H]),
            "sle_GetMeasurementCount": (i32, [H]),
            "sle_GetMeterCount": (i32, [H]),
            "sle_GetGPUMemoryUsage": (C.c_size_t, [H]),
            "sle_GetLastSolveTime": (R, [H]),
            "sle_LoadIEEE14Bus": (C.c_int, [H]),
            "sle_GetVersion": (C.c_char_p, []),
            "sle_GetCUDAVersion": (i32, []),
            "sle_IsGPUAvailable": (i32, []),
            "sle_ext_Solve": (C.c_int, [H, P(S["SolverConfig"]), P(ExtResult)]),
            "sle_ext_SolveNoAlloc": (C.c_int, [H, P(S["SolverConfig"]), P(ExtResult)]),
            "sle_ext_NormalizedResiduals": (C.c_int, [H, dp, dp, i32]),
            "sle_ext_SetMeasurementActive": (C.c_int, [H, i32, i32]),
            "sle_ext_GainRhs": (C.c_int, [H, i32, dp]),
            "sle_ext_UpdateTelemetryIndexed": (C.c_int, [H, ip, dp, i32]),
            "sle_ext_UpdateTelemetryF64": (C.c_int, [H, dp, i32]),
            "sle_ext_SetState": (C.c_int, [H, dp, dp, i32]),
            "sle_ext_GetState": (C.c_int, [H, dp, dp, i32]),
            "sle_ext_GetResults": (C.c_int, [H] + [C.c_void_p] * 8),
            "sle_ext_SetBranchStatus": (C.c_int, [H, i32, C.c_int]),
            "sle_ext_GetSizes": (C.c_int, [H, lp]),
            "sle_ext_GetYbus": (C.c_int, [H, ip, ip, C.c_void_p, C.c_void_p]),
            "sle_ext_GetHPattern": (C.c_int, [H, ip, ip]),
            "sle_ext_GetGPattern": (C.c_int, [H, ip, ip]),
            "sle_ext_GetOrdering": (C.c_int, [H, ip, ip, ip, ip, ip]),
            "sle_ext_Evaluate": (C.c_int, [H, dp, dp, dp, dp, dp, dp, dp]),
            "sle_ext_JacobianGain": (C.c_int, [H, dp, dp]),
            "sle_ext_Factor": (C.c_int, [H, lp, ip, dp]),
            "sle_ext_SolveFactor": (C.c_int, [H, dp, dp]),
            "sle_ext_StageTelemetry": (C.c_int, [H, i32, ip, dp, i32]),
            "sle_ext_ApplyStagedTelemetry": (C.c_int, [H, i32]),
            "sle_ext_GetKernelTimes": (C.c_int, [H, dp, lp, lp, i32]),
            "sle_ext_ResetKernelTimes": (C.c_int, [H]),
            "sle_ext_SetOption": (C.c_int, [H, C.c_char_p, C.c_double]),
            "sle_ext_Describe": (C.c_char_p, [H]),
            "sle_ext_AutotuneSolve": (C.c_int, [H, i32]),
            "sle_ext_GetSolvePlanSizes": (C.c_int, [H, lp, i32]),
        }
        missing = [n for n in API_FUNCTIONS + EXT_FUNCTIONS if n not in sig]
        assert not missing, missing
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    def version(self) -> str:
        return self.L.sle_GetVersion().decode()

    def gpu_available(self) -> bool:
        return bool(self.L.sle_IsGPUAvailable())


_LIBS: dict[bool, Library] = {}


def library(double: bool = True) -> Library:
    if double not in _LIBS:
        _LIBS[double] = Library(double)
    return _LIBS[double]


def _id_buffer(prefix: str, n: int):
    """Contiguous fixed-width NUL-terminated ids "<prefix><i>" and the address of each.

    The digits come from integer arithmetic on whole arrays (no string
    operations per element), so a large model's ids take about a millisecond
    and other Python threads, such as a serving loop's, are not held up."""
    pb = prefix.encode()
    nd = len(str(max(n - 1, 0)))
    w = len(pb) + nd + 1
    arr = np.zeros((max(n, 1), w), np.uint8)
    arr[:, : len(pb)] = np.frombuffer(pb, np.uint8)
    for d in range(1, nd + 1):  # the numbers with d digits form one range
        lo, hi = (0 if d == 1 else 10 ** (d - 1)), min(10**d, n)
        if lo >= hi:
            continue
        x = np.arange(lo, hi, dtype=np.int32)[:, None]
        pw = (10 ** np.arange(d - 1, -1, -1)).astype(np.int32)[None, :]
        arr[lo:hi, len(pb) : len(pb) + d] = (48 + (x // pw) % 10).astype(np.uint8)
    base = arr.ctypes.data
    return arr, base + np.arange(n, dtype=np.uint64) * np.uint64(w)


class Engine:
    """One SLE_Handle (sle_Create + sle_Initialize)."""

    def __init__(self, double: bool = True, config: dict | None = None, initialize: bool = True):
        self.lib = library(double)
        self.L = self.lib.L
        self.S = self.lib.S
        self.h = C.c_void_p()
        if config is not None:
            cfg = self.default_config()
            for k, v in config.items():
                setattr(cfg, k, v)
            self._check(self.L.sle_CreateWithConfig(C.byref(self.h), C.byref(cfg)), "sle_CreateWithConfig")
        else:
            self._check(self.L.sle_Create(C.byref(self.h)), "sle_Create")
        self._keep = []
        if initialize:
            self._check(self.L.sle_Initialize(self.h), "sle_Initialize")

    def default_config(self):
        cfg = self.S["Config"]()
        cfg.device_id = 0
        cfg.max_buses, cfg.max_branches, cfg.max_measurements, cfg.max_switching_devices = (
            400_000, 800_000, 4_000_000, 100_000)
        cfg.enable_robust = 0
        cfg.huber_gamma = 1.5
        cfg.convergence_tolerance = 1e-4
        cfg.max_realtime_iterations = 5
        cfg.max_precision_iterations = 100
        cfg.realtime_time_limit_ms = 20.0
        cfg.use_flat_start_default = 0
        cfg.enable_profiling = 0
        cfg.use_cuda_graphs = 1
        cfg.block_size_standard = 256
        cfg.block_size_reduction = 512
        cfg.use_double_precision = 0
        cfg.voltage_min_pu = 0.5
        cfg.voltage_max_pu = 1.5
        cfg.bad_data_threshold = 3.0
        cfg.auto_bad_data_rejection = 0
        return cfg

    def close(self):
        if self.h:
            self.L.sle_Destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self) -> str:
        s = self.L.sle_GetLastError(self.h)
        return s.decode() if s else ""

    def _check(self, rc, where):
        if rc != OK:
            raise SLEError(rc, where, self.last_error() if self.h else "")
        return rc

    # -- model -------------------------------------------------------------
    def load_grid(self, g, switches: bool = False):
        """Add buses, branches and measurements of a gridgen.Grid in insertion order.

        Ids: buses "B<i>", branches "L<k>", measurements "M<i>"; with
        ``switches`` every branch gets a switching device "S<k>"."""
        real = self.lib.np_real
        n, nb, m = g.n, g.nb, g.m
        cfg = self.get_config()
        need = {"max_buses": n, "max_branches": nb, "max_measurements": m,
                "max_switching_devices": nb if switches else 0}
        grow = {k: v for k, v in need.items() if v > cfg[k]}
        if grow:
            self.set_config(**grow)  # capacity limits of SLE_Config
        bid, bptr = _id_buffer("B", n)
        lid, lptr = _id_buffer("L", nb)
        mid, mptr = _id_buffer("M", m)
        self._keep += [bid, lid, mid]
        # the record arrays are numpy allocations passed by address (a ctypes
        # array of that many structs is built with the interpreter lock held)
        B = self.S["BusInfo"]
        arr = self._records(B, n)
        arr["id"] = bptr
        arr["base_kv"] = 100.0
        arr["type"] = g.bus_type
        arr["v_setpoint"] = g.vset.astype(real)
        arr["p_scheduled"] = 0
        arr["q_scheduled"] = 0
        self._check(self.L.sle_AddBuses(self.h, C.c_void_p(arr.ctypes.data), n), "sle_AddBuses")
        if nb:
            Br = self.S["BranchInfo"]
            a = self._records(Br, nb)
            a["id"] = lptr
            a["from_bus"] = bptr[g.br_from]
            a["to_bus"] = bptr[g.br_to]
            a["resistance"] = g.r
            a["reactance"] = g.x
            a["susceptance"] = g.b
            a["conductance"] = 0
            a["tap_ratio"] = g.tap
            a["phase_shift"] = g.shift
            a["is_transformer"] = g.is_xfmr
            a["sd_id"] = 0
            self._check(self.L.sle_AddBranches(self.h, C.c_void_p(a.ctypes.data), nb), "sle_AddBranches")
        M = self.S["MeasurementInfo"]
        a = self._records(M, m)
        a["id"] = mptr
        a["type"] = g.m_type
        bus_meas = np.isin(g.m_type, (0, 1, 2, 3, 7, 8))
        loc = np.where(bus_meas, bptr[np.minimum(g.m_loc, n - 1)], lptr[np.minimum(g.m_loc, max(nb - 1, 0))] if nb else 0)
        a["location"] = loc
        a["branch_end"] = g.m_end
        a["sigma"] = g.sigma
        a["pt_ratio"] = 1
        a["ct_ratio"] = 1
        a["is_pseudo"] = np.isin(g.m_type, (7, 8)).astype(np.int32)
        self._check(self.L.sle_AddMeasurements(self.h, C.c_void_p(a.ctypes.data), m), "sle_AddMeasurements")
        if switches and nb:
            sid, sptr = _id_buffer("S", nb)
            self._keep.append(sid)
            SW = self.S["SwitchInfo"]
            for k in range(nb):
                info = SW(C.c_char_p(int(sptr[k])) if False else None, None, 1)
                info.id = C.cast(C.c_void_p(int(sptr[k])), C.c_char_p)
                info.branch_id = C.cast(C.c_void_p(int(lptr[k])), C.c_char_p)
                info.status = int(g.status[k])
                self._check(self.L.sle_AddSwitch(self.h, C.byref(info)), "sle_AddSwitch")
        self.n, self.nb, self.m = n, nb, m

    @classmethod
    def _records(cls, struct, count):
        """Zeroed array of `count` records laid out as the C struct (size and
        field offsets of the ctypes mirror)."""
        return np.zeros(max(count, 1), dtype=cls._dtype(struct))[:count]

    @staticmethod
    def _dtype(struct):
        names, formats, offsets = [], [], []
        for name, ct in struct._fields_:
            names.append(name)
            off = getattr(struct, name).offset
            offsets.append(off)
            if ct is C.c_char_p:
                formats.append(np.uint64)
            elif ct is C.c_double:
                formats.append(np.float64)
            elif ct is C.c_float:
                formats.append(np.float32)
            else:
                formats.append(np.int32)
        return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": C.sizeof(struct)})

    def upload(self):
        self._check(self.L.sle_UploadModel(self.h), "sle_UploadModel")

    def set_telemetry(self, z):
        z = np.ascontiguousarray(z, dtype=self.lib.np_real)
        self._check(self.L.sle_UpdateTelemetry(self.h, z.ctypes.data, int(z.shape[0])), "sle_UpdateTelemetry")

    def set_telemetry_f64(self, z):
        z = np.ascontiguousarray(z, dtype=np.float64)
        self._check(self.L.sle_ext_UpdateTelemetryF64(self.h, z, int(z.shape[0])), "sle_ext_UpdateTelemetryF64")

    def set_telemetry_indexed(self, idx, val):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        self._check(self.L.sle_ext_UpdateTelemetryIndexed(self.h, idx, val, int(idx.shape[0])),
                    "sle_ext_UpdateTelemetryIndexed")

    def set_config(self, **kw):
        cfg = self.S["Config"]()
        self._check(self.L.sle_GetConfig(self.h, C.byref(cfg)), "sle_GetConfig")
        for k, v in kw.items():
            setattr(cfg, k, v)
        self._check(self.L.sle_SetConfig(self.h, C.byref(cfg)), "sle_SetConfig")

    def get_config(self):
        cfg = self.S["Config"]()
        self._check(self.L.sle_GetConfig(self.h, C.byref(cfg)), "sle_GetConfig")
        return {k: getattr(cfg, k) for k, _ in cfg._fields_}

    # -- solve ---------------------------------------------------------------
    def solve(self, mode=MODE_PRECISION, max_iterations=100, tolerance=1e-6, robust=False, gamma=1.5,
              flat_start=False, time_limit_ms=0.0, check=False, no_alloc=False):
        """sle_ext_Solve, or with no_alloc its allocation-free variant sle_ext_SolveNoAlloc."""
        sc = self.S["SolverConfig"](int(mode), int(max_iterations), tolerance, gamma, int(bool(robust)),
                                    int(bool(flat_start)), time_limit_ms)
        r = ExtResult()
        fn = self.L.sle_ext_SolveNoAlloc if no_alloc else self.L.sle_ext_Solve
        rc = fn(self.h, C.byref(sc), C.byref(r))
        if check:
            self._check(rc, "sle_ext_SolveNoAlloc" if no_alloc else "sle_ext_Solve")
        return dict(rc=rc, status=r.status, iterations=r.iterations, refactors=r.refactors,
                    fp64_refactors=r.fp64_refactors, max_mismatch=r.max_mismatch, objective=r.objective,
                    largest_residual=r.largest_residual, largest_residual_idx=r.largest_residual_idx, flags=r.flags,
                    pivots_substituted=r.pivots_substituted, pcg_steps=r.pcg_steps,
                    graph_captures=r.graph_captures, seq_fallbacks=r.seq_fallbacks, retries=r.retries,
                    ahead=r.ahead, ahead_idle=r.ahead_idle, gpu_ms=r.gpu_ms, host_ms=r.host_ms)

    def normalized_residuals(self):
        """(rn, omega) per measurement: rn = |r| / sqrt(Omega_ii), Omega = R - H G^-1 H^T."""
        rn = np.zeros(self.m)
        om = np.zeros(self.m)
        self._check(self.L.sle_ext_NormalizedResiduals(self.h, rn, om, self.m), "sle_ext_NormalizedResiduals")
        return rn, om

    def set_measurement_active(self, i, active):
        self._check(self.L.sle_ext_SetMeasurementActive(self.h, int(i), int(bool(active))),
                    "sle_ext_SetMeasurementActive")

    def identify_bad_data(self, threshold=0.0, max_count=None):
        n = self.m if max_count is None else int(max_count)
        idx = np.zeros(max(n, 1), np.int32)
        cnt = C.c_int32(0)
        self._check(self.L.sle_IdentifyBadData(self.h, threshold, idx.ctypes.data_as(C.c_void_p), n,
                                                C.byref(cnt)), "sle_IdentifyBadData")
        return idx[:min(cnt.value, n)]

    def solve_mode(self, mode):
        res = self.S["Result"]()
        rc = self.L.sle_Solve(self.h, int(mode), C.byref(res))
        return rc, res

    def state(self, out=None):
        """(V, theta) in insertion order; out=(V, theta) fills caller arrays instead of new ones."""
        V, th = out if out is not None else (np.empty(self.n), np.empty(self.n))
        self._check(self.L.sle_ext_GetState(self.h, V, th, self.n), "sle_ext_GetState")
        return V, th

    def set_state(self, V, th):
        self._check(self.L.sle_ext_SetState(self.h, np.ascontiguousarray(V, np.float64),
                                            np.ascontiguousarray(th, np.float64), self.n), "sle_ext_SetState")

    def results(self):
        out = {k: np.zeros(self.n) for k in ("P", "Q")}
        out.update({k: np.zeros(self.nb) for k in ("Pf", "Qf", "Pt", "Qt")})
        out.update({k: np.zeros(self.m) for k in ("h", "r")})
        ptrs = [out[k].ctypes.data for k in ("P", "Q", "Pf", "Qf", "Pt", "Qt", "h", "r")]
        self._check(self.L.sle_ext_GetResults(self.h, *ptrs), "sle_ext_GetResults")
        return out

    def describe(self) -> str:
        d = self.L.sle_ext_Describe(self.h)
        return d.decode() if d else ""

    def gpu_bytes(self) -> int:
        """Device memory held by this engine (sle_GetGPUMemoryUsage)."""
        return int(self.L.sle_GetGPUMemoryUsage(self.h))

    def autotune(self, repetitions=5):
        """Time the solve plan candidates on the current factor and keep the fastest."""
        self._check(self.L.sle_ext_AutotuneSolve(self.h, int(repetitions)), "sle_ext_AutotuneSolve")
        return self.describe()

    def set_option(self, key, value):
        self._check(self.L.sle_ext_SetOption(self.h, key.encode(), float(value)), "sle_ext_SetOption")

    def stage_telemetry(self, slot, idx, val):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        self._check(self.L.sle_ext_StageTelemetry(self.h, int(slot), idx, val, int(idx.shape[0])),
                    "sle_ext_StageTelemetry")

    def apply_staged(self, slot):
        self._check(self.L.sle_ext_ApplyStagedTelemetry(self.h, int(slot)), "sle_ext_ApplyStagedTelemetry")

    def kernel_times(self):
        n = len(KERNEL_CLASSES)
        ms = np.zeros(n)
        la = np.zeros(n, np.int64)
        ca = np.zeros(n, np.int64)
        self._check(self.L.sle_ext_GetKernelTimes(self.h, ms, la, ca, n), "sle_ext_GetKernelTimes")
        return {k: dict(ms=float(ms[i]), launches=int(la[i]), calls=int(ca[i])) for i, k in enumerate(KERNEL_CLASSES)}

    def solve_plan_sizes(self):
        out = np.zeros(len(PLAN_SIZES), np.int64)
        self._check(self.L.sle_ext_GetSolvePlanSizes(self.h, out, len(PLAN_SIZES)), "sle_ext_GetSolvePlanSizes")
        return {k: int(out[i]) for i, k in enumerate(PLAN_SIZES)}

    def reset_kernel_times(self):
        self._check(self.L.sle_ext_ResetKernelTimes(self.h), "sle_ext_ResetKernelTimes")

    def set_branch_status(self, k, status):
        self._check(self.L.sle_ext_SetBranchStatus(self.h, int(k), int(status)), "sle_ext_SetBranchStatus")

    # -- introspection ---------------------------------------------------------
    def sizes(self):
        s = np.zeros(12, np.int64)
        self._check(self.L.sle_ext_GetSizes(self.h, s), "sle_ext_GetSizes")
        keys = ["ns", "nnz_y", "nnz_h", "nnz_g", "nnz_l", "n_snodes", "n_levels", "max_front", "g_terms",
                "front_elems", "flops", "symbolic_us"]
        return dict(zip(keys, (int(v) for v in s)))

    def ybus(self, values=True):
        sz = self.sizes()
        rp = np.zeros(self.n + 1, np.int32)
        col = np.zeros(sz["nnz_y"], np.int32)
        if not values:
            self._check(self.L.sle_ext_GetYbus(self.h, rp, col, None, None), "sle_ext_GetYbus")
            return rp, col
        G = np.zeros(sz["nnz_y"])
        B = np.zeros(sz["nnz_y"])
        self._check(self.L.sle_ext_GetYbus(self.h, rp, col, G.ctypes.data, B.ctypes.data), "sle_ext_GetYbus")
        return rp, col, G, B

    def h_pattern(self):
        sz = self.sizes()
        rp = np.zeros(self.m + 1, np.int32)
        col = np.zeros(sz["nnz_h"], np.int32)
        self._check(self.L.sle_ext_GetHPattern(self.h, rp, col), "sle_ext_GetHPattern")
        return rp, col

    def g_pattern(self):
        sz = self.sizes()
        cp = np.zeros(sz["ns"] + 1, np.int32)
        row = np.zeros(sz["nnz_g"], np.int32)
        self._check(self.L.sle_ext_GetGPattern(self.h, cp, row), "sle_ext_GetGPattern")
        return cp, row

    def ordering(self):
        sz = self.sizes()
        bp = np.zeros(self.n, np.int32)
        perm = np.zeros(sz["ns"], np.int32)
        par = np.zeros(sz["ns"], np.int32)
        cc = np.zeros(sz["ns"], np.int32)
        sf = np.zeros(sz["n_snodes"] + 1, np.int32)
        self._check(self.L.sle_ext_GetOrdering(self.h, bp, perm, par, cc, sf), "sle_ext_GetOrdering")
        return dict(bus_perm=bp, perm=perm, parent=par, colcount=cc, snode_ptr=sf)

    def evaluate(self):
        out = dict(h=np.zeros(self.m), P=np.zeros(self.n), Q=np.zeros(self.n))
        for k in ("Pf", "Qf", "Pt", "Qt"):
            out[k] = np.zeros(max(self.nb, 1))
        self._check(self.L.sle_ext_Evaluate(self.h, out["h"], out["P"], out["Q"], out["Pf"], out["Qf"], out["Pt"],
                                            out["Qt"]), "sle_ext_Evaluate")
        return out

    def gain_rhs(self, by_location=True):
        ns = self.sizes()["ns"]
        b = np.zeros(ns)
        self._check(self.L.sle_ext_GainRhs(self.h, int(bool(by_location)), b), "sle_ext_GainRhs")
        return b

    def jacobian_gain(self):
        sz = self.sizes()
        hv = np.zeros(max(sz["nnz_h"], 1))
        gv = np.zeros(max(sz["nnz_g"], 1))
        self._check(self.L.sle_ext_JacobianGain(self.h, hv, gv), "sle_ext_JacobianGain")
        return hv[: sz["nnz_h"]], gv[: sz["nnz_g"]]

    def factor(self):
        sz = self.sizes()
        cp = np.zeros(sz["ns"] + 1, np.int64)
        row = np.zeros(sz["nnz_l"], np.int32)
        val = np.zeros(sz["nnz_l"])
        rc = self.L.sle_ext_Factor(self.h, cp, row, val)
        return rc, cp, row, val

    def solve_factor(self, b):
        b = np.ascontiguousarray(b, np.float64)
        dx = np.zeros_like(b)
        self._check(self.L.sle_ext_SolveFactor(self.h, b, dx), "sle_ext_SolveFactor")
        return dx


def gpu_available() -> bool:
    try:
        return library(True).gpu_available()
    except ImportError:
        return False
