"""ctypes wrapper of the FP64 oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The shipped library never uses
it.  Build with ``make -f oracle/Makefile`` from the repository root.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OracleModel(C.Structure):
    _fields_ = [
        ("n_bus", C.c_int32),
        ("n_branch", C.c_int32),
        ("n_meas", C.c_int32),
        ("slack", C.c_int32),
        ("bus_type", C.c_void_p),
        ("vset", C.c_void_p),
        ("br_from", C.c_void_p),
        ("br_to", C.c_void_p),
        ("br_r", C.c_void_p),
        ("br_x", C.c_void_p),
        ("br_b", C.c_void_p),
        ("br_tap", C.c_void_p),
        ("br_shift", C.c_void_p),
        ("br_status", C.c_void_p),
        ("m_type", C.c_void_p),
        ("m_loc", C.c_void_p),
        ("m_end", C.c_void_p),
        ("m_sigma", C.c_void_p),
        ("m_z", C.c_void_p),
    ]


class SolveParams(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("max_iter", C.c_int32),
        ("tol", C.c_double),
        ("robust", C.c_int32),
        ("gamma", C.c_double),
        ("flat_start", C.c_int32),
        ("vmin", C.c_double),
        ("vmax", C.c_double),
    ]


class SolveResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("iterations", C.c_int32),
        ("refactors", C.c_int32),
        ("max_mismatch", C.c_double),
        ("objective", C.c_double),
        ("seconds", C.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-f", os.path.join(_HERE, "Makefile")], check=True, cwd=os.path.dirname(_HERE))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(OracleModel)]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_set_z.argtypes = [C.c_void_p, f64p]
        L.or_set_branch_status.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.or_set_state.argtypes = [C.c_void_p, f64p, f64p]
        L.or_get_state.argtypes = [C.c_void_p, f64p, f64p]
        L.or_sizes.argtypes = [C.c_void_p, i64p]
        L.or_get_ybus.argtypes = [C.c_void_p, i32p, i32p, f64p, f64p]
        L.or_get_h_pattern.argtypes = [C.c_void_p, i32p, i32p]
        L.or_get_g_pattern.argtypes = [C.c_void_p, i32p, i32p]
        L.or_get_symbolic.argtypes = [C.c_void_p, i32p, i32p, i32p, i32p, i32p]
        L.or_eval.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p]
        L.or_jacobian.argtypes = [C.c_void_p, f64p, f64p, f64p]
        L.or_gain.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p]
        L.or_rhs.argtypes = [C.c_void_p, f64p, f64p, f64p]
        L.or_factor.restype = C.c_int32
        L.or_factor.argtypes = [C.c_void_p, f64p, i64p, i32p, f64p]
        L.or_solve.argtypes = [C.c_void_p, C.POINTER(SolveParams), C.POINTER(SolveResult)]
        L.or_get_residuals.argtypes = [C.c_void_p, f64p]
        L.or_normalized_residuals.restype = C.c_int32
        L.or_normalized_residuals.argtypes = [C.c_void_p, f64p, f64p]
        L.or_num_threads.restype = C.c_int32
        L.or_set_num_threads.argtypes = [C.c_int32]
        _lib = L
    return _lib


class Oracle:
    """One oracle instance (model + symbolic analysis + cached factor + state)."""

    def __init__(self, g):
        L = lib()
        self.g = g
        self._keep = [
            np.ascontiguousarray(g.bus_type, np.int32),
            np.ascontiguousarray(g.vset, np.float64),
            np.ascontiguousarray(g.br_from, np.int32),
            np.ascontiguousarray(g.br_to, np.int32),
            np.ascontiguousarray(g.r, np.float64),
            np.ascontiguousarray(g.x, np.float64),
            np.ascontiguousarray(g.b, np.float64),
            np.ascontiguousarray(g.tap, np.float64),
            np.ascontiguousarray(g.shift, np.float64),
            np.ascontiguousarray(g.status, np.uint8),
            np.ascontiguousarray(g.m_type, np.int32),
            np.ascontiguousarray(g.m_loc, np.int32),
            np.ascontiguousarray(g.m_end, np.int32),
            np.ascontiguousarray(g.sigma, np.float64),
            np.ascontiguousarray(g.z, np.float64),
        ]
        k = self._keep
        md = OracleModel(g.n, g.nb, g.m, g.slack, *[a.ctypes.data for a in k])
        self.h = C.c_void_p(L.or_create(C.byref(md)))
        self.n, self.nb, self.m = g.n, g.nb, g.m
        self._refresh_sizes()

    def _refresh_sizes(self):
        s = np.zeros(6, np.int64)
        lib().or_sizes(self.h, s)
        self.ns, self.nnz_y, self.nnz_h, self.nnz_g, self.nnz_l, self.n_snodes = (int(v) for v in s)

    def __del__(self):
        try:
            if self.h:
                lib().or_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # -- data ---------------------------------------------------------------
    def set_z(self, z):
        lib().or_set_z(self.h, np.ascontiguousarray(z, np.float64))

    def set_branch_status(self, k, st):
        lib().or_set_branch_status(self.h, int(k), int(st))
        self._refresh_sizes()

    def set_state(self, V, th):
        lib().or_set_state(self.h, np.ascontiguousarray(V, np.float64), np.ascontiguousarray(th, np.float64))

    def state(self):
        V = np.zeros(self.n)
        th = np.zeros(self.n)
        lib().or_get_state(self.h, V, th)
        return V, th

    # -- symbolic -----------------------------------------------------------
    def ybus(self):
        rp = np.zeros(self.n + 1, np.int32)
        col = np.zeros(self.nnz_y, np.int32)
        G = np.zeros(self.nnz_y)
        B = np.zeros(self.nnz_y)
        lib().or_get_ybus(self.h, rp, col, G, B)
        return rp, col, G, B

    def h_pattern(self):
        rp = np.zeros(self.m + 1, np.int32)
        col = np.zeros(self.nnz_h, np.int32)
        lib().or_get_h_pattern(self.h, rp, col)
        return rp, col

    def g_pattern(self):
        cp = np.zeros(self.ns + 1, np.int32)
        row = np.zeros(self.nnz_g, np.int32)
        lib().or_get_g_pattern(self.h, cp, row)
        return cp, row

    def symbolic(self):
        bp = np.zeros(self.n, np.int32)
        perm = np.zeros(self.ns, np.int32)
        par = np.zeros(self.ns, np.int32)
        cc = np.zeros(self.ns, np.int32)
        sp = np.zeros(self.n_snodes + 1, np.int32)
        lib().or_get_symbolic(self.h, bp, perm, par, cc, sp)
        return dict(bus_perm=bp, perm=perm, parent=par, colcount=cc, snode_ptr=sp)

    # -- numerics -----------------------------------------------------------
    def eval(self, V, th):
        V = np.ascontiguousarray(V, np.float64)
        th = np.ascontiguousarray(th, np.float64)
        hx = np.zeros(self.m)
        P = np.zeros(self.n)
        Q = np.zeros(self.n)
        Pf, Qf, Pt, Qt = (np.zeros(self.nb) for _ in range(4))
        lib().or_eval(self.h, V, th, hx, P, Q, Pf, Qf, Pt, Qt)
        return dict(h=hx, P=P, Q=Q, Pf=Pf, Qf=Qf, Pt=Pt, Qt=Qt)

    def jacobian(self, V, th):
        hv = np.zeros(self.nnz_h)
        lib().or_jacobian(self.h, np.ascontiguousarray(V, np.float64), np.ascontiguousarray(th, np.float64), hv)
        return hv

    def rhs(self, V, th):
        """Gauss-Newton right-hand side H^T W r at (V, th), natural state order."""
        b = np.zeros(self.ns)
        lib().or_rhs(self.h, np.ascontiguousarray(V, np.float64), np.ascontiguousarray(th, np.float64), b)
        return b

    def gain(self, V, th, w):
        gv = np.zeros(self.nnz_g)
        lib().or_gain(
            self.h,
            np.ascontiguousarray(V, np.float64),
            np.ascontiguousarray(th, np.float64),
            np.ascontiguousarray(w, np.float64),
            gv,
        )
        return gv

    def factor(self, gv):
        lp = np.zeros(self.ns + 1, np.int64)
        lr = np.zeros(self.nnz_l, np.int32)
        lv = np.zeros(self.nnz_l)
        rc = lib().or_factor(self.h, np.ascontiguousarray(gv, np.float64), lp, lr, lv)
        if rc != 0:
            raise np.linalg.LinAlgError("oracle factor: non-positive pivot")
        return lp, lr, lv

    def solve(self, mode=1, max_iter=100, tol=1e-6, robust=False, gamma=1.5, flat_start=False, vmin=0.5, vmax=1.5):
        # robust: False/0 WLS, True/1 Huber IRLS (the reference's weights), 2 Huber with Newton gain weights
        p = SolveParams(int(mode), int(max_iter), float(tol), int(robust), float(gamma), int(bool(flat_start)),
                        float(vmin), float(vmax))
        r = SolveResult()
        lib().or_solve(self.h, C.byref(p), C.byref(r))
        return dict(
            status=r.status,
            iterations=r.iterations,
            refactors=r.refactors,
            max_mismatch=r.max_mismatch,
            objective=r.objective,
            seconds=r.seconds,
        )

    def normalized_residuals(self):
        """(rn, omega) at the current state: rn = |r| / sqrt(Omega_ii), Omega = R - H G^-1 H^T."""
        rn = np.zeros(self.m)
        om = np.zeros(self.m)
        rc = lib().or_normalized_residuals(self.h, rn, om)
        if rc != 0:
            raise np.linalg.LinAlgError("oracle normalized residuals: singular gain")
        return rn, om

    def residuals(self):
        out = np.zeros(self.m)
        lib().or_get_residuals(self.h, out)
        return out


def num_threads() -> int:
    return int(lib().or_num_threads())


def set_num_threads(n: int) -> None:
    lib().or_set_num_threads(int(n))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"
