"""Python binding of the FP64 CPU oracle (test infrastructure only).

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
checker of the HIP library; the product never imports this module.
"""
import ctypes
import os

import numpy as np

from . import build as _build

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

DBL_P = ctypes.POINTER(ctypes.c_double)
INT_P = ctypes.POINTER(ctypes.c_int)
FLT_P = ctypes.POINTER(ctypes.c_float)


def load(rebuild=False):
    global _lib
    if _lib is not None and not rebuild:
        return _lib
    path = _build.LIB
    if rebuild or not os.path.exists(path):
        _build.build()
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    sig = {
        "oracle_create": (ctypes.c_void_p, [P]),
        "oracle_destroy": (None, [ctypes.c_void_p]),
        "oracle_ibm": (None, [P, DBL_P, INT_P, INT_P, DBL_P, INT_P]),
        "oracle_rhs": (None, [P, DBL_P, DBL_P, DBL_P]),
        "oracle_turbulence": (None, [P, DBL_P, DBL_P]),
        "oracle_boundary_conditions": (None, [P, DBL_P]),
        "oracle_phi": (None, [ctypes.c_void_p, DBL_P, DBL_P]),
        "oracle_bounds": (None, [P, DBL_P]),
        "oracle_velocity_correction": (None, [P, DBL_P, DBL_P, DBL_P]),
        "oracle_pressure": (None, [ctypes.c_void_p, DBL_P, DBL_P, DBL_P, DBL_P, INT_P, INT_P]),
        "oracle_step": (ctypes.c_int, [ctypes.c_void_p, DBL_P, DBL_P, ctypes.c_int]),
        "oracle_run": (ctypes.c_int, [ctypes.c_void_p, DBL_P, DBL_P, ctypes.c_int, ctypes.c_int]),
        "oracle_pressure_solves": (ctypes.c_long, [ctypes.c_void_p]),
        "oracle_fft": (None, [DBL_P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "oracle_tridiagonal_coefficients": (None, [P, ctypes.c_int, ctypes.c_int, DBL_P, DBL_P, DBL_P]),
        "oracle_thomas": (None, [ctypes.c_int, DBL_P, DBL_P, DBL_P, DBL_P, DBL_P]),
        "oracle_threads": (ctypes.c_int, []),
        "oracle_set_closure": (None, [FLT_P]),
        "oracle_closure_eval": (None, [FLT_P, DBL_P, ctypes.c_long, DBL_P, INT_P]),
        "oracle_closure_gradient": (None, [FLT_P, DBL_P, DBL_P, ctypes.c_long, ctypes.c_double, FLT_P, DBL_P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DBL_P)


def _ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(INT_P)


def _pp(params):
    """Address of a ctypes Parameters structure (or pointer to one)."""
    if isinstance(params, ctypes._Pointer):
        return ctypes.cast(params, ctypes.c_void_p)
    return ctypes.cast(ctypes.pointer(params), ctypes.c_void_p)


class Oracle:
    """Workspace-owning oracle for one parameter set."""

    def __init__(self, params):
        self.lib = load()
        self.params = params
        self._pptr = _pp(params)
        self.ctx = self.lib.oracle_create(self._pptr)

    def close(self):
        if self.ctx:
            self.lib.oracle_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def phi(self, R_old):
        R_new = np.zeros_like(R_old)
        self.lib.oracle_phi(self.ctx, _dp(R_old), _dp(R_new))
        return R_new

    def rhs(self, R_old, R_new, R_turb):
        self.lib.oracle_rhs(self._pptr, _dp(R_old), _dp(R_new), _dp(R_turb))

    def turbulence(self, R_turb, R_new):
        self.lib.oracle_turbulence(self._pptr, _dp(R_turb), _dp(R_new))

    def boundary_conditions(self, R):
        self.lib.oracle_boundary_conditions(self._pptr, _dp(R))

    def bounds(self, y):
        self.lib.oracle_bounds(self._pptr, _dp(y))

    def velocity_correction(self, y_np1, y_n, p):
        self.lib.oracle_velocity_correction(self._pptr, _dp(y_np1), _dp(y_n), _dp(p))

    def pressure(self, y_np1, y_n, p):
        err = ctypes.c_double(0.0)
        it = ctypes.c_int(0)
        ns = ctypes.c_int(0)
        self.lib.oracle_pressure(self.ctx, _dp(y_np1), _dp(y_n), _dp(p), ctypes.byref(err), ctypes.byref(it),
                                 ctypes.byref(ns))
        return err.value, it.value, ns.value

    def step(self, y, p, n):
        rc = self.lib.oracle_step(self.ctx, _dp(y), _dp(p), n)
        if rc != 0:
            raise RuntimeError("oracle_step failed (%d)" % rc)

    def run(self, y, p, step0, n_steps):
        rc = self.lib.oracle_run(self.ctx, _dp(y), _dp(p), step0, n_steps)
        if rc != 0:
            raise RuntimeError("oracle_run failed (%d)" % rc)

    def pressure_solves(self):
        return self.lib.oracle_pressure_solves(self.ctx)


def ibm(params, topography):
    lib = load()
    Nx, Ny, Nz = params.Nx, params.Ny, params.Nz
    cut = np.zeros(Nx * Ny, dtype=np.int32)
    nzy = np.zeros(Nx * Ny, dtype=np.int32)
    zibm = np.zeros(Nx * Ny * Nz, dtype=np.float64)
    m = ctypes.c_int(0)
    topo = np.ascontiguousarray(topography, dtype=np.float64)
    lib.oracle_ibm(_pp(params), _dp(topo), _ip(cut), _ip(nzy), _dp(zibm), ctypes.byref(m))
    return cut, nzy, zibm, m.value


def fft(x, sign=-1):
    """Complex DFT of a 1-D array with the oracle's mixed-radix FFT."""
    lib = load()
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    buf = a.view(np.float64)
    lib.oracle_fft(_dp(buf), a.shape[0], 1, sign)
    return a


def tridiagonal_coefficients(params, r, s):
    lib = load()
    n = params.Nz - 1
    a = np.zeros(n, dtype=np.float64)
    b = np.zeros(n, dtype=np.float64)
    c = np.zeros(n, dtype=np.float64)
    lib.oracle_tridiagonal_coefficients(_pp(params), r, s, _dp(a), _dp(b), _dp(c))
    return a[: n - 1], b, c[: n - 1]


def thomas(a, b, c, d):
    lib = load()
    n = b.shape[0]
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    dd = np.ascontiguousarray(d, dtype=np.complex128).copy()
    x = np.zeros(n, dtype=np.complex128)
    lib.oracle_thomas(n, _dp(a), _dp(b), _dp(c), _dp(dd.view(np.float64)), _dp(x.view(np.float64)))
    return x


def threads():
    return load().oracle_threads()


_closure_keep = None


def set_closure(closure_params):
    """Learned closure in every later oracle RHS (None: Smagorinsky).  The
    oracle keeps a pointer: the array is held here until replaced."""
    global _closure_keep
    lib = load()
    if closure_params is None:
        _closure_keep = None
        lib.oracle_set_closure(None)
        return
    _closure_keep = np.ascontiguousarray(closure_params, dtype=np.float32)
    lib.oracle_set_closure(_closure_keep.ctypes.data_as(FLT_P))


def closure_eval(closure_params, features):
    prm = np.ascontiguousarray(closure_params, dtype=np.float32)
    f = np.ascontiguousarray(features, dtype=np.float64).reshape(-1, 12)
    nu = np.zeros(f.shape[0])
    learned = np.zeros(f.shape[0], dtype=np.int32)
    load().oracle_closure_eval(prm.ctypes.data_as(FLT_P), _dp(f), f.shape[0], _dp(nu), _ip(learned))
    return nu, learned


def closure_gradient(closure_params, features, nu_target, weight=None):
    """Loss weight * sum (softplus(z) - nu/out_scale)^2 and its gradient
    (parameter-block-shaped float32; weight defaults to 1/n)."""
    prm = np.ascontiguousarray(closure_params, dtype=np.float32)
    f = np.ascontiguousarray(features, dtype=np.float64).reshape(-1, 12)
    t = np.ascontiguousarray(nu_target, dtype=np.float64)
    w = 1.0 / len(t) if weight is None else weight
    grad = np.zeros(prm.size, dtype=np.float32)
    loss = ctypes.c_double(0.0)
    load().oracle_closure_gradient(prm.ctypes.data_as(FLT_P), _dp(f), _dp(t), len(t), w, grad.ctypes.data_as(FLT_P),
                                   ctypes.byref(loss))
    return grad, loss.value
