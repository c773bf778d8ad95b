"""ctypes wrapper of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")
# The threaded CPU baseline: the same FP32 restatement with sorted per-axis
# candidate lists and SWIFT's production flags (timed, never the checker).
_PORT = os.path.join(_HERE, "libcpuport.so")
_lib = None
_port = {}


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def port_lib(path: str = _PORT):
    """The CPU-baseline build at `path` (default: the prebuilt one)."""
    if path not in _port:
        if path == _PORT and not os.path.exists(_PORT):
            build()
        _port[path] = C.CDLL(path)
    return _port[path]


def _load(path):
    L = C.CDLL(path)
    L.orc_kernel_gamma.restype = C.c_float
    L.orc_kernel_gamma2.restype = C.c_float
    L.orc_kernel_root.restype = C.c_float
    L.orc_kernel_norm.restype = C.c_float
    L.orc_cbrt32.restype = C.c_float
    L.orc_cbrt32.argtypes = [C.c_float]
    L.orc_kernel_deval32.argtypes = [C.c_float, C.POINTER(C.c_float),
                                     C.POINTER(C.c_float)]
    L.orc_kernel_deval64.argtypes = [C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
    for fn in ("orc_iact_density32", "orc_iact_gradient32",
               "orc_iact_force32"):
        getattr(L, fn).argtypes = [C.c_float, C.c_void_p, C.c_float,
                                   C.c_float, C.c_void_p, C.c_void_p, C.c_int]
    return L


@contextlib.contextmanager
def using(path: str):
    """Temporarily route every call of this module to another build of the
    oracle (the planted-mistake builds of the pinning tests)."""
    global _lib
    saved = _lib
    _lib = _load(path)
    try:
        yield
    finally:
        _lib = saved


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = _load(_LIB)
    return _lib


class OrcConfig(C.Structure):
    _fields_ = [("dim", C.c_double * 3), ("ncell", C.c_int * 3),
                ("eta", C.c_float), ("h_tolerance", C.c_float),
                ("h_max", C.c_float), ("h_min", C.c_float),
                ("max_iter", C.c_int), ("alpha_max", C.c_float),
                ("alpha_min", C.c_float), ("visc_length", C.c_float),
                ("diff_beta", C.c_float), ("diff_alpha_max", C.c_float),
                ("diff_alpha_min", C.c_float), ("time_base", C.c_double),
                ("max_active_bin", C.c_int), ("nthreads", C.c_int)]


_PART_IN = ["x", "v", "m", "u", "time_bin"]
_PART_IO = ["h", "rho", "pressure", "soundspeed", "f", "balsara",
            "alpha_visc", "alpha_diff", "div_v_prev"]


class OrcParts(C.Structure):
    _fields_ = [("n", C.c_int)] + [(k, C.c_void_p) for k in _PART_IN + _PART_IO]


OUT64 = ["rho", "rho_dh", "wcount", "wcount_dh", "div_v", "rot_v",
         "pressure", "soundspeed", "f", "balsara",
         "v_sig", "laplace_u", "alpha_visc_max_ngb",
         "alpha_visc", "alpha_diff", "div_v_dt",
         "a_hydro", "u_dt", "h_dt",
         "n_density", "n_gradient", "n_force", "min_ngb_time_bin",
         "s_rho_dh", "s_wcount_dh", "s_div_v", "s_rot_v", "s_laplace_u",
         "s_a_hydro", "s_u_dt", "s_h_dt"]
OUT32 = ["rho_dh", "wcount", "wcount_dh", "div_v", "rot_v",
         "v_sig", "laplace_u", "alpha_visc_max_ngb", "div_v_dt",
         "a_hydro", "u_dt", "h_dt",
         "n_density", "n_gradient", "n_force", "min_ngb_time_bin",
         "ghost_iters"]
VEC = {"rot_v", "a_hydro", "s_rot_v", "s_a_hydro"}


class OrcOut64(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in OUT64]


class OrcOut32(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in OUT32]


def make_config(props, nthreads: int = 0) -> OrcConfig:
    c = OrcConfig()
    c.dim[:] = list(props.dim)
    c.ncell[:] = list(props.ncell)
    for k in ("eta", "h_tolerance", "h_max", "h_min", "max_iter", "alpha_max",
              "alpha_min", "visc_length", "diff_beta", "diff_alpha_max",
              "diff_alpha_min", "time_base", "max_active_bin"):
        setattr(c, k, getattr(props, k))
    c.nthreads = nthreads
    return c


def _parts(st: dict):
    keep = {}
    p = OrcParts()
    p.n = st["h"].shape[0]
    for k in _PART_IN + _PART_IO:
        a = st[k]
        assert a.flags["C_CONTIGUOUS"], k
        keep[k] = a
        setattr(p, k, a.ctypes.data)
    return p, keep


def _out_dtype(k: str, fp64: bool):
    if k.startswith("n_") or k == "ghost_iters":
        return np.int32
    if k == "min_ngb_time_bin":
        return np.int8
    return np.float64 if fp64 else np.float32


def pass_fp64(props, st: dict, nthreads: int = 0) -> dict:
    """FP64 rerun of density -> gradient -> force at the h in ``st``."""
    n = st["h"].shape[0]
    out = {}
    o = OrcOut64()
    for k in OUT64:
        shape = (n, 3) if k in VEC else (n,)
        out[k] = np.zeros(shape, _out_dtype(k, True))
        setattr(o, k, out[k].ctypes.data)
    cfg = make_config(props, nthreads)
    p, _keep = _parts(st)
    rc = lib().orc_pass_fp64(C.byref(cfg), C.byref(p), C.byref(o))
    if rc != 0:
        raise RuntimeError(f"orc_pass_fp64 failed ({rc})")
    return out


def pass_fp32(props, st: dict, nthreads: int = 0, port: bool = False) -> dict:
    """FP32 restatement of the full pass including the ghost; updates the
    (io) fields of ``st`` in place (h converged) and returns the outputs.
    ``port`` runs the CPU-baseline build (sorted candidate lists, -O3
    -ffast-math) instead of the oracle proper: True for the prebuilt one, or
    the path of another build of it."""
    n = st["h"].shape[0]
    out = {}
    o = OrcOut32()
    for k in OUT32:
        shape = (n, 3) if k in VEC else ((1,) if k == "ghost_iters" else (n,))
        out[k] = np.zeros(shape, _out_dtype(k, False))
        setattr(o, k, out[k].ctypes.data)
    cfg = make_config(props, nthreads)
    p, _keep = _parts(st)
    if port:
        L = port_lib(port) if isinstance(port, str) else port_lib()
    else:
        L = lib()
    rc = L.orc_pass_fp32(C.byref(cfg), C.byref(p), C.byref(o))
    if rc != 0:
        raise RuntimeError(f"orc_pass_fp32 failed ({rc})")
    for k in ("rho", "pressure", "soundspeed", "f", "balsara", "alpha_visc",
              "alpha_diff", "h"):
        out[k] = st[k].copy()
    return out


def count_bruteforce(props, st: dict, mode: int) -> np.ndarray:
    n = st["h"].shape[0]
    counts = np.zeros(n, np.int32)
    cfg = make_config(props)
    p, _keep = _parts(st)
    lib().orc_count_bruteforce(C.byref(cfg), C.byref(p), mode,
                               counts.ctypes.data_as(C.c_void_p))
    return counts


def contract_flips(props, st: dict) -> dict:
    """Density-mask decisions on which the reference's pair contract and a
    single rounding of the FP64 difference disagree (27-cell candidates)."""
    out = np.zeros(4, np.int64)
    cfg = make_config(props)
    p, _keep = _parts(st)
    lib().orc_contract_flips(C.byref(cfg), C.byref(p),
                             out.ctypes.data_as(C.c_void_p))
    return {"ref_only": int(out[0]), "fp64_only": int(out[1]),
            "candidates": int(out[2]), "both": int(out[3])}


class OrcFrame(C.Structure):
    _fields_ = [("self", C.c_int), ("oi", C.c_double * 3),
                ("oj", C.c_double * 3)]


def pair_frame(loc_i, loc_j, dim) -> OrcFrame:
    f = OrcFrame()
    a = (C.c_double * 3)(*loc_i)
    b = (C.c_double * 3)(*loc_j)
    d = (C.c_double * 3)(*dim)
    lib().orc_pair_frame(a, b, d, C.byref(f))
    return f


def pair_dx32(xi, xj, frame: OrcFrame):
    dx = (C.c_float * 3)()
    r2 = C.c_float()
    lib().orc_pair_dx32((C.c_double * 3)(*xi), (C.c_double * 3)(*xj),
                        C.byref(frame), dx, C.byref(r2))
    return np.array(dx[:], np.float32), np.float32(r2.value)


def kernel_deval32(u: float):
    w, dw = C.c_float(), C.c_float()
    lib().orc_kernel_deval32(u, C.byref(w), C.byref(dw))
    return w.value, dw.value


def kernel_deval64(u: float):
    w, dw = C.c_double(), C.c_double()
    lib().orc_kernel_deval64(u, C.byref(w), C.byref(dw))
    return w.value, dw.value


def constants() -> dict:
    L = lib()
    return {"kernel_gamma": L.orc_kernel_gamma(),
            "kernel_gamma2": L.orc_kernel_gamma2(),
            "kernel_root": L.orc_kernel_root(),
            "kernel_norm": L.orc_kernel_norm()}


PART_FIELDS = ["V0", "V1", "V2", "M", "H", "U", "RHO", "RHO_DH", "WC", "WC_DH",
               "DIVV", "ROT0", "ROT1", "ROT2", "F", "P", "C", "BAL", "VSIG",
               "LAPU", "AVMAX", "ALPHA", "ADIFF", "A0", "A1", "A2", "UDT", "HDT",
               "NDEN", "NGRAD", "NFORCE", "TB", "MINTB"]


def iact(kind: str, r2: float, dx, hi: float, hj: float, pi: np.ndarray,
         pj: np.ndarray, symmetric: bool) -> None:
    """Single FP32 pair interaction on flat float32 particle records."""
    dxa = np.ascontiguousarray(dx, dtype=np.float32)
    fn = getattr(lib(), f"orc_iact_{kind}32")
    fn(r2, dxa.ctypes.data, hi, hj, pi.ctypes.data, pj.ctypes.data,
       int(symmetric))


# ---- time integration ------------------------------------------------------
class OrcGas(C.Structure):
    _fields_ = [("n", C.c_int)] + [(k, C.c_void_p) for k in (
        "x", "v", "v_full", "a_hydro", "u", "u_full", "u_dt", "h", "h_dt",
        "rho", "pressure", "soundspeed", "v_sig", "time_bin",
        "min_ngb_time_bin")]


class OrcInteg(C.Structure):
    _fields_ = [("time_base", C.c_double), ("cfl", C.c_float),
                ("log_max_h_change", C.c_float), ("min_u", C.c_float),
                ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("fixed_bins", C.c_int)]


GAS_TYPES = {"x": np.float64, "v": np.float32, "v_full": np.float32,
             "a_hydro": np.float32, "u": np.float32, "u_full": np.float32,
             "u_dt": np.float32, "h": np.float32, "h_dt": np.float32,
             "rho": np.float32, "pressure": np.float32,
             "soundspeed": np.float32, "v_sig": np.float32,
             "time_bin": np.int8, "min_ngb_time_bin": np.int8}


def _gas(g: dict) -> OrcGas:
    s = OrcGas()
    s.n = g["h"].shape[0]
    for k, dt in GAS_TYPES.items():
        a = g[k]
        assert a.dtype == dt and a.flags["C_CONTIGUOUS"], k
        setattr(s, k, a.ctypes.data)
    return s


def integ(time_base: float, p) -> OrcInteg:
    """OrcInteg from a time base and an object with cfl, log_max_h_change,
    min_u, dt_min, dt_max, fixed_bins attributes."""
    return OrcInteg(time_base, p.cfl, p.log_max_h_change, p.min_u, p.dt_min,
                    p.dt_max, int(p.fixed_bins))


def max_active_bin(ti: int) -> int:
    return lib().orc_max_active_bin(C.c_longlong(ti))


def make_integer_timestep(new_dt: float, old_bin: int, min_ngb_bin: int,
                          ti_current: int, time_base_inv: float) -> int:
    f = lib().orc_make_integer_timestep
    f.restype = C.c_longlong
    f.argtypes = [C.c_float, C.c_int, C.c_int, C.c_longlong, C.c_double]
    return f(new_dt, old_bin, min_ngb_bin, ti_current, time_base_inv)


def drift(g: dict, c: OrcInteg, ti_old: int, ti_current: int) -> None:
    lib().orc_drift(C.byref(_gas(g)), C.byref(c), C.c_longlong(ti_old),
                    C.c_longlong(ti_current))


def kick(g: dict, c: OrcInteg, ti_current: int, mode: int) -> None:
    lib().orc_kick(C.byref(_gas(g)), C.byref(c), C.c_longlong(ti_current),
                   C.c_int(mode))


def timestep(g: dict, c: OrcInteg, ti_current: int):
    """Returns (ti_end_min, n_active, n_bad)."""
    te, na = C.c_longlong(), C.c_longlong()
    bad = lib().orc_timestep(C.byref(_gas(g)), C.byref(c),
                             C.c_longlong(ti_current), C.byref(te), C.byref(na))
    return te.value, na.value, bad
