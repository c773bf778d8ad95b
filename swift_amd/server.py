"""Serving form of the SPHENIX pass: a request server for independent boxes.

Python mirror of the swsph_server_* C ABI (include/swsph.h): requests are
admitted first come, first served by an HBM budget and a slot count, run
concurrently on their own engines and HIP streams, and their results stay on
the host until released.  No counterpart in the reference, whose engine
steps one box at a time.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import SwsphError, lib
from .engine import FIELDS, PASS_FIELDS, _HostParts, _props
from .ics import HydroProps


class _Opts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_double), ("max_inflight", C.c_int),
                ("use_graphs", C.c_int), ("fields", C.c_uint)]


class _Info(C.Structure):
    _fields_ = [("box", C.c_int64), ("ghost_rounds", C.c_int),
                ("n_density", C.c_int64), ("n_force", C.c_int64),
                ("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double),
                ("t_start_ms", C.c_double), ("t_done_ms", C.c_double),
                ("t_uploaded_ms", C.c_double), ("t_passed_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().swsph_server_last_error().decode(errors="replace")
        raise SwsphError(f"{what}: status {rc}: {msg}")


class Server:
    """Request server; use as a context manager."""

    def __init__(self, max_inflight: int = 8, hbm_budget_bytes: float = 0.0,
                 use_graphs: bool = False, fields=None):
        """`fields`: names of the result fields to return (None: all)."""
        mask = 0
        for k in fields or ():
            if k not in PASS_FIELDS:
                raise ValueError(f"{k} is not an output of a pass")
            mask |= 1 << FIELDS[k][0]
        o = _Opts(float(hbm_budget_bytes), int(max_inflight), int(use_graphs),
                  mask)
        self._fields = list(fields) if fields else None
        self._s = C.c_void_p()
        _check(lib().swsph_server_create(C.byref(o), C.byref(self._s)),
               "swsph_server_create")
        self._n = {}

    def close(self) -> None:
        if self._s:
            lib().swsph_server_destroy(self._s)
            self._s = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, props: HydroProps, st: dict, keep: bool = False) -> int:
        """Queue one density+force step of the box `st`; returns its id."""
        hp = _HostParts()
        n = st["h"].shape[0]
        hp.n = n
        keepalive = []
        want = {"x": np.float64, "v": np.float32, "m": np.float32,
                "u": np.float32, "time_bin": np.int8, "h": np.float32,
                "rho": np.float32, "pressure": np.float32,
                "soundspeed": np.float32, "f": np.float32,
                "balsara": np.float32, "alpha_visc": np.float32,
                "alpha_diff": np.float32, "div_v_prev": np.float32}
        for k, dt in want.items():
            a = np.ascontiguousarray(st[k], dtype=dt)
            keepalive.append(a)
            setattr(hp, k, a.ctypes.data)
        rid = C.c_int64()
        _check(lib().swsph_server_submit(self._s, C.byref(_props(props)),
                                         C.byref(hp), int(keep), C.byref(rid)),
               "swsph_server_submit")
        self._n[rid.value] = n
        return rid.value

    def step(self, box: int) -> int:
        """Queue another step of a resident box."""
        rid = C.c_int64()
        _check(lib().swsph_server_step(self._s, C.c_int64(box), C.byref(rid)),
               "swsph_server_step")
        self._n[rid.value] = self._n[box]
        return rid.value

    def wait(self, rid: int) -> dict:
        info = _Info()
        _check(lib().swsph_server_wait(self._s, C.c_int64(rid), C.byref(info)),
               "swsph_server_wait")
        return info.as_dict()

    def get(self, rid: int, name: str) -> np.ndarray:
        fid, dt, w = FIELDS[name]
        n = self._n[rid]
        out = np.empty((n, 3) if w == 3 else (n,), dt)
        _check(lib().swsph_server_get(self._s, C.c_int64(rid), fid,
                                      out.ctypes.data_as(C.c_void_p)),
               f"swsph_server_get({name})")
        return out

    def results(self, rid: int, names=None) -> dict:
        return {k: self.get(rid, k) for k in (names or self._fields or PASS_FIELDS)}

    def release(self, rid: int) -> None:
        _check(lib().swsph_server_release(self._s, C.c_int64(rid)),
               "swsph_server_release")
        self._n.pop(rid, None)

    def close_box(self, box: int) -> None:
        _check(lib().swsph_server_close_box(self._s, C.c_int64(box)),
               "swsph_server_close_box")

    def stats(self) -> dict:
        b = C.c_double()
        i = C.c_int()
        q = C.c_int()
        _check(lib().swsph_server_stats(self._s, C.byref(b), C.byref(i),
                                        C.byref(q)), "swsph_server_stats")
        return {"bytes_in_use": b.value, "inflight": i.value, "queued": q.value}


__all__ = ["Server"]
