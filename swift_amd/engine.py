"""Host-side mirror of SWIFT's task interface for the SPHENIX neighbour pass.

``Engine`` owns a device-resident particle store (libswsph.so) and exposes
the reference's task functions under their SWIFT names, with the same
argument meaning: a top-level cell index where SWIFT passes a ``struct cell``
(src/runner_doiact_hydro.h:177-210, src/runner.h:94-121).  Errors raise
``SwsphError`` where SWIFT calls error().
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import SwsphError, check, lib
from .ics import HydroProps

FIELDS = {
    "h": (0, np.float32, 1), "rho": (1, np.float32, 1),
    "rho_dh": (2, np.float32, 1), "wcount": (3, np.float32, 1),
    "wcount_dh": (4, np.float32, 1), "div_v": (5, np.float32, 1),
    "rot_v": (6, np.float32, 3), "pressure": (7, np.float32, 1),
    "soundspeed": (8, np.float32, 1), "f": (9, np.float32, 1),
    "balsara": (10, np.float32, 1), "v_sig": (11, np.float32, 1),
    "laplace_u": (12, np.float32, 1), "alpha_visc_max_ngb": (13, np.float32, 1),
    "alpha_visc": (14, np.float32, 1), "alpha_diff": (15, np.float32, 1),
    "div_v_dt": (16, np.float32, 1), "div_v_prev": (17, np.float32, 1),
    "a_hydro": (18, np.float32, 3), "u_dt": (19, np.float32, 1),
    "h_dt": (20, np.float32, 1), "n_density": (21, np.int32, 1),
    "n_gradient": (22, np.int32, 1), "n_force": (23, np.int32, 1),
    "min_ngb_time_bin": (24, np.int8, 1),
    # particle state advanced by the time integration
    "x": (32, np.float64, 3), "v": (33, np.float32, 3),
    "v_full": (34, np.float32, 3), "u": (35, np.float32, 1),
    "u_full": (36, np.float32, 1), "time_bin": (37, np.int8, 1),
}
PASS_FIELDS = [k for k, v in FIELDS.items() if v[0] < 25]


class _FieldDst(C.Structure):
    _fields_ = [("field", C.c_int), ("base", C.c_void_p), ("stride", C.c_int64)]


class SchedStats(C.Structure):
    _fields_ = [("active_cells", C.c_int64), ("tasks", C.c_int64),
                ("bundles", C.c_int64), ("levels", C.c_int),
                ("wall_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class _TuneResult(C.Structure):
    _fields_ = [("schedule", C.c_int), ("chunk", C.c_int * 3),
                ("loop_ms", C.c_double * 3), ("pass_ms", C.c_double),
                ("trials", C.c_int), ("chunk_j", C.c_int * 3)]


class _Props(C.Structure):
    _fields_ = [("dim", C.c_double * 3), ("ncell", C.c_int * 3),
                ("eta", C.c_float), ("h_tolerance", C.c_float),
                ("h_max", C.c_float), ("h_min", C.c_float),
                ("max_iter", C.c_int), ("alpha_max", C.c_float),
                ("alpha_min", C.c_float), ("visc_length", C.c_float),
                ("diff_beta", C.c_float), ("diff_alpha_max", C.c_float),
                ("diff_alpha_min", C.c_float), ("time_base", C.c_double),
                ("max_active_bin", C.c_int)]


class _HostParts(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in (
        "x", "v", "m", "u", "time_bin", "h", "rho", "pressure", "soundspeed",
        "f", "balsara", "alpha_visc", "alpha_diff", "div_v_prev")]


class PassStats(C.Structure):
    _fields_ = [("ghost_rounds", C.c_int), ("n_density", C.c_int64),
                ("n_gradient", C.c_int64), ("n_force", C.c_int64),
                ("launches", C.c_int64), ("list_rebuilt", C.c_int64),
                ("list_fallback", C.c_int64), ("n_density_all", C.c_int64),
                ("regrids", C.c_int64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Integration(C.Structure):
    """swsph_integration: time-step parameters (SPH:CFL_condition,
    hydro_props.log_max_h_change, minimal_internal_energy, dt_min, dt_max)
    and whether the particles keep the time bins they were given."""
    _fields_ = [("cfl", C.c_float), ("log_max_h_change", C.c_float),
                ("min_u", C.c_float), ("dt_min", C.c_double),
                ("dt_max", C.c_double), ("fixed_bins", C.c_int)]


class StepInfo(C.Structure):
    _fields_ = [("ti_current", C.c_int64), ("ti_end_min", C.c_int64),
                ("max_active_bin", C.c_int), ("rebuilt", C.c_int),
                ("n_active", C.c_int64), ("dx_max", C.c_float),
                ("launches", C.c_int64), ("pass_", PassStats),
                ("sparse", C.c_int)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "pass_"}
        d["pass"] = self.pass_.as_dict()
        return d


def integration(cfl: float = 0.1, max_volume_change: float = 1.4,
                min_u: float = 0.0, dt_min: float = 0.0, dt_max: float = 1e30,
                fixed_bins: bool = False) -> Integration:
    """Parameters with SWIFT's defaults (src/hydro_properties.c:42, 122)."""
    third = np.float32(0.3333333333)
    lmhc = np.log(np.power(np.float32(max_volume_change), third,
                           dtype=np.float32), dtype=np.float32)
    return Integration(cfl, float(lmhc), min_u, dt_min, dt_max, int(fixed_bins))


def _props(p: HydroProps) -> _Props:
    s = _Props()
    s.dim[:] = list(p.dim)
    s.ncell[:] = list(p.ncell)
    for k in ("eta", "h_tolerance", "h_max", "h_min", "max_iter", "alpha_max",
              "alpha_min", "visc_length", "diff_beta", "diff_alpha_max",
              "diff_alpha_min", "time_base", "max_active_bin"):
        setattr(s, k, getattr(p, k))
    return s


class Engine:
    """A device-resident particle store plus the SWIFT task entry points."""

    def __init__(self, props: HydroProps):
        self.props = props
        self._h = C.c_void_p()
        check(lib().swsph_engine_create(C.byref(_props(props)), C.byref(self._h)),
              "swsph_engine_create")
        self.n = 0

    def close(self) -> None:
        if self._h:
            lib().swsph_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_props(self, props: HydroProps) -> None:
        """New run-time parameters; a new grid needs set_particles next
        (the device buffers are reused when large enough)."""
        check(lib().swsph_set_props(self._h, C.byref(_props(props))),
              "swsph_set_props")
        self.props = props

    # -- particles ---------------------------------------------------------
    def set_particles(self, st: dict) -> None:
        hp = _HostParts()
        keep = []
        n = st["h"].shape[0]
        hp.n = n
        want = {"x": np.float64, "v": np.float32, "m": np.float32,
                "u": np.float32, "time_bin": np.int8, "h": np.float32,
                "rho": np.float32, "pressure": np.float32,
                "soundspeed": np.float32, "f": np.float32,
                "balsara": np.float32, "alpha_visc": np.float32,
                "alpha_diff": np.float32, "div_v_prev": np.float32}
        for k, dt in want.items():
            a = np.ascontiguousarray(st[k], dtype=dt)
            keep.append(a)
            setattr(hp, k, a.ctypes.data)
        check(lib().swsph_set_particles(self._h, C.byref(hp)),
              "swsph_set_particles")
        self.n = n

    def get(self, name: str) -> np.ndarray:
        fid, dt, w = FIELDS[name]
        out = np.empty((self.n, 3) if w == 3 else (self.n,), dt)
        check(lib().swsph_get_field(self._h, fid, out.ctypes.data_as(C.c_void_p)),
              f"swsph_get_field({name})")
        return out

    def write_back(self, ci: int, dst: dict) -> None:
        """swsph_cell_write_back: waits for the engine and writes the fields
        of cell ci's particles (ci = -1: all) into the caller's arrays,
        dst[name] = an array indexed by the caller's particle index (rows of
        a structured array work too: its field views carry the stride)."""
        recs = (_FieldDst * max(1, len(dst)))()
        for k, (name, arr) in enumerate(dst.items()):
            fid, dt, w = FIELDS[name]
            if arr.dtype != dt:
                raise SwsphError(f"write_back({name}): dtype {arr.dtype}, "
                                 f"want {np.dtype(dt)}")
            recs[k].field = fid
            recs[k].base = arr.ctypes.data
            recs[k].stride = arr.strides[0]
        check(lib().swsph_cell_write_back(self._h, ci, recs, len(dst)),
              "swsph_cell_write_back")

    def results(self, names=None) -> dict:
        return {k: self.get(k) for k in (names or PASS_FIELDS)}

    def num_cells(self) -> int:
        n = C.c_int()
        check(lib().swsph_num_cells(self._h, C.byref(n)), "swsph_num_cells")
        return n.value

    def cell_count(self, c: int) -> int:
        n = C.c_int()
        check(lib().swsph_cell_count(self._h, c, C.byref(n)), "swsph_cell_count")
        return n.value

    def sync(self) -> None:
        check(lib().swsph_sync(self._h), "swsph_sync")

    # -- SWIFT task entry points (src/runner_doiact_hydro.h, src/runner.h) --
    def hydro_init_cell(self, ci: int) -> None:
        check(lib().swsph_init_cell(self._h, ci), "swsph_init_cell")

    def runner_doself1_branch_density(self, ci: int) -> None:
        check(lib().swsph_doself1_density(self._h, ci), "runner_doself1_branch_density")

    def runner_dopair1_branch_density(self, ci: int, cj: int) -> None:
        check(lib().swsph_dopair1_density(self._h, ci, cj),
              "runner_dopair1_branch_density")

    def runner_doself_subset_branch_density(self, ci: int, ind,
                                            first: int = 0) -> None:
        """Density of the caller's particles first + ind[k] (all in cell ci)
        against ci (runner_doself_subset_branch_density(r, ci, parts, ind,
        count) with parts = the caller's array + first)."""
        ind = np.ascontiguousarray(ind, dtype=np.int32)
        check(lib().swsph_doself_subset_density(
            self._h, ci, C.c_int64(first), ind.ctypes.data_as(C.c_void_p),
            int(ind.size)), "runner_doself_subset_branch_density")

    def runner_dopair_subset_branch_density(self, ci: int, ind, cj: int,
                                            first: int = 0) -> None:
        ind = np.ascontiguousarray(ind, dtype=np.int32)
        check(lib().swsph_dopair_subset_density(
            self._h, ci, C.c_int64(first), ind.ctypes.data_as(C.c_void_p),
            int(ind.size), cj), "runner_dopair_subset_branch_density")

    def runner_do_ghost(self, ci: int, offset: int = 0, ntasks: int = 1) -> None:
        check(lib().swsph_do_ghost(self._h, ci, offset, ntasks),
              "runner_do_ghost")

    def runner_doself1_branch_gradient(self, ci: int) -> None:
        check(lib().swsph_doself1_gradient(self._h, ci),
              "runner_doself1_branch_gradient")

    def runner_dopair1_branch_gradient(self, ci: int, cj: int) -> None:
        check(lib().swsph_dopair1_gradient(self._h, ci, cj),
              "runner_dopair1_branch_gradient")

    def runner_do_extra_ghost(self, ci: int) -> None:
        check(lib().swsph_do_extra_ghost(self._h, ci), "runner_do_extra_ghost")

    def runner_doself2_branch_force(self, ci: int) -> None:
        check(lib().swsph_doself2_force(self._h, ci), "runner_doself2_branch_force")

    def runner_dopair2_branch_force(self, ci: int, cj: int) -> None:
        check(lib().swsph_dopair2_force(self._h, ci, cj),
              "runner_dopair2_branch_force")

    def runner_do_end_hydro_force(self, ci: int) -> None:
        check(lib().swsph_do_end_hydro_force(self._h, ci),
              "runner_do_end_hydro_force")

    # -- bundles of cell tasks and the scheduler-driven pass -----------------
    @staticmethod
    def _cells(cells):
        return np.ascontiguousarray(cells, dtype=np.int32)

    def init_cells(self, cells) -> None:
        c = self._cells(cells)
        check(lib().swsph_init_cells(self._h, c.ctypes.data_as(C.c_void_p),
                                     int(c.size)), "swsph_init_cells")

    def runner_do_ghost_cells(self, cells) -> int:
        c = self._cells(cells)
        r = C.c_int(0)
        check(lib().swsph_do_ghost_cells(self._h, c.ctypes.data_as(C.c_void_p),
                                         int(c.size), C.byref(r)),
              "swsph_do_ghost_cells")
        return r.value

    def runner_do_extra_ghost_cells(self, cells) -> None:
        c = self._cells(cells)
        check(lib().swsph_do_extra_ghost_cells(
            self._h, c.ctypes.data_as(C.c_void_p), int(c.size)),
            "swsph_do_extra_ghost_cells")

    def runner_do_end_hydro_force_cells(self, cells) -> None:
        c = self._cells(cells)
        check(lib().swsph_do_end_hydro_force_cells(
            self._h, c.ctypes.data_as(C.c_void_p), int(c.size)),
            "swsph_do_end_hydro_force_cells")

    def cell_active(self) -> np.ndarray:
        out = np.zeros(self.num_cells(), np.uint8)
        check(lib().swsph_cell_active(self._h, out.ctypes.data_as(C.c_void_p)),
              "swsph_cell_active")
        return out

    def sched_pass(self) -> tuple:
        """One pass through the task graph: (scheduler stats, pass stats)."""
        ss, ps = SchedStats(), PassStats()
        check(lib().swsph_sched_pass(self._h, C.byref(ss), C.byref(ps)),
              "swsph_sched_pass")
        return ss.as_dict(), ps.as_dict()

    def set_regrid(self, on: bool) -> None:
        check(lib().swsph_set_regrid(self._h, 1 if on else 0),
              "swsph_set_regrid")

    def get_grid(self) -> tuple:
        nc = (C.c_int * 3)()
        check(lib().swsph_get_grid(self._h, nc), "swsph_get_grid")
        return tuple(nc)

    def set_step_sched(self, on: bool) -> None:
        check(lib().swsph_set_step_sched(self._h, 1 if on else 0),
              "swsph_set_step_sched")

    # -- task bundles (scheduler -> device export) ----------------------------
    TASK_SELF, TASK_PAIR = 0, 1

    def run_tasks(self, loop: str, tasks) -> None:
        """Runs a bundle of released tasks of one loop in one launch: rows
        (type, ci, cj) with type TASK_SELF (cj ignored) or TASK_PAIR; the
        same interactions as the per-task entry points, merged per i-cell."""
        t = np.ascontiguousarray(np.asarray(tasks, np.int32).reshape(-1, 3))
        check(lib().swsph_run_tasks(self._h, self.LOOPS[loop],
                                    t.ctypes.data_as(C.c_void_p), t.shape[0]),
              "swsph_run_tasks")

    # -- whole passes --------------------------------------------------------
    def pass_plain(self) -> dict:
        s = PassStats()
        check(lib().swsph_pass_plain(self._h, C.byref(s)), "swsph_pass_plain")
        return s.as_dict()

    def pass_batched(self) -> dict:
        s = PassStats()
        check(lib().swsph_pass_batched(self._h, C.byref(s)), "swsph_pass_batched")
        return s.as_dict()

    def pass_batched_noalloc(self) -> dict:
        """The batched pass on the library's fixed device scratch: no device
        allocation (at most 2^21 particles on 2^16 cells)."""
        s = PassStats()
        check(lib().swsph_pass_batched_noalloc(self._h, C.byref(s)),
              "swsph_pass_batched_noalloc")
        return s.as_dict()

    def autotune(self, repeats: int = 2) -> dict:
        """Time the schedules and chunk capacities on the current particles
        and apply the fastest (the pass state is restored afterwards)."""
        r = _TuneResult()
        check(lib().swsph_autotune(self._h, int(repeats), C.byref(r)),
              "swsph_autotune")
        return {"schedule": r.schedule, "chunk": list(r.chunk),
                "chunk_j": list(r.chunk_j), "loop_ms": list(r.loop_ms),
                "pass_ms": r.pass_ms, "trials": r.trials}

    def pass_graph(self) -> dict:
        """The batched pass replayed from captured HIP graphs."""
        s = PassStats()
        check(lib().swsph_pass_graph(self._h, C.byref(s)), "swsph_pass_graph")
        return s.as_dict()

    LOOPS = {"density": 0, "gradient": 1, "force": 2}

    def set_chunk(self, loop: str, C: int) -> None:
        """Tile shape of one neighbour loop: chunks of C particles, 64 / C
        candidate streams per particle (C in 8, 16, 32, 64)."""
        check(lib().swsph_set_chunk(self._h, self.LOOPS[loop], int(C)),
              "swsph_set_chunk")

    def set_tile_shape(self, loop: str, CI: int, CJ: int) -> None:
        """Tiles of CI i-particles against candidate chunks of CJ particles
        in the work-queue schedules (CJ = CI, or CI / 2 for CI = 16, 32, the
        narrow tiles (4, 16), (4, 32), (8, 16) for sparse activity, or
        (16, 32))."""
        check(lib().swsph_set_tile_shape(self._h, self.LOOPS[loop], int(CI),
                                         int(CJ)), "swsph_set_tile_shape")

    def set_bin_order(self, on: bool) -> None:
        """Order the fields inside each cell by time bin, then Morton code
        (the active particles of a step become a prefix of every cell; the
        loops' records keep the Morton order).  On by default."""
        check(lib().swsph_set_bin_order(self._h, int(bool(on))),
              "swsph_set_bin_order")

    def set_sparse_shape(self, CI: int, CJ: int, per_cell: int) -> None:
        """Tile shape of step() passes with fewer than per_cell active
        particles per top-level cell (0 turns it off)."""
        check(lib().swsph_set_sparse_shape(self._h, int(CI), int(CJ),
                                           int(per_cell)),
              "swsph_set_sparse_shape")

    def get_tile_shape(self, loop: str) -> tuple:
        ci, cj = C.c_int(), C.c_int()
        check(lib().swsph_get_tile_shape(self._h, self.LOOPS[loop],
                                         C.byref(ci), C.byref(cj)),
              "swsph_get_tile_shape")
        return ci.value, cj.value

    def get_chunk(self, loop: str) -> int:
        c = C.c_int()
        check(lib().swsph_get_chunk(self._h, self.LOOPS[loop], C.byref(c)),
              "swsph_get_chunk")
        return c.value

    def set_schedule(self, mode: int = 1, waves: int = 0) -> None:
        """Batched loops: 0 one wave per chunk (grid), 1 one wave per tile
        of a device-built tile queue, 2 persistent waves draining that queue
        (`waves` 0: fill every CU)."""
        check(lib().swsph_set_schedule(self._h, int(mode), int(waves)),
              "swsph_set_schedule")

    def set_lists(self, enable: bool = True, skin: float = 1.2,
                  slots: int = 48) -> None:
        """Neighbour lists of the batched pass: candidates within `skin`
        times the search radius, `slots` entries per candidate stream."""
        check(lib().swsph_set_lists(self._h, int(bool(enable)), C.c_float(skin),
                                    int(slots)), "swsph_set_lists")

    def save_state(self) -> None:
        check(lib().swsph_save_state(self._h), "swsph_save_state")

    def restore_state(self) -> None:
        check(lib().swsph_restore_state(self._h), "swsph_restore_state")

    def count_interactions(self) -> tuple:
        out = (C.c_int64 * 3)()
        check(lib().swsph_count_interactions(self._h, out),
              "swsph_count_interactions")
        return tuple(out)

    def set_timing(self, on: bool) -> None:
        check(lib().swsph_set_timing(self._h, int(on)), "swsph_set_timing")

    def timing(self) -> tuple:
        ms = (C.c_double * 3)()
        nl = (C.c_int64 * 3)()
        check(lib().swsph_get_timing(self._h, ms, nl), "swsph_get_timing")
        return list(ms), list(nl)

    # -- time integration (runner_do_drift_part, kick1, kick2, timestep) --
    def set_integration(self, p: Integration) -> None:
        check(lib().swsph_set_integration(self._h, C.byref(p)),
              "swsph_set_integration")

    def runner_do_drift_part(self, ci: int, ti_old: int, ti_current: int) -> None:
        check(lib().swsph_do_drift_part(self._h, ci, C.c_int64(ti_old),
                                        C.c_int64(ti_current)),
              "swsph_do_drift_part")

    def runner_do_kick1(self, ci: int, ti_current: int) -> None:
        check(lib().swsph_do_kick1(self._h, ci, C.c_int64(ti_current)),
              "swsph_do_kick1")

    def runner_do_kick2(self, ci: int, ti_current: int) -> None:
        check(lib().swsph_do_kick2(self._h, ci, C.c_int64(ti_current)),
              "swsph_do_kick2")

    def runner_do_timestep(self, ci: int, ti_current: int) -> int:
        out = C.c_int64()
        check(lib().swsph_do_timestep(self._h, ci, C.c_int64(ti_current),
                                      C.byref(out)), "swsph_do_timestep")
        return out.value

    def rebuild(self) -> None:
        check(lib().swsph_rebuild(self._h), "swsph_rebuild")

    def step(self) -> dict:
        """One engine_step of the gas (see swsph_engine_step)."""
        info = StepInfo()
        check(lib().swsph_engine_step(self._h, C.byref(info)),
              "swsph_engine_step")
        return info.as_dict()

    def cell_coords(self, c: int) -> tuple:
        nx, ny, nz = self.props.ncell
        return c // (ny * nz), (c // nz) % ny, c % nz

    def cell_id(self, cx: int, cy: int, cz: int) -> int:
        nx, ny, nz = self.props.ncell
        return ((cx % nx) * ny + (cy % ny)) * nz + (cz % nz)


__all__ = ["Engine", "SwsphError", "FIELDS", "PASS_FIELDS", "PassStats",
           "Integration", "StepInfo", "integration"]
