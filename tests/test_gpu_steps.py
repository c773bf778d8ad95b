"""Time integration and device re-binning on the GPU against the oracle.

Every entry point of a step is compared with the oracle applied to the same
device state: drift (src/drift.h:138), kick1 / kick2
(src/runner_time_integration.c:87, 360) and the time-step task (:639) --
positions, time bins and synchronisation points exactly, float fields to
two ulp.  The pass inside each step is compared with the FP64 rerun at the
GPU's smoothing lengths (neighbour counts exactly, 1e-5 / 1e-4), as in the
single-pass parity tests.  swsph_engine_step is checked to be exactly the
sequence of those entry points.  Re-binning is checked to be a permutation
that carries every field, and a pass after it to equal a pass on a fresh
upload of the same state.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from parity import DENSITY_FIELDS, FORCE_FIELDS, assert_alpha, assert_parity

pytestmark = pytest.mark.gpu

STATE = ["x", "v", "v_full", "a_hydro", "u", "u_full", "u_dt", "h", "h_dt",
         "rho", "pressure", "soundspeed", "v_sig", "time_bin",
         "min_ngb_time_bin"]
IO = ["rho", "pressure", "soundspeed", "f", "balsara", "alpha_visc",
      "alpha_diff", "div_v_prev"]
EXACT = ["x", "time_bin"]


def _state(eng):
    return {k: np.ascontiguousarray(eng.get(k)) for k in STATE}


def _pass_input(eng, m):
    """The inputs of a pass in the caller's order, from the device."""
    s = {"x": eng.get("x"), "v": eng.get("v"), "m": m, "u": eng.get("u"),
         "time_bin": eng.get("time_bin"), "h": eng.get("h")}
    for k in IO:
        s[k] = eng.get(k)
    return {k: np.ascontiguousarray(v) for k, v in s.items()}


def _same(a: dict, b: dict, keys, rtol=2.5e-7, what=""):
    for k in keys:
        if k in EXACT or rtol == 0:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what} {k}")
        else:
            np.testing.assert_allclose(a[k], b[k], rtol=rtol, atol=1e-30,
                                       err_msg=f"{what} {k}")


def _box(n=16, seed=3, bulk=(0.0, 0.0, 0.0)):
    from swift_amd import ics
    st = ics.modulated_box(n, seed=seed)
    st["v"] += np.asarray(bulk, np.float32)
    ics.set_bin_hierarchy(st, {1: 0.1, 2: 0.2, 3: 0.7}, seed=seed)
    return st


def _integ(fixed=True):
    from swift_amd.engine import integration
    return integration(fixed_bins=fixed)


def _checked_pass(orc, eng, props, m):
    pre = _pass_input(eng, m)
    eng.set_props(props)
    stats = eng.pass_batched()
    res = eng.results()
    chk = {k: np.array(v, copy=True) for k, v in pre.items()}
    chk["h"] = res["h"]
    ref = orc.pass_fp64(props, chk)
    act = pre["time_bin"] <= props.max_active_bin
    # alpha_visc / alpha_diff divide cancelling sums by dt_alpha > 0: they
    # are held to the propagated bound instead
    fields = {**DENSITY_FIELDS, **{k: v for k, v in FORCE_FIELDS.items()
                                   if not k.startswith("alpha")}}
    assert_parity(res, ref, act, fields=fields)
    pre["h"] = res["h"]
    assert_alpha(res, ref, pre, props, props.time_base, act)
    return stats, act


def _manual_step(orc, eng, props, m, oi, ti_old, ti, first):
    """One step through the entry points, each checked against the oracle."""
    if not first:
        g = _state(eng)
        orc.drift(g, oi, ti_old, ti)
        eng.runner_do_drift_part(-1, ti_old, ti)
        _same(_state(eng), g, STATE, what="drift")
    mab = orc.max_active_bin(ti)
    _checked_pass(orc, eng, dataclasses.replace(props, max_active_bin=mab), m)
    if not first:
        g = _state(eng)
        orc.kick(g, oi, ti, 2)
        eng.runner_do_kick2(-1, ti)
        _same(_state(eng), g, STATE, what="kick2")
    g = _state(eng)
    te_o, na_o, bad = orc.timestep(g, oi, ti)
    assert bad == 0
    te = eng.runner_do_timestep(-1, ti)
    assert te == te_o
    _same(_state(eng), g, STATE, what="timestep")
    g = _state(eng)
    orc.kick(g, oi, ti, 1)
    eng.runner_do_kick1(-1, ti)
    _same(_state(eng), g, STATE, what="kick1")
    return te, na_o


def test_steps_match_oracle_and_engine_step(orc):
    from swift_amd import Engine, HydroProps
    st = _box()
    props = HydroProps(ncell=(4, 4, 4), time_base=1e-5)
    integ = _integ(fixed=True)
    oi = orc.integ(props.time_base, integ)
    with Engine(props) as a, Engine(props) as b:
        for e in (a, b):
            e.set_particles(st)
            e.set_integration(integ)
        # the passes of the entry points use the engine's tile shapes
        a.set_sparse_shape(4, 32, 0)
        ti, te = 0, 0
        seen = set()
        for k in range(6):
            info = a.step()
            ti_old, ti = ti, (0 if k == 0 else te)
            te, na = _manual_step(orc, b, props, st["m"], oi, ti_old, ti, k == 0)
            assert info["ti_current"] == ti and info["ti_end_min"] == te
            assert info["n_active"] == na and info["rebuilt"] == 0
            seen.add(info["max_active_bin"])
            _same(_state(a), _state(b), STATE, rtol=0, what=f"step {k}")
        # the hierarchy was exercised: steps with 1, 2 and all 3 bins active
        assert {1, 2, 3} <= seen or 56 in seen


def test_timestep_rule_matches_oracle(orc):
    """SWIFT's CFL / h-change / neighbour-limited bins, bit for bit."""
    from swift_amd import Engine, HydroProps, ics
    st = ics.modulated_box(16, seed=5, amp=0.7)
    # bin 0 (no step yet): no growth limit, so the CFL condition decides
    st["time_bin"][:] = 0
    props = HydroProps(ncell=(4, 4, 4), time_base=1e-7)
    integ = _integ(fixed=False)
    oi = orc.integ(props.time_base, integ)
    with Engine(props) as e:
        e.set_particles(st)
        e.set_integration(integ)
        e.pass_batched()
        g = _state(e)
        te_o, na, bad = orc.timestep(g, oi, 0)
        te = e.runner_do_timestep(-1, 0)
        assert bad == 0 and te == te_o
        tb = e.get("time_bin")
        np.testing.assert_array_equal(tb, g["time_bin"])
        assert len(np.unique(tb)) >= 2  # the rule spreads the bins


def test_engine_step_physical_bins(orc):
    """engine_step with SWIFT's time-step rule advances through the
    synchronisation points the bins define."""
    from swift_amd import Engine, HydroProps, ics
    st = ics.modulated_box(16, seed=6)
    props = HydroProps(ncell=(4, 4, 4), time_base=1e-7)
    with Engine(props) as e:
        e.set_particles(st)
        e.set_integration(_integ(fixed=False))
        prev = -1
        for _ in range(4):
            info = e.step()
            assert info["ti_current"] > prev and info["ti_end_min"] > info["ti_current"]
            prev = info["ti_current"]
            tb = e.get("time_bin")
            assert info["n_active"] >= 1 and tb.min() >= 1


def _wrap(x):
    return np.where(x >= 1.0, x - 1.0, np.where(x < 0.0, x + 1.0, x))


def test_rebuild_carries_every_field(orc):
    from swift_amd import Engine, HydroProps
    from swift_amd.engine import FIELDS
    st = _box(bulk=(0.9, -0.6, 0.3))
    props = HydroProps(ncell=(4, 4, 4), time_base=1e-3)
    with Engine(props) as e:
        e.set_particles(st)
        e.set_integration(_integ())
        e.pass_batched()
        # move every particle by ~0.3 of a cell: many change cells
        e.runner_do_drift_part(-1, 0, 64)
        before = {k: e.get(k) for k in FIELDS}
        e.rebuild()
        after = {k: e.get(k) for k in FIELDS}
        for k in FIELDS:
            if k == "x":
                np.testing.assert_array_equal(after[k], _wrap(before[k]))
            else:
                np.testing.assert_array_equal(after[k], before[k], err_msg=k)
        x = after["x"]
        assert np.all((x >= 0) & (x < 1))
        cells = np.floor(x * 4).astype(int)
        cid = (cells[:, 0] * 4 + cells[:, 1]) * 4 + cells[:, 2]
        counts = np.bincount(cid, minlength=64)
        assert [e.cell_count(c) for c in range(64)] == counts.tolist()
        moved = np.floor(_wrap(before["x"]) * 4) != np.floor(st["x"] * 4)
        assert moved.any()
        # a pass on the re-binned particles equals one on a fresh upload
        inp = _pass_input(e, st["m"])
        e.pass_batched()
        r1 = e.results()
    with Engine(props) as f:
        f.set_particles(inp)
        f.pass_batched()
        r2 = f.results()
    for k in r1:
        np.testing.assert_array_equal(r1[k], r2[k], err_msg=k)


def test_engine_step_rebuilds_when_particles_leave_the_stencil(orc):
    from swift_amd import Engine, HydroProps
    st = _box(bulk=(1.5, 0.8, -0.4))
    props = HydroProps(ncell=(4, 4, 4), time_base=4e-3)
    integ = _integ(fixed=True)
    oi = orc.integ(props.time_base, integ)
    with Engine(props) as a, Engine(props) as b:
        for e in (a, b):
            e.set_particles(st)
            e.set_integration(integ)
        # the passes of the entry points use the engine's tile shapes
        a.set_sparse_shape(4, 32, 0)
        ti = te = 0
        rebuilt = 0
        for k in range(4):
            info = a.step()
            ti_old, ti = ti, (0 if k == 0 else te)
            if info["rebuilt"]:
                rebuilt += 1
                g = _state(b)
                orc.drift(g, oi, ti_old, ti)
                b.runner_do_drift_part(-1, ti_old, ti)
                b.rebuild()
                _same(_state(b), {**g, "x": _wrap(g["x"])}, STATE, what="drift")
                te, _ = _rest_of_step(orc, b, props, st["m"], oi, ti)
            else:
                te, _ = _manual_step(orc, b, props, st["m"], oi, ti_old, ti, k == 0)
            assert info["ti_end_min"] == te
            _same(_state(a), _state(b), STATE, rtol=0, what=f"step {k}")
        assert rebuilt >= 1


def _rest_of_step(orc, eng, props, m, oi, ti):
    """A step after its drift (and re-binning): pass, kick2, time step,
    kick1, each checked against the oracle."""
    mab = orc.max_active_bin(ti)
    _checked_pass(orc, eng, dataclasses.replace(props, max_active_bin=mab), m)
    g = _state(eng)
    orc.kick(g, oi, ti, 2)
    eng.runner_do_kick2(-1, ti)
    _same(_state(eng), g, STATE, what="kick2")
    g = _state(eng)
    te_o, na, bad = orc.timestep(g, oi, ti)
    te = eng.runner_do_timestep(-1, ti)
    assert bad == 0 and te == te_o
    g = _state(eng)
    orc.kick(g, oi, ti, 1)
    eng.runner_do_kick1(-1, ti)
    _same(_state(eng), g, STATE, what="kick1")
    return te, na


def test_records_follow_fields_outside_passes(orc):
    """Passes rebuild only the records of the particles they update; every
    change of fields outside a pass (kick2's predicted values, a restore
    that brings back other values for particles the next pass leaves alone)
    must still reach the records.  Each case equals a fresh upload."""
    from swift_amd import Engine, HydroProps, ics
    st = ics.modulated_box(16, seed=9)
    ics.set_bin_hierarchy(st, {1: 0.5, 3: 0.5}, seed=9)
    all_on = HydroProps(ncell=(4, 4, 4), time_base=1e-5)
    half = dataclasses.replace(all_on, max_active_bin=1)
    # (a) restore, then a pass over fewer particles than the last one
    with Engine(all_on) as e:
        e.set_particles(st)
        e.save_state()
        e.pass_batched()
        e.restore_state()
        e.set_props(half)
        e.pass_batched()
        ra = e.results()
    with Engine(half) as f:
        f.set_particles(st)
        f.pass_batched()
        rb = f.results()
    # (outputs of particles a pass leaves alone keep older values)
    act = st["time_bin"] <= 1
    for k in ra:
        np.testing.assert_array_equal(ra[k][act], rb[k][act],
                                      err_msg=f"restore {k}")
    # (b) kick2 changes v, u, P, c of the active particles; a pass that
    # follows without a drift sees them
    with Engine(half) as e:
        e.set_particles(st)
        e.set_integration(_integ())
        e.pass_batched()
        e.runner_do_kick2(-1, 4)
        inp = _pass_input(e, st["m"])
        e.pass_batched()
        ra = e.results()
    with Engine(half) as f:
        f.set_particles(inp)
        f.pass_batched()
        rb = f.results()
    for k in ra:
        np.testing.assert_array_equal(ra[k][act], rb[k][act],
                                      err_msg=f"kick2 {k}")


def test_sparse_tile_shape_steps():
    """Steps whose passes run with narrow tiles (4 particles against chunks
    of 32) -- always, never, or by the default rule (fewer than 16 active
    particles per cell: here the 10 % steps) -- find the same neighbours
    (pass counts equal) and the same state to FP32 summation order."""
    from swift_amd import Engine, HydroProps
    st = _box()
    props = HydroProps(ncell=(4, 4, 4), time_base=1e-5)
    with Engine(props) as a, Engine(props) as b, Engine(props) as c:
        for e in (a, b, c):
            e.set_particles(st)
            e.set_integration(_integ(fixed=True))
        a.set_sparse_shape(4, 32, 1 << 30)  # every step
        b.set_sparse_shape(4, 32, 0)        # none
        kinds = set()
        for k in range(8):
            ia, ib, ic = a.step(), b.step(), c.step()
            assert ia["sparse"] == 1 and ib["sparse"] == 0
            kinds.add(ic["sparse"])
            for i in (ia, ic):
                for key in ("n_density", "n_gradient", "n_force", "ghost_rounds"):
                    assert i["pass"][key] == ib["pass"][key], (k, key)
                assert i["ti_end_min"] == ib["ti_end_min"]
            sb = _state(b)
            for e in (a, c):
                se = _state(e)
                np.testing.assert_array_equal(se["time_bin"], sb["time_bin"])
                for key in ("x", "v", "v_full", "u", "u_full", "h", "rho"):
                    # (velocity components through zero: absolute to 1e-7 of
                    # the field's scale)
                    np.testing.assert_allclose(
                        se[key], sb[key], rtol=2e-5,
                        atol=1e-7 * float(np.abs(sb[key]).max()),
                        err_msg=f"{k} {key}")
                for key in ("a_hydro", "u_dt"):
                    scale = np.abs(sb[key]).max()
                    assert np.abs(se[key] - sb[key]).max() <= 1e-4 * scale, (k, key)
        # the default rule switched between the shapes
        assert kinds == {0, 1}


def test_clustered_four_bins_steps_rebuild(orc):
    """The multi-step clustered construction at a size the oracle checks:
    24^3 particles, 1000 Plummer spheres (scale radii in [0.002, 0.02]) and
    a uniform background, u log-uniform in [0.1, 100], on a fixed 4-level
    time-bin hierarchy (1.5 / 4.5 / 19 / 75 % in bins 1-4, so 1.5, 6, 25 or
    100 % of the particles are active), boosted by a bulk velocity so that
    particles leave their cells' stencils within a few steps.  Every step is
    replayed through the entry points against the oracle, re-binning
    included, and engine_step must equal that sequence bit for bit."""
    from swift_amd import Engine, HydroProps, ics
    st = ics.clustered_box(24 ** 3, seed=4, n_spheres=1000, a_min=0.002,
                           a_max=0.02)
    st["v"] += np.asarray((150.0, 80.0, -40.0), np.float32)
    ics.set_bin_hierarchy(st, {1: 0.015, 2: 0.045, 3: 0.19, 4: 0.75}, seed=5)
    # 6^3 cells of width 0.167 against a largest kernel support of ~0.15:
    # a drift of ~0.007 outside a cell leaves the stencil
    props = HydroProps(ncell=(6, 6, 6), time_base=2e-6)
    integ = _integ(fixed=True)
    oi = orc.integ(props.time_base, integ)
    with Engine(props) as a, Engine(props) as b:
        for e in (a, b):
            e.set_particles(st)
            e.set_integration(integ)
        a.set_sparse_shape(4, 32, 0)
        ti = te = 0
        rebuilt = 0
        fractions = set()
        for k in range(10):
            info = a.step()
            ti_old, ti = ti, (0 if k == 0 else te)
            if info["rebuilt"]:
                rebuilt += 1
                g = _state(b)
                orc.drift(g, oi, ti_old, ti)
                b.runner_do_drift_part(-1, ti_old, ti)
                b.rebuild()
                _same(_state(b), {**g, "x": _wrap(g["x"])}, STATE, what="drift")
                te, na = _rest_of_step(orc, b, props, st["m"], oi, ti)
            else:
                te, na = _manual_step(orc, b, props, st["m"], oi, ti_old, ti,
                                      k == 0)
            assert info["ti_end_min"] == te and info["n_active"] == na
            fractions.add(round(na / st["h"].shape[0], 2))
            _same(_state(a), _state(b), STATE, rtol=0, what=f"step {k}")
        assert rebuilt >= 1
        # every level of the hierarchy was active at some step
        assert len(fractions) >= 4, fractions


def test_bin_order_steps():
    """Steps with the particles ordered by time bin inside each cell, with a
    bulk flow that re-bins them on the way: the same time line, neighbour
    counts and state (to FP32 summation order) as the Morton order."""
    from swift_amd import Engine, HydroProps
    st = _box(bulk=(1.5, 0.8, -0.4))
    props = HydroProps(ncell=(4, 4, 4), time_base=4e-3)
    with Engine(props) as a, Engine(props) as b:
        a.set_bin_order(True)
        for e in (a, b):
            e.set_particles(st)
            e.set_integration(_integ(fixed=True))
        rebuilt = 0
        for k in range(6):
            ia, ib = a.step(), b.step()
            rebuilt += ia["rebuilt"]
            assert ia["ti_end_min"] == ib["ti_end_min"], k
            assert ia["n_active"] == ib["n_active"], k
            for key in ("n_density", "n_gradient", "n_force"):
                assert ia["pass"][key] == ib["pass"][key], (k, key)
            sa, sb = _state(a), _state(b)
            np.testing.assert_array_equal(sa["time_bin"], sb["time_bin"])
            # (the kicks of these long steps carry the FP32 summation-order
            # differences of the accelerations into v, x and u step by step)
            for key in ("x", "v", "u", "h", "rho"):
                np.testing.assert_allclose(
                    sa[key], sb[key], rtol=2e-4,
                    atol=1e-6 * float(np.abs(sb[key]).max()),
                    err_msg=f"{k} {key}")
        assert rebuilt >= 1
