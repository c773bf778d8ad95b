"""The pass driven by the reference's task graph (swsph_sched_pass) and the
bundles of cell tasks it releases (needs an MI355X).

The scheduler builds the hydro tasks of the cells with active particles and
their dependencies (engine_make_hydroloop_tasks_mapper and engine_make_-
hydro_loops_dependencies, src/engine_maketasks.c:3501, 2546-2573) and
releases every set of ready tasks as one bundle per task type
(scheduler_done, src/scheduler.c:1203-1222).  Its results are held to the
oracle's bar like the other passes: neighbour counts exactly, density
fields within 1e-5 and force fields within 1e-4 of the FP64 rerun at the
converged smoothing lengths; the graph's size and release rounds are
checked against a count of the tasks from the active cells.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from parity import DENSITY_FIELDS, FORCE_FIELDS, assert_alpha, assert_parity

pytestmark = pytest.mark.gpu

IO = ("h", "rho", "pressure", "soundspeed", "f", "balsara", "alpha_visc",
      "alpha_diff")


def _check(orc, props, st, res):
    chk = {k: np.array(v, copy=True) for k, v in st.items()}
    chk["h"] = res["h"]
    ref = orc.pass_fp64(props, chk)
    assert_parity(res, ref, st["time_bin"] <= props.max_active_bin)


def _expected_tasks(act, nc):
    """Tasks of the graph from the cells' activity: 4 cell tasks per active
    cell, and per loop a self task per active cell and a pair task per
    adjacent pair with an active side."""
    a = act.reshape(nc)
    pairs = 0
    for o in range(14, 27):
        sh = (-(o // 9 - 1), -((o // 3) % 3 - 1), -(o % 3 - 1))
        b = np.roll(a, sh, axis=(0, 1, 2))
        pairs += int(np.count_nonzero(a | b))
    nact = int(a.sum())
    return 4 * nact + 3 * (nact + pairs), nact


def _partial_box(seed, frac):
    from swift_amd import Engine, HydroProps, ics
    props = HydroProps(ncell=(5, 5, 5))
    st = ics.clustered_box(20000, seed=seed, n_spheres=8, a_min=0.02,
                           a_max=0.05)
    with Engine(props) as eng:
        eng.set_particles(st)
        eng.pass_batched()
        pre = eng.results(list(IO))
    for k in IO:
        st[k] = pre[k].astype(st[k].dtype)
    ics.set_active_fraction(st, frac, seed=seed + 1)
    return props, st


@pytest.mark.parametrize("case", ["modulated", "clustered_30", "clustered_2"])
def test_sched_pass_matches_oracle(orc, case):
    from swift_amd import Engine, HydroProps, ics
    if case == "modulated":
        props = HydroProps(ncell=(4, 4, 4))
        st = ics.modulated_box(18, seed=31, amp=0.5)
    else:
        props, st = _partial_box(33, 0.3 if case.endswith("30") else 0.02)
    with Engine(props) as eng:
        eng.set_particles(st)
        act_cells = eng.cell_active()
        ss, ps = eng.sched_pass()
        res = eng.results()
    _check(orc, props, st, res)
    assert ps["n_density"] == int(res["n_density"][st["time_bin"] <= props.max_active_bin].sum())
    want, nact = _expected_tasks(act_cells.astype(bool), props.ncell)
    assert ss["active_cells"] == nact
    assert ss["tasks"] == want
    # init -> density -> ghost -> gradient -> extra ghost -> force -> end
    assert ss["levels"] == 7 and ss["bundles"] == 7


def test_cell_active_matches_time_bins():
    from swift_amd import Engine
    props, st = _partial_box(35, 0.02)
    inv_w = 1.0 / (np.array(props.dim) / np.array(props.ncell))
    cc = np.clip((st["x"] * inv_w).astype(int), 0, np.array(props.ncell) - 1)
    cid = (cc[:, 0] * props.ncell[1] + cc[:, 1]) * props.ncell[2] + cc[:, 2]
    want = np.zeros(np.prod(props.ncell), np.uint8)
    want[cid[st["time_bin"] <= props.max_active_bin]] = 1
    with Engine(props) as eng:
        eng.set_particles(st)
        got = eng.cell_active()
    np.testing.assert_array_equal(got, want)
    assert 0 < want.sum() < want.size


def test_cell_bundles_match_per_cell_tasks(orc):
    """The task sequence with every cell task bundled over all cells (and
    the loop tasks as per-cell calls) against the oracle."""
    from swift_amd import Engine, HydroProps, ics
    props = HydroProps(ncell=(3, 3, 3))
    st = ics.modulated_box(14, seed=37, amp=0.6)
    with Engine(props) as eng:
        eng.set_particles(st)
        n = eng.num_cells()
        cells = np.arange(n)
        eng.init_cells(cells)
        loops = [(eng.runner_doself1_branch_density,
                  eng.runner_dopair1_branch_density),
                 (eng.runner_doself1_branch_gradient,
                  eng.runner_dopair1_branch_gradient),
                 (eng.runner_doself2_branch_force,
                  eng.runner_dopair2_branch_force)]
        rounds = 0
        for k, (self_fn, pair_fn) in enumerate(loops):
            for c in range(n):
                self_fn(c)
                cx, cy, cz = eng.cell_coords(c)
                for o in range(14, 27):
                    pair_fn(c, eng.cell_id(cx + o // 9 - 1,
                                           cy + (o // 3) % 3 - 1,
                                           cz + o % 3 - 1))
            if k == 0:
                rounds = eng.runner_do_ghost_cells(cells)
            elif k == 1:
                eng.runner_do_extra_ghost_cells(cells)
            else:
                eng.runner_do_end_hydro_force_cells(cells)
        res = eng.results()
    assert rounds >= 2
    _check(orc, props, st, res)


def test_cell_bundle_errors():
    from swift_amd import Engine, HydroProps, ics
    from swift_amd._lib import SwsphError
    st = ics.modulated_box(12, seed=2)
    with Engine(HydroProps(ncell=(3, 3, 3))) as eng:
        eng.set_particles(st)
        with pytest.raises(SwsphError, match="out of range"):
            eng.runner_do_ghost_cells([0, 27])
        with pytest.raises(SwsphError, match="cells listed"):
            eng.runner_do_extra_ghost_cells(np.zeros(28, np.int32))
        eng.runner_do_end_hydro_force_cells([])   # no cells, no work
        stats = eng.pass_batched()                # usable afterwards
        assert stats["n_force"] > 0


def test_engine_step_through_the_task_graph(orc):
    """Steps of a 3-level hierarchy whose passes run through the task graph:
    each step's pass against the oracle (drift and kicks through the entry
    points, as the step-by-step test of the device queue does), and
    swsph_engine_step in scheduler mode on the same time line as in
    work-queue mode."""
    from swift_amd import Engine, HydroProps, ics
    from swift_amd.engine import integration
    st = ics.modulated_box(16, seed=3)
    ics.set_bin_hierarchy(st, {1: 0.1, 2: 0.2, 3: 0.7}, seed=3)
    props = HydroProps(ncell=(4, 4, 4), time_base=1e-4)
    m = st["m"]
    lines = []
    for sched in (False, True):
        with Engine(props) as eng:
            eng.set_particles(st)
            eng.set_integration(integration(fixed_bins=True))
            eng.set_step_sched(sched)
            line = []
            for _ in range(5):
                info = eng.step()
                line.append((info["ti_current"], info["n_active"],
                             info["pass"]["n_density"] > 0))
            lines.append(line)
            x = eng.get("x")
        if sched:
            assert np.all(np.isfinite(x))
    assert lines[0] == lines[1]
    # the pass at the second synchronisation point through the graph,
    # checked against the oracle
    with Engine(props) as eng:
        eng.set_particles(st)
        eng.set_integration(integration(fixed_bins=True))
        first = eng.step()
        ti = first["ti_end_min"]
        eng.runner_do_drift_part(-1, 0, ti)
        mab = orc.max_active_bin(ti)
        p2 = dataclasses.replace(props, max_active_bin=mab)
        pre = {"x": eng.get("x"), "v": eng.get("v"), "m": m, "u": eng.get("u"),
               "time_bin": eng.get("time_bin")}
        for k in ("rho", "pressure", "soundspeed", "f", "balsara",
                  "alpha_visc", "alpha_diff", "div_v_prev"):
            pre[k] = eng.get(k)
        pre["h"] = eng.get("h")
        eng.set_props(p2)
        ss, ps = eng.sched_pass()
        res = eng.results()
    act = pre["time_bin"] <= mab
    assert 0 < act.sum() < act.size
    chk = {k: np.array(v, copy=True) for k, v in pre.items()}
    chk["h"] = res["h"]
    ref = orc.pass_fp64(p2, chk)
    fields = {**DENSITY_FIELDS, **{k: v for k, v in FORCE_FIELDS.items()
                                   if not k.startswith("alpha")}}
    assert_parity(res, ref, act, fields=fields)
    pre["h"] = res["h"]
    assert_alpha(res, ref, pre, p2, p2.time_base, act)
