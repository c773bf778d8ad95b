"""The drop-in boundary beyond whole passes (needs an MI355X): the
caller-subset density entry points, the ghost split in shares, the
write-back into the caller's particle memory, and the C program that
replays the test27cells call sequence through the C ABI.

The subset functions take the reference's (parts, ind, count) arguments
(src/runner_doiact_hydro.h:197-210): particle k of the subset is the
caller's particle first + ind[k].  Driven as the ghost drives them
(runner_ghost.c:1548-1580) -- a subset of a cell against the cell and its
26 neighbours -- they must give those particles their full density sums:
checked against the oracle's FP64 rerun.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from parity import DENSITY_FIELDS, assert_parity, rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cells_of(props, x):
    # the reference's cell rule: (int)(x * iwidth), iwidth = 1 / (dim / cdim)
    inv_w = 1.0 / (np.array(props.dim) / np.array(props.ncell))
    return np.clip((x * inv_w).astype(int), 0, np.array(props.ncell) - 1)


def _neighbours(eng, c):
    cx, cy, cz = eng.cell_coords(c)
    return [eng.cell_id(cx + o // 9 - 1, cy + (o // 3) % 3 - 1, cz + o % 3 - 1)
            for o in range(27) if o != 13]


def test_subset_density_matches_oracle(orc):
    """Half of a cell's particles through the subset entry points (self +
    26 pairs, indices relative to an offset as SWIFT's parts/ind), the
    ghost at an h tolerance so loose that h stays put: the subset has its
    full density sums, the rest of the cell only the self term."""
    from swift_amd import Engine, HydroProps, ics
    st = ics.modulated_box(16, seed=21, amp=0.4)
    props = HydroProps(ncell=(4, 4, 4), h_tolerance=1.0)
    cc = _cells_of(props, st["x"])
    with Engine(props) as eng:
        eng.set_particles(st)
        main = eng.cell_id(2, 1, 3)
        cx, cy, cz = eng.cell_coords(main)
        in_main = np.flatnonzero((cc[:, 0] == cx) & (cc[:, 1] == cy) &
                                 (cc[:, 2] == cz))
        assert in_main.size == eng.cell_count(main)
        rng = np.random.default_rng(3)
        sub = np.sort(rng.choice(in_main, in_main.size // 2, replace=False))
        first = int(sub[0])
        ind = (sub - first).astype(np.int32)
        for c in range(eng.num_cells()):
            eng.hydro_init_cell(c)
        eng.runner_doself_subset_branch_density(main, ind, first=first)
        for cj in _neighbours(eng, main):
            eng.runner_dopair_subset_branch_density(main, ind, cj, first=first)
        eng.runner_do_ghost(main)
        res = eng.results(["h", "rho", "wcount", "div_v", "rot_v",
                           "n_density"])
    assert np.array_equal(res["h"], st["h"])
    ref = orc.pass_fp64(props, {k: np.array(v, copy=True) for k, v in st.items()})
    mask = np.zeros(st["h"].shape[0], bool)
    mask[sub] = True
    assert np.array_equal(res["n_density"][mask], ref["n_density"][mask])
    for k, (tol, sk) in DENSITY_FIELDS.items():
        if k in res:
            e = rel_err(res[k], ref[k], ref[sk] if sk else None, mask)
            assert e <= tol, (k, e)
    rest = np.setdiff1d(in_main, sub)
    assert rest.size > 0 and np.all(res["n_density"][rest] == 0)


def test_subset_rejects_foreign_and_out_of_range_particles():
    from swift_amd import Engine, HydroProps, ics
    from swift_amd._lib import SwsphError
    st = ics.lattice_box(12, seed=2)
    props = HydroProps(ncell=(3, 3, 3))
    cc = _cells_of(props, st["x"])
    with Engine(props) as eng:
        eng.set_particles(st)
        c0 = eng.cell_id(0, 0, 0)
        other = np.flatnonzero((cc[:, 0] == 1))[:3].astype(np.int32)
        with pytest.raises(SwsphError, match="not in cell"):
            eng.runner_doself_subset_branch_density(c0, other)
        with pytest.raises(SwsphError, match="out of range"):
            eng.runner_doself_subset_branch_density(
                c0, np.array([st["h"].shape[0]], np.int32))
        with pytest.raises(SwsphError, match="share"):
            eng.runner_do_ghost(c0, 2, 2)
        # the engine is usable afterwards
        stats = eng.pass_batched()
        assert stats["n_density"] > 0


def _task_graph(eng, shares):
    """One step's task graph over every cell, the ghost of each cell in
    `shares` tasks."""
    n = eng.num_cells()
    for c in range(n):
        eng.hydro_init_cell(c)
    loops = [(eng.runner_doself1_branch_density,
              eng.runner_dopair1_branch_density),
             (eng.runner_doself1_branch_gradient,
              eng.runner_dopair1_branch_gradient),
             (eng.runner_doself2_branch_force,
              eng.runner_dopair2_branch_force)]
    for k, (self_fn, pair_fn) in enumerate(loops):
        for c in range(n):
            self_fn(c)
            cx, cy, cz = eng.cell_coords(c)
            for o in range(14, 27):
                pair_fn(c, eng.cell_id(cx + o // 9 - 1, cy + (o // 3) % 3 - 1,
                                       cz + o % 3 - 1))
        for c in range(n):
            if k == 0:
                for off in range(shares):
                    eng.runner_do_ghost(c, off, shares)
            elif k == 1:
                eng.runner_do_extra_ghost(c)
            else:
                eng.runner_do_end_hydro_force(c)


def test_ghost_shares_equal_one_ghost(orc):
    """runner_do_ghost's offset / ntasks split: three shares per cell give
    bit for bit the results of one ghost per cell, and both match the
    oracle."""
    from swift_amd import Engine, HydroProps, ics
    st = ics.modulated_box(14, seed=5, amp=0.6)
    props = HydroProps(ncell=(3, 3, 3))
    out = []
    for shares in (1, 3):
        with Engine(props) as eng:
            eng.set_particles(st)
            _task_graph(eng, shares)
            out.append(eng.results())
    for k in out[0]:
        assert out[0][k].tobytes() == out[1][k].tobytes(), k
    chk = {k: np.array(v, copy=True) for k, v in st.items()}
    chk["h"] = out[1]["h"]
    ref = orc.pass_fp64(props, chk)
    assert_parity(out[1], ref, np.ones(st["h"].shape[0], bool))


def test_write_back_into_an_array_of_structs():
    """swsph_cell_write_back fills the fields of a struct-part-like array
    (strided, mixed element types) for one cell, then for all particles,
    leaving every other byte alone."""
    from swift_amd import Engine, HydroProps, ics
    st = ics.modulated_box(12, seed=9, amp=0.3)
    props = HydroProps(ncell=(3, 3, 3))
    n = st["h"].shape[0]
    part = np.dtype([("x", "f8", 3), ("pad0", "i4"), ("rho", "f4"),
                     ("a_hydro", "f4", 3), ("n_force", "i4"),
                     ("min_ngb_time_bin", "i1"), ("pad1", "i1", 3),
                     ("h", "f4")])
    arr = np.zeros(n, part)
    arr["pad0"] = -7
    arr["pad1"] = 5
    arr["rho"] = -1.0
    cc = _cells_of(props, st["x"])
    names = ["x", "rho", "a_hydro", "n_force", "min_ngb_time_bin", "h"]
    with Engine(props) as eng:
        eng.set_particles(st)
        eng.pass_batched()
        ref = eng.results(names)
        c = eng.cell_id(1, 1, 1)
        eng.write_back(c, {k: arr[k] for k in names})
        cx, cy, cz = eng.cell_coords(c)
        inc = (cc[:, 0] == cx) & (cc[:, 1] == cy) & (cc[:, 2] == cz)
        assert inc.sum() == eng.cell_count(c) > 0
        for k in names:
            assert np.array_equal(arr[k][inc], ref[k][inc]), k
        assert np.all(arr["rho"][~inc] == -1.0)
        eng.write_back(-1, {k: arr[k] for k in names})
    for k in names:
        assert np.array_equal(arr[k], ref[k]), k
    assert np.all(arr["pad0"] == -7) and np.all(arr["pad1"] == 5)


def test_c_program_replays_test27cells():
    """tests/c/abi_27cells.c, compiled with gcc against include/swsph.h:
    the 27-cell task sequence through the C ABI against the oracle."""
    exe = os.path.join(ROOT, "tests", "c", "abi_27cells")
    assert os.path.exists(exe), "build it with make -C tests/c"
    r = subprocess.run([exe, "15"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "0 failures" in r.stdout
