"""Edge cases of the pass on the GPU, against the oracle: an empty particle
set, particles exactly on cell faces and box edges (the binning rule and the
periodic images), and a box whose particles all sit in one top-level cell
(every other cell empty), as the reference's tests exercise empty cells and
ragged occupancy (tests/testActivePair.c, tests/test27cells.c)."""
from __future__ import annotations

import numpy as np
import pytest

from parity import assert_parity

pytestmark = pytest.mark.gpu


def _check(orc, props, st, res):
    chk = {k: np.array(v, copy=True) for k, v in st.items()}
    chk["h"] = res["h"]
    ref = orc.pass_fp64(props, chk)
    return assert_parity(res, ref, st["time_bin"] <= props.max_active_bin)


def test_empty_particle_set():
    from swift_amd import Engine, HydroProps, SwsphError, ics
    st = ics.lattice_box(4, seed=0)
    empty = {k: v[:0].copy() for k, v in st.items()}
    with Engine(HydroProps(ncell=(4, 4, 4))) as eng:
        eng.set_particles(empty)
        for run in (eng.pass_batched, eng.pass_plain, eng.pass_graph):
            s = run()
            assert s["n_density"] == 0 and s["n_force"] == 0
        assert eng.get("rho").shape == (0,)
        assert all(eng.cell_count(c) == 0 for c in range(64))
        with pytest.raises(SwsphError):
            eng.step()
        # the engine takes particles again afterwards
        full = ics.lattice_box(12, seed=1)
        eng.set_particles(full)
        s = eng.pass_batched()
        assert s["n_density"] > 0


def test_particles_on_cell_faces_and_box_edges(orc):
    from swift_amd import Engine, HydroProps, ics
    n = 16
    st = ics.lattice_box(n, seed=2, jitter=0.0)
    # lattice points k / 16: every fourth plane lies exactly on the faces of
    # the 4^3 cells, including the box corner 0; some of the points on an
    # inner face move to the last representable position below it (the cell
    # before), which leaves the spacing unchanged to 1e-16
    g = np.arange(n, dtype=np.float64) / n
    ix, iy, iz = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)
    on_face = np.all((x * 4) == np.floor(x * 4), axis=1) & np.all(x > 0, axis=1)
    rng = np.random.default_rng(4)
    idx = rng.choice(np.flatnonzero(on_face), 20, replace=False)
    x[idx] = np.nextafter(x[idx], 0.0)
    st["x"] = x
    props = HydroProps(ncell=(4, 4, 4))
    with Engine(props) as eng:
        eng.set_particles(st)
        eng.pass_batched()
        res = eng.results()
        cells = np.clip(np.floor(x * 4).astype(int), 0, 3)
        cid = (cells[:, 0] * 4 + cells[:, 1]) * 4 + cells[:, 2]
        assert [eng.cell_count(c) for c in range(64)] == \
            np.bincount(cid, minlength=64).tolist()
    _check(orc, props, st, res)


def test_one_occupied_cell(orc):
    """All particles in top-level cell (1, 1, 1) of 4^3; 63 cells empty."""
    from swift_amd import Engine, HydroProps, ics
    k = 10
    st = ics.lattice_box(k, seed=6)
    st["x"] = 0.25 + st["x"] * 0.25           # a 10^3 lattice in one cell
    st["m"] = st["m"] / 64.0
    st["h"] = st["h"] / 4.0
    props = HydroProps(ncell=(4, 4, 4))
    with Engine(props) as eng:
        eng.set_particles(st)
        eng.pass_batched()
        res = eng.results()
        assert eng.cell_count(21) == k ** 3
    _check(orc, props, st, res)


def _pass(props, st):
    from swift_amd import Engine
    with Engine(props) as eng:
        eng.set_particles(st)
        stats = eng.pass_batched()
        res = eng.results()
    return res, stats


def test_no_neighbour_branch_at_h_max(orc):
    """A smoothing length capped so low that no particle finds a neighbour:
    the ghost doubles h (runner_ghost.c:1217-1227) until it reaches h_max,
    clamps it there and ends the particle in hydro_part_has_no_neighbours
    (:1432-1442, hydro.h:791-825).  The jittered lattice keeps every pair
    at least 0.8 spacings apart, beyond gamma_K h_max."""
    from swift_amd import HydroProps, ics
    st = ics.lattice_box(16, seed=3, jitter=0.1)
    st["h"][:] = np.float32(0.004)
    props = HydroProps(ncell=(4, 4, 4), h_max=0.02)
    res, stats = _pass(props, st)
    assert np.all(res["h"] == np.float32(0.02))
    assert np.all(res["n_density"] == 0) and stats["n_force"] == 0
    assert stats["ghost_rounds"] >= 3
    m = st["m"].astype(np.float64)
    kroot = orc.constants()["kernel_root"]
    np.testing.assert_allclose(res["rho"], m * kroot / 0.02 ** 3, rtol=1e-6)
    assert np.all(res["rho_dh"] == 0) and np.all(res["wcount_dh"] == 0)
    _check(orc, props, st, res)


def _density_at(orc, props, st, res, h_eval):
    """Density outputs of the pass against the FP64 rerun at the smoothing
    length of the ghost's last density evaluation (a clamp changes h after
    that evaluation and keeps its sums, runner_ghost.c:1429-1442)."""
    from parity import DENSITY_FIELDS
    chk = {k: np.array(v, copy=True) for k, v in st.items()}
    chk["h"] = np.full_like(st["h"], np.float32(h_eval))
    ref = orc.pass_fp64(props, chk)
    act = st["time_bin"] <= props.max_active_bin
    np.testing.assert_array_equal(res["n_density"], ref["n_density"])
    from parity import compare
    errs = compare(res, ref, act, DENSITY_FIELDS)
    for k, (tol, _) in DENSITY_FIELDS.items():
        assert errs[k] <= tol, (k, errs[k])


# A jittered 16^3 lattice converges to h ~ 0.0772 (eta x spacing); at
# h = 0.05 every particle has a neighbour (the nearest lies within 1.2
# spacings), so the clamps below are reached through the ordinary branch.
def test_h_max_caps_the_bracket(orc):
    """From h = 0.05 the Newton step wants ~0.095; the bracket's upper end
    starts at h_max = 0.07 (runner_ghost.c:1182), so the step lands on it
    and the oscillation rule bisects in h^3 between the last h and h_max
    (:1345-1352) until the change is within the tolerance: every particle
    converges just below h_max, its sums evaluated there."""
    from swift_amd import HydroProps, ics
    st = ics.lattice_box(16, seed=13, jitter=0.1)
    st["h"][:] = np.float32(0.05)
    props = HydroProps(ncell=(4, 4, 4), h_max=0.07)
    res, stats = _pass(props, st)
    assert np.all(res["h"] < np.float32(0.07))
    assert np.all(res["h"] > np.float32(0.07 * (1 - 2 * props.h_tolerance)))
    assert stats["ghost_rounds"] > 5
    _check(orc, props, st, res)


def test_h_max_bound(orc):
    """Starting at h_max with too few neighbours ends the iteration there
    at once (runner_ghost.c:1284-1289): the whole pass at h_max."""
    from swift_amd import HydroProps, ics
    st = ics.lattice_box(16, seed=14, jitter=0.1)
    st["h"][:] = np.float32(0.07)
    props = HydroProps(ncell=(4, 4, 4), h_max=0.07)
    res, stats = _pass(props, st)
    assert np.all(res["h"] == np.float32(0.07)) and stats["ghost_rounds"] == 1
    _check(orc, props, st, res)


def test_h_min_clamp_after_a_step(orc):
    """From h = 0.105 the Newton step goes below h_min = 0.1: the particle
    ends at h_min with the sums of its evaluation at 0.105."""
    from swift_amd import HydroProps, ics
    st = ics.lattice_box(16, seed=15, jitter=0.1)
    st["h"][:] = np.float32(0.105)
    props = HydroProps(ncell=(4, 4, 4), h_min=0.1)
    res, stats = _pass(props, st)
    assert np.all(res["h"] == np.float32(0.1))
    assert stats["ghost_rounds"] == 1
    _density_at(orc, props, st, res, 0.105)


def test_h_min_bound(orc):
    """Starting at h_min with too many neighbours ends the iteration there
    at once: the whole pass at h_min."""
    from swift_amd import HydroProps, ics
    st = ics.lattice_box(16, seed=16, jitter=0.1)
    st["h"][:] = np.float32(0.1)
    props = HydroProps(ncell=(4, 4, 4), h_min=0.1)
    res, stats = _pass(props, st)
    assert np.all(res["h"] == np.float32(0.1)) and stats["ghost_rounds"] == 1
    _check(orc, props, st, res)


def test_unconverged_is_an_error_and_engine_recovers(orc):
    """max_ghost_iterations = 1 with a poor initial h: status 3 (the
    reference's error(), runner_ghost.c:1592); the same engine then runs a
    pass with the default limit and matches the oracle."""
    from swift_amd import Engine, HydroProps, SwsphError, ics
    st = ics.modulated_box(16, seed=9, amp=0.5)
    st["h"] = (st["h"] * np.float32(0.6)).astype(np.float32)
    props = HydroProps(ncell=(4, 4, 4), max_iter=1)
    with Engine(props) as eng:
        eng.set_particles(st)
        with pytest.raises(SwsphError, match="status 3.*converge"):
            eng.pass_batched()
        ok = HydroProps(ncell=(4, 4, 4))
        eng.set_props(ok)
        eng.set_particles(st)
        eng.pass_batched()
        res = eng.results()
    _check(orc, ok, st, res)


def test_coincident_particles(orc):
    """Pairs of particles at exactly the same position: r = 0, where the
    interactions take r_inv = 0 (hydro_iact.h:94, 169, 549)."""
    from swift_amd import HydroProps, ics
    st = ics.modulated_box(16, seed=10, amp=0.3)
    rng = np.random.default_rng(11)
    src = rng.choice(st["h"].shape[0], 40, replace=False)
    st["x"][src[20:]] = st["x"][src[:20]]
    props = HydroProps(ncell=(4, 4, 4))
    res, _ = _pass(props, st)
    errs = _check(orc, props, st, res)
    assert errs["n_density"] == 0


def test_zero_energy_pair(orc):
    """Two neighbouring particles with u = 0 (so P_i + P_j = 0): the
    reference divides by P_i + P_j in the diffusion coefficient
    (hydro_iact.h:610-611) and both particles get u_dt = NaN; everything
    else, and every other particle, matches the oracle as usual."""
    from parity import FORCE_FIELDS, DENSITY_FIELDS, compare
    from swift_amd import HydroProps, ics
    st = ics.modulated_box(16, seed=12, amp=0.3)
    props = HydroProps(ncell=(4, 4, 4))
    # two particles that are neighbours of each other: a close pair
    x = st["x"]
    d2 = ((x[1:200, None, :] - x[None, :200, :]) ** 2).sum(-1)
    i, j = np.unravel_index(np.argmin(np.where(d2 > 0, d2, 9.0)), d2.shape)
    i += 1
    st["u"][[i, j]] = 0.0
    res, _ = _pass(props, st)
    chk = {k: np.array(v, copy=True) for k, v in st.items()}
    chk["h"] = res["h"]
    ref = orc.pass_fp64(props, chk)
    nan_g = np.isnan(res["u_dt"])
    nan_r = np.isnan(ref["u_dt"])
    assert nan_g[i] and nan_g[j]
    np.testing.assert_array_equal(nan_g, nan_r)
    ok = ~nan_r
    fields = {k: v for k, v in {**DENSITY_FIELDS, **FORCE_FIELDS}.items()}
    errs = compare(res, ref, ok, fields)
    for k, (tol, _) in fields.items():
        assert errs[k] <= tol, (k, errs[k])
