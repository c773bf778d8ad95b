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
