"""The allocation-free batched pass (swsph_pass_batched_noalloc) against the
batched pass it is a variant of (needs an MI355X).

swsph_pass_batched_noalloc runs the same kernels as swsph_pass_batched on
the library's fixed device scratch instead of the engine's work queue and
counters, so every output must be bit-identical; the scratch keeps the last
call's contents, so two calls in a row, the second on a smaller box, check
that nothing carries over.  The outputs are also held to the oracle's bar.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from parity import assert_parity

pytestmark = pytest.mark.gpu

MAX_PARTS = 1 << 21


def _fields(res):
    return {k: np.array(v, copy=True) for k, v in res.items()}


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert a[k].tobytes() == b[k].tobytes(), k


def _oracle(orc, props, st, res):
    chk = {k: np.array(v, copy=True) for k, v in st.items()}
    chk["h"] = res["h"]
    ref = orc.pass_fp64(props, chk)
    active = st["time_bin"] <= props.max_active_bin
    assert_parity(res, ref, active)


IO_FIELDS = ("h", "rho", "pressure", "soundspeed", "f", "balsara",
             "alpha_visc", "alpha_diff")


def _boxes():
    """A 24^3 box with half of its particles active (the inactive ones
    carry the state of an earlier all-active pass, as in testActivePair)
    and a smaller all-active 14^3 box."""
    from swift_amd import Engine, HydroProps, ics
    pb = HydroProps(ncell=(4, 4, 4))
    big = ics.modulated_box(24, seed=11, amp=0.5)
    with Engine(pb) as eng:
        eng.set_particles(big)
        eng.pass_batched()
        pre = eng.results(list(IO_FIELDS))
    for k in IO_FIELDS:
        big[k] = pre[k].astype(big[k].dtype)
    ics.set_active_fraction(big, 0.5, seed=12)
    small = ics.modulated_box(14, seed=13, amp=0.3)
    return ((pb, big), (HydroProps(ncell=(3, 3, 3)), small))


def test_noalloc_matches_batched_twice(orc):
    """Two allocation-free passes in a row on one engine (the second on a
    smaller box) against swsph_pass_batched on another engine."""
    from swift_amd import Engine
    from swift_amd._lib import lib
    from swift_amd.engine import PassStats
    cases = _boxes()
    got = []
    with Engine(cases[0][0]) as eng:
        for props, st in cases:
            eng.set_props(props)
            eng.set_particles(st)
            s = PassStats()
            # through the C ABI
            rc = lib().swsph_pass_batched_noalloc(eng._h, C.byref(s))
            assert rc == 0, lib().swsph_last_error()
            got.append((s.as_dict(), _fields(eng.results())))
    for (props, st), (stats, res) in zip(cases, got):
        with Engine(props) as ref_eng:
            ref_eng.set_particles(st)
            ref_stats = ref_eng.pass_batched()
            ref = _fields(ref_eng.results())
        _same(res, ref)
        for k in ("ghost_rounds", "n_density", "n_gradient", "n_force",
                  "n_density_all"):
            assert stats[k] == ref_stats[k], k
        _oracle(orc, props, st, res)


def test_noalloc_at_the_particle_limit():
    """2^21 particles (the limit) on 32^3 cells: bit-identical to the
    batched pass on the same engine and state."""
    from swift_amd import Engine, HydroProps, ics
    st = ics.modulated_box(128, seed=0, amp=0.5)
    assert st["h"].shape[0] == MAX_PARTS
    with Engine(HydroProps(ncell=(32, 32, 32))) as eng:
        eng.set_particles(st)
        eng.save_state()
        a = eng.pass_batched()
        ra = _fields(eng.results())
        eng.restore_state()
        b = eng.pass_batched_noalloc()
        rb = _fields(eng.results())
    _same(ra, rb)
    assert a["n_density"] == b["n_density"] and a["n_force"] == b["n_force"]


def test_noalloc_refusals_leave_the_engine_usable():
    """Sizes beyond the limit, neighbour lists and kernel timing are
    refused before any launch; the engine runs normally afterwards."""
    from swift_amd import Engine, HydroProps, ics
    from swift_amd._lib import SwsphError
    big = ics.modulated_box(129, seed=1, amp=0.2)
    assert big["h"].shape[0] > MAX_PARTS
    props = HydroProps(ncell=(32, 32, 32))
    with Engine(props) as eng:
        eng.set_particles(big)
        with pytest.raises(SwsphError, match="at most"):
            eng.pass_batched_noalloc()
    (props, st), _ = _boxes()
    with Engine(props) as eng:
        eng.set_particles(st)
        eng.save_state()
        eng.set_lists(True)
        with pytest.raises(SwsphError, match="lists"):
            eng.pass_batched_noalloc()
        eng.set_lists(False)
        eng.set_timing(True)
        with pytest.raises(SwsphError, match="timing"):
            eng.pass_batched_noalloc()
        eng.set_timing(False)
        a = eng.pass_batched_noalloc()
        ra = _fields(eng.results())
        eng.restore_state()
        b = eng.pass_batched()
        rb = _fields(eng.results())
    _same(ra, rb)
    assert a["n_force"] == b["n_force"]
