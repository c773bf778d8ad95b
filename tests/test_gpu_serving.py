"""Serving form: the request server gives each box the same result alone
and in a mix of concurrent requests, admits by HBM budget, and keeps
resident boxes on the device between steps (needs an MI355X)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _box(n, seed, amp):
    from swift_amd import HydroProps, ics
    st = ics.modulated_box(n, seed=seed, amp=amp)
    props = HydroProps(ncell=(max(3, n // 4),) * 3)
    return props, st


def test_alone_equals_in_mix(orc):
    from swift_amd import Server
    boxes = [_box(12, 1, 0.5), _box(16, 2, 0.2), _box(10, 3, 0.4),
             _box(20, 4, 0.4)]
    with Server(max_inflight=4) as srv:
        rid = srv.submit(*boxes[0])
        info_alone = srv.wait(rid)
        alone = srv.results(rid)
        srv.release(rid)
        ids = [srv.submit(p, s) for p, s in boxes + boxes[:1]]
        infos = [srv.wait(i) for i in ids]
        mixed = [srv.results(i) for i in ids]
    for k in alone:
        assert np.array_equal(alone[k], mixed[0][k]), k
        assert np.array_equal(alone[k], mixed[-1][k]), k
    assert infos[0]["n_density"] == info_alone["n_density"]
    # the requests overlapped: some started before others finished
    starts = sorted(i["t_start_ms"] for i in infos)
    dones = sorted(i["t_done_ms"] for i in infos)
    assert starts[1] < dones[0]
    # and each result is the oracle's
    from parity import assert_parity
    for (props, st), res in zip(boxes, mixed):
        chk = {k: np.array(v, copy=True) for k, v in st.items()}
        chk["h"] = res["h"]
        ref = orc.pass_fp64(props, chk)
        assert_parity(res, ref, st["time_bin"] <= props.max_active_bin)


def test_budget_admits_one_at_a_time():
    from swift_amd import Server
    boxes = [_box(12, s, 0.5) for s in range(4)]
    # a budget below two boxes: admitted one after the other
    with Server(max_inflight=4, hbm_budget_bytes=7e6) as srv:
        ids = [srv.submit(p, s) for p, s in boxes]
        infos = [srv.wait(i) for i in ids]
    for a, b in zip(infos, infos[1:]):
        assert b["t_admit_ms"] >= a["t_done_ms"] - 1e-6


def test_resident_box_steps():
    """A kept box is stepped again on the device; two steps equal two
    consecutive passes of one engine."""
    from swift_amd import Engine, Server
    props, st = _box(12, 7, 0.5)
    with Server(max_inflight=2) as srv:
        r1 = srv.submit(props, st, keep=True)
        i1 = srv.wait(r1)
        box = i1["box"]
        assert box == r1
        r2 = srv.step(box)
        srv.wait(r2)
        a = srv.results(r2)
        srv.close_box(box)
    with Engine(props) as eng:
        eng.set_particles(st)
        eng.pass_batched()
        eng.pass_batched()
        b = eng.results()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
