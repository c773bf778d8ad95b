"""Pins the oracle's time integration (drift, kicks, time steps) on the CPU.

The checks come from what the mathematics and the time line fix, not from
a second copy of the oracle's code:
  * get_max_active_bin: the active bins at time ti are those whose step
    2^(b+1) divides ti (src/timeline.h), i.e. bin = ctz(ti) - 1;
  * make_integer_timestep (src/timestep.h:49-89): the step is the largest
    power of two allowed by the requested dt, by the neighbour limiter
    (bin <= min_ngb_bin + 2) and by the growth limit (<= 2 x the current
    step), falling back to the current step when the larger one would not
    start on its own grid -- found here by search over powers of two;
  * drift (src/drift.h:138, hydro.h:1029): x advances by v_full dt exactly
    in FP64; h and rho follow exp(+-w) to the order of the expansion, and
    rho h^3 is conserved; P and c follow the ideal-gas EOS;
  * kicks (src/kick.h, hydro.h:1120): two half-kicks over a step add a dt
    to v_full; u_full never drops below half per kick nor below min_u;
  * the CFL step (hydro.h:475-487) puts each particle in the bin whose step
    is the largest power of two below 2 gamma_K CFL h / v_sig.
"""
from __future__ import annotations

import numpy as np
import pytest

GAMMA = 5.0 / 3.0
KG = 1.936492


def _ctz(v: int) -> int:
    return (v & -v).bit_length() - 1


def _gas(n, rng, bins=(1, 2, 3)):
    f32 = np.float32
    g = {
        "x": rng.uniform(0.1, 0.9, (n, 3)),
        "v": rng.normal(0, 0.3, (n, 3)).astype(f32),
        "a_hydro": rng.normal(0, 2.0, (n, 3)).astype(f32),
        "u": rng.uniform(0.5, 1.5, n).astype(f32),
        "u_dt": rng.normal(0, 0.5, n).astype(f32),
        "h": rng.uniform(0.02, 0.04, n).astype(f32),
        "h_dt": rng.normal(0, 0.05, n).astype(f32),
        "rho": rng.uniform(0.5, 2.0, n).astype(f32),
        "pressure": np.zeros(n, f32), "soundspeed": np.zeros(n, f32),
        "v_sig": rng.uniform(1.0, 3.0, n).astype(f32),
        "time_bin": rng.choice(np.array(bins, np.int8), n),
        "min_ngb_time_bin": np.full(n, 57, np.int8),
    }
    g["v_full"] = g["v"].copy()
    g["u_full"] = g["u"].copy()
    return g


class _P:
    cfl = 0.1
    log_max_h_change = float(np.log(np.float32(1.4) ** np.float32(1 / 3)))
    min_u = 0.0
    dt_min = 0.0
    dt_max = 1e30
    fixed_bins = False


def test_max_active_bin_is_the_trailing_zero_rule(orc):
    assert orc.max_active_bin(0) == 56
    for ti in range(4, 1 << 14, 4):
        assert orc.max_active_bin(ti) == _ctz(ti) - 1, ti


def test_make_integer_timestep_is_the_largest_allowed_power_of_two(orc):
    rng = np.random.default_rng(3)
    top = 1 << 57
    for _ in range(4000):
        tbi = float(rng.choice([1.0, 1e3, 1e6]))
        new_dt = np.float32(10 ** rng.uniform(-6, 3) / tbi * 1e3)
        old_bin = int(rng.integers(0, 12))
        ngb = int(rng.choice([57, int(rng.integers(1, 12))]))
        cur = 0 if old_bin <= 0 else 1 << (old_bin + 1)
        # a time on the grid of the old step (the particle is active)
        ti = cur * int(rng.integers(1, 50)) if cur else 0
        got = orc.make_integer_timestep(float(new_dt), old_bin, ngb, ti, tbi)
        raw = int(np.float64(new_dt) * tbi)
        if raw <= 0:
            assert got == 0
            continue
        cand = [1 << k for k in range(58)
                if (1 << k) <= raw and k <= ngb + 3 and k >= 2
                and (old_bin <= 0 or (1 << k) <= 2 * cur)]
        # the time bin of a step 2^(b+1) is b, so steps below 4 map to bins
        # <= 0 (step 0): no candidate
        want = max(cand) if cand else 0
        # the particle is active: its current step ends at ti
        if want > cur and (top - ti) % want:
            want = cur
        assert got == want, (new_dt, old_bin, ngb, ti, tbi, got, want)


def test_drift_closed_forms(orc):
    rng = np.random.default_rng(4)
    g = _gas(5000, rng)
    c = orc.integ(1e-4, _P)
    g0 = {k: v.copy() for k, v in g.items()}
    orc.drift(g, c, 16, 48)
    dt = 32 * 1e-4
    # positions: one FP64 multiply-add per component
    np.testing.assert_array_equal(g["x"], g0["x"] + g0["v_full"].astype(np.float64) * dt)
    # predicted velocities
    v_exp = (g0["v"].astype(np.float64) + g0["a_hydro"].astype(np.float64) * dt).astype(np.float32)
    np.testing.assert_array_equal(g["v"], v_exp)
    # h follows exp(w), rho exp(-3w): 4th-order expansion error + rounding
    w = g0["h_dt"].astype(np.float64) / g0["h"] * dt
    assert np.all(np.abs(w) < 0.2)
    rel_h = np.abs(g["h"] / (g0["h"] * np.exp(w)) - 1)
    assert rel_h.max() < np.abs(w).max() ** 5 / 120 + 4e-7
    vol = (g["rho"].astype(np.float64) * g["h"].astype(np.float64) ** 3 /
           (g0["rho"].astype(np.float64) * g0["h"].astype(np.float64) ** 3))
    assert np.abs(vol - 1).max() < 2e-6
    # internal energy and EOS
    u = np.maximum(g0["u"] + g0["u_dt"] * np.float32(dt), 0)
    np.testing.assert_allclose(g["u"], u, rtol=1e-6, atol=1e-7)
    P = (GAMMA - 1) * g["u"].astype(np.float64) * g["rho"]
    np.testing.assert_allclose(g["pressure"], P, rtol=3e-7, atol=1e-12)
    np.testing.assert_allclose(g["soundspeed"], np.sqrt(GAMMA * P / g["rho"]),
                               rtol=3e-7)
    assert np.all(g["v_sig"] >= 2 * g["soundspeed"] * (1 - 1e-7))
    # a large step takes the exact exponential
    g2 = _gas(100, rng)
    g2["h_dt"][:] = 0.05
    h0 = g2["h"].copy()
    orc.drift(g2, c, 0, 4000)
    w2 = 0.05 / h0.astype(np.float64) * 0.4
    assert np.all(w2 > 0.2)
    np.testing.assert_allclose(g2["h"], h0 * np.exp(w2), rtol=1e-6)


def test_kicks_add_one_step_of_acceleration(orc):
    rng = np.random.default_rng(5)
    g = _gas(3000, rng, bins=(2,))
    g["u_dt"][:] = np.abs(g["u_dt"])  # no floor / halving in play
    tb = 1e-3
    c = orc.integ(tb, _P)
    g0 = {k: v.copy() for k, v in g.items()}
    orc.kick(g, c, 0, 1)    # kick1 at t = 0: half of the step of 8
    orc.kick(g, c, 8, 2)    # kick2 at the end of the step
    dt = 8 * tb
    np.testing.assert_allclose(g["v_full"], g0["v_full"] + g0["a_hydro"] * dt,
                               rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(g["u_full"], g0["u_full"] + g0["u_dt"] * dt,
                               rtol=1e-6)
    # kick2 resets the predicted values to the kicked ones
    np.testing.assert_array_equal(g["v"], g["v_full"])
    np.testing.assert_array_equal(g["u"], g["u_full"])
    # particles in other bins are left alone by kick2 at t = 8 if inactive
    g = _gas(1000, rng, bins=(3,))
    g0 = {k: v.copy() for k, v in g.items()}
    orc.kick(g, c, 8, 2)  # bin 3 (step 16) is not active at t = 8
    np.testing.assert_array_equal(g["v_full"], g0["v_full"])
    # a strongly negative du/dt halves u at most; the floor zeroes du/dt
    g = _gas(1000, rng, bins=(2,))
    g["u_dt"][:] = -1e6
    u0 = g["u_full"].copy()
    orc.kick(g, c, 0, 1)
    np.testing.assert_allclose(g["u_full"], 0.5 * u0, rtol=1e-7)
    P = _P()
    P.min_u = 0.9
    c2 = orc.integ(tb, P)
    g = _gas(1000, rng, bins=(2,))
    g["u_dt"][:] = -1e6
    orc.kick(g, c2, 0, 1)
    assert np.all(g["u_full"] >= np.float32(0.9))
    assert np.all(g["u_dt"][g["u_full"] == np.float32(0.9)] == 0)


def test_cfl_bins(orc):
    rng = np.random.default_rng(6)
    g = _gas(4000, rng, bins=(0,))
    g["h_dt"][:] = 0.0          # only the CFL condition
    tb = 1e-7
    c = orc.integ(tb, _P)
    te, na, bad = orc.timestep(g, c, 0)
    assert bad == 0 and na == 4000
    dt_cfl = 2 * np.float32(KG) * 0.1 * g["h"].astype(np.float64) / g["v_sig"]
    step = (2.0 ** (g["time_bin"].astype(np.float64) + 1)) * tb
    assert np.all(step <= dt_cfl * (1 + 1e-6))
    assert np.all(2 * step > dt_cfl * (1 - 1e-6))
    assert te == int(np.min(2 ** (g["time_bin"].astype(np.int64) + 1)))
    # the h-change limit takes over when |h_dt| is large
    g2 = _gas(500, rng, bins=(0,))
    g2["h_dt"][:] = 1e3
    orc.timestep(g2, c, 0)
    lim = _P.log_max_h_change * g2["h"].astype(np.float64) / 1e3
    step2 = (2.0 ** (g2["time_bin"].astype(np.float64) + 1)) * tb
    assert np.all(step2 <= lim * (1 + 1e-6))
    # dt below dt_min is reported
    P = _P()
    P.dt_min = 1.0
    assert orc.timestep(_gas(10, rng, bins=(0,)), orc.integ(tb, P), 0)[2] == 10


@pytest.mark.parametrize("bins", [(1,), (1, 2, 3, 4)])
def test_fixed_bins_schedule(orc, bins):
    """With fixed bins the synchronisation points walk the grid of the
    smallest step and the active set at ti is exactly the bins <= ctz-1."""
    rng = np.random.default_rng(7)
    g = _gas(2000, rng, bins=bins)
    P = _P()
    P.fixed_bins = True
    c = orc.integ(1e-6, P)
    te, na, bad = orc.timestep(g, c, 0)
    assert bad == 0 and na == 2000 and te == 1 << (min(bins) + 1)
    ti = te
    for _ in range(8):
        te, na, bad = orc.timestep(g, c, ti)
        mab = orc.max_active_bin(ti)
        assert na == int(np.count_nonzero(g["time_bin"] <= mab))
        assert te == ti + (1 << (min(bins) + 1))
        ti = te
