"""Pinning the CPU oracle (runs without a GPU).

Each check ties the oracle to something it cannot simply restate: the
closed-form Wendland C2 kernel and its normalisation, the reference's kernel
invariants (tests/testKernel.c:60-80), the symmetric == nonsym(i,j) +
nonsym(j,i) identity (tests/testSymmetry.c:113-167), an O(N^2) brute force
(tools.c:207-763), the analytic fields of tests/test27cells.c:460-463 and
tests/test125cells.c:148-207, and the Newton-Raphson stopping rule of the
ghost (runner_ghost.c:1250-1388) checked in FP64 plus an independent FP64
bisection for h.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from swift_amd import HydroProps, ics

GAMMA = 1.936492
K = 21.0 / (2.0 * math.pi)


def test_constants(orc):
    c = orc.constants()
    g = np.float32(GAMMA)
    assert c["kernel_gamma"] == g
    assert c["kernel_gamma2"] == np.float32(g * g)
    assert abs(c["kernel_root"] - K / GAMMA ** 3) < 2e-7
    assert abs(c["kernel_norm"] - 4.0 * math.pi / 3.0 * GAMMA ** 3) < 1e-4


def test_kernel_normalisation(orc):
    """int W d^3u = 1 and int (3W + u W') d^3u = 0 (dW/dh of the norm)."""
    n = 20000
    u = (np.arange(n) + 0.5) / n * GAMMA
    du = GAMMA / n
    w = np.empty(n)
    dw = np.empty(n)
    for k in range(n):
        w[k], dw[k] = orc.kernel_deval64(u[k])
    norm = np.sum(4 * math.pi * u ** 2 * w) * du
    assert abs(norm - 1.0) < 1e-6
    dnorm = np.sum(4 * math.pi * u ** 2 * (3 * w + u * dw)) * du
    assert abs(dnorm) < 1e-6
    # W' is the derivative of W (central differences)
    eps = 1e-6
    for uu in (0.1, 0.5, 1.0, 1.5, 1.9):
        wp, _ = orc.kernel_deval64(uu + eps)
        wm, _ = orc.kernel_deval64(uu - eps)
        _, d = orc.kernel_deval64(uu)
        assert abs((wp - wm) / (2 * eps) - d) < 1e-6


def test_kernel_fp32_invariants(orc):
    """testKernel.c: W >= 0, dW <= 0; W(0) = kernel_root; FP32 Horner form
    within its cancellation bound of the FP64 closed form; zero outside."""
    rng = np.random.default_rng(0)
    us = np.concatenate([rng.uniform(0, 2.2, 4000), [0.0, GAMMA * 0.999999,
                                                     GAMMA, 2.5]])
    w0 = orc.constants()["kernel_root"]
    for uu in us:
        w32, d32 = orc.kernel_deval32(float(uu))
        w64, d64 = orc.kernel_deval64(float(np.float32(uu)))
        assert w32 >= 0.0 and d32 <= 0.0
        assert abs(w32 - w64) <= 4e-6 * w0
        assert abs(d32 - d64) <= 4e-6 * w0 * 10
        if uu >= GAMMA:
            assert w64 == 0.0 and d64 == 0.0
    assert orc.kernel_deval32(0.0)[0] == pytest.approx(w0, rel=1e-7)


def _rand_part(rng, orc):
    p = rng.uniform(0.1, 2.0, len(orc.PART_FIELDS)).astype(np.float32)
    F = {k: i for i, k in enumerate(orc.PART_FIELDS)}
    for k in ("RHO_DH", "WC", "WC_DH", "DIVV", "ROT0", "ROT1", "ROT2", "LAPU",
              "A0", "A1", "A2", "UDT", "HDT", "NDEN", "NGRAD", "NFORCE"):
        p[F[k]] = 0.0
    p[F["TB"]] = np.float32(rng.integers(1, 10))
    p[F["MINTB"]] = np.float32(57)
    return p


@pytest.mark.parametrize("kind", ["density", "gradient", "force"])
def test_symmetric_equals_two_nonsym(orc, kind):
    """runner_iact_X(i, j) == nonsym(i, j) then nonsym(j, i), byte for byte
    (tests/testSymmetry.c:113-167)."""
    rng = np.random.default_rng({"density": 1, "gradient": 2, "force": 3}[kind])
    for _ in range(200):
        pi = _rand_part(rng, orc)
        pj = _rand_part(rng, orc)
        hi = np.float32(rng.uniform(0.2, 0.3))
        hj = np.float32(rng.uniform(0.2, 0.3))
        dx = rng.uniform(-0.15, 0.15, 3).astype(np.float32)
        r2 = np.float32(np.fma(dx[2], dx[2], np.fma(dx[1], dx[1], dx[0] * dx[0]))
                        if hasattr(np, "fma") else float(np.dot(dx, dx)))
        a_i, a_j = pi.copy(), pj.copy()
        orc.iact(kind, r2, dx, hi, hj, a_i, a_j, True)
        b_i, b_j = pi.copy(), pj.copy()
        orc.iact(kind, r2, dx, hi, hj, b_i, b_j, False)
        orc.iact(kind, r2, -dx, hj, hi, b_j, b_i, False)
        assert a_i.tobytes() == b_i.tobytes()
        assert a_j.tobytes() == b_j.tobytes()


@pytest.mark.parametrize("seed", [0, 1])
def test_counts_bruteforce_vs_cells(orc, seed):
    """Cell-based search (27-cell stencil with image translations) against an
    O(N^2) loop with the minimal-image rule, h up to the cell width."""
    rng = np.random.default_rng(seed)
    st = ics.lattice_box(9, seed=seed, jitter=0.3)
    n = st["h"].shape[0]
    w = 1.0 / 3.0
    st["h"] = rng.uniform(0.3, 1.0, n).astype(np.float32) * np.float32(w / GAMMA)
    props = HydroProps(ncell=(3, 3, 3))
    ref = orc.pass_fp64(props, st)
    for mode, key in ((0, "n_density"), (1, "n_force")):
        bf = orc.count_bruteforce(props, st, mode)
        assert np.array_equal(bf, ref[key])
    assert ref["n_force"].min() >= ref["n_density"].min()


def _interior(x, margin):
    return np.all((x > margin) & (x < 1 - margin), axis=1)


def _lattice_sums(n, h):
    """Independent FP64 lattice sums of the Wendland C2 kernel around one
    particle of a perfect n^3 lattice in the unit box (spacing s = 1/n)."""
    s = 1.0 / n
    g = np.arange(-4, 5)
    d = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * s
    r = np.sqrt((d ** 2).sum(1))
    q = r / (GAMMA * h)
    inside = q < 1
    d, r, q = d[inside], r[inside], q[inside]
    w = (1 - q) ** 4 * (1 + 4 * q) * K / GAMMA ** 3
    dw = -20 * q * (1 - q) ** 3 * K / GAMMA ** 4
    m = 1.0 / n ** 3
    rho = m * w.sum() / h ** 3
    nz = r > 0
    div = -(m / rho) * np.sum(dw[nz] * r[nz]) / h ** 4
    rotz = (m / rho) * np.sum(dw[nz] / r[nz] * (d[nz, 0] ** 2 + d[nz, 1] ** 2)) / h ** 4
    return rho, div, rotz


def test_analytic_density_fields(orc):
    """test27cells analytic targets on a perfect lattice: the oracle's sums
    equal independent FP64 lattice sums to round-off, and those are the SPH
    estimates of rho = 1, div v = 3 (v = x - c) and curl v = (0, 0, -2)
    (v = (y, -x, 0)) within the kernel's discretisation error."""
    n = 24
    st = ics.lattice_box(n, seed=0, jitter=0.0)
    props = HydroProps(ncell=(6, 6, 6))
    rho_l, div_l, rotz_l = _lattice_sums(n, float(st["h"][0]))
    assert abs(rho_l - 1) < 0.03 and abs(div_l - 3) < 0.1 and abs(rotz_l + 2) < 0.07
    ref = orc.pass_fp64(props, st)
    # (the oracle uses the reference's float-rounded constants: ~1e-7)
    assert np.allclose(ref["rho"], rho_l, rtol=1e-6)
    x = st["x"]
    inner = _interior(x, 0.2)
    st["v"] = (x - 0.5).astype(np.float32)
    ref = orc.pass_fp64(props, st)
    assert np.allclose(ref["div_v"][inner], div_l, rtol=1e-6)
    assert np.allclose(ref["rot_v"][inner], 0.0, atol=1e-6)
    st["v"] = np.stack([x[:, 1], -x[:, 0], 0 * x[:, 0]], 1).astype(np.float32)
    ref = orc.pass_fp64(props, st)
    assert np.allclose(ref["div_v"][inner], 0.0, atol=1e-6)
    assert np.allclose(ref["rot_v"][inner, 2], rotz_l, rtol=1e-6)
    assert np.allclose(ref["rot_v"][inner, :2], 0.0, atol=1e-6)


def test_analytic_force_fields(orc):
    """test125cells get_solution (tests/test125cells.c:148-207): for
    P = 1.5 x and v = x - c: a = -grad P / rho, du/dt = -(P/rho) div v,
    h_dt = h div v / 3, v_sig = 2c + 3 |mu| (converging pairs only)."""
    n = 24
    st = ics.lattice_box(n, seed=0, jitter=0.0)
    props = HydroProps(ncell=(6, 6, 6))
    x = st["x"]
    rho0 = 1.0
    P = 1.5 * x[:, 0] + 1.0
    st["u"] = (P / (2.0 / 3.0 * rho0)).astype(np.float32)
    st["v"] = (x - 0.5).astype(np.float32)
    ref = orc.pass_fp64(props, st)
    inner = _interior(x, 0.25)
    rho = ref["rho"][inner]
    a = ref["a_hydro"][inner]
    # SPH estimates of the analytic values, within the discretisation error
    # of the kernel on this lattice (a few per cent).
    assert np.allclose(a[:, 0], -1.5 / rho, rtol=4e-2)
    assert np.allclose(a[:, 1:], 0.0, atol=1e-8)
    Pi = ref["pressure"][inner]
    assert np.allclose(ref["u_dt"][inner], -(Pi / rho) * 3.0, rtol=4e-2)
    h = st["h"][inner]
    assert np.allclose(ref["h_dt"][inner], h * 3.0 / 3.0, rtol=4e-2)
    # diverging flow (mu = 0): v_sig = c_i + max_j c_j over neighbours, at
    # least the 2c of hydro_reset_gradient and within the pressure spread
    c = ref["soundspeed"][inner]
    assert np.all(ref["v_sig"][inner] >= 2 * c * (1 - 1e-12))
    assert np.all(ref["v_sig"][inner] <= 2 * c * 1.05)


def test_ghost_stopping_rule_fp64(orc):
    """After the FP32 ghost, one Newton step from the final h, evaluated in
    FP64, moves h by at most ~h_tolerance (runner_ghost.c:1250-1388)."""
    st = ics.modulated_box(16, seed=3, amp=0.5)
    props = HydroProps(ncell=(4, 4, 4))
    o32 = orc.pass_fp32(props, st)
    assert o32["ghost_iters"][0] >= 2
    ref = orc.pass_fp64(props, st)
    h = st["h"].astype(np.float64)
    eta3 = np.float32(props.eta) ** 3
    f = ref["wcount"] * h ** 3 - eta3
    fp = ref["wcount_dh"] * h ** 3 + 3 * ref["wcount"] * h ** 2
    step = np.abs(f / fp) / h
    assert step.max() <= 1.2 * props.h_tolerance, step.max()


def test_ghost_against_bisection(orc):
    """Independent FP64 bisection for h (wcount(h) h^3 = eta^3) on a few
    particles of a tiny box agrees with the FP32 ghost within 2 tolerances."""
    st = ics.modulated_box(12, seed=9, amp=0.5)
    props = HydroProps(ncell=(3, 3, 3))
    st32 = ics.copy_state(st)
    orc.pass_fp32(props, st32)
    eta3 = float(np.float32(props.eta) ** 3)
    x = st["x"]
    m = st["m"]
    for i in (0, 17, 400, 1000, 1727):
        d = x - x[i]
        d -= np.round(d)
        r = np.sqrt((d ** 2).sum(axis=1))

        def nsum(h):
            u = r / h
            q = u / GAMMA
            w = np.where(q < 1, (1 - q) ** 4 * (1 + 4 * q), 0.0) * K / GAMMA ** 3
            return w.sum() - eta3  # includes the self term (r = 0)

        lo, hi = 0.3 / 12, 3.0 / 12
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            if nsum(mid) < 0:
                lo = mid
            else:
                hi = mid
        assert abs(st32["h"][i] / lo - 1) <= 2 * props.h_tolerance
    del m


def test_magnitude_scales_bound_values(orc):
    st = ics.modulated_box(12, seed=5, amp=0.5)
    props = HydroProps(ncell=(3, 3, 3))
    orc.pass_fp32(props, st)
    ref = orc.pass_fp64(props, st)
    for k in ("div_v", "rot_v", "rho_dh", "wcount_dh", "laplace_u",
              "a_hydro", "u_dt", "h_dt"):
        assert np.all(ref["s_" + k] * (1 + 1e-12) >= np.abs(ref[k])), k


def test_fp32_restatement_close_to_fp64(orc):
    """The FP32 restatement (Horner kernel, reference operation order) stays
    within FP32 rounding of the FP64 rerun: this is the error budget a
    faithful FP32 implementation of the reference has."""
    from parity import DENSITY_FIELDS, rel_err
    st = ics.modulated_box(16, seed=2, amp=0.5)
    props = HydroProps(ncell=(4, 4, 4))
    o32 = orc.pass_fp32(props, st)
    ref = orc.pass_fp64(props, st)
    assert np.array_equal(o32["n_density"], ref["n_density"])
    assert np.array_equal(o32["n_force"], ref["n_force"])
    for k, (tol, sk) in DENSITY_FIELDS.items():
        assert rel_err(o32[k], ref[k], ref[sk] if sk else None) <= tol, k
    e = rel_err(o32["a_hydro"], ref["a_hydro"], ref["s_a_hydro"])
    assert e <= 1e-4


@pytest.mark.parametrize("nc,dim", [(3, 1.0), (8, 1.0), (96, 1.0), (5, 1e3)])
def test_pair_contract_is_the_reference_expression(orc, nc, dim):
    """The oracle's pair frames (the getsid rule on cell corners,
    space_getsid.h:47-80) and contract give, bit for bit, the expression of
    the reference's brute-force checkers restated independently in numpy
    from the stencil offset (tests/parity.py contract_dx), for pairs in
    every one of the 27 neighbour positions, periodic images included."""
    from parity import contract_dx
    rng = np.random.default_rng(nc)
    n = 3000
    w = dim / nc
    ci = rng.integers(0, nc, (n, 3))
    o = rng.integers(-1, 2, (n, 3))
    v = ci + o
    cj = np.mod(v, nc)
    t = np.where(v < 0, -dim, np.where(v >= nc, dim, 0.0))
    xi = (ci + rng.uniform(0, 1, (n, 3))) * w
    xj = (cj + rng.uniform(0, 1, (n, 3))) * w
    # particles on cell faces too
    face = rng.uniform(size=n) < 0.2
    xj[face] = (cj[face] + rng.integers(0, 2, (face.sum(), 3)) * 0.999999) * w
    want = contract_dx(xi, xj, ci, cj, w, t)
    for k in range(n):
        fr = orc.pair_frame(ci[k] * w, cj[k] * w, (dim,) * 3)
        assert fr.self == int(np.all(o[k] == 0))
        dx, r2 = orc.pair_dx32(xi[k], xj[k], fr)
        assert dx.tobytes() == want[k].tobytes(), (k, o[k], dx, want[k])


def test_contract_flips_are_rare(orc):
    """The reference's expression and a single rounding of the FP64
    difference, (float)((x_i - t) - x_j), disagree on a handful of mask
    decisions at r ~ H: count them on the 32^3 lattice (8^3 cells) at its
    converged h.  The neighbour sets are bit-exact only under the contract
    the oracle and the GPU share."""
    st = ics.lattice_box(32, seed=0)
    props = HydroProps(ncell=(8, 8, 8))
    orc.pass_fp32(props, st)
    f = orc.contract_flips(props, st)
    assert f["both"] > 1_800_000
    assert f["ref_only"] + f["fp64_only"] <= 1e-5 * f["both"]
