"""Pins of the oracle's viscosity, diffusion, gradient-loop and ghost terms
(runs without a GPU).

Each check ties a part of the oracle to something fixed outside its own
code: conservation laws the SPHENIX equations of motion satisfy exactly
(momentum and total energy, with artificial viscosity and conduction on),
the direction of heat flow, closed forms on a perfect lattice (Brookshaw's
1985 SPH Laplacian estimate of del^2 (x^2) = 2, the signal velocity of a
uniform compression), hand-derived values for single pairs
(src/hydro/SPHENIX/hydro_iact.h:211-344, 358-632), the viscosity and
diffusion switches of hydro_prepare_force (hydro.h:840-954) evaluated from
the pass's own, separately pinned, div v, del^2 u and v_sig, and the cube
root of the ghost's bisection (dimension.h:98-102) against FP64 cbrt.

Every check is also run on builds of the oracle with one planted mistake
each (oracle/Makefile `mutants`, sphenix_oracle.c MUT(k, ...)): the test
``test_check_bites`` asserts that the check fails there, so none of them is
vacuous.
"""
from __future__ import annotations

import math
import os
import subprocess

import numpy as np
import pytest

from swift_amd import HydroProps, ics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GAMMA_K = 1.936492           # kernel_gamma, Wendland C2 3D
K = 21.0 / (2.0 * math.pi)   # kernel_constant
GAS_GAMMA = 5.0 / 3.0
GAMMA_KF = float(np.float32(GAMMA_K))  # the FP32 constant the oracle uses


def _kernel_dw(u, h=1.0):
    """dW/du of the Wendland C2 kernel, closed form, for u = r / h."""
    q = u / GAMMA_K
    return np.where(q < 1, -20.0 * q * (1 - q) ** 3, 0.0) * K / GAMMA_K ** 4


# ------------------------------------------------------------------ checks --
# Each returns (ok, detail).

def check_conservation(orc):
    """Sum_i m_i a_i = 0 and sum_i m_i (v_i . a_i + du_i/dt) = 0 for an
    all-active box, with viscosity (alpha in [0.5, 1], random converging
    flows) and conduction (alpha_diff in [0.1, 0.4], varying u) on: the pair
    terms of runner_iact_force are antisymmetric (hydro_iact.h:358-493)."""
    st = ics.modulated_box(12, seed=4, amp=0.5, sigma_v=0.5)
    props = HydroProps(ncell=(3, 3, 3))
    orc.pass_fp32(props, st)  # converged h
    rng = np.random.default_rng(8)
    n = st["h"].shape[0]
    st["alpha_visc"] = rng.uniform(0.5, 1.0, n).astype(np.float32)
    st["alpha_diff"] = rng.uniform(0.1, 0.4, n).astype(np.float32)
    ref = orc.pass_fp64(props, st)
    m = st["m"].astype(np.float64)[:, None]
    a = ref["a_hydro"]
    mom = np.abs((m * a).sum(0)).max() / (m * ref["s_a_hydro"]).sum(0).max()
    v = st["v"].astype(np.float64)
    e = (m[:, 0] * ((v * a).sum(1) + ref["u_dt"])).sum()
    es = (m[:, 0] * ((np.abs(v) * ref["s_a_hydro"]).sum(1) + ref["s_u_dt"])).sum()
    ok = mom < 1e-12 and abs(e) / es < 1e-12
    return ok, dict(momentum=mom, energy=abs(e) / es)


def _lattice(n, ncell):
    st = ics.lattice_box(n, seed=0, jitter=0.0)
    return st, HydroProps(ncell=(ncell,) * 3)


def check_diffusion_direction(orc):
    """At rest (v = 0) only the conduction term changes u; heat flows from
    hot to cold: du/dt has the opposite sign of u - <u> across a sinusoidal
    temperature profile (hydro_iact.h:607-626)."""
    st, props = _lattice(16, 4)
    x = st["x"][:, 0]
    s = np.sin(2 * np.pi * x)
    st["u"] = (1.0 + 0.5 * s).astype(np.float32)
    st["alpha_diff"][:] = 0.5
    st["alpha_visc"][:] = 0.0
    ref = orc.pass_fp64(props, st)
    sel = np.abs(s) > 0.3
    ok = bool(np.all(np.sign(ref["u_dt"][sel]) == -np.sign(s[sel])))
    return ok, dict(agree=float(np.mean(np.sign(ref["u_dt"][sel]) == -np.sign(s[sel]))))


def _brookshaw_lattice(n, h, fn):
    """Brookshaw's (1985) SPH estimate 2 sum_j (m_j / rho_j) (A_i - A_j)
    W'(r_ij) / r_ij of del^2 A at the origin of a perfect n^3 lattice of unit
    box mass, for A = fn(offset), in FP64 with the closed-form kernel (rho
    is the lattice's own SPH density, self term included)."""
    g = np.arange(-5, 6)
    d = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) / n
    r = np.sqrt((d ** 2).sum(1))
    q = r / (GAMMA_K * h)
    m = 1.0 / n ** 3
    w = np.where(q < 1, (1 - q) ** 4 * (1 + 4 * q), 0.0) * K / GAMMA_K ** 3
    rho = m * w.sum() / h ** 3
    nz = (r > 0) & (q < 1)
    d, r = d[nz], r[nz]
    dWdr = _kernel_dw(r / h) / h ** 4
    return 2.0 * np.sum(m / rho * (fn(np.zeros((3, 1))) - fn(d.T)) * dWdr / r)


def check_laplace_u(orc):
    """del^2 u for u = 1 + x^2 on a perfect lattice: the oracle's laplace_u
    equals Brookshaw's estimate (an independent FP64 lattice sum) to
    round-off, and that estimate is the analytic 2 within the kernel's
    discretisation error (hydro_iact.h:246-250, hydro.h:762-780)."""
    n = 24
    st, props = _lattice(n, 6)
    x = st["x"][:, 0]
    st["u"] = (1.0 + x * x).astype(np.float32)
    ref = orc.pass_fp64(props, st)
    inner = (x > 0.25) & (x < 0.75)
    h = float(st["h"][0])
    est = _brookshaw_lattice(n, h, lambda d: d[0] ** 2)
    lap = ref["laplace_u"][inner]
    # (u is stored in FP32: its rounding limits the agreement to ~1e-5 of
    # the differences u_i - u_j)
    ok = abs(est - 2.0) < 0.08 and np.allclose(lap, est, rtol=1e-4)
    return ok, dict(est=est, lap_min=lap.min(), lap_max=lap.max())


def check_v_sig(orc):
    """Uniform compression v = -lambda (x - c) on a perfect lattice: every
    pair closes at mu_ij = -lambda r_ij, so the gradient loop's signal
    velocity is v_sig = 2c + 3 lambda r_max, r_max the farthest lattice
    neighbour inside the kernel (hydro_iact.h:229-237, beta = 3 of
    hydro_parameters.h:53)."""
    n = 24
    st, props = _lattice(n, 6)
    lam = 0.5
    x = st["x"]
    st["v"] = (-lam * (x - 0.5)).astype(np.float32)
    ref = orc.pass_fp64(props, st)
    inner = np.all((x > 0.25) & (x < 0.75), axis=1)
    h = float(st["h"][0])
    g = np.arange(-4, 5)
    d = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) / n
    r = np.sqrt((d ** 2).sum(1))
    rmax = r[(r > 0) & (r < GAMMA_K * h)].max()
    c = ref["soundspeed"][inner]
    want = 2 * c + 3 * lam * rmax
    err = np.max(np.abs(ref["v_sig"][inner] / want - 1))
    return err < 1e-6, dict(err=err)


def _two_passes(orc, st, props, prev_shift, time_base, tb=1):
    """A pass at time_base 0 for div v, then the pass under test with
    div_v_previous_step = div v + prev_shift and the time line set."""
    ref0 = orc.pass_fp64(props, ics.copy_state(st))
    st["div_v_prev"] = (ref0["div_v"] + prev_shift).astype(np.float32)
    st["time_bin"][:] = tb
    props.time_base = time_base
    ref = orc.pass_fp64(props, st)
    dt = float(np.float32((2 ** (tb + 1)) * time_base))
    return ref, dt


def check_alpha_switch(orc):
    """The viscosity switch of hydro_prepare_force (hydro.h:856-905, the
    Cullen & Dehnen form of SPHENIX): in a compression whose div v keeps
    falling, alpha jumps to alpha_loc = alpha_max S / (c^2 + S) with
    S = H^2 max(0, -d(div v)/dt), H = gamma_K h the kernel support; in an
    expansion (S = 0) it decays as alpha / (1 + dt c ell / H)."""
    errs = []
    # (a) converging, div v falling
    st, props = _lattice(16, 4)
    x = st["x"]
    st["v"] = (-1.0 * (x - 0.5)).astype(np.float32)
    st["alpha_visc"][:] = 0.01
    ref, dt = _two_passes(orc, st, props, 0.01, 1e-3)
    # inside the compression (the flow diverges across the periodic faces)
    inner = np.all((x > 0.25) & (x < 0.75), axis=1)
    assert np.all(ref["div_v"][inner] < 0)
    H = GAMMA_KF * st["h"].astype(np.float64)
    dvdt = (ref["div_v"] - st["div_v_prev"]) / dt
    S = H * H * np.maximum(0.0, -dvdt)
    c = ref["soundspeed"]
    want = props.alpha_max * S / (c * c + S)
    assert np.all(want[inner] > 0.01)
    errs.append(np.max(np.abs(ref["alpha_visc"][inner] / want[inner] - 1)))
    # (b) expanding: decay towards alpha_loc = 0
    st, props = _lattice(16, 4)
    st["v"] = (1.0 * (st["x"] - 0.5)).astype(np.float32)
    st["alpha_visc"][:] = 1.0
    ref, dt = _two_passes(orc, st, props, 0.0, 1e-3)
    inner = np.all((st["x"] > 0.25) & (st["x"] < 0.75), axis=1)
    assert np.all(ref["div_v"][inner] > 0)
    H = GAMMA_KF * st["h"].astype(np.float64)
    tau = dt * ref["soundspeed"] / H * float(np.float32(props.visc_length))
    want = 1.0 / (1.0 + tau)
    errs.append(np.max(np.abs(ref["alpha_visc"][inner] / want[inner] - 1)))
    err = max(errs)
    return err < 1e-9, dict(err=err)


def check_alpha_diff(orc):
    """The conduction switch of hydro_prepare_force (hydro.h:907-953):
    alpha_diff grows at beta_diff H del^2 u / sqrt(u) and decays at
    (alpha_diff - alpha_diff_min) v_sig / H, bounded by
    alpha_diff_max (1 - alpha_visc_max_ngb / alpha_max), from the pass's own
    del^2 u and v_sig."""
    st, props = _lattice(16, 4)
    x = st["x"][:, 0]
    st["u"] = (1.0 + x * x).astype(np.float32)
    st["alpha_diff"][:] = 0.1
    st["alpha_visc"][:] = 0.0
    ref, dt = _two_passes(orc, st, props, 0.0, 1e-3)
    H = GAMMA_KF * st["h"].astype(np.float64)
    u = st["u"].astype(np.float64)
    rate = props.diff_beta * H * ref["laplace_u"] / np.sqrt(u)
    a0 = float(np.float32(0.1))  # the stored FP32 alpha_diff
    rate -= (a0 - props.diff_alpha_min) * ref["v_sig"] / H
    want = np.clip(a0 + rate * dt, props.diff_alpha_min,
                   props.diff_alpha_max * (1 - ref["alpha_visc_max_ngb"] / props.alpha_max))
    # away from the jump of u across the periodic boundary in x
    inner = (x > 0.25) & (x < 0.75)
    assert np.all(want[inner] > 0.05)
    err = np.max(np.abs(ref["alpha_diff"][inner] / want[inner] - 1))
    return err < 1e-9, dict(err=err)


def _pair_record(orc, **kv):
    F = {k: i for i, k in enumerate(orc.PART_FIELDS)}
    p = np.zeros(len(F), np.float32)
    p[F["M"]] = 1.0
    p[F["RHO"]] = 1.0
    p[F["MINTB"]] = 57.0
    for k, v in kv.items():
        p[F[k]] = v
    return p, F


def check_pair_closed_forms(orc):
    """Single pairs of the FP32 interaction functions against values derived
    by hand from the SPHENIX equations (hydro_iact.h), unit masses,
    densities and smoothing lengths, f = 0, at separation dx = (0.5, 0, 0):
    (a) head-on gradient pair, v_i = -1, v_j = +1 (mu = -2), c_i = 1,
        c_j = 1.5: v_sig = c_i + c_j - 3 mu = 8.5; receding: 2.5.
    (b) head-on force pair with alpha = 1, Balsara = 1, c = 1, P ~ 0:
        visc = -1/4 (2)(8)(-2)(2)/(2) = 8, a_i = -8 W'(0.5) along x,
        du_i/dt = -8 W'(0.5) (viscous heating).
    (c) conduction at rest: P_i = 2, P_j = 1, alpha_diff = 0.5, u_i = 3,
        u_j = 1: v_diff = 0.5 * 0.5 * sqrt(2 * 1 / 2) = 0.25 and
        du_i/dt = 0.25 * 2 * 2 W'(0.5) = W'(0.5)."""
    dx = np.array([0.5, 0.0, 0.0], np.float32)
    r2 = np.float32(0.25)
    w1 = _kernel_dw(0.5)
    errs = []
    for vi, want in ((-1.0, 8.5), (1.0, 2.5)):
        pi, F = _pair_record(orc, V0=vi, C=1.0)
        pj, _ = _pair_record(orc, V0=-vi, C=1.5)
        orc.iact("gradient", r2, dx, 1.0, 1.0, pi, pj, False)
        errs.append(abs(pi[F["VSIG"]] / want - 1))
    pi, F = _pair_record(orc, V0=-1.0, C=1.0, P=1e-20, ALPHA=1.0, BAL=1.0, U=1.0)
    pj, _ = _pair_record(orc, V0=1.0, C=1.0, P=1e-20, ALPHA=1.0, BAL=1.0, U=1.0)
    orc.iact("force", r2, dx, 1.0, 1.0, pi, pj, False)
    errs.append(abs(pi[F["A0"]] / (-8 * w1) - 1))
    errs.append(abs(pi[F["UDT"]] / (-8 * w1) - 1))
    errs.append(abs(pi[F["A1"]]) + abs(pi[F["A2"]]))
    pi, F = _pair_record(orc, C=1.0, P=2.0, ADIFF=0.5, U=3.0)
    pj, _ = _pair_record(orc, C=1.0, P=1.0, ADIFF=0.5, U=1.0)
    orc.iact("force", r2, dx, 1.0, 1.0, pi, pj, False)
    errs.append(abs(pi[F["UDT"]] / w1 - 1))
    err = max(errs)
    return err < 1e-5, dict(err=err)


def check_fp32_matches_fp64(orc):
    """The FP32 restatement (whose force loop is the pair function pinned
    above) and the FP64 rerun agree on every gradient and force output at
    the GPU tolerance, with viscosity and conduction on: the FP64 checker
    computes the same physics as the pinned pair functions."""
    from parity import FORCE_FIELDS, rel_err
    st = ics.modulated_box(16, seed=2, amp=0.5, sigma_v=0.5)
    st["alpha_visc"][:] = 1.0
    st["alpha_diff"][:] = 0.3
    props = HydroProps(ncell=(4, 4, 4))
    o32 = orc.pass_fp32(props, st)
    ref = orc.pass_fp64(props, st)
    errs = {}
    for k in ("v_sig", "laplace_u", "a_hydro", "u_dt", "h_dt"):
        tol, sk = FORCE_FIELDS[k]
        errs[k] = rel_err(o32[k], ref[k], ref[sk] if sk else None)
    ok = all(e <= 1e-4 for e in errs.values())
    return ok, errs


def check_cbrt(orc):
    """The cube root of the ghost's bisection step (runner_ghost.c:1397,
    pow_inv_dimension, dimension.h:98-102) is within 1 ulp of the correctly
    rounded FP64 cube root over 12 decades and at exact cubes."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([10 ** rng.uniform(-6, 6, 20000),
                         np.arange(1, 200, dtype=np.float64) ** 3]).astype(np.float32)
    L = orc.lib()
    worst = 0.0
    for x in xs:
        got = np.float32(L.orc_cbrt32(float(x)))
        exact = np.cbrt(np.float64(x))
        ulp = np.spacing(np.float32(exact))
        worst = max(worst, abs(float(got) - exact) / float(ulp))
    return worst <= 1.0, dict(worst_ulp=worst)


CHECKS = {
    "conservation": check_conservation,
    "diffusion_direction": check_diffusion_direction,
    "laplace_u": check_laplace_u,
    "v_sig": check_v_sig,
    "alpha_switch": check_alpha_switch,
    "alpha_diff": check_alpha_diff,
    "pair_closed_forms": check_pair_closed_forms,
    "fp32_matches_fp64": check_fp32_matches_fp64,
    "cbrt": check_cbrt,
}

# planted mistake (sphenix_oracle.c MUT(k, ...)) -> the check that must fail
MUTANTS = {
    1: "conservation",          # viscous du/dt with 1 instead of 1/2
    2: "conservation",          # viscosity over 2 rho_i instead of rho_i + rho_j
    3: "conservation",          # pressure-weighted alpha_diff uses alpha_i twice
    15: "conservation",         # SPH du/dt with W'_j instead of W'_i
    4: "diffusion_direction",   # conduction with (u_j - u_i)
    5: "laplace_u",             # end_gradient factor 1 instead of 2
    6: "laplace_u",             # laplace_u term without 1 / r
    7: "v_sig",                 # signal velocity with beta = 2
    8: "v_sig",                 # mu from receding instead of closing pairs
    9: "alpha_switch",          # source term with h^2 instead of H^2
    10: "alpha_switch",         # decay time scale without the viscosity length
    11: "alpha_diff",           # 1 / u instead of 1 / sqrt(u)
    12: "pair_closed_forms",    # viscosity coefficient -1/2 instead of -1/4
    13: "pair_closed_forms",    # gradient-loop beta = 2
    14: "pair_closed_forms",    # v_diff with 1 instead of 1/2
    16: "fp32_matches_fp64",    # FP64 force-loop signal velocity with beta = 2
    17: "cbrt",                 # cube root as exp(log(x) / 3)
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_check(orc, name):
    ok, detail = CHECKS[name](orc)
    assert ok, detail


@pytest.fixture(scope="module")
def mutants():
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "oracle"),
                    "mutants"], check=True)
    return os.path.join(ROOT, "oracle", "_mut")


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_check_bites(orc, mutants, k):
    """The check named for planted mistake k fails on that build."""
    with orc.using(os.path.join(mutants, f"liborc_m{k}.so")):
        try:
            ok, detail = CHECKS[MUTANTS[k]](orc)
        except (AssertionError, RuntimeError, FloatingPointError):
            ok, detail = False, "raised"
    assert not ok, (k, MUTANTS[k], detail)
