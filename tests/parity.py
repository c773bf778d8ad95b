"""Comparison helpers shared by the parity tests.

Error measure (the tolerance stated by the task: 1e-5 for density
quantities, 1e-4 for force quantities, against an FP64 rerun of the same
pass at the same smoothing lengths):

    err_i = |x_i - x64_i| / max(|x64_i|, S_i)

where S_i is the FP64 magnitude sum sum_j |term_ij| of the accumulation
(0 for quantities that are not sums).  For sums without cancellation
(rho, wcount) S_i = |x64_i| and this is the plain relative error; for
cancelling sums (div_v, curl, a_hydro, u_dt, ...) it is the error relative
to the size of the contributions, the scale FP32 summation can resolve.
Neighbour counts are compared exactly.
"""
from __future__ import annotations

import numpy as np

DENSITY_TOL = 1e-5
FORCE_TOL = 1e-4

# field -> (tolerance, scale key in the FP64 oracle output or None)
DENSITY_FIELDS = {
    "rho": (DENSITY_TOL, None),
    "wcount": (DENSITY_TOL, None),
    "div_v": (DENSITY_TOL, "s_div_v"),
    "rot_v": (DENSITY_TOL, "s_rot_v"),
    "rho_dh": (DENSITY_TOL, "s_rho_dh"),
    "wcount_dh": (DENSITY_TOL, "s_wcount_dh"),
}
FORCE_FIELDS = {
    "pressure": (FORCE_TOL, None),
    "soundspeed": (FORCE_TOL, None),
    "v_sig": (FORCE_TOL, None),
    "laplace_u": (FORCE_TOL, "s_laplace_u"),
    "a_hydro": (FORCE_TOL, "s_a_hydro"),
    "u_dt": (FORCE_TOL, "s_u_dt"),
    "h_dt": (FORCE_TOL, "s_h_dt"),
    "alpha_visc": (FORCE_TOL, None),
    "alpha_diff": (FORCE_TOL, None),
}
COUNT_FIELDS = ["n_density", "n_gradient", "n_force", "min_ngb_time_bin"]


def rel_err(x, ref, scale=None, mask=None) -> float:
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.abs(ref)
    if scale is not None:
        den = np.maximum(den, np.asarray(scale, np.float64))
    den = np.where(den > 0, den, 1.0)
    e = np.abs(x - ref) / den
    if mask is not None:
        e = e[mask] if e.ndim == 1 else e[mask, :]
    return float(e.max()) if e.size else 0.0


def compare(res: dict, ref: dict, active: np.ndarray, fields=None) -> dict:
    """Max error per field over active particles; counts as mismatch counts."""
    out = {}
    fields = fields or {**DENSITY_FIELDS, **FORCE_FIELDS}
    for k, (tol, sk) in fields.items():
        out[k] = rel_err(res[k], ref[k], ref[sk] if sk else None, active)
    for k in COUNT_FIELDS:
        out[k] = int(np.count_nonzero(res[k][active] != ref[k][active]))
    return out


def assert_parity(res: dict, ref: dict, active: np.ndarray, fields=None,
                  tol_scale: float = 1.0) -> dict:
    fields = fields or {**DENSITY_FIELDS, **FORCE_FIELDS}
    errs = compare(res, ref, active, fields)
    bad = []
    for k, (tol, _) in fields.items():
        if not errs[k] <= tol * tol_scale:
            bad.append(f"{k}: {errs[k]:.3e} > {tol * tol_scale:.1e}")
    for k in COUNT_FIELDS:
        if errs[k] != 0:
            bad.append(f"{k}: {errs[k]} mismatching particles")
    assert not bad, "; ".join(bad) + f"  (all: {errs})"
    return errs


# The viscosity and diffusion switches (hydro_prepare_force,
# src/hydro/SPHENIX/hydro.h:840-954) divide cancelling sums by the step
# dt_alpha: alpha_loc = alpha_max S / (c^2 + S) with
# S = (gamma_K h)^2 max(0, -(div_v - div_v_prev) / dt_alpha), and
# alpha_diff += (beta gamma_K h laplace_u / sqrt(u) - ...) dt_alpha.  Their
# error is the error of div_v and laplace_u (within the tolerances above,
# relative to the sums' magnitudes) carried through those expressions; the
# update itself has a sensitivity <= 1 to alpha_loc in both of its branches.
KERNEL_GAMMA = 1.936492


def alpha_bounds(ref: dict, pre: dict, props, time_base: float):
    """Per-particle bounds on |alpha - alpha64| for alpha_visc and
    alpha_diff, given the FP64 rerun `ref` and the pass inputs `pre`."""
    tb = pre["time_bin"].astype(np.int64)
    dt_a = np.where(tb > 0, np.ldexp(1.0, np.clip(tb, 0, 60) + 1) * time_base, 0.0)
    ksp = pre["h"].astype(np.float64) * KERNEL_GAMMA
    c2 = np.asarray(ref["soundspeed"], np.float64) ** 2
    ddv = DENSITY_TOL * np.maximum(np.abs(ref["div_v"]), ref["s_div_v"])
    with np.errstate(divide="ignore", invalid="ignore"):
        b_visc = np.where(dt_a > 0,
                          props.alpha_max * ksp ** 2 * ddv / (dt_a * c2), 0.0)
    dl = FORCE_TOL * np.maximum(np.abs(ref["laplace_u"]), ref["s_laplace_u"])
    b_diff = (props.diff_beta * ksp * dl / np.sqrt(pre["u"].astype(np.float64))
              + FORCE_TOL * np.abs(pre["alpha_diff"] - props.diff_alpha_min)
              * np.asarray(ref["v_sig"], np.float64) / ksp) * dt_a
    b_visc = b_visc + FORCE_TOL * np.abs(ref["alpha_visc"]) + 1e-7
    b_diff = b_diff + FORCE_TOL * np.abs(ref["alpha_diff"]) + 1e-7
    return b_visc, b_diff


def assert_alpha(res: dict, ref: dict, pre: dict, props, time_base: float,
                 active: np.ndarray) -> None:
    bv, bd = alpha_bounds(ref, pre, props, time_base)
    for k, b in (("alpha_visc", bv), ("alpha_diff", bd)):
        e = np.abs(np.asarray(res[k], np.float64) - ref[k])
        bad = active & (e > b)
        assert not bad.any(), (k, int(bad.sum()), float((e / b)[active].max()))


def contract_dx(xi, xj, ci, cj, w, t=None):
    """The reference's pair expression (tools.c:218-246, 617-619;
    runner_doiact_functions_hydro.h:1327-1346, 2459) for i in cell ci and j
    in the adjacent cell cj (integer cell coordinates, j's image translated
    by t): self pairs round the FP64 difference once; otherwise both
    particles are taken relative to the corner of the upper cell (the one
    whose offset from the other has its first non-zero component positive),
    with the translation on the lower cell's side."""
    if t is None:
        t = np.zeros_like(xi)
    o = cj - ci
    o = np.where(o > 1, -1, np.where(o < -1, 1, o))  # wrapped offset
    first = np.where(o[:, 0] != 0, o[:, 0], np.where(o[:, 1] != 0, o[:, 1], o[:, 2]))
    self_ = np.all(o == 0, axis=1)
    up_j = first > 0
    li, lj = ci * w, cj * w
    oi = np.where(up_j[:, None], lj + t, li)
    oj = np.where(up_j[:, None], lj, li + (-t))
    pair = (xi - oi).astype(np.float32) - (xj - oj).astype(np.float32)
    return np.where(self_[:, None], (xi - xj).astype(np.float32), pair).astype(np.float32)
