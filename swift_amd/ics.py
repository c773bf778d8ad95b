"""Synthetic initial conditions for the SPHENIX neighbour pass.

The workloads are named by what they are:

* ``lattice_box``   -- periodic unit cube, n^3 particles on a cubic lattice,
  each coordinate jittered by U(-jitter, +jitter) x spacing; m = 1/N, u = 1,
  v = 0, h0 = eta x spacing (the 32^3 / 8^3-cell case uses n = 32).
* ``modulated_box`` -- the lattice box with x displaced so that
  rho(x) is proportional to 1 + A sin(2 pi x), Gaussian velocities (sigma 0.1
  per component), u ~ U[0.5, 1.5] (the 128^3 / 32^3-cell case uses n = 128).
* ``clustered_box`` -- Plummer spheres on a uniform background with a random
  time-bin activity mask.

Particle state is a dict of numpy arrays with the field names of SWIFT's
SPHENIX ``struct part`` (src/hydro/SPHENIX/hydro_part.h:105-324): positions
are float64 ``x`` (n, 3); everything else is float32, ``time_bin`` is int8.
Defaults follow hydro_convert_quantities (src/hydro/SPHENIX/hydro.h): the
viscosity alpha starts at 0.1 and the diffusion alpha at 0.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

ETA_DEFAULT = 1.2348          # hydro_properties.c:355 (hydro_props_init_no_hydro)
KERNEL_GAMMA = np.float32(1.936492)   # kernel_hydro.h:137 (Wendland C2, 3D)
NUM_TIME_BINS = 56            # timeline.h:36
TIME_BIN_INHIBITED = NUM_TIME_BINS + 2


@dataclasses.dataclass
class HydroProps:
    """Run-time parameters of the pass (SWIFT's hydro_props + engine fields)."""

    dim: tuple = (1.0, 1.0, 1.0)
    ncell: tuple = (8, 8, 8)
    eta: float = ETA_DEFAULT
    h_tolerance: float = 1e-4
    h_max: float = float(np.finfo(np.float32).max)
    h_min: float = 0.0
    max_iter: int = 30
    alpha_max: float = 2.0      # hydro_parameters.h:71
    alpha_min: float = 0.0      # hydro_parameters.h:68
    visc_length: float = 0.05   # hydro_parameters.h:74
    diff_beta: float = 1.0      # hydro_parameters.h:92
    diff_alpha_max: float = 1.0  # hydro_parameters.h:95
    diff_alpha_min: float = 0.0  # hydro_parameters.h:98
    time_base: float = 0.0      # dt_alpha = 2^(bin+1) * time_base
    max_active_bin: int = NUM_TIME_BINS


def _base_state(x: np.ndarray, h0: float, m: float) -> dict:
    n = x.shape[0]
    f32 = np.float32
    st = {
        "x": np.ascontiguousarray(x, dtype=np.float64),
        "v": np.zeros((n, 3), f32),
        "m": np.full(n, m, f32),
        "u": np.ones(n, f32),
        "h": np.full(n, h0, f32),
        "time_bin": np.ones(n, np.int8),
        "rho": np.zeros(n, f32),
        "pressure": np.zeros(n, f32),
        "soundspeed": np.zeros(n, f32),
        "f": np.zeros(n, f32),
        "balsara": np.zeros(n, f32),
        "alpha_visc": np.full(n, 0.1, f32),
        "alpha_diff": np.zeros(n, f32),
        "div_v_prev": np.zeros(n, f32),
    }
    return st


def _wrap(x: np.ndarray, box: float = 1.0) -> np.ndarray:
    x = np.mod(x, box)
    x[x >= box] = 0.0
    return x


def lattice_box(n: int, seed: int = 0, jitter: float = 0.1,
                eta: float = ETA_DEFAULT) -> dict:
    """n^3 jittered lattice in the periodic unit cube."""
    rng = np.random.default_rng(seed)
    g = (np.arange(n, dtype=np.float64) + 0.5)
    ix, iy, iz = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)
    x += rng.uniform(-jitter, jitter, size=x.shape)
    x = _wrap(x / n)
    N = n ** 3
    return _base_state(x, eta / n, 1.0 / N)


def _invert_modulation(q: np.ndarray, amp: float) -> np.ndarray:
    """Solve F(x) = q for F(x) = x + amp/(2 pi) (1 - cos 2 pi x)."""
    x = q.copy()
    tp = 2.0 * math.pi
    for _ in range(50):
        fx = x + amp / tp * (1.0 - np.cos(tp * x)) - q
        dfx = 1.0 + amp * np.sin(tp * x)
        dx = fx / dfx
        x -= dx
        if np.max(np.abs(dx)) < 1e-15:
            break
    return x


def modulated_box(n: int, seed: int = 0, amp: float = 0.5,
                  sigma_v: float = 0.1, jitter: float = 0.1,
                  eta: float = ETA_DEFAULT) -> dict:
    """Lattice box displaced in x so that rho ~ 1 + amp sin(2 pi x)."""
    st = lattice_box(n, seed=seed, jitter=jitter, eta=eta)
    rng = np.random.default_rng(seed + 7919)
    x = st["x"]
    x[:, 0] = _wrap(_invert_modulation(x[:, 0], amp))
    st["v"] = rng.normal(0.0, sigma_v, size=(x.shape[0], 3)).astype(np.float32)
    st["u"] = rng.uniform(0.5, 1.5, size=x.shape[0]).astype(np.float32)
    return st


def clustered_box(N: int, seed: int = 0, n_spheres: int = 1000,
                  frac_clustered: float = 0.7, a_min: float = 0.002,
                  a_max: float = 0.02, trunc: float = 10.0,
                  eta: float = ETA_DEFAULT) -> dict:
    """Plummer spheres (log-uniform scale radii) on a uniform background.

    Member counts are Poisson draws renormalised to frac_clustered * N; radii
    are drawn from the Plummer cumulative mass profile truncated at
    trunc * a; u is log-uniform in [0.1, 100].  h0 is a per-particle estimate
    from the local sphere density (the ghost converges it)."""
    rng = np.random.default_rng(seed)
    n_cl = int(round(frac_clustered * N))
    lam = rng.poisson(n_cl / n_spheres, size=n_spheres).astype(np.float64) + 1.0
    counts = np.floor(lam / lam.sum() * n_cl).astype(np.int64)
    counts[: n_cl - counts.sum()] += 1
    centres = rng.uniform(0.0, 1.0, size=(n_spheres, 3))
    scale = np.exp(rng.uniform(math.log(a_min), math.log(a_max), n_spheres))
    sid = np.repeat(np.arange(n_spheres), counts)
    # Plummer: M(<r)/M = r^3/(r^2+a^2)^(3/2); truncate the mass fraction.
    mt = trunc ** 3 / (trunc ** 2 + 1.0) ** 1.5
    u = rng.uniform(0.0, mt, size=n_cl)
    r = scale[sid] / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    d = rng.normal(size=(n_cl, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    xc = centres[sid] + r[:, None] * d
    xb = rng.uniform(0.0, 1.0, size=(N - n_cl, 3))
    x = _wrap(np.concatenate([xc, xb], axis=0))
    # Local density estimate: Plummer density of the host sphere + background.
    rho_bg = (N - n_cl) / N
    rho = np.full(N, rho_bg)
    mass_sph = counts / N
    a = scale[sid]
    rho[:n_cl] += 3.0 * mass_sph[sid] / (4.0 * math.pi * a ** 3) * (
        1.0 + (r / a) ** 2) ** (-2.5)
    h0 = eta * (1.0 / (N * rho)) ** (1.0 / 3.0)
    st = _base_state(x, 0.0, 1.0 / N)
    st["h"] = h0.astype(np.float32)
    st["u"] = np.exp(rng.uniform(math.log(0.1), math.log(100.0), N)).astype(np.float32)
    perm = rng.permutation(N)
    for k in st:
        st[k] = np.ascontiguousarray(st[k][perm])
    return st


def set_active_fraction(st: dict, fraction: float, seed: int = 1,
                        active_bin: int = 1) -> None:
    """Random time-bin mask: active particles get ``active_bin``, the others
    NUM_TIME_BINS + 1, as in testActivePair.c:148-152."""
    rng = np.random.default_rng(seed)
    act = rng.uniform(0.0, 1.0, size=st["h"].shape[0]) < fraction
    tb = np.full(st["h"].shape[0], NUM_TIME_BINS + 1, np.int8)
    tb[act] = active_bin
    st["time_bin"] = tb


def copy_state(st: dict) -> dict:
    return {k: np.array(v, copy=True) for k, v in st.items()}


def set_bin_hierarchy(st: dict, fractions: dict, seed: int = 2) -> None:
    """Random time bins with the given fraction of particles per bin
    (e.g. {1: 0.015, 2: 0.045, 3: 0.19, 4: 0.75}: at times that are odd
    multiples of the smallest step 1.5 % of the particles are active, 6 %,
    25 % and 100 % at multiples of 2, 4 and 8 of it)."""
    rng = np.random.default_rng(seed)
    bins = np.array(sorted(fractions), np.int8)
    p = np.array([fractions[b] for b in sorted(fractions)], np.float64)
    st["time_bin"] = rng.choice(bins, size=st["h"].shape[0], p=p / p.sum())
