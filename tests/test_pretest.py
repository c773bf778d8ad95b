"""The FP32 candidate pre-test of the neighbour loops never rejects a pair
that the pair contract accepts.

The loops form a cheap distance from float positions relative to top-level
cell corners, and compare it with a widened radius (pretest_radius2 in
swift_amd/csrc/sph_device.h); the exact FP64 contract is applied afterwards
to every pair that passes.  This test emulates both computations with IEEE
float32/float64 arithmetic in numpy on adversarial pairs -- separations
within 1e-6 of the search radius, particles at cell corners, radii from
1e-3 to 1 cell width, boxes from 1e-3 to 1e3 -- and checks the implication
r^2 < R  =>  r_f^2 < R'.  It also checks that the bound is not vacuous
(a pre-test radius without the margin does reject some contract pairs).
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def _fma(a, b, c):
    # single rounding of a*b + c (a*b is exact in float64)
    return (a.astype(np.float64) * b.astype(np.float64) + c).astype(f32)


def _r2(dx, dy, dz):
    return _fma(dz, dz, _fma(dy, dy, (dx * dx).astype(f32)))


def _pretest_radius2(R, eps):
    # fmaf(eps, fmaf(4, sqrtf(R), 4 eps), R * 1.000002f)
    inner = _fma(f32(4.0) * np.ones_like(R), np.sqrt(R).astype(f32),
                 (f32(4.0) * eps).astype(f32))
    return _fma(np.full_like(R, eps), inner, (R * f32(1.000002)).astype(f32))


def _case(rng, dim, nc, hfrac, n=200000):
    w = dim / nc
    eps = f32(w * 2.0 ** -19)
    H = hfrac * w
    R = np.full(n, f32(H * H), dtype=f32)
    # i anywhere in cell (1,1,1); j at distance r ~ H (just inside or just
    # outside) in a random direction, in any of the 27 neighbour cells
    ci = np.ones((n, 3), dtype=np.int64)
    xi = (ci + rng.uniform(0, 1, (n, 3))) * w
    # push some i right onto cell corners
    corner = rng.uniform(size=n) < 0.3
    xi[corner] = (ci[corner] + rng.integers(0, 2, (corner.sum(), 3)) * (1 - 1e-12)) * w
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    r = H * (1.0 + rng.uniform(-1e-6, 1e-6, n))
    xj = xi + u * r[:, None]
    cj = np.floor(xj / w).astype(np.int64)
    ok = np.all((cj >= 0) & (cj <= 2), axis=1)
    xi, xj, cj, R = xi[ok], xj[ok], cj[ok], R[ok]
    ci = ci[ok]
    # contract: dx = (float)((x_i - t) - x_j), t = 0 here (no wrap)
    d = (xi - xj).astype(f32)
    r2 = _r2(d[:, 0], d[:, 1], d[:, 2])
    # pre-test: i relative to the j-cell corner, j relative to its corner
    oj = cj * w
    xs = (xi - oj).astype(f32)
    pj = (xj - oj).astype(f32)
    df = (xs - pj).astype(f32)
    r2f = _r2(df[:, 0], df[:, 1], df[:, 2])
    Rm = _pretest_radius2(R, eps)
    inside = r2 < R
    return inside, r2f < Rm, r2f < R


def test_pretest_never_rejects_a_contract_neighbour():
    rng = np.random.default_rng(0)
    missed_without_margin = 0
    for dim in (1e-3, 1.0, 1e3):
        for nc in (3, 16, 64):
            for hfrac in (1e-3, 0.05, 0.6, 1.0):
                inside, passed, passed_nomargin = _case(rng, dim, nc, hfrac)
                assert inside.sum() > 1000
                assert np.all(passed[inside]), (dim, nc, hfrac)
                missed_without_margin += int((inside & ~passed_nomargin).sum())
    assert missed_without_margin > 0
