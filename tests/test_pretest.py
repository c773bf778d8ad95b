"""The FP32 candidate pre-test of the neighbour loops never rejects a pair
that the pair contract accepts.

The loops form a cheap distance from float positions relative to the
candidate's top-level cell corner, and compare it with a widened radius
(pretest_radius2 in swift_amd/csrc/sph_device.h); the exact pair contract --
the reference's expression: the FP64 difference rounded once inside a cell,
cell-frame offsets of the pair's upper cell rounded and subtracted in float
across cells -- is applied afterwards to every pair that passes.  This test emulates both computations with IEEE
float32/float64 arithmetic in numpy on adversarial pairs -- separations
within 1e-6 of the search radius, particles at cell corners, radii from
1e-3 to 1 cell width, boxes from 1e-3 to 1e3 -- and checks the implication
r^2 < R  =>  r_f^2 < R'.  It also checks that the bound is not vacuous
(a pre-test radius without the margin does reject some contract pairs).
"""
from __future__ import annotations

import numpy as np

from parity import contract_dx

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
    # contract (t = 0 here: no wrap)
    d = contract_dx(xi, xj, ci, cj, w)
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


def _f16_pretest_case(rng, n_tiles=150):
    """The packed-FP16 pre-test of the density / gradient loops
    (swift_amd/csrc/sph_gather.h): positions relative to the centre of the
    tile's box in units of its largest search radius R, rounded to half; the
    squared distance in half arithmetic (with and without fused
    multiply-adds); each radius widened to (r + 2e)^2 (1 + 8u)(1 + 2^-10),
    e = 2 (V + 1.1) u + 1e-6, u = 2^-11, V = 1.001 max |i coordinate|."""
    f16, f32 = np.float16, np.float32
    u = 2.0 ** -11
    missed, missed_bare, total = 0, 0, 0
    for _ in range(n_tiles):
        w = 10 ** rng.uniform(-3, 3)
        H = w * 10 ** rng.uniform(-2.5, 0)
        a = H * 10 ** rng.uniform(-1, 1.5)          # tile extent
        ct = rng.uniform(0, w, 3) + w
        xi = ct + rng.uniform(-a / 2, a / 2, (16, 3))
        Hi = H * rng.uniform(0.5, 1.0, 16)
        Hi[0] = H
        hg2 = (f32(Hi) * f32(Hi)).astype(f32)
        invR = f32(1.0) / np.sqrt(hg2.max()).astype(f32)
        fx = ((xi - ct).astype(f32) * invR).astype(f32)
        irel = fx.astype(f16)
        vb = f32(1.001) * np.abs(fx).max()
        e = f32(2.0) * (vb + f32(1.1)) * f32(u) + f32(1e-6)
        n = 3000
        k = rng.integers(0, 16, n)
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        r = Hi[k] * (1 - rng.uniform(0, 2e-3, n))
        xj = xi[k] + d * r[:, None]
        corner = np.floor(xj / w) * w
        pj = (xj - corner).astype(f32)
        qoff = (corner - ct).astype(f32)
        jrel = ((pj + qoff).astype(f32) * invR).astype(f16)
        dd = (irel[k] - jrel).astype(f16).astype(np.float64)
        # separate roundings, and fused multiply-adds
        t = f16(dd[:, 0] ** 2).astype(np.float64)
        t = f16(t + f16(dd[:, 1] ** 2)).astype(np.float64)
        r2a = f16(t + f16(dd[:, 2] ** 2)).astype(np.float64)
        t = f16(dd[:, 0] ** 2).astype(np.float64)
        t = f16(dd[:, 1] ** 2 + t).astype(np.float64)
        r2b = f16(dd[:, 2] ** 2 + t).astype(np.float64)
        dc = contract_dx(xi[k], xj, np.floor(xi[k] / w).astype(np.int64),
                         np.floor(xj / w).astype(np.int64), w)
        inside = _r2(dc[:, 0], dc[:, 1], dc[:, 2]) < hg2[k]
        rn = np.sqrt(hg2[k] * invR * invR) + 2 * e
        bnd = f16(np.minimum(rn * rn * (1 + 8 * u) * (1 + 2.0 ** -10),
                             65504)).astype(np.float64)
        bare = f16(hg2[k] * invR * invR).astype(np.float64)
        missed += int((inside & ~((r2a < bnd) & (r2b < bnd))).sum())
        missed_bare += int((inside & ~(r2a < bare)).sum())
        total += int(inside.sum())
    return missed, missed_bare, total


def test_f16_pretest_never_rejects_a_contract_neighbour():
    missed, missed_bare, total = _f16_pretest_case(np.random.default_rng(11))
    assert total > 100000
    assert missed == 0
    assert missed_bare > 0  # the widening is needed
