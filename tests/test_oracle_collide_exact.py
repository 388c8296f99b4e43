"""CPU pin of the oracle's fused stream-collide step, bit for bit, against an independent numpy
restatement (src/tgv3d_uniform.cu:41-61, the BGK collision of :55-60).

The restatement pulls with np.roll, sums the moments in q order in float32, and evaluates the
collision's multiply-adds as single-rounding fused operations: `fma32` forms the exact product in
float64 (two 24-bit significands fit in 53 bits), adds with an error-free two-sum, and rounds the exact
sum once to float32, resolving the one case where rounding the float64 sum first would differ (a sum
that lands exactly on a float32 midpoint) by the sign of the two-sum's error. The plain and the shifted
(FP16, BF16 deviation) forms are checked from random states at 16^3. A contraction point written
differently, a weight, a sign or a moment order that differs from the collision's definition changes
the bits of some population and fails here.
"""
import numpy as np
import pytest

import oracle as orc

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
F32 = np.float32
W = [F32(1) / F32(3)] + [F32(1) / F32(18)] * 6 + [F32(1) / F32(36)] * 12


def fma32(a, b, c):
    """a * b + c rounded once to float32 (a, b, c float32 arrays or scalars)."""
    a, b, c = (np.asarray(x, dtype=np.float32).astype(np.float64) for x in (a, b, c))
    p = a * b  # exact
    s = p + c
    bb = s - p
    e = (p - (s - bb)) + (c - bb)  # p + c == s + e exactly
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    # neighbour of r on the side of s; s exactly between r and it is the only double-rounding case
    other = np.nextafter(r, np.where(s > r64, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32))
    mid = (r64 + other.astype(np.float64)) * 0.5
    tie = (s != r64) & (s == mid)
    up, down = np.maximum(r, other), np.minimum(r, other)
    r = np.where(tie & (e > 0), up, np.where(tie & (e < 0), down, r))
    return r.astype(np.float32)


def test_fma32_against_exact_rationals():
    from fractions import Fraction

    rng = np.random.default_rng(7)
    a = rng.standard_normal(4000).astype(np.float32)
    b = rng.standard_normal(4000).astype(np.float32)
    c = (rng.standard_normal(4000) * 10.0 ** rng.integers(-9, 3, 4000)).astype(np.float32)
    # planted exact midpoints: a*b + c = float32 midpoint + tiny, where rounding in float64 first loses the tiny
    a[:8] = np.float32(1 + 2.0 ** -12)
    b[:8] = np.float32(1 + 2.0 ** -12)
    c[:8] = np.float32(2.0 ** -80) * np.array([1, -1, 3, -3, 0, -1, 5, -5], dtype=np.float32)
    got = fma32(a, b, c)

    def round32(x):
        r = np.float32(float(x))
        cands = [r, np.nextafter(r, np.float32(np.inf)), np.nextafter(r, np.float32(-np.inf))]
        best = min(cands, key=lambda v: (abs(Fraction(float(v)) - x), int(np.float32(v).view(np.uint32)) & 1))
        return best

    for i in range(len(a)):
        x = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        assert got[i] == round32(x), (i, a[i], b[i], c[i])


def tiles_to_grid(f, n):
    g = n // 8
    return f.reshape(g, g, g, 19, 8, 8, 8).transpose(3, 0, 4, 1, 5, 2, 6).reshape(19, n, n, n)


def grid_to_tiles(a, n):
    g = n // 8
    return a.reshape(19, g, 8, g, 8, g, 8).transpose(1, 3, 5, 0, 2, 4, 6).reshape(-1)


def restated_step(fin, om, shift):
    """One pull stream-collide of the form's working values fin[q][z][y][x] (float32)."""
    f = [np.roll(fin[q], (CZ[q], CY[q], CX[q]), axis=(0, 1, 2)) for q in range(19)]
    rho = np.zeros_like(f[0])
    ux, uy, uz = rho.copy(), rho.copy(), rho.copy()
    for q in range(19):
        rho = rho + f[q]
        ux = ux + F32(CX[q]) * f[q]
        uy = uy + F32(CY[q]) * f[q]
        uz = uz + F32(CZ[q]) * f[q]
    drho = rho
    if shift:
        rho = F32(1) + drho
    ux, uy, uz = ux / rho, uy / rho, uz / rho
    t15 = F32(1.5) * fma32(uz, uz, fma32(uy, uy, ux * ux))
    omt = F32(1) - t15
    out = np.empty((19,) + rho.shape, dtype=np.float32)
    for q in range(19):
        if q == 0:
            d = fma32(-W[0], fma32(rho, -t15, drho), f[0]) if shift else fma32(-(W[0] * rho), omt, f[0])
        else:
            cu = F32(3) * (F32(CX[q]) * ux + F32(CY[q]) * uy + F32(CZ[q]) * uz)
            hc = F32(0.5) * cu
            if shift:
                d = fma32(-W[q], fma32(rho, fma32(hc, cu, -t15) + cu, drho), f[q])
            else:
                d = fma32(-(W[q] * rho), fma32(hc, cu, omt) + cu, f[q])
        out[q] = fma32(-F32(om), d, f[q])
    return out


def bf16_bits(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("om", [1.0, 1.7, 1.9886])
def test_fp32_step_bits(om):
    n = 16
    rng = np.random.default_rng(11)
    w = np.array(W, dtype=np.float32)[:, None, None, None]
    g = (w * (1 + 0.15 * rng.standard_normal((19, n, n, n)))).astype(np.float32)
    u = orc.Uniform(n, prec=orc.PREC_FP32)
    u.set_raw(grid_to_tiles(g, n))
    u.set_omega(om)
    u.step(1)
    got = tiles_to_grid(u.raw(), n)
    want = restated_step(g, om, shift=False)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("prec", [orc.PREC_FP16, orc.PREC_BF16])
def test_shifted_step_bits(prec):
    n = 16
    om = 1.9886
    rng = np.random.default_rng(5 + prec)
    u = orc.Uniform(n, prec=prec)
    sc = F32(u.scale)
    w = np.array(W, dtype=np.float32)[:, None, None, None]
    dev = (w * 0.15 * rng.standard_normal((19, n, n, n))).astype(np.float32)
    if prec == orc.PREC_FP16:
        raw = (dev * sc).astype(np.float16)
        d = raw.astype(np.float32) * (F32(1) / sc)
    else:
        raw = bf16_bits(dev * sc)
        d = bf16_value(raw) * (F32(1) / sc)
    u.set_raw(grid_to_tiles(raw.view(np.uint16), n))
    u.set_omega(om)
    u.step(1)
    got = tiles_to_grid(u.raw(), n)
    out = restated_step(d, om, shift=True) * sc
    want = out.astype(np.float16).view(np.uint16) if prec == orc.PREC_FP16 else bf16_bits(out)
    assert np.array_equal(got.view(np.uint16), want)
