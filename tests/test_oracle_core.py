"""CPU pins of the oracle's lattice core and storage forms (no GPU).

Each check is fixed by mathematics or by an independent implementation, not by retyping the oracle:
lattice isotropy, moments of the equilibrium, Hermite-projection identities, numpy's float16 rounding,
nearest-value property of the bfloat16 rounding.
"""
import numpy as np
import pytest

import oracle as orc

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
C = np.stack([CX, CY, CZ], 1)


def feq(q, rho, ux, uy, uz):
    out = np.zeros(1, dtype=np.float32)
    orc.lib().orc_feq(q, rho, ux, uy, uz, out.ctypes.data)
    return float(out[0])


def test_velocity_set_is_d3q19_and_isotropic():
    # 19 distinct velocities: rest, 6 axis, 12 face diagonals; each has its opposite
    assert len({tuple(c) for c in C}) == 19
    assert all(tuple(-c) in {tuple(d) for d in C} for c in C)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(w @ C, 0)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, C, C), np.eye(3) / 3)


def test_equilibrium_moments_reproduce_rho_and_momentum():
    rng = np.random.default_rng(1)
    for _ in range(20):
        rho = 1 + 0.01 * rng.standard_normal()
        u = 0.05 * rng.standard_normal(3)
        f = np.array([feq(q, rho, *u) for q in range(19)], dtype=np.float64)
        assert abs(f.sum() - rho) < 1e-6
        assert np.allclose(f @ C, rho * u, atol=1e-7)
        # second moment: rho (cs^2 delta + u u)
        P = np.einsum("q,qa,qb->ab", f, C, C)
        assert np.allclose(P, rho * (np.eye(3) / 3 + np.outer(u, u)), atol=1e-6)


def test_regularize_is_a_projection_that_keeps_the_stress():
    rng = np.random.default_rng(2)
    for _ in range(20):
        n = rng.standard_normal(19)
        a = n.copy()
        orc.lib().orc_regularize64(a.ctypes.data)
        b = a.copy()
        orc.lib().orc_regularize64(b.ctypes.data)
        assert np.allclose(a, b, atol=1e-13)  # idempotent
        # keeps the traceless part of the second moment of n, zero mass and momentum
        Pn = np.einsum("q,qa,qb->ab", n, C, C)
        Pa = np.einsum("q,qa,qb->ab", a, C, C)
        tl = lambda P: P - np.trace(P) / 3 * np.eye(3)
        assert np.allclose(tl(Pa), tl(Pn), atol=1e-12)
        assert np.allclose(a @ C, 0, atol=1e-12)


def test_regularize_sum_rule():
    # Hermite projection onto H2: sum_q n_q = 4.5 * sum_ab (sum_q w_q Q_ab) P_ab = 0 since sum_q w_q Q = 0
    rng = np.random.default_rng(3)
    n = rng.standard_normal(19)
    orc.lib().orc_regularize64(n.ctypes.data)
    assert abs(n.sum()) < 1e-12


def test_float16_rounding_matches_numpy():
    rng = np.random.default_rng(4)
    x = np.concatenate([
        rng.standard_normal(200000).astype(np.float32) * np.float32(2.0) ** rng.integers(-30, 17, 200000).astype(np.float32),
        np.array([0.0, -0.0, 65504, 65519.99, 65520, 1e6, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8, 1e-9,
                  np.inf, -np.inf], dtype=np.float32),
    ]).astype(np.float32)
    # every binary16 value and the midpoints between neighbours (exact ties)
    h = np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float32)
    mids = ((h[:-1].astype(np.float64) + h[1:]) / 2).astype(np.float32)
    x = np.concatenate([x, h, -h, mids, -mids])
    ours = np.empty(x.size, dtype=np.uint16)
    orc.lib().orc_f32_to_f16(x.ctypes.data, ours.ctypes.data, x.size)
    ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)
    back = np.empty(x.size, dtype=np.float32)
    orc.lib().orc_f16_to_f32(ours.ctypes.data, back.ctypes.data, x.size)
    assert np.array_equal(back, ours.view(np.float16).astype(np.float32))


def test_bfloat16_rounding_is_nearest_even():
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(100000) * 10.0 ** rng.integers(-20, 20, 100000)).astype(np.float32)
    bits = x.view(np.uint32)
    # add exact ties: low 16 bits = 0x8000
    ties = ((bits[:2000] & 0xFFFF0000) | 0x8000).view(np.float32)
    x = np.concatenate([x, ties])
    out = np.empty(x.size, dtype=np.uint16)
    orc.lib().orc_f32_to_bf16(x.ctypes.data, out.ctypes.data, x.size)
    dec = lambda h: (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    d0, dm, dp = dec(out), dec(out - 1), dec(out + 1)
    xd = x.astype(np.float64)
    e0 = np.abs(xd - d0)
    assert np.all(e0 <= np.abs(xd - dm)) and np.all(e0 <= np.abs(xd - dp))
    tie = (e0 == np.abs(xd - dm)) | (e0 == np.abs(xd - dp))
    assert np.all(out[tie] % 2 == 0)


@pytest.mark.parametrize("prec", [orc.PREC_FP16, orc.PREC_BF16])
def test_shifted_store_error_bound(prec):
    # decode(encode(f)) error is at most half an ulp of the stored (f - w) * scale, divided by scale
    u = orc.Uniform(16, prec=prec)
    u.init()
    f = u.f64().reshape(-1, 19, 512)
    u32 = orc.Uniform(16, prec=orc.PREC_FP32)
    u32.init()
    g = u32.f64().reshape(-1, 19, 512)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)[None, :, None]
    mant = 11 if prec == orc.PREC_FP16 else 8
    bound = np.abs(g - w) * 2.0 ** -mant + 1e-7 * np.abs(g) + 1e-12
    assert np.all(np.abs(f - g) <= bound)
