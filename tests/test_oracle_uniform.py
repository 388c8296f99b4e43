"""CPU pins of the oracle's uniform step (src/tgv3d_uniform.cu): closed forms of the TGV initial
diagnostics, conservation, the omega = 0 pure-shift permutation, equilibrium fixed point, and the
viscous (Stokes) decay rate."""
import math

import numpy as np
import pytest

import oracle as orc

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])


def tiles_to_grid(f, n):
    """tile-major [b][q][cell] -> [q][z][y][x]"""
    g = n // 8
    a = f.reshape(g, g, g, 19, 8, 8, 8)  # bz by bx q k j i
    return a.transpose(3, 0, 4, 1, 5, 2, 6).reshape(19, n, n, n)


@pytest.mark.parametrize("n", [16, 32, 64])
def test_initial_diagnostics_closed_forms(n):
    u0 = 0.05
    u = orc.Uniform(n, u0=u0, prec=orc.PREC_FP64)
    u.init()
    E, ens, mass = u.diag()
    kf = 2 * math.pi / n
    assert E == pytest.approx(u0 ** 2 / 8, rel=1e-12)
    # central differences of sin: (sin(x+kf)-sin(x-kf))/2 = cos x sin kf -> enstrophy 3/8 u0^2 sin^2 kf
    assert ens == pytest.approx(3 / 8 * u0 ** 2 * math.sin(kf) ** 2, rel=1e-10)
    assert mass == pytest.approx(1.0, abs=1e-14)


def test_initial_diagnostics_fp32_near_closed_forms():
    u = orc.Uniform(64)
    u.init()
    E, ens, mass = u.diag()
    assert E == pytest.approx(0.05 ** 2 / 8, rel=1e-6)
    assert ens == pytest.approx(9.006900e-6, rel=1e-5)
    assert mass == pytest.approx(1.0, abs=1e-6)


def test_omega_zero_is_a_periodic_shift():
    n = 16
    u = orc.Uniform(n, prec=orc.PREC_FP64)
    rng = np.random.default_rng(7)
    f0 = rng.random(n ** 3 * 19)
    u.set_raw(f0)
    u.set_omega(0.0)
    u.step(1)
    f1 = tiles_to_grid(u.raw(), n)
    g0 = tiles_to_grid(f0, n)
    for q in range(19):
        # pull: f1(x) = f0(x - c)  <=>  roll by +c
        expect = np.roll(g0[q], shift=(CZ[q], CY[q], CX[q]), axis=(0, 1, 2))
        assert np.array_equal(f1[q], expect), q


def test_collision_conserves_mass_and_momentum_fp64():
    n = 16
    u = orc.Uniform(n, re=100, prec=orc.PREC_FP64)
    rng = np.random.default_rng(8)
    u.init()
    f = u.raw() * (1 + 0.01 * rng.standard_normal(n ** 3 * 19))
    u.set_raw(f)
    g0 = tiles_to_grid(u.raw(), n)
    u.step(1)
    g1 = tiles_to_grid(u.raw(), n)
    assert g1.sum() == pytest.approx(g0.sum(), rel=1e-13)
    for c in (CX, CY, CZ):
        assert np.tensordot(c, g1, axes=(0, 0)).sum() == pytest.approx(np.tensordot(c, g0, axes=(0, 0)).sum(), abs=1e-11)


def test_uniform_equilibrium_is_a_fixed_point():
    n = 16
    u = orc.Uniform(n, prec=orc.PREC_FP64)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    f = np.broadcast_to(w[None, :, None], (n ** 3 // 512, 19, 512)).copy().ravel()
    u.set_raw(f)
    u.step(3)
    assert np.allclose(u.raw(), f, rtol=0, atol=2e-15)


def _stokes_rate(n, re=0.5, u0=0.01):
    u = orc.Uniform(n, re=re, u0=u0, prec=orc.PREC_FP64)
    u.init()
    spt = u0 * 2 * math.pi / n  # t* per step
    s1, s2 = int(round(0.05 / spt)), int(round(0.25 / spt))
    u.step(s1)
    E1 = u.diag()[0]
    u.step(s2 - s1)
    E2 = u.diag()[0]
    return -math.log(E2 / E1) / ((s2 - s1) * spt)


def test_stokes_decay_rate():
    # Stokes limit: the TGV mode has |k|^2 = 3 kf^2, so the energy decays like exp(-6 nu kf^2 t), i.e.
    # exp(-6 t*/Re) in t* = u0 kf t. The lattice rate converges to it with resolution.
    re = 0.5
    e16 = abs(_stokes_rate(16, re) / (6 / re) - 1)
    e32 = abs(_stokes_rate(32, re) / (6 / re) - 1)
    assert e32 < 0.015
    assert e32 < e16 / 3


def test_fp32_and_fp64_agree_over_short_run():
    n = 32
    a = orc.Uniform(n)
    b = orc.Uniform(n, prec=orc.PREC_FP64)
    a.init()
    b.init()
    a.step(20)
    b.step(20)
    assert np.allclose(a.f64(), b.f64(), rtol=0, atol=2e-6)
    assert a.diag() == pytest.approx(b.diag(), rel=1e-4)


W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def test_fp32_mass_drift_is_the_weight_sum_bias():
    """The FP32 step keeps the reference's arithmetic, whose only systematic mass change is
    omega (sum of the FP32 weights - 1) per cell and step."""
    n = 16
    a = orc.Uniform(n)
    a.init()
    m0 = a.diag()[2]
    a.step(200)
    drift = (a.diag()[2] - m0) / m0 / 200
    om = np.float32(1.0 / (3 * 0.05 * n / (2 * math.pi * 1600) + 0.5))  # src/tgv3d_uniform.cu:151-152
    bias = float(om) * (float(np.sum(W.astype(np.float32).astype(np.float64))) - 1.0)
    assert drift == pytest.approx(bias, rel=0.05)


@pytest.mark.parametrize("prec", [orc.PREC_FP16, orc.PREC_BF16])
def test_shifted_forms_hold_mass(prec):
    """The shifted forms collide the deviations f - w_q, so their mass changes only by unbiased rounding
    (FP16: ~1e-7 over 400 steps at 32^3, against 1.2e-5 for the FP32 form's bias and 1.7e-4 for an FP32
    collision of w_q + d_q on the stored values). BF16's 8-bit rounding of the slowly varying
    deviations stays correlated over many steps (~5e-8 per step, measured), hence its looser bound."""
    n = 32
    a = orc.Uniform(n, prec=prec)
    a.init()
    m0 = a.diag()[2]
    a.step(400)
    assert abs(a.diag()[2] - m0) / m0 < (1e-6 if prec == orc.PREC_FP16 else 6e-5)


@pytest.mark.parametrize("prec", [orc.PREC_FP16, orc.PREC_BF16])
def test_shifted_step_is_the_fp64_step_within_storage_rounding(prec):
    """One shifted-form step from a stored state equals the FP64 step of the decoded state, up to the
    rounding of the stored deviation (half a unit in its last place) plus FP32 collision error."""
    n = 16
    h = orc.Uniform(n, prec=prec)
    h.init()
    h.step(10)
    d = orc.Uniform(n, prec=orc.PREC_FP64)
    d.set_raw(h.f64())
    h.step(1)
    d.step(1)
    exact, got = d.f64(), h.f64()
    scale = orc.default_scale(0.05)
    w = np.repeat(W, 512)[None, :].repeat(n ** 3 // 512, axis=0).reshape(-1)
    dev = (exact - w) * scale
    if prec == orc.PREC_FP16:
        half = np.spacing(np.abs(dev).astype(np.float16)).astype(np.float64) / 2
    else:
        e = np.floor(np.log2(np.maximum(np.abs(dev), 1e-30)))
        half = 2.0 ** (e - 8)
    assert np.all(np.abs(got - exact) <= half / scale * 1.0001 + 3e-8)
