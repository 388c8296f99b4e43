"""CPU pins of the oracle's output field (src/tgv3d_uniform.cu:96-119; src/tgv3d_amr.cu:553-564): Q-criterion
from periodic central differences, speed and velocity. Pinned by the exact discrete gradient of the TGV
field (central differences of sin/cos are sin(kf) times the shifted function), by the Q of a rigid rotation
(+Omega^2) and of a pure strain (-a^2), and by speed and velocity against numpy."""
import math

import numpy as np
import pytest

import oracle as orc


def tgv_field(n, u0=0.05):
    kf = 2 * math.pi / n
    c = (np.arange(n) + 0.5) * kf
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")  # [k, j, i]
    F = np.zeros((n, n, n, 4))
    F[..., 0] = 1.0
    F[..., 1] = u0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    F[..., 2] = -u0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    return F, X, Y, Z, kf


@pytest.mark.parametrize("diff_in_field", [True, False])
def test_q_of_the_tgv_field_is_the_exact_discrete_q(diff_in_field):
    n, u0 = 32, 0.05
    F, X, Y, Z, kf = tgv_field(n, u0)
    s = math.sin(kf)  # (f(x + 1) - f(x - 1)) / 2 of sin/cos at cell centres
    g = np.zeros((3, 3) + X.shape)
    g[0, 0] = u0 * s * np.cos(X) * np.cos(Y) * np.cos(Z)
    g[0, 1] = -u0 * s * np.sin(X) * np.sin(Y) * np.cos(Z)
    g[0, 2] = -u0 * s * np.sin(X) * np.cos(Y) * np.sin(Z)
    g[1, 0] = u0 * s * np.sin(X) * np.sin(Y) * np.cos(Z)
    g[1, 1] = -u0 * s * np.cos(X) * np.cos(Y) * np.cos(Z)
    g[1, 2] = u0 * s * np.cos(X) * np.sin(Y) * np.sin(Z)
    q_exact = -0.5 * np.einsum("ab...,ba...->...", g, g)
    out = orc.qspeed_field(F.astype(np.float32), diff_in_field)
    scale = np.abs(q_exact).max()
    assert np.max(np.abs(out["Q"] - q_exact)) < 2e-6 * scale  # FP32 field rounding only
    assert np.allclose(out["speed"], np.sqrt(F[..., 1] ** 2 + F[..., 2] ** 2), rtol=1e-6, atol=0)


def test_rotation_and_strain_signs():
    n, om, a = 16, 0.01, 0.02
    idx = np.arange(n) - n / 2
    Zc, Yc, Xc = np.meshgrid(idx, idx, idx, indexing="ij")
    rot = np.zeros((n, n, n, 4), np.float32)
    rot[..., 0] = 1
    rot[..., 1] = -om * Yc
    rot[..., 2] = om * Xc
    strain = np.zeros((n, n, n, 4), np.float32)
    strain[..., 0] = 1
    strain[..., 1] = a * Xc
    strain[..., 2] = -a * Yc
    inner = (slice(1, -1),) * 3  # away from the periodic wrap of the linear fields
    qr = orc.qspeed_field(rot, False)["Q"][inner]
    qs = orc.qspeed_field(strain, False)["Q"][inner]
    assert np.allclose(qr, om ** 2, rtol=1e-5)   # vorticity-dominated: Q = Omega^2 > 0
    assert np.allclose(qs, -a ** 2, rtol=1e-5)   # strain-dominated: Q = -a^2 < 0


def test_speed_and_velocity_pass_through():
    rng = np.random.default_rng(4)
    F = rng.standard_normal((8, 8, 8, 4)).astype(np.float32)
    out = orc.qspeed_field(F)
    u = F[..., 1:].astype(np.float64)
    assert np.array_equal(out["speed"], np.sqrt((u * u).sum(-1)).astype(np.float32))
    assert np.array_equal(out["velocity"], F[..., 1:])


def test_float_and_double_differences_agree_to_rounding():
    rng = np.random.default_rng(9)
    # values of mixed magnitude, so that float differences round (nearby values would subtract exactly)
    F = (rng.standard_normal((12, 12, 12, 4)) * np.exp(3 * rng.standard_normal((12, 12, 12, 4)))).astype(np.float32)
    a = orc.qspeed_field(F, True)["Q"].astype(np.float64)
    b = orc.qspeed_field(F, False)["Q"].astype(np.float64)
    assert np.max(np.abs(a - b) / (np.abs(b) + 1e-3 * np.abs(b).max())) < 1e-5
    assert not np.array_equal(a, b)  # the two reference paths round differently


def test_uniform_oracle_field_matches_its_macro_field():
    o = orc.Uniform(16)
    o.init()
    o.step(3)
    out = o.field()
    assert out["Q"].shape == (16, 16, 16) and np.all(np.isfinite(out["Q"]))
    assert np.all(out["speed"] >= 0) and np.max(out["speed"]) < 0.1
