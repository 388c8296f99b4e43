"""The device's velocity quotients u = j / rho (quot3 in jp-solver_amd/csrc/lattice.hpp) against IEEE binary32
division, on the host (tests/quot_check.c built with gcc, hardware fma).

The reference divides (src/tgv3d_uniform.cu:58, src/tgv3d_amr.cu:151, 165, 217, 271, 291, 353, 374); the device computes the
same correctly rounded quotients from one reciprocal and two binary64 Newton steps. This pins the routine's
arithmetic over every density in [0.9, 1.1], random operands and the quotients closest to a rounding
midpoint, with the reciprocal's start moved by up to 2 ulps; the GPU tests pin the device's bits against
the oracle, which divides as the reference does.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("quot") / "quot_check")
    subprocess.run(["gcc", "-O2", "-mfma", "-std=c11", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "quot_check.c"), "-lm"], check=True)
    return exe


def run(exe, newton):
    out = subprocess.run([exe, str(newton)], check=True, capture_output=True, text=True).stdout.split()
    return int(out[1]), int(out[3])


def test_two_newton_steps_give_ieee_quotients(checker):
    cases, bad = run(checker, 2)
    assert cases > 300_000_000
    assert bad == 0


def test_one_newton_step_is_caught(checker):
    """Planted mistake: one Newton step leaves the reciprocal within 2^-46, too coarse for the quotients
    next to a midpoint; the checker must see it."""
    _, bad = run(checker, 1)
    assert bad > 0
