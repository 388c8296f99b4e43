"""CPU pins of the oracle's coarse-fine interface: seeding, ghost fill (with and without absent corners),
the fine steps' ghost reads, restriction, migration and the sensor, each against a closed form or a brute
force (tests/interface_pins.py), in the FP64 and FP32 forms."""
import pytest

import interface_pins as ip
import oracle as orc

FORMS = {"fp64": orc.PREC_FP64, "fp32": orc.PREC_FP32}


@pytest.fixture(scope="module", params=list(FORMS))
def central(request):
    return ip.build_central(orc, FORMS[request.param])


@pytest.mark.parametrize("pin", ip.CENTRAL_PINS, ids=lambda p: p.__name__)
def test_interface_pin(central, pin):
    pin(central)


@pytest.mark.parametrize("form", list(FORMS))
def test_renormalised_fill(form):
    ip.pin_renormalised_fill(orc, FORMS[form])


@pytest.mark.parametrize("form", list(FORMS))
def test_migration(form):
    ip.pin_migrate(orc, FORMS[form])
