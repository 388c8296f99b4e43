"""The allocation-free diagnostic entry point (jp_diag_noalloc) through the C ABI: on two calls in a row,
the second on a smaller lattice, it returns exactly what jp_diag returns (the same kernels and summation
order), and both agree with the oracle; it refuses AMR contexts."""
import numpy as np
import pytest

import jpabi
import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [jpabi.PREC_FP32, jpabi.PREC_FP16, jpabi.PREC_BF16, jpabi.PREC_FP64])
def test_noalloc_diag_equals_jp_diag(prec):
    for n in (136, 40):  # the second call on a smaller lattice, sharing the same scratch array
        o = orc.Uniform(n, prec=prec)
        o.init()
        o.step(4)
        g = jpabi.Uniform(n, prec=prec)
        g.set_raw(o.raw())
        a = g.diag_noalloc()
        b = g.diag()
        assert np.array_equal(a, b), (n, a, b)
        assert a == pytest.approx(o.diag(), rel=1e-12)
        g.step(2)
        assert np.array_equal(g.diag_noalloc(), g.diag())
        g.close()


def test_noalloc_diag_refuses_amr():
    a = jpabi.Amr(16, l1=jpabi.MODE_SLAB, l2=jpabi.MODE_ALL)
    a.init()
    with pytest.raises(jpabi.JPError) as e:
        a.diag_noalloc()
    assert e.value.code == jpabi.JP_ERR_STATE
    a.close()
