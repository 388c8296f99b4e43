"""The C-ABI library loads and exports every entry point include/jp.h declares; argument errors are
reported before any device is touched (no GPU needed)."""
import ctypes
import os

import pytest

import jpabi


def test_library_exports_every_declared_symbol():
    L = jpabi.lib()
    names = jpabi.declared_symbols()
    assert "jp_uniform_create" in names and "jp_step" in names
    for name in names:
        assert hasattr(L, name), name
    assert b"gfx950" in L.jp_version()


def test_header_is_plain_c():
    txt = open(os.path.join(jpabi.HERE, "..", "include", "jp.h")).read()
    for bad in ("torch", "std::", "hip_runtime", "class "):
        assert bad not in txt


@pytest.mark.parametrize("n", [0, 12, 7, 2000])
def test_bad_lattice_edge_is_an_argument_error(n):
    # src/tgv3d_uniform.cu:148 exits with status 2 on N % MB; the library returns JP_ERR_ARG instead
    with pytest.raises(jpabi.JPError) as e:
        jpabi.Uniform(n)
    assert e.value.code == jpabi.JP_ERR_ARG


def test_bad_precision_is_an_argument_error():
    with pytest.raises(jpabi.JPError) as e:
        jpabi.Uniform(16, prec=9)
    assert e.value.code == jpabi.JP_ERR_ARG
