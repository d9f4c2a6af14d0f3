"""The C ABI library loads on a machine without a GPU and exports every symbol that
include/ibx.h declares; its host-side kernel matches the oracle bit for bit.  No device calls."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from suguru_amd import _lib


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _lib.declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_and_binding_agree():
    names = set(_lib.declared_symbols())
    lib = _lib.load()
    for n in names:
        assert getattr(lib, n).argtypes is not None or n in ("ibx_version",), n


def test_host_kernel_bitwise_vs_oracle():
    lib = _lib.load()
    xs = np.concatenate([np.linspace(-2, 2, 4001), np.random.default_rng(1).uniform(-2, 2, 5000)])
    for x in xs:
        assert lib.ibx_kernel(float(x)) == O.kernel(float(x))


def test_grid_struct_layout():
    # include/ibx.h: 4 int32 then 11 doubles
    assert ctypes.sizeof(_lib.IbxGrid) == 16 + 11 * 8
    assert ctypes.sizeof(_lib.IbxBatchBufs) == 16 * 8


def test_create_rejects_bad_grid_without_device():
    lib = _lib.load()
    g = _lib.IbxGrid(2, 2, 2, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0)
    h = ctypes.c_void_p()
    assert lib.ibx_create(ctypes.byref(h), ctypes.byref(g), 0) == _lib.IBX_E_ARG
    assert not h.value


def test_fluid_library_exports_every_declared_symbol():
    lib = _lib.load_fluid()
    names = _lib.declared_fluid_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
        assert getattr(lib, n).argtypes is not None, n


def test_fluid_create_rejects_bad_grid_without_device():
    """itot and jtot must be powers of two (radix-2 transforms), ktot >= 2: refused on the
    host before any device call."""
    lib = _lib.load_fluid()
    h = ctypes.c_void_p()
    for itot, jtot, ktot in ((24, 16, 8), (16, 12, 8), (16, 16, 1), (2, 16, 8)):
        g = _lib.IbxGrid(itot, jtot, ktot, 0, itot * 0.1, jtot * 0.1, ktot * 0.1, 10, 10, 10,
                         0.1, 0.1, 0.1, 0, 0)
        assert lib.fld_create(ctypes.byref(h), ctypes.byref(g), 0.0016, 1e-4, 0) == -1
        assert not h.value
