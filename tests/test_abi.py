"""CPU checks of the C ABI: the in-tree library loads and exports every entry point that
include/ogp.h declares; the CLI binary exists. No compute calls (no GPU here)."""
import ctypes
import os

import pytest

from opengpop_amd import _capi


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_capi.LIB_PATH)
    names = _capi.declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_declares_signatures():
    L = _capi.lib()
    for n in _capi.declared_symbols():
        assert getattr(L, n).argtypes is not None, n


def test_default_config_is_reference():
    cfg = _capi.OgpConfig()
    _capi.lib().ogp_default_config(ctypes.byref(cfg))
    assert (cfg.degree, cfg.n_corrections, cfg.quirks, cfg.force) == (-1, 7, 3, 0)
    assert (cfg.year, cfg.month, cfg.day) == (2025, 1, 1)


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    cfg = _capi.OgpConfig()
    _capi.lib().ogp_default_config(ctypes.byref(cfg))
    h = ctypes.c_void_p()
    rc = _capi.lib().ogp_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == _capi.ERR_DEVICE
    assert b"device" in _capi.lib().ogp_last_error(None).lower()
