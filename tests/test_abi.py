"""The C ABI library loads, exports exactly what include/h4hip.h declares,
and rejects bad arguments with error codes (no GPU needed: argument checks
run before any device call)."""
import ctypes
import os
import re
import subprocess

import pytest

from gravidy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "h4hip.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(h4h_\w+)\s*\(", src)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (h4h_\w+)", out))
    assert exported == set(_declared())


def test_abi_version_and_error_text():
    lib = _lib.load()
    assert lib.h4h_abi_version() == 107
    assert isinstance(lib.h4h_last_error(), bytes)


@pytest.mark.parametrize("n,e2,eta,dtmin,dtmax,prec", [
    (0, 1e-8, 0.01, 2.0 ** -23, 0.125, 0),        # n = 0
    (16, -1.0, 0.01, 2.0 ** -23, 0.125, 0),       # negative e2
    (16, float("nan"), 0.01, 2.0 ** -23, 0.125, 0),
    (16, 1e-8, 0.0, 2.0 ** -23, 0.125, 0),        # eta = 0
    (16, 1e-8, 0.01, 0.0, 0.125, 0),              # dt_min = 0
    (16, 1e-8, 0.01, 0.5, 0.125, 0),              # dt_min > dt_max
    (16, 1e-8, 0.01, 2.0 ** -23, 0.125, 7),       # unknown precision
])
def test_create_rejects_bad_arguments(n, e2, eta, dtmin, dtmax, prec):
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    rc = lib.h4h_create(ctypes.byref(ctx), n, e2, eta, dtmin, dtmax, prec)
    assert rc == _lib.H4H_EINVAL
    assert not ctx.value
    assert lib.h4h_last_error() != b""


def test_null_context_is_an_error():
    lib = _lib.load()
    assert lib.h4h_destroy(None) == _lib.H4H_OK
    assert lib.h4h_init_acc_jrk(None) == _lib.H4H_EINVAL
    assert lib.h4h_step(None, 0.0, None, None) == _lib.H4H_EINVAL
    assert lib.h4h_get_energy(None, None, None) == _lib.H4H_EINVAL
    assert lib.h4h_set_small_path(None, 64) == _lib.H4H_EINVAL
    assert lib.h4h_ensemble_energy(None, None, None) == _lib.H4H_EINVAL
    assert lib.h4h_create(None, 16, 1e-8, 0.01, 2.0 ** -23, 0.125, 0) == _lib.H4H_EINVAL


def test_static_update_rejects_bad_sizes():
    """h4h_update_acc_jrk_static refuses, before any device call, sizes
    beyond its static buffer and missing pointers."""
    lib = _lib.load()
    f = lib.h4h_update_acc_jrk_static
    assert f(None, None, _lib.H4H_STATIC_MAX_N + 1, None, 1, 1e-8, 0, None, None) == _lib.H4H_EINVAL
    assert b"65536" in lib.h4h_last_error()
    assert f(None, None, 16, None, 17, 1e-8, 0, None, None) == _lib.H4H_EINVAL
    assert f(None, None, 16, None, 4, 1e-8, 0, None, None) == _lib.H4H_EINVAL
    assert f(None, None, 16, None, 4, -1.0, 0, None, None) == _lib.H4H_EINVAL
    assert f(None, None, 16, None, 4, 1e-8, 9, None, None) == _lib.H4H_EINVAL
