"""C ABI checks that need no GPU: the library loads, exports every entry
point include/lulesh_hip.h declares, validates arguments before touching the
device, and carries the reference's error texts."""
import ctypes
import os
import re

import pytest

from lulesh_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "lulesh_hip.h")

REFERENCE_TEXTS = {  # src/domain.jl error() messages, in program order
    1: "Early Volume Error", 2: "Mid Volume Error", 3: "Volume Error: Volume is negative",
    4: "Volume Error: calcKinematicsForElems", 5: "Volume Error :2157", 6: "Error",
    7: "QStopError", 8: "Volume Error :2887", 9: "Volume Error :2908",
}


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lulesh_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    names = declared()
    assert len(names) >= 20
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_texts_are_the_reference_messages():
    lib = _lib.load()
    for code, text in REFERENCE_TEXTS.items():
        assert lib.lulesh_error_string(code).decode() == text
    assert lib.lulesh_error_string(0).decode() == "ok"


def test_argument_validation_before_device_use():
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    assert lib.lulesh_create(0, 11, 1, 1, 7, 0, ctypes.byref(ctx)) == 100
    assert lib.lulesh_create(8, 11, 1, 1, 9, 0, ctypes.byref(ctx)) == 100
    assert not ctx.value
    assert lib.lulesh_math_check(7, None, None, None, 1, 0) == 100
    assert lib.lulesh_destroy(None) == 100
    assert lib.lulesh_set_nodal_tiles(None, 16, 32) == 100
    assert lib.lulesh_nodal_tile(None, 0, None) == 100
    assert lib.lulesh_kernel_name(0).decode() == "k_nodal"
    # batch limits: instance count, and nodes of all instances below 2^31
    big = 1 << 20
    assert lib.lulesh_batch_create(big + 1, (ctypes.c_int32 * 1)(8), None, None, None, 7, 0,
                                   ctypes.byref(ctx)) == 100
    nx = (ctypes.c_int32 * 5)(*[800] * 5)  # 5 x 801^3 nodes > 2^31
    assert lib.lulesh_batch_create(5, nx, None, None, None, 7, 0, ctypes.byref(ctx)) == 100
    assert not ctx.value
    assert lib.lulesh_error_string(103).decode() == "host memory allocation failed"


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/liblulesh_hip.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.load()


def test_vjp_grad_params_size_checked_before_the_library():
    """The library writes three parameter gradients; a shorter buffer is
    refused on the host before any call."""
    from lulesh_amd import vjp_cycle
    for bad in ([1.0, 2.0], [0.0] * 4, [[0.0] * 3]):
        with pytest.raises(ValueError):
            vjp_cycle(object(), None, bad)


def test_batch_parameter_lengths_checked_before_the_library():
    """Per-instance parameter lists must match the instance count (a short
    hgcoef list would otherwise reach the library zero-padded)."""
    from lulesh_amd import Batch
    with pytest.raises(ValueError):
        Batch([8, 9], hgcoef=[3.0])
    with pytest.raises(ValueError):
        Batch([8, 9], einit_scale=[1.0, 1.0, 1.0])
    with pytest.raises(ValueError):
        Batch([8], nr=[1, 5])
