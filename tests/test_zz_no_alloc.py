"""lulesh_lagrange_leapfrog_noalloc (the one entry point that allocates
nothing) against lulesh_lagrange_leapfrog on the same Domain: every field
and the two time constraints bit for bit, cycle after cycle.  The caller's
fields live in device buffers the test allocates; the library uses only its
fixed scratch array.  Two domains in a row, the second smaller, so a stale
table or count left in the shared scratch by the first would show."""
import ctypes

import numpy as np
import pytest

from lulesh_amd import (SCALARS, Domain, LuleshProblem, _lib, lagrangeLeapFrog,
                        timeIncrement)

FIELDS_E = ("e", "p", "q", "v", "volo", "delv", "vdov", "arealg", "ss", "elemMass")
FIELDS_N = ("x", "y", "z", "xd", "yd", "zd", "nodalMass")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


_hip = None


def hip():
    """The HIP runtime, for the caller-owned device buffers (the caller may
    allocate; the entry point under test may not)."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DeviceFields:
    """The Domain's fields copied into device buffers the test allocates,
    and the lulesh_device_domain that points at them."""

    def __init__(self, dom, quirks):
        h = hip()
        self.buf, self.len = {}, {}
        host = {n: dom.get(n) for n in FIELDS_E + FIELDS_N}
        host["elemBC"] = dom.get("elemBC")
        for n, a in host.items():
            p = ctypes.c_void_p()
            assert h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
            assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
            self.buf[n], self.len[n] = p, a.size
        d = _lib.DeviceDomain()
        d.nx, d.quirks = dom.sizeX, quirks
        for n in FIELDS_E + FIELDS_N:
            setattr(d, n, ctypes.cast(self.buf[n], _lib.P_double))
        d.elemBC = ctypes.cast(self.buf["elemBC"], _lib.P_uint32)
        self.d = d

    def get(self, n):
        out = np.empty(self.len[n], dtype=np.float64)
        assert hip().hipMemcpy(out.ctypes.data, self.buf[n], out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.buf.values():
            hip().hipFree(p)
        self.buf = {}


def scalars_of(dom):
    return np.array([float(dom.get(n)) for n in SCALARS], dtype=np.float64)


def run_both(nx, cycles, quirks=7):
    lib = _lib.load()
    print(f"[noalloc] nx={nx} cycles={cycles} quirks={quirks}", flush=True)
    dom = Domain(LuleshProblem(10, True, nx, 11, 1, 1), quirks=quirks)
    df = DeviceFields(dom, quirks)
    for c in range(cycles):
        timeIncrement(dom)
        sc = scalars_of(dom)  # the cycle's inputs, dt included
        rc = lib.lulesh_lagrange_leapfrog_noalloc(ctypes.byref(df.d),
                                                  sc.ctypes.data_as(_lib.P_double), None, 0)
        assert rc == 0, (nx, c, rc)
        lagrangeLeapFrog(dom)
        want = scalars_of(dom)
        assert bits(sc).tolist() == bits(want).tolist(), (nx, c, sc, want)
    for n in FIELDS_E + FIELDS_N:
        g, o = df.get(n), dom.get(n)
        assert np.array_equal(bits(g), bits(o)), (nx, n)
    df.close()
    dom.close()


@pytest.mark.gpu
def test_noalloc_matches_context_entry_point():
    # the largest domain the scratch holds, then a smaller one
    run_both(_lib.NOALLOC_MAX_NX, 12)
    run_both(12, 60)
    run_both(9, 40, quirks=0)


@pytest.mark.gpu
def test_noalloc_reports_reference_errors():
    lib = _lib.load()
    print("[noalloc] QStop", flush=True)
    dom = Domain(LuleshProblem(10, True, 8, 11, 1, 1))
    df = DeviceFields(dom, 7)
    timeIncrement(dom)
    sc = scalars_of(dom)
    sc[SCALARS.index("qstop")] = -1.0  # q = 0 > -1 at the first cycle
    rc = lib.lulesh_lagrange_leapfrog_noalloc(ctypes.byref(df.d), sc.ctypes.data_as(_lib.P_double),
                                              None, 0)
    assert rc == 7
    # the next call starts with a clear error code
    sc[SCALARS.index("qstop")] = 1.0e12
    rc = lib.lulesh_lagrange_leapfrog_noalloc(ctypes.byref(df.d), sc.ctypes.data_as(_lib.P_double),
                                              None, 0)
    assert rc == 0
    df.close()
    dom.close()


def test_noalloc_rejects_sizes_and_arguments_before_the_device():
    """Checked on the host before any GPU call (runs without a GPU)."""
    lib = _lib.load()
    d = _lib.DeviceDomain()
    sc = np.zeros(len(SCALARS))
    d.nx, d.quirks = _lib.NOALLOC_MAX_NX + 1, 7
    assert lib.lulesh_lagrange_leapfrog_noalloc(ctypes.byref(d), sc.ctypes.data_as(_lib.P_double),
                                                None, 0) == _lib.LULESH_ERR_SCRATCH
    d.nx = 0
    assert lib.lulesh_lagrange_leapfrog_noalloc(ctypes.byref(d), sc.ctypes.data_as(_lib.P_double),
                                                None, 0) == _lib.LULESH_ERR_INVALID_ARG
    d.nx = 8  # null field pointers
    assert lib.lulesh_lagrange_leapfrog_noalloc(ctypes.byref(d), sc.ctypes.data_as(_lib.P_double),
                                                None, 0) == _lib.LULESH_ERR_INVALID_ARG
    assert lib.lulesh_lagrange_leapfrog_noalloc(None, sc.ctypes.data_as(_lib.P_double), None,
                                                0) == _lib.LULESH_ERR_INVALID_ARG
    assert b"scratch" in lib.lulesh_error_string(_lib.LULESH_ERR_SCRATCH)
