"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, always as the checker / CPU reference, never as the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblulesh_oracle.so")

Q1, Q2, Q3, QUIRKS_REFERENCE = 1, 2, 4, 7

_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, pi64 = ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)
    sig = {
        "oracle_create": (vp, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "oracle_create_regions": (vp, [ctypes.c_int] * 6),
        "oracle_julia_rng_state": (None, [ctypes.POINTER(ctypes.c_uint64)]),
        "oracle_destroy": (None, [vp]),
        "oracle_time_increment": (ctypes.c_int, [vp]),
        "oracle_lagrange_leapfrog": (ctypes.c_int, [vp]),
        "oracle_lagrange_nodal": (ctypes.c_int, [vp]),
        "oracle_lagrange_elements": (ctypes.c_int, [vp]),
        "oracle_calc_time_constraints": (ctypes.c_int, [vp]),
        "oracle_run": (ctypes.c_int, [vp, i64, _PD]),
        "oracle_verify": (None, [vp, _PD]),
        "oracle_dfield": (_PD, [vp, ctypes.c_char_p, pi64]),
        "oracle_ifield": (pi64, [vp, ctypes.c_char_p, pi64]),
        "oracle_elemBC": (ctypes.POINTER(ctypes.c_uint32), [vp, pi64]),
        "oracle_get_scalar": (ctypes.c_int, [vp, ctypes.c_char_p, _PD]),
        "oracle_set_scalar": (ctypes.c_int, [vp, ctypes.c_char_p, _D]),
        "oracle_cbrt": (_D, [_D]),
        "oracle_calc_elem_volume": (_D, [_PD, _PD, _PD]),
        "oracle_calc_elem_node_normals": (None, [_PD, _PD, _PD, ctypes.c_int, _PD]),
        "oracle_calc_elem_shape_function_derivatives": (_D, [_PD, _PD, _PD, _PD]),
        "oracle_calc_elem_volume_derivative": (None, [_PD, _PD, _PD, ctypes.c_int, _PD, _PD, _PD]),
        "oracle_calc_elem_characteristic_length": (_D, [_PD, _PD, _PD, _D]),
        "oracle_calc_elem_velocity_gradient": (None, [_PD, _PD, _PD, _PD, _D, _PD]),
        "oracle_calc_elem_fb_hourglass": (None, [_PD, _PD, _PD, _PD, _PD, _PD, _D, _D, _D, _D,
                                                 ctypes.c_int, _PD]),
        "oracle_eval_eos_elem": (None, [_D] * 8 + [_PD]),
        "oracle_time_increment_scalars": (None, [_PD]),
        "oracle_monoq_gradients_elem": (None, [_PD] * 6 + [_D, _D, _PD]),
        "oracle_monoq_region_elem": (None, [_PD] * 4 + [_D] * 6 + [_PD]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(_PD)


def arr(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleDomain:
    """Domain(prob) of the CPU restatement (one structured rank)."""

    def __init__(self, nx, quirks=QUIRKS_REFERENCE, nthreads=1, nr=11, balance=1, cost=1):
        self.L = load()
        self.nx = nx
        self.h = self.L.oracle_create_regions(nx, quirks, nthreads, nr, balance, cost)
        if not self.h:
            raise ValueError("oracle_create failed")

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, name):
        n = ctypes.c_int64()
        p = self.L.oracle_dfield(self.h, name.encode(), ctypes.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def set_field(self, name, values):
        n = ctypes.c_int64()
        p = self.L.oracle_dfield(self.h, name.encode(), ctypes.byref(n))
        if not p:
            raise KeyError(name)
        a = np.ctypeslib.as_array(p, shape=(n.value,))
        a[:] = values

    def ifield(self, name):
        n = ctypes.c_int64()
        p = self.L.oracle_ifield(self.h, name.encode(), ctypes.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def elemBC(self):
        n = ctypes.c_int64()
        p = self.L.oracle_elemBC(self.h, ctypes.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def set_elemBC(self, values):
        n = ctypes.c_int64()
        p = self.L.oracle_elemBC(self.h, ctypes.byref(n))
        np.ctypeslib.as_array(p, shape=(n.value,))[:] = values

    def get(self, name):
        v = _D()
        if self.L.oracle_get_scalar(self.h, name.encode(), ctypes.byref(v)) != 0:
            raise KeyError(name)
        return int(v.value) if name in ("cycle", "numElem", "numNode") else v.value

    def set(self, name, value):
        if self.L.oracle_set_scalar(self.h, name.encode(), float(value)) != 0:
            raise KeyError(name)

    def time_increment(self):
        return self.L.oracle_time_increment(self.h)

    def leapfrog(self):
        return self.L.oracle_lagrange_leapfrog(self.h)

    def nodal(self):
        return self.L.oracle_lagrange_nodal(self.h)

    def elements(self):
        return self.L.oracle_lagrange_elements(self.h)

    def constraints(self):
        return self.L.oracle_calc_time_constraints(self.h)

    def cycle_once(self):
        self.time_increment()
        return self.leapfrog()

    def run(self, num_iters):
        el = _D()
        rc = self.L.oracle_run(self.h, int(num_iters), ctypes.byref(el))
        return rc, el.value

    def verify(self):
        out = (_D * 4)()
        self.L.oracle_verify(self.h, out)
        return tuple(out)


# element-level functions
def julia_rng_state():
    out = (ctypes.c_uint64 * 4)()
    load().oracle_julia_rng_state(out)
    return list(out)


def cbrt(x):
    return load().oracle_cbrt(float(x))


def elem_volume(x, y, z):
    x, y, z = arr(x), arr(y), arr(z)
    return load().oracle_calc_elem_volume(_p(x), _p(y), _p(z))


def node_normals(x, y, z, quirks=QUIRKS_REFERENCE):
    x, y, z = arr(x), arr(y), arr(z)
    B = np.zeros((8, 3))
    load().oracle_calc_elem_node_normals(_p(x), _p(y), _p(z), quirks, _p(B))
    return B


def shape_derivatives(x, y, z):
    x, y, z = arr(x), arr(y), arr(z)
    b = np.zeros((8, 3))
    v = load().oracle_calc_elem_shape_function_derivatives(_p(x), _p(y), _p(z), _p(b))
    return b, v


def volume_derivative(x, y, z, quirks=QUIRKS_REFERENCE):
    x, y, z = arr(x), arr(y), arr(z)
    dx, dy, dz = np.zeros(8), np.zeros(8), np.zeros(8)
    load().oracle_calc_elem_volume_derivative(_p(x), _p(y), _p(z), quirks, _p(dx), _p(dy), _p(dz))
    return dx, dy, dz


def char_length(x, y, z, volume):
    x, y, z = arr(x), arr(y), arr(z)
    return load().oracle_calc_elem_characteristic_length(_p(x), _p(y), _p(z), float(volume))


def velocity_gradient(xv, yv, zv, b, detJ):
    xv, yv, zv, b = arr(xv), arr(yv), arr(zv), arr(b)
    d = np.zeros(6)
    load().oracle_calc_elem_velocity_gradient(_p(xv), _p(yv), _p(zv), _p(b), float(detJ), _p(d))
    return d


def fb_hourglass(x, y, z, xd, yd, zd, determ, ss, mass, hgcoef, quirks=QUIRKS_REFERENCE):
    x, y, z, xd, yd, zd = (arr(a) for a in (x, y, z, xd, yd, zd))
    f = np.zeros((8, 3))
    load().oracle_calc_elem_fb_hourglass(_p(x), _p(y), _p(z), _p(xd), _p(yd), _p(zd),
                                          float(determ), float(ss), float(mass), float(hgcoef),
                                          quirks, _p(f))
    return f


def eos_elem(vnew, delv, e_old, p_old, q_old, ql, qq):
    out = np.zeros(4)
    load().oracle_eval_eos_elem(float(vnew), 0.0, float(delv), float(e_old), float(p_old),
                                float(q_old), float(ql), float(qq), _p(out))
    return tuple(out)  # p, e, q, ss


def time_increment_scalars(time, deltatime, cycle, dtcourant, dthydro, stoptime=1e-2,
                           dtmax=1e-2, dtfixed=-1e-6, lb=1.1, ub=1.2):
    st = np.array([time, deltatime, cycle, dtcourant, dthydro, stoptime, dtmax, dtfixed, lb, ub],
                  dtype=np.float64)
    load().oracle_time_increment_scalars(_p(st))
    return st[0], st[1], int(st[2])


def monoq_gradients(x, y, z, xv, yv, zv, volo, vnew):
    """{delx_xi, delx_eta, delx_zeta, delv_xi, delv_eta, delv_zeta} of one element."""
    x, y, z, xv, yv, zv = (arr(a) for a in (x, y, z, xv, yv, zv))
    out = np.zeros(6)
    load().oracle_monoq_gradients_elem(_p(x), _p(y), _p(z), _p(xv), _p(yv), _p(zv), float(volo),
                                       float(vnew), _p(out))
    return out


def monoq_region(dv, dm, dp, dx, vdov, elem_mass, volo, vnew, qlc=0.5, qqc=2.0 / 3.0):
    """(qlin, qquad) of one element; dv/dm/dp/dx per direction xi, eta, zeta."""
    dv, dm, dp, dx = (arr(a) for a in (dv, dm, dp, dx))
    q = np.zeros(2)
    load().oracle_monoq_region_elem(_p(dv), _p(dm), _p(dp), _p(dx), float(vdov), float(elem_mass),
                                    float(volo), float(vnew), float(qlc), float(qqc), _p(q))
    return q[0], q[1]
