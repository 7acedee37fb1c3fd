"""Host-side mirror of LULESH.jl's Domain and driver API, backed by the
MI355X kernels of liblulesh_hip.

Names, argument meaning and error behaviour follow the reference:
  LuleshProblem        src/LULESH.jl:31-49
  Domain(prob)         src/domain.jl:467-771   (built on the device)
  timeIncrement!       src/domain.jl:920-961   -> timeIncrement(domain)
  lagrangeLeapFrog     src/domain.jl:3067-3080
  verifyAndWriteFinalOutput  src/utils.jl:124-164
  main (driver loop)   examples/benchmark.jl:22-111
A physics error raises LuleshError carrying the reference's error() text.
Fields are read and written by their reference names (domain.e,
domain.x, domain.time ...); each access copies between host and device.
"""
import ctypes
import time as _time

import numpy as np

from . import _lib

QUIRK_FACE5_NORMAL = 1
QUIRK_DVDX_SWAP = 2
QUIRK_XIP_SELF = 4
QUIRKS_REFERENCE = 7

ELEM_FIELDS = ("e", "p", "q", "ql", "qq", "v", "volo", "delv", "vdov", "arealg", "ss",
               "elemMass", "vnew", "delv_xi", "delv_eta", "delv_zeta", "delx_xi", "delx_eta",
               "delx_zeta", "dxx", "dyy", "dzz")
NODE_FIELDS = ("x", "y", "z", "xd", "yd", "zd", "xdd", "ydd", "zdd", "fx", "fy", "fz",
               "nodalMass")
SCALARS = ("time", "deltatime", "cycle", "dtcourant", "dthydro", "stoptime", "dtmax",
           "dtfixed", "deltatimemultlb", "deltatimemultub", "u_cut", "hgcoef", "qstop",
           "monoq_max_slope", "monoq_limiter_mult", "e_cut", "p_cut", "ss4o3", "q_cut",
           "v_cut", "qlc_monoq", "qqc_monoq", "qqc", "eosvmax", "eosvmin", "pmin", "emin",
           "dvovmax", "refdens")

KERNELS = ("k_nodal", "k_kin", "k_qeos", "k_dt")


class LuleshError(Exception):
    """Raised where the reference calls error(msg); str(exc) is msg."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class LuleshProblem:
    """LuleshProblem(num_iters, structured, nx, nr, balance, cost, devicetype,
    floattype, comm) of src/LULESH.jl:31-49, for one rank (comm=None)."""

    def __init__(self, num_iters, structured, nx, nr=11, balance=1, cost=1, devicetype="hip",
                 floattype=np.float64, comm=None):
        if comm is not None:
            raise NotImplementedError("one single-GPU domain: comm must be None")
        if floattype not in (np.float64, float):
            raise ValueError("the path is FP64 only")
        self.num_iters = int(num_iters)
        self.structured = bool(structured)
        self.nx = int(nx)
        self.nr, self.balance, self.cost = int(nr), int(balance), int(cost)
        self.devicetype, self.floattype, self.comm = devicetype, np.float64, None
        # initMeshDecomp for one rank (src/LULESH.jl:62-97)
        self.col = self.row = self.plane = 0
        self.side = 1


def _check(lib, ctx, rc):
    if rc != 0:
        msg = lib.lulesh_last_error(ctx) if ctx else lib.lulesh_error_string(rc)
        msg = msg.decode() if isinstance(msg, bytes) else str(msg)
        raise LuleshError(rc, msg)


class Domain:
    """GPU-resident Domain{Float64}; Domain(prob) of src/domain.jl:467-771.

    A Domain is also the per-instance view of a Batch (see Batch.instance);
    field and scalar access then address that instance, while the
    cycle-level functions (timeIncrement, lagrangeLeapFrog, run ...) always
    advance every instance of the batch."""

    def __init__(self, prob, quirks=QUIRKS_REFERENCE, device=0):
        if not prob.structured:
            # src/domain.jl:611-613
            raise LuleshError(0, "Reading unstructured mesh is currently missing in the Julia "
                                 "version of LULESH.")
        lib = _lib.load()
        self._lib = lib
        ctx = ctypes.c_void_p()
        rc = lib.lulesh_create(prob.nx, prob.nr, prob.balance, prob.cost, int(quirks),
                               int(device), ctypes.byref(ctx))
        if rc != 0:
            raise LuleshError(rc, lib.lulesh_error_string(rc).decode())
        self._ctx = ctx
        self._inst = 0
        self._owner = True
        self.prob = prob
        self.quirks = int(quirks)
        self.sizeX = self.sizeY = self.sizeZ = prob.nx
        self.numElem = int(lib.lulesh_num_elem(ctx))
        self.numNode = int(lib.lulesh_num_node(ctx))
        self.numReg, self.balance, self.cost = prob.nr, prob.balance, prob.cost
        self.comm = None

    @classmethod
    def _view(cls, batch, inst, nx):
        d = object.__new__(cls)
        object.__setattr__(d, "_lib", batch._lib)
        object.__setattr__(d, "_ctx", batch._ctx)
        object.__setattr__(d, "_inst", int(inst))
        object.__setattr__(d, "_owner", False)
        object.__setattr__(d, "_batch", batch)
        object.__setattr__(d, "prob", None)
        object.__setattr__(d, "quirks", batch.quirks)
        for k in ("sizeX", "sizeY", "sizeZ"):
            object.__setattr__(d, k, int(nx))
        object.__setattr__(d, "numElem", int(batch._lib.lulesh_batch_num_elem(batch._ctx, inst)))
        object.__setattr__(d, "numNode", int(batch._lib.lulesh_batch_num_node(batch._ctx, inst)))
        object.__setattr__(d, "numReg", int(batch.nr[inst]))
        object.__setattr__(d, "comm", None)
        return d

    # -- lifetime --
    def close(self):
        if getattr(self, "_ctx", None) and getattr(self, "_owner", False):
            self._lib.lulesh_destroy(self._ctx)
        object.__setattr__(self, "_ctx", None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        _check(self._lib, self._ctx, self._lib.lulesh_reset(self._ctx))

    # -- k_nodal geometry (no reference counterpart; whole context) --
    def autotune(self):
        """Time k_nodal's z-march depths on the live state and keep the
        fastest; returns (depth, ms per launch)."""
        ch, ms = ctypes.c_int32(), ctypes.c_double()
        _check(self._lib, self._ctx,
               self._lib.lulesh_autotune(self._ctx, ctypes.byref(ch), ctypes.byref(ms)))
        return ch.value, ms.value

    def set_nodal_depth(self, depth):
        _check(self._lib, self._ctx, self._lib.lulesh_set_nodal_depth(self._ctx, int(depth)))

    def nodal_depth(self):
        return int(self._lib.lulesh_nodal_depth(self._ctx))

    def set_nodal_tiles(self, lo=0, hi=0):
        """Allow k_nodal tiles etx in [lo, hi] (0, 0: the default 16..32);
        results do not depend on it."""
        _check(self._lib, self._ctx, self._lib.lulesh_set_nodal_tiles(self._ctx, int(lo), int(hi)))

    def nodal_tile(self):
        """This instance's k_nodal tile: (etx, ety, tiles along x, tiles along y)."""
        out = (ctypes.c_int32 * 4)()
        _check(self._lib, self._ctx, self._lib.lulesh_nodal_tile(self._ctx, self._inst, out))
        return tuple(int(v) for v in out)

    # -- field access by reference name --
    def _len(self, name):
        if name in ELEM_FIELDS:
            return self.numElem
        if name in NODE_FIELDS:
            return self.numNode
        raise AttributeError(name)

    def get(self, name):
        if name in SCALARS:
            v = ctypes.c_double()
            _check(self._lib, self._ctx, self._lib.lulesh_batch_get_scalar(
                self._ctx, self._inst, name.encode(), ctypes.byref(v)))
            return int(v.value) if name == "cycle" else v.value
        if name == "elemBC":
            out = np.empty(self.numElem, dtype=np.uint32)
            _check(self._lib, self._ctx, self._lib.lulesh_batch_download_elem_bc(
                self._ctx, self._inst, out.ctypes.data_as(_lib.P_uint32), out.size))
            return out
        n = self._len(name)
        out = np.empty(n, dtype=np.float64)
        _check(self._lib, self._ctx, self._lib.lulesh_batch_download(
            self._ctx, self._inst, name.encode(), out.ctypes.data_as(_lib.P_double), n))
        return out

    def set(self, name, value):
        if name in SCALARS:
            _check(self._lib, self._ctx, self._lib.lulesh_batch_set_scalar(
                self._ctx, self._inst, name.encode(), float(value)))
            return
        if name == "elemBC":
            a = np.ascontiguousarray(value, dtype=np.uint32)
            _check(self._lib, self._ctx, self._lib.lulesh_batch_upload_elem_bc(
                self._ctx, self._inst, a.ctypes.data_as(_lib.P_uint32), a.size))
            return
        a = np.ascontiguousarray(value, dtype=np.float64)
        _check(self._lib, self._ctx, self._lib.lulesh_batch_upload(
            self._ctx, self._inst, name.encode(), a.ctypes.data_as(_lib.P_double), a.size))

    def __getattr__(self, name):
        if name.startswith("_") or name in ("prob", "quirks"):
            raise AttributeError(name)
        if name in SCALARS or name in ELEM_FIELDS or name in NODE_FIELDS or name == "elemBC":
            return self.get(name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in SCALARS or name in ELEM_FIELDS or name in NODE_FIELDS or name == "elemBC":
            self.set(name, value)
        else:
            object.__setattr__(self, name, value)

    def keep_aux(self, on=True):
        """Also write fx, xdd, dxx, ql, qq ... (off by default for speed)."""
        _check(self._lib, self._ctx, self._lib.lulesh_set_keep_aux(self._ctx, 1 if on else 0))

    # -- regions (createRegionIndexSets!, src/domain.jl:338-465) --
    def region(self, name):
        """regNumList / regElemlist (numElem entries) or regCSR / regReps /
        regSorted (numReg entries), as the reference stores them (1-based
        region numbers and element ids)."""
        n = self.numElem if name in ("regNumList", "regElemlist") else self.numReg
        out = np.empty(n, dtype=np.int64)
        _check(self._lib, self._ctx, self._lib.lulesh_region_download(
            self._ctx, self._inst, name.encode(), out.ctypes.data_as(_lib.P_int64), n))
        return out

    def set_cost_mode(self, on=True):
        """LULESH 2.0's load-imbalance model: each element's energy update
        is evaluated regReps times for its region (whole context; results
        unchanged)."""
        _check(self._lib, self._ctx, self._lib.lulesh_set_cost_mode(self._ctx, 1 if on else 0))

    # -- measurement --
    def set_timing(self, on=True):
        _check(self._lib, self._ctx, self._lib.lulesh_set_timing(self._ctx, 1 if on else 0))

    def kernel_times(self):
        out = {}
        for k, name in enumerate(KERNELS):
            ms, cnt = ctypes.c_double(), ctypes.c_int64()
            _check(self._lib, self._ctx, self._lib.lulesh_kernel_time(
                self._ctx, k, ctypes.byref(ms), ctypes.byref(cnt)))
            out[name] = (ms.value, cnt.value)
        return out


# -- the hot path, by the reference's names --
def timeIncrement(domain):
    """timeIncrement!(domain), src/domain.jl:920-961."""
    _check(domain._lib, domain._ctx, domain._lib.lulesh_time_increment(domain._ctx))


def lagrangeLeapFrog(domain):
    """lagrangeLeapFrog(domain), src/domain.jl:3067-3080."""
    _check(domain._lib, domain._ctx, domain._lib.lulesh_lagrange_leapfrog(domain._ctx))


def lagrangeNodal(domain):
    """lagrangeNodal(domain), src/domain.jl:1840-1872."""
    _check(domain._lib, domain._ctx, domain._lib.lulesh_lagrange_nodal(domain._ctx))


def lagrangeElements(domain):
    """lagrangeElements(domain), src/domain.jl:2932-2962."""
    _check(domain._lib, domain._ctx, domain._lib.lulesh_lagrange_elements(domain._ctx))


def calcTimeConstraintsForElems(domain):
    """calcTimeConstraintsForElems(domain), src/domain.jl:3058-3064."""
    _check(domain._lib, domain._ctx, domain._lib.lulesh_calc_time_constraints(domain._ctx))


def run(domain, num_iters):
    """The loop of examples/benchmark.jl:82-98 on the device; returns cycles run."""
    n = ctypes.c_int64()
    _check(domain._lib, domain._ctx,
           domain._lib.lulesh_run(domain._ctx, int(num_iters), ctypes.byref(n)))
    return n.value


def verify(domain):
    """(e[1], MaxAbsDiff, TotalAbsDiff, MaxRelDiff) of src/utils.jl:131-152."""
    out = (ctypes.c_double * 4)()
    _check(domain._lib, domain._ctx,
           domain._lib.lulesh_batch_verify(domain._ctx, domain._inst, out))
    return tuple(out)


def grad_run(domain, ncycles, checkpoint_every=20):
    """Gradient of the total internal energy sum(e * elemMass) after
    `ncycles` cycles with respect to the current e field and to hgcoef,
    qlc_monoq and qqc_monoq, with the dt schedule of the forward pass held
    fixed (replaces the Enzyme call of examples/benchmark.jl:85-86).  The
    domain is left in the forward pass's end state."""
    E = domain.numElem
    ge = np.zeros(E)
    gp = np.zeros(3)
    loss = ctypes.c_double()
    dt = np.zeros(int(ncycles))
    _check(domain._lib, domain._ctx, domain._lib.lulesh_grad_run(
        domain._ctx, int(ncycles), int(checkpoint_every), ge.ctypes.data_as(_lib.P_double),
        gp.ctypes.data_as(_lib.P_double), ctypes.byref(loss), dt.ctypes.data_as(_lib.P_double)))
    return {"loss": loss.value, "grad_e": ge, "grad_hgcoef": gp[0], "grad_qlc_monoq": gp[1],
            "grad_qqc_monoq": gp[2], "dt": dt}


def grad_run_revolve(domain, ncycles, snapshots):
    """grad_run in bounded memory: binomial checkpointing (Griewank and
    Walther's revolve) with at most `snapshots` saved states, re-running
    cycles as the optimal schedule requires.  Same results as grad_run, bit
    for bit, plus "forward_cycles", the cycles run in all."""
    E = domain.numElem
    ge = np.zeros(E)
    gp = np.zeros(3)
    loss = ctypes.c_double()
    dt = np.zeros(int(ncycles))
    fwd = ctypes.c_int64()
    _check(domain._lib, domain._ctx, domain._lib.lulesh_grad_run_revolve(
        domain._ctx, int(ncycles), int(snapshots), ge.ctypes.data_as(_lib.P_double),
        gp.ctypes.data_as(_lib.P_double), ctypes.byref(loss), dt.ctypes.data_as(_lib.P_double),
        ctypes.byref(fwd)))
    return {"loss": loss.value, "grad_e": ge, "grad_hgcoef": gp[0], "grad_qlc_monoq": gp[1],
            "grad_qqc_monoq": gp[2], "dt": dt, "forward_cycles": fwd.value}


def replay(domain, dt):
    """Advance with a prescribed dt per cycle instead of timeIncrement!."""
    a = np.ascontiguousarray(dt, dtype=np.float64)
    _check(domain._lib, domain._ctx,
           domain._lib.lulesh_replay(domain._ctx, a.ctypes.data_as(_lib.P_double), a.size))


SHADOW_FIELDS = ("x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss")


def set_shadow(domain, field, values):
    """Load one field of the context's shadow (adjoint of that state field)."""
    a = np.ascontiguousarray(values, dtype=np.float64)
    _check(domain._lib, domain._ctx, domain._lib.lulesh_shadow_upload(
        domain._ctx, field.encode(), a.ctypes.data_as(_lib.P_double), a.size))


def get_shadow(domain, field):
    n = domain.numNode if field in ("x", "y", "z", "xd", "yd", "zd") else domain.numElem
    out = np.empty(n, dtype=np.float64)
    _check(domain._lib, domain._ctx, domain._lib.lulesh_shadow_download(
        domain._ctx, field.encode(), out.ctypes.data_as(_lib.P_double), n))
    return out


def vjp_cycle(domain, shadow=None, grad_params=None):
    """One lagrangeLeapFrog under reverse mode, the counterpart of
    `Enzyme.autodiff(lagrangeLeapFrog, Duplicated(domain, shadowDomain))`
    (examples/benchmark.jl:86): the domain advances one cycle and the shadow
    maps from the adjoint of the state after the cycle to the adjoint of the
    state before it.  `shadow` (dict field -> array over SHADOW_FIELDS, or
    None to use the device-resident shadow as it stands) is loaded first and
    the returned dict holds the new shadow; `grad_params` (3 floats: hgcoef,
    qlc_monoq, qqc_monoq) accumulates."""
    gp = np.zeros(3) if grad_params is None else np.asarray(grad_params, dtype=np.float64)
    if gp.shape != (3,):  # the library writes exactly three values
        raise ValueError("grad_params must hold 3 values (hgcoef, qlc_monoq, qqc_monoq)")
    buf = np.ascontiguousarray(gp)
    if shadow is not None:
        for f in SHADOW_FIELDS:
            if f in shadow:
                set_shadow(domain, f, shadow[f])
    _check(domain._lib, domain._ctx,
           domain._lib.lulesh_vjp_cycle(domain._ctx, buf.ctypes.data_as(_lib.P_double)))
    if grad_params is not None:
        grad_params[:] = buf
    return {f: get_shadow(domain, f) for f in SHADOW_FIELDS}, buf


def total_internal_energy(domain):
    e, m = domain.e, domain.elemMass
    total = 0.0
    for a, b in zip(e.tolist(), m.tolist()):
        total += a * b
    return total


class Batch:
    """Many independent Sedov domains advanced together (no reference
    counterpart: the reference runs one Domain per process).

    Instance i has nx[i] elements per edge, deposited energy
    einit_scale[i] x the reference's einit (dt0 follows from it), initial
    hgcoef[i] and region count nr[i].  All instances share one ragged SoA
    layout; each cycle is one launch per kernel over every instance, with
    per-instance dt, and an instance whose driver loop has ended is skipped."""

    def __init__(self, nx, einit_scale=None, hgcoef=None, nr=None, quirks=QUIRKS_REFERENCE,
                 device=0):
        n = len(nx)
        for name, arr in (("einit_scale", einit_scale), ("hgcoef", hgcoef), ("nr", nr)):
            if arr is not None and len(arr) != n:
                raise ValueError(f"{name} must have one entry per instance ({n})")
        lib = _lib.load()
        self._lib = lib
        nx_a = (ctypes.c_int32 * n)(*[int(v) for v in nx])
        sc_a = (ctypes.c_double * n)(*[float(v) for v in einit_scale]) if einit_scale is not None else None
        hg_a = (ctypes.c_double * n)(*[float(v) for v in hgcoef]) if hgcoef is not None else None
        nr_a = (ctypes.c_int32 * n)(*[int(v) for v in nr]) if nr is not None else None
        ctx = ctypes.c_void_p()
        rc = lib.lulesh_batch_create(n, nx_a, sc_a, hg_a, nr_a, int(quirks), int(device),
                                     ctypes.byref(ctx))
        if rc != 0:
            raise LuleshError(rc, lib.lulesh_error_string(rc).decode())
        self._ctx = ctx
        self._inst = 0
        self.quirks = int(quirks)
        self.n = n
        self.nx = [int(v) for v in nx]
        self.nr = [int(v) for v in nr] if nr is not None else [1] * n
        self.views = [Domain._view(self, i, self.nx[i]) for i in range(n)]

    def instance(self, i):
        return self.views[i]

    def __len__(self):
        return self.n

    def close(self):
        if getattr(self, "_ctx", None):
            # the instance views hold the same context: drop it there too, so
            # a later call through a view fails with an argument error
            # instead of reaching freed memory
            for v in getattr(self, "views", ()):
                object.__setattr__(v, "_ctx", None)
            self._lib.lulesh_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, num_iters, check=True):
        """Driver loop for every instance; returns cycles run per instance.
        An instance that raises a reference error stops while the others run
        on; with check=True the lowest such error is raised afterwards,
        otherwise `errors()` tells which instances stopped that way."""
        out = (ctypes.c_int64 * self.n)()
        rc = self._lib.lulesh_batch_run(self._ctx, int(num_iters), out)
        if check or not (1 <= rc <= 9):
            _check(self._lib, self._ctx, rc)
        return np.array(out[:], dtype=np.int64)

    def errors(self):
        """Per instance: the reference error code that stopped it in the last
        run (0 if none)."""
        return [int(self._lib.lulesh_batch_error(self._ctx, i)) for i in range(self.n)]

    def verify(self, i):
        return verify(self.views[i])

    def autotune(self):
        """Pick the k_nodal z-march depth by timing it on the live state."""
        ch, ms = ctypes.c_int32(), ctypes.c_double()
        _check(self._lib, self._ctx,
               self._lib.lulesh_autotune(self._ctx, ctypes.byref(ch), ctypes.byref(ms)))
        return ch.value, ms.value

    def set_nodal_depth(self, depth):
        _check(self._lib, self._ctx, self._lib.lulesh_set_nodal_depth(self._ctx, int(depth)))

    def set_nodal_tiles(self, lo=0, hi=0):
        _check(self._lib, self._ctx, self._lib.lulesh_set_nodal_tiles(self._ctx, int(lo), int(hi)))

    def set_timing(self, on=True):
        _check(self._lib, self._ctx, self._lib.lulesh_set_timing(self._ctx, 1 if on else 0))

    kernel_times = Domain.kernel_times

    def keep_aux(self, on=True):
        _check(self._lib, self._ctx, self._lib.lulesh_set_keep_aux(self._ctx, 1 if on else 0))

    def set_cost_mode(self, on=True):
        _check(self._lib, self._ctx, self._lib.lulesh_set_cost_mode(self._ctx, 1 if on else 0))


def c5_instances(n=64, seed=0):
    """The ragged Sedov workload: n instances drawn from numpy's PCG64
    default_rng(seed), per instance in this order: s ~ U{16..64},
    einit factor = exp(U[ln 0.5, ln 2]), hgcoef ~ U[1, 5],
    nr ~ U{1, 5, 11, 16}."""
    rng = np.random.default_rng(seed)
    nx, scale, hg, nr = [], [], [], []
    for _ in range(n):
        nx.append(int(rng.integers(16, 65)))
        scale.append(float(np.exp(rng.uniform(np.log(0.5), np.log(2.0)))))
        hg.append(float(rng.uniform(1.0, 5.0)))
        nr.append(int(rng.choice([1, 5, 11, 16])))
    return nx, scale, hg, nr


def _julia_num(fmt, x):
    """x formatted as Julia's @printf writes it: non-finite values as Inf,
    -Inf and NaN, the rest as C's printf."""
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "Inf" if x > 0 else "-Inf"
    return fmt % x


def verifyAndWriteFinalOutput(elapsed_time, domain, nx, numRanks=1, file=None):
    """Prints the reference's final report (src/utils.jl:124-164)."""
    import sys
    f = file or sys.stdout
    cycle = domain.cycle
    e1, maxabs, totabs, maxrel = verify(domain)
    # IEEE division, as Julia's (a run of 0 cycles gives Inf or NaN, not an exception)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.float64(elapsed_time * 1e6) / np.float64(cycle)
        grind1 = float(t / np.float64(nx * nx * nx))
        grind2 = float(t / np.float64(nx * nx * nx * numRanks))
        fom = float(np.float64(1000.0) / np.float64(grind2))
    jl = _julia_num
    print("Run completed:", file=f)
    print("   Problem size        =  %d" % nx, file=f)
    print("   MPI tasks           =  %d" % numRanks, file=f)
    print("   Iteration count     =  %d" % cycle, file=f)
    print("   Final Origin Energy =  %s" % jl("%.6e", e1), file=f)
    print("   Testing Plane 0 of Energy Array on rank 0:", file=f)
    print("        MaxAbsDiff   = %s" % jl("%.6e", maxabs), file=f)
    print("        TotalAbsDiff = %s" % jl("%.6e", totabs), file=f)
    print("        MaxRelDiff   = %s" % jl("%.6e", maxrel), file=f)
    print("\nElapsed time         = %.2f (s)" % elapsed_time, file=f)
    print("Grind time (us/z/c)  =  %s  (per dom)  (%s overall)"
          % (jl("%.8f", grind1), jl("%.8f", elapsed_time)), file=f)
    print("FOM                  = %s (z/s)\n" % jl("%.8f", fom), file=f)
    return {"cycle": cycle, "e1": e1, "MaxAbsDiff": maxabs, "TotalAbsDiff": totabs,
            "MaxRelDiff": maxrel, "elapsed": elapsed_time}


def main(nx, structured=True, num_iters=10, quirks=QUIRKS_REFERENCE, device=0, per_cycle=False,
         file=None):
    """examples/benchmark.jl main() for one rank; the cycle loop runs on the
    device unless per_cycle=True, which steps it from the host exactly as
    the reference's while-loop does."""
    prob = LuleshProblem(num_iters, structured, nx, 1, 1, 1)
    domain = Domain(prob, quirks=quirks, device=device)
    start = _time.time()
    if per_cycle:
        while domain.time < domain.stoptime:
            timeIncrement(domain)
            lagrangeLeapFrog(domain)
            if domain.cycle >= num_iters:
                break
    else:
        run(domain, num_iters)
    elapsed = _time.time() - start
    report = verifyAndWriteFinalOutput(elapsed, domain, nx, 1, file=file)
    domain.close()
    return report
