"""The per-cycle VJP (Enzyme.autodiff(lagrangeLeapFrog, Duplicated(domain,
shadow)) of examples/benchmark.jl:86) against the multi-cycle gradient driver
and against central finite differences of one lagrangeLeapFrog, field by
field."""
import numpy as np
import pytest

from lulesh_amd import (SHADOW_FIELDS, Domain, LuleshProblem, grad_run, lagrangeLeapFrog, run,
                        timeIncrement, vjp_cycle)

pytestmark = pytest.mark.gpu

NODE = ("x", "y", "z", "xd", "yd", "zd")


def fresh(nx, cycles=0):
    d = Domain(LuleshProblem(10, True, nx, 11, 1, 1))
    if cycles:
        run(d, cycles)
    return d


def test_vjp_cycle_equals_one_cycle_grad_run():
    """timeIncrement + one VJP seeded with dL/de = elemMass is the same
    computation as a one-cycle gradient run: bit-identical results."""
    print("[vjp] nx=12 vs grad_run(1)", flush=True)
    a = fresh(12, 25)
    g = grad_run(a, 1, 1)
    b = fresh(12, 25)
    timeIncrement(b)
    seed = {f: np.zeros(b.numNode if f in NODE else b.numElem) for f in SHADOW_FIELDS}
    seed["e"] = b.elemMass.copy()
    gp = np.zeros(3)
    out, _ = vjp_cycle(b, seed, gp)
    assert np.array_equal(out["e"], g["grad_e"])
    assert gp[0] == g["grad_hgcoef"] and gp[1] == g["grad_qlc_monoq"]
    assert gp[2] == g["grad_qqc_monoq"]
    # both leave the forward state after the cycle
    for f in ("e", "p", "x", "xd", "v", "ss"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.cycle == b.cycle and a.time == b.time
    a.close()
    b.close()


def _weights(d, rng):
    """A dense seed: the adjoint of the state after the cycle."""
    w = {}
    for f in SHADOW_FIELDS:
        n = d.numNode if f in NODE else d.numElem
        w[f] = rng.standard_normal(n)
    w["e"] *= d.elemMass  # O(1) contributions from every field
    return w


def _loss(d, w):
    total = 0.0
    for f in SHADOW_FIELDS:
        total += float(np.dot(w[f], d.get(f)))
    return total


def _loss_after(nx, cycles, dt, w, field, idx, step):
    d = fresh(nx, cycles)
    timeIncrement(d)
    assert d.deltatime == dt
    a = d.get(field)
    a[idx] += step
    d.set(field, a)
    lagrangeLeapFrog(d)
    L = _loss(d, w)
    d.close()
    return L


@pytest.mark.parametrize("field", SHADOW_FIELDS)
def test_vjp_cycle_matches_finite_differences(field):
    nx, cycles = 8, 30
    print(f"[vjp] FD field {field}", flush=True)
    d = fresh(nx, cycles)
    timeIncrement(d)
    dt = d.deltatime
    rng = np.random.default_rng(7)
    w = _weights(d, rng)
    base = d.get(field)
    d_e, d_xd, d_yd, d_zd = d.e, d.xd, d.yd, d.zd
    out, _ = vjp_cycle(d, w)
    d.close()
    grad = out[field]
    n = base.size
    # Entries where the blast has arrived.  Ahead of it the state sits on
    # flat spots of the cycle map -- the EOS energy cut, the u_cut velocity
    # snap and the v_cut volume snap hold e, xd and v at exactly 0 or 1 --
    # where the local derivative is exactly 0 while a finite step crosses the
    # cut; those entries are not finite-difference checkable.
    if field in ("xd", "yd", "zd"):
        cand = np.flatnonzero(np.abs(base) > 1e-3 * np.abs(base).max())
    elif field in NODE:
        sp = np.sqrt(d_xd ** 2 + d_yd ** 2 + d_zd ** 2)
        cand = np.flatnonzero(sp > 1e-3 * sp.max())
    else:
        cand = np.flatnonzero(np.abs(d_e) > 1e-3 * np.abs(d_e).max())
    assert cand.size >= 6
    idx = sorted(set([int(i) for i in (0, 1, 2, nx + 1, nx * nx) if i in set(cand.tolist())]
                     + [int(i) for i in rng.choice(cand, 4, replace=False)]))
    scale = {"x": 1.0, "y": 1.0, "z": 1.0, "xd": max(1.0, float(np.abs(base).max())),
             "yd": max(1.0, float(np.abs(base).max())), "zd": max(1.0, float(np.abs(base).max())),
             "v": 1.0}.get(field, max(1e-12, float(np.abs(base).max())))
    for i in idx:
        h = 1e-6 * max(abs(float(base[i])), scale)
        fd = (_loss_after(nx, cycles, dt, w, field, i, h)
              - _loss_after(nx, cycles, dt, w, field, i, -h)) / (2 * h)
        tol = 1e-5 * abs(fd) + 1e-7 * float(np.abs(grad).max())
        assert abs(grad[i] - fd) <= tol, (field, i, grad[i], fd)


def test_vjp_params_match_finite_differences():
    nx, cycles = 8, 30
    d = fresh(nx, cycles)
    timeIncrement(d)
    dt = d.deltatime
    rng = np.random.default_rng(11)
    w = _weights(d, rng)
    gp = np.zeros(3)
    vjp_cycle(d, w, gp)
    d.close()
    for k, name in enumerate(("hgcoef", "qlc_monoq", "qqc_monoq")):
        vals = []
        for sgn in (1, -1):
            e = fresh(nx, cycles)
            timeIncrement(e)
            assert e.deltatime == dt
            p0 = e.get(name)
            h = 1e-6 * p0
            e.set(name, p0 + sgn * h)
            lagrangeLeapFrog(e)
            vals.append(_loss(e, w))
            e.close()
        fd = (vals[0] - vals[1]) / (2 * h)
        assert gp[k] == pytest.approx(fd, rel=1e-5, abs=1e-9 * max(1.0, abs(fd))), (name, gp[k], fd)
