"""Gradients from the hand-written adjoint kernels against central finite
differences of the forward path with the same (recorded) dt schedule."""
import numpy as np
import pytest

from lulesh_amd import Domain, LuleshProblem, grad_run, replay, run, total_internal_energy

pytestmark = pytest.mark.gpu


# Gradients are taken from the state after PRE cycles, when the blast has
# spread energy over many elements.  At the Sedov initial state every element
# but the first has e = 0, and the EOS energy cut (|e_new| < e_cut -> 0,
# src/domain.jl:2661-2663) makes the loss locally flat in those entries: the
# local derivative there is exactly 0, while a finite difference with a step
# far above e_cut sees through the cut.
PRE = 20


def fresh(nx, pre=0):
    d = Domain(LuleshProblem(10, True, nx, 11, 1, 1))
    if pre:
        run(d, pre)
    return d


def loss_with(nx, dt, e_index=None, e_step=0.0, scalar=None, s_step=0.0):
    d = fresh(nx, PRE)
    if e_index is not None:
        e = d.e
        e[e_index] += e_step
        d.e = e
    if scalar is not None:
        d.set(scalar, d.get(scalar) + s_step)
    replay(d, dt)
    L = total_internal_energy(d)
    d.close()
    return L


def fd_e(nx, dt, i, h):
    return (loss_with(nx, dt, i, h) - loss_with(nx, dt, i, -h)) / (2 * h)


def fd_scalar(nx, dt, name, h):
    return (loss_with(nx, dt, scalar=name, s_step=h) - loss_with(nx, dt, scalar=name, s_step=-h)) / (2 * h)


def test_gradient_at_the_initial_state_is_local():
    """From the Sedov initial state only e[1] carries energy; every other
    entry of e sits inside the energy cut after cycle 1 (all forces are zero
    in cycle 1, so e_new = e_old there) and its local derivative is 0."""
    d = fresh(8)
    g = grad_run(d, 10, 5)
    assert g["grad_e"][0] != 0.0
    assert np.count_nonzero(g["grad_e"][1:]) == 0
    # a step below e_cut (1e-7) leaves the loss unchanged, as the adjoint says
    dt = g["dt"]
    assert loss_with_pre(8, dt, 0, 5, 5e-8) == loss_with_pre(8, dt, 0, 5, -5e-8)
    d.close()


def loss_with_pre(nx, dt, pre, i, step):
    d = fresh(nx, pre)
    e = d.e
    e[i] += step
    d.e = e
    replay(d, dt)
    L = total_internal_energy(d)
    d.close()
    return L


def test_gradient_checkpoint_independent_and_forward_state():
    print("[adjoint] nx=12 ckpt independence", flush=True)
    outs = []
    for every in (1, 7, 60):
        d = fresh(12)
        outs.append(grad_run(d, 60, every))
        ref = fresh(12)
        run(ref, 60)
        for f in ("e", "p", "x", "xd", "ss", "v"):
            assert np.array_equal(d.get(f), ref.get(f)), f  # left in the end state
        assert d.cycle == 60
        d.close()
        ref.close()
    for o in outs[1:]:
        assert np.array_equal(o["grad_e"], outs[0]["grad_e"])
        assert o["grad_hgcoef"] == outs[0]["grad_hgcoef"]
        assert o["loss"] == outs[0]["loss"] and np.array_equal(o["dt"], outs[0]["dt"])


def test_replay_reproduces_forward():
    d = fresh(12)
    g = grad_run(d, 40, 10)
    r = fresh(12)
    replay(r, g["dt"])
    assert total_internal_energy(r) == g["loss"]
    assert np.array_equal(r.e, d.e)


@pytest.mark.parametrize("nx,ncyc", [(8, 40), (12, 80)])
def test_gradient_matches_finite_differences_small(nx, ncyc):
    print(f"[adjoint] FD nx={nx} cycles={ncyc}", flush=True)
    d = fresh(nx, PRE)
    e0 = d.e
    g = grad_run(d, ncyc, 9)
    dt = g["dt"]
    rng = np.random.default_rng(0)
    hot = np.flatnonzero(np.abs(e0) > 1e-3 * np.abs(e0).max())
    assert hot.size >= 8
    idx = sorted(set([0, 1, nx, nx * nx] + [int(v) for v in rng.choice(hot, 8, replace=False)]))
    for i in idx:
        if abs(e0[i]) <= 1e-3 * np.abs(e0).max():
            continue
        h = 1e-6 * abs(e0[i])
        fd = fd_e(nx, dt, int(i), h)
        ad = g["grad_e"][i]
        assert ad == pytest.approx(fd, rel=1e-4, abs=1e-7 * abs(g["grad_e"]).max()), (i, ad, fd)
    for name, key in (("hgcoef", "grad_hgcoef"), ("qlc_monoq", "grad_qlc_monoq"),
                      ("qqc_monoq", "grad_qqc_monoq")):
        base = fresh(nx).get(name)
        fd = fd_scalar(nx, dt, name, 1e-6 * base)
        assert g[key] == pytest.approx(fd, rel=1e-4, abs=1e-9 * abs(g["loss"])), (name, g[key], fd)
    d.close()


def test_gradient_quirk_free_matches_finite_differences():
    """The LULESH 2.0 arithmetic (quirks cleared) runs the other compiled
    variant of the force adjoint; same finite-difference check on hgcoef
    and a few e entries behind the blast front."""
    print("[adjoint] quirk-free FD", flush=True)
    nx, ncyc = 8, 30

    def make():
        d = Domain(LuleshProblem(10, True, nx, 11, 1, 1), quirks=0)
        run(d, PRE)
        return d

    d = make()
    e0 = d.e
    g = grad_run(d, ncyc, 7)
    d.close()
    dt = g["dt"]

    def loss(mod):
        x = make()
        mod(x)
        replay(x, dt)
        L = total_internal_energy(x)
        x.close()
        return L

    hot = np.flatnonzero(np.abs(e0) > 1e-3 * np.abs(e0).max())
    for i in [int(hot[0]), int(hot[len(hot) // 2]), int(hot[-1])]:
        h = 1e-6 * abs(e0[i])

        def bump(step):
            def f(x):
                e = x.e
                e[i] += step
                x.e = e
            return f
        fd = (loss(bump(h)) - loss(bump(-h))) / (2 * h)
        assert g["grad_e"][i] == pytest.approx(fd, rel=1e-4, abs=1e-7 * abs(g["grad_e"]).max()), i
    hg = 3.0
    fd = (loss(lambda x: x.set("hgcoef", hg * (1 + 1e-6))) -
          loss(lambda x: x.set("hgcoef", hg * (1 - 1e-6)))) / (2e-6 * hg)
    assert g["grad_hgcoef"] == pytest.approx(fd, rel=1e-4, abs=1e-9 * abs(g["loss"]))


def test_c4_configuration_gradient():
    """The adjoint benchmark's own configuration: s=48, 200 cycles from the
    Sedov initial state, checkpoints every 20, loss = total internal energy.
    d loss / d e[1] and d loss / d {hgcoef, qlc_monoq, qqc_monoq} against
    central finite differences with a 1e-6 relative step; the 15 other
    elements drawn by numpy's PCG64 default_rng(0) start at e = 0, inside
    the EOS energy cut, where the adjoint is exactly 0 and a step below the
    cut (5e-8 < e_cut = 1e-7) leaves the loss unchanged."""
    nx, ncyc = 48, 200
    print("[adjoint] C4: s=48, 200 cycles, checkpoint 20", flush=True)
    d = fresh(nx)
    e1 = float(d.e[0])
    g = grad_run(d, ncyc, 20)
    d.close()
    dt = g["dt"]

    def loss(mod):
        x = fresh(nx)
        mod(x)
        replay(x, dt)
        L = total_internal_energy(x)
        x.close()
        return L

    def bump(i, step):
        def f(x):
            e = x.e
            e[i] += step
            x.e = e
        return f

    h = 1e-6 * e1
    fd = (loss(bump(0, h)) - loss(bump(0, -h))) / (2 * h)
    assert g["grad_e"][0] == pytest.approx(fd, rel=1e-4), (g["grad_e"][0], fd)
    for name, key in (("hgcoef", "grad_hgcoef"), ("qlc_monoq", "grad_qlc_monoq"),
                      ("qqc_monoq", "grad_qqc_monoq")):
        p0 = fresh(nx).get(name)
        hp = 1e-6 * p0
        fd = (loss(lambda x: x.set(name, p0 + hp)) - loss(lambda x: x.set(name, p0 - hp))) / (2 * hp)
        assert g[key] == pytest.approx(fd, rel=1e-4, abs=1e-9 * abs(g["loss"])), (name, g[key], fd)
    rng = np.random.default_rng(0)
    others = [int(v) for v in rng.choice(np.arange(1, nx ** 3), 15, replace=False)]
    for i in others:
        assert g["grad_e"][i] == 0.0, i
        assert loss(bump(i, 5e-8)) == loss(bump(i, -5e-8)), i


def test_c4_size_hot_elements_match_finite_differences():
    """At the adjoint benchmark's size (s=48, checkpoints every 20) from a
    state in which the blast has spread (after 100 cycles): d loss / d e of
    element 1 and of 8 hot elements drawn by numpy's PCG64 default_rng(1),
    over 100 more cycles, against central finite differences with a 1e-6
    relative step."""
    nx, pre, ncyc = 48, 100, 100
    print("[adjoint] s=48 hot elements", flush=True)
    d = fresh(nx, pre)
    e0 = d.e
    g = grad_run(d, ncyc, 20)
    d.close()
    dt = g["dt"]
    hot = np.flatnonzero(np.abs(e0) > 1e-3 * np.abs(e0).max())
    assert hot.size >= 8
    rng = np.random.default_rng(1)
    idx = sorted(set([0] + [int(v) for v in rng.choice(hot, 8, replace=False)]))
    scale = np.abs(g["grad_e"]).max()
    for i in idx:
        h = 1e-6 * abs(e0[i])
        fd = (loss_with_pre(nx, dt, pre, i, h) - loss_with_pre(nx, dt, pre, i, -h)) / (2 * h)
        assert g["grad_e"][i] == pytest.approx(fd, rel=1e-4, abs=1e-7 * scale), (i, g["grad_e"][i], fd)
