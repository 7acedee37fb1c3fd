"""Binomial checkpointing (lulesh_grad_run_revolve) against the two-level
checkpointing of lulesh_grad_run: the adjoint kernels run on identical stored
states in the same reverse order, so the loss, the gradients, the dt
schedule and the state left behind agree bit for bit, whatever the number
of snapshots; and the cycles run are the planned optimum."""
import numpy as np
import pytest

from lulesh_amd import Domain, LuleshError, LuleshProblem, _lib, grad_run, grad_run_revolve, run

pytestmark = pytest.mark.gpu

PRE = 20  # start where the blast has spread (every e entry carries a gradient)


def fresh(nx):
    d = Domain(LuleshProblem(10, True, nx, 11, 1, 1))
    run(d, PRE)
    return d


@pytest.mark.parametrize("ncycles,snapshots", [(1, 1), (7, 1), (7, 2), (12, 3), (30, 4),
                                               (30, 30), (30, 100)])
def test_revolve_equals_two_level_checkpointing(ncycles, snapshots):
    print(f"[revolve] n={ncycles} snapshots={snapshots}", flush=True)
    a = fresh(8)
    ref = grad_run(a, ncycles, 5)
    b = fresh(8)
    got = grad_run_revolve(b, ncycles, snapshots)
    assert got["loss"] == ref["loss"]
    assert np.array_equal(got["grad_e"], ref["grad_e"])
    for k in ("grad_hgcoef", "grad_qlc_monoq", "grad_qqc_monoq"):
        assert got[k] == ref[k], k
    assert np.array_equal(got["dt"], ref["dt"])
    assert got["forward_cycles"] == _lib.load().lulesh_revolve_cycles(ncycles, snapshots)
    # both leave the forward pass's end state
    for f in ("e", "p", "q", "v", "ss", "x", "y", "zd"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.cycle == b.cycle and a.time == b.time
    assert np.count_nonzero(got["grad_e"]) > 0
    a.close()
    b.close()


def test_revolve_rejects_bad_arguments():
    d = fresh(4)
    with pytest.raises(LuleshError):
        grad_run_revolve(d, 0, 3)
    with pytest.raises(LuleshError):
        grad_run_revolve(d, 5, 0)
    d.close()
