"""The largest domain the C ABI accepts (s=800: 512,000,000 elements,
513,922,401 nodes, every byte offset of an array just under 4 GiB) through
its first cycle, whose result has a closed form: no node moves (every force
is zero at the Sedov initial state), the origin element keeps its energy
and takes the EOS of it, every other element stays cold (e = p = q = 0,
ss = TINY3, v = 1).  The last element and node are checked with the first,
so the highest offsets are exercised."""
import math

import numpy as np
import pytest

from lulesh_amd import Domain, LuleshError, LuleshProblem, lagrangeLeapFrog, timeIncrement

pytestmark = pytest.mark.gpu

TINY3 = 0.333333e-18  # src/domain.jl:2742-2744


def test_largest_domain_first_cycle_closed_form():
    nx = 800
    print(f"[limits] nx={nx}: create + 1 cycle", flush=True)
    d = Domain(LuleshProblem(1, True, nx, 11, 1, 1))
    E, N = d.numElem, d.numNode
    assert E == nx ** 3 and N == (nx + 1) ** 3
    e0 = d.e
    einit = float(e0[0])
    assert einit > 0.0 and np.count_nonzero(e0) == 1
    del e0
    x0 = d.x
    timeIncrement(d)
    lagrangeLeapFrog(d)
    assert d.cycle == 1
    x1 = d.x
    assert np.array_equal(x0, x1)  # no force anywhere: nothing moves
    del x0, x1
    xd = d.xd
    assert not xd.any()
    del xd
    e = d.e
    assert e[0] == einit and not e[1:].any()
    del e
    c1s = 2.0 / 3.0  # calcPressureForElems at compression 0 (vnew = 1)
    p = d.p
    p_origin = c1s * einit
    assert p[0] == p_origin and not p[1:].any()
    del p
    ss = d.ss
    assert ss[0] == math.sqrt((c1s * einit) + ((1.0 * c1s) * p_origin))
    assert ss[-1] == TINY3 and ss[E // 2] == TINY3 and np.all(ss[1:] == TINY3)
    del ss
    v = d.v
    assert np.all(v == 1.0)
    del v
    assert d.dtcourant == 1.0e20 and d.dthydro == 1.0e20
    d.close()


def test_sizes_beyond_the_limit_are_refused():
    with pytest.raises(LuleshError):
        Domain(LuleshProblem(1, True, 801, 11, 1, 1))
    with pytest.raises(LuleshError):
        Domain(LuleshProblem(1, True, 0, 11, 1, 1))
