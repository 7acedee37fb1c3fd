"""The binomial checkpointing schedule of lulesh_grad_run_revolve, checked on
the host (no GPU): the number of cycles the library plans to run equals the
optimum found by exhaustive search over every checkpoint placement.

Search: to reverse n cycles from a saved start state with s further state
buffers free, either (s = 0) re-run from the start up to each cycle in turn,
n(n-1)/2 advancing cycles, or advance l cycles, save that state and reverse
the n-l cycles after it with s-1 buffers, then the l before it with s.
Every cycle is also run once more, keeping its temporaries for its adjoint
(n runs).  The library computes the same optimum in closed form (Griewank's
r*n - C(c+r, c+1)); a wrong repetition count, off-by-one in the snapshot
count or a non-optimal split would change the numbers."""
from functools import lru_cache

import pytest

from lulesh_amd import _lib


@lru_cache(None)
def advances(n, spare):
    if n <= 1:
        return 0
    if spare == 0:
        return n * (n - 1) // 2
    return min(l + advances(n - l, spare - 1) + advances(l, spare) for l in range(1, n))


@pytest.mark.parametrize("snapshots", [1, 2, 3, 4, 6, 9])
def test_planned_cycles_are_the_exhaustive_optimum(snapshots):
    lib = _lib.load()
    for n in range(1, 70):
        c = min(snapshots, n)
        assert lib.lulesh_revolve_cycles(n, snapshots) == n + advances(n, c - 1), (n, snapshots)


def test_known_counts_and_argument_checks():
    lib = _lib.load()
    # one snapshot: every cycle re-run from the start; all snapshots: each
    # cycle twice (first pass, then with its temporaries kept)
    assert lib.lulesh_revolve_cycles(10, 1) == 10 + 45
    assert lib.lulesh_revolve_cycles(10, 10) == 2 * 10 - 1
    assert lib.lulesh_revolve_cycles(10, 50) == 2 * 10 - 1
    assert lib.lulesh_revolve_cycles(1, 1) == 1
    assert lib.lulesh_revolve_cycles(0, 3) == -1
    assert lib.lulesh_revolve_cycles(5, 0) == -1
    # the logarithmic regime: 10,000 cycles with 20 states
    assert lib.lulesh_revolve_cycles(10000, 20) < 6 * 10000
