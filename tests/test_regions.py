"""Region index sets of Domain(prob) (createRegionIndexSets! and
sortRegions, src/domain.jl:313-465) and the LULESH-2.0 cost mode.

Pins of the oracle (CPU):
* nr = 1 draws nothing, so every array has a closed form: all elements in
  region 1, regElemlist = 1..E, regCSR = [E] (the fill loop leaves end
  offsets), regReps = [10(1+cost)] (r = 1 is not below nr - (nr+15) div 20
  = 1, src/domain.jl:433-439), regSorted = [1];
* for any nr: regElemlist is a permutation of 1..E grouped by sorted region
  in ascending element order, regCSR are the groups' end offsets, regSorted
  is a permutation of 1..nr and the rep of region n, regReps[regSorted[n]],
  is the formula's value for n;
* sortRegions' inner loop stops one pair short (j in 1:numReg-i-1,
  src/domain.jl:317): for nr = 2 nothing is compared, so regReps stays
  [1+cost, 10(1+cost)], and for nr = 11, cost = 1 the second 20 never
  moves off the end: [20, 2, 2, 2, 2, 2, 1, 1, 1, 1, 20] (a full descending
  sort would put both 20s first);
* the generator's starting state is the SHA-256 digest of the seed word
  0 (Julia's make_seed(0) = UInt32[0]) read as four little-endian words.
The draw stream itself (xoshiro256++ as Julia 1.7's Random.seed!(0) sets
it up) is parity unpinned: no Julia run is available to compare with.

GPU: the library's arrays equal the oracle's, single domains and every
instance of a ragged batch; the cost mode leaves every field bit-identical
while its k_qeos does the repeated work (it takes longer)."""
import hashlib
import struct

import numpy as np
import pytest

import oracle as O

NAMES_E = ("regNumList", "regElemlist")
NAMES_R = ("regCSR", "regReps", "regSorted")


def rep_formula(r, nr, cost):
    if r < nr // 2:
        return 1
    if r < nr - (nr + 15) // 20:
        return 1 + cost
    return 10 * (1 + cost)


@pytest.mark.parametrize("nx,cost", [(1, 1), (4, 0), (8, 3)])
def test_oracle_single_region_closed_form(nx, cost):
    od = O.OracleDomain(nx, nr=1, balance=1, cost=cost)
    E = nx ** 3
    assert np.all(od.ifield("regNumList") == 1)
    assert np.array_equal(od.ifield("regElemlist"), np.arange(1, E + 1))
    assert od.ifield("regCSR").tolist() == [E]
    assert od.ifield("regReps").tolist() == [10 * (1 + cost)]
    assert od.ifield("regSorted").tolist() == [1]


def check_invariants(get, E, nr, cost):
    num, elems = get("regNumList"), get("regElemlist")
    csr, reps, srt = get("regCSR"), get("regReps"), get("regSorted")
    assert num.min() >= 1 and num.max() <= nr
    assert sorted(srt.tolist()) == list(range(1, nr + 1))
    assert np.array_equal(np.sort(elems), np.arange(1, E + 1))
    assert csr[-1] == E and np.all(np.diff(csr) >= 0)
    start = 0
    for slot in range(1, nr + 1):
        grp = elems[start:csr[slot - 1]]
        assert np.all(np.diff(grp) > 0)
        assert np.all(srt[num[grp - 1] - 1] == slot)
        start = csr[slot - 1]
    for n in range(1, nr + 1):
        assert reps[srt[n - 1] - 1] == rep_formula(n, nr, cost)


@pytest.mark.parametrize("nx,nr,balance,cost", [(4, 2, 1, 1), (8, 5, 0, 2), (8, 11, 1, 1),
                                                 (13, 16, 2, 1), (6, 32, 1, 0)])
def test_oracle_region_invariants(nx, nr, balance, cost):
    od = O.OracleDomain(nx, nr=nr, balance=balance, cost=cost)
    check_invariants(od.ifield, nx ** 3, nr, cost)


def test_oracle_partial_bubble_sort():
    assert O.OracleDomain(4, nr=2, cost=1).ifield("regReps").tolist() == [2, 20]
    assert O.OracleDomain(4, nr=11, cost=1).ifield("regReps").tolist() == [20, 2, 2, 2, 2, 2, 1, 1,
                                                                             1, 1, 20]


def test_generator_seed_is_sha256_of_seed_word():
    words = struct.unpack("<4Q", hashlib.sha256(struct.pack("<I", 0)).digest())
    assert O.julia_rng_state() == list(words)


def test_oracle_rejects_bad_region_arguments():
    for kw in ({"nr": 0}, {"balance": -1}, {"cost": -1}, {"nr": 16, "balance": 40}):
        with pytest.raises(ValueError):
            O.OracleDomain(2, **kw)


def test_create_rejects_bad_region_arguments():
    """lulesh_create validates nr/balance/cost on the host before touching
    the device (runs without a GPU)."""
    import ctypes

    from lulesh_amd import _lib
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    for nr, bal, cost in ((0, 1, 1), (4097, 1, 1), (11, -1, 1), (11, 1, -1), (16, 40, 1)):
        assert lib.lulesh_create(8, nr, bal, cost, 7, 0, ctypes.byref(ctx)) == 100, (nr, bal, cost)


@pytest.mark.gpu
@pytest.mark.parametrize("nx,nr,balance,cost", [(8, 1, 1, 1), (8, 11, 1, 1), (12, 5, 0, 2),
                                                 (16, 16, 2, 1), (30, 11, 1, 1), (5, 2, 1, 3)])
def test_gpu_regions_match_oracle(nx, nr, balance, cost):
    from lulesh_amd import Domain, LuleshProblem
    print(f"[regions] nx={nx} nr={nr} balance={balance} cost={cost}", flush=True)
    dom = Domain(LuleshProblem(10, True, nx, nr, balance, cost))
    od = O.OracleDomain(nx, nr=nr, balance=balance, cost=cost)
    for n in NAMES_E + NAMES_R:
        assert np.array_equal(dom.region(n), od.ifield(n)), n
    check_invariants(dom.region, nx ** 3, nr, cost)
    dom.close()


@pytest.mark.gpu
def test_gpu_batch_regions_match_oracle():
    from lulesh_amd import Batch, c5_instances
    print("[regions] c5 batch", flush=True)
    nx, scale, hg, nr = c5_instances(12, seed=3)
    b = Batch(nx, scale, hg, nr)
    for i in range(len(nx)):
        od = O.OracleDomain(int(nx[i]), nr=int(nr[i]), balance=1, cost=1)
        for n in NAMES_E + NAMES_R:
            assert np.array_equal(b.instance(i).region(n), od.ifield(n)), (i, n)
    b.close()


@pytest.mark.gpu
def test_gpu_cost_mode_is_bit_identical_and_does_the_work():
    from lulesh_amd import Domain, LuleshProblem, lagrangeLeapFrog, run, timeIncrement
    print("[cost-mode] s=30 stepped, s=48 driver loop", flush=True)
    a = Domain(LuleshProblem(10, True, 30, 11, 1, 1))
    b = Domain(LuleshProblem(10, True, 30, 11, 1, 1))
    b.set_cost_mode(True)
    for _ in range(60):
        for d in (a, b):
            timeIncrement(d)
            lagrangeLeapFrog(d)
    for n in ("e", "p", "q", "v", "ss", "x", "xd"):
        assert np.array_equal(a.get(n).view(np.uint64), b.get(n).view(np.uint64)), n
    a.close()
    b.close()
    # the driver loop (captured graph), timed per kernel
    times = {}
    fields = {}
    for mode in (False, True):
        d = Domain(LuleshProblem(10, True, 48, 11, 1, 1))
        d.set_cost_mode(mode)
        assert run(d, 50) == 50
        d.set_timing(True)
        assert run(d, 100) == 50  # num_iters bounds the cycle count, as in the reference
        times[mode] = d.kernel_times()["k_qeos"][0]
        fields[mode] = {n: d.get(n) for n in ("e", "p", "q", "x", "xd")}
        d.close()
    for n in fields[False]:
        assert np.array_equal(fields[False][n].view(np.uint64), fields[True][n].view(np.uint64)), n
    print(f"k_qeos 50 cycles: {times[False]:.3f} ms, cost mode {times[True]:.3f} ms", flush=True)
    assert times[True] > 1.2 * times[False]
