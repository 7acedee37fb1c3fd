"""Ragged batches on the GPU: every instance of a batch must end bit-identical
to the same domain run alone (and to the oracle), whatever its size, dt
history or finishing cycle."""
import math

import numpy as np
import pytest

import oracle as O
from lulesh_amd import Batch, c5_instances, verify

pytestmark = pytest.mark.gpu

FIELDS = ("e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd", "delv", "vdov", "arealg")
SCAL = ("time", "deltatime", "cycle", "dtcourant", "dthydro", "hgcoef")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def same(da, db, tag):
    for f in FIELDS:
        a, b = da.get(f), db.get(f)
        assert np.array_equal(bits(a), bits(b)), f"{tag}: field {f}"
    for s in SCAL:
        assert bits([da.get(s)])[0] == bits([db.get(s)])[0], f"{tag}: scalar {s}"


def oracle_instance(nx, scale, hg, nthreads=4):
    """Oracle domain with the batch's per-instance deposit and hgcoef: e[1] =
    scale * einit and dt0 = (0.5*cbrt(volo[1]))/sqrt(2*e[1])."""
    od = O.OracleDomain(nx, O.QUIRKS_REFERENCE, nthreads)
    sc = nx / 45.0
    einit = (((3.948746e+7 * sc) * sc) * sc) * scale
    e = od.field("e")
    e[0] = einit
    od.set_field("e", e)
    od.set("deltatime", (0.5 * O.cbrt(od.field("volo")[0])) / math.sqrt(2.0 * einit))
    od.set("hgcoef", hg)
    return od


NX = [8, 12, 5, 16, 9]
SCALE = [1.0, 0.7, 1.5, 1.2, 0.5]
HG = [3.0, 1.5, 4.2, 2.0, 3.3]


def test_batch_instances_equal_lone_runs_and_oracle():
    print("[batch] 5 instances, 120 cycles", flush=True)
    b = Batch(NX, SCALE, HG, [11, 1, 5, 16, 11])
    cyc = b.run(120)
    # s=5 with 1.5x the deposit reaches the stop time at cycle 91 (oracle too)
    assert list(cyc) == [120, 120, 91, 120, 120]
    for i, nx in enumerate(NX):
        lone = Batch([nx], [SCALE[i]], [HG[i]])
        assert lone.run(120)[0] == cyc[i]
        same(b.instance(i), lone.instance(0), f"instance {i} vs lone")
        od = oracle_instance(nx, SCALE[i], HG[i])
        assert od.run(120)[0] == 0
        assert od.get("cycle") == cyc[i]
        for f in ("e", "p", "q", "v", "ss", "x", "xd"):
            assert np.array_equal(bits(b.instance(i).get(f)), bits(od.field(f))), (i, f)
        lone.close()
    b.close()


def test_batch_instances_finish_at_their_own_cycle():
    """Run to stop time: instances end at different cycles (188 and 352 for
    the reference problem at s=8 and s=12); finished ones are skipped and
    keep exactly the state a lone run ends with."""
    print("[batch] run to stoptime", flush=True)
    b = Batch([12, 8])
    cyc = b.run(10 ** 9)
    assert list(cyc) == [352, 188]
    for i, nx in enumerate([12, 8]):
        lone = Batch([nx])
        lone.run(10 ** 9)
        same(b.instance(i), lone.instance(0), f"to-stop instance {i}")
        assert verify(b.instance(i)) == verify(lone.instance(0))
        lone.close()
    b.close()


def test_c5_batch_matches_lone_runs():
    nx, scale, hg, nr = c5_instances(64, seed=0)
    print(f"[c5] 64 instances, {sum(s ** 3 for s in nx)} elements, 40 cycles", flush=True)
    b = Batch(nx, scale, hg, nr)
    assert list(b.run(40)) == [40] * 64
    for i in (0, 7, 31, 63):
        lone = Batch([nx[i]], [scale[i]], [hg[i]], [nr[i]])
        lone.run(40)
        same(b.instance(i), lone.instance(0), f"c5 instance {i}")
        lone.close()
    b.close()


def test_autotune_leaves_state_unchanged_and_results_identical():
    print("[autotune]", flush=True)
    b = Batch(NX, SCALE, HG)
    b.instance(0).keep_aux(True)
    b.run(30)
    aux = ("fx", "fy", "fz", "xdd", "ydd", "zdd", "ql", "qq", "dxx")
    before = {(i, f): b.instance(i).get(f) for i in range(len(NX)) for f in FIELDS + aux}
    depth, ms = b.autotune()
    assert depth >= 1 and ms > 0  # instance 0's depth: the model's or a fixed one
    for i in range(len(NX)):  # a tile of one of the tuned width ranges
        etx, ety, tx, ty = b.instance(i).nodal_tile()
        assert 8 <= etx <= 32 and ety == 256 // etx
    for (i, f), a in before.items():
        assert np.array_equal(bits(a), bits(b.instance(i).get(f))), (i, f)
    b.run(60)
    ref = Batch(NX, SCALE, HG)
    ref.set_nodal_depth(2 if depth != 2 else 16)
    ref.run(60)
    for i in range(len(NX)):
        same(b.instance(i), ref.instance(i), f"depth-independence {i}")
    b.close()
    ref.close()


def test_autotune_keeps_error_codes():
    """An instance stopped by a reference error before the tuning is still
    reported so after it, and the others stay clean."""
    print("[autotune-errors]", flush=True)
    b = Batch(NX, SCALE, HG)
    b.instance(2).set("qstop", -1.0)  # q = 0 > -1 at the first cycle
    b.run(5, check=False)
    assert b.errors() == [0, 0, 7, 0, 0]
    b.autotune()
    assert b.errors() == [0, 0, 7, 0, 0]
    b.close()


@pytest.mark.parametrize("lo,hi", [(16, 16), (12, 21), (8, 32), (24, 32), (2, 128)])
def test_tile_widths_do_not_change_results(lo, hi):
    """k_nodal with other tile shapes (down to 2-wide tiles and up to 128
    wide) ends bit-identical to the default shapes: every node sums its
    corners in the same order whatever the tiling."""
    print(f"[tiles] widths {lo}..{hi}", flush=True)
    b = Batch(NX, SCALE, HG)
    b.set_nodal_tiles(lo, hi)
    for i in range(len(NX)):
        etx, ety, tx, ty = b.instance(i).nodal_tile()
        assert lo <= etx <= hi and ety == 256 // etx and tx >= 1 and ty >= 1
    b.run(40)
    ref = Batch(NX, SCALE, HG)
    ref.run(40)
    for i in range(len(NX)):
        same(b.instance(i), ref.instance(i), f"tiles {lo}..{hi} instance {i}")
    b.close()
    ref.close()


def test_tile_width_arguments():
    b = Batch([8])
    for lo, hi in ((1, 16), (20, 16), (16, 129)):
        with pytest.raises(Exception):
            b.set_nodal_tiles(lo, hi)
    b.set_nodal_tiles(0, 0)
    assert b.instance(0).nodal_tile()[0] >= 16
    b.close()


def test_error_stops_only_its_instance():
    """An invalid elemBC mask in instance 1 raises the reference's "Error"
    there (code 6) in the first cycle; the other instances run on and end
    exactly as lone runs."""
    print("[batch] per-instance error", flush=True)
    b = Batch([8, 10, 12])
    bc = b.instance(1).get("elemBC")
    bc[5] |= 0x1 | 0x2  # XI_M_SYMM | XI_M_FREE: no valid case
    b.instance(1).set("elemBC", bc)
    cyc = b.run(50, check=False)
    assert b.errors() == [0, 6, 0]
    assert cyc[0] == 50 and cyc[2] == 50 and cyc[1] == 1
    for i, nx in ((0, 8), (2, 12)):
        lone = Batch([nx])
        lone.run(50)
        same(b.instance(i), lone.instance(0), f"instance {i} next to a failing one")
        lone.close()
    with pytest.raises(Exception) as ei:
        b2 = Batch([8, 10])
        bc = b2.instance(0).get("elemBC")
        bc[5] |= 0x1 | 0x2
        b2.instance(0).set("elemBC", bc)
        b2.run(5)
    assert "Error" in str(ei.value)
    b.close()


def test_closed_batch_and_domain_refuse_calls():
    """After close() the instance views and the domain raise LuleshError
    (the library's null-context check) rather than reach freed memory."""
    from lulesh_amd import Domain, LuleshError, LuleshProblem, lagrangeLeapFrog, timeIncrement
    print("[close] views after close", flush=True)
    b = Batch([4, 5])
    v = b.instance(1)
    assert v.get("e").size == 125
    b.close()
    with pytest.raises(LuleshError):
        v.get("e")
    with pytest.raises(LuleshError):
        timeIncrement(v)
    b.close()  # a second close is a no-op
    d = Domain(LuleshProblem(10, True, 4, 11, 1, 1))
    d.close()
    with pytest.raises(LuleshError):
        lagrangeLeapFrog(d)
    with pytest.raises(LuleshError):
        d.get("x")
