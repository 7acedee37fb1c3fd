"""Closed-form pins of three pieces that a GPU == oracle comparison alone
cannot pin, because the product and the oracle could share a slip:

* the plane-0 symmetry numbers of verifyAndWriteFinalOutput
  (src/utils.jl:134-152): hand-computed MaxAbsDiff / TotalAbsDiff /
  MaxRelDiff of an uploaded asymmetric energy plane, where dividing by the
  other member of a pair, taking abs() of the divisor or reading another
  plane each gives a different answer;
* the face-neighbour lists of setupElementConnectivities
  (src/domain.jl:793-820), restated below loop by loop in 1-based form,
  including the odd entries lxim[1] = 1, the last row of letap (it wraps to
  row 0) and the last plane of lzetap (itself);
* the neighbour reads of calcMonotonicQRegionForElems
  (src/domain.jl:2339-2350 and the eta/zeta cases after it), through a
  velocity field with exact binary values on a cube of edge 1/8: uniform
  compression on one side of a layer and rest on the other.  At the last
  compressed layer the limiter sees a resting neighbour (phi = 0, so
  ql = 2^-7 and qq = (2/3) 2^-12) unless the read lands on the element
  itself -- which is exactly quirk Q3 for xi-plus (src/domain.jl:2349:
  delv_xi[lxip[i]] with the 0-based id used as a 1-based index), giving
  phi = 1 and ql = qq = 0.

Each closed form runs on the oracle (CPU) and, marked gpu, through the
C ABI on the HIP path.
"""
import numpy as np
import pytest

import oracle as O

# ---------------------------------------------------------------------------
# verify
# ---------------------------------------------------------------------------

S_VER = 4


def verify_plane(with_inf):
    """Energy field of an s=4 domain: plane 0 holds the pairs below, the
    other planes hold large values that a read of the wrong plane would
    pick up.  (row, col) of plane 0 is element row*4 + col; reference pairs
    are e[(j-1)s+k] (row j-1, col k-1, upper triangle) against its mirror
    e[(k-1)s+j] (lower triangle), which is the divisor."""
    s = S_VER
    e = 1000.0 + np.arange(s ** 3, dtype=np.float64)
    e[:s * s] = 0.0
    P = e[:s * s].reshape(s, s)  # view: P[row, col]
    P[0, 0] = 7.0                 # e[1], reported as is
    P[0, 1], P[1, 0] = 3.0, 1.0   # |3-1| = 2,    2/1 = 2
    P[0, 2], P[2, 0] = 1.0, 4.0   # |1-4| = 3,    3/4 = 0.75
    P[1, 2], P[2, 1] = 1.0, -0.5  # 1.5,          1.5/-0.5 = -3
    P[2, 3], P[3, 2] = 0.25, 0.125  # 0.125,      0.125/0.125 = 1
    # (0,3)/(3,0) and (1,3)/(3,1) are 0/0 = NaN: never the maximum
    if with_inf:
        P[1, 3], P[3, 1] = 5.0, 0.0  # 5/0 = Inf is the maximum
    return e


# {e[1], MaxAbsDiff, TotalAbsDiff, MaxRelDiff}; all sums exact in binary
VERIFY_FINITE = (7.0, 3.0, 6.625, 2.0)
VERIFY_INF = (7.0, 5.0, 11.625, np.inf)


@pytest.mark.parametrize("with_inf", [False, True])
def test_oracle_verify_closed_form(with_inf):
    od = O.OracleDomain(S_VER)
    od.set_field("e", verify_plane(with_inf))
    assert od.verify() == (VERIFY_INF if with_inf else VERIFY_FINITE)


def test_verify_closed_form_discriminates():
    """The hand values differ from what the plausible slips give: the
    other member of each pair as divisor (MaxRel 3), abs() of the divisor
    (MaxRel 3), or the mirror pairs of plane 1."""
    e = verify_plane(False)
    P = e[:16].reshape(4, 4)
    rel_swapped = max(abs(P[r, c] - P[c, r]) / P[r, c] for r in range(4) for c in range(r + 1, 4)
                      if P[r, c] != 0)
    rel_abs = max(abs(P[r, c] - P[c, r]) / abs(P[c, r]) for r in range(4) for c in range(r + 1, 4)
                  if P[c, r] != 0)
    assert rel_swapped == 3.0 and rel_abs == 3.0 != VERIFY_FINITE[3]
    Q = e[16:32].reshape(4, 4)
    assert max(abs(Q[r, c] - Q[c, r]) for r in range(4) for c in range(r + 1, 4)) != VERIFY_FINITE[1]


@pytest.mark.gpu
@pytest.mark.parametrize("with_inf", [False, True])
def test_gpu_verify_closed_form(with_inf):
    from lulesh_amd import Domain, LuleshProblem, verify
    print(f"[verify-pin] s=4 with_inf={with_inf}", flush=True)
    dom = Domain(LuleshProblem(10, True, S_VER, 11, 1, 1))
    dom.set("e", verify_plane(with_inf))
    assert verify(dom) == (VERIFY_INF if with_inf else VERIFY_FINITE)
    dom.close()


@pytest.mark.gpu
def test_gpu_batch_verify_closed_form_second_instance():
    """lulesh_batch_verify on instance 1 of a ragged batch reads that
    instance's plane 0 (its element offset), not instance 0's."""
    from lulesh_amd import Batch, verify
    print("[verify-pin] batch (6, 4)", flush=True)
    b = Batch([6, S_VER])
    d1 = b.instance(1)
    d1.set("e", verify_plane(False))
    assert verify(d1) == VERIFY_FINITE
    assert verify(b.instance(0))[1:] == (0.0, 0.0, 0.0)
    b.close()


# ---------------------------------------------------------------------------
# face-neighbour lists
# ---------------------------------------------------------------------------

def connectivities_1based(s):
    """setupElementConnectivities (src/domain.jl:793-820) restated with
    1-based arrays (index 0 unused); returns the six lists 0-based in
    memory, holding the stored values."""
    E = s ** 3
    lxim, lxip, letam, letap, lzetam, lzetap = ([None] * (E + 1) for _ in range(6))
    lxim[1] = 1
    for i in range(1, E):
        lxim[i + 1] = i - 1
        lxip[i] = i
    lxip[E] = E - 1
    for i in range(0, s):
        letam[i + 1] = i
        letap[E - s + i + 1] = i
    for i in range(s, E):
        letam[i + 1] = i - s
        letap[i - s + 1] = i
    for i in range(0, s * s):
        lzetam[i + 1] = i
        lzetap[E - s * s + i + 1] = E - s * s + i
    for i in range(s * s, E):
        lzetam[i + 1] = i - s * s
        lzetap[i - s * s + 1] = i
    return {k: np.array(v[1:], dtype=np.int64) for k, v in
            (("lxim", lxim), ("lxip", lxip), ("letam", letam), ("letap", letap),
             ("lzetam", lzetam), ("lzetap", lzetap))}


@pytest.mark.parametrize("nx", [1, 2, 3, 5, 8])
def test_oracle_neighbour_lists(nx):
    od = O.OracleDomain(nx)
    ref = connectivities_1based(nx)
    for name, want in ref.items():
        assert np.array_equal(od.ifield(name), want), name


def test_neighbour_list_odd_entries():
    """The entries that differ from a plain face neighbour, spelled out."""
    s = 3
    E = s ** 3
    ref = connectivities_1based(s)
    assert ref["lxim"][0] == 1               # element 1's xi-minus is element 2
    assert ref["lxip"][E - 1] == E - 1
    last_row = np.arange(E - s, E)
    assert np.array_equal(ref["letap"][last_row], np.arange(s))        # wraps to row 0
    last_plane = np.arange(E - s * s, E)
    assert np.array_equal(ref["lzetap"][last_plane], last_plane)       # itself
    assert np.array_equal(ref["letam"][:s], np.arange(s))               # itself
    assert np.array_equal(ref["lzetam"][:s * s], np.arange(s * s))      # itself


# ---------------------------------------------------------------------------
# MonoQ neighbour reads
# ---------------------------------------------------------------------------

S_Q = 8
H = 0.125            # edge of the cube's elements: every quantity below is exact
DU = -2.0 ** -6      # velocity jump across a compressed element
M = 4                # layer where compression stops / starts
QL_LAYER = 2.0 ** -7                  # -qlc * rho * (delv * delx) = -0.5 * (-2^-6)
QQ_LAYER = (2.0 / 3.0) * 2.0 ** -12   # qqc * rho * (delv * delx)^2
AXES = {"xi": 0, "eta": 1, "zeta": 2}
VEL = ("xd", "yd", "zd")


def exact_cube(s):
    n = np.arange((s + 1) ** 3)
    s1 = s + 1
    return {"x": (n % s1) * H, "y": ((n // s1) % s1) * H, "z": (n // (s1 * s1)) * H}


def layer_state(axis, side, s=S_Q):
    """Inputs of lagrangeElements: the exact cube, element volume 2^-9 in
    volo and elemMass, v = 1, no energy/pressure/q, and velocity along
    `axis` with jump DU across the compressed elements.  side "plus":
    elements 0..M-1 along the axis are compressed (node t moves with
    DU*min(t, M)), the target layer M-1 has a resting plus-neighbour.
    side "minus": elements M..s-1 are compressed (DU*max(t-M, 0)), the
    target layer M has a resting minus-neighbour."""
    s1 = s + 1
    n = np.arange(s1 ** 3)
    t = ((n % s1), (n // s1) % s1, n // (s1 * s1))[AXES[axis]]
    u = DU * (np.minimum(t, M) if side == "plus" else np.maximum(t - M, 0))
    st = dict(exact_cube(s))
    for k, name in enumerate(VEL):
        st[name] = u.astype(np.float64) if k == AXES[axis] else np.zeros(s1 ** 3)
    E = s ** 3
    st.update(volo=np.full(E, 2.0 ** -9), elemMass=np.full(E, 2.0 ** -9), v=np.ones(E),
              e=np.zeros(E), p=np.zeros(E), q=np.zeros(E))
    return st


def expected_q(axis, side, q3, s=S_Q):
    """Closed forms of ql, qq over the domain.  Only a compressed element
    whose limiter sees a zero ratio on one side is nonzero:
      plus  -- layer M-1 (resting plus-neighbour), except xi under Q3, where
               the xi-plus read is the element's own delv_xi (phi = 1);
      minus -- layer M (resting minus-neighbour) and layer s-1, whose plus
               face is FREE (delvp = 0).
    Everywhere else the compressed neighbours give phi = 1 (so no q) or the
    element is at rest (no q)."""
    E = s ** 3
    e = np.arange(E)
    c = ((e % s), (e // s) % s, e // (s * s))[AXES[axis]]
    if side == "plus":
        hot = (c == M - 1) & ~(q3 and axis == "xi")
    else:
        hot = (c == M) | (c == s - 1)
    ql = np.where(hot, QL_LAYER, 0.0)
    qq = np.where(hot, QQ_LAYER, 0.0)
    return ql, qq, c


def check_layer(get, axis, side, q3, s=S_Q):
    ql_want, qq_want, c = expected_q(axis, side, q3, s)
    compressed = (c < M) if side == "plus" else (c >= M)
    dv_name = "delv_" + axis
    assert np.all(get("vnew") == 1.0)
    assert np.array_equal(get(dv_name), np.where(compressed, DU / H, 0.0))
    assert np.all(get("delx_" + axis) == H)
    for other in AXES:
        if other != axis:
            assert not get("delv_" + other).any()
    ql, qq = get("ql"), get("qq")
    bad = np.nonzero(ql != ql_want)[0]
    assert bad.size == 0, (axis, side, q3, bad[:8], ql[bad[:8]], ql_want[bad[:8]])
    bad = np.nonzero(qq != qq_want)[0]
    assert bad.size == 0, (axis, side, q3, bad[:8], qq[bad[:8]], qq_want[bad[:8]])


CASES = [(a, sd, q) for a in AXES for sd in ("plus", "minus") for q in (0, O.Q3)]


@pytest.mark.parametrize("axis,side,quirks", CASES)
def test_oracle_monoq_neighbour_reads(axis, side, quirks):
    od = O.OracleDomain(S_Q, quirks)
    for name, val in layer_state(axis, side).items():
        od.set_field(name, val)
    od.set("deltatime", 2.0 ** -10)
    assert od.elements() == 0
    check_layer(od.field, axis, side, bool(quirks & O.Q3))


@pytest.mark.gpu
@pytest.mark.parametrize("axis,side,quirks", CASES)
def test_gpu_monoq_neighbour_reads(axis, side, quirks):
    from lulesh_amd import Domain, LuleshProblem, lagrangeElements
    print(f"[monoq-pin] {axis} {side} quirks={quirks}", flush=True)
    dom = Domain(LuleshProblem(10, True, S_Q, 11, 1, 1), quirks=quirks)
    dom.keep_aux(True)
    for name, val in layer_state(axis, side).items():
        dom.set(name, val)
    dom.set("deltatime", 2.0 ** -10)
    lagrangeElements(dom)
    check_layer(dom.get, axis, side, bool(quirks & O.Q3))
    dom.close()


@pytest.mark.gpu
@pytest.mark.parametrize("nx,quirks", [(2, 7), (3, 0), (8, 7)])
def test_gpu_interior_elem_bc_follows_neighbour_lists(nx, quirks):
    """With every elemBC mask cleared, every MonoQ read follows the stored
    face-neighbour lists -- the odd entries included (element 1's xi-minus
    is element 2, the last eta row reads row 0, the last zeta plane reads
    itself).  The HIP path computes those ids arithmetically; cycles of the
    Sedov problem must match the oracle, which indexes the stored lists
    (pinned against src/domain.jl:793-820 above), bit for bit."""
    from lulesh_amd import Domain, LuleshProblem, lagrangeLeapFrog, timeIncrement
    print(f"[elembc-0] nx={nx} quirks={quirks}", flush=True)
    dom = Domain(LuleshProblem(10, True, nx, 11, 1, 1), quirks=quirks)
    od = O.OracleDomain(nx, quirks)
    zero = np.zeros(nx ** 3, dtype=np.uint32)
    dom.set("elemBC", zero)
    od.set_elemBC(zero)
    for _ in range(30):
        timeIncrement(dom)
        lagrangeLeapFrog(dom)
        assert od.cycle_once() == 0
    for n in ("e", "p", "q", "v", "ss", "x", "xd", "zd"):
        g, o = dom.get(n), od.field(n)
        assert np.array_equal(g.view(np.uint64), o.view(np.uint64)), n
    dom.close()
