"""k_nodal's tiling rule, restated on the CPU (lulesh_kernels.h nodal_span /
nodal_tile and the index arithmetic at the top of k_nodal in
lulesh_kernels.hip): for every tile shape the selection may pick, every node
column of an s x s element plane is owned by exactly one tile, and every
element a node column folds in is computed by a lane of the same tile, at
the LDS slot the node reads.  The GPU parity tests check the compiled
kernel bit for bit against the oracle; this test guards the invariants the
kernel's reads rely on (no LDS read outside the tile, no node left out or
written twice) over far more sizes than those can run."""
import pytest

BLOCK = 256


def span(s, a):
    """tiles along an axis of s elements for a lanes per tile"""
    m = s + 1
    if m <= a:
        return 1
    rest = m - 2 * a
    return 2 + (max(rest, 0) + a - 2) // (a - 1)


def tile_nodes(k, a, ntiles, s):
    """first node column of tile k, the lane offset of its first element,
    and its node end"""
    i0 = 0 if k == 0 else a + (k - 1) * (a - 1)
    oi = 0 if k == 0 else 1
    iend = s + 1 if k == ntiles - 1 else i0 + a - oi
    return i0, oi, iend


def check_plane(s, etx, ety):
    tx, ty = span(s, etx), span(s, ety)
    owner = {}
    for ky in range(ty):
        j0, oj, jend = tile_nodes(ky, ety, ty, s)
        for kx in range(tx):
            i0, oi, iend = tile_nodes(kx, etx, tx, s)
            lane_elem = {}
            for t in range(BLOCK):
                ti, tj = t % etx, t // etx
                ei, ej = i0 - oi + ti, j0 - oj + tj
                if tj < ety and ei < s and ej < s:
                    assert ei >= 0 and ej >= 0
                    lane_elem[tj * etx + ti] = (ei, ej)
            for t in range(BLOCK):
                ti, tj = t % etx, t // etx
                ni, nj = i0 + ti, j0 + tj
                if not (tj < ety and ni < iend and nj < jend):
                    continue
                assert (ni, nj) not in owner, (s, etx, ni, nj)
                owner[(ni, nj)] = (kx, ky)
                l00 = (tj + oj - 1) * etx + (ti + oi - 1)
                for need, slot, elem in (
                        (ni >= 1 and nj >= 1, l00, (ni - 1, nj - 1)),
                        (ni < s and nj >= 1, l00 + 1, (ni, nj - 1)),
                        (ni >= 1 and nj < s, l00 + etx, (ni - 1, nj)),
                        (ni < s and nj < s, l00 + etx + 1, (ni, nj))):
                    if need:
                        assert 0 <= slot < etx * ety, (s, etx, kx, ky, ti, tj)
                        assert lane_elem.get(slot) == elem, (s, etx, kx, ky, ti, tj)
    assert len(owner) == (s + 1) ** 2


@pytest.mark.parametrize("etx", range(16, 33))
def test_every_node_owned_once_and_reads_stay_in_tile(etx):
    ety = BLOCK // etx
    for s in list(range(1, 98)) + [127, 191, 192]:
        check_plane(s, etx, ety)


@pytest.mark.parametrize("etx", [2, 3, 5, 8, 12, 13, 64, 128])
def test_other_tile_widths(etx):
    """widths the tile-range setting and the autotuner may allow"""
    ety = BLOCK // etx
    for s in list(range(1, 42)) + [63]:
        check_plane(s, etx, ety)


def test_span_counts_cover_the_axis():
    """the tiles of nodal_span cover s+1 node columns exactly, each owning
    between 1 and a columns (the kernel's lanes per row)"""
    for a in list(range(2, 33)) + [64, 128]:
        for s in range(1, 900):
            n = span(s, a)
            i0, oi, iend = tile_nodes(n - 1, a, n, s)
            owned = iend - i0
            assert 1 <= owned <= a, (a, s, owned)
            bounds = [tile_nodes(k, a, n, s) for k in range(n)]
            assert bounds[0][0] == 0 and bounds[-1][2] == s + 1
            for k in range(n - 1):
                assert bounds[k][2] == bounds[k + 1][0]
                assert 1 <= bounds[k][2] - bounds[k][0] <= a
