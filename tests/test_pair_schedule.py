"""The pass schedule of the pair-tile exchange (mattohf_amd/csrc/tdhf_pairtile.h)
covers every ordered pair of rows of a tile (8 W x 8 W rows for W = 8, 4, 2
waves) exactly once, and the J-side sums of one window are touched by one
wave only.

The schedule is restated here from the kernel's comment: group g = 8 w + gw
owns row g of block bi and in pass (p1, p0) pairs with row
8 ((w + p1) % W) + (gw + p0) % 8 of block bj; off-diagonal tiles run all 8 W
passes two-sided, diagonal tiles the offsets o <= -o (mod W, mod 8),
two-sided unless o == -o.  A wrong schedule also fails the GPU parity tests
of the pair tiles against the oracle; this test names the broken property.
"""
import pytest


def diag_pass(p1, p0, W=8):
    m1, m0 = (W - p1) % W, (8 - p0) % 8
    if (p1, p0) == (m1, m0):
        return True, False
    return (p1 < m1 or (p1 == m1 and p0 < m0)), True


def passes(diag, W=8):
    """(I-side pairs, J-side pairs, window owners) of one tile of W waves; a
    pair (r, s) is row r of the tile's first block with row s of its second."""
    iside, jside, owners = {}, {}, {}
    for p1 in range(W):
        for p0 in range(8):
            run, both = (True, True) if not diag else diag_pass(p1, p0, W)
            if not run:
                continue
            for w in range(W):
                for gw in range(8):
                    r = 8 * w + gw
                    s = 8 * ((w + p1) % W) + (gw + p0) % 8
                    iside[(r, s)] = iside.get((r, s), 0) + 1
                    if both:
                        jside[(s, r)] = jside.get((s, r), 0) + 1
                        owners.setdefault((p1, s), set()).add(w)
    return iside, jside, owners


@pytest.mark.parametrize("W", [8, 4, 2])
def test_off_diagonal_tile_covers_both_orders_once(W):
    iside, jside, owners = passes(False, W)
    full = {(a, b) for a in range(8 * W) for b in range(8 * W)}
    assert set(iside) == full and set(iside.values()) == {1}
    assert set(jside) == full and set(jside.values()) == {1}
    # a J-side row is written by one wave per window (no atomics needed)
    assert all(len(ws) == 1 for ws in owners.values())


@pytest.mark.parametrize("W", [8, 4, 2])
def test_diagonal_tile_covers_each_ordered_pair_once(W):
    iside, jside, owners = passes(True, W)
    both = dict(iside)
    for k, v in jside.items():
        both[k] = both.get(k, 0) + v
    assert set(both) == {(a, b) for a in range(8 * W) for b in range(8 * W)}
    assert set(both.values()) == {1}
    assert all(len(ws) == 1 for ws in owners.values())
    assert sum(diag_pass(p1, p0, W)[0] for p1 in range(W) for p0 in range(8)) == 4 * W + 2
