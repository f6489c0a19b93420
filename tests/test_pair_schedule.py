"""The pass schedule of the pair tiles covers every ordered pair of rows of a
tile exactly once, and the J-side sums of one window are touched by one wave
only: the forward exchange tiles (mattohf_amd/csrc/tdhf_pairtile.h, 8 groups
per wave at NG = 64, tiles of 8 W rows for W = 8, 4, 2 waves) and the reverse
stage tiles (mattohf_amd/csrc/tdhf_adjtile.h, 4 groups per wave at NG = 128,
tiles of 4 W rows for W = 8, 4).

The schedule is restated here from the kernels' comments: group g = G w + gw
owns row g of block bi and in pass (p1, p0) pairs with row
G ((w + p1) % W) + (gw + p0) % G of block bj; off-diagonal tiles run all G W
passes two-sided, diagonal tiles the offsets o <= -o (mod W, mod G),
two-sided unless o == -o.  A wrong schedule also fails the GPU parity tests
of the pair tiles against the oracle; this test names the broken property.
"""
import pytest


def diag_pass(p1, p0, W=8, G=8):
    m1, m0 = (W - p1) % W, (G - p0) % G
    if (p1, p0) == (m1, m0):
        return True, False
    return (p1 < m1 or (p1 == m1 and p0 < m0)), True


def passes(diag, W=8, G=8):
    """(I-side pairs, J-side pairs, window owners) of one tile of W waves of G
    groups; a pair (r, s) is row r of the tile's first block with row s of its
    second."""
    iside, jside, owners = {}, {}, {}
    for p1 in range(W):
        for p0 in range(G):
            run, both = (True, True) if not diag else diag_pass(p1, p0, W, G)
            if not run:
                continue
            for w in range(W):
                for gw in range(G):
                    r = G * w + gw
                    s = G * ((w + p1) % W) + (gw + p0) % G
                    iside[(r, s)] = iside.get((r, s), 0) + 1
                    if both:
                        jside[(s, r)] = jside.get((s, r), 0) + 1
                        owners.setdefault((p1, s), set()).add(w)
    return iside, jside, owners


GEOMETRIES = [(8, 8), (4, 8), (2, 8), (8, 4), (4, 4)]  # (W, G)


@pytest.mark.parametrize("W,G", GEOMETRIES)
def test_off_diagonal_tile_covers_both_orders_once(W, G):
    iside, jside, owners = passes(False, W, G)
    full = {(a, b) for a in range(G * W) for b in range(G * W)}
    assert set(iside) == full and set(iside.values()) == {1}
    assert set(jside) == full and set(jside.values()) == {1}
    # a J-side row is written by one wave per window (no atomics needed)
    assert all(len(ws) == 1 for ws in owners.values())


@pytest.mark.parametrize("W,G", GEOMETRIES)
def test_diagonal_tile_covers_each_ordered_pair_once(W, G):
    iside, jside, owners = passes(True, W, G)
    both = dict(iside)
    for k, v in jside.items():
        both[k] = both.get(k, 0) + v
    assert set(both) == {(a, b) for a in range(G * W) for b in range(G * W)}
    assert set(both.values()) == {1}
    assert all(len(ws) == 1 for ws in owners.values())
    # 4 W + 2 passes for 8 groups per wave, 2 W + 2 for 4
    npass = sum(diag_pass(p1, p0, W, G)[0] for p1 in range(W) for p0 in range(G))
    assert npass == (G // 2) * W + 2
