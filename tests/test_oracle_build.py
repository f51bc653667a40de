"""Pins the oracle's Morton / sort / LBVH restatement (bvh.h) against closed
forms, brute force and invariants -- no GPU."""
import numpy as np
import pytest

import oracle_py as O


def interleave_loop(x, y, z):
    """Brute-force Morton: bit b of x -> 3b, of y -> 3b+1, of z -> 3b+2."""
    k = 0
    for b in range(10):
        k |= ((x >> b) & 1) << (3 * b)
        k |= ((y >> b) & 1) << (3 * b + 1)
        k |= ((z >> b) & 1) << (3 * b + 2)
    return k


def test_left_shift3_known_answers():
    assert O.lib.orc_left_shift3(1023) == 0x09249249
    assert O.lib.orc_left_shift3(1024) == 0x09249249  # clamped (bvh.h:15)
    assert O.lib.orc_left_shift3(0) == 0
    assert O.lib.orc_left_shift3(1) == 1
    assert O.lib.orc_left_shift3(2) == 8
    assert O.lib.orc_left_shift3(3) == 9
    assert O.lib.orc_left_shift3(0x2AA) == 0x08208208


def test_left_shift3_matches_magic_mask_form():
    # the mask-and-shift form of the reference (bvh.h:16-23), evaluated in numpy
    x = np.arange(0, 1025, dtype=np.uint64)
    y = np.where(x == 1024, 1023, x)
    y = (y | (y << 16)) & 0x30000ff
    y = (y | (y << 8)) & 0x300f00f
    y = (y | (y << 4)) & 0x30c30c3
    y = (y | (y << 2)) & 0x9249249
    got = np.array([O.lib.orc_left_shift3(int(v)) for v in x], np.uint64)
    assert np.array_equal(got, y)


def test_encode_morton3_known_answers():
    assert O.lib.orc_encode_morton3(1023, 1023, 1023) == 0x3FFFFFFF
    assert O.lib.orc_encode_morton3(1024, 1024, 1024) == 0x3FFFFFFF
    assert O.lib.orc_encode_morton3(1023.999, 0, 0) == 0x09249249
    assert O.lib.orc_encode_morton3(1, 0, 0) == 1
    assert O.lib.orc_encode_morton3(0, 1, 0) == 2
    assert O.lib.orc_encode_morton3(0, 0, 1) == 4
    assert O.lib.orc_encode_morton3(0.99, 0.99, 0.99) == 0
    rng = np.random.default_rng(3)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(0, 1024, 3))
        assert O.lib.orc_encode_morton3(x + 0.5, y + 0.25, z + 0.75) == interleave_loop(x, y, z)


def grid_soup(cells):
    """Degenerate triangles (3 equal vertices) at cell centres of a 1024^3 grid
    spanning [0,1]; two corner triangles pin the centroid bounds to [0,1]^3."""
    pts = [(0.0, 0.0, 0.0), (1.0, 1.0, 1.0)] + [((x + 0.5) / 1024, (y + 0.5) / 1024, (z + 0.5) / 1024)
                                                for x, y, z in cells]
    pos = np.repeat(np.asarray(pts, np.float32), 3, axis=0).reshape(-1, 9)
    return pos


def test_morton_keys_quantize_against_bounds():
    rng = np.random.default_rng(5)
    cells = [tuple(int(v) for v in rng.integers(0, 1024, 3)) for _ in range(500)]
    keys = O.morton_keys(grid_soup(cells))
    assert keys[0] == 0 and keys[1] == 0x3FFFFFFF
    for (x, y, z), k in zip(cells, keys[2:]):
        assert k == interleave_loop(x, y, z)


@pytest.mark.parametrize("flat", [(2,), (1, 2), (0, 1, 2)])
def test_morton_keys_flat_axes_quantize_to_zero(flat):
    # a flat axis of the centroid bounds makes (c - min) / (max - min) 0/0 = NaN:
    # its coordinate is defined as 0, the other axes quantize as usual
    rng = np.random.default_rng(len(flat))
    cells = [tuple(int(v) for v in rng.integers(0, 1024, 3)) for _ in range(300)]
    pos = grid_soup(cells).reshape(-1, 3, 3)
    for a in flat:
        pos[:, :, a] = 0.25
    keys = O.morton_keys(pos.reshape(-1, 9))
    for (x, y, z), k in zip([(0, 0, 0), (1023, 1023, 1023)] + cells, keys):
        q = [0 if a in flat else v for a, v in enumerate((x, y, z))]
        assert k == interleave_loop(*q)
    nodes, sk, si = O.lbvh(pos.reshape(-1, 9))
    assert len(nodes) == 2 * len(keys) - 1
    assert np.array_equal(sk, np.sort(keys)) and np.array_equal(keys[si], sk)


def test_morton_keys_use_bounding_box_centroid():
    # a triangle whose vertex mean differs from its box centre
    tri = np.array([[0, 0, 0, 1, 0, 0, 1, 1, 0]], np.float32)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    pos = np.concatenate([tri, np.array([[2, 2, 2] * 3], np.float32), np.array([[-2, -2, -2] * 3], np.float32)])
    O.lib.orc_centroid_bounds(O.ptr(pos), 3, O.ptr(lo), O.ptr(hi))
    assert np.array_equal(lo, [-2, -2, -2]) and np.array_equal(hi, [2, 2, 2])
    k = O.morton_keys(pos)[0]
    # box centre (0.5, 0.5, 0) -> q = (2.5/4*1024, 2.5/4*1024, 2/4*1024) = (640, 640, 512)
    assert k == interleave_loop(640, 640, 512)


def test_sort_pairs_sorted_permutation_stable():
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 64, 5000).astype(np.uint32)
    sk, si = O.sort_pairs(keys)
    assert np.all(np.diff(sk.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(si), np.arange(5000))
    assert np.array_equal(keys[si], sk)
    same = sk[1:] == sk[:-1]
    assert np.all(si[1:][same] > si[:-1][same])  # ties in index order


def split_by_bit(keys, l, r):
    """Independent statement of the reference split: highest differing bit of
    the end keys, split before the first key with that bit set; equal end keys
    split at the midpoint."""
    if keys[l] == keys[r - 1]:
        return (l + r) // 2
    b = int(keys[l] ^ keys[r - 1]).bit_length() - 1
    for i in range(l, r):
        if (int(keys[i]) >> b) & 1:
            return i
    raise AssertionError


def ref_ranges(keys):
    out = []

    def rec(l, r):
        out.append((l, r))
        if r - l == 1:
            return
        m = split_by_bit(keys, l, r)
        rec(l, m)
        rec(m, r)

    rec(0, len(keys))
    return out


@pytest.mark.parametrize("n,span", [(1, 4), (2, 1), (7, 3), (300, 40), (513, 2), (1000, 1 << 30), (64, 1)])
def test_topology_matches_bit_split_definition(n, span):
    rng = np.random.default_rng(n)
    keys = np.sort(rng.integers(0, span, n).astype(np.uint32))
    lo, hi = O.ranges(keys)
    assert list(zip(lo.tolist(), hi.tolist())) == ref_ranges(keys)


def tri_boxes(pos):
    p = pos.reshape(-1, 3, 3)
    return p.min(axis=1), p.max(axis=1)


@pytest.mark.parametrize("n", [1, 2, 3, 36, 257, 2000])
def test_lbvh_layout_boxes_and_links(n):
    rng = np.random.default_rng(11 + n)
    c = rng.random((n, 3)).astype(np.float32)
    pos = (np.repeat(c, 3, axis=0) + rng.normal(0, 0.01, (3 * n, 3)).astype(np.float32)).reshape(n, 9)
    nodes, sk, si = O.lbvh(pos)
    assert len(nodes) == 2 * n - 1
    lo, hi = O.ranges(sk)
    blo, bhi = tri_boxes(pos)
    # DFS order: left child = i+1, right child = idx; leaves hold the sorted triangle ids
    for i, nd in enumerate(nodes):
        a, b = lo[i], hi[i]
        members = si[a:b]
        assert np.array_equal(nd["bmin"], blo[members].min(axis=0))
        assert np.array_equal(nd["bmax"], bhi[members].max(axis=0))
        if nd["leaf"]:
            assert b - a == 1 and nd["idx"] == si[a]
        else:
            assert lo[i + 1] == a and hi[nd["idx"]] == b and hi[i + 1] == lo[nd["idx"]]
