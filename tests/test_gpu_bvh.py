"""GPU LBVH build vs the oracle: Morton keys, sorted (key, id) pairs and the
reference DFS layout must be bit-exact (bvh.h:186-221)."""
import numpy as np
import pytest

import gpr
import gpu_scenes as G
import oracle_py as O

pytestmark = pytest.mark.gpu

CASES = {
    "n1": lambda: G.soup(1),
    "n2": lambda: G.soup(2),
    "n3": lambda: G.soup(3),
    "soup4096": lambda: G.soup(4096),
    "soup_odd": lambda: G.soup(77777, seed=5),
    "dups": lambda: G.duplicate_heavy(20000),
    "identical": lambda: G.identical(1000),
    "identical_big": lambda: G.identical(9000),  # one bucket larger than the LDS tile
    "planar": lambda: G.planar_grid(90),  # zero z extent of the centroid bounds
    "collinear": lambda: G.collinear(5000),  # zero y and z extent
}


def check_build(pos, scene):
    n = len(pos)
    keys, sk, si = scene.keys()
    okeys = O.morton_keys(pos)
    print(f"[bvh] n={n} keys", flush=True)
    assert np.array_equal(keys, okeys)
    osk, osi = O.sort_pairs(okeys)
    assert np.array_equal(sk, osk)
    assert np.array_equal(si, osi)
    nodes = scene.ref_layout()
    onodes, _, _ = O.lbvh(pos)
    for f in ("bmin", "bmax", "idx", "leaf"):
        assert np.array_equal(nodes[f], onodes[f]), f


@pytest.mark.parametrize("name", list(CASES))
def test_build_matches_oracle(name):
    pos = CASES[name]()
    s = G.scene_from_soup(pos)
    check_build(pos, s)
    s.build()  # a rebuild must reproduce the same arrays
    check_build(pos, s)


def test_build_cornell():
    s = gpr.Scene.cornell(64, 64)
    f = s.flat()
    check_build(f["pos"], s)


def test_build_1m_soup():
    pos = G.soup(1 << 20)
    s = G.scene_from_soup(pos)
    check_build(pos, s)
    t = s.timing()
    print("[bvh] 1M build ms (morton, sort, tree, layouts, total):", t.tolist(), flush=True)


def test_sort_device_entry_point():
    import torch
    rng = np.random.default_rng(4)
    # sizes around the 4096-key scatter blocks and 8192-key histogram tiles;
    # 4M keys: 2048 top-digit buckets of ~2K keys, above the bucket pass's LDS fast path
    for n in (1, 5, 4096, 4097, 8191, 8192, 8193, 12289, 300001, 1 << 22):
        keys = rng.integers(0, 1 << 30, n).astype(np.uint32)
        keys[: n // 3] = keys[0]  # a long run of equal keys
        dk = torch.from_numpy(keys.view(np.int32)).cuda()
        sk = torch.empty_like(dk)
        si = torch.empty_like(dk)
        gpr.sort_pairs_device(dk.data_ptr(), sk.data_ptr(), si.data_ptr(), n)
        torch.cuda.synchronize()
        osk, osi = O.sort_pairs(keys)
        assert np.array_equal(sk.cpu().numpy().view(np.uint32), osk), n
        assert np.array_equal(si.cpu().numpy().view(np.uint32), osi), n
    # keys that do not start on a 16-byte boundary (the histogram pass then
    # reads them one by one instead of as 16-byte vectors)
    keys = rng.integers(0, 1 << 30, 300002).astype(np.uint32)
    dk = torch.from_numpy(keys.view(np.int32)).cuda()
    sk = torch.empty(300001, dtype=torch.int32, device="cuda")
    si = torch.empty_like(sk)
    assert (dk.data_ptr() + 4) % 16 != 0
    gpr.sort_pairs_device(dk.data_ptr() + 4, sk.data_ptr(), si.data_ptr(), 300001)
    torch.cuda.synchronize()
    osk, osi = O.sort_pairs(keys[1:])
    assert np.array_equal(sk.cpu().numpy().view(np.uint32), osk)
    assert np.array_equal(si.cpu().numpy().view(np.uint32), osi)


def test_sort_graph_replays_interleaved_with_eager_sorts():
    # three sorts captured into one graph (an odd count) and replayed three
    # times, with an eager sort of other keys (another digit count) on the same
    # stream and synchronisations between the replays: every replay and every
    # eager sort matches the oracle (each sort leaves its digit totals cleared
    # for the next one in stream order)
    import torch
    rng = np.random.default_rng(21)
    ka = rng.integers(0, 1 << 30, 300001).astype(np.uint32)
    kb = rng.integers(0, 1 << 30, 100003).astype(np.uint32)
    osa, osb = O.sort_pairs(ka), O.sort_pairs(kb)
    d_a = torch.from_numpy(ka.view(np.int32)).cuda()
    d_b = torch.from_numpy(kb.view(np.int32)).cuda()
    sa, ia, sb, ib = (torch.empty_like(d_a), torch.empty_like(d_a), torch.empty_like(d_b), torch.empty_like(d_b))

    def check(sk, si, ref, label):
        assert np.array_equal(sk.cpu().numpy().view(np.uint32), ref[0]), label
        assert np.array_equal(si.cpu().numpy().view(np.uint32), ref[1]), label

    st = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    print("[bvh] sort graph replays", flush=True)
    try:
        with torch.cuda.stream(st):
            gpr.set_stream(st.cuda_stream)
            # allocates this stream's scratch for the larger n, outside the capture
            gpr.sort_pairs_device(d_a.data_ptr(), sa.data_ptr(), ia.data_ptr(), len(ka))
            torch.cuda.synchronize()
            with torch.cuda.graph(graph, stream=st):
                for _ in range(3):
                    gpr.sort_pairs_device(d_a.data_ptr(), sa.data_ptr(), ia.data_ptr(), len(ka))
            for rep in range(3):
                sa.zero_()
                ia.zero_()
                graph.replay()
                torch.cuda.synchronize()
                check(sa, ia, osa, f"replay {rep}")
                gpr.sort_pairs_device(d_b.data_ptr(), sb.data_ptr(), ib.data_ptr(), len(kb))
                torch.cuda.synchronize()
                check(sb, ib, osb, f"eager {rep}")
    finally:
        gpr.set_stream(0)


def test_empty_scene():
    s = gpr.Scene().commit()
    assert s.info().ntri == 0
    rays = gpr.gen_rays(1, 1, 100)
    h = s.trace(rays)
    assert np.all(h["id"] == -1)


def skewed_keys(seed=12):
    """Clustered 30-bit keys: most top-digit buckets empty, a few far above the
    LDS bucket size, crowded sub-buckets, one sub-bucket of 5000 keys (above
    the rank limit) and runs of equal keys."""
    rng = np.random.default_rng(seed)
    parts = []
    for c in range(48):  # 48 clusters of 2^12-wide key ranges
        base = int(rng.integers(0, 1 << 30 - 12)) << 12
        parts.append(base + rng.integers(0, 1 << 12, int(rng.integers(500, 9000))))
    parts.append((7 << 21) + rng.integers(0, 1 << 9, 5000))  # one sub-bucket of 5000
    parts.append(np.full(2500, 12345678))  # equal keys
    parts.append(rng.integers(0, 1 << 30, 20000))  # some uniform background
    k = np.concatenate(parts).astype(np.uint32)
    return k[rng.permutation(len(k))]


def test_sort_skewed_keys_device_entry_point():
    import torch
    for seed in (12, 13):
        keys = skewed_keys(seed)
        n = len(keys)
        print(f"[bvh] skewed sort n={n}", flush=True)
        dk = torch.from_numpy(keys.view(np.int32)).cuda()
        sk = torch.empty_like(dk)
        si = torch.empty_like(dk)
        gpr.sort_pairs_device(dk.data_ptr(), sk.data_ptr(), si.data_ptr(), n)
        torch.cuda.synchronize()
        osk, osi = O.sort_pairs(keys)
        assert np.array_equal(sk.cpu().numpy().view(np.uint32), osk)
        assert np.array_equal(si.cpu().numpy().view(np.uint32), osi)


@pytest.mark.parametrize("name", ["sphere_grid", "mix_1m"])
def test_build_skewed_scenes(name):
    # the sphere grid of the full render (253,958 triangles: 88 % of its keys
    # fall in 32 top-digit buckets far above the LDS bucket size) and a
    # 1M-triangle sphere grid + soup scene
    s = gpr.Scene.sphere_grid(64, 36, 32) if name == "sphere_grid" else gpr.Scene.mix(1011, 1 << 20, 0.5)
    f = s.flat()
    print(f"[bvh] {name}: {len(f['pos'])} triangles", flush=True)
    check_build(f["pos"], s)
