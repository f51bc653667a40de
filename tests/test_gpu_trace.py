"""GPU traversal vs the oracle (trace / trace_shadow, pathtracer_kernel.h:175-321).

* reference order: bit-exact hits against the oracle's restatement;
* fast BVH4 traversal: triangle id exact except ties (|dt| <= 1e-6 t), t
  bit-equal to the oracle whenever the id agrees, t within 1e-5 relative of an
  FP64 brute-force intersector;
* any-hit: occlusion flag exact.
"""
import numpy as np
import pytest

import gpr
import gpu_scenes as G
import oracle_py as O

pytestmark = pytest.mark.gpu


def rays_for(n, seed=2):
    return gpr.gen_rays(seed, 1, n)


def compare_fast(h, ref, pos, rays, label):
    same = h["id"] == ref["id"]
    bad = np.nonzero(~same)[0]
    for i in bad:
        # a disagreement must be a tie: both triangles hit at (nearly) the same t
        assert h["id"][i] != -1 and ref["id"][i] != -1, (label, i, h[i], ref[i])
        assert abs(float(h["t"][i]) - float(ref["t"][i])) <= 1e-6 * abs(float(ref["t"][i])), (label, i)
    assert len(bad) <= max(2, len(h) // 10000), (label, len(bad))
    assert np.array_equal(h["t"][same], ref["t"][same])
    assert np.array_equal(h["u"][same], ref["u"][same])
    assert np.array_equal(h["v"][same], ref["v"][same])
    return len(bad)


@pytest.mark.parametrize("n", [1, 3, 4096, 100000])
def test_closest_and_any(n):
    check_closest_and_any(G.soup(n), f"n={n}")


@pytest.mark.parametrize("name", ["planar", "collinear"])
def test_closest_and_any_flat_geometry(name):
    # every box of the planar grid is flat in z (zero-thickness slabs); the
    # collinear scene's centroid bounds are flat in y and z
    pos = G.planar_grid(90) if name == "planar" else G.collinear(5000)
    check_closest_and_any(pos, name)


def check_closest_and_any(pos, label):
    s = G.scene_from_soup(pos)
    onodes, _, _ = O.lbvh(pos)
    rays = rays_for(20000)
    print(f"[trace] {label}", flush=True)
    ref, _, _ = O.trace(onodes, pos, rays)
    href = s.trace(rays, ref_order=True)
    assert np.array_equal(href.view(np.uint32), ref.view(np.uint32))  # bit-exact, misses included
    hfast = s.trace(rays)
    compare_fast(hfast, ref, pos, rays, f"closest {label}")
    sh = gpr.gen_shadow_rays(2, 2, rays)
    aref, _, _ = O.trace(onodes, pos, sh, any_hit=True)
    a_ref_gpu = s.trace(sh, any_hit=True, ref_order=True)
    assert np.array_equal(a_ref_gpu.view(np.uint32), aref.view(np.uint32))
    afast = s.trace(sh, any_hit=True)
    assert np.array_equal(afast["id"] != -1, aref["id"] != -1)
    hit = afast["id"] != -1
    assert np.all((afast["t"][hit] >= 0) & (afast["t"][hit] <= np.float32(0.9999)))


def test_host_buffers_in_chunks_equal_device_buffers():
    # host-buffer batches above 2^20 rays are traced in pipelined chunks; the
    # result must be that of one launch over device buffers.  Closest hit is
    # bit-equal (each ray's visit order is fixed); for any hit the occluder
    # reported may depend on which rays share a wave, so the flag is compared.
    import torch
    s = G.scene_from_soup(G.soup(1 << 16))
    for n in ((1 << 21) + 13, 3 << 20):  # a ragged last chunk; whole chunks
        rays = rays_for(n, seed=5)
        print(f"[trace] host chunks n={n}", flush=True)
        for any_hit in (False, True):
            rr = gpr.gen_shadow_rays(3, 3, rays) if any_hit else rays
            h = s.trace(rr, any_hit=any_hit)
            d_r = torch.from_numpy(rr.view(np.uint8)).cuda()
            d_h = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
            s.trace_device(d_r.data_ptr(), d_h.data_ptr(), n, any_hit=any_hit)
            torch.cuda.synchronize()
            dh = d_h.cpu().numpy().view(gpr.HIT_DTYPE)
            if any_hit:
                assert np.array_equal(h["id"] != -1, dh["id"] != -1)
            else:
                assert np.array_equal(h.view(np.uint32), dh.view(np.uint32))
            assert (h["id"] != -1).mean() > 0.05


def test_concurrent_streams_trace_and_sort_as_alone():
    # two host threads, each on its own stream, trace different scenes and
    # sort different keys at the same time (the traversal cursor / spill area
    # and the sort scratch are kept per stream); every result equals the same
    # work done alone
    import threading
    import torch
    n = 1 << 18
    scenes = [G.scene_from_soup(G.soup(1 << 15, seed=k)) for k in (11, 12)]
    d_r = [torch.from_numpy(rays_for(n, seed=k).view(np.uint8)).cuda() for k in (11, 12)]
    d_h = [torch.empty(n * 16, dtype=torch.uint8, device="cuda") for _ in range(2)]
    rng = np.random.default_rng(8)
    keys = [rng.integers(0, 1 << 30, 300001 + 77 * k).astype(np.uint32) for k in range(2)]
    d_k = [torch.from_numpy(k.view(np.int32)).cuda() for k in keys]
    d_sk = [torch.empty_like(k) for k in d_k]
    d_si = [torch.empty_like(k) for k in d_k]
    alone = []
    for k in range(2):
        scenes[k].trace_device(d_r[k].data_ptr(), d_h[k].data_ptr(), n)
        torch.cuda.synchronize()
        alone.append(d_h[k].cpu().numpy().copy())
    streams = [torch.cuda.Stream() for _ in range(2)]
    torch.cuda.synchronize()
    errors = []

    def work(k):
        try:
            gpr.set_stream(streams[k].cuda_stream)
            for _ in range(12):
                scenes[k].trace_device(d_r[k].data_ptr(), d_h[k].data_ptr(), n)
                gpr.sort_pairs_device(d_k[k].data_ptr(), d_sk[k].data_ptr(), d_si[k].data_ptr(), len(keys[k]))
            streams[k].synchronize()
        except Exception as e:  # reported by the main thread
            errors.append(e)

    print("[trace] two streams from two threads", flush=True)
    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k in range(2):
        assert np.array_equal(d_h[k].cpu().numpy(), alone[k])
        osk, osi = O.sort_pairs(keys[k])
        assert np.array_equal(d_sk[k].cpu().numpy().view(np.uint32), osk)
        assert np.array_equal(d_si[k].cpu().numpy().view(np.uint32), osi)


def test_fast_vs_fp64_bruteforce_1m():
    pos = G.soup(1 << 20)
    s = G.scene_from_soup(pos)
    rays = rays_for(1 << 22)
    h = s.trace(rays)
    sel = np.random.default_rng(0).choice(len(rays), 512, replace=False)
    bf, t64 = O.brute(pos, rays[sel], fp64=True)
    hs = h[sel]
    agree = hs["id"] == bf["id"]
    for i in np.nonzero(~agree)[0]:
        # only near-ties / edge grazes may differ
        assert hs["id"][i] != -1 and bf["id"][i] != -1
        assert abs(float(hs["t"][i]) - t64[i]) <= 1e-6 * t64[i]
    hitm = bf["id"] != -1
    assert np.array_equal(hs["id"] != -1, hitm)
    rel = np.abs(hs["t"][hitm].astype(np.float64) - t64[hitm]) / t64[hitm]
    assert rel.max() <= 1e-5
    # reference order on all 4M rays: bit-exact with the oracle on a slice
    href = s.trace(rays, ref_order=True)
    nv, nt = s.ref_stats()
    print(f"[trace] 1M: mean nodes {nv / len(rays):.2f} tris {nt / len(rays):.2f}", flush=True)
    onodes, _, _ = O.lbvh(pos)
    ref, onv, ont = O.trace(onodes, pos, rays[:65536])
    assert np.array_equal(href[:65536].view(np.uint32), ref.view(np.uint32))
    compare_fast(h[:65536], ref, pos, rays[:65536], "closest 1M")
    # the second pass of the workload: any-hit shadow rays to random in-cube
    # targets (t in [0, 0.9999]) on all 4M rays; reference order bit-exact and
    # the fast traversal's occlusion flag exact on a 65,536-ray slice, and the
    # occlusion flag equal to a brute-force intersector on a sample
    sh = gpr.gen_shadow_rays(2, 2, rays)
    print("[trace] 1M any-hit shadow pass", flush=True)
    a_ref_gpu = s.trace(sh, any_hit=True, ref_order=True)
    a_fast = s.trace(sh, any_hit=True)
    aref, _, _ = O.trace(onodes, pos, sh[:65536], any_hit=True)
    assert np.array_equal(a_ref_gpu[:65536].view(np.uint32), aref.view(np.uint32))
    assert np.array_equal(a_fast["id"][:65536] != -1, aref["id"] != -1)
    occ = a_fast["id"] != -1
    # a dense soup: almost every segment is blocked (99.8 %), but not all
    assert occ.mean() > 0.5 and (~occ).sum() > 1000
    hit = a_fast[occ]
    assert np.all((hit["t"] >= 0) & (hit["t"] <= np.float32(0.9999)))
    bfa, _ = O.brute(pos, sh[sel], any_hit=True)
    assert np.array_equal(a_fast["id"][sel] != -1, bfa["id"] != -1)


def test_axis_parallel_and_degenerate_rays():
    pos = G.soup(4096)
    s = G.scene_from_soup(pos)
    onodes, _, _ = O.lbvh(pos)
    rng = np.random.default_rng(3)
    n = 3000
    rays = np.zeros(n, gpr.RAY_DTYPE)
    rays["o"] = rng.random((n, 3)).astype(np.float32)
    d = np.zeros((n, 3), np.float32)
    ax = rng.integers(0, 3, n)
    d[np.arange(n), ax] = rng.choice([-1.0, 1.0], n)
    d[: n // 3, (ax[: n // 3] + 1) % 3] = 0.5  # some in a coordinate plane
    rays["d"] = d
    rays["tmin"] = 0
    rays["tmax"] = np.inf
    rays["tmax"][:50] = 0  # empty interval
    ref, _, _ = O.trace(onodes, pos, rays)
    assert np.array_equal(s.trace(rays, ref_order=True).view(np.uint32), ref.view(np.uint32))
    compare_fast(s.trace(rays), ref, pos, rays, "axis")


def test_launch_configurations_give_identical_hits():
    import torch
    pos = gpr.gen_soup(5, 0, 20000, 0.01)
    rays = gpr.gen_rays(5, 1, 50000)
    s = gpr.Scene.soup(pos).commit()
    d_rays = torch.from_numpy(rays.view(np.uint8)).cuda()
    d_hits = torch.empty(len(rays) * 16, dtype=torch.uint8, device="cuda")
    ref = None
    print("[trace] launch configurations", flush=True)
    for cfg in [(256, 16, 64, 0), (64, 8, 64, 0), (512, 32, 256, 0), (128, 16, 128, 0),
                (256, 16, 64, 1), (64, 8, 128, 1), (512, 32, 64, 1)]:
        s.trace_set_cfg(*cfg)
        s.trace_device(d_rays.data_ptr(), d_hits.data_ptr(), len(rays))
        torch.cuda.synchronize()
        h = d_hits.cpu().numpy().copy()
        if ref is None:
            ref = h
        assert np.array_equal(h, ref), cfg
    # any hit: occlusion identical across configurations
    occ = None
    for cfg in [(256, 16, 64, 0), (128, 8, 64, 1)]:
        s.trace_set_cfg(*cfg)
        s.trace_device(d_rays.data_ptr(), d_hits.data_ptr(), len(rays), any_hit=True)
        torch.cuda.synchronize()
        o = d_hits.cpu().numpy().view(gpr.HIT_DTYPE)["id"] != -1
        if occ is None:
            occ = o
        assert np.array_equal(o, occ), cfg
    s.trace_set_cfg(256, 16, 64, 0)
    res = s.trace_tune(d_rays.data_ptr(), len(rays), cap=128)
    assert len(res) == 72 and all(res[k].ms <= res[k + 1].ms for k in range(71))
    s.trace_device(d_rays.data_ptr(), d_hits.data_ptr(), len(rays))
    torch.cuda.synchronize()
    assert np.array_equal(d_hits.cpu().numpy(), ref)


def test_bvh4_stats_overflow_count_and_stream_release():
    import torch
    pos = G.soup(1 << 16)
    s = G.scene_from_soup(pos)
    n = 1 << 16
    rays = rays_for(n)
    d_r = torch.from_numpy(rays.view(np.uint8)).cuda()
    d_h = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    nf, nt = s.bvh4_stats(d_r.data_ptr(), n)
    af, at = s.bvh4_stats(d_r.data_ptr(), n, any_hit=True)
    nodes4 = (len(pos) + 2) // 3  # BVH4 nodes are at least ntri / 3
    print(f"[trace] 64K soup: {nf / n:.2f} BVH4 nodes, {nt / n:.2f} triangle tests per ray", flush=True)
    assert n <= nf < n * nodes4 and 0 < nt < n * len(pos)
    assert af <= nf  # any hit stops at its first occluder on the same walk
    # one triangle: the root is a leaf, every ray tests it once and fetches no node
    one = G.scene_from_soup(G.soup(1))
    assert one.bvh4_stats(d_r.data_ptr(), n) == (0, n)
    s.trace_device(d_r.data_ptr(), d_h.data_ptr(), n)
    assert gpr.trace_overflow_count() == 0
    h0 = d_h.cpu().numpy().copy()
    # release this stream's work buffers: the next trace allocates them again
    st = torch.cuda.Stream()
    try:
        gpr.set_stream(st.cuda_stream)
        s.trace_device(d_r.data_ptr(), d_h.data_ptr(), n)
        gpr.release_stream(st.cuda_stream)
        gpr.release_stream(st.cuda_stream)  # nothing left to release: no error
        d_h.zero_()
        torch.cuda.synchronize()
        s.trace_device(d_r.data_ptr(), d_h.data_ptr(), n)
        st.synchronize()
        assert np.array_equal(d_h.cpu().numpy(), h0)
        assert gpr.trace_overflow_count() == 0
    finally:
        gpr.set_stream(0)
