"""Batched and served forms on the GPU.

A job rendered by the multi-job engine must equal, bit for bit, the same job
rendered alone by the single-scene renderer (PCG mode, BVH4 traversal, same
seed, depth, roulette start and camera), whatever other jobs share its passes:
each path keeps its own RNG stream and each pixel receives its terms in the
same kernel order.  The server must give the same image alone and in a mix.
"""
import numpy as np
import pytest

import gpr

pytestmark = pytest.mark.gpu


def plain(scene, j):
    """The single-scene renderer on the job's camera."""
    scene.set_camera(np.array(j.c2w[:], np.float32), j.fov_deg, j.width, j.height)
    r = gpr.Renderer(scene, max_depth=j.max_depth, rng_mode=gpr.RNG_PCG, rr_start=j.rr_start,
                     seed=j.seed, traversal=gpr.TRAVERSAL_FAST)
    r.render(j.spp)
    rgb, _ = r.read()
    r.free()
    return rgb


@pytest.fixture(scope="module")
def scenes():
    c = gpr.Scene.cornell(32, 32)
    g = gpr.Scene.sphere_grid(32, 32, 6)
    return [c, g]


def jobs_for(ids):
    cam_c = gpr.lookat((0, 0, 0), (0.3, 0.1, 3.9))
    cam_g = gpr.lookat((0, 0, 0), (1.0, 6.0, 10.0))
    return [
        gpr.job(ids[0], 40, 24, 3, cam_c, fov=40, max_depth=5, rr_start=3, seed=11),
        gpr.job(ids[1], 33, 17, 2, cam_g, fov=35, max_depth=8, rr_start=3, seed=7),
        gpr.job(ids[0], 16, 16, 4, cam_c, fov=50, max_depth=2, rr_start=1, seed=3),
        gpr.job(ids[1], 64, 8, 1, cam_g, fov=30, max_depth=6, rr_start=2, seed=5),
    ]


def test_batch_equals_plain_per_job(scenes):
    arena = sum(gpr.arena_bytes(s) for s in scenes)
    b = gpr.Batch(arena, capacity=4096, max_jobs=8)
    ids = [b.add_scene(s) for s in scenes]
    js = jobs_for(ids)
    for j in js:
        b.add_job(j)
    print("[serve] batch of 4 jobs, capacity 4096", flush=True)
    b.run()
    st = b.stats()
    assert st.passes >= 4 and st.paths == sum(j.width * j.height * j.spp for j in js)
    for k, j in enumerate(js):
        got = b.read(k)
        want = plain(scenes[j.scene], j)
        bad = np.nonzero(np.any(got.view(np.uint32) != want.view(np.uint32), axis=1))[0]
        assert len(bad) == 0, (k, len(bad), bad[:8])
    b.free()


def test_batch_capacity_forces_many_passes_same_result(scenes):
    # capacity of one job at a time: every pass holds a single job
    arena = sum(gpr.arena_bytes(s) for s in scenes)
    js = jobs_for([0, 1])
    outs = []
    for cap in (1024, 1 << 16):
        b = gpr.Batch(arena, capacity=cap, max_jobs=8)
        for s in scenes:
            b.add_scene(s)
        for j in js:
            b.add_job(j)
        b.run()
        outs.append([b.read(k) for k in range(len(js))])
        b.free()
    for a, c in zip(*outs):
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_tuner_configs_agree(scenes):
    arena = sum(gpr.arena_bytes(s) for s in scenes)
    b = gpr.Batch(arena, capacity=8192, max_jobs=8)
    ids = [b.add_scene(s) for s in scenes]
    js = jobs_for(ids)
    for j in js:
        b.add_job(j)
    print("[serve] autotune", flush=True)
    res = b.tune()
    assert len(res) == 4 * 3 * 6 + 1
    assert all(res[k].ms <= res[k + 1].ms for k in range(len(res) - 1))
    ref = None
    for cfg in [(64, 8, 64, 0, 0), (512, 32, 512, 0, 0), (128, 16, 64, 0, 1), (512, 8, 128, 0, 1),
                (256, 16, 128, 1, 0), (res[0].block, res[0].stack, res[0].chunk, res[0].layout, res[0].refill)]:
        b.set_cfg(*cfg)
        b.run()
        got = [b.read(k) for k in range(len(js))]
        if ref is None:
            ref = got
        for a, c in zip(ref, got):
            if cfg[3] == 1:
                # the reference layout visits triangles in another order: ties of t may differ
                assert np.mean(np.all(np.abs(a - c) <= 1e-4 * np.maximum(np.abs(c), 1e-6) + 1e-7, axis=1)) >= 0.99
            else:
                assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
    b.free()


def test_server_alone_equals_mix_equals_plain(scenes):
    arena = sum(gpr.arena_bytes(s) for s in scenes)
    js = jobs_for([0, 1])
    budget = arena + (1 << 30)
    srv = gpr.Server(budget, arena, capacity=4096, max_jobs=16)
    for s in scenes:
        srv.add_scene(s)
    res_t = srv.tune(js[:2])
    assert len(res_t) == 73 and all(res_t[k].ms <= res_t[k + 1].ms for k in range(len(res_t) - 1))
    srv.set_cfg(64, 16, 128, 0)
    print("[serve] server alone", flush=True)
    alone, _ = srv.wait(srv.submit(js[1]))
    print("[serve] server mix", flush=True)
    tickets = [srv.submit(j) for j in js] + [srv.submit(js[1])]
    res = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.free()
    assert st.completed == 1 + len(tickets)
    assert st.p50_ms > 0 and st.p99_ms >= st.p50_ms and st.renders_per_s > 0
    assert st.mean_jobs_per_pass > 1.0
    for r in (res[1][0], res[-1][0]):
        assert np.array_equal(r.view(np.uint32), alone.view(np.uint32))
    for k, j in enumerate(js):
        want = plain(scenes[j.scene], j)
        assert np.array_equal(res[k][0].view(np.uint32), want.view(np.uint32)), k


def test_server_budget_admission(scenes):
    arena = sum(gpr.arena_bytes(s) for s in scenes)
    cam = gpr.lookat((0, 0, 0), (0, 0, 3.9))
    # two job slots: later requests wait for admission
    small = gpr.job(0, 64, 64, 2, cam, seed=9)
    with pytest.raises(gpr.GprError) as e:
        gpr.Server(arena + 4096, arena, capacity=4096, max_jobs=2)
    assert e.value.code == gpr.ERR_BUDGET
    srv = gpr.Server(arena + (1 << 30), arena, capacity=4096, max_jobs=2)
    srv.add_scene(scenes[0])
    srv.add_scene(scenes[1])
    ts = [srv.submit(small) for _ in range(5)]
    outs = [srv.wait(t)[0] for t in ts]
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))
    with pytest.raises(gpr.GprError) as e:
        srv.submit(gpr.job(0, 1 << 15, 1 << 15, 1, cam))
    assert e.value.code in (gpr.ERR_INVALID, gpr.ERR_BUDGET)
    srv.free()


def test_server_cancel_queued_active_and_finished(scenes):
    # one job slot: the second request stays queued while the first renders
    import time
    arena = sum(gpr.arena_bytes(x) for x in scenes)
    srv = gpr.Server(arena + (1 << 30), arena, capacity=4096, max_jobs=1)
    ids = [srv.add_scene(x) for x in scenes]
    cam = gpr.lookat((0, 0, 0), (0.3, 0.1, 3.9))
    long_job = gpr.job(ids[0], 48, 40, 4000, cam, fov=40, max_depth=5, rr_start=3, seed=1)
    t_active = srv.submit(long_job)
    t_queued = srv.submit(gpr.job(ids[1], 32, 32, 4, gpr.lookat((0, 0, 0), (1.0, 6.0, 10.0)), fov=35, max_depth=6,
                                  rr_start=3, seed=2))
    srv.cancel(t_queued)
    with pytest.raises(gpr.GprError) as e:
        srv.wait(t_queued)
    assert e.value.code == gpr.ERR_CANCELLED
    time.sleep(0.05)  # the long request is on the GPU
    srv.cancel(t_active)
    with pytest.raises(gpr.GprError) as e:
        srv.wait(t_active)
    assert e.value.code == gpr.ERR_CANCELLED
    # the slot and the frame-buffer range are free again, and a new request is unaffected
    j = gpr.job(ids[0], 40, 24, 3, cam, fov=40, max_depth=5, rr_start=3, seed=11)
    t = srv.submit(j)
    got, _ = srv.wait(t)
    want = plain(scenes[0], j)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    with pytest.raises(gpr.GprError):
        srv.cancel(t)  # finished and collected: the ticket is gone
    st = srv.stats()
    assert st.cancelled == 2 and st.completed == 1
    srv.free()


def test_batch_keeps_its_own_copy_of_added_scenes(scenes):
    # add_scene copies the scene into the arena: the caller may free its scene
    # straight away and the batch still renders it exactly
    tmp = gpr.Scene.cornell(32, 32)
    arena = gpr.arena_bytes(tmp)
    b = gpr.Batch(arena, capacity=4096, max_jobs=4)
    sid = b.add_scene(tmp)
    tmp.free()
    j = jobs_for([sid, sid])[0]
    b.add_job(j)
    b.run()
    got = b.read(0)
    want = plain(scenes[0], j)  # the same procedural Cornell box, built again
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    b.free()
