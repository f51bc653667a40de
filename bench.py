#!/usr/bin/env python3
"""Benchmark of the wavefront path-tracing core on one MI355X.

Workload (one "step" = one closest-hit pass over the whole ray batch):
  LBVH build + ray query microbenchmark, FP32.  1,048,576 triangles with
  centroids uniform in the unit cube and N(0, 0.005^2) vertex jitter (PCG32
  seed 2 stream 0); 4,194,304 rays with origins uniform in the cube and
  directions uniform on the sphere, tmax = inf (seed 2 stream 1); a second
  any-hit pass of shadow rays to uniform in-cube targets (seed 2 stream 2).

The metric is closest-hit Mrays/s with rays and hits resident in HBM.  The
JSON line also carries the any-hit rate, the LBVH build time, the radix-sort
time against its HBM roofline, the traversal roofline, a CPU baseline (the
CPU restatement of the reference traversal and LBVH build, timed on this
host) and the served half of the metric: 64 device-resident scenes and a
60-s Poisson request trace at 200 requests/s served with merged queues,
against a seeded subset of the same requests rendered one at a time by the
single-scene renderer.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-renderer_amd"))

HBM_PEAK_GBS = 8000.0
N_TRI = 1 << 20
N_RAYS = 1 << 22


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--sort-iters", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 20, help="rays in the CPU baseline sample")
    ap.add_argument("--cpu-sample-1core", type=int, default=1 << 16, help="rays in the 1-core CPU baseline sample")
    ap.add_argument("--only", choices=["all", "trace", "sort", "copy", "gather"], default="all",
                    help="profiling runs: restrict the GPU work to one kernel family")
    ap.add_argument("--workload", choices=["raycast", "render", "batch", "serve", "build"], default="raycast",
                    help="raycast: 1M-triangle LBVH + ray query (the headline metric, with a served-stream "
                         "segment); render: 1920x1080 sphere-grid path tracer; batch: 256 scenes in one "
                         "run; serve: the Poisson request stream; build: LBVH builds of the skewed scenes")
    ap.add_argument("--requests", type=int, default=0,
                    help="cap on the served requests (default: every arrival of the trace)")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the Poisson arrival trace (200 requests/s) of the served segment")
    ap.add_argument("--scenes", type=int, default=0, help="scenes of the batch / serve workloads (default 256 / 64)")
    ap.add_argument("--no-host-buffers", action="store_true",
                    help="raycast run without the host-buffer (PCIe-inclusive) timing, so that a kernel "
                         "trace holds only the resident 4M-ray traversal launches")
    ap.add_argument("--no-serve", action="store_true", help="raycast run without the served-stream segment")
    ap.add_argument("--no-skewed-builds", action="store_true",
                    help="raycast run without the skewed-scene LBVH builds (so a kernel trace holds only the "
                         "1M-triangle soup's build and sorts)")
    ap.add_argument("--capacity", type=int, default=1 << 26,
                    help="paths per pass of the served / batched engine (64 M: 554 against 513 Mpaths/s at 16 M "
                         "on the served trace)")
    ap.add_argument("--trace-cfg", default="64,16,64,1",
                    help="block,stack,chunk,refill of the --only trace profiling run (the tuned configuration)")
    return ap.parse_args()


def pmc_traffic(key, field="bytes"):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes (scaled by calibration kernels of known size), committed under
    profiles/; None if absent."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic_pmc.json")) as f:
            return json.load(f)[key][field]
    except (OSError, KeyError, ValueError):
        return None


def copy_calibration(torch, nbytes=1 << 30, reps=5):
    """Plain device copy of known size: the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters of a profiling run to bytes."""
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


def gather_calibration(torch, rows=1 << 24, n=1 << 23, reps=5):
    """Random 128-B row reads of known size from a 2 GiB table (past the 256 MB
    Infinity Cache): the reference point for FETCH_SIZE on a kernel whose
    reads are scattered whole lines, like the traversal's node fetches."""
    src = torch.empty(rows, 32, dtype=torch.float32, device="cuda")
    idx = torch.randint(0, rows, (n,), device="cuda")
    dst = torch.empty(n, 32, dtype=torch.float32, device="cuda")
    for _ in range(reps):
        torch.index_select(src, 0, idx, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"gather_rows": n, "row_bytes": 128, "index_bytes": 8, "reps": reps}), flush=True)


def host_cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def native_oracle():
    """The CPU restatement built with -O3 -march=native for this host (a
    temporary copy; the committed build stays portable).  Returns its path."""
    import subprocess
    import tempfile
    out = os.path.join(tempfile.mkdtemp(prefix="gpr_oracle_"), "liboracle_native.so")
    subprocess.check_call(["g++", "-O3", "-march=native", "-std=c++17", "-fPIC", "-ffp-contract=off",
                           "-fno-fast-math", "-fopenmp", "-shared", "-o", out,
                           os.path.join(ROOT, "oracle", "oracle.cpp")])
    return out


def median_time(fn, runs=5):
    """One warm-up call, then the median of `runs` timed calls (seconds)."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(pos, rays_np, sample, sample_1core, gpu_ref_hits, cpu_nodes):
    """The CPU restatement of the reference path, -O3 -march=native on this
    host: the LBVH build single-threaded as the reference's LBVH_build runs
    (bvh.h:186-221), and the reference-order closest-hit traversal on the
    host's CPU share and on one core; median of 5 after a warm-up each."""
    os.environ["GPR_ORACLE_LIB"] = native_oracle()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    info = host_cpu_info()
    # the GPU's share of the host: at most 16 threads per GPU
    threads = max(1, min(info["affinity"], 16))
    box = {}
    build_s, _ = median_time(lambda: box.__setitem__("nodes", O.lbvh(pos)[0]))
    nodes = box["nodes"]
    rays = rays_np[:sample]
    trav_s, runs = median_time(lambda: box.__setitem__("hits", O.trace(nodes, pos, rays, nthreads=threads)[0]))
    same = bool(np.array_equal(box["hits"].view(np.uint32), gpu_ref_hits[:sample].view(np.uint32)))
    r1 = rays_np[:sample_1core]
    one_s, _ = median_time(lambda: O.trace(nodes, pos, r1, nthreads=1))
    cpu_nodes.append(nodes)
    return {"value": round(len(rays) / trav_s / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(rays)} of the {len(rays_np)} closest-hit rays, reference-order traversal of the "
                      f"CPU restatement (-O3 -march=native), {threads} threads, median of 5 after a warm-up "
                      f"({trav_s:.3f} s)",
            "one_core_mrays_s": round(len(r1) / one_s / 1e6, 4),
            "one_core_sample": f"{len(r1)} rays, 1 thread, median of 5 ({one_s:.3f} s)",
            "lbvh_build_ms_1thread": round(build_s * 1e3, 2),
            "lbvh_build": "Morton keys, std::sort, recursive split, DFS flatten on one thread (the reference's "
                          "LBVH_build path), 1,048,576 triangles, median of 5 after a warm-up",
            "hits_equal_gpu_reference_order": same,
            "host": info, "thread_cap": "16 threads: the host share of one GPU on this machine",
            "runs_s": [round(x, 4) for x in runs]}


# ------------------------------------------------------------------ render --
def run_render(a, gpr, np, torch):
    """Full microfacet path tracer: 8x8 UV spheres (64x32 segments) + ground,
    two quad lights, 1920x1080, 64 spp, depth 8, roulette from bounce 3, PCG
    seed 7.  A step is one sample of every pixel."""
    W, H, SPP = 1920, 1080, 64
    scene = gpr.Scene.sphere_grid(W, H, 32)
    info = scene.info()
    r = gpr.Renderer(scene, max_depth=8, rng_mode=gpr.RNG_PCG, rr_start=3, seed=7,
                     traversal=gpr.TRAVERSAL_FAST)
    steps = a.steps if a.steps != 20 else SPP
    r.render(max(1, a.warmup))
    r.reset()
    torch.cuda.synchronize()
    sh, sc = r.queue_sizes(8)
    rays_per_sample = W * H + int(sh.sum()) + int(sc.sum())
    t0 = time.perf_counter()
    r.render(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = 1000 * dt / steps
    out = {"metric": "Mpaths/s, 1920x1080 sphere-grid path tracer", "value": round(W * H / (ms * 1e-3) / 1e6, 2),
           "unit": "Mpaths/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (procedural sphere grid)",
           "config": {"workload": "sphere_grid_1920x1080_64spp_depth8", "triangles": info.ntri, "spp": steps,
                      "max_depth": 8, "rr_start": 3, "seed": 7},
           "rays_per_sample": rays_per_sample,
           "mrays_per_s_all_rays": round(rays_per_sample / (ms * 1e-3) / 1e6, 2)}
    if not a.no_cpu_baseline:
        os.environ["GPR_ORACLE_LIB"] = native_oracle()
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_py as O
        cores = max(1, min(len(os.sched_getaffinity(0)), 16))
        # a bounded sample: 1 spp of the same scene at 1920x1080
        sw, sh_ = 1920, 1080
        small = gpr.Scene.sphere_grid(sw, sh_, 32)
        f = small.flat()
        inf = f["info"]
        osc = O.make_scene(f, np.array(inf.c2w[:], np.float32), inf.fov_deg, sw, sh_)
        nodes, _, _ = O.lbvh(f["pos"])
        t0 = time.perf_counter()
        O.render(osc, nodes, 1, 8, gpr.RNG_PCG, seed=7, rr_start=3, nthreads=cores)
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(sw * sh_ / cdt / 1e6, 4), "unit": "Mpaths/s", "cores": cores,
                               "kind": "port",
                               "sample": f"1 spp of the same scene at {sw}x{sh_}, depth 8, PCG, CPU restatement "
                                         f"(-O3 -march=native, depth-first per path, {cores} threads), {cdt:.2f} s",
                               "host": host_cpu_info()}
    return out


# --------------------------------------------------------- batch / serve --
def loguniform(rng, lo, hi):
    return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def orbit_camera(rng, radius):
    """cameraToWorld on a sphere around the scene centre, looking at it."""
    z = rng.uniform(-0.2, 1.0)
    phi = rng.uniform(0, 2 * np.pi)
    rxy = np.sqrt(max(0.0, 1 - z * z))
    o = np.array([radius * rxy * np.cos(phi), 0.5 + radius * z, radius * rxy * np.sin(phi)])
    return o


def make_scenes(gpr, n, lo, hi, soup, seed):
    rng = np.random.default_rng(seed)
    scenes, tris = [], []
    for k in range(n):
        t = int(loguniform(rng, lo, hi))
        s = gpr.Scene.mix(1000 + k, t, soup)
        scenes.append(s)
        tris.append(s.info().ntri)
    return scenes, tris


def run_batch(a, gpr, np_, torch):
    """256 scenes (sphere grid + random soup, 1e3..2e6 triangles log-uniform),
    each with its own camera and resolution in {128^2, 256^2, 512^2, 1024^2},
    16 spp, depth 6, in one autotuned batch."""
    nsc = a.scenes or 256
    t0 = time.perf_counter()
    scenes, tris = make_scenes(gpr, nsc, 1e3, 2e6, 0.5, seed=41)
    setup_s = time.perf_counter() - t0
    arena = sum(gpr.arena_bytes(s) for s in scenes)
    b = gpr.Batch(arena, capacity=a.capacity, max_jobs=256)
    ids = [b.add_scene(s) for s in scenes]
    del scenes
    rng = np.random.default_rng(43)
    res_choices = [128, 256, 512, 1024]
    total_px = 0
    for k in range(nsc):
        res = int(rng.choice(res_choices))
        o = orbit_camera(rng, 9.0)
        j = gpr.job(ids[k], res, res, 16, gpr.lookat((0, 0.5, 0), o), fov=40, max_depth=6, rr_start=3,
                    seed=100 + k)
        b.add_job(j)
        total_px += res * res
    t1 = time.perf_counter()
    tuned = b.tune()
    tune_s = time.perf_counter() - t1
    torch.cuda.synchronize()
    b.run()
    st = b.stats()
    paths = st.paths
    out = {"metric": "renders/s, 256-scene batch", "value": round(nsc / (st.ms * 1e-3), 3), "unit": "renders/s",
           "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(st.ms, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scenes)",
           "config": {"workload": "batch_256_scenes_16spp_depth6", "scenes": nsc, "triangles_total": int(sum(tris)),
                      "pixels_total": total_px, "spp": 16, "max_depth": 6},
           "mpaths_per_s": round(paths / (st.ms * 1e-3) / 1e6, 2), "passes": st.passes,
           "tuned": {"block": tuned[0].block, "stack": tuned[0].stack, "chunk": tuned[0].chunk,
                     "layout": tuned[0].layout, "probe_ms": round(tuned[0].ms, 3),
                     "worst_probe_ms": round(tuned[-1].ms, 3), "configs": len(tuned), "tune_s": round(tune_s, 2)},
           "setup_s": round(setup_s, 1)}
    b.free()
    return out


def server_budget(capacity):
    """Device memory the server may use beyond its scene arena: the merged
    queues (3 x 84 B per path of the pass capacity) and the requests' frame
    buffers (16 B per pixel), with 4 GiB for frame buffers in flight."""
    return 3 * 84 * capacity + (4 << 30)


def c5_requests(gpr, ids, seconds, rate=200.0, seed=53, max_requests=0):
    """The request trace: Poisson arrivals at `rate`/s over `seconds`; each
    request a random scene, a camera on a sphere around it, resolution in
    {256^2, 512^2, 1024^2} and spp in {4, 16, 64}, depth 8, roulette from 3."""
    rng = np.random.default_rng(seed)
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds or (max_requests and len(arrivals) >= max_requests):
            break
        arrivals.append(t)
    reqs = []
    for k in range(len(arrivals)):
        sc = int(rng.integers(0, len(ids)))
        res = int(rng.choice([256, 512, 1024]))
        spp = int(rng.choice([4, 16, 64]))
        o = orbit_camera(rng, 9.0)
        reqs.append(gpr.job(ids[sc], res, res, spp, gpr.lookat((0, 0.5, 0), o), fov=40, max_depth=8, rr_start=3,
                            seed=200 + k))
    return np.array(arrivals), reqs


def serve_stream(a, gpr, np_, torch, seconds, n_scenes, plain_n=96, max_requests=0):
    """The served form on the request trace (open loop: requests are submitted
    at their arrival times whatever the backlog), then a seeded subset of the
    same requests one at a time through the single-scene renderer, one
    renderer per (scene, resolution) reused with new camera and options per
    request.  Setup (scene upload, server creation, autotuning, renderer
    creation) is timed apart from both legs."""
    import threading
    t0 = time.perf_counter()
    scenes, tris = make_scenes(gpr, n_scenes, 1e4, 4e6, 0.0, seed=51)
    scene_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    arena = sum(gpr.arena_bytes(s) for s in scenes)
    srv = gpr.Server(arena + server_budget(a.capacity), arena, capacity=a.capacity, max_jobs=256)
    ids = [srv.add_scene(s) for s in scenes]
    arrivals, reqs = c5_requests(gpr, ids, seconds, max_requests=max_requests)
    n_req = len(reqs)
    tuned = srv.tune(reqs[: min(8, n_req)])
    server_setup_s = time.perf_counter() - t0
    # the plain leg's subset (seeded) and the requests whose images are kept:
    # the subset, plus the first 256^2 request of each spp class (RMSE)
    sub = sorted(np.random.default_rng(59).choice(n_req, min(plain_n, n_req), replace=False).tolist())
    rmse_k = {}
    for k, j in enumerate(reqs):
        # a camera well above the ground plane (one below it sees the plane's back: black)
        if j.width == 256 and j.spp not in rmse_k and j.c2w[7] > 2.0:
            rmse_k[j.spp] = k
    keep = set(sub) | set(rmse_k.values())
    tickets = [None] * n_req

    def client():
        start = time.perf_counter()
        for k in range(n_req):
            dt = arrivals[k] - (time.perf_counter() - start)
            if dt > 0:
                time.sleep(dt)
            tickets[k] = srv.submit(reqs[k])

    th = threading.Thread(target=client)
    torch.cuda.synchronize()
    print(f"[bench] serving {n_req} requests ({seconds:g} s of Poisson arrivals at 200/s)", file=sys.stderr,
          flush=True)
    ts = time.perf_counter()
    th.start()
    served, lat = {}, np.zeros(n_req)
    for k in range(n_req):
        while tickets[k] is None:
            time.sleep(0.0005)
        img, lat[k] = srv.wait(tickets[k], want_image=k in keep)
        if k in keep:
            served[k] = img
        if (k + 1) % 1000 == 0:
            print(f"[bench] served {k + 1}/{n_req} in {time.perf_counter() - ts:.1f} s", file=sys.stderr, flush=True)
    th.join()
    served_s = time.perf_counter() - ts
    st = srv.stats()
    srv.free()
    # ---- the same requests (seeded subset) one at a time, plain renderer
    print(f"[bench] plain renderer on {len(sub)} of the requests", file=sys.stderr, flush=True)
    rend, plain_setup_s, plain_s = {}, 0.0, 0.0
    identical, max_rmse_plain = True, 0.0
    for k in sub:
        j = reqs[k]
        s = scenes[j.scene]
        s.set_camera(np.array(j.c2w[:], np.float32), j.fov_deg, j.width, j.height)
        key = (j.scene, j.width)
        if key not in rend:
            t1 = time.perf_counter()
            rend[key] = gpr.Renderer(s, max_depth=j.max_depth, rng_mode=gpr.RNG_PCG, rr_start=j.rr_start,
                                     seed=j.seed, traversal=gpr.TRAVERSAL_FAST)
            plain_setup_s += time.perf_counter() - t1
        r = rend[key]
        t1 = time.perf_counter()
        r.set_opts(max_depth=j.max_depth, rng_mode=gpr.RNG_PCG, rr_start=j.rr_start, seed=j.seed,
                   traversal=gpr.TRAVERSAL_FAST)
        r.render(j.spp)
        plain_s += time.perf_counter() - t1
        img = r.read()[0]
        identical = identical and bool(np.array_equal(img.view(np.uint32), served[k].view(np.uint32)))
        max_rmse_plain = max(max_rmse_plain, float(np.sqrt(np.mean((img / j.spp - served[k] / j.spp) ** 2))))
    for r in rend.values():
        r.free()
    # ---- RMSE of served images against 4096-sample renders of the same request
    rmse = {"served_vs_plain_max": round(max_rmse_plain, 9), "per_spp_vs_4096spp": []}
    for spp in sorted(rmse_k):
        k = rmse_k[spp]
        j = reqs[k]
        s = scenes[j.scene]
        s.set_camera(np.array(j.c2w[:], np.float32), j.fov_deg, j.width, j.height)
        r = gpr.Renderer(s, max_depth=j.max_depth, rng_mode=gpr.RNG_PCG, rr_start=j.rr_start, seed=j.seed + 100000,
                         traversal=gpr.TRAVERSAL_FAST)
        r.render(4096)
        ref = r.read()[0] / 4096.0
        r.free()
        rmse["per_spp_vs_4096spp"].append({"request": k, "resolution": j.width, "spp": spp,
                                           "rmse": round(float(np.sqrt(np.mean((served[k] / spp - ref) ** 2))), 6),
                                           "reference_mean": round(float(ref.mean()), 6)})
    sub_paths = sum(reqs[k].width * reqs[k].height * reqs[k].spp for k in sub)
    paths = sum(j.width * j.height * j.spp for j in reqs)
    plain_rate = len(sub) / plain_s
    return {"renders_per_s": round(n_req / served_s, 3), "p50_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_ms": round(float(np.percentile(lat, 99)), 1),
            "plain_renders_per_s": round(plain_rate, 3),
            "speedup_vs_plain": round((n_req / served_s) / plain_rate, 3),
            "speedup_path_weighted": round((paths / served_s) / (sub_paths / plain_s), 3),
            "requests": n_req, "trace_seconds": seconds, "scenes": n_scenes,
            "scene_triangles_total": int(sum(tris)),
            "served_s": round(served_s, 2), "plain_requests": len(sub), "plain_s": round(plain_s, 3),
            "mpaths_per_s": round(paths / served_s / 1e6, 2),
            "plain_mpaths_per_s": round(sub_paths / plain_s / 1e6, 2),
            "mean_jobs_per_pass": round(st.mean_jobs_per_pass, 2), "passes": st.passes,
            "images_identical_to_plain": identical,
            "rmse": rmse,
            "setup_s": {"scene_build": round(scene_s, 2), "server_create_upload_tune": round(server_setup_s, 2),
                        "plain_renderers_create": round(plain_setup_s, 3)},
            "tuned": {"block": tuned[0].block, "stack": tuned[0].stack, "chunk": tuned[0].chunk,
                      "layout": tuned[0].layout},
            "offered_load": "Poisson 200 requests/s, open loop",
            "latency": "submit -> image readable on the host"}


def serve_latency_under_capacity(a, gpr, np_, torch, rate=30.0, seconds=10.0, n_scenes=64):
    """p50 / p99 latency of the served form when the offered load is below
    what the GPU sustains (the trace above is a backlog at 200/s, where
    latency measures the queue): the same kind of requests at `rate`/s."""
    import threading
    scenes, _ = make_scenes(gpr, n_scenes, 1e4, 4e6, 0.0, seed=51)
    arena = sum(gpr.arena_bytes(s) for s in scenes)
    srv = gpr.Server(arena + server_budget(a.capacity), arena, capacity=a.capacity, max_jobs=256)
    ids = [srv.add_scene(s) for s in scenes]
    del scenes
    arrivals, reqs = c5_requests(gpr, ids, seconds, rate=rate, seed=61)
    srv.tune(reqs[: min(8, len(reqs))])
    tickets = [None] * len(reqs)

    def client():
        start = time.perf_counter()
        for k in range(len(reqs)):
            dt = arrivals[k] - (time.perf_counter() - start)
            if dt > 0:
                time.sleep(dt)
            tickets[k] = srv.submit(reqs[k])

    th = threading.Thread(target=client)
    ts = time.perf_counter()
    th.start()
    lat = np.zeros(len(reqs))
    for k in range(len(reqs)):
        while tickets[k] is None:
            time.sleep(0.0005)
        lat[k] = srv.wait(tickets[k], want_image=False)[1]
    th.join()
    wall = time.perf_counter() - ts
    srv.free()
    return {"offered_load": f"Poisson {rate:g} requests/s for {seconds:g} s", "requests": len(reqs),
            "renders_per_s": round(len(reqs) / wall, 3), "p50_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_ms": round(float(np.percentile(lat, 99)), 1)}


def skewed_builds(gpr, reps=10):
    """LBVH build stage times (ms, median of `reps` rebuilds after 2) of the
    scenes whose Morton keys are clustered: the 1920x1080 render's sphere grid
    (253,958 triangles) and a 4M-triangle sphere-grid scene of the request
    stream."""
    out = {}
    for name, mk in (("sphere_grid_253958_tris", lambda: gpr.Scene.sphere_grid(1920, 1080, 32)),
                     ("sphere_grid_4032006_tris", lambda: gpr.Scene.mix(2005, 4000000, 0.0))):
        s = mk()
        ts = []
        for _ in range(reps + 2):
            s.build()
            ts.append(s.timing().tolist())
        m = np.median(np.array(ts[2:]), axis=0)
        out[name] = {"bounds_morton": round(float(m[0]), 4), "sort": round(float(m[1]), 4),
                     "tree_refit": round(float(m[2]), 4), "layouts": round(float(m[3]), 4),
                     "total": round(float(m[4]), 4)}
        s.free()
    return out


def run_build(a, gpr, np_, torch):
    res = skewed_builds(gpr, reps=max(1, a.steps))
    return {"metric": "LBVH build ms (skewed scenes)", "value": res["sphere_grid_253958_tris"]["total"],
            "unit": "ms", "n_gpus": 1, "steps": max(1, a.steps), "warmup": 2,
            "ms_per_step": res["sphere_grid_253958_tris"]["total"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scenes)",
            "config": {"workload": "lbvh_build_sphere_grids"}, "builds": res}


def run_serve(a, gpr, np_, torch):
    res = serve_stream(a, gpr, np_, torch, a.serve_seconds, a.scenes or 64, max_requests=a.requests)
    out = {"metric": "served renders/s (Poisson request stream)", "value": res["renders_per_s"],
           "unit": "renders/s", "n_gpus": 1, "steps": res["requests"], "warmup": 0,
           "ms_per_step": round(1000.0 / max(res["renders_per_s"], 1e-9), 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scenes)",
           "config": {"workload": "serve_64_scenes_poisson_200_per_s", "requests": res["requests"]},
           "serving": res}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    import gpr

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    gpr.set_stream(stream.cuda_stream)
    if a.workload != "raycast":
        fn = {"render": run_render, "batch": run_batch, "serve": run_serve, "build": run_build}[a.workload]
        print(json.dumps(fn(a, gpr, np, torch)), flush=True)
        return
    if a.only == "copy":
        copy_calibration(torch)
        return
    if a.only == "gather":
        gather_calibration(torch)
        return

    print("[bench] generating inputs", file=sys.stderr, flush=True)
    pos = gpr.gen_soup(2, 0, N_TRI, 0.005)
    rays_np = gpr.gen_rays(2, 1, N_RAYS)
    shadow_np = gpr.gen_shadow_rays(2, 2, rays_np)

    print("[bench] building LBVH", file=sys.stderr, flush=True)
    scene = gpr.Scene.soup(pos).commit()
    builds = []
    for _ in range(5):
        scene.build()
        builds.append(scene.timing().tolist())
    build_ms = [float(np.median([b[k] for b in builds])) for k in range(5)]

    # device-resident rays / hits
    d_rays = torch.from_numpy(rays_np.view(np.uint8)).cuda()
    d_shadow = torch.from_numpy(shadow_np.view(np.uint8)).cuda()
    d_hits = torch.empty(N_RAYS * 16, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()

    if a.only == "trace":
        scene.trace_set_cfg(*[int(x) for x in a.trace_cfg.split(",")])
        for _ in range(a.warmup + a.steps):
            scene.trace_device(d_rays.data_ptr(), d_hits.data_ptr(), N_RAYS)
        torch.cuda.synchronize()
        print(json.dumps({"profiled": "closest-hit", "launches": a.warmup + a.steps}), flush=True)
        return
    # ---- radix sort: mean over back-to-back sorts of the scene's Morton keys,
    # captured as one graph of sort_iters sorts and replayed
    keys, _, _ = scene.keys()
    d_keys = torch.from_numpy(keys.view(np.int32)).cuda()
    d_sk = torch.empty_like(d_keys)
    d_si = torch.empty_like(d_keys)
    gpr.sort_pairs_device(d_keys.data_ptr(), d_sk.data_ptr(), d_si.data_ptr(), N_TRI)
    torch.cuda.synchronize()
    ref_sk, ref_si = d_sk.clone(), d_si.clone()
    iters = max(1, a.sort_iters)
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    with torch.cuda.stream(cap):
        gpr.set_stream(cap.cuda_stream)
        gpr.sort_pairs_device(d_keys.data_ptr(), d_sk.data_ptr(), d_si.data_ptr(), N_TRI)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=cap):
            for _ in range(iters):
                gpr.sort_pairs_device(d_keys.data_ptr(), d_sk.data_ptr(), d_si.data_ptr(), N_TRI)
    gpr.set_stream(stream.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    graph.replay()
    torch.cuda.synchronize()
    sort_ok = bool(torch.equal(d_sk, ref_sk) and torch.equal(d_si, ref_si))
    # five timed replays, each bracketed by device synchronisation (a replay
    # queued behind other work measured ~0.7 us per sort more), the median kept
    sort_replays_ms = []
    for _ in range(5):
        d_sk.zero_()
        torch.cuda.synchronize()
        e0.record()
        graph.replay()
        e1.record()
        torch.cuda.synchronize()
        sort_replays_ms.append(e0.elapsed_time(e1) / iters)
        sort_ok = sort_ok and bool(torch.equal(d_sk, ref_sk) and torch.equal(d_si, ref_si))
    sort_ms = float(np.median(sort_replays_ms))
    # the same sorts launched one by one (no graph), for reference
    torch.cuda.synchronize()
    e0.record()
    for _ in range(iters):
        gpr.sort_pairs_device(d_keys.data_ptr(), d_sk.data_ptr(), d_si.data_ptr(), N_TRI)
    e1.record()
    torch.cuda.synchronize()
    sort_ms_eager = e0.elapsed_time(e1) / iters
    sort_bytes = 68.0 * N_TRI
    if a.only == "sort":
        print(json.dumps({"profiled": "sort", "iters": iters, "ms_per_sort_graph": sort_ms,
                          "replays_ms_per_sort": sort_replays_ms,
                          "ms_per_sort_eager": sort_ms_eager, "graph_result_ok": sort_ok}), flush=True)
        return

    # ---- traversal statistics.  Algorithmic bytes of the BVH4 walk the timed
    # kernel runs: 32-B ray + 16-B hit + one 128-B node per node fetch + one
    # 48-B triangle record per triangle test, counted on the device.  The
    # reference visit order's counts (32-B nodes, 36-B triangles) are kept as
    # a diagnostic: BVH4 fetches far fewer, larger nodes.
    print("[bench] traversal statistics", file=sys.stderr, flush=True)
    nf4, nt4 = scene.bvh4_stats(d_rays.data_ptr(), N_RAYS)
    af4, at4 = scene.bvh4_stats(d_shadow.data_ptr(), N_RAYS, any_hit=True)
    n_node4, n_tri4 = nf4 / N_RAYS, nt4 / N_RAYS
    bytes_per_ray = 32 + 16 + 128 * n_node4 + 48 * n_tri4
    bytes_per_ray_any = 32 + 16 + 128 * af4 / N_RAYS + 48 * at4 / N_RAYS
    scene.trace_device(d_rays.data_ptr(), d_hits.data_ptr(), N_RAYS, ref_order=True)
    torch.cuda.synchronize()
    nv, nt = scene.ref_stats()
    n_node, n_tri = nv / N_RAYS, nt / N_RAYS
    bytes_per_ray_ref = 32 + 16 + 32 * n_node + 36 * n_tri
    ref_hits = d_hits[: min(a.cpu_sample, N_RAYS) * 16].cpu().numpy().view(gpr.HIT_DTYPE).copy()

    # ---- traversal launch configuration: autotuned on the batch itself
    print("[bench] traversal autotune", file=sys.stderr, flush=True)
    tuned_any = scene.trace_tune(d_shadow.data_ptr(), N_RAYS, any_hit=True, cap=128)
    tuned = scene.trace_tune(d_rays.data_ptr(), N_RAYS, cap=128)

    # ---- closest hit: warmup, then exactly K timed steps
    print("[bench] closest-hit timing", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        scene.trace_device(d_rays.data_ptr(), d_hits.data_ptr(), N_RAYS)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(a.steps):
        scene.trace_device(d_rays.data_ptr(), d_hits.data_ptr(), N_RAYS)
    e1.record()
    torch.cuda.synchronize()
    ms_closest = e0.elapsed_time(e1) / a.steps
    overflow = gpr.trace_overflow_count()  # rays whose stack overflowed in the last launch (must be 0)
    hits = d_hits.cpu().numpy().view(gpr.HIT_DTYPE)
    hit_frac = float((hits["id"] != -1).mean())

    # ---- any-hit shadow pass (its own tuned configuration)
    scene.trace_set_cfg(tuned_any[0].block, tuned_any[0].stack, tuned_any[0].chunk, tuned_any[0].refill)
    for _ in range(a.warmup):
        scene.trace_device(d_shadow.data_ptr(), d_hits.data_ptr(), N_RAYS, any_hit=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(a.steps):
        scene.trace_device(d_shadow.data_ptr(), d_hits.data_ptr(), N_RAYS, any_hit=True)
    e1.record()
    torch.cuda.synchronize()
    ms_any = e0.elapsed_time(e1) / a.steps

    # ---- the same closest-hit batch handed over in host memory, as the
    # reference's host wrappers take it: each call stages the rays in and the
    # hits out over PCIe (pinned host buffers, then pageable numpy arrays)
    host = None
    if not a.no_host_buffers:
        import ctypes as C
        scene.trace_set_cfg(tuned[0].block, tuned[0].stack, tuned[0].chunk, tuned[0].refill)
        h_rays = torch.from_numpy(rays_np.view(np.uint8)).pin_memory()
        h_hits = torch.empty(N_RAYS * 16, dtype=torch.uint8).pin_memory()

        def trace_pinned():
            gpr.check(gpr.lib.gpr_trace(scene.h, C.c_void_p(h_rays.data_ptr()), C.c_void_p(h_hits.data_ptr()),
                                        N_RAYS, 0))

        trace_pinned()
        t0 = time.perf_counter()
        for _ in range(3):
            trace_pinned()
        ms_pinned = (time.perf_counter() - t0) / 3 * 1e3
        host_ok = bool(np.array_equal(h_hits.numpy().view(gpr.HIT_DTYPE)["id"], hits["id"]))
        t0 = time.perf_counter()
        for _ in range(2):
            scene.trace(rays_np)
        ms_pageable = (time.perf_counter() - t0) / 2 * 1e3
        del h_rays, h_hits
        host = {"note": "PCIe-inclusive: rays copied in and hits out inside each call",
                "pinned_mrays_s": round(N_RAYS / (ms_pinned * 1e-3) / 1e6, 2),
                "pinned_ms": round(ms_pinned, 3),
                "pageable_mrays_s": round(N_RAYS / (ms_pageable * 1e-3) / 1e6, 2),
                "pageable_ms": round(ms_pageable, 3), "hits_equal_device": host_ok}

    value = N_RAYS / (ms_closest * 1e-3) / 1e6
    achieved = bytes_per_ray * N_RAYS / (ms_closest * 1e-3) / 1e9
    out = {
        "metric": "Mrays/s closest-hit @1M tris",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_closest, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded PCG32 triangle soup and rays)",
        "config": {"workload": "lbvh_build_ray_query_1M_tris_4M_rays", "triangles": N_TRI,
                   "rays": N_RAYS, "tmax": "inf", "traversal": "bvh4_128B_lds_stack",
                   "tuned": {"block": tuned[0].block, "stack": tuned[0].stack, "chunk": tuned[0].chunk,
                             "refill": tuned[0].refill,
                             "any_hit": [tuned_any[0].block, tuned_any[0].stack, tuned_any[0].chunk,
                                         tuned_any[0].refill],
                             "default_256_16_64_ms": round(next(c.ms for c in tuned if (c.block, c.stack, c.chunk, c.refill)
                                                                == (256, 16, 64, 0)), 4)}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": pmc_traffic("closest_hit_tuned"),
                     "traffic_kernel": pmc_traffic("closest_hit_tuned", "kernel"),
                     "traffic_note": "FETCH_SIZE x 1.063 (a random 128-B line gather of known size) + WRITE_SIZE; "
                                     "L2-to-fabric bytes, most of them Infinity Cache hits",
                     "kernel": f"k_trace_{'refill' if tuned[0].refill else 'fast'}<false, {tuned[0].block}, {tuned[0].stack}>",
                     "bytes_per_ray": round(bytes_per_ray, 2),
                     "bytes_formula": "32 ray + 16 hit + 128 * BVH4 node fetches + 48 * triangle tests, per ray, "
                                      "counted on the device (gpr_trace_bvh4_stats)",
                     "mean_bvh4_nodes": round(n_node4, 3), "mean_tri_tests": round(n_tri4, 3),
                     "limiter": "VALU issue (79 % busy, 39 % of lanes active per instruction) and node-fetch "
                                "latency (L2 hit 29.6 %); 2.9 TB/s of L2-to-fabric traffic, mostly from the "
                                "Infinity Cache: not HBM bandwidth",
                     "reference_order_diagnostic": {"bytes_per_ray": round(bytes_per_ray_ref, 2),
                                                    "mean_nodes": round(n_node, 3), "mean_tris": round(n_tri, 3),
                                                    "formula": "32 + 16 + 32 * nodes + 36 * tris of the "
                                                               "reference's binary visit order"}},
        "roofline_sort": {"bound": "hbm", "achieved": round(sort_bytes / (sort_ms * 1e-3) / 1e9, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(sort_bytes / (sort_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "ms_per_sort": round(sort_ms, 5), "ms_per_sort_eager": round(sort_ms_eager, 5),
                          "keys": N_TRI, "bytes_per_key": 68, "iters": iters,
                          "traffic": pmc_traffic("sort_1M"),
                          "timing": "mean over back-to-back sorts replayed from one hipGraph; "
                                    "median of 5 replays, each between device synchronisations",
                          "replays_ms_per_sort": [round(x, 5) for x in sort_replays_ms],
                          "graph_result_ok": sort_ok},
        "anyhit_mrays_s": round(N_RAYS / (ms_any * 1e-3) / 1e6, 2),
        "anyhit_ms": round(ms_any, 4),
        "anyhit_roofline": {"bytes_per_ray": round(bytes_per_ray_any, 2),
                            "achieved_gbs": round(bytes_per_ray_any * N_RAYS / (ms_any * 1e-3) / 1e9, 1),
                            "frac": round(bytes_per_ray_any * N_RAYS / (ms_any * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "note": "node / triangle counts of the immediate-leaf walk; the timed kernel "
                                    "postpones leaves into wave-wide rounds"},
        "closest_hit_fraction": round(hit_frac, 4),
        "stack_overflow_rays": int(overflow),
        "build_ms": {"bounds_morton": round(build_ms[0], 4), "sort": round(build_ms[1], 4),
                     "tree_refit": round(build_ms[2], 4), "layouts": round(build_ms[3], 4),
                     "total": round(build_ms[4], 4)},
    }
    if host is not None:
        out["host_buffers"] = host
    if not a.no_skewed_builds:
        print("[bench] skewed-scene builds", file=sys.stderr, flush=True)
        out["build_ms_skewed_scenes"] = skewed_builds(gpr)
    if not a.no_cpu_baseline:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        cpu_nodes = []
        out["cpu_baseline"] = cpu_baseline(pos, rays_np, min(a.cpu_sample, N_RAYS),
                                           min(a.cpu_sample_1core, N_RAYS), ref_hits, cpu_nodes)
        out["build_ms"]["cpu_reference_path_1thread"] = out["cpu_baseline"]["lbvh_build_ms_1thread"]
        # the tree the GPU build timed above is the reference's tree, node for node
        out["build_ms"]["tree_equal_cpu_reference"] = bool(
            np.array_equal(scene.ref_layout().view(np.uint32), cpu_nodes[0].view(np.uint32)))
    if not a.no_serve:
        print("[bench] served request stream", file=sys.stderr, flush=True)
        del scene, d_rays, d_shadow, d_hits
        torch.cuda.synchronize()
        out["served"] = serve_stream(a, gpr, np, torch, a.serve_seconds, a.scenes or 64, max_requests=a.requests)
        out["served"]["config"] = "serve_64_scenes_poisson_200_per_s (64 sphere-grid scenes, 1e4..4e6 triangles)"
        print("[bench] served latency below capacity", file=sys.stderr, flush=True)
        out["served"]["below_capacity"] = serve_latency_under_capacity(a, gpr, np, torch)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
