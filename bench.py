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
time against its HBM roofline, the traversal roofline and a CPU baseline (the
CPU restatement of the reference traversal, timed on this host).
"""
import argparse
import json
import os
import sys
import time

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
    ap.add_argument("--only", choices=["all", "trace", "sort", "copy"], default="all",
                    help="profiling runs: restrict the GPU work to one kernel family")
    return ap.parse_args()


def pmc_traffic(key):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes (copy-calibrated), committed under profiles/; None if absent."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic_pmc.json")) as f:
            return json.load(f)[key]["bytes"]
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


def cpu_baseline(pos, rays_np, sample):
    """Reference-order traversal of the CPU restatement on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    cores = max(1, min(os.cpu_count() or 1, 16))
    nodes, _, _ = O.lbvh(pos)
    rays = rays_np[:sample]
    O.trace(nodes, pos, rays[: min(len(rays), 65536)], nthreads=cores)  # warm
    t0 = time.perf_counter()
    O.trace(nodes, pos, rays, nthreads=cores)
    dt = time.perf_counter() - t0
    return {"value": round(len(rays) / dt / 1e6, 3), "unit": "Mrays/s", "cores": cores,
            "kind": "port",
            "sample": f"{len(rays)} of the {len(rays_np)} closest-hit rays, reference-order "
                      f"traversal of the CPU restatement, {cores} threads, {dt:.2f} s"}


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
    if a.only == "copy":
        copy_calibration(torch)
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
        for _ in range(a.warmup + a.steps):
            scene.trace_device(d_rays.data_ptr(), d_hits.data_ptr(), N_RAYS)
        torch.cuda.synchronize()
        print(json.dumps({"profiled": "closest-hit", "launches": a.warmup + a.steps}), flush=True)
        return
    # ---- radix sort: mean over back-to-back sorts of the scene's Morton keys,
    # captured as one graph of sort_iters sorts (an even count keeps the
    # double-buffered digit totals in phase) and replayed
    keys, _, _ = scene.keys()
    d_keys = torch.from_numpy(keys.view(np.int32)).cuda()
    d_sk = torch.empty_like(d_keys)
    d_si = torch.empty_like(d_keys)
    gpr.sort_pairs_device(d_keys.data_ptr(), d_sk.data_ptr(), d_si.data_ptr(), N_TRI)
    torch.cuda.synchronize()
    ref_sk, ref_si = d_sk.clone(), d_si.clone()
    iters = a.sort_iters + (a.sort_iters & 1)
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
    d_sk.zero_()
    e0.record()
    graph.replay()
    e1.record()
    torch.cuda.synchronize()
    sort_ms = e0.elapsed_time(e1) / iters
    sort_ok = sort_ok and bool(torch.equal(d_sk, ref_sk) and torch.equal(d_si, ref_si))
    # the same sorts launched one by one (no graph), for reference
    e0.record()
    for _ in range(iters):
        gpr.sort_pairs_device(d_keys.data_ptr(), d_sk.data_ptr(), d_si.data_ptr(), N_TRI)
    e1.record()
    torch.cuda.synchronize()
    sort_ms_eager = e0.elapsed_time(e1) / iters
    sort_bytes = 68.0 * N_TRI
    if a.only == "sort":
        print(json.dumps({"profiled": "sort", "iters": iters, "ms_per_sort_graph": sort_ms,
                          "ms_per_sort_eager": sort_ms_eager, "graph_result_ok": sort_ok}), flush=True)
        return

    # ---- traversal statistics of the reference visit order (algorithmic bytes)
    print("[bench] reference-order statistics", file=sys.stderr, flush=True)
    scene.trace_device(d_rays.data_ptr(), d_hits.data_ptr(), N_RAYS, ref_order=True)
    torch.cuda.synchronize()
    nv, nt = scene.ref_stats()
    n_node, n_tri = nv / N_RAYS, nt / N_RAYS
    bytes_per_ray = 32 + 16 + 32 * n_node + 36 * n_tri

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
    t_wall0 = time.perf_counter()
    ms_closest = e0.elapsed_time(e1) / a.steps
    hits = d_hits.cpu().numpy().view(gpr.HIT_DTYPE)
    hit_frac = float((hits["id"] != -1).mean())

    # ---- any-hit shadow pass
    for _ in range(a.warmup):
        scene.trace_device(d_shadow.data_ptr(), d_hits.data_ptr(), N_RAYS, any_hit=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(a.steps):
        scene.trace_device(d_shadow.data_ptr(), d_hits.data_ptr(), N_RAYS, any_hit=True)
    e1.record()
    torch.cuda.synchronize()
    ms_any = e0.elapsed_time(e1) / a.steps
    del t_wall0

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
                   "rays": N_RAYS, "tmax": "inf", "traversal": "bvh4_128B_lds_stack"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": pmc_traffic("k_trace_fast<false>"),
                     "kernel": "k_trace_fast<false>",
                     "bytes_per_ray": round(bytes_per_ray, 2),
                     "mean_nodes": round(n_node, 3), "mean_tris": round(n_tri, 3)},
        "roofline_sort": {"bound": "hbm", "achieved": round(sort_bytes / (sort_ms * 1e-3) / 1e9, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(sort_bytes / (sort_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "ms_per_sort": round(sort_ms, 5), "ms_per_sort_eager": round(sort_ms_eager, 5),
                          "keys": N_TRI, "bytes_per_key": 68, "iters": iters,
                          "traffic": pmc_traffic("sort_1M"),
                          "timing": "mean over back-to-back sorts replayed from one hipGraph",
                          "graph_result_ok": sort_ok},
        "anyhit_mrays_s": round(N_RAYS / (ms_any * 1e-3) / 1e6, 2),
        "anyhit_ms": round(ms_any, 4),
        "closest_hit_fraction": round(hit_frac, 4),
        "build_ms": {"bounds_morton": round(build_ms[0], 4), "sort": round(build_ms[1], 4),
                     "tree_refit": round(build_ms[2], 4), "layouts": round(build_ms[3], 4),
                     "total": round(build_ms[4], 4)},
    }
    if not a.no_cpu_baseline:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(pos, rays_np, min(a.cpu_sample, N_RAYS))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
