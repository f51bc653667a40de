"""Host logic of the benchmark (no GPU): the served request trace and the
CPU-baseline bookkeeping."""
import numpy as np

import bench
import gpr


def test_request_trace_is_poisson_200_per_s_over_60_s():
    ids = list(range(64))
    arrivals, reqs = bench.c5_requests(gpr, ids, 60.0)
    n = len(reqs)
    # Poisson count with mean 12,000: within 4 standard deviations
    assert abs(n - 12000) < 4 * np.sqrt(12000)
    assert np.all(np.diff(arrivals) > 0) and 0 < arrivals[0] and arrivals[-1] < 60.0
    gaps = np.diff(arrivals)
    assert abs(gaps.mean() - 1 / 200) < 0.1 / 200
    assert {r.width for r in reqs} == {256, 512, 1024} and all(r.width == r.height for r in reqs)
    assert {r.spp for r in reqs} == {4, 16, 64}
    assert {r.scene for r in reqs} == set(ids)
    assert len({r.seed for r in reqs}) == n and all(r.max_depth == 8 and r.rr_start == 3 for r in reqs)
    # seeded: the same trace every time, a prefix under a request cap
    a2, r2 = bench.c5_requests(gpr, ids, 60.0, max_requests=100)
    assert len(r2) == 100 and np.array_equal(a2, arrivals[:100])
    assert [r.seed for r in r2] == [r.seed for r in reqs[:100]]


def test_server_budget_covers_queues_of_the_capacity():
    for cap in (1 << 24, 1 << 26):
        assert bench.server_budget(cap) >= 3 * 84 * cap + (1 << 30)


def test_host_cpu_info_fields():
    info = bench.host_cpu_info()
    assert info["nproc"] >= 1 and 1 <= info["affinity"] <= info["nproc"] and isinstance(info["model"], str)
