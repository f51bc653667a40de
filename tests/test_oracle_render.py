"""Pins the oracle's integrator (pathtracer.cu:41-150 restated) with the
closed-box furnace: all walls Lambert (albedo rho) and emitting Le inwards,
so L = Le * sum_{k<=D} rho^k exactly in expectation (no GPU)."""
import numpy as np
import pytest

import oracle_py as O


def closed_box(rho, le):
    # 12 triangles of the cube [-1,1]^3 wound so the geometric normal points inwards
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, c, d in quads:
        for t in ((a, b, c), (a, c, d)):
            p = v[list(t)]
            n = np.cross(p[0] - p[2], p[1] - p[2])
            centre = p.mean(axis=0)
            if np.dot(n, -centre) < 0:  # make the normal face the box centre
                p = p[[1, 0, 2]]
            tris.append(p.ravel())
    pos = np.array(tris, np.float32)
    n = len(pos)
    mats = np.zeros(1, O.MATERIAL_DTYPE)
    mats["type"] = 0
    mats["color"] = rho
    mats["emission"] = le
    flat = dict(pos=pos, nrm=np.zeros_like(pos), uv=np.zeros((n, 6), np.float32),
                flags=np.zeros(n, np.uint8), mat=np.zeros(n, np.uint32), mats=mats)
    c2w = np.eye(4, dtype=np.float32)  # camera at the centre looking down +z
    return flat, c2w


@pytest.mark.parametrize("rho", [0.3, 0.6])
def test_furnace_pcg_unbiased(rho):
    flat, c2w = closed_box(rho, 1.0)
    s = O.make_scene(flat, c2w, 60.0, 16, 16)
    nodes, _, _ = O.lbvh(flat["pos"])
    depth = 40
    rgb, cnt = O.render(s, nodes, 64, depth, 1, seed=5, rr_start=3)
    est = rgb.sum(axis=0) / cnt.sum()
    want = sum(rho ** k for k in range(depth + 1))
    assert np.allclose(est, want, rtol=0.03), (est, want)


def test_reference_mode_roulette_is_biased_low():
    # the reference terminates paths at bounce > 3 without reweighting the survivors
    rho, depth = 0.6, 40
    flat, c2w = closed_box(rho, 1.0)
    s = O.make_scene(flat, c2w, 60.0, 16, 16)
    nodes, _, _ = O.lbvh(flat["pos"])
    rgb, cnt = O.render(s, nodes, 64, depth, 0)
    est = float(rgb.sum() / cnt.sum() / 3)
    want = sum(rho ** k for k in range(depth + 1))
    assert est < 0.97 * want


def test_deterministic():
    flat, c2w = closed_box(0.5, 1.0)
    s = O.make_scene(flat, c2w, 60.0, 8, 8)
    nodes, _, _ = O.lbvh(flat["pos"])
    a, _ = O.render(s, nodes, 4, 8, 1, seed=3)
    b, _ = O.render(s, nodes, 4, 8, 1, seed=3, nthreads=1)
    c, _ = O.render(s, nodes, 4, 8, 1, seed=4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
