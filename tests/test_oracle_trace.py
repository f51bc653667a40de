"""Pins the oracle's Moller-Trumbore and reference-order traversal
(pathtracer_kernel.h:119-308) against closed forms and brute force (no GPU)."""
import numpy as np

import oracle_py as O


def one_ray(o, d, tmin=0.0, tmax=np.inf):
    r = np.zeros(1, O.RAY_DTYPE)
    r["o"], r["d"], r["tmin"], r["tmax"] = o, d, tmin, tmax
    return r


def test_intersect_known_answer():
    p0, p1, p2 = (np.array(v, np.float32) for v in ([0, 0, 0], [1, 0, 0], [0, 1, 0]))
    r = one_ray([0.25, 0.25, 1], [0, 0, -1])
    t, u, v = (np.zeros(1, np.float32) for _ in range(3))
    assert O.lib.orc_intersect_ref(O.ptr(p0), O.ptr(p1), O.ptr(p2), O.ptr(r), O.ptr(t), O.ptr(u), O.ptr(v))
    assert (t[0], u[0], v[0]) == (1.0, 0.25, 0.25)
    # no back-face culling, t == tmax accepted, t < tmin rejected, parallel rejected
    r = one_ray([0.25, 0.25, -1], [0, 0, 1], 0, 1.0)
    assert O.lib.orc_intersect_ref(O.ptr(p0), O.ptr(p1), O.ptr(p2), O.ptr(r), O.ptr(t), O.ptr(u), O.ptr(v))
    r = one_ray([0.25, 0.25, -1], [0, 0, 1], 1.5, 9)
    assert not O.lib.orc_intersect_ref(O.ptr(p0), O.ptr(p1), O.ptr(p2), O.ptr(r), O.ptr(t), O.ptr(u), O.ptr(v))
    r = one_ray([0.25, 0.25, 1], [1, 0, 0])
    assert not O.lib.orc_intersect_ref(O.ptr(p0), O.ptr(p1), O.ptr(p2), O.ptr(r), O.ptr(t), O.ptr(u), O.ptr(v))
    r = one_ray([0.75, 0.75, 1], [0, 0, -1])  # u+v > 1
    assert not O.lib.orc_intersect_ref(O.ptr(p0), O.ptr(p1), O.ptr(p2), O.ptr(r), O.ptr(t), O.ptr(u), O.ptr(v))


def soup(n, seed=3):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3)).astype(np.float32)
    return (np.repeat(c, 3, axis=0) + rng.normal(0, 0.03, (3 * n, 3)).astype(np.float32)).reshape(n, 9)


def rays(n, seed=4, tmax=np.inf):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, O.RAY_DTYPE)
    r["o"] = rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    r["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmin"] = 0
    r["tmax"] = tmax
    return r


def test_reference_traversal_equals_brute_force():
    pos = soup(1500)
    nodes, _, _ = O.lbvh(pos)
    r = rays(3000)
    h, nv, nt = O.trace(nodes, pos, r)
    b, _ = O.brute(pos, r)
    same = h["id"] == b["id"]
    # disagreements are exact ties of t only
    assert np.all(h["t"][~same] == b["t"][~same])
    assert same.mean() > 0.999
    assert np.array_equal(h["t"], b["t"])
    assert nv > 0 and nt > 0


def test_any_hit_occlusion_equals_closest_hit():
    pos = soup(1500)
    nodes, _, _ = O.lbvh(pos)
    r = rays(3000, tmax=0.3)
    a, _, _ = O.trace(nodes, pos, r, any_hit=True)
    c, _, _ = O.trace(nodes, pos, r)
    assert np.array_equal(a["id"] != -1, c["id"] != -1)


def test_fp64_brute_force_close_to_reference_arithmetic():
    pos = soup(800)
    r = rays(2000)
    b32, _ = O.brute(pos, r)
    b64, t64 = O.brute(pos, r, fp64=True)
    hit = (b32["id"] != -1) & (b64["id"] != -1)
    assert (b32["id"] == b64["id"]).mean() > 0.999
    rel = np.abs(b32["t"][hit] - t64[hit]) / t64[hit]
    assert rel.max() < 1e-5
