"""Pins the oracle's LookAt / Perspective / raster-to-camera transform
(transform.cpp, camera.h) against pinhole geometry (no GPU)."""
import math

import numpy as np
import pytest

import oracle_py as O


def lookat(target, origin, up):
    m = np.zeros(16, np.float32)
    mi = np.zeros(16, np.float32)
    O.lib.orc_lookat(O.ptr(np.asarray(target, np.float32)), O.ptr(np.asarray(origin, np.float32)),
                     O.ptr(np.asarray(up, np.float32)), O.ptr(m), O.ptr(mi))
    return m.reshape(4, 4), mi.reshape(4, 4)


def test_lookat_frame_and_inverse():
    m, mi = lookat([0, 0, 0], [1, 2, 5], [0, 1, 0])
    fwd = -np.array([1, 2, 5], np.float64) / math.sqrt(30)
    assert np.allclose(m[:3, 2], fwd, atol=1e-6)
    assert np.allclose(m[:3, 3], [1, 2, 5])
    left = np.cross([0, 1, 0], fwd)
    left /= np.linalg.norm(left)
    assert np.allclose(m[:3, 0], left, atol=1e-6)
    assert np.allclose(m[:3, 1], np.cross(fwd, left), atol=1e-6)
    assert np.allclose(mi.astype(np.float64), np.linalg.inv(m.astype(np.float64)), atol=1e-5)


def test_matrix_inverse_general():
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = rng.normal(size=(4, 4)).astype(np.float32)
        out = np.zeros(16, np.float32)
        O.lib.orc_matrix_inverse(O.ptr(a.ravel().copy()), O.ptr(out))
        assert np.allclose(out.reshape(4, 4) @ a, np.eye(4), atol=1e-3)


def cam(w, h, fov, target=(0, 0, 0), origin=(0, 0, 5), up=(0, 1, 0)):
    m, _ = lookat(target, origin, up)
    s = O.Scene()
    s.c2w[:] = [float(x) for x in m.ravel()]
    s.fov_deg = fov
    s.width, s.height = w, h
    return s, m


def ray(s, x, y):
    r = np.zeros(1, O.RAY_DTYPE)
    O.lib.orc_camera_ray(s, x, y, O.ptr(r))
    return r[0]


@pytest.mark.parametrize("w,h,fov", [(256, 256, 40.0), (320, 200, 30.0), (64, 128, 60.0)])
def test_pinhole_geometry(w, h, fov):
    s, m = cam(w, h, fov)
    left, up, fwd = m[:3, 0].astype(np.float64), m[:3, 1].astype(np.float64), m[:3, 2].astype(np.float64)
    c = ray(s, w / 2, h / 2)
    assert np.allclose(c["o"], [0, 0, 5], atol=1e-6)
    assert np.allclose(c["d"], fwd, atol=1e-6)
    # raster x = 0 is camera +x (the left vector) at half the field of view
    d = ray(s, 0, h / 2)["d"].astype(np.float64)
    assert abs(math.degrees(math.acos(np.dot(d, fwd))) - fov / 2) < 1e-3
    assert np.dot(d, left) > 0 and abs(np.dot(d, up)) < 1e-6
    # raster y = 0 looks down; vertical half-angle atan(tan(fov/2)/aspect)
    d = ray(s, w / 2, 0)["d"].astype(np.float64)
    want = math.degrees(math.atan(math.tan(math.radians(fov / 2)) / (w / h)))
    assert abs(math.degrees(math.acos(np.dot(d, fwd))) - want) < 1e-3
    assert np.dot(d, up) < 0
    assert abs(np.linalg.norm(d) - 1) < 1e-6
