"""The C-ABI library loads and exports every entry point include/gpr.h
declares; host-only entry points behave without a GPU."""
import ctypes as C
import os
import re

import numpy as np

import gpr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "gpr.h")).read()
    return sorted(set(re.findall(r"\b(gpr_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) > 25
    missing = [n for n in names if not hasattr(gpr.lib, n)]
    assert not missing, missing


def test_generators_deterministic_and_in_range():
    a = gpr.gen_soup(2, 0, 1000)
    b = gpr.gen_soup(2, 0, 1000)
    c = gpr.gen_soup(2, 1, 1000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    centres = a.reshape(-1, 3, 3).mean(axis=1)
    assert np.all((centres > -0.05) & (centres < 1.05))
    jitter = (a.reshape(-1, 3, 3) - centres[:, None, :]).ravel()
    assert 0.003 < jitter.std() < 0.007
    r = gpr.gen_rays(2, 1, 20000)
    assert np.allclose(np.linalg.norm(r["d"], axis=1), 1, atol=1e-6)
    assert abs(r["d"].mean()) < 0.02 and np.all(r["tmin"] == 0) and np.all(np.isinf(r["tmax"]))
    s = gpr.gen_shadow_rays(2, 2, r)
    tgt = s["o"] + s["d"]
    assert np.all((tgt >= 0) & (tgt < 1)) and np.all(s["tmax"] == np.float32(0.9999))


def test_mesh_validation_errors_without_gpu():
    s = gpr.Scene()
    p = np.zeros((3, 3), np.float32)
    bad = np.array([[[0, -1, -1], [1, -1, -1], [5, -1, -1]]], np.int32)
    try:
        s.add_mesh(p, corners=bad)
        raise AssertionError("out-of-range corner accepted")
    except gpr.GprError as e:
        assert e.code == gpr.ERR_INVALID
    mixed = np.array([[[0, -1, 0], [1, -1, -1], [2, -1, 0]]], np.int32)
    try:
        s.add_mesh(p, corners=mixed, n=np.ones((1, 3), np.float32))
        raise AssertionError("inconsistent normal indices accepted")
    except gpr.GprError as e:
        assert e.code == gpr.ERR_INVALID
    s.add_mesh(p)  # valid
    assert gpr.lib.gpr_last_error() is not None


def test_set_device_rejects_indices_outside_the_tables():
    # checked before any HIP call, so this runs without a GPU
    for bad in (-1, 16, 1 << 20):
        assert gpr.lib.gpr_set_device(bad) == gpr.ERR_INVALID
        assert b"0..15" in gpr.lib.gpr_last_error()
