"""ctypes binding of oracle/liboracle.so (the CPU restatement used as checker)."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# GPR_ORACLE_LIB names another build of the same source (the benchmark's CPU
# baseline compiles one with -O3 -march=native for the host it runs on)
PATH = os.environ.get("GPR_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")

NODE_DTYPE = np.dtype([("bmin", "<f4", 3), ("bmax", "<f4", 3), ("idx", "<u4"), ("leaf", "<u4")])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("tmin", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("id", "<i4"), ("u", "<f4"), ("v", "<f4")])
MATERIAL_DTYPE = np.dtype([("type", "<u4"), ("color", "<f4", 3), ("color1", "<f4", 3),
                           ("emission", "<f4", 3), ("ior", "<f4"), ("alpha_x", "<f4"),
                           ("alpha_y", "<f4"), ("eta_i", "<f4", 3), ("eta_t", "<f4", 3),
                           ("k", "<f4", 3), ("tex", "<u4", 8)])


def _build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


if not os.path.exists(PATH):
    _build()
lib = C.CDLL(PATH)
P = C.c_void_p


class Scene(C.Structure):
    _fields_ = [("ntri", C.c_uint32), ("pos", P), ("nrm", P), ("uv", P), ("flags", P), ("mat", P),
                ("mats", P), ("nmat", C.c_uint32), ("c2w", C.c_float * 16), ("c2w_inv", C.c_float * 16),
                ("fov_deg", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32),
                ("has_env", C.c_uint32), ("env", C.c_float * 3),
                ("cam_free", C.c_uint32), ("eye", C.c_float * 3), ("dir", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_x", C.c_float), ("fov_y", C.c_float),
                ("tex", P), ("texels", P), ("ntex", C.c_uint32), ("env_tex", C.c_uint32),
                ("env_w2o", C.c_float * 9), ("env_center", C.c_float * 3)]


class RenderOpts(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("max_depth", C.c_uint32), ("rng_mode", C.c_uint32),
                ("rr_start", C.c_uint32), ("seed", C.c_uint64), ("first_sample", C.c_uint32)]


def ptr(a):
    return None if a is None else a.ctypes.data_as(P)


lib.orc_sinf.restype = C.c_float
lib.orc_sinf.argtypes = [C.c_float]
lib.orc_cosf.restype = C.c_float
lib.orc_cosf.argtypes = [C.c_float]
lib.orc_left_shift3.restype = C.c_uint32
lib.orc_left_shift3.argtypes = [C.c_uint32]
lib.orc_encode_morton3.restype = C.c_uint32
lib.orc_encode_morton3.argtypes = [C.c_float] * 3
lib.orc_centroid_bounds.argtypes = [P, C.c_uint32, P, P]
lib.orc_morton_keys.argtypes = [P, C.c_uint32, P]
lib.orc_sort_pairs.argtypes = [P, C.c_uint32, P, P]
lib.orc_lbvh_build.restype = C.c_uint32
lib.orc_lbvh_build.argtypes = [P, C.c_uint32, P, P, P]
lib.orc_lbvh_ranges.argtypes = [P, C.c_uint32, P, P]
lib.orc_intersect_ref.restype = C.c_int
lib.orc_intersect_ref.argtypes = [P, P, P, P, P, P, P]
lib.orc_trace.argtypes = [P, P, C.c_uint32, P, C.c_uint64, C.c_int, P, P, P, C.c_int]
lib.orc_brute.argtypes = [P, C.c_uint32, P, C.c_uint64, C.c_int, C.c_int, P, P, C.c_int]
lib.orc_tea_keyed.argtypes = [P, P, P, C.c_uint32]
lib.orc_init_random.restype = C.c_uint32
lib.orc_init_random.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
lib.orc_next_random.restype = C.c_float
lib.orc_next_random.argtypes = [P]
lib.orc_pcg32_seed.argtypes = [P, P, C.c_uint64, C.c_uint64]
lib.orc_pcg32_next.restype = C.c_uint32
lib.orc_pcg32_next.argtypes = [P, C.c_uint64]
lib.orc_lookat.argtypes = [P, P, P, P, P]
lib.orc_matrix_inverse.argtypes = [P, P]
lib.orc_camera_ray.argtypes = [C.POINTER(Scene), C.c_float, C.c_float, P]
lib.orc_bsdf_eval.argtypes = [P, P, P, P, P, P]
lib.orc_bsdf_sample.restype = C.c_int
lib.orc_bsdf_sample.argtypes = [P, P, P, C.c_float, C.c_float, C.c_float, P, P, P, P]
lib.orc_fr_dielectric.restype = C.c_float
lib.orc_fr_dielectric.argtypes = [C.c_float] * 3
lib.orc_fr_conductor.argtypes = [C.c_float, P, P, P, P]
lib.orc_ggx_d.restype = C.c_float
lib.orc_ggx_d.argtypes = [P, C.c_float, C.c_float]
lib.orc_ggx_g.restype = C.c_float
lib.orc_ggx_g.argtypes = [P, P, C.c_float, C.c_float]
lib.orc_ggx_lambda.restype = C.c_float
lib.orc_ggx_lambda.argtypes = [P, C.c_float, C.c_float]
lib.orc_ggx_d_n.restype = None
lib.orc_ggx_d_n.argtypes = [P, C.c_uint32, C.c_float, C.c_float, P]
lib.orc_ggx_sample_wh.restype = None
lib.orc_ggx_sample_wh.argtypes = [P, P, C.c_uint32, C.c_float, C.c_float, P]
lib.orc_tex_eval.restype = None
lib.orc_tex_eval.argtypes = [C.POINTER(Scene), C.c_uint32, C.c_float, C.c_float, P]
lib.orc_env_le.restype = None
lib.orc_env_le.argtypes = [C.POINTER(Scene), P, P, P]
lib.orc_render.restype = C.c_int
lib.orc_render.argtypes = [C.POINTER(Scene), P, C.POINTER(RenderOpts), P, P, C.c_int]
lib.orc_render_ties.restype = C.c_int
lib.orc_render_ties.argtypes = [C.POINTER(Scene), P, C.POINTER(RenderOpts), P, P, P, C.c_int]

NTHREADS = min(8, os.cpu_count() or 1)


def morton_keys(pos):
    pos = np.ascontiguousarray(pos, np.float32)
    n = len(pos)
    k = np.empty(n, np.uint32)
    lib.orc_morton_keys(ptr(pos), n, ptr(k))
    return k


def sort_pairs(keys):
    keys = np.ascontiguousarray(keys, np.uint32)
    sk = np.empty_like(keys)
    si = np.empty_like(keys)
    lib.orc_sort_pairs(ptr(keys), len(keys), ptr(sk), ptr(si))
    return sk, si


def lbvh(pos):
    pos = np.ascontiguousarray(pos, np.float32)
    n = len(pos)
    nodes = np.zeros(max(2 * n - 1, 0), NODE_DTYPE)
    sk = np.empty(n, np.uint32)
    si = np.empty(n, np.uint32)
    cnt = lib.orc_lbvh_build(ptr(pos), n, ptr(nodes), ptr(sk), ptr(si))
    assert cnt == len(nodes)
    return nodes, sk, si


def ranges(skeys):
    skeys = np.ascontiguousarray(skeys, np.uint32)
    n = len(skeys)
    lo = np.empty(2 * n - 1, np.uint32)
    hi = np.empty(2 * n - 1, np.uint32)
    lib.orc_lbvh_ranges(ptr(skeys), n, ptr(lo), ptr(hi))
    return lo, hi


def trace(nodes, pos, rays, any_hit=False, nthreads=NTHREADS):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.empty(len(rays), HIT_DTYPE)
    nv = C.c_uint64()
    nt = C.c_uint64()
    pos = np.ascontiguousarray(pos, np.float32)
    lib.orc_trace(ptr(nodes), ptr(pos), len(pos), ptr(rays), len(rays), int(any_hit), ptr(hits),
                  C.byref(nv), C.byref(nt), nthreads)
    return hits, nv.value, nt.value


def brute(pos, rays, any_hit=False, fp64=False, nthreads=NTHREADS):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    pos = np.ascontiguousarray(pos, np.float32)
    hits = np.empty(len(rays), HIT_DTYPE)
    t64 = np.empty(len(rays), np.float64)
    lib.orc_brute(ptr(pos), len(pos), ptr(rays), len(rays), int(any_hit), int(fp64), ptr(hits),
                  ptr(t64), nthreads)
    return hits, t64


def make_scene(flat, c2w, fov, w, h, env=None, info=None):
    """orc_scene from the flat arrays of gpr.Scene.flat(); keeps the arrays alive.
    With `info` (gpr_scene_info) the free-camera state is taken over as well."""
    s = Scene()
    keep = {}
    for k in ("pos", "nrm", "uv", "flags", "mat"):
        keep[k] = np.ascontiguousarray(flat[k])
    keep["mats"] = np.ascontiguousarray(flat["mats"], MATERIAL_DTYPE)
    s.ntri = len(keep["mat"])
    s.pos, s.nrm, s.uv = ptr(keep["pos"]), ptr(keep["nrm"]), ptr(keep["uv"])
    s.flags, s.mat, s.mats = ptr(keep["flags"]), ptr(keep["mat"]), ptr(keep["mats"])
    s.nmat = len(keep["mats"])
    s.c2w[:] = [float(x) for x in np.asarray(c2w, np.float32).ravel()]
    s.fov_deg = fov
    s.width, s.height = w, h
    if env is not None:
        s.has_env = 1
        s.env[:] = env
    if "tex" in flat and len(flat["tex"]):
        keep["tex"] = np.ascontiguousarray(flat["tex"])
        keep["texels"] = np.ascontiguousarray(flat["texels"], np.float32)
        s.tex, s.texels, s.ntex = ptr(keep["tex"]), ptr(keep["texels"]), len(keep["tex"])
    if info is not None:
        s.env_tex = info.env_tex
        s.env_w2o[:] = list(info.env_w2o)
        s.env_center[:] = list(info.env_center)
    if info is not None and info.cam_free:
        s.cam_free = 1
        s.eye[:] = list(info.eye)
        s.dir[:] = list(info.dir)
        s.up[:] = list(info.up)
        s.fov_x, s.fov_y = info.fov_x, info.fov_y
    s._keep = keep
    return s


def render(scene, nodes, spp, max_depth, rng_mode, seed=1, rr_start=3, first_sample=0,
           nthreads=NTHREADS):
    n = scene.width * scene.height
    rgb = np.zeros((n, 3), np.float32)
    cnt = np.zeros(n, np.uint32)
    o = RenderOpts(spp, max_depth, rng_mode, rr_start, seed, first_sample)
    rc = lib.orc_render(C.byref(scene), ptr(nodes), C.byref(o), ptr(rgb), ptr(cnt), nthreads)
    assert rc == 0
    return rgb, cnt


def render_ties(scene, nodes, spp, max_depth, seed=1, rr_start=3, first_sample=0, nthreads=NTHREADS):
    """pcg-mode render plus a per-pixel flag: some closest hit on the pixel's
    paths had another triangle within 1e-6 relative of its t."""
    n = scene.width * scene.height
    rgb = np.zeros((n, 3), np.float32)
    cnt = np.zeros(n, np.uint32)
    ties = np.zeros(n, np.uint8)
    o = RenderOpts(spp, max_depth, 1, rr_start, seed, first_sample)
    rc = lib.orc_render_ties(C.byref(scene), ptr(nodes), C.byref(o), ptr(rgb), ptr(cnt), ptr(ties), nthreads)
    assert rc == 0
    return rgb, cnt, ties.astype(bool)
