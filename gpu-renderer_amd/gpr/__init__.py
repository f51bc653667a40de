"""Python binding (ctypes) of the gpr C ABI declared in include/gpr.h.

The native library gpr/libgpr.so (hand-written HIP for gfx950 + C++ host code)
is the product; this module only marshals numpy arrays through the C ABI, the
way a maintainer of the reference would bind it.  Importing fails loudly if the
native library has not been built: there is no Python or CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpr.so")

OK = 0
ERR_INVALID, ERR_OOM, ERR_HIP, ERR_IO, ERR_PARSE, ERR_UNSUPPORTED, ERR_BUDGET, ERR_CANCELLED = -1, -2, -3, -4, -5, -6, -7, -8

MAT_DIFFUSE, MAT_SPECULAR, MAT_ROUGH_CONDUCTOR, MAT_ROUGH_DIELECTRIC, MAT_ROUGH_PLASTIC = 0, 1, 2, 3, 4
TRACE_ANY, TRACE_REF_ORDER, TRACE_DEVICE_PTRS = 1, 2, 4
RNG_REFERENCE, RNG_PCG = 0, 1
TRAVERSAL_FAST, TRAVERSAL_REF = 0, 1

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("tmin", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("id", "<i4"), ("u", "<f4"), ("v", "<f4")])
NODE_DTYPE = np.dtype([("bmin", "<f4", 3), ("bmax", "<f4", 3), ("idx", "<u4"), ("leaf", "<u4")])


class Material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("color", C.c_float * 3), ("color1", C.c_float * 3),
                ("emission", C.c_float * 3), ("ior", C.c_float), ("alpha_x", C.c_float),
                ("alpha_y", C.c_float), ("eta_i", C.c_float * 3), ("eta_t", C.c_float * 3),
                ("k", C.c_float * 3), ("tex", C.c_uint32 * 8)]


MATERIAL_DTYPE = np.dtype([("type", "<u4"), ("color", "<f4", 3), ("color1", "<f4", 3),
                           ("emission", "<f4", 3), ("ior", "<f4"), ("alpha_x", "<f4"),
                           ("alpha_y", "<f4"), ("eta_i", "<f4", 3), ("eta_t", "<f4", 3),
                           ("k", "<f4", 3), ("tex", "<u4", 8)])
assert MATERIAL_DTYPE.itemsize == C.sizeof(Material)
# texture slots of Material.tex (a handle from Scene.add_texture; 0 = the constant field)
TEX_COLOR, TEX_COLOR1, TEX_IOR, TEX_ALPHA_X, TEX_ALPHA_Y, TEX_ETA_I, TEX_ETA_T, TEX_K = range(8)
RESOLVE_DEVICE_PTRS = 1
TEXTURE_DTYPE = np.dtype([("offset", "<u4"), ("width", "<u4"), ("height", "<u4"), ("pad", "<u4"),
                          ("uv_offset", "<f4", 2), ("uv_scale", "<f4", 2)])


class SceneInfo(C.Structure):
    _fields_ = [("ntri", C.c_uint32), ("nlights", C.c_uint32), ("nmat", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("fov_deg", C.c_float),
                ("c2w", C.c_float * 16), ("has_env", C.c_uint32), ("env", C.c_float * 3),
                ("bvh4_nodes", C.c_uint32), ("committed", C.c_uint32),
                ("cam_free", C.c_uint32), ("eye", C.c_float * 3), ("dir", C.c_float * 3),
                ("up", C.c_float * 3), ("fov_x", C.c_float), ("fov_y", C.c_float),
                ("yaw", C.c_float), ("pitch", C.c_float),
                ("ntex", C.c_uint32), ("ntexel", C.c_uint32), ("env_tex", C.c_uint32),
                ("env_w2o", C.c_float * 9), ("env_center", C.c_float * 3)]


class RenderOpts(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("rng_mode", C.c_uint32), ("rr_start", C.c_uint32),
                ("seed", C.c_uint64), ("traversal", C.c_uint32)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("max_depth", C.c_uint32), ("rr_start", C.c_uint32),
                ("seed", C.c_uint64), ("fov_deg", C.c_float), ("c2w", C.c_float * 16)]


class TraceCfg(C.Structure):
    _fields_ = [("block", C.c_uint32), ("stack", C.c_uint32), ("chunk", C.c_uint32),
                ("layout", C.c_uint32), ("refill", C.c_uint32), ("ms", C.c_float)]

    def __repr__(self):
        return (f"TraceCfg(block={self.block}, stack={self.stack}, chunk={self.chunk}, "
                f"layout={self.layout}, refill={self.refill}, ms={self.ms:.3f})")


class BatchOpts(C.Structure):
    _fields_ = [("arena_bytes", C.c_uint64), ("capacity", C.c_uint32), ("max_jobs", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("passes", C.c_uint64), ("paths", C.c_uint64), ("ms", C.c_double)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("arena_bytes", C.c_uint64), ("capacity", C.c_uint32),
                ("max_jobs", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("passes", C.c_uint64),
                ("renders_per_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_jobs_per_pass", C.c_double), ("cancelled", C.c_uint64)]


class GprError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gpr error {code}: {msg}")
        self.code = code


def _load():
    # One HIP runtime per process: when PyTorch is installed its bundled HIP /
    # HSA runtime is loaded first, and libgpr's libamdhip64.so.7 dependency
    # then binds to that same copy (matching soname) instead of a second one.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"native library {LIB_PATH} is not built; run __graft_entry__.build()")
    return C.CDLL(LIB_PATH)


lib = _load()
_P = C.c_void_p
_SIG = {
    "gpr_last_error": (C.c_char_p, []),
    "gpr_set_device": (C.c_int, [C.c_int]),
    "gpr_set_stream": (C.c_int, [_P]),
    "gpr_synchronize": (C.c_int, []),
    "gpr_scene_create": (C.c_int, [C.POINTER(_P)]),
    "gpr_scene_add_mesh": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _P,
                                     C.c_uint32, C.POINTER(Material)]),
    "gpr_scene_set_camera": (C.c_int, [_P, _P, C.c_float, C.c_uint32, C.c_uint32]),
    "gpr_scene_set_camera_lookat": (C.c_int, [_P, _P, _P, _P, C.c_float, C.c_uint32, C.c_uint32]),
    "gpr_scene_set_camera_free": (C.c_int, [_P, _P, _P, _P, C.c_float, C.c_uint32, C.c_uint32]),
    "gpr_renderer_zoom": (C.c_int, [_P, C.c_float]),
    "gpr_renderer_translate": (C.c_int, [_P, C.c_float, C.c_float]),
    "gpr_renderer_rotate": (C.c_int, [_P, C.c_float, C.c_float]),
    "gpr_renderer_resize": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "gpr_renderer_get_resolution": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "gpr_scene_set_env": (C.c_int, [_P, _P]),
    "gpr_scene_add_texture": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, _P, C.POINTER(C.c_uint32)]),
    "gpr_scene_add_checkerboard": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(C.c_uint32)]),
    "gpr_scene_set_env_map": (C.c_int, [_P, C.c_uint32, _P]),
    "gpr_scene_export_textures": (C.c_int, [_P, _P, _P]),
    "gpr_image_read": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]),
    "gpr_image_free": (None, [C.POINTER(C.c_float)]),
    "gpr_scene_commit": (C.c_int, [_P]),
    "gpr_scene_free": (C.c_int, [_P]),
    "gpr_scene_get_info": (C.c_int, [_P, C.POINTER(SceneInfo)]),
    "gpr_scene_export_flat": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "gpr_scene_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "gpr_scene_parse": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "gpr_scene_save": (C.c_int, [_P, C.c_char_p]),
    "gpr_scene_flatten": (C.c_int, [_P]),
    "gpr_bvh_build": (C.c_int, [_P]),
    "gpr_bvh_export_keys": (C.c_int, [_P, _P, _P, _P]),
    "gpr_bvh_export_ref_layout": (C.c_int, [_P, _P, C.c_size_t]),
    "gpr_bvh_last_timing": (C.c_int, [_P, _P]),
    "gpr_sort_pairs_device": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "gpr_sort_pairs_device_noalloc": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "gpr_trace": (C.c_int, [_P, _P, _P, C.c_uint64, C.c_uint32]),
    "gpr_trace_tune": (C.c_int, [_P, _P, C.c_uint64, C.c_uint32, C.POINTER(TraceCfg), C.c_uint32,
                                 C.POINTER(C.c_uint32)]),
    "gpr_trace_set_cfg": (C.c_int, [_P, C.POINTER(TraceCfg)]),
    "gpr_trace_ref_stats": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "gpr_trace_bvh4_stats": (C.c_int, [_P, _P, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]),
    "gpr_trace_overflow_count": (C.c_int, [C.POINTER(C.c_uint64)]),
    "gpr_release_stream": (C.c_int, [_P]),
    "gpr_renderer_create": (C.c_int, [_P, C.POINTER(RenderOpts), C.POINTER(_P)]),
    "gpr_render": (C.c_int, [_P, C.c_uint32]),
    "gpr_renderer_read": (C.c_int, [_P, _P, _P]),
    "gpr_renderer_write_png": (C.c_int, [_P, C.c_char_p]),
    "gpr_renderer_write_image": (C.c_int, [_P, C.c_char_p]),
    "gpr_renderer_restore": (C.c_int, [_P, _P, C.c_uint32]),
    "gpr_renderer_resolve_rgba8": (C.c_int, [_P, _P, C.c_uint32]),
    "gpr_renderer_reset": (C.c_int, [_P]),
    "gpr_renderer_set_opts": (C.c_int, [_P, C.POINTER(RenderOpts)]),
    "gpr_renderer_queue_sizes": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "gpr_renderer_free": (C.c_int, [_P]),
    "gpr_scene_arena_bytes": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "gpr_batch_create": (C.c_int, [C.POINTER(BatchOpts), C.POINTER(_P)]),
    "gpr_batch_add_scene": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "gpr_batch_add_job": (C.c_int, [_P, C.POINTER(Job), C.POINTER(C.c_uint32)]),
    "gpr_batch_tune": (C.c_int, [_P, C.POINTER(TraceCfg), C.c_uint32, C.POINTER(C.c_uint32)]),
    "gpr_batch_set_cfg": (C.c_int, [_P, C.POINTER(TraceCfg)]),
    "gpr_batch_run": (C.c_int, [_P]),
    "gpr_batch_read": (C.c_int, [_P, C.c_uint32, _P]),
    "gpr_batch_get_stats": (C.c_int, [_P, C.POINTER(BatchStats)]),
    "gpr_batch_free": (C.c_int, [_P]),
    "gpr_server_create": (C.c_int, [C.POINTER(ServerOpts), C.POINTER(_P)]),
    "gpr_server_add_scene": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "gpr_server_submit": (C.c_int, [_P, C.POINTER(Job), C.POINTER(C.c_uint64)]),
    "gpr_server_wait": (C.c_int, [_P, C.c_uint64, _P, C.POINTER(C.c_double)]),
    "gpr_server_cancel": (C.c_int, [_P, C.c_uint64]),
    "gpr_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "gpr_server_tune": (C.c_int, [_P, C.POINTER(Job), C.c_uint32, C.POINTER(TraceCfg), C.c_uint32,
                                  C.POINTER(C.c_uint32)]),
    "gpr_server_set_cfg": (C.c_int, [_P, C.POINTER(TraceCfg)]),
    "gpr_server_free": (C.c_int, [_P]),
    "gpr_gen_soup": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_float, _P]),
    "gpr_gen_rays": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, _P]),
    "gpr_gen_shadow_rays": (C.c_int, [C.c_uint64, C.c_uint64, _P, C.c_uint64, _P]),
    "gpr_scene_cornell": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    "gpr_scene_sphere_grid": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    "gpr_scene_mix": (C.c_int, [C.c_uint64, C.c_uint32, C.c_float, C.POINTER(_P)]),
}
for _name, (_res, _args) in _SIG.items():
    if hasattr(lib, _name):
        _f = getattr(lib, _name)
        _f.restype = _res
        _f.argtypes = _args


def check(rc):
    if rc != OK:
        raise GprError(rc, lib.gpr_last_error().decode(errors="replace"))
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def material(type=MAT_DIFFUSE, color=(0.5, 0.5, 0.5), color1=(1, 1, 1), emission=(0, 0, 0),
             ior=1.5, alpha=0.1, alpha_y=None, eta_i=(1, 1, 1), eta_t=(0.200438, 0.924033, 1.10221),
             k=(3.91295, 2.45285, 2.14219), tex=None):
    """tex: {slot: texture handle} with slots TEX_COLOR ... TEX_K."""
    m = Material()
    m.type = type
    m.color[:] = color
    m.color1[:] = color1
    m.emission[:] = emission
    m.ior = ior
    m.alpha_x = alpha
    m.alpha_y = alpha if alpha_y is None else alpha_y
    m.eta_i[:] = eta_i
    m.eta_t[:] = eta_t
    m.k[:] = k
    for slot, h in (tex or {}).items():
        m.tex[slot] = h
    return m


def read_image(path):
    """ReadImage: (height, width, 3) linear float32, row 0 = bottom row of the file."""
    p = C.POINTER(C.c_float)()
    w, h = C.c_uint32(), C.c_uint32()
    check(lib.gpr_image_read(os.fsencode(path), C.byref(p), C.byref(w), C.byref(h)))
    try:
        n = w.value * h.value * 3
        out = np.ctypeslib.as_array(p, shape=(n,)).copy().reshape(h.value, w.value, 3)
    finally:
        lib.gpr_image_free(p)
    return out


# ------------------------------------------------------------ generators --
def gen_soup(seed, stream, n, sigma=0.005):
    pos = np.empty((n, 9), np.float32)
    check(lib.gpr_gen_soup(seed, stream, n, sigma, _ptr(pos)))
    return pos


def gen_rays(seed, stream, n):
    rays = np.empty(n, RAY_DTYPE)
    check(lib.gpr_gen_rays(seed, stream, n, _ptr(rays)))
    return rays


def gen_shadow_rays(seed, stream, rays):
    out = np.empty(len(rays), RAY_DTYPE)
    check(lib.gpr_gen_shadow_rays(seed, stream, _ptr(rays), len(rays), _ptr(out)))
    return out


# ----------------------------------------------------------------- scene --
class Scene:
    """Owns a gpr_scene handle."""

    def __init__(self, handle=None):
        if handle is None:
            h = _P()
            check(lib.gpr_scene_create(C.byref(h)))
            handle = h
        self.h = handle

    @classmethod
    def soup(cls, pos, mat=None):
        """One mesh of unindexed triangles: pos is (ntri, 9) float32."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 9)
        s = cls()
        s.add_mesh(pos.reshape(-1, 3), corners=None, mat=mat)
        return s

    @classmethod
    def cornell(cls, w=256, h=256):
        hh = _P()
        check(lib.gpr_scene_cornell(w, h, C.byref(hh)))
        return cls(hh)

    @classmethod
    def sphere_grid(cls, w=1920, h=1080, subdiv=32):
        hh = _P()
        check(lib.gpr_scene_sphere_grid(w, h, subdiv, C.byref(hh)))
        return cls(hh)

    @classmethod
    def mix(cls, seed, ntri, soup_fraction=0.0):
        hh = _P()
        check(lib.gpr_scene_mix(seed, ntri, soup_fraction, C.byref(hh)))
        return cls(hh)

    @classmethod
    def load(cls, path):
        """Load and commit a .xml / .obj / .gprs scene."""
        hh = _P()
        check(lib.gpr_scene_load(os.fsencode(path), C.byref(hh)))
        return cls(hh)

    @classmethod
    def parse(cls, path):
        """Load a scene file on the host only (flattened, not committed)."""
        hh = _P()
        check(lib.gpr_scene_parse(os.fsencode(path), C.byref(hh)))
        return cls(hh)

    def save(self, path):
        check(lib.gpr_scene_save(self.h, os.fsencode(path)))

    def flatten(self):
        check(lib.gpr_scene_flatten(self.h))
        return self

    def add_mesh(self, p, corners=None, n=None, uv=None, mat=None):
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        if corners is None:
            ntri = len(p) // 3
            corners = np.full((ntri, 3, 3), -1, np.int32)
            corners[:, :, 0] = np.arange(3 * ntri, dtype=np.int32).reshape(ntri, 3)
        corners = np.ascontiguousarray(corners, np.int32).reshape(-1, 3, 3)
        n = None if n is None else np.ascontiguousarray(n, np.float32).reshape(-1, 3)
        uv = None if uv is None else np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        m = mat if mat is not None else material()
        check(lib.gpr_scene_add_mesh(self.h, _ptr(p), len(p), _ptr(n), 0 if n is None else len(n),
                                     _ptr(uv), 0 if uv is None else len(uv), _ptr(corners),
                                     len(corners), C.byref(m)))

    def set_camera(self, c2w, fov, w, h):
        m = np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(16))
        check(lib.gpr_scene_set_camera(self.h, _ptr(m), fov, w, h))

    def set_camera_free(self, eye, direction, up, fov, w, h):
        e, d, u = (np.ascontiguousarray(np.asarray(v, np.float32)) for v in (eye, direction, up))
        check(lib.gpr_scene_set_camera_free(self.h, _ptr(e), _ptr(d), _ptr(u), fov, w, h))

    def set_camera_lookat(self, target, origin, up, fov, w, h):
        t = np.asarray(target, np.float32)
        o = np.asarray(origin, np.float32)
        u = np.asarray(up, np.float32)
        check(lib.gpr_scene_set_camera_lookat(self.h, _ptr(t), _ptr(o), _ptr(u), fov, w, h))

    def set_env(self, rgb):
        a = np.asarray(rgb, np.float32)
        check(lib.gpr_scene_set_env(self.h, _ptr(a)))

    def add_texture(self, rgb, uv_offset=None, uv_scale=None):
        """rgb: (height, width, 3) linear texels, row 0 at v = 0.  Returns the handle."""
        a = np.ascontiguousarray(rgb, np.float32)
        if a.ndim != 3 or a.shape[2] != 3:
            raise ValueError("texture must be (height, width, 3)")
        o = None if uv_offset is None else np.ascontiguousarray(uv_offset, np.float32)
        sc = None if uv_scale is None else np.ascontiguousarray(uv_scale, np.float32)
        out = C.c_uint32()
        check(lib.gpr_scene_add_texture(self.h, _ptr(a), a.shape[1], a.shape[0], _ptr(o), _ptr(sc), C.byref(out)))
        return out.value

    def add_checkerboard(self, c0=(0.4, 0.4, 0.4), c1=(0.2, 0.2, 0.2), uv_offset=None, uv_scale=None):
        a, b = (np.ascontiguousarray(c, np.float32) for c in (c0, c1))
        o = None if uv_offset is None else np.ascontiguousarray(uv_offset, np.float32)
        sc = None if uv_scale is None else np.ascontiguousarray(uv_scale, np.float32)
        out = C.c_uint32()
        check(lib.gpr_scene_add_checkerboard(self.h, _ptr(a), _ptr(b), _ptr(o), _ptr(sc), C.byref(out)))
        return out.value

    def set_env_map(self, tex, o2w=None):
        m = None if o2w is None else np.ascontiguousarray(np.asarray(o2w, np.float32).reshape(16))
        check(lib.gpr_scene_set_env_map(self.h, tex, _ptr(m)))

    def textures(self):
        """(records TEXTURE_DTYPE, texels (ntexel, 4) float32)"""
        i = self.info()
        recs = np.zeros(max(i.ntex, 1), TEXTURE_DTYPE)
        texels = np.zeros((max(i.ntexel, 1), 4), np.float32)
        check(lib.gpr_scene_export_textures(self.h, _ptr(recs), _ptr(texels)))
        return recs[:i.ntex], texels[:i.ntexel]

    def commit(self):
        check(lib.gpr_scene_commit(self.h))
        return self

    def info(self):
        i = SceneInfo()
        check(lib.gpr_scene_get_info(self.h, C.byref(i)))
        return i

    def flat(self):
        i = self.info()
        n = i.ntri
        pos = np.empty((n, 9), np.float32)
        nrm = np.empty((n, 9), np.float32)
        uv = np.empty((n, 6), np.float32)
        flags = np.empty(n, np.uint8)
        mat = np.empty(n, np.uint32)
        mats = np.empty(max(i.nmat, 1), MATERIAL_DTYPE)
        check(lib.gpr_scene_export_flat(self.h, _ptr(pos), _ptr(nrm), _ptr(uv), _ptr(flags),
                                        _ptr(mat), _ptr(mats)))
        recs, texels = self.textures()
        return dict(pos=pos, nrm=nrm, uv=uv, flags=flags, mat=mat, mats=mats[:i.nmat], info=i,
                    tex=recs, texels=texels)

    def build(self):
        check(lib.gpr_bvh_build(self.h))

    def keys(self):
        n = self.info().ntri
        k = np.empty(n, np.uint32)
        sk = np.empty(n, np.uint32)
        si = np.empty(n, np.uint32)
        check(lib.gpr_bvh_export_keys(self.h, _ptr(k), _ptr(sk), _ptr(si)))
        return k, sk, si

    def ref_layout(self):
        n = self.info().ntri
        nodes = np.empty(max(2 * n - 1, 0), NODE_DTYPE)
        if n:
            check(lib.gpr_bvh_export_ref_layout(self.h, _ptr(nodes), len(nodes)))
        return nodes

    def timing(self):
        t = np.zeros(5, np.float32)
        check(lib.gpr_bvh_last_timing(self.h, _ptr(t)))
        return t

    def trace(self, rays, any_hit=False, ref_order=False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.empty(len(rays), HIT_DTYPE)
        flags = (TRACE_ANY if any_hit else 0) | (TRACE_REF_ORDER if ref_order else 0)
        check(lib.gpr_trace(self.h, _ptr(rays), _ptr(hits), len(rays), flags))
        return hits

    def trace_device(self, rays_ptr, hits_ptr, n, any_hit=False, ref_order=False):
        flags = TRACE_DEVICE_PTRS | (TRACE_ANY if any_hit else 0) | (TRACE_REF_ORDER if ref_order else 0)
        check(lib.gpr_trace(self.h, C.c_void_p(rays_ptr), C.c_void_p(hits_ptr), n, flags))

    def trace_tune(self, rays_ptr, n, any_hit=False, cap=64):
        out = (TraceCfg * cap)()
        k = C.c_uint32()
        fl = TRACE_DEVICE_PTRS | (TRACE_ANY if any_hit else 0)
        check(lib.gpr_trace_tune(self.h, C.c_void_p(rays_ptr), n, fl, out, cap, C.byref(k)))
        return [out[i] for i in range(k.value)]

    def trace_set_cfg(self, block, stack, chunk, refill=0):
        c = TraceCfg(block, stack, chunk, 0, refill, 0.0)
        check(lib.gpr_trace_set_cfg(self.h, C.byref(c)))

    def ref_stats(self):
        nv, nt = C.c_uint64(), C.c_uint64()
        check(lib.gpr_trace_ref_stats(self.h, C.byref(nv), C.byref(nt)))
        return nv.value, nt.value

    def bvh4_stats(self, rays_ptr, n, any_hit=False):
        """(BVH4 node fetches, triangle tests) summed over n device rays."""
        nf, nt = C.c_uint64(), C.c_uint64()
        fl = TRACE_DEVICE_PTRS | (TRACE_ANY if any_hit else 0)
        check(lib.gpr_trace_bvh4_stats(self.h, C.c_void_p(rays_ptr), n, fl, C.byref(nf), C.byref(nt)))
        return nf.value, nt.value

    def free(self):
        if self.h:
            check(lib.gpr_scene_free(self.h))  # refused while renderers still use the scene
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Renderer:
    def __init__(self, scene, max_depth=5, rng_mode=RNG_PCG, rr_start=3, seed=1,
                 traversal=TRAVERSAL_FAST):
        self.scene = scene
        o = RenderOpts(max_depth, rng_mode, rr_start, seed, traversal)
        h = _P()
        check(lib.gpr_renderer_create(scene.h, C.byref(o), C.byref(h)))
        self.h = h
        i = scene.info()
        self.w, self.h_px = i.width, i.height

    def render(self, spp):
        check(lib.gpr_render(self.h, spp))

    def set_opts(self, max_depth=5, rng_mode=RNG_PCG, rr_start=3, seed=1, traversal=TRAVERSAL_FAST):
        """New options for the next accumulation (resets it)."""
        o = RenderOpts(max_depth, rng_mode, rr_start, seed, traversal)
        check(lib.gpr_renderer_set_opts(self.h, C.byref(o)))

    def zoom(self, d):
        check(lib.gpr_renderer_zoom(self.h, d))

    def translate(self, x, y):
        check(lib.gpr_renderer_translate(self.h, x, y))

    def rotate(self, yaw, pitch):
        check(lib.gpr_renderer_rotate(self.h, yaw, pitch))

    def resize(self, w, h):
        check(lib.gpr_renderer_resize(self.h, w, h))
        self.w, self.h_px = w, h

    def resolution(self):
        w, h = C.c_uint32(), C.c_uint32()
        check(lib.gpr_renderer_get_resolution(self.h, C.byref(w), C.byref(h)))
        return w.value, h.value

    def read(self):
        n = self.w * self.h_px
        rgb = np.empty((n, 3), np.float32)
        cnt = np.empty(n, np.uint32)
        check(lib.gpr_renderer_read(self.h, _ptr(rgb), _ptr(cnt)))
        return rgb, cnt

    def write_png(self, path):
        check(lib.gpr_renderer_write_png(self.h, os.fsencode(path)))

    def restore(self, rgb_sum, spp):
        """continue from a checkpoint (rgb_sum as read() returns it, spp samples)"""
        a = np.ascontiguousarray(rgb_sum, np.float32).reshape(-1)
        if a.size != 3 * self.w * self.h_px:
            raise ValueError("checkpoint size does not match the image")
        check(lib.gpr_renderer_restore(self.h, _ptr(a), spp))

    def write_image(self, path):
        """.png or .hdr by extension"""
        check(lib.gpr_renderer_write_image(self.h, os.fsencode(path)))

    def resolve_rgba8(self, device_ptr=None):
        """RGBA8 of the mean (frame-buffer order).  With device_ptr the bytes go to
        that device buffer (no synchronisation) and None is returned."""
        if device_ptr is not None:
            check(lib.gpr_renderer_resolve_rgba8(self.h, C.c_void_p(device_ptr), RESOLVE_DEVICE_PTRS))
            return None
        out = np.empty(self.w * self.h_px, np.uint32)
        check(lib.gpr_renderer_resolve_rgba8(self.h, _ptr(out), 0))
        return out

    def reset(self):
        check(lib.gpr_renderer_reset(self.h))

    def queue_sizes(self, cap=64):
        sh = np.zeros(cap, np.uint32)
        sc = np.zeros(cap, np.uint32)
        check(lib.gpr_renderer_queue_sizes(self.h, _ptr(sh), _ptr(sc), cap))
        return sh, sc

    def free(self):
        if self.h:
            lib.gpr_renderer_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def job(scene, width, height, spp, c2w, fov=40.0, max_depth=5, rr_start=3, seed=1):
    """A render request: scene id, resolution, samples, camera (row-major 4x4 cameraToWorld)."""
    j = Job()
    j.scene, j.width, j.height, j.spp = scene, width, height, spp
    j.max_depth, j.rr_start, j.seed, j.fov_deg = max_depth, rr_start, seed, fov
    j.c2w[:] = [float(x) for x in np.asarray(c2w, np.float32).ravel()]
    return j


def lookat(target, origin, up=(0, 1, 0)):
    """cameraToWorld of LookAt (row-major), as gpr_scene_set_camera_lookat builds it."""
    t, o, u = (np.asarray(v, np.float64) for v in (target, origin, up))
    fwd = (t - o) / np.linalg.norm(t - o)
    left = np.cross(u, fwd)
    left /= np.linalg.norm(left)
    rup = np.cross(fwd, left)
    m = np.eye(4, dtype=np.float64)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = left, rup, fwd, o
    return m.astype(np.float32)


def arena_bytes(scene):
    b = C.c_uint64()
    check(lib.gpr_scene_arena_bytes(scene.h, C.byref(b)))
    return b.value


class Batch:
    """Many jobs over many scenes in one multi-job wavefront (gpr_batch_*)."""

    def __init__(self, arena_bytes, capacity=1 << 22, max_jobs=256):
        o = BatchOpts(arena_bytes, capacity, max_jobs)
        h = _P()
        check(lib.gpr_batch_create(C.byref(o), C.byref(h)))
        self.h = h
        self.jobs = []

    def add_scene(self, scene):
        i = C.c_uint32()
        check(lib.gpr_batch_add_scene(self.h, scene.h, C.byref(i)))
        return i.value

    def add_job(self, j):
        i = C.c_uint32()
        check(lib.gpr_batch_add_job(self.h, C.byref(j), C.byref(i)))
        self.jobs.append((j.width, j.height))
        return i.value

    def tune(self, cap=128):
        out = (TraceCfg * cap)()
        n = C.c_uint32()
        check(lib.gpr_batch_tune(self.h, out, cap, C.byref(n)))
        return [out[k] for k in range(n.value)]

    def set_cfg(self, block, stack, chunk, layout=0, refill=0):
        c = TraceCfg(block, stack, chunk, layout, refill, 0.0)
        check(lib.gpr_batch_set_cfg(self.h, C.byref(c)))

    def run(self):
        check(lib.gpr_batch_run(self.h))

    def read(self, job):
        w, h = self.jobs[job]
        rgb = np.empty((w * h, 3), np.float32)
        check(lib.gpr_batch_read(self.h, job, _ptr(rgb)))
        return rgb

    def stats(self):
        st = BatchStats()
        check(lib.gpr_batch_get_stats(self.h, C.byref(st)))
        return st

    def free(self):
        if self.h:
            lib.gpr_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Request server with HBM-budget admission and merged per-bounce queues (gpr_server_*)."""

    def __init__(self, hbm_budget, arena_bytes, capacity=1 << 22, max_jobs=256):
        o = ServerOpts(hbm_budget, arena_bytes, capacity, max_jobs)
        h = _P()
        check(lib.gpr_server_create(C.byref(o), C.byref(h)))
        self.h = h
        self.sizes = {}

    def add_scene(self, scene):
        i = C.c_uint32()
        check(lib.gpr_server_add_scene(self.h, scene.h, C.byref(i)))
        return i.value

    def tune(self, jobs, cap=128):
        arr = (Job * len(jobs))(*jobs)
        out = (TraceCfg * cap)()
        n = C.c_uint32()
        check(lib.gpr_server_tune(self.h, arr, len(jobs), out, cap, C.byref(n)))
        return [out[k] for k in range(n.value)]

    def set_cfg(self, block, stack, chunk, layout=0, refill=0):
        c = TraceCfg(block, stack, chunk, layout, refill, 0.0)
        check(lib.gpr_server_set_cfg(self.h, C.byref(c)))

    def submit(self, j):
        t = C.c_uint64()
        check(lib.gpr_server_submit(self.h, C.byref(j), C.byref(t)))
        self.sizes[t.value] = j.width * j.height
        return t.value

    def cancel(self, ticket):
        check(lib.gpr_server_cancel(self.h, ticket))

    def wait(self, ticket, want_image=True):
        n = self.sizes.pop(ticket)
        rgb = np.empty((n, 3), np.float32) if want_image else None
        lat = C.c_double()
        check(lib.gpr_server_wait(self.h, ticket, _ptr(rgb), C.byref(lat)))
        return rgb, lat.value

    def stats(self):
        st = ServerStats()
        check(lib.gpr_server_get_stats(self.h, C.byref(st)))
        return st

    def free(self):
        if self.h:
            lib.gpr_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sort_pairs_device(keys_ptr, skeys_ptr, sidx_ptr, n):
    check(lib.gpr_sort_pairs_device(C.c_void_p(keys_ptr), C.c_void_p(skeys_ptr), C.c_void_p(sidx_ptr), n))


def sort_pairs_device_noalloc(keys_ptr, skeys_ptr, sidx_ptr, n):
    """gpr_sort_pairs_device with the library's fixed scratch array (no
    allocation; n at most 1,048,576)."""
    check(lib.gpr_sort_pairs_device_noalloc(C.c_void_p(keys_ptr), C.c_void_p(skeys_ptr), C.c_void_p(sidx_ptr), n))


def trace_overflow_count():
    """Rays that ran out of traversal stack in the last fast trace on this
    thread's stream (device-pointer traces do not raise it themselves)."""
    v = C.c_uint64()
    check(lib.gpr_trace_overflow_count(C.byref(v)))
    return v.value


def release_stream(ptr):
    """Free the trace / sort work buffers kept for this thread's device and stream."""
    check(lib.gpr_release_stream(None if not ptr else C.c_void_p(ptr)))


def set_stream(ptr):
    check(lib.gpr_set_stream(None if not ptr else C.c_void_p(ptr)))


def synchronize():
    check(lib.gpr_synchronize())
