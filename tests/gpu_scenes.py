"""Scene inputs shared by the GPU parity tests (seeded, deterministic)."""
import numpy as np

import gpr


def soup(n, seed=2, sigma=0.005):
    return gpr.gen_soup(seed, 0, n, sigma)


def duplicate_heavy(n, seed=9):
    """Many triangles share centroid cells: exercises the equal-key midpoint rule."""
    rng = np.random.default_rng(seed)
    base = rng.random((max(n // 8, 1), 3)).astype(np.float32)
    c = base[rng.integers(0, len(base), n)]
    jitter = rng.normal(0, 1e-4, (n, 3, 3)).astype(np.float32)
    return (c[:, None, :] + jitter).reshape(n, 9).astype(np.float32)


def identical(n):
    t = np.array([0.1, 0.2, 0.3, 0.5, 0.2, 0.3, 0.1, 0.6, 0.3], np.float32)
    return np.tile(t, (n, 1))


def planar_grid(m, z=0.5, seed=3):
    """2 m^2 triangles tiling [0,1]^2 in the plane z = const, interior vertices
    jittered by under half a cell (no overlaps, so no exact-t ties): the
    centroid bounds have zero z extent (0/0 in the key quantizer) and every
    BVH box is flat in z."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0, 1, m + 1, dtype=np.float32)
    x, y = np.meshgrid(g, g, indexing="ij")
    j = rng.uniform(-0.3 / m, 0.3 / m, (2, m + 1, m + 1)).astype(np.float32)
    j[:, [0, -1], :] = 0
    j[:, :, [0, -1]] = 0
    v = np.stack([x + j[0], y + j[1], np.full_like(x, z)], -1)
    a, b, c, d = v[:-1, :-1], v[1:, :-1], v[1:, 1:], v[:-1, 1:]
    t0 = np.concatenate([a, b, c], -1).reshape(-1, 9)
    t1 = np.concatenate([a, c, d], -1).reshape(-1, 9)
    return np.concatenate([t0, t1]).astype(np.float32)


def collinear(n, seed=4):
    """Tilted triangles whose box centres all lie on the line y = z = 0.5
    (two flat axes of the centroid bounds); all coordinates but x are exact
    binary fractions, so the centres are exact."""
    rng = np.random.default_rng(seed)
    x0 = rng.random(n).astype(np.float32)
    w = np.float32(0.01)
    t = np.zeros((n, 3, 3), np.float32)
    t[:, 0] = np.stack([x0, np.full(n, 0.375), np.full(n, 0.5)], -1)
    t[:, 1] = np.stack([x0 + w, np.full(n, 0.625), np.full(n, 0.375)], -1)
    t[:, 2] = np.stack([x0 + w / 2, np.full(n, 0.5), np.full(n, 0.625)], -1)
    return t.reshape(n, 9)


def scene_from_soup(pos):
    return gpr.Scene.soup(pos).commit()


def _quad(p, uv):
    """Two triangles (0,1,2) (0,2,3) with per-corner uvs."""
    corners = np.array([[[0, 0, -1], [1, 1, -1], [2, 2, -1]], [[0, 0, -1], [2, 2, -1], [3, 3, -1]]], np.int32)
    return np.asarray(p, np.float32), np.asarray(uv, np.float32), corners


def textured(w=40, h=32, env=True, seed=5):
    """Every textured parameter kind in one scene: a checkerboard diffuse floor
    (uv scale 6 x 4, negative offset), a rough plastic back wall with a bitmap
    diffuse colour and a roughness texture, a conductor and a glass without
    mesh uvs (barycentric lookups) with textured k / eta and ior, an area light
    and, optionally, a rotated latitude-longitude environment map."""
    rng = np.random.default_rng(seed)
    s = gpr.Scene()
    chk = s.add_checkerboard((0.8, 0.2, 0.1), (0.1, 0.3, 0.7), uv_offset=(-0.35, 0.1), uv_scale=(6, 4))
    bmp = s.add_texture(rng.random((8, 16, 3)) * 0.8 + 0.1)
    rough = s.add_texture(rng.random((4, 4, 3)) * 0.4 + 0.05, uv_scale=(2, 3))
    kmap = s.add_checkerboard((3.9, 2.4, 2.1), (0.5, 0.6, 0.9), uv_scale=(3, 3))
    iors = s.add_texture(np.array([[[1.3, 0, 0], [1.8, 0, 0]]], np.float32))
    p, uv, c = _quad([(-3, -1, -3), (3, -1, -3), (3, -1, 3), (-3, -1, 3)], [(0, 0), (1, 0), (1, 1), (0, 1)])
    s.add_mesh(p, corners=c, uv=uv, mat=gpr.material(gpr.MAT_DIFFUSE, tex={gpr.TEX_COLOR: chk}))
    p, uv, c = _quad([(-3, -1, -2), (3, -1, -2), (3, 3, -2), (-3, 3, -2)], [(0, 0), (1, 0), (1, 1), (0, 1)])
    s.add_mesh(p, corners=c, uv=uv, mat=gpr.material(gpr.MAT_ROUGH_PLASTIC, eta_i=(1.5, 1.5, 1.5), eta_t=(1, 1, 1),
                                                     tex={gpr.TEX_COLOR: bmp, gpr.TEX_ALPHA_X: rough,
                                                          gpr.TEX_ALPHA_Y: rough}))
    cond = gpr.gen_soup(seed, 0, 60, 0.25).reshape(-1, 3) * 1.2 + np.float32([-1.0, 0.0, 0.0])
    s.add_mesh(cond, mat=gpr.material(gpr.MAT_ROUGH_CONDUCTOR, alpha=0.2, color=(0.9, 0.9, 0.9),
                                      tex={gpr.TEX_K: kmap, gpr.TEX_ETA_T: kmap}))
    glass = gpr.gen_soup(seed + 1, 0, 40, 0.25).reshape(-1, 3) * 1.2 + np.float32([0.8, 0.0, 0.5])
    s.add_mesh(glass, mat=gpr.material(gpr.MAT_SPECULAR, color=(1, 1, 1), tex={gpr.TEX_IOR: iors}))
    p, uv, c = _quad([(-0.6, 2.9, -0.6), (0.6, 2.9, -0.6), (0.6, 2.9, 0.6), (-0.6, 2.9, 0.6)],
                     [(0, 0), (1, 0), (1, 1), (0, 1)])
    s.add_mesh(p[::-1].copy(), corners=c, uv=uv, mat=gpr.material(gpr.MAT_DIFFUSE, emission=(12, 11, 9)))
    if env:
        e = s.add_texture(rng.random((16, 32, 3)) * 1.5)
        ang = np.deg2rad(30.0)
        o2w = np.eye(4, dtype=np.float32)
        o2w[0, 0], o2w[0, 2], o2w[2, 0], o2w[2, 2] = np.cos(ang), np.sin(ang), -np.sin(ang), np.cos(ang)
        s.set_env_map(e, o2w)
    s.set_camera_lookat((0, 0.5, 0), (0.5, 1.5, 6.0), (0, 1, 0), 50, w, h)
    return s
