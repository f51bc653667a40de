"""Textures, the latitude-longitude environment light and image input (CPU).

* Image input (ReadImage, src/renderer/imageio.cpp:45-57): files written here
  by an independent encoder of the PNG / Radiance / PNM specifications; every
  decoded value is checked against what stbi_loadf defines -- (c/255)^2.2 for
  8-bit samples, m * 2^(e-136) for RGBE -- and the rows come back bottom first.
* Oracle texture lookup (TextureView, texture.h:37-52): known texels at
  hand-picked uvs, including the negative-uv wrap of int % uint32.
* Oracle environment lookup (environment.h:16-28): the axis directions land
  on the texels that theta = acos(z), phi = atan2(y, x) fix in closed form; a
  rotation about z shifts the map by a quarter turn; random directions agree
  with a float64 evaluation except at texel borders.
* Loader: <texture> and envmap emitters, .gprs round trip.
"""
import ctypes as C
import os
import struct
import zlib

import numpy as np
import pytest

import gpr
import oracle_py as O


def lin(c):
    """stbi's LDR -> linear: powf(c / 255, 2.2) as a float."""
    x = np.float32(np.float32(c) / np.float32(255.0))
    return np.float32(np.float64(x) ** np.float64(np.float32(2.2)))


# ------------------------------------------------------------ encoders ---
def png_chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xffffffff)


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def png_filter(rows, bpp, ftypes):
    """PNG specification section 9: filter each scanline with the given type."""
    out = b""
    prev = bytes(len(rows[0]))
    for r, f in zip(rows, ftypes):
        enc = bytearray()
        for x in range(len(r)):
            a = r[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            pred = [0, a, b, (a + b) // 2, paeth(a, b, c)][f]
            enc.append((r[x] - pred) & 255)
        out += bytes([f]) + bytes(enc)
        prev = r
    return out


def write_png(path, w, h, depth, ctype, rows, ftypes=None, palette=None):
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bpp = max(1, chans * depth // 8)
    ftypes = ftypes or [0] * h
    raw = png_filter(rows, bpp, ftypes)
    data = b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if palette is not None:
        data += png_chunk(b"PLTE", bytes(palette))
    comp = zlib.compress(raw)
    data += png_chunk(b"IDAT", comp[:7]) + png_chunk(b"IDAT", comp[7:])  # split over two chunks
    data += png_chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def rgbe_rle_line(px):
    """New-style Radiance scanline: 2 2 hi lo, then each component run-length coded."""
    w = len(px)
    out = bytes([2, 2, w >> 8, w & 255])
    for c in range(4):
        comp = [p[c] for p in px]
        i = 0
        while i < w:
            run = 1
            while i + run < w and run < 127 and comp[i + run] == comp[i]:
                run += 1
            if run >= 3:
                out += bytes([128 + run, comp[i]])
                i += run
            else:
                j = i
                while j < w and j - i < 128 and not (j + 2 < w and comp[j] == comp[j + 1] == comp[j + 2]):
                    j += 1
                out += bytes([j - i]) + bytes(comp[i:j])
                i = j
    return out


def write_hdr(path, px, rle=True):
    h, w = len(px), len(px[0])
    data = b"#?RADIANCE\n# test\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode()
    for row in px:
        data += rgbe_rle_line(row) if rle else b"".join(bytes(p) for p in row)
    with open(path, "wb") as f:
        f.write(data)


# ----------------------------------------------------------- image input ---
@pytest.mark.parametrize("ftype", [0, 1, 2, 3, 4])
def test_png_rgb8_filters_and_flip(tmp_path, ftype):
    rng = np.random.default_rng(ftype)
    w, h = 7, 5
    img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    p = str(tmp_path / "a.png")
    write_png(p, w, h, 8, 2, [img[y].tobytes() for y in range(h)], [ftype] * h)
    got = gpr.read_image(p)
    assert got.shape == (h, w, 3)
    want = np.vectorize(lin)(img[::-1]).astype(np.float32)  # bottom row first
    np.testing.assert_array_equal(got, want)


def test_png_mixed_filters_rgba16_grey_palette(tmp_path):
    rng = np.random.default_rng(9)
    w, h = 6, 5
    # RGBA 16-bit: alpha dropped, each sample keeps its high byte
    img16 = rng.integers(0, 65536, (h, w, 4), dtype=np.uint16)
    p = str(tmp_path / "b.png")
    write_png(p, w, h, 16, 6, [img16[y].astype(">u2").tobytes() for y in range(h)], [0, 1, 2, 3, 4])
    want = np.vectorize(lin)((img16[::-1, :, :3] >> 8).astype(np.uint8)).astype(np.float32)
    np.testing.assert_array_equal(gpr.read_image(p), want)
    # grey + alpha 8-bit: grey replicated
    ga = rng.integers(0, 256, (h, w, 2), dtype=np.uint8)
    p = str(tmp_path / "c.png")
    write_png(p, w, h, 8, 4, [ga[y].tobytes() for y in range(h)], [4, 3, 2, 1, 0])
    g = np.vectorize(lin)(ga[::-1, :, 0]).astype(np.float32)
    np.testing.assert_array_equal(gpr.read_image(p), np.repeat(g[:, :, None], 3, axis=2))
    # 2-bit grey: 0..3 scaled by 0x55
    vals = rng.integers(0, 4, (h, 8), dtype=np.uint8)
    rows = [bytes([(r[0] << 6) | (r[1] << 4) | (r[2] << 2) | r[3], (r[4] << 6) | (r[5] << 4) | (r[6] << 2) | r[7]])
            for r in vals]
    p = str(tmp_path / "d.png")
    write_png(p, 8, h, 2, 0, rows)
    g = np.vectorize(lin)(vals[::-1] * 0x55).astype(np.float32)
    np.testing.assert_array_equal(gpr.read_image(p)[:, :, 1], g)
    # 4-bit palette
    pal = rng.integers(0, 256, (16, 3), dtype=np.uint8)
    idx = rng.integers(0, 16, (h, 4), dtype=np.uint8)
    rows = [bytes([(r[0] << 4) | r[1], (r[2] << 4) | r[3]]) for r in idx]
    p = str(tmp_path / "e.png")
    write_png(p, 4, h, 4, 3, rows, palette=pal.ravel())
    np.testing.assert_array_equal(gpr.read_image(p), np.vectorize(lin)(pal[idx[::-1]]).astype(np.float32))


def test_png_known_values(tmp_path):
    p = str(tmp_path / "k.png")
    write_png(p, 3, 1, 8, 2, [bytes([0, 0, 0, 255, 255, 255, 128, 64, 1])])
    got = gpr.read_image(p)
    assert got[0, 0].tolist() == [0.0, 0.0, 0.0]
    assert got[0, 1].tolist() == [1.0, 1.0, 1.0]
    np.testing.assert_allclose(got[0, 2], [(128 / 255) ** 2.2, (64 / 255) ** 2.2, (1 / 255) ** 2.2], rtol=1e-6)


@pytest.mark.parametrize("rle", [True, False])
def test_radiance_hdr(tmp_path, rle):
    w, h = 20, 3
    rng = np.random.default_rng(4)
    px = [[(int(a), int(b), int(c), int(e)) for a, b, c, e in
           zip(rng.integers(0, 256, w), np.full(w, 64), rng.integers(0, 3, w), rng.choice([0, 120, 128, 136, 140], w))]
          for _ in range(h)]
    px[0][0] = (128, 64, 32, 129)  # 1.0, 0.5, 0.25
    p = str(tmp_path / "m.hdr")
    write_hdr(p, px, rle)
    got = gpr.read_image(p)
    assert got.shape == (h, w, 3)
    np.testing.assert_array_equal(got[h - 1, 0], np.float32([1.0, 0.5, 0.25]))
    for y in range(h):
        for x in range(w):
            r, g, b, e = px[y][x]
            want = [0.0, 0.0, 0.0] if e == 0 else [c * 2.0 ** (e - 136) for c in (r, g, b)]
            assert got[h - 1 - y, x].tolist() == want, (y, x)


def test_pnm(tmp_path):
    img = np.arange(2 * 3 * 3, dtype=np.uint8).reshape(2, 3, 3) * 13
    p = str(tmp_path / "a.ppm")
    with open(p, "wb") as f:
        f.write(b"P6\n# comment\n3 2\n255\n" + img.tobytes())
    np.testing.assert_array_equal(gpr.read_image(p), np.vectorize(lin)(img[::-1]).astype(np.float32))
    p = str(tmp_path / "a.pgm")
    with open(p, "wb") as f:
        f.write(b"P5 3 2 255\n" + img[:, :, 0].tobytes())
    g = np.vectorize(lin)(img[::-1, :, 0]).astype(np.float32)
    np.testing.assert_array_equal(gpr.read_image(p), np.repeat(g[:, :, None], 3, axis=2))


def test_image_errors(tmp_path):
    p = str(tmp_path / "x.png")
    with open(p, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", 2, 2, 8, 2, 0, 0, 1)) +
                png_chunk(b"IEND", b""))
    with pytest.raises(gpr.GprError, match="interlaced"):
        gpr.read_image(p)
    with open(p, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", 2, 2, 8, 2, 0, 0, 0)) +
                png_chunk(b"IDAT", zlib.compress(b"\x00" * 5)) + png_chunk(b"IEND", b""))
    with pytest.raises(gpr.GprError, match="corrupt"):
        gpr.read_image(p)
    with pytest.raises(gpr.GprError, match="cannot open"):
        gpr.read_image(str(tmp_path / "none.png"))


# ---------------------------------------------------- oracle lookups ---
def tex_scene(recs, texels, env_tex=0, w2o=None, center=(0, 0, 0)):
    s = O.Scene()
    keep = {"tex": np.ascontiguousarray(recs), "texels": np.ascontiguousarray(texels, np.float32)}
    s.tex, s.texels, s.ntex = O.ptr(keep["tex"]), O.ptr(keep["texels"]), len(recs)
    s.env_tex = env_tex
    s.env_w2o[:] = list(np.eye(3, dtype=np.float32).ravel() if w2o is None else np.asarray(w2o, np.float32).ravel())
    s.env_center[:] = list(center)
    s._keep = keep
    return s


def index_texture(w, h, scale=(1, 1), offset=(0, 0)):
    """texel (x, y) holds (x, y, 7)"""
    recs = np.zeros(1, gpr.TEXTURE_DTYPE)
    recs[0]["width"], recs[0]["height"] = w, h
    recs[0]["uv_scale"], recs[0]["uv_offset"] = scale, offset
    tx = np.zeros((h, w, 4), np.float32)
    tx[:, :, 0] = np.arange(w)[None, :]
    tx[:, :, 1] = np.arange(h)[:, None]
    tx[:, :, 2] = 7
    return recs, tx.reshape(-1, 4)


def tex_at(s, u, v, h=1):
    out = np.zeros(3, np.float32)
    O.lib.orc_tex_eval(C.byref(s), h, u, v, O.ptr(out))
    return tuple(int(x) for x in out[:2])


def test_oracle_texture_lookup_known_texels():
    s = tex_scene(*index_texture(4, 2))
    assert tex_at(s, 0.26, 0.51) == (1, 1)
    assert tex_at(s, 0.0, 0.0) == (0, 0)
    assert tex_at(s, 0.999, 0.49) == (3, 0)
    assert tex_at(s, 1.0, 1.0) == (0, 0)  # int(4) % 4, int(2) % 2
    assert tex_at(s, 1.3, 2.6) == (1, 1)  # 5 % 4, 5 % 2
    # negative: int(-0.4) = 0; int(-1.2) = -1 -> 0xffffffff % 4 = 3, % 2 = 1
    assert tex_at(s, -0.1, -0.2) == (0, 0)
    assert tex_at(s, -0.3, -0.6) == (3, 1)
    s3 = tex_scene(*index_texture(3, 5))
    assert tex_at(s3, -0.4, 0.5) == (0, 2)  # 0xffffffff % 3 = 0 (2^32 - 1 = 3 * 1431655765)
    assert tex_at(s3, -0.7, -0.1) == (2, 0)  # int(-2.1) = -2 -> 0xfffffffe % 3 = 2; int(-0.5) = 0
    # scale and offset: uv' = uv * scale + offset
    s2 = tex_scene(*index_texture(4, 4, scale=(2, 3), offset=(0.25, -0.5)))
    assert tex_at(s2, 0.1, 0.3) == (1, 1)  # (0.45 * 4 = 1.8, 0.4 * 4 = 1.6)
    assert tex_at(s2, 0.5, 0.1) == (1, 0)  # (1.25 * 4 = 5 -> 1, -0.2 * 4 = -0.8 -> 0)
    assert tex_at(s2, 0.5, 0.0) == (1, 2)  # v: -0.5 * 4 = -2 -> 0xfffffffe % 4 = 2


def test_oracle_checkerboard_via_scene():
    sc = gpr.Scene()
    h = sc.add_checkerboard((1, 0, 0), (0, 0, 1), uv_scale=(2, 2))
    recs, texels = sc.textures()
    assert h == 1 and len(recs) == 1 and recs[0]["width"] == 2 and recs[0]["height"] == 2
    s = tex_scene(recs, texels)
    out = np.zeros(3, np.float32)
    # texel (x, y) = c0 where x == y (texture.cpp:50-62), uv scaled by 2
    for (u, v), want in [((0.1, 0.1), (1, 0, 0)), ((0.3, 0.1), (0, 0, 1)), ((0.1, 0.3), (0, 0, 1)),
                         ((0.3, 0.3), (1, 0, 0)), ((0.6, 0.6), (1, 0, 0)), ((0.8, 0.6), (0, 0, 1))]:
        O.lib.orc_tex_eval(C.byref(s), 1, u, v, O.ptr(out))
        assert tuple(out) == want, (u, v)
    sc.free()


def env_texel(s, d, o=(0, 0, 0)):
    out = np.zeros(3, np.float32)
    O.lib.orc_env_le(C.byref(s), O.ptr(np.float32(o)), O.ptr(np.float32(d)), O.ptr(out))
    return int(out[0]), int(out[1])


def test_oracle_env_axis_directions():
    w, h = 16, 8
    s = tex_scene(*index_texture(w, h), env_tex=1)
    # theta = acos(z) -> v = theta / pi; phi = atan2(y, x) in [0, 2 pi) -> u = phi / (2 pi)
    assert env_texel(s, (1, 0, 0)) == (0, h // 2)
    assert env_texel(s, (0, 1, 0)) == (w // 4, h // 2)
    assert env_texel(s, (-1, 0, 0)) == (w // 2, h // 2)
    assert env_texel(s, (0, -1, 0)) == (3 * w // 4, h // 2)
    assert env_texel(s, (0, 0, 1)) == (0, 0)
    assert env_texel(s, (0, 0, -1)) == (0, h - 1)  # v = 1 clamps to the last row
    # the direction is taken from the world centre to o + 1e5 d: an origin below
    # the centre tips +x just under the equator (theta > pi/2: row h/2), one
    # above it just over (row h/2 - 1)
    assert env_texel(s, (1, 0, 0), o=(3, 2, -1)) == (0, h // 2)
    assert env_texel(s, (1, 0, 0), o=(3, 2, 1)) == (0, h // 2 - 1)
    # w2o = rotation by +90 degrees about z: world +x maps to object +y (u = 1/4)
    sr = tex_scene(*index_texture(w, h), env_tex=1, w2o=[[0, -1, 0], [1, 0, 0], [0, 0, 1]])
    assert env_texel(sr, (1, 0, 0)) == (w // 4, h // 2)
    assert env_texel(sr, (0, 1, 0)) == (w // 2, h // 2)


def test_oracle_env_random_directions_match_float64():
    w, h = 64, 32
    s = tex_scene(*index_texture(w, h), env_tex=1)
    rng = np.random.default_rng(1)
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    got = np.array([env_texel(s, x) for x in d.astype(np.float32)])
    theta = np.arccos(np.clip(d[:, 2], -1, 1))
    phi = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * np.pi)
    want = np.stack([np.minimum(w - 1, np.floor(phi / (2 * np.pi) * w)),
                     np.minimum(h - 1, np.floor(theta / np.pi * h))], 1).astype(int)
    diff = np.abs(got - want)
    diff[:, 0] = np.minimum(diff[:, 0], w - diff[:, 0])  # u wraps
    assert diff.max() <= 1
    assert (diff.sum(1) != 0).mean() < 0.01  # only float32 rounding at texel borders


# --------------------------------------------------------------- loader ---
def test_xml_textures_and_envmap(tmp_path):
    img = np.array([[[255, 0, 0], [0, 255, 0]], [[0, 0, 255], [255, 255, 255]]], np.uint8)
    write_png(str(tmp_path / "t.png"), 2, 2, 8, 2, [img[0].tobytes(), img[1].tobytes()])
    write_hdr(str(tmp_path / "env.hdr"), [[(128, 128, 128, 129)] * 8] * 4)
    xml = """<?xml version="1.0"?>
<scene version="0.5.0">
  <sensor type="perspective"><float name="fov" value="45"/>
    <film type="hdrfilm"><integer name="width" value="32"/><integer name="height" value="24"/></film></sensor>
  <bsdf type="diffuse" id="chk">
    <texture type="checkerboard" name="reflectance">
      <rgb name="color0" value="0.9 0.1 0.1"/><rgb name="color1" value="0.1 0.9 0.1"/>
      <float name="uscale" value="4"/><float name="vscale" value="2"/>
    </texture>
  </bsdf>
  <bsdf type="roughplastic" id="bmp">
    <texture type="bitmap" name="diffuseReflectance">
      <string name="filename" value="t.png"/><float name="uoffset" value="0.5"/>
    </texture>
    <texture type="checkerboard" name="alphaU"/>
  </bsdf>
  <shape type="rectangle"><ref id="chk"/></shape>
  <shape type="cube"><ref id="bmp"/></shape>
  <shape type="rectangle"><bsdf type="diffuse"/></shape>
  <emitter type="envmap">
    <string name="filename" value="env.hdr"/>
    <transform name="toWorld"><rotate y="1" angle="90"/></transform>
  </emitter>
</scene>"""
    p = tmp_path / "s.xml"
    p.write_text(xml)
    s = gpr.Scene.parse(str(p))
    f = s.flat()
    info = f["info"]
    assert info.ntex == 4 and info.env_tex == 4 and info.has_env == 1
    recs, texels = f["tex"], f["texels"]
    # checkerboard: 2x2, scales, colours c0 c1 / c1 c0
    assert (recs[0]["width"], recs[0]["height"]) == (2, 2)
    np.testing.assert_array_equal(recs[0]["uv_scale"], np.float32([4, 2]))
    np.testing.assert_array_equal(texels[0:4, :3], np.float32([[0.9, 0.1, 0.1], [0.1, 0.9, 0.1],
                                                                [0.1, 0.9, 0.1], [0.9, 0.1, 0.1]]))
    # bitmap: decoded, flipped, offset
    np.testing.assert_array_equal(recs[1]["uv_offset"], np.float32([0.5, 0]))
    got = texels[recs[1]["offset"]:recs[1]["offset"] + 4, :3].reshape(2, 2, 3)
    np.testing.assert_array_equal(got, np.vectorize(lin)(img[::-1]).astype(np.float32))
    # default checkerboard colours 0.4 / 0.2 (mitsubaloader.cpp:549-550)
    np.testing.assert_array_equal(texels[recs[2]["offset"], :3], np.float32([0.4] * 3))
    np.testing.assert_array_equal(texels[recs[2]["offset"] + 1, :3], np.float32([0.2] * 3))
    # env map texels 1.0, toWorld = rotate 90 degrees about y; worldToObject is its inverse
    np.testing.assert_array_equal(texels[recs[3]["offset"]:, :3], np.ones((32, 3), np.float32))
    r = np.array(info.env_w2o, np.float32).reshape(3, 3)
    np.testing.assert_allclose(r, [[0, 0, -1], [0, 1, 0], [1, 0, 0]], atol=1e-6)
    # world box centroid: unit rectangles and the [-1,1]^3 cube
    np.testing.assert_array_equal(np.array(info.env_center, np.float32), np.float32([0, 0, 0]))
    mats = f["mats"]
    assert mats[0]["tex"][gpr.TEX_COLOR] == 1 and mats[0]["tex"].sum() == 1
    assert mats[1]["tex"][gpr.TEX_COLOR] == 2
    assert mats[1]["tex"][gpr.TEX_ALPHA_X] == 3 and mats[1]["tex"][gpr.TEX_ALPHA_Y] == 3  # both read alphaU
    assert mats[1]["tex"][gpr.TEX_COLOR1] == 0
    assert mats[2]["tex"].sum() == 0  # a new diffuse material drops the texture
    s.free()


def test_gprs_round_trip_textures_and_free_camera(tmp_path):
    s = gpr.Scene()
    t1 = s.add_texture(np.random.default_rng(0).random((3, 5, 3)), uv_offset=(0.1, 0.2), uv_scale=(3, 4))
    t2 = s.add_checkerboard()
    s.set_env_map(t1, o2w=np.diag([1, 1, 1, 1]).astype(np.float32))
    s.add_mesh(gpr.gen_soup(1, 0, 4, 0.5).reshape(-1, 3),
               mat=gpr.material(gpr.MAT_ROUGH_CONDUCTOR, tex={gpr.TEX_K: t2, gpr.TEX_COLOR: t1}))
    s.set_camera_free((1, 2, 3), (0, 0, -1), (0, 1, 0), 50, 64, 48)
    p = str(tmp_path / "s.gprs")
    s.save(p)
    s2 = gpr.Scene.parse(p)
    s.flatten()
    a, b = s.flat(), s2.flat()
    for k in ("pos", "nrm", "uv", "flags", "mat", "texels"):
        np.testing.assert_array_equal(a[k], b[k])
    assert a["mats"].tobytes() == b["mats"].tobytes()
    assert a["tex"].tobytes() == b["tex"].tobytes()
    ia, ib = a["info"], b["info"]
    for k in ("env_tex", "cam_free", "fov_x", "fov_y", "yaw", "pitch", "width", "height"):
        assert getattr(ia, k) == getattr(ib, k), k
    for k in ("eye", "dir", "up", "env_w2o", "env_center"):
        assert list(getattr(ia, k)) == list(getattr(ib, k)), k
    s.free()
    s2.free()


def test_texture_handle_validation():
    s = gpr.Scene()
    with pytest.raises(gpr.GprError, match="texture handle"):
        s.add_mesh(gpr.gen_soup(1, 0, 2, 0.5).reshape(-1, 3), mat=gpr.material(tex={gpr.TEX_COLOR: 1}))
    with pytest.raises(gpr.GprError):
        s.set_env_map(3)
    s.free()


def test_published_cbox_render_fixture_decodes():
    # the reference's images/3006spp.png (an 8-bit RGB PNG with row filters):
    # 800 x 600, every value one of the 256 ReadImage levels pow(c / 255.f, 2.2f)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cbox", "cbox_sphere_3006spp.png")
    lin = gpr.read_image(path)
    assert lin.shape == (600, 800, 3) and lin.dtype == np.float32
    lut = np.array([np.float32(np.float64(np.float32(c) / np.float32(255)) ** np.float64(np.float32(2.2)))
                    for c in range(256)], np.float32)
    assert np.isin(np.unique(lin), lut).all()
    assert float(lin.min()) >= 0.0 and float(lin.max()) <= 1.0
