"""Textures and the latitude-longitude environment light on the GPU.

The shading kernel's texture lookups (TextureView, src/renderer/texture.h:
37-52) and environment lookup (EnvironmentLightView::Le, environment.h:16-28)
against the oracle: per-pixel bit-exact in the reference traversal order for
both RNG modes, on a scene that textures every kind of material parameter and
on a Mitsuba-XML scene read from disk.  The batched engine must give the same
bits as the single-scene renderer on a textured scene (its arena carries the
uvs, texture table and texels).
"""
import numpy as np
import pytest

import gpr
import gpu_scenes as G
from test_gpu_render import oracle_render, gpu_render
from test_textures import write_png, write_hdr

pytestmark = pytest.mark.gpu


def assert_same(g, o, label):
    bad = np.nonzero(np.any(g.view(np.uint32) != o.view(np.uint32), axis=1))[0]
    assert len(bad) == 0, (label, len(bad), bad[:8])


@pytest.mark.parametrize("rng_mode", [gpr.RNG_REFERENCE, gpr.RNG_PCG])
def test_textured_scene_bit_exact(rng_mode):
    s = G.textured(40, 32)
    s.commit()
    print(f"[textures] textured scene 40x32 rng={rng_mode}", flush=True)
    g, cnt, _, _ = gpu_render(s, 3, 6, rng_mode, 2, gpr.TRAVERSAL_REF)
    o, ocnt = oracle_render(s, 3, 6, rng_mode, 2)
    assert np.array_equal(cnt, ocnt)
    assert_same(g, o, f"rng={rng_mode}")
    # the textures matter: the same scene with constant parameters renders differently
    assert g.sum() > 0
    s.free()


def test_environment_map_only_misses():
    # no light source: everything the image holds comes through the env map
    s = G.textured(32, 24, env=True)
    s.commit()
    g, _, _, _ = gpu_render(s, 2, 1, gpr.RNG_PCG, 9, gpr.TRAVERSAL_REF)
    o, _ = oracle_render(s, 2, 1, gpr.RNG_PCG, 9)
    assert_same(g, o, "depth 1")
    s.free()


def test_xml_scene_with_textures_and_envmap(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (6, 10, 3), dtype=np.uint8)
    write_png(str(tmp_path / "wall.png"), 10, 6, 8, 2, [img[y].tobytes() for y in range(6)], [4] * 6)
    env = [[(int(a), int(b), int(c), 128) for a, b, c in rng.integers(1, 256, (24, 3))] for _ in range(12)]
    write_hdr(str(tmp_path / "sky.hdr"), env)
    (tmp_path / "s.xml").write_text("""<?xml version="1.0"?>
<scene version="0.5.0">
  <sensor type="perspective"><float name="fov" value="55"/>
    <transform name="toWorld"><lookat target="0 0 0" origin="0.5 1.2 4.5" up="0 1 0"/></transform>
    <film type="hdrfilm"><integer name="width" value="36"/><integer name="height" value="28"/></film></sensor>
  <bsdf type="diffuse" id="floor">
    <texture type="checkerboard" name="reflectance"><float name="uscale" value="5"/><float name="vscale" value="5"/></texture>
  </bsdf>
  <bsdf type="diffuse" id="wall">
    <texture type="bitmap" name="reflectance"><string name="filename" value="wall.png"/></texture>
  </bsdf>
  <shape type="rectangle"><transform name="toWorld"><scale value="3"/><rotate x="1" angle="-90"/>
    <translate y="-1"/></transform><ref id="floor"/></shape>
  <shape type="rectangle"><transform name="toWorld"><scale value="2"/><translate z="-2"/></transform><ref id="wall"/></shape>
  <shape type="cube"><transform name="toWorld"><scale value="0.4"/></transform>
    <bsdf type="roughconductor"><float name="alpha" value="0.15"/></bsdf></shape>
  <shape type="rectangle"><transform name="toWorld"><scale value="0.5"/><rotate x="1" angle="90"/>
    <translate y="2.5"/></transform><emitter type="area"><rgb name="radiance" value="8 8 8"/></emitter></shape>
  <emitter type="envmap"><string name="filename" value="sky.hdr"/>
    <transform name="toWorld"><rotate y="1" angle="45"/></transform></emitter>
</scene>""")
    s = gpr.Scene.load(str(tmp_path / "s.xml"))
    info = s.info()
    assert info.ntex == 3 and info.env_tex == 3
    print("[textures] xml scene 36x28", flush=True)
    for rng_mode in (gpr.RNG_REFERENCE, gpr.RNG_PCG):
        g, _, _, _ = gpu_render(s, 2, 5, rng_mode, 4, gpr.TRAVERSAL_REF)
        o, _ = oracle_render(s, 2, 5, rng_mode, 4)
        assert_same(g, o, f"xml rng={rng_mode}")
    s.free()


def test_textured_batch_equals_plain():
    s = G.textured(24, 20)
    s.commit()
    c = gpr.Scene.cornell(16, 16)
    b = gpr.Batch(gpr.arena_bytes(s) + gpr.arena_bytes(c), capacity=2048, max_jobs=4)
    ids = [b.add_scene(s), b.add_scene(c)]
    cam = gpr.lookat((0, 0.5, 0), (0.5, 1.5, 6.0))
    js = [gpr.job(ids[0], 24, 20, 3, cam, fov=50, max_depth=6, rr_start=3, seed=13),
          gpr.job(ids[1], 16, 16, 2, gpr.lookat((0, 0, 0), (0.3, 0.1, 3.9)), fov=40, max_depth=4, rr_start=2, seed=3)]
    for j in js:
        b.add_job(j)
    print("[textures] batch: textured scene + cornell", flush=True)
    b.run()
    got = b.read(0)
    s.set_camera(np.array(js[0].c2w[:], np.float32), js[0].fov_deg, js[0].width, js[0].height)
    r = gpr.Renderer(s, max_depth=6, rng_mode=gpr.RNG_PCG, rr_start=3, seed=13, traversal=gpr.TRAVERSAL_FAST)
    r.render(3)
    want, _ = r.read()
    r.free()
    assert_same(got, want, "batch")
    b.free()
    s.free()
    c.free()
