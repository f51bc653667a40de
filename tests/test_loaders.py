"""Scene ingest on the host: Mitsuba-XML subset, OBJ and the .gprs format.

The XML follows ParseRecord (src/renderer/loaders/mitsubaloader.cpp:262-805),
OBJ corners follow load_obj_file (objloader.cpp:12-121).  Checked against the
reference's own scene files (tests/golden/cbox, copied from scene/CBox by
tests/golden/copy_cbox_fixture.sh), an independent numpy OBJ reader, the
geometry the transforms must produce, and the defaults of createMaterial.
No GPU is used: gpr_scene_parse flattens on the host only.
"""
import math
import os
import struct

import numpy as np
import pytest

import gpr

HERE = os.path.dirname(os.path.abspath(__file__))
CBOX = os.path.join(HERE, "golden", "cbox")


def read_obj(path):
    """Independent OBJ reader: triangles as (v, vt, vn) corners, 0-based, -1 absent."""
    v, vt, vn, faces = [], [], [], []
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "v":
            v.append([np.float32(x) for x in t[1:4]])
        elif t[0] == "vt":
            vt.append([np.float32(x) for x in t[1:3]])
        elif t[0] == "vn":
            vn.append([np.float32(x) for x in t[1:4]])
        elif t[0] == "f":
            cs = []
            for tok in t[1:]:
                parts = tok.split("/") + ["", ""]
                fix = lambda s, n: -1 if s == "" else (int(s) - 1 if int(s) > 0 else n + int(s))
                cs.append((fix(parts[0], len(v)), fix(parts[1], len(vt)), fix(parts[2], len(vn))))
            for k in range(2, len(cs)):
                faces.append([cs[0], cs[k - 1], cs[k]])
    return (np.array(v, np.float32).reshape(-1, 3), np.array(vt, np.float32).reshape(-1, 2),
            np.array(vn, np.float32).reshape(-1, 3), np.array(faces, np.int32).reshape(-1, 3, 3))


def expected_flat(objs):
    pos, nrm = [], []
    for path in objs:
        v, vt, vn, f = read_obj(path)
        pos.append(v[f[:, :, 0]].reshape(-1, 9))
        has_n = f[:, 0, 2] >= 0
        n = np.zeros((len(f), 3, 3), np.float32)
        if has_n.any():
            n[has_n] = vn[f[has_n][:, :, 2]]
        nrm.append(n.reshape(-1, 9))
    return np.concatenate(pos), np.concatenate(nrm)


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def parse(path):
    s = gpr.Scene.parse(path)
    return s, s.flat()


# ------------------------------------------------------- reference scenes --
def test_cbox_xml_geometry_materials_camera():
    s, f = parse(os.path.join(CBOX, "cbox.xml"))
    order = ["sphere1", "sphere2", "walls", "rightwall", "leftwall", "light"]
    pos, nrm = expected_flat([os.path.join(CBOX, "meshes", m + ".obj") for m in order])
    info = f["info"]
    assert info.ntri == len(pos) == 3968 * 2 + 6 + 2 + 2 + 2
    np.testing.assert_array_equal(f["pos"], pos)  # identity toWorld: exact
    mats = f["mats"]
    assert len(mats) == 6
    # mirror: SPECULAR, reflectance 1, ior 100; glass: SPECULAR, ior 1.5 (mitsubaloader.cpp:298-320)
    assert mats[0]["type"] == gpr.MAT_SPECULAR and mats[0]["ior"] == np.float32(100)
    assert mats[1]["type"] == gpr.MAT_SPECULAR and mats[1]["ior"] == np.float32(1.5)
    np.testing.assert_array_equal(mats[2]["color"], np.float32([0.725, 0.71, 0.68]))
    np.testing.assert_array_equal(mats[3]["color"], np.float32([0.161, 0.133, 0.427]))
    np.testing.assert_array_equal(mats[4]["color"], np.float32([0.630, 0.065, 0.05]))
    # light: diffuse with the 0.3 default, emission 40 (createLight "area", :454-459)
    np.testing.assert_array_equal(mats[5]["color"], np.float32([0.3] * 3))
    np.testing.assert_array_equal(mats[5]["emission"], np.float32([40] * 3))
    for k in range(5):
        np.testing.assert_array_equal(mats[k]["emission"], np.zeros(3, np.float32))
    assert info.nlights == 2
    lights_first = int(info.ntri) - 2
    np.testing.assert_array_equal(f["mat"][lights_first:], [5, 5])
    # film inside the sensor sets the resolution the camera sees (createFilm, :282-287)
    assert (info.width, info.height) == (800, 600)
    assert info.fov_deg == pytest.approx(27.7856, abs=1e-6)
    c2w = np.array(info.c2w, np.float64).reshape(4, 4)
    origin = np.array([0, 0.919769, 5.41159])
    target = np.array([0, 0.893051, 4.41198])
    up = np.array([0, 1, 0.0])
    fwd = (target - origin) / np.linalg.norm(target - origin)
    left = np.cross(up, fwd)
    left /= np.linalg.norm(left)
    np.testing.assert_allclose(c2w[:3, 3], origin, atol=1e-6)
    np.testing.assert_allclose(c2w[:3, 2], fwd, atol=1e-6)
    np.testing.assert_allclose(c2w[:3, 0], left, atol=1e-6)
    np.testing.assert_allclose(c2w[:3, 1], np.cross(fwd, left), atol=1e-6)
    np.testing.assert_array_equal(c2w[3], [0, 0, 0, 1])


def test_cbox_normals_and_flags():
    s, f = parse(os.path.join(CBOX, "cbox.xml"))
    order = ["sphere1", "sphere2", "walls", "rightwall", "leftwall", "light"]
    objs = [os.path.join(CBOX, "meshes", m + ".obj") for m in order]
    pos, nrm = expected_flat(objs)
    flags = np.concatenate([
        np.where(read_obj(p)[3][:, 0, 2] >= 0, 1, 0) | np.where(read_obj(p)[3][:, 0, 1] >= 0, 2, 0)
        for p in objs])
    np.testing.assert_array_equal(f["flags"], flags.astype(np.uint8))
    has_n = (flags & 1) == 1
    # normalize(transformNormal(n)): the file's 4-digit normals come back unit length
    want = nrm[has_n].reshape(-1, 3).astype(np.float64)
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    np.testing.assert_allclose(f["nrm"][has_n].reshape(-1, 3), want, atol=1e-6)


def test_cbox_sphere_variant_uv_spheres():
    s, f = parse(os.path.join(CBOX, "cbox_sphere.xml"))
    info = f["info"]
    sub = 100
    per = (sub - 1) * 2 * sub * 2  # (nLat-1)*nLong*2 (mitsubaloader.cpp:218)
    assert info.ntri == 2 * per + 6 + 2 + 2 + 2
    r = np.float32(0.3261)
    for k, c in enumerate([(-0.4219, 0.3301, -0.2796), (0.4456, 0.3301, 0.3767)]):
        p = f["pos"][k * per:(k + 1) * per].reshape(-1, 3).astype(np.float64)
        d = np.linalg.norm(p - np.array(c), axis=1)
        np.testing.assert_allclose(d, r, rtol=3e-6)
        # the top pole (0,0,r) is the third corner of the first cap triangle
        np.testing.assert_allclose(f["pos"][k * per][6:9], np.float32(c) + np.float32([0, 0, r]), atol=1e-6)
        nz = f["nrm"][k * per:(k + 1) * per].reshape(-1, 3)
        np.testing.assert_allclose(np.linalg.norm(nz, axis=1), 1, atol=1e-6)
        np.testing.assert_allclose(nz, (p - np.array(c)) / r, atol=2e-5)
        # outward winding: geometric normal agrees with the shading normal
        t = p.reshape(-1, 3, 3)
        g = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
        mid = t.mean(axis=1) - np.array(c)
        assert (np.einsum("ij,ij->i", g, mid) > 0).mean() > 0.99
    # the light is listed after the bsdf in this variant: same emission
    np.testing.assert_array_equal(f["mats"][-1]["emission"], np.float32([40] * 3))


def test_cbox_free_camera_defaults():
    # cbox(mesh).xml sets <bool name="free" value="true"/>: the GUI camera of
    # Camera::Camera (eye (0,2,7), direction (0,-0.2,-1), up (0,1,0)) with the
    # sensor's fov on both film axes (camera.h:9-37)
    s = gpr.Scene.parse(os.path.join(CBOX, "cbox_mesh.xml"))
    info = s.info()
    assert info.cam_free == 1
    np.testing.assert_array_equal(np.array(info.eye, np.float32), np.float32([0, 2, 7]))
    np.testing.assert_array_equal(np.array(info.dir, np.float32), np.float32([0, -0.2, -1]))
    np.testing.assert_array_equal(np.array(info.up, np.float32), np.float32([0, 1, 0]))
    fov = np.float32(27.7856) / np.float32(180) * np.float32(np.pi)
    assert info.fov_x == info.fov_y and abs(info.fov_x - fov) < 1e-7
    assert abs(info.yaw - np.arctan2(0, -1)) < 1e-6 and abs(info.pitch - np.arcsin(-0.2)) < 1e-6


# ------------------------------------------------------------ XML subset ----
SCENE_HEAD = """<?xml version="1.0"?>
<!-- comment -->
<scene version="0.5.0">
  <sensor type="perspective">
    <float name="fov" value="45"/>
    <transform name="toWorld"><lookat target="0,0,0" origin="0,0,5" up="0,1,0"/></transform>
    <film type="hdrfilm"><integer name="width" value="64"/><integer name="height" value="32"/></film>
  </sensor>
"""


def test_transforms_compose_in_document_order(tmp_path):
    # m = I * T * R * S: scale first, then rotate 90 deg about z, then translate
    xml = SCENE_HEAD + """
  <shape type="rectangle">
    <transform name="toWorld">
      <translate x="1" y="2" z="3"/>
      <rotate z="1" angle="90"/>
      <scale value="2, 3, 4"/>
    </transform>
    <bsdf type="diffuse"><rgb name="reflectance" value="0.2, 0.4, 0.6"/></bsdf>
  </shape>
</scene>"""
    s, f = parse(write(tmp_path, "t.xml", xml))
    rect = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [1, 1, 0], [-1, 1, 0], [-1, -1, 0]], np.float64)
    sc = rect * [2, 3, 4]
    rot = np.stack([-sc[:, 1], sc[:, 0], sc[:, 2]], 1)
    want = rot + [1, 2, 3]
    np.testing.assert_allclose(f["pos"].reshape(-1, 3), want, atol=2e-6)
    # the normal (0,0,1) maps through the inverse transpose: stays +z
    np.testing.assert_allclose(f["nrm"].reshape(-1, 3), np.tile([0, 0, 1.0], (6, 1)), atol=1e-6)
    np.testing.assert_allclose(f["uv"].reshape(-1, 2),
                               [[0, 0], [1, 0], [1, 1], [1, 1], [0, 1], [0, 0]])
    np.testing.assert_array_equal(f["mats"][0]["color"], np.float32([0.2, 0.4, 0.6]))
    info = f["info"]
    assert (info.width, info.height, info.fov_deg) == (64, 32, 45)


def test_matrix_and_normal_transform(tmp_path):
    xml = SCENE_HEAD + """
  <shape type="cube">
    <transform name="toWorld"><matrix value="2 0 0 0  0 1 0 0  0 0 1 0  0 0 0 1"/></transform>
  </shape>
</scene>"""
    s, f = parse(write(tmp_path, "m.xml", xml))
    p = f["pos"].reshape(-1, 3)
    assert f["info"].ntri == 12
    np.testing.assert_array_equal(np.abs(p[:, 0]), 2)  # x stretched
    np.testing.assert_array_equal(np.abs(p[:, 1:]), 1)
    n = f["nrm"].reshape(-1, 3)
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    # each face's normal is axis-aligned and points away from the centre
    t = p.reshape(-1, 3, 3).mean(axis=1)
    assert (np.einsum("ij,ij->i", n.reshape(-1, 3, 3)[:, 0], t) > 0).all()


def test_material_defaults_and_named_refs(tmp_path):
    xml = SCENE_HEAD + """
  <bsdf type="roughconductor" id="gold">
    <float name="alpha" value="0.25"/>
    <rgb name="eta" value="0.2, 0.9, 1.1"/>
    <rgb name="k" value="3.9, 2.4, 2.1"/>
  </bsdf>
  <bsdf type="twosided" id="plastic2">
    <bsdf type="roughplastic"><float name="alphaU" value="0.3"/><float name="alphaV" value="0.7"/></bsdf>
  </bsdf>
  <shape type="rectangle"><ref id="gold"/></shape>
  <shape type="rectangle"><ref id="plastic2"/></shape>
  <shape type="rectangle"><bsdf type="roughdielectric"><float name="alpha" value="0.3"/></bsdf></shape>
  <shape type="rectangle"><bsdf type="conductor"/></shape>
  <shape type="rectangle"><bsdf type="something-else"/></shape>
  <emitter type="constant"><rgb name="radiance" value="0.5, 0.25, 0.125"/></emitter>
</scene>"""
    s, f = parse(write(tmp_path, "mat.xml", xml))
    m = f["mats"]
    assert m[0]["type"] == gpr.MAT_ROUGH_CONDUCTOR
    assert m[0]["alpha_x"] == m[0]["alpha_y"] == np.float32(0.25)
    np.testing.assert_array_equal(m[0]["eta_t"], np.float32([0.2, 0.9, 1.1]))
    np.testing.assert_array_equal(m[0]["eta_i"], np.float32([1, 1, 1]))
    np.testing.assert_array_equal(m[0]["color"], np.float32([0.3] * 3))
    # roughplastic: both roughness textures read "alphaU" (mitsubaloader.cpp:397-398)
    assert m[1]["type"] == gpr.MAT_ROUGH_PLASTIC
    assert m[1]["alpha_x"] == m[1]["alpha_y"] == np.float32(0.3)
    np.testing.assert_array_equal(m[1]["eta_i"], np.float32([1.5] * 3))
    np.testing.assert_array_equal(m[1]["eta_t"], np.float32([1] * 3))
    np.testing.assert_array_equal(m[1]["color"], np.float32([0.5] * 3))
    np.testing.assert_array_equal(m[1]["color1"], np.float32([1] * 3))
    # roughdielectric is the smooth dielectric in the reference's loader
    assert m[2]["type"] == gpr.MAT_SPECULAR and m[2]["ior"] == np.float32(1.5)
    # conductor defaults: Cu-like eta / k
    np.testing.assert_array_equal(m[3]["eta_t"], np.float32([0.200438, 0.924033, 1.10221]))
    np.testing.assert_array_equal(m[3]["k"], np.float32([3.91295, 2.45285, 2.14219]))
    assert m[3]["alpha_x"] == np.float32(0.1)
    # unknown type: diffuse 0.2
    assert m[4]["type"] == gpr.MAT_DIFFUSE
    np.testing.assert_array_equal(m[4]["color"], np.float32([0.2] * 3))
    info = f["info"]
    assert info.has_env == 1
    np.testing.assert_array_equal(np.array(info.env, np.float32), np.float32([0.5, 0.25, 0.125]))
    assert info.nlights == 0


def test_emitter_applies_to_next_shape_only(tmp_path):
    xml = SCENE_HEAD + """
  <shape type="rectangle">
    <emitter type="area"><rgb name="radiance" value="2, 4, 8"/><float name="scale" value="0.5"/></emitter>
  </shape>
  <shape type="rectangle"/>
</scene>"""
    s, f = parse(write(tmp_path, "e.xml", xml))
    np.testing.assert_array_equal(f["mats"][0]["emission"], np.float32([1, 2, 4]))
    np.testing.assert_array_equal(f["mats"][1]["emission"], np.float32([0, 0, 0]))
    assert f["info"].nlights == 2


@pytest.mark.parametrize("body,code,needle", [
    ("<shape type=\"rectangle\"><spectrum name=\"x\" value=\"400:1\"/></shape>", gpr.ERR_UNSUPPORTED, "spectrum"),
    ("<shape type=\"disk\"/>", gpr.ERR_UNSUPPORTED, "disk"),
    ("<bogus/>", gpr.ERR_PARSE, "bogus"),
    ("<shape type=\"rectangle\"><ref id=\"nope\"/></shape>", gpr.ERR_PARSE, "nope"),
    ("<shape type=\"rectangle\"></shap>", gpr.ERR_PARSE, "mismatched"),
    ("<shape type=\"rectangle\"><float name=\"r\" value=\"1x\"/></shape>", gpr.ERR_PARSE, "float"),
    ("<bsdf type=\"diffuse\"><texture name=\"reflectance\" type=\"bitmap\"/></bsdf>", gpr.ERR_PARSE, "filename"),
    ("<bsdf type=\"diffuse\"><texture name=\"reflectance\" type=\"wood\"/></bsdf>", gpr.ERR_UNSUPPORTED, "wood"),
    ("<bsdf type=\"diffuse\"><texture name=\"reflectance\" type=\"bitmap\">"
     "<string name=\"filename\" value=\"missing.png\"/></texture></bsdf>", gpr.ERR_IO, "missing.png"),
    ("<shape type=\"obj\"><string name=\"filename\" value=\"missing.obj\"/></shape>", gpr.ERR_IO, "missing.obj"),
    ("<emitter type=\"envmap\"/>", gpr.ERR_PARSE, "filename"),
    ("<shape type=\"sphere\"><integer name=\"subdiv\" value=\"100000\"/></shape>", gpr.ERR_UNSUPPORTED, "subdiv"),
    ("<shape type=\"sphere\"><integer name=\"subdiv\" value=\"1\"/></shape>", gpr.ERR_UNSUPPORTED, "subdiv"),
    ("<emitter type=\"envmap\"><string name=\"filename\" value=\"bad.xml\"/></emitter>", gpr.ERR_UNSUPPORTED,
     "not recognised"),
])
def test_xml_errors(tmp_path, body, code, needle):
    p = write(tmp_path, "bad.xml", SCENE_HEAD + body + "\n</scene>")
    with pytest.raises(gpr.GprError) as e:
        gpr.Scene.parse(p)
    assert e.value.code == code
    assert needle in str(e.value)


def test_xml_nesting_limit(tmp_path):
    # parsing, the scene walk and teardown recurse per level: a pathological
    # nesting depth is a parse error, not a stack overflow
    deep = "<hide>" * 100000 + "</hide>" * 100000
    p = write(tmp_path, "deep.xml", SCENE_HEAD + deep + "\n</scene>")
    with pytest.raises(gpr.GprError) as e:
        gpr.Scene.parse(p)
    assert e.value.code == gpr.ERR_PARSE and "nested deeper" in str(e.value)
    ok = "<hide>" * 200 + "<shape type=\"cube\"/>" + "</hide>" * 200
    gpr.Scene.parse(write(tmp_path, "ok.xml", SCENE_HEAD + ok + "\n<shape type=\"rectangle\"/>\n</scene>"))


def test_hide_skips_subtree(tmp_path):
    xml = SCENE_HEAD + """
  <hide><shape type="cube"/></hide>
  <shape type="rectangle"/>
</scene>"""
    s, f = parse(write(tmp_path, "h.xml", xml))
    assert f["info"].ntri == 2


# -------------------------------------------------------------------- OBJ ---
def test_obj_polygons_negative_indices_and_corners(tmp_path):
    obj = """# quad with uvs and normals, then a triangle by negative indices
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 1
f 1/1/1 2/2/1 3/3/1 4/4/1
v 0 0 1
f -5//1 -4//1 -1//1
f 1 2 5
"""
    p = write(tmp_path, "q.obj", obj)
    s, f = parse(p)
    v, vt, vn, faces = read_obj(p)
    assert f["info"].ntri == len(faces) == 4
    np.testing.assert_array_equal(f["pos"], v[faces[:, :, 0]].reshape(-1, 9))
    np.testing.assert_array_equal(f["flags"], [3, 3, 1, 0])
    np.testing.assert_array_equal(f["uv"][:2], vt[faces[:2, :, 1]].reshape(-1, 6))
    # one Lambert mesh, 0.5 reflectance, camera on +z looking at the bounds centre
    assert f["mats"][0]["type"] == gpr.MAT_DIFFUSE
    np.testing.assert_array_equal(f["mats"][0]["color"], np.float32([0.5] * 3))
    c2w = np.array(f["info"].c2w).reshape(4, 4)
    np.testing.assert_allclose(c2w[:3, 2], [0, 0, -1], atol=1e-6)


@pytest.mark.parametrize("obj,needle", [
    ("v 0 0 0\nv 1 0 0\nf 1 2 3\n", "out of range"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\n", "< 3"),
    ("v 0 0\n", "bad v"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2 3\n", "normals"),
])
def test_obj_errors(tmp_path, obj, needle):
    with pytest.raises(gpr.GprError) as e:
        gpr.Scene.parse(write(tmp_path, "bad.obj", obj))
    assert e.value.code == gpr.ERR_PARSE
    assert needle in str(e.value)


# ------------------------------------------------------------------ .gprs ---
def test_gprs_round_trip(tmp_path):
    s, f = parse(os.path.join(CBOX, "cbox.xml"))
    out = str(tmp_path / "cbox.gprs")
    s.save(out)
    s2, f2 = parse(out)
    for k in ("pos", "nrm", "uv", "flags", "mat"):
        np.testing.assert_array_equal(f[k], f2[k])
    assert f["mats"].tobytes() == f2["mats"].tobytes()
    a, b = f["info"], f2["info"]
    for k in ("ntri", "nlights", "nmat", "width", "height", "fov_deg", "has_env"):
        assert getattr(a, k) == getattr(b, k)
    assert list(a.c2w) == list(b.c2w)


def test_gprs_rejects_truncated(tmp_path):
    s, f = parse(os.path.join(CBOX, "cbox.xml"))
    out = tmp_path / "t.gprs"
    s.save(str(out))
    data = out.read_bytes()
    (tmp_path / "cut.gprs").write_bytes(data[: len(data) // 2])
    with pytest.raises(gpr.GprError) as e:
        gpr.Scene.parse(str(tmp_path / "cut.gprs"))
    assert e.value.code == gpr.ERR_PARSE
    (tmp_path / "magic.gprs").write_bytes(b"XXXX" + data[4:])
    with pytest.raises(gpr.GprError):
        gpr.Scene.parse(str(tmp_path / "magic.gprs"))
    # a header that claims ~2^30 triangles in a small file is rejected before
    # anything is allocated for them (first mesh header after the 208-byte v2
    # scene header of a texture-free scene: material, np, nn, nuv, nt)
    nt_at = 208 + gpr.MATERIAL_DTYPE.itemsize + 12
    big = bytearray(data)
    big[nt_at:nt_at + 4] = struct.pack("<I", (1 << 30) - 1)
    (tmp_path / "big.gprs").write_bytes(bytes(big))
    with pytest.raises(gpr.GprError) as e:
        gpr.Scene.parse(str(tmp_path / "big.gprs"))
    assert e.value.code == gpr.ERR_PARSE and "implausible mesh size" in str(e.value)


def test_unknown_extension(tmp_path):
    with pytest.raises(gpr.GprError) as e:
        gpr.Scene.parse(write(tmp_path, "a.ply", "ply"))
    assert e.value.code == gpr.ERR_UNSUPPORTED


def test_cli_help_and_load_error(capsys):
    from gpr.__main__ import main
    with pytest.raises(SystemExit) as e:
        main(["--help"])
    assert e.value.code == 0
    assert main([os.path.join(CBOX, "no_such_scene.xml"), "--spp", "1"]) == 1
    assert "cannot open" in capsys.readouterr().err
