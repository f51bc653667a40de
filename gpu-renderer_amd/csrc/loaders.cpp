// loaders.cpp -- scene ingest on the host: Mitsuba-XML subset, Wavefront OBJ
// and the binary .gprs format (load_scene, src/renderer/loaders/loader.h:11-29).
//
// The XML subset follows ParseRecord (src/renderer/loaders/mitsubaloader.cpp):
// children are handled before their parent; value tags fill the parent's
// property list; <transform> starts from identity and its lookat / translate /
// scale / rotate / matrix children compose onto it; a <bsdf> sets the current
// material (fields it does not set are kept from the previous one), an area
// <emitter> sets the radiance the next shape receives; each <shape> takes the
// current material and is moved to world space by its toWorld transform.
// OBJ follows tinyobjloader's reading of v / vt / vn / f (1-based or negative
// indices, polygons fanned into triangles, corners as (v, vt, vn) with -1 for
// an absent index).  The reference's top-level .obj path is dead code
// (loader.h:17-29); here a top-level .obj becomes one diffuse mesh with a
// camera framing its bounds.
#include <hip/hip_runtime.h>

#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "gpr_internal.h"

namespace gpr {
namespace {

// ------------------------------------------------------------------ XML ---
struct XNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<XNode>> kids;
    std::string attr(const std::string& k, const std::string& def = "") const {
        for (auto& a : attrs)
            if (a.first == k) return a.second;
        return def;
    }
    bool has(const std::string& k) const {
        for (auto& a : attrs)
            if (a.first == k) return true;
        return false;
    }
};

struct XmlParser {
    static constexpr int kMaxDepth = 256;  // element nesting (parsing, the scene walk and teardown recurse per level)
    const std::string& s;
    size_t i = 0;
    int depth = 0;
    std::string err;
    explicit XmlParser(const std::string& src) : s(src) {}
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool starts(const char* t) const { return s.compare(i, std::strlen(t), t) == 0; }
    bool skip_misc() {
        // comments, declarations, processing instructions, text
        while (true) {
            ws();
            if (starts("<!--")) {
                size_t e = s.find("-->", i);
                if (e == std::string::npos) return fail_at("unterminated comment");
                i = e + 3;
            } else if (starts("<?")) {
                size_t e = s.find("?>", i);
                if (e == std::string::npos) return fail_at("unterminated declaration");
                i = e + 2;
            } else if (starts("<!")) {
                size_t e = s.find('>', i);
                if (e == std::string::npos) return fail_at("unterminated directive");
                i = e + 1;
            } else if (i < s.size() && s[i] != '<') {
                while (i < s.size() && s[i] != '<') ++i;  // character data is ignored
            } else {
                return true;
            }
        }
    }
    bool fail_at(const std::string& m) {
        size_t line = 1;
        for (size_t k = 0; k < i && k < s.size(); ++k) line += s[k] == '\n';
        err = m + " (line " + std::to_string(line) + ")";
        return false;
    }
    std::string name() {
        size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == ':' || s[i] == '.')) ++i;
        return s.substr(b, i - b);
    }
    std::unique_ptr<XNode> element() {
        struct Level {
            int& d;
            explicit Level(int& x) : d(++x) {}
            ~Level() { --d; }
        } level(depth);
        if (depth > kMaxDepth) {
            fail_at("elements nested deeper than " + std::to_string(kMaxDepth) + " levels");
            return nullptr;
        }
        if (!skip_misc()) return nullptr;
        if (i >= s.size() || s[i] != '<') {
            fail_at("expected an element");
            return nullptr;
        }
        ++i;
        auto n = std::make_unique<XNode>();
        n->name = name();
        if (n->name.empty()) {
            fail_at("missing element name");
            return nullptr;
        }
        while (true) {
            ws();
            if (i >= s.size()) {
                fail_at("unterminated tag <" + n->name + ">");
                return nullptr;
            }
            if (starts("/>")) {
                i += 2;
                return n;
            }
            if (s[i] == '>') {
                ++i;
                break;
            }
            std::string k = name();
            ws();
            if (k.empty() || i >= s.size() || s[i] != '=') {
                fail_at("malformed attribute in <" + n->name + ">");
                return nullptr;
            }
            ++i;
            ws();
            if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) {
                fail_at("unquoted attribute value in <" + n->name + ">");
                return nullptr;
            }
            const char q = s[i++];
            size_t e = s.find(q, i);
            if (e == std::string::npos) {
                fail_at("unterminated attribute value");
                return nullptr;
            }
            n->attrs.emplace_back(k, s.substr(i, e - i));
            i = e + 1;
        }
        while (true) {
            if (!skip_misc()) return nullptr;
            if (starts("</")) {
                i += 2;
                std::string close = name();
                ws();
                if (close != n->name || i >= s.size() || s[i] != '>') {
                    fail_at("mismatched closing tag </" + close + "> for <" + n->name + ">");
                    return nullptr;
                }
                ++i;
                return n;
            }
            auto k = element();
            if (!k) return nullptr;
            n->kids.push_back(std::move(k));
        }
    }
};

// --------------------------------------------------------- transforms ---
struct M4 {
    float m[4][4];
};
M4 ident() {
    M4 r;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) r.m[a][b] = a == b ? 1.f : 0.f;
    return r;
}
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q)
            r.m[p][q] = a.m[p][0] * b.m[0][q] + a.m[p][1] * b.m[1][q] + a.m[p][2] * b.m[2][q] + a.m[p][3] * b.m[3][q];
    return r;
}
// Inverse (transform.cpp:30-90): the Gauss-Jordan elimination shared with
// the scene API (api.cpp).
M4 inverse(const M4& a) {
    M4 r;
    (void)invert4(&a.m[0][0], &r.m[0][0]);
    return r;
}
struct Xform {
    M4 m = ident(), inv = ident();
};
Xform compose(const Xform& a, const Xform& b) {  // Transform::operator*= (transform.h:110-114)
    Xform r;
    r.m = mul(a.m, b.m);
    r.inv = mul(b.inv, a.inv);
    return r;
}
struct F3 {
    float x, y, z;
};
F3 nrm(F3 v) {
    const float inv = 1.0f / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return F3{v.x * inv, v.y * inv, v.z * inv};
}
F3 crs(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
Xform xf_translate(F3 d) {
    Xform t;
    t.m.m[0][3] = d.x; t.m.m[1][3] = d.y; t.m.m[2][3] = d.z;
    t.inv.m[0][3] = -d.x; t.inv.m[1][3] = -d.y; t.inv.m[2][3] = -d.z;
    return t;
}
Xform xf_scale(F3 s) {
    Xform t;
    t.m.m[0][0] = s.x; t.m.m[1][1] = s.y; t.m.m[2][2] = s.z;
    t.inv.m[0][0] = (float)(1. / s.x); t.inv.m[1][1] = (float)(1. / s.y); t.inv.m[2][2] = (float)(1. / s.z);
    return t;
}
// Rotate(theta, axis) (transform.cpp:195-218): inverse = transpose
Xform xf_rotate(float deg, F3 axis) {
    const F3 a = nrm(axis);
    const float th = deg / 180 * 3.14159265358979323846f;
    const float s = std::sin(th), c = std::cos(th);
    Xform t;
    M4& m = t.m;
    m.m[0][0] = a.x * a.x + (1 - a.x * a.x) * c;
    m.m[0][1] = a.x * a.y * (1 - c) - a.z * s;
    m.m[0][2] = a.x * a.z * (1 - c) + a.y * s;
    m.m[1][0] = a.x * a.y * (1 - c) + a.z * s;
    m.m[1][1] = a.y * a.y + (1 - a.y * a.y) * c;
    m.m[1][2] = a.y * a.z * (1 - c) - a.x * s;
    m.m[2][0] = a.x * a.z * (1 - c) - a.y * s;
    m.m[2][1] = a.y * a.z * (1 - c) + a.x * s;
    m.m[2][2] = a.z * a.z + (1 - a.z * a.z) * c;
    for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) t.inv.m[p][q] = m.m[q][p];
    return t;
}
Xform xf_lookat(F3 target, F3 origin, F3 up) {
    float m[16];
    const float tg[3] = {target.x, target.y, target.z}, og[3] = {origin.x, origin.y, origin.z}, u[3] = {up.x, up.y, up.z};
    lookat_matrix(tg, og, u, m);
    Xform t;
    std::memcpy(t.m.m, m, sizeof(m));
    t.inv = inverse(t.m);
    return t;
}

// strtof-based readers of "a, b, c" / "a b c" (mitsubaloader.cpp:15-51)
bool parse_floats(const std::string& s, float* out, int n, int* got) {
    const char* p = s.c_str();
    int k = 0;
    while (k < n) {
        while (*p && (std::isspace((unsigned char)*p) || *p == ',')) ++p;
        if (!*p) break;
        char* e = nullptr;
        out[k] = std::strtof(p, &e);
        if (e == p) return false;
        ++k;
        p = e;
    }
    while (*p && (std::isspace((unsigned char)*p) || *p == ',')) ++p;
    *got = k;
    return *p == 0;
}

// ------------------------------------------------------------- props ----
struct Props {
    std::map<std::string, float> f;
    std::map<std::string, int> i;
    std::map<std::string, bool> b;
    std::map<std::string, std::string> s;
    std::map<std::string, F3> v;
    std::map<std::string, F3> c;
    std::map<std::string, Xform> t;
    std::map<std::string, uint32_t> tex;  // texture handles of nested <texture> objects
    float getf(const std::string& k, float d) const {
        auto it = f.find(k);
        if (it != f.end()) return it->second;
        auto ii = i.find(k);
        return ii != i.end() ? (float)ii->second : d;
    }
    int geti(const std::string& k, int d) const {
        auto it = i.find(k);
        return it != i.end() ? it->second : d;
    }
    bool getb(const std::string& k, bool d) const {
        auto it = b.find(k);
        return it != b.end() ? it->second : d;
    }
    F3 getc(const std::string& k, F3 d) const {
        auto it = c.find(k);
        if (it != c.end()) return it->second;
        auto fi = f.find(k);  // a float given for a colour is grey
        return fi != f.end() ? F3{fi->second, fi->second, fi->second} : d;
    }
    bool hasf(const std::string& k) const { return f.count(k) || i.count(k); }
};

struct MeshData {
    std::vector<float> p, n, uv;
    std::vector<int32_t> c;
    uint32_t ntri = 0;
};

// ------------------------------------------------------------- OBJ -------
int load_obj(const std::string& path, MeshData* out) {
    std::ifstream in(path);
    if (!in) return fail(GPR_ERR_IO, "cannot open " + path);
    MeshData m;
    std::string line;
    size_t lineno = 0;
    auto fix = [](long idx, size_t count) -> long {
        if (idx > 0) return idx - 1;
        if (idx < 0) return (long)count + idx;
        return -2;  // 0 is invalid in OBJ
    };
    while (std::getline(in, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag) || tag[0] == '#') continue;
        if (tag == "v" || tag == "vn") {
            float x, y, z;
            if (!(ss >> x >> y >> z)) return fail(GPR_ERR_PARSE, path + ":" + std::to_string(lineno) + ": bad " + tag);
            auto& dst = tag == "v" ? m.p : m.n;
            dst.insert(dst.end(), {x, y, z});
        } else if (tag == "vt") {
            float u, v = 0;
            if (!(ss >> u)) return fail(GPR_ERR_PARSE, path + ":" + std::to_string(lineno) + ": bad vt");
            ss >> v;
            m.uv.insert(m.uv.end(), {u, v});
        } else if (tag == "f") {
            std::vector<std::array<long, 3>> face;
            std::string tok;
            while (ss >> tok) {
                long vi = 0, ti = 0, ni = 0;
                int parts = 0;
                size_t a = tok.find('/');
                vi = std::strtol(tok.c_str(), nullptr, 10);
                parts = 1;
                if (a != std::string::npos) {
                    size_t b = tok.find('/', a + 1);
                    std::string t1 = tok.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
                    if (!t1.empty()) ti = std::strtol(t1.c_str(), nullptr, 10);
                    if (b != std::string::npos) ni = std::strtol(tok.c_str() + b + 1, nullptr, 10);
                    parts = 3;
                }
                (void)parts;
                const long v = fix(vi, m.p.size() / 3);
                const long t = ti ? fix(ti, m.uv.size() / 2) : -1;
                const long nn = ni ? fix(ni, m.n.size() / 3) : -1;
                if (v < 0 || v >= (long)(m.p.size() / 3) || t < -1 || t >= (long)(m.uv.size() / 2) || nn < -1 ||
                    nn >= (long)(m.n.size() / 3))
                    return fail(GPR_ERR_PARSE, path + ":" + std::to_string(lineno) + ": face index out of range");
                face.push_back({v, t, nn});
            }
            if (face.size() < 3) return fail(GPR_ERR_PARSE, path + ":" + std::to_string(lineno) + ": face with < 3 vertices");
            for (size_t k = 2; k < face.size(); ++k) {  // fan triangulation
                for (const auto& cv : {face[0], face[k - 1], face[k]}) {
                    m.c.push_back((int32_t)cv[0]);
                    m.c.push_back((int32_t)cv[1]);
                    m.c.push_back((int32_t)cv[2]);
                }
                ++m.ntri;
            }
        }  // o, g, s, usemtl, mtllib and other records carry no geometry
    }
    // corners of one triangle must agree on having uvs / normals
    for (uint32_t t = 0; t < m.ntri; ++t) {
        const int32_t* c = &m.c[9ull * t];
        const bool hn = c[2] != -1, ht = c[1] != -1;
        for (int k = 1; k < 3; ++k) {
            if ((c[3 * k + 2] != -1) != hn) return fail(GPR_ERR_PARSE, path + ": face mixes corners with and without normals");
            if ((c[3 * k + 1] != -1) != ht) return fail(GPR_ERR_PARSE, path + ": face mixes corners with and without uvs");
        }
    }
    *out = std::move(m);
    return GPR_OK;
}

// ----------------------------------------------------------- shapes ------
// create_cube_triangles / create_rectangle_triangles / create_sphere_triangles
// (mitsubaloader.cpp:53-221)
void make_rectangle(MeshData& m) {
    m.p = {-1, -1, 0, 1, -1, 0, 1, 1, 0, -1, 1, 0};
    m.uv = {0, 0, 1, 0, 1, 1, 0, 1};
    m.n = {0, 0, 1};
    m.c = {0, 0, 0, 1, 1, 0, 2, 2, 0, 2, 2, 0, 3, 3, 0, 0, 0, 0};
    m.ntri = 2;
}
void make_cube(MeshData& m) {
    const float P[8][3] = {{1, 1, -1}, {1, -1, -1}, {1, 1, 1}, {1, -1, 1}, {-1, 1, -1}, {-1, -1, -1}, {-1, 1, 1}, {-1, -1, 1}};
    const float N[6][3] = {{0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {1, 0, 0}, {0, 0, -1}};
    const float T[14][2] = {{0.625f, 0.5f}, {0.875f, 0.5f}, {0.875f, 0.75f}, {0.625f, 0.75f}, {0.375f, 0.75f},
                            {0.625f, 1.f}, {0.375f, 1.f}, {0.375f, 0.f}, {0.625f, 0.f}, {0.625f, 0.25f},
                            {0.375f, 0.25f}, {0.125f, 0.5f}, {0.375f, 0.5f}, {0.125f, 0.75f}};
    // corner (v, vt, vn) of the 4 corners of each face
    const int Q[6][4][3] = {{{0, 0, 0}, {4, 1, 0}, {6, 2, 0}, {2, 3, 0}},
                            {{3, 4, 1}, {2, 3, 1}, {6, 5, 1}, {7, 6, 1}},
                            {{7, 7, 2}, {6, 8, 2}, {4, 9, 2}, {5, 10, 2}},
                            {{5, 11, 3}, {1, 12, 3}, {3, 4, 3}, {7, 13, 3}},
                            {{1, 12, 4}, {0, 0, 4}, {2, 3, 4}, {3, 4, 4}},
                            {{5, 10, 5}, {4, 9, 5}, {0, 0, 5}, {1, 12, 5}}};
    for (auto& v : P) m.p.insert(m.p.end(), v, v + 3);
    for (auto& v : N) m.n.insert(m.n.end(), v, v + 3);
    for (auto& v : T) m.uv.insert(m.uv.end(), v, v + 2);
    for (int f = 0; f < 6; ++f)
        for (int k : {0, 1, 2, 0, 2, 3}) m.c.insert(m.c.end(), {Q[f][k][0], Q[f][k][1], Q[f][k][2]});
    m.ntri = 12;
}
void make_sphere(MeshData& m, float radius, int subdiv) {
    const int nlat = subdiv, nlong = 2 * subdiv;
    const float dlat = 3.14159265358979323846f / nlat, dlong = 2 * 3.14159265358979323846f / nlong;
    auto vert = [&](float x, float y, float z) {
        m.p.insert(m.p.end(), {x * radius, y * radius, z * radius});
        m.n.insert(m.n.end(), {x, y, z});
    };
    vert(0, 0, 1);
    for (int i = 1; i < nlat; ++i)
        for (int j = 0; j < nlong; ++j) {
            const float th = i * dlat, ph = j * dlong;
            vert(std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th));
        }
    vert(0, 0, -1);
    auto tri = [&](int a, int b, int c) {
        for (int v : {a, b, c}) m.c.insert(m.c.end(), {v, -1, v});
        ++m.ntri;
    };
    for (int i = 0; i < nlong; ++i) tri(i + 1, i == nlong - 1 ? 1 : i + 2, 0);
    for (int i = 2; i < nlat; ++i)
        for (int j = 0; j < nlong; ++j) {
            const int a = (i - 1) * nlong + j + 1;
            int b = a + 1, c = a - nlong, d = c + 1;
            if (j == nlong - 1) {
                b = (i - 1) * nlong + 1;
                d = b - nlong;
            }
            tri(a, b, c);
            tri(c, b, d);
        }
    const int bottom = nlong * (nlat - 1) + 1;
    for (int i = 0; i < nlong; ++i) {
        const int a = (nlat - 2) * nlong + i + 1;
        tri(a, bottom, i == nlong - 1 ? (nlat - 2) * nlong + 1 : a + 1);
    }
}

// ------------------------------------------------------- Mitsuba XML -----
struct Loader {
    std::string dir;
    gpr_scene* scene = nullptr;
    Xform cur_xf;
    gpr_material cur_mat{};
    F3 cur_light{0, 0, 0};
    std::map<std::string, gpr_material> named;
    uint32_t film_w = 256, film_h = 256;
    bool cam = false, cam_free = false;
    float cam_c2w[16];
    float cam_fov = 30.f;
    int err = GPR_OK;

    Loader() {
        cur_mat.type = GPR_MAT_DIFFUSE;
        cur_mat.color[0] = cur_mat.color[1] = cur_mat.color[2] = 0.5f;
    }
    int bad(int code, const std::string& m) {
        err = fail(code, m);
        return err;
    }
    void set_color(float* dst, F3 c) {
        dst[0] = c.x;
        dst[1] = c.y;
        dst[2] = c.z;
    }
    // createTexture (mitsubaloader.cpp:530-554): bitmap (ReadImage) or 2x2 checkerboard
    int texture(const std::string& type, const Props& L, uint32_t* id) {
        const float off[2] = {L.getf("uoffset", 0.f), L.getf("voffset", 0.f)};
        const float sc[2] = {L.getf("uscale", 1.f), L.getf("vscale", 1.f)};
        if (type == "bitmap") {
            auto it = L.s.find("filename");
            if (it == L.s.end()) return bad(GPR_ERR_PARSE, "bitmap texture without filename");
            const std::string f = it->second.size() && it->second[0] == '/' ? it->second : dir + it->second;
            uint32_t w = 0, h = 0;
            std::vector<float> px;
            int rc = read_image(f, &w, &h, &px);
            if (rc == GPR_OK) rc = gpr_scene_add_texture(scene, px.data(), w, h, off, sc, id);
            return rc == GPR_OK ? GPR_OK : (err = rc);
        } else if (type == "checkerboard") {
            const F3 a = L.getc("color0", {0.4f, 0.4f, 0.4f}), b = L.getc("color1", {0.2f, 0.2f, 0.2f});
            const float c0[3] = {a.x, a.y, a.z}, c1[3] = {b.x, b.y, b.z};
            const int rc = gpr_scene_add_checkerboard(scene, c0, c1, off, sc, id);
            return rc == GPR_OK ? GPR_OK : (err = rc);
        }
        return bad(GPR_ERR_UNSUPPORTED, "texture type '" + type + "'");
    }
    // createMaterial (mitsubaloader.cpp:289-441).  Every parameter a type sets
    // is a texture when the property is a nested <texture>, else a constant.
    int material(const std::string& type, const Props& L) {
        gpr_material& m = cur_mat;
        auto tx = [&](int slot, const char* key) {
            auto it = L.tex.find(key);
            m.tex[slot] = it == L.tex.end() ? 0u : it->second;
        };
        // Parameters a type does not read keep their previous values in the
        // reference's current material without effect; their textures are
        // dropped here so that the shading kernel does not look them up.
        if (type != "disney") std::memset(m.tex, 0, sizeof(m.tex));
        if (type == "diffuse") {
            set_color(m.color, L.getc("reflectance", {0.3f, 0.3f, 0.3f}));
            tx(GPR_TEX_COLOR, "reflectance");
            m.type = GPR_MAT_DIFFUSE;
        } else if (type == "mirror") {
            set_color(m.color, L.getc("reflectance", {1, 1, 1}));
            m.ior = L.getf("ior", 100.f);
            tx(GPR_TEX_COLOR, "reflectance");
            tx(GPR_TEX_IOR, "ior");
            m.type = GPR_MAT_SPECULAR;
        } else if (type == "glass") {
            set_color(m.color, L.getc("reflectance", {1, 1, 1}));
            m.ior = L.getf("ior", 1.5f);
            tx(GPR_TEX_COLOR, "reflectance");
            tx(GPR_TEX_IOR, "ior");
            m.type = GPR_MAT_SPECULAR;
        } else if (type == "dielectric" || type == "roughdielectric") {
            // Both map to the smooth dielectric and intIOR/extIOR are ignored: the
            // reference computes their ratio but stores the "ior" value
            // (mitsubaloader.cpp:321-342).  The rough dielectric is reachable
            // through gpr_scene_add_mesh with GPR_MAT_ROUGH_DIELECTRIC.
            set_color(m.color, L.getc("reflectance", {1, 1, 1}));
            m.ior = L.getf("ior", 1.5f);
            tx(GPR_TEX_COLOR, "reflectance");
            tx(GPR_TEX_IOR, "ior");
            m.type = GPR_MAT_SPECULAR;
        } else if (type == "conductor" || type == "roughconductor") {
            set_color(m.color, L.getc("specularReflectance", {0.3f, 0.3f, 0.3f}));
            const bool has_eta_i = L.c.count("etaI") != 0 || L.tex.count("etaI") != 0;  // findColor: the key exists
            set_color(m.eta_i, L.getc("etaI", {1, 1, 1}));
            set_color(m.eta_t, L.getc(has_eta_i ? "etaT" : "eta", {0.200438f, 0.924033f, 1.10221f}));
            set_color(m.k, L.getc("k", {3.91295f, 2.45285f, 2.14219f}));
            float au, av;
            if (L.hasf("alphaU")) {
                au = L.getf("alphaU", 0.1f);
                av = L.getf("alphaV", 0.1f);
            } else {
                au = av = L.getf("alpha", 0.1f);
            }
            // both roughness textures read the "alphaU" key (mitsubaloader.cpp:368-369)
            m.alpha_x = L.hasf("alphaU") ? L.getf("alphaU", au) : au;
            m.alpha_y = L.hasf("alphaU") ? L.getf("alphaU", av) : av;
            tx(GPR_TEX_COLOR, "specularReflectance");
            tx(GPR_TEX_ETA_I, "etaI");
            tx(GPR_TEX_ETA_T, has_eta_i ? "etaT" : "eta");
            tx(GPR_TEX_K, "k");
            tx(GPR_TEX_ALPHA_X, "alphaU");
            tx(GPR_TEX_ALPHA_Y, "alphaU");
            m.type = GPR_MAT_ROUGH_CONDUCTOR;
        } else if (type == "roughplastic") {
            const bool has_eta_i = L.c.count("etaI") != 0 || L.tex.count("etaI") != 0;
            set_color(m.eta_i, L.getc(has_eta_i ? "etaI" : "eta", {1.5f, 1.5f, 1.5f}));
            set_color(m.eta_t, L.getc(has_eta_i ? "etaT" : "etaI", {1, 1, 1}));
            float au, av;
            if (L.hasf("alphaU")) {
                au = L.getf("alphaU", 0.1f);
                av = L.getf("alphaV", 0.1f);
            } else {
                au = av = L.getf("alpha", 0.1f);
            }
            m.alpha_x = L.hasf("alphaU") ? L.getf("alphaU", au) : au;
            m.alpha_y = L.hasf("alphaU") ? L.getf("alphaU", av) : av;
            set_color(m.color, L.getc("diffuseReflectance", {0.5f, 0.5f, 0.5f}));
            set_color(m.color1, L.getc("specularReflectance", {1, 1, 1}));
            tx(GPR_TEX_ETA_I, has_eta_i ? "etaI" : "eta");
            tx(GPR_TEX_ETA_T, has_eta_i ? "etaT" : "etaI");
            tx(GPR_TEX_ALPHA_X, "alphaU");
            tx(GPR_TEX_ALPHA_Y, "alphaU");
            tx(GPR_TEX_COLOR, "diffuseReflectance");
            tx(GPR_TEX_COLOR1, "specularReflectance");
            m.type = GPR_MAT_ROUGH_PLASTIC;
        } else if (type == "disney") {
            // parsed but without effect in the reference (mitsubaloader.cpp:429-435)
        } else {
            set_color(m.color, {0.2f, 0.2f, 0.2f});
            m.tex[GPR_TEX_COLOR] = 0;
            m.type = GPR_MAT_DIFFUSE;
        }
        return GPR_OK;
    }
    int shape(const std::string& type, const Props& L) {
        MeshData m;
        if (type == "obj") {
            auto it = L.s.find("filename");
            if (it == L.s.end()) return bad(GPR_ERR_PARSE, "obj shape without filename");
            const std::string f = it->second.size() && it->second[0] == '/' ? it->second : dir + it->second;
            int rc = load_obj(f, &m);
            if (rc != GPR_OK) return err = rc;
        } else if (type == "sphere") {
            // 2 .. 4096 latitude bands (4096: 33.5 M triangles); outside that the
            // tessellation has no valid mesh or its int indices would overflow
            const int subdiv = L.geti("subdiv", 30);
            if (subdiv < 2 || subdiv > 4096)
                return bad(GPR_ERR_UNSUPPORTED, "sphere subdiv " + std::to_string(subdiv) + " outside 2..4096");
            make_sphere(m, L.getf("radius", 1.f), subdiv);
        } else if (type == "rectangle") {
            make_rectangle(m);
        } else if (type == "cube") {
            make_cube(m);
        } else {
            return bad(GPR_ERR_UNSUPPORTED, "shape type '" + type + "'");
        }
        Xform o2w;
        auto it = L.t.find("toWorld");
        if (it != L.t.end()) o2w = it->second;
        const M4& X = o2w.m;
        for (size_t k = 0; k + 2 < m.p.size(); k += 3) {  // transformPoint (transform.h:162-174)
            const float px = m.p[k], py = m.p[k + 1], pz = m.p[k + 2];
            float x = X.m[0][0] * px + X.m[0][1] * py + X.m[0][2] * pz + X.m[0][3];
            float y = X.m[1][0] * px + X.m[1][1] * py + X.m[1][2] * pz + X.m[1][3];
            float z = X.m[2][0] * px + X.m[2][1] * py + X.m[2][2] * pz + X.m[2][3];
            const float w = X.m[3][0] * px + X.m[3][1] * py + X.m[3][2] * pz + X.m[3][3];
            if (!(w == 1)) {
                x = x / w;
                y = y / w;
                z = z / w;
            }
            m.p[k] = x;
            m.p[k + 1] = y;
            m.p[k + 2] = z;
        }
        const M4& I = o2w.inv;
        for (size_t k = 0; k + 2 < m.n.size(); k += 3) {  // normalize(transformNormal) (transform.h:185-192)
            const float nx = m.n[k], ny = m.n[k + 1], nz = m.n[k + 2];
            const F3 r = nrm(F3{I.m[0][0] * nx + I.m[1][0] * ny + I.m[2][0] * nz,
                                I.m[0][1] * nx + I.m[1][1] * ny + I.m[2][1] * nz,
                                I.m[0][2] * nx + I.m[1][2] * ny + I.m[2][2] * nz});
            m.n[k] = r.x;
            m.n[k + 1] = r.y;
            m.n[k + 2] = r.z;
        }
        cur_mat.emission[0] = cur_light.x;
        cur_mat.emission[1] = cur_light.y;
        cur_mat.emission[2] = cur_light.z;
        int rc = gpr_scene_add_mesh(scene, m.p.data(), (uint32_t)(m.p.size() / 3), m.n.empty() ? nullptr : m.n.data(),
                                    (uint32_t)(m.n.size() / 3), m.uv.empty() ? nullptr : m.uv.data(),
                                    (uint32_t)(m.uv.size() / 2), m.c.data(), m.ntri, &cur_mat);
        cur_light = F3{0, 0, 0};
        return rc == GPR_OK ? GPR_OK : (err = rc);
    }
    int handle(const XNode& n, Props& mine, Props& parent) {
        const std::string& tag = n.name;
        const std::string name = n.attr("name"), type = n.attr("type"), id = n.attr("id");
        const std::string value = n.attr("value");
        auto floats = [&](int want, F3* out) -> int {
            float v[3] = {0, 0, 0};
            int got = 0;
            if (!parse_floats(value, v, 3, &got) || (got != want && !(want == 3 && got == 1)))
                return bad(GPR_ERR_PARSE, "<" + tag + " name=\"" + name + "\"> bad value '" + value + "'");
            if (got == 1) v[1] = v[2] = v[0];
            *out = F3{v[0], v[1], v[2]};
            return GPR_OK;
        };
        if (tag == "scene" || tag == "mode" || tag == "integrator" || tag == "sampler" || tag == "rfilter" ||
            tag == "medium" || tag == "volume") {
            return GPR_OK;  // accepted, no effect (mitsubaloader.cpp:621-676)
        } else if (tag == "sensor" || tag == "camera") {
            Xform c2w;
            auto it = mine.t.find("toWorld");
            if (it != mine.t.end()) c2w = it->second;
            cam_free = mine.getb("free", false);  // GUI camera: the default eye / direction of Camera
            std::memcpy(cam_c2w, c2w.m.m, sizeof(cam_c2w));
            cam_fov = mine.getf("fov", 30.f);
            cam = true;
        } else if (tag == "film") {
            film_w = (uint32_t)mine.geti("width", 256);
            film_h = (uint32_t)mine.geti("height", 256);
        } else if (tag == "shape") {
            return shape(type, mine);
        } else if (tag == "emitter") {
            if (type == "area") {
                const float sc = mine.getf("scale", 1.f);
                const F3 r = mine.getc("radiance", {1, 1, 1});
                cur_light = F3{r.x * sc, r.y * sc, r.z * sc};
            } else if (type == "constant") {
                const F3 r = mine.getc("radiance", {1, 1, 1});
                const float rgb[3] = {r.x, r.y, r.z};
                gpr_scene_set_env(scene, rgb);
            } else if (type == "envmap") {
                auto it = mine.s.find("filename");
                if (it == mine.s.end()) return bad(GPR_ERR_PARSE, "envmap emitter without filename");
                const std::string f = it->second.size() && it->second[0] == '/' ? it->second : dir + it->second;
                Xform o2w;
                auto xt = mine.t.find("toWorld");
                if (xt != mine.t.end()) o2w = xt->second;
                uint32_t w = 0, h = 0, id = 0;
                std::vector<float> px;
                int rc = read_image(f, &w, &h, &px);
                if (rc == GPR_OK) rc = gpr_scene_add_texture(scene, px.data(), w, h, nullptr, nullptr, &id);
                if (rc == GPR_OK) rc = gpr_scene_set_env_map(scene, id, &o2w.m.m[0][0]);
                if (rc != GPR_OK) return err = rc;
                // Inverse(Transform) swaps m and mInv (transform.h:137-139): the
                // worldToObject matrix is the inverse accumulated with toWorld
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) scene->env_w2o[3 * r + c] = o2w.inv.m[r][c];
            } else {
                return bad(GPR_ERR_UNSUPPORTED, "emitter type '" + type + "'");
            }
        } else if (tag == "bsdf") {
            if (type == "twosided") {
                if (!id.empty()) named[id] = cur_mat;
            } else {
                if (material(type, mine) != GPR_OK) return err;
                if (!id.empty()) named[id] = cur_mat;
            }
        } else if (tag == "ref") {
            auto it = named.find(id);
            if (it == named.end()) return bad(GPR_ERR_PARSE, "<ref id=\"" + id + "\"> names no bsdf");
            cur_mat = it->second;
        } else if (tag == "bool" || tag == "boolean") {
            if (value != "true" && value != "false") return bad(GPR_ERR_PARSE, "boolean '" + value + "'");
            parent.b[name] = value == "true";
        } else if (tag == "integer") {
            char* e = nullptr;
            const long v = std::strtol(value.c_str(), &e, 10);
            if (value.empty() || *e) return bad(GPR_ERR_PARSE, "integer '" + value + "'");
            parent.i[name] = (int)v;
        } else if (tag == "float") {
            char* e = nullptr;
            const float v = std::strtof(value.c_str(), &e);
            if (value.empty() || *e) return bad(GPR_ERR_PARSE, "float '" + value + "'");
            parent.f[name] = v;
        } else if (tag == "string") {
            parent.s[name] = value;
        } else if (tag == "vector") {
            F3 v;
            if (floats(3, &v) != GPR_OK) return err;
            parent.v[name] = v;
        } else if (tag == "rgb" || tag == "color") {
            F3 v;
            if (floats(3, &v) != GPR_OK) return err;
            parent.c[name] = v;
        } else if (tag == "spectrum") {
            return bad(GPR_ERR_UNSUPPORTED, "<spectrum> values");
        } else if (tag == "transform") {
            parent.t[name] = cur_xf;
        } else if (tag == "lookat" || tag == "lookAt") {
            float t[3], o[3], u[3];
            int g1 = 0, g2 = 0, g3 = 0;
            if (!parse_floats(n.attr("target"), t, 3, &g1) || !parse_floats(n.attr("origin"), o, 3, &g2) ||
                !parse_floats(n.attr("up", "0 1 0"), u, 3, &g3) || g1 != 3 || g2 != 3 || g3 != 3)
                return bad(GPR_ERR_PARSE, "malformed <lookat>");
            cur_xf = compose(cur_xf, xf_lookat(F3{t[0], t[1], t[2]}, F3{o[0], o[1], o[2]}, F3{u[0], u[1], u[2]}));
        } else if (tag == "translate" || tag == "scale") {
            F3 v;
            const bool sc = tag == "scale";
            if (value.empty()) {
                const char* d = sc ? "1" : "0";
                v = F3{std::strtof(n.attr("x", d).c_str(), nullptr), std::strtof(n.attr("y", d).c_str(), nullptr),
                       std::strtof(n.attr("z", d).c_str(), nullptr)};
            } else if (floats(3, &v) != GPR_OK) {
                return err;
            }
            cur_xf = compose(cur_xf, sc ? xf_scale(v) : xf_translate(v));
        } else if (tag == "rotate") {
            const F3 ax{std::strtof(n.attr("x", "0").c_str(), nullptr), std::strtof(n.attr("y", "0").c_str(), nullptr),
                        std::strtof(n.attr("z", "0").c_str(), nullptr)};
            cur_xf = compose(cur_xf, xf_rotate(std::strtof(n.attr("angle", "0").c_str(), nullptr), ax));
        } else if (tag == "matrix") {
            float v[16];
            int got = 0;
            if (!parse_floats(value, v, 16, &got) || got != 16) return bad(GPR_ERR_PARSE, "matrix needs 16 values");
            Xform x;
            std::memcpy(x.m.m, v, sizeof(v));
            x.inv = inverse(x.m);
            cur_xf = compose(cur_xf, x);
        } else if (tag == "texture") {
            uint32_t tid = 0;
            if (texture(type, mine, &tid) != GPR_OK) return err;
            parent.tex[name] = tid;
        } else {
            return bad(GPR_ERR_PARSE, "unexpected tag <" + tag + ">");
        }
        return GPR_OK;
    }
    // ParseTag (mitsubaloader.cpp:774-805)
    int walk(const XNode& n, Props& parent) {
        if (n.name == "hide") return GPR_OK;
        if (n.name == "transform") cur_xf = Xform();
        Props mine;
        for (auto& k : n.kids) {
            if (walk(*k, mine) != GPR_OK) return err ? err : GPR_ERR_PARSE;
        }
        return handle(n, mine, parent);
    }
};

int read_file(const std::string& path, std::string* out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(GPR_ERR_IO, "cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    *out = ss.str();
    return GPR_OK;
}

std::string dir_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

std::string ext_of(const std::string& p) {
    const size_t k = p.find_last_of('.');
    std::string e = k == std::string::npos ? std::string() : p.substr(k);
    for (auto& ch : e) ch = (char)std::tolower((unsigned char)ch);
    return e;
}

int load_xml(const std::string& path, gpr_scene* s) {
    std::string text;
    GPR_TRY(read_file(path, &text));
    XmlParser px(text);
    auto root = px.element();
    if (!root) return fail(GPR_ERR_PARSE, path + ": " + px.err);
    Loader L;
    L.dir = dir_of(path);
    L.scene = s;
    Props top;
    if (L.walk(*root, top) != GPR_OK) return L.err ? L.err : GPR_ERR_PARSE;
    if (L.cam && L.cam_free) {
        const float eye[3] = {0.f, 2.f, 7.f}, dir[3] = {0.f, -0.2f, -1.f}, up[3] = {0.f, 1.f, 0.f};
        GPR_TRY(gpr_scene_set_camera_free(s, eye, dir, up, L.cam_fov, L.film_w, L.film_h));
    } else if (L.cam) GPR_TRY(gpr_scene_set_camera(s, L.cam_c2w, L.cam_fov, L.film_w, L.film_h));
    else {
        s->width = L.film_w;
        s->height = L.film_h;
    }
    return GPR_OK;
}

int load_obj_scene(const std::string& path, gpr_scene* s) {
    MeshData m;
    GPR_TRY(load_obj(path, &m));
    gpr_material mat{};
    mat.type = GPR_MAT_DIFFUSE;
    mat.color[0] = mat.color[1] = mat.color[2] = 0.5f;
    GPR_TRY(gpr_scene_add_mesh(s, m.p.data(), (uint32_t)(m.p.size() / 3), m.n.empty() ? nullptr : m.n.data(),
                               (uint32_t)(m.n.size() / 3), m.uv.empty() ? nullptr : m.uv.data(),
                               (uint32_t)(m.uv.size() / 2), m.c.data(), m.ntri, &mat));
    // camera on +z framing the bounds
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t k = 0; k + 2 < m.p.size(); k += 3)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], m.p[k + a]);
            hi[a] = std::max(hi[a], m.p[k + a]);
        }
    if (m.p.empty()) lo[0] = lo[1] = lo[2] = hi[0] = hi[1] = hi[2] = 0;
    const float c[3] = {(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
    const float r = std::max(1e-3f, 0.5f * std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                                      (hi[2] - lo[2]) * (hi[2] - lo[2])));
    const float o[3] = {c[0], c[1], c[2] + 3.f * r}, up[3] = {0, 1, 0};
    return gpr_scene_set_camera_lookat(s, c, o, up, 40.f, 256, 256);
}

// .gprs: little-endian binary, "GPRS" magic, u32 version (1 or 2), then
//   u32 nmeshes, u32 width, u32 height, f32 fov, f32 c2w[16], u32 has_env, f32 env[3]
//   version 2 only:
//     u32 cam_free, f32 eye[3], dir[3], up[3], fov_x, fov_y, yaw, pitch
//     u32 ntex, per texture: u32 width, height, f32 uv_offset[2], uv_scale[2], f32 rgb[3 w h]
//     u32 env_tex, f32 env_w2o[9]
//   per mesh: gpr_material (version 1: without tex[]), u32 np, nn, nuv, ntri,
//             f32 p[3np], n[3nn], uv[2nuv], i32 corners[9ntri]
template <class T>
bool rd(FILE* f, T* v, size_t n = 1) {
    return std::fread(v, sizeof(T), n, f) == n;
}
int load_gprs(const std::string& path, gpr_scene* s) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return fail(GPR_ERR_IO, "cannot open " + path);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    // every array's claimed size is checked against the bytes left in the
    // file before anything is allocated for it
    if (std::fseek(f, 0, SEEK_END) != 0) return fail(GPR_ERR_IO, "cannot seek " + path);
    const long fsize = std::ftell(f);
    if (fsize < 0 || std::fseek(f, 0, SEEK_SET) != 0) return fail(GPR_ERR_IO, "cannot seek " + path);
    auto fits = [&](uint64_t bytes) {
        const long at = std::ftell(f);
        return at >= 0 && at <= fsize && bytes <= (uint64_t)(fsize - at);
    };
    char magic[4];
    uint32_t ver = 0, nm = 0, w = 0, h = 0, has_env = 0;
    float fov = 0, c2w[16], env[3];
    if (!rd(f, magic, 4) || std::memcmp(magic, "GPRS", 4) != 0 || !rd(f, &ver) || (ver != 1 && ver != 2))
        return fail(GPR_ERR_PARSE, path + ": not a version-1/2 .gprs file");
    if (!rd(f, &nm) || !rd(f, &w) || !rd(f, &h) || !rd(f, &fov) || !rd(f, c2w, 16) || !rd(f, &has_env) || !rd(f, env, 3))
        return fail(GPR_ERR_PARSE, path + ": truncated header");
    uint32_t cam_free = 0, env_tex = 0;
    float cam[14], w2o[9];
    if (ver == 2) {
        uint32_t ntex = 0;
        if (!rd(f, &cam_free) || !rd(f, cam, 14) || !rd(f, &ntex)) return fail(GPR_ERR_PARSE, path + ": truncated header");
        if (ntex > (1u << 20)) return fail(GPR_ERR_PARSE, path + ": implausible texture count");
        for (uint32_t k = 0; k < ntex; ++k) {
            uint32_t tw = 0, th = 0, id = 0;
            float uvo[2], uvs[2];
            if (!rd(f, &tw) || !rd(f, &th) || !rd(f, uvo, 2) || !rd(f, uvs, 2))
                return fail(GPR_ERR_PARSE, path + ": truncated texture header");
            if (tw == 0 || th == 0 || (uint64_t)tw * th > (1ull << 28) || !fits(12ull * tw * th))
                return fail(GPR_ERR_PARSE, path + ": implausible texture size");
            std::vector<float> px(3ull * tw * th);
            if (!rd(f, px.data(), px.size())) return fail(GPR_ERR_PARSE, path + ": truncated texels");
            GPR_TRY(gpr_scene_add_texture(s, px.data(), tw, th, uvo, uvs, &id));
        }
        if (!rd(f, &env_tex) || !rd(f, w2o, 9)) return fail(GPR_ERR_PARSE, path + ": truncated header");
        if (env_tex > ntex) return fail(GPR_ERR_PARSE, path + ": environment texture out of range");
    }
    const size_t mat_bytes = ver == 1 ? offsetof(gpr_material, tex) : sizeof(gpr_material);
    for (uint32_t k = 0; k < nm; ++k) {
        gpr_material m{};
        uint32_t np, nn, nuv, nt;
        if (std::fread(&m, 1, mat_bytes, f) != mat_bytes || !rd(f, &np) || !rd(f, &nn) || !rd(f, &nuv) || !rd(f, &nt))
            return fail(GPR_ERR_PARSE, path + ": truncated mesh header");
        if (np > (1u << 30) || nn > (1u << 30) || nuv > (1u << 30) || nt > (1u << 30) ||
            !fits(12ull * np + 12ull * nn + 8ull * nuv + 36ull * nt))
            return fail(GPR_ERR_PARSE, path + ": implausible mesh size");
        std::vector<float> p(3ull * np), n(3ull * nn), uv(2ull * nuv);
        std::vector<int32_t> c(9ull * nt);
        if ((np && !rd(f, p.data(), p.size())) || (nn && !rd(f, n.data(), n.size())) ||
            (nuv && !rd(f, uv.data(), uv.size())) || (nt && !rd(f, c.data(), c.size())))
            return fail(GPR_ERR_PARSE, path + ": truncated mesh data");
        GPR_TRY(gpr_scene_add_mesh(s, p.data(), np, nn ? n.data() : nullptr, nn, nuv ? uv.data() : nullptr, nuv,
                                   c.data(), nt, &m));
    }
    if (has_env) GPR_TRY(gpr_scene_set_env(s, env));
    GPR_TRY(gpr_scene_set_camera(s, c2w, fov, w, h));
    if (env_tex) {
        s->has_env = 1;
        s->env_tex = env_tex;
        std::memcpy(s->env_w2o, w2o, sizeof(w2o));
    }
    if (cam_free) {
        s->cam_free = true;
        std::memcpy(s->eye, cam, 12);
        std::memcpy(s->dir, cam + 3, 12);
        std::memcpy(s->up, cam + 6, 12);
        s->fovx = cam[9];
        s->fovy = cam[10];
        s->yaw = cam[11];
        s->pitch = cam[12];
    }
    return GPR_OK;
}

}  // namespace
}  // namespace gpr

using namespace gpr;

extern "C" int gpr_scene_load(const char* path, gpr_scene** out) {
    if (!path || !out) return fail(GPR_ERR_INVALID, "gpr_scene_load: null argument");
    gpr_scene* s = nullptr;
    GPR_TRY(gpr_scene_create(&s));
    const int rc = guarded("gpr_scene_load", [&] {
        const std::string p(path), e = ext_of(p);
        int r;
        if (e == ".xml") r = load_xml(p, s);
        else if (e == ".obj") r = load_obj_scene(p, s);
        else if (e == ".gprs") r = load_gprs(p, s);
        else r = fail(GPR_ERR_UNSUPPORTED, "unknown scene format '" + e + "' (expected .xml, .obj or .gprs)");
        return r == GPR_OK ? gpr_scene_commit(s) : r;
    });
    if (rc != GPR_OK) {
        gpr_scene_free(s);
        return rc;
    }
    *out = s;
    return GPR_OK;
}

// Parse and assemble without committing (no device work): lets the host path
// be checked on machines without a GPU.
extern "C" int gpr_scene_parse(const char* path, gpr_scene** out) {
    if (!path || !out) return fail(GPR_ERR_INVALID, "gpr_scene_parse: null argument");
    gpr_scene* s = nullptr;
    GPR_TRY(gpr_scene_create(&s));
    const int rc = guarded("gpr_scene_parse", [&] {
        const std::string p(path), e = ext_of(p);
        int r;
        if (e == ".xml") r = load_xml(p, s);
        else if (e == ".obj") r = load_obj_scene(p, s);
        else if (e == ".gprs") r = load_gprs(p, s);
        else r = fail(GPR_ERR_UNSUPPORTED, "unknown scene format '" + e + "' (expected .xml, .obj or .gprs)");
        return r == GPR_OK ? gpr_scene_flatten(s) : r;
    });
    if (rc != GPR_OK) {
        gpr_scene_free(s);
        return rc;
    }
    *out = s;
    return GPR_OK;
}

extern "C" int gpr_scene_save(const gpr_scene* s, const char* path) {
    if (!s || !path) return fail(GPR_ERR_INVALID, "gpr_scene_save: null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(GPR_ERR_IO, std::string("cannot open ") + path);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    const uint32_t ver = 2, nm = (uint32_t)s->meshes.size();
    auto put = [&](const void* p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes; };
    bool ok = put("GPRS", 4) && put(&ver, 4) && put(&nm, 4) && put(&s->width, 4) && put(&s->height, 4) &&
              put(&s->fov, 4) && put(s->c2w, 64) && put(&s->has_env, 4) && put(s->env, 12);
    const uint32_t cam_free = s->cam_free ? 1u : 0u, ntex = (uint32_t)s->textures.size();
    const float cam[14] = {s->eye[0], s->eye[1], s->eye[2], s->dir[0], s->dir[1], s->dir[2], s->up[0], s->up[1], s->up[2],
                           s->fovx, s->fovy, s->yaw, s->pitch, 0.f};
    ok = ok && put(&cam_free, 4) && put(cam, sizeof(cam)) && put(&ntex, 4);
    for (const gpr_texture& t : s->textures) {
        ok = ok && put(&t.width, 4) && put(&t.height, 4) && put(t.uv_offset, 8) && put(t.uv_scale, 8);
        std::vector<float> px(3ull * t.width * t.height);
        for (size_t i = 0; i < (size_t)t.width * t.height; ++i)
            for (int c = 0; c < 3; ++c) px[3 * i + c] = s->texels[4 * (t.offset + i) + c];
        ok = ok && put(px.data(), 4 * px.size());
    }
    const uint32_t env_tex = s->has_env ? s->env_tex : 0u;
    ok = ok && put(&env_tex, 4) && put(s->env_w2o, 36);
    for (const Mesh& m : s->meshes) {
        const uint32_t np = (uint32_t)(m.p.size() / 3), nn = (uint32_t)(m.n.size() / 3), nuv = (uint32_t)(m.uv.size() / 2);
        ok = ok && put(&m.mat, sizeof(gpr_material)) && put(&np, 4) && put(&nn, 4) && put(&nuv, 4) && put(&m.ntri, 4);
        ok = ok && put(m.p.data(), 4 * m.p.size()) && put(m.n.data(), 4 * m.n.size()) && put(m.uv.data(), 4 * m.uv.size()) &&
             put(m.corners.data(), 4 * m.corners.size());
    }
    if (!ok) return fail(GPR_ERR_IO, std::string("short write to ") + path);
    return GPR_OK;
}
