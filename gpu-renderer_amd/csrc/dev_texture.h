// dev_texture.h -- texture and environment lookups of the shading kernels.
//
// TextureView::evalSpectrum / evalFloat (src/renderer/texture.h:37-52):
// nearest texel of uv * scale + offset, indices wrapped as the reference's
// int % uint32 does.  EnvironmentLightView::Le (src/renderer/environment.h:
// 16-28): latitude-longitude lookup of the direction from the world-box
// centre.  acos / atan2 are evaluated in double with only IEEE add, mul, div
// and a Newton-refined sqrt, so the chosen texel is the same on host and
// device (the float result is within one ulp of acosf / atan2f).
#pragma once

#include "dev_math.h"
#include "gpr_internal.h"

namespace gpr {

// ---------------------------------------------------- deterministic trig --
GPR_HD double det_sqrt(double x) {
    double y = (double)sqrtf((float)x);
    if (!(y > 0.0)) return 0.0;
    y = 0.5 * (y + x / y);
    y = 0.5 * (y + x / y);
    return y;
}
// atan(t), t in [0, 1]: t = c + reduction around c = k/8, series to r^17 (|r| <= 1/16)
GPR_HD double det_atan01(double t) {
    const int k = (int)(t * 8.0 + 0.5);
    double c0;
    switch (k) {
        case 0: c0 = 0.0; break;
        case 1: c0 = 0.12435499454676144; break;
        case 2: c0 = 0.24497866312686414; break;
        case 3: c0 = 0.35877067027057225; break;
        case 4: c0 = 0.4636476090008061; break;
        case 5: c0 = 0.5585993153435624; break;
        case 6: c0 = 0.6435011087932844; break;
        case 7: c0 = 0.7188299996216245; break;
        default: c0 = 0.7853981633974483; break;
    }
    const double c = (double)k * 0.125;
    const double r = (t - c) / (1.0 + t * c);
    const double r2 = r * r;
    double p = 1.0 / 17.0;
    p = -1.0 / 15.0 + r2 * p;
    p = 1.0 / 13.0 + r2 * p;
    p = -1.0 / 11.0 + r2 * p;
    p = 1.0 / 9.0 + r2 * p;
    p = -1.0 / 7.0 + r2 * p;
    p = 1.0 / 5.0 + r2 * p;
    p = -1.0 / 3.0 + r2 * p;
    p = 1.0 + r2 * p;
    return c0 + r * p;
}
// atan2 with the IEEE signed-zero conventions
GPR_HD double det_atan2(double y, double x) {
    if (x != x || y != y) return x + y;
    const double pi = 3.141592653589793, pi_2 = 1.5707963267948966;
    const double ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
    double a;
    if (ay == 0.0) {
        a = __builtin_signbit(x) ? pi : 0.0;
    } else if (ax == 0.0) {
        a = pi_2;
    } else {
        a = ay <= ax ? det_atan01(ay / ax) : pi_2 - det_atan01(ax / ay);
        if (x < 0) a = pi - a;
    }
    return __builtin_signbit(y) ? -a : a;
}
GPR_HD double det_acos(double z) { return det_atan2(det_sqrt((1.0 - z) * (1.0 + z)), z); }

// float -> int as the conversion instruction does it: truncation, saturation
// at the int range, NaN -> 0 (the C++ cast leaves these cases undefined)
GPR_HD int to_int_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.f) return 2147483647;
    if (x < -2147483648.f) return (int)0x80000000u;
    return (int)x;
}

// ------------------------------------------------------------- textures --
// TextureView lookup (texture.h:37-52): u2 = int(uv'.x * w) % w with the
// int converted to uint32 by the modulo, idx = v2 * w + u2.
// Texture record k as two global 16-byte loads (the table starts on a 256-byte
// boundary and its records are 32 bytes)
GPR_D gpr_texture load_texture(const SceneView& sv, uint32_t k) {
    static_assert(sizeof(gpr_texture) == 32, "texture record is two 16-byte words");
    const float4* p = reinterpret_cast<const float4*>(sv.tex + k);
    const float4 a = ldg(p), b = ldg(p + 1);
    gpr_texture t;
    t.offset = __float_as_uint(a.x);
    t.width = __float_as_uint(a.y);
    t.height = __float_as_uint(a.z);
    t.pad = __float_as_uint(a.w);
    t.uv_offset[0] = b.x;
    t.uv_offset[1] = b.y;
    t.uv_scale[0] = b.z;
    t.uv_scale[1] = b.w;
    return t;
}
GPR_D float4 tex_eval(const SceneView& sv, uint32_t handle, float u, float v) {
    const gpr_texture t = load_texture(sv, handle - 1);
    const float nu = u * t.uv_scale[0] + t.uv_offset[0];
    const float nv = v * t.uv_scale[1] + t.uv_offset[1];
    const uint32_t u2 = (uint32_t)to_int_sat(nu * (float)t.width) % t.width;
    const uint32_t v2 = (uint32_t)to_int_sat(nv * (float)t.height) % t.height;
    return ldg(sv.texels + (t.offset + v2 * t.width + u2));
}
GPR_D bool has_textures(const gpr_material& m) {
    uint32_t any = 0;
#pragma unroll
    for (int k = 0; k < GPR_TEX_SLOTS; ++k) any |= m.tex[k];
    return any != 0;
}
// Replace the textured fields of m by their lookups at the hit (the BxDF
// factories, bsdfs/bxdf.h:160-218).  uv: Triangle::setupDifferential
// (triangle.h:126-134) -- interpolated corner uvs, else the barycentrics.
GPR_D void resolve_textures(const SceneView& sv, gpr_material& m, uint32_t t, uint32_t flags, float bu,
                            float bv) {
    float u = bu, v = bv;
    if (flags & 2u) {
        // the corner uvs of the triangle's shading record
        const float4* r = reinterpret_cast<const float4*>(sv.shade + t);
        const float4 q = ldg(r + 5), q2 = ldg(r + 6);
        const float w0 = 1 - bu - bv;
        u = q.x * w0 + q.z * bu + q2.x * bv;
        v = q.y * w0 + q.w * bu + q2.y * bv;
    }
#pragma unroll
    for (int k = 0; k < GPR_TEX_SLOTS; ++k) {
        const uint32_t h = m.tex[k];
        if (h == 0) continue;
        const float4 c = tex_eval(sv, h, u, v);
        float* dst = k == GPR_TEX_COLOR ? m.color
                   : k == GPR_TEX_COLOR1 ? m.color1
                   : k == GPR_TEX_ETA_I ? m.eta_i
                   : k == GPR_TEX_ETA_T ? m.eta_t
                   : k == GPR_TEX_K ? m.k
                   : k == GPR_TEX_IOR ? &m.ior
                   : k == GPR_TEX_ALPHA_X ? &m.alpha_x : &m.alpha_y;
        dst[0] = c.x;
        if (k == GPR_TEX_COLOR || k == GPR_TEX_COLOR1 || k >= GPR_TEX_ETA_I) {
            dst[1] = c.y;
            dst[2] = c.z;
        }
    }
}

// ---------------------------------------------------------- environment --
// EnvironmentLightView::Le (environment.h:16-28)
GPR_D c3 env_le(const SceneView& sv, v3 o, v3 d) {
    v3 w = norm3(o + d * 100000.f - load3(sv.env_center));
    const float* m = sv.env_w2o;
    w = norm3(V(m[0] * w.x + m[1] * w.y + m[2] * w.z, m[3] * w.x + m[4] * w.y + m[5] * w.z,
                m[6] * w.x + m[7] * w.y + m[8] * w.z));
    const float z = fmaxf(-1.f, fminf(w.z, 1.f));  // clamp of helper_math (NaN -> 1)
    const float theta = (float)det_acos((double)z);
    float phi = (float)det_atan2((double)w.y, (double)w.x);
    if (phi < 0) phi = phi + 2 * 3.14159265358979323846f;
    const float u = phi * 0.15915494309189533577f, v = theta * 0.31830988618379067154f;
    const gpr_texture t = load_texture(sv, sv.env_tex - 1);
    const int iw = (int)t.width, ih = (int)t.height;
    const int u2 = min(iw - 1, to_int_sat(u * (float)iw)), v2 = min(ih - 1, to_int_sat(v * (float)ih));
    const float4 c = ldg(sv.texels + (t.offset + (uint32_t)(v2 * iw + u2)));
    return C(c.x, c.y, c.z);
}

}  // namespace gpr
