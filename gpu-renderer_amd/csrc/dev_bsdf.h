// dev_bsdf.h -- device BSDFs for the shading kernels.
//
// Lambert (bsdfs/lambert.h), FresnelSpecular smooth dielectric / mirror
// (bsdfs/specular.h), GGX microfacet reflection with conductor or dielectric
// Fresnel (bsdfs/microfacet.h, bsdfs/optical.h), rough plastic (Lambert + GGX,
// bsdf.h:33-40) and a GGX rough dielectric (Walter et al. 2007, a build
// addition: the reference maps rough dielectrics to the smooth case).
// A BSDF is evaluated directly from the material record; the shading frame and
// the f*|cos|, pdf and sampling conventions follow bsdf.h:9-190.
#pragma once

#include "dev_math.h"
#include "gpr_internal.h"

namespace gpr {

// Shading frame (sampling.h:124-163)
struct Frame {
    v3 s, t, n;
};
GPR_D Frame make_frame(v3 n) {
    Frame f;
    f.n = n;
    if (fabsf(n.x) > fabsf(n.y)) {
        const float inv = 1.f / sqrtf(n.x * n.x + n.z * n.z);
        f.s = V(n.z * inv, 0, -n.x * inv);
    } else {
        const float inv = 1.f / sqrtf(n.y * n.y + n.z * n.z);
        f.s = V(0, n.z * inv, -n.y * inv);
    }
    f.t = cross3(f.s, n);
    return f;
}
GPR_D v3 to_world(const Frame& f, v3 v) { return norm3(f.s * v.x + f.t * v.y + f.n * v.z); }
GPR_D v3 to_local(const Frame& f, v3 v) { return norm3(V(dot3(v, f.s), dot3(v, f.t), dot3(v, f.n))); }

// spherical helpers of the local frame (sampling.h:144-181)
GPR_D float cos2t(v3 w) { return w.z * w.z; }
GPR_D float sin2t(v3 w) { return fmax_h(0.f, 1.f - cos2t(w)); }
GPR_D float sint(v3 w) { return sqrtf(sin2t(w)); }
GPR_D float clampr(float a, float l, float r) { return a > r ? r : (a < l ? l : a); }
GPR_D float cosp(v3 w) {
    const float st = sint(w);
    return st == 0 ? 1.f : clampr(w.x / st, -1, 1);
}
GPR_D float sinp(v3 w) {
    const float st = sint(w);
    return st == 0 ? 0.f : clampr(w.y / st, -1, 1);
}
GPR_D bool same_side(v3 a, v3 b) { return a.z * b.z > 0; }

// Fresnel terms (bsdfs/optical.h:43-100)
GPR_D float fresnel_diel(float ci, float ei, float et) {
    ci = fmax_h(-1.f, fmin_h(ci, 1.f));
    if (!(ci > 0.f)) {
        const float t = ei;
        ei = et;
        et = t;
        ci = fabsf(ci);
    }
    const float si = sqrtf(fmax_h(0.f, 1 - ci * ci));
    const float st = ei / et * si;
    if (st >= 1) return 1.f;
    const float ct = sqrtf(fmax_h(0.f, 1 - st * st));
    const float rpa = ((et * ci) - (ei * ct)) / ((et * ci) + (ei * ct));
    const float rpe = ((ei * ci) - (et * ct)) / ((ei * ci) + (et * ct));
    return (rpa * rpa + rpe * rpe) * 0.5f;
}
GPR_D c3 fresnel_cond(float ci, c3 etai, c3 etat, c3 k) {
    ci = fmax_h(-1.f, fmin_h(ci, 1.f));
    const c3 eta = etat / etai, etak = k / etai;
    const float c2 = ci * ci, s2 = 1.f - c2;
    const c3 eta2 = eta * eta, etak2 = etak * etak;
    const c3 t0 = eta2 - etak2 - C(s2);
    const c3 a2b2 = csqrt(t0 * t0 + eta2 * etak2 * 4.f);
    const c3 t1 = a2b2 + C(c2);
    const c3 a = csqrt((a2b2 + t0) * 0.5f);
    const c3 t2 = a * 2.f * ci;
    const c3 rs = (t1 - t2) / (t1 + t2);
    const c3 t3 = a2b2 * c2 + C(s2 * s2);
    const c3 t4 = t2 * s2;
    const c3 rp = rs * (t3 - t4) / (t3 + t4);
    return (rp + rs) * 0.5f;
}

// Trowbridge-Reitz (bsdfs/microfacet.h:5-156)
GPR_D float ggx_lambda(v3 w, float ax, float ay) {
    const float at = fabsf(sint(w) / w.z);
    if (isinf(at)) return 0.f;
    const float cp = cosp(w), sp = sinp(w);
    const float alpha = sqrtf(cp * cp * ax * ax + sp * sp * ay * ay);
    const float a2t2 = (alpha * at) * (alpha * at);
    return (-1 + sqrtf(1.f + a2t2)) / 2;
}
GPR_D float ggx_g(v3 wo, v3 wi, float ax, float ay) {
    return 1 / (1 + ggx_lambda(wo, ax, ay) + ggx_lambda(wi, ax, ay));
}
GPR_D float ggx_d(v3 wh, float ax, float ay) {
    const float t2 = sin2t(wh) / cos2t(wh);
    if (isinf(t2)) return 0.f;
    const float c4 = cos2t(wh) * cos2t(wh);
    const float cp = cosp(wh), sp = sinp(wh);
    const float e = (cp * cp / (ax * ax) + sp * sp / (ay * ay)) * t2;
    return 1 / (kPi * ax * ay * c4 * (1 + e) * (1 + e));
}
GPR_D void ggx_sample11(float ct, float U1, float U2, float& sx, float& sy) {
    if (ct > .9999) {
        const float r = sqrtf(U1 / (1 - U1));
        const float phi = (float)(6.28318530718 * (double)U2);
        sx = r * cos_f(phi);
        sy = r * sin_f(phi);
        return;
    }
    const float st = sqrtf(fmax_h(0.f, 1.f - ct * ct));
    const float tt = st / ct;
    const float a = 1 / tt;
    const float G1 = 2 / (1 + sqrtf(1.f + 1.f / (a * a)));
    const float A = 2 * U1 / G1 - 1;
    float tmp = 1.f / (A * A - 1.f);
    if (tmp > 1e10) tmp = 1e10f;
    const float B = tt;
    const float D = sqrtf(fmax_h(float(B * B * tmp * tmp - (A * A - B * B) * tmp), 0.f));
    const float x1 = B * tmp - D, x2 = B * tmp + D;
    sx = (A < 0 || x2 > 1.f / tt) ? x1 : x2;
    float S;
    if (U2 > 0.5f) {
        S = 1.f;
        U2 = 2.f * (U2 - .5f);
    } else {
        S = -1.f;
        U2 = 2.f * (.5f - U2);
    }
    const float z = (U2 * (U2 * (U2 * 0.27385f - 0.73369f) + 0.46341f)) /
                    (U2 * (U2 * (U2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    sy = S * z * sqrtf(1.f + sx * sx);
}
// visible-normal sample of wh (TrowbridgeReitzSample + Sample_wh)
GPR_D v3 ggx_sample_wh(v3 wo, float u0, float u1, float ax, float ay) {
    const bool flip = wo.z < 0;
    const v3 w = flip ? -wo : wo;
    const v3 ws = norm3(V(ax * w.x, ay * w.y, w.z));
    float sx, sy;
    ggx_sample11(ws.z, u0, u1, sx, sy);
    const float cp = cosp(ws), sp = sinp(ws);
    const float tmp = cp * sx - sp * sy;
    sy = sp * sx + cp * sy;
    sx = tmp;
    sx = ax * sx;
    sy = ay * sy;
    const v3 wh = norm3(V(-sx, -sy, 1.f));
    return flip ? -wh : wh;
}
GPR_D v3 reflect_v(v3 wo, v3 n) { return -wo + n * 2.f * dot3(wo, n); }
GPR_D bool refract_v(v3 wi, v3 n, float eta, v3& wt) {
    const float ci = dot3(n, wi);
    const float s2i = fmax_h(0.f, 1 - ci * ci);
    const float s2t = eta * eta * s2i;
    if (s2t >= 1) return false;
    const float ct = sqrtf(1 - s2t);
    wt = -wi * eta + n * (eta * ci - ct);
    return true;
}

// --------------------------------------------------------------- lobes ---
GPR_D c3 lambert_f(c3 col, v3 wo, v3 wi) {
    if (!same_side(wo, wi)) return C(0.f);
    return col * kInvPi * fabsf(wi.z);
}
GPR_D float lambert_pdf(v3 wo, v3 wi) {
    if (!same_side(wo, wi)) return 0.f;
    return fabsf(wi.z) * kInvPi;
}
GPR_D c3 fresnel_term(const gpr_material& m, bool conductor, float c) {
    if (conductor) return fresnel_cond(fabsf(c), loadc(m.eta_i), loadc(m.eta_t), loadc(m.k));
    return C(fresnel_diel(c, m.eta_i[0], m.eta_t[0]));
}
GPR_D c3 micro_f(c3 col, const gpr_material& m, bool conductor, v3 wo, v3 wi) {
    const float co = fabsf(wo.z), ci = fabsf(wi.z);
    v3 wh = wi + wo;
    if (ci == 0 || co == 0) return C(0.f);
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return C(0.f);
    wh = norm3(wh);
    const c3 F = fresnel_term(m, conductor, dot3(wi, wh));
    return col * ggx_d(wh, m.alpha_x, m.alpha_y) * ggx_g(wo, wi, m.alpha_x, m.alpha_y) * F / (4 * co);
}
// pdf D|cos h| / (4 wo.wh) (microfacet.h:199-200)
GPR_D float micro_pdf(const gpr_material& m, v3 wo, v3 wi) {
    if (!same_side(wo, wi)) return 0.f;
    const v3 wh = norm3(wo + wi);
    return ggx_d(wh, m.alpha_x, m.alpha_y) * fabsf(wh.z) / (4 * dot3(wo, wh));
}
// rough dielectric: f*|cos i| and pdf for a given pair (isotropic alpha_x)
GPR_D void rdiel_eval(const gpr_material& m, v3 wo, v3 wi, c3& f, float& pdf) {
    f = C(0.f);
    pdf = 0.f;
    const float co = wo.z, ci = wi.z;
    if (co == 0 || ci == 0) return;
    const bool refl = co * ci > 0;
    const float eta = co > 0 ? m.ior : 1.f / m.ior;
    v3 wh;
    if (refl) {
        wh = wi + wo;
        if (wh.x == 0 && wh.y == 0 && wh.z == 0) return;
        wh = norm3(wh);
    } else {
        wh = norm3(wo + wi * eta);
    }
    if (wh.z < 0) wh = -wh;
    const float dwo = dot3(wo, wh), dwi = dot3(wi, wh);
    const float F = fresnel_diel(dwo, 1.f, m.ior);
    const float a = m.alpha_x;
    const float D = ggx_d(wh, a, a);
    const float G = ggx_g(wo, wi, a, a);
    const float pwh = D * fabsf(wh.z);
    const c3 col = loadc(m.color);
    if (refl) {
        if (dwo * co <= 0 || dwi * ci <= 0) return;
        f = col * (D * G * F / (4 * fabsf(co)));
        pdf = F * pwh / (4 * fabsf(dwo));
    } else {
        if (dwo * dwi > 0) return;
        if (dwo * co <= 0 || dwi * ci <= 0) return;
        const float sd = dwo + eta * dwi;
        const float jac = fabsf(dwi) / (sd * sd);
        f = col * ((1 - F) * D * G * fabsf(dwo) * jac / fabsf(co));
        pdf = (1 - F) * pwh * eta * eta * jac;
    }
}

// BSDF::eval + BSDF::pdf (bsdf.h:58-105) on local directions
GPR_D void bsdf_eval_local(const gpr_material& m, v3 lo, v3 li, c3& f, float& pdf) {
    f = C(0.f);
    pdf = 0.f;
    if (lo.z == 0) return;
    const bool refl = lo.z * li.z > 0;
    switch (m.type) {
        case GPR_MAT_DIFFUSE:
            f = f + (refl ? lambert_f(loadc(m.color), lo, li) : C(0.f));
            pdf = pdf + lambert_pdf(lo, li);
            break;
        case GPR_MAT_SPECULAR:
            f = f + C(0.f);
            pdf = pdf + 0.f;
            break;
        case GPR_MAT_ROUGH_CONDUCTOR:
            f = f + (refl ? micro_f(loadc(m.color), m, true, lo, li) : C(0.f));
            pdf = pdf + micro_pdf(m, lo, li);
            break;
        case GPR_MAT_ROUGH_DIELECTRIC: {
            c3 fi;
            float pi;
            rdiel_eval(m, lo, li, fi, pi);
            f = f + fi;
            pdf = pdf + pi;
            break;
        }
        default: {  // rough plastic: the two lobe pdfs are summed (bsdf.h:102)
            f = f + (refl ? lambert_f(loadc(m.color), lo, li) : C(0.f));
            f = f + (refl ? micro_f(loadc(m.color1), m, false, lo, li) : C(0.f));
            pdf = pdf + lambert_pdf(lo, li);
            pdf = pdf + micro_pdf(m, lo, li);
            break;
        }
    }
}

GPR_D bool bsdf_is_delta(const gpr_material& m) { return m.type == GPR_MAT_SPECULAR; }

struct BSample {
    v3 wi;
    c3 f;
    float pdf;
    bool specular;
};

// BSDF::sample (bsdf.h:112-167); u2 is the lobe uniform of the rough dielectric.
GPR_D BSample bsdf_sample(const gpr_material& m, const Frame& fr, v3 wo, float u0, float u1, float u2) {
    BSample s;
    s.f = C(0.f);
    s.pdf = 0.f;
    s.wi = V(0, 0, 0);
    s.specular = false;
    const v3 lo = to_local(fr, wo);
    const uint32_t nlobes = m.type == GPR_MAT_ROUGH_PLASTIC ? 2u : 1u;
    uint32_t idx = (uint32_t)floorf(u0 * (float)nlobes);
    if (idx > nlobes - 1) idx = nlobes - 1;
    const float v0 = u0 * (float)nlobes - (float)idx;
    if (lo.z == 0) return s;
    v3 li = V(0, 0, 0);
    const int lobe = m.type == GPR_MAT_ROUGH_PLASTIC ? (idx == 0 ? 0 : 2) : (int)m.type;
    // lobe ids: 0 lambert, 1 fresnel-specular, 2 microfacet, 3 rough dielectric
    const int kind = lobe == GPR_MAT_DIFFUSE ? 0 : lobe == GPR_MAT_SPECULAR ? 1 : lobe == GPR_MAT_ROUGH_DIELECTRIC ? 3 : 2;
    if (kind == 0) {
        const float a = sqrtf(v0);
        const float b = kPi * 2 * u1;
        const float z = sqrtf(fmax_h(0.f, 1 - v0));
        li = V(a * cos_f(b), a * sin_f(b), z);
        if (lo.z < 0) li.z *= -1;
        s.f = loadc(m.color) * kInvPi * fabsf(li.z);
        s.pdf = fabsf(li.z) * kInvPi;
    } else if (kind == 1) {
        const float F = fresnel_diel(lo.z, 1, m.ior);
        if (v0 < F) {
            li = V(-lo.x, -lo.y, lo.z);
            s.pdf = F;
            s.f = loadc(m.color) * F;
        } else {
            const bool entering = lo.z > 0;
            const float ei = entering ? 1 : m.ior;
            const float et = entering ? m.ior : 1;
            v3 n = V(0, 0, 1);
            if (lo.z < 0) n = -n;
            if (!refract_v(lo, n, ei / et, li)) {
                s.f = C(0.f);
                s.specular = true;
                return s;
            }
            c3 ft = loadc(m.color) * (1 - F);
            ft = ft * ((ei * ei) / (et * et));
            s.pdf = 1 - F;
            s.f = ft;
        }
    } else if (kind == 2) {
        const bool conductor = m.type == GPR_MAT_ROUGH_CONDUCTOR;
        const c3 col = conductor ? loadc(m.color) : loadc(m.color1);
        const v3 wh = ggx_sample_wh(lo, v0, u1, m.alpha_x, m.alpha_y);
        li = reflect_v(lo, wh);
        if (same_side(lo, li)) {
            s.pdf = ggx_d(wh, m.alpha_x, m.alpha_y) * fabsf(wh.z) / (4 * dot3(lo, wh));
            s.f = micro_f(col, m, conductor, lo, li);
        }
    } else {
        const float a = m.alpha_x;
        const float t2 = a * a * v0 / (1 - v0);
        const float ct = 1 / sqrtf(1 + t2);
        const float st = sqrtf(fmax_h(0.f, 1 - ct * ct));
        const float phi = kPi * 2 * u1;
        const v3 wh = V(st * cos_f(phi), st * sin_f(phi), ct);
        const float dwo = dot3(lo, wh);
        // only microfacets facing wo scatter it; a direction on the wrong side
        // for its lobe has no counterpart in eval/pdf and is dropped
        if (dwo * lo.z > 0) {
            const float F = fresnel_diel(dwo, 1.f, m.ior);
            bool ok = true;
            if (u2 < F) {
                li = reflect_v(lo, wh);
                ok = li.z * lo.z > 0;
            } else {
                const float eta = lo.z > 0 ? m.ior : 1.f / m.ior;
                const v3 n = dwo < 0 ? -wh : wh;
                ok = refract_v(lo, n, 1.f / eta, li) && li.z * lo.z < 0;
            }
            if (ok) rdiel_eval(m, lo, li, s.f, s.pdf);
        }
    }
    s.specular = kind == 1;
    if (s.pdf == 0.f) return s;
    s.wi = to_world(fr, li);
    if (m.type == GPR_MAT_ROUGH_PLASTIC) {
        // add the other lobe's pdf and re-evaluate f over both lobes (bsdf.h:137-166)
        const bool refl = lo.z * li.z > 0;
        if (idx == 0) s.pdf += micro_pdf(m, lo, li);
        else s.pdf += lambert_pdf(lo, li);
        c3 f = C(0.f);
        f = f + (refl ? lambert_f(loadc(m.color), lo, li) : C(0.f));
        f = f + (refl ? micro_f(loadc(m.color1), m, false, lo, li) : C(0.f));
        s.f = f;
    } else if (kind != 1) {
        // single non-delta lobe: f re-evaluated through the lobe mask
        const bool refl = lo.z * li.z > 0;
        c3 f = C(0.f);
        if (kind == 0) f = f + (refl ? lambert_f(loadc(m.color), lo, li) : C(0.f));
        else if (kind == 2) f = f + (refl ? micro_f(loadc(m.color), m, true, lo, li) : C(0.f));
        else {
            c3 fi;
            float pi;
            rdiel_eval(m, lo, li, fi, pi);
            f = f + fi;
        }
        s.f = f;
    }
    return s;
}

}  // namespace gpr
