"""Pins the oracle's BSDFs (bsdf.h, bsdfs/) with closed forms, normalisation,
reciprocity and sampling/pdf consistency -- no GPU."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_py as O

N_UP = np.array([0, 0, 1], np.float32)


def mat(type, color=(1, 1, 1), ior=1.5, alpha=0.1, alpha_y=None, eta_t=(0.200438, 0.924033, 1.10221),
        k=(3.91295, 2.45285, 2.14219), color1=(1, 1, 1), eta_i=(1, 1, 1)):
    m = np.zeros(1, O.MATERIAL_DTYPE)
    m["type"], m["color"], m["color1"], m["ior"] = type, color, color1, ior
    m["alpha_x"], m["alpha_y"] = alpha, alpha if alpha_y is None else alpha_y
    m["eta_i"], m["eta_t"], m["k"] = eta_i, eta_t, k
    return m


def ev(m, wo, wi):
    f = np.zeros(3, np.float32)
    p = C.c_float()
    O.lib.orc_bsdf_eval(O.ptr(m), O.ptr(N_UP), O.ptr(np.asarray(wo, np.float32)),
                        O.ptr(np.asarray(wi, np.float32)), O.ptr(f), C.byref(p))
    return f.astype(np.float64), p.value


def sm(m, wo, u0, u1, u2=0.5):
    wi = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    p = C.c_float()
    sp = C.c_int()
    O.lib.orc_bsdf_sample(O.ptr(m), O.ptr(N_UP), O.ptr(np.asarray(wo, np.float32)), u0, u1, u2,
                          O.ptr(wi), O.ptr(f), C.byref(p), C.byref(sp))
    return wi.astype(np.float64), f.astype(np.float64), p.value, sp.value


def unit(theta, phi=0.3):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)], np.float32)


def test_fresnel_dielectric_closed_forms():
    assert abs(O.lib.orc_fr_dielectric(1.0, 1.0, 1.5) - 0.04) < 1e-7
    assert O.lib.orc_fr_dielectric(-0.3, 1.0, 1.5) == 1.0       # total internal reflection
    assert abs(O.lib.orc_fr_dielectric(-1.0, 1.0, 1.5) - 0.04) < 1e-7  # leaving at normal incidence
    # Brewster angle: parallel component vanishes, F = Rperp^2 / 2
    tb = math.atan(1.5)
    ci = math.cos(tb)
    ct = math.sqrt(1 - (math.sin(tb) / 1.5) ** 2)
    rperp = (ci - 1.5 * ct) / (ci + 1.5 * ct)
    assert abs(O.lib.orc_fr_dielectric(ci, 1.0, 1.5) - rperp ** 2 / 2) < 1e-6


def test_fresnel_conductor_normal_incidence():
    eta = np.array([0.200438, 0.924033, 1.10221], np.float32)
    k = np.array([3.91295, 2.45285, 2.14219], np.float32)
    out = np.zeros(3, np.float32)
    O.lib.orc_fr_conductor(1.0, O.ptr(np.ones(3, np.float32)), O.ptr(eta), O.ptr(k), O.ptr(out))
    want = ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)
    assert np.allclose(out, want, rtol=1e-5)
    O.lib.orc_fr_conductor(1e-6, O.ptr(np.ones(3, np.float32)), O.ptr(eta), O.ptr(k), O.ptr(out))
    assert np.all(out > 0.99)


@pytest.mark.parametrize("alpha", [0.1, 0.25, 0.5])
def test_ggx_d_projected_area_normalisation(alpha):
    # integral of D(wh) cos(theta_h) over the hemisphere is 1 (the last 1e-4 rad
    # before grazing are left out: there cos^4 underflows in float and D is inf,
    # as in the reference, while the true contribution is ~alpha^2 * 1e-8)
    t = np.concatenate([np.linspace(0, 1e-3, 2000, endpoint=False), np.geomspace(1e-3, math.pi / 2 - 1e-4, 6000)])
    vals = np.array([O.lib.orc_ggx_d(O.ptr(unit(x)), alpha, alpha) for x in t])
    integ = 2 * math.pi * np.trapezoid(vals * np.cos(t) * np.sin(t), t)
    assert abs(integ - 1) < 2e-3


def test_ggx_d_anisotropic_normalisation():
    ax, ay = 0.2, 0.45
    th = np.linspace(0, math.pi / 2 - 1e-4, 700)
    ph = np.linspace(0, 2 * math.pi, 360, endpoint=False)
    tot = 0.0
    for p in ph:
        v = np.array([O.lib.orc_ggx_d(O.ptr(unit(x, p)), ax, ay) for x in th])
        tot += np.trapezoid(v * np.cos(th) * np.sin(th), th) * (2 * math.pi / len(ph))
    assert abs(tot - 1) < 5e-3


def test_smith_g_range():
    rng = np.random.default_rng(1)
    for _ in range(300):
        a = float(rng.uniform(0.01, 1))
        wo = unit(rng.uniform(0, 1.55), rng.uniform(0, 6.28))
        wi = unit(rng.uniform(0, 1.55), rng.uniform(0, 6.28))
        g = O.lib.orc_ggx_g(O.ptr(wo), O.ptr(wi), a, a)
        assert 0 < g <= 1


def test_lambert_closed_form_and_sampling_weight():
    m = mat(0, color=(0.5, 0.25, 0.75))
    wo, wi = unit(0.4), unit(1.0, 2.0)
    f, p = ev(m, wo, wi)
    assert np.allclose(f, np.array([0.5, 0.25, 0.75]) / math.pi * math.cos(1.0), rtol=1e-5)
    assert abs(p - math.cos(1.0) / math.pi) < 1e-6
    f, p = ev(m, wo, -wi)  # transmission side
    assert np.all(f == 0) and p == 0
    rng = np.random.default_rng(2)
    for _ in range(200):
        wi, f, p, spec = sm(m, wo, rng.random(), rng.random())
        assert spec == 0 and p > 0 and wi[2] >= 0
        assert np.allclose(f / p, [0.5, 0.25, 0.75], rtol=1e-5)  # f/pdf = albedo


@pytest.mark.parametrize("type_", [0, 2, 4])
def test_reciprocity(type_):
    m = mat(type_, alpha=0.3, color=(0.6, 0.5, 0.4), eta_i=(1.5, 1.5, 1.5), eta_t=(1, 1, 1)) if type_ == 4 \
        else mat(type_, alpha=0.3, color=(0.6, 0.5, 0.4))
    rng = np.random.default_rng(type_)
    for _ in range(100):
        a = unit(rng.uniform(0.05, 1.5), rng.uniform(0, 6.28))
        b = unit(rng.uniform(0.05, 1.5), rng.uniform(0, 6.28))
        fab, _ = ev(m, a, b)  # f(wo=a, wi=b) * cos(b)
        fba, _ = ev(m, b, a)
        assert np.allclose(fab / b[2], fba / a[2], rtol=2e-4, atol=1e-6)


def sphere_uniform(n, rng):
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * math.pi, n)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1).astype(np.float32)


def albedo_uniform(m, wo, n, rng):
    """integral over the sphere of f*|cos i| by uniform sampling (pdf 1/4pi)."""
    tot = np.zeros(3)
    for wi in sphere_uniform(n, rng):
        f, _ = ev(m, wo, wi)
        tot += f
    return tot * 4 * math.pi / n


def albedo_sampled(m, wo, n, rng):
    tot = np.zeros(3)
    for _ in range(n):
        _, f, p, _ = sm(m, wo, rng.random(), rng.random(), rng.random())
        if p > 0:
            tot += f / p
    return tot / n


def test_rough_dielectric_sampling_matches_pdf_and_energy():
    # broad lobes so that uniform sphere sampling resolves them
    m = mat(3, ior=1.5, alpha=0.6)
    rng = np.random.default_rng(7)
    for theta in (0.3, 1.0):
        for side in (1, -1):
            wo = unit(theta) * np.float32(side)
            a_u = albedo_uniform(m, wo, 120000, rng)
            a_s = albedo_sampled(m, wo, 20000, rng)
            assert np.allclose(a_s, a_u, rtol=0.05), (theta, side, a_s, a_u)
            # energy: reflection + transmission scaled by (eta_i/eta_t)^2 stays bounded
            bound = 1.0 if side > 0 else 1.5 ** 2
            assert a_s[0] <= bound * 1.02


def test_rough_dielectric_smooth_limit():
    # alpha -> 0: the reflected fraction of sampled directions is F(cos theta)
    m = mat(3, ior=1.5, alpha=0.001)
    rng = np.random.default_rng(9)
    wo = unit(0.9)
    n, refl = 4000, 0
    for _ in range(n):
        wi, f, p, _ = sm(m, wo, rng.random(), rng.random(), rng.random())
        refl += wi[2] > 0
    F = O.lib.orc_fr_dielectric(math.cos(0.9), 1.0, 1.5)
    assert abs(refl / n - F) < 0.02


def test_specular_is_delta_and_reflects():
    m = mat(1, ior=1.5)
    wo = unit(0.5)
    f, p = ev(m, wo, unit(0.5, 0.3 + math.pi))
    assert np.all(f == 0) and p == 0
    wi, f, p, spec = sm(m, wo, 0.0, 0.5)  # u < F: mirror direction
    assert spec == 1 and np.allclose(wi, [-wo[0], -wo[1], wo[2]], atol=1e-6)
    assert np.allclose(f / p, [1, 1, 1])
    wi, f, p, spec = sm(m, wo, 0.99, 0.5)  # refraction: Snell's law, weight (1/eta)^2
    assert wi[2] < 0 and abs(math.sqrt(wi[0] ** 2 + wi[1] ** 2) * 1.5 - math.sin(0.5)) < 1e-5
    assert np.allclose(f / p, [1 / 2.25] * 3, rtol=1e-5)


# ---- Trowbridge-Reitz sampling and Smith Lambda (microfacet.h:7-130, 203-226)
# These checks are independent of the oracle's formulas: the visible-normal
# density and the Smith identity are evaluated from the geometry of the
# microsurface (Heitz 2014) with the oracle's D as the only input.
def _slope_grid(ns, nphi, s_min=1e-4):
    """Midpoint grid over the GGX slope plane in (s, phi), where the scaled
    slope radius r = sqrt(rho / (1 - rho)) and rho = 1 - s^2: rho is the
    radial CDF of the alpha = 1 distribution, and s tames its tail."""
    s = s_min + (np.arange(ns) + 0.5) * (1 - s_min) / ns
    ph = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    S, PH = np.meshgrid(s, ph, indexing="ij")
    rho = 1 - S * S
    r = np.sqrt(rho / (1 - rho))
    w = 2 * S * ((1 - s_min) / ns) * (2 * math.pi / nphi)  # d rho d phi
    return r, PH, w


def _half_vectors(r, ph, ax, ay):
    """wh for scaled slopes (r cos phi, r sin phi), and domega_h / (d rho d phi)."""
    x, y = ax * r * np.cos(ph), ay * r * np.sin(ph)
    inv = 1 / np.sqrt(1 + x * x + y * y)  # cos(theta_h)
    wh = np.stack([-x * inv, -y * inv, inv], -1)
    # d omega = cos^3 dx dy, dx dy = ax ay r dr dphi, r dr = (1 + r^2)^2 / 2 d rho
    jac = inv ** 3 * ax * ay * (1 + r * r) ** 2 / 2
    return wh, jac


def _d_at(wh, ax, ay):
    flat = np.ascontiguousarray(wh.reshape(-1, 3), np.float32)
    out = np.zeros(len(flat), np.float32)
    O.lib.orc_ggx_d_n(O.ptr(flat), len(flat), ax, ay, O.ptr(out))
    return out.reshape(wh.shape[:-1]).astype(np.float64)


GGX_CASES = [(0.05, 0.05), (0.3, 0.3), (0.8, 0.8), (0.15, 0.6)]
WO_ANGLES = [(0.2, 0.4), (0.9, 1.1), (1.3, 2.5), (1.5, -2.0)]


@pytest.mark.parametrize("ax,ay", GGX_CASES)
def test_smith_lambda_projected_area_identity(ax, ay):
    # integral of max(0, wo.wh) D(wh) over the microsurface = cos(theta_o) (1 + Lambda(wo)):
    # the projected area of the visible microfacets, which defines Smith's Lambda
    r, ph, w = _slope_grid(1600, 720)
    wh, jac = _half_vectors(r, ph, ax, ay)
    d = _d_at(wh, ax, ay)
    ok = np.isfinite(d)
    for th, phi in WO_ANGLES:
        wo = unit(th, phi).astype(np.float64)
        proj = np.maximum(0.0, wh @ wo)
        integ = float(np.sum((proj * d * jac * w)[ok]))
        lam = O.lib.orc_ggx_lambda(O.ptr(unit(th, phi)), ax, ay)
        want = wo[2] * (1 + lam)
        assert abs(integ - want) < 2e-3 * want, (ax, ay, th, phi, integ, want, lam)
        g1 = 1 / (1 + lam)
        assert abs(g1 * integ - wo[2]) < 2e-3 * wo[2]  # G1 normalises the visible normals
    # Lambda vanishes at normal incidence and is symmetric under wo -> -wo
    assert O.lib.orc_ggx_lambda(O.ptr(np.array([0, 0, 1], np.float32)), ax, ay) == 0
    w1 = unit(1.1, 0.7)
    assert O.lib.orc_ggx_lambda(O.ptr(w1), ax, ay) == O.lib.orc_ggx_lambda(O.ptr(-w1), ax, ay)


def test_ggx_d_isotropic_textbook_form():
    # Walter et al. 2007, eq. 33: D = alpha^2 / (pi cos^4 (alpha^2 + tan^2)^2)
    for a in (0.05, 0.3, 0.8):
        for th in (0.0, 0.3, 0.9, 1.3):
            for phi in (0.0, 1.0, 2.5):
                wh = unit(th, phi)
                c = math.cos(th)
                want = a * a / (math.pi * c ** 4 * (a * a + math.tan(th) ** 2) ** 2)
                got = O.lib.orc_ggx_d(O.ptr(wh), a, a)
                assert abs(got - want) < 2e-5 * want, (a, th, phi, got, want)


def _chi2_pvalue(obs, exp):
    from scipy.stats import chi2
    keep = exp >= 5
    o = np.append(obs[keep], obs[~keep].sum())
    e = np.append(exp[keep], exp[~keep].sum())
    if e[-1] < 5:  # fold a tiny remainder into the largest bin
        o, e = o[:-1].copy(), e[:-1].copy()
        o[np.argmax(e)] += obs[~keep].sum()
        e[np.argmax(e)] += exp[~keep].sum()
    stat = float(np.sum((o - e) ** 2 / e))
    return chi2.sf(stat, len(e) - 1), stat, len(e)


@pytest.mark.parametrize("ax,ay", GGX_CASES + [(0.01, 0.01)])
@pytest.mark.parametrize("side", [1, -1])
def test_ggx_visible_normal_sampling_chi2(ax, ay, side):
    # Sample_wh draws wh from the visible-normal density
    #   D_wo(wh) = G1(wo) max(0, wo.wh) D(wh) / cos(theta_o)            (Heitz 2014),
    # mirrored for wo below the surface.  In scaled-slope coordinates (rho, phi)
    # that density is proportional to max(0, wo.z - r (wo.x ax cos phi + wo.y ay sin phi)):
    # the expected bin counts come from that geometry alone.
    n = 120000
    nr, nphi = 16, 24
    rng = np.random.default_rng(int(ax * 1000) + (7 if side > 0 else 11))
    for th, phi in [(0.5, 1.0), (1.2, -2.2), (1.45, 0.6)]:
        wo = unit(th, phi) * np.float32(side)
        u = rng.random(2 * n, dtype=np.float64).astype(np.float32)
        wh = np.zeros(3 * n, np.float32)
        O.lib.orc_ggx_sample_wh(O.ptr(wo), O.ptr(u), n, ax, ay, O.ptr(wh))
        wh = wh.reshape(-1, 3).astype(np.float64) * side  # back to the upper hemisphere
        wp = wo.astype(np.float64) * side
        assert np.all(wh[:, 2] > 0) and np.allclose(np.linalg.norm(wh, axis=1), 1, atol=1e-5)
        x = -wh[:, 0] / wh[:, 2] / ax
        y = -wh[:, 1] / wh[:, 2] / ay
        r2 = x * x + y * y
        rho = r2 / (1 + r2)
        pph = np.mod(np.arctan2(y, x), 2 * math.pi)
        obs, _, _ = np.histogram2d(rho, pph, bins=[nr, nphi], range=[[0, 1], [0, 2 * math.pi]])
        # expected mass per bin: fine midpoint sub-grid in (s, phi) with rho = 1 - s^2
        r, ph, w = _slope_grid(nr * 60, nphi * 20, s_min=0.0)
        dens = np.maximum(0.0, wp[2] - r * (wp[0] * ax * np.cos(ph) + wp[1] * ay * np.sin(ph))) * w
        rho_g = r * r / (1 + r * r)
        exp, _, _ = np.histogram2d(rho_g.ravel(), ph.ravel(), bins=[nr, nphi], range=[[0, 1], [0, 2 * math.pi]],
                                   weights=dens.ravel())
        exp = exp / exp.sum() * n
        p, stat, dof = _chi2_pvalue(obs.ravel(), exp.ravel())
        assert p > 1e-4, (ax, ay, side, th, phi, stat, dof)


@pytest.mark.parametrize("type_", [2, 4])
def test_microfacet_reflect_sample_is_mirror_of_sampled_normal(type_):
    # MicrofacetReflectSample (microfacet.h:203-226): wi = reflect(wo, wh) with wh
    # from Sample_wh; f is eval(wo, wi); the pdf is D |cos h| / (4 wo.wh) -- the
    # reference's form, not the visible-normal one (its sampler and pdf differ).
    # For the rough plastic (Lambert + GGX) u0 picks the lobe and is rescaled,
    # and f / pdf are the sums over both lobes (bsdf.h:102, 149).
    ax, ay = 0.3, 0.12
    m = mat(type_, alpha=ax, alpha_y=ay, color=(0.7, 0.6, 0.5), eta_i=(1.5, 1.5, 1.5), eta_t=(1, 1, 1)) \
        if type_ == 4 else mat(type_, alpha=ax, alpha_y=ay, color=(0.7, 0.6, 0.5))
    rng = np.random.default_rng(40 + type_)
    glossy = 0
    for _ in range(600):
        wo = unit(rng.uniform(0.05, 1.5), rng.uniform(-3, 3)) * np.float32(rng.choice([1, -1]))
        u0, u1, u2 = (float(np.float32(v)) for v in rng.random(3))
        wi, f, p, spec = sm(m, wo, u0, u1, u2)
        assert spec == 0
        if type_ == 4 and u0 < 0.5:
            continue  # the Lambert lobe
        v0 = u0 * 2 - 1 if type_ == 4 else u0
        # the shading frame of n = +z has s = +y, t = +x (sampling.h:130-144):
        # local (x, y, z) = world (y, x, z)
        wh = np.zeros(3, np.float32)
        O.lib.orc_ggx_sample_wh(O.ptr(np.ascontiguousarray(wo[[1, 0, 2]])), O.ptr(np.array([v0, u1], np.float32)),
                                1, ax, ay, O.ptr(wh))
        wh = wh.astype(np.float64)[[1, 0, 2]]
        wod = wo.astype(np.float64)
        mirror = -wod + 2 * np.dot(wod, wh) * wh
        if p == 0:
            assert mirror[2] * wod[2] <= 1e-6  # reflected below the surface: no sample
            continue
        glossy += 1
        assert np.allclose(wi, mirror, atol=2e-5), (wi, mirror)
        fe, pe = ev(m, wo, wi)
        # eval re-derives the local directions from world ones: a few float ulps apart
        assert np.allclose(f, fe, rtol=1e-4, atol=1e-7) and abs(p - pe) <= 1e-4 * pe
        if type_ == 2:
            hh = (wh * np.sign(wh[2]))[[1, 0, 2]]
            d = O.lib.orc_ggx_d(O.ptr(hh.astype(np.float32)), ax, ay)
            want = d * abs(hh[2]) / (4 * np.dot(wod, wh))
            assert abs(p - want) <= 2e-4 * want
    assert glossy > 200
