"""Pins the oracle's BSDFs (Disney principled BRDF and Lambert).

The reference's Disney closure is compiled but unused and has no tests
(Tracer/CUDA/GPU/Shade/BSDF/Disney.h, PathTracing.h:127-129), so the BSDF is pinned by
what the mathematics fixes: closed forms at normal incidence, reciprocity, the
consistency of bsdf_sample with bsdf_pdf (integral of the pdf over the upper
hemisphere = acceptance rate of the sampler; chi-square of sample histograms against
the pdf), energy bounds, and the exact Lambert sampling weight.
"""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import oracle_py as O
from trayracer_amd.scenes import MODEL_DISNEY, MODEL_LAMBERT, material

MATS = {
    "lambert": material((0.6, 0.3, 0.9), roughness=1.0, model=MODEL_LAMBERT),
    "diffuse": material((0.7, 0.5, 0.2), roughness=0.8, metallic=0.0),
    "plastic": material((0.3, 0.6, 0.8), roughness=0.3, metallic=0.0, specular=0.5),
    "metal_rough": material((0.9, 0.6, 0.3), roughness=0.6, metallic=1.0),
    "metal_shiny": material((0.95, 0.9, 0.8), roughness=0.2, metallic=1.0),
    "aniso": material((0.5, 0.5, 0.5), roughness=0.5, metallic=1.0, anisotropic=0.8),
    "full": material((0.4, 0.7, 0.3), roughness=0.45, metallic=0.3, sheen=0.7, clearcoat=0.8, clearcoat_gloss=0.6,
                     subsurface=0.5, specular_tint=0.4),
}


def _uniform_hemisphere(n, rng):
    u = rng.random((n, 2))
    z = u[:, 0]
    r = np.sqrt(np.maximum(0, 1 - z * z))
    phi = 2 * np.pi * u[:, 1]
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def _dir(theta_deg, phi_deg=20.0):
    t, p = math.radians(theta_deg), math.radians(phi_deg)
    return np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)], np.float32)


def test_lambert_weight_is_albedo():
    rng = np.random.default_rng(0)
    m = MATS["lambert"]
    wo = np.tile(_dir(30), (10000, 1))
    wi, pdf = O.bsdf_sample(m, wo, rng.random((10000, 3)))
    f, pdf2 = O.bsdf_eval(m, wo, wi)
    ok = pdf > 0
    assert ok.mean() > 0.999
    np.testing.assert_allclose(pdf[ok], pdf2[ok], rtol=1e-6)
    np.testing.assert_allclose(f[ok] * wi[ok, 2:3] / pdf[ok, None], np.tile(m[0:3], (ok.sum(), 1)), rtol=2e-6)


@pytest.mark.parametrize("name", ["diffuse", "plastic", "metal_rough", "metal_shiny"])
def test_disney_normal_incidence_closed_form(name):
    m = MATS[name]
    z = np.array([[0, 0, 1]], np.float32)
    f, _ = O.bsdf_eval(m, z, z)
    base, r, met, spec, stint = m[0:3].astype(np.float64), m[3], m[7], m[8], m[9]
    lum = 0.2126729 * base[0] + 0.7151522 * base[1] + 0.0721750 * base[2]
    tint = base / lum
    c0 = (1 - met) * 0.08 * spec * ((1 - stint) + stint * tint) + met * base
    alpha = max(r * r, 1e-3)
    # wo = wi = n: Schlick weights vanish, D(n) = 1/(pi a^2), Smith G2 = 1
    expect = base * (1 - met) / math.pi + c0 / (4 * math.pi * alpha * alpha)
    np.testing.assert_allclose(f[0], expect, rtol=2e-6)


@pytest.mark.parametrize("name", list(MATS))
def test_reciprocity(name):
    rng = np.random.default_rng(1)
    a = _uniform_hemisphere(5000, rng).astype(np.float32)
    b = _uniform_hemisphere(5000, rng).astype(np.float32)
    f1, _ = O.bsdf_eval(MATS[name], a, b)
    f2, _ = O.bsdf_eval(MATS[name], b, a)
    np.testing.assert_allclose(f1, f2, rtol=5e-5, atol=1e-7)


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("theta", [5.0, 45.0, 75.0])
def test_pdf_integrates_to_sampler_acceptance(name, theta):
    """int_{H+} pdf(wo, wi) dwi (uniform-hemisphere MC) equals the fraction of
    bsdf_sample draws that are accepted: a wrong Jacobian, D normalisation or G1
    would break the equality."""
    rng = np.random.default_rng(2)
    m = MATS[name]
    n = 400000
    wo = np.tile(_dir(theta), (n, 1))
    wi, pdf = O.bsdf_sample(m, wo, rng.random((n, 3)))
    accept = (pdf > 0).mean()
    u = _uniform_hemisphere(n, rng).astype(np.float32)
    _, p = O.bsdf_eval(m, wo, u)
    integral = p.astype(np.float64).mean() * 2 * np.pi
    se = p.astype(np.float64).std() * 2 * np.pi / math.sqrt(n)
    assert abs(integral - accept) < 5 * se + 2e-3, (integral, accept, se)
    assert accept <= 1.0 and integral <= 1.0 + 5 * se + 2e-3


@pytest.mark.parametrize("name", ["plastic", "metal_shiny", "full", "aniso"])
def test_sample_histogram_matches_pdf(name):
    """Chi-square of sampled directions against the pdf integrated per (cos, phi) bin."""
    rng = np.random.default_rng(3)
    m = MATS[name]
    n = 300000
    wo = np.tile(_dir(40, 30), (n, 1))
    wi, pdf = O.bsdf_sample(m, wo, rng.random((n, 3)))
    ok = pdf > 0
    nb_t, nb_p = 10, 16
    def cell(w):
        ct = np.clip((w[:, 2] * nb_t).astype(int), 0, nb_t - 1)
        ph = ((np.arctan2(w[:, 1], w[:, 0]) / (2 * np.pi)) % 1.0 * nb_p).astype(int) % nb_p
        return ct * nb_p + ph
    obs = np.bincount(cell(wi[ok]), minlength=nb_t * nb_p).astype(np.float64)
    # expected: n * int_cell pdf dw, by stratified MC over each cell (uniform in cos and phi)
    k = 8000
    ct = np.repeat(np.arange(nb_t), nb_p * k)
    ph = np.tile(np.repeat(np.arange(nb_p), k), nb_t)
    uu = rng.random((ct.size, 2))
    z = (ct + uu[:, 0]) / nb_t
    phi = 2 * np.pi * (ph + uu[:, 1]) / nb_p
    r = np.sqrt(np.maximum(0, 1 - z * z))
    w = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32)
    _, p = O.bsdf_eval(m, np.tile(_dir(40, 30), (w.shape[0], 1)), w)
    expct = p.astype(np.float64).reshape(nb_t * nb_p, k).mean(axis=1) * (2 * np.pi / (nb_t * nb_p)) * n
    mask = expct > 20
    scale = obs[mask].sum() / expct[mask].sum()
    assert abs(scale - 1) < 0.02
    chi2 = ((obs[mask] - expct[mask] * scale) ** 2 / (expct[mask] * scale)).sum()
    dof = mask.sum() - 1
    # the expected counts carry their own MC error: allow 2x slack
    assert chi2 / dof < 2.0, (chi2, dof)


def _ggx_albedo_normal_incidence(alpha, n=20000):
    """Textbook single-scattering albedo of a white GGX conductor seen along the normal:
    E = int D(h) G2(i) / 4 dw_i with theta_h = theta_i / 2 and Lambda(o) = 0."""
    ti = (np.arange(n) + 0.5) / n * (math.pi / 2)
    d = alpha ** 2 / (math.pi * (np.cos(ti / 2) ** 2 * (alpha ** 2 - 1) + 1) ** 2)
    g2 = 1 / (1 + (-1 + np.sqrt(1 + alpha ** 2 * np.tan(ti) ** 2)) / 2)
    return float((d * g2 / 4 * np.sin(ti)).sum() * (math.pi / 2 / n) * 2 * math.pi)


@pytest.mark.parametrize("roughness", [0.1, 0.4, 0.8, 1.0])
def test_white_metal_albedo(roughness):
    """White furnace of a white metal (F = 1): the importance-sampled albedo never
    exceeds 1 and, along the normal, equals the textbook integral (1 - ln 2 at alpha = 1)."""
    rng = np.random.default_rng(4)
    m = material((1.0, 1.0, 1.0), roughness=roughness, metallic=1.0)
    n = 200000
    if roughness == 1.0:
        assert abs(_ggx_albedo_normal_incidence(1.0) - (1 - math.log(2))) < 1e-6
    for theta in (0.0, 60.0, 85.0):
        wo = np.tile(_dir(theta), (n, 1))
        wi, pdf = O.bsdf_sample(m, wo, rng.random((n, 3)))
        ok = pdf > 0
        f, _ = O.bsdf_eval(m, wo[ok], wi[ok])
        w = f[:, 0].astype(np.float64) * wi[ok, 2] / pdf[ok]
        albedo = w.sum() / n
        se = np.concatenate([w, np.zeros(n - ok.sum())]).std() / math.sqrt(n)
        assert albedo <= 1.0 + 4 * se + 1e-4
        if theta == 0.0:
            expect = _ggx_albedo_normal_incidence(roughness ** 2)
            assert abs(albedo - expect) < 4 * se + 2e-3, (albedo, expect)


def _black(**kw):
    """Disney material with no diffuse, no specular reflectance at normal incidence and
    no other lobe unless given: base colour black, specular 0, metallic 0."""
    args = dict(roughness=0.5, metallic=0.0, specular=0.0, sheen=0.0, clearcoat=0.0)
    args.update(kw)
    return material((0.0, 0.0, 0.0), **args)


@pytest.mark.parametrize("gloss", [0.0, 0.4, 1.0])
def test_clearcoat_normal_incidence(gloss):
    """Burley 2012 clearcoat at wo = wi = n: 0.25 c D_GTR1(1) F(0.04) G(1)^2 / 4 with
    D_GTR1(cos 0) = (a^2 - 1) / (pi ln(a^2) a^2), a = lerp(0.1, 0.001, gloss), and the
    Smith factors equal to 1; the other lobes vanish for this material."""
    cc = 0.7
    z = np.array([[0, 0, 1]], np.float32)
    f, _ = O.bsdf_eval(_black(clearcoat=cc, clearcoat_gloss=gloss), z, z)
    a = 0.1 + (0.001 - 0.1) * gloss
    d = (a * a - 1) / (math.pi * math.log(a * a) * a * a)
    np.testing.assert_allclose(f[0], 0.25 * cc * d * 0.04 / 4, rtol=2e-5)


def _hemisphere_grid(nt=1500, nphi=720):
    """Half vectors on a (theta, phi) grid dense near the pole, with quadrature weights
    (theta = pi/2 u^3)."""
    u = (np.arange(nt) + 0.5) / nt
    th = 0.5 * math.pi * u ** 3
    jac = 0.5 * math.pi * 3 * u ** 2
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * math.pi
    T, P = np.meshgrid(th, phi, indexing="ij")
    m = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1).reshape(-1, 3)
    w = (np.sin(T) * jac[:, None]).reshape(-1) / nt * (2 * math.pi / nphi)
    return m, np.cos(T).reshape(-1), w


@pytest.mark.parametrize("a", [0.1, 0.03, 0.001])
def test_gtr1_is_normalised(a):
    """The clearcoat distribution integrates to one against cos(theta_m) over the
    hemisphere (quadrature of the oracle's own D); at a = 0.001 this needs its
    denominator formed without cancellation."""
    m, ct, w = _hemisphere_grid(nt=4000 if a < 0.01 else 1500)
    D = O.gtr1_d(a, m).astype(np.float64)
    assert abs((D * ct * w).sum() - 1.0) < 2e-3


def test_sheen_at_known_half_angle():
    """Burley 2012 sheen = sheen x C_sheen x (1 - cos theta_d)^5 (x (1 - metallic)),
    C_sheen = 1 for sheen tint 0, theta_d the angle between wi and the half vector; with
    a near-zero roughness the specular lobe is negligible away from the mirror
    direction, and the diffuse and clearcoat lobes vanish."""
    sheen = 0.8
    m = _black(sheen=sheen, sheen_tint=0.0, roughness=0.0)
    wo = np.array([[math.sin(1.1), 0.0, math.cos(1.1)]], np.float32)
    wi = np.array([[0.0, math.sin(0.9), math.cos(0.9)]], np.float32)
    f, _ = O.bsdf_eval(m, wo, wi)
    h = (wo[0].astype(np.float64) + wi[0]) / np.linalg.norm(wo[0].astype(np.float64) + wi[0])
    cd = float(np.dot(wi[0], h))
    np.testing.assert_allclose(f[0], sheen * (1.0 - cd) ** 5, rtol=1e-5)
    # metallic 0.5 halves it
    f2, _ = O.bsdf_eval(_black(sheen=sheen, sheen_tint=0.0, roughness=0.0, metallic=0.5), wo, wi)
    np.testing.assert_allclose(f2[0], 0.5 * sheen * (1.0 - cd) ** 5, rtol=1e-5)


@pytest.mark.parametrize("sub", [0.0, 0.4, 1.0])
def test_subsurface_normal_incidence(sub):
    """At wo = wi = n Burley's diffuse is base / pi and the Hanrahan-Krueger subsurface
    term 1.25 (F_ss (1 / (cos_i + cos_o) - 0.5) + 0.5) base / pi with F_ss = 1 is
    0.625 base / pi; the BRDF blends them by `subsurface` (specular 0: no other lobe)."""
    base = np.array([0.5, 0.3, 0.2])
    m = material(tuple(base), roughness=0.6, metallic=0.0, specular=0.0, subsurface=sub)
    z = np.array([[0, 0, 1]], np.float32)
    f, _ = O.bsdf_eval(m, z, z)
    np.testing.assert_allclose(f[0], base / math.pi * ((1 - sub) + 0.625 * sub), rtol=2e-6)


@pytest.mark.parametrize("ax,ay", [(0.3, 0.08), (0.5, 0.5), (0.05, 0.4)])
def test_anisotropic_ggx_is_normalised(ax, ay):
    """The anisotropic GGX distribution integrates to one against cos(theta_m) over the
    hemisphere (quadrature of the oracle's own D on a grid dense near the peak)."""
    m, ct, w = _hemisphere_grid()
    D = O.ggx_d(ax, ay, m).astype(np.float64)
    integral = (D * ct * w).sum()
    assert abs(integral - 1.0) < 2e-3, integral
    # the aspect: D falls off faster along the axis of the smaller alpha
    d = O.ggx_d(ax, ay, np.array([[math.sin(0.1), 0, math.cos(0.1)], [0, math.sin(0.1), math.cos(0.1)]]))
    assert (d[0] > d[1]) == (ax > ay) or ax == ay


@pytest.mark.parametrize("name", sorted(MATS))
def test_one_pass_evaluation_equals_separate_functions(name):
    """The integrator's bsdf_f_pdf and bsdf_sample_dir (used when evaluation and sampling
    parameters coincide) give bsdf_f, bsdf_pdf and bsdf_sample's results bit for bit,
    including directions below the horizon and grazing ones."""
    rng = np.random.default_rng(7)
    m = MATS[name]
    n = 20000
    wo = rng.normal(size=(n, 3)).astype(np.float32)
    wi = rng.normal(size=(n, 3)).astype(np.float32)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wo[:200, 2] = 0.0
    wi[200:400, 2] = 1e-7
    wi[400:600] = wo[400:600]  # wi = wo: half vector = wo
    f, pdf = O.bsdf_eval(m, wo, wi)
    f1, pdf1 = O.bsdf_eval_fused(m, wo, wi)
    assert np.array_equal(f.view(np.uint32), f1.view(np.uint32))
    assert np.array_equal(pdf.view(np.uint32), pdf1.view(np.uint32))
    assert (pdf > 0).sum() > n // 4
    u = rng.random((n, 3)).astype(np.float32)
    s_wi, s_pdf = O.bsdf_sample(m, np.abs(wo), u)
    d_wi, d_pdf = O.bsdf_sample_fused(m, np.abs(wo), u)
    ok = s_pdf > 0
    assert ok.sum() > n // 2
    assert np.array_equal(s_pdf.view(np.uint32), d_pdf.view(np.uint32))
    assert np.array_equal(s_wi[ok].view(np.uint32), d_wi[ok].view(np.uint32))
