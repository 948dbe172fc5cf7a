"""Pins the oracle's Disney principled BSDF and Marschner hair (CPU only).

Hair: the pdf integrates to 1 over the sphere (tests/test_hair_pdf.cpp:6-39 and
test_pdf_normalization.cpp:78-116, tolerance 0.02 as there), no negative pdf,
sample/pdf consistency.  Disney: lobe closed forms and mixture identities --
front faces evaluate the glass lobe only, the metal lobe of the mixture uses
C0 = 0.5*baseColor, sheen is in f but not in the pdf, the clearcoat GTR1
distribution is normalised.  The remaining exact values of the mixture have no
closed form and no reference run exists in this image (parity unpinned beyond
these checks).
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import hemi_dirs, random_dirs, unit

DISNEY, HAIR, DIEL, LAMB = 7, 8, 4, 1
MARSCH_HAIR = [0.419, 0.697, 1.37, 0.0, 1.55, 0.7, 0.5, np.radians(2.0)]  # commonMaterials.hpp:34-35


def disney(base=(0.8, 0.3, 0.2), spec_trans=0.0, metallic=0.0, subsurface=0.0, specular=0.5, roughness=0.4,
           specular_tint=0.0, anisotropic=0.0, sheen=0.0, sheen_tint=0.0, clearcoat=0.0, clearcoat_gloss=0.5,
           eta=1.5):
    return [*base, spec_trans, metallic, subsurface, specular, roughness, specular_tint, anisotropic, sheen,
            sheen_tint, clearcoat, clearcoat_gloss, eta]


def sphere_grid(nt=600, nph=1200):
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3).astype(np.float32)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nph)).reshape(-1)
    return d, dw


@pytest.mark.parametrize("wo,params", [((0, 1, 0), [0.419, 0.697, 1.37, 0.0, 1.55, 0.7, 0.5, np.radians(2.0)]),
                                       ((0.3, 0.5, 0.4), MARSCH_HAIR)])
def test_hair_pdf_normalised(wo, params):
    d, dw = sphere_grid()
    n = d.shape[0]
    wo_a = np.repeat(unit(np.array([wo])), n, 0)
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    _, pdf = O.bsdf_eval(HAIR, params, wo_a, d, nrm)
    assert np.all(pdf >= 0)
    assert abs(np.sum(pdf * dw) - 1.0) < 0.02


def test_hair_sample_pdf_consistency():
    n = 300_000
    rng = np.random.default_rng(4)
    wo = np.repeat(unit(np.array([[0.3, 0.5, 0.4]])), n, 0)
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wi, f, pdf, fl, ok = O.bsdf_sample(HAIR, MARSCH_HAIR, wo, nrm, rng.random(n, np.float32),
                                       rng.random((n, 2), np.float32))
    assert np.all(ok == 1) and np.all(fl == 9)
    assert np.allclose(np.linalg.norm(wi, axis=1), 1, atol=1e-5)
    # returned pdf and f are scatter_pdf and f at the sampled direction
    f2, p2 = O.bsdf_eval(HAIR, MARSCH_HAIR, wo[:1000], wi[:1000], nrm[:1000])
    assert np.allclose(p2, pdf[:1000], rtol=1e-4, atol=1e-7)
    assert np.allclose(f2, f[:1000], rtol=1e-4, atol=1e-7)
    # E[1/pdf] over the sphere is 4*pi up to the mass of the p = pMax term the pdf leaves out
    est = np.mean(1.0 / np.maximum(pdf, 1e-12)[pdf > 1e-6])
    assert 0.5 * 4 * np.pi < est < 1.5 * 4 * np.pi


def test_disney_front_face_is_glass_only():
    rng = np.random.default_rng(1)
    n = 2000
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wo = hemi_dirs(rng, n, nrm)
    wi = hemi_dirs(rng, n, nrm)
    p = disney(base=(0.6, 0.5, 0.4), roughness=0.5, metallic=0.3, spec_trans=0.4, clearcoat=0.7, sheen=0.5)
    f_in, pdf_in = O.bsdf_eval(DISNEY, p, wo, wi, nrm, np.ones(n, np.uint8))
    # DisneyGlass reflection = baseColor * the rough dielectric's reflection (alpha = roughness^2)
    fd, pd = O.bsdf_eval(DIEL, [1.5, 0.25, 0.25], wo, wi, nrm)
    assert np.allclose(f_in, fd[:, :1] * np.array([0.6, 0.5, 0.4]), rtol=1e-5, atol=1e-8)
    assert np.allclose(pdf_in, pd, rtol=1e-5, atol=1e-8)


def test_disney_metal_lobe_uses_default_variant():
    """metallic = 1, no transmission, no sheen or clearcoat: the mixture is the
    metal lobe's f_variant with C0 = 0.5*baseColor, not its f with C0 = baseColor."""
    rng = np.random.default_rng(2)
    n = 3000
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wo = hemi_dirs(rng, n, nrm)
    wi = hemi_dirs(rng, n, nrm)
    base = np.array([0.9, 0.6, 0.3])
    f_full, _ = O.bsdf_eval(DISNEY, disney(base=base, metallic=1.0, roughness=0.5), wo, wi, nrm, np.zeros(n, np.uint8))
    f_half, _ = O.bsdf_eval(DISNEY, disney(base=base * 0.5, metallic=1.0, roughness=0.5), wo, wi, nrm,
                            np.zeros(n, np.uint8))
    # Schlick is affine in C0: F(0.5 b) = 0.5 F(b) + 0.5 (1-c)^5 ; so f(b) - 2 f(b/2) depends only on geometry
    g = f_full - 2 * f_half
    assert np.allclose(g[:, 0], g[:, 1], rtol=1e-3, atol=1e-6) and np.allclose(g[:, 1], g[:, 2], rtol=1e-3, atol=1e-6)
    assert np.all(g[:, 0] <= 1e-6)


def test_disney_sheen_in_f_not_in_pdf():
    rng = np.random.default_rng(3)
    n = 2000
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wo = hemi_dirs(rng, n, nrm)
    wi = hemi_dirs(rng, n, nrm)
    z = np.zeros(n, np.uint8)
    a_f, a_p = O.bsdf_eval(DISNEY, disney(sheen=0.0), wo, wi, nrm, z)
    b_f, b_p = O.bsdf_eval(DISNEY, disney(sheen=1.0), wo, wi, nrm, z)
    assert np.array_equal(a_p, b_p)
    # sheen lobe, tint 0: (1 - |wh.wi|)^5 times weight (1 - metallic) * sheen
    wh = unit(wo + wi)
    sh = (1 - np.abs(np.sum(wh * wi, 1))) ** 5
    assert np.allclose(b_f - a_f, sh[:, None], rtol=1e-3, atol=1e-6)


def test_disney_diffuse_normal_incidence():
    # wo = wi = n, roughness 0, no subsurface: Fd90 = 0.5 and both Schlick factors are 1,
    # so the diffuse lobe is baseColor/pi; metallic 0 and specular transmission 0 make the
    # metal weight 1 as well (disneyPrincipledBSDF.hpp:58-59)
    nrm = np.array([[0, 0, 1]], np.float32)
    w = np.array([[0, 0, 1]], np.float32)
    p = disney(base=(0.5, 0.5, 0.5), roughness=0.0)
    f, _ = O.bsdf_eval(DISNEY, p, w, w, nrm, np.zeros(1, np.uint8))
    pm = disney(base=(0.5, 0.5, 0.5), roughness=0.0, metallic=1.0)  # diffuse weight 0, metal weight 1
    fm, _ = O.bsdf_eval(DISNEY, pm, w, w, nrm, np.zeros(1, np.uint8))
    assert np.allclose(f - fm, 0.5 / np.pi, rtol=1e-5)


@pytest.mark.parametrize("gloss", [0.0, 0.5, 0.9])
def test_clearcoat_gtr1_mass(gloss):
    """Clearcoat-only mixture at normal incidence: the pdf D(h)|h.z|/(4|wi.h|) over the
    upper hemisphere is the GTR1 mass of half-vectors with h.z > 1/sqrt(2) (those that
    reflect wo = n above the horizon).  GTR1's CDF is closed-form:
    P(theta_h < t) = 1 - ln(1 - cos^2 t (1 - a^2)) / ln(a^2)."""
    nt, nph = 4000, 8
    s = (np.arange(nt) + 0.5) / nt
    th = 0.5 * np.pi * s ** 3                      # cells refined towards the pole
    dth = 0.5 * np.pi * 3 * s ** 2 / nt
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3).astype(np.float32)
    dw = (np.sin(T) * dth[:, None] * (2 * np.pi / nph)).reshape(-1)
    n = d.shape[0]
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wo = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    # metallic 1 with roughness 0 makes the metal lobe a delta (pdf 0): only clearcoat remains
    p = disney(metallic=1.0, roughness=0.0, clearcoat=1.0, clearcoat_gloss=gloss)
    _, pdf = O.bsdf_eval(DISNEY, p, wo, d, nrm, np.zeros(n, np.uint8))
    cw, mw = 0.25, 1.0
    a = (1 - gloss) * 0.1 + gloss * 0.001
    c2 = 0.5
    mass = 1 - np.log(1 - c2 * (1 - a * a)) / np.log(a * a)
    assert abs(np.sum(pdf * dw) / (cw / (cw + mw)) / mass - 1.0) < 1e-2


def test_disney_sample_pdf_consistency():
    n = 300_000
    rng = np.random.default_rng(5)
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wo = np.repeat(unit(np.array([[0.2, 0.1, 0.95]])), n, 0)
    p = disney(roughness=0.5, metallic=0.3, clearcoat=0.5, sheen=0.3, subsurface=0.3)
    wi, f, pdf, fl, ok = O.bsdf_sample(DISNEY, p, wo, nrm, rng.random(n, np.float32), rng.random((n, 2), np.float32),
                                       None, np.zeros(n, np.uint8))
    m = ok == 1
    assert m.mean() > 0.9
    f2, p2 = O.bsdf_eval(DISNEY, p, wo[:1000], wi[:1000], nrm[:1000], np.zeros(1000, np.uint8))
    mm = m[:1000]
    assert np.allclose(p2[mm], pdf[:1000][mm], rtol=1e-5)
    assert np.allclose(f2[mm], f[:1000][mm], rtol=1e-5)
    g = np.maximum(wi[:, 2], 0) ** 2
    est = np.where(m, g / np.maximum(pdf, 1e-30), 0).mean()
    assert 0.95 * 2 * np.pi / 3 < est < 1.01 * 2 * np.pi / 3


# ---------------------------------------------------------------- hair: closed-form identities
def _fresnel_dielectric(ci, eta):
    """Textbook unpolarised Fresnel reflectance of a dielectric interface (float64)."""
    if ci < 0:
        eta, ci = 1.0 / eta, -ci
    st2 = (1.0 - ci * ci) / eta ** 2
    if st2 >= 1.0:
        return 1.0
    ct = np.sqrt(1.0 - st2)
    r_par = (eta * ci - ct) / (eta * ci + ct)
    r_perp = (ci - eta * ct) / (ci + eta * ct)
    return 0.5 * (r_par ** 2 + r_perp ** 2)


def _fibre_sphere_grid(n_s=1024, n_phi=1024):
    """Midpoint grid over the sphere in the fibre's coordinates for the normal
    n = (0,0,1): the fibre axis is OrthoBasis(n).u = (-1,0,0), sin(theta) is
    the component along it and phi the angle about it, so dw = dsin(theta) dphi."""
    s = (np.arange(n_s) + 0.5) / n_s * 2 - 1
    ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * np.pi - np.pi
    S, P = np.meshgrid(s, ph, indexing="ij")
    c = np.sqrt(1 - S ** 2)
    d = np.stack([-S, c * np.cos(P), c * np.sin(P)], -1).reshape(-1, 3).astype(np.float32)
    return d, S.reshape(-1), P.reshape(-1), (2.0 / n_s) * (2 * np.pi / n_phi)


def _fibre_dir(sin_t, phi):
    c = np.sqrt(1 - sin_t ** 2)
    return np.array([-sin_t, c * np.cos(phi), c * np.sin(phi)], np.float32)


@pytest.mark.parametrize("beta_m", [0.2, 0.4, 0.6])
def test_hair_white_furnace(beta_m):
    """sigma_a = 0: the lobes' energies A_p sum to 1 (the p = pMax term is the
    geometric remainder), Mp is normalised in theta_i for any theta_o (so the
    unrotated theta_o of evalMarschner keeps the identity, hairMarsch.cpp:67-92)
    and Np in phi, hence the integral of f |cos| over the sphere is 1.  The
    reference's fastExp / I0 approximations leave it within 1.5% below."""
    d, _, _, dw = _fibre_sphere_grid()
    n = d.shape[0]
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    for h in (0.0, 0.5, -0.7):
        for sin_o, phi_o in ((0.0, 1.5707964), (0.6, 0.0), (0.3, 1.03)):
            wo = np.repeat(_fibre_dir(sin_o, phi_o)[None], n, 0)
            f, _ = O.bsdf_eval(HAIR, [0, 0, 0, h, 1.55, beta_m, 0.3, np.radians(2.0)], wo, d, nrm)
            e = (f * np.abs(d[:, 2:3])).sum(0) * dw
            assert np.all(e > 0.985) and np.all(e < 1.005), (h, sin_o, e)


def test_hair_r_lobe_energy_and_direction():
    """A strongly absorbing fibre (sigma_a = 50) leaves the R lobe alone: its
    energy is A0 = Fresnel(cos theta_o cos gamma_o, eta) with cos gamma_o =
    sqrt(1 - h^2) (hairMarsch.hpp:96-103), its peak lies on the specular cone
    sin theta_i = -sin theta_o (Mp's exp(-sin_i sin_o / v), :81-86), and in
    azimuth at Phi(0) = -2 gamma_o (:89-92) with the reference's gamma_o =
    asin(sin theta_o) (hairMarsch.cpp:38), not pbrt's asin(h)."""
    d, S, P, dw = _fibre_sphere_grid()
    n = d.shape[0]
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    for h in (0.0, 0.5, -0.7):
        for sin_o, phi_o in ((0.3, 0.4), (-0.5, 1.2)):
            wo = np.repeat(_fibre_dir(sin_o, phi_o)[None], n, 0)
            f, _ = O.bsdf_eval(HAIR, [50.0, 50.0, 50.0, h, 1.55, 0.2, 0.3, np.radians(2.0)], wo, d, nrm)
            g = f[:, 0].astype(np.float64) * np.abs(d[:, 2])
            a0 = _fresnel_dielectric(np.sqrt(1 - sin_o ** 2) * np.sqrt(1 - h * h), 1.55)
            assert abs(g.sum() * dw / a0 - 1.0) < 0.01, (h, sin_o)
            k = int(np.argmax(g))
            assert abs(S[k] + sin_o) < 0.012
            dphi = (P[k] - phi_o + np.pi) % (2 * np.pi) - np.pi
            assert abs(dphi + 2 * np.arcsin(sin_o)) < 0.01
