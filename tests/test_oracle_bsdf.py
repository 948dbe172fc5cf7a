"""Pins the oracle's materials (CPU only).

Known answers of the reference's own tests (tests/test_dielectric.cpp:80-300,
tests/test_refract.cpp:8-86, tests/test_pdf_normalization.cpp:48-76) and
closed forms: Fresnel at normal incidence and at Brewster's angle, the
conductor's normal-incidence reflectance ((n-1)^2+k^2)/((n+1)^2+k^2), GGX
normalisation of D and of the visible-normal distribution, Lambert, Snell's
law, reciprocity, and Monte-Carlo consistency of sampling against pdf.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle_py as O
from tests.helpers import hemi_dirs, random_dirs, unit

LAMB, COND, DIEL, THIN, PHONG, KK = 1, 3, 4, 5, 6, 9
GOLD = [0.14, 0.43, 1.38, 4.54, 2.455, 1.914]
F3 = ctypes.c_float * 3

lib = O.lib
lib.or_fresnel_dielectric.restype = ctypes.c_float
lib.or_fresnel_dielectric.argtypes = [ctypes.c_float, ctypes.c_float]
lib.or_trd_D.restype = ctypes.c_float
lib.or_trd_D.argtypes = [ctypes.c_float, ctypes.c_float, F3]
lib.or_trd_Dw.restype = ctypes.c_float
lib.or_trd_Dw.argtypes = [ctypes.c_float, ctypes.c_float, F3, F3]
lib.or_refract.restype = ctypes.c_int
lib.or_refract.argtypes = [F3, F3, ctypes.c_float, F3, ctypes.POINTER(ctypes.c_float)]
lib.or_fresnel_complex.argtypes = [ctypes.c_float, F3, F3, F3]


def smooth_scatter(kind, params, wo, uc, n=(0.0, 0.0, 1.0)):
    wo = unit(np.array([wo], np.float64))
    return O.bsdf_sample(kind, params, wo, np.array([n], np.float32), np.array([uc], np.float32),
                         np.array([[0.5, 0.5]], np.float32))


# ---------------------------------------------------------------- test_dielectric.cpp
def test_smooth_air_to_glass():
    R = ((1.5 - 1) / (1.5 + 1)) ** 2
    wi, f, pdf, fl, ok = smooth_scatter(DIEL, [1.5, 0, 0], (0, 0, 1), 0.01)
    assert ok[0] and fl[0] == 16 | 1 and abs(pdf[0] - R) < 1e-6
    assert np.allclose(wi[0], [0, 0, 1], atol=1e-6)
    wi, f, pdf, fl, ok = smooth_scatter(DIEL, [1.5, 0, 0], (0, 0, 1), 0.99)
    assert ok[0] and fl[0] == 16 | 2 and abs(pdf[0] - (1 - R)) < 1e-6
    assert np.allclose(wi[0], [0, 0, -1], atol=1e-6)
    assert np.allclose(f[0], 1.0)  # TransportMode::Importance: no 1/eta^2


def test_smooth_glass_to_air():
    e = 1 / 1.5
    R = ((e - 1) / (e + 1)) ** 2
    wi, f, pdf, fl, ok = smooth_scatter(DIEL, [1.5, 0, 0], (0, 0, -1), 0.01)
    assert ok[0] and fl[0] == 17 and abs(pdf[0] - R) < 1e-6
    assert np.allclose(wi[0], [0, 0, -1], atol=1e-6)
    wi, f, pdf, fl, ok = smooth_scatter(DIEL, [1.5, 0, 0], (0, 0, -1), 0.99)
    assert ok[0] and fl[0] == 18 and abs(pdf[0] - (1 - R)) < 1e-6
    assert np.allclose(wi[0], [0, 0, 1], atol=1e-6)


def test_smooth_total_internal_reflection():
    cos_crit = np.sqrt(1 - (1 / 1.5) ** 2)
    c = cos_crit * 0.8
    s = np.sqrt(1 - c * c)
    wi, f, pdf, fl, ok = smooth_scatter(DIEL, [1.5, 0, 0], (s, 0, -c), 0.5)
    assert ok[0] and fl[0] == 17 and abs(pdf[0] - 1.0) < 1e-6
    assert np.allclose(wi[0], [-s, 0, -c], atol=1e-6)


def test_very_low_roughness_is_smooth():
    R = ((1.5 - 1) / (1.5 + 1)) ** 2
    wi, f, pdf, fl, ok = smooth_scatter(DIEL, [1.5, 1e-5, 1e-5], (0, 0, 1), 0.01)
    assert fl[0] == 17 and abs(pdf[0] - R) < 1e-5 and abs(wi[0][2] - 1) < 1e-3
    wi, f, pdf, fl, ok = smooth_scatter(DIEL, [1.5, 1e-5, 1e-5], (0, 0, 1), 0.99)
    assert fl[0] == 18 and abs(wi[0][2] + 1) < 1e-3


@pytest.mark.parametrize("wo", [(0.99, 0.1, 0.01), (0.01, 0.01, 0.99), (0.01, 0.01, -0.99)])
def test_rough_dielectric_invariants(wo):
    # test_dielectric.cpp:175-261 (with BIASED_SAMPLING, as the renderer builds)
    n = 100_000
    rng = np.random.default_rng(7)
    wo_a = np.repeat(unit(np.array([wo])), n, 0)
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wi, f, pdf, fl, ok = O.bsdf_sample(DIEL, [1.5, 0.3, 0.3], wo_a, nrm, rng.random(n, np.float32),
                                       rng.random((n, 2), np.float32))
    m = ok == 1
    assert m.mean() > 0.9
    assert np.all(np.isfinite(pdf[m])) and np.all(pdf[m] >= 0)
    assert np.all(np.isfinite(f[m])) and np.all(f[m] >= 0)
    assert np.allclose(np.sum(wi[m] ** 2, 1), 1.0, atol=1e-5)
    assert np.all((fl[m] == 8 | 1) | (fl[m] == 8 | 2))


# ---------------------------------------------------------------- test_refract.cpp / Snell
def test_refract_snell_law():
    ang = np.pi / 6
    I = F3(np.sin(ang), 0, np.cos(ang))
    wt, etap = F3(), ctypes.c_float()
    assert lib.or_refract(I, F3(0, 0, 1), 1.5, wt, ctypes.byref(etap))
    assert abs(etap.value - 1.5) < 1e-7
    sin_t = np.hypot(wt[0], wt[1])
    assert abs(sin_t - np.sin(ang) / 1.5) < 1e-6          # sin(theta_t) = sin(theta_i)/eta
    assert wt[2] < 0 and abs(np.linalg.norm(list(wt)) - 1) < 1e-6


def test_refract_identity_and_tir():
    wt, etap = F3(), ctypes.c_float()
    assert lib.or_refract(F3(0, 0, 1), F3(0, 0, 1), 1.0, wt, ctypes.byref(etap))
    assert np.allclose(list(wt), [0, 0, -1], atol=1e-6)    # eta = 1 passes straight through
    a = np.pi / 4      # 45 degrees with relative index 1/2: sin^2(t) = 0.5 * 4 >= 1, total reflection
    assert not lib.or_refract(F3(np.sin(a), 0, np.cos(a)), F3(0, 0, 1), 1 / 2.0, wt, ctypes.byref(etap))


# ---------------------------------------------------------------- Fresnel closed forms
def test_fresnel_dielectric_closed_forms():
    for eta in (1.33, 1.5, 2.4):
        assert abs(lib.or_fresnel_dielectric(1.0, eta) - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
        brewster = np.arctan(eta)
        ci, ct = np.cos(brewster), np.sqrt(1 - np.sin(brewster) ** 2 / eta ** 2)
        rs = ((ci - eta * ct) / (ci + eta * ct)) ** 2
        assert abs(lib.or_fresnel_dielectric(float(ci), eta) - rs / 2) < 1e-5  # r_parallel = 0
    assert lib.or_fresnel_dielectric(0.1, 1 / 1.5) == 1.0 or lib.or_fresnel_dielectric(-0.1, 1.5) == 1.0


def test_conductor_normal_incidence():
    out = F3()
    lib.or_fresnel_complex(1.0, F3(*GOLD[:3]), F3(*GOLD[3:]), out)
    n, k = np.array(GOLD[:3]), np.array(GOLD[3:])
    expect = ((n - 1) ** 2 + k ** 2) / ((n + 1) ** 2 + k ** 2)
    assert np.allclose(list(out), expect, rtol=1e-5)
    # k = 0 reduces to the dielectric formula
    lib.or_fresnel_complex(0.6, F3(1.5, 1.5, 1.5), F3(0, 0, 0), out)
    assert abs(out[0] - lib.or_fresnel_dielectric(0.6, 1.5)) < 1e-6
    # the smooth conductor returns F as its path weight (conductor.hpp:117-133)
    wi, f, pdf, fl, ok = smooth_scatter(COND, GOLD + [1e-4, 1e-4], (0, 0, 1), 0.5)
    assert fl[0] == 17 and np.allclose(f[0], expect, rtol=1e-5) and np.allclose(wi[0], [0, 0, 1])


# ---------------------------------------------------------------- GGX
def _hemisphere_grid(nt=1500, nph=600):
    th = (np.arange(nt) + 0.5) / nt * (np.pi / 2)
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / 2 / nt) * (2 * np.pi / nph)).reshape(-1)
    return d.astype(np.float32), dw


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.2, 0.5), (0.6, 0.6)])
def test_ggx_normalisation(ax, ay):
    d, dw = _hemisphere_grid()
    D = np.array([lib.or_trd_D(ax, ay, F3(*m)) for m in d[::7]])
    # int D(m) cos(theta_m) dw = 1 (projected microfacet area)
    assert abs(np.sum(D * d[::7, 2] * dw[::7]) * 7 - 1.0) < 5e-3
    # visible normals: int G1(w)/cos(w) D(m) max(0, w.m) dm = 1
    w = F3(*unit(np.array([0.4, 0.2, 0.8])))
    Dw = np.array([lib.or_trd_Dw(ax, ay, w, F3(*m)) if np.dot(m, list(w)) > 0 else 0.0 for m in d[::7]])
    assert abs(np.sum(Dw * dw[::7]) * 7 - 1.0) < 5e-3


def test_microfacet_pdf_expectation():
    # test_pdf_normalization.cpp:48-76: alpha 0.05, wo = (1,1,1)/sqrt(3): E[1/p] = 2*pi.
    # With BIASED_SAMPLING the samples that leave the hemisphere end the path, so
    # the estimator is E[1{ok}/p] over all samples.
    n = 1_000_000
    rng = np.random.default_rng(3)
    wo = np.repeat(unit(np.array([[1, 1, 1]])), n, 0)
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wi, f, pdf, fl, ok = O.bsdf_sample(COND, GOLD + [0.05, 0.05], wo, nrm, rng.random(n, np.float32),
                                       rng.random((n, 2), np.float32))
    est = np.where(ok == 1, 1.0 / np.maximum(pdf, 1e-30), 0.0).mean()
    assert abs(est - 2 * np.pi) < 0.02 * 2 * np.pi


@pytest.mark.parametrize("kind,params", [(LAMB, [0.5, 0.5, 0.5]), (COND, GOLD + [0.3, 0.3]),
                                         (COND, GOLD + [0.1, 0.4])])
def test_sample_pdf_consistency(kind, params):
    """E over samples of g(wi)/pdf(wi) equals the integral of g over the hemisphere
    (g = cos theta): pdf is the density the sampler draws from."""
    n = 400_000
    rng = np.random.default_rng(11)
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wo = np.repeat(unit(np.array([[0.3, -0.2, 0.9]])), n, 0)
    wi, f, pdf, fl, ok = O.bsdf_sample(kind, params, wo, nrm, rng.random(n, np.float32),
                                       rng.random((n, 2), np.float32))
    g = np.maximum(wi[:, 2], 0) ** 2
    est = np.where(ok == 1, g / np.maximum(pdf, 1e-30), 0.0).mean()
    # int cos^2 dw over the hemisphere = 2*pi/3; the sampler covers the whole
    # hemisphere except what reflects below it
    assert est <= 2 * np.pi / 3 * 1.01
    assert est > 2 * np.pi / 3 * (0.97 if kind == COND else 0.99)
    # pdf value returned by scatter equals scatter_pdf(wo, wi)
    f2, p2 = O.bsdf_eval(kind, params, wo[:2000], wi[:2000], nrm[:2000])
    m = ok[:2000] == 1
    assert np.allclose(p2[m], pdf[:2000][m], rtol=1e-5)
    assert np.allclose(f2[m], f[:2000][m], rtol=1e-5)


def test_lambert_closed_form():
    rng = np.random.default_rng(5)
    n = 1000
    nrm = random_dirs(rng, n)
    wo = hemi_dirs(rng, n, nrm)
    wi = hemi_dirs(rng, n, nrm)
    f, pdf = O.bsdf_eval(LAMB, [0.2, 0.4, 0.6], wo, wi, nrm)
    assert np.allclose(f, np.array([0.2, 0.4, 0.6]) / np.pi, rtol=1e-6)
    assert np.allclose(pdf, np.sum(wi * nrm, 1) / np.pi, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("kind,params", [(COND, GOLD + [0.3, 0.3]), (DIEL, [1.5, 0.3, 0.3])])
def test_reflection_reciprocity(kind, params):
    rng = np.random.default_rng(9)
    n = 5000
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wo = hemi_dirs(rng, n, nrm)
    wi = hemi_dirs(rng, n, nrm)
    f1, _ = O.bsdf_eval(kind, params, wo, wi, nrm)
    f2, _ = O.bsdf_eval(kind, params, wi, wo, nrm)
    assert np.allclose(f1, f2, rtol=1e-4, atol=1e-7)


def test_dielectric_pdf_integrates_to_one():
    """scatter_pdf of the rough dielectric over the whole sphere sums its
    reflection and transmission parts to 1 (near-normal incidence, where
    neither lobe leaks across the horizon)."""
    nt, nph = 1400, 700
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3).astype(np.float32)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nph)).reshape(-1)
    n = d.shape[0]
    wo = np.repeat(unit(np.array([[0.1, 0.05, 0.99]])), n, 0)
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    _, pdf = O.bsdf_eval(DIEL, [1.5, 0.3, 0.3], wo, d, nrm)
    assert abs(np.sum(pdf * dw) - 1.0) < 1e-2


def test_phong_lobe_normalised():
    d, dw = _hemisphere_grid(1000, 400)
    n = d.shape[0]
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wo = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    # pure specular Phong (diffuse 0) at normal incidence: lobe around the normal
    _, pdf = O.bsdf_eval(PHONG, [0, 0, 0, 0.5, 0.5, 0.5, 0, 0, 0, 20.0], wo, d, nrm)
    assert abs(np.sum(pdf * dw) - 1.0) < 5e-3


# ---------------------------------------------------------------- ThinDielectric
def _fresnel(ci, eta):
    """Textbook unpolarised Fresnel reflectance of a dielectric interface (float64)."""
    st2 = (1.0 - ci * ci) / eta ** 2
    if st2 >= 1.0:
        return 1.0
    ct = np.sqrt(1.0 - st2)
    r_par = (eta * ci - ct) / (eta * ci + ct)
    r_perp = (ci - eta * ct) / (ci + eta * ct)
    return 0.5 * (r_par ** 2 + r_perp ** 2)


@pytest.mark.parametrize("cos_o", [1.0, 0.8, 0.5, 0.2, 0.05])
def test_thin_dielectric_closed_form(cos_o):
    """ThinDielectric (thinDielectric.hpp:26-33): a thin slab reflects
    R' = R + T^2 R / (1 - R^2), which sums to 2R / (1 + R), and passes the rest
    straight through (wi = -wo).  The branch is chosen by uc against R', its
    pdf is R' or 1 - R', and the path weight is 1 either way."""
    R = _fresnel(cos_o, 1.5)
    Rt = 2 * R / (1 + R)
    s = np.sqrt(1 - cos_o ** 2)
    wo = (s, 0.0, cos_o)
    wi, f, pdf, fl, ok = smooth_scatter(THIN, [1.5, 0.0], wo, Rt * (1 - 1e-4))
    assert ok[0] and fl[0] == 16 | 1
    assert abs(pdf[0] - Rt) < 1e-5 * Rt
    assert np.allclose(wi[0], [-s, 0, cos_o], atol=1e-6) and np.allclose(f[0], 1.0)
    wi, f, pdf, fl, ok = smooth_scatter(THIN, [1.5, 0.0], wo, min(Rt * (1 + 1e-4), 0.9999))
    assert ok[0] and fl[0] == 16 | 2
    assert abs(pdf[0] - (1 - Rt)) < 1e-5
    assert np.allclose(wi[0], [-s, 0, -cos_o], atol=1e-6) and np.allclose(f[0], 1.0)


def test_thin_dielectric_index_matched_passes_through():
    """eta = 1: no reflection at all (R = 0), every sample passes straight through."""
    for uc in (0.0, 0.5, 0.999):
        wi, f, pdf, fl, ok = smooth_scatter(THIN, [1.0, 0.0], (0.3, 0.2, 0.9), uc)
        assert ok[0] and fl[0] == 16 | 2 and abs(pdf[0] - 1.0) < 1e-6
        assert np.allclose(wi[0], -unit(np.array([0.3, 0.2, 0.9])), atol=1e-6)


# ---------------------------------------------------------------- Kajiya-Kay
KK_PARAMS = [0.4, 0.3, 0.2, 0.3, 0.25, 0.2, 30.0]


def _kk_expect(sin_i, spec_cos):
    """Kajiya & Kay (1989): diffuse kd/pi sin(t, wi); specular ks (p+2)/(2pi) cos^p
    of the angle to the mirror cone, or 0 where that cosine is negative."""
    kd, ks, p = np.array(KK_PARAMS[0:3]), np.array(KK_PARAMS[3:6]), KK_PARAMS[6]
    return kd / np.pi * sin_i + ks * (p + 2) / (2 * np.pi) * max(spec_cos, 0.0) ** p


def test_kajiya_kay_closed_forms():
    """On a surface the fibre tangent is OrthoBasis(n).u: (-1,0,0) for n = +z and
    (0,-1,0) for n = +x (orthoBasis.hpp:15-21).  Values at normal incidence, on
    the mirror cone t.wo = -t.wi (the specular peak), at retro-reflection
    (cos of the cone angle = cos 2a) and along the tangent (no diffuse term)."""
    a = 0.6
    sa, ca = np.sin(a), np.cos(a)
    z = (0.0, 0.0, 1.0)
    cases = [  # (normal, wo, wi, expected f)
        (z, z, z, _kk_expect(1.0, 1.0)),
        (z, (-sa, 0, ca), (sa, 0, ca), _kk_expect(ca, 1.0)),                 # mirror cone
        (z, (sa, 0, ca), (sa, 0, ca), _kk_expect(ca, np.cos(2 * a))),        # retro-reflection
        (z, (np.sin(0.9), 0, np.cos(0.9)), (np.sin(0.9), 0, np.cos(0.9)), _kk_expect(np.cos(0.9), np.cos(1.8))),
        (z, (0, 0.6, 0.8), (0, -0.6, 0.8), _kk_expect(1.0, 1.0)),            # plane normal to the fibre
        ((1.0, 0, 0), (1.0, 0, 0), (1.0, 0, 0), _kk_expect(1.0, 1.0)),
        ((1.0, 0, 0), (1.0, 0, 0), (0.0, 1.0, 0), _kk_expect(0.0, 0.0)),     # wi along the tangent
    ]
    for nrm, wo, wi, want in cases:
        f, pdf = O.bsdf_eval(KK, KK_PARAMS, np.array([wo], np.float32), np.array([wi], np.float32),
                             np.array([nrm], np.float32))
        assert np.allclose(f[0], want, rtol=1e-5, atol=1e-7), (nrm, wo, wi, f[0], want)
        assert abs(pdf[0] - max(np.dot(wi, nrm), 0.0) / np.pi) < 1e-6


def test_kajiya_kay_sampling():
    """Cosine-lobe sampling: returned f and pdf are f and scatter_pdf at the
    sampled direction, and E[cos^2 / pdf] is the hemisphere integral 2pi/3."""
    n = 200_000
    rng = np.random.default_rng(13)
    nrm = np.repeat(np.array([[0, 0, 1]], np.float32), n, 0)
    wo = np.repeat(unit(np.array([[0.3, -0.2, 0.9]])), n, 0)
    wi, f, pdf, fl, ok = O.bsdf_sample(KK, KK_PARAMS, wo, nrm, rng.random(n, np.float32),
                                       rng.random((n, 2), np.float32))
    assert np.all(ok == 1) and np.all(fl == 8 | 1)
    est = np.mean(wi[:, 2] ** 2 / pdf)
    assert abs(est - 2 * np.pi / 3) < 0.01 * 2 * np.pi / 3
    f2, p2 = O.bsdf_eval(KK, KK_PARAMS, wo[:2000], wi[:2000], nrm[:2000])
    assert np.allclose(p2, pdf[:2000], rtol=1e-5) and np.allclose(f2, f[:2000], rtol=1e-5)
    assert np.all(f >= 0)
