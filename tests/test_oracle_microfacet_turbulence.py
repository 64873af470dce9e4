"""Pins of the oracle's microfacet reflection and turbulence against published closed forms (no GPU).

* Trowbridge-Reitz (PBRT-v3 section 8.4; Microfacet.cuh:124-229): for an isotropic alpha,
  D(wh) = 1 / (pi a^2 cos^4 th (1 + tan^2 th / a^2)^2), Lambda(w) = (-1 + sqrt(1 + a^2 tan^2 th)) / 2,
  G = 1 / (1 + Lambda(wo) + Lambda(wi)); MicrofacetReflection with FresnelNoOp has
  f = R D G / (4 cos th_o cos th_i) and pdf = D(wh) G1(wo) |wo . wh| / |cos th_o| / (4 wo . wh).
  f is reciprocal and reflects at most the incident energy; D is normalised (checked here on the
  closed form itself, which the oracle values are then held to). RoughnessToAlpha is PBRT-v3's
  quartic in log(roughness).
* Turbulence (Procedural.cuh:181-205) is PBRT-v3's Turbulence: |noise| summed over the full
  octaves, the partial octave lerped from 0.2 towards |noise|, then 0.2 per octave up to maxOctaves;
  the octave count comes from the differentials. Checked against a Python restatement of that
  published algorithm over Ken Perlin's improved noise (the permutation table fixture).
"""
import math

import numpy as np

import oracle_lib as O
from test_oracle_pins import _perlin_ref, _perm_table


def _bxdf(kind, param, mode, wo, wi=(0, 0, 1), u=(0.5, 0.5)):
    out = np.zeros(6, np.float32)
    a, b, c = (np.asarray(x, np.float32) for x in (wo, wi, u))
    O.lib().oracle_bxdf.argtypes = [O.C.c_int32, O.C.c_float, O.C.c_int32] + [O.C.c_void_p] * 4
    O.lib().oracle_bxdf(kind, param, mode, O._p(a), O._p(b), O._p(c), O._p(out))
    return out


def _alpha(rough):
    x = math.log(max(rough, 1e-3))
    return 1.62142 + 0.819955 * x + 0.1734 * x * x + 0.0171201 * x ** 3 + 0.000640711 * x ** 4


def _D(wh, a):
    c2 = wh[2] ** 2
    t2 = (1 - c2) / c2
    return 1 / (math.pi * a * a * c2 * c2 * (1 + t2 / (a * a)) ** 2)


def _Lambda(w, a):
    c2 = w[2] ** 2
    t2 = (1 - c2) / c2
    return (-1 + math.sqrt(1 + a * a * t2)) / 2


def _f(wo, wi, a):
    wh = (wo + wi) / np.linalg.norm(wo + wi)
    G = 1 / (1 + _Lambda(wo, a) + _Lambda(wi, a))
    return _D(wh, a) * G / (4 * abs(wo[2]) * abs(wi[2]))


def _pdf(wo, wi, a):
    wh = (wo + wi) / np.linalg.norm(wo + wi)
    G1 = 1 / (1 + _Lambda(wo, a))
    return _D(wh, a) * G1 * abs(wo @ wh) / abs(wo[2]) / (4 * (wo @ wh))


def _hemi(rng, n):
    z = rng.uniform(0.05, 1, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32).astype(np.float64)


def test_roughness_to_alpha_is_pbrt_polynomial():
    fn = O.lib().oracle_roughness_to_alpha
    fn.restype = O.C.c_float
    fn.argtypes = [O.C.c_float]
    for r in (1e-4, 1e-3, 0.05, 0.1, 0.3, 0.5, 0.9):
        assert abs(fn(r) - _alpha(r)) < 2e-6 * max(1, _alpha(r)), r


def test_closed_form_D_is_normalised():
    # integral of D(wh) cos(th) over the hemisphere = 1, midpoint rule in (cos th, phi)
    for a in (0.2, 0.5, 1.0):
        n = 200000
        mu = (np.arange(n) + 0.5) / n   # cos(theta); D is isotropic, so phi integrates to 2 pi
        c2 = mu * mu
        D = 1 / (math.pi * a * a * c2 * c2 * (1 + (1 - c2) / c2 / (a * a)) ** 2)
        assert abs((D * mu).sum() / n * 2 * math.pi - 1) < 2e-3, a


def test_microfacet_f_and_pdf_match_trowbridge_reitz():
    rng = np.random.Generator(np.random.PCG64(17))
    for rough in (0.05, 0.2, 0.5):
        a = _alpha(rough)
        wos, wis = _hemi(rng, 150), _hemi(rng, 150)
        for wo, wi in zip(wos, wis):
            f = _bxdf(4, rough, 0, wo, wi)[0]
            ref = _f(wo, wi, a)
            assert abs(f - ref) <= 2e-4 * ref + 1e-7, (rough, wo, wi, f, ref)
            p = _bxdf(4, rough, 1, wo, wi)[0]
            pref = _pdf(wo, wi, a)
            assert abs(p - pref) <= 2e-4 * pref + 1e-7, (rough, p, pref)
            # reciprocity (up to the order of the float products and sums)
            assert abs(_bxdf(4, rough, 0, wi, wo)[0] - f) <= 1e-6 * f


def test_microfacet_reflects_at_most_the_incident_energy():
    """Albedo integral of f cos over the hemisphere, by stratified uniform directions: at most 1,
    and equal to the same quadrature of the closed form (a single-scattering GGX lobe loses more
    energy to masking the rougher it is: about 0.86 at alpha 0.31, 0.39 at alpha 0.86). Exactly at
    the pole cos(phi) is 0 / 0 (the reference's sinTheta test never fires, BxDF.cuh:144-154), so the
    head-on direction is tilted by 0.01."""
    for rough, wo in ((0.05, [0.3, 0.0, 0.95]), (0.3, [0.01, 0.0, 1.0]), (0.3, [0.6, 0.0, 0.8]), (0.5, [0.8, 0.0, 0.6])):
        wo = np.asarray(wo) / np.linalg.norm(wo)
        a = _alpha(rough)
        n_mu, n_phi = 120, 120
        tot, ref = 0.0, 0.0
        for i in range(n_mu):
            mu = (i + 0.5) / n_mu
            r = math.sqrt(1 - mu * mu)
            for j in range(n_phi):
                phi = (j + 0.5) / n_phi * 2 * math.pi
                wi = np.array([r * math.cos(phi), r * math.sin(phi), mu])
                tot += _bxdf(4, rough, 0, wo, wi)[0] * mu
                ref += _f(wo, wi, a) * mu
        albedo = tot * 2 * math.pi / (n_mu * n_phi)
        ref *= 2 * math.pi / (n_mu * n_phi)
        assert albedo <= 1.0, (rough, wo, albedo)
        assert abs(albedo - ref) < 1e-3 * ref, (rough, wo, albedo, ref)


def _turbulence_ref(p, dpdx, dpdy, omega, max_oct, perm):
    len2 = max(float(np.dot(dpdx, dpdx)), float(np.dot(dpdy, dpdy)))
    n = -1 - 0.5 * math.log2(len2) if len2 > 0 else math.inf
    n = min(max(n, 0.0), float(max_oct))
    n_int = int(math.floor(n))
    total, lam, o = 0.0, 1.0, 1.0
    for _ in range(n_int):
        total += o * abs(_perlin_ref([float(x) * lam for x in p], perm))
        lam *= 1.99
        o *= omega
    part = n - n_int
    t = min(max((part - 0.3) / (0.7 - 0.3), 0.0), 1.0)
    s = t * t * (3 - 2 * t)
    total += o * ((1 - s) * 0.2 + s * abs(_perlin_ref([float(x) * lam for x in p], perm)))
    for _ in range(n_int, max_oct):
        total += o * 0.2
        o *= omega
    return total


def test_turbulence_matches_pbrt():
    L = O.lib()
    L.oracle_turbulence.restype = O.C.c_float
    L.oracle_turbulence.argtypes = [O.C.c_void_p] * 3 + [O.C.c_float, O.C.c_int32]
    perm = _perm_table()
    rng = np.random.Generator(np.random.PCG64(23))
    z = np.zeros(3, np.float32)
    for _ in range(60):
        p = rng.uniform(-5, 5, 3).astype(np.float32)
        omega = float(rng.uniform(0.3, 0.7))
        oct_ = int(rng.integers(2, 9))
        for dx in (z, np.array([2.0 ** -3, 0, 0], np.float32), rng.uniform(-0.05, 0.05, 3).astype(np.float32)):
            dy = (dx * np.float32(0.5)).astype(np.float32)
            got = L.oracle_turbulence(O._p(p), O._p(dx), O._p(dy), omega, oct_)
            ref = _turbulence_ref(p, dx.astype(np.float64), dy.astype(np.float64), omega, oct_, perm)
            assert abs(got - ref) < 2e-4 * max(1.0, abs(ref)), (p, dx, omega, oct_, got, ref)


# ----------------------------------------------------------------------------- glass, marble, transform
def test_fresnel_specular_sampling_closed_form():
    """FresnelSpecular::sampleF (Specular.cuh:84-136) with the clamp quirk: F = ((eta - 1) / (eta + 1))^2
    whatever the angle; u < F keeps -wo (the 'reflection' passes straight on) with pdf F and f = F R / |cos|;
    otherwise Snell refraction about the normal facing wo, pdf 1 - F, f = T (1 - F) (eta_i / eta_t)^2 / |cos|."""
    eta = 1.5
    F = ((eta - 1) / (eta + 1)) ** 2
    rng = np.random.Generator(np.random.PCG64(41))
    for _ in range(100):
        wo = _hemi(rng, 1)[0] * (1 if rng.random() < 0.5 else -1)
        r = _bxdf(3, eta, 2, wo, u=(F * 0.5, 0.3))
        assert np.allclose(r[:3], -wo, atol=0) and abs(r[3] - F) < 1e-7
        assert abs(r[4] - F / abs(wo[2])) < 1e-5 * r[4]
        r = _bxdf(3, eta, 2, wo, u=(0.5, 0.3))
        entering = wo[2] > 0
        ei, et = (1.0, eta) if entering else (eta, 1.0)
        n = np.array([0, 0, 1.0 if entering else -1.0])
        e = ei / et
        ci = wo @ n
        if 1 - e * e * (1 - ci * ci) <= 0:  # total internal reflection: no transmitted direction
            assert r[4] == 0
            continue
        ct = math.sqrt(1 - e * e * (1 - ci * ci))
        wt = -wo * e + n * (e * ci - ct)
        assert np.allclose(r[:3], wt, atol=2e-6), (wo, r[:3], wt)
        assert abs(r[3] - (1 - F)) < 1e-6
        assert abs(r[4] - (1 - F) * e * e / abs(wt[2])) < 2e-5 * r[4]


MARBLE_COLOURS = np.array([[.58, .58, .6], [.58, .58, .6], [.58, .58, .6], [.5, .5, .5], [.6, .59, .58],
                           [.58, .58, .6], [.58, .58, .6], [.2, .2, .33], [.58, .58, .6]])


def test_marble_is_a_cubic_bezier_over_its_colour_table():
    """MarbleTexture (Procedural.cuh:267-311) without variation: t = 0.5 + 0.5 sin(scale p.y), segment
    floor(6 t) of the 9-colour table, and inside it the de Casteljau evaluation, i.e. the Bernstein
    form (1-s)^3 c0 + 3 (1-s)^2 s c1 + 3 (1-s) s^2 c2 + s^3 c3, times 1.5."""
    L = O.lib()
    L.oracle_marble.argtypes = [O.C.c_void_p, O.C.c_int32, O.C.c_float, O.C.c_float, O.C.c_float, O.C.c_void_p]
    out = np.zeros(3, np.float32)
    for y in np.linspace(-3.0, 3.0, 97):
        scale = 1.7
        p = np.array([0.3, y, -0.8], np.float32)
        L.oracle_marble(O._p(p), 4, 0.5, scale, 0.0, O._p(out))
        t = 0.5 + 0.5 * math.sin(float(np.float32(p[1] * np.float32(scale))))
        first = min(int(math.floor(t * 6)), 6)
        s = t * 6 - first
        c = [MARBLE_COLOURS[first + k] if first + k < 9 else np.zeros(3) for k in range(4)]
        ref = 1.5 * ((1 - s) ** 3 * c[0] + 3 * (1 - s) ** 2 * s * c[1] + 3 * (1 - s) * s * s * c[2] + s ** 3 * c[3])
        assert np.allclose(out, ref, atol=2e-6), (y, out, ref)


def test_subject_transform_is_a_rotation_about_the_axis():
    """The .scene subject transform (translation, unit axis, angle in radians, scale; SceneLoader.cuh:59-77)
    applied by quaternion conjugation (Transform.cuh:131-240): p -> R(axis, angle) p * scale + t by
    Rodrigues' formula. Normals are conjugated the other way round, conj(q) n q (Transform.cuh:218-223):
    they turn by the inverse rotation R^T (and are renormalised), without the scale or translation, a
    quirk of the reference that the oracle keeps."""
    fn = O.lib().oracle_transform
    fn.argtypes = [O.C.c_void_p, O.C.c_void_p, O.C.c_float, O.C.c_float, O.C.c_void_p, O.C.c_void_p, O.C.c_void_p]
    rng = np.random.Generator(np.random.PCG64(43))
    out = np.zeros(6, np.float32)
    for _ in range(50):
        t = rng.uniform(-3, 3, 3).astype(np.float32)
        ax = rng.normal(size=3)
        ax = (ax / np.linalg.norm(ax)).astype(np.float32)
        ang = np.float32(rng.uniform(-3, 3))
        sc = np.float32(rng.uniform(0.5, 2))
        p = rng.uniform(-2, 2, 3).astype(np.float32)
        n = rng.normal(size=3).astype(np.float32)
        fn(O._p(t), O._p(ax), ang, sc, O._p(p), O._p(n), O._p(out))
        a = ax.astype(np.float64)
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K
        assert np.allclose(out[:3], R @ p * sc + t, atol=2e-5 * (1 + np.abs(p).max() * sc)), (out[:3], R @ p * sc + t)
        rn = R.T @ n
        assert np.allclose(out[3:], rn / np.linalg.norm(rn), atol=3e-6)
