"""Principled BSDF building blocks of the oracle against the reference's BxDFTest
invariants (tests/unit_test/nova/material/BxDFTest.cpp:27-272) and closed forms."""
import math

import numpy as np
import pytest

UC = [(0.0, 0.0), (0.1, 0.5), (0.2, 0.25), (0.3, 0.75), (0.4, 0.125), (0.5, 0.625), (0.6, 0.375), (0.7, 0.875), (0.8, 0.0625), (0.9, 0.5625)]
ROUGH = [1e-4, 0.1, 0.45, 0.75, 1.0]


def test_specular_normal_correction_keeps_reflection_above_ng(oracle):
    wo = np.array([0.8, 0.6, 0.0]) / 1.0
    ng = np.array([0.0, 1.0, 0.0])
    ns = np.array([0.9, 0.1, 0.0]) / np.linalg.norm([0.9, 0.1, 0.0])
    c = oracle.correct_shading_normal(ng, wo / np.linalg.norm(wo), ns)
    R = -wo / np.linalg.norm(wo) - 2 * np.dot(c, -wo / np.linalg.norm(wo)) * c
    assert np.dot(ng, R) > 0 and np.dot(c, ng) > 0


def test_fresnel_real(oracle):
    assert 1 - oracle.fresnel_real(1.5, 0.5) > oracle.fresnel_real(1.5, 0.5)
    assert oracle.fresnel_real(1.5, 0.001) > 1 - oracle.fresnel_real(1.5, 0.001)
    assert oracle.fresnel_real(2.417, 0.1) > 1 - oracle.fresnel_real(2.417, 0.1)
    # closed form at normal incidence: ((eta - 1) / (eta + 1))^2
    for eta in (1.309, 1.333, 1.5, 2.417):
        assert oracle.fresnel_real(eta, 1.0) == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-6)
    assert oracle.fresnel_real(1.0, 0.5) == 0.0
    # total internal reflection from inside glass beyond the critical angle
    assert oracle.fresnel_real(1.5, -0.2) == 1.0


@pytest.mark.parametrize("rough", ROUGH)
def test_vndf_samples_are_normalized_and_visible(oracle, rough):
    wo = np.array([1.2, 0.4, 1.0], dtype=np.float32)
    for u in UC[:5]:
        wm = oracle.vndf_sample(0.0, rough, wo, u)
        assert abs(np.dot(wm, wm) - 1) < 1e-4
        assert wm[2] * wo[2] > 0


def test_g1_bounds(oracle):
    g = oracle.vndf_G1(0.0, 0.75, [0, 0, 1])
    assert 0 < g <= 1 and abs(g - 1) < 1e-3
    g = oracle.vndf_G1(0.0, 0.75, [1, 1, 1e-4])
    assert 0 <= g < 0.1


@pytest.mark.parametrize("rough", [0.3, 0.6, 0.9])
def test_ggx_d_integrates_projected_area_to_one(oracle, rough):
    """int D(wm) cos(theta_m) dwm = 1 (isotropic GGX, alpha = r^2)."""
    n = 400
    th = (np.arange(n) + 0.5) / n * (math.pi / 2)
    tot = 0.0
    for t in th:
        wm = [math.sin(t), 0.0, math.cos(t)]
        tot += oracle.vndf_D(0.0, rough, wm) * math.cos(t) * math.sin(t) * (math.pi / 2 / n) * 2 * math.pi
    assert tot == pytest.approx(1.0, rel=2e-2)


def test_disney_diffuse_lobe_sample_is_consistent(oracle):
    """Metal 0, no transmission: the diffuse lobe samples the upper hemisphere with pdf = p_d cos/pi."""
    wo = np.array([0.3, -0.2, 0.9])
    wo /= np.linalg.norm(wo)
    seen = 0
    for i in range(64):
        u = [(i * 0.618034) % 1.0, 0.01, (i * 0.754877) % 1.0, (i * 0.569840) % 1.0]
        lobe, ok, f, pdf, wi, ct = oracle.principled_sample(wo, [0.7, 0.5, 0.3], 0.6, 0.0, 0.0, u)
        if lobe != 1 or not ok:
            continue
        seen += 1
        assert wi[2] > 0 and abs(np.linalg.norm(wi) - 1) < 1e-5 and ct == pytest.approx(wi[2], rel=1e-6)
        assert (f > 0).all() and f[0] > f[1] > f[2]
    assert seen > 30


def test_dielectric_lobe_reflects_or_transmits(oracle):
    wo = np.array([0.2, 0.1, 0.97])
    wo /= np.linalg.norm(wo)
    refl = trans = 0
    for i in range(200):
        u = [0.999, (i * 0.618034) % 1.0, (i * 0.754877) % 1.0, (i * 0.3819) % 1.0]
        lobe, ok, f, pdf, wi, ct = oracle.principled_sample(wo, [1, 1, 1], 0.1, 0.0, 1.5, u)
        assert lobe == 3
        if ok:
            refl += wi[2] > 0
            trans += wi[2] < 0
            assert pdf > 0 and abs(np.linalg.norm(wi) - 1) < 1e-5
    assert trans > refl > 0


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def _halton(i, b):
    f, r = 1.0, 0.0
    while i > 0:
        f /= b
        r += f * (i % b)
        i //= b
    return r


# lobe selectors for principled_sample / principled_eval with ns = ng = +z:
# metal 0, eta 0: diffuse for uc below 1 - F, specular above; metal 1: specular only;
# eta > 0 (transmission 1, metal 0): specular below F / (1 + F), dielectric above.
DIFFUSE = dict(metal=0.0, eta=0.0, uc=0.01)
SPECULAR = dict(metal=1.0, eta=0.0, uc=0.5)
DIELECTRIC = dict(metal=0.0, eta=1.5, uc=0.999)


@pytest.mark.parametrize("lobe,sel", [(1, DIFFUSE), (2, SPECULAR), (3, DIELECTRIC)])
@pytest.mark.parametrize("rough", [0.1, 0.45, 0.75])
def test_sampled_f_and_pdf_match_evaluation(oracle, lobe, sel, rough):
    """BxDFTest DielectricBxDF.f / .pdf (BxDFTest.cpp:227-272): the f and pdf sample_f returns
    equal the lobe's f(wo, wi) and pdf(wo, wi), written separately (BxDF.h:257-286, 343-364,
    497-551), at the sampled wi. The evaluation recomputes the microfacet normal from wi, so
    agreement is to rounding, not bit for bit (the reference test allows 1e-2 on f and 1e-1
    on pdf, absolute)."""
    albedo = [0.8, 0.5, 0.2]
    wo = _unit([-0.3, 0.5, 1.0])
    n = 0
    for i in range(1, 65):
        u = [sel["uc"], _halton(i, 2), _halton(i, 3), _halton(i, 5)]
        got, ok, f, pdf, wi, ct = oracle.principled_sample(wo, albedo, rough, sel["metal"], sel["eta"], u)
        assert got == lobe
        if not ok:
            continue
        n += 1
        # sample_f returns wi normalized in the (identity) shading frame
        el, ef, epdf = oracle.principled_eval(wo, wi, albedo, rough, sel["metal"], sel["eta"], sel["uc"])
        assert el == lobe
        # at roughness 0.1 (alpha 0.01) D moves by ~0.3% for the rounding of a recomputed
        # normal; a dropped or wrong term is off by O(1)
        np.testing.assert_allclose(ef, f, rtol=1e-2, atol=1e-7)
        assert epdf == pytest.approx(pdf, rel=1e-2, abs=1e-7)
        assert ct == pytest.approx(abs(wi[2]), rel=1e-6)
    assert n > 40


@pytest.mark.parametrize("lobe,sel", [(1, DIFFUSE), (2, SPECULAR)])
def test_reflection_lobes_are_reciprocal(oracle, lobe, sel):
    """BxDFTest reciprocity (BxDFTest.cpp:118-125): f(wo, wi) = f(wi, wo) for the reflection
    lobes (the lobe selection probability depends on wo only through pdf, not f)."""
    albedo = [0.9, 0.6, 0.3]
    pairs = [([0.5, 0.5, 0.5], [0.0, 0.0, 1.0]), ([0.2, -0.4, 0.8], [-0.6, 0.1, 0.7]), ([0.9, 0.1, 0.3], [0.1, 0.8, 0.4])]
    for rough in (0.2, 0.6, 1.0):
        for a, b in pairs:
            a, b = _unit(a), _unit(b)
            _, fab, _ = oracle.principled_eval(a, b, albedo, rough, sel["metal"], sel["eta"], sel["uc"])
            _, fba, _ = oracle.principled_eval(b, a, albedo, rough, sel["metal"], sel["eta"], sel["uc"])
            assert (fab > 0).all()
            np.testing.assert_allclose(fab, fba, rtol=1e-5)


def test_diffuse_lobe_pdf_closed_form(oracle):
    """BxDFTest DiffuseBxDF.pdf (BxDFTest.cpp:170-182): pdf = |cos(wi)| / pi for the diffuse
    lobe, here times its selection probability 1 - F with F = Schlick(|wo.z|, 0.04); zero
    for wi below the surface."""
    wo = _unit([0.5, 0.5, 0.5])
    F = 0.04 + 0.96 * (1.0 - float(wo[2])) ** 5
    for wi in ([0.0, 0.0, 1.0], [0.3, 0.2, 0.5], [-0.7, 0.1, 0.2]):
        wi = _unit(wi)
        _, _, pdf = oracle.principled_eval(wo, wi, [1, 1, 1], 0.5, 0.0, 0.0, DIFFUSE["uc"])
        assert pdf == pytest.approx((1.0 - F) * float(wi[2]) / math.pi, rel=1e-5)
    _, f, pdf = oracle.principled_eval(wo, _unit([0.0, 0.0, -1.0]), [1, 1, 1], 0.5, 0.0, 0.0, DIFFUSE["uc"])
    assert pdf == 0.0 and not f.any()


@pytest.mark.parametrize("rough", [0.05, 0.3, 0.7, 1.0])
@pytest.mark.parametrize("cos_o", [0.95, 0.6, 0.25])
def test_directional_albedo_does_not_exceed_one(oracle, rough, cos_o):
    """BxDFTest energy conservation (BxDFTest.cpp:127-151, rho = mean f |cos| / pdf): the metal
    GGX lobe with F0 = 1 has albedo E[G / G1(wo)] <= 1, close to 1 when nearly smooth; the
    glass lobe seen from outside moves at most R + T / eta^2 <= 1 of the energy."""
    wo = _unit([math.sqrt(1 - cos_o * cos_o), 0.0, cos_o])
    n = 1024
    rho_m = np.zeros(3)
    rho_g = np.zeros(3)
    for i in range(1, n + 1):
        u = [_halton(i, 2), _halton(i, 3), _halton(i, 5), _halton(i, 7)]
        _, ok, f, pdf, wi, ct = oracle.principled_sample(wo, [1, 1, 1], rough, 1.0, 0.0, u)
        if ok:
            rho_m += f * ct / pdf
        lobe, ok, f, pdf, wi, ct = oracle.principled_sample(wo, [1, 1, 1], rough, 0.0, 1.5, u)
        if ok and lobe == 3:
            rho_g += f * ct / pdf
    rho_m /= n
    rho_g /= n
    assert rho_m.max() <= 1.0 + 1e-3
    if rough <= 0.05:
        assert rho_m.min() > 0.97
    assert 0.0 < rho_g.max() <= 1.0 + 1e-3


def test_unit_index_transmission_goes_straight_through(oracle):
    """BxDFTest DielectricBxDF.sample_f (BxDFTest.cpp:207-225): with eta 1 and roughness 0 the
    dielectric lobe transmits along -wo, |cos| = 1, f = albedo <= 1."""
    wo = np.array([0.0, 0.0, 1.0], dtype=np.float32)
    lobe, ok, f, pdf, wi, ct = oracle.principled_sample(wo, [1, 1, 1], 0.0, 0.0, 1.0, [0.999, 0.33, 0.35, 0.533])
    assert lobe == 3 and ok
    assert ct == pytest.approx(1.0, abs=1e-5)
    np.testing.assert_allclose(wi, [0.0, 0.0, -1.0], atol=1e-6)
    assert (f <= 1.0).all() and (f > 0).all()
