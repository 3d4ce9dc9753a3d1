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
