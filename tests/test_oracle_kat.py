"""Pins the CPU oracle with known answers that do not come from its own code.

Each check names its source: a published table or constant, a closed form
that the mathematics fixes, or an invariant (normalisation, energy
conservation, moments of a sampling distribution) chosen so that a dropped
term, a wrong sign or a wrong constant fails it.
"""

import math

import numpy as np
import pytest

import oracle

# ---------------------------------------------------------------- RNG ----


def _inverse_lowbias32(y: int) -> int:
    """Inverse of lowbias32 (xorshift-multiply steps are bijections)."""

    def unxorshift(x, s):
        r = x
        for _ in range(32 // s + 1):
            r = x ^ (r >> s)
        return r & 0xFFFFFFFF

    inv1 = pow(0x846CA68B, -1, 1 << 32)
    inv2 = pow(0x7FEB352D, -1, 1 << 32)
    x = unxorshift(y, 16)
    x = (x * inv1) & 0xFFFFFFFF
    x = unxorshift(x, 15)
    x = (x * inv2) & 0xFFFFFFFF
    return unxorshift(x, 16)


def test_hash_is_a_bijection_with_known_inverse():
    # lowbias32 (C. Wellons, "hash prospector"): each step is invertible, so
    # hash(inverse(y)) == y for every y; zero is a fixed point.
    assert oracle.hash_state(0) == 0
    rng = np.random.default_rng(0)
    for y in rng.integers(0, 1 << 32, size=200, dtype=np.uint64):
        assert oracle.hash_state(_inverse_lowbias32(int(y))) == int(y)


def test_hash_avalanche():
    # Flipping one input bit flips ~16 of 32 output bits on average.
    rng = np.random.default_rng(1)
    flips = []
    for x in rng.integers(0, 1 << 32, size=300, dtype=np.uint64):
        for b in range(0, 32, 4):
            flips.append(bin(oracle.hash_state(int(x)) ^ oracle.hash_state(int(x) ^ (1 << b))).count("1"))
    assert abs(np.mean(flips) - 16.0) < 0.5


def test_random_float_rounds_to_one_at_the_top():
    # float(uint) rounds to nearest: hash outputs >= 2^32 - 128 give exactly
    # 1.0 (the reference's randomFloat can return 1.0).
    f, st = oracle.random_float(_inverse_lowbias32(0xFFFFFFC0))
    assert st == 0xFFFFFFC0 and f == 1.0
    f, st = oracle.random_float(_inverse_lowbias32(0xFFFFFF7F))
    assert f < 1.0
    f, _ = oracle.random_float(_inverse_lowbias32(0))
    assert f == 0.0


# ------------------------------------------------------------ spectral ----

# CIE 1931 2-degree standard observer (CIE 015, 5 nm table).
CIE_TABLE = {
    440: (0.34828, 0.02300, 1.74706),
    450: (0.33620, 0.03800, 1.77211),
    500: (0.00490, 0.32300, 0.27200),
    550: (0.43345, 0.99500, 0.00875),
    555: (0.51205, 1.00000, 0.00575),
    580: (0.91630, 0.87000, 0.00165),
    600: (1.06220, 0.63100, 0.00080),
    650: (0.28350, 0.10700, 0.00000),
}


@pytest.mark.parametrize("w", sorted(CIE_TABLE))
def test_cie_fit_tracks_the_cie_1931_table(w):
    # Wyman, Sloan, Shirley (JCGT 2013) multi-lobe fit: within 0.03 of the
    # tabulated colour-matching functions.
    got = oracle.cie_xyz(w)
    assert np.max(np.abs(got - np.array(CIE_TABLE[w]))) < 0.03


def test_sensor_matrix_is_the_srgb_inverse():
    # Recover M from sensor(w) = M @ xyz(w) * 400/106.856895 and check it
    # inverts the IEC 61966-2-1 linear-sRGB -> XYZ matrix.
    ws = np.linspace(390, 770, 97)
    X = np.array([oracle.cie_xyz(w) * (400.0 / 106.856895) for w in ws])
    S = np.array([oracle.sensor(w) for w in ws])
    M = np.linalg.lstsq(X, S, rcond=None)[0].T
    srgb_to_xyz = np.array([[0.4124564, 0.3575761, 0.1804375],
                            [0.2126729, 0.7151522, 0.0721750],
                            [0.0193339, 0.1191920, 0.9503041]])
    assert np.allclose(M @ srgb_to_xyz, np.eye(3), atol=2e-3)


def _integrate(f, lo=380.0, hi=780.0, n=40001):
    w = np.linspace(lo, hi, n)
    v = np.array([f(x) for x in w])
    return np.trapezoid(v, w, axis=0)


def test_cie_fit_integrals_and_sensor_scale():
    # Integrals of the fit over 380-780 nm: 106.714 / 106.944 / 106.806; a
    # flat unit spectrum therefore has Y = 106.944 / 106.856895.
    xyz = _integrate(oracle.cie_xyz, n=8001)
    assert np.allclose(xyz, [106.714, 106.944, 106.806], atol=0.01)
    y_flat = _integrate(lambda w: oracle.sensor(w), n=8001) / 400.0
    # Y row of the sRGB->XYZ matrix applied to the sensor's RGB.
    Y = 0.2126729 * y_flat[0] + 0.7151522 * y_flat[1] + 0.0721750 * y_flat[2]
    assert abs(Y - 106.944 / 106.856895) < 2e-4


def test_white_light_integrates_to_the_6504k_white_point():
    # White RGB under the 6504 K illuminant spectrum integrates to linear
    # sRGB ~ (1.0346, 0.9796, 1.0289), not (1, 1, 1).
    k = _integrate(lambda w: oracle.sensor(w) * oracle.illuminant([1, 1, 1], w), n=8001) / 400.0
    assert np.allclose(k, [1.0346, 0.9796, 1.0289], atol=2e-3)


def test_illuminant_is_planck_normalised_at_560nm():
    assert abs(oracle.illuminant([1, 1, 1], 560.0) - 1.0) < 1e-12
    c2, T = 1.438776877e7, 6504.0  # CODATA second radiation constant, nm K

    def planck(w):
        return w ** -5 / (math.exp(c2 / (w * T)) - 1.0)

    for w in (420.0, 500.0, 650.0, 700.0):
        expected = oracle.reconstruct_rgb([1, 1, 1], w) * planck(w) / planck(560.0)
        assert abs(oracle.illuminant([1, 1, 1], w) / expected - 1.0) < 1e-6


def test_reconstruct_rgb_closed_forms():
    def g(w, c, s):
        return math.exp(-0.5 * ((w - c) / s) ** 2)

    for w in (400.0, 500.0, 620.0, 700.0):
        b = (g(w, 620, 30), g(w, 535, 25), g(w, 450, 20))
        assert abs(oracle.reconstruct_rgb([1, 0, 0], w) - b[0] / sum(b)) < 1e-12
        assert abs(oracle.reconstruct_rgb([0.3, 0.3, 0.3], w) - 0.3) < 1e-12
    # Equal RGB is flat until the 1e-5 floor of the basis sum takes over.
    b780 = g(780, 620, 30) + g(780, 535, 25) + g(780, 450, 20)
    assert b780 < 1e-5
    assert abs(oracle.reconstruct_rgb([1, 1, 1], 780.0) - b780 / 1e-5) < 1e-12
    assert oracle.reconstruct_rgb([-1, -1, -1], 550.0) == 0.0


def test_dispersed_ior_is_a_cauchy_fit_with_the_abbe_number():
    nd, abbe = 1.5168, 64.17
    assert oracle.dispersed_ior(nd, 1.0, abbe, 587.56) == pytest.approx(nd, abs=1e-12)
    nf, nc = oracle.dispersed_ior(nd, 1.0, abbe, 486.13), oracle.dispersed_ior(nd, 1.0, abbe, 656.27)
    assert (nd - 1.0) / (nf - nc) == pytest.approx(abbe, rel=1e-9)
    assert oracle.dispersed_ior(nd, 0.0, abbe, 450.0) == nd
    half = oracle.dispersed_ior(nd, 0.5, abbe, 450.0)
    assert half == pytest.approx(0.5 * (nd + oracle.dispersed_ior(nd, 1.0, abbe, 450.0)), abs=1e-12)


# ---------------------------------------------------------------- BSDFs ----


def _hemisphere_grid(n_theta=400, n_phi=96, power=1.0):
    s = (np.arange(n_theta) + 0.5) / n_theta
    theta = 0.5 * math.pi * s ** power
    dtheta = 0.5 * math.pi * power * s ** (power - 1) / n_theta
    phi = (np.arange(n_phi) + 0.5) / n_phi * 2 * math.pi
    dirs, weights = [], []
    for t, dt in zip(theta, dtheta):
        for p in phi:
            dirs.append((math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)))
            weights.append(math.sin(t) * dt * 2 * math.pi / n_phi)
    return np.array(dirs), np.array(weights)


DIRS, WEIGHTS = _hemisphere_grid()


def _view(mu):
    return [math.sqrt(1 - mu * mu), 0.0, mu]


def test_lambert_white_furnace():
    v = _view(0.6)
    total = sum(w * oracle.eval_diffuse(v, l, 1.0, 0.0, 0) for l, w in zip(DIRS, WEIGHTS))
    assert total == pytest.approx(1.0, abs=2e-3)


def test_oren_nayar_reduces_to_lambert_at_zero_roughness():
    rng = np.random.default_rng(2)
    for _ in range(20):
        l = rng.normal(size=3)
        l[2] = abs(l[2])
        l /= np.linalg.norm(l)
        v = _view(rng.uniform(0.1, 1.0))
        assert oracle.eval_diffuse(v, l, 0.5, 0.0, 1) == pytest.approx(l[2] / math.pi, rel=1e-12)


def test_burley_is_lambert_when_the_grazing_factor_is_one():
    # F_D90 = 0.5 + 2 r (l.h)^2 = 1 at r = 0.25 for v = l = n.
    assert oracle.eval_diffuse([0, 0, 1], [0, 0, 1], 0.5, 0.25, 2) == pytest.approx(1.0 / math.pi, rel=1e-12)


def test_eon_constants():
    # EON (Portsmouth et al.): c1 = 1/2 - 2/(3 pi) exactly. The average
    # albedo constant c2 is 2/3 - 28/(15 pi) = 0.0724882 by the closed form
    # (and by quadrature of E_FON below); the reference's literal
    # 0.0726760455 (src/gpu/vulkan_path_tracer.cpp:989) is 0.26% off and is
    # kept as is by the restatement.
    assert abs(0.2877934092 - (0.5 - 2 / (3 * math.pi))) < 1e-10
    exact_c2 = 2 / 3 - 28 / (15 * math.pi)
    assert abs(0.0726760455 / exact_c2 - 1.0) < 3e-3
    mu = (np.arange(4000) + 0.5) / 4000
    for s in (0.5, 1.0):
        e_avg = 2 * np.mean([oracle.eon_albedo(m, s) * m for m in mu])
        implied = (e_avg * (1 + 0.2877934092 * s) - 1) / s
        assert abs(implied - exact_c2) < 1e-3


@pytest.mark.parametrize("sigma", [0.3, 1.0])
@pytest.mark.parametrize("mu", [0.25, 0.8])
def test_eon_albedo_is_the_exact_single_scatter_albedo(sigma, mu):
    # With rho = 0 only the single-scatter FON lobe remains; its integral
    # over the hemisphere must equal the closed-form albedo E_FON(mu).
    v = _view(mu)
    total = sum(w * oracle.eval_diffuse(v, l, 0.0, sigma, 3) for l, w in zip(DIRS, WEIGHTS))
    assert total == pytest.approx(oracle.eon_albedo(mu, sigma), abs=3e-3)


@pytest.mark.parametrize("sigma", [0.3, 1.0])
def test_eon_conserves_energy_for_white_albedo(sigma):
    v = _view(0.5)
    total = sum(w * oracle.eval_diffuse(v, l, 1.0, sigma, 3) for l, w in zip(DIRS, WEIGHTS))
    assert total == pytest.approx(1.0, abs=0.02)


def test_ggx_reflection_pdf_closed_form_at_normal_incidence():
    # For v = n, pdf(l) = D(h) (n.h) / (4 v.h) integrates over l above the
    # horizon to P(theta_h < 45 deg) = 1 / (1 + alpha^2) (isotropic GGX).
    dirs, weights = _hemisphere_grid(1200, 48, power=2.0)
    for r in (0.3, 0.5):
        alpha = r * r
        total = sum(w * oracle.ggx_pdf([0, 0, 1], l, r, r) for l, w in zip(dirs, weights))
        assert total == pytest.approx(1.0 / (1.0 + alpha * alpha), abs=3e-3)


def test_ggx_sampler_follows_its_pdf():
    # Mean sample weight (f cos / pdf) equals the directional albedo by
    # quadrature, plus the mirror fallback mass P(theta_h > 45) w(mirror).
    r, f0 = 0.45, 0.8
    v = [0.0, 0.0, 1.0]
    dirs, weights = _hemisphere_grid(1200, 48, power=2.0)
    albedo = sum(w * oracle.ggx_eval(v, l, r, r, f0) * l[2] for l, w in zip(dirs, weights))
    alpha = r * r
    p_fallback = alpha * alpha / (1.0 + alpha * alpha)
    _, w_mirror, _ = oracle.ggx_sample([0, 0, -1], 1e-4, 1e-4, f0, 7)
    state = 12345
    ws = []
    for _ in range(60000):
        _, w, state = oracle.ggx_sample([0, 0, -1], r, r, f0, state)
        ws.append(w)
    ws = np.array(ws)
    expected = albedo + p_fallback * oracle.ggx_eval(v, [0, 0, 1], r, r, f0) * 1.0 / max(
        oracle.ggx_pdf(v, [0, 0, 1], r, r), 1e-6)
    assert abs(ws.mean() - expected) < 4 * ws.std() / math.sqrt(len(ws)) + 2e-3
    assert w_mirror > 0


def test_ggx_sample_weight_is_eval_cos_over_pdf():
    state = 99
    for _ in range(200):
        inc = np.array([0.3, -0.2, -0.9])
        inc /= np.linalg.norm(inc)
        l, w, state = oracle.ggx_sample(inc, 0.4, 0.6, 0.5, state)
        v = -inc
        pdf = oracle.ggx_pdf(v, l, 0.4, 0.6)
        if pdf > 1e-3 and l[2] > 1e-3:
            assert w == pytest.approx(oracle.ggx_eval(v, l, 0.4, 0.6, 0.5) * l[2] / pdf, rel=1e-9)


@pytest.mark.parametrize("mu", [0.2, 0.6, 0.95])
@pytest.mark.parametrize("rough", [0.2, 0.5, 0.9])
def test_conty_kulla_sheen_albedo_fit_matches_its_brdf(mu, rough):
    # The fitted albedo approximates the integral of the Imageworks sheen
    # BRDF x cos it was fitted to.
    v = _view(mu)
    total = sum(w * oracle.sheen_response(v, l, rough, False) for l, w in zip(DIRS, WEIGHTS))
    assert oracle.sheen_albedo(mu, rough, False) == pytest.approx(total, abs=0.03)


def test_zeltner_sheen_response_integrates_to_its_albedo():
    v = _view(0.7)
    total = sum(w * oracle.sheen_response(v, l, 0.4, True) for l, w in zip(DIRS, WEIGHTS))
    assert total == pytest.approx(oracle.sheen_albedo(0.7, 0.4, True), abs=2e-3)
    for mu in np.linspace(0, 1, 11):
        for r in (0.01, 0.3, 1.0):
            assert 0.0 <= oracle.sheen_albedo(mu, r, True) <= 1.0


@pytest.mark.parametrize("g", [-0.6, 0.0, 0.3, 0.8])
def test_henyey_greenstein_mean_cosine_is_g(g):
    fwd = np.array([0.2, 0.5, -0.7])
    fwd /= np.linalg.norm(fwd)
    state = 5
    cos = []
    for _ in range(40000):
        d, state = oracle.sample_hg(fwd, g, state)
        cos.append(float(np.dot(d, fwd)))
    cos = np.array(cos)
    assert abs(cos.mean() - g) < 4 * cos.std() / math.sqrt(len(cos))


def test_cosine_hemisphere_moments():
    n = np.array([0.0, 0.6, 0.8])
    state = 3
    c = []
    for _ in range(40000):
        d, state = oracle.cosine_hemisphere(n, state)
        c.append(float(np.dot(d, n)))
    c = np.array(c)
    assert c.min() >= 0.0
    assert abs(c.mean() - 2 / 3) < 0.01
    assert abs((c * c).mean() - 0.5) < 0.01


def test_lat_long_round_trip():
    rng = np.random.default_rng(4)
    for u, v in rng.uniform(0.01, 0.99, size=(100, 2)):
        d = oracle.lat_long_direction(u, v)
        assert np.allclose(oracle.dome_uv(d, 1), [u, v], atol=1e-9)
    assert np.allclose(oracle.dome_uv([0, -1, 0], 1)[1], 0.0)  # row 0 = nadir
