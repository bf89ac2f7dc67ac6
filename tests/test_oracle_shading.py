"""Pins for the oracle's shading restatement (src/sampling.cpp, src/image.cpp:615-659)
that do not retype its formulas: symmetry, closed-form invariants and
distribution tests of the samplers."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O

f32 = np.float32
RNG = np.random.default_rng(123)


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def params(N, V, base=(0.5, 0.4, 0.3), rough=0.5, metal=0.0, opacity=1.0, eta=1.0, entering=1.0, shapeN=None):
    shapeN = N if shapeN is None else shapeN
    return np.array([*N, *shapeN, *V, *base, rough, metal, opacity, eta, entering], dtype=np.float32)


def brdf(p, L):
    out = np.empty(3, np.float32)
    O.lib().oracle_bsdf_eval(O.fp(p), O.fp(np.asarray(L, np.float32)), 0, O.fp(out))
    return out


def hemi(n):
    v = RNG.normal(size=3)
    v = v / np.linalg.norm(v)
    return unit(v if np.dot(v, n) > 0 else -v)


def test_brdf_reciprocity():
    # getBRDF returns f(L,V) * N.L with a Disney-style f symmetric in L and V
    for _ in range(300):
        N = unit(RNG.normal(size=3))
        V, L = hemi(N), hemi(N)
        rough = float(RNG.uniform(0.05, 1.0))
        metal = float(RNG.uniform(0, 1))
        base = RNG.uniform(0.05, 0.95, 3)
        a = brdf(params(N, V, base, rough, metal), L) / np.dot(N, L)
        b = brdf(params(N, L, base, rough, metal), V) / np.dot(N, V)
        assert np.allclose(a, b, rtol=2e-4, atol=1e-6)


def test_brdf_zero_below_horizon_and_nonnegative():
    N = unit([0, 0, 1])
    V = unit([0.3, 0.1, 0.9])
    assert (brdf(params(N, V), unit([0.2, 0.1, -0.5])) == 0).all()
    for _ in range(200):
        L = hemi(N)
        assert (brdf(params(N, V, rough=float(RNG.uniform(0.01, 1))), L) >= 0).all()


def test_brdf_lambert_limit():
    # metallic 0: diffuse lobe is base/pi * Fd with Fd -> 1 at normal incidence,
    # plus specular + clearcoat; with V = L = N the diffuse part is exactly base/pi
    N = unit([0, 0, 1])
    p = params(N, N, base=(0.5, 0.5, 0.5), rough=0.5)
    p_black = params(N, N, base=(0.0, 0.0, 0.0), rough=0.5)
    diff = brdf(p, N) - brdf(p_black, N)
    # the black base also changes Cspec0 only through metallic (0 here): the
    # difference is the diffuse lobe (Fd = 1 at NdotL = NdotV = 1, LdotH = 1 -> Fd90 = 1.5,
    # schlick(1) = 0 -> Fd = 1)
    assert np.allclose(diff, 0.5 / np.pi, rtol=1e-5)


def accumulate(base, bsdf, light, w):
    d = np.zeros(4, np.float32)
    s = np.zeros(4, np.float32)
    O.lib().oracle_accumulate(O.fp(np.asarray(base, np.float32)), O.fp(np.asarray(bsdf, np.float32)),
                              O.fp(np.asarray(light, np.float32)), f32(w), O.fp(d), O.fp(s))
    return d, s


def test_accumulate_demodulation_invariant():
    # base * d + s == w * L * bsdf, and d == w * L * B / |base| where B is the
    # coefficient of base/|base| when bsdf's projection onto span{white, base}
    # is written in that (non-orthogonal) basis
    for _ in range(300):
        base = RNG.uniform(0.05, 1.0, 3).astype(np.float32)
        bsdf = RNG.uniform(0.0, 2.0, 3).astype(np.float32)
        L = RNG.uniform(0.1, 5.0, 3).astype(np.float32)
        w = f32(RNG.uniform(0.1, 1.0))
        d, s = accumulate(base, bsdf, L, w)
        b0 = base / np.linalg.norm(base)
        wh = np.ones(3) / np.sqrt(3)
        if np.dot(b0, wh) > 0.99:
            assert np.allclose(base * d[:3], w * L * bsdf, rtol=1e-4)
            assert (s == 0).all()
            continue
        Bm = np.stack([wh, b0], 1)
        coef, *_ = np.linalg.lstsq(Bm, bsdf.astype(np.float64), rcond=None)
        assert np.allclose(base * d[:3] + s[:3], w * L * bsdf, rtol=1e-3, atol=1e-5)
        B = coef[1]
        assert np.allclose(d[:3], w * L * B / np.linalg.norm(base), rtol=2e-3, atol=1e-4)
        # Var parts: w * |inradiance|^2 per buffer
        assert d[3] >= 0 and s[3] >= 0


def test_accumulate_special_cases():
    # black base -> everything specular; grey base -> everything diffuse;
    # tiny light -> dropped; NaN light -> dropped
    d, s = accumulate([0, 0, 0], [1, 2, 3], [1, 1, 1], 0.5)
    assert (d == 0).all() and np.allclose(s[:3], [0.5, 1.0, 1.5])
    d, s = accumulate([0.5, 0.5, 0.5], [1, 2, 3], [1, 1, 1], 1.0)
    assert (s == 0).all() and np.allclose(d[:3], [2, 4, 6])
    d, s = accumulate([0.5, 0.2, 0.1], [1, 1, 1], [1e-5, 0, 0], 1.0)
    assert (d == 0).all() and (s == 0).all()
    d, s = accumulate([0.5, 0.2, 0.1], [1, 1, 1], [np.nan] * 3, 1.0)
    assert (d == 0).all() and (s == 0).all()


def sample(p, which, n, seed=1):
    out = np.empty((n, 8), np.float32)
    for i in range(n):
        O.lib().oracle_bsdf_sample(O.fp(p), which, seed, i, 0, O.fp(out[i]))
    return out


def chi2_uniform(counts, expected):
    m = expected > 5
    return float((((counts - expected) ** 2)[m] / expected[m]).sum()), int(m.sum()) - 1


def test_cosine_sampler_distribution():
    # sampleCos draws cos(theta) ~ sqrt(U): P(cos < c) = c^2
    N = unit([0, 0, 1])
    p = params(N, unit([0.2, 0.3, 0.9]), rough=0.8)
    s = sample(p, 0, 20000)
    ok = np.isfinite(s[:, 0])
    cos = s[ok, 2]
    bins = np.linspace(0, 1, 21)
    counts, _ = np.histogram(cos ** 2, bins)
    chi2, dof = chi2_uniform(counts, np.full(20, ok.sum() / 20))
    assert chi2 < dof + 5 * np.sqrt(2 * dof)
    # pdf returned = cos / pi
    assert np.allclose(s[ok, 6], cos / np.pi, rtol=1e-5)


def test_ggx_sampler_halfvector_distribution():
    # the GGX half vector has cos(theta_h) with CDF  F(c) = (1 - c^2)... inverted
    # from u: cos^2 = (1-u) / (1 + (a^2-1) u)  ->  u = (1 - c^2) / (1 + (a^2-1) c^2) * ... ;
    # check via the inverse: u(c) is uniform
    N = unit([0, 0, 1])
    # near-normal incidence (V == N exactly gives the reference a NaN tangent
    # frame: getTangentSpaceWithInDir normalizes cross(N, V)); H = normalize(V + L)
    V = unit([0.01, 0.0, 1.0])
    a = 0.4
    p = params(N, V, rough=a)
    s = sample(p, 1, 20000)
    ok = np.isfinite(s[:, 0])
    L = s[ok, :3].astype(np.float64)
    H = (L + V) / np.linalg.norm(L + V, axis=1, keepdims=True)
    c2 = H[:, 2] ** 2
    u = (1 - c2) / (1 + (a * a - 1) * c2)
    counts, _ = np.histogram(u, np.linspace(0, 1, 21))
    # samples with L below the surface are rejected, so the top of u is thinned:
    # test only the part that can never be rejected (u < 0.5 here)
    chi2, dof = chi2_uniform(counts[:10], np.full(10, counts[:10].sum() / 10))
    assert chi2 < dof + 5 * np.sqrt(2 * dof)


def test_btdf_sampler_refracts_through_surface():
    N = unit([0, 0, 1])
    V = unit([0.3, 0.0, 0.95])
    p = params(N, V, rough=0.05, opacity=0.0, eta=1 / 1.5, entering=1.0)
    s = sample(p, 3, 2000)
    ok = np.isfinite(s[:, 0])
    assert ok.mean() > 0.9
    assert (s[ok, 2] < 0).all()          # below the surface
    # near-specular: Snell's law on average: sin_t = eta * sin_i
    sin_i = np.sqrt(1 - V[2] ** 2)
    sin_t = np.sqrt(1 - s[ok, 2].astype(np.float64) ** 2)
    assert abs(np.median(sin_t) - sin_i / 1.5) < 0.02
