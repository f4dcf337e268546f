"""Oracle pins for the sampling layer: Halton (halton.h:11-87), the reproducible sin/cos,
UniformSampleSphere (sampling.h:38-47) and the AWR antenna pattern (emitter.h:109-137)."""
from fractions import Fraction

import numpy as np
import pytest


def exact_radical_inverse(i: int, b: int) -> Fraction:
    r, f = Fraction(0), Fraction(1, b)
    while i:
        r += (i % b) * f
        f /= b
        i //= b
    return r


def test_halton_known_values(oracle):
    assert oracle.halton(0, 8, 0).tolist() == [0.0, 0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875]
    b3 = oracle.halton(0, 4, 1)
    assert b3[0] == 0.0 and b3[1] == np.float32(1 / 3) and b3[2] == np.float32(2 / 3)
    assert oracle.halton(0, 6, 2)[1:].tolist() == pytest.approx([0.2, 0.4, 0.6, 0.8, 0.04], abs=1e-7)


def test_halton_base2_is_correctly_rounded(oracle):
    # the reversed-digit integer times 2^-n: exact scaling, one rounding of the integer
    idx = np.concatenate([np.arange(0, 4096), np.random.default_rng(1).integers(0, 1 << 40, 2048)]).astype(np.uint64)
    L = oracle.lib()
    for i in idx:
        got = L.orc_radical_inverse(int(i), 2)
        want = np.float32(float(exact_radical_inverse(int(i), 2)))
        want = min(want, np.nextafter(np.float32(1), np.float32(0)))
        assert got == want, (int(i), got, want)


@pytest.mark.parametrize("base", [3, 5, 7])
def test_halton_matches_definition(oracle, base):
    rng = np.random.default_rng(base)
    idx = np.concatenate([np.arange(0, 500), rng.integers(0, 1 << 27, 500)])
    L = oracle.lib()
    for i in idx:
        got = L.orc_radical_inverse(int(i), base)
        want = float(exact_radical_inverse(int(i), base))
        # float(reversed) * invBase^n with invBase and invBase^n rounded: <= (digits + 2) ulp
        assert abs(got - want) <= 64 * np.finfo(np.float32).eps * max(want, 1e-30) + 1e-30, (int(i), got, want)


def test_halton_stratification(oracle):
    # the first 3^7 points of base 3 fall one per stratum [j/3^7, (j+1)/3^7)
    n = 3 ** 7
    v = oracle.halton(0, n, 1).astype(np.float64)
    strata = np.floor(v * n + 1e-6).astype(int)
    assert sorted(strata.tolist()) == list(range(n))


def test_sincos_f32_matches_correct_rounding(oracle):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(0, 2 * np.pi, 100000), rng.uniform(-40, 40, 20000),
                        np.float32(np.pi / 2) * np.arange(-8, 9)]).astype(np.float32)
    s, c = oracle.sincos_f32(x)
    np.testing.assert_array_equal(s, np.sin(x.astype(np.float64)).astype(np.float32))
    np.testing.assert_array_equal(c, np.cos(x.astype(np.float64)).astype(np.float32))


@pytest.mark.parametrize("prec", [0, 2])
def test_uniform_sphere_closed_form(oracle, prec):
    u = np.stack([oracle.halton(1, 20000, 2), oracle.halton(1, 20000, 1)], axis=1)
    d = oracle.sphere_sample(u, prec)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-7)
    # y = 2 u1 - 1 and phi = 2 pi u0 with d = (-r sin phi, y, -r cos phi)
    np.testing.assert_allclose(d[:, 1], 2.0 * u[:, 1].astype(np.float64) - 1.0, atol=1e-7)
    phi = np.mod(np.arctan2(-d[:, 0], -d[:, 2]), 2 * np.pi)
    dphi = np.angle(np.exp(1j * (phi - 2 * np.pi * u[:, 0].astype(np.float64))))
    r = np.sqrt(np.maximum(0, 1 - d[:, 1] ** 2))
    assert np.all(np.abs(dphi[r > 1e-3]) < 5e-6)
    # uniform on the sphere: E[d] = 0, E[d_k^2] = 1/3 (Halton QMC, small error)
    np.testing.assert_allclose(d.mean(0), 0.0, atol=2e-3)
    np.testing.assert_allclose((d ** 2).mean(0), 1 / 3, atol=2e-3)


def test_awr_pattern_closed_form(oracle):
    # boresight: cos_h = cos_e = 1 -> H = (2.622/0.8^6 * 0.625/0.5^4, 0, 0)
    H = oracle.awr_evaluate(np.array([[0.0, 0.0, 1.0]]), prec=2)[0]
    np.testing.assert_allclose(H.real, [2.622 / 0.8 ** 6 * 0.625 / 0.5 ** 4, 0, 0], rtol=1e-12)
    np.testing.assert_array_equal(H.imag, 0.0)
    # along +y the field e_y x d vanishes
    assert np.all(oracle.awr_evaluate(np.array([[0.0, 1.0, 0.0]]))[0] == 0)


def test_awr_pattern_invariants(oracle):
    rng = np.random.default_rng(3)
    d = rng.normal(size=(5000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    H = oracle.awr_evaluate(d, prec=2)
    # H is orthogonal to d (AssertOrthogonal in Ray::setH, geometry.h:123-126) and real
    np.testing.assert_allclose(np.einsum("ij,ij->i", H.real, d), 0.0, atol=1e-9 * np.abs(H).max())
    # odd in d: H(-d) = -H(d); lobe symmetric in the sign of dx and dy separately
    np.testing.assert_allclose(oracle.awr_evaluate(-d, prec=2), -H, rtol=1e-12, atol=0)
    m = d * np.array([-1.0, -1.0, 1.0])
    Hm = oracle.awr_evaluate(m, prec=2)
    np.testing.assert_allclose(np.abs(Hm), np.abs(H), rtol=1e-12)
    # FP32 instance agrees with FP64 to float precision
    H32 = oracle.awr_evaluate(d, prec=0)
    np.testing.assert_allclose(H32, H, rtol=2e-6, atol=2e-6 * np.abs(H).max())
