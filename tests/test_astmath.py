"""The project's elementary functions (astmath.h) against mpmath, and the
random stream (astrng.h) against the published Philox4x32-10 known-answer
vectors (Random123 kat_vectors)."""
import os
import subprocess
import sys

import mpmath as mp
import numpy as np
import pytest

import oracle

mp.mp.dps = 40


def _ulps(got, ref_mp):
    ref = np.array([float(r) for r in ref_mp])
    err = np.abs(got - ref) / np.spacing(np.abs(ref))
    return err


CASES = [
    ("sin", lambda r: np.concatenate([r.uniform(-4, 4, 3000), r.uniform(-2e4, 2e4, 3000)]), mp.sin, 1.0),
    ("cos", lambda r: np.concatenate([r.uniform(-4, 4, 3000), r.uniform(-2e4, 2e4, 3000)]), mp.cos, 1.0),
    ("tan", lambda r: r.uniform(-1.5, 1.5, 4000), mp.tan, 3.0),
    ("log", lambda r: np.exp(r.uniform(-700, 700, 4000)), mp.log, 1.0),
    ("log10", lambda r: np.exp(r.uniform(-700, 700, 4000)), mp.log10, 1.5),
    ("exp", lambda r: r.uniform(-700, 700, 4000), mp.exp, 1.0),
    ("asin", lambda r: r.uniform(-1, 1, 4000), mp.asin, 3.0),
    ("acos", lambda r: r.uniform(-1, 1, 4000), mp.acos, 3.0),
    ("atan", lambda r: r.normal(scale=10, size=4000), mp.atan, 2.0),
]


@pytest.mark.parametrize("name,gen,ref,tol", CASES, ids=[c[0] for c in CASES])
def test_accuracy_vs_mpmath(name, gen, ref, tol):
    x = gen(np.random.default_rng(1))
    got = oracle.math(name, x)
    err = _ulps(got, [ref(mp.mpf(float(v))) for v in x])
    assert err.max() <= tol, (name, err.max())


def test_atan2_all_quadrants():
    r = np.random.default_rng(2)
    y, x = r.normal(size=4000), r.normal(size=4000)
    got = oracle.math("atan2", y, x)
    err = _ulps(got, [mp.atan2(mp.mpf(float(a)), mp.mpf(float(b))) for a, b in zip(y, x)])
    assert err.max() <= 2.0


def test_special_values():
    assert np.isnan(oracle.math("sin", np.array([np.inf, np.nan]))).all()
    lg = oracle.math("log", np.array([0.0, -1.0, np.inf, 5e-324, 1.0]))
    assert lg[0] == -np.inf and np.isnan(lg[1]) and lg[2] == np.inf and lg[4] == 0.0
    assert lg[3] == np.log(5e-324)  # subnormal input
    ex = oracle.math("exp", np.array([-746.0, 710.0, 0.0, -740.0]))
    assert ex[0] == 0.0 and ex[1] == np.inf and ex[2] == 1.0 and ex[3] == np.exp(-740.0)
    at = oracle.math("atan2", np.array([0.0, -0.0, 1.0, 0.0]), np.array([-1.0, -1.0, 0.0, 0.0]))
    assert at[0] == np.pi and at[1] == -np.pi and at[2] == np.pi / 2 and at[3] == 0.0
    assert np.isnan(oracle.math("asin", np.array([1.5]))).all()


def test_expf_is_rounded_double_exp():
    x = np.random.default_rng(3).uniform(-1.2, 0.0, 1000).astype(np.float32).astype(np.float64)
    got = oracle.math("expf", x)
    want = np.array([np.float32(float(mp.exp(mp.mpf(v)))) for v in x], np.float64)
    assert np.max(np.abs(got - want) / np.spacing(np.float32(want))) <= 1.0


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds.
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_stream_layout():
    seed, fit = 0x1234567890ABCDEF, 77
    u = oracle.uniform(seed, fit, 0, 8)
    assert np.all(u > 0) and np.all(u <= 1)
    # draw k is word k%4 of Philox(counter=(k/4, 0, fit_lo, fit_hi), key=(seed_lo, seed_hi))
    words = oracle.philox([0, 0, fit, 0], [seed & 0xFFFFFFFF, seed >> 32]) + \
        oracle.philox([1, 0, fit, 0], [seed & 0xFFFFFFFF, seed >> 32])
    want = np.array([((w >> 8) + 1) * 2.0 ** -24 for w in words], np.float32)
    assert np.array_equal(u, want)


def test_normal_deviates():
    """normal01: Box-Muller on draws k, k+1 of the stream: the transform of
    the stream's own uniforms, and standard normal in distribution."""
    from scipy import stats
    seed, fit = 99, 5
    z = oracle.normal(seed, fit, 0, 4)
    u = oracle.uniform(seed, fit, 0, 8).astype(np.float64)
    want = np.sqrt(-2 * np.log(u[0::2])) * np.cos(2 * np.pi * u[1::2])
    np.testing.assert_allclose(z, want.astype(np.float32), rtol=2e-7, atol=1e-7)
    big = np.concatenate([oracle.normal(seed, f, 10, 2000) for f in range(50)]).astype(np.float64)
    assert abs(big.mean()) < 0.02 and abs(big.std() - 1) < 0.02
    assert stats.kstest(big, "norm").pvalue > 1e-3


def test_coefficients_regenerate():
    here = os.path.join(os.path.dirname(__file__), "..", "asteroid_amd", "csrc")
    out = subprocess.run([sys.executable, os.path.join(here, "gen_math_coeffs.py")], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    with open(os.path.join(here, "astmath_coeffs.h")) as f:
        assert f.read().strip() == out.strip()



@pytest.mark.parametrize("name,ref", [("sin_t", mp.sin), ("cos_t", mp.cos)])
def test_table_driven_sincos(name, ref):
    """The evaluator's table-driven sin/cos (Tang's method, 1024-entry table of
    k pi/512) within 1 ulp of mpmath on small and large arguments (the Euler
    angles reach ~1e4 rad after days of fast rotation), at the table points
    and half-way between them (where the reduction's |r| is largest and k's
    rounding decides), NaN for non-finite arguments."""
    rng = np.random.default_rng(21)
    x = np.concatenate([rng.uniform(-4, 4, 3000), rng.uniform(-2e4, 2e4, 3000), rng.uniform(-1e-3, 1e-3, 500)])
    got = oracle.math(name, x)
    err = _ulps(got, [ref(mp.mpf(float(v))) for v in x])
    assert err.max() <= 1.0, err.max()
    k = np.arange(-3000, 3000)
    pts = k * (np.pi / 512)  # the rounded k pi/512: within 1 ulp of the true values there too
    err = _ulps(oracle.math(name, pts), [ref(mp.mpf(float(v))) for v in pts])
    assert err.max() <= 1.0
    half = (k + 0.5) * (np.pi / 512)
    pts = np.concatenate([half, np.nextafter(half, np.inf), np.nextafter(half, -np.inf)])
    err = _ulps(oracle.math(name, pts), [ref(mp.mpf(float(v))) for v in pts])
    assert err.max() <= 1.0
    assert np.isnan(oracle.math(name, np.array([np.inf, -np.inf, np.nan]))).all()


def test_powf_is_rounded_double_pow():
    """NUDGE's powf: exp(y log x) in double rounded to float, within one float
    ulp of the exact power over the merit's range (x in [1e-40, 1], y in
    (0, 1]); 1 at x = 1, 0 at x = 0."""
    r = np.random.default_rng(4)
    x = np.exp(r.uniform(np.log(1e-38), 0.0, 2000)).astype(np.float32).astype(np.float64)
    y = r.uniform(0.05, 1.0, 2000).astype(np.float32).astype(np.float64)
    got = oracle.math("powf", x, y)
    want = np.array([float(mp.power(mp.mpf(a), mp.mpf(b))) for a, b in zip(x, y)])
    assert np.all(got == got.astype(np.float32))
    assert np.max(np.abs(got - want) / np.spacing(np.float32(want))) <= 1.0
    assert oracle.math("powf", np.array([1.0, 0.0]), np.array([0.3, 0.5])).tolist() == [1.0, 0.0]
