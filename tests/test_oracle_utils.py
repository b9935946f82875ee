"""Pin the numeric utilities (utils.jl) and the portable libm (nbg_libm.h).

* libm: ulp error against libquadmath on dense samples, special values.
* comp_sum (utils.jl:16-23): closed-form cancellation cases.
* cubic1 (utils.jl:103-122): residual of the cubic it solves.
* G3/H1/H2/H3/H5/H6 (utils.jl:124-396): series (gamma < 0.5) against the
  closed forms on both sides of the switch, the identities the reference
  states (H3 = G1 G2 - 3 G3, H5 = H2 - 2 G3, H6 = 2 G2^2 - 3 G1 G3), and the
  binary128 evaluation of the same template.
"""
import numpy as np
import pytest

import oracle_lib as O


@pytest.mark.parametrize("name,lo,hi,bound", [
    ("sin", -20.0, 20.0, 1.0), ("cos", -20.0, 20.0, 1.0), ("sin", -1e5, 1e5, 1.0),
    ("exp", -700.0, 700.0, 1.0), ("exp", -1.0, 1.0, 1.0),
    ("sinh", -30.0, 30.0, 2.0), ("sinh", -1.0, 1.0, 1.0), ("cosh", -30.0, 30.0, 2.0),
    ("cbrt", 0.4, 9.0, 1.0),
])
def test_libm_ulp(name, lo, hi, bound):
    x = np.random.default_rng(7).uniform(lo, hi, 100000)
    _, ulp = O.libm_eval(name, x)
    assert ulp.max() <= bound, (name, ulp.max())


@pytest.mark.parametrize("lo,hi", [(6.6e5, 1e7), (1e7, 1e20), (1e20, 1e100), (1e100, 1.79e308)])
def test_libm_sincos_large_arguments(lo, hi):
    """|x * 2/pi| >= 2^20 takes the Payne-Hanek reduction."""
    x = 10 ** np.random.default_rng(11).uniform(np.log10(lo), np.log10(hi), 40000)
    x[::2] *= -1
    for name in ("sin", "cos"):
        _, ulp = O.libm_eval(name, x)
        assert ulp.max() <= 1.0, (name, lo, ulp.max())


def test_libm_sincos_hardest_reduction():
    """The double closest to a multiple of pi/2 (x mod pi/2 ~ 4.7e-19), the
    largest doubles and the first reductions past the Cody-Waite range."""
    big = np.finfo(float).max
    x = np.array([6381956970095103 * 2.0 ** 797, 2.0 ** 1023, big, -big, 2.0 ** 20 * np.pi / 2, -(2.0 ** 20) * np.pi / 2])
    for name in ("sin", "cos"):
        _, ulp = O.libm_eval(name, x)
        assert ulp.max() <= 1.0, (name, ulp)
    c, _ = O.libm_eval("cos", x[:1])
    assert abs(c[0] + 4.687165924254628e-19) < 1e-33


def test_libm_cbrt_wide_range_and_specials():
    rng = np.random.default_rng(3)
    x = np.concatenate([10 ** rng.uniform(-300, 300, 50000), -(10 ** rng.uniform(-300, 300, 5000)), [5e-324, 2.0 ** -1060]])
    _, ulp = O.libm_eval("cbrt", x)
    assert ulp.max() <= 1.0
    out, _ = O.libm_eval("cbrt", np.array([0.0, -0.0, 8.0, -27.0, np.inf, -np.inf]))
    assert list(out[:2]) == [0.0, -0.0] and np.signbit(out[1])
    assert out[2] == 2.0 and out[3] == -3.0 and out[4] == np.inf and out[5] == -np.inf
    out, _ = O.libm_eval("exp", np.array([0.0, 710.0, -746.0, 1.0]))
    assert out[0] == 1.0 and out[1] == np.inf and out[2] == 0.0
    s, _ = O.libm_eval("sin", np.array([0.0, -0.0, 1e-300]))
    assert s[0] == 0.0 and np.signbit(s[1]) and s[2] == 1e-300
    assert np.isnan(O.libm_eval("sin", np.array([np.inf]))[0][0])


def test_comp_sum_cancellation():
    # Adding 1e-16 ten thousand times to 1.0: plain summation loses it all,
    # the compensated pair keeps value + error = 1 + 1e-12 to rounding.
    sv = np.array([1.0])
    se = np.array([0.0])
    for _ in range(10000):
        sv, se = O.comp_sum(sv, se, np.array([1e-16]))
    assert sv[0] + se[0] == pytest.approx(1.0 + 1e-12, rel=0, abs=1e-18)
    assert abs((sv[0] - 1.0) + se[0] - 1e-12) < 1e-20
    # an exact cancellation restores the original value bit for bit
    sv, se = O.comp_sum(np.array([1.0]), np.array([0.0]), np.array([1e-17]))
    sv, se = O.comp_sum(sv, se, np.array([-1e-17]))
    assert sv[0] == 1.0 and se[0] == 0.0


def test_cubic1_root():
    rng = np.random.default_rng(11)
    a, b, c = rng.normal(size=(3, 20000))
    x = O.utils_eval("cubic1", a, b, c)
    a3 = a / 3
    Q = a3 ** 2 - b / 3
    R = a3 ** 3 + 0.5 * (-a3 * b + c)
    ok = R ** 2 >= Q ** 3  # one real root branch; otherwise the reference returns -c/b
    res = x ** 3 + a * x ** 2 + b * x + c
    scale = np.abs(x) ** 3 + np.abs(a * x ** 2) + np.abs(b * x) + np.abs(c)
    # Cardano cancels when the root is near zero; the result only seeds Newton.
    assert np.max(np.abs(res[ok]) / scale[ok]) < 1e-11
    assert np.array_equal(x[~ok], -c[~ok] / b[~ok])


def _gh_inputs(beta_sign, gamma):
    beta = beta_sign * np.full_like(gamma, 0.37)
    sqb = np.sqrt(np.abs(beta))
    return gamma, beta, sqb


@pytest.mark.parametrize("beta_sign", [1.0, -1.0])
def test_series_meet_closed_forms_at_switch(beta_sign):
    # evaluate the series just above 0.5 where the closed form is used
    g = np.linspace(0.5, 0.7, 41)
    gamma, beta, sqb = _gh_inputs(beta_sign, g)
    for name in ("G3", "H2", "H3", "H5"):
        ser = O.utils_eval(name + "_series", gamma, beta, sqb)
        cf = O.utils_eval(name, gamma, beta, sqb)
        assert np.max(np.abs(ser - cf) / np.abs(cf)) < 1e-12, name
    for name in ("H1", "H6"):
        ser = O.utils_eval(name + "_series", gamma, beta)
        cf = O.utils_eval(name, gamma, beta)
        assert np.max(np.abs(ser - cf) / np.abs(cf)) < 1e-9, name


@pytest.mark.parametrize("beta_sign", [1.0, -1.0])
def test_gh_identities_and_quad(beta_sign):
    g = np.concatenate([np.linspace(0.01, 0.49, 25), np.linspace(0.5, 2.5, 25)])
    gamma, beta, sqb = _gh_inputs(beta_sign, g)
    if beta_sign > 0:
        s, c = np.sin(gamma), np.cos(gamma)
    else:
        s, c = np.sinh(gamma), np.cosh(gamma)
    G1 = s / sqb
    G2 = beta_sign * (1 - c) / np.abs(beta) if beta_sign > 0 else (c - 1) / np.abs(beta)
    G3 = O.utils_eval("G3", gamma, beta, sqb)
    H2 = O.utils_eval("H2", gamma, beta, sqb)
    H3 = O.utils_eval("H3", gamma, beta, sqb)
    H5 = O.utils_eval("H5", gamma, beta, sqb)
    H6 = O.utils_eval("H6", gamma, beta)
    # identities stated in utils.jl:257, :306, :352 (tolerances scale with the
    # cancellation in the right-hand sides at small gamma)
    big = g >= 0.5
    assert np.max(np.abs(H3 - (G1 * G2 - 3 * G3))[big] / np.abs(H3[big])) < 1e-11
    assert np.max(np.abs(H5 - (H2 - 2 * G3))[big] / np.abs(H5[big])) < 1e-11
    assert np.max(np.abs(H6 - (2 * G2 ** 2 - 3 * G1 * G3))[big] / np.abs(H6[big])) < 1e-9
    # double template against the binary128 template
    for name in ("G3", "H2", "H3", "H5"):
        d = O.utils_eval(name, gamma, beta, sqb)
        q = O.utils_eval(name, gamma, beta, sqb, quad=True)
        assert np.max(np.abs(d - q) / np.abs(q)) < 1e-13, name
    for name in ("H1", "H6"):
        d = O.utils_eval(name, gamma, beta)
        q = O.utils_eval(name, gamma, beta, quad=True)
        assert np.max(np.abs(d - q) / np.abs(q)) < 2e-10, name
