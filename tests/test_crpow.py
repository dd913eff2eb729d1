"""The correctly rounded pow of the exact-order kernels (canary_pow_cr, the
same source the device compiles) and the oracle's (quad-precision powq
rounded once), each against mpmath at 250 bits, and against each other.

Correct rounding is what lets the device and the oracle agree bit for bit:
two correctly rounded functions of the same (x, y) return the same double
whatever their method.  The draws cover the pressure's own arguments
(x = R rho theta / p0 near 1, y = cp/cv), wide ranges of x, large |y|,
results near the overflow and underflow limits of the claim, and x within
a few ulp of 1 (where log x cancels)."""
import ctypes as C

import mpmath as mp
import numpy as np
import pytest

from canary_amd._lib import lib
from oracle import oracle as O

mp.mp.prec = 250
KAPPA = 100467 / 100 / (7175 / 10)  # cp / cv as the kernels form it


def cr(x, y):
    return float(mp.power(mp.mpf(x), mp.mpf(y)))


def draws(seed, n):
    rng = np.random.default_rng(seed)
    out = {
        "pressure": [(x, KAPPA) for x in rng.uniform(0.85, 1.05, n)],
        "wide_x": [(x, KAPPA) for x in np.exp(rng.uniform(-30, 30, n))],
        "mid_y": list(zip(np.exp(rng.uniform(-5, 5, n)), rng.uniform(-60, 60, n))),
        "big_y": list(zip(rng.uniform(0.5, 2.0, n), rng.uniform(-900, 900, n))),
        "near_limits": list(zip(np.exp(rng.uniform(-690, 690, n)), rng.uniform(-1, 1, n))),
        "near_one": [(1.0 + d, y) for d, y in
                     zip(rng.uniform(-1, 1, n) * 2.0 ** rng.uniform(-52, -6, n),
                         rng.uniform(-3, 3, n))],
    }
    out["near_one"] += [(float(np.nextafter(1.0, 2.0 * s) if k == 0 else 1.0 + s * 2.0 ** -k),
                         KAPPA) for s in (-1, 1) for k in range(0, 53)]
    return out


@pytest.mark.parametrize("category", ["pressure", "wide_x", "mid_y", "big_y", "near_limits",
                                      "near_one"])
def test_both_pows_correctly_rounded(category):
    L = lib()
    cases = draws(2024, 700)[category]
    bad_lib, bad_oracle = [], []
    for x, y in cases:
        x, y = float(x), float(y)
        want = cr(x, y)
        if L.canary_pow_cr(x, y) != want:
            bad_lib.append((x, y))
        if O.opow(x, y) != want:
            bad_oracle.append((x, y))
    assert not bad_lib, bad_lib[:5]
    assert not bad_oracle, bad_oracle[:5]


@pytest.mark.parametrize("kappa", [KAPPA, 1.4, 1.0, 2.0, 1.9999, 5.0 / 3.0])
def test_fixed_exponent_table_path(kappa):
    """The kernels' cheaper path for a gas's kappa (r^-kappa per cell times
    the binomial series of (1 + t)^kappa): the same correctly rounded bits,
    over [1/sqrt2, sqrt2) where it applies and outside it (general path)."""
    L = lib()
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(0.7071, 1.4143, 1500), 1.0 + rng.uniform(-1, 1, 300)
                         * 2.0 ** rng.uniform(-52, -6, 300), np.exp(rng.uniform(-6, 6, 200))])
    bad = [x for x in xs if L.canary_pow_cr_table(float(x), kappa) != cr(float(x), kappa)]
    assert not bad, bad[:5]


def test_library_and_oracle_agree_on_many_pressures():
    """2e5 pressure arguments of a stratified atmosphere and 1e5 generic
    ones: the two independent implementations return the same bits."""
    L = lib()
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(0.9, 1.0, 200000), np.exp(rng.uniform(-8, 8, 100000))])
    f = L.canary_pow_cr
    diff = [x for x in xs if f(float(x), KAPPA) != O.opow(float(x), KAPPA)]
    assert not diff, diff[:5]


def test_special_values():
    L = lib()
    f = L.canary_pow_cr
    assert f(1.0, KAPPA) == 1.0
    assert f(1.0, -123.5) == 1.0
    assert f(4.0, 0.5) == 2.0            # exact result
    assert f(2.0, 10.0) == 1024.0        # exact result
    assert f(0.25, -1.5) == 8.0          # exact result
    sub = 5e-324
    assert f(sub, 0.001) == cr(sub, 0.001)     # subnormal x
    # outside the claim: the C library's pow
    for x, y in ((0.0, KAPPA), (-2.0, 3.0), (float("inf"), 2.0), (2.0, 2000.0), (1e-300, 3.0)):
        with np.errstate(all="ignore"):
            a, b = f(x, y), np.power(x, y)
        assert (a == b) or (np.isnan(a) and np.isnan(b)), (x, y, a, b)


def test_a_planted_one_ulp_error_is_caught():
    """The bar is the exact double: a pow one ulp off in one case out of
    many fails the comparison above (mpmath, not the implementation under
    test, decides)."""
    x = 0.97531
    good = cr(x, KAPPA)
    off = float(np.nextafter(good, 2.0))
    assert off != good and lib().canary_pow_cr(x, KAPPA) == good
