"""ROCK2's cubic coefficient, pinned by what the mathematics fixes.

The examples' ROCK2(eigen_est=eigen_estimator) (examples/single_borehole_basic.jl:133)
takes its stage count from the degree rule mdeg = floor(sqrt((1.5 + dt lambda)
/ 0.811)) + 1 with lambda from eigen_estimator_pre_calculation
(src/solver.jl:474-488).  While max(mdeg, 3) - 2 == 1 -- i.e. dt lambda <
9 * 0.811 - 1.5 = 5.799 -- it applies its 3-stage member (one orthogonal
stage and the two-stage finish), R(z) = (1 + mu z)(1 + 2 sigma z + tau z^2) =
1 + z + z^2/2 + c3 z^3, second order by construction; c3 = mu tau comes from
coefficient tables of OrdinaryDiffEqStabilizedRK that are not in this image.
The degree rule exists to pick a member that is stable for the estimate, so
the member it keeps must satisfy |R(-x)| <= 1 on [0, 5.799].  That fixes

  R(-x) <= 1   for all x <= 5.799  <=>  c3 >= max_x (x/2 - 1)/x^2 = 1/16 (x = 4)
  R(-x) >= -1  for all x <= 5.799  <=>  c3 <= (2 - x + x^2/2)/x^3 at x = 5.799
                                        (decreasing in x) = 0.0667408...

The library's default c3 = 1/16 lies in this interval (at its lower end);
the GPU runs at both ends (tests/test_gpu_rock2_spread.py) bound what the
unpinned value can change.
"""
import numpy as np
import pytest

import helpers as H

X_MAX = 9 * 0.811 - 1.5


def R(x, c3):
    """R(-x) of the 3-stage member."""
    return 1 - x + x * x / 2 - c3 * x ** 3


def stable_on(c3, x_max, n=200001):
    x = np.linspace(0.0, x_max, n)
    return bool(np.all(np.abs(R(x, c3)) <= 1 + 1e-12))


def interval():
    return 1.0 / 16.0, (2 - X_MAX + X_MAX ** 2 / 2) / X_MAX ** 3


def test_degree_rule_threshold_matches_oracle():
    """The oracle's restatement of the degree rule switches exactly at 5.799."""
    w = H.small_case()
    orc = H.oracle_for(w)
    lam = orc.eigen_estimate
    assert orc.rock2_minimum_degree(0.999999 * X_MAX / lam)
    assert not orc.rock2_minimum_degree(1.000001 * X_MAX / lam)
    assert X_MAX == pytest.approx(5.799, abs=1e-12)


def test_admissible_interval_closed_form():
    lo, hi = interval()
    assert lo == 0.0625
    assert hi == pytest.approx(0.0667408, abs=1e-7)
    # every c3 inside is stable on the whole interval, just outside it is not
    for c3 in np.linspace(lo, hi, 9):
        assert stable_on(c3, X_MAX), c3
    assert not stable_on(lo - 1e-4, X_MAX)
    assert not stable_on(hi + 1e-4, X_MAX)
    # the binding points: tangency R(-4) = 1 at c3 = 1/16, R(-5.799) = -1 at hi
    assert R(4.0, lo) == pytest.approx(1.0, abs=1e-15)
    assert R(X_MAX, hi) == pytest.approx(-1.0, abs=1e-12)


def test_library_default_inside_interval():
    from geothermalwells_jl_amd import cache
    lo, hi = interval()
    assert lo <= cache.DEFAULT_ROCK2_C3 <= hi
    # the staged coefficients of both ends exist (real sigma) and reproduce c3
    O = H.O()
    for c3 in (lo, hi):
        mu, sigma, tau = O.rock2_coef(c3)
        assert mu * tau == pytest.approx(c3, abs=1e-15)
        assert mu + 2 * sigma == pytest.approx(1.0, abs=1e-15)
