"""Block statistics of the driver (host numpy) against closed forms and hand cases."""
import numpy as np
import pytest

from ad_afqmc_prototype_amd.stats import _plateau, blocking_analysis_ratio, jackknife_ratios, reject_outliers


def test_block_size_one_is_the_standard_error_of_the_mean():
    """Equal weights, B = 1: the leave-one-out jackknife variance is exactly
    s^2 / n (mu_i - mu_bar = -(x_i - x_bar) / (n - 1))."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal(400) * 0.3 + 2.0
    st = blocking_analysis_ratio(x, np.ones_like(x), block_grid=[1, 2, 4], print_q=False)
    assert st["block_sizes"][0] == 1
    np.testing.assert_allclose(st["se_curve"][0], x.std(ddof=1) / np.sqrt(x.size), rtol=1e-12)
    np.testing.assert_allclose(st["mu"], x.mean(), rtol=1e-15)


def test_weighted_ratio_and_block_sums():
    """mu is sum(w e) / sum(w); with blocks of B the SE is the B = 1 formula
    applied to the block ratios when the block weights are equal."""
    rng = np.random.default_rng(1)
    e = rng.standard_normal(300)
    w = rng.uniform(0.5, 1.5, 300)
    st = blocking_analysis_ratio(e, w, block_grid=[3], print_q=False)
    np.testing.assert_allclose(st["mu"], (w * e).sum() / w.sum(), rtol=1e-14)
    w_eq = np.ones(300)
    st = blocking_analysis_ratio(e, w_eq, block_grid=[3], print_q=False)
    blocks = e.reshape(100, 3).mean(axis=1)
    np.testing.assert_allclose(st["se_curve"][0], blocks.std(ddof=1) / np.sqrt(100), rtol=1e-12)


def test_correlated_series_error_is_recovered():
    """AR(1) with rho = 0.8: for block sizes well past the correlation time
    (tau ~ 9) and >= 100 blocks the SE curve sits at the asymptotic
    sigma sqrt((1 + rho) / ((1 - rho) n)) to 10%; the selected SE (which may
    fall back to a large block size with ~20 blocks, +-16% noise) is within
    40% and well above the naive SE."""
    rng = np.random.default_rng(2)
    n, rho = 20000, 0.8
    x = np.empty(n)
    x[0] = rng.standard_normal()
    eps = rng.standard_normal(n) * np.sqrt(1 - rho ** 2)
    for i in range(1, n):
        x[i] = rho * x[i - 1] + eps[i]
    st = blocking_analysis_ratio(x, np.ones(n), print_q=False)
    target = np.sqrt((1 + rho) / ((1 - rho) * n))
    sel = (st["block_sizes"] >= 30) & (st["n_blocks"] >= 100)
    assert sel.sum() >= 3
    np.testing.assert_allclose(st["se_curve"][sel], target, rtol=0.10)
    assert abs(st["se_star"] / target - 1) < 0.40, (st["se_star"], target)
    assert st["se_star"] > 2.0 * x.std() / np.sqrt(n)
    assert st["B_star"] > 1


def test_plateau_rules():
    sizes = np.array([1, 2, 4, 8, 16, 32, 64])
    counts = np.array([640, 320, 160, 80, 40, 20, 10])
    # rises by > 20 % and is then flat within 3 % for k = 3 steps: first such point
    ses = np.array([1.0, 1.1, 1.3, 1.31, 1.32, 1.33, 1.5])
    assert _plateau(sizes, ses, counts) == (4, 1.3, 160)
    # never flat: the first size within 5 % of the largest SE (counts >= 20 only)
    ses = np.array([1.0, 1.2, 1.5, 1.9, 2.4, 3.0, 9.0])
    assert _plateau(sizes, ses, counts)[0] == 32
    # nothing with enough blocks: the first entry
    assert _plateau(sizes, ses, np.full(7, 5)) == (1, 1.0, 5)


def test_estimator_scale_samples_have_the_se_as_spread():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(2000)
    st = blocking_analysis_ratio(x, np.ones_like(x), print_q=False)
    smp = st["estimator_scale_samples"]
    np.testing.assert_allclose(smp.mean(), st["mu"], atol=1e-12)
    # (G-1)/sqrt(G) * LOO deviations: population spread = SE * sqrt(G / (G - 1))... ~ SE
    assert abs(smp.std() / st["se_star"] - 1) < 0.05


def test_jackknife_ratios_constant_denominator():
    """denom == 1: the ratio is the mean and sigma = s / sqrt(n)."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal(500) + 1.0
    m, s = jackknife_ratios(x, np.ones_like(x))
    np.testing.assert_allclose(m, x.mean(), rtol=1e-13)
    np.testing.assert_allclose(s, x.std(ddof=1) / np.sqrt(x.size), rtol=1e-12)


def test_reject_outliers():
    rng = np.random.default_rng(5)
    e = rng.standard_normal(200) * 0.01 - 10.0
    e[[7, 50]] = [5.0, -40.0]
    data = np.column_stack((e, np.ones(200)))
    clean, mask = reject_outliers(data, obs=0)
    assert clean.shape == (198, 2) and not mask[7] and not mask[50] and mask.sum() == 198
    # a constant column keeps everything through the dispersion floor
    data = np.column_stack((np.full(10, 3.0), np.arange(10.0)))
    clean, mask = reject_outliers(data, obs=0)
    assert mask.all()


def test_few_samples_fall_back_to_block_size_one():
    """Fewer than 10 samples leave no block size with >= 5 blocks (the
    reference raises there): the B = 1 jackknife, NaN for a single sample."""
    x = np.array([1.0, 2.0, 4.0])
    st = blocking_analysis_ratio(x, np.ones(3), print_q=False)
    assert st["B_star"] == 1
    np.testing.assert_allclose(st["se_star"], x.std(ddof=1) / np.sqrt(3), rtol=1e-13)
    st = blocking_analysis_ratio(np.array([1.0]), np.ones(1), print_q=False)
    assert np.isnan(st["se_star"]) and st["mu"] == 1.0
