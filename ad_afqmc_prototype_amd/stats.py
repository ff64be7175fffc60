"""Block statistics for the driver: weighted ratio estimator with a blocking
(leave-one-block-out) standard error and a plateau rule for the block size,
and median/IQR outlier rejection.  Same estimator and selection rule as
stat_utils.py:6-184 of the reference (host-side numpy; no device work).
"""
from __future__ import annotations

import numpy as np


def _plateau(sizes, ses, counts, *, min_blocks=20, min_rise=0.20, flat_tol=0.03, k=3):
    """Pick the block size where the SE curve has risen by min_rise over its
    first value and then stays flat (relative steps <= flat_tol) for k steps;
    otherwise the first size within 5% of the maximum SE."""
    sizes, ses, counts = (np.asarray(a) for a in (sizes, ses, counts))
    keep = counts >= min_blocks
    if not np.any(keep):
        return int(sizes[0]), float(ses[0]), int(counts[0])
    b, s, c = sizes[keep], ses[keep], counts[keep]
    for i in range(0, b.size - k):
        if s[i] < (1.0 + min_rise) * s[0]:
            continue
        win = s[i:i + k + 1]
        if np.all(np.abs(np.diff(win)) <= flat_tol * win[:-1]):
            return int(b[i]), float(s[i]), int(c[i])
    j = int(np.nonzero(s >= 0.95 * s.max())[0][0])
    return int(b[j]), float(s[j]), int(c[j])


def blocking_analysis_ratio(ene, wt, block_grid=None, *, min_blocks=20, min_rise=0.20,
                            flat_tol=0.03, k=3, print_q=True, exact=None):
    """mu = sum(w e) / sum(w) with its blocking standard error."""
    ene = np.asarray(ene, float).ravel()
    wt = np.asarray(wt, float).ravel()
    n = ene.size
    if wt.size != n:
        raise ValueError("energy and weight series differ in length")
    num = wt * ene
    mu = num.sum() / wt.sum()
    if block_grid is None:
        cand = np.unique(np.rint(np.geomspace(1, max(2, n // min_blocks), 18)).astype(int))
        block_grid = [int(b) for b in cand if b >= 1 and n // b >= min_blocks]
        if n // cand[-1] >= 5 and cand[-1] not in block_grid:
            block_grid.append(int(cand[-1]))
    sizes, ses, counts, loo = [], [], [], {}
    if not any(n // bsz >= 5 for bsz in block_grid):
        # too few samples for any block size (the reference raises IndexError
        # here): the block-size-1 jackknife, NaN below two samples
        block_grid = [1]
    for bsz in block_grid:
        nb = n // bsz
        if nb < 5 and not (bsz == 1 and len(block_grid) == 1 and nb >= 2):
            continue
        used = nb * bsz
        sn = num[:used].reshape(nb, bsz).sum(axis=1)
        sd = wt[:used].reshape(nb, bsz).sum(axis=1)
        tn, td = sn.sum(), sd.sum()
        rest = td - sd
        ok = np.abs(rest) > 1e-18
        mu_i = np.where(ok, (tn - sn) / np.where(ok, rest, 1.0), tn / td)
        mbar = mu_i.mean()
        var = (nb - 1) / nb * np.sum((mu_i - mbar) ** 2)
        sizes.append(bsz)
        ses.append(float(np.sqrt(max(var, 0.0))))
        counts.append(nb)
        loo[bsz] = (mu_i, mbar, nb)
    if not sizes:
        return {"mu": float(mu), "block_sizes": np.zeros(0, int), "se_curve": np.zeros(0),
                "n_blocks": np.zeros(0, int), "B_star": 1, "se_star": float("nan"),
                "ci95_star": (float("nan"), float("nan")), "estimator_scale_samples": np.zeros(0),
                "bias": None, "z_score": None}
    b_star, se_star, _ = _plateau(sizes, ses, counts, min_blocks=min_blocks,
                                  min_rise=min_rise, flat_tol=flat_tol, k=k)
    mu_i, mbar, nb = loo[b_star]
    out = {"mu": float(mu), "block_sizes": np.array(sizes, int), "se_curve": np.array(ses),
           "n_blocks": np.array(counts, int), "B_star": b_star, "se_star": se_star,
           "ci95_star": (float(mu - 1.96 * se_star), float(mu + 1.96 * se_star)),
           # leave-one-block-out estimates rescaled to the spread of the estimator
           "estimator_scale_samples": mu + (nb - 1) / np.sqrt(nb) * (mu_i - mbar),
           "bias": None, "z_score": None}
    if exact is not None and np.isfinite(se_star) and se_star > 0:
        out["bias"] = float(mu - exact)
        out["z_score"] = float((mu - exact) / se_star)
    if print_q:
        print(f"mu: {out['mu']:.16g}  SE*: {out['se_star']:.16g}  95% CI: {out['ci95_star']}")
        print("\nBlocking SE curve (ratio LOO):")
        print(f" {'B':>6s} {'G':>6s} {'SE':>14s} {'SE/SE(B=1)':>12s}")
        se0 = ses[0] if ses else float("nan")
        for bsz, nb, se in zip(sizes, counts, ses):
            mark = "*" if bsz == b_star else " "
            print(f"{mark}{bsz:6d} {nb:6d} {se:14.6e} {se / se0 if se0 > 0 else float('nan'):12.3f}")
        print("")
    return out


def reject_outliers(data, obs, m=10.0, min_threshold=1e-5):
    """Rows whose column ``obs`` lies more than m robust dispersions from the
    median; dispersion = max(MAD, IQR/1.349, min_threshold)."""
    col = data[:, obs]
    med = np.median(col)
    dev = np.abs(col - med)
    q1, q3 = np.percentile(col, [25, 75])
    disp = max(np.median(dev), (q3 - q1) / 1.349, min_threshold)
    mask = dev / disp < m
    return data[mask], mask


def jackknife_ratios(num, denom):
    """Jackknife mean and standard deviation of mean(num) / mean(denom)
    (stat_utils.py:187-213): leave-one-out ratios r_i, sigma = sqrt((n-1) var(r))."""
    num = np.asarray(num)
    denom = np.asarray(denom)
    n = num.size
    tot_n, tot_d = num.sum(), denom.sum()
    r = ((tot_n - num) / (n - 1) / ((tot_d - denom) / (n - 1))).real
    return np.mean(r), np.sqrt((n - 1) * np.var(r))
