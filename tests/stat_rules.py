"""The acceptance rule of every Monte Carlo comparison in these tests.

Each compared quantity (a tally cell, or a summary of a run) gives one
z-score: a difference of means divided by its standard error. The error
comes either from independent replicas on both sides (Welch's t, turned into
the normal z with the same two-sided tail probability, so that few replicas
do not inflate the tails) or from an exactly known variance.

Agreement is asked "within 3 sigma". With K compared quantities a correct
implementation still puts about K * 0.27 % of them beyond 3 sigma, so a
family of z-scores passes when both hold:

(a) the number of |z| > 3 is at most the (1 - alpha) quantile of
    Binomial(K, 2 Phi(-3));
(b) max |z| is at most the Bonferroni bound Phi^-1(1 - alpha / (2 K)).

(b) holds at family level alpha whatever the correlation between the
quantities; (a) is exact for independent ones and conservative (it fails
more often than alpha) for positively correlated ones, such as the cells of
one run. alpha = 1e-3 throughout.
"""
import numpy as np
from scipy import stats

P3 = 2.0 * stats.norm.sf(3.0)


def welch_z(a, b):
    """a (Ra, ...), b (Rb, ...): replicas -> normal-equivalent z of the
    difference of means (Welch's t with Welch-Satterthwaite degrees of
    freedom, mapped through its two-sided tail probability)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    na, nb = a.shape[0], b.shape[0]
    va = a.var(0, ddof=1) / na
    vb = b.var(0, ddof=1) / nb
    d = a.mean(0) - b.mean(0)
    s2 = va + vb
    with np.errstate(divide="ignore", invalid="ignore"):
        t = d / np.sqrt(s2)
        df = s2 ** 2 / (va ** 2 / (na - 1) + vb ** 2 / (nb - 1))
        z = np.sign(t) * stats.norm.isf(stats.t.sf(np.abs(t), df))
    # equal constants on both sides agree; different constants do not
    z = np.where(s2 > 0, z, np.where(d == 0, 0.0, np.inf))
    return z


def exact_z(x, mean, sigma):
    """z of an estimate against a known mean and standard deviation"""
    x, mean, sigma = np.broadcast_arrays(*(np.asarray(v, float) for v in (x, mean, sigma)))
    with np.errstate(divide="ignore", invalid="ignore"):
        z = (x - mean) / sigma
    return np.where(sigma > 0, z, np.where(x == mean, 0.0, np.inf))


def three_sigma_verdict(z, alpha=1e-3):
    """(ok, message) for a family of z-scores under rules (a) and (b); z is
    an array or a list of arrays and scalars (one family)"""
    if isinstance(z, (list, tuple)):
        z = np.concatenate([np.ravel(np.asarray(v, float)) for v in z])
    z = np.abs(np.ravel(np.asarray(z, float)))
    K = z.size
    n3 = int((z > 3.0).sum())
    kmax = int(stats.binom.ppf(1.0 - alpha, K, P3))
    zmax = float(stats.norm.isf(alpha / (2.0 * K)))
    worst = float(z.max()) if K else 0.0
    ok = n3 <= kmax and worst <= zmax
    msg = f"{K} compared, {n3} beyond 3 sigma (at most {kmax}), max |z| {worst:.2f} (at most {zmax:.2f})"
    return ok, msg


def assert_3sigma(z, what="", alpha=1e-3):
    ok, msg = three_sigma_verdict(z, alpha)
    assert ok, f"{what}: {msg}"
    return msg
