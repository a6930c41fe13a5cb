"""FP64 host model math of the native library, pinned by closed forms,
invariants, scipy and the oracle (no GPU needed)."""

import math

import numpy as np
import pytest
import scipy.linalg
import scipy.special

from crux_amd import _native as nat
from oracle import gamma as og
from oracle.oracle import p_matrix, q_decomp as oq_decomp, q_matrix


def _rng():
    return np.random.default_rng(7)


def test_jc_closed_form():
    V, Vi, w, pin, qn = nat.q_decomp([1.0] * 6, [0.25] * 4)
    assert qn == pytest.approx(4.0 / 3.0)
    for t in (0.0, 0.1, 0.5, 2.0):
        P = nat.pt(V, Vi, w, t * qn)
        same = 0.25 + 0.75 * math.exp(-4.0 * t / 3.0)
        diff = 0.25 - 0.25 * math.exp(-4.0 * t / 3.0)
        assert np.allclose(np.diag(P), same, atol=1e-14)
        off = P[~np.eye(4, dtype=bool)]
        assert np.allclose(off, diff, atol=1e-14)


def test_k80_closed_form():
    kappa = 3.0
    # rclass [0,1,0,0,1,0]: transitions AG and CT
    rtri = [1.0, kappa, 1.0, 1.0, kappa, 1.0]
    V, Vi, w, pin, qn = nat.q_decomp(rtri, [0.25] * 4)
    t = 0.3
    P = nat.pt(V, Vi, w, t * qn)
    # K80 with branch length t in substitutions per site
    a = math.exp(-4.0 * t / (kappa + 2.0))
    b = math.exp(-2.0 * t * (kappa + 1.0) / (kappa + 2.0))
    p_ts = 0.25 + 0.25 * a - 0.5 * b
    p_tv = 0.25 - 0.25 * a
    assert P[0, 2] == pytest.approx(p_ts, abs=1e-13)
    assert P[1, 3] == pytest.approx(p_ts, abs=1e-13)
    assert P[0, 1] == pytest.approx(p_tv, abs=1e-13)
    assert P[0, 0] == pytest.approx(1 - p_ts - 2 * p_tv, abs=1e-13)


def test_f81_closed_form():
    pi = np.array([0.4, 0.3, 0.2, 0.1])
    V, Vi, w, pin, qn = nat.q_decomp([1.0] * 6, pi)
    beta = 1.0 / (1.0 - (pi ** 2).sum())
    assert qn == pytest.approx(beta)
    t = 0.7
    P = nat.pt(V, Vi, w, t * qn)
    e = math.exp(-beta * t)
    expect = np.outer(np.ones(4), pi) * (1 - e) + np.eye(4) * e
    assert np.allclose(P, expect, atol=1e-13)


@pytest.mark.parametrize("n", [4, 20, 22])
def test_gtr_matches_expm_and_oracle(n):
    rng = _rng()
    rtri = rng.uniform(0.5, 2.0, n * (n - 1) // 2) if n == 4 else rng.lognormal(0, 1, n * (n - 1) // 2)
    pi = rng.dirichlet(np.full(n, 5.0 if n == 4 else 1.0))
    V, Vi, w, pin, qn = nat.q_decomp(rtri, pi)
    Q, opin, oqn = q_matrix(rtri, pi)
    assert np.allclose(pin, opin, rtol=0, atol=1e-15)
    assert qn == pytest.approx(oqn, rel=1e-13)
    assert np.allclose(V @ np.diag(w) @ Vi, Q, atol=1e-12)
    cube, vals, _, _ = oq_decomp(rtri, pi)
    for t in (0.0, 0.05, 0.4, 3.0):
        P = nat.pt(V, Vi, w, t)
        E = scipy.linalg.expm(Q * t)
        assert np.allclose(P, np.maximum(E, 0), atol=1e-12)
        assert np.allclose(P, p_matrix(cube, vals, t), atol=1e-12)
        assert np.allclose(P.sum(1), 1.0, atol=1e-12)
        # detailed balance pi_i P_ij = pi_j P_ji
        D = pin[:, None] * P
        assert np.allclose(D, D.T, atol=1e-13)
    # Chapman-Kolmogorov
    assert np.allclose(nat.pt(V, Vi, w, 0.3) @ nat.pt(V, Vi, w, 0.5), nat.pt(V, Vi, w, 0.8),
                       atol=1e-12)


def test_zero_rate_model_is_identity():
    V, Vi, w, pin, qn = nat.q_decomp([0.0] * 6, [0.25] * 4)
    assert math.isinf(qn)
    assert np.allclose(nat.pt(V, Vi, w, 1.7), np.eye(4), atol=1e-15)


@pytest.mark.parametrize("median", [False, True])
@pytest.mark.parametrize("alpha", [0.05, 0.179062, 0.5, 1.0, 5.0, 300.0])
@pytest.mark.parametrize("ncat", [2, 4, 8])
def test_gamma_rates(alpha, ncat, median):
    got = nat.gamma_rates(alpha, ncat, median)
    ora = og.gamma_cmult(alpha, ncat, median)
    assert np.allclose(got, ora, rtol=1e-12, atol=1e-14)
    assert np.mean(got) == pytest.approx(1.0, rel=1e-9)
    # Yang (1994) from scipy: the textbook construction.
    if not median:
        b = scipy.special.gammaincinv(alpha, np.arange(1, ncat) / ncat) / alpha
        cum = np.append(scipy.special.gammainc(alpha + 1, b * alpha), 1.0)
        ref = np.diff(np.concatenate([[0.0], cum])) * ncat
    else:
        m = scipy.special.gammaincinv(alpha, (2 * np.arange(ncat) + 1) / (2 * ncat)) / alpha
        ref = m * ncat / m.sum()
    assert np.allclose(got, ref, rtol=2e-6, atol=1e-9)


def test_gamma_helpers_against_scipy():
    for x in (0.1, 0.5, 1.0, 3.3, 7.0, 25.0):
        assert og.ln_gamma(x) == pytest.approx(math.lgamma(x), abs=1e-10)
    for p in (1e-10, 0.01, 0.3, 0.5, 0.8, 0.999):
        assert og.pt_norm(p) == pytest.approx(scipy.special.ndtri(p), abs=1e-12)
    for a, x in ((0.5, 0.2), (2.0, 3.0), (10.0, 4.0), (1.5, 12.0)):
        assert og.gamma_inc(x, a, math.lgamma(a)) == pytest.approx(
            scipy.special.gammainc(a, x), abs=1e-9)
    for p, v in ((0.25, 1.0), (0.5, 4.0), (0.75, 0.6), (0.9, 10.0), (0.1, 0.2)):
        q = og.pt_chi2(p, v, math.lgamma(v / 2))
        assert q == pytest.approx(scipy.special.gammaincinv(v / 2, p) * 2, rel=2e-6)


def test_bad_arguments_raise():
    with pytest.raises(ValueError):
        nat.q_decomp([1.0] * 6, [0.25, 0.25, 0.5, 0.0])
    with pytest.raises(ValueError):
        nat.gamma_rates(-1.0, 4)


@pytest.mark.parametrize("n", [4, 20])
def test_oracle_symmetric_decomposition_fallback(n):
    # The oracle's fallback for a failed general eigensolve: the same P(t)
    # as scipy's matrix exponential, rows summing to 1.
    from oracle.oracle import q_decomp_symmetric
    rng = np.random.default_rng(11 + n)
    rtri = rng.lognormal(0, 1, n * (n - 1) // 2)
    pi = rng.dirichlet(np.ones(n))
    Q, pin, _ = q_matrix(rtri, pi)
    w, V, Vi = q_decomp_symmetric(Q, pin)
    assert np.allclose(V @ Vi, np.eye(n), atol=1e-12)
    for t in (0.05, 0.7):
        P = V @ np.diag(np.exp(w * t)) @ Vi
        assert np.allclose(P, scipy.linalg.expm(Q * t), atol=1e-12)
