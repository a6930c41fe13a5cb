"""Discrete-Gamma rate categories, FP64, for the oracle.  TEST INFRASTRUCTURE
ONLY (imported by tests/, smoke() and bench.py's cpu_baseline leg).

Restates the published algorithms the reference uses
(lib/Crux/CxMath.pxd:125-368) and the category construction of
pkg/Crux/Tree/Lik.pyx:1212-1265:

* ln_gamma: Pike & Hill (1966), CACM Algorithm 291 -- shift the argument up to
  7 with a running product, then Stirling's series with 4 terms.
* gamma_inc: Bhattacharjee (1970), AS 32 -- series expansion for x <= 1 or
  x < p, otherwise the continued fraction with rescaling on overflow;
  accuracy 1e-10.
* pt_norm: Wichura (1988), AS 241 (PPND16) -- rational approximations in
  three regions.
* pt_chi2: Best & Roberts (1975), AS 91, with Shea's (1991) remark AS R85 --
  starting approximation, then a 7-term Taylor refinement, tolerance 5e-7.
"""

import math

__all__ = ["ln_gamma", "gamma_inc", "pt_norm", "pt_chi2", "gamma_cmult"]


def ln_gamma(x):
    assert x > 0.0
    if x < 7.0:
        f = 1.0
        z = x
        while z < 7.0:
            f *= z
            z += 1.0
        x = z
        f = -math.log(f)
    else:
        f = 0.0
    z = 1.0 / (x * x)
    series = (((-0.000595238095238 * z + 0.000793650793651) * z
               - 0.002777777777778) * z + 0.083333333333333) / x
    return f + (x - 0.5) * math.log(x) - x + 0.918938533204673 + series


def gamma_inc(x, p, ln_gamma_p):
    """Regularised lower incomplete gamma P(p, x)."""
    assert p > 0.0 and x >= 0.0
    if x == 0.0:
        return 0.0
    acu = 1.0e-10
    oflo = 1.0e30
    factor = math.exp(p * math.log(x) - x - ln_gamma_p)
    if x <= 1.0 or x < p:
        gin = 1.0
        term = 1.0
        rn = p
        while True:
            rn += 1.0
            term *= x / rn
            gin += term
            if term <= acu:
                return gin * factor / p
    a = 1.0 - p
    b = a + x + 1.0
    term = 0.0
    pn = [1.0, x, x + 1.0, x * b, 0.0, 0.0]
    gin = pn[2] / pn[3]
    while True:
        a += 1.0
        b += 2.0
        term += 1.0
        an = a * term
        pn[4] = b * pn[2] - an * pn[0]
        pn[5] = b * pn[3] - an * pn[1]
        if pn[5] != 0.0:
            rn = pn[4] / pn[5]
            dif = abs(gin - rn)
            if dif <= acu and dif <= acu * rn:
                return 1.0 - factor * gin
            gin = rn
        pn[0], pn[1], pn[2], pn[3] = pn[2], pn[3], pn[4], pn[5]
        if abs(pn[4]) >= oflo:
            for i in range(4):
                pn[i] /= oflo


_A = (3.3871328727963666080e0, 1.3314166789178437745e2,
      1.9715909503065514427e3, 1.3731693765509461125e4,
      4.5921953931549871457e4, 6.7265770927008700853e4,
      3.3430575583588128105e4, 2.5090809287301226727e3)
_B = (1.0, 4.2313330701600911252e1, 6.8718700749205790830e2,
      5.3941960214247511077e3, 2.1213794301586595867e4,
      3.9307895800092710610e4, 2.8729085735721942674e4,
      5.2264952788528545610e3)
_C = (1.42343711074968357734e0, 4.63033784615654529590e0,
      5.76949722146069140550e0, 3.64784832476320460504e0,
      1.27045825245236838258e0, 2.41780725177450611770e-1,
      2.27238449892691845833e-2, 7.74545014278341407640e-4)
_D = (1.0, 2.05319162663775882187e0, 1.67638483018380384940e0,
      6.89767334985100004550e-1, 1.48103976427480074590e-1,
      1.51986665636164571966e-2, 5.47593808499534494600e-4,
      1.05075007164441684324e-9)
_E = (6.65790464350110377720e0, 5.46378491116411436990e0,
      1.78482653991729133580e0, 2.96560571828504891230e-1,
      2.65321895265761230930e-2, 1.24266094738807843860e-3,
      2.71155556874348757815e-5, 2.01033439929228813265e-7)
_F = (1.0, 5.99832206555887937690e-1, 1.36929880922735805310e-1,
      1.48753612908506148525e-2, 7.86869131145613259100e-4,
      1.84631831751005468180e-5, 1.42151175831644588870e-7,
      2.04426310338993978564e-15)


def _horner(c, r):
    acc = 0.0
    for v in reversed(c):
        acc = acc * r + v
    return acc


def pt_norm(p):
    """Standard normal quantile."""
    assert 0.0 < p < 1.0
    q = p - 0.5
    if abs(q) <= 0.425:
        r = 0.180625 - q * q
        return q * _horner(_A, r) / _horner(_B, r)
    r = p if q < 0.0 else 1.0 - p
    r = math.sqrt(-math.log(r))
    if r <= 5.0:
        r -= 1.6
        ret = _horner(_C, r) / _horner(_D, r)
    else:
        r -= 5.0
        ret = _horner(_E, r) / _horner(_F, r)
    return -ret if q < 0.0 else ret


def pt_chi2(p, v, ln_gamma_half_v):
    """Chi-square quantile with v degrees of freedom; -1.0 on failure."""
    assert 0.0 <= p < 1.0 and v > 0.0
    e = 5.0e-7
    aa = 0.6931471805
    xx = 0.5 * v
    c = xx - 1.0
    if v < -1.24 * math.log(p):
        ch = (p * xx * math.exp(ln_gamma_half_v + xx * aa)) ** (1.0 / xx)
        if ch - e < 0.0:
            return ch
    elif v > 0.32:
        x = pt_norm(p)
        p1 = 0.222222 / v
        ch = v * (x * math.sqrt(p1) + 1.0 - p1) ** 3
        if ch > 2.2 * v + 6.0:
            ch = -2.0 * (math.log(1.0 - p) - c * math.log(0.5 * ch)
                         + ln_gamma_half_v)
    else:
        ch = 0.4
        a = math.log(1.0 - p)
        while True:
            q = ch
            p1 = 1.0 + ch * (4.67 + ch)
            p2 = ch * (6.73 + ch * (6.66 + ch))
            t = -0.5 + (4.67 + 2.0 * ch) / p1 - \
                (6.73 + ch * (13.32 + 3.0 * ch)) / p2
            ch -= (1.0 - math.exp(a + ln_gamma_half_v + 0.5 * ch + c * aa)
                   * p2 / p1) / t
            if abs(q / ch - 1.0) - 0.01 <= 0.0:
                break
    for _ in range(20):
        q = ch
        p1 = 0.5 * ch
        if p1 < 0.0:
            return -1.0
        p2 = p - gamma_inc(p1, xx, ln_gamma_half_v)
        t = p2 * math.exp(xx * aa + ln_gamma_half_v + p1 - c * math.log(ch))
        b = t / ch
        a = 0.5 * t - b * c
        s1 = (210.0 + a * (140.0 + a * (105.0 + a * (84.0 + a * (70.0 + 60.0
              * a))))) / 420.0
        s2 = (420.0 + a * (735.0 + a * (966.0 + a * (1141.0 + 1278.0 * a)))) \
            / 2520.0
        s3 = (210.0 + a * (462.0 + a * (707.0 + 932.0 * a))) / 2520.0
        s4 = (252.0 + a * (672.0 + 1182.0 * a) + c * (294.0 + a * (889.0
              + 1740.0 * a))) / 5040.0
        s5 = (84.0 + 264.0 * a + c * (175.0 + 606.0 * a)) / 2520.0
        s6 = (120.0 + c * (346.0 + 127.0 * c)) / 5040.0
        ch += t * (1.0 + 0.5 * t * s1 - b * c * (s1 - b * (s2 - b * (s3 - b
                   * (s4 - b * (s5 - b * s6))))))
        if abs(q / ch - 1.0) <= e:
            break
    return ch


def gamma_cmult(alpha, ncat, median=False):
    """Rate multipliers of ncat equal-weight Gamma(alpha) categories with
    mean 1 (pkg/Crux/Tree/Lik.pyx:1212-1257)."""
    lga = ln_gamma(alpha)
    cm = [0.0] * ncat
    if not median:
        lga1 = ln_gamma(alpha + 1.0)
        for i in range(ncat - 1):
            pt = pt_chi2((i + 1.0) / ncat, alpha * 2.0, lga)
            if pt == -1.0:
                raise OverflowError("Error discretizing gamma (ncat=%d, "
                                    "alpha=%e)" % (ncat, alpha))
            cm[i] = gamma_inc(pt / 2.0, alpha + 1.0, lga1)
        cm[ncat - 1] = 1.0
        for i in range(ncat - 1, 0, -1):
            cm[i] -= cm[i - 1]
            cm[i] *= float(ncat)
        cm[0] *= float(ncat)
    else:
        s = 0.0
        for i in range(ncat):
            pt = pt_chi2((i * 2 + 1.0) / (ncat * 2), alpha * 2.0, lga)
            if pt == -1.0:
                raise OverflowError("Error discretizing gamma (ncat=%d, "
                                    "alpha=%e)" % (ncat, alpha))
            pt /= alpha * 2.0
            s += pt
            cm[i] = pt
        for i in range(ncat):
            cm[i] *= float(ncat) / s
    return cm
