// FP64 model setup on the host: Q, its eigensystem, P(t), Gamma categories.
//
// Q follows CxLikQ (lib/Crux/Tree/CxLik.c:59-141): Q[i][j] = R_ij * pi_j for
// i != j with pi normalised to sum 1, rows summing to 0, and
// qNorm = 1 / sum_{i!=j} pi_i Q_ij.  The reference factors Q with LAPACK
// dgeev + dgetrf/dgetri (CxLik.c:177-250).  Here Q is reversible, so it is
// similar to the symmetric matrix B = Pi^1/2 Q Pi^-1/2 (B_ij =
// sqrt(pi_i pi_j) R_ij off the diagonal); B = U L U^T is found with cyclic
// Jacobi rotations, and V = Pi^-1/2 U, V^-1 = U^T Pi^1/2.  P(t) is
// invariant to how V is normalised, so this gives the same P as the
// reference's factorisation up to rounding.
//
// Gamma categories restate the published algorithms the reference uses
// (lib/Crux/CxMath.pxd:125-368): Pike & Hill 1966 (ln Gamma),
// Bhattacharjee 1970 AS 32 (incomplete Gamma ratio), Wichura 1988 AS 241
// (normal quantile), Best & Roberts 1975 AS 91 with Shea 1991 AS R85
// (chi-square quantile); categories as in Lik.setAlpha (Lik.pyx:1212-1257).

#include <cmath>
#include <cstring>
#include <vector>

#include "crux_lik.h"
#include "internal.h"

namespace crux {

static inline int tri_index(int n, int x, int y)
{
    // Index of (x, y), x < y, in the row-by-row upper triangle.
    return n * x + y - (((x + 3) * x) >> 1) - 1;
}

// Cyclic Jacobi eigensolver for a symmetric n x n matrix a (row-major,
// destroyed).  On return w holds eigenvalues and u (row-major) the
// eigenvectors in columns.  Returns false if it fails to converge.
static bool jacobi_eigen(int n, std::vector<double> &a, std::vector<double> &w,
                         std::vector<double> &u)
{
    u.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        u[(size_t)i * n + i] = 1.0;
    auto A = [&](int i, int j) -> double & { return a[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; i++) {
            diag += A(i, i) * A(i, i);
            for (int j = i + 1; j < n; j++)
                off += A(i, j) * A(i, j);
        }
        if (off == 0.0 || off <= 1e-34 * diag)
            break;
        for (int p = 0; p < n - 1; p++) {
            for (int q = p + 1; q < n; q++) {
                double apq = A(p, q);
                if (apq == 0.0)
                    continue;
                double app = A(p, p), aqq = A(q, q);
                // Rotation angle that zeroes (p, q): tan(2 phi) = 2apq/(aqq-app).
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0.0 ? 1.0 : -1.0) /
                           (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - s * akq;
                    A(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A(p, k), aqk = A(q, k);
                    A(p, k) = c * apk - s * aqk;
                    A(q, k) = s * apk + c * aqk;
                }
                A(p, q) = A(q, p) = 0.0;
                for (int k = 0; k < n; k++) {
                    double ukp = u[(size_t)k * n + p], ukq = u[(size_t)k * n + q];
                    u[(size_t)k * n + p] = c * ukp - s * ukq;
                    u[(size_t)k * n + q] = s * ukp + c * ukq;
                }
            }
        }
        if (sweep == 99)
            return false;
    }
    w.resize(n);
    for (int i = 0; i < n; i++)
        w[i] = A(i, i);
    return true;
}

int q_decomp(int n, const double *rtri, const double *pi, double *pi_norm,
             double *q_norm, double *evec, double *evec_inv, double *evals)
{
    if (n < 2 || n > kMaxStates)
        return set_error(CRUX_EINVAL, "q_decomp: n=%d out of range", n);
    double psum = 0.0;
    for (int i = 0; i < n; i++) {
        if (!(pi[i] > 0.0))
            return set_error(CRUX_EINVAL, "q_decomp: frequency %d is not > 0", i);
        psum += pi[i];
    }
    for (int i = 0; i < n; i++)
        pi_norm[i] = pi[i] / psum;
    // Off-diagonal Q, qScale = sum_{i != j} pi_i Q_ij (pairs counted twice).
    double qscale = 0.0;
    std::vector<double> b((size_t)n * n, 0.0);
    std::vector<double> rowsum(n, 0.0);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            if (i == j)
                continue;
            double r = i < j ? rtri[tri_index(n, i, j)] : rtri[tri_index(n, j, i)];
            if (r < 0.0)
                return set_error(CRUX_EINVAL, "q_decomp: negative rate");
            double qij = r * pi_norm[j];
            rowsum[i] += qij;
            if (j < i)
                qscale += qij * pi_norm[i] * 2.0;
            b[(size_t)i * n + j] = std::sqrt(pi_norm[i] * pi_norm[j]) * r;
        }
    }
    for (int i = 0; i < n; i++)
        b[(size_t)i * n + i] = -rowsum[i];
    *q_norm = qscale != 0.0 ? 1.0 / qscale : INFINITY;

    std::vector<double> w, u;
    if (!jacobi_eigen(n, b, w, u))
        return set_error(CRUX_EDECOMP, "Error decomposing Q");
    for (int i = 0; i < n; i++) {
        double rs = 1.0 / std::sqrt(pi_norm[i]), sq = std::sqrt(pi_norm[i]);
        for (int k = 0; k < n; k++) {
            evec[(size_t)i * n + k] = rs * u[(size_t)i * n + k];
            evec_inv[(size_t)k * n + i] = u[(size_t)i * n + k] * sq;
        }
    }
    for (int k = 0; k < n; k++) {
        if (!std::isfinite(w[k]))
            return set_error(CRUX_EDECOMP, "Error decomposing Q");
        evals[k] = w[k];
    }
    return CRUX_OK;
}

int pt(int n, const double *evec, const double *evec_inv, const double *evals,
       double t, double *P)
{
    double e[kMaxStates];
    for (int k = 0; k < n; k++)
        e[k] = std::exp(evals[k] * t);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            double p = 0.0;
            for (int k = 0; k < n; k++)
                p += evec[(size_t)i * n + k] * e[k] * evec_inv[(size_t)k * n + j];
            P[(size_t)i * n + j] = p < 0.0 ? 0.0 : p;
        }
    }
    return CRUX_OK;
}

// ---- Gamma categories ------------------------------------------------------

// ln Gamma(x): raise x to >= 7 by a running product, then Stirling's series.
static double ln_gamma(double x)
{
    double shift = 0.0;
    if (x < 7.0) {
        double prod = 1.0;
        for (; x < 7.0; x += 1.0)
            prod *= x;
        shift = -std::log(prod);
    }
    const double z = 1.0 / (x * x);
    const double series =
        (((-0.000595238095238 * z + 0.000793650793651) * z - 0.002777777777778) * z +
         0.083333333333333) / x;
    return shift + (x - 0.5) * std::log(x) - x + 0.918938533204673 + series;
}

// Regularised lower incomplete Gamma ratio P(p, x) (AS 32), accuracy 1e-10.
static double inc_gamma(double x, double p, double lg_p)
{
    if (x == 0.0)
        return 0.0;
    const double acu = 1.0e-10, oflo = 1.0e30;
    const double factor = std::exp(p * std::log(x) - x - lg_p);
    if (x <= 1.0 || x < p) {
        // Series: sum_{k>=0} x^k / ((p+1)...(p+k)).
        double sum = 1.0, term = 1.0, den = p;
        do {
            den += 1.0;
            term *= x / den;
            sum += term;
        } while (term > acu);
        return sum * factor / p;
    }
    // Continued fraction, with the recurrence rescaled when it grows large.
    double a = 1.0 - p, b = a + x + 1.0, k = 0.0;
    double p0 = 1.0, p1 = x, p2 = x + 1.0, p3 = x * b;
    double g = p2 / p3;
    for (;;) {
        a += 1.0;
        b += 2.0;
        k += 1.0;
        const double an = a * k;
        const double p4 = b * p2 - an * p0, p5 = b * p3 - an * p1;
        if (p5 != 0.0) {
            const double r = p4 / p5, d = std::fabs(g - r);
            if (d <= acu && d <= acu * r)
                return 1.0 - factor * g;
            g = r;
        }
        p0 = p2;
        p1 = p3;
        p2 = p4;
        p3 = p5;
        if (std::fabs(p4) >= oflo) {
            p0 /= oflo;
            p1 /= oflo;
            p2 /= oflo;
            p3 /= oflo;
        }
    }
}

static double poly(const double *c, int n, double r)
{
    double acc = 0.0;
    for (int i = n - 1; i >= 0; i--)
        acc = acc * r + c[i];
    return acc;
}

// Normal quantile (AS 241, PPND16): three rational approximations.
static double norm_quantile(double p)
{
    static const double a[8] = {3.3871328727963666080e0, 1.3314166789178437745e2,
                                1.9715909503065514427e3, 1.3731693765509461125e4,
                                4.5921953931549871457e4, 6.7265770927008700853e4,
                                3.3430575583588128105e4, 2.5090809287301226727e3};
    static const double b[8] = {1.0, 4.2313330701600911252e1, 6.8718700749205790830e2,
                                5.3941960214247511077e3, 2.1213794301586595867e4,
                                3.9307895800092710610e4, 2.8729085735721942674e4,
                                5.2264952788528545610e3};
    static const double c[8] = {1.42343711074968357734e0, 4.63033784615654529590e0,
                                5.76949722146069140550e0, 3.64784832476320460504e0,
                                1.27045825245236838258e0, 2.41780725177450611770e-1,
                                2.27238449892691845833e-2, 7.74545014278341407640e-4};
    static const double d[8] = {1.0, 2.05319162663775882187e0, 1.67638483018380384940e0,
                                6.89767334985100004550e-1, 1.48103976427480074590e-1,
                                1.51986665636164571966e-2, 5.47593808499534494600e-4,
                                1.05075007164441684324e-9};
    static const double e[8] = {6.65790464350110377720e0, 5.46378491116411436990e0,
                                1.78482653991729133580e0, 2.96560571828504891230e-1,
                                2.65321895265761230930e-2, 1.24266094738807843860e-3,
                                2.71155556874348757815e-5, 2.01033439929228813265e-7};
    static const double f[8] = {1.0, 5.99832206555887937690e-1, 1.36929880922735805310e-1,
                                1.48753612908506148525e-2, 7.86869131145613259100e-4,
                                1.84631831751005468180e-5, 1.42151175831644588870e-7,
                                2.04426310338993978564e-15};
    const double q = p - 0.5;
    if (std::fabs(q) <= 0.425) {
        const double r = 0.180625 - q * q;
        return q * poly(a, 8, r) / poly(b, 8, r);
    }
    double r = std::sqrt(-std::log(q < 0.0 ? p : 1.0 - p));
    double v;
    if (r <= 5.0) {
        r -= 1.6;
        v = poly(c, 8, r) / poly(d, 8, r);
    } else {
        r -= 5.0;
        v = poly(e, 8, r) / poly(f, 8, r);
    }
    return q < 0.0 ? -v : v;
}

// Chi-square quantile with nu degrees of freedom (AS 91 + AS R85); returns -1
// when the refinement leaves the domain.
static double chi2_quantile(double p, double nu, double lg_half_nu)
{
    const double eps = 5.0e-7, ln2 = 0.6931471805;
    const double xx = 0.5 * nu, c = xx - 1.0;
    double ch;
    if (nu < -1.24 * std::log(p)) {
        ch = std::pow(p * xx * std::exp(lg_half_nu + xx * ln2), 1.0 / xx);
        if (ch - eps < 0.0)
            return ch;
    } else if (nu > 0.32) {
        // Wilson-Hilferty start, with the p -> 1 correction.
        const double x = norm_quantile(p), p1 = 0.222222 / nu;
        ch = nu * std::pow(x * std::sqrt(p1) + 1.0 - p1, 3.0);
        if (ch > 2.2 * nu + 6.0)
            ch = -2.0 * (std::log(1.0 - p) - c * std::log(0.5 * ch) + lg_half_nu);
    } else {
        // Small nu: Newton-like iteration on the log of the upper tail.
        ch = 0.4;
        const double a = std::log(1.0 - p);
        for (;;) {
            const double prev = ch;
            const double p1 = 1.0 + ch * (4.67 + ch);
            const double p2 = ch * (6.73 + ch * (6.66 + ch));
            const double t = -0.5 + (4.67 + 2.0 * ch) / p1 -
                             (6.73 + ch * (13.32 + 3.0 * ch)) / p2;
            ch -= (1.0 - std::exp(a + lg_half_nu + 0.5 * ch + c * ln2) * p2 / p1) / t;
            if (std::fabs(prev / ch - 1.0) - 0.01 <= 0.0)
                break;
        }
    }
    // Seven-term Taylor refinement.
    for (int it = 0; it < 20; it++) {
        const double prev = ch, half = 0.5 * ch;
        if (half < 0.0)
            return -1.0;
        const double p2 = p - inc_gamma(half, xx, lg_half_nu);
        const double t = p2 * std::exp(xx * ln2 + lg_half_nu + half - c * std::log(ch));
        const double b = t / ch, a = 0.5 * t - b * c;
        const double s1 = (210.0 + a * (140.0 + a * (105.0 + a * (84.0 + a * (70.0 + 60.0 * a))))) / 420.0;
        const double s2 = (420.0 + a * (735.0 + a * (966.0 + a * (1141.0 + 1278.0 * a)))) / 2520.0;
        const double s3 = (210.0 + a * (462.0 + a * (707.0 + 932.0 * a))) / 2520.0;
        const double s4 = (252.0 + a * (672.0 + 1182.0 * a) + c * (294.0 + a * (889.0 + 1740.0 * a))) / 5040.0;
        const double s5 = (84.0 + 264.0 * a + c * (175.0 + 606.0 * a)) / 2520.0;
        const double s6 = (120.0 + c * (346.0 + 127.0 * c)) / 5040.0;
        ch += t * (1.0 + 0.5 * t * s1 - b * c * (s1 - b * (s2 - b * (s3 - b * (s4 - b * (s5 - b * s6))))));
        if (std::fabs(prev / ch - 1.0) <= eps)
            break;
    }
    return ch;
}

int gamma_rates(double alpha, int ncat, int median, double *cmult)
{
    if (!(alpha > 0.0) || ncat < 1)
        return set_error(CRUX_EINVAL, "gamma_rates: alpha=%g ncat=%d", alpha, ncat);
    if (std::isinf(alpha) || ncat == 1) {
        for (int i = 0; i < ncat; i++)
            cmult[i] = 1.0;
        return CRUX_OK;
    }
    const double lg = ln_gamma(alpha);
    if (!median) {
        // Equal-probability boundaries from the chi-square quantile, then the
        // mean of each slice from the incomplete Gamma of shape alpha+1.
        const double lg1 = ln_gamma(alpha + 1.0);
        std::vector<double> cum(ncat);
        for (int i = 0; i < ncat - 1; i++) {
            const double b = chi2_quantile((i + 1.0) / ncat, alpha * 2.0, lg);
            if (b == -1.0)
                return set_error(CRUX_EINVAL, "Error discretizing gamma (ncat=%d, alpha=%e)",
                                 ncat, alpha);
            cum[i] = inc_gamma(b / 2.0, alpha + 1.0, lg1);
        }
        cum[ncat - 1] = 1.0;
        for (int i = ncat - 1; i > 0; i--)
            cmult[i] = (cum[i] - cum[i - 1]) * (double)ncat;
        cmult[0] = cum[0] * (double)ncat;
    } else {
        double sum = 0.0;
        for (int i = 0; i < ncat; i++) {
            double m = chi2_quantile((i * 2 + 1.0) / (ncat * 2), alpha * 2.0, lg);
            if (m == -1.0)
                return set_error(CRUX_EINVAL, "Error discretizing gamma (ncat=%d, alpha=%e)",
                                 ncat, alpha);
            m /= alpha * 2.0;
            sum += m;
            cmult[i] = m;
        }
        for (int i = 0; i < ncat; i++)
            cmult[i] *= (double)ncat / sum;
    }
    return CRUX_OK;
}

}  // namespace crux

extern "C" int crux_q_decomp(int n, const double *rtri, const double *pi, double *pi_norm,
                             double *q_norm, double *evec, double *evec_inv, double *evals)
{
    return crux::q_decomp(n, rtri, pi, pi_norm, q_norm, evec, evec_inv, evals);
}

extern "C" int crux_pt(int n, const double *evec, const double *evec_inv,
                       const double *evals, double t, double *P)
{
    if (n < 1 || n > crux::kMaxStates)
        return crux::set_error(CRUX_EINVAL, "crux_pt: n=%d out of range", n);
    return crux::pt(n, evec, evec_inv, evals, t, P);
}

extern "C" int crux_gamma_rates(double alpha, int ncat, int median, double *cmult)
{
    return crux::gamma_rates(alpha, ncat, median, cmult);
}
