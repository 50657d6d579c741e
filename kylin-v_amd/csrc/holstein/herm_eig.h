// Small dense Hermitian eigenproblem and the Krylov-space exponential.  Host-only C++.
//
// The reference diagonalises the <= 14 x 14 Lanczos matrix with LAPACKE_zheevd('V','U')
// (util/allocation.h:223-230, called from holstein/exponential.h:175), reading the
// upper triangle and the real part of the diagonal.  Here: cyclic complex Jacobi
// rotations (each rotation = a phase that makes a_pq real, then the classical real
// Jacobi angle), which is accurate to a few ulps for matrices this small.
#pragma once

#include <cmath>
#include <complex>
#include <vector>

namespace kv {

using zd = std::complex<double>;

// a: n x n row-major; only the upper triangle and Re(diag) are read.  On return w holds
// the eigenvalues and v (row-major, column j = eigenvector j) the eigenvectors.
inline void herm_eig_jacobi(int n, const std::vector<zd>& a_in, std::vector<double>& w,
                            std::vector<zd>& v) {
    std::vector<zd> a((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (i == j)
                a[i * n + j] = zd(a_in[i * n + i].real(), 0.0);
            else if (j > i)
                a[i * n + j] = a_in[i * n + j];
            else
                a[i * n + j] = std::conj(a_in[j * n + i]);
        }
    v.assign((size_t)n * n, zd(0, 0));
    for (int i = 0; i < n; ++i) v[i * n + i] = zd(1, 0);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; ++i) {
            diag += std::norm(a[i * n + i]);
            for (int j = i + 1; j < n; ++j) off += std::norm(a[i * n + j]);
        }
        if (off <= 1e-34 * (diag + 1e-300) || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const zd apq = a[p * n + q];
                const double b = std::abs(apq);
                if (b == 0.0) continue;
                const zd e = apq / b;  // a_pq = b e
                const double app = a[p * n + p].real(), aqq = a[q * n + q].real();
                const double theta = (aqq - app) / (2.0 * b);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                // G = diag(1, conj(e)) * [[c, s], [-s, c]]
                const zd gpp(c, 0), gpq(s, 0), gqp = -s * std::conj(e), gqq = c * std::conj(e);
                for (int k = 0; k < n; ++k) {  // A <- A G (columns p, q)
                    const zd akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = akp * gpp + akq * gqp;
                    a[k * n + q] = akp * gpq + akq * gqq;
                }
                for (int k = 0; k < n; ++k) {  // A <- G^H A (rows p, q)
                    const zd apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = std::conj(gpp) * apk + std::conj(gqp) * aqk;
                    a[q * n + k] = std::conj(gpq) * apk + std::conj(gqq) * aqk;
                }
                a[p * n + q] = zd(0, 0);
                a[q * n + p] = zd(0, 0);
                a[p * n + p] = zd(a[p * n + p].real(), 0);
                a[q * n + q] = zd(a[q * n + q].real(), 0);
                for (int k = 0; k < n; ++k) {  // V <- V G
                    const zd vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = vkp * gpp + vkq * gqp;
                    v[k * n + q] = vkp * gpq + vkq * gqq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = a[i * n + i].real();
}

// Krylov-space exponential coefficients (holstein/exponential.h:162-183):
// T = diag(alpha) + upper(beta_upper), c = V exp(img * Lambda) V^H e_0 with img = -i
// forward and +i for the reverse (backward) step.
inline std::vector<zd> krylov_coefficients(const std::vector<zd>& alpha,
                                           const std::vector<zd>& beta_upper, bool reverse) {
    const int n = (int)alpha.size();
    std::vector<zd> T((size_t)n * n, zd(0, 0));
    for (int i = 0; i < n; ++i) {
        T[i * n + i] = alpha[i];
        if (i < n - 1) T[i * n + i + 1] = beta_upper[i];
    }
    std::vector<double> w;
    std::vector<zd> v;
    herm_eig_jacobi(n, T, w, v);
    const zd img = reverse ? zd(0, 1) : zd(0, -1);
    std::vector<zd> c(n, zd(0, 0));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) c[i] += v[i * n + j] * std::exp(img * w[j]) * std::conj(v[j]);
    return c;
}

}  // namespace kv
