// Oracle: smooth particle-mesh Ewald for the reciprocal-space term. Test
// infrastructure only (see oracle.hpp).
//
// The reference names PME as future work for Ewald's O(K N) reciprocal sum
// (forcefield/ewald.cpp:112-172); it has no PME code. This restates the
// published method (U. Essmann, L. Perera, M. L. Berkowitz, T. Darden, H. Lee
// and L. G. Pedersen, J. Chem. Phys. 103, 8577 (1995)) over the reference's
// own k-vector set, so that it approximates exactly the sum the reference
// evaluates:
//
//   E = sum_{n in Z^3, 0 < |n|^2 < n2max} C (2 pi / V) e^{-k^2/4a^2} / k^2 |S(n)|^2,
//   S(n) = sum_i q_i e^{i k.r_i},  k = 2 pi n / L  (ewald.cpp:112-172).
//
// With u_i = K frac(x_i / L + 1/2) (scaled fractional coordinates) and the
// cardinal B-spline M_p of order p, S(n) = (-1)^(n.1) sum_i q_i
// e^{2 pi i n.u_i/K} ~ (-1)^(n.1) b(n) F[Q](n), where
//   Q(k)   = sum_i q_i prod_d M_p(u_id - k_d) (periodic in k),
//   F[Q](n) = sum_k Q(k) e^{2 pi i n.k/K},
//   |b_d(m)|^2 = 1 / |sum_{j=0}^{p-2} M_p(j+1) e^{2 pi i m j/K_d}|^2,
// so E ~ sum_n G(n) |F[Q](n)|^2 with G = C (2 pi/V) e^{-k^2/4a^2}/k^2 |b|^2.
// Forces are the exact gradient of that energy:
//   phi(k) = dE/dQ(k) = 2 Re sum_n G(n) conj(F[Q](n)) e^{2 pi i n.k/K},
//   F_i = -q_i sum_k phi(k) grad_i prod_d M_p(u_id - k_d),
//   dM_p(t)/dt = M_{p-1}(t) - M_{p-1}(t - 1),  du_d/dx_d = K_d / L_d.
// Only |n_d| <= kmax enters, so F[Q] and phi are evaluated as separable
// sums over one axis at a time restricted to those n (no FFT).
#include <algorithm>
#include <complex>

#include "oracle.hpp"

namespace oracle {

namespace {
constexpr double kPi = 3.14159265358979323846;
using cplx = std::complex<double>;

// M_p(w + j) for j = 0 .. p-1 and w in [0, 1): the recursion
// M_n(t) = (t M_{n-1}(t) + (n - t) M_{n-1}(t - 1)) / (n - 1) from
// M_2(t) = 1 - |t - 1| on [0, 2]. Also M_{p-1}(w + j), for the derivative.
void bspline(double w, int p, std::vector<double>& m, std::vector<double>& m1) {
    std::vector<double> cur(static_cast<size_t>(p) + 1, 0.0), nxt(cur.size(), 0.0);
    cur[0] = w;        // M_2(w)
    cur[1] = 1.0 - w;  // M_2(w + 1)
    for (int n = 3; n <= p; ++n) {
        if (n == p) m1.assign(cur.begin(), cur.begin() + p);
        for (int j = 0; j < n; ++j) {
            const double t = w + j;
            const double a = j < n - 1 ? cur[j] : 0.0;   // M_{n-1}(t)
            const double b = j > 0 ? cur[j - 1] : 0.0;   // M_{n-1}(t - 1)
            nxt[j] = (t * a + (n - t) * b) / (n - 1);
        }
        std::copy(nxt.begin(), nxt.begin() + n, cur.begin());
    }
    m.assign(cur.begin(), cur.begin() + p);
    if (p == 2) m1.assign({1.0, 0.0});  // M_1 = 1 on [0, 1)
}

// M_p(t) at the integer points t = 1 .. p - 1 (for |b(m)|^2).
double bspline_at(int p, int t) {
    std::vector<double> m, m1;
    bspline(0.0, p, m, m1);  // M_p(0 + j), j = 0 .. p-1
    return m[static_cast<size_t>(t)];
}

// e^{2 pi i m k / K} with the exponent reduced in integers first.
cplx twiddle(long m, long k, int K) {
    long r = (m * k) % K;
    if (r < 0) r += K;
    const double ang = 2.0 * kPi * static_cast<double>(r) / static_cast<double>(K);
    return cplx(std::cos(ang), std::sin(ang));
}
}  // namespace

double pme_bsq(int p, int K, int m) {
    cplx s(0.0, 0.0);
    for (int j = 0; j <= p - 2; ++j) s += bspline_at(p, j + 1) * twiddle(m, j, K);
    return 1.0 / std::norm(s);
}

double pme_recip(const State& s, int kmax, double alpha_in, const int K[3], int p,
                 bool accumulate_forces, State* sink) {
    if (p < 3 || p > 12) throw std::invalid_argument("pme: order must be 3 .. 12");
    for (int d = 0; d < 3; ++d)
        if (K[d] < 2 * kmax + 1) throw std::invalid_argument("pme: grid smaller than 2 kmax + 1");
    const std::size_t n = s.n();
    const double minL = std::min({s.L[0], s.L[1], s.L[2]});
    const double alpha = alpha_in > 0.0 ? alpha_in : 5.6 / minL;
    const double C = coulomb_prefactor();
    const double V = s.volume();
    const int nb = 2 * kmax + 1;
    const long K0 = K[0], K1 = K[1], K2 = K[2];

    // spline weights and the charge grid
    std::vector<double> W[3], D[3];
    std::vector<long> base[3];
    for (int d = 0; d < 3; ++d) {
        W[d].resize(n * p);
        D[d].resize(n * p);
        base[d].resize(n);
    }
    std::vector<double> Q(static_cast<size_t>(K0 * K1 * K2), 0.0);
    for (std::size_t i = 0; i < n; ++i) {
        const double r[3] = {s.x[i], s.y[i], s.z[i]};
        for (int d = 0; d < 3; ++d) {
            double f = r[d] / s.L[d] + 0.5;
            f -= std::floor(f);
            double u = f * K[d];
            if (u >= K[d]) u -= K[d];
            const double fl = std::floor(u);
            std::vector<double> m, m1;
            bspline(u - fl, p, m, m1);
            base[d][i] = static_cast<long>(fl);
            for (int j = 0; j < p; ++j) {
                W[d][i * p + j] = m[j];
                D[d][i * p + j] = m1[j] - (j > 0 ? m1[j - 1] : 0.0);
            }
        }
        for (int j2 = 0; j2 < p; ++j2) {
            const long k2 = ((base[2][i] - j2) % K2 + K2) % K2;
            for (int j1 = 0; j1 < p; ++j1) {
                const long k1 = ((base[1][i] - j1) % K1 + K1) % K1;
                for (int j0 = 0; j0 < p; ++j0) {
                    const long k0 = ((base[0][i] - j0) % K0 + K0) % K0;
                    Q[(k2 * K1 + k1) * K0 + k0] +=
                        s.q[i] * W[0][i * p + j0] * W[1][i * p + j1] * W[2][i * p + j2];
                }
            }
        }
    }

    // F[Q](n) for |n_d| <= kmax, one axis at a time
    std::vector<cplx> A(static_cast<size_t>(K2 * K1 * nb));
    for (long k2 = 0; k2 < K2; ++k2)
        for (long k1 = 0; k1 < K1; ++k1)
            for (int a = 0; a < nb; ++a) {
                cplx acc(0.0, 0.0);
                for (long k0 = 0; k0 < K0; ++k0)
                    acc += Q[(k2 * K1 + k1) * K0 + k0] * twiddle(a - kmax, k0, K[0]);
                A[(k2 * K1 + k1) * nb + a] = acc;
            }
    std::vector<cplx> B(static_cast<size_t>(K2 * nb * nb));
    for (long k2 = 0; k2 < K2; ++k2)
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b) {
                cplx acc(0.0, 0.0);
                for (long k1 = 0; k1 < K1; ++k1)
                    acc += A[(k2 * K1 + k1) * nb + a] * twiddle(b - kmax, k1, K[1]);
                B[(k2 * nb + a) * nb + b] = acc;
            }
    std::vector<cplx> P(static_cast<size_t>(nb * nb * nb));
    const double k0v[3] = {2.0 * kPi / s.L[0], 2.0 * kPi / s.L[1], 2.0 * kPi / s.L[2]};
    const int n2max = kmax * kmax + 2;
    double E = 0.0;
    for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nb; ++c) {
                cplx fq(0.0, 0.0);
                for (long k2 = 0; k2 < K2; ++k2)
                    fq += B[(k2 * nb + a) * nb + b] * twiddle(c - kmax, k2, K[2]);
                const int m0 = a - kmax, m1 = b - kmax, m2 = c - kmax;
                const int n2 = m0 * m0 + m1 * m1 + m2 * m2;
                double g = 0.0;
                if (n2 != 0 && n2 < n2max) {
                    const double kx = k0v[0] * m0, ky = k0v[1] * m1, kz = k0v[2] * m2;
                    const double k2 = kx * kx + ky * ky + kz * kz;
                    g = C * (2.0 * kPi / V) * std::exp(-k2 / (4.0 * alpha * alpha)) / k2 *
                        pme_bsq(p, K[0], m0) * pme_bsq(p, K[1], m1) * pme_bsq(p, K[2], m2);
                }
                E += g * std::norm(fq);
                P[(a * nb + b) * nb + c] = g * std::conj(fq);
            }
    if (!accumulate_forces) return E;

    // phi(k) = 2 Re sum_n P(n) e^{2 pi i n.k/K}, again one axis at a time
    std::vector<cplx> Cz(static_cast<size_t>(K2 * nb * nb));
    for (long k2 = 0; k2 < K2; ++k2)
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b) {
                cplx acc(0.0, 0.0);
                for (int c = 0; c < nb; ++c)
                    acc += P[(a * nb + b) * nb + c] * twiddle(c - kmax, k2, K[2]);
                Cz[(k2 * nb + a) * nb + b] = acc;
            }
    std::vector<cplx> Dy(static_cast<size_t>(K2 * K1 * nb));
    for (long k2 = 0; k2 < K2; ++k2)
        for (long k1 = 0; k1 < K1; ++k1)
            for (int a = 0; a < nb; ++a) {
                cplx acc(0.0, 0.0);
                for (int b = 0; b < nb; ++b)
                    acc += Cz[(k2 * nb + a) * nb + b] * twiddle(b - kmax, k1, K[1]);
                Dy[(k2 * K1 + k1) * nb + a] = acc;
            }
    std::vector<double> phi(Q.size());
    for (long k2 = 0; k2 < K2; ++k2)
        for (long k1 = 0; k1 < K1; ++k1)
            for (long k0 = 0; k0 < K0; ++k0) {
                cplx acc(0.0, 0.0);
                for (int a = 0; a < nb; ++a)
                    acc += Dy[(k2 * K1 + k1) * nb + a] * twiddle(a - kmax, k0, K[0]);
                phi[(k2 * K1 + k1) * K0 + k0] = 2.0 * acc.real();
            }

    for (std::size_t i = 0; i < n; ++i) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int j2 = 0; j2 < p; ++j2) {
            const long k2 = ((base[2][i] - j2) % K2 + K2) % K2;
            for (int j1 = 0; j1 < p; ++j1) {
                const long k1 = ((base[1][i] - j1) % K1 + K1) % K1;
                for (int j0 = 0; j0 < p; ++j0) {
                    const long k0 = ((base[0][i] - j0) % K0 + K0) % K0;
                    const double ph = phi[(k2 * K1 + k1) * K0 + k0];
                    const double w0 = W[0][i * p + j0], w1 = W[1][i * p + j1], w2 = W[2][i * p + j2];
                    gx += ph * D[0][i * p + j0] * w1 * w2;
                    gy += ph * w0 * D[1][i * p + j1] * w2;
                    gz += ph * w0 * w1 * D[2][i * p + j2];
                }
            }
        }
        sink->fx[i] += -s.q[i] * gx * (K[0] / s.L[0]);
        sink->fy[i] += -s.q[i] * gy * (K[1] / s.L[1]);
        sink->fz[i] += -s.q[i] * gz * (K[2] / s.L[2]);
    }
    return E;
}

}  // namespace oracle
