// Oracle restatement of the trajectory analyses: g(r) pair histograms and
// the MSD / VACF time correlations over every time origin. Test
// infrastructure only (see oracle.hpp).
//
// Sums run over atoms in index order, one FP64 rounding per term, as the
// reference accumulates them, so this restatement and the reference agree
// bit for bit on the same frames (checked against oracle/_ref where built).
#include <algorithm>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

// analyze/rdf.cpp:16-17 (bin width rMax / bins).
Rdf::Rdf(double rmax, long nbins)
    : r_max(rmax), bins(nbins), width(rmax / static_cast<double>(nbins)),
      hist(static_cast<std::size_t>(nbins), 0.0) {}

// analyze/rdf.cpp:19-46: every i < j pair of the frame, minimum image of
// r_i - r_j, r = sqrt(|d|^2), counted in bin floor(r / width) when r < rMax;
// rMax beyond half the smallest box length is refused.
void Rdf::accumulate(const State& s) {
    const double smallest = std::min(s.L[0], std::min(s.L[1], s.L[2]));
    if (r_max > 0.5 * smallest)
        throw std::invalid_argument("Rdf: rMax exceeds half the smallest box length");
    const std::size_t n = s.n();
    for (std::size_t i = 0; i < n; ++i) {
        const V3 ri = s.r(i);
        for (std::size_t j = i + 1; j < n; ++j) {
            const V3 rj = s.r(j);
            const V3 d = min_image(s.L, V3{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z});
            const double r = std::sqrt(dot3(d, d));
            if (!(r < r_max)) continue;
            const std::size_t k = static_cast<std::size_t>(r / width);
            if (k < static_cast<std::size_t>(bins)) hist[k] += 1.0;
        }
    }
    natoms = static_cast<long>(n);
    sum_rho += static_cast<double>(n) / s.volume();
    ++frames;
}

// analyze/rdf.cpp:48-54.
std::vector<double> Rdf::centres() const {
    std::vector<double> c(static_cast<std::size_t>(bins));
    for (long k = 0; k < bins; ++k) c[k] = (static_cast<double>(k) + 0.5) * width;
    return c;
}

// analyze/rdf.cpp:56-71: hist / (frames * N/2 * <rho> * shell volume).
std::vector<double> Rdf::g() const {
    std::vector<double> out(static_cast<std::size_t>(bins), 0.0);
    if (frames == 0) return out;
    const double pi = 3.14159265358979323846;
    const double rho = sum_rho / static_cast<double>(frames);
    for (long k = 0; k < bins; ++k) {
        const double a = static_cast<double>(k) * width;
        const double b = static_cast<double>(k + 1) * width;
        const double shell = (4.0 / 3.0) * pi * (b * b * b - a * a * a);
        const double denom =
            static_cast<double>(frames) * 0.5 * static_cast<double>(natoms) * rho * shell;
        out[k] = denom > 0.0 ? hist[k] / denom : 0.0;
    }
    return out;
}

Corr::Corr(long natoms, long maxlag, bool displacement)
    : n(natoms), max_lag(maxlag), period(maxlag + 1), msd(displacement),
      prev(3 * static_cast<std::size_t>(natoms), 0.0),
      cur(3 * static_cast<std::size_t>(natoms), 0.0),
      ring(static_cast<std::size_t>(maxlag + 1) * 3 * natoms, 0.0),
      sum(static_cast<std::size_t>(maxlag + 1), 0.0), count(static_cast<std::size_t>(maxlag + 1), 0) {}

// analyze/msd.cpp:22-52 (displacement) and vacf.cpp:19-40 (velocity).
// The MSD's current vector is the running unwrapped position: the first
// frame as given, then + minimum image of (wrapped - previous wrapped). The
// VACF's is the velocity. The frame goes into ring slot (pushes mod
// period); every lag up to min(pushes, maxLag) adds its sum over atoms.
void Corr::push(const State& s) {
    const std::size_t N = static_cast<std::size_t>(n);
    for (std::size_t i = 0; i < N; ++i) {
        if (msd) {
            const V3 w = s.r(i);
            if (pushed == 0) {
                cur[3 * i] = w.x; cur[3 * i + 1] = w.y; cur[3 * i + 2] = w.z;
            } else {
                const V3 d = min_image(s.L, V3{w.x - prev[3 * i], w.y - prev[3 * i + 1],
                                               w.z - prev[3 * i + 2]});
                cur[3 * i] += d.x; cur[3 * i + 1] += d.y; cur[3 * i + 2] += d.z;
            }
            prev[3 * i] = w.x; prev[3 * i + 1] = w.y; prev[3 * i + 2] = w.z;
        } else {
            const V3 v = s.v(i);
            cur[3 * i] = v.x; cur[3 * i + 1] = v.y; cur[3 * i + 2] = v.z;
        }
    }
    const long slot = pushed % period;
    std::copy(cur.begin(), cur.end(), ring.begin() + slot * 3 * n);
    const long top = std::min(pushed, max_lag);
    for (long lag = 0; lag <= top; ++lag) {
        const double* o = ring.data() + ((slot + period - lag) % period) * 3 * n;
        double acc = 0.0;
        for (std::size_t i = 0; i < N; ++i) {
            const V3 a{cur[3 * i], cur[3 * i + 1], cur[3 * i + 2]};
            const V3 b{o[3 * i], o[3 * i + 1], o[3 * i + 2]};
            if (msd) {
                const V3 d{a.x - b.x, a.y - b.y, a.z - b.z};
                acc += dot3(d, d);
            } else {
                acc += dot3(a, b);
            }
        }
        sum[lag] += acc;
        count[lag] += n;
    }
    ++pushed;
}

// msd.cpp:54-62 / vacf.cpp:42-50: sum / count where counted, else 0.
std::vector<double> Corr::mean() const {
    std::vector<double> m(static_cast<std::size_t>(period), 0.0);
    for (long k = 0; k < period; ++k)
        if (count[k] > 0) m[k] = sum[k] / static_cast<double>(count[k]);
    return m;
}

// msd.cpp:64-92: least-squares slope of msd against t = k dt over
// [from, min(to, period)), divided by 6; 0 with fewer than 2 points or a
// zero denominator.
double Corr::msd_diffusion(double dt, long from, long to) const {
    const std::vector<double> m = mean();
    const long upper = std::min(to, period);
    double st = 0.0, sm = 0.0, stt = 0.0, stm = 0.0;
    long cnt = 0;
    for (long k = from; k < upper; ++k) {
        const double t = static_cast<double>(k) * dt;
        st += t;
        sm += m[k];
        stt += t * t;
        stm += t * m[k];
        ++cnt;
    }
    if (cnt < 2) return 0.0;
    const double nd = static_cast<double>(cnt);
    const double den = nd * stt - st * st;
    if (den == 0.0) return 0.0;
    return (nd * stm - st * sm) / den / 6.0;
}

// vacf.cpp:52-63: C(lag) / C(0); all zero when C(0) is 0.
std::vector<double> Corr::normalised() const {
    const std::vector<double> u = mean();
    std::vector<double> c(u.size(), 0.0);
    if (u.empty() || u[0] == 0.0) return c;
    for (std::size_t k = 0; k < u.size(); ++k) c[k] = u[k] / u[0];
    return c;
}

// vacf.cpp:65-85: trapezoid integral of the unnormalised C over the leading
// run of counted lags, divided by 3; 0 with fewer than 2 such lags.
double Corr::vacf_diffusion(double dt) const {
    const std::vector<double> u = mean();
    long valid = 0;
    for (long k = 0; k < period; ++k) {
        if (count[k] == 0) break;
        valid = k + 1;
    }
    if (valid < 2) return 0.0;
    double integral = 0.0;
    for (long k = 0; k + 1 < valid; ++k) integral += 0.5 * (u[k] + u[k + 1]) * dt;
    return integral / 3.0;
}

}  // namespace oracle
