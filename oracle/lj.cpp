// Oracle: Lennard-Jones pair kernel, pair sources (brute force, linked
// cells, Verlet list) and tail corrections. Test infrastructure only (see
// oracle.hpp).
#include <algorithm>

#include "oracle.hpp"

namespace oracle {

namespace {
constexpr double kPi = 3.14159265358979323846;

// Neumaier's compensated summation (Neumaier 1974, "Rundungsfehleranalyse
// einiger Verfahren zur Summation endlicher Summen", ZAMM 54:39-51): the
// rounding error of every addition is carried in c.
struct CompSum {
    double s = 0.0, c = 0.0;
    void add(double x) {
        const double t = s + x;
        if (std::fabs(s) >= std::fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    double value() const { return s + c; }
};
}  // namespace

// u(rc) = 4 eps (s6^2 - s6), du/dr(rc) = -w(rc)/rc with s = sigma/rc
// (lennard_jones.cpp:32-48).
LJ::LJ(double sig, double eps, double rc, bool linear_force_shift)
    : sigma(sig), epsilon(eps), cutoff(rc), lfs(linear_force_shift) {
    const double s2 = (sigma * sigma) / (cutoff * cutoff);
    const double s6 = s2 * s2 * s2;
    u_rc = 4.0 * epsilon * (s6 * s6 - s6);
    const double w_rc = 24.0 * epsilon * (2.0 * s6 * s6 - s6);
    dudr_rc = -w_rc / cutoff;
}

// Every i<j, i outer (lennard_jones.cpp:18-29).
void brute_pairs(std::size_t n, const PairFn& fn) {
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = i + 1; j < n; ++j) fn(i, j);
}

// Per pair: d = minimum image of r_i - r_j; skip unless r2 < rc2; s2 = sig2/r2,
// s6 = s2^3, u = 4 eps (s6^2 - s6), w = 24 eps (2 s6^2 - s6); LFS shifts
// u -= u(rc) + (r - rc) u'(rc) and w += r u'(rc); E += u, W += w; f = (w/r2) d
// added to i and subtracted from j (lennard_jones.cpp:50-88).
EV lj_pair_sum(const LJ& lj, const State& s, State* sink,
               const std::function<void(const PairFn&)>& source, EV* compensated, long* pairs) {
    EV ev;
    CompSum ce, cw;
    long np = 0;
    const double sig2 = lj.sigma * lj.sigma;
    const double rc2 = lj.cutoff * lj.cutoff;
    source([&](std::size_t i, std::size_t j) {
        const V3 ri = s.r(i), rj = s.r(j);
        const V3 d = min_image(s.L, V3{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z});
        const double r2 = dot3(d, d);
        if (!(r2 < rc2)) return;
        const double s2 = sig2 / r2;
        const double s6 = s2 * s2 * s2;
        double u = 4.0 * lj.epsilon * (s6 * s6 - s6);
        double w = 24.0 * lj.epsilon * (2.0 * s6 * s6 - s6);
        if (lj.lfs) {
            const double r = std::sqrt(r2);
            u = u - lj.u_rc - (r - lj.cutoff) * lj.dudr_rc;
            w = w + r * lj.dudr_rc;
        }
        ev.energy += u;
        ev.virial += w;
        if (compensated != nullptr) {
            ce.add(u);
            cw.add(w);
        }
        ++np;
        if (sink != nullptr) {
            const double c = w / r2;
            const V3 f{c * d.x, c * d.y, c * d.z};
            sink->fx[i] += f.x;
            sink->fy[i] += f.y;
            sink->fz[i] += f.z;
            // the reference adds (-1.0 * f) to j
            sink->fx[j] += -1.0 * f.x;
            sink->fy[j] += -1.0 * f.y;
            sink->fz[j] += -1.0 * f.z;
        }
    });
    if (compensated != nullptr) {
        compensated->energy = ce.value();
        compensated->virial = cw.value();
    }
    if (pairs != nullptr) *pairs = np;
    return ev;
}

EV lj_energy_virial(const LJ& lj, const State& s) {
    return lj_pair_sum(lj, s, nullptr, [&](const PairFn& fn) { brute_pairs(s.n(), fn); });
}

EV lj_forces_brute(const LJ& lj, State& s) {
    s.clear_forces();
    return lj_pair_sum(lj, s, &s, [&](const PairFn& fn) { brute_pairs(s.n(), fn); });
}

// Refuses an unusable cell list (lennard_jones.cpp:99-114).
EV lj_forces_cells(const LJ& lj, State& s, const Cells& c) {
    if (!c.usable())
        throw std::invalid_argument("lj_forces_cells: cell list is not usable (< 3 cells per side)");
    s.clear_forces();
    return lj_pair_sum(lj, s, &s, [&](const PairFn& fn) { c.pairs(fn); });
}

EV lj_forces_verlet(const LJ& lj, State& s, const Verlet& v) {
    s.clear_forces();
    return lj_pair_sum(lj, s, &s, [&](const PairFn& fn) { v.pairs(fn); });
}

// U_LRC = (8/3) pi N rho eps sig^3 [ (s/rc)^9 / 3 - (s/rc)^3 ], 0 under LFS
// (lennard_jones.cpp:121-135).
double lj_lrc_energy(const LJ& lj, const State& s) {
    if (lj.lfs) return 0.0;
    const double n = static_cast<double>(s.n());
    const double rho = n / s.volume();
    const double t = lj.sigma / lj.cutoff;
    const double t3 = t * t * t;
    const double t9 = t3 * t3 * t3;
    const double sig3 = lj.sigma * lj.sigma * lj.sigma;
    return (8.0 / 3.0) * kPi * n * rho * lj.epsilon * sig3 * ((1.0 / 3.0) * t9 - t3);
}

// P_LRC = (16/3) pi rho^2 eps sig^3 [ (2/3)(s/rc)^9 - (s/rc)^3 ], 0 under LFS
// (lennard_jones.cpp:137-151).
double lj_lrc_pressure(const LJ& lj, const State& s) {
    if (lj.lfs) return 0.0;
    const double n = static_cast<double>(s.n());
    const double rho = n / s.volume();
    const double t = lj.sigma / lj.cutoff;
    const double t3 = t * t * t;
    const double t9 = t3 * t3 * t3;
    const double sig3 = lj.sigma * lj.sigma * lj.sigma;
    return (16.0 / 3.0) * kPi * rho * rho * lj.epsilon * sig3 * ((2.0 / 3.0) * t9 - t3);
}

// ---- linked cells -----------------------------------------------------------

// n_d = max(1, floor(L_d / cutoff)) (cell_list.cpp:8-14).
Cells::Cells(const double box[3], double cutoff) : width_req(cutoff) {
    for (int d = 0; d < 3; ++d) {
        L[d] = box[d];
        nc[d] = std::max(1, static_cast<int>(std::floor(L[d] / cutoff)));
    }
}

// Cell of the folded position, floor((r + L/2)/width) clamped to [0, n-1];
// each atom is pushed onto the front of its cell's chain in index order
// (cell_list.cpp:16-43).
void Cells::build(const State& s) {
    const std::size_t ncell =
        static_cast<std::size_t>(nc[0]) * static_cast<std::size_t>(nc[1]) * static_cast<std::size_t>(nc[2]);
    head.assign(ncell, -1);
    next.assign(s.n(), -1);
    const double w[3] = {L[0] / nc[0], L[1] / nc[1], L[2] / nc[2]};
    for (std::size_t i = 0; i < s.n(); ++i) {
        const V3 r = min_image(L, s.r(i));
        const double c[3] = {r.x, r.y, r.z};
        int idx[3];
        for (int d = 0; d < 3; ++d) {
            int k = static_cast<int>(std::floor((c[d] + 0.5 * L[d]) / w[d]));
            idx[d] = k < 0 ? 0 : (k >= nc[d] ? nc[d] - 1 : k);
        }
        const int cell = index(idx[0], idx[1], idx[2]);
        next[i] = head[static_cast<std::size_t>(cell)];
        head[static_cast<std::size_t>(cell)] = static_cast<int>(i);
    }
}

// Cells visited with ix outermost and iz innermost; each cell pairs its own
// chain (an entry with every later entry) and then its chain against each of
// the 13 forward neighbours with periodic index wrap; pairs are reported as
// (min, max) (cell_list.hpp:55-108).
void Cells::pairs(const PairFn& fn) const {
    static const int fwd[13][3] = {{1, 0, 0},  {1, 1, 0},  {0, 1, 0},   {-1, 1, 0}, {0, 0, 1},
                                   {1, 0, 1},  {-1, 0, 1}, {0, 1, 1},   {0, -1, 1}, {1, 1, 1},
                                   {1, -1, 1}, {-1, 1, 1}, {-1, -1, 1}};
    auto emit = [&](int a, int b) {
        if (a < b)
            fn(static_cast<std::size_t>(a), static_cast<std::size_t>(b));
        else
            fn(static_cast<std::size_t>(b), static_cast<std::size_t>(a));
    };
    auto wrapi = [](int v, int n) {
        const int r = v % n;
        return r < 0 ? r + n : r;
    };
    for (int ix = 0; ix < nc[0]; ++ix) {
        for (int iy = 0; iy < nc[1]; ++iy) {
            for (int iz = 0; iz < nc[2]; ++iz) {
                const int c = index(ix, iy, iz);
                for (int a = head[static_cast<std::size_t>(c)]; a != -1; a = next[static_cast<std::size_t>(a)])
                    for (int b = next[static_cast<std::size_t>(a)]; b != -1; b = next[static_cast<std::size_t>(b)])
                        emit(a, b);
                for (const auto& o : fwd) {
                    const int nb = index(wrapi(ix + o[0], nc[0]), wrapi(iy + o[1], nc[1]),
                                         wrapi(iz + o[2], nc[2]));
                    for (int a = head[static_cast<std::size_t>(c)]; a != -1; a = next[static_cast<std::size_t>(a)])
                        for (int b = head[static_cast<std::size_t>(nb)]; b != -1; b = next[static_cast<std::size_t>(b)])
                            emit(a, b);
                }
            }
        }
    }
}

// ---- Verlet list ------------------------------------------------------------

// Cells sized at cutoff + skin (verlet_list.cpp:5-6).
Verlet::Verlet(const double box[3], double rc, double sk)
    : cutoff(rc), skin(sk), cells(box, rc + sk) {
    for (int d = 0; d < 3; ++d) L[d] = box[d];
}

// Keep candidate pairs with |minimum image|^2 < (rc+skin)^2, from the cell
// list when usable and from all i<j otherwise; record reference positions;
// count the build (verlet_list.cpp:8-42).
void Verlet::build(const State& s) {
    list.clear();
    const double lc2 = (cutoff + skin) * (cutoff + skin);
    auto keep = [&](std::size_t i, std::size_t j) {
        const V3 ri = s.r(i), rj = s.r(j);
        const V3 d = min_image(L, V3{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z});
        if (dot3(d, d) < lc2) list.emplace_back(i, j);
    };
    if (cells.usable()) {
        cells.build(s);
        cells.pairs(keep);
    } else {
        brute_pairs(s.n(), keep);
    }
    ref.resize(s.n());
    for (std::size_t i = 0; i < s.n(); ++i) ref[i] = s.r(i);
    ++builds;
}

// Two largest minimum-image displacements since the build; rebuild when their
// sum exceeds the skin (verlet_list.cpp:44-62).
bool Verlet::needs_rebuild(const State& s) const {
    double first = 0.0, second = 0.0;
    for (std::size_t i = 0; i < s.n(); ++i) {
        const V3 r = s.r(i);
        const V3 d = min_image(L, V3{r.x - ref[i].x, r.y - ref[i].y, r.z - ref[i].z});
        const double len = std::sqrt(dot3(d, d));
        if (len > first) {
            second = first;
            first = len;
        } else if (len > second) {
            second = len;
        }
    }
    return (first + second) > skin;
}

}  // namespace oracle
