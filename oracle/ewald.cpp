// Oracle: Ewald electrostatics, the SPC/E force field, rigid-water SHAKE /
// RATTLE, and the per-type-pair LJ used for parameter derivatives. Test
// infrastructure only (see oracle.hpp).
#include <algorithm>
#include <map>
#include <string>

#include "oracle.hpp"

namespace oracle {

namespace {
constexpr double kPi = 3.14159265358979323846;
}

// e^2/(4 pi eps0 kB) * 1e10 from the 2010 CODATA constants the reference
// names (forcefield/ewald.hpp:10-13, ewald.cpp:17-21).
double coulomb_prefactor() {
    const double e = 1.602176565e-19, eps0 = 8.854187817e-12, kb = 1.3806488e-23;
    return (e * e) / (4.0 * kPi * eps0) * 1e10 / kb;
}

// Tin-foil Ewald (forcefield/ewald.cpp:33-175):
//  real   C qq erfc(a r)/r over intermolecular i<j with r < rc (minimum image),
//         w = C qq (2a/sqrt(pi) e^{-a^2 r^2} + erfc(a r)/r), f = (w/r^2) d;
//  self   -C a/sqrt(pi) sum q^2;
//  intra  -C qq erf(a r)/r over pairs of each molecule (minimum image),
//         w = C qq (2a/sqrt(pi) e^{-a^2 r^2} - erf(a r)/r);
//  recip  sum over n in [-kmax,kmax]^3, 0 < |n|^2 < kmax^2 + 2, of
//         C (2 pi/V) e^{-k^2/4a^2}/k^2 |S(k)|^2, S = sum q e^{i k.r},
//         f_i += 2 factor q_i (Re sin - Im cos) k.
EwaldTerms ewald(const State& s, double rc, int kmax, double alpha_in, bool accumulate_forces,
                 State* sink) {
    EwaldTerms t;
    const std::size_t n = s.n();
    const double minL = std::min({s.L[0], s.L[1], s.L[2]});
    const double alpha = alpha_in > 0.0 ? alpha_in : 5.6 / minL;
    const double C = coulomb_prefactor();
    const double rc2 = rc * rc;
    const double c2a = 2.0 * alpha / std::sqrt(kPi);

    for (std::size_t i = 0; i < n; ++i) {
        for (std::size_t j = i + 1; j < n; ++j) {
            if (s.mol[i] == s.mol[j]) continue;
            const V3 ri = s.r(i), rj = s.r(j);
            const V3 d = min_image(s.L, V3{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z});
            const double r2 = dot3(d, d);
            if (!(r2 < rc2)) continue;
            const double r = std::sqrt(r2);
            const double qq = s.q[i] * s.q[j];
            const double ec = std::erfc(alpha * r);
            t.real += C * qq * ec / r;
            if (accumulate_forces) {
                const double w = C * qq * (c2a * std::exp(-alpha * alpha * r2) + ec / r);
                const double c = w / r2;
                sink->fx[i] += c * d.x; sink->fy[i] += c * d.y; sink->fz[i] += c * d.z;
                sink->fx[j] += -1.0 * (c * d.x); sink->fy[j] += -1.0 * (c * d.y); sink->fz[j] += -1.0 * (c * d.z);
            }
        }
    }

    double q2 = 0.0;
    for (std::size_t i = 0; i < n; ++i) q2 += s.q[i] * s.q[i];
    t.self = -C * alpha / std::sqrt(kPi) * q2;

    // Group atoms by molecule id in index order. The reference iterates an
    // unordered_map, whose molecule order is unspecified, so only the set of
    // per-molecule terms (not their summation order) is reproduced.
    std::map<long, std::vector<std::size_t>> groups;
    for (std::size_t i = 0; i < n; ++i) groups[s.mol[i]].push_back(i);
    for (const auto& g : groups) {
        const auto& at = g.second;
        for (std::size_t a = 0; a < at.size(); ++a) {
            for (std::size_t b = a + 1; b < at.size(); ++b) {
                const std::size_t i = at[a], j = at[b];
                const V3 ri = s.r(i), rj = s.r(j);
                const V3 d = min_image(s.L, V3{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z});
                const double r2 = dot3(d, d);
                const double r = std::sqrt(r2);
                const double qq = s.q[i] * s.q[j];
                const double ef = std::erf(alpha * r);
                t.intra += -C * qq * ef / r;
                if (accumulate_forces) {
                    const double w = C * qq * (c2a * std::exp(-alpha * alpha * r2) - ef / r);
                    const double c = w / r2;
                    sink->fx[i] += c * d.x; sink->fy[i] += c * d.y; sink->fz[i] += c * d.z;
                    sink->fx[j] += -1.0 * (c * d.x); sink->fy[j] += -1.0 * (c * d.y); sink->fz[j] += -1.0 * (c * d.z);
                }
            }
        }
    }

    const double V = s.volume();
    const double k0[3] = {2.0 * kPi / s.L[0], 2.0 * kPi / s.L[1], 2.0 * kPi / s.L[2]};
    const int n2max = kmax * kmax + 2;
    std::vector<double> cs(n), sn(n);
    for (int nx = -kmax; nx <= kmax; ++nx) {
        for (int ny = -kmax; ny <= kmax; ++ny) {
            for (int nz = -kmax; nz <= kmax; ++nz) {
                const int n2 = nx * nx + ny * ny + nz * nz;
                if (n2 == 0 || n2 >= n2max) continue;
                const double kx = k0[0] * nx, ky = k0[1] * ny, kz = k0[2] * nz;
                const double k2 = kx * kx + ky * ky + kz * kz;
                double re = 0.0, im = 0.0;
                for (std::size_t i = 0; i < n; ++i) {
                    const double ph = kx * s.x[i] + ky * s.y[i] + kz * s.z[i];
                    cs[i] = std::cos(ph);
                    sn[i] = std::sin(ph);
                    re += s.q[i] * cs[i];
                    im += s.q[i] * sn[i];
                }
                const double factor = C * (2.0 * kPi / V) * std::exp(-k2 / (4.0 * alpha * alpha)) / k2;
                t.recip += factor * (re * re + im * im);
                if (accumulate_forces) {
                    for (std::size_t i = 0; i < n; ++i) {
                        const double c = factor * 2.0 * s.q[i] * (re * sn[i] - im * cs[i]);
                        sink->fx[i] += c * kx; sink->fy[i] += c * ky; sink->fz[i] += c * kz;
                    }
                }
            }
        }
    }
    return t;
}

namespace {
// Oxygen copies (every third site) in the same box (spce_forcefield.cpp:8-20).
State oxygens_of(const State& s, double mO) {
    if (s.n() % 3 != 0) throw std::invalid_argument("spce: atom count is not a multiple of 3");
    State o;
    for (int d = 0; d < 3; ++d) o.L[d] = s.L[d];
    o.resize(s.n() / 3);
    for (std::size_t i = 0, k = 0; i < s.n(); i += 3, ++k) {
        o.set_r(k, s.r(i));
        o.m[k] = mO;
    }
    return o;
}
}  // namespace

// Dispersion on gathered oxygens (Truncated LJ + LRC), forces zeroed and the
// oxygen forces scattered back, then Ewald accumulated on top
// (spce_forcefield.cpp:26-41).
double spce_forces(const Spce& p, double rc, int kmax, double alpha, State& s) {
    State o = oxygens_of(s, p.mO);
    const LJ lj(p.sigmaO, p.epsilonO, rc, false);
    const EV disp = lj_forces_brute(lj, o);
    const double lrc = lj_lrc_energy(lj, o);
    s.clear_forces();
    for (std::size_t i = 0, k = 0; i < s.n(); i += 3, ++k) {
        s.fx[i] += o.fx[k]; s.fy[i] += o.fy[k]; s.fz[i] += o.fz[k];
    }
    const EwaldTerms e = ewald(s, rc, kmax, alpha, true, &s);
    return disp.energy + lrc + e.total();
}

double spce_energy(const Spce& p, double rc, int kmax, double alpha, const State& s) {
    const State o = oxygens_of(s, p.mO);
    const LJ lj(p.sigmaO, p.epsilonO, rc, false);
    return lj_energy_virial(lj, o).energy + lj_lrc_energy(lj, o) +
           ewald(s, rc, kmax, alpha, false, nullptr).total();
}

// d_HH = 2 r_OH sin(theta/2); argument checks as the reference
// (constrain/rigid_water.cpp:30-52).
Rigid::Rigid(double oh_dist, double hoh_deg, double tolerance, int max_iterations)
    : oh(oh_dist), hh(0.0), tol(tolerance), max_iter(max_iterations) {
    if (!(oh_dist > 0.0)) throw std::invalid_argument("Rigid: oh distance must be positive");
    if (!(hoh_deg > 0.0) || !(hoh_deg < 180.0))
        throw std::invalid_argument("Rigid: angle must lie in (0, 180)");
    if (!(tolerance > 0.0)) throw std::invalid_argument("Rigid: tolerance must be positive");
    if (max_iterations < 1) throw std::invalid_argument("Rigid: max iterations must be >= 1");
    hh = 2.0 * oh_dist * std::sin(0.5 * hoh_deg * kPi / 180.0);
}

// 6 per molecule, minus 3 with the centre of mass removed
// (rigid_water.cpp:54-59).
std::size_t Rigid::dof(std::size_t molecules, bool com_removed) {
    const std::size_t total = 6 * molecules;
    return com_removed ? (total >= 3 ? total - 3 : 0) : total;
}

// Gauss-Seidel SHAKE over (O,H1), (O,H2), (H1,H2) per molecule until every
// relative violation |r^2 - d^2|/d^2 <= tol; correction along the reference
// bond, g = s / (2 (1/mi + 1/mj) r.rref), r_i -= rref g/mi, r_j += rref g/mj
// (constrain/rigid_water.cpp:61-116).
int Rigid::shake(State& s, const std::vector<V3>& reference) const {
    const std::size_t n = s.n();
    if (n % 3 != 0) throw std::runtime_error("Rigid: atom count is not a multiple of 3");
    if (reference.size() != n) throw std::runtime_error("Rigid: reference size mismatch");
    static const std::size_t cp[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    const double target[3] = {oh, oh, hh};
    for (int it = 0; it < max_iter; ++it) {
        double worst = 0.0;
        for (std::size_t m0 = 0; m0 < n; m0 += 3) {
            for (int c = 0; c < 3; ++c) {
                const std::size_t i = m0 + cp[c][0], j = m0 + cp[c][1];
                const double d2 = target[c] * target[c];
                const V3 ri = s.r(i), rj = s.r(j);
                const V3 rij = min_image(s.L, V3{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z});
                const double sdev = dot3(rij, rij) - d2;
                const double viol = std::abs(sdev) / d2;
                worst = std::max(worst, viol);
                if (viol <= tol) continue;
                const V3 a = reference[i], b = reference[j];
                const V3 rr = min_image(s.L, V3{a.x - b.x, a.y - b.y, a.z - b.z});
                const double inv = 1.0 / s.m[i] + 1.0 / s.m[j];
                const double den = 2.0 * inv * dot3(rij, rr);
                if (std::abs(den) < 1e-30)
                    throw std::runtime_error("Rigid: SHAKE constraint became singular");
                const double g = sdev / den;
                const double gi = g / s.m[i], gj = g / s.m[j];
                s.set_r(i, V3{ri.x - rr.x * gi, ri.y - rr.y * gi, ri.z - rr.z * gi});
                const V3 rj2 = s.r(j);
                s.set_r(j, V3{rj2.x + rr.x * gj, rj2.y + rr.y * gj, rj2.z + rr.z * gj});
            }
        }
        if (worst <= tol) return it + 1;
    }
    throw std::runtime_error("Rigid: SHAKE failed to converge in " + std::to_string(max_iter) +
                             " iterations");
}

// RATTLE: remove the relative velocity along each bond until every
// |r.v|/d^2 <= tol; k = r.v / ((1/mi + 1/mj) |r|^2)
// (constrain/rigid_water.cpp:118-161).
int Rigid::rattle(State& s) const {
    const std::size_t n = s.n();
    if (n % 3 != 0) throw std::runtime_error("Rigid: atom count is not a multiple of 3");
    static const std::size_t cp[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    const double target[3] = {oh, oh, hh};
    for (int it = 0; it < max_iter; ++it) {
        double worst = 0.0;
        for (std::size_t m0 = 0; m0 < n; m0 += 3) {
            for (int c = 0; c < 3; ++c) {
                const std::size_t i = m0 + cp[c][0], j = m0 + cp[c][1];
                const double d2 = target[c] * target[c];
                const V3 ri = s.r(i), rj = s.r(j);
                const V3 rij = min_image(s.L, V3{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z});
                const V3 vi = s.v(i), vj = s.v(j);
                const V3 vij{vi.x - vj.x, vi.y - vj.y, vi.z - vj.z};
                const double rv = dot3(rij, vij);
                const double sc = std::abs(rv) / d2;
                worst = std::max(worst, sc);
                if (sc <= tol) continue;
                const double inv = 1.0 / s.m[i] + 1.0 / s.m[j];
                const double k = rv / (inv * dot3(rij, rij));
                const double ki = k / s.m[i], kj = k / s.m[j];
                s.set_v(i, V3{vi.x - rij.x * ki, vi.y - rij.y * ki, vi.z - rij.z * ki});
                const V3 vj2 = s.v(j);
                s.set_v(j, V3{vj2.x + rij.x * kj, vj2.y + rij.y * kj, vj2.z + rij.z * kj});
            }
        }
        if (worst <= tol) return it + 1;
    }
    throw std::runtime_error("Rigid: RATTLE failed to converge in " + std::to_string(max_iter) +
                             " iterations");
}

// Per-type-pair LJ over all i<j (minimum image): pair (i,j) uses the
// sigma/epsilon/cutoff of its unordered type pair and the single-type
// arithmetic of lennard_jones.cpp:50-88. Also returns per type pair the
// energy E_ab, the virial W_ab and Y_ab = sum 24 eps (24 s^12 - 6 s^6), from
// which dE/deps_ab = E_ab/eps_ab, dE/dsig_ab = W_ab/sig_ab,
// dW/deps_ab = W_ab/eps_ab, dW/dsig_ab = Y_ab/sig_ab at fixed cutoff
// (Truncated scheme).
MixEV mix_lj(const MixLJ& p, const State& s, const std::vector<int>& type, bool forces,
             State* sink) {
    MixEV out;
    const int nt = p.ntypes;
    out.e_ab.assign(static_cast<std::size_t>(nt * nt), 0.0);
    out.w_ab = out.e_ab;
    out.y_ab = out.e_ab;
    if (forces) sink->clear_forces();
    std::vector<LJ> pot(static_cast<std::size_t>(nt * nt));
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) {
            const std::size_t k = static_cast<std::size_t>(a * nt + b);
            pot[k] = LJ(p.sigma[k], p.epsilon[k], p.cutoff[k], p.lfs);
        }
    for (std::size_t i = 0; i < s.n(); ++i) {
        for (std::size_t j = i + 1; j < s.n(); ++j) {
            const int ta = std::min(type[i], type[j]), tb = std::max(type[i], type[j]);
            const std::size_t k = static_cast<std::size_t>(ta * nt + tb);
            const LJ& lj = pot[k];
            const V3 ri = s.r(i), rj = s.r(j);
            const V3 d = min_image(s.L, V3{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z});
            const double r2 = dot3(d, d);
            if (!(r2 < lj.cutoff * lj.cutoff)) continue;
            const double s2 = (lj.sigma * lj.sigma) / r2;
            const double s6 = s2 * s2 * s2;
            double u = 4.0 * lj.epsilon * (s6 * s6 - s6);
            double w = 24.0 * lj.epsilon * (2.0 * s6 * s6 - s6);
            const double y = 24.0 * lj.epsilon * (24.0 * s6 * s6 - 6.0 * s6);
            if (lj.lfs) {
                const double r = std::sqrt(r2);
                u = u - lj.u_rc - (r - lj.cutoff) * lj.dudr_rc;
                w = w + r * lj.dudr_rc;
            }
            out.energy += u;
            out.virial += w;
            out.e_ab[k] += u;
            out.w_ab[k] += w;
            out.y_ab[k] += y;
            if (forces) {
                const double c = w / r2;
                sink->fx[i] += c * d.x; sink->fy[i] += c * d.y; sink->fz[i] += c * d.z;
                sink->fx[j] -= c * d.x; sink->fy[j] -= c * d.y; sink->fz[j] -= c * d.z;
            }
        }
    }
    return out;
}

std::vector<double> mix_minimize(const MixLJ& p, State& s, const std::vector<int>& type,
                                 int nsteps, double step, double max_disp) {
    std::vector<double> e;
    e.push_back(mix_lj(p, s, type, true, &s).energy);
    for (int k = 0; k < nsteps; ++k) {
        double fmax = 0.0;
        for (std::size_t i = 0; i < s.n(); ++i) {
            const V3 f{s.fx[i], s.fy[i], s.fz[i]};
            fmax = std::max(fmax, std::sqrt(dot3(f, f)));
        }
        double g = step;
        if (fmax > 0.0 && max_disp / fmax < g) g = max_disp / fmax;
        for (std::size_t i = 0; i < s.n(); ++i) {
            s.x[i] = fold(std::fma(g, s.fx[i], s.x[i]), s.L[0]);
            s.y[i] = fold(std::fma(g, s.fy[i], s.y[i]), s.L[1]);
            s.z[i] = fold(std::fma(g, s.fz[i], s.z[i]), s.L[2]);
        }
        e.push_back(mix_lj(p, s, type, true, &s).energy);
    }
    return e;
}

void mix_vv(const MixLJ& p, State& s, const std::vector<int>& type, double dt, int nsteps,
            std::vector<double>& energy, std::vector<double>& virial) {
    energy.clear();
    virial.clear();
    auto eval = [&](State& st) {
        const MixEV r = mix_lj(p, st, type, true, &st);
        energy.push_back(r.energy);
        virial.push_back(r.virial);
        return r.energy;
    };
    eval(s);
    for (int k = 0; k < nsteps; ++k) vv_step(s, dt, eval);
}

}  // namespace oracle
