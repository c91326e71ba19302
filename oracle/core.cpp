// Oracle: particle state, reductions, seeded inputs. Test infrastructure only
// (see oracle.hpp).
#include "oracle.hpp"

namespace oracle {

void State::resize(std::size_t n) {
    for (auto* a : {&x, &y, &z, &vx, &vy, &vz, &fx, &fy, &fz, &q}) a->assign(n, 0.0);
    m.assign(n, 1.0);
    mol.resize(n);
    for (std::size_t i = 0; i < n; ++i) mol[i] = static_cast<long>(i);
}

void State::clear_forces() {
    std::fill(fx.begin(), fx.end(), 0.0);
    std::fill(fy.begin(), fy.end(), 0.0);
    std::fill(fz.begin(), fz.end(), 0.0);
}

// Sequential sum of 0.5*m*v2 with v2 = (vx^2 + vy^2) + vz^2
// (core/system.cpp:124-131).
double kinetic_energy(const State& s) {
    double acc = 0.0;
    for (std::size_t i = 0; i < s.n(); ++i) {
        double v2 = s.vx[i] * s.vx[i];
        v2 = v2 + s.vy[i] * s.vy[i];
        v2 = v2 + s.vz[i] * s.vz[i];
        acc += (0.5 * s.m[i]) * v2;
    }
    return acc;
}

// core/system.cpp:137-145
V3 total_momentum(const State& s) {
    V3 p{0.0, 0.0, 0.0};
    for (std::size_t i = 0; i < s.n(); ++i) {
        p.x += s.m[i] * s.vx[i];
        p.y += s.m[i] * s.vy[i];
        p.z += s.m[i] * s.vz[i];
    }
    return p;
}

// Subtract P/M from every velocity; no-op for zero total mass
// (core/system.cpp:147-164).
void remove_com_motion(State& s) {
    double mass = 0.0;
    for (double mi : s.m) mass += mi;
    if (mass == 0.0) return;
    const V3 p = total_momentum(s);
    const V3 vc{p.x / mass, p.y / mass, p.z / mass};
    for (std::size_t i = 0; i < s.n(); ++i) {
        s.vx[i] -= vc.x;
        s.vy[i] -= vc.y;
        s.vz[i] -= vc.z;
    }
}

// Marsaglia polar method over uniform() in (0,1]; rejects s >= 1 and s == 0;
// returns u*mult and keeps v*mult (util/random.hpp:52-69).
double Rng::normal() {
    if (spare_ok_) {
        spare_ok_ = false;
        return spare_;
    }
    double a, b, s;
    for (;;) {
        a = 2.0 * uniform() - 1.0;
        b = 2.0 * uniform() - 1.0;
        s = a * a + b * b;
        if (s < 1.0 && s != 0.0) break;
    }
    const double mult = std::sqrt(-2.0 * std::log(s) / s);
    spare_ = b * mult;
    spare_ok_ = true;
    return a * mult;
}

// 4*cells^3 atoms, L = cbrt(N/rho), lattice constant a = L/cells, basis
// {0, (0,.5,.5), (.5,0,.5), (.5,.5,0)}, position a*(i+b) - L/2 folded into the
// box, velocities sqrt(T/m)*normal() drawn x,y,z per atom in (ix,iy,iz,basis)
// loop order, mass 1, then centre-of-mass motion removed
// (util/lattice.cpp:15-64).
State fcc_lattice(int cells, double density, double temperature, std::uint64_t seed) {
    const int n = 4 * cells * cells * cells;
    const double len = std::cbrt(static_cast<double>(n) / density);
    const double a = len / static_cast<double>(cells);
    static const double basis[4][3] = {
        {0.0, 0.0, 0.0}, {0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}};

    State s;
    s.L[0] = s.L[1] = s.L[2] = len;
    s.resize(static_cast<std::size_t>(n));
    Rng rng(seed);
    const double sv = std::sqrt(temperature / 1.0);
    std::size_t k = 0;
    for (int ix = 0; ix < cells; ++ix) {
        for (int iy = 0; iy < cells; ++iy) {
            for (int iz = 0; iz < cells; ++iz) {
                for (const auto& b : basis) {
                    V3 r{a * (static_cast<double>(ix) + b[0]) - 0.5 * len,
                         a * (static_cast<double>(iy) + b[1]) - 0.5 * len,
                         a * (static_cast<double>(iz) + b[2]) - 0.5 * len};
                    s.set_r(k, min_image(s.L, r));
                    const double gx = rng.normal();
                    const double gy = rng.normal();
                    const double gz = rng.normal();
                    s.set_v(k, V3{sv * gx, sv * gy, sv * gz});
                    s.m[k] = 1.0;
                    ++k;
                }
            }
        }
    }
    remove_com_motion(s);
    return s;
}

}  // namespace oracle
