// Oracle: velocity Verlet, BAOAB Langevin, Nose-Hoover chain and the
// RATTLE-form constrained velocity Verlet. Test infrastructure only (see
// oracle.hpp).
#include <array>

#include "oracle.hpp"

namespace oracle {

namespace {
// v <- fma(h/m, f, v) per component (integrate/velocity_verlet.hpp:17-23, 55-58).
void half_kick(State& s, double half_dt) {
    for (std::size_t i = 0; i < s.n(); ++i) {
        const double c = half_dt * (1.0 / s.m[i]);
        s.vx[i] = std::fma(c, s.fx[i], s.vx[i]);
        s.vy[i] = std::fma(c, s.fy[i], s.vy[i]);
        s.vz[i] = std::fma(c, s.fz[i], s.vz[i]);
    }
}
// r <- fold(fma(h, v, r)) per component (velocity_verlet.hpp:61-64).
void drift(State& s, double h) {
    for (std::size_t i = 0; i < s.n(); ++i) {
        const V3 r{std::fma(h, s.vx[i], s.x[i]), std::fma(h, s.vy[i], s.y[i]),
                   std::fma(h, s.vz[i], s.z[i])};
        s.set_r(i, min_image(s.L, r));
    }
}
}  // namespace

// B(dt/2) A(dt) forces B(dt/2) (integrate/velocity_verlet.hpp:48-74).
void vv_step(State& s, double dt, const ForceFn& forces) {
    const double h = 0.5 * dt;
    half_kick(s, h);
    drift(s, dt);
    forces(s);
    half_kick(s, h);
}

// B A O A forces B with c1 = exp(-gamma dt), c2 = sqrt(1 - c1^2); the O step
// draws normal() for x, y, z of each atom in index order and sets
// v = c1 v + (c2 sqrt(T/m)) xi (integrate/langevin.hpp:36-88).
void Baoab::step(State& s, const ForceFn& forces) {
    const double h = 0.5 * dt;
    const double c1 = std::exp(-gamma * dt);
    const double c2 = std::sqrt(1.0 - c1 * c1);
    half_kick(s, h);
    drift(s, h);
    for (std::size_t i = 0; i < s.n(); ++i) {
        const double sd = std::sqrt(T / s.m[i]);
        const double g0 = rng.normal();
        const double g1 = rng.normal();
        const double g2 = rng.normal();
        const double k = c2 * sd;
        s.vx[i] = c1 * s.vx[i] + g0 * k;
        s.vy[i] = c1 * s.vy[i] + g1 * k;
        s.vz[i] = c1 * s.vz[i] + g2 * k;
    }
    drift(s, h);
    forces(s);
    half_kick(s, h);
}

// Q_1 = dof T tau^2, Q_k = T tau^2; chain length >= 1 (nose_hoover.cpp:32-50).
NoseHoover::NoseHoover(double dt_, double T_, double tau_, int M_, std::size_t dof_)
    : dt(dt_), T(T_), tau(tau_), M(M_), dof(dof_) {
    if (M_ < 1) throw std::invalid_argument("NoseHoover: chain length must be >= 1");
    Q.assign(static_cast<std::size_t>(M), T * tau * tau);
    Q[0] = static_cast<double>(dof) * T * tau * tau;
    xi.assign(static_cast<std::size_t>(M), 0.0);
    p.assign(static_cast<std::size_t>(M), 0.0);
}

// sum p^2/(2Q) + dof T xi_1 + T sum_{k>1} xi_k (nose_hoover.cpp:52-61).
double NoseHoover::conserved_contribution() const {
    double acc = static_cast<double>(dof) * T * xi[0];
    for (std::size_t k = 1; k < xi.size(); ++k) acc += T * xi[k];
    for (std::size_t k = 0; k < p.size(); ++k) acc += 0.5 * p[k] * p[k] / Q[k];
    return acc;
}

// One chain propagation of length delta: three Suzuki-Yoshida sub-steps with
// w1 = 1/(2 - 2^(1/3)), w2 = 1 - 2 w1; each sub-step is a backward sweep of
// momentum updates (M-1..0), a drift of every xi, a velocity rescale by
// exp(-d p_1/Q_1) and a forward sweep (0..M-1). A momentum update damps p_j
// by exp(-(d/4) p_{j+1}/Q_{j+1}) on both sides of the force kick (d/2) G_j,
// G_0 = 2 KE - dof T, G_j = p_{j-1}^2/Q_{j-1} - T; KE is recomputed from
// the velocities each time (nose_hoover.cpp:63-124).
void NoseHoover::half_chain(State& s, double delta) {
    const double w1 = 1.0 / (2.0 - std::cbrt(2.0));
    const std::array<double, 3> ws = {w1, 1.0 - 2.0 * w1, w1};
    const std::size_t m = static_cast<std::size_t>(M);
    auto link = [&](std::size_t j, double d2, double d4) {
        const bool next = j + 1 < m;
        if (next) p[j] *= std::exp(-d4 * p[j + 1] / Q[j + 1]);
        const double g = (j == 0) ? 2.0 * kinetic_energy(s) - static_cast<double>(dof) * T
                                  : p[j - 1] * p[j - 1] / Q[j - 1] - T;
        p[j] += d2 * g;
        if (next) p[j] *= std::exp(-d4 * p[j + 1] / Q[j + 1]);
    };
    for (double w : ws) {
        const double d = w * delta;
        const double d2 = 0.5 * d, d4 = 0.25 * d;
        for (std::size_t j = m; j-- > 0;) link(j, d2, d4);
        for (std::size_t k = 0; k < m; ++k) xi[k] += d * p[k] / Q[k];
        const double factor = std::exp(-d * p[0] / Q[0]);
        for (std::size_t i = 0; i < s.n(); ++i) {
            s.vx[i] = s.vx[i] * factor;
            s.vy[i] = s.vy[i] * factor;
            s.vz[i] = s.vz[i] * factor;
        }
        for (std::size_t j = 0; j < m; ++j) link(j, d2, d4);
    }
}

// chain(dt/2) . VV(dt) . chain(dt/2) (nose_hoover.hpp:37-42).
void NoseHoover::step(State& s, const ForceFn& forces) {
    half_chain(s, 0.5 * dt);
    vv_step(s, dt, forces);
    half_chain(s, 0.5 * dt);
}

// Constrained velocity Verlet: kick v + f (dt/2/m) (plain multiply-add),
// drift without wrap, SHAKE against the start-of-step positions, velocity
// correction by the minimum-imaged SHAKE shift over dt, forces, kick, RATTLE
// (integrate/constrained_velocity_verlet.hpp:27-57).
namespace {
template <class PosFn, class VelFn>
void cvv_impl(State& s, double dt, const PosFn& constrain_positions, const VelFn& constrain_velocities,
              const ForceFn& forces) {
    const std::size_t n = s.n();
    std::vector<V3> start(n), free_pos(n);
    for (std::size_t i = 0; i < n; ++i) start[i] = s.r(i);
    for (std::size_t i = 0; i < n; ++i) {
        const double c = 0.5 * dt / s.m[i];
        s.vx[i] = s.vx[i] + s.fx[i] * c;
        s.vy[i] = s.vy[i] + s.fy[i] * c;
        s.vz[i] = s.vz[i] + s.fz[i] * c;
    }
    for (std::size_t i = 0; i < n; ++i) {
        s.x[i] = s.x[i] + s.vx[i] * dt;
        s.y[i] = s.y[i] + s.vy[i] * dt;
        s.z[i] = s.z[i] + s.vz[i] * dt;
        free_pos[i] = s.r(i);
    }
    constrain_positions(s, start);
    const double inv_dt = 1.0 / dt;
    for (std::size_t i = 0; i < n; ++i) {
        const V3 r = s.r(i);
        const V3 d = min_image(s.L, V3{r.x - free_pos[i].x, r.y - free_pos[i].y, r.z - free_pos[i].z});
        s.vx[i] = s.vx[i] + d.x * inv_dt;
        s.vy[i] = s.vy[i] + d.y * inv_dt;
        s.vz[i] = s.vz[i] + d.z * inv_dt;
    }
    forces(s);
    for (std::size_t i = 0; i < n; ++i) {
        const double c = 0.5 * dt / s.m[i];
        s.vx[i] = s.vx[i] + s.fx[i] * c;
        s.vy[i] = s.vy[i] + s.fy[i] * c;
        s.vz[i] = s.vz[i] + s.fz[i] * c;
    }
    constrain_velocities(s);
}
}  // namespace

void cvv_step(State& s, double dt, const Rigid& rw, const ForceFn& forces) {
    cvv_impl(
        s, dt, [&](State& st, const std::vector<V3>& ref) { rw.shake(st, ref); },
        [&](State& st) { rw.rattle(st); }, forces);
}

// The same step with SETTLE in place of SHAKE and RATTLE.
void cvv_step(State& s, double dt, const Settle& c, const ForceFn& forces) {
    cvv_impl(
        s, dt, [&](State& st, const std::vector<V3>& ref) { c.positions(st, ref); },
        [&](State& st) { c.velocities(st); }, forces);
}

}  // namespace oracle
