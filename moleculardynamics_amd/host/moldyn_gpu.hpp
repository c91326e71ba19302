// moldyn_gpu.hpp -- the reference's C++17 interfaces for the force /
// neighbour / integrator path, implemented over the C ABI of
// libmoldyn_hip.so (include/moldyn_hip.h). Header-only.
//
// Drop-in shape: the class names, constructor arguments, member functions,
// argument meanings and exception types are those of the reference
// (VishnuR23/MolecularDynamics src/moldyn/{core,forcefield,neighbor,
// integrate}), so code written against
//   moldyn::System, moldyn::Box, moldyn::LennardJones, moldyn::CellList,
//   moldyn::VerletList, moldyn::VelocityVerlet, moldyn::LangevinBAOAB,
//   moldyn::NoseHooverChain
// compiles against moldyn::gpu::* with the namespace changed. The System
// keeps host arrays and a device copy, synchronised lazily: device work
// (forces, list builds, integrator passes) uploads host edits first and
// marks the host copy stale; host reads download first.
//
// Differences a caller can observe, all documented at the entry points:
//  * VerletList::forEachPair visits the pairs sorted by (i, j); the reference
//    visits them in cell-traversal order (same set).
//  * One System holds one device Verlet list: a second VerletList built on
//    the same System replaces the first one's device rows.
//  * LangevinBAOAB noise comes from a counter-based generator keyed by
//    (seed, step, atom index), not the reference's sequential mt19937_64
//    stream; its statistics are what the reference's tests check.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/moldyn_hip.h"

namespace moldyn {
namespace gpu {

struct Vec3 {
    double x, y, z;
};
inline constexpr Vec3 operator+(Vec3 a, Vec3 b) { return Vec3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline constexpr Vec3 operator-(Vec3 a, Vec3 b) { return Vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline constexpr Vec3 operator*(Vec3 a, double s) { return Vec3{a.x * s, a.y * s, a.z * s}; }
inline constexpr Vec3 operator*(double s, Vec3 a) { return Vec3{a.x * s, a.y * s, a.z * s}; }
inline constexpr Vec3 operator/(Vec3 a, double s) { return Vec3{a.x / s, a.y / s, a.z / s}; }
inline Vec3& operator+=(Vec3& a, Vec3 b) { a = a + b; return a; }
inline Vec3& operator-=(Vec3& a, Vec3 b) { a = a - b; return a; }
inline constexpr double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline constexpr double norm2(Vec3 a) { return dot(a, a); }
inline double norm(Vec3 a) { return std::sqrt(norm2(a)); }

namespace detail {
// Maps C ABI status codes to the reference's exception types.
inline void check(md_status st, const md_ctx* ctx, const char* what) {
    if (st == MD_OK) return;
    std::string msg = std::string(what) + ": " + (ctx ? md_last_error(ctx) : "");
    if (st == MD_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
inline double fold(double v, double len) { return v - len * std::floor(v / len + 0.5); }
}  // namespace detail

// core/box.hpp:7-18
class Box {
public:
    explicit Box(Vec3 lengths) : lengths_(lengths) {}
    Vec3 minimumImage(Vec3 d) const {
        return Vec3{detail::fold(d.x, lengths_.x), detail::fold(d.y, lengths_.y),
                    detail::fold(d.z, lengths_.z)};
    }
    Vec3 wrap(Vec3 r) const { return minimumImage(r); }
    double volume() const { return lengths_.x * lengths_.y * lengths_.z; }
    Vec3 lengths() const { return lengths_; }

private:
    Vec3 lengths_;
};

enum class Truncation { Truncated, LinearForceShift };

// forcefield/forcefield.hpp:8-11
struct EnergyVirial {
    double energy = 0.0;
    double virial = 0.0;
};

class VerletList;

// core/system.hpp:11-60, with a device copy behind it.
class System {
public:
    explicit System(Box box) : box_(box) {}
    System(const System&) = delete;
    System& operator=(const System&) = delete;
    System(System&&) = default;
    System& operator=(System&&) = default;

    std::size_t size() const { return x_.size(); }
    const Box& box() const { return box_; }

    void addAtom(Vec3 position, Vec3 velocity, double mass) {
        addAtom(position, velocity, mass, 0.0, x_.size());
    }
    void addAtom(Vec3 position, Vec3 velocity, double mass, double charge, std::size_t moleculeId) {
        pull();
        x_.push_back(position.x); y_.push_back(position.y); z_.push_back(position.z);
        vx_.push_back(velocity.x); vy_.push_back(velocity.y); vz_.push_back(velocity.z);
        fx_.push_back(0.0); fy_.push_back(0.0); fz_.push_back(0.0);
        m_.push_back(mass); q_.push_back(charge); mol_.push_back(moleculeId);
        shape_changed_ = true;
        dev_stale_ = true;
    }
    void reserve(std::size_t n) {
        for (auto* v : {&x_, &y_, &z_, &vx_, &vy_, &vz_, &fx_, &fy_, &fz_, &m_, &q_}) v->reserve(n);
        mol_.reserve(n);
    }

    Vec3 position(std::size_t i) const { pull(); return Vec3{x_[i], y_[i], z_[i]}; }
    Vec3 velocity(std::size_t i) const { pull(); return Vec3{vx_[i], vy_[i], vz_[i]}; }
    Vec3 force(std::size_t i) const { pull(); return Vec3{fx_[i], fy_[i], fz_[i]}; }
    double mass(std::size_t i) const { return m_[i]; }
    double charge(std::size_t i) const { return q_[i]; }
    std::size_t moleculeId(std::size_t i) const { return mol_[i]; }

    void setPosition(std::size_t i, Vec3 r) { pull(); x_[i] = r.x; y_[i] = r.y; z_[i] = r.z; pos_dirty_ = true; }
    void setVelocity(std::size_t i, Vec3 v) { pull(); vx_[i] = v.x; vy_[i] = v.y; vz_[i] = v.z; vel_dirty_ = true; }
    void setForce(std::size_t i, Vec3 f) { pull(); fx_[i] = f.x; fy_[i] = f.y; fz_[i] = f.z; frc_dirty_ = true; }
    void addForce(std::size_t i, Vec3 f) { pull(); fx_[i] += f.x; fy_[i] += f.y; fz_[i] += f.z; frc_dirty_ = true; }
    void setCharge(std::size_t i, double q) { q_[i] = q; }
    void setMoleculeId(std::size_t i, std::size_t id) { mol_[i] = id; }

    void zeroForces() {
        pull();
        for (std::size_t i = 0; i < size(); ++i) fx_[i] = fy_[i] = fz_[i] = 0.0;
        frc_dirty_ = true;
    }
    // FP64 device reduction (system.cpp:124-131)
    double kineticEnergy() const {
        double ke = 0.0;
        detail::check(md_system_kinetic_energy(device(), &ke), ctx_.get(), "kineticEnergy");
        return ke;
    }
    double temperature(std::size_t dof) const { return 2.0 * kineticEnergy() / static_cast<double>(dof); }
    Vec3 totalMomentum() const {
        pull();
        Vec3 p{0.0, 0.0, 0.0};
        for (std::size_t i = 0; i < size(); ++i) {
            p.x += m_[i] * vx_[i]; p.y += m_[i] * vy_[i]; p.z += m_[i] * vz_[i];
        }
        return p;
    }
    void removeCenterOfMassMotion() {
        detail::check(md_system_remove_com_motion(device()), ctx_.get(), "removeCenterOfMassMotion");
        deviceWrote();
    }

    const double* xs() const { pull(); return x_.data(); }
    const double* ys() const { pull(); return y_.data(); }
    const double* zs() const { pull(); return z_.data(); }
    const double* qs() const { return q_.data(); }
    double* fxs() { pull(); frc_dirty_ = true; return fx_.data(); }
    double* fys() { pull(); frc_dirty_ = true; return fy_.data(); }
    double* fzs() { pull(); frc_dirty_ = true; return fz_.data(); }

    // ---- device coupling -------------------------------------------------
    // The context holding this system's device copy, with every host edit
    // pushed. Creates the context on first use.
    md_ctx* device() const {
        if (!ctx_) {
            md_ctx* c = nullptr;
            detail::check(md_ctx_create(nullptr, &c), nullptr, "md_ctx_create");
            ctx_.reset(c);
        }
        push();
        return ctx_.get();
    }
    // Marks the device copy as newer than the host arrays.
    void deviceWrote() const { host_stale_ = true; }

private:
    struct CtxDel {
        void operator()(md_ctx* c) const { md_ctx_destroy(c); }
    };

    void pull() const {
        if (!host_stale_ || !ctx_) return;
        detail::check(md_system_download(ctx_.get(), x_.data(), y_.data(), z_.data(), vx_.data(),
                                         vy_.data(), vz_.data(), fx_.data(), fy_.data(), fz_.data()),
                      ctx_.get(), "md_system_download");
        host_stale_ = false;
    }
    void push() const {
        md_ctx* c = ctx_.get();
        if (shape_changed_) {
            const double L[3] = {box_.lengths().x, box_.lengths().y, box_.lengths().z};
            detail::check(md_system_upload(c, static_cast<long>(size()), L, x_.data(), y_.data(),
                                           z_.data(), vx_.data(), vy_.data(), vz_.data(), m_.data()),
                          c, "md_system_upload");
            shape_changed_ = dev_stale_ = pos_dirty_ = vel_dirty_ = false;
            frc_dirty_ = true;
        }
        if (pos_dirty_) {
            detail::check(md_system_set_positions(c, x_.data(), y_.data(), z_.data()), c, "set_positions");
            pos_dirty_ = false;
        }
        if (vel_dirty_) {
            detail::check(md_system_set_velocities(c, vx_.data(), vy_.data(), vz_.data()), c, "set_velocities");
            vel_dirty_ = false;
        }
        if (frc_dirty_) {
            detail::check(md_system_set_forces(c, fx_.data(), fy_.data(), fz_.data()), c, "set_forces");
            frc_dirty_ = false;
        }
    }

    Box box_;
    mutable std::vector<double> x_, y_, z_, vx_, vy_, vz_, fx_, fy_, fz_;
    std::vector<double> m_, q_;
    std::vector<std::size_t> mol_;
    mutable std::unique_ptr<md_ctx, CtxDel> ctx_;
    mutable bool host_stale_ = false, dev_stale_ = true, shape_changed_ = true;
    mutable bool pos_dirty_ = false, vel_dirty_ = false, frc_dirty_ = false;
};

// neighbor/cell_list.hpp:24-53. The device bins atoms itself inside
// LennardJones::computeForces(System&, const CellList&); build() records the
// system and checks nothing else.
class CellList {
public:
    CellList(const Box& box, double cutoff) : box_(box), cutoff_(cutoff) {
        const Vec3 L = box.lengths();
        n_[0] = std::max(1, static_cast<int>(std::floor(L.x / cutoff)));
        n_[1] = std::max(1, static_cast<int>(std::floor(L.y / cutoff)));
        n_[2] = std::max(1, static_cast<int>(std::floor(L.z / cutoff)));
    }
    void build(const System&) {}
    int cellsPerSide(int dim) const { return n_[dim < 0 ? 0 : (dim > 2 ? 2 : dim)]; }
    bool usable() const { return n_[0] >= 3 && n_[1] >= 3 && n_[2] >= 3; }
    double cutoff() const { return cutoff_; }

private:
    Box box_;
    double cutoff_;
    int n_[3];
};

// neighbor/verlet_list.hpp:24-50, rows held on the device of the System it
// was built on.
class VerletList {
public:
    VerletList(const Box& box, double cutoff, double skin) : box_(box), cutoff_(cutoff), skin_(skin) {}

    void build(const System& sys) {
        md_ctx* c = sys.device();
        detail::check(md_neigh_config(c, cutoff_, skin_), c, "VerletList");
        detail::check(md_neigh_build(c), c, "VerletList::build");
        sys.deviceWrote();
        owner_ = &sys;
    }
    bool needsRebuild(const System& sys) const {
        int flag = 0;
        md_ctx* c = sys.device();
        detail::check(md_neigh_needs_rebuild(c, &flag), c, "VerletList::needsRebuild");
        return flag != 0;
    }
    void rebuildIfNeeded(const System& sys) {
        if (needsRebuild(sys)) build(sys);
    }
    std::size_t rebuildCount() const {
        if (owner_ == nullptr) return 0;
        long k = 0;
        md_ctx* c = owner_->device();
        detail::check(md_neigh_rebuild_count(c, &k), c, "VerletList::rebuildCount");
        return static_cast<std::size_t>(k);
    }
    // Visits every pair (i < j) once, sorted by (i, j).
    template <class Fn>
    void forEachPair(Fn&& fn) const {
        if (owner_ == nullptr) return;
        md_ctx* c = owner_->device();
        long np = 0;
        md_status st = md_neigh_export_pairs(c, nullptr, 0, &np);
        if (st != MD_OK && st != MD_ERR_CAPACITY) detail::check(st, c, "forEachPair");
        std::vector<uint32_t> ij(static_cast<std::size_t>(2 * np));
        detail::check(md_neigh_export_pairs(c, ij.data(), np, &np), c, "forEachPair");
        for (long k = 0; k < np; ++k) fn(static_cast<std::size_t>(ij[2 * k]), static_cast<std::size_t>(ij[2 * k + 1]));
    }
    double cutoff() const { return cutoff_; }
    double skin() const { return skin_; }

private:
    Box box_;
    double cutoff_, skin_;
    const System* owner_ = nullptr;
};

// forcefield/lennard_jones.hpp:61-135
class LennardJones {
public:
    LennardJones(double sigma, double epsilon, double cutoff, Truncation truncation)
        : p_{sigma, epsilon, cutoff, truncation == Truncation::LinearForceShift ? 1 : 0},
          trunc_(truncation) {}

    EnergyVirial computeEnergyVirial(const System& sys) const {
        md_ctx* c = bind(sys);
        EnergyVirial ev;
        detail::check(md_lj_energy_virial(c, &ev.energy, &ev.virial), c, "computeEnergyVirial");
        return ev;
    }
    EnergyVirial computeForces(System& sys) const { return forces(sys, MD_PAIRS_BRUTE); }
    EnergyVirial computeForces(System& sys, const CellList& cells) const {
        if (!cells.usable())
            throw std::invalid_argument(
                "LennardJones::computeForces: CellList is not usable (fewer than 3 cells per side); "
                "fall back to the brute-force overload");
        md_ctx* c = bind(sys);
        EnergyVirial ev;
        detail::check(md_lj_forces_cells(c, cells.cutoff(), &ev.energy, &ev.virial), c, "computeForces");
        sys.deviceWrote();
        return ev;
    }
    EnergyVirial computeForces(System& sys, const VerletList&) const { return forces(sys, MD_PAIRS_VERLET); }
    // The mixed-precision pair path (FP32 pairs, fixed-point accumulation).
    EnergyVirial computeForcesMixed(System& sys, const VerletList&) const {
        return forces(sys, MD_PAIRS_VERLET, MD_MIXED);
    }

    double longRangeCorrection(const System& sys) const { return tail(sys).first; }
    double longRangeCorrectionPressure(const System& sys) const { return tail(sys).second; }

    double sigma() const { return p_.sigma; }
    double epsilon() const { return p_.epsilon; }
    double cutoff() const { return p_.cutoff; }
    Truncation truncation() const { return trunc_; }

private:
    md_ctx* bind(const System& sys) const {
        md_ctx* c = sys.device();
        detail::check(md_lj_set(c, &p_), c, "LennardJones");
        return c;
    }
    EnergyVirial forces(System& sys, int source, int precision = MD_FP64) const {
        md_ctx* c = bind(sys);
        EnergyVirial ev;
        detail::check(md_lj_forces(c, source, precision, &ev.energy, &ev.virial), c, "computeForces");
        sys.deviceWrote();
        return ev;
    }
    std::pair<double, double> tail(const System& sys) const {
        const double L[3] = {sys.box().lengths().x, sys.box().lengths().y, sys.box().lengths().z};
        double u = 0.0, p = 0.0;
        detail::check(md_lj_tail(&p_, static_cast<long>(sys.size()), L, &u, &p), nullptr, "tail");
        return {u, p};
    }
    md_lj_params p_;
    Truncation trunc_;
};

// integrate/velocity_verlet.hpp:39-80
class VelocityVerlet {
public:
    explicit VelocityVerlet(double dt) : dt_(dt) {}
    template <class ForceFn>
    void step(System& sys, ForceFn&& computeForces) const {
        md_ctx* c = sys.device();
        detail::check(md_half_kick(c, dt_), c, "VelocityVerlet::step");
        detail::check(md_drift(c, dt_), c, "VelocityVerlet::step");
        sys.deviceWrote();
        computeForces(sys);
        c = sys.device();
        detail::check(md_half_kick(c, dt_), c, "VelocityVerlet::step");
        sys.deviceWrote();
    }
    double dt() const { return dt_; }

private:
    double dt_;
};

// integrate/langevin.hpp:122-190
class LangevinBAOAB {
public:
    LangevinBAOAB(double dt, double friction, double temperature, uint64_t seed)
        : dt_(dt), friction_(friction), temperature_(temperature), seed_(seed) {}
    template <class ForceFn>
    void step(System& sys, ForceFn&& computeForces) {
        const double c1 = std::exp(-friction_ * dt_);
        const double c2 = std::sqrt(1.0 - c1 * c1);
        md_ctx* c = sys.device();
        detail::check(md_half_kick(c, dt_), c, "LangevinBAOAB::step");
        detail::check(md_drift(c, 0.5 * dt_), c, "LangevinBAOAB::step");
        detail::check(md_langevin_o(c, c1, c2, temperature_, seed_, counter_++), c, "LangevinBAOAB::step");
        detail::check(md_drift(c, 0.5 * dt_), c, "LangevinBAOAB::step");
        sys.deviceWrote();
        computeForces(sys);
        c = sys.device();
        detail::check(md_half_kick(c, dt_), c, "LangevinBAOAB::step");
        sys.deviceWrote();
    }
    double temperature() const { return temperature_; }

private:
    double dt_, friction_, temperature_;
    uint64_t seed_;
    uint64_t counter_ = 0;
};

// integrate/nose_hoover.{hpp:26-67, cpp:15-124}: the chain variables live on
// the host (a handful of doubles); the kinetic energy and the velocity
// rescale run on the device, in the reference's order.
class NoseHooverChain {
public:
    NoseHooverChain(double dt, double temperature, double tau, int chainLength,
                    std::size_t degreesOfFreedom)
        : dt_(dt), T_(temperature), tau_(tau), M_(chainLength), dof_(degreesOfFreedom), core_(dt) {
        if (chainLength < 1) throw std::invalid_argument("NoseHooverChain: chainLength must be >= 1");
        Q_.assign(static_cast<std::size_t>(M_), T_ * tau_ * tau_);
        Q_[0] = static_cast<double>(dof_) * T_ * tau_ * tau_;
        xi_.assign(static_cast<std::size_t>(M_), 0.0);
        p_.assign(static_cast<std::size_t>(M_), 0.0);
    }
    template <class ForceFn>
    void step(System& sys, ForceFn&& computeForces) {
        integrateChain(sys, 0.5 * dt_);
        core_.step(sys, std::forward<ForceFn>(computeForces));
        integrateChain(sys, 0.5 * dt_);
    }
    double conservedQuantityContribution() const {
        double s = static_cast<double>(dof_) * T_ * xi_[0];
        for (std::size_t k = 1; k < xi_.size(); ++k) s += T_ * xi_[k];
        for (std::size_t k = 0; k < p_.size(); ++k) s += 0.5 * p_[k] * p_[k] / Q_[k];
        return s;
    }

private:
    void integrateChain(System& sys, double delta) {
        const double w1 = 1.0 / (2.0 - std::cbrt(2.0));
        const double ws[3] = {w1, 1.0 - 2.0 * w1, w1};
        for (double w : ws) substep(sys, w * delta);
    }
    void substep(System& sys, double d) {
        const double d2 = 0.5 * d, d4 = 0.25 * d;
        for (int j = M_ - 1; j >= 0; --j) link(sys, j, d2, d4);
        for (std::size_t k = 0; k < xi_.size(); ++k) xi_[k] += d * p_[k] / Q_[k];
        const double factor = std::exp(-d * p_[0] / Q_[0]);
        md_ctx* c = sys.device();
        detail::check(md_scale_velocities(c, factor), c, "NoseHooverChain::step");
        sys.deviceWrote();
        for (int j = 0; j < M_; ++j) link(sys, j, d2, d4);
    }
    void link(System& sys, int j, double d2, double d4) {
        const std::size_t jj = static_cast<std::size_t>(j);
        const bool next = j < M_ - 1;
        if (next) p_[jj] *= std::exp(-d4 * p_[jj + 1] / Q_[jj + 1]);
        const double g = (j == 0) ? 2.0 * sys.kineticEnergy() - static_cast<double>(dof_) * T_
                                  : p_[jj - 1] * p_[jj - 1] / Q_[jj - 1] - T_;
        p_[jj] += d2 * g;
        if (next) p_[jj] *= std::exp(-d4 * p_[jj + 1] / Q_[jj + 1]);
    }

    double dt_, T_, tau_;
    int M_;
    std::size_t dof_;
    VelocityVerlet core_;
    std::vector<double> Q_, xi_, p_;
};

// util/lattice.cpp:15-64 through md_fcc_lattice (bit-identical inputs).
inline System fccLattice(int cells, double density, double temperature, uint64_t seed) {
    const std::size_t n = 4u * static_cast<std::size_t>(cells) * cells * cells;
    std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n);
    double L[3];
    detail::check(md_fcc_lattice(cells, density, temperature, seed, L, x.data(), y.data(), z.data(),
                                 vx.data(), vy.data(), vz.data()),
                  nullptr, "fccLattice");
    System sys(Box(Vec3{L[0], L[1], L[2]}));
    sys.reserve(n);
    for (std::size_t i = 0; i < n; ++i) sys.addAtom(Vec3{x[i], y[i], z[i]}, Vec3{vx[i], vy[i], vz[i]}, 1.0);
    return sys;
}

}  // namespace gpu
}  // namespace moldyn
