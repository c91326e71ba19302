// moldyn_gpu.hpp -- the reference's C++17 interfaces for the force /
// neighbour / integrator path, implemented over the C ABI of
// libmoldyn_hip.so (include/moldyn_hip.h). Header-only.
//
// Drop-in shape: the class names, constructor arguments, member functions,
// argument meanings and exception types are those of the reference
// (VishnuR23/MolecularDynamics src/moldyn/{core,forcefield,neighbor,
// integrate,constrain,analyze,util}), so code written against
//   moldyn::System, moldyn::Box, moldyn::ForceField, moldyn::LennardJones,
//   moldyn::CellList, moldyn::VerletList, moldyn::VelocityVerlet,
//   moldyn::LangevinBAOAB, moldyn::NoseHooverChain, moldyn::Rng,
//   moldyn::units, moldyn::Ewald, moldyn::SpceForceField,
//   moldyn::buildSpceSystem, moldyn::RigidWater,
//   moldyn::ConstrainedVelocityVerlet, moldyn::Rdf, moldyn::Msd, moldyn::Vacf
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
//  * Ewald, SpceForceField and RigidWater run on rigid three-site water only
//    (O, H, H triples with charges -2 qH, qH, qH, as buildSpceSystem makes
//    them); other charge layouts are refused with std::invalid_argument. They
//    keep a device copy of the water (md_water) next to the System's own and
//    refresh it from the System on every call.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <fstream>
#include <memory>
#include <random>
#include <sstream>
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
class System;

// forcefield/forcefield.hpp:17-35: the force-field interface the reference's
// call sites are written against.
class ForceField {
public:
    virtual ~ForceField() = default;
    virtual EnergyVirial computeEnergyVirial(const System& sys) const = 0;
    virtual double longRangeCorrection(const System& sys) const = 0;
    virtual double longRangeCorrectionPressure(const System& sys) const = 0;
};

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

// neighbor/cell_list.hpp:24-53. The GPU force path bins atoms itself inside
// LennardJones::computeForces(System&, const CellList&). build() also bins
// them on the host (cell of the folded position, floor((r + L/2) / width)
// clamped, each atom prepended to its cell's chain: cell_list.cpp:16-43), so
// that forEachPair can offer the candidate pairs to host callers.
class CellList {
public:
    CellList(const Box& box, double cutoff) : box_(box), cutoff_(cutoff) {
        const Vec3 L = box.lengths();
        n_[0] = std::max(1, static_cast<int>(std::floor(L.x / cutoff)));
        n_[1] = std::max(1, static_cast<int>(std::floor(L.y / cutoff)));
        n_[2] = std::max(1, static_cast<int>(std::floor(L.z / cutoff)));
    }
    void build(const System& sys);
    int cellsPerSide(int dim) const { return n_[dim < 0 ? 0 : (dim > 2 ? 2 : dim)]; }
    bool usable() const { return n_[0] >= 3 && n_[1] >= 3 && n_[2] >= 3; }
    double cutoff() const { return cutoff_; }

    // cell_list.hpp:33-34, 55-108: fn(i, j), i < j, once for every pair in
    // the same cell or in one of the 13 forward neighbour cells (periodic),
    // cells in (x, y, z) order, chains in the order build() made them. The
    // caller applies the cutoff. Valid after build() on a usable list.
    template <class Fn>
    void forEachPair(Fn&& fn) const {
        static constexpr int kForward[13][3] = {{1, 0, 0},  {1, 1, 0},  {0, 1, 0},   {-1, 1, 0},
                                                {0, 0, 1},  {1, 0, 1},  {-1, 0, 1},  {0, 1, 1},
                                                {0, -1, 1}, {1, 1, 1},  {1, -1, 1},  {-1, 1, 1},
                                                {-1, -1, 1}};
        auto ordered = [&](int a, int b) {
            fn(static_cast<std::size_t>(std::min(a, b)), static_cast<std::size_t>(std::max(a, b)));
        };
        for (int ix = 0; ix < n_[0]; ++ix)
            for (int iy = 0; iy < n_[1]; ++iy)
                for (int iz = 0; iz < n_[2]; ++iz) {
                    const int c = index(ix, iy, iz);
                    for (int a = head_[c]; a >= 0; a = next_[a])
                        for (int b = next_[a]; b >= 0; b = next_[b]) ordered(a, b);
                    for (const auto& o : kForward) {
                        const int nb = index(wrap(ix + o[0], n_[0]), wrap(iy + o[1], n_[1]),
                                             wrap(iz + o[2], n_[2]));
                        for (int a = head_[c]; a >= 0; a = next_[a])
                            for (int b = head_[nb]; b >= 0; b = next_[b]) ordered(a, b);
                    }
                }
    }

private:
    static int wrap(int v, int n) { return ((v % n) + n) % n; }
    int index(int ix, int iy, int iz) const { return ix + n_[0] * (iy + n_[1] * iz); }

    Box box_;
    double cutoff_;
    int n_[3];
    std::vector<int> head_, next_;
};

inline void CellList::build(const System& sys) {
    const std::size_t n = sys.size();
    const Vec3 L = box_.lengths();
    const double Ld[3] = {L.x, L.y, L.z};
    const double w[3] = {L.x / n_[0], L.y / n_[1], L.z / n_[2]};
    head_.assign(static_cast<std::size_t>(n_[0]) * n_[1] * n_[2], -1);
    next_.assign(n, -1);
    const double* r[3] = {sys.xs(), sys.ys(), sys.zs()};
    for (std::size_t i = 0; i < n; ++i) {
        int k[3];
        for (int d = 0; d < 3; ++d) {
            const double f = detail::fold(r[d][i], Ld[d]);
            const int q = static_cast<int>(std::floor((f + 0.5 * Ld[d]) / w[d]));
            k[d] = q < 0 ? 0 : (q >= n_[d] ? n_[d] - 1 : q);
        }
        const int c = index(k[0], k[1], k[2]);
        next_[i] = head_[c];
        head_[c] = static_cast<int>(i);
    }
}

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

// forcefield/lennard_jones.hpp:24-98
class LennardJones : public ForceField {
public:
    LennardJones(double sigma, double epsilon, double cutoff, Truncation truncation)
        : p_{sigma, epsilon, cutoff, truncation == Truncation::LinearForceShift ? 1 : 0},
          trunc_(truncation) {}

    EnergyVirial computeEnergyVirial(const System& sys) const override {
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

    double longRangeCorrection(const System& sys) const override { return tail(sys).first; }
    double longRangeCorrectionPressure(const System& sys) const override { return tail(sys).second; }

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

// integrate/langevin.hpp:29-97
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

namespace detail {
// Status of an analysis handle call, with the handle's own message.
template <class H>
inline void check_handle(md_status st, const H* h, const char* (*last)(const H*), const char* what) {
    if (st == MD_OK) return;
    const std::string msg = std::string(what) + ": " + (h ? last(h) : "");
    if (st == MD_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
}  // namespace detail

// analyze/rdf.hpp:16-34: g(r) accumulated on the device from the System's
// device copy (md_rdf_*). Pair counts are exact, so g equals the
// reference's bit for bit on the same frames.
class Rdf {
public:
    Rdf(double rMax, std::size_t bins) : bins_(bins) {
        md_rdf* h = nullptr;
        detail::check_handle<md_rdf>(md_rdf_create(rMax, static_cast<long>(bins), &h), nullptr,
                                     md_rdf_last_error, "Rdf");
        h_.reset(h);
    }
    // rdf.cpp:19-46; throws std::invalid_argument when rMax exceeds half the
    // smallest box length
    void accumulate(const System& sys) {
        detail::check_handle<md_rdf>(md_rdf_accumulate(h_.get(), sys.device(), MD_RDF_AUTO), h_.get(),
                                     md_rdf_last_error, "Rdf::accumulate");
    }
    std::vector<double> binCentres() const { return fetch(true); }
    std::vector<double> g() const { return fetch(false); }
    std::size_t frames() const {
        long f = 0;
        detail::check_handle<md_rdf>(md_rdf_result(h_.get(), nullptr, nullptr, nullptr, &f), h_.get(),
                                     md_rdf_last_error, "Rdf::frames");
        return static_cast<std::size_t>(f);
    }

private:
    std::vector<double> fetch(bool centres) const {
        std::vector<double> v(bins_);
        detail::check_handle<md_rdf>(md_rdf_result(h_.get(), centres ? v.data() : nullptr,
                                                   centres ? nullptr : v.data(), nullptr, nullptr),
                                     h_.get(), md_rdf_last_error, "Rdf::g");
        return v;
    }
    struct Del {
        void operator()(md_rdf* h) const { md_rdf_destroy(h); }
    };
    std::size_t bins_;
    std::unique_ptr<md_rdf, Del> h_;
};

// analyze/msd.hpp:19-42: ring buffer and unwrapping on the device (md_msd_*).
class Msd {
public:
    Msd(std::size_t natoms, std::size_t maxLag) {
        md_msd* h = nullptr;
        detail::check_handle<md_msd>(md_msd_create(static_cast<long>(natoms), static_cast<long>(maxLag), &h),
                                     nullptr, md_msd_last_error, "Msd");
        h_.reset(h);
        lags_ = maxLag + 1;
    }
    void push(const System& sys) {
        detail::check_handle<md_msd>(md_msd_push(h_.get(), sys.device()), h_.get(), md_msd_last_error,
                                     "Msd::push");
    }
    std::vector<double> msd() const {
        std::vector<double> v(lags_);
        detail::check_handle<md_msd>(md_msd_result(h_.get(), v.data()), h_.get(), md_msd_last_error,
                                     "Msd::msd");
        return v;
    }
    double diffusionCoefficient(double dt, std::size_t fitFrom, std::size_t fitTo) const {
        double d = 0.0;
        detail::check_handle<md_msd>(md_msd_diffusion(h_.get(), dt, static_cast<long>(fitFrom),
                                                      static_cast<long>(fitTo), &d),
                                     h_.get(), md_msd_last_error, "Msd::diffusionCoefficient");
        return d;
    }

private:
    struct Del {
        void operator()(md_msd* h) const { md_msd_destroy(h); }
    };
    std::size_t lags_ = 0;
    std::unique_ptr<md_msd, Del> h_;
};

// analyze/vacf.hpp:151-174 on the device (md_vacf_*).
class Vacf {
public:
    Vacf(std::size_t natoms, std::size_t maxLag) {
        md_vacf* h = nullptr;
        detail::check_handle<md_vacf>(md_vacf_create(static_cast<long>(natoms), static_cast<long>(maxLag), &h),
                                      nullptr, md_vacf_last_error, "Vacf");
        h_.reset(h);
        lags_ = maxLag + 1;
    }
    void push(const System& sys) {
        detail::check_handle<md_vacf>(md_vacf_push(h_.get(), sys.device()), h_.get(), md_vacf_last_error,
                                      "Vacf::push");
    }
    std::vector<double> vacf() const {
        std::vector<double> v(lags_);
        detail::check_handle<md_vacf>(md_vacf_result(h_.get(), 1, v.data()), h_.get(),
                                      md_vacf_last_error, "Vacf::vacf");
        return v;
    }
    double diffusionCoefficient(double dt) const {
        double d = 0.0;
        detail::check_handle<md_vacf>(md_vacf_diffusion(h_.get(), dt, &d), h_.get(), md_vacf_last_error,
                                      "Vacf::diffusionCoefficient");
        return d;
    }

private:
    struct Del {
        void operator()(md_vacf* h) const { md_vacf_destroy(h); }
    };
    std::size_t lags_ = 0;
    std::unique_ptr<md_vacf, Del> h_;
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

// util/random.hpp:37-76: mt19937_64, uniform deviates in (0, 1] as
// 1 - U[0, 1), standard normals by Marsaglia's polar method with the second
// deviate of each pair kept for the next call (the reference's stream, draw
// for draw).
class Rng {
public:
    explicit Rng(uint64_t seed) : engine_(seed) {}
    double uniform() { return 1.0 - unit_(engine_); }
    double normal() {
        if (spare_ready_) {
            spare_ready_ = false;
            return spare_;
        }
        double a, b, q;
        do {
            a = 2.0 * uniform() - 1.0;
            b = 2.0 * uniform() - 1.0;
            q = a * a + b * b;
        } while (q >= 1.0 || q == 0.0);
        const double k = std::sqrt(-2.0 * std::log(q) / q);
        spare_ = b * k;
        spare_ready_ = true;
        return a * k;
    }

private:
    std::mt19937_64 engine_;
    std::uniform_real_distribution<double> unit_{0.0, 1.0};
    double spare_ = 0.0;
    bool spare_ready_ = false;
};

// core/units.hpp:33-40: the internal time unit (A sqrt(amu / kB)) in ps.
namespace units {
inline constexpr double kTauInPicoseconds = 1.0966875353;
inline constexpr double picoseconds(double ps) { return ps / kTauInPicoseconds; }
inline constexpr double femtoseconds(double fs) { return picoseconds(fs * 1.0e-3); }
inline constexpr double toPicoseconds(double tau) { return tau * kTauInPicoseconds; }
}  // namespace units

// ---- rigid SPC/E water with Ewald ------------------------------------------

// forcefield/ewald.hpp:10-19: e^2 / (4 pi eps0 kB) in K A / e^2 from the
// CODATA 2010 constants (167100.956632).
inline double coulombPrefactorKelvinAngstrom() {
    const double e = 1.602176565e-19, eps0 = 8.854187817e-12, kB = 1.3806488e-23;
    const double pi = 3.14159265358979323846;
    return e * e / (4.0 * pi * eps0 * kB) * 1e10;
}

// forcefield/ewald.hpp:24-31 (K)
struct EwaldTerms {
    double real = 0.0;
    double reciprocal = 0.0;
    double self = 0.0;
    double intra = 0.0;
    double total() const { return real + reciprocal + self + intra; }
};

// forcefield/spce.hpp:16-24
struct SpceParameters {
    double qH = 0.42380;
    double sigmaO = 3.16555789;
    double epsilonO = 78.19743111;
    double rOH = 1.0;
    double angleHOH = 109.47;
    double massOxygen = 15.9994;
    double massHydrogen = 1.008;
};

// io/nist_config.hpp:20-26 and the reader's format (io/nist_config.cpp:30-86):
// box line, count line (not trusted as an atom count), "index x y z [type]"
// lines to the end of the file; mass 1, zero velocity.
struct NistConfig {
    System system;
    std::vector<std::string> types;
};

inline NistConfig readNistConfig(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("readNistConfig: cannot open " + path);
    std::string line;
    double lx = 0, ly = 0, lz = 0;
    long count = 0;
    if (!std::getline(in, line) || !(std::istringstream(line) >> lx >> ly >> lz))
        throw std::runtime_error("readNistConfig: malformed box line in " + path);
    if (!std::getline(in, line) || !(std::istringstream(line) >> count))
        throw std::runtime_error("readNistConfig: malformed count line in " + path);
    NistConfig cfg{System(Box(Vec3{lx, ly, lz})), {}};
    while (std::getline(in, line)) {
        std::istringstream is(line);
        long idx;
        double x, y, z;
        if (!(is >> idx)) continue;  // blank line
        if (!(is >> x >> y >> z)) throw std::runtime_error("readNistConfig: malformed atom line in " + path);
        std::string type;
        if (is >> type) cfg.types.push_back(type);
        cfg.system.addAtom(Vec3{x, y, z}, Vec3{0, 0, 0}, 1.0);
    }
    if (!cfg.types.empty() && cfg.types.size() != cfg.system.size())
        throw std::runtime_error("readNistConfig: types on some lines only in " + path);
    return cfg;
}

// forcefield/spce.hpp:29-40
struct SpceSystem {
    System full;
    System oxygens;
    std::size_t moleculeCount = 0;
};

// forcefield/spce.cpp:8-40: molecules of three consecutive atoms O, H, H;
// charges -2 qH, qH, qH; molecule id = index / 3.
inline SpceSystem buildSpceSystem(const NistConfig& cfg, const SpceParameters& params = SpceParameters()) {
    const std::size_t n = cfg.system.size();
    if (n == 0 || n % 3 != 0)
        throw std::invalid_argument("buildSpceSystem: atom count must be a positive multiple of 3 (O, H, H per molecule)");
    if (cfg.types.size() != n) throw std::invalid_argument("buildSpceSystem: config has no per-atom types");
    SpceSystem out{System(cfg.system.box()), System(cfg.system.box()), n / 3};
    out.full.reserve(n);
    for (std::size_t i = 0; i < n; ++i) {
        const Vec3 r = cfg.system.position(i), v = cfg.system.velocity(i);
        if (cfg.types[i] == "O") {
            out.full.addAtom(r, v, params.massOxygen, -2.0 * params.qH, i / 3);
            out.oxygens.addAtom(r, v, params.massOxygen);
        } else if (cfg.types[i] == "H") {
            out.full.addAtom(r, v, params.massHydrogen, params.qH, i / 3);
        } else {
            throw std::invalid_argument("buildSpceSystem: unrecognised atom type '" + cfg.types[i] +
                                        "' (expected \"O\" or \"H\")");
        }
    }
    return out;
}

namespace detail {
inline void wcheck(md_status st, const md_water* w, const char* what) {
    if (st == MD_OK) return;
    std::string msg = std::string(what) + ": " + (w ? md_water_last_error(w) : "md_water_create failed");
    if (st == MD_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// The device water (md_water) a water class works on, made from the System
// on first use and again when its size, box or the parameters change; every
// call refreshes positions and velocities from the System. The System must
// hold O, H, H triples with charges -2 qH, qH, qH and masses (mO, mH, mH).
class WaterBinding {
public:
    md_water* bind(const System& sys, md_water_params p) const {
        const std::size_t n = sys.size();
        if (n == 0 || n % 3 != 0)
            throw std::invalid_argument("water: atom count must be a positive multiple of 3 (O, H, H)");
        const double qh = sys.charge(1), mo = sys.mass(0), mh = sys.mass(1);
        for (std::size_t i = 0; i < n; ++i) {
            const bool o = i % 3 == 0;
            if (sys.charge(i) != (o ? -2.0 * qh : qh) || sys.mass(i) != (o ? mo : mh) ||
                sys.moleculeId(i) != i / 3)
                throw std::invalid_argument(
                    "water: the GPU path needs O, H, H triples with charges -2 qH, qH, qH (buildSpceSystem)");
        }
        p.qH = qh;
        p.mass_o = mo;
        p.mass_h = mh;
        const Vec3 L = sys.box().lengths();
        const bool same = w_ && n_ == n && L_.x == L.x && L_.y == L.y && L_.z == L.z && equal(p_, p);
        if (!same) {
            const double Lb[3] = {L.x, L.y, L.z};
            md_water* w = nullptr;
            const md_status st = md_water_create(nullptr, static_cast<long>(n), Lb, sys.xs(), sys.ys(),
                                                 sys.zs(), nullptr, nullptr, nullptr, &p, &w);
            w_.reset(w);
            wcheck(st, w, "md_water_create");
            n_ = n;
            L_ = L;
            p_ = p;
        }
        std::vector<double> vx(n), vy(n), vz(n);
        for (std::size_t i = 0; i < n; ++i) {
            const Vec3 v = sys.velocity(i);
            vx[i] = v.x;
            vy[i] = v.y;
            vz[i] = v.z;
        }
        wcheck(md_water_set_state(w_.get(), sys.xs(), sys.ys(), sys.zs(), vx.data(), vy.data(), vz.data()),
               w_.get(), "md_water_set_state");
        return w_.get();
    }
    // Per-atom arrays (x, y, z) of the device water's positions, velocities
    // or forces (which = 0, 1, 2).
    void download(int which, std::vector<Vec3>& out) const {
        const std::size_t n = n_;
        std::vector<double> a(n), b(n), c(n);
        double* p[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        p[3 * which] = a.data();
        p[3 * which + 1] = b.data();
        p[3 * which + 2] = c.data();
        wcheck(md_water_download(w_.get(), p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8]), w_.get(),
               "md_water_download");
        out.resize(n);
        for (std::size_t i = 0; i < n; ++i) out[i] = Vec3{a[i], b[i], c[i]};
    }

private:
    static bool equal(const md_water_params& a, const md_water_params& b) {
        return a.qH == b.qH && a.sigma_o == b.sigma_o && a.epsilon_o == b.epsilon_o && a.r_oh == b.r_oh &&
               a.angle_hoh == b.angle_hoh && a.mass_o == b.mass_o && a.mass_h == b.mass_h &&
               a.cutoff == b.cutoff && a.kmax == b.kmax && a.n2max == b.n2max && a.alpha == b.alpha &&
               a.skin == b.skin && a.tolerance == b.tolerance && a.max_iterations == b.max_iterations;
    }
    struct Del {
        void operator()(md_water* w) const { md_water_destroy(w); }
    };
    mutable std::unique_ptr<md_water, Del> w_;
    mutable std::size_t n_ = 0;
    mutable Vec3 L_{0, 0, 0};
    mutable md_water_params p_{};
};

inline md_water_params water_params(double cutoff, int kmax) {
    md_water_params p{};
    const SpceParameters d;
    p.qH = d.qH;
    p.sigma_o = d.sigmaO;
    p.epsilon_o = d.epsilonO;
    p.r_oh = d.rOH;
    p.angle_hoh = d.angleHOH;
    p.mass_o = d.massOxygen;
    p.mass_h = d.massHydrogen;
    p.cutoff = cutoff;
    p.kmax = kmax;
    p.n2max = 0;     // |n|^2 < kmax^2 + 2 (ewald.cpp:119)
    p.alpha = 0.0;   // 5.6 / min(L) (ewald.cpp:40)
    p.skin = 1.0;
    p.tolerance = 1e-12;
    p.max_iterations = 500;
    return p;
}
}  // namespace detail

// forcefield/ewald.hpp:49-76: real space (intermolecular pairs within the
// cutoff), reciprocal space (|n|^2 < kmax^2 + 2), self and intramolecular
// terms, alpha = 5.6 / min(L); energies in K.
class Ewald {
public:
    Ewald(double cutoff, int kmax) : cutoff_(cutoff), kmax_(kmax) {}
    EwaldTerms computeEnergies(const System& sys) const { return run(sys, nullptr, false); }
    // Zeroes the forces, then accumulates the electrostatic ones.
    EwaldTerms computeForces(System& sys) const {
        sys.zeroForces();
        return run(sys, &sys, true);
    }
    // Adds the electrostatic forces to those already in sys.
    EwaldTerms accumulateForces(System& sys) const { return run(sys, &sys, true); }
    double cutoff() const { return cutoff_; }
    int kmax() const { return kmax_; }

private:
    EwaldTerms run(const System& sys, System* sink, bool) const {
        md_water_params p = detail::water_params(cutoff_, kmax_);
        p.epsilon_o = 0.0;  // electrostatics only
        md_water* w = water_.bind(sys, p);
        double t[6] = {0, 0, 0, 0, 0, 0}, pe = 0.0;
        detail::wcheck(md_water_forces(w, MD_FP64, t, &pe), w, "Ewald");
        if (sink != nullptr) {
            std::vector<Vec3> f;
            water_.download(2, f);
            for (std::size_t i = 0; i < f.size(); ++i) sink->addForce(i, f[i]);
        }
        EwaldTerms e;
        e.real = t[2];
        e.reciprocal = t[3];
        e.self = t[4];
        e.intra = t[5];
        return e;
    }
    double cutoff_;
    int kmax_;
    detail::WaterBinding water_;
};

// forcefield/spce_forcefield.hpp:23-51: oxygen Lennard-Jones (Truncated, with
// its tail correction) plus Ewald over all sites, in K.
class SpceForceField {
public:
    explicit SpceForceField(const SpceParameters& params = SpceParameters(), double cutoff = 10.0, int kmax = 5)
        : params_(params),
          lj_(params.sigmaO, params.epsilonO, cutoff, Truncation::Truncated),
          ewald_(cutoff, kmax),
          cutoff_(cutoff),
          kmax_(kmax) {}
    // Zeroes the forces, then dispersion plus electrostatics; returns the
    // potential energy including the dispersion tail.
    double computeForces(System& sys) const {
        double pe = 0.0;
        run(sys, &pe);
        std::vector<Vec3> f;
        water_.download(2, f);
        for (std::size_t i = 0; i < f.size(); ++i) sys.setForce(i, f[i]);
        return pe;
    }
    double computeEnergy(const System& sys) const {
        double pe = 0.0;
        run(sys, &pe);
        return pe;
    }
    const Ewald& ewald() const { return ewald_; }
    const LennardJones& dispersion() const { return lj_; }

private:
    void run(const System& sys, double* pe) const {
        md_water_params p = detail::water_params(cutoff_, kmax_);
        p.sigma_o = params_.sigmaO;
        p.epsilon_o = params_.epsilonO;
        md_water* w = water_.bind(sys, p);
        double t[6];
        detail::wcheck(md_water_forces(w, MD_FP64, t, pe), w, "SpceForceField");
    }
    SpceParameters params_;
    LennardJones lj_;
    Ewald ewald_;
    double cutoff_;
    int kmax_;
    detail::WaterBinding water_;
};

// constrain/rigid_water.hpp:28-66: SHAKE (positions, along the reference
// bonds) and RATTLE (velocities) on the constraints O-H1, O-H2, H1-H2 of every
// molecule, to a relative tolerance; std::runtime_error when they do not
// converge. The sweeps run on the device, one thread per molecule.
class RigidWater {
public:
    RigidWater(double ohDistance, double hohAngleDegrees, double tolerance = 1e-12, int maxIterations = 500)
        : oh_(ohDistance), angle_(hohAngleDegrees), tol_(tolerance), maxit_(maxIterations) {
        if (!(ohDistance > 0.0)) throw std::invalid_argument("RigidWater: ohDistance must be positive");
        if (!(hohAngleDegrees > 0.0) || !(hohAngleDegrees < 180.0))
            throw std::invalid_argument("RigidWater: hohAngleDegrees must lie in (0, 180)");
        if (!(tolerance > 0.0)) throw std::invalid_argument("RigidWater: tolerance must be positive");
        if (maxIterations < 1) throw std::invalid_argument("RigidWater: maxIterations must be at least 1");
        hh_ = 2.0 * oh_ * std::sin(0.5 * hohAngleDegrees * 3.14159265358979323846 / 180.0);
    }
    RigidWater(const RigidWater& o) : oh_(o.oh_), angle_(o.angle_), tol_(o.tol_), maxit_(o.maxit_), hh_(o.hh_) {}

    int constrainPositions(System& sys, const std::vector<Vec3>& reference) const {
        md_water* w = water_.bind(sys, params());
        const std::size_t n = sys.size();
        if (reference.size() != n) throw std::invalid_argument("RigidWater: reference size differs from the system");
        std::vector<double> rx(n), ry(n), rz(n);
        for (std::size_t i = 0; i < n; ++i) {
            rx[i] = reference[i].x;
            ry[i] = reference[i].y;
            rz[i] = reference[i].z;
        }
        int it = 0;
        detail::wcheck(md_water_shake(w, rx.data(), ry.data(), rz.data(), &it), w, "RigidWater::constrainPositions");
        std::vector<Vec3> r;
        water_.download(0, r);
        for (std::size_t i = 0; i < n; ++i) sys.setPosition(i, r[i]);
        return it;
    }
    int constrainVelocities(System& sys) const {
        md_water* w = water_.bind(sys, params());
        int it = 0;
        detail::wcheck(md_water_rattle(w, &it), w, "RigidWater::constrainVelocities");
        std::vector<Vec3> v;
        water_.download(1, v);
        for (std::size_t i = 0; i < v.size(); ++i) sys.setVelocity(i, v[i]);
        return it;
    }
    double ohDistance() const { return oh_; }
    double hhDistance() const { return hh_; }
    // 6 per rigid molecule, 3 fewer with the centre-of-mass motion removed
    static std::size_t degreesOfFreedom(std::size_t molecules, bool comRemoved) {
        return 6 * molecules - (comRemoved ? 3 : 0);
    }

private:
    md_water_params params() const {
        md_water_params p = detail::water_params(10.0, 5);
        p.r_oh = oh_;
        p.angle_hoh = angle_;
        p.tolerance = tol_;
        p.max_iterations = maxit_;
        return p;
    }
    double oh_, angle_, tol_;
    int maxit_;
    double hh_ = 0.0;
    detail::WaterBinding water_;
};

// SETTLE (Miyamoto & Kollman 1992): the closed-form replacement of
// RigidWater that the reference plans (constrain/rigid_water.hpp:20-24; no
// class in its sources). Same interface and the same fixed points as
// RigidWater's SHAKE and RATTLE, solved exactly in one pass per molecule on
// the device (both calls return 1); std::runtime_error when a molecule's step
// has no solution. The two hydrogen masses of a molecule must be equal.
class Settle {
public:
    Settle(double ohDistance, double hohAngleDegrees) : oh_(ohDistance), angle_(hohAngleDegrees) {
        if (!(ohDistance > 0.0)) throw std::invalid_argument("Settle: ohDistance must be positive");
        if (!(hohAngleDegrees > 0.0) || !(hohAngleDegrees < 180.0))
            throw std::invalid_argument("Settle: hohAngleDegrees must lie in (0, 180)");
        hh_ = 2.0 * oh_ * std::sin(0.5 * hohAngleDegrees * 3.14159265358979323846 / 180.0);
    }
    Settle(const Settle& o) : oh_(o.oh_), angle_(o.angle_), hh_(o.hh_) {}

    int constrainPositions(System& sys, const std::vector<Vec3>& reference) const {
        md_water* w = water_.bind(sys, params());
        const std::size_t n = sys.size();
        if (reference.size() != n) throw std::invalid_argument("Settle: reference size differs from the system");
        std::vector<double> rx(n), ry(n), rz(n);
        for (std::size_t i = 0; i < n; ++i) {
            rx[i] = reference[i].x;
            ry[i] = reference[i].y;
            rz[i] = reference[i].z;
        }
        detail::wcheck(md_water_settle(w, rx.data(), ry.data(), rz.data()), w, "Settle::constrainPositions");
        std::vector<Vec3> r;
        water_.download(0, r);
        for (std::size_t i = 0; i < n; ++i) sys.setPosition(i, r[i]);
        return 1;
    }
    int constrainVelocities(System& sys) const {
        md_water* w = water_.bind(sys, params());
        detail::wcheck(md_water_settle_velocities(w), w, "Settle::constrainVelocities");
        std::vector<Vec3> v;
        water_.download(1, v);
        for (std::size_t i = 0; i < v.size(); ++i) sys.setVelocity(i, v[i]);
        return 1;
    }
    double ohDistance() const { return oh_; }
    double hhDistance() const { return hh_; }
    static std::size_t degreesOfFreedom(std::size_t molecules, bool comRemoved) {
        return RigidWater::degreesOfFreedom(molecules, comRemoved);
    }

private:
    md_water_params params() const {
        md_water_params p = detail::water_params(10.0, 5);
        p.r_oh = oh_;
        p.angle_hoh = angle_;
        return p;
    }
    double oh_, angle_;
    double hh_ = 0.0;
    detail::WaterBinding water_;
};

// integrate/constrained_velocity_verlet.hpp:22-68: kick, drift (no wrap),
// SHAKE against the step's starting positions, the velocity corrected by the
// SHAKE displacement over dt, forces, kick, RATTLE. The kicks are the
// reference's plain multiply-add v + f (dt / 2 / m). Forces must be current
// on entry. The constraint is RigidWater, as in the reference, or Settle
// (class template argument deduction picks it from the constructor, so
// `ConstrainedVelocityVerlet integrator(dt, rigid);` reads as the reference).
template <class Constraint = RigidWater>
class ConstrainedVelocityVerlet {
public:
    ConstrainedVelocityVerlet(double dt, Constraint constraint) : dt_(dt), constraint_(constraint) {}
    template <class ForceFn>
    void step(System& sys, ForceFn&& computeForces) {
        const std::size_t n = sys.size();
        reference_.resize(n);
        unconstrained_.resize(n);
        for (std::size_t i = 0; i < n; ++i) reference_[i] = sys.position(i);
        for (std::size_t i = 0; i < n; ++i)
            sys.setVelocity(i, sys.velocity(i) + sys.force(i) * (0.5 * dt_ / sys.mass(i)));
        for (std::size_t i = 0; i < n; ++i) {
            sys.setPosition(i, sys.position(i) + sys.velocity(i) * dt_);
            unconstrained_[i] = sys.position(i);
        }
        constraint_.constrainPositions(sys, reference_);
        for (std::size_t i = 0; i < n; ++i) {
            const Vec3 shift = sys.box().minimumImage(sys.position(i) - unconstrained_[i]);
            sys.setVelocity(i, sys.velocity(i) + shift * (1.0 / dt_));
        }
        computeForces(sys);
        for (std::size_t i = 0; i < n; ++i)
            sys.setVelocity(i, sys.velocity(i) + sys.force(i) * (0.5 * dt_ / sys.mass(i)));
        constraint_.constrainVelocities(sys);
    }
    double dt() const { return dt_; }
    const Constraint& constraint() const { return constraint_; }

private:
    double dt_;
    Constraint constraint_;
    std::vector<Vec3> reference_, unconstrained_;
};

}  // namespace gpu
}  // namespace moldyn
