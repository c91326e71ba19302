// CPU oracle for the moldyn force / neighbour / integrator path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
// path is compared against. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it, and only as the checker or as the
// timed CPU baseline. Nothing in the product (moleculardynamics_amd/,
// include/) may call into it.
//
// It is an independent C++17 restatement, written from the reference's
// semantics (VishnuR23/MolecularDynamics, src/moldyn/...), not a copy: state
// is a flat struct of arrays, every pass is a free function, and the pair
// sources are plain index enumerators. Each function cites the reference
// file:line whose behaviour it reproduces. Arithmetic is FP64 in the same
// operation order as the reference (g++ ISO mode: no FMA contraction), so
// the neighbour pair set is reproduced bit-exactly and E/W/F to rounding.
//
// Pins (see tests/test_oracle_*.py): NIST LJ 12-case table, NIST SPC/E 28
// values, the Coulomb prefactor, the exp02 108-atom NVE trajectories, LRC
// hand values, finite-difference gradients, Newton's third law, brute force
// vs cell/Verlet lists, the dt^2 slope, time reversibility, NHC H'
// conservation, BAOAB statistics, SHAKE/RATTLE invariants, and (when built)
// the reference compiled from its own sources (oracle/_ref).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <random>
#include <stdexcept>
#include <utility>
#include <vector>

namespace oracle {

struct V3 {
    double x, y, z;
};

// (a.x*b.x + a.y*b.y) + a.z*b.z, left to right, no contraction
// (core/vec3.hpp:59-65).
inline double dot3(V3 a, V3 b) {
    double s = a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    return s;
}

// Flat particle state. Positions, velocities and forces are separate arrays
// (core/system.hpp:52-59); charge defaults to 0 and molecule id to the atom's
// own index (core/system.cpp:17-21).
struct State {
    double L[3] = {0.0, 0.0, 0.0};
    std::vector<double> x, y, z, vx, vy, vz, fx, fy, fz, m, q;
    std::vector<long> mol;

    std::size_t n() const { return x.size(); }
    void resize(std::size_t n);
    V3 r(std::size_t i) const { return V3{x[i], y[i], z[i]}; }
    V3 v(std::size_t i) const { return V3{vx[i], vy[i], vz[i]}; }
    V3 f(std::size_t i) const { return V3{fx[i], fy[i], fz[i]}; }
    void set_r(std::size_t i, V3 a) { x[i] = a.x; y[i] = a.y; z[i] = a.z; }
    void set_v(std::size_t i, V3 a) { vx[i] = a.x; vy[i] = a.y; vz[i] = a.z; }
    double volume() const { return L[0] * L[1] * L[2]; }
    void clear_forces();
};

// ---- periodic box (core/box.cpp:12-33) -------------------------------------
// v - L*floor(v/L + 0.5): maps to [-L/2, L/2), with -L/2 kept and +L/2 sent to
// -L/2 (core/box.cpp:8-14).
inline double fold(double v, double len) {
    const double k = std::floor(v / len + 0.5);
    const double shift = len * k;
    return v - shift;
}
inline V3 min_image(const double L[3], V3 d) {
    return V3{fold(d.x, L[0]), fold(d.y, L[1]), fold(d.z, L[2])};
}

// ---- state reductions (core/system.cpp:124-164) ----------------------------
double kinetic_energy(const State& s);
V3 total_momentum(const State& s);
void remove_com_motion(State& s);

// ---- seeded inputs (util/random.hpp:37-76, util/lattice.cpp:15-64) ---------
// mt19937_64 with std::uniform_real_distribution<double>(0,1); uniform() is
// 1 - U so it lies in (0,1]; normal() is Marsaglia's polar method, keeping
// the second deviate as a spare.
class Rng {
public:
    explicit Rng(std::uint64_t seed) : eng_(seed) {}
    double uniform() { return 1.0 - unif_(eng_); }
    double normal();

private:
    std::mt19937_64 eng_;
    std::uniform_real_distribution<double> unif_{0.0, 1.0};
    bool spare_ok_ = false;
    double spare_ = 0.0;
};

State fcc_lattice(int cells, double density, double temperature, std::uint64_t seed);

// ---- Lennard-Jones (forcefield/lennard_jones.cpp:32-151) -------------------
struct EV {
    double energy = 0.0;
    double virial = 0.0;  // raw sum of w(r), not divided by 3
};

struct LJ {
    double sigma = 1.0, epsilon = 1.0, cutoff = 2.5;
    bool lfs = false;  // LinearForceShift when true, Truncated otherwise
    double u_rc = 0.0, dudr_rc = 0.0;
    LJ() = default;
    LJ(double sig, double eps, double rc, bool linear_force_shift);
};

// Pair enumerations; every callback receives (i, j) with i < j.
using PairFn = std::function<void(std::size_t, std::size_t)>;
void brute_pairs(std::size_t n, const PairFn& fn);  // lennard_jones.cpp:18-29

// Linked cells (neighbor/cell_list.{hpp:55-108, cpp:8-68}).
struct Cells {
    double L[3] = {0, 0, 0};
    double width_req = 0.0;
    int nc[3] = {1, 1, 1};
    std::vector<int> head, next;
    Cells() = default;
    Cells(const double box[3], double cutoff);
    bool usable() const { return nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3; }
    void build(const State& s);
    void pairs(const PairFn& fn) const;
    int index(int ix, int iy, int iz) const { return ix + nc[0] * (iy + nc[1] * iz); }
};

// Verlet list with skin (neighbor/verlet_list.cpp:5-73).
struct Verlet {
    double L[3] = {0, 0, 0};
    double cutoff = 0.0, skin = 0.0;
    Cells cells;
    std::vector<std::pair<std::size_t, std::size_t>> list;
    std::vector<V3> ref;
    std::size_t builds = 0;
    Verlet() = default;
    Verlet(const double box[3], double rc, double sk);
    void build(const State& s);
    bool needs_rebuild(const State& s) const;
    void rebuild_if_needed(const State& s) {
        if (needs_rebuild(s)) build(s);
    }
    void pairs(const PairFn& fn) const {
        for (const auto& p : list) fn(p.first, p.second);
    }
};

// Shared pair kernel (lennard_jones.cpp:50-88): one evaluation per offered
// pair, cutoff test r2 < rc2, optional force accumulation. When
// `compensated` is given it also receives E and W summed with Neumaier's
// compensated summation over the same pair terms (a checker for sums of
// millions of terms, where the reference's sequential FP64 sum itself carries
// rounding error of order n * 2^-53 * sum |u|); `pairs` (or null) counts the
// pairs inside the cutoff.
EV lj_pair_sum(const LJ& lj, const State& s, State* sink,
               const std::function<void(const PairFn&)>& source, EV* compensated = nullptr,
               long* pairs = nullptr);
EV lj_energy_virial(const LJ& lj, const State& s);                 // :90-92
EV lj_forces_brute(const LJ& lj, State& s);                       // :94-97
EV lj_forces_cells(const LJ& lj, State& s, const Cells& c);       // :99-114
EV lj_forces_verlet(const LJ& lj, State& s, const Verlet& v);     // :116-119
double lj_lrc_energy(const LJ& lj, const State& s);               // :121-135
double lj_lrc_pressure(const LJ& lj, const State& s);             // :137-151

// ---- integrators ------------------------------------------------------------
using ForceFn = std::function<double(State&)>;

// Kick-drift-kick with single-rounding fma (integrate/velocity_verlet.hpp:48-74).
void vv_step(State& s, double dt, const ForceFn& forces);

// BAOAB Langevin (integrate/langevin.hpp:36-88).
struct Baoab {
    double dt, gamma, T;
    Rng rng;
    Baoab(double dt_, double gamma_, double T_, std::uint64_t seed)
        : dt(dt_), gamma(gamma_), T(T_), rng(seed) {}
    void step(State& s, const ForceFn& forces);
};

// Nose-Hoover chain, MTK96 with Suzuki-Yoshida weights
// (integrate/nose_hoover.{hpp:26-67, cpp:15-124}).
struct NoseHoover {
    double dt, T, tau;
    int M;
    std::size_t dof;
    std::vector<double> Q, xi, p;
    NoseHoover(double dt_, double T_, double tau_, int M_, std::size_t dof_);
    void step(State& s, const ForceFn& forces);
    double conserved_contribution() const;
    void half_chain(State& s, double delta);
};

// ---- electrostatics and rigid water (forcefield/ewald.cpp:17-188,
// forcefield/spce*.cpp, constrain/rigid_water.cpp:30-161) --------------------
double coulomb_prefactor();  // ewald.cpp:17-21
struct EwaldTerms {
    double real = 0, recip = 0, self = 0, intra = 0;
    double total() const { return real + recip + self + intra; }
};
// alpha <= 0 selects the reference convention alpha = 5.6 / min(L)
// (ewald.cpp:40); n2max <= 0 selects |n|^2 < kmax^2 + 2 (ewald.cpp:119).
EwaldTerms ewald(const State& s, double rc, int kmax, double alpha, bool accumulate_forces,
                 State* sink);

// Smooth particle-mesh Ewald (Essmann et al. 1995) for the reciprocal term
// over the same k-vector set as ewald() (pme.cpp): K[d] grid points per
// axis (at least 2 kmax + 1), B-spline order p (3 .. 12). Returns the
// reciprocal energy; forces are added to sink when asked. pme_bsq is the
// spline correction |b_d(m)|^2 of one axis.
double pme_recip(const State& s, int kmax, double alpha, const int K[3], int p,
                 bool accumulate_forces, State* sink);
double pme_bsq(int p, int K, int m);

struct Spce {
    double qH = 0.42380, sigmaO = 3.16555789, epsilonO = 78.19743111;
    double rOH = 1.0, angleHOH = 109.47, mO = 15.9994, mH = 1.008;
};
// Dispersion on oxygens (Truncated, + LRC) plus Ewald on every site; forces
// zeroed first; returns PE in Kelvin (spce_forcefield.cpp:26-41).
double spce_forces(const Spce& p, double rc, int kmax, double alpha, State& s);
double spce_energy(const Spce& p, double rc, int kmax, double alpha, const State& s);

struct Rigid {
    double oh, hh, tol;
    int max_iter;
    Rigid(double oh_dist, double hoh_deg, double tolerance = 1e-12, int max_iterations = 500);
    int shake(State& s, const std::vector<V3>& reference) const;
    int rattle(State& s) const;
    static std::size_t dof(std::size_t molecules, bool com_removed);
};
// Closed-form rigid three-site constraints (SETTLE, Miyamoto & Kollman 1992),
// the replacement of Rigid's SHAKE/RATTLE that the reference plans
// (constrain/rigid_water.hpp:20-24): the same fixed points, solved exactly
// (settle.cpp). Molecules are O, H, H triples with equal hydrogen masses.
// Both return 1 (one pass); degenerate or over-large steps throw.
struct Settle {
    double oh, hh;
    Settle(double oh_dist, double hoh_deg);
    int positions(State& s, const std::vector<V3>& reference) const;
    int velocities(State& s) const;
};
// RATTLE-form constrained velocity Verlet
// (integrate/constrained_velocity_verlet.hpp:27-57), with SHAKE/RATTLE or
// with SETTLE as the constraint.
void cvv_step(State& s, double dt, const Rigid& rw, const ForceFn& forces);
void cvv_step(State& s, double dt, const Settle& st, const ForceFn& forces);

// ---- per-type-pair LJ (no reference symbol; built from the single-type
// kernel, lennard_jones.cpp:50-88, by giving each unordered type pair its own
// sigma, epsilon and cutoff) -------------------------------------------------
struct MixLJ {
    int ntypes = 1;
    std::vector<double> sigma, epsilon, cutoff;  // ntypes*ntypes, symmetric
    bool lfs = false;
};
struct MixEV {
    double energy = 0, virial = 0;
    // per type pair (a<=b, row-major ntypes*ntypes, symmetric): E_ab, W_ab and
    // Y_ab = sum 24 eps (24 s^12 - 6 s^6) (the sigma-derivative of w times sigma)
    std::vector<double> e_ab, w_ab, y_ab;
};
MixEV mix_lj(const MixLJ& p, const State& s, const std::vector<int>& type, bool forces,
             State* sink);
// Steepest descent (no reference symbol): nsteps moves
// x <- wrap(fma(gamma, F, x)), gamma = min(step, max_disp / max_i |F_i|)
// (gamma = step when every force is zero), brute-force forces. Returns the
// energies before the first move and after each one (nsteps + 1 values).
std::vector<double> mix_minimize(const MixLJ& p, State& s, const std::vector<int>& type,
                                 int nsteps, double step, double max_disp);
// nsteps of vv_step with per-type-pair forces; E_t and W_t for t = 0..nsteps.
void mix_vv(const MixLJ& p, State& s, const std::vector<int>& type, double dt, int nsteps,
            std::vector<double>& energy, std::vector<double>& virial);

// ---- trajectory analysis (analyze/rdf.cpp:16-71, msd.cpp:12-92,
// vacf.cpp:11-85) -------------------------------------------------------------
struct Rdf {
    double r_max;
    long bins;
    double width;
    std::vector<double> hist;  // pair counts, exact integers in FP64
    long frames = 0, natoms = 0;
    double sum_rho = 0.0;
    Rdf(double rmax, long nbins);
    void accumulate(const State& s);  // throws std::invalid_argument
    std::vector<double> centres() const;
    std::vector<double> g() const;
};
// One ring-buffered time correlation: the MSD (displacement = true) or the
// VACF. Per-atom vectors are packed x,y,z per atom.
struct Corr {
    long n, max_lag, period;
    bool msd;
    long pushed = 0;
    std::vector<double> prev, cur, ring, sum;
    std::vector<long> count;
    Corr(long natoms, long maxlag, bool displacement);
    void push(const State& s);
    std::vector<double> mean() const;        // msd() / the unnormalised VACF
    std::vector<double> normalised() const;  // vacf()
    double msd_diffusion(double dt, long from, long to) const;
    double vacf_diffusion(double dt) const;
};

}  // namespace oracle
