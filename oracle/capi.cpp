// Oracle C entry points for the Python checker (ctypes). Test infrastructure
// only (see oracle.hpp). oracle/ref_shim.cpp exports the same symbols over the
// reference's own classes, so one Python wrapper drives either library.
#include <cstring>
#include <memory>
#include <string>

#include "oracle.hpp"

using namespace oracle;

namespace {
thread_local std::string g_err;

State make_state(long n, const double* L, const double* x, const double* y, const double* z) {
    State s;
    for (int d = 0; d < 3; ++d) s.L[d] = L[d];
    s.resize(static_cast<std::size_t>(n));
    for (long i = 0; i < n; ++i) {
        s.x[i] = x[i];
        s.y[i] = y[i];
        s.z[i] = z[i];
    }
    return s;
}

void copy_out(const std::vector<double>& v, double* dst) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return -1;
    } catch (const std::runtime_error& e) {
        g_err = e.what();
        return -2;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -3;
    }
}

struct Sim {
    State s;
    LJ lj;
    bool use_list = false;
    Verlet verlet;
    int kind = 0;  // 0 VV, 1 BAOAB, 2 NHC
    double dt = 0.0;
    std::unique_ptr<Baoab> baoab;
    std::unique_ptr<NoseHoover> nhc;
    EV last;
};

double sim_forces(Sim& h, State& s) {
    if (h.use_list) {
        h.verlet.rebuild_if_needed(s);
        h.last = lj_forces_verlet(h.lj, s, h.verlet);
    } else {
        h.last = lj_forces_brute(h.lj, s);
    }
    return h.last.energy;
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
int orc_is_reference() { return 0; }

int orc_fcc_lattice(int cells, double rho, double T, unsigned long long seed, double* L, double* x,
                    double* y, double* z, double* vx, double* vy, double* vz) {
    State s = fcc_lattice(cells, rho, T, seed);
    for (int d = 0; d < 3; ++d) L[d] = s.L[d];
    copy_out(s.x, x); copy_out(s.y, y); copy_out(s.z, z);
    copy_out(s.vx, vx); copy_out(s.vy, vy); copy_out(s.vz, vz);
    return 0;
}

int orc_lj_ev_brute(long n, const double* L, const double* x, const double* y, const double* z,
                    double sig, double eps, double rc, int lfs, double* ev) {
    const State s = make_state(n, L, x, y, z);
    const EV r = lj_energy_virial(LJ(sig, eps, rc, lfs != 0), s);
    ev[0] = r.energy;
    ev[1] = r.virial;
    return 0;
}

// mode 0: brute force, 1: linked cells at rc, 2: Verlet list at rc + skin.
int orc_lj_forces(long n, const double* L, const double* x, const double* y, const double* z,
                  double sig, double eps, double rc, int lfs, int mode, double skin, double* fx,
                  double* fy, double* fz, double* ev) {
    return guarded([&] {
        State s = make_state(n, L, x, y, z);
        const LJ lj(sig, eps, rc, lfs != 0);
        EV r;
        if (mode == 0) {
            r = lj_forces_brute(lj, s);
        } else if (mode == 1) {
            Cells c(s.L, rc);
            c.build(s);
            r = lj_forces_cells(lj, s, c);
        } else {
            Verlet v(s.L, rc, skin);
            v.build(s);
            r = lj_forces_verlet(lj, s, v);
        }
        copy_out(s.fx, fx); copy_out(s.fy, fy); copy_out(s.fz, fz);
        ev[0] = r.energy;
        ev[1] = r.virial;
        return 0;
    });
}

// The pair terms of orc_lj_forces' linked-cell (mode 1) or brute-force
// (mode 0) evaluation summed with compensated summation: out = {E, W, pairs
// inside the cutoff}. A checker for systems of millions of pairs.
int orc_lj_ev_compensated(long n, const double* L, const double* x, const double* y,
                          const double* z, double sig, double eps, double rc, int lfs, int mode,
                          double* out) {
    return guarded([&] {
        const State s = make_state(n, L, x, y, z);
        const LJ lj(sig, eps, rc, lfs != 0);
        EV comp;
        long np = 0;
        if (mode == 1) {
            Cells c(s.L, rc);
            c.build(s);
            if (!c.usable()) throw std::invalid_argument("cell list is not usable");
            lj_pair_sum(lj, s, nullptr, [&](const PairFn& fn) { c.pairs(fn); }, &comp, &np);
        } else {
            lj_pair_sum(lj, s, nullptr, [&](const PairFn& fn) { brute_pairs(s.n(), fn); }, &comp,
                        &np);
        }
        out[0] = comp.energy;
        out[1] = comp.virial;
        out[2] = static_cast<double>(np);
        return 0;
    });
}

int orc_lj_lrc(long n, const double* L, double sig, double eps, double rc, int lfs, double* out) {
    State s;
    for (int d = 0; d < 3; ++d) s.L[d] = L[d];
    s.resize(static_cast<std::size_t>(n));
    const LJ lj(sig, eps, rc, lfs != 0);
    out[0] = lj_lrc_energy(lj, s);
    out[1] = lj_lrc_pressure(lj, s);
    return 0;
}

int orc_cells_usable(const double* L, double cutoff) {
    return Cells(L, cutoff).usable() ? 1 : 0;
}

// Pairs (i<j) in build order, flattened; returns the pair count, or -1 when
// `cap` pairs do not fit (npairs still reports the count).
long orc_verlet_pairs(long n, const double* L, const double* x, const double* y, const double* z,
                      double rc, double skin, unsigned int* out, long cap) {
    const State s = make_state(n, L, x, y, z);
    Verlet v(s.L, rc, skin);
    v.build(s);
    const long np = static_cast<long>(v.list.size());
    if (np > cap) return -np - 1;
    for (long k = 0; k < np; ++k) {
        out[2 * k] = static_cast<unsigned int>(v.list[static_cast<std::size_t>(k)].first);
        out[2 * k + 1] = static_cast<unsigned int>(v.list[static_cast<std::size_t>(k)].second);
    }
    return np;
}

int orc_needs_rebuild(long n, const double* L, const double* rx, const double* ry, const double* rz,
                      const double* x, const double* y, const double* z, double rc, double skin) {
    const State ref = make_state(n, L, rx, ry, rz);
    Verlet v(ref.L, rc, skin);
    v.build(ref);
    return v.needs_rebuild(make_state(n, L, x, y, z)) ? 1 : 0;
}

// Simulation handle. kind 0: velocity Verlet(dt); 1: BAOAB(dt, p1 = gamma,
// p2 = T, seed); 2: Nose-Hoover chain(dt, p1 = T, p2 = tau, M, dof).
// skin < 0 evaluates forces by brute force (the reference's unit tests),
// otherwise through a Verlet list with rebuild-if-needed before each force
// call (the reference driver's force callback). Forces are primed once.
void* orc_sim_create(long n, const double* L, const double* x, const double* y, const double* z,
                     const double* vx, const double* vy, const double* vz, const double* m,
                     double sig, double eps, double rc, int lfs, double skin, int kind, double dt,
                     double p1, double p2, int M, long dof, unsigned long long seed) {
    try {
        auto* h = new Sim();
        h->s = make_state(n, L, x, y, z);
        for (long i = 0; i < n; ++i) {
            h->s.vx[i] = vx[i]; h->s.vy[i] = vy[i]; h->s.vz[i] = vz[i];
            if (m) h->s.m[i] = m[i];
        }
        h->lj = LJ(sig, eps, rc, lfs != 0);
        h->kind = kind;
        h->dt = dt;
        if (skin >= 0.0) {
            h->use_list = true;
            h->verlet = Verlet(h->s.L, rc, skin);
            h->verlet.build(h->s);
        }
        if (kind == 1) h->baoab.reset(new Baoab(dt, p1, p2, seed));
        if (kind == 2) h->nhc.reset(new NoseHoover(dt, p1, p2, M, static_cast<std::size_t>(dof)));
        sim_forces(*h, h->s);
        return h;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_sim_destroy(void* p) { delete static_cast<Sim*>(p); }

// Runs nsteps; after each step records PE (pair energy of the step's force
// call), KE, the virial and the chain's conserved-quantity contribution.
int orc_sim_run(void* p, long nsteps, double* pe, double* ke, double* vir, double* hc) {
    auto& h = *static_cast<Sim*>(p);
    ForceFn fn = [&h](State& s) { return sim_forces(h, s); };
    return guarded([&] {
        for (long k = 0; k < nsteps; ++k) {
            if (h.kind == 0) vv_step(h.s, h.dt, fn);
            else if (h.kind == 1) h.baoab->step(h.s, fn);
            else h.nhc->step(h.s, fn);
            if (pe) pe[k] = h.last.energy;
            if (vir) vir[k] = h.last.virial;
            if (ke) ke[k] = kinetic_energy(h.s);
            if (hc) hc[k] = h.nhc ? h.nhc->conserved_contribution() : 0.0;
        }
        return 0;
    });
}

int orc_sim_get(void* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                double* fx, double* fy, double* fz) {
    auto& s = static_cast<Sim*>(p)->s;
    copy_out(s.x, x); copy_out(s.y, y); copy_out(s.z, z);
    copy_out(s.vx, vx); copy_out(s.vy, vy); copy_out(s.vz, vz);
    copy_out(s.fx, fx); copy_out(s.fy, fy); copy_out(s.fz, fz);
    return 0;
}

int orc_sim_set_velocities(void* p, const double* vx, const double* vy, const double* vz) {
    auto& s = static_cast<Sim*>(p)->s;
    for (std::size_t i = 0; i < s.n(); ++i) {
        s.vx[i] = vx[i]; s.vy[i] = vy[i]; s.vz[i] = vz[i];
    }
    return 0;
}

long orc_sim_rebuilds(void* p) { return static_cast<long>(static_cast<Sim*>(p)->verlet.builds); }

double orc_sim_conserved(void* p) {
    auto& h = *static_cast<Sim*>(p);
    return h.nhc ? h.nhc->conserved_contribution() : 0.0;
}

double orc_coulomb_prefactor() { return coulomb_prefactor(); }

// Ewald terms (real, recip, self, intra); alpha <= 0 selects 5.6/min(L).
int orc_ewald(long n, const double* L, const double* x, const double* y, const double* z,
              const double* q, const long* mol, double rc, int kmax, double alpha, int forces,
              double* fx, double* fy, double* fz, double* terms) {
    State s = make_state(n, L, x, y, z);
    for (long i = 0; i < n; ++i) {
        s.q[i] = q[i];
        s.mol[i] = mol[i];
    }
    const EwaldTerms t = ewald(s, rc, kmax, alpha, forces != 0, &s);
    if (forces) {
        copy_out(s.fx, fx); copy_out(s.fy, fy); copy_out(s.fz, fz);
    }
    terms[0] = t.real; terms[1] = t.recip; terms[2] = t.self; terms[3] = t.intra;
    return 0;
}

// PME reciprocal energy (and forces, when asked) over the Ewald k-set.
int orc_pme(long n, const double* L, const double* x, const double* y, const double* z,
            const double* q, int kmax, double alpha, const int* K, int order, int forces,
            double* fx, double* fy, double* fz, double* recip) {
    return guarded([&] {
        State s = make_state(n, L, x, y, z);
        for (long i = 0; i < n; ++i) s.q[i] = q[i];
        *recip = pme_recip(s, kmax, alpha, K, order, forces != 0, &s);
        if (forces) {
            copy_out(s.fx, fx); copy_out(s.fy, fy); copy_out(s.fz, fz);
        }
        return 0;
    });
}

double orc_pme_bsq(int order, int K, int m) { return pme_bsq(order, K, m); }

namespace {
State water_state(long n, const double* L, const double* x, const double* y, const double* z) {
    const Spce p;
    State s = make_state(n, L, x, y, z);
    for (long i = 0; i < n; ++i) {
        const bool oxy = (i % 3) == 0;
        s.m[i] = oxy ? p.mO : p.mH;
        s.q[i] = oxy ? -2.0 * p.qH : p.qH;
        s.mol[i] = i / 3;
    }
    return s;
}
}  // namespace

// SPC/E (O,H,H triples): forces (zeroed first) and PE in Kelvin.
int orc_spce_forces(long n, const double* L, const double* x, const double* y, const double* z,
                    double rc, int kmax, double alpha, double* fx, double* fy, double* fz,
                    double* pe) {
    return guarded([&] {
        State s = water_state(n, L, x, y, z);
        *pe = spce_forces(Spce(), rc, kmax, alpha, s);
        copy_out(s.fx, fx); copy_out(s.fy, fy); copy_out(s.fz, fz);
        return 0;
    });
}

// SHAKE positions (x,y,z updated in place) against reference positions;
// returns the iteration count or a negative status.
int orc_rigid_shake(long n, const double* L, double* x, double* y, double* z, const double* m,
                    const double* rx, const double* ry, const double* rz, double oh, double angle,
                    double tol, int maxit) {
    return guarded([&] {
        State s = make_state(n, L, x, y, z);
        for (long i = 0; i < n; ++i) s.m[i] = m[i];
        std::vector<V3> ref(static_cast<std::size_t>(n));
        for (long i = 0; i < n; ++i) ref[i] = V3{rx[i], ry[i], rz[i]};
        const int it = Rigid(oh, angle, tol, maxit).shake(s, ref);
        copy_out(s.x, x); copy_out(s.y, y); copy_out(s.z, z);
        return it;
    });
}

int orc_rigid_rattle(long n, const double* L, const double* x, const double* y, const double* z,
                     const double* m, double* vx, double* vy, double* vz, double oh, double angle,
                     double tol, int maxit) {
    return guarded([&] {
        State s = make_state(n, L, x, y, z);
        for (long i = 0; i < n; ++i) {
            s.m[i] = m[i]; s.vx[i] = vx[i]; s.vy[i] = vy[i]; s.vz[i] = vz[i];
        }
        const int it = Rigid(oh, angle, tol, maxit).rattle(s);
        copy_out(s.vx, vx); copy_out(s.vy, vy); copy_out(s.vz, vz);
        return it;
    });
}

// SETTLE positions (x,y,z updated in place) against reference positions;
// returns 1 or a negative status. No reference symbol (the reference plans
// SETTLE, constrain/rigid_water.hpp:20-24).
int orc_settle_positions(long n, const double* L, double* x, double* y, double* z, const double* m,
                         const double* rx, const double* ry, const double* rz, double oh,
                         double angle) {
    return guarded([&] {
        State s = make_state(n, L, x, y, z);
        for (long i = 0; i < n; ++i) s.m[i] = m[i];
        std::vector<V3> ref(static_cast<std::size_t>(n));
        for (long i = 0; i < n; ++i) ref[i] = V3{rx[i], ry[i], rz[i]};
        const int it = Settle(oh, angle).positions(s, ref);
        copy_out(s.x, x); copy_out(s.y, y); copy_out(s.z, z);
        return it;
    });
}

int orc_settle_velocities(long n, const double* L, const double* x, const double* y,
                          const double* z, const double* m, double* vx, double* vy, double* vz,
                          double oh, double angle) {
    return guarded([&] {
        State s = make_state(n, L, x, y, z);
        for (long i = 0; i < n; ++i) {
            s.m[i] = m[i]; s.vx[i] = vx[i]; s.vy[i] = vy[i]; s.vz[i] = vz[i];
        }
        const int it = Settle(oh, angle).velocities(s);
        copy_out(s.vx, vx); copy_out(s.vy, vy); copy_out(s.vz, vz);
        return it;
    });
}

}  // extern "C"

namespace {
// Rigid SPC/E dynamics with constraint C (Rigid or Settle).
template <class Constraint>
int water_run(long n, const double* L, double* x, double* y, double* z, double* vx, double* vy,
              double* vz, double rc, int kmax, double alpha, double dt, long nsteps, double* pe,
              double* ke) {
    return guarded([&] {
        State s = water_state(n, L, x, y, z);
        for (long i = 0; i < n; ++i) {
            s.vx[i] = vx[i]; s.vy[i] = vy[i]; s.vz[i] = vz[i];
        }
        const Spce p;
        const Constraint c(p.rOH, p.angleHOH);
        double last = 0.0;
        ForceFn fn = [&](State& st) {
            last = spce_forces(p, rc, kmax, alpha, st);
            return last;
        };
        fn(s);
        for (long k = 0; k < nsteps; ++k) {
            cvv_step(s, dt, c, fn);
            if (pe) pe[k] = last;
            if (ke) ke[k] = kinetic_energy(s);
        }
        copy_out(s.x, x); copy_out(s.y, y); copy_out(s.z, z);
        copy_out(s.vx, vx); copy_out(s.vy, vy); copy_out(s.vz, vz);
        return 0;
    });
}
}  // namespace

extern "C" {

// Rigid SPC/E dynamics: constrained velocity Verlet with the full SPC/E
// force field; state updated in place; PE (K) and KE per step.
int orc_water_run(long n, const double* L, double* x, double* y, double* z, double* vx, double* vy,
                  double* vz, double rc, int kmax, double alpha, double dt, long nsteps, double* pe,
                  double* ke) {
    return water_run<Rigid>(n, L, x, y, z, vx, vy, vz, rc, kmax, alpha, dt, nsteps, pe, ke);
}

// The same with SETTLE in place of SHAKE / RATTLE (no reference symbol).
int orc_water_run_settle(long n, const double* L, double* x, double* y, double* z, double* vx,
                         double* vy, double* vz, double rc, int kmax, double alpha, double dt,
                         long nsteps, double* pe, double* ke) {
    return water_run<Settle>(n, L, x, y, z, vx, vy, vz, rc, kmax, alpha, dt, nsteps, pe, ke);
}

// Per-type-pair LJ; sig/eps/rc are ntypes*ntypes row-major symmetric tables;
// out[0]=E, out[1]=W; e_ab/w_ab/y_ab are ntypes*ntypes tables.
int orc_mix_lj(long n, const double* L, const double* x, const double* y, const double* z,
               const int* type, int ntypes, const double* sig, const double* eps, const double* rc,
               int lfs, int forces, double* fx, double* fy, double* fz, double* out, double* e_ab,
               double* w_ab, double* y_ab) {
    State s = make_state(n, L, x, y, z);
    MixLJ p;
    p.ntypes = ntypes;
    const std::size_t nt2 = static_cast<std::size_t>(ntypes * ntypes);
    p.sigma.assign(sig, sig + nt2);
    p.epsilon.assign(eps, eps + nt2);
    p.cutoff.assign(rc, rc + nt2);
    p.lfs = lfs != 0;
    std::vector<int> t(type, type + n);
    const MixEV r = mix_lj(p, s, t, forces != 0, &s);
    if (forces) {
        copy_out(s.fx, fx); copy_out(s.fy, fy); copy_out(s.fz, fz);
    }
    out[0] = r.energy;
    out[1] = r.virial;
    copy_out(r.e_ab, e_ab); copy_out(r.w_ab, w_ab); copy_out(r.y_ab, y_ab);
    return 0;
}

namespace {
MixLJ make_mix(int ntypes, const double* sig, const double* eps, const double* rc, int lfs) {
    MixLJ p;
    p.ntypes = ntypes;
    const std::size_t nt2 = static_cast<std::size_t>(ntypes * ntypes);
    p.sigma.assign(sig, sig + nt2);
    p.epsilon.assign(eps, eps + nt2);
    p.cutoff.assign(rc, rc + nt2);
    p.lfs = lfs != 0;
    return p;
}
}  // namespace

// Steepest descent on the mixture; positions updated in place, energies[nsteps+1].
int orc_mix_minimize(long n, const double* L, double* x, double* y, double* z, const int* type,
                     int ntypes, const double* sig, const double* eps, const double* rc, int lfs,
                     int nsteps, double step, double max_disp, double* energies) {
    State s = make_state(n, L, x, y, z);
    const std::vector<int> t(type, type + n);
    const std::vector<double> e = mix_minimize(make_mix(ntypes, sig, eps, rc, lfs), s, t, nsteps,
                                               step, max_disp);
    copy_out(s.x, x); copy_out(s.y, y); copy_out(s.z, z);
    copy_out(e, energies);
    return 0;
}

// Velocity Verlet on the mixture (masses m or 1); state updated in place,
// energies / virials[nsteps+1].
int orc_mix_vv(long n, const double* L, double* x, double* y, double* z, double* vx, double* vy,
               double* vz, const double* m, const int* type, int ntypes, const double* sig,
               const double* eps, const double* rc, int lfs, double dt, int nsteps,
               double* energies, double* virials) {
    State s = make_state(n, L, x, y, z);
    for (long i = 0; i < n; ++i) {
        s.vx[i] = vx[i];
        s.vy[i] = vy[i];
        s.vz[i] = vz[i];
        if (m) s.m[i] = m[i];
    }
    const std::vector<int> t(type, type + n);
    std::vector<double> e, w;
    mix_vv(make_mix(ntypes, sig, eps, rc, lfs), s, t, dt, nsteps, e, w);
    copy_out(s.x, x); copy_out(s.y, y); copy_out(s.z, z);
    copy_out(s.vx, vx); copy_out(s.vy, vy); copy_out(s.vz, vz);
    copy_out(e, energies);
    copy_out(w, virials);
    return 0;
}

// Trajectory analyses over `frames` frames. Frame f holds box lengths
// L[3f..3f+2] and, for each of x, y, z, n values at xyz[(3f + d) n ...].
// orc_rdf writes g, the bin centres and the raw pair counts (hist, oracle
// only); orc_msd / orc_vacf write the max_lag + 1 correlation values and D.
int orc_rdf(long frames, long n, const double* L, const double* xyz, double r_max, long bins,
            double* g, double* centres, double* hist) {
    return guarded([&] {
        oracle::Rdf rdf(r_max, bins);
        for (long f = 0; f < frames; ++f) {
            const double* b = xyz + 3 * f * n;
            rdf.accumulate(make_state(n, L + 3 * f, b, b + n, b + 2 * n));
        }
        copy_out(rdf.g(), g);
        copy_out(rdf.centres(), centres);
        copy_out(rdf.hist, hist);
        return 0;
    });
}

int orc_msd(long frames, long n, long max_lag, const double* L, const double* xyz, double dt,
            long fit_from, long fit_to, double* msd, double* D) {
    return guarded([&] {
        Corr c(n, max_lag, true);
        for (long f = 0; f < frames; ++f) {
            const double* b = xyz + 3 * f * n;
            c.push(make_state(n, L + 3 * f, b, b + n, b + 2 * n));
        }
        copy_out(c.mean(), msd);
        if (D) *D = c.msd_diffusion(dt, fit_from, fit_to);
        return 0;
    });
}

int orc_vacf(long frames, long n, long max_lag, const double* vxyz, double dt, double* vacf,
             double* D) {
    return guarded([&] {
        Corr c(n, max_lag, false);
        const double L[3] = {1.0, 1.0, 1.0};
        for (long f = 0; f < frames; ++f) {
            const double* b = vxyz + 3 * f * n;
            State s = make_state(n, L, b, b + n, b + 2 * n);
            s.vx = s.x; s.vy = s.y; s.vz = s.z;
            c.push(s);
        }
        copy_out(c.normalised(), vacf);
        if (D) *D = c.vacf_diffusion(dt);
        return 0;
    });
}

}  // extern "C"
