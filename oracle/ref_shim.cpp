// Checker-side driver for the reference compiled from its own sources
// (oracle/_ref). Exports the same orc_* C entry points as oracle/capi.cpp,
// implemented over the reference's classes (System, LennardJones, CellList,
// VerletList, VelocityVerlet, LangevinBAOAB, NoseHooverChain, Ewald,
// SpceForceField, RigidWater, ConstrainedVelocityVerlet), so the same Python
// wrapper can cross-check the oracle against the reference and time the
// reference as the CPU baseline. Test infrastructure only; it is built only
// where the reference sources are present, and nothing in the product loads
// it.
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "moldyn/analyze/msd.hpp"
#include "moldyn/analyze/rdf.hpp"
#include "moldyn/analyze/vacf.hpp"
#include "moldyn/constrain/rigid_water.hpp"
#include "moldyn/core/box.hpp"
#include "moldyn/core/system.hpp"
#include "moldyn/forcefield/ewald.hpp"
#include "moldyn/forcefield/lennard_jones.hpp"
#include "moldyn/forcefield/spce.hpp"
#include "moldyn/forcefield/spce_forcefield.hpp"
#include "moldyn/integrate/constrained_velocity_verlet.hpp"
#include "moldyn/integrate/langevin.hpp"
#include "moldyn/integrate/nose_hoover.hpp"
#include "moldyn/integrate/velocity_verlet.hpp"
#include "moldyn/neighbor/cell_list.hpp"
#include "moldyn/neighbor/verlet_list.hpp"
#include "moldyn/util/lattice.hpp"

using namespace moldyn;

namespace {
thread_local std::string g_err;

System make_system(long n, const double* L, const double* x, const double* y, const double* z) {
    System s(Box(Vec3{L[0], L[1], L[2]}));
    s.reserve(static_cast<std::size_t>(n));
    for (long i = 0; i < n; ++i) s.addAtom(Vec3{x[i], y[i], z[i]}, Vec3{0, 0, 0}, 1.0);
    return s;
}

void out_forces(const System& s, double* fx, double* fy, double* fz) {
    for (std::size_t i = 0; i < s.size(); ++i) {
        const Vec3 f = s.force(i);
        fx[i] = f.x; fy[i] = f.y; fz[i] = f.z;
    }
}

Truncation trunc(int lfs) { return lfs ? Truncation::LinearForceShift : Truncation::Truncated; }

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
    System s;
    LennardJones lj;
    std::unique_ptr<VerletList> verlet;
    int kind = 0;
    std::unique_ptr<VelocityVerlet> vv;
    std::unique_ptr<LangevinBAOAB> baoab;
    std::unique_ptr<NoseHooverChain> nhc;
    EnergyVirial last;
    Sim(System sys, LennardJones l) : s(std::move(sys)), lj(l) {}
    double forces(System& sys) {
        if (verlet) {
            verlet->rebuildIfNeeded(sys);
            last = lj.computeForces(sys, *verlet);
        } else {
            last = lj.computeForces(sys);
        }
        return last.energy;
    }
};
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
int orc_is_reference() { return 1; }

int orc_fcc_lattice(int cells, double rho, double T, unsigned long long seed, double* L, double* x,
                    double* y, double* z, double* vx, double* vy, double* vz) {
    const System s = fccLattice(cells, rho, T, seed);
    const Vec3 len = s.box().lengths();
    L[0] = len.x; L[1] = len.y; L[2] = len.z;
    for (std::size_t i = 0; i < s.size(); ++i) {
        const Vec3 r = s.position(i), v = s.velocity(i);
        x[i] = r.x; y[i] = r.y; z[i] = r.z;
        vx[i] = v.x; vy[i] = v.y; vz[i] = v.z;
    }
    return 0;
}

int orc_lj_ev_brute(long n, const double* L, const double* x, const double* y, const double* z,
                    double sig, double eps, double rc, int lfs, double* ev) {
    const System s = make_system(n, L, x, y, z);
    const EnergyVirial r = LennardJones(sig, eps, rc, trunc(lfs)).computeEnergyVirial(s);
    ev[0] = r.energy;
    ev[1] = r.virial;
    return 0;
}

int orc_lj_forces(long n, const double* L, const double* x, const double* y, const double* z,
                  double sig, double eps, double rc, int lfs, int mode, double skin, double* fx,
                  double* fy, double* fz, double* ev) {
    return guarded([&] {
        System s = make_system(n, L, x, y, z);
        const LennardJones lj(sig, eps, rc, trunc(lfs));
        EnergyVirial r;
        if (mode == 0) {
            r = lj.computeForces(s);
        } else if (mode == 1) {
            CellList c(s.box(), rc);
            c.build(s);
            r = lj.computeForces(s, c);
        } else {
            VerletList v(s.box(), rc, skin);
            v.build(s);
            r = lj.computeForces(s, v);
        }
        out_forces(s, fx, fy, fz);
        ev[0] = r.energy;
        ev[1] = r.virial;
        return 0;
    });
}

int orc_lj_lrc(long n, const double* L, double sig, double eps, double rc, int lfs, double* out) {
    System s(Box(Vec3{L[0], L[1], L[2]}));
    for (long i = 0; i < n; ++i) s.addAtom(Vec3{0, 0, 0}, Vec3{0, 0, 0}, 1.0);
    const LennardJones lj(sig, eps, rc, trunc(lfs));
    out[0] = lj.longRangeCorrection(s);
    out[1] = lj.longRangeCorrectionPressure(s);
    return 0;
}

int orc_cells_usable(const double* L, double cutoff) {
    return CellList(Box(Vec3{L[0], L[1], L[2]}), cutoff).usable() ? 1 : 0;
}

long orc_verlet_pairs(long n, const double* L, const double* x, const double* y, const double* z,
                      double rc, double skin, unsigned int* out, long cap) {
    const System s = make_system(n, L, x, y, z);
    VerletList v(s.box(), rc, skin);
    v.build(s);
    long np = 0;
    v.forEachPair([&](std::size_t i, std::size_t j) {
        if (np < cap) {
            out[2 * np] = static_cast<unsigned int>(i);
            out[2 * np + 1] = static_cast<unsigned int>(j);
        }
        ++np;
    });
    return np > cap ? -np - 1 : np;
}

int orc_needs_rebuild(long n, const double* L, const double* rx, const double* ry, const double* rz,
                      const double* x, const double* y, const double* z, double rc, double skin) {
    const System ref = make_system(n, L, rx, ry, rz);
    VerletList v(ref.box(), rc, skin);
    v.build(ref);
    return v.needsRebuild(make_system(n, L, x, y, z)) ? 1 : 0;
}

void* orc_sim_create(long n, const double* L, const double* x, const double* y, const double* z,
                     const double* vx, const double* vy, const double* vz, const double* m,
                     double sig, double eps, double rc, int lfs, double skin, int kind, double dt,
                     double p1, double p2, int M, long dof, unsigned long long seed) {
    try {
        System s(Box(Vec3{L[0], L[1], L[2]}));
        for (long i = 0; i < n; ++i)
            s.addAtom(Vec3{x[i], y[i], z[i]}, Vec3{vx[i], vy[i], vz[i]}, m ? m[i] : 1.0);
        auto* h = new Sim(std::move(s), LennardJones(sig, eps, rc, trunc(lfs)));
        h->kind = kind;
        if (skin >= 0.0) {
            h->verlet.reset(new VerletList(h->s.box(), rc, skin));
            h->verlet->build(h->s);
        }
        if (kind == 0) h->vv.reset(new VelocityVerlet(dt));
        if (kind == 1) h->baoab.reset(new LangevinBAOAB(dt, p1, p2, seed));
        if (kind == 2)
            h->nhc.reset(new NoseHooverChain(dt, p1, p2, M, static_cast<std::size_t>(dof)));
        h->forces(h->s);
        return h;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_sim_destroy(void* p) { delete static_cast<Sim*>(p); }

int orc_sim_run(void* p, long nsteps, double* pe, double* ke, double* vir, double* hc) {
    auto& h = *static_cast<Sim*>(p);
    auto fn = [&h](System& s) { return h.forces(s); };
    return guarded([&] {
        for (long k = 0; k < nsteps; ++k) {
            if (h.kind == 0) h.vv->step(h.s, fn);
            else if (h.kind == 1) h.baoab->step(h.s, fn);
            else h.nhc->step(h.s, fn);
            if (pe) pe[k] = h.last.energy;
            if (vir) vir[k] = h.last.virial;
            if (ke) ke[k] = h.s.kineticEnergy();
            if (hc) hc[k] = h.nhc ? h.nhc->conservedQuantityContribution() : 0.0;
        }
        return 0;
    });
}

int orc_sim_get(void* p, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                double* fx, double* fy, double* fz) {
    const System& s = static_cast<Sim*>(p)->s;
    for (std::size_t i = 0; i < s.size(); ++i) {
        const Vec3 r = s.position(i), v = s.velocity(i), f = s.force(i);
        if (x) { x[i] = r.x; y[i] = r.y; z[i] = r.z; }
        if (vx) { vx[i] = v.x; vy[i] = v.y; vz[i] = v.z; }
        if (fx) { fx[i] = f.x; fy[i] = f.y; fz[i] = f.z; }
    }
    return 0;
}

int orc_sim_set_velocities(void* p, const double* vx, const double* vy, const double* vz) {
    System& s = static_cast<Sim*>(p)->s;
    for (std::size_t i = 0; i < s.size(); ++i) s.setVelocity(i, Vec3{vx[i], vy[i], vz[i]});
    return 0;
}

long orc_sim_rebuilds(void* p) {
    auto& h = *static_cast<Sim*>(p);
    return h.verlet ? static_cast<long>(h.verlet->rebuildCount()) : 0;
}

double orc_sim_conserved(void* p) {
    auto& h = *static_cast<Sim*>(p);
    return h.nhc ? h.nhc->conservedQuantityContribution() : 0.0;
}

double orc_coulomb_prefactor() { return coulombPrefactorKelvinAngstrom(); }

// The reference fixes alpha = 5.6/min(L); a positive alpha is refused here.
int orc_ewald(long n, const double* L, const double* x, const double* y, const double* z,
              const double* q, const long* mol, double rc, int kmax, double alpha, int forces,
              double* fx, double* fy, double* fz, double* terms) {
    if (alpha > 0.0) {
        g_err = "reference Ewald has a fixed alpha";
        return -1;
    }
    System s(Box(Vec3{L[0], L[1], L[2]}));
    for (long i = 0; i < n; ++i)
        s.addAtom(Vec3{x[i], y[i], z[i]}, Vec3{0, 0, 0}, 1.0, q[i], static_cast<std::size_t>(mol[i]));
    const Ewald e(rc, kmax);
    const EwaldTerms t = forces ? e.computeForces(s) : e.computeEnergies(s);
    if (forces) out_forces(s, fx, fy, fz);
    terms[0] = t.real; terms[1] = t.reciprocal; terms[2] = t.self; terms[3] = t.intra;
    return 0;
}

namespace {
System water_system(long n, const double* L, const double* x, const double* y, const double* z) {
    const SpceParameters p;
    System s(Box(Vec3{L[0], L[1], L[2]}));
    for (long i = 0; i < n; ++i) {
        const bool oxy = (i % 3) == 0;
        s.addAtom(Vec3{x[i], y[i], z[i]}, Vec3{0, 0, 0}, oxy ? p.massOxygen : p.massHydrogen,
                  oxy ? -2.0 * p.qH : p.qH, static_cast<std::size_t>(i / 3));
    }
    return s;
}
}  // namespace

int orc_spce_forces(long n, const double* L, const double* x, const double* y, const double* z,
                    double rc, int kmax, double alpha, double* fx, double* fy, double* fz,
                    double* pe) {
    if (alpha > 0.0) {
        g_err = "reference Ewald has a fixed alpha";
        return -1;
    }
    return guarded([&] {
        System s = water_system(n, L, x, y, z);
        const SpceForceField ff(SpceParameters(), rc, kmax);
        *pe = ff.computeForces(s);
        out_forces(s, fx, fy, fz);
        return 0;
    });
}

int orc_rigid_shake(long n, const double* L, double* x, double* y, double* z, const double* m,
                    const double* rx, const double* ry, const double* rz, double oh, double angle,
                    double tol, int maxit) {
    return guarded([&] {
        System s(Box(Vec3{L[0], L[1], L[2]}));
        for (long i = 0; i < n; ++i) s.addAtom(Vec3{x[i], y[i], z[i]}, Vec3{0, 0, 0}, m[i]);
        std::vector<Vec3> ref(static_cast<std::size_t>(n));
        for (long i = 0; i < n; ++i) ref[static_cast<std::size_t>(i)] = Vec3{rx[i], ry[i], rz[i]};
        const int it = RigidWater(oh, angle, tol, maxit).constrainPositions(s, ref);
        for (long i = 0; i < n; ++i) {
            const Vec3 r = s.position(static_cast<std::size_t>(i));
            x[i] = r.x; y[i] = r.y; z[i] = r.z;
        }
        return it;
    });
}

int orc_rigid_rattle(long n, const double* L, const double* x, const double* y, const double* z,
                     const double* m, double* vx, double* vy, double* vz, double oh, double angle,
                     double tol, int maxit) {
    return guarded([&] {
        System s(Box(Vec3{L[0], L[1], L[2]}));
        for (long i = 0; i < n; ++i)
            s.addAtom(Vec3{x[i], y[i], z[i]}, Vec3{vx[i], vy[i], vz[i]}, m[i]);
        const int it = RigidWater(oh, angle, tol, maxit).constrainVelocities(s);
        for (long i = 0; i < n; ++i) {
            const Vec3 v = s.velocity(static_cast<std::size_t>(i));
            vx[i] = v.x; vy[i] = v.y; vz[i] = v.z;
        }
        return it;
    });
}

int orc_water_run(long n, const double* L, double* x, double* y, double* z, double* vx, double* vy,
                  double* vz, double rc, int kmax, double alpha, double dt, long nsteps, double* pe,
                  double* ke) {
    if (alpha > 0.0) {
        g_err = "reference Ewald has a fixed alpha";
        return -1;
    }
    return guarded([&] {
        System s = water_system(n, L, x, y, z);
        for (long i = 0; i < n; ++i) s.setVelocity(static_cast<std::size_t>(i), Vec3{vx[i], vy[i], vz[i]});
        const SpceParameters p;
        const SpceForceField ff(p, rc, kmax);
        double last = 0.0;
        auto fn = [&](System& st) {
            last = ff.computeForces(st);
            return last;
        };
        fn(s);
        ConstrainedVelocityVerlet cvv(dt, RigidWater(p.rOH, p.angleHOH));
        for (long k = 0; k < nsteps; ++k) {
            cvv.step(s, fn);
            if (pe) pe[k] = last;
            if (ke) ke[k] = s.kineticEnergy();
        }
        for (long i = 0; i < n; ++i) {
            const Vec3 r = s.position(static_cast<std::size_t>(i)), v = s.velocity(static_cast<std::size_t>(i));
            x[i] = r.x; y[i] = r.y; z[i] = r.z;
            vx[i] = v.x; vy[i] = v.y; vz[i] = v.z;
        }
        return 0;
    });
}

int orc_rdf(long frames, long n, const double* L, const double* xyz, double r_max, long bins,
            double* g, double* centres, double* /*hist: not exposed by Rdf*/) {
    return guarded([&] {
        Rdf rdf(r_max, static_cast<std::size_t>(bins));
        for (long f = 0; f < frames; ++f) {
            const double* b = xyz + 3 * f * n;
            rdf.accumulate(make_system(n, L + 3 * f, b, b + n, b + 2 * n));
        }
        const std::vector<double> gv = rdf.g(), cv = rdf.binCentres();
        if (g) std::memcpy(g, gv.data(), gv.size() * sizeof(double));
        if (centres) std::memcpy(centres, cv.data(), cv.size() * sizeof(double));
        return 0;
    });
}

int orc_msd(long frames, long n, long max_lag, const double* L, const double* xyz, double dt,
            long fit_from, long fit_to, double* msd, double* D) {
    return guarded([&] {
        Msd m(static_cast<std::size_t>(n), static_cast<std::size_t>(max_lag));
        for (long f = 0; f < frames; ++f) {
            const double* b = xyz + 3 * f * n;
            m.push(make_system(n, L + 3 * f, b, b + n, b + 2 * n));
        }
        const std::vector<double> v = m.msd();
        if (msd) std::memcpy(msd, v.data(), v.size() * sizeof(double));
        if (D)
            *D = m.diffusionCoefficient(dt, static_cast<std::size_t>(fit_from),
                                        static_cast<std::size_t>(fit_to));
        return 0;
    });
}

int orc_vacf(long frames, long n, long max_lag, const double* vxyz, double dt, double* vacf,
             double* D) {
    return guarded([&] {
        Vacf c(static_cast<std::size_t>(n), static_cast<std::size_t>(max_lag));
        for (long f = 0; f < frames; ++f) {
            const double* b = vxyz + 3 * f * n;
            System s(Box(Vec3{1.0, 1.0, 1.0}));
            for (long i = 0; i < n; ++i)
                s.addAtom(Vec3{0, 0, 0}, Vec3{b[i], b[n + i], b[2 * n + i]}, 1.0);
            c.push(s);
        }
        const std::vector<double> v = c.vacf();
        if (vacf) std::memcpy(vacf, v.data(), v.size() * sizeof(double));
        if (D) *D = c.diffusionCoefficient(dt);
        return 0;
    });
}

}  // extern "C"
