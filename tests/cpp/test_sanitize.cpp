// AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU oracle and of
// the library's host-side C ABI paths (built host-only, without device code:
// nothing here needs a GPU). Every oracle pass the tests use is exercised on
// small inputs, and every entry point that the library answers on the host
// (lattices, tail terms, capacities, recipe draws, Ewald parameters) is
// called with valid and invalid arguments, together with the argument
// checks of the handle-based entry points (null handles, null arrays, bad
// sizes), which must return an error code before touching the device. Any
// sanitizer report aborts the run with a non-zero status; the last line is
// "sanitize ok" on success. tests/test_sanitize.py builds and runs it.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/moldyn_hip.h"
#include "../../oracle/oracle.hpp"

static int g_fail = 0;
#define EXPECT(cond)                                                            \
    do {                                                                        \
        if (!(cond)) {                                                          \
            ++g_fail;                                                           \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);         \
        }                                                                       \
    } while (0)

static void oracle_passes() {
    using namespace oracle;
    State s = fcc_lattice(3, 0.8, 1.0, 12345);
    LJ lj(1.0, 1.0, 2.5, true);
    const EV a = lj_forces_brute(lj, s);
    Cells c(s.L, 2.5);
    c.build(s);
    State s2 = s;
    const EV b = c.usable() ? lj_forces_cells(lj, s2, c) : a;
    Verlet v(s.L, 2.5, 0.3);
    v.build(s);
    State s3 = s;
    const EV d = lj_forces_verlet(lj, s3, v);
    EXPECT(std::fabs(a.energy - b.energy) <= 1e-10 * std::fabs(a.energy));
    EXPECT(std::fabs(a.energy - d.energy) <= 1e-10 * std::fabs(a.energy));
    EXPECT(std::isfinite(lj_lrc_energy(LJ(1.0, 1.0, 2.5, false), s)));
    // a cell grid that is usable
    State big = fcc_lattice(6, 0.8442, 0.722, 7);
    Cells cb(big.L, 2.5);
    cb.build(big);
    EXPECT(cb.usable());
    State big2 = big;
    lj_forces_cells(lj, big2, cb);
    // integrators
    auto rebuild = [&](State& st) {
        v.rebuild_if_needed(st);
        return lj_forces_verlet(lj, st, v).energy;
    };
    for (int k = 0; k < 20; ++k) vv_step(s3, 0.005, rebuild);
    Baoab bo(0.005, 1.0, 1.0, 9);
    for (int k = 0; k < 20; ++k) bo.step(s3, rebuild);
    NoseHoover nh(0.005, 1.0, 0.5, 3, 3 * s3.n() - 3);
    for (int k = 0; k < 20; ++k) nh.step(s3, rebuild);
    EXPECT(std::isfinite(nh.conserved_contribution()));
    remove_com_motion(s3);
    // analysis
    Rdf r(2.4, 24);
    r.accumulate(s3);
    EXPECT(r.g().size() == 24u);
    Corr m(static_cast<long>(s3.n()), 5, true), va(static_cast<long>(s3.n()), 5, false);
    for (int k = 0; k < 8; ++k) {
        m.push(s3);
        va.push(s3);
        vv_step(s3, 0.005, rebuild);
    }
    EXPECT(std::isfinite(m.msd_diffusion(0.005, 1, 6)));
    EXPECT(std::isfinite(va.vacf_diffusion(0.005)));
    // mixture
    MixLJ mx;
    mx.ntypes = 2;
    mx.sigma = {1.0, 0.8, 0.8, 0.88};
    mx.epsilon = {1.0, 1.5, 1.5, 0.5};
    mx.cutoff = {2.5, 2.0, 2.0, 2.2};
    std::vector<int> type(s.n(), 0);
    for (std::size_t i = 0; i < s.n(); i += 5) type[i] = 1;
    State sm = fcc_lattice(3, 0.8, 1.0, 3);
    const MixEV me = mix_lj(mx, sm, type, true, &sm);
    EXPECT(me.e_ab.size() == 4u);
    std::vector<double> en = mix_minimize(mx, sm, type, 5, 1e-3, 0.05), ee, ww;
    EXPECT(en.size() == 6u);
    mix_vv(mx, sm, type, 0.002, 5, ee, ww);
    EXPECT(ee.size() == 6u);
}

static void water_passes() {
    using namespace oracle;
    // two rigid SPC/E molecules in a 20 A box
    State s;
    s.L[0] = s.L[1] = s.L[2] = 20.0;
    s.resize(6);
    const double q[3] = {-0.8476, 0.4238, 0.4238};
    const double m[3] = {15.9994, 1.008, 1.008};
    const double pos[6][3] = {{0, 0, 0},       {1, 0, 0},       {-0.333807, 0.942641, 0},
                              {3.1, 0.4, 0.2}, {4.1, 0.4, 0.2}, {2.766193, 1.342641, 0.2}};
    for (int i = 0; i < 6; ++i) {
        s.x[i] = pos[i][0];
        s.y[i] = pos[i][1];
        s.z[i] = pos[i][2];
        s.q[i] = q[i % 3];
        s.m[i] = m[i % 3];
        s.mol[i] = i / 3;
    }
    Spce p;
    const double pe = spce_forces(p, 9.0, 3, 0.0, s);
    EXPECT(std::isfinite(pe));
    Rigid rw(1.0, 109.47);
    std::vector<V3> ref;
    for (int i = 0; i < 6; ++i) ref.push_back(s.r(i));
    EXPECT(rw.shake(s, ref) >= 0);
    EXPECT(rw.rattle(s) >= 0);
    cvv_step(s, 0.001, rw, [&](State& st) { return spce_forces(p, 9.0, 3, 0.0, st); });
    Settle sw(1.0, 109.47);
    for (int i = 0; i < 6; ++i) ref[i] = s.r(i);
    EXPECT(sw.positions(s, ref) == 1);
    EXPECT(sw.velocities(s) == 1);
    cvv_step(s, 0.001, sw, [&](State& st) { return spce_forces(p, 9.0, 3, 0.0, st); });
}

static void library_host_paths() {
    // seeded lattice, bit for bit the oracle's
    const int cells = 4;
    const long n = 4L * cells * cells * cells;
    std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n);
    double L[3];
    EXPECT(md_fcc_lattice(cells, 0.8442, 0.722, 99, L, x.data(), y.data(), z.data(), vx.data(),
                          vy.data(), vz.data()) == MD_OK);
    const oracle::State o = oracle::fcc_lattice(cells, 0.8442, 0.722, 99);
    EXPECT(std::memcmp(x.data(), o.x.data(), n * sizeof(double)) == 0);
    EXPECT(std::memcmp(vz.data(), o.vz.data(), n * sizeof(double)) == 0);
    EXPECT(md_fcc_lattice(0, 0.8, 1.0, 1, L, x.data(), y.data(), z.data(), vx.data(), vy.data(),
                          vz.data()) != MD_OK);
    // tail terms and host arithmetic
    md_lj_params p = {1.0, 1.0, 2.5, 0};
    double u = 0, pr = 0;
    EXPECT(md_lj_tail(&p, n, L, &u, &pr) == MD_OK && u < 0.0);
    EXPECT(md_lj_tail(nullptr, n, L, &u, &pr) == MD_ERR_INVALID_ARGUMENT);
    EXPECT(md_lj_tail(&p, n, nullptr, &u, &pr) == MD_ERR_INVALID_ARGUMENT);
    EXPECT(md_cells_usable(L, 2.5) == 0);  // 6.72 sigma: two cells of 2.5
    const double L10[3] = {10.0, 10.0, 10.0};
    EXPECT(md_cells_usable(L10, 2.5) == 1);
    EXPECT(md_cells_usable(nullptr, 2.5) == 0);
    EXPECT(md_cells_usable(L, -1.0) == 0);
    EXPECT(md_neigh_auto_capacity(n, L, 2.5, 0.3) > 0);
    EXPECT(md_neigh_auto_capacity(1, L, 2.5, 0.3) >= 0);
    EXPECT(md_neigh_auto_capacity(-5, L, 2.5, 0.3) == 0);
    EXPECT(md_neigh_auto_capacity(n, nullptr, 2.5, 0.3) == 0);
    const double zero[3] = {0.0, 1.0, 1.0};
    EXPECT(md_neigh_auto_capacity(n, zero, 2.5, 0.3) == 0);
    // the replica recipe draw
    std::vector<int> bc(16);
    std::vector<double> brho(16), bT(16);
    std::vector<uint64_t> bseed(16);
    EXPECT(md_batch_draw(16, 4000, bc.data(), brho.data(), bT.data(), bseed.data()) == MD_OK);
    EXPECT(md_batch_draw(-1, 4000, bc.data(), brho.data(), bT.data(), bseed.data()) != MD_OK);
    // Ewald parameters
    double alpha = 0;
    int kmax = 0;
    const double Lw[3] = {49.71853, 49.71853, 49.71853};
    EXPECT(md_ewald_parameters(12288, Lw, 4096 * 1.077, 9.0, 1e-5, &alpha, &kmax) == MD_OK);
    EXPECT(alpha > 0.0 && kmax > 0);
    EXPECT(md_ewald_parameters(12288, Lw, 4096 * 1.077, -9.0, 1e-5, &alpha, &kmax) != MD_OK);
    // argument checks of handle-based entry points: no device is touched
    double e = 0, w = 0;
    EXPECT(md_lj_forces(nullptr, MD_PAIRS_BRUTE, MD_FP64, &e, &w) != MD_OK);
    EXPECT(md_run(nullptr, 10, MD_FP64, nullptr) != MD_OK);
    EXPECT(md_system_upload(nullptr, n, L, x.data(), y.data(), z.data(), nullptr, nullptr,
                            nullptr, nullptr) != MD_OK);
    EXPECT(md_neigh_export_pairs(nullptr, nullptr, 0, nullptr) != MD_OK);
    EXPECT(md_batch_run(nullptr, 10, 10, 0, nullptr) != MD_OK);
    EXPECT(md_water_forces(nullptr, MD_FP64, nullptr, nullptr) != MD_OK);
    EXPECT(md_water_settle(nullptr, x.data(), y.data(), z.data()) != MD_OK);
    EXPECT(md_water_settle_velocities(nullptr) != MD_OK);
    EXPECT(md_water_set_constraint(nullptr, MD_RIGID_SETTLE) != MD_OK);
    EXPECT(md_mix_forces(nullptr, nullptr, nullptr, nullptr) != MD_OK);
    EXPECT(md_rdf_accumulate(nullptr, nullptr, MD_RDF_AUTO) != MD_OK);
    EXPECT(md_msd_push(nullptr, nullptr) != MD_OK);
    EXPECT(md_vacf_result(nullptr, 1, nullptr) != MD_OK);
    EXPECT(md_rdf_create(-1.0, 10, nullptr) != MD_OK);
    EXPECT(std::strlen(md_last_error(nullptr)) > 0);
    // the allocation-free entry point refuses before any device call
    EXPECT(md_lj_forces_noalloc(MD_PAIRS_CELLS, &p, MD_NOALLOC_MAX_ATOMS + 1, L, x.data(),
                                y.data(), z.data(), nullptr, nullptr, nullptr, &e, &w,
                                nullptr) == MD_ERR_CAPACITY);
    EXPECT(md_lj_forces_noalloc(MD_PAIRS_VERLET, &p, 10, L, x.data(), y.data(), z.data(), nullptr,
                                nullptr, nullptr, &e, &w, nullptr) == MD_ERR_INVALID_ARGUMENT);
    EXPECT(md_lj_forces_noalloc(MD_PAIRS_BRUTE, nullptr, 10, L, x.data(), y.data(), z.data(),
                                nullptr, nullptr, nullptr, &e, &w, nullptr) ==
           MD_ERR_INVALID_ARGUMENT);
    EXPECT(std::strstr(md_last_error(nullptr), "") != nullptr);
}

int main() {
    oracle_passes();
    water_passes();
    library_host_paths();
    if (g_fail) {
        std::printf("%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("sanitize ok\n");
    return 0;
}
