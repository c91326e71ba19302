// moldyn_run_gpu: the reference driver's command line (apps/moldyn_run.cpp)
// on the MI355X path. The flags, their meaning and checks, the NIST table it
// prints, and the thermo / g(r) / MSD+VACF / XYZ files it writes (columns,
// provenance header, number format) are the reference's, so its experiment
// scripts read the output unchanged. Underneath, every step runs on the
// device through the C ABI (md_run: rebuild check, Verlet-list LJ forces,
// the integrator), and the analyses accumulate from the device copy
// (md_rdf_*, md_msd_*, md_vacf_*); a step copies nothing to the host unless
// a trajectory frame is written.
//
// One flag beyond the reference's closed list: --precision fp64|mixed
// (default fp64, the reference's arithmetic).
//
// Observable differences from the reference driver: Langevin noise comes
// from the counter-based stream of md_run (keyed by seed, step and atom),
// not the reference's mt19937_64 sequence, so stochastic trajectories differ
// draw for draw (velocity-Verlet and Nose-Hoover runs from the same lattice
// follow the reference's to rounding); the provenance's git_sha is this
// build's.
#include <sys/utsname.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/moldyn_hip.h"

#ifndef MOLDYN_GPU_GIT_SHA
#define MOLDYN_GPU_GIT_SHA "unknown"
#endif

namespace {

const char* kUsage =
    "Usage:\n"
    "  moldyn_run_gpu --nist-config PATH\n"
    "  moldyn_run_gpu --cells N --density RHO --temperature T --cutoff RC\n"
    "                 --skin S --dt DT --equilibrate STEPS --production STEPS\n"
    "                 --thermostat none|langevin|nose-hoover --seed S\n"
    "                 [--truncation truncated|linear-force-shift]\n"
    "                 [--rdf-out PATH] [--msd-out PATH] [--thermo-out PATH]\n"
    "                 [--traj-out PATH] [--precision fp64|mixed]\n"
    "\n"
    "  --nist-config PATH   directory holding the NIST LJ configuration files;\n"
    "                       prints the 12-case energy / virial / tail table,\n"
    "                       each value checked against half a unit in NIST's\n"
    "                       fifth published significant figure, and exits\n"
    "  --cells N            fcc cells per side (N_atoms = 4*cells^3)\n"
    "  --density RHO        reduced number density\n"
    "  --temperature T      target temperature\n"
    "  --cutoff RC          interaction cutoff\n"
    "  --skin S             Verlet skin\n"
    "  --dt DT              timestep\n"
    "  --equilibrate STEPS  steps of Langevin equilibration\n"
    "  --production STEPS   steps of NVE or NVT production\n"
    "  --thermostat KIND    none|langevin|nose-hoover (production phase)\n"
    "  --seed S             RNG seed\n"
    "  --truncation KIND    truncated (default; tail-corrected thermo) or\n"
    "                       linear-force-shift (continuous, no tail term)\n"
    "  --rdf-out PATH       write g(r) as CSV\n"
    "  --msd-out PATH       write MSD and VACF as CSV\n"
    "  --thermo-out PATH    write per-step energy, temperature, pressure as CSV\n"
    "  --traj-out PATH      write an XYZ trajectory\n"
    "  --precision KIND     fp64 (default) or mixed pair arithmetic\n";

[[noreturn]] void usage_exit(const std::string& msg) {
    std::cerr << "moldyn_run_gpu: " << msg << "\n\n" << kUsage;
    std::exit(1);
}

[[noreturn]] void device_exit(const std::string& what, md_status st, const md_ctx* ctx) {
    std::cerr << "moldyn_run_gpu: " << what << " failed (status " << static_cast<int>(st)
              << "): " << (ctx ? md_last_error(ctx) : "") << "\n";
    std::exit(2);
}

void check(md_status st, const md_ctx* ctx, const char* what) {
    if (st != MD_OK) device_exit(what, st, ctx);
}

const char* const kFlags[] = {"--nist-config", "--cells",       "--density",    "--temperature",
                              "--cutoff",      "--skin",        "--dt",         "--equilibrate",
                              "--production",  "--thermostat",  "--seed",       "--truncation",
                              "--rdf-out",     "--msd-out",     "--thermo-out", "--traj-out",
                              "--precision"};

using Args = std::map<std::string, std::string>;

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        const std::string t = argv[i];
        if (std::find(std::begin(kFlags), std::end(kFlags), t) == std::end(kFlags))
            usage_exit("unknown argument '" + t + "'");
        if (i + 1 >= argc) usage_exit("missing value for '" + t + "'");
        a[t] = argv[++i];
    }
    return a;
}

std::string need(const Args& a, const std::string& f) {
    const auto it = a.find(f);
    if (it == a.end()) usage_exit("missing required flag " + f);
    return it->second;
}

std::optional<std::string> opt(const Args& a, const std::string& f) {
    const auto it = a.find(f);
    if (it == a.end()) return std::nullopt;
    return it->second;
}

double num(const std::string& f, const std::string& v) {
    std::size_t used = 0;
    double x = 0.0;
    try {
        x = std::stod(v, &used);
    } catch (const std::exception&) {
        usage_exit("invalid number for " + f + ": '" + v + "'");
    }
    if (used != v.size()) usage_exit("invalid number for " + f + ": '" + v + "'");
    return x;
}

long integer(const std::string& f, const std::string& v) {
    std::size_t used = 0;
    long x = 0;
    try {
        x = std::stol(v, &used);
    } catch (const std::exception&) {
        usage_exit("invalid integer for " + f + ": '" + v + "'");
    }
    if (used != v.size()) usage_exit("invalid integer for " + f + ": '" + v + "'");
    return x;
}

uint64_t seed_of(const std::string& v) {
    if (v.empty() || v.find('-') != std::string::npos)
        usage_exit("invalid --seed '" + v + "' (must be a non-negative integer)");
    std::size_t used = 0;
    unsigned long long x = 0;
    try {
        x = std::stoull(v, &used);
    } catch (const std::exception&) {
        usage_exit("invalid --seed '" + v + "'");
    }
    if (used != v.size()) usage_exit("invalid --seed '" + v + "'");
    return static_cast<uint64_t>(x);
}

// ---- provenance ----------------------------------------------------------

std::string host() {
    struct utsname u {};
    if (uname(&u) == 0) return std::string(u.sysname) + " " + u.machine;
    return "unknown";
}

std::string command_line(int argc, char** argv) {
    std::string s;
    for (int i = 0; i < argc; ++i) s += (i ? " " : "") + std::string(argv[i]);
    return s;
}

std::ofstream open_csv(const std::string& path, const std::string& cmd) {
    std::ofstream f(path);
    if (!f.is_open()) {
        std::cerr << "moldyn_run_gpu: cannot open '" << path << "' for writing\n";
        std::exit(1);
    }
    f << "# command: " << cmd << "\n"
      << "# git_sha: " << MOLDYN_GPU_GIT_SHA << "\n"
      << "# host: " << host() << "\n"
      << "# units: reduced Lennard-Jones units (sigma = epsilon = mass = kB = 1)\n";
    f << std::setprecision(10);
    return f;
}

// ---- NIST table ----------------------------------------------------------

// NIST Standard Reference Simulation Website, LJ fluid, cuboid cell:
// configuration, truncation, cutoff, U, W, U_LRC (5 significant figures;
// the same 12 cases the reference checks).
struct NistCase {
    int cfg;
    int lfs;
    double rc, u, w, ulrc;
};
const NistCase kNist[] = {
    {1, 0, 3.0, -4.3515e3, -5.6867e2, -1.9849e2}, {1, 0, 4.0, -4.4675e3, -1.2639e3, -8.3769e1},
    {1, 1, 3.0, -3.8709e3, 3.1754e2, 0.0},        {2, 0, 3.0, -6.9000e2, -5.6846e2, -2.4230e1},
    {2, 0, 4.0, -7.0460e2, -6.5599e2, -1.0226e1}, {2, 1, 3.0, -6.2012e2, -4.4533e2, 0.0},
    {3, 0, 3.0, -1.1467e3, -1.1649e3, -4.9622e1}, {3, 0, 4.0, -1.1754e3, -1.3371e3, -2.0942e1},
    {3, 1, 3.0, -1.0210e3, -9.3578e2, 0.0},       {4, 0, 3.0, -1.6790e1, -4.6249e1, -5.4517e-1},
    {4, 0, 4.0, -1.7060e1, -4.7869e1, -2.3008e-1}, {4, 1, 3.0, -1.5001e1, -4.3096e1, 0.0},
};

// Half a unit in the fifth significant figure of a published value.
double half_width(double v) {
    const double m = std::fabs(v);
    return m == 0.0 ? 0.0 : 0.5 * std::pow(10.0, std::floor(std::log10(m)) - 4.0);
}

// NIST SRS configuration: box line, count line, "index x y z [type]" lines
// read to the end of the file (the count line is not trusted).
bool read_nist(const std::string& path, double L[3], std::vector<double>& x,
               std::vector<double>& y, std::vector<double>& z) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    if (!std::getline(in, line)) return false;
    std::istringstream(line) >> L[0] >> L[1] >> L[2];
    if (!std::getline(in, line)) return false;
    while (std::getline(in, line)) {
        std::istringstream is(line);
        long idx;
        double a, b, c;
        if (!(is >> idx >> a >> b >> c)) continue;
        x.push_back(a);
        y.push_back(b);
        z.push_back(c);
    }
    return !x.empty();
}

int nist_table(const std::string& dir) {
    md_ctx* ctx = nullptr;
    check(md_ctx_create(nullptr, &ctx), nullptr, "md_ctx_create");
    bool all = true;
    double worst = 0.0;
    std::cout << std::left << std::setw(4) << "cfg" << std::setw(5) << "trunc" << std::right
              << std::setw(5) << "rc" << std::setw(14) << "u_calc" << std::setw(14) << "u_ref"
              << std::setw(12) << "u_dev" << std::setw(12) << "u_tol" << std::setw(14) << "w_calc"
              << std::setw(14) << "w_ref" << std::setw(12) << "w_dev" << std::setw(12) << "w_tol"
              << std::setw(14) << "ulrc_calc" << std::setw(14) << "ulrc_ref" << std::setw(12)
              << "ulrc_dev" << std::setw(12) << "ulrc_tol" << "  status\n";
    std::cout << std::scientific << std::setprecision(4);
    for (const NistCase& c : kNist) {
        const std::string path = dir + "/lj_sample_config_periodic" + std::to_string(c.cfg) + ".txt";
        double L[3];
        std::vector<double> x, y, z;
        if (!read_nist(path, L, x, y, z)) {
            std::cerr << "moldyn_run_gpu: failed to read '" << path << "'\n";
            md_ctx_destroy(ctx);
            return 1;
        }
        const long n = static_cast<long>(x.size());
        check(md_system_upload(ctx, n, L, x.data(), y.data(), z.data(), nullptr, nullptr, nullptr,
                               nullptr),
              ctx, "md_system_upload");
        const md_lj_params p = {1.0, 1.0, c.rc, c.lfs};
        check(md_lj_set(ctx, &p), ctx, "md_lj_set");
        double e = 0.0, w = 0.0, lrc = 0.0;
        check(md_lj_energy_virial(ctx, &e, &w), ctx, "md_lj_energy_virial");
        check(md_lj_tail(&p, n, L, &lrc, nullptr), ctx, "md_lj_tail");
        const double u_dev = std::fabs(e - c.u) / std::fabs(c.u);
        const double w_dev = std::fabs(w - c.w) / std::fabs(c.w);
        const double u_tol = half_width(c.u) / std::fabs(c.u);
        const double w_tol = half_width(c.w) / std::fabs(c.w);
        // an LFS tail is a structural zero, not a rounded value: absolute bound
        const bool zero = c.ulrc == 0.0;
        const double l_dev = zero ? std::fabs(lrc) : std::fabs(lrc - c.ulrc) / std::fabs(c.ulrc);
        const double l_tol = zero ? 1e-9 : half_width(c.ulrc) / std::fabs(c.ulrc);
        const bool pass = u_dev <= u_tol && w_dev <= w_tol && l_dev <= l_tol;
        all = all && pass;
        worst = std::max({worst, u_dev / u_tol, w_dev / w_tol, l_dev / l_tol});
        std::cout << std::left << std::setw(4) << c.cfg << std::setw(5) << (c.lfs ? "LFS" : "LRC")
                  << std::right << std::setw(5) << c.rc << std::setw(14) << e << std::setw(14)
                  << c.u << std::setw(12) << u_dev << std::setw(12) << u_tol << std::setw(14) << w
                  << std::setw(14) << c.w << std::setw(12) << w_dev << std::setw(12) << w_tol
                  << std::setw(14) << lrc << std::setw(14) << c.ulrc << std::setw(12) << l_dev
                  << std::setw(12) << l_tol << "  " << (pass ? "PASS" : "FAIL") << "\n";
    }
    md_ctx_destroy(ctx);
    std::cout << "\nworst case: " << worst << " of the per-value bound\n";
    std::cout << (all ? "ALL 12 CASES MATCH EVERY DIGIT NIST PUBLISHES (within each value's "
                        "5-significant-figure rounding half-width): PASS\n"
                      : "ONE OR MORE CASES FELL OUTSIDE NIST'S OWN 5-SIGNIFICANT-FIGURE "
                        "ROUNDING HALF-WIDTH: FAIL\n");
    return all ? 0 : 1;
}

// ---- simulation ------------------------------------------------------------

enum Thermostat { kNone, kLangevin, kNoseHoover };

struct Run {
    int cells;
    double rho, T, rc, skin, dt;
    long equil, prod;
    Thermostat th;
    uint64_t seed;
    int lfs = 0;
    int precision = MD_FP64;
    std::optional<std::string> rdf_out, msd_out, thermo_out, traj_out;
};

Run parse_run(const Args& a) {
    Run r{};
    const long cells = integer("--cells", need(a, "--cells"));
    if (cells < 1) usage_exit("--cells must be a positive integer");
    r.cells = static_cast<int>(cells);
    r.rho = num("--density", need(a, "--density"));
    if (r.rho <= 0.0) usage_exit("--density must be positive");
    r.T = num("--temperature", need(a, "--temperature"));
    if (r.T < 0.0) usage_exit("--temperature must be non-negative");
    r.rc = num("--cutoff", need(a, "--cutoff"));
    if (r.rc <= 0.0) usage_exit("--cutoff must be positive");
    r.skin = num("--skin", need(a, "--skin"));
    if (r.skin < 0.0) usage_exit("--skin must be non-negative");
    r.dt = num("--dt", need(a, "--dt"));
    if (r.dt <= 0.0) usage_exit("--dt must be positive");
    r.equil = integer("--equilibrate", need(a, "--equilibrate"));
    if (r.equil < 0) usage_exit("--equilibrate must be non-negative");
    r.prod = integer("--production", need(a, "--production"));
    if (r.prod < 0) usage_exit("--production must be non-negative");
    const std::string th = need(a, "--thermostat");
    if (th == "none")
        r.th = kNone;
    else if (th == "langevin")
        r.th = kLangevin;
    else if (th == "nose-hoover")
        r.th = kNoseHoover;
    else
        usage_exit("--thermostat must be one of none|langevin|nose-hoover, got '" + th + "'");
    r.seed = seed_of(need(a, "--seed"));
    if (const auto t = opt(a, "--truncation")) {
        if (*t == "truncated")
            r.lfs = 0;
        else if (*t == "linear-force-shift")
            r.lfs = 1;
        else
            usage_exit("--truncation must be one of truncated|linear-force-shift, got '" + *t + "'");
    }
    if (const auto p = opt(a, "--precision")) {
        if (*p == "fp64")
            r.precision = MD_FP64;
        else if (*p == "mixed")
            r.precision = MD_MIXED;
        else
            usage_exit("--precision must be one of fp64|mixed, got '" + *p + "'");
    }
    r.rdf_out = opt(a, "--rdf-out");
    r.msd_out = opt(a, "--msd-out");
    r.thermo_out = opt(a, "--thermo-out");
    r.traj_out = opt(a, "--traj-out");
    return r;
}

// Thermostat constants of the reference driver (not flags there either).
constexpr double kFriction = 1.0;
constexpr double kTau = 0.5;
constexpr int kChain = 3;
constexpr long kRdfBins = 200;
constexpr double kRdfFraction = 0.98;  // of half the smallest box length
constexpr long kMaxLag = 2000;

struct Driver {
    const Run r;
    std::string cmd;
    md_ctx* ctx = nullptr;
    long n = 0;
    double L[3] = {0, 0, 0};
    md_lj_params lj{};
    double u_lrc = 0.0, p_lrc = 0.0;
    long dof = 0;
    std::optional<std::ofstream> thermo, traj;
    md_rdf* rdf = nullptr;
    md_msd* msd = nullptr;
    md_vacf* vacf = nullptr;
    long max_lag = 0;
    std::vector<double> hx, hy, hz;  // trajectory download buffers

    Driver(Run run, std::string c) : r(std::move(run)), cmd(std::move(c)) {}

    // One thermo row (apps/moldyn_run.cpp:497-510): PE and P include the
    // tail terms (zero under linear-force-shift).
    void row(long step, const char* phase, double e, double w, double ke) {
        if (!thermo) return;
        const double V = L[0] * L[1] * L[2];
        const double t = 2.0 * ke / static_cast<double>(dof);
        const double P = (static_cast<double>(n) / V) * t + w / (3.0 * V) + p_lrc;
        const double pe = e + u_lrc;
        *thermo << step << ',' << static_cast<double>(step) * r.dt << ',' << phase << ',' << pe
                << ',' << ke << ',' << pe + ke << ',' << t << ',' << P << '\n';
    }

    void frame(long step, const char* phase) {
        check(md_system_download(ctx, hx.data(), hy.data(), hz.data(), nullptr, nullptr, nullptr,
                                 nullptr, nullptr, nullptr),
              ctx, "md_system_download");
        std::ofstream& f = *traj;
        f << n << '\n'
          << "step=" << step << " time=" << static_cast<double>(step) * r.dt << " phase=" << phase
          << " git_sha=" << MOLDYN_GPU_GIT_SHA << " host=\"" << host() << "\" command=\"" << cmd
          << "\"" << '\n';
        for (long i = 0; i < n; ++i) f << "Ar " << hx[i] << ' ' << hy[i] << ' ' << hz[i] << '\n';
    }

    void set_integrator(int kind, uint64_t seed) {
        md_integrator it{};
        it.kind = kind;
        it.dt = r.dt;
        it.friction = kFriction;
        it.temperature = r.T;
        it.tau = kTau;
        it.chain_length = kChain;
        it.dof = 3 * n - 3;
        it.seed = seed;
        check(md_integrator_set(ctx, &it), ctx, "md_integrator_set");
    }

    // nsteps of the current integrator, a thermo row each; with `sample`,
    // the analyses and trajectory frame after every step.
    void steps(long nsteps, long& step, const char* phase, bool sample) {
        const bool per_step = sample && (rdf || msd || traj);
        const long chunk = per_step ? 1 : 4096;
        std::vector<md_thermo> th(static_cast<size_t>(std::min(nsteps, chunk)));
        for (long done = 0; done < nsteps;) {
            const long k = std::min(chunk, nsteps - done);
            check(md_run(ctx, k, r.precision, th.data()), ctx, "md_run");
            for (long j = 0; j < k; ++j) {
                ++step;
                row(step, phase, th[j].potential, th[j].virial, th[j].kinetic);
            }
            if (per_step) {
                if (rdf && md_rdf_accumulate(rdf, ctx, MD_RDF_AUTO) != MD_OK) {
                    std::cerr << "moldyn_run_gpu: " << md_rdf_last_error(rdf) << "\n";
                    std::exit(2);
                }
                if (msd && md_msd_push(msd, ctx) != MD_OK) {
                    std::cerr << "moldyn_run_gpu: " << md_msd_last_error(msd) << "\n";
                    std::exit(2);
                }
                if (vacf && md_vacf_push(vacf, ctx) != MD_OK) {
                    std::cerr << "moldyn_run_gpu: " << md_vacf_last_error(vacf) << "\n";
                    std::exit(2);
                }
                if (traj) frame(step, phase);
            }
            done += k;
        }
    }

    int run() {
        n = 4L * r.cells * r.cells * r.cells;
        std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n);
        check(md_fcc_lattice(r.cells, r.rho, r.T, r.seed, L, x.data(), y.data(), z.data(),
                             vx.data(), vy.data(), vz.data()),
              nullptr, "md_fcc_lattice");
        check(md_ctx_create(nullptr, &ctx), nullptr, "md_ctx_create");
        check(md_system_upload(ctx, n, L, x.data(), y.data(), z.data(), vx.data(), vy.data(),
                               vz.data(), nullptr),
              ctx, "md_system_upload");
        lj = {1.0, 1.0, r.rc, r.lfs};
        check(md_lj_set(ctx, &lj), ctx, "md_lj_set");
        check(md_lj_tail(&lj, n, L, &u_lrc, &p_lrc), ctx, "md_lj_tail");
        check(md_neigh_config(ctx, r.rc, r.skin), ctx, "md_neigh_config");
        check(md_neigh_build(ctx), ctx, "md_neigh_build");
        double e = 0.0, w = 0.0, ke = 0.0;
        check(md_lj_forces(ctx, MD_PAIRS_VERLET, r.precision, &e, &w), ctx, "md_lj_forces");
        check(md_system_kinetic_energy(ctx, &ke), ctx, "md_system_kinetic_energy");
        const long dof_lv = 3 * n, dof_nve = 3 * n - 3;
        if (r.thermo_out) {
            thermo = open_csv(*r.thermo_out, cmd);
            const bool tail = !r.lfs;
            *thermo << "# truncation: " << (tail ? "truncated" : "linear-force-shift") << "\n"
                    << "# pressure_includes_tail_correction: " << (tail ? "yes" : "no") << "\n"
                    << "# potential_energy_includes_tail_correction: " << (tail ? "yes" : "no")
                    << "\n"
                    << "# temperature_dof: 3N=" << dof_lv
                    << " while Langevin drives (equilibration, and Langevin production), 3N-3="
                    << dof_nve
                    << " under the momentum-conserving integrators (NVE, Nose-Hoover), for which "
                       "the centre-of-mass momentum is re-zeroed at the equilibration/production "
                       "boundary\n"
                    << "step,time,phase,potential_energy,kinetic_energy,total_energy,temperature,"
                       "pressure\n";
        }
        if (r.traj_out) {
            traj.emplace(*r.traj_out);
            if (!traj->is_open()) {
                std::cerr << "moldyn_run_gpu: cannot open '" << *r.traj_out << "' for writing\n";
                return 1;
            }
            hx.resize(n);
            hy.resize(n);
            hz.resize(n);
        }
        if (r.rdf_out) {
            const double rmax = kRdfFraction * 0.5 * std::min({L[0], L[1], L[2]});
            check(md_rdf_create(rmax, kRdfBins, &rdf), ctx, "md_rdf_create");
        }
        if (r.msd_out && r.prod > 0) {
            max_lag = std::max(1L, std::min(r.prod / 2, kMaxLag));
            check(md_msd_create(n, max_lag, &msd), ctx, "md_msd_create");
            check(md_vacf_create(n, max_lag, &vacf), ctx, "md_vacf_create");
        }
        long step = 0;
        dof = dof_nve;  // the lattice starts at zero momentum
        row(0, "start", e, w, ke);
        dof = dof_lv;
        if (r.equil > 0) {
            set_integrator(MD_LANGEVIN_BAOAB, r.seed);
            steps(r.equil, step, "equilibrate", false);
        }
        if (r.prod > 0) {
            check(md_system_remove_com_motion(ctx), ctx, "md_system_remove_com_motion");
            if (r.th == kNone) {
                dof = dof_nve;
                set_integrator(MD_VELOCITY_VERLET, 0);
            } else if (r.th == kLangevin) {
                dof = dof_lv;
                set_integrator(MD_LANGEVIN_BAOAB, r.seed + 1);
            } else {
                dof = dof_nve;
                set_integrator(MD_NOSE_HOOVER, 0);
            }
            steps(r.prod, step, "production", true);
        }
        if (rdf) {
            std::ofstream f = open_csv(*r.rdf_out, cmd);
            std::vector<double> c(kRdfBins), g(kRdfBins);
            long frames = 0;
            check(md_rdf_result(rdf, c.data(), g.data(), nullptr, &frames), ctx, "md_rdf_result");
            f << "# frames: " << frames << "\n" << "r,g\n";
            for (long k = 0; k < kRdfBins; ++k) f << c[k] << ',' << g[k] << '\n';
        }
        if (msd && vacf) {
            std::ofstream f = open_csv(*r.msd_out, cmd);
            // Einstein fit over the middle half of the lags
            long from = max_lag / 4, to = max_lag - max_lag / 4 + 1;
            if (to <= from + 1) {
                from = 0;
                to = max_lag + 1;
            }
            double d_msd = 0.0, d_vacf = 0.0;
            check(md_msd_diffusion(msd, r.dt, from, to, &d_msd), ctx, "md_msd_diffusion");
            check(md_vacf_diffusion(vacf, r.dt, &d_vacf), ctx, "md_vacf_diffusion");
            std::vector<double> m(max_lag + 1), c(max_lag + 1);
            check(md_msd_result(msd, m.data()), ctx, "md_msd_result");
            check(md_vacf_result(vacf, 1, c.data()), ctx, "md_vacf_result");
            f << "# D_msd (Einstein, fit lags [" << from << "," << to << ")): " << d_msd << "\n"
              << "# D_vacf (Green-Kubo): " << d_vacf << "\n"
              << "lag,time,msd,vacf\n";
            for (long k = 0; k <= max_lag; ++k)
                f << k << ',' << static_cast<double>(k) * r.dt << ',' << m[k] << ',' << c[k] << '\n';
        }
        md_rdf_destroy(rdf);
        md_msd_destroy(msd);
        md_vacf_destroy(vacf);
        md_ctx_destroy(ctx);
        return 0;
    }
};

}  // namespace

int main(int argc, char** argv) {
    const std::string cmd = command_line(argc, argv);
    const Args a = parse(argc, argv);
    if (const auto dir = opt(a, "--nist-config")) return nist_table(*dir);
    Driver d(parse_run(a), cmd);
    return d.run();
}
