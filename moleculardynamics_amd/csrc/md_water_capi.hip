// C ABI of the rigid SPC/E water path (include/moldyn_hip.h, md_water_*):
// the SpceForceField (forcefield/spce_forcefield.cpp:8-48) and Ewald
// (forcefield/ewald.cpp:17-188) forces, RigidWater SHAKE/RATTLE
// (constrain/rigid_water.cpp:30-161) and the constrained velocity Verlet
// (integrate/constrained_velocity_verlet.hpp:27-57) on the device, plus the
// config-3 input builder and the Ewald parameter estimate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/moldyn_hip.h"
#include "md_water.hpp"

namespace {
constexpr double kPi = 3.14159265358979323846;

// e^2/(4 pi eps0 kB) * 1e10 in K*A from the 2010 CODATA constants
// (forcefield/ewald.hpp:10-13, ewald.cpp:17-21).
double coulomb_prefactor() {
    const double e = 1.602176565e-19, eps0 = 8.854187817e-12, kb = 1.3806488e-23;
    return (e * e) / (4.0 * kPi * eps0) * 1e10 / kb;
}
}  // namespace

struct md_water {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    long n = 0;
    double L[3] = {0, 0, 0};
    md_water_params p{};
    double alpha = 0.0;
    mdk::WaterArgs a{};
    mdk::Grid grid{};
    // device buffers
    double4 *pos = nullptr, *vel = nullptr, *frc = nullptr, *ref = nullptr, *start = nullptr,
            *freep = nullptr;
    uint32_t* nbr = nullptr;
    int *ncnt = nullptr, *cell_of = nullptr, *cell_count = nullptr, *cell_start = nullptr,
        *cell_fill = nullptr, *cell_atoms = nullptr, *cell_tmp = nullptr, *status = nullptr,
        *flags = nullptr;
    double *part = nullptr, *part2 = nullptr, *kfac = nullptr;
    int3* kvec = nullptr;
    double2 *eik = nullptr, *skpart = nullptr, *sk = nullptr, *skf = nullptr;
    int4* kcol = nullptr;
    int* kcol0 = nullptr;
    int4* kseg = nullptr;
    int* kseg0 = nullptr;
    int nseg = 0;
    double4* fpart = nullptr;
    long builds = 0;
    bool list_built = false;
    double* eterm = nullptr;  // per-step device sums of md_water_run: 5 energy sums + KE
    long eterm_steps = 0;
    int constraint = MD_RIGID_SHAKE;  // constraint algorithm of md_water_run
    // reciprocal term: k-vector sums (MD_KSPACE_EWALD) or PME
    int kspace = MD_KSPACE_EWALD;
    int n2max = 0;
    long part2_alloc = 0;
    double* pme_q = nullptr;
    double2 *pme_A = nullptr, *pme_B = nullptr, *pme_P = nullptr, *pme_C = nullptr, *pme_D = nullptr,
            *pme_tw = nullptr;
    double *pme_phi = nullptr, *pme_G = nullptr, *pme_W = nullptr, *pme_DW = nullptr;
    int* pme_base = nullptr;
};

namespace {

md_status wfail(md_water* w, md_status s, const std::string& m) {
    if (w) w->err = m;
    return s;
}

#define WHIP(w, call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return wfail(w, MD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t walloc(T*& p, long count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (count <= 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(&p), static_cast<size_t>(count) * sizeof(T));
}

long wblocks(long n) { return (n + mdk::kBlock - 1) / mdk::kBlock; }

md_status build_list(md_water* w) {
    const mdk::Grid& g = w->grid;
    if (g.usable) {
        mdk::launch_bin_index(w->n, w->pos, g, w->cell_of, w->cell_count, w->cell_start,
                              w->cell_fill, w->cell_tmp, w->cell_atoms, nullptr, w->stream);
    }
    mdk::water_launch_list(w->a, g, w->cell_of, w->cell_start, w->cell_atoms, w->stream);
    WHIP(w, hipGetLastError());
    int ov = 0;
    WHIP(w, hipStreamSynchronize(w->stream));
    WHIP(w, hipMemcpy(&ov, w->flags, sizeof(int), hipMemcpyDeviceToHost));
    if (ov) return wfail(w, MD_ERR_CAPACITY, "water neighbour row capacity exceeded");
    w->builds += 1;
    w->list_built = true;
    return MD_OK;
}

// Forces on the current positions; terms = {disp, lrc, real, recip, self, intra}.
void terms_of(const md_water* w, const double* sums, double* terms);

md_status compute(md_water* w, bool mixed, double* terms) {
    if (!w->list_built) {
        md_status st = build_list(w);
        if (st != MD_OK) return st;
    }
    mdk::water_launch_forces(w->a, mixed, w->stream);
    mdk::water_launch_energy_sum(w->a, w->eterm, w->stream);
    WHIP(w, hipGetLastError());
    double sums[5];
    WHIP(w, hipMemcpyAsync(sums, w->eterm, sizeof(sums), hipMemcpyDeviceToHost, w->stream));
    WHIP(w, hipStreamSynchronize(w->stream));
    terms_of(w, sums, terms);
    return MD_OK;
}

// The six terms {disp, lrc, real, recip, self, intra} from the device sums
// {real, lj, intra, virial, recip} of one force evaluation.
void terms_of(const md_water* w, const double* sums, double* terms) {
    double real = sums[0], lj = sums[1], intra = sums[2];
    const double recip = sums[4];
    // every intermolecular and intramolecular pair was seen from both sites
    real *= 0.5;
    lj *= 0.5;
    intra *= 0.5;
    // self term and the oxygen LJ tail (lennard_jones.cpp:121-135 on the oxygens)
    const double C = coulomb_prefactor();
    const double qO = -2.0 * w->p.qH, qH = w->p.qH;
    const long nm = w->n / 3;
    const double q2 = static_cast<double>(nm) * (qO * qO + 2.0 * qH * qH);
    const double self = -C * w->alpha / std::sqrt(kPi) * q2;
    const double V = w->L[0] * w->L[1] * w->L[2];
    const double no = static_cast<double>(nm);
    const double rho = no / V;
    const double t = w->p.sigma_o / w->p.cutoff;
    const double t3 = t * t * t, t9 = t3 * t3 * t3;
    const double sig3 = w->p.sigma_o * w->p.sigma_o * w->p.sigma_o;
    const double lrc = (8.0 / 3.0) * kPi * no * rho * w->p.epsilon_o * sig3 * ((1.0 / 3.0) * t9 - t3);
    if (terms) {
        terms[0] = lj; terms[1] = lrc; terms[2] = real; terms[3] = recip; terms[4] = self; terms[5] = intra;
    }
}

// A SETTLE status: -2 for a molecule whose step has no closed-form solution.
md_status settle_check(md_water* w, const char* what) {
    const long nm = w->n / 3;
    std::vector<int> st(static_cast<size_t>(nm));
    WHIP(w, hipStreamSynchronize(w->stream));
    WHIP(w, hipMemcpy(st.data(), w->status, nm * sizeof(int), hipMemcpyDeviceToHost));
    for (long m = 0; m < nm; ++m)
        if (st[m] < 0)
            return wfail(w, MD_ERR_RUNTIME,
                         std::string(what) + ": molecule " + std::to_string(m) +
                             " has no rigid solution (degenerate, or rotated by 90 degrees or more in "
                             "one step); the timestep is far too large");
    return MD_OK;
}

md_status status_check(md_water* w, const char* what, int* iters) {
    const long nm = w->n / 3;
    std::vector<int> st(static_cast<size_t>(nm));
    WHIP(w, hipStreamSynchronize(w->stream));
    WHIP(w, hipMemcpy(st.data(), w->status, nm * sizeof(int), hipMemcpyDeviceToHost));
    int worst = 0;
    for (long m = 0; m < nm; ++m) {
        if (st[m] == -2) return wfail(w, MD_ERR_RUNTIME, std::string(what) + ": constraint became singular");
        if (st[m] < 0)
            return wfail(w, MD_ERR_RUNTIME, std::string(what) + " failed to converge in " +
                                                std::to_string(w->p.max_iterations) + " iterations");
        worst = std::max(worst, st[m]);
    }
    if (iters) *iters = worst;
    return MD_OK;
}

}  // namespace

extern "C" {

// Kolafa & Perram (Mol. Simul. 9, 351 (1992)) RMS force-error estimates:
//   real  dF_r = 2 Q2 exp(-a^2 rc^2) / sqrt(N rc V)
//   recip dF_k = 2 Q2 a / (pi L) * sqrt(1/(pi kmax N)) * exp(-(pi kmax/(a L))^2)
// with Q2 = sum q^2; both set to accuracy * C (the force between two unit
// charges 1 A apart): a from the real-space estimate, then the smallest kmax
// meeting the reciprocal one.
md_status md_ewald_parameters(long n, const double* L, double sum_q2, double cutoff,
                              double accuracy, double* alpha, int* kmax) {
    if (n <= 0 || L == nullptr || !(cutoff > 0.0) || !(accuracy > 0.0) || alpha == nullptr ||
        kmax == nullptr)
        return MD_ERR_INVALID_ARGUMENT;
    const double C = coulomb_prefactor();
    const double target = accuracy * C;
    const double V = L[0] * L[1] * L[2];
    const double Lmin = std::min({L[0], L[1], L[2]});
    const double scale = 2.0 * sum_q2 * C / std::sqrt(static_cast<double>(n) * cutoff * V);
    // solve scale * exp(-a^2 rc^2) = target for a
    double a = std::sqrt(std::max(1e-12, std::log(scale / target))) / cutoff;
    int km = 1;
    for (; km <= mdk::kMaxK; ++km) {
        const double est = 2.0 * sum_q2 * C * a / (kPi * Lmin) *
                           std::sqrt(1.0 / (kPi * km * static_cast<double>(n))) *
                           std::exp(-std::pow(kPi * km / (a * Lmin), 2.0));
        if (est <= target) break;
    }
    if (km > mdk::kMaxK) return MD_ERR_INVALID_ARGUMENT;
    *alpha = a;
    *kmax = km;
    return MD_OK;
}

// molecules_per_side^3 SPC/E molecules on a cubic lattice of spacing
// L/molecules_per_side (oxygen on the lattice point), each rotated by a
// uniform random rotation (Shoemake's quaternion from three uniforms of the
// mt19937_64 stream, uniform in (0,1] as util/random.hpp:43-48), hydrogens at
// r_OH = rOH with the H-O-H angle angleHOH; Maxwell-Boltzmann velocities at
// temperature (Marsaglia polar normals of the same stream), centre-of-mass
// motion removed. Sites are O, H, H per molecule; x..vz hold 3 m^3 values.
md_status md_water_lattice(int molecules_per_side, double box_length, uint64_t seed,
                           double temperature, const md_water_params* p, double* x, double* y,
                           double* z, double* vx, double* vy, double* vz) {
    if (molecules_per_side < 1 || !(box_length > 0.0) || p == nullptr || x == nullptr)
        return MD_ERR_INVALID_ARGUMENT;
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    auto uniform = [&]() { return 1.0 - unif(eng); };
    bool have = false;
    double spare = 0.0;
    auto normal = [&]() {
        if (have) {
            have = false;
            return spare;
        }
        double a, b, s;
        do {
            a = 2.0 * uniform() - 1.0;
            b = 2.0 * uniform() - 1.0;
            s = a * a + b * b;
        } while (s >= 1.0 || s == 0.0);
        const double k = std::sqrt(-2.0 * std::log(s) / s);
        spare = b * k;
        have = true;
        return a * k;
    };
    const int m = molecules_per_side;
    const double a0 = box_length / m;
    const double half = 0.5 * p->angle_hoh * kPi / 180.0;
    // body frame: O at the origin, H's in the xz plane
    const double hb[2][3] = {{p->r_oh * std::sin(half), 0.0, p->r_oh * std::cos(half)},
                             {-p->r_oh * std::sin(half), 0.0, p->r_oh * std::cos(half)}};
    long s = 0;
    for (int ix = 0; ix < m; ++ix)
        for (int iy = 0; iy < m; ++iy)
            for (int iz = 0; iz < m; ++iz) {
                const double o[3] = {a0 * (ix + 0.5) - 0.5 * box_length,
                                     a0 * (iy + 0.5) - 0.5 * box_length,
                                     a0 * (iz + 0.5) - 0.5 * box_length};
                const double u1 = uniform(), u2 = uniform(), u3 = uniform();
                const double qx = std::sqrt(1.0 - u1) * std::sin(2.0 * kPi * u2);
                const double qy = std::sqrt(1.0 - u1) * std::cos(2.0 * kPi * u2);
                const double qz = std::sqrt(u1) * std::sin(2.0 * kPi * u3);
                const double qw = std::sqrt(u1) * std::cos(2.0 * kPi * u3);
                const double R[3][3] = {
                    {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                    {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                    {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
                x[s] = o[0]; y[s] = o[1]; z[s] = o[2];
                for (int h = 0; h < 2; ++h) {
                    double r[3];
                    for (int d = 0; d < 3; ++d)
                        r[d] = o[d] + R[d][0] * hb[h][0] + R[d][1] * hb[h][1] + R[d][2] * hb[h][2];
                    x[s + 1 + h] = r[0]; y[s + 1 + h] = r[1]; z[s + 1 + h] = r[2];
                }
                s += 3;
            }
    const long n = s;
    double M = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    for (long i = 0; i < n; ++i) {
        const double mi = (i % 3 == 0) ? p->mass_o : p->mass_h;
        const double sv = std::sqrt(temperature / mi);
        const double g0 = normal(), g1 = normal(), g2 = normal();
        vx[i] = sv * g0; vy[i] = sv * g1; vz[i] = sv * g2;
        M += mi;
        px += mi * vx[i]; py += mi * vy[i]; pz += mi * vz[i];
    }
    for (long i = 0; i < n; ++i) {
        vx[i] -= px / M;
        vy[i] -= py / M;
        vz[i] -= pz / M;
    }
    return MD_OK;
}

md_status md_water_create(void* hip_stream, long n, const double* L, const double* x,
                          const double* y, const double* z, const double* vx, const double* vy,
                          const double* vz, const md_water_params* p, md_water** out) {
    if (out == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (p == nullptr || L == nullptr || x == nullptr || y == nullptr || z == nullptr)
        return MD_ERR_INVALID_ARGUMENT;
    md_water* w = new (std::nothrow) md_water();
    if (w == nullptr) return MD_ERR_RUNTIME;
    *out = w;
    if (n <= 0 || n % 3 != 0)
        return wfail(w, MD_ERR_INVALID_ARGUMENT, "atom count must be a positive multiple of 3 (O, H, H)");
    if (!(p->r_oh > 0.0)) return wfail(w, MD_ERR_INVALID_ARGUMENT, "RigidWater: ohDistance must be positive");
    if (!(p->angle_hoh > 0.0) || !(p->angle_hoh < 180.0))
        return wfail(w, MD_ERR_INVALID_ARGUMENT, "RigidWater: hohAngleDegrees must lie in (0, 180)");
    if (!(p->tolerance > 0.0)) return wfail(w, MD_ERR_INVALID_ARGUMENT, "RigidWater: tolerance must be positive");
    if (p->max_iterations < 1) return wfail(w, MD_ERR_INVALID_ARGUMENT, "RigidWater: maxIterations must be at least 1");
    if (p->kmax < 1 || p->kmax > mdk::kMaxK)
        return wfail(w, MD_ERR_INVALID_ARGUMENT, "kmax out of range (1..24)");
    if (hip_stream) {
        w->stream = static_cast<hipStream_t>(hip_stream);
    } else {
        WHIP(w, hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
        w->own_stream = true;
    }
    w->n = n;
    w->p = *p;
    for (int d = 0; d < 3; ++d) w->L[d] = L[d];
    const double minL = std::min({L[0], L[1], L[2]});
    w->alpha = p->alpha > 0.0 ? p->alpha : 5.6 / minL;
    const double rl = p->cutoff + p->skin;
    w->grid = mdk::Grid{};
    for (int d = 0; d < 3; ++d) {
        w->grid.L[d] = L[d];
        // cells of at least rl/2 (a 1e-9 relative margin keeps rounding from
        // putting a pair within rl three cells apart), searched 5 x 5 x 5
        w->grid.nc[d] = std::max(1, static_cast<int>(std::floor(2.0 * L[d] / (rl * (1.0 + 1e-9)))));
        w->grid.w[d] = L[d] / w->grid.nc[d];
    }
    w->grid.usable = (w->grid.nc[0] >= 5 && w->grid.nc[1] >= 5 && w->grid.nc[2] >= 5) ? 1 : 0;
    const long ncell = static_cast<long>(w->grid.nc[0]) * w->grid.nc[1] * w->grid.nc[2];
    const double rho = n / (L[0] * L[1] * L[2]);
    long cap = static_cast<long>(std::ceil(1.5 * 4.0 / 3.0 * kPi * rl * rl * rl * rho + 64.0));
    cap = std::min<long>((cap + 7) / 8 * 8, n - 1);
    // k-vectors of the half space with 0 < |n|^2 < n2max
    const int km = p->kmax;
    const int n2max = p->n2max > 0 ? p->n2max : km * km + 2;
    std::vector<int3> kv;
    std::vector<double> kf;
    const double V = L[0] * L[1] * L[2];
    const double C = coulomb_prefactor();
    const double k0[3] = {2.0 * kPi / L[0], 2.0 * kPi / L[1], 2.0 * kPi / L[2]};
    for (int nx = 0; nx <= km; ++nx)
        for (int ny = -km; ny <= km; ++ny)
            for (int nz = -km; nz <= km; ++nz) {
                const int n2 = nx * nx + ny * ny + nz * nz;
                if (n2 == 0 || n2 >= n2max) continue;
                if (nx == 0 && (ny < 0 || (ny == 0 && nz < 0))) continue;
                const double kx = k0[0] * nx, ky = k0[1] * ny, kz = k0[2] * nz;
                const double k2 = kx * kx + ky * ky + kz * kz;
                kv.push_back(make_int3(nx, ny, nz));
                kf.push_back(C * (2.0 * kPi / V) * std::exp(-k2 / (4.0 * w->alpha * w->alpha)) / k2);
            }
    const int nk = static_cast<int>(kv.size());
    const int groups = mdk::water_site_groups(n);
    // columns: runs of consecutive k-vectors with the same (nx, ny) and nz rising by one
    std::vector<int4> cols;
    std::vector<int> col0;
    for (int k = 0; k < nk; ++k) {
        if (k > 0 && kv[k].x == kv[k - 1].x && kv[k].y == kv[k - 1].y && kv[k].z == kv[k - 1].z + 1) {
            cols.back().w += 1;
        } else {
            cols.push_back(make_int4(kv[k].x, kv[k].y, kv[k].z, 1));
            col0.push_back(k);
        }
    }
    const int ncol = static_cast<int>(cols.size());
    const int nchunk = mdk::water_recip_chunks(n, std::max(1, ncol));
    // segments of at most 8 k-vectors per column (the S(k) kernel's unit)
    std::vector<int4> segs;
    std::vector<int> seg0;
    for (int c = 0; c < ncol; ++c)
        for (int m = 0; m < cols[c].w; m += 8) {
            segs.push_back(make_int4(cols[c].x, cols[c].y, cols[c].z + m, std::min(8, cols[c].w - m)));
            seg0.push_back(col0[c] + m);
        }
    w->nseg = static_cast<int>(segs.size());
    WHIP(w, walloc(w->pos, n)); WHIP(w, walloc(w->vel, n)); WHIP(w, walloc(w->frc, n));
    WHIP(w, walloc(w->ref, n)); WHIP(w, walloc(w->start, n)); WHIP(w, walloc(w->freep, n));
    const long npad = (n + 63) / 64 * 64;
    WHIP(w, walloc(w->nbr, cap * npad)); WHIP(w, walloc(w->ncnt, n));
    WHIP(w, walloc(w->cell_of, n)); WHIP(w, walloc(w->cell_count, ncell));
    WHIP(w, walloc(w->cell_start, ncell + 1)); WHIP(w, walloc(w->cell_fill, ncell));
    WHIP(w, walloc(w->cell_atoms, n)); WHIP(w, walloc(w->cell_tmp, n)); WHIP(w, walloc(w->status, n / 3)); WHIP(w, walloc(w->flags, 4));
    WHIP(w, walloc(w->part, 4 * mdk::water_real_blocks(n) + 8)); WHIP(w, walloc(w->part2, wblocks(std::max<long>(n, nk)) + 8));
    w->part2_alloc = wblocks(std::max<long>(n, nk)) + 8;
    w->n2max = n2max;
    WHIP(w, walloc(w->kvec, std::max(1, nk))); WHIP(w, walloc(w->kfac, std::max(1, nk)));
    WHIP(w, walloc(w->eik, 3L * (km + 1) * n));
    WHIP(w, walloc(w->skpart, static_cast<long>(groups) * std::max(1, nk)));
    WHIP(w, walloc(w->sk, std::max(1, nk)));
    WHIP(w, walloc(w->skf, std::max(1, nk)));
    WHIP(w, walloc(w->kcol, std::max(1, ncol)));
    WHIP(w, walloc(w->kcol0, std::max(1, ncol)));
    WHIP(w, walloc(w->kseg, std::max(1, w->nseg)));
    WHIP(w, walloc(w->kseg0, std::max(1, w->nseg)));
    WHIP(w, walloc(w->fpart, static_cast<long>(nchunk) * n));
    WHIP(w, hipMemset(w->flags, 0, 4 * sizeof(int)));
    WHIP(w, walloc(w->eterm, 6));
    w->eterm_steps = 1;
    if (nk > 0) {
        WHIP(w, hipMemcpy(w->kvec, kv.data(), nk * sizeof(int3), hipMemcpyHostToDevice));
        WHIP(w, hipMemcpy(w->kfac, kf.data(), nk * sizeof(double), hipMemcpyHostToDevice));
        WHIP(w, hipMemcpy(w->kcol, cols.data(), ncol * sizeof(int4), hipMemcpyHostToDevice));
        WHIP(w, hipMemcpy(w->kcol0, col0.data(), ncol * sizeof(int), hipMemcpyHostToDevice));
        WHIP(w, hipMemcpy(w->kseg, segs.data(), w->nseg * sizeof(int4), hipMemcpyHostToDevice));
        WHIP(w, hipMemcpy(w->kseg0, seg0.data(), w->nseg * sizeof(int), hipMemcpyHostToDevice));
    }
    std::vector<double4> hp(static_cast<size_t>(n)), hv(static_cast<size_t>(n)),
        hf(static_cast<size_t>(n), make_double4(0, 0, 0, 0));
    for (long i = 0; i < n; ++i) {
        const bool oxy = (i % 3) == 0;
        hp[i] = make_double4(x[i], y[i], z[i], oxy ? -2.0 * p->qH : p->qH);
        hv[i] = make_double4(vx ? vx[i] : 0.0, vy ? vy[i] : 0.0, vz ? vz[i] : 0.0,
                             oxy ? p->mass_o : p->mass_h);
    }
    WHIP(w, hipMemcpy(w->pos, hp.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    WHIP(w, hipMemcpy(w->vel, hv.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    WHIP(w, hipMemcpy(w->frc, hf.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    mdk::WaterArgs& a = w->a;
    a.n = n;
    a.box = mdk::make_box(L);
    a.pos = w->pos; a.vel = w->vel; a.frc = w->frc; a.ref = w->ref;
    a.nbr = w->nbr; a.ncnt = w->ncnt; a.cap = static_cast<int>(cap);
    a.rl2 = rl * rl;
    a.rc2 = p->cutoff * p->cutoff;
    a.skin = p->skin;
    a.alpha = w->alpha;
    a.c2a = 2.0 * w->alpha / std::sqrt(kPi);
    a.coulomb = C;
    a.sig2 = p->sigma_o * p->sigma_o;
    a.eps = p->epsilon_o;
    a.part = w->part; a.part2 = w->part2;
    a.kmax = km; a.nk = nk; a.kvec = w->kvec; a.kfac = w->kfac;
    for (int d = 0; d < 3; ++d) a.k0[d] = k0[d];
    a.eik = w->eik; a.skpart = w->skpart; a.sk = w->sk; a.skf = w->skf;
    a.kcol = w->kcol; a.kcol0 = w->kcol0; a.ncol = ncol; a.nchunk = nk > 0 ? nchunk : 0;
    a.kseg = w->kseg; a.kseg0 = w->kseg0; a.nseg = nk > 0 ? w->nseg : 0;
    a.fpart = w->fpart;
    a.nkpart = static_cast<int>(wblocks(nk));
    a.pme = mdk::PmeArgs{};
    a.overflow = w->flags;
    a.oh = p->r_oh;
    a.hh = 2.0 * p->r_oh * std::sin(0.5 * p->angle_hoh * kPi / 180.0);
    a.tol = p->tolerance;
    a.max_iter = p->max_iterations;
    a.settle_rc = 0.5 * a.hh;
    a.settle_h = std::sqrt(a.oh * a.oh - a.settle_rc * a.settle_rc);
    return MD_OK;
}

md_status md_water_destroy(md_water* w) {
    if (w == nullptr) return MD_OK;
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    for (void* q : {static_cast<void*>(w->pos), static_cast<void*>(w->vel), static_cast<void*>(w->frc),
                    static_cast<void*>(w->ref), static_cast<void*>(w->start), static_cast<void*>(w->freep),
                    static_cast<void*>(w->nbr), static_cast<void*>(w->ncnt), static_cast<void*>(w->cell_of),
                    static_cast<void*>(w->cell_count), static_cast<void*>(w->cell_start),
                    static_cast<void*>(w->cell_fill), static_cast<void*>(w->cell_atoms),
                    static_cast<void*>(w->cell_tmp),
                    static_cast<void*>(w->status), static_cast<void*>(w->flags), static_cast<void*>(w->part),
                    static_cast<void*>(w->part2), static_cast<void*>(w->kvec), static_cast<void*>(w->kfac),
                    static_cast<void*>(w->eik), static_cast<void*>(w->skpart), static_cast<void*>(w->sk),
                    static_cast<void*>(w->skf), static_cast<void*>(w->kcol), static_cast<void*>(w->kcol0),
                    static_cast<void*>(w->kseg), static_cast<void*>(w->kseg0),
                    static_cast<void*>(w->fpart), static_cast<void*>(w->eterm),
                    static_cast<void*>(w->pme_q), static_cast<void*>(w->pme_A), static_cast<void*>(w->pme_B),
                    static_cast<void*>(w->pme_P), static_cast<void*>(w->pme_C), static_cast<void*>(w->pme_D),
                    static_cast<void*>(w->pme_tw), static_cast<void*>(w->pme_phi), static_cast<void*>(w->pme_G),
                    static_cast<void*>(w->pme_W), static_cast<void*>(w->pme_DW),
                    static_cast<void*>(w->pme_base)})
        if (q) (void)hipFree(q);
    if (w->own_stream && w->stream) (void)hipStreamDestroy(w->stream);
    delete w;
    return MD_OK;
}

const char* md_water_last_error(const md_water* w) { return w ? w->err.c_str() : "null water"; }

md_status md_water_forces(md_water* w, int precision, double* terms, double* pe) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    double t[6];
    md_status st = compute(w, precision == MD_MIXED, t);
    if (st != MD_OK) return st;
    if (terms)
        for (int k = 0; k < 6; ++k) terms[k] = t[k];
    if (pe) *pe = t[0] + t[1] + t[2] + t[3] + t[4] + t[5];
    return MD_OK;
}

namespace {
// Cardinal B-spline M_p at the integer points 1 .. p-1 (M_p(t) = (t M_{p-1}(t)
// + (p - t) M_{p-1}(t - 1)) / (p - 1) from M_2 = 1 - |t - 1|), and the PME
// spline correction of one axis, |b(m)|^2 = 1 / |sum_{j=0}^{p-2} M_p(j+1)
// e^{2 pi i m j / K}|^2 (Essmann et al. 1995, eq. 4.4).
std::vector<double> bspline_integers(int p) {
    std::vector<double> c(static_cast<size_t>(p) + 1, 0.0);  // M_n(j), j = 0 .. p
    c[1] = 1.0;                                              // M_2(1) = 1
    for (int n = 3; n <= p; ++n) {
        std::vector<double> d(c.size(), 0.0);
        for (int j = 1; j < n; ++j) d[j] = (j * c[j] + (n - j) * c[j - 1]) / (n - 1);
        c = d;
    }
    return c;
}

double pme_bsq(const std::vector<double>& mp, int p, int K, int m) {
    double re = 0.0, im = 0.0;
    for (int j = 0; j <= p - 2; ++j) {
        long r = (static_cast<long>(m) * j) % K;
        if (r < 0) r += K;
        const double ang = 2.0 * kPi * static_cast<double>(r) / K;
        re += mp[j + 1] * std::cos(ang);
        im += mp[j + 1] * std::sin(ang);
    }
    return 1.0 / (re * re + im * im);
}
}  // namespace

md_status md_water_set_kspace(md_water* w, int method, int grid, int order) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    mdk::WaterArgs& a = w->a;
    if (method == MD_KSPACE_EWALD) {
        a.pme = mdk::PmeArgs{};
        a.nkpart = static_cast<int>(wblocks(a.nk));
        w->kspace = method;
        return MD_OK;
    }
    if (method != MD_KSPACE_PME) return wfail(w, MD_ERR_INVALID_ARGUMENT, "unknown k-space method");
    const int km = w->p.kmax;
    const int nb = 2 * km + 1;
    if (order == 0) order = 8;
    if (order < 3 || order > mdk::kPmeMaxOrder)
        return wfail(w, MD_ERR_INVALID_ARGUMENT, "PME spline order must be 3 .. 10");
    int K = grid;
    if (K == 0) K = std::max(nb, 2 * ((21 * km + 9) / 10));  // 4.2 kmax rounded up to even
    if (K < nb || K < order || K > 1024)
        return wfail(w, MD_ERR_INVALID_ARGUMENT,
                     "PME grid must hold max(2 kmax + 1, order) .. 1024 points per axis");
    mdk::PmeArgs pm{};
    pm.on = 1;
    pm.K[0] = pm.K[1] = pm.K[2] = K;
    pm.p = order;
    pm.kmax = km;
    pm.nb = nb;
    pm.ng = static_cast<long>(K) * K * K;
    // tables: twiddles per axis and the influence function on the k-set
    const std::vector<double> mp = bspline_integers(order);
    std::vector<double2> tw(3L * nb * K), twT(3L * nb * K);
    std::vector<double> bsq(3L * nb);
    for (int d = 0; d < 3; ++d)
        for (int c = 0; c < nb; ++c) {
            const long m = c - km;
            for (int k = 0; k < K; ++k) {
                long r = (m * k) % K;
                if (r < 0) r += K;
                const double ang = 2.0 * kPi * static_cast<double>(r) / K;
                tw[(static_cast<long>(d) * nb + c) * K + k] = make_double2(std::cos(ang), std::sin(ang));
                twT[(static_cast<long>(d) * K + k) * nb + c] = tw[(static_cast<long>(d) * nb + c) * K + k];
            }
            bsq[d * nb + c] = pme_bsq(mp, order, K, static_cast<int>(m));
        }
    std::vector<double> G(static_cast<size_t>(nb) * nb * nb, 0.0);
    const double V = w->L[0] * w->L[1] * w->L[2];
    const double C = coulomb_prefactor();
    const double k0[3] = {2.0 * kPi / w->L[0], 2.0 * kPi / w->L[1], 2.0 * kPi / w->L[2]};
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j)
            for (int l = 0; l < nb; ++l) {
                const int m0 = i - km, m1 = j - km, m2 = l - km;
                const int n2 = m0 * m0 + m1 * m1 + m2 * m2;
                if (n2 == 0 || n2 >= w->n2max) continue;
                const double kx = k0[0] * m0, ky = k0[1] * m1, kz = k0[2] * m2;
                const double k2 = kx * kx + ky * ky + kz * kz;
                G[(static_cast<size_t>(i) * nb + j) * nb + l] =
                    C * (2.0 * kPi / V) * std::exp(-k2 / (4.0 * w->alpha * w->alpha)) / k2 *
                    bsq[0 * nb + i] * bsq[1 * nb + j] * bsq[2 * nb + l];
            }
    WHIP(w, walloc(w->pme_q, pm.ng));
    WHIP(w, walloc(w->pme_A, static_cast<long>(K) * K * nb));
    WHIP(w, walloc(w->pme_B, static_cast<long>(K) * nb * nb));
    WHIP(w, walloc(w->pme_P, static_cast<long>(nb) * nb * nb));
    WHIP(w, walloc(w->pme_C, static_cast<long>(K) * nb * nb));
    WHIP(w, walloc(w->pme_D, static_cast<long>(K) * K * nb));
    WHIP(w, walloc(w->pme_phi, pm.ng));
    WHIP(w, walloc(w->pme_tw, 6L * nb * K));
    WHIP(w, walloc(w->pme_G, static_cast<long>(nb) * nb * nb));
    WHIP(w, walloc(w->pme_W, 3L * order * w->n));
    WHIP(w, walloc(w->pme_DW, 3L * order * w->n));
    WHIP(w, walloc(w->pme_base, 3L * w->n));
    WHIP(w, hipMemcpy(w->pme_tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
    WHIP(w, hipMemcpy(w->pme_tw + tw.size(), twT.data(), twT.size() * sizeof(double2),
                      hipMemcpyHostToDevice));
    WHIP(w, hipMemcpy(w->pme_G, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice));
    pm.q = w->pme_q;
    pm.A = w->pme_A;
    pm.B = w->pme_B;
    pm.P = w->pme_P;
    pm.Cz = w->pme_C;
    pm.Dy = w->pme_D;
    pm.phi = w->pme_phi;
    for (int d = 0; d < 3; ++d) {
        pm.tw[d] = w->pme_tw + static_cast<long>(d) * nb * K;
        pm.twT[d] = w->pme_tw + 3L * nb * K + static_cast<long>(d) * nb * K;
    }
    pm.G = w->pme_G;
    pm.W = w->pme_W;
    pm.DW = w->pme_DW;
    pm.base = w->pme_base;
    const long nkp = mdk::pme_energy_blocks(pm);
    if (nkp + 8 > w->part2_alloc) {
        WHIP(w, walloc(w->part2, nkp + 8));
        w->part2_alloc = nkp + 8;
        a.part2 = w->part2;
    }
    a.pme = pm;
    a.nkpart = static_cast<int>(nkp);
    w->kspace = method;
    return MD_OK;
}

md_status md_water_alpha(const md_water* w, double* alpha, int* nk) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (alpha) *alpha = w->alpha;
    if (nk) *nk = w->a.nk;
    return MD_OK;
}

md_status md_water_build_list(md_water* w) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    return build_list(w);
}

md_status md_water_shake(md_water* w, const double* rx, const double* ry, const double* rz,
                         int* iterations) {
    if (w == nullptr || rx == nullptr || ry == nullptr || rz == nullptr) return MD_ERR_INVALID_ARGUMENT;
    std::vector<double4> r(static_cast<size_t>(w->n));
    for (long i = 0; i < w->n; ++i) r[i] = make_double4(rx[i], ry[i], rz[i], 0.0);
    WHIP(w, hipMemcpy(w->start, r.data(), w->n * sizeof(double4), hipMemcpyHostToDevice));
    mdk::water_launch_shake(w->a, w->start, w->status, w->stream);
    WHIP(w, hipGetLastError());
    return status_check(w, "RigidWater: SHAKE", iterations);
}

md_status md_water_rattle(md_water* w, int* iterations) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    mdk::water_launch_rattle(w->a, w->status, w->stream);
    WHIP(w, hipGetLastError());
    return status_check(w, "RigidWater: RATTLE", iterations);
}

md_status md_water_set_constraint(md_water* w, int method) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (method != MD_RIGID_SHAKE && method != MD_RIGID_SETTLE)
        return wfail(w, MD_ERR_INVALID_ARGUMENT, "constraint must be MD_RIGID_SHAKE or MD_RIGID_SETTLE");
    w->constraint = method;
    return MD_OK;
}

md_status md_water_settle(md_water* w, const double* rx, const double* ry, const double* rz) {
    if (w == nullptr || rx == nullptr || ry == nullptr || rz == nullptr) return MD_ERR_INVALID_ARGUMENT;
    std::vector<double4> r(static_cast<size_t>(w->n));
    for (long i = 0; i < w->n; ++i) r[i] = make_double4(rx[i], ry[i], rz[i], 0.0);
    WHIP(w, hipMemcpy(w->start, r.data(), w->n * sizeof(double4), hipMemcpyHostToDevice));
    mdk::water_launch_settle(w->a, w->start, w->status, w->stream);
    WHIP(w, hipGetLastError());
    return settle_check(w, "SETTLE positions");
}

md_status md_water_settle_velocities(md_water* w) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    mdk::water_launch_settle_vel(w->a, w->status, w->stream);
    WHIP(w, hipGetLastError());
    return settle_check(w, "SETTLE velocities");
}

// nsteps of constrained velocity Verlet with the SPC/E force field; forces
// must be current on entry (md_water_forces once). The Verlet list is
// rebuilt when the two largest displacements exceed the skin.
md_status md_water_run(md_water* w, long nsteps, double dt, int precision, double* pe, double* ke) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (nsteps <= 0) return MD_OK;
    if (!w->list_built) {
        md_status st = build_list(w);
        if (st != MD_OK) return st;
    }
    if (nsteps > w->eterm_steps) {
        WHIP(w, walloc(w->eterm, 6 * nsteps));
        w->eterm_steps = nsteps;
    }
    // The whole run is enqueued without a host round trip: the rebuild is
    // decided and gated on the device (flags[1]), SHAKE/RATTLE failures are
    // collected by atomic minimum (flags[2], flags[3]) and the step's
    // energies and kinetic energy are summed on the device; all of it is
    // read once, after the last step. A failure is reported for the run.
    const bool mixed = precision == MD_MIXED;
    mdk::WaterArgs a = w->a;
    a.fail = w->flags + 2;
    a.build_flag = w->flags + 1;
    WHIP(w, hipMemsetAsync(w->flags + 2, 0, 2 * sizeof(int), w->stream));
    const mdk::Grid& g = w->grid;
    const bool settle = w->constraint == MD_RIGID_SETTLE;
    for (long k = 0; k < nsteps; ++k) {
        if (settle) {
            mdk::water_launch_cvv_settle_first(a, dt, w->status, w->stream);
        } else {
            mdk::water_launch_cvv_first(a, dt, w->start, w->freep, w->stream);
            mdk::water_launch_shake(a, w->start, w->status, w->stream);
            mdk::water_launch_cvv_correct(a, dt, w->freep, w->stream);
        }
        mdk::water_launch_check(a, w->flags + 1, w->stream);
        if (g.usable)
            mdk::launch_bin_index(w->n, w->pos, g, w->cell_of, w->cell_count, w->cell_start,
                                  w->cell_fill, w->cell_tmp, w->cell_atoms, w->flags + 1, w->stream);
        mdk::water_launch_list(a, g, w->cell_of, w->cell_start, w->cell_atoms, w->stream);
        mdk::water_launch_forces(a, mixed, w->stream);
        mdk::water_launch_energy_sum(a, w->eterm + 6 * k, w->stream);
        if (settle) {
            mdk::water_launch_cvv_settle_second(a, dt, w->status, w->stream);
        } else {
            mdk::water_launch_cvv_kick(a, dt, w->stream);
            mdk::water_launch_rattle(a, w->status, w->stream);
        }
        if (ke) {
            mdk::water_launch_kinetic(a, w->stream);
            mdk::water_launch_kinetic_sum(a, w->eterm + 6 * k + 5, w->stream);
        }
    }
    WHIP(w, hipGetLastError());
    int f[4] = {0, 0, 0, 0};
    WHIP(w, hipMemcpyAsync(f, w->flags, sizeof(f), hipMemcpyDeviceToHost, w->stream));
    std::vector<double> e;
    if (pe || ke) {
        e.resize(static_cast<size_t>(6 * nsteps));
        WHIP(w, hipMemcpyAsync(e.data(), w->eterm, e.size() * sizeof(double), hipMemcpyDeviceToHost,
                               w->stream));
    }
    WHIP(w, hipStreamSynchronize(w->stream));
    for (int c = 0; c < 2 && settle; ++c)
        if (f[2 + c] < 0)
            return wfail(w, MD_ERR_RUNTIME,
                         std::string(c == 0 ? "SETTLE positions" : "SETTLE velocities") +
                             ": a molecule has no rigid solution (rotated by 90 degrees or more in one "
                             "step); the timestep is far too large");
    for (int c = 0; c < 2; ++c) {
        const char* what = c == 0 ? "RigidWater: SHAKE" : "RigidWater: RATTLE";
        if (f[2 + c] == -2) return wfail(w, MD_ERR_RUNTIME, std::string(what) + ": constraint became singular");
        if (f[2 + c] < 0)
            return wfail(w, MD_ERR_RUNTIME, std::string(what) + " failed to converge in " +
                                                std::to_string(w->p.max_iterations) + " iterations");
    }
    if (f[0]) return wfail(w, MD_ERR_CAPACITY, "water neighbour row capacity exceeded");
    for (long k = 0; k < nsteps; ++k) {
        if (pe) {
            double t[6];
            terms_of(w, &e[6 * k], t);
            pe[k] = t[0] + t[1] + t[2] + t[3] + t[4] + t[5];
        }
        if (ke) ke[k] = e[6 * k + 5];
    }
    return MD_OK;
}

md_status md_water_download(md_water* w, double* x, double* y, double* z, double* vx, double* vy,
                            double* vz, double* fx, double* fy, double* fz) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    std::vector<double4> t(static_cast<size_t>(w->n));
    double* outs[3][3] = {{x, y, z}, {vx, vy, vz}, {fx, fy, fz}};
    double4* src[3] = {w->pos, w->vel, w->frc};
    WHIP(w, hipStreamSynchronize(w->stream));
    for (int k = 0; k < 3; ++k) {
        if (!outs[k][0]) continue;
        WHIP(w, hipMemcpy(t.data(), src[k], w->n * sizeof(double4), hipMemcpyDeviceToHost));
        for (long i = 0; i < w->n; ++i) {
            outs[k][0][i] = t[i].x;
            outs[k][1][i] = t[i].y;
            outs[k][2][i] = t[i].z;
        }
    }
    return MD_OK;
}

md_status md_water_set_state(md_water* w, const double* x, const double* y, const double* z,
                             const double* vx, const double* vy, const double* vz) {
    if (w == nullptr) return MD_ERR_INVALID_ARGUMENT;
    std::vector<double4> t(static_cast<size_t>(w->n));
    WHIP(w, hipStreamSynchronize(w->stream));
    if (x) {
        WHIP(w, hipMemcpy(t.data(), w->pos, w->n * sizeof(double4), hipMemcpyDeviceToHost));
        for (long i = 0; i < w->n; ++i) { t[i].x = x[i]; t[i].y = y[i]; t[i].z = z[i]; }
        WHIP(w, hipMemcpy(w->pos, t.data(), w->n * sizeof(double4), hipMemcpyHostToDevice));
        w->list_built = false;
    }
    if (vx) {
        WHIP(w, hipMemcpy(t.data(), w->vel, w->n * sizeof(double4), hipMemcpyDeviceToHost));
        for (long i = 0; i < w->n; ++i) { t[i].x = vx[i]; t[i].y = vy[i]; t[i].z = vz[i]; }
        WHIP(w, hipMemcpy(w->vel, t.data(), w->n * sizeof(double4), hipMemcpyHostToDevice));
    }
    return MD_OK;
}

}  // extern "C"
