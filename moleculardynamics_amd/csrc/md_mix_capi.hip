// C ABI of the differentiable per-type-pair LJ path (include/moldyn_hip.h,
// md_mix_*): forces, energy and virial of a multi-type LJ system with their
// parameter derivatives, the force vector-Jacobian product, a steepest-descent
// relaxation, velocity-Verlet trajectories with checkpoints and the adjoint
// of such a trajectory. The reference has no counterpart (its LJ is
// single-type, forcefield/lennard_jones.hpp:26); the pair terms, the Verlet
// list with skin (neighbor/verlet_list.cpp:8-73) and the velocity-Verlet step
// (integrate/velocity_verlet.hpp:48-74) follow the reference.
//
// Adjoint of velocity Verlet (kick a = (dt/2)/m, F = F(x; theta)):
//   u_t = v_t + a F(x_t),  x_{t+1} = x_t + dt u_t,  v_{t+1} = u_t + a F(x_{t+1})
// for L = sum_t [wE E(x_t) + wW W(x_t)] + gX . x_T + gV . v_T, sweeping frames
// t = T .. 0 with lx = gX, lv = gV, lu = 0 at t = T:
//   frame t:  lx += J(x_t)^T [a (lv + lu)] + wE grad E + wW grad W   (frame 0: a lu)
//             ltheta += (dF/dtheta)^T [same vector] + wE dE/dtheta + wW dW/dtheta
//   step t-1: lu = lv + dt lx;  lv = lu
// ending with dL/dx_0 = lx, dL/dv_0 = lv. Frames are recomputed segment by
// segment from checkpoints taken every K steps; at a checkpoint the forward
// run rebuilds the list and re-evaluates the forces, so the recomputation
// reproduces the forward trajectory bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/moldyn_hip.h"
#include "md_mix.hpp"

struct md_mix {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    long n = 0;
    double L[3] = {0, 0, 0};
    mdk::Box box{};
    mdk::Grid grid{};
    // parameters
    bool have_params = false;
    int nt = 0;
    int lfs = 0;
    double sigma[mdk::kTypePairs] = {}, epsilon[mdk::kTypePairs] = {}, cutoff[mdk::kTypePairs] = {};
    mdk::MixTab tab{};
    double skin = 0.3;
    double rl = 0.0;
    // device state
    double4 *pos = nullptr, *vel = nullptr, *frc = nullptr, *ref = nullptr;
    uint32_t* nbr = nullptr;
    int* ncnt = nullptr;
    int cap = 0;
    long nbr_alloc = 0;
    int *cell_of = nullptr, *cell_count = nullptr, *cell_start = nullptr, *cell_fill = nullptr,
        *cell_atoms = nullptr, *cell_tmp = nullptr;
    long ncell_alloc = 0;
    double *part_ew = nullptr, *part_q = nullptr, *part_d = nullptr, *gamma = nullptr,
           *hist = nullptr, *tables = nullptr, *ew = nullptr;
    long hist_alloc = 0;
    mdk::Scalars* sc = nullptr;
    int* flag_host = nullptr;  // pinned
    // trajectory / adjoint
    double4 *ckpt = nullptr, *frames = nullptr, *lx = nullptr, *lv = nullptr, *lu = nullptr,
            *vec = nullptr, *fin = nullptr;
    long ckpt_alloc = 0, frames_alloc = 0;
    int T = 0, K = 0;
    double dt = 0.0;
    bool have_run = false;
    bool list_ok = false;
    long builds = 0;
};

namespace {

constexpr double kPi = 3.14159265358979323846;

md_status mfail(md_mix* m, md_status s, const std::string& msg) {
    if (m) m->err = msg;
    return s;
}

#define MHIP(m, call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return mfail(m, MD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define MTRY(call)                    \
    do {                              \
        md_status s_ = (call);        \
        if (s_ != MD_OK) return s_;   \
    } while (0)

template <class T>
hipError_t malloc_dev(T*& p, long count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (count <= 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(&p), static_cast<size_t>(count) * sizeof(T));
}

template <class T>
void free_dev(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

long nblocks(long n) { return (n + mdk::kBlock - 1) / mdk::kBlock; }
int npairs(int nt) { return nt * (nt + 1) / 2; }

// tables[(a*nt+b)*q + s] (ordered pairs, lower index of type a) -> canonical
// unordered pairs k over (a <= b) row-major.
void canonical(const std::vector<double>& t, int nt, int q, double* out) {
    int k = 0;
    for (int a = 0; a < nt; ++a)
        for (int b = a; b < nt; ++b, ++k)
            for (int s = 0; s < q; ++s) {
                double v = t[static_cast<size_t>((a * nt + b) * q + s)];
                if (a != b) v += t[static_cast<size_t>((b * nt + a) * q + s)];
                out[k * q + s] = v;
            }
}

md_status need_system(md_mix* m) {
    if (m == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (m->n <= 0) return mfail(m, MD_ERR_STATE, "no system uploaded");
    if (!m->have_params) return mfail(m, MD_ERR_STATE, "no pair parameters set");
    return MD_OK;
}

// u(rc), u'(rc) = -w(rc)/rc as lennard_jones.cpp:32-48 computes them; y(rc)/rc
// for the sigma-derivatives of the shifted terms.
void make_tab(md_mix* m) {
    mdk::MixTab& t = m->tab;
    t = mdk::MixTab{};
    t.nt = m->nt;
    t.lfs = m->lfs;
    for (int k = 0; k < m->nt * m->nt; ++k) {
        const double sg = m->sigma[k], ep = m->epsilon[k], rc = m->cutoff[k];
        const double s2 = (sg * sg) / (rc * rc);
        const double s6 = s2 * s2 * s2;
        const double u_rc = 4.0 * ep * (s6 * s6 - s6);
        const double w_rc = 24.0 * ep * (2.0 * s6 * s6 - s6);
        const double y_rc = 24.0 * ep * (24.0 * s6 * s6 - 6.0 * s6);
        t.sig2[k] = sg * sg;
        t.eps[k] = ep;
        t.rc[k] = rc;
        t.rc2[k] = rc * rc;
        t.inv_eps[k] = 1.0 / ep;
        t.inv_sig[k] = 1.0 / sg;
        t.urc[k] = u_rc;
        t.dudr[k] = -w_rc / rc;
        t.wrc[k] = w_rc;
        t.yrc_rc[k] = y_rc / rc;
    }
}

md_status alloc_list(md_mix* m) {
    double rcmax = 0.0;
    for (int k = 0; k < m->nt * m->nt; ++k) rcmax = std::max(rcmax, m->cutoff[k]);
    m->rl = rcmax + m->skin;
    mdk::Grid& g = m->grid;
    g = mdk::Grid{};
    for (int d = 0; d < 3; ++d) {
        g.L[d] = m->L[d];
        // cells of at least rl/2 (a 1e-9 relative margin keeps rounding from
        // putting a pair within rl three cells apart), searched 5 x 5 x 5
        g.nc[d] = std::max(1, static_cast<int>(std::floor(2.0 * m->L[d] / (m->rl * (1.0 + 1e-9)))));
        g.w[d] = m->L[d] / static_cast<double>(g.nc[d]);
    }
    g.usable = (g.nc[0] >= 5 && g.nc[1] >= 5 && g.nc[2] >= 5) ? 1 : 0;
    const long ncell = g.usable ? static_cast<long>(g.nc[0]) * g.nc[1] * g.nc[2] : 1;
    if (ncell > m->ncell_alloc) {
        MHIP(m, malloc_dev(m->cell_count, ncell));
        MHIP(m, malloc_dev(m->cell_start, ncell + 1));
        MHIP(m, malloc_dev(m->cell_fill, ncell));
        m->ncell_alloc = ncell;
    }
    if (m->cap == 0) {
        const double rho = static_cast<double>(m->n) / (m->L[0] * m->L[1] * m->L[2]);
        long cap = static_cast<long>(std::ceil(2.0 * 4.0 / 3.0 * kPi * m->rl * m->rl * m->rl * rho)) + 32;
        m->cap = static_cast<int>(std::min<long>(std::max<long>(cap, 8), std::max<long>(m->n - 1, 1)));
    }
    const long need = static_cast<long>(m->cap) * ((m->n + 63) / 64 * 64);
    if (need > m->nbr_alloc) {
        MHIP(m, malloc_dev(m->nbr, need));
        m->nbr_alloc = need;
    }
    m->list_ok = false;
    return MD_OK;
}

// Builds the list on positions p (upload order); ref receives p when given.
md_status build_list(md_mix* m, const double4* p, double4* ref) {
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (m->grid.usable)
            mdk::launch_bin_index(m->n, p, m->grid, m->cell_of, m->cell_count, m->cell_start,
                                  m->cell_fill, m->cell_tmp, m->cell_atoms, nullptr, m->stream);
        mdk::MixListArgs a{};
        a.n = m->n;
        a.grid = m->grid;
        a.box = m->box;
        a.rl2 = m->rl * m->rl;
        a.pos = p;
        a.ref = ref;
        a.cell_of = m->cell_of;
        a.cell_start = m->cell_start;
        a.cell_atoms = m->cell_atoms;
        a.nbr = m->nbr;
        a.ncnt = m->ncnt;
        a.cap = m->cap;
        a.sc = m->sc;
        mdk::launch_mix_list(a, m->stream);
        MHIP(m, hipGetLastError());
        int ov = 0;
        MHIP(m, hipMemcpyAsync(&ov, &m->sc->overflow, sizeof(int), hipMemcpyDeviceToHost, m->stream));
        MHIP(m, hipStreamSynchronize(m->stream));
        if (!ov) {
            m->builds += 1;
            return MD_OK;
        }
        // grow the rows and rebuild
        const int zero = 0;
        MHIP(m, hipMemcpy(&m->sc->overflow, &zero, sizeof(int), hipMemcpyHostToDevice));
        m->cap = static_cast<int>(std::min<long>(2L * m->cap, std::max<long>(m->n - 1, 1)));
        const long need = static_cast<long>(m->cap) * ((m->n + 63) / 64 * 64);
        if (need > m->nbr_alloc) {
            MHIP(m, malloc_dev(m->nbr, need));
            m->nbr_alloc = need;
        }
    }
    return mfail(m, MD_ERR_CAPACITY, "neighbour row capacity could not be met");
}

mdk::MixArgs margs(md_mix* m, const double4* p) {
    mdk::MixArgs a{};
    a.n = m->n;
    a.box = m->box;
    a.tab = m->tab;
    a.pos = p;
    a.frc = m->frc;
    a.nbr = m->nbr;
    a.ncnt = m->ncnt;
    a.cap = m->cap;
    a.part_ew = m->part_ew;
    a.part_q = m->part_q;
    return a;
}

md_status ensure_list(md_mix* m) {
    if (m->list_ok) return MD_OK;
    MTRY(build_list(m, m->pos, m->ref));
    m->list_ok = true;
    return MD_OK;
}

// Forces on m->pos; E and W into e_dev[0], e_dev[1] (device).
md_status force_eval(md_mix* m, int deriv, double* e_dev) {
    mdk::launch_mix_force(margs(m, m->pos), deriv, m->stream);
    mdk::launch_mix_sum_ew(m->n, m->part_ew, e_dev, e_dev + 1, m->stream);
    MHIP(m, hipGetLastError());
    return MD_OK;
}

mdk::StepArgs step_args(md_mix* m, double dt) {
    mdk::StepArgs a{};
    a.n = m->n;
    a.box = m->box;
    a.pos = m->pos;
    a.vel = m->vel;
    a.frc = m->frc;
    a.ref = m->ref;
    a.kind = MD_VELOCITY_VERLET;
    a.dt = dt;
    a.skin = m->skin;
    a.check = 1;
    a.part_d = m->part_d;
    a.sc = m->sc;
    return a;
}

// Rebuilds the list when the last displacement check asked for it.
md_status rebuild_if_flagged(md_mix* m) {
    MHIP(m, hipMemcpyAsync(m->flag_host, &m->sc->rebuild_flag, sizeof(int), hipMemcpyDeviceToHost,
                           m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    if (*m->flag_host) {
        MTRY(build_list(m, m->pos, m->ref));
        m->list_ok = true;
    }
    return MD_OK;
}

// One velocity-Verlet step (kick, drift, displacement check, rebuild if
// needed, forces, kick); E and W of the new frame into e_dev.
md_status vv_step(md_mix* m, double dt, double* e_dev) {
    mdk::launch_pre_step(step_args(m, dt), m->stream);
    MTRY(rebuild_if_flagged(m));
    MTRY(force_eval(m, 0, e_dev));
    mdk::launch_half_kick(m->n, m->vel, m->frc, 0.5 * dt, m->stream);
    MHIP(m, hipGetLastError());
    return MD_OK;
}

md_status ensure_hist(md_mix* m, long frames) {
    if (2 * frames > m->hist_alloc) {
        MHIP(m, malloc_dev(m->hist, 2 * frames));
        m->hist_alloc = 2 * frames;
    }
    return MD_OK;
}

md_status copy_out(md_mix* m, long frames, double* e, double* w) {
    std::vector<double> h(static_cast<size_t>(2 * frames));
    MHIP(m, hipMemcpyAsync(h.data(), m->hist, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                           m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    for (long t = 0; t < frames; ++t) {
        if (e) e[t] = h[static_cast<size_t>(2 * t)];
        if (w) w[t] = h[static_cast<size_t>(2 * t + 1)];
    }
    return MD_OK;
}

md_status upload_vec3(md_mix* m, double4* dst, const double* x, const double* y, const double* z) {
    std::vector<double4> h(static_cast<size_t>(m->n));
    for (long i = 0; i < m->n; ++i)
        h[static_cast<size_t>(i)] = make_double4(x ? x[i] : 0.0, y ? y[i] : 0.0, z ? z[i] : 0.0, 0.0);
    MHIP(m, hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double4), hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return MD_OK;
}

md_status download_vec3(md_mix* m, const double4* src, double* x, double* y, double* z) {
    if (!x && !y && !z) return MD_OK;
    std::vector<double4> h(static_cast<size_t>(m->n));
    MHIP(m, hipMemcpyAsync(h.data(), src, h.size() * sizeof(double4), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    for (long i = 0; i < m->n; ++i) {
        const double4& v = h[static_cast<size_t>(i)];
        if (x) x[i] = v.x;
        if (y) y[i] = v.y;
        if (z) z[i] = v.z;
    }
    return MD_OK;
}

// Adjoint pass on positions p with cotangent m->vec; adds to m->lx and
// accumulates the parameter cotangent into m->tables (kMixG rows).
md_status adjoint_frame(md_mix* m, const double4* p, double wE, double wW) {
    MTRY(build_list(m, p, nullptr));
    mdk::MixArgs a = margs(m, p);
    a.vec = m->vec;
    a.out = m->lx;
    a.w_energy = wE;
    a.w_virial = wW;
    a.accumulate = 1;
    mdk::launch_mix_adjoint(a, m->stream);
    mdk::launch_mix_tables(m->n, p, m->nt, mdk::kMixG, m->part_q, m->tables, 1, m->stream);
    MHIP(m, hipGetLastError());
    return MD_OK;
}

}  // namespace

extern "C" {

md_status md_mix_create(void* hip_stream, md_mix** out) {
    if (out == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    md_mix* m = new (std::nothrow) md_mix();
    if (m == nullptr) return MD_ERR_RUNTIME;
    *out = m;
    if (hip_stream) {
        m->stream = static_cast<hipStream_t>(hip_stream);
    } else {
        MHIP(m, hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        m->own_stream = true;
    }
    MHIP(m, hipMalloc(reinterpret_cast<void**>(&m->sc), sizeof(mdk::Scalars)));
    mdk::Scalars s{};
    s.vscale = 1.0;
    MHIP(m, hipMemcpy(m->sc, &s, sizeof(s), hipMemcpyHostToDevice));
    MHIP(m, hipHostMalloc(reinterpret_cast<void**>(&m->flag_host), sizeof(int)));
    MHIP(m, malloc_dev(m->gamma, 1));
    MHIP(m, malloc_dev(m->ew, 2));
    MHIP(m, malloc_dev(m->tables, mdk::kTypePairs * mdk::kMixQ));
    return MD_OK;
}

md_status md_mix_destroy(md_mix* m) {
    if (m == nullptr) return MD_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_dev(m->pos); free_dev(m->vel); free_dev(m->frc); free_dev(m->ref);
    free_dev(m->nbr); free_dev(m->ncnt);
    free_dev(m->cell_of); free_dev(m->cell_count); free_dev(m->cell_start); free_dev(m->cell_fill);
    free_dev(m->cell_atoms);
    free_dev(m->cell_tmp);
    free_dev(m->part_ew); free_dev(m->part_q); free_dev(m->part_d); free_dev(m->gamma);
    free_dev(m->hist); free_dev(m->tables); free_dev(m->ew); free_dev(m->sc);
    free_dev(m->ckpt); free_dev(m->frames); free_dev(m->lx); free_dev(m->lv); free_dev(m->lu);
    free_dev(m->vec); free_dev(m->fin);
    if (m->flag_host) (void)hipHostFree(m->flag_host);
    if (m->own_stream && m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return MD_OK;
}

const char* md_mix_last_error(const md_mix* m) { return m ? m->err.c_str() : "null handle"; }

md_status md_mix_upload(md_mix* m, long n, const double* L, const double* x, const double* y,
                        const double* z, const double* vx, const double* vy, const double* vz,
                        const double* mass, const int* type) {
    if (m == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (n <= 1 || n >= 0x7fffffffL) return mfail(m, MD_ERR_INVALID_ARGUMENT, "atom count out of range");
    if (L == nullptr || x == nullptr || y == nullptr || z == nullptr || type == nullptr)
        return mfail(m, MD_ERR_INVALID_ARGUMENT, "box, positions and types are required");
    for (int d = 0; d < 3; ++d)
        if (!(L[d] > 0.0)) return mfail(m, MD_ERR_INVALID_ARGUMENT, "box lengths must be positive");
    for (long i = 0; i < n; ++i) {
        if (type[i] < 0 || type[i] >= mdk::kMaxTypes)
            return mfail(m, MD_ERR_INVALID_ARGUMENT, "atom type out of range (0..3)");
        if (mass && !(mass[i] > 0.0)) return mfail(m, MD_ERR_INVALID_ARGUMENT, "masses must be positive");
    }
    MHIP(m, hipStreamSynchronize(m->stream));
    m->n = n;
    for (int d = 0; d < 3; ++d) m->L[d] = L[d];
    m->box = mdk::make_box(L);
    MHIP(m, malloc_dev(m->pos, n)); MHIP(m, malloc_dev(m->vel, n)); MHIP(m, malloc_dev(m->frc, n));
    MHIP(m, malloc_dev(m->ref, n)); MHIP(m, malloc_dev(m->ncnt, n));
    MHIP(m, malloc_dev(m->cell_of, n)); MHIP(m, malloc_dev(m->cell_atoms, n));
    MHIP(m, malloc_dev(m->cell_tmp, n));
    MHIP(m, malloc_dev(m->part_ew, 2 * n));
    MHIP(m, malloc_dev(m->part_q, static_cast<long>(mdk::kMaxTypes) * mdk::kMixQ * n));
    MHIP(m, malloc_dev(m->part_d, 4 * nblocks(n) + 8));
    MHIP(m, malloc_dev(m->lx, n)); MHIP(m, malloc_dev(m->lv, n)); MHIP(m, malloc_dev(m->lu, n));
    MHIP(m, malloc_dev(m->vec, n)); MHIP(m, malloc_dev(m->fin, 3 * n));
    std::vector<double4> p(static_cast<size_t>(n)), v(static_cast<size_t>(n));
    for (long i = 0; i < n; ++i) {
        p[static_cast<size_t>(i)] = make_double4(x[i], y[i], z[i], static_cast<double>(type[i]));
        v[static_cast<size_t>(i)] = make_double4(vx ? vx[i] : 0.0, vy ? vy[i] : 0.0, vz ? vz[i] : 0.0,
                                                 mass ? mass[i] : 1.0);
    }
    MHIP(m, hipMemcpy(m->pos, p.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    MHIP(m, hipMemcpy(m->ref, p.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    MHIP(m, hipMemcpy(m->vel, v.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    MHIP(m, hipMemset(m->frc, 0, n * sizeof(double4)));
    m->cap = 0;
    m->nbr_alloc = 0;
    free_dev(m->nbr);
    m->have_run = false;
    m->list_ok = false;
    if (m->have_params) MTRY(alloc_list(m));
    return MD_OK;
}

md_status md_mix_set_params(md_mix* m, int ntypes, const double* sigma, const double* epsilon,
                            const double* cutoff, int linear_force_shift) {
    if (m == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (ntypes < 1 || ntypes > mdk::kMaxTypes)
        return mfail(m, MD_ERR_INVALID_ARGUMENT, "ntypes must lie in 1..4");
    if (sigma == nullptr || epsilon == nullptr || cutoff == nullptr)
        return mfail(m, MD_ERR_INVALID_ARGUMENT, "sigma, epsilon and cutoff tables are required");
    for (int a = 0; a < ntypes; ++a)
        for (int b = 0; b < ntypes; ++b) {
            const int k = a * ntypes + b, kt = b * ntypes + a;
            if (!(sigma[k] > 0.0) || !(epsilon[k] > 0.0) || !(cutoff[k] > 0.0))
                return mfail(m, MD_ERR_INVALID_ARGUMENT, "sigma, epsilon and cutoff must be positive");
            if (sigma[k] != sigma[kt] || epsilon[k] != epsilon[kt] || cutoff[k] != cutoff[kt])
                return mfail(m, MD_ERR_INVALID_ARGUMENT, "parameter tables must be symmetric");
        }
    m->nt = ntypes;
    m->lfs = linear_force_shift ? 1 : 0;
    for (int k = 0; k < ntypes * ntypes; ++k) {
        m->sigma[k] = sigma[k];
        m->epsilon[k] = epsilon[k];
        m->cutoff[k] = cutoff[k];
    }
    make_tab(m);
    m->have_params = true;
    m->have_run = false;
    if (m->n > 0) MTRY(alloc_list(m));
    return MD_OK;
}

md_status md_mix_set_skin(md_mix* m, double skin) {
    if (m == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (!(skin >= 0.0)) return mfail(m, MD_ERR_INVALID_ARGUMENT, "skin must be non-negative");
    m->skin = skin;
    m->cap = 0;
    if (m->n > 0 && m->have_params) MTRY(alloc_list(m));
    return MD_OK;
}

md_status md_mix_forces(md_mix* m, double* energy, double* virial, double* tables) {
    MTRY(need_system(m));
    MTRY(ensure_list(m));
    MTRY(force_eval(m, tables != nullptr ? 1 : 0, m->ew));
    double ew[2];
    MHIP(m, hipMemcpyAsync(ew, m->ew, sizeof(ew), hipMemcpyDeviceToHost, m->stream));
    if (tables) {
        mdk::launch_mix_tables(m->n, m->pos, m->nt, mdk::kMixQ, m->part_q, m->tables, 0, m->stream);
        MHIP(m, hipGetLastError());
        std::vector<double> t(static_cast<size_t>(m->nt * m->nt * mdk::kMixQ));
        MHIP(m, hipMemcpyAsync(t.data(), m->tables, t.size() * sizeof(double), hipMemcpyDeviceToHost,
                               m->stream));
        MHIP(m, hipStreamSynchronize(m->stream));
        canonical(t, m->nt, mdk::kMixQ, tables);
    }
    MHIP(m, hipStreamSynchronize(m->stream));
    if (energy) *energy = ew[0];
    if (virial) *virial = ew[1];
    return MD_OK;
}

md_status md_mix_force_vjp(md_mix* m, double w_energy, double w_virial, const double* gx,
                           const double* gy, const double* gz, double* ox, double* oy, double* oz,
                           double* dtheta) {
    MTRY(need_system(m));
    MTRY(ensure_list(m));
    MTRY(upload_vec3(m, m->vec, gx, gy, gz));
    mdk::MixArgs a = margs(m, m->pos);
    a.vec = m->vec;
    a.out = m->lx;
    a.w_energy = w_energy;
    a.w_virial = w_virial;
    a.accumulate = 0;
    mdk::launch_mix_adjoint(a, m->stream);
    mdk::launch_mix_tables(m->n, m->pos, m->nt, mdk::kMixG, m->part_q, m->tables, 0, m->stream);
    MHIP(m, hipGetLastError());
    MTRY(download_vec3(m, m->lx, ox, oy, oz));
    if (dtheta) {
        std::vector<double> t(static_cast<size_t>(m->nt * m->nt * mdk::kMixG));
        MHIP(m, hipMemcpyAsync(t.data(), m->tables, t.size() * sizeof(double), hipMemcpyDeviceToHost,
                               m->stream));
        MHIP(m, hipStreamSynchronize(m->stream));
        canonical(t, m->nt, mdk::kMixG, dtheta);
    }
    return MD_OK;
}

md_status md_mix_minimize(md_mix* m, int nsteps, double step, double max_disp, double* energies) {
    MTRY(need_system(m));
    if (nsteps < 0 || !(step > 0.0) || !(max_disp > 0.0))
        return mfail(m, MD_ERR_INVALID_ARGUMENT, "minimize: nsteps >= 0, step > 0 and max_disp > 0");
    MTRY(ensure_hist(m, nsteps + 1));
    MTRY(ensure_list(m));
    MTRY(force_eval(m, 0, m->hist));
    for (int k = 0; k < nsteps; ++k) {
        mdk::launch_mix_sd(m->n, m->box, m->pos, m->frc, step, max_disp, m->part_d, m->gamma, m->stream);
        mdk::StepArgs sa = step_args(m, 0.0);
        mdk::launch_check(sa, m->stream);
        MHIP(m, hipGetLastError());
        MTRY(rebuild_if_flagged(m));
        MTRY(force_eval(m, 0, m->hist + 2 * (k + 1)));
    }
    m->have_run = false;
    return copy_out(m, nsteps + 1, energies, nullptr);
}

md_status md_mix_run(md_mix* m, double dt, int nsteps, int checkpoint_every, double* energies,
                     double* virials) {
    MTRY(need_system(m));
    if (nsteps < 0 || !(dt > 0.0) || checkpoint_every < 1)
        return mfail(m, MD_ERR_INVALID_ARGUMENT, "run: nsteps >= 0, dt > 0, checkpoint_every >= 1");
    const long n = m->n;
    const long ncp = nsteps / checkpoint_every + 1;
    if (2 * ncp * n > m->ckpt_alloc) {
        MHIP(m, malloc_dev(m->ckpt, 2 * ncp * n));
        m->ckpt_alloc = 2 * ncp * n;
    }
    MTRY(ensure_hist(m, nsteps + 1));
    for (int t = 0; t <= nsteps; ++t) {
        if (t % checkpoint_every == 0) {
            // checkpoint: fresh list and forces, so a recomputation from here is exact
            const long s = t / checkpoint_every;
            MHIP(m, hipMemcpyAsync(m->ckpt + 2 * s * n, m->pos, n * sizeof(double4),
                                   hipMemcpyDeviceToDevice, m->stream));
            MHIP(m, hipMemcpyAsync(m->ckpt + (2 * s + 1) * n, m->vel, n * sizeof(double4),
                                   hipMemcpyDeviceToDevice, m->stream));
            MTRY(build_list(m, m->pos, m->ref));
            m->list_ok = true;
            MTRY(force_eval(m, 0, m->hist + 2 * t));
        }
        if (t == nsteps) break;
        MTRY(vv_step(m, dt, m->hist + 2 * (t + 1)));
    }
    MHIP(m, hipMemcpyAsync(m->fin, m->pos, n * sizeof(double4), hipMemcpyDeviceToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->fin + n, m->vel, n * sizeof(double4), hipMemcpyDeviceToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->fin + 2 * n, m->frc, n * sizeof(double4), hipMemcpyDeviceToDevice, m->stream));
    m->T = nsteps;
    m->K = checkpoint_every;
    m->dt = dt;
    m->have_run = true;
    return copy_out(m, nsteps + 1, energies, virials);
}

md_status md_mix_adjoint(md_mix* m, double w_energy, double w_virial, const double* gx,
                         const double* gy, const double* gz, const double* gvx, const double* gvy,
                         const double* gvz, double* dtheta, double* dx, double* dy, double* dz,
                         double* dvx, double* dvy, double* dvz) {
    MTRY(need_system(m));
    if (!m->have_run) return mfail(m, MD_ERR_STATE, "adjoint needs a preceding md_mix_run");
    const long n = m->n;
    const int T = m->T, K = m->K;
    const double dt = m->dt;
    const long fr = std::min(K, std::max(T, 1)) + 1;
    if (fr * n > m->frames_alloc) {
        MHIP(m, malloc_dev(m->frames, fr * n));
        m->frames_alloc = fr * n;
    }
    MTRY(ensure_hist(m, T + 1));
    MTRY(upload_vec3(m, m->lx, gx, gy, gz));
    MTRY(upload_vec3(m, m->lv, gvx, gvy, gvz));
    MHIP(m, hipMemsetAsync(m->lu, 0, n * sizeof(double4), m->stream));
    MHIP(m, hipMemsetAsync(m->tables, 0, mdk::kTypePairs * mdk::kMixG * sizeof(double), m->stream));
    const int nseg = T > 0 ? (T + K - 1) / K : 0;
    if (T == 0) {
        // a single frame: only the direct terms and the frame-0 vector (lu = 0)
        mdk::launch_mix_vec(n, m->vel, m->lv, m->lu, 0.5 * dt, 1, m->vec, m->stream);
        MTRY(adjoint_frame(m, m->pos, w_energy, w_virial));
    }
    for (int s = nseg - 1; s >= 0; --s) {
        const int t0 = s * K, t1 = std::min(T, t0 + K);
        // recompute frames t0..t1 from checkpoint s exactly as the forward run did
        MHIP(m, hipMemcpyAsync(m->pos, m->ckpt + 2L * s * n, n * sizeof(double4),
                               hipMemcpyDeviceToDevice, m->stream));
        MHIP(m, hipMemcpyAsync(m->vel, m->ckpt + (2L * s + 1) * n, n * sizeof(double4),
                               hipMemcpyDeviceToDevice, m->stream));
        MTRY(build_list(m, m->pos, m->ref));
        MTRY(force_eval(m, 0, m->ew));
        MHIP(m, hipMemcpyAsync(m->frames, m->pos, n * sizeof(double4), hipMemcpyDeviceToDevice, m->stream));
        for (int t = t0; t < t1; ++t) {
            MTRY(vv_step(m, dt, m->ew));
            MHIP(m, hipMemcpyAsync(m->frames + static_cast<long>(t + 1 - t0) * n, m->pos,
                                   n * sizeof(double4), hipMemcpyDeviceToDevice, m->stream));
        }
        // backward sweep over frames t1 .. t0+1 (and frame 0 in the first segment)
        for (int t = t1; t > t0; --t) {
            mdk::launch_mix_vec(n, m->vel, m->lv, m->lu, 0.5 * dt, 0, m->vec, m->stream);
            MTRY(adjoint_frame(m, m->frames + static_cast<long>(t - t0) * n, w_energy, w_virial));
            mdk::launch_mix_back(n, dt, m->lx, m->lv, m->lu, m->stream);
            MHIP(m, hipGetLastError());
        }
        if (s == 0) {
            mdk::launch_mix_vec(n, m->vel, m->lv, m->lu, 0.5 * dt, 1, m->vec, m->stream);
            MTRY(adjoint_frame(m, m->frames, w_energy, w_virial));
        }
    }
    if (dtheta) {
        std::vector<double> t(static_cast<size_t>(m->nt * m->nt * mdk::kMixG));
        MHIP(m, hipMemcpyAsync(t.data(), m->tables, t.size() * sizeof(double), hipMemcpyDeviceToHost,
                               m->stream));
        MHIP(m, hipStreamSynchronize(m->stream));
        canonical(t, m->nt, mdk::kMixG, dtheta);
    }
    MTRY(download_vec3(m, m->lx, dx, dy, dz));
    MTRY(download_vec3(m, m->lv, dvx, dvy, dvz));
    // back to the final state of the run
    MHIP(m, hipMemcpyAsync(m->pos, m->fin, n * sizeof(double4), hipMemcpyDeviceToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->vel, m->fin + n, n * sizeof(double4), hipMemcpyDeviceToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->frc, m->fin + 2 * n, n * sizeof(double4), hipMemcpyDeviceToDevice, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    m->list_ok = false;
    return MD_OK;
}

md_status md_mix_download(md_mix* m, double* x, double* y, double* z, double* vx, double* vy,
                          double* vz, double* fx, double* fy, double* fz) {
    if (m == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (m->n <= 0) return mfail(m, MD_ERR_STATE, "no system uploaded");
    MTRY(download_vec3(m, m->pos, x, y, z));
    MTRY(download_vec3(m, m->vel, vx, vy, vz));
    MTRY(download_vec3(m, m->frc, fx, fy, fz));
    return MD_OK;
}

md_status md_mix_set_state(md_mix* m, const double* x, const double* y, const double* z,
                           const double* vx, const double* vy, const double* vz) {
    if (m == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (m->n <= 0) return mfail(m, MD_ERR_STATE, "no system uploaded");
    const long n = m->n;
    std::vector<double4> p(static_cast<size_t>(n)), v(static_cast<size_t>(n));
    MHIP(m, hipMemcpy(p.data(), m->pos, n * sizeof(double4), hipMemcpyDeviceToHost));
    MHIP(m, hipMemcpy(v.data(), m->vel, n * sizeof(double4), hipMemcpyDeviceToHost));
    for (long i = 0; i < n; ++i) {
        double4& a = p[static_cast<size_t>(i)];
        double4& b = v[static_cast<size_t>(i)];
        if (x) a.x = x[i];
        if (y) a.y = y[i];
        if (z) a.z = z[i];
        if (vx) b.x = vx[i];
        if (vy) b.y = vy[i];
        if (vz) b.z = vz[i];
    }
    MHIP(m, hipMemcpy(m->pos, p.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    MHIP(m, hipMemcpy(m->vel, v.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    m->list_ok = false;
    m->have_run = false;
    return MD_OK;
}

md_status md_mix_list_builds(const md_mix* m, long* builds) {
    if (m == nullptr || builds == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *builds = m->builds;
    return MD_OK;
}

// Uniform random positions (u - 1/2) L per coordinate, atom by atom (the
// recipe of tests/test_analysis.cpp:24-28), then Maxwell-Boltzmann
// velocities from the same stream (util/random.hpp:37-76: mt19937_64,
// u = 1 - U[0,1), Marsaglia polar normals with a cached spare), each
// component normal() * sqrt(T/m), and the centre-of-mass motion removed
// (core/system.cpp:147-164).
md_status md_mix_random_state(long n, double box_length, uint64_t seed, double temperature,
                              const double* mass, double* x, double* y, double* z, double* vx,
                              double* vy, double* vz) {
#pragma clang fp contract(off)
    if (n < 1 || !(box_length > 0.0) || !(temperature >= 0.0) || x == nullptr || y == nullptr ||
        z == nullptr || vx == nullptr || vy == nullptr || vz == nullptr)
        return MD_ERR_INVALID_ARGUMENT;
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    bool have_spare = false;
    double spare = 0.0;
    auto uniform = [&]() { return 1.0 - unif(eng); };
    auto normal = [&]() {
        if (have_spare) {
            have_spare = false;
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
        have_spare = true;
        return a * k;
    };
    for (long i = 0; i < n; ++i) {
        x[i] = (uniform() - 0.5) * box_length;
        y[i] = (uniform() - 0.5) * box_length;
        z[i] = (uniform() - 0.5) * box_length;
    }
    double mtot = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    for (long i = 0; i < n; ++i) {
        const double mi = mass ? mass[i] : 1.0;
        if (!(mi > 0.0)) return MD_ERR_INVALID_ARGUMENT;
        const double sv = std::sqrt(temperature / mi);
        const double g0 = normal();
        const double g1 = normal();
        const double g2 = normal();
        vx[i] = sv * g0;
        vy[i] = sv * g1;
        vz[i] = sv * g2;
        mtot += mi;
        px += mi * vx[i];
        py += mi * vy[i];
        pz += mi * vz[i];
    }
    const double cx = px / mtot, cy = py / mtot, cz = pz / mtot;
    for (long i = 0; i < n; ++i) {
        vx[i] -= cx;
        vy[i] -= cy;
        vz[i] -= cz;
    }
    return MD_OK;
}

}  // extern "C"
